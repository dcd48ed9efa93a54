// awp_kernels.hip -- CDNA4 kernels of the AWP-ODC-SGT time step.
//
// Every arithmetic statement follows the reference CPU solver's operation
// order (AWP-ODC-SGT/src/operator.f, viscoop.f, cerjan.f, source.f,
// pmcl3d.f) and the library is compiled with -ffp-contract=off, so each
// FP32 result is the IEEE result of the same sequence of +,-,*,/ as the
// reference.  No MFMA: the stencil is not a contraction.
//
// Kernel shape: a block is one wave64 row in x times tile_y rows in y and
// marches a chunk of z levels bottom-up.  Values a point needs from the
// levels below/above its own are carried in registers (z queues), so each
// level of each array is fetched once per column; x and y neighbours come
// from the lines the neighbouring lanes and rows of the block just loaded.
//
// The stress kernel is a template on the storage policy (StoreF32 /
// StoreMixed); the mixed variant differs only in how r1..r6 and qp/qs are
// loaded and stored and in the exact power-of-two rescaling of qp/qs.

#include "awp_kernels.hpp"

#include <algorithm>

namespace awp {

// ---------------------------------------------------------------------------
// velocity: uxx1/vyy1/wzz1 (operator.f:92-244) fused with addcrjvel
// (cerjan.f:292-328)
// ---------------------------------------------------------------------------
template <class P, int U>
__device__ __forceinline__ void velocity_tile(const Geo &g, const State &s0, const Coef &c, const int kc, const int bx,
                                     const int by, const int bz)
{
    const GState s = as_global(s0);
    using M = typename P::M;
    const int i = (bx * blockDim.z + threadIdx.z) * TX + threadIdx.x + 1;
    const int j = by * blockDim.y + threadIdx.y + 1;
    const int kb = bz * kc + 1;
    const int ke = min(kb + kc - 1, g.nz);
    if (i > g.nx || j > g.ny) return;
    const long long sy = g.sy, sz = g.sz;
    long long o = g.base + (long long)kb * sz + (long long)j * sy + i;

    const G<float> *__restrict__ xx = s.xx;
    const G<float> *__restrict__ yy = s.yy;
    const G<float> *__restrict__ zz = s.zz;
    const G<float> *__restrict__ xy = s.xy;
    const G<float> *__restrict__ xz = s.xz;
    const G<float> *__restrict__ yz = s.yz;
    const G<M> *__restrict__ d1 = (const G<M> *)s.d1;
    const float c1 = c.c1, c2 = c.c2, dth = c.dth;
    const float fxy = s.fx[i] * s.fy[j];

    // z queues at the column (i, j)
    float xz_m2 = xz[o - 2 * sz], xz_m1 = xz[o - sz], xz_0 = xz[o];
    float yz_m2 = yz[o - 2 * sz], yz_m1 = yz[o - sz], yz_0 = yz[o];
    float zz_m1 = zz[o - sz], zz_0 = zz[o], zz_p1 = zz[o + sz];
    // density at level k-1: (i,j), (i,j-1), (i+1,j)
    const float db_m = P::mixed ? s.mbar[mbar_idx(g, 0, kb - 1)] : 0.0f;
    float dA_m = mdec(d1[o - sz], db_m), dB_m = mdec(d1[o - sz - sy], db_m), dC_m = mdec(d1[o - sz + 1], db_m);

#pragma unroll U
    for (int k = kb; k <= ke; ++k, o += sz) {
        const float xz_p1 = xz[o + sz];
        const float yz_p1 = yz[o + sz];
        const float zz_p2 = nt_ld(&zz[o + 2 * sz]);   // zz: read at (i, j) only
        const float db = P::mixed ? s.mbar[mbar_idx(g, 0, k)] : 0.0f;
        const float dA = mdec(d1[o], db), dB = mdec(d1[o - sy], db), dC = mdec(d1[o + 1], db),
                    dD = mdec(d1[o + 1 - sy], db);

        const float xx_0 = xx[o], xx_m1 = xx[o - 1], xx_p1 = xx[o + 1], xx_m2 = xx[o - 2];
        const float xy_0 = xy[o], xy_jm1 = xy[o - sy], xy_jp1 = xy[o + sy], xy_jm2 = xy[o - 2 * sy];
        const float xy_ip1 = xy[o + 1], xy_ip2 = xy[o + 2], xy_im1 = xy[o - 1];
        const float yy_0 = yy[o], yy_jp1 = yy[o + sy], yy_jp2 = yy[o + 2 * sy], yy_jm1 = yy[o - sy];
        const float xz_ip1 = xz[o + 1], xz_ip2 = xz[o + 2], xz_im1 = xz[o - 1];
        const float yz_jm1 = yz[o - sy], yz_jp1 = yz[o + sy], yz_jm2 = yz[o - 2 * sy];

        float d, q, u, v, w;
        d = 0.25f * ((dA + dB) + (dA_m + dB_m));
        q = dth / d;
        u = nt_ld(&s.u1[o]) + q * (
            c1 * (xx_0 - xx_m1) + c2 * (xx_p1 - xx_m2) +
            c1 * (xy_0 - xy_jm1) + c2 * (xy_jp1 - xy_jm2) +
            c1 * (xz_0 - xz_m1) + c2 * (xz_p1 - xz_m2));
        d = 0.25f * ((dA + dC) + (dA_m + dC_m));
        q = dth / d;
        v = nt_ld(&s.v1[o]) + q * (
            c1 * (xy_ip1 - xy_0) + c2 * (xy_ip2 - xy_im1) +
            c1 * (yy_jp1 - yy_0) + c2 * (yy_jp2 - yy_jm1) +
            c1 * (yz_0 - yz_m1) + c2 * (yz_p1 - yz_m2));
        d = 0.25f * ((dA + dC) + (dB + dD));
        q = dth / d;
        w = nt_ld(&s.w1[o]) + q * (
            c1 * (xz_ip1 - xz_0) + c2 * (xz_ip2 - xz_im1) +
            c1 * (yz_0 - yz_jm1) + c2 * (yz_jp1 - yz_jm2) +
            c1 * (zz_p1 - zz_0) + c2 * (zz_p2 - zz_m1));
        const float D = fxy * s.fz[k];
        nt_st(&s.u1[o], u * D);
        nt_st(&s.v1[o], v * D);
        nt_st(&s.w1[o], w * D);

        xz_m2 = xz_m1; xz_m1 = xz_0; xz_0 = xz_p1;
        yz_m2 = yz_m1; yz_m1 = yz_0; yz_0 = yz_p1;
        zz_m1 = zz_0; zz_0 = zz_p1; zz_p1 = zz_p2;
        dA_m = dA; dB_m = dB; dC_m = dC;
    }
}

// The same statements through a 32-bit byte offset (Cursor32: uniform array
// bases, one offset register, neighbours as immediates); needs every array
// below 4 GiB.
template <class P, int U>
__device__ __forceinline__ void velocity_tile_b(const Geo &g, const State &s0, const Coef &c, const int kc, const int bx,
                                       const int by, const int bz)
{
    const GState s = as_global(s0);
    const int i = (bx * blockDim.z + threadIdx.z) * TX + threadIdx.x + 1;
    const int j = by * blockDim.y + threadIdx.y + 1;
    const int kb = bz * kc + 1;
    const int ke = min(kb + kc - 1, g.nz);
    if (i > g.nx || j > g.ny) return;
    Cursor32 p(g, g.base + (long long)kb * g.sz + (long long)j * g.sy + i);

    const G<float> *__restrict__ xx = s.xx;
    const G<float> *__restrict__ yy = s.yy;
    const G<float> *__restrict__ zz = s.zz;
    const G<float> *__restrict__ xy = s.xy;
    const G<float> *__restrict__ xz = s.xz;
    const G<float> *__restrict__ yz = s.yz;
    using M = typename P::M;
    const G<M> *__restrict__ d1 = (const G<M> *)s.d1;
    const float c1 = c.c1, c2 = c.c2, dth = c.dth;
    const float fxy = s.fx[i] * s.fy[j];

    float xz_m2 = p.ld(xz, 0, 0, -2), xz_m1 = p.ld(xz, 0, 0, -1), xz_0 = p.ld(xz);
    float yz_m2 = p.ld(yz, 0, 0, -2), yz_m1 = p.ld(yz, 0, 0, -1), yz_0 = p.ld(yz);
    float zz_m1 = p.ld(zz, 0, 0, -1), zz_0 = p.ld(zz), zz_p1 = p.ld(zz, 0, 0, 1);
    const float db_m = P::mixed ? s.mbar[mbar_idx(g, 0, kb - 1)] : 0.0f;
    float dA_m = mdec(p.ld(d1, 0, 0, -1), db_m), dB_m = mdec(p.ld(d1, 0, -1, -1), db_m),
          dC_m = mdec(p.ld(d1, 1, 0, -1), db_m);

#pragma unroll U
    for (int k = kb; k <= ke; ++k, p.up()) {
        const float xz_p1 = p.ld(xz, 0, 0, 1);
        const float yz_p1 = p.ld(yz, 0, 0, 1);
        const float zz_p2 = gln(zz, p.at(0, 2));   // zz: read at (i, j) only
        const float db = P::mixed ? s.mbar[mbar_idx(g, 0, k)] : 0.0f;
        const float dA = mdec(p.ld(d1), db), dB = mdec(p.ld(d1, 0, -1), db), dC = mdec(p.ld(d1, 1), db),
                    dD = mdec(p.ld(d1, 1, -1), db);

        const float xx_0 = p.ld(xx), xx_m1 = p.ld(xx, -1), xx_p1 = p.ld(xx, 1), xx_m2 = p.ld(xx, -2);
        const float xy_0 = p.ld(xy), xy_jm1 = p.ld(xy, 0, -1), xy_jp1 = p.ld(xy, 0, 1), xy_jm2 = p.ld(xy, 0, -2);
        const float xy_ip1 = p.ld(xy, 1), xy_ip2 = p.ld(xy, 2), xy_im1 = p.ld(xy, -1);
        const float yy_0 = p.ld(yy), yy_jp1 = p.ld(yy, 0, 1), yy_jp2 = p.ld(yy, 0, 2), yy_jm1 = p.ld(yy, 0, -1);
        const float xz_ip1 = p.ld(xz, 1), xz_ip2 = p.ld(xz, 2), xz_im1 = p.ld(xz, -1);
        const float yz_jm1 = p.ld(yz, 0, -1), yz_jp1 = p.ld(yz, 0, 1), yz_jm2 = p.ld(yz, 0, -2);

        float d, q, u, v, w;
        d = 0.25f * ((dA + dB) + (dA_m + dB_m));
        q = dth / d;
        u = p.ldn(s.u1) + q * (
            c1 * (xx_0 - xx_m1) + c2 * (xx_p1 - xx_m2) +
            c1 * (xy_0 - xy_jm1) + c2 * (xy_jp1 - xy_jm2) +
            c1 * (xz_0 - xz_m1) + c2 * (xz_p1 - xz_m2));
        d = 0.25f * ((dA + dC) + (dA_m + dC_m));
        q = dth / d;
        v = p.ldn(s.v1) + q * (
            c1 * (xy_ip1 - xy_0) + c2 * (xy_ip2 - xy_im1) +
            c1 * (yy_jp1 - yy_0) + c2 * (yy_jp2 - yy_jm1) +
            c1 * (yz_0 - yz_m1) + c2 * (yz_p1 - yz_m2));
        d = 0.25f * ((dA + dC) + (dB + dD));
        q = dth / d;
        w = p.ldn(s.w1) + q * (
            c1 * (xz_ip1 - xz_0) + c2 * (xz_ip2 - xz_im1) +
            c1 * (yz_0 - yz_jm1) + c2 * (yz_jp1 - yz_jm2) +
            c1 * (zz_p1 - zz_0) + c2 * (zz_p2 - zz_m1));
        const float D = fxy * s.fz[k];
        p.stn(s.u1, u * D);
        p.stn(s.v1, v * D);
        p.stn(s.w1, w * D);

        xz_m2 = xz_m1; xz_m1 = xz_0; xz_0 = xz_p1;
        yz_m2 = yz_m1; yz_m1 = yz_0; yz_0 = yz_p1;
        zz_m1 = zz_0; zz_0 = zz_p1; zz_p1 = zz_p2;
        dA_m = dA; dB_m = dB; dC_m = dC;
    }
}

// ---------------------------------------------------------------------------
// free-surface velocity images: fvelxy then fvelz (operator.f:246-306).
// u(i+1), v(j-1) at nz+1 are recomputed from the same inputs fvelxy uses,
// so no thread reads another thread's output.
// ---------------------------------------------------------------------------
template <class P>
__device__ __forceinline__ void fs_velocity_block(const Geo &g, const State &s0, const int bx, const int by)
{
    const GState s = as_global(s0);
    const int i = bx * TX + threadIdx.x + 1;
    const int j = by * blockDim.y + threadIdx.y + 1;
    if (i > g.nx || j > g.ny) return;
    const long long sy = g.sy, sz = g.sz;
    const long long p = gidx(g, i, j, g.nz);
    const long long q = p + sz;
    const G<float> *u1 = s.u1, *v1 = s.v1, *w1 = s.w1;
    const float w0 = w1[p];
    const float ut = u1[p] - (w0 - w1[p - 1]);
    const float vt = v1[p] - (w1[p + sy] - w0);
    const float ut_ip1 = (i < g.nx) ? (u1[p + 1] - (w1[p + 1] - w0)) : 0.0f;
    const float vt_jm1 = (j > 1) ? (v1[p - sy] - (w0 - w1[p - sy])) : 0.0f;
    using M = typename P::M;
    const float mb = P::mixed ? s.mbar[mbar_idx(g, 1, g.nz)] : 0.0f;
    const float lb = P::mixed ? s.mbar[mbar_idx(g, 2, g.nz)] : 0.0f;
    float xl, xm;
    if constexpr (P::mixed) {
        const unsigned w = ((const G<unsigned> *)s.mu)[p];   // (mu, lam) pair
        xl = 1.0f / mdec(h_hi(w), lb);
        xm = 2.0f / mdec(h_lo(w), mb);
    } else {
        xl = 1.0f / mdec(((const G<M> *)s.lam)[p], lb);
        xm = 2.0f / mdec(((const G<M> *)s.mu)[p], mb);
    }
    const float xl2m = xm + xl;
    const float wt = w1[p - sz] - (xl / xl2m) * (
        (ut_ip1 - ut) +
        (u1[p + 1] - u1[p]) +
        (vt - vt_jm1) +
        (v1[p] - v1[p - sy]));
    s.u1[q] = ut;
    s.v1[q] = vt;
    s.w1[q] = wt;
}

// ---------------------------------------------------------------------------
// stress + memory variables: allq (viscoop.f:49-191) fused with addcrjstr
// (cerjan.f:249-290) and, in the block that reaches k = nz, fstr
// (operator.f:575-612).
// ---------------------------------------------------------------------------
// Two forms of the same statements: stress_tile_e addresses the arrays with
// 64-bit element indices (any size), stress_tile_b through a 32-bit byte
// offset cursor (arrays below 4 GiB).  The compiler schedules them
// differently: the element-indexed form keeps more loads in flight (189
// registers, one 512-thread block per CU) and is the faster FP32 kernel; the
// byte-offset form needs 125 registers in mixed storage, where two blocks per
// CU hide the latency of the exponent round trip, and is the faster mixed
// kernel.
template <class P, int U>
__device__ __forceinline__ void stress_tile_e(const Geo &g, const State &s0, const Coef &c, const int kc, const int bx,
                                   const int by, const int bz)
{
    const GState s = as_global(s0);
    using Q = typename P::Q;
    using R = typename P::R;
    const int seg = bx * blockDim.z + threadIdx.z;   // the wave's 64-point row segment
    const int ir = seg * TX + threadIdx.x + 1;
    const int jr = by * blockDim.y + threadIdx.y + 1;
    const int kb = bz * kc + 1;
    const int ke = min(kb + kc - 1, g.nz);
    // mixed storage needs every lane of a wave for the row maximum of the
    // memory variables: lanes outside the domain compute an in-range
    // duplicate point and store nothing
    const bool act = ir <= g.nx && jr <= g.ny;
    if constexpr (!P::mixed) {
        if (!act) return;
    } else {
        // a wave whose whole row segment lies beyond nx (the second wave of
        // a 128-point tile at the grid's edge) has no point and no exponent
        if (seg * TX >= g.nx || jr > g.ny) return;
    }
    const int i = min(ir, g.nx), j = min(jr, g.ny);
    const long long sy = g.sy, sz = g.sz;
    long long o = g.base + (long long)kb * sz + (long long)j * sy + i;

    const G<float> *__restrict__ u1 = s.u1;
    const G<float> *__restrict__ v1 = s.v1;
    const G<float> *__restrict__ w1 = s.w1;
    using M = typename P::M;
    const G<M> *__restrict__ lam = (const G<M> *)s.lam;
    const G<M> *__restrict__ mu = (const G<M> *)s.mu;
    const G<unsigned> *__restrict__ mlq = (const G<unsigned> *)s.mu;   // mixed: (mu, lam) pairs
    const G<Q> *__restrict__ qp = (const G<Q> *)s.qp;
    const G<Q> *__restrict__ qs = (const G<Q> *)s.qs;
    const G<unsigned> *__restrict__ qq = (const G<unsigned> *)s.qp;   // mixed: (qp, qs) pairs
    G<R> *__restrict__ R1 = (G<R> *)s.r1;
    G<R> *__restrict__ R2 = (G<R> *)s.r2;
    G<R> *__restrict__ R3 = (G<R> *)s.r3;
    G<R> *__restrict__ R4 = (G<R> *)s.r4;
    G<R> *__restrict__ R5 = (G<R> *)s.r5;
    G<R> *__restrict__ R6 = (G<R> *)s.r6;
    const float c1 = c.c1, c2 = c.c2, dth = c.dth, dt = c.dt, dh1 = c.dh1;
    const float fxy = s.fx[i] * s.fy[j];
    // relaxation weights of this column for even / odd depth parity
    // (selected per level: no per-lane indexing of the argument block)
    const int itxy = ((i & 1) * 2 + (j & 1)) * 2;
    const float t1e = c.tau1[itxy], t1o = c.tau1[itxy + 1];
    const float t2e = c.tau2[itxy], t2o = c.tau2[itxy + 1];

    // z queues of the velocities at the column
    float u_m1 = u1[o - sz], u_0 = u1[o], u_p1 = u1[o + sz];
    float v_m1 = v1[o - sz], v_0 = v1[o], v_p1 = v1[o + sz];
    float w_m2 = w1[o - 2 * sz], w_m1 = w1[o - sz], w_0 = w1[o];
    // media at level k-1 on the 2x2 (x,y) cell: A=(i,j) B=(i+1,j) C=(i,j-1) D=(i+1,j-1)
    long long om = o - sz;
    const float mb_m = P::mixed ? s.mbar[mbar_idx(g, 1, kb - 1)] : 0.0f;
    const float lb_m = P::mixed ? s.mbar[mbar_idx(g, 2, kb - 1)] : 0.0f;
    float lA_m, lB_m, lC_m, lD_m, mA_m, mB_m, mC_m, mD_m;
    if constexpr (P::mixed) {
        const unsigned wA = mlq[om], wB = mlq[om + 1], wC = mlq[om - sy], wD = mlq[om + 1 - sy];
        lA_m = mdec(h_hi(wA), lb_m); lB_m = mdec(h_hi(wB), lb_m); lC_m = mdec(h_hi(wC), lb_m); lD_m = mdec(h_hi(wD), lb_m);
        mA_m = mdec(h_lo(wA), mb_m); mB_m = mdec(h_lo(wB), mb_m); mC_m = mdec(h_lo(wC), mb_m); mD_m = mdec(h_lo(wD), mb_m);
    } else {
        lA_m = mdec(lam[om], lb_m); lB_m = mdec(lam[om + 1], lb_m); lC_m = mdec(lam[om - sy], lb_m);
        lD_m = mdec(lam[om + 1 - sy], lb_m);
        mA_m = mdec(mu[om], mb_m); mB_m = mdec(mu[om + 1], mb_m); mC_m = mdec(mu[om - sy], mb_m);
        mD_m = mdec(mu[om + 1 - sy], mb_m);
    }
    float pA_m, pB_m, pC_m, pD_m, sA_m, sB_m, sC_m, sD_m;
    if constexpr (P::mixed) {
        const unsigned wA = qq[om], wB = qq[om + 1], wC = qq[om - sy], wD = qq[om + 1 - sy];
        pA_m = q_lo(wA); pB_m = q_lo(wB); pC_m = q_lo(wC); pD_m = q_lo(wD);
        sA_m = q_hi(wA); sB_m = q_hi(wB); sC_m = q_hi(wC); sD_m = q_hi(wD);
    } else {
        pA_m = ld(qp[om]); pB_m = ld(qp[om + 1]); pC_m = ld(qp[om - sy]); pD_m = ld(qp[om + 1 - sy]);
        sA_m = ld(qs[om]); sB_m = ld(qs[om + 1]); sC_m = ld(qs[om - sy]); sD_m = ld(qs[om + 1 - sy]);
    }
    // final xz/yz/zz of the two levels below (for the free-surface images)
    float xz_f1 = 0.f, xz_f2 = 0.f, yz_f1 = 0.f, yz_f2 = 0.f, zz_f1 = 0.f;

#pragma unroll U
    for (int k = kb; k <= ke; ++k, o += sz) {
        const float u_p2 = u1[o + 2 * sz];
        const float v_p2 = v1[o + 2 * sz];
        const float w_p1 = w1[o + sz];
        const float mb = P::mixed ? s.mbar[mbar_idx(g, 1, k)] : 0.0f;
        const float lb = P::mixed ? s.mbar[mbar_idx(g, 2, k)] : 0.0f;
        float lA, lB, lC, lD, mA, mB, mC, mD;
        if constexpr (P::mixed) {
            const unsigned wA = mlq[o], wB = mlq[o + 1], wC = mlq[o - sy], wD = mlq[o + 1 - sy];
            lA = mdec(h_hi(wA), lb); lB = mdec(h_hi(wB), lb); lC = mdec(h_hi(wC), lb); lD = mdec(h_hi(wD), lb);
            mA = mdec(h_lo(wA), mb); mB = mdec(h_lo(wB), mb); mC = mdec(h_lo(wC), mb); mD = mdec(h_lo(wD), mb);
        } else {
            lA = mdec(lam[o], lb); lB = mdec(lam[o + 1], lb); lC = mdec(lam[o - sy], lb); lD = mdec(lam[o + 1 - sy], lb);
            mA = mdec(mu[o], mb); mB = mdec(mu[o + 1], mb); mC = mdec(mu[o - sy], mb); mD = mdec(mu[o + 1 - sy], mb);
        }
        float pA, pB, pC, pD, sA, sB, sC, sD;
        if constexpr (P::mixed) {
            const unsigned wA = qq[o], wB = qq[o + 1], wC = qq[o - sy], wD = qq[o + 1 - sy];
            pA = q_lo(wA); pB = q_lo(wB); pC = q_lo(wC); pD = q_lo(wD);
            sA = q_hi(wA); sB = q_hi(wB); sC = q_hi(wC); sD = q_hi(wD);
        } else {
            pA = ld(qp[o]); pB = ld(qp[o + 1]); pC = ld(qp[o - sy]); pD = ld(qp[o + 1 - sy]);
            sA = ld(qs[o]); sB = ld(qs[o + 1]); sC = ld(qs[o - sy]); sD = ld(qs[o + 1 - sy]);
        }

        const float u_ip1 = u1[o + 1], u_ip2 = u1[o + 2], u_im1 = u1[o - 1];
        const float u_jp1 = u1[o + sy], u_jp2 = u1[o + 2 * sy], u_jm1 = u1[o - sy];
        const float v_jm1 = v1[o - sy], v_jp1 = v1[o + sy], v_jm2 = v1[o - 2 * sy];
        const float v_im1 = v1[o - 1], v_ip1 = v1[o + 1], v_im2 = v1[o - 2];
        const float w_im1 = w1[o - 1], w_ip1 = w1[o + 1], w_im2 = w1[o - 2];
        const float w_jp1 = w1[o + sy], w_jp2 = w1[o + 2 * sy], w_jm1 = w1[o - sy];

        const bool odd = (((g.nz + 1) - k) & 1) != 0;
        const float vx1 = odd ? t1o : t1e;
        const float vx2 = odd ? t2o : t2e;

        float vxl = 8.0f / (lA + lB + lC + lD + lA_m + lB_m + lC_m + lD_m);
        const float vxm = 8.0f / (mA + mB + mC + mD + mA_m + mB_m + mC_m + mD_m);
        vxl = vxl + 2.0f * vxm;
        float vqpa = 0.125f * (pA + pB + pC + pD + pA_m + pB_m + pC_m + pD_m);
        float vqsa = 0.125f * (sA + sB + sC + sD + sA_m + sB_m + sC_m + sD_m);
        if constexpr (P::mixed) {
            vqpa = vqpa * s.ms.iqp;
            vqsa = vqsa * s.ms.iqs;
        }

        const float vt1 = c1 * (u_ip1 - u_0) + c2 * (u_ip2 - u_im1);
        const float vt2 = c1 * (v_0 - v_jm1) + c2 * (v_jp1 - v_jm2);
        const float vt3 = c1 * (w_0 - w_m1) + c2 * (w_p1 - w_m2);
        const float vt4 = c1 * (u_jp1 - u_0) + c2 * (u_jp2 - u_jm1) +
                          c1 * (v_0 - v_im1) + c2 * (v_ip1 - v_im2);
        const float vt5 = c1 * (u_p1 - u_0) + c2 * (u_p2 - u_m1) +
                          c1 * (w_0 - w_im1) + c2 * (w_ip1 - w_im2);
        const float vt6 = c1 * (v_p1 - v_0) + c2 * (v_p2 - v_m1);
        const float vt7 = c1 * (w_jp1 - w_0) + c2 * (w_jp2 - w_jm1);

        const float vh1 = -vxm * vqsa;
        const float vtmp = vxl * (vt1 + vt2 + vt3);
        const float va1 = -vqpa * vxl * dh1 * (vt1 + vt2 + vt3) / 2.0f;

        float r1, r2, r3, r4, r5, r6;
        unsigned long long ew = 0;   // mixed: the row segment's exponents before this update
        if constexpr (P::mixed) {
            ew = rexp_load(s.rexp, rexp_idx(g, s.nseg, seg, j, k));
            r1 = r_decode(R1[o], rexp_of(ew, 0)); r2 = r_decode(R2[o], rexp_of(ew, 1));
            r3 = r_decode(R3[o], rexp_of(ew, 2)); r4 = r_decode(R4[o], rexp_of(ew, 3));
            r5 = r_decode(R5[o], rexp_of(ew, 4)); r6 = r_decode(R6[o], rexp_of(ew, 5));
        } else {
            r1 = R1[o]; r2 = R2[o]; r3 = R3[o]; r4 = R4[o]; r5 = R5[o]; r6 = R6[o];
        }
        float xx = s.xx[o], yy = s.yy[o], zz = s.zz[o], xy = s.xy[o], xz = s.xz[o], yz = s.yz[o];

        xx = xx + dth * (vtmp - 2.0f * vxm * (vt2 + vt3)) + dt * r1;
        yy = yy + dth * (vtmp - 2.0f * vxm * (vt1 + vt3)) + dt * r2;
        zz = zz + dth * (vtmp - 2.0f * vxm * (vt1 + vt2)) + dt * r3;

        float vxmu = 2.0f / (mA + mA_m);
        float vaas = 0.5f * (sA + sA_m);
        if constexpr (P::mixed) vaas = vaas * s.ms.iqs;
        const float vh2_r4 = -vaas * vxmu / 2.0f;
        xy = xy + vxmu * dth * vt4 + dt * r4;
        vxmu = 2.0f / (mA + mC);
        vaas = 0.5f * (sA + sC);
        if constexpr (P::mixed) vaas = vaas * s.ms.iqs;
        const float vh2_r5 = -vaas * vxmu / 2.0f;
        xz = xz + vxmu * dth * vt5 + dt * r5;
        vxmu = 2.0f / (mA + mB);
        vaas = 0.5f * (sA + sB);
        if constexpr (P::mixed) vaas = vaas * s.ms.iqs;
        const float vh2_r6 = -vaas * vxmu / 2.0f;
        yz = yz + vxmu * (dth * vt6 + dth * vt7) + dt * r6;

        r1 = (vx2 * r1 - vh1 * (vt2 + vt3) * dh1 + va1) * vx1;
        r2 = (vx2 * r2 - vh1 * (vt1 + vt3) * dh1 + va1) * vx1;
        r3 = (vx2 * r3 - vh1 * (vt1 + vt2) * dh1 + va1) * vx1;
        r4 = (vx2 * r4 + vh2_r4 * dh1 * vt4) * vx1;
        r5 = (vx2 * r5 + vh2_r5 * dh1 * vt5) * vx1;
        r6 = (vx2 * r6 + vh2_r6 * dh1 * (vt6 + vt7)) * vx1;

        xx = xx + dt * r1;
        yy = yy + dt * r2;
        zz = zz + dt * r3;
        xy = xy + dt * r4;
        xz = xz + dt * r5;
        yz = yz + dt * r6;

        const float D = fxy * s.fz[k];
        xx = xx * D; yy = yy * D; zz = zz * D;
        xz = xz * D; yz = yz * D; xy = xy * D;

        // the new memory variables enter this step's stresses in FP32 and
        // are rounded to the storage format once, on the store
        if constexpr (P::mixed) {
            const float rv[6] = {r1, r2, r3, r4, r5, r6};
            G<R> *RR[6] = {R1, R2, R3, R4, R5, R6};
            const unsigned long long en = row_exponents6(rv, act, ew);
#pragma unroll
            for (int q = 0; q < 6; ++q)
                if (act) RR[q][o] = r_encode(rv[q], rexp_of(en, q));
            if (act && threadIdx.x == 0)
                ((G<unsigned long long> *)s.rexp)[rexp_idx(g, s.nseg, seg, j, k)] = en;
        } else {
            R1[o] = r1; R2[o] = r2; R3[o] = r3;
            R4[o] = r4; R5[o] = r5; R6[o] = r6;
        }
        if (act) {
        s.xx[o] = xx; s.yy[o] = yy; s.xy[o] = xy;
        if (k == g.nz) {
            // fstr: images above the surface, xz = yz = 0 on it
            s.zz[o] = zz;
            s.zz[o + sz] = -zz;
            s.zz[o + 2 * sz] = -zz_f1;
            s.xz[o] = 0.0f;
            s.yz[o] = 0.0f;
            s.xz[o + sz] = -xz_f1;
            s.xz[o + 2 * sz] = -xz_f2;
            s.yz[o + sz] = -yz_f1;
            s.yz[o + 2 * sz] = -yz_f2;
        } else {
            s.zz[o] = zz; s.xz[o] = xz; s.yz[o] = yz;
        }
        }
        xz_f2 = xz_f1; xz_f1 = xz;
        yz_f2 = yz_f1; yz_f1 = yz;
        zz_f1 = zz;

        u_m1 = u_0; u_0 = u_p1; u_p1 = u_p2;
        v_m1 = v_0; v_0 = v_p1; v_p1 = v_p2;
        w_m2 = w_m1; w_m1 = w_0; w_0 = w_p1;
        lA_m = lA; lB_m = lB; lC_m = lC; lD_m = lD;
        mA_m = mA; mB_m = mB; mC_m = mC; mD_m = mD;
        pA_m = pA; pB_m = pB; pC_m = pC; pD_m = pD;
        sA_m = sA; sB_m = sB; sC_m = sC; sD_m = sD;
    }
}

template <class P, int U>
__device__ __forceinline__ void stress_tile_b(const Geo &g, const State &s0, const Coef &c, const int kc, const int bx,
                                   const int by, const int bz)
{
    const GState s = as_global(s0);
    using Q = typename P::Q;
    using R = typename P::R;
    const int seg = bx * blockDim.z + threadIdx.z;   // the wave's 64-point row segment
    const int ir = seg * TX + threadIdx.x + 1;
    const int jr = by * blockDim.y + threadIdx.y + 1;
    const int kb = bz * kc + 1;
    const int ke = min(kb + kc - 1, g.nz);
    // mixed storage needs every lane of a wave for the row maximum of the
    // memory variables: lanes outside the domain compute an in-range
    // duplicate point and store nothing
    const bool act = ir <= g.nx && jr <= g.ny;
    if constexpr (!P::mixed) {
        if (!act) return;
    } else {
        // a wave whose whole row segment lies beyond nx (the second wave of
        // a 128-point tile at the grid's edge) has no point and no exponent
        if (seg * TX >= g.nx || jr > g.ny) return;
    }
    const int i = min(ir, g.nx), j = min(jr, g.ny);
    Cursor32 p(g, g.base + (long long)kb * g.sz + (long long)j * g.sy + i);
    using O = unsigned;

    const G<float> *__restrict__ u1 = s.u1;
    const G<float> *__restrict__ v1 = s.v1;
    const G<float> *__restrict__ w1 = s.w1;
    using M = typename P::M;
    const G<M> *__restrict__ lam = (const G<M> *)s.lam;
    const G<M> *__restrict__ mu = (const G<M> *)s.mu;
    const G<unsigned> *__restrict__ mlq = (const G<unsigned> *)s.mu;   // mixed: (mu, lam) pairs
    const G<Q> *__restrict__ qp = (const G<Q> *)s.qp;
    const G<Q> *__restrict__ qs = (const G<Q> *)s.qs;
    const G<unsigned> *__restrict__ qq = (const G<unsigned> *)s.qp;   // mixed: (qp, qs) pairs
    G<R> *__restrict__ R1 = (G<R> *)s.r1;
    G<R> *__restrict__ R2 = (G<R> *)s.r2;
    G<R> *__restrict__ R3 = (G<R> *)s.r3;
    G<R> *__restrict__ R4 = (G<R> *)s.r4;
    G<R> *__restrict__ R5 = (G<R> *)s.r5;
    G<R> *__restrict__ R6 = (G<R> *)s.r6;
    const float c1 = c.c1, c2 = c.c2, dth = c.dth, dt = c.dt, dh1 = c.dh1;
    const float fxy = s.fx[i] * s.fy[j];
    // relaxation weights of this column for even / odd depth parity
    // (selected per level: no per-lane indexing of the argument block)
    const int itxy = ((i & 1) * 2 + (j & 1)) * 2;
    const float t1e = c.tau1[itxy], t1o = c.tau1[itxy + 1];
    const float t2e = c.tau2[itxy], t2o = c.tau2[itxy + 1];

    // z queues of the velocities at the column
    float u_m1 = p.ld(u1, 0, 0, -1), u_0 = p.ld(u1), u_p1 = p.ld(u1, 0, 0, 1);
    float v_m1 = p.ld(v1, 0, 0, -1), v_0 = p.ld(v1), v_p1 = p.ld(v1, 0, 0, 1);
    float w_m2 = p.ld(w1, 0, 0, -2), w_m1 = p.ld(w1, 0, 0, -1), w_0 = p.ld(w1);
    // media at level k-1 on the 2x2 (x,y) cell: A=(i,j) B=(i+1,j) C=(i,j-1) D=(i+1,j-1)
    const float mb_m = P::mixed ? s.mbar[mbar_idx(g, 1, kb - 1)] : 0.0f;
    const float lb_m = P::mixed ? s.mbar[mbar_idx(g, 2, kb - 1)] : 0.0f;
    float lA_m, lB_m, lC_m, lD_m, mA_m, mB_m, mC_m, mD_m;
    if constexpr (P::mixed) {
        const unsigned wA = p.ld(mlq, 0, 0, -1), wB = p.ld(mlq, 1, 0, -1), wC = p.ld(mlq, 0, -1, -1), wD = p.ld(mlq, 1, -1, -1);
        lA_m = mdec(h_hi(wA), lb_m); lB_m = mdec(h_hi(wB), lb_m); lC_m = mdec(h_hi(wC), lb_m); lD_m = mdec(h_hi(wD), lb_m);
        mA_m = mdec(h_lo(wA), mb_m); mB_m = mdec(h_lo(wB), mb_m); mC_m = mdec(h_lo(wC), mb_m); mD_m = mdec(h_lo(wD), mb_m);
    } else {
        lA_m = mdec(p.ld(lam, 0, 0, -1), lb_m); lB_m = mdec(p.ld(lam, 1, 0, -1), lb_m);
        lC_m = mdec(p.ld(lam, 0, -1, -1), lb_m); lD_m = mdec(p.ld(lam, 1, -1, -1), lb_m);
        mA_m = mdec(p.ld(mu, 0, 0, -1), mb_m); mB_m = mdec(p.ld(mu, 1, 0, -1), mb_m);
        mC_m = mdec(p.ld(mu, 0, -1, -1), mb_m); mD_m = mdec(p.ld(mu, 1, -1, -1), mb_m);
    }
    float pA_m, pB_m, pC_m, pD_m, sA_m, sB_m, sC_m, sD_m;
    if constexpr (P::mixed) {
        const unsigned wA = p.ld(qq, 0, 0, -1), wB = p.ld(qq, 1, 0, -1);
        const unsigned wC = p.ld(qq, 0, -1, -1), wD = p.ld(qq, 1, -1, -1);
        pA_m = q_lo(wA); pB_m = q_lo(wB); pC_m = q_lo(wC); pD_m = q_lo(wD);
        sA_m = q_hi(wA); sB_m = q_hi(wB); sC_m = q_hi(wC); sD_m = q_hi(wD);
    } else {
        pA_m = ld(p.ld(qp, 0, 0, -1)); pB_m = ld(p.ld(qp, 1, 0, -1));
        pC_m = ld(p.ld(qp, 0, -1, -1)); pD_m = ld(p.ld(qp, 1, -1, -1));
        sA_m = ld(p.ld(qs, 0, 0, -1)); sB_m = ld(p.ld(qs, 1, 0, -1));
        sC_m = ld(p.ld(qs, 0, -1, -1)); sD_m = ld(p.ld(qs, 1, -1, -1));
    }
    // final xz/yz/zz of the two levels below (for the free-surface images)
    float xz_f1 = 0.f, xz_f2 = 0.f, yz_f1 = 0.f, yz_f2 = 0.f, zz_f1 = 0.f;
    // mixed storage: offset of this row segment's exponent at level k in
    // each exponent array, and its step per level
    O xo = P::mixed ? (O)rexp_idx(g, s.nseg, seg, j, kb) * 8 : (O)0;
    const O xstep = (O)g.ny * (O)s.nseg * 8;

#pragma unroll U
    for (int k = kb; k <= ke; ++k, p.up(), xo += xstep) {
        const float u_p2 = p.ld(u1, 0, 0, 2);
        const float v_p2 = p.ld(v1, 0, 0, 2);
        const float w_p1 = p.ld(w1, 0, 0, 1);
        const float mb = P::mixed ? s.mbar[mbar_idx(g, 1, k)] : 0.0f;
        const float lb = P::mixed ? s.mbar[mbar_idx(g, 2, k)] : 0.0f;
        float lA, lB, lC, lD, mA, mB, mC, mD;
        if constexpr (P::mixed) {
            const unsigned wA = p.ld(mlq), wB = p.ld(mlq, 1), wC = p.ld(mlq, 0, -1), wD = p.ld(mlq, 1, -1);
            lA = mdec(h_hi(wA), lb); lB = mdec(h_hi(wB), lb); lC = mdec(h_hi(wC), lb); lD = mdec(h_hi(wD), lb);
            mA = mdec(h_lo(wA), mb); mB = mdec(h_lo(wB), mb); mC = mdec(h_lo(wC), mb); mD = mdec(h_lo(wD), mb);
        } else {
            lA = mdec(p.ld(lam), lb); lB = mdec(p.ld(lam, 1), lb); lC = mdec(p.ld(lam, 0, -1), lb);
            lD = mdec(p.ld(lam, 1, -1), lb);
            mA = mdec(p.ld(mu), mb); mB = mdec(p.ld(mu, 1), mb); mC = mdec(p.ld(mu, 0, -1), mb);
            mD = mdec(p.ld(mu, 1, -1), mb);
        }
        float pA, pB, pC, pD, sA, sB, sC, sD;
        if constexpr (P::mixed) {
            const unsigned wA = p.ld(qq), wB = p.ld(qq, 1), wC = p.ld(qq, 0, -1), wD = p.ld(qq, 1, -1);
            pA = q_lo(wA); pB = q_lo(wB); pC = q_lo(wC); pD = q_lo(wD);
            sA = q_hi(wA); sB = q_hi(wB); sC = q_hi(wC); sD = q_hi(wD);
        } else {
            pA = ld(p.ld(qp)); pB = ld(p.ld(qp, 1)); pC = ld(p.ld(qp, 0, -1)); pD = ld(p.ld(qp, 1, -1));
            sA = ld(p.ld(qs)); sB = ld(p.ld(qs, 1)); sC = ld(p.ld(qs, 0, -1)); sD = ld(p.ld(qs, 1, -1));
        }

        const float u_ip1 = p.ld(u1, 1), u_ip2 = p.ld(u1, 2), u_im1 = p.ld(u1, -1);
        const float u_jp1 = p.ld(u1, 0, 1), u_jp2 = p.ld(u1, 0, 2), u_jm1 = p.ld(u1, 0, -1);
        const float v_jm1 = p.ld(v1, 0, -1), v_jp1 = p.ld(v1, 0, 1), v_jm2 = p.ld(v1, 0, -2);
        const float v_im1 = p.ld(v1, -1), v_ip1 = p.ld(v1, 1), v_im2 = p.ld(v1, -2);
        const float w_im1 = p.ld(w1, -1), w_ip1 = p.ld(w1, 1), w_im2 = p.ld(w1, -2);
        const float w_jp1 = p.ld(w1, 0, 1), w_jp2 = p.ld(w1, 0, 2), w_jm1 = p.ld(w1, 0, -1);

        const bool odd = (((g.nz + 1) - k) & 1) != 0;
        const float vx1 = odd ? t1o : t1e;
        const float vx2 = odd ? t2o : t2e;

        float vxl = 8.0f / (lA + lB + lC + lD + lA_m + lB_m + lC_m + lD_m);
        const float vxm = 8.0f / (mA + mB + mC + mD + mA_m + mB_m + mC_m + mD_m);
        vxl = vxl + 2.0f * vxm;
        float vqpa = 0.125f * (pA + pB + pC + pD + pA_m + pB_m + pC_m + pD_m);
        float vqsa = 0.125f * (sA + sB + sC + sD + sA_m + sB_m + sC_m + sD_m);
        if constexpr (P::mixed) {
            vqpa = vqpa * s.ms.iqp;
            vqsa = vqsa * s.ms.iqs;
        }

        const float vt1 = c1 * (u_ip1 - u_0) + c2 * (u_ip2 - u_im1);
        const float vt2 = c1 * (v_0 - v_jm1) + c2 * (v_jp1 - v_jm2);
        const float vt3 = c1 * (w_0 - w_m1) + c2 * (w_p1 - w_m2);
        const float vt4 = c1 * (u_jp1 - u_0) + c2 * (u_jp2 - u_jm1) +
                          c1 * (v_0 - v_im1) + c2 * (v_ip1 - v_im2);
        const float vt5 = c1 * (u_p1 - u_0) + c2 * (u_p2 - u_m1) +
                          c1 * (w_0 - w_im1) + c2 * (w_ip1 - w_im2);
        const float vt6 = c1 * (v_p1 - v_0) + c2 * (v_p2 - v_m1);
        const float vt7 = c1 * (w_jp1 - w_0) + c2 * (w_jp2 - w_jm1);

        const float vh1 = -vxm * vqsa;
        const float vtmp = vxl * (vt1 + vt2 + vt3);
        const float va1 = -vqpa * vxl * dh1 * (vt1 + vt2 + vt3) / 2.0f;

        float r1, r2, r3, r4, r5, r6;
        unsigned long long ew = 0;   // mixed: the row segment's exponents before this update
        if constexpr (P::mixed) {
            // the exponents of this row segment: one 8-byte slot at byte offset xo
            ew = rexp_load(s.rexp, (long long)(xo >> 3));
            r1 = r_decode(p.ldn(R1), rexp_of(ew, 0)); r2 = r_decode(p.ldn(R2), rexp_of(ew, 1));
            r3 = r_decode(p.ldn(R3), rexp_of(ew, 2)); r4 = r_decode(p.ldn(R4), rexp_of(ew, 3));
            r5 = r_decode(p.ldn(R5), rexp_of(ew, 4)); r6 = r_decode(p.ldn(R6), rexp_of(ew, 5));
        } else {
            r1 = p.ldn(R1); r2 = p.ldn(R2); r3 = p.ldn(R3);
            r4 = p.ldn(R4); r5 = p.ldn(R5); r6 = p.ldn(R6);
        }
        float xx = p.ldn(s.xx), yy = p.ldn(s.yy), zz = p.ldn(s.zz);
        float xy = p.ldn(s.xy), xz = p.ldn(s.xz), yz = p.ldn(s.yz);

        xx = xx + dth * (vtmp - 2.0f * vxm * (vt2 + vt3)) + dt * r1;
        yy = yy + dth * (vtmp - 2.0f * vxm * (vt1 + vt3)) + dt * r2;
        zz = zz + dth * (vtmp - 2.0f * vxm * (vt1 + vt2)) + dt * r3;

        float vxmu = 2.0f / (mA + mA_m);
        float vaas = 0.5f * (sA + sA_m);
        if constexpr (P::mixed) vaas = vaas * s.ms.iqs;
        const float vh2_r4 = -vaas * vxmu / 2.0f;
        xy = xy + vxmu * dth * vt4 + dt * r4;
        vxmu = 2.0f / (mA + mC);
        vaas = 0.5f * (sA + sC);
        if constexpr (P::mixed) vaas = vaas * s.ms.iqs;
        const float vh2_r5 = -vaas * vxmu / 2.0f;
        xz = xz + vxmu * dth * vt5 + dt * r5;
        vxmu = 2.0f / (mA + mB);
        vaas = 0.5f * (sA + sB);
        if constexpr (P::mixed) vaas = vaas * s.ms.iqs;
        const float vh2_r6 = -vaas * vxmu / 2.0f;
        yz = yz + vxmu * (dth * vt6 + dth * vt7) + dt * r6;

        r1 = (vx2 * r1 - vh1 * (vt2 + vt3) * dh1 + va1) * vx1;
        r2 = (vx2 * r2 - vh1 * (vt1 + vt3) * dh1 + va1) * vx1;
        r3 = (vx2 * r3 - vh1 * (vt1 + vt2) * dh1 + va1) * vx1;
        r4 = (vx2 * r4 + vh2_r4 * dh1 * vt4) * vx1;
        r5 = (vx2 * r5 + vh2_r5 * dh1 * vt5) * vx1;
        r6 = (vx2 * r6 + vh2_r6 * dh1 * (vt6 + vt7)) * vx1;

        xx = xx + dt * r1;
        yy = yy + dt * r2;
        zz = zz + dt * r3;
        xy = xy + dt * r4;
        xz = xz + dt * r5;
        yz = yz + dt * r6;

        const float D = fxy * s.fz[k];
        xx = xx * D; yy = yy * D; zz = zz * D;
        xz = xz * D; yz = yz * D; xy = xy * D;

        // the new memory variables enter this step's stresses in FP32 and
        // are rounded to the storage format once, on the store
        if constexpr (P::mixed) {
            const float rv[6] = {r1, r2, r3, r4, r5, r6};
            G<R> *RR[6] = {R1, R2, R3, R4, R5, R6};
            const unsigned long long en = row_exponents6(rv, act, ew);
#pragma unroll
            for (int q = 0; q < 6; ++q)
                if (act) p.stn(RR[q], r_encode(rv[q], rexp_of(en, q)));
            if (act && threadIdx.x == 0) gsn((G<unsigned long long> *)s.rexp, xo, en);
        } else {
            p.stn(R1, r1); p.stn(R2, r2); p.stn(R3, r3);
            p.stn(R4, r4); p.stn(R5, r5); p.stn(R6, r6);
        }
        if (act) {
            p.stn(s.xx, xx); p.stn(s.yy, yy); p.stn(s.xy, xy);
            if (k == g.nz) {
                // fstr: images above the surface, xz = yz = 0 on it
                p.st(s.zz, zz);
                p.st(s.zz, -zz, 0, 0, 1);
                p.st(s.zz, -zz_f1, 0, 0, 2);
                p.st(s.xz, 0.0f);
                p.st(s.yz, 0.0f);
                p.st(s.xz, -xz_f1, 0, 0, 1);
                p.st(s.xz, -xz_f2, 0, 0, 2);
                p.st(s.yz, -yz_f1, 0, 0, 1);
                p.st(s.yz, -yz_f2, 0, 0, 2);
            } else {
                p.stn(s.zz, zz); p.stn(s.xz, xz); p.stn(s.yz, yz);
            }
        }
        xz_f2 = xz_f1; xz_f1 = xz;
        yz_f2 = yz_f1; yz_f1 = yz;
        zz_f1 = zz;

        u_m1 = u_0; u_0 = u_p1; u_p1 = u_p2;
        v_m1 = v_0; v_0 = v_p1; v_p1 = v_p2;
        w_m2 = w_m1; w_m1 = w_0; w_0 = w_p1;
        lA_m = lA; lB_m = lB; lC_m = lC; lD_m = lD;
        mA_m = mA; mB_m = mB; mC_m = mC; mD_m = mD;
        pA_m = pA; pB_m = pB; pC_m = pC; pD_m = pD;
        sA_m = sA; sB_m = sB; sC_m = sC; sD_m = sD;
    }
}

template <class P, int U, bool O32>
__device__ __forceinline__ void stress_tile(const Geo &g, const State &s0, const Coef &c, const int kc, const int bx,
                                   const int by, const int bz)
{
    if constexpr (O32)
        stress_tile_b<P, U>(g, s0, c, kc, bx, by, bz);
    else
        stress_tile_e<P, U>(g, s0, c, kc, bx, by, bz);
}

// ---------------------------------------------------------------------------
// Two-stage step kernels over a job table of one or more runs: one launch per
// stage covers the concatenated tile lists of all runs (one run: the table
// has one entry).
//
// Block -> tile map.  The hardware hands consecutive workgroup ids to the 8
// XCDs in turn, and each XCD has its own L2.  Each XCD gets one contiguous
// range of the tile list (patches of tiles, x fastest inside a patch, within
// each z chunk), so the tiles it holds at once are x and y neighbours whose
// halo lines it reads from its own L2.
// ---------------------------------------------------------------------------
// Run of global tile `tile` (jobs ascending in tile0) and the tile's
// (bx, by, bz) in that run's patch order.
__device__ inline const SJob &split_tile(const SJob *__restrict__ jobs, const int njobs, const int tile, int &bx,
                                         int &by, int &bz)
{
    int jb = 0;
    while (jb + 1 < njobs && jobs[jb + 1].tile0 <= tile) ++jb;
    const SJob &J = jobs[jb];
    patch_tile(tile - J.tile0, J.nbx, J.nby, J.pw, J.ph, bx, by, bz);
    return J;
}

// The next tile of a persistent stage launch, the same in every thread of
// the block, or -1 when none is left: claimed (claim_tile) when `claim` is
// set, else the static walk -- block n stays on XCD n % 8 and walks that
// XCD's contiguous range of the global tile list with stride gridDim.x / 8
// (k: the block's tiles so far), so the blocks of one XCD work on
// neighbouring tiles of one run at a time and read each other's halo lines
// from that XCD's L2.  `next`: the block's LDS broadcast slot.
__device__ inline int stage_tile(unsigned long long *claim, int per, int total, int rot, int k, int *next)
{
    if (claim) {
        __syncthreads();   // every thread has read the previous tile
        if (threadIdx.x == 0 && threadIdx.y == 0 && threadIdx.z == 0) *next = claim_tile(claim, per, total, rot);
        __syncthreads();
        return *next;
    }
    const int xcd = blockIdx.x & 7, slot = (int)(blockIdx.x >> 3) + k * (int)(gridDim.x >> 3);
    const int tile = xcd * per + slot;
    return (slot < per && tile < total) ? tile : -1;
}

// Persistent stage kernels over the job table (stage_tile assigns tiles).
// MAXT: most threads per block -- 512 for tiles of up to 8 rows, 1024 for
// 16-row tiles (whose 16 waves leave 128 registers a lane)
template <class P, int U, int MAXT, bool O32>
__global__ void __launch_bounds__(MAXT)
k_velocity_batch(const SJob *__restrict__ jobs, const int njobs, const int total, const int per,
                 unsigned long long *claim, const int rot)
{
    __shared__ int next;
    for (int k = 0;; ++k) {
        const int tile = stage_tile(claim, per, total, rot, k, &next);
        if (tile < 0) break;
        int bx, by, bz;
        const SJob &J = split_tile(jobs, njobs, tile, bx, by, bz);
        if constexpr (O32)
            velocity_tile_b<P, U>(J.g, J.s, J.c, J.kc, bx, by, bz);
        else
            velocity_tile<P, U>(J.g, J.s, J.c, J.kc, bx, by, bz);
    }
    if (claim && threadIdx.x == 0 && threadIdx.y == 0 && threadIdx.z == 0) claim_done(claim, (int)gridDim.x);
}

template <class P, int U, bool O32, int MAXT>
__global__ void __launch_bounds__(MAXT)
k_stress_batch(const SJob *__restrict__ jobs, const int njobs, const int total, const int per,
               unsigned long long *claim, const int rot)
{
    __shared__ int next;
    for (int k = 0;; ++k) {
        const int tile = stage_tile(claim, per, total, rot, k, &next);
        if (tile < 0) break;
        int bx, by, bz;
        const SJob &J = split_tile(jobs, njobs, tile, bx, by, bz);
        stress_tile<P, U, O32>(J.g, J.s, J.c, J.kc, bx, by, bz);
    }
    if (claim && threadIdx.x == 0 && threadIdx.y == 0 && threadIdx.z == 0) claim_done(claim, (int)gridDim.x);
}

// free-surface images of every run: block b of the launch is block
// b - fs0 of its run's (nx/64) x (ny/4) surface grid
template <class P>
__global__ void __launch_bounds__(256)
k_fs_velocity_batch(const SJob *__restrict__ jobs, const int njobs)
{
    const int b = blockIdx.x;
    int jb = 0;
    while (jb + 1 < njobs && jobs[jb + 1].fs0 <= b) ++jb;
    const SJob &J = jobs[jb];
    const int lb = b - J.fs0;
    const int fnx = (J.g.nx + TX - 1) / TX;
    fs_velocity_block<P>(J.g, J.s, lb % fnx, lb / fnx);
}

// ---------------------------------------------------------------------------
// point source, addsrc (source.f:812-913).  One thread walks the sources in
// file order so coincident sources accumulate in the reference's order.
// ---------------------------------------------------------------------------
__device__ void apply_sources(const State &s, const int igreen, const int npsrc, const long long *psrc_idx,
                              const long long *psrc_top, const long long sz, const float *stf, const int nst,
                              const int istep, const float vtst, const float vtst1)
{
    for (int n = 0; n < npsrc; ++n) {
        const float *a = stf + ((size_t)n * nst + (istep - 1)) * 6;
        const float axx = a[0], ayy = a[1], azz = a[2];
        const long long p = psrc_idx[n];
        const long long t = psrc_top[n];   // (x, y, nz)
        float add = 0.0f;
        if (igreen >= 1 && igreen <= 3) {
            const float amp = igreen == 1 ? axx : (igreen == 2 ? ayy : azz);
            // density at the source: FP32, or (mixed storage, s.mbar set)
            // FP16 under the profile of its level p / sz - 1
            const float dd = s.mbar ? mdec(((const _Float16 *)s.d1)[p], s.mbar[p / sz]) : ((const float *)s.d1)[p];
            add = amp * vtst / dd;
        }
        switch (igreen) {
        case 1: s.u1[p] = s.u1[p] + add; break;
        case 2: s.v1[p] = s.v1[p] + add; break;
        case 3: s.w1[p] = s.w1[p] + add; break;
        case 4:
            s.xz[t] = axx * vtst1;
            s.xz[t + sz] = 2.0f * axx * vtst1 - s.xz[t - sz];
            s.xz[t + 2 * sz] = 2.0f * axx * vtst1 - s.xz[t - 2 * sz];
            break;
        case 5:
            s.yz[t] = ayy * vtst1;
            s.yz[t + sz] = 2.0f * ayy * vtst1 - s.yz[t - sz];
            s.yz[t + 2 * sz] = 2.0f * ayy * vtst1 - s.yz[t - 2 * sz];
            break;
        case 6:
            s.zz[t + sz] = 2.0f * azz * vtst1 - s.zz[t];
            s.zz[t + 2 * sz] = 2.0f * azz * vtst1 - s.zz[t - sz];
            break;
        default: break;
        }
    }
}

// The job-value trace the reference's master rank writes to CHKJ after
// every step (pmcl3d.f:1476-1479): u1, v1, w1 at node tpos, as row `row`
// of trace[][3].
__device__ inline void record_trace(const State &s, float *trace, const long long tpos, const int row)
{
    trace[3 * row + 0] = s.u1[tpos];
    trace[3 * row + 1] = s.v1[tpos];
    trace[3 * row + 2] = s.w1[tpos];
}

// Source sample `istep` (when `apply`), then the trace row of the step just
// completed (istep - 2) when `trace` is set.
__global__ void k_source(const State s, const int igreen, const int npsrc, const long long *psrc_idx,
                         const long long *psrc_top, const long long sz, const float *stf, const int nst,
                         const int istep, const float vtst, const float vtst1, const int apply, float *trace,
                         const long long tpos)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (apply) apply_sources(s, igreen, npsrc, psrc_idx, psrc_top, sz, stf, nst, istep, vtst, vtst1);
    if (trace) record_trace(s, trace, tpos, istep - 2);
}

// ---------------------------------------------------------------------------
// SGT: moduli coefficients at the output nodes (pmcl3d.f:1158-1170), from the
// stored media (mixed storage, s.mbar set: decoded under the profile of the
// node's level kpos / sz - 1), and the strain gather (pmcl3d.f:1546-1557).
// out is [npts][6][nsamp_max].
// ---------------------------------------------------------------------------
__global__ void k_sgt_coef(const State s, const long long sz, const int pz, const int nkept, const long long *kpos,
                           float *sg)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nkept) return;
    const long long p = kpos[n];
    float lam, mu;
    if (s.mbar) {
        const long long q = p / sz;   // profile index of the level
        const unsigned w = ((const unsigned *)s.mu)[p];   // (mu, lam) pair
        mu = mdec(h_lo(w), s.mbar[pz + q]);
        lam = mdec(h_hi(w), s.mbar[2 * pz + q]);
    } else {
        mu = ((const float *)s.mu)[p];
        lam = ((const float *)s.lam)[p];
    }
    sg[3 * n + 0] = (1.0f / lam + 1.0f / mu) / (1.0f / mu * (3.0f / lam + 2.0f / mu));
    sg[3 * n + 1] = -((1.0f / lam) / (2.0f / mu * (3.0f / lam + 2.0f / mu)));
    sg[3 * n + 2] = mu;
}

__device__ inline void gather_point(const State &s, const int n, const long long *kpos, const int *kidx,
                                    const float *sg, float *out, const int nsamp_max, const int sample)
{
    const long long p = kpos[n];
    const float sg1 = sg[3 * n], sg2 = sg[3 * n + 1], sg3 = sg[3 * n + 2];
    const float xx = s.xx[p], yy = s.yy[p], zz = s.zz[p];
    float *dst = out + (size_t)kidx[n] * 6 * nsamp_max + sample;
    dst[0 * (size_t)nsamp_max] = sg1 * xx + sg2 * yy + sg2 * zz;
    dst[1 * (size_t)nsamp_max] = sg2 * xx + sg1 * yy + sg2 * zz;
    dst[2 * (size_t)nsamp_max] = sg2 * xx + sg2 * yy + sg1 * zz;
    dst[3 * (size_t)nsamp_max] = sg3 * s.xy[p];
    dst[4 * (size_t)nsamp_max] = sg3 * s.xz[p];
    dst[5 * (size_t)nsamp_max] = sg3 * s.yz[p];
}

__global__ void k_sgt_gather(const State s, const int nkept, const long long *kpos, const int *kidx,
                             const float *sg, float *out, const int nsamp_max, const int sample)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nkept) return;
    gather_point(s, n, kpos, kidx, sg, out, nsamp_max, sample);
}

// The source and SGT work of one step for every run of a batch (the step's
// number comes from the run's device counter, so the launches are the same
// every step and can be replayed from a graph).  The sources come first, in
// a launch of their own (thread 0 of block r: run r's sources), because the
// surface sources (igreen 4..6) write stresses at k = nz that an SGT point
// on the surface reads: the reference applies the source sample, then takes
// the SGT sample (pmcl3d.f:1381-1383, 1536-1557).  Then block (x, r) of the
// gather launch handles run r's SGT points.
__global__ void __launch_bounds__(64) k_src_batch(const AuxJob *__restrict__ jobs, const int par)
{
    const AuxJob &J = jobs[blockIdx.x];
    const int i = *J.counter + 1;    // this step
    if (threadIdx.x != 0) return;
    if (i + 1 <= J.nst)
        apply_sources(J.s[par ^ 1], J.igreen, J.npsrc, J.psrc, J.ptop, J.g.sz, J.stf, J.nst, i + 1, J.vtst,
                      J.vtst1);
    if (J.trace) record_trace(J.s[par ^ 1], J.trace, J.tpos, i - 1);
}

__global__ void __launch_bounds__(256) k_aux_batch(const AuxJob *__restrict__ jobs, const int par)
{
    const AuxJob &J = jobs[blockIdx.y];
    const int i = *J.counter + 1;    // this step
    const State &s = J.s[par ^ 1];   // the state after it
    if (i % J.ntiskp == 0) {
        const int sample = i / J.ntiskp - 1;
        if (sample < J.nsamp_max)
            for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < J.nkept; n += gridDim.x * blockDim.x)
                gather_point(s, n, J.kpos, J.kidx, J.sg, J.sgt, J.nsamp_max, sample);
    }
}

__global__ void k_count_batch(const AuxJob *__restrict__ jobs, const int njobs)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < njobs) *jobs[r].counter += 1;
}

// ---------------------------------------------------------------------------
// media set-up (FP32): structure.f:449-524 (transform of one file record into
// the compliances, density and Q weights at k = nz+1-z), then the ghost ring.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_media_setup(const Geo g, const Media32 m, const float *media, const MediaConst mc)
{
    const int i = blockIdx.x * TX + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const int zf = g.nz - k;
    const int nvar = mc.nvar;
    const int voff = (nvar == 8) ? 3 : 0;
    const float *rec = media + (((size_t)zf * g.ny + (j - 1)) * g.nx + (i - 1)) * nvar;
    float vp = rec[voff + 0], vs = rec[voff + 1], dd = rec[voff + 2];
    float pq, sq;
    if (nvar == 3) {
        sq = 0.05f * vs;
        pq = 2.0f * sq;
    } else {
        pq = rec[voff + 3];
        sq = rec[voff + 4];
    }
    const float pi = mc.pi, lw = mc.lw;
    if (mc.socalq == 1) {
        vs = vs * (1.0f + (lw) / (pi * sq));
        vp = vp * (1.0f + (lw) / (pi * pq));
        const float vpvs = vp / vs;
        if (vpvs < 1.45f) vs = vp / 1.45f;
    }
    const long long p = gidx(g, i, j, k);
    m.mu[p] = 1.0f / (dd * vs * vs);
    m.lam[p] = 1.0f / (dd * (vp * vp - 2.0f * vs * vs));
    m.d1[p] = dd;
    float qpinv, qsinv;
    if (pq <= 0.0f) {
        qpinv = 0.0f; qsinv = 0.0f;
    } else {
        qpinv = 1.0f / pq;
        qsinv = 1.0f / sq;
    }
    m.qp[p] = mc.tmp1 * qpinv / (1.0f - mc.tmp2 * qpinv);
    m.qs[p] = mc.tmp1 * qsinv / (1.0f - mc.tmp2 * qsinv);
}

// Ghost ring of the media, i in 0..nx+1, j in 0..ny+1, k in 0..nz+1 with at
// least one index on the ring.  lam/mu/d1: padmedia (structure.f:563-676)
// and afsmedia (structure.f:679-707) replicate the nearest interior value;
// qp/qs are never padded (zero) except the afsmedia copy of the surface to
// k = nz+1 under the interior.
__global__ void __launch_bounds__(256)
k_media_ghosts(const Geo g, const Media32 m)
{
    const int i = blockIdx.x * TX + threadIdx.x;         // 0..nx+1
    const int j = blockIdx.y * blockDim.y + threadIdx.y; // 0..ny+1
    const int k = blockIdx.z;                            // 0..nz+1
    if (i > g.nx + 1 || j > g.ny + 1 || k > g.nz + 1) return;
    const bool ring = (i == 0 || i == g.nx + 1 || j == 0 || j == g.ny + 1 || k == 0 || k == g.nz + 1);
    if (!ring) return;
    const int ci = min(max(i, 1), g.nx), cj = min(max(j, 1), g.ny), ck = min(max(k, 1), g.nz);
    const long long p = gidx(g, i, j, k);
    const long long q = gidx(g, ci, cj, ck);
    m.lam[p] = m.lam[q];
    m.mu[p] = m.mu[q];
    m.d1[p] = m.d1[q];
    const bool top_interior = (k == g.nz + 1 && i >= 1 && i <= g.nx && j >= 1 && j <= g.ny);
    m.qp[p] = top_interior ? m.qp[q] : 0.0f;
    m.qs[p] = top_interior ? m.qs[q] : 0.0f;
}

// padmedia's corner (nxt+1,0,0) takes lam(nxt,1,1) for lam, mu and d1
// (structure.f:653-655); applied after the ring.
__global__ void k_media_corner(const Geo g, const Media32 m)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const long long p = gidx(g, g.nx + 1, 0, 0);
    const float v = m.lam[gidx(g, g.nx, 1, 1)];
    m.lam[p] = v;
    m.mu[p] = v;
    m.d1[p] = v;
}

// Mixed storage of d1, mu, lam.  First, per level k = 1..nz, the smallest
// and the largest nonzero value of the level's interior (i = 1..nx,
// j = 1..ny): order-preserving integer keys of the floats meet in atomics,
// so the result does not depend on the order of the blocks.  Block (y, k)
// takes rows j = 16y+1 .. 16y+16 of level k.
__device__ inline unsigned float_key(float f)
{
    const unsigned b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__global__ void __launch_bounds__(256) k_level_range(const Geo g, const float *a, unsigned *range)
{
    __shared__ unsigned red[2][256];
    const int k = blockIdx.y + 1;
    const int j0 = blockIdx.x * 16 + 1, j1 = min(j0 + 15, g.ny);
    unsigned lo = 0xFFFFFFFFu, hi = 0u;
    for (int j = j0; j <= j1; ++j)
        for (int i = threadIdx.x + 1; i <= g.nx; i += blockDim.x) {
            const float v = a[gidx(g, i, j, k)];
            if (v != 0.0f) {
                const unsigned key = float_key(v);
                lo = min(lo, key);
                hi = max(hi, key);
            }
        }
    red[0][threadIdx.x] = lo;
    red[1][threadIdx.x] = hi;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] = min(red[0][threadIdx.x], red[0][threadIdx.x + w]);
            red[1][threadIdx.x] = max(red[1][threadIdx.x], red[1][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicMin(range + 2 * (k - 1), red[0][0]);
        atomicMax(range + 2 * (k - 1) + 1, red[1][0]);
    }
}

// Then every cell of the padded volume (ghosts included) as the FP16 (round
// to nearest even) of (float)(x / b - 1) in double, b = bar[z] the profile
// value of its level plane z = blockIdx.y; a level without a profile value
// (b = 0) holds zeros.
__global__ void __launch_bounds__(256)
k_media_encode(const float *src, _Float16 *dst, const float *bar, const long long sz)
{
    const float b = bar[blockIdx.y];
    const long long z0 = (long long)blockIdx.y * sz;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < sz; t += (long long)gridDim.x * blockDim.x) {
        const float x = src[z0 + t];
        dst[z0 + t] = b != 0.0f ? (_Float16)(float)((double)x / (double)b - 1.0) : (_Float16)0.0f;
    }
}

// mu and lam as (mu, lam) pairs, each half rounded as k_media_encode rounds it
__global__ void __launch_bounds__(256)
k_media_encode2(const float *mu, const float *lam, unsigned *dst, const float *bar_mu, const float *bar_lam,
                const long long sz)
{
    const float bm = bar_mu[blockIdx.y], bl = bar_lam[blockIdx.y];
    const long long z0 = (long long)blockIdx.y * sz;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < sz; t += (long long)gridDim.x * blockDim.x) {
        const float x = mu[z0 + t], y = lam[z0 + t];
        const _Float16 hm = bm != 0.0f ? (_Float16)(float)((double)x / (double)bm - 1.0) : (_Float16)0.0f;
        const _Float16 hl = bl != 0.0f ? (_Float16)(float)((double)y / (double)bl - 1.0) : (_Float16)0.0f;
        dst[z0 + t] = (unsigned)__builtin_bit_cast(unsigned short, hm) |
                      ((unsigned)__builtin_bit_cast(unsigned short, hl) << 16);
    }
}

__global__ void __launch_bounds__(256) k_absmax(const float *a, const long long n, float *part)
{
    __shared__ float red[256];
    float mx = 0.0f;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        mx = fmaxf(mx, fabsf(a[t]));
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// FP16 SGT series: one wave per (kept point, component).  scale = 2^e with
// the series' largest magnitude below 2^15 * scale; values stored as the
// FP16 nearest (ties to even) to x / scale.
__global__ void __launch_bounds__(64)
k_sgt_fp16(const float *sgt, const int *kidx, const int nkept, const int nsamp_max, const int nsamp,
           float *scales, _Float16 *out)
{
    const int item = blockIdx.x;  // kept point * 6 + component
    if (item >= nkept * 6) return;
    const int n = kidx[item / 6], c = item % 6;
    const float *src = sgt + ((size_t)n * 6 + c) * nsamp_max;
    float mx = 0.0f;
    for (int t = threadIdx.x; t < nsamp; t += 64) mx = fmaxf(mx, fabsf(src[t]));
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    int ex = 0;
    if (mx > 0.0f) frexpf(mx, &ex);   // mx = f * 2^ex, f in [0.5, 1)
    ex = max(ex, -134);               // scale 2^(ex-15) >= 2^-149 stays representable
    const float scale = mx > 0.0f ? ldexpf(1.0f, ex - 15) : 0.0f;
    if (threadIdx.x == 0) scales[(size_t)n * 6 + c] = scale;
    _Float16 *dst = out + ((size_t)n * 6 + c) * nsamp;
    // ldexpf scales exactly, also where 2^(15-ex) itself would overflow FP32
    for (int t = threadIdx.x; t < nsamp; t += 64)
        dst[t] = (_Float16)(mx > 0.0f ? ldexpf(src[t], 15 - ex) : 0.0f);
}

hipError_t launch_sgt_fp16(const float *sgt, const int *kidx, int nkept, int nsamp_max, int nsamp, float *scales,
                           _Float16 *out, hipStream_t st)
{
    if (nkept <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sgt_fp16, dim3(nkept * 6), dim3(64), 0, st, sgt, kidx, nkept, nsamp_max, nsamp, scales, out);
    return hipGetLastError();
}

// RWG component order and signs of SgtHead/src/reformat_awp.c:38-62 for
// records g0 .. g0+nrec-1 of the raw file: output component co takes input
// component (1 0 2 3 5 4)[co]; the Z source first multiplies every value by
// -2, XZ/YZ are then negated.  Dropped points and samples beyond the full
// WRITE_STEP buffers are holes (zeros) in the raw file.
__global__ void __launch_bounds__(256)
k_sgt_rwg(const float *sgt, const int *keep, const int g0, const int nrec, const int nsamp_max, const int nt_out,
          const int navail, const int zsrc, float *out)
{
    const long long total = (long long)nrec * 6 * nt_out;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (long long)gridDim.x * blockDim.x) {
        const int t = (int)(q % nt_out);
        const long long r = q / nt_out;
        const int co = (int)(r % 6);
        const int g = (int)(r / 6);
        const int ci = co == 0 ? 1 : co == 1 ? 0 : co == 4 ? 5 : co == 5 ? 4 : co;
        float v = (keep[g] && t < navail) ? sgt[((size_t)(g0 + g) * 6 + ci) * nsamp_max + t] : 0.0f;
        if (zsrc) v = -2.0f * v;
        if (co >= 4) v = -1.0f * v;
        out[q] = v;
    }
}

hipError_t launch_sgt_rwg(const float *sgt, const int *keep, int g0, int nrec, int nsamp_max, int nt_out,
                          int navail, int zsrc, float *out, hipStream_t st)
{
    if (nrec <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sgt_rwg, dim3(2048), dim3(256), 0, st, sgt, keep, g0, nrec, nsamp_max, nt_out, navail,
                       zsrc, out);
    return hipGetLastError();
}

// SgtHead/src/reformat_velocity.c:62-72 on the device: the RWG arrays
// (x fastest, then z, then y; km/s, g/cm^3) are, in AWP terms, [X][Z][Y]
// with AWP y fastest; the AWP media file is [Z][Y][X][3] (vp vs rho, m/s,
// kg/m^3).  32x32 (x, y) tiles of one z plane go through LDS so both the
// reads (along y) and the writes (along x) are contiguous.
__global__ void __launch_bounds__(256)
k_rwg_media(const float *p, const float *s, const float *d, const int NX, const int NY, const int NZ, float *out)
{
    __shared__ float tile[3][32][33];
    const int x0 = blockIdx.x * 32, y0 = blockIdx.y * 32, k = blockIdx.z;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const int x = x0 + r, y = y0 + tx;
        if (x < NX && y < NY) {
            const size_t q = ((size_t)x * NZ + k) * NY + y;
            tile[0][r][tx] = p[q];
            tile[1][r][tx] = s[q];
            tile[2][r][tx] = d[q];
        }
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int y = y0 + r, x = x0 + tx;
        if (x < NX && y < NY) {
            float *o = out + (((size_t)k * NY + y) * NX + x) * 3;
            o[0] = tile[0][tx][r] * 1000.0f;
            o[1] = tile[1][tx][r] * 1000.0f;
            o[2] = tile[2][tx][r] * 1000.0f;
        }
    }
}

hipError_t launch_rwg_media(const float *p, const float *s, const float *d, int NX, int NY, int NZ, float *out,
                            hipStream_t st)
{
    dim3 grd((NX + 31) / 32, (NY + 31) / 32, NZ);
    hipLaunchKernelGGL(k_rwg_media, grd, dim3(256), 0, st, p, s, d, NX, NY, NZ, out);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_to_half(const float *src, _Float16 *dst, const float scale, const long long n)
{
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        dst[t] = (_Float16)(src[t] * scale);
}

// qp and qs of every point as one word of two FP16 halves (qp low, qs high),
// each rounded as k_to_half rounds it
__global__ void __launch_bounds__(256)
k_to_qpair(const float *qp, const float *qs, unsigned *dst, const float sp, const float ss, const long long n)
{
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const _Float16 a = (_Float16)(qp[t] * sp), b = (_Float16)(qs[t] * ss);
        dst[t] = (unsigned)__builtin_bit_cast(unsigned short, a) | ((unsigned)__builtin_bit_cast(unsigned short, b) << 16);
    }
}

__global__ void __launch_bounds__(256)
k_pack_r(const Geo g, const _Float16 *src, const signed char *ex, const int q, const int nseg, const int kext,
         const int k0, float *dst)
{
    const int i = blockIdx.x * TX + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1 + k0;
    if (i > g.nx || j > g.ny || k > g.nz + kext) return;
    // levels above the surface are never written: zero
    const float v = k <= g.nz ? r_decode(src[gidx(g, i, j, k)], ex[rexp_idx(g, nseg, (i - 1) / TX, j, k) * 8 + q])
                              : 0.0f;
    dst[((size_t)(k - 1 - k0) * g.ny + (j - 1)) * g.nx + (i - 1)] = v;
}

// levels k0+1 .. k0+gridDim.z of d1, mu or lam in mixed storage (FP16
// deviations under the level profile bar[k + 1]), as [k - 1 - k0][j - 1][i - 1]
__global__ void __launch_bounds__(256)
k_pack_media(const Geo g, const _Float16 *src, const float *bar, const int kext, const int k0, float *dst)
{
    const int i = blockIdx.x * TX + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1 + k0;
    if (i > g.nx || j > g.ny || k > g.nz + kext) return;
    dst[((size_t)(k - 1 - k0) * g.ny + (j - 1)) * g.nx + (i - 1)] = mdec(src[gidx(g, i, j, k)], bar[k + 1]);
}

// mu (half 0) or lam (half 1) of the mixed-storage (mu, lam) pairs
__global__ void __launch_bounds__(256)
k_pack_media2(const Geo g, const unsigned *src, const int half, const float *bar, const int kext, const int k0,
              float *dst)
{
    const int i = blockIdx.x * TX + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1 + k0;
    if (i > g.nx || j > g.ny || k > g.nz + kext) return;
    const unsigned w = src[gidx(g, i, j, k)];
    dst[((size_t)(k - 1 - k0) * g.ny + (j - 1)) * g.nx + (i - 1)] = mdec(half ? h_hi(w) : h_lo(w), bar[k + 1]);
}

// levels k0+1 .. k0+gridDim.z of one field, as [k - 1 - k0][j - 1][i - 1]
template <class T>
__global__ void __launch_bounds__(256)
k_pack_field(const Geo g, const T *src, const float inv_scale, const int kext, const int k0, float *dst)
{
    const int i = blockIdx.x * TX + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1 + k0;
    if (i > g.nx || j > g.ny || k > g.nz + kext) return;
    dst[((size_t)(k - 1 - k0) * g.ny + (j - 1)) * g.nx + (i - 1)] = ld(src[gidx(g, i, j, k)]) * inv_scale;
}

// one half (0 qp, 1 qs) of the mixed-storage (qp, qs) pairs, unscaled
__global__ void __launch_bounds__(256)
k_pack_qpair(const Geo g, const unsigned *src, const int half, const float inv_scale, const int kext, const int k0,
             float *dst)
{
    const int i = blockIdx.x * TX + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1 + k0;
    if (i > g.nx || j > g.ny || k > g.nz + kext) return;
    const unsigned w = src[gidx(g, i, j, k)];
    dst[((size_t)(k - 1 - k0) * g.ny + (j - 1)) * g.nx + (i - 1)] = (half ? q_hi(w) : q_lo(w)) * inv_scale;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_source(const Geo &g, const State &s, int igreen, int npsrc, const long long *psrc_idx,
                         const long long *psrc_top, const float *stf, int nst, int istep, float vtst,
                         float vtst1, bool apply, float *trace, long long tpos, hipStream_t st)
{
    hipLaunchKernelGGL(k_source, dim3(1), dim3(64), 0, st, s, igreen, npsrc, psrc_idx, psrc_top, g.sz, stf, nst,
                       istep, vtst, vtst1, apply ? 1 : 0, trace, tpos);
    return hipGetLastError();
}

hipError_t launch_sgt_coef(const Geo &g, const State &s, int nkept, const long long *kpos, float *sg, hipStream_t st)
{
    if (nkept <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sgt_coef, dim3((nkept + 255) / 256), dim3(256), 0, st, s, g.sz, g.pz, nkept, kpos, sg);
    return hipGetLastError();
}

hipError_t launch_sgt_gather(const State &s, int nkept, const long long *kpos, const int *kidx,
                             const float *sg, float *out, int nsamp_max, int sample, hipStream_t st)
{
    if (nkept <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sgt_gather, dim3((nkept + 255) / 256), dim3(256), 0, st, s, nkept, kpos, kidx, sg, out,
                       nsamp_max, sample);
    return hipGetLastError();
}

void set_block_caps(Launch &l, int device)
{
    int cu = 0, blk = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) != hipSuccess) cu = 0;
    if (hipDeviceGetAttribute(&blk, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) blk = 0;
    l.lds_v = lds_cap(l.bpc_v, cu, blk);
    l.lds_s = lds_cap(l.bpc_s, cu, blk);
}

// blocks of more than 512 threads (64 x 16 tiles, 128 x 8 tiles): the
// 1024-thread variants (byte-offset stress form, no z unroll; choose_launch
// and the tuner keep to that)
static bool tall(const Launch &l) { return l.tile_y * l.tile_xw > 8; }

template <class P, int U>
static int split_grid_t(int device, const Launch &l, int *gs)
{
    int nv = 0, ns = 0, ncu = 0;
    const int nt = TX * l.tile_y * l.tile_xw;
    // blocks of the variants in use that fit on a CU at once (a run whose
    // arrays need 64-bit offsets walks more tiles per block)
    hipError_t ev, es;
    if (tall(l)) {
        ev = l.o32_v ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nv, k_velocity_batch<P, 1, 1024, true>, nt, l.lds_v)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nv, k_velocity_batch<P, 1, 1024, false>, nt,
                                                                    l.lds_v);
        es = hipOccupancyMaxActiveBlocksPerMultiprocessor(&ns, k_stress_batch<P, 1, true, 1024>, nt, l.lds_s);
    } else {
        ev = l.o32_v ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nv, k_velocity_batch<P, U, 512, true>, nt, l.lds_v)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nv, k_velocity_batch<P, U, 512, false>, nt, l.lds_v);
        es = l.o32_s ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&ns, k_stress_batch<P, U, true, 512>, nt, l.lds_s)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&ns, k_stress_batch<P, U, false, 512>, nt,
                                                                    l.lds_s);
    }
    if (ev != hipSuccess) nv = 1;
    if (es != hipSuccess) ns = 1;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) ncu = 256;
    ncu = std::max(8, ncu);
    *gs = std::max(1, ns) * ncu;
    return std::max(1, nv) * ncu;
}

void split_batch_grids(bool mixed, const Launch &l, int device, int *grid_v, int *grid_s)
{
    const bool u2 = l.unroll == 2;
    if (mixed)
        *grid_v = u2 ? split_grid_t<StoreMixed, 2>(device, l, grid_s)
                     : split_grid_t<StoreMixed, 1>(device, l, grid_s);
    else
        *grid_v = u2 ? split_grid_t<StoreF32, 2>(device, l, grid_s)
                     : split_grid_t<StoreF32, 1>(device, l, grid_s);
}

// whole XCD rounds of persistent blocks, no more than the tiles need (the
// static walk); claimed launches: no more blocks than tiles
static int persistent_grid(int grid, int per, int total, bool claimed)
{
    if (claimed) return std::max(1, std::min(grid, total));
    return std::max(8, std::min(grid, 8 * per)) / 8 * 8;
}

template <class P, int U, int MAXT>
static void vel_batch_t(const Launch &l, bool ov, const SJob *d_jobs, int njobs, int total, int per, int grid_v,
                        unsigned long long *claim, hipStream_t st)
{
    const dim3 grid(persistent_grid(grid_v, per, total, claim)), block(TX, l.tile_y, l.tile_xw);
    if (ov)
        hipLaunchKernelGGL((k_velocity_batch<P, U, MAXT, true>), grid, block, l.lds_v, st, d_jobs, njobs, total,
                           per, claim, l.xrot & 7);
    else
        hipLaunchKernelGGL((k_velocity_batch<P, U, MAXT, false>), grid, block, l.lds_v, st, d_jobs, njobs, total,
                           per, claim, l.xrot & 7);
}

template <class P>
static hipError_t velocity_stage(const Launch &l, bool ov, const SJob *d_jobs, int njobs, int total, int per,
                                 int grid_v, unsigned long long *claim, hipStream_t st)
{
    if (tall(l)) {
        if (l.unroll == 2) return hipErrorInvalidValue;   // no such variant
        vel_batch_t<P, 1, 1024>(l, ov, d_jobs, njobs, total, per, grid_v, claim, st);
    } else if (l.unroll == 2) {
        vel_batch_t<P, 2, 512>(l, ov, d_jobs, njobs, total, per, grid_v, claim, st);
    } else {
        vel_batch_t<P, 1, 512>(l, ov, d_jobs, njobs, total, per, grid_v, claim, st);
    }
    return hipSuccess;
}

template <class P, int U, bool O32, int MAXT>
static void str_batch_t(const Launch &l, const SJob *d_jobs, int njobs, int total, int per, int grid_s,
                        unsigned long long *claim, hipStream_t st)
{
    hipLaunchKernelGGL((k_stress_batch<P, U, O32, MAXT>), dim3(persistent_grid(grid_s, per, total, claim)),
                       dim3(TX, l.tile_y, l.tile_xw), l.lds_s, st, d_jobs, njobs, total, per, claim, l.xrot & 7);
}

template <class P>
static hipError_t stress_stage(const Launch &l, bool os, const SJob *d_jobs, int njobs, int total, int per,
                               int grid_s, unsigned long long *claim, hipStream_t st)
{
    if (tall(l)) {
        if (!os || l.unroll == 2) return hipErrorInvalidValue;   // no such variant
        str_batch_t<P, 1, true, 1024>(l, d_jobs, njobs, total, per, grid_s, claim, st);
        return hipSuccess;
    }
    if (l.unroll == 2) {
        if (os) str_batch_t<P, 2, true, 512>(l, d_jobs, njobs, total, per, grid_s, claim, st);
        else str_batch_t<P, 2, false, 512>(l, d_jobs, njobs, total, per, grid_s, claim, st);
    } else {
        if (os) str_batch_t<P, 1, true, 512>(l, d_jobs, njobs, total, per, grid_s, claim, st);
        else str_batch_t<P, 1, false, 512>(l, d_jobs, njobs, total, per, grid_s, claim, st);
    }
    return hipSuccess;
}

hipError_t launch_split_velocity(bool mixed, const Launch &l, bool o32, const SJob *d_jobs, int njobs, int total,
                                 int grid_v, unsigned long long *claim, hipStream_t st)
{
    if (njobs <= 0 || total <= 0) return hipSuccess;
    const int per = (total + 7) / 8;
    const bool ov = o32 && l.o32_v;
    const hipError_t e = mixed ? velocity_stage<StoreMixed>(l, ov, d_jobs, njobs, total, per, grid_v, claim, st)
                               : velocity_stage<StoreF32>(l, ov, d_jobs, njobs, total, per, grid_v, claim, st);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

hipError_t launch_split_fs(bool mixed, const SJob *d_jobs, int njobs, int fs_total, hipStream_t st)
{
    if (njobs <= 0 || fs_total <= 0) return hipSuccess;
    if (mixed)
        hipLaunchKernelGGL(k_fs_velocity_batch<StoreMixed>, dim3(fs_total), dim3(TX, 4), 0, st, d_jobs, njobs);
    else
        hipLaunchKernelGGL(k_fs_velocity_batch<StoreF32>, dim3(fs_total), dim3(TX, 4), 0, st, d_jobs, njobs);
    return hipGetLastError();
}

hipError_t launch_split_stress(bool mixed, const Launch &l, bool o32, const SJob *d_jobs, int njobs, int total,
                               int grid_s, unsigned long long *claim, hipStream_t st)
{
    if (njobs <= 0 || total <= 0) return hipSuccess;
    const int per = (total + 7) / 8;
    const bool os = o32 && l.o32_s;
    const hipError_t e = mixed ? stress_stage<StoreMixed>(l, os, d_jobs, njobs, total, per, grid_s, claim, st)
                               : stress_stage<StoreF32>(l, os, d_jobs, njobs, total, per, grid_s, claim, st);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

hipError_t launch_step_split_batch(bool mixed, const Launch &l, bool o32, const SJob *d_jobs, int njobs,
                                   int total, int fs_total, int grid_v, int grid_s, unsigned long long *claim,
                                   hipStream_t st)
{
    hipError_t e = launch_split_velocity(mixed, l, o32, d_jobs, njobs, total, grid_v, claim, st);
    if (e == hipSuccess) e = launch_split_fs(mixed, d_jobs, njobs, fs_total, st);
    if (e == hipSuccess) e = launch_split_stress(mixed, l, o32, d_jobs, njobs, total, grid_s, claim, st);
    return e;
}

hipError_t launch_aux_batch(const AuxJob *d_jobs, int njobs, int max_kept, int par, hipStream_t st)
{
    if (njobs <= 0) return hipSuccess;
    const int gx = std::max(1, std::min((max_kept + 255) / 256, 1024));
    hipLaunchKernelGGL(k_src_batch, dim3(njobs), dim3(64), 0, st, d_jobs, par);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_aux_batch, dim3(gx, njobs), dim3(256), 0, st, d_jobs, par);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_count_batch, dim3((njobs + 63) / 64), dim3(64), 0, st, d_jobs, njobs);
    return hipGetLastError();
}

hipError_t launch_media_setup(const Geo &g, const Media32 &m, const float *media, const MediaConst &mc, hipStream_t st)
{
    dim3 blk(TX, 4);
    dim3 grd((g.nx + TX - 1) / TX, (g.ny + 3) / 4, g.nz);
    hipLaunchKernelGGL(k_media_setup, grd, blk, 0, st, g, m, media, mc);
    return hipGetLastError();
}

hipError_t launch_media_ghosts(const Geo &g, const Media32 &m, hipStream_t st)
{
    dim3 blk(TX, 4);
    dim3 grd((g.nx + 2 + TX - 1) / TX, (g.ny + 2 + 3) / 4, g.nz + 2);
    hipLaunchKernelGGL(k_media_ghosts, grd, blk, 0, st, g, m);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_media_corner, dim3(1), dim3(64), 0, st, g, m);
    return hipGetLastError();
}

hipError_t launch_level_range(const Geo &g, const float *a, unsigned *range, hipStream_t st)
{
    hipLaunchKernelGGL(k_level_range, dim3((g.ny + 15) / 16, g.nz), dim3(256), 0, st, g, a, range);
    return hipGetLastError();
}

hipError_t launch_media_encode(const Geo &g, const float *src, _Float16 *dst, const float *bar, hipStream_t st)
{
    const int bx = (int)std::min<long long>((g.sz + 255) / 256, 1024);
    hipLaunchKernelGGL(k_media_encode, dim3(bx, g.pz), dim3(256), 0, st, src, dst, bar, g.sz);
    return hipGetLastError();
}

hipError_t launch_media_encode2(const Geo &g, const float *mu, const float *lam, unsigned *dst, const float *bar_mu,
                                const float *bar_lam, hipStream_t st)
{
    const int bx = (int)std::min<long long>((g.sz + 255) / 256, 1024);
    hipLaunchKernelGGL(k_media_encode2, dim3(bx, g.pz), dim3(256), 0, st, mu, lam, dst, bar_mu, bar_lam, g.sz);
    return hipGetLastError();
}

hipError_t launch_absmax(const float *a, long long n, float *part, int nblk, hipStream_t st)
{
    hipLaunchKernelGGL(k_absmax, dim3(nblk), dim3(256), 0, st, a, n, part);
    return hipGetLastError();
}

hipError_t launch_to_half(const float *src, _Float16 *dst, float scale, long long n, hipStream_t st)
{
    hipLaunchKernelGGL(k_to_half, dim3(4096), dim3(256), 0, st, src, dst, scale, n);
    return hipGetLastError();
}

hipError_t launch_to_qpair(const float *qp, const float *qs, unsigned *dst, float sp, float ss, long long n,
                           hipStream_t st)
{
    hipLaunchKernelGGL(k_to_qpair, dim3(4096), dim3(256), 0, st, qp, qs, dst, sp, ss, n);
    return hipGetLastError();
}

hipError_t launch_pack_field(bool mixed, const Geo &g, const State &s, int which, int kext, float *dst,
                             hipStream_t st)
{
    return launch_pack_levels(mixed, g, s, which, kext, 0, g.nz + kext, dst, st);
}

hipError_t launch_pack_levels(bool mixed, const Geo &g, const State &s, int which, int kext, int k0, int nk,
                              float *dst, hipStream_t st)
{
    if (nk <= 0) return hipSuccess;
    dim3 blk(TX, 4);
    dim3 grd((g.nx + TX - 1) / TX, (g.ny + 3) / 4, nk);
    const float *f32[9] = {s.u1, s.v1, s.w1, s.xx, s.yy, s.zz, s.xy, s.xz, s.yz};
    void *rr[6] = {s.r1, s.r2, s.r3, s.r4, s.r5, s.r6};
    const void *mm[3] = {s.d1, s.mu, s.lam};
    void *qq[2] = {s.qp, s.qs};
    const float iqs[2] = {s.ms.iqp, s.ms.iqs};
    if (which < 9) {
        hipLaunchKernelGGL(k_pack_field<float>, grd, blk, 0, st, g, f32[which], 1.0f, kext, k0, dst);
    } else if (which >= 15 && which < 18) {
        if (mixed)
        {
            if (which == 15)
                hipLaunchKernelGGL(k_pack_media, grd, blk, 0, st, g, (const _Float16 *)mm[0], s.mbar, kext, k0, dst);
            else
                hipLaunchKernelGGL(k_pack_media2, grd, blk, 0, st, g, (const unsigned *)s.mu, which - 16,
                                   s.mbar + (size_t)(which - 15) * g.pz, kext, k0, dst);
        }
        else
            hipLaunchKernelGGL(k_pack_field<float>, grd, blk, 0, st, g, (const float *)mm[which - 15], 1.0f, kext,
                               k0, dst);
    } else if (which < 15) {
        if (mixed)
            hipLaunchKernelGGL(k_pack_r, grd, blk, 0, st, g, (const _Float16 *)rr[which - 9],
                               s.rexp, which - 9, s.nseg, kext, k0, dst);
        else
            hipLaunchKernelGGL(k_pack_field<float>, grd, blk, 0, st, g, (const float *)rr[which - 9], 1.0f, kext, k0,
                               dst);
    } else {
        if (mixed)
            hipLaunchKernelGGL(k_pack_qpair, grd, blk, 0, st, g, (const unsigned *)s.qp, which - 18,
                               iqs[which - 18], kext, k0, dst);
        else
            hipLaunchKernelGGL(k_pack_field<float>, grd, blk, 0, st, g, (const float *)qq[which - 18], 1.0f, kext,
                               k0, dst);
    }
    return hipGetLastError();
}

}  // namespace awp
