// awp_fused.hip -- the whole time step in one sweep of the volume.
//
// The two-stage step (awp_kernels.hip) streams the stresses and velocities
// through HBM twice per step: once to update the velocities, once more to
// update the stresses from them.  Here one block marches a column tile of the
// volume bottom-up and, at every level m, first computes the new velocities
// of level m, then the new stresses of level m-2 (the highest velocity level
// a stress at s needs is s+2).  The old state is read from buffer `a` and the
// new state written to buffer `b` (the caller swaps them every step), so
// every array is read once and written once per step:
//   FP32 storage: 140 B per point-update (176 B for the two-stage step)
//   mixed storage: 116 B (FP16 r1..r6 under per-row exponents, FP16 qp/qs).
//
// Stresses at the tile's edge need new velocities of the neighbouring tiles
// (two columns / rows on each side).  The block recomputes them: its waves
// are TY "main" rows (velocity and stress of the tile), 4 halo rows (the two
// rows below and above the tile) and one wave for the 4 halo columns, all of
// which compute velocities of level m into a ring of six LDS planes.  The
// stress of level s reads the new velocities of levels s-2..s+2 from the ring
// (x, y and z neighbours alike).  One barrier per level separates writing
// level m from reading levels m-4..m; the sixth plane keeps the next level's
// writes apart from them without a second barrier.  The halo velocities are
// the same IEEE operations on the same inputs as the owner tile's, so they
// are bit-identical to it.
//
// The old stresses and density along z at the thread's own column are
// carried in registers (z queues); x/y neighbours of the old state and the
// media come from the lines neighbouring lanes just loaded.  Arithmetic statements and their order are those of the
// two-stage kernels, which follow the reference CPU solver (operator.f:92-306,
// viscoop.f:49-191, cerjan.f:249-328), so results are bit-identical to them.

#include "awp_kernels.hpp"

#include <algorithm>

namespace awp {

namespace {

// PFC: the variant code of step_tile (3: two levels per barrier)
template <int TY, int PFC>
struct FShape {
    static constexpr int RW = TX + 4;          // region columns i0-2 .. i0+65
    static constexpr int RH = TY + 4;          // region rows j0-2 .. j0+TY+1
    static constexpr int PLANE = RW * RH;
    static constexpr int SLOT = 3 * PLANE;     // u, v, w of one level
    static constexpr int NSLOT = PFC == 3 ? 8 : 6;   // ring of levels
    static constexpr int NW = TY + 5;          // waves per block
};

// Same XCD-contiguous tile order as the two-stage kernels.
__device__ inline bool ftile_of(const FTiles &t, int &bx, int &by, int &bz)
{
    const int n = blockIdx.x;
    const int tile = (n & 7) * t.per + (n >> 3);
    if (tile >= t.nbx * t.nby * t.nbz) return false;
    patch_tile(tile, t.nbx, t.nby, t.pw, t.ph, bx, by, bz);
    return true;
}

}  // namespace

// Old-state values the velocity of one level needs besides the z queues.
struct VelIn {
    float xz_p1, yz_p1, zz_p2;                 // queue heads (levels m+1, m+2)
    float dA, dB, dC, dD;                      // d1 at m: (i,j) (i,j-1) (i+1,j) (i+1,j-1)
    float xx_0, xx_m1, xx_p1, xx_m2;
    float xy_0, xy_jm1, xy_jp1, xy_jm2, xy_ip1, xy_ip2, xy_im1;
    float yy_0, yy_jp1, yy_jp2, yy_jm1;
    float xz_ip1, xz_ip2, xz_im1;
    float yz_jm1, yz_jp1, yz_jm2;
    float u, v, w;                             // old velocities of the point
    float fz;                                  // sponge profile at m
};

// Values the stress of one level needs from HBM (new velocities come from LDS).
struct StrIn {
    float lA, lB, lC, lD, mA, mB, mC, mD;      // lam, mu on the 2x2 cell
    float pA, pB, pC, pD, sA, sB, sC, sD;      // qp, qs
    float r1, r2, r3, r4, r5, r6;
    float xx, yy, xy;
    float fz;                                  // sponge profile at s
    unsigned long long ew;                     // mixed: the row segment's exponents of r1..r6
};

// PF selects how far loads run ahead of their use: 0 = each stage loads
// what it needs; 1 = the stress stage's loads are issued before the velocity
// stage's arithmetic; 2 = as 1, and the next level's velocity loads are
// issued before the stress arithmetic.
// PFC = 3: no prefetch, and the velocities of two levels per barrier (then
// the stresses of two levels), which halves the block's barriers.
// One column tile (bx, by, z chunk bz) of one run: the whole block calls it.
// The volume arrays of s advanced by `e` elements (the sponge profiles and
// the exponent slots keep their own indexing).
template <class P>
__device__ __forceinline__ GState rebased(const State &s0, const long long e)
{
    GState s = as_global(s0);
    G<float> **f[9] = {&s.u1, &s.v1, &s.w1, &s.xx, &s.yy, &s.zz, &s.xy, &s.xz, &s.yz};
#pragma unroll
    for (int q = 0; q < 9; ++q) *f[q] += e;
    using R = typename P::R;
    using Q = typename P::Q;
    using M = typename P::M;
    G<void> **r[6] = {&s.r1, &s.r2, &s.r3, &s.r4, &s.r5, &s.r6};
#pragma unroll
    for (int q = 0; q < 6; ++q) *r[q] = (G<R> *)*r[q] + e;
    if constexpr (P::mixed) {
        // d1 FP16; (mu, lam) pairs, one 4-byte word a point
        s.d1 = (G<M> *)s.d1 + e;
        s.mu = (G<unsigned> *)s.mu + e;
        s.lam = s.mu;
    } else {
        G<void> **m[3] = {&s.d1, &s.mu, &s.lam};
#pragma unroll
        for (int q = 0; q < 3; ++q) *m[q] = (G<M> *)*m[q] + e;
    }
    if constexpr (P::mixed) {
        // (qp, qs) pairs, one 4-byte word a point
        s.qp = (G<unsigned> *)s.qp + e;
        s.qs = s.qp;
    } else {
        s.qp = (G<Q> *)s.qp + e;
        s.qs = (G<Q> *)s.qs + e;
    }
    return s;
}

template <class P, int TY, int PFC>
__device__ __forceinline__ void step_tile(const Geo &g, const State &a0, const State &b0, const Coef &c,
                                          const int kc, const int bx, const int by, const int bz, float *lds)
{
    // 32-bit byte offsets count from level lz (the chunk's lowest level read)
    const int lz = max(bz * kc + 1 - 4, -1);
    const long long rb = (long long)lz * g.sz;
    const GState a = rebased<P>(a0, rb), b = rebased<P>(b0, rb);
    using Sh = FShape<TY, PFC>;
    constexpr int PF = PFC == 3 ? 0 : PFC;   // prefetch depth
    constexpr int LV = PFC == 3 ? 2 : 1;     // levels per barrier
    using Q = typename P::Q;
    using R = typename P::R;
    const int lane = threadIdx.x, wv = threadIdx.y;
    const int i0 = bx * TX + 1, j0 = by * TY + 1;
    const int kb = bz * kc + 1;
    const int ke = min(kb + kc - 1, g.nz);

    // this thread's velocity point, as region column rc / row rr
    int rc, rr;
    bool active = true;
    if (wv < TY) {                 // tile rows: velocity and stress
        rc = lane + 2;
        rr = wv + 2;
    } else if (wv < TY + 4) {      // halo rows j0-2, j0-1, j0+TY, j0+TY+1
        const int h = wv - TY;
        rc = lane + 2;
        rr = h < 2 ? h : TY + h;
    } else {                       // halo columns i0-2, i0-1, i0+64, i0+65 of the tile rows
        const int q = lane & 3;
        rc = q < 2 ? q : TX + q;
        rr = 2 + (lane >> 2);
        active = lane < 4 * TY;
    }
    const bool is_main = wv < TY;
    const int i = i0 - 2 + rc, j = j0 - 2 + rr;
    const bool dom = active && i >= 1 && i <= g.nx && j >= 1 && j <= g.ny;
    const bool smain = is_main && dom;
    const long long sy = g.sy, sz = g.sz;
    // byte offset of (i, j, level 0) in a 4-byte array from the rebased
    // pointers, modulo 2^32; level l adds l*sz4 (the offsets of the levels
    // the tile reads, lz .. lz + kc + 8, are exact: below 2^32)
    const unsigned ob = (unsigned)((g.base + (long long)j * sy + i - rb) * 4);
    const unsigned sy4 = (unsigned)(sy * 4), sz4 = (unsigned)(sz * 4);
    using M = typename P::M;
    constexpr int QSH = sizeof(Q) == 2 ? 1 : 0;   // byte offset shift of qp/qs
    constexpr int RSH = sizeof(R) == 2 ? 1 : 0;   // and of r1..r6
    constexpr int MSH = sizeof(M) == 2 ? 1 : 0;   // and of d1, mu, lam
    constexpr int QB = (int)sizeof(Q);
    constexpr int MB = (int)sizeof(M);
    const float c1 = c.c1, c2 = c.c2, dth = c.dth, dt = c.dt, dh1 = c.dh1;
    const float fxy = dom ? a.fx[i] * a.fy[j] : 0.0f;
    // relaxation weights of this column for even / odd depth parity
    // (selected per level: no per-lane indexing of the argument block)
    const int itxy = ((i & 1) * 2 + (j & 1)) * 2;
    const float t1e = c.tau1[itxy], t1o = c.tau1[itxy + 1];
    const float t2e = c.tau2[itxy], t2o = c.tau2[itxy + 1];
    const int lpos = rr * Sh::RW + rc;
    constexpr int PL = Sh::PLANE;
    const G<M> *__restrict__ d1 = (const G<M> *)a.d1;
    const G<M> *__restrict__ lam = (const G<M> *)a.lam;
    const G<M> *__restrict__ mu = (const G<M> *)a.mu;
    const G<unsigned> *__restrict__ mlq = (const G<unsigned> *)a.mu;   // mixed: (mu, lam) pairs
    // profile values of level l (mixed storage; 0 unused in FP32)
    auto mbar = [&](int arr, int l) -> float { return P::mixed ? a.mbar[mbar_idx(g, arr, l)] : 0.0f; };
    const G<Q> *__restrict__ qp = (const G<Q> *)a.qp;
    const G<Q> *__restrict__ qs = (const G<Q> *)a.qs;
    const G<unsigned> *__restrict__ qq = (const G<unsigned> *)a.qp;   // mixed: (qp, qs) pairs
    G<R> *__restrict__ R1 = (G<R> *)a.r1;
    G<R> *__restrict__ R2 = (G<R> *)a.r2;
    G<R> *__restrict__ R3 = (G<R> *)a.r3;
    G<R> *__restrict__ R4 = (G<R> *)a.r4;
    G<R> *__restrict__ R5 = (G<R> *)a.r5;
    G<R> *__restrict__ R6 = (G<R> *)a.r6;

    // loads of the velocity stage at level m (dom threads, -1 <= m <= nz;
    // every address lies in the allocation, the ghost levels hold zeros)
    auto load_vel = [&](int m, VelIn &x) {
        const unsigned o = ob + (unsigned)m * sz4;
        const unsigned ojm1 = o - sy4, ojp1 = o + sy4, ojm2 = ojm1 - sy4, ojp2 = ojp1 + sy4;
        x.xz_p1 = gl(a.xz, o + sz4);
        x.yz_p1 = gl(a.yz, o + sz4);
        x.zz_p2 = gl(a.zz, o + 2 * sz4);
        const float db = mbar(0, m);
        const unsigned om = o >> MSH, omj = ojm1 >> MSH;
        x.dA = mdec(gl(d1, om), db); x.dB = mdec(gl(d1, omj), db);
        x.dC = mdec(gl(d1, om, MB), db); x.dD = mdec(gl(d1, omj, MB), db);
        x.xx_0 = gl(a.xx, o); x.xx_m1 = gl(a.xx, o, -4); x.xx_p1 = gl(a.xx, o, 4); x.xx_m2 = gl(a.xx, o, -8);
        x.xy_0 = gl(a.xy, o); x.xy_jm1 = gl(a.xy, ojm1); x.xy_jp1 = gl(a.xy, ojp1); x.xy_jm2 = gl(a.xy, ojm2);
        x.xy_ip1 = gl(a.xy, o, 4); x.xy_ip2 = gl(a.xy, o, 8); x.xy_im1 = gl(a.xy, o, -4);
        x.yy_0 = gl(a.yy, o); x.yy_jp1 = gl(a.yy, ojp1); x.yy_jp2 = gl(a.yy, ojp2); x.yy_jm1 = gl(a.yy, ojm1);
        x.xz_ip1 = gl(a.xz, o, 4); x.xz_ip2 = gl(a.xz, o, 8); x.xz_im1 = gl(a.xz, o, -4);
        x.yz_jm1 = gl(a.yz, ojm1); x.yz_jp1 = gl(a.yz, ojp1); x.yz_jm2 = gl(a.yz, ojm2);
        x.u = gl(a.u1, o); x.v = gl(a.v1, o); x.w = gl(a.w1, o);
        x.fz = a.fz[m < 0 ? 0 : m];
    };
    // loads of the stress stage at level s (main dom threads)
    auto load_str = [&](int s, StrIn &x) {
        const unsigned o = ob + (unsigned)s * sz4, oj = o - sy4;
        {
            const float lb = mbar(2, s), mb = mbar(1, s);
            if constexpr (P::mixed) {
                const unsigned wA = gl(mlq, o), wB = gl(mlq, o, 4), wC = gl(mlq, oj), wD = gl(mlq, oj, 4);
                x.lA = mdec(h_hi(wA), lb); x.lB = mdec(h_hi(wB), lb); x.lC = mdec(h_hi(wC), lb); x.lD = mdec(h_hi(wD), lb);
                x.mA = mdec(h_lo(wA), mb); x.mB = mdec(h_lo(wB), mb); x.mC = mdec(h_lo(wC), mb); x.mD = mdec(h_lo(wD), mb);
            } else {
                const unsigned mo = o >> MSH, mj = oj >> MSH;
                x.lA = mdec(gl(lam, mo), lb); x.lB = mdec(gl(lam, mo, MB), lb);
                x.lC = mdec(gl(lam, mj), lb); x.lD = mdec(gl(lam, mj, MB), lb);
                x.mA = mdec(gl(mu, mo), mb); x.mB = mdec(gl(mu, mo, MB), mb);
                x.mC = mdec(gl(mu, mj), mb); x.mD = mdec(gl(mu, mj, MB), mb);
            }
        }
        if constexpr (P::mixed) {
            const unsigned wA = gl(qq, o), wB = gl(qq, o, 4), wC = gl(qq, oj), wD = gl(qq, oj, 4);
            x.pA = q_lo(wA); x.pB = q_lo(wB); x.pC = q_lo(wC); x.pD = q_lo(wD);
            x.sA = q_hi(wA); x.sB = q_hi(wB); x.sC = q_hi(wC); x.sD = q_hi(wD);
        } else {
            const unsigned qo = o >> QSH, qj = oj >> QSH;
            x.pA = ld(gl(qp, qo)); x.pB = ld(gl(qp, qo, QB)); x.pC = ld(gl(qp, qj)); x.pD = ld(gl(qp, qj, QB));
            x.sA = ld(gl(qs, qo)); x.sB = ld(gl(qs, qo, QB)); x.sC = ld(gl(qs, qj)); x.sD = ld(gl(qs, qj, QB));
        }
        const unsigned orr = o >> RSH;
        if constexpr (P::mixed) {
            // FP16 values under the row segment's exponent (one per wave)
            const unsigned long long ew = rexp_load(a.rexp, rexp_idx(g, a.nseg, bx, j, s));
            x.ew = ew;
            x.r1 = r_decode(gln(R1, orr), rexp_of(ew, 0)); x.r2 = r_decode(gln(R2, orr), rexp_of(ew, 1));
            x.r3 = r_decode(gln(R3, orr), rexp_of(ew, 2)); x.r4 = r_decode(gln(R4, orr), rexp_of(ew, 3));
            x.r5 = r_decode(gln(R5, orr), rexp_of(ew, 4)); x.r6 = r_decode(gln(R6, orr), rexp_of(ew, 5));
        } else {
            x.r1 = gln(R1, orr); x.r2 = gln(R2, orr); x.r3 = gln(R3, orr);
            x.r4 = gln(R4, orr); x.r5 = gln(R5, orr); x.r6 = gln(R6, orr);
        }
        x.xx = gl(a.xx, o); x.yy = gl(a.yy, o); x.xy = gl(a.xy, o);
        x.fz = a.fz[s];
    };

    // --- z queues of the old stresses (velocity stage) at this column -----
    // levels below -1 lie outside the allocation and are zero
    const int m0 = kb - 2;
    float xz_m2 = 0.f, xz_m1 = 0.f, xz_0 = 0.f;     // xz at m-2, m-1, m
    float yz_m2 = 0.f, yz_m1 = 0.f, yz_0 = 0.f;
    float zz_m2 = 0.f, zz_m1 = 0.f, zz_0 = 0.f, zz_p1 = 0.f;   // zz at m-2 .. m+1
    float dA_m = 0.f, dB_m = 0.f, dC_m = 0.f;       // d1 at m-1: (i,j), (i,j-1), (i+1,j)
    if (dom) {
        for (int l = m0 - 2; l <= m0; ++l) {
            if (l < -1) continue;   // leading levels: the queue already holds their zeros
            const unsigned o = ob + (unsigned)l * sz4;
            const float x = gl(a.xz, o), y = gl(a.yz, o);
            xz_m2 = xz_m1; xz_m1 = xz_0; xz_0 = x;
            yz_m2 = yz_m1; yz_m1 = yz_0; yz_0 = y;
        }
        for (int l = m0 - 1; l <= m0 + 1; ++l) {
            if (l < -1) continue;
            const float z = gl(a.zz, ob + (unsigned)l * sz4);
            zz_m1 = zz_0; zz_0 = zz_p1; zz_p1 = z;
        }
        if (m0 - 1 >= -1) {
            const unsigned o = ob + (unsigned)(m0 - 1) * sz4;
            const float db = mbar(0, m0 - 1);
            dA_m = mdec(gl(d1, o >> MSH), db); dB_m = mdec(gl(d1, (o - sy4) >> MSH), db);
            dC_m = mdec(gl(d1, o >> MSH, MB), db);
        }
    }

    // media of level s-1 on the 2x2 (x,y) cell: A=(i,j) B=(i+1,j) C=(i,j-1) D=(i+1,j-1)
    float lA_m = 0.f, lB_m = 0.f, lC_m = 0.f, lD_m = 0.f;
    float mA_m = 0.f, mB_m = 0.f, mC_m = 0.f, mD_m = 0.f;
    float pA_m = 0.f, pB_m = 0.f, pC_m = 0.f, pD_m = 0.f;
    float sA_m = 0.f, sB_m = 0.f, sC_m = 0.f, sD_m = 0.f;
    if (smain) {
        const unsigned om = ob + (unsigned)(kb - 1) * sz4, omj = om - sy4;
        {
            const float lb = mbar(2, kb - 1), mb = mbar(1, kb - 1);
            if constexpr (P::mixed) {
                const unsigned wA = gl(mlq, om), wB = gl(mlq, om, 4), wC = gl(mlq, omj), wD = gl(mlq, omj, 4);
                lA_m = mdec(h_hi(wA), lb); lB_m = mdec(h_hi(wB), lb); lC_m = mdec(h_hi(wC), lb); lD_m = mdec(h_hi(wD), lb);
                mA_m = mdec(h_lo(wA), mb); mB_m = mdec(h_lo(wB), mb); mC_m = mdec(h_lo(wC), mb); mD_m = mdec(h_lo(wD), mb);
            } else {
                const unsigned mo = om >> MSH, mj = omj >> MSH;
                lA_m = mdec(gl(lam, mo), lb); lB_m = mdec(gl(lam, mo, MB), lb);
                lC_m = mdec(gl(lam, mj), lb); lD_m = mdec(gl(lam, mj, MB), lb);
                mA_m = mdec(gl(mu, mo), mb); mB_m = mdec(gl(mu, mo, MB), mb);
                mC_m = mdec(gl(mu, mj), mb); mD_m = mdec(gl(mu, mj, MB), mb);
            }
        }
        if constexpr (P::mixed) {
            const unsigned wA = gl(qq, om), wB = gl(qq, om, 4), wC = gl(qq, omj), wD = gl(qq, omj, 4);
            pA_m = q_lo(wA); pB_m = q_lo(wB); pC_m = q_lo(wC); pD_m = q_lo(wD);
            sA_m = q_hi(wA); sB_m = q_hi(wB); sC_m = q_hi(wC); sD_m = q_hi(wD);
        } else {
            const unsigned qo = om >> QSH, qj = omj >> QSH;
            pA_m = ld(gl(qp, qo)); pB_m = ld(gl(qp, qo, QB)); pC_m = ld(gl(qp, qj)); pD_m = ld(gl(qp, qj, QB));
            sA_m = ld(gl(qs, qo)); sB_m = ld(gl(qs, qo, QB)); sC_m = ld(gl(qs, qj)); sD_m = ld(gl(qs, qj, QB));
        }
    }
    // final xz/yz/zz of the two levels below (free-surface images)
    float xz_f1 = 0.f, xz_f2 = 0.f, yz_f1 = 0.f, yz_f2 = 0.f, zz_f1 = 0.f;

    VelIn vin = {};
    if constexpr (PF >= 2) {
        if (dom && m0 <= g.nz) load_vel(m0, vin);
    }

    // Old xz / yz / zz of the level a stress stage updates: the bottom of the
    // velocity stage's z queues when that stage computed level s + 2.
    struct Old {
        float xz, yz, zz;
    };
    // velocity stage of level m into the LDS ring, then the z queues advance
    auto vel_level = [&](const int m, Old &old) {
        // ring slot of level l (l >= -1)
        float *slot = lds + ((m + 6) % Sh::NSLOT) * Sh::SLOT;
            // ================= velocity of level m (operator.f:92-244) ==========
            float un = 0.f, vn = 0.f, wn = 0.f;
            float xz_p1 = 0.f, yz_p1 = 0.f, zz_p2 = 0.f;
            if (dom && m <= g.nz) {
                if constexpr (PF == 0) load_vel(m, vin);
                const VelIn &x = vin;
                xz_p1 = x.xz_p1; yz_p1 = x.yz_p1; zz_p2 = x.zz_p2;
                if (m >= 1) {
                    float d, q, u, v, w;
                    d = 0.25f * ((x.dA + x.dB) + (dA_m + dB_m));
                    q = dth / d;
                    u = x.u + q * (
                        c1 * (x.xx_0 - x.xx_m1) + c2 * (x.xx_p1 - x.xx_m2) +
                        c1 * (x.xy_0 - x.xy_jm1) + c2 * (x.xy_jp1 - x.xy_jm2) +
                        c1 * (xz_0 - xz_m1) + c2 * (x.xz_p1 - xz_m2));
                    d = 0.25f * ((x.dA + x.dC) + (dA_m + dC_m));
                    q = dth / d;
                    v = x.v + q * (
                        c1 * (x.xy_ip1 - x.xy_0) + c2 * (x.xy_ip2 - x.xy_im1) +
                        c1 * (x.yy_jp1 - x.yy_0) + c2 * (x.yy_jp2 - x.yy_jm1) +
                        c1 * (yz_0 - yz_m1) + c2 * (x.yz_p1 - yz_m2));
                    d = 0.25f * ((x.dA + x.dC) + (x.dB + x.dD));
                    q = dth / d;
                    w = x.w + q * (
                        c1 * (x.xz_ip1 - xz_0) + c2 * (x.xz_ip2 - x.xz_im1) +
                        c1 * (yz_0 - x.yz_jm1) + c2 * (x.yz_jp1 - x.yz_jm2) +
                        c1 * (zz_p1 - zz_0) + c2 * (x.zz_p2 - zz_m1));
                    const float D = fxy * x.fz;       // addcrjvel, cerjan.f:292-328
                    un = u * D;
                    vn = v * D;
                    wn = w * D;
                    if (is_main && m >= kb && m <= ke) {
                        const unsigned o = ob + (unsigned)m * sz4;
                        gsn(b.u1, o, un);
                        gsn(b.v1, o, vn);
                        gsn(b.w1, o, wn);
                    }
                }
                dA_m = x.dA; dB_m = x.dB; dC_m = x.dC;
            }
            if (m <= g.nz) {
                if (active) {
                    slot[lpos] = un;
                    slot[PL + lpos] = vn;
                    slot[2 * PL + lpos] = wn;
                }
            } else if (is_main) {
                if (m == g.nz + 1 && smain) {
                    // free-surface velocity images fvelxy / fvelz (operator.f:246-306)
                    // from the level-nz plane (and w at nz-1)
                    const float *pl = lds + ((g.nz + 6) % Sh::NSLOT) * Sh::SLOT;
                    const float *pl1 = lds + ((g.nz - 1 + 6) % Sh::NSLOT) * Sh::SLOT;
                    const float *pu = pl, *pv = pl + PL, *pw = pl + 2 * PL;
                    const float w0 = pw[lpos], u0 = pu[lpos], v0 = pv[lpos];
                    const float ut = u0 - (w0 - pw[lpos - 1]);
                    const float vt = v0 - (pw[lpos + Sh::RW] - w0);
                    const float ut_ip1 = (i < g.nx) ? (pu[lpos + 1] - (pw[lpos + 1] - w0)) : 0.0f;
                    const float vt_jm1 = (j > 1) ? (pv[lpos - Sh::RW] - (w0 - pw[lpos - Sh::RW])) : 0.0f;
                    const unsigned o = ob + (unsigned)g.nz * sz4;
                    float xl, xm;
                    if constexpr (P::mixed) {
                        const unsigned w = gl(mlq, o);   // (mu, lam) pair
                        xl = 1.0f / mdec(h_hi(w), mbar(2, g.nz));
                        xm = 2.0f / mdec(h_lo(w), mbar(1, g.nz));
                    } else {
                        xl = 1.0f / mdec(gl(lam, o >> MSH), mbar(2, g.nz));
                        xm = 2.0f / mdec(gl(mu, o >> MSH), mbar(1, g.nz));
                    }
                    const float xl2m = xm + xl;
                    const float wt = pl1[2 * PL + lpos] - (xl / xl2m) * (
                        (ut_ip1 - ut) +
                        (pu[lpos + 1] - u0) +
                        (vt - vt_jm1) +
                        (v0 - pv[lpos - Sh::RW]));
                    un = ut;
                    vn = vt;
                    wn = wt;
                    if (ke == g.nz) {
                        gsn(b.u1, o + sz4, un);
                        gsn(b.v1, o + sz4, vn);
                        gsn(b.w1, o + sz4, wn);
                    }
                }
                // level nz+1: the images; level nz+2: zero (the reference never sets it)
                slot[lpos] = un;
                slot[PL + lpos] = vn;
                slot[2 * PL + lpos] = wn;
            }

        old = {xz_m2, yz_m2, zz_m2};
        // advance the old-stress queues to level m+1
        xz_m2 = xz_m1; xz_m1 = xz_0; xz_0 = xz_p1;
        yz_m2 = yz_m1; yz_m1 = yz_0; yz_0 = yz_p1;
        zz_m2 = zz_m1; zz_m1 = zz_0; zz_0 = zz_p1; zz_p1 = zz_p2;
    };
    // stress stage of level s = m - 2 from the ring (sin_: its loads, when
    // issued ahead)
    auto str_level = [&](const int m, const Old &old, StrIn &sin_) {
        const int s = m - 2;
        const bool do_str = smain && s >= kb;
            // ================= stress of level s = m-2 (viscoop.f:49-191) =======
            float rn[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};   // new memory variables (mixed: stored below)
            unsigned long long ew_old = 0;                     // their row exponents before the update
            if (do_str) {
                if constexpr (PF < 1) load_str(s, sin_);
                const StrIn &x = sin_;
                const unsigned o = ob + (unsigned)s * sz4;
                const float *ps = lds + ((s + 6) % Sh::NSLOT) * Sh::SLOT;          // level s
                const float *pm1 = lds + ((s - 1 + 6) % Sh::NSLOT) * Sh::SLOT;     // s-1
                const float *pm2 = lds + ((s - 2 + 6) % Sh::NSLOT) * Sh::SLOT;     // s-2
                const float *pp1 = lds + ((s + 1 + 6) % Sh::NSLOT) * Sh::SLOT;     // s+1
                const float *pp2 = lds + ((s + 2 + 6) % Sh::NSLOT) * Sh::SLOT;     // s+2
                const float *pu = ps, *pv = ps + PL, *pw = ps + 2 * PL;
                const float uc = pu[lpos], vc = pv[lpos], wc = pw[lpos];
                const float u_ip1 = pu[lpos + 1], u_ip2 = pu[lpos + 2], u_im1 = pu[lpos - 1];
                const float u_jp1 = pu[lpos + Sh::RW], u_jp2 = pu[lpos + 2 * Sh::RW], u_jm1 = pu[lpos - Sh::RW];
                const float v_jm1 = pv[lpos - Sh::RW], v_jp1 = pv[lpos + Sh::RW], v_jm2 = pv[lpos - 2 * Sh::RW];
                const float v_im1 = pv[lpos - 1], v_ip1 = pv[lpos + 1], v_im2 = pv[lpos - 2];
                const float w_im1 = pw[lpos - 1], w_ip1 = pw[lpos + 1], w_im2 = pw[lpos - 2];
                const float w_jp1 = pw[lpos + Sh::RW], w_jp2 = pw[lpos + 2 * Sh::RW], w_jm1 = pw[lpos - Sh::RW];
                // own column along z
                const float uc_p1 = pp1[lpos], uc_p2 = pp2[lpos], uc_m1 = pm1[lpos];
                const float vc_p1 = pp1[PL + lpos], vc_p2 = pp2[PL + lpos], vc_m1 = pm1[PL + lpos];
                const float wc_p1 = pp1[2 * PL + lpos], wc_m1 = pm1[2 * PL + lpos], wc_m2 = pm2[2 * PL + lpos];

                const bool odd = (((g.nz + 1) - s) & 1) != 0;
                const float vx1 = odd ? t1o : t1e;
                const float vx2 = odd ? t2o : t2e;

                float vxl = 8.0f / (x.lA + x.lB + x.lC + x.lD + lA_m + lB_m + lC_m + lD_m);
                const float vxm = 8.0f / (x.mA + x.mB + x.mC + x.mD + mA_m + mB_m + mC_m + mD_m);
                vxl = vxl + 2.0f * vxm;
                float vqpa = 0.125f * (x.pA + x.pB + x.pC + x.pD + pA_m + pB_m + pC_m + pD_m);
                float vqsa = 0.125f * (x.sA + x.sB + x.sC + x.sD + sA_m + sB_m + sC_m + sD_m);
                if constexpr (P::mixed) {
                    vqpa = vqpa * a.ms.iqp;
                    vqsa = vqsa * a.ms.iqs;
                }

                const float vt1 = c1 * (u_ip1 - uc) + c2 * (u_ip2 - u_im1);
                const float vt2 = c1 * (vc - v_jm1) + c2 * (v_jp1 - v_jm2);
                const float vt3 = c1 * (wc - wc_m1) + c2 * (wc_p1 - wc_m2);
                const float vt4 = c1 * (u_jp1 - uc) + c2 * (u_jp2 - u_jm1) +
                                  c1 * (vc - v_im1) + c2 * (v_ip1 - v_im2);
                const float vt5 = c1 * (uc_p1 - uc) + c2 * (uc_p2 - uc_m1) +
                                  c1 * (wc - w_im1) + c2 * (w_ip1 - w_im2);
                const float vt6 = c1 * (vc_p1 - vc) + c2 * (vc_p2 - vc_m1);
                const float vt7 = c1 * (w_jp1 - wc) + c2 * (w_jp2 - w_jm1);

                const float vh1 = -vxm * vqsa;
                const float vtmp = vxl * (vt1 + vt2 + vt3);
                const float va1 = -vqpa * vxl * dh1 * (vt1 + vt2 + vt3) / 2.0f;

                float r1 = x.r1, r2 = x.r2, r3 = x.r3, r4 = x.r4, r5 = x.r5, r6 = x.r6;
                float xx = x.xx, yy = x.yy, xy = x.xy;
                float zz = old.zz, xz = old.xz, yz = old.yz;   // old values of level s

                xx = xx + dth * (vtmp - 2.0f * vxm * (vt2 + vt3)) + dt * r1;
                yy = yy + dth * (vtmp - 2.0f * vxm * (vt1 + vt3)) + dt * r2;
                zz = zz + dth * (vtmp - 2.0f * vxm * (vt1 + vt2)) + dt * r3;

                float vxmu = 2.0f / (x.mA + mA_m);
                float vaas = 0.5f * (x.sA + sA_m);
                if constexpr (P::mixed) vaas = vaas * a.ms.iqs;
                const float vh2_r4 = -vaas * vxmu / 2.0f;
                xy = xy + vxmu * dth * vt4 + dt * r4;
                vxmu = 2.0f / (x.mA + x.mC);
                vaas = 0.5f * (x.sA + x.sC);
                if constexpr (P::mixed) vaas = vaas * a.ms.iqs;
                const float vh2_r5 = -vaas * vxmu / 2.0f;
                xz = xz + vxmu * dth * vt5 + dt * r5;
                vxmu = 2.0f / (x.mA + x.mB);
                vaas = 0.5f * (x.sA + x.sB);
                if constexpr (P::mixed) vaas = vaas * a.ms.iqs;
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

                const float D = fxy * x.fz;       // addcrjstr, cerjan.f:249-290
                xx = xx * D; yy = yy * D; zz = zz * D;
                xz = xz * D; yz = yz * D; xy = xy * D;

                // memory variables in place (own point only); in mixed storage
                // they are encoded below, after the row segment's maximum
                if constexpr (!P::mixed) {
                    gsn(R1, o, r1); gsn(R2, o, r2); gsn(R3, o, r3);
                    gsn(R4, o, r4); gsn(R5, o, r5); gsn(R6, o, r6);
                } else {
                    rn[0] = r1; rn[1] = r2; rn[2] = r3; rn[3] = r4; rn[4] = r5; rn[5] = r6;
                    ew_old = x.ew;
                }
                gsn(b.xx, o, xx); gsn(b.yy, o, yy); gsn(b.xy, o, xy);
                if (s == g.nz) {
                    // fstr (operator.f:575-612): images above the surface, xz = yz = 0 on it
                    const unsigned o1 = o + sz4, o2 = o1 + sz4;
                    gsn(b.zz, o, zz);
                    gsn(b.zz, o1, -zz);
                    gsn(b.zz, o2, -zz_f1);
                    gsn(b.xz, o, 0.0f);
                    gsn(b.yz, o, 0.0f);
                    gsn(b.xz, o1, -xz_f1);
                    gsn(b.xz, o2, -xz_f2);
                    gsn(b.yz, o1, -yz_f1);
                    gsn(b.yz, o2, -yz_f2);
                } else {
                    gsn(b.zz, o, zz); gsn(b.xz, o, xz); gsn(b.yz, o, yz);
                }
                xz_f2 = xz_f1; xz_f1 = xz;
                yz_f2 = yz_f1; yz_f1 = yz;
                zz_f1 = zz;
                lA_m = x.lA; lB_m = x.lB; lC_m = x.lC; lD_m = x.lD;
                mA_m = x.mA; mB_m = x.mB; mC_m = x.mC; mD_m = x.mD;
                pA_m = x.pA; pB_m = x.pB; pC_m = x.pC; pD_m = x.pD;
                sA_m = x.sA; sB_m = x.sB; sC_m = x.sC; sD_m = x.sD;
            }
            if constexpr (P::mixed) {
                // whole waves (one row segment each): lanes outside the domain hold 0
                if (is_main && s >= kb) {
                    const unsigned orr = (ob + (unsigned)s * sz4) >> RSH;
                    G<R> *RR[6] = {R1, R2, R3, R4, R5, R6};
                    // the row's exponents before the update: lane 0 (i = i0,
                    // inside the domain whenever the row is) loaded the slot
                    const unsigned long long ew_row =
                        (unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)ew_old) |
                        ((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(ew_old >> 32))
                         << 32);
                    const unsigned long long en = row_exponents6(rn, do_str, ew_row);
    #pragma unroll
                    for (int q = 0; q < 6; ++q)
                        if (do_str) gsn(RR[q], orr, r_encode(rn[q], rexp_of(en, q)));
                    if (do_str && lane == 0)
                        ((G<unsigned long long> *)a.rexp)[rexp_idx(g, a.nseg, bx, j, s)] = en;
                }
            }

    };

    if constexpr (LV == 1) {
        for (int m = m0; m <= ke + 2; ++m) {
            const int s = m - 2;
            const bool do_str = smain && s >= kb;
            StrIn sin_ = {};
            if constexpr (PF >= 1) {
                if (do_str) load_str(s, sin_);
            }
            if constexpr (PF == 1) {
                if (dom && m <= g.nz) load_vel(m, vin);
                // keep this level's loads together, ahead of all arithmetic, so
                // they are in flight at once
                __builtin_amdgcn_sched_barrier(0);
            }
            Old old;
            vel_level(m, old);
            __syncthreads();

            if constexpr (PF >= 2) {
                if (dom && m + 1 <= g.nz) load_vel(m + 1, vin);
                __builtin_amdgcn_sched_barrier(0);
            }

            str_level(m, old, sin_);
        }
    } else {
        // two levels per barrier: velocities of m and m+1, one barrier, the
        // stresses of m-2 and m-1 (reading levels m-4 .. m+1; the 8-level
        // ring keeps the next pair's writes, m+2 and m+3, off them)
        for (int m = m0; m <= ke + 2; m += 2) {
            const bool two = m + 1 <= ke + 2;   // uniform over the block
            Old old0, old1;
            vel_level(m, old0);
            if (two) {
                // the free-surface images at nz+1 read level nz of every lane
                if (m + 1 == g.nz + 1) __syncthreads();
                vel_level(m + 1, old1);
            }
            __syncthreads();
            StrIn sin0 = {}, sin1 = {};
            str_level(m, old0, sin0);
            if (two) str_level(m + 1, old1, sin1);
        }
    }
}

template <class P, int TY, int PF>
__global__ void __launch_bounds__(64 * (TY + 5))
k_step(const Geo g, const State a, const State b, const Coef c, const int kc, const FTiles tl)
{
    using Sh = FShape<TY, PF>;
    __shared__ float lds[Sh::NSLOT * Sh::SLOT];
    int bx, by, bz;
    if (!ftile_of(tl, bx, by, bz)) return;   // uniform over the block
    step_tile<P, TY, PF>(g, a, b, c, kc, bx, by, bz, lds);
}

// Runs in one launch (a single run is a one-entry table).  Runs' tiles are
// numbered consecutively (run r owns [tile0, tile0 + ntiles)); the grid is
// persistent, each block claiming tiles (claim_tile: a contiguous range of
// the list per XCD, spent ranges taken over from the back) until none is
// left.  Each tile's run is looked up in the table, so the kernel keeps no
// run's arguments live across tiles.  `par` selects each run's (read,
// write) buffer pair.
template <class P, int TY, int PF>
__global__ void __launch_bounds__(64 * (TY + 5))
k_step_batch(const FJob *__restrict__ jobs, const int njobs, const int total, const int per, const int par,
             unsigned long long *claim, const int rot)
{
    using Sh = FShape<TY, PF>;
    __shared__ float lds[Sh::NSLOT * Sh::SLOT];
    __shared__ int next;
    const bool lead = threadIdx.x == 0 && threadIdx.y == 0;
    for (;;) {
        if (lead) next = claim_tile(claim, per, total, rot);
        __syncthreads();
        const int tile = next;
        if (tile < 0) break;   // uniform over the block
        int jb = 0;
        while (jb + 1 < njobs && jobs[jb + 1].tile0 <= tile) ++jb;
        const FJob &J = jobs[jb];
        int bx, by, bz;
        patch_tile(tile - J.tile0, J.t.nbx, J.t.nby, J.t.pw, J.t.ph, bx, by, bz);
        step_tile<P, TY, PF>(J.g, J.s[par], J.s[par ^ 1], J.c, J.kc, bx, by, bz, lds);
        __syncthreads();   // the next tile reuses the LDS ring and `next`
    }
    if (lead) claim_done(claim, (int)gridDim.x);
}

FTiles fused_tiles(const Geo &g, const Launch &l)
{
    FTiles t;
    t.nbx = (g.nx + TX - 1) / TX;
    t.nby = (g.ny + l.tile_y - 1) / l.tile_y;
    t.nbz = l.nchunk;
    t.pw = l.pw > 0 ? l.pw : 1;
    t.ph = l.ph > 0 ? l.ph : 1;
    const long long n = (long long)t.nbx * t.nby * t.nbz;
    t.per = (int)((n + 7) / 8);
    return t;
}

template <class P, int TY, int PF>
static hipError_t step_t(const Geo &g, const State &a, const State &b, const Coef &c, const Launch &l,
                         hipStream_t st)
{
    const FTiles t = fused_tiles(g, l);
    hipLaunchKernelGGL((k_step<P, TY, PF>), dim3(8 * t.per), dim3(TX, TY + 5), 0, st, g, a, b, c, l.kc, t);
    return hipGetLastError();
}

template <class P, int TY, int PF>
static hipError_t batch_t(const FJob *d_jobs, int njobs, int total, int grid, int par, unsigned long long *claim,
                          int rot, hipStream_t st)
{
    const int per = (total + 7) / 8;
    grid = std::max(1, std::min(grid, total));
    hipLaunchKernelGGL((k_step_batch<P, TY, PF>), dim3(grid), dim3(TX, TY + 5), 0, st, d_jobs, njobs, total, per,
                       par, claim, rot & 7);
    return hipGetLastError();
}

template <class P, int TY, int PF>
static int grid_t(int device)
{
    int nb = 0, ncu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_step_batch<P, TY, PF>, TX * (TY + 5), 0) != hipSuccess)
        nb = 1;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) ncu = 256;
    return std::max(1, nb) * std::max(8, ncu);
}

// kernel variant of (storage, tile rows, prefetch depth): CALL is run with
// P, TY, PF of the matching template instance.  64 x 8 and 64 x 11 tiles
// come without prefetch only: their 13- and 16-wave blocks leave 128
// registers a lane, and the prefetching variants spill there.  64 x 11
// tiles fill a CU's 16 wave slots (4 per SIMD at up to 128 registers) with
// one block, where a 13-wave block leaves 3 of them empty.
#define AWP_FUSED_DISPATCH(mixed, l, CALL)                                           \
    do {                                                                             \
        const int pf_ = (l).unroll;                                                  \
        if (mixed) {                                                                 \
            using P = StoreMixed;                                                    \
            if ((l).tile_y == 7) {                                                   \
                if (pf_ == 0) { constexpr int TY = 7, PF = 0; CALL; }                \
                else if (pf_ == 2) { constexpr int TY = 7, PF = 2; CALL; }           \
                else { constexpr int TY = 7, PF = 1; CALL; }                         \
            } else if ((l).tile_y == 4) {                                            \
                if (pf_ == 0) { constexpr int TY = 4, PF = 0; CALL; }                \
                else if (pf_ == 2) { constexpr int TY = 4, PF = 2; CALL; }           \
                else { constexpr int TY = 4, PF = 1; CALL; }                         \
            } else if ((l).tile_y == 11) {                                           \
                if (pf_ == 3) { constexpr int TY = 11, PF = 3; CALL; }               \
                else { constexpr int TY = 11, PF = 0; CALL; }                        \
            } else {                                                                 \
                constexpr int TY = 8, PF = 0; CALL;                                  \
            }                                                                        \
        } else {                                                                     \
            using P = StoreF32;                                                      \
            if ((l).tile_y == 7) {                                                   \
                if (pf_ == 0) { constexpr int TY = 7, PF = 0; CALL; }                \
                else if (pf_ == 2) { constexpr int TY = 7, PF = 2; CALL; }           \
                else { constexpr int TY = 7, PF = 1; CALL; }                         \
            } else if ((l).tile_y == 4) {                                            \
                if (pf_ == 0) { constexpr int TY = 4, PF = 0; CALL; }                \
                else if (pf_ == 2) { constexpr int TY = 4, PF = 2; CALL; }           \
                else { constexpr int TY = 4, PF = 1; CALL; }                         \
            } else if ((l).tile_y == 11) {                                           \
                if (pf_ == 3) { constexpr int TY = 11, PF = 3; CALL; }               \
                else { constexpr int TY = 11, PF = 0; CALL; }                        \
            } else {                                                                 \
                constexpr int TY = 8, PF = 0; CALL;                                  \
            }                                                                        \
        }                                                                            \
    } while (0)

hipError_t launch_step_fused(bool mixed, const Geo &g, const State &a, const State &b, const Coef &c,
                             const Launch &l, hipStream_t st)
{
    AWP_FUSED_DISPATCH(mixed, l, return (step_t<P, TY, PF>(g, a, b, c, l, st)));
    return hipErrorInvalidValue;
}

hipError_t launch_step_fused_batch(bool mixed, const Launch &l, const FJob *d_jobs, int njobs, int total,
                                   int grid, int par, unsigned long long *claim, hipStream_t st)
{
    if (njobs <= 0 || total <= 0) return hipSuccess;
    if (!claim) return hipErrorInvalidValue;   // the kernel takes its tiles through the counters
    AWP_FUSED_DISPATCH(mixed, l, return (batch_t<P, TY, PF>(d_jobs, njobs, total, grid, par, claim, l.xrot, st)));
    return hipErrorInvalidValue;
}

int fused_batch_grid(bool mixed, const Launch &l, int device)
{
    AWP_FUSED_DISPATCH(mixed, l, return (grid_t<P, TY, PF>(device)));
    return 256;
}

#undef AWP_FUSED_DISPATCH

}  // namespace awp
