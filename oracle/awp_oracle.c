/*
 * awp_oracle.c -- CPU restatement of the AWP-ODC-SGT time step (the reference
 * CPU Fortran solver, AWP-ODC-SGT/src/) as a single-domain C program.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path: it is
 * loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
 * leg, and nothing in the product links, loads or calls it.
 *
 * Compiled twice (see oracle/build_oracle.sh): REAL=float gives the FP32
 * restatement (symbols *_f32) whose operation order follows the Fortran
 * statement by statement, so it agrees with the reference binary to rounding
 * level; REAL=double gives the FP64 companion (*_f64) used to state the error
 * of the FP32 and mixed-storage device paths.
 *
 * Domain semantics are those of the reference run on >= 2 ranks per
 * horizontal axis (each rank at least ND cells wide): the Cerjan sponge is the
 * separable product fx*fy*fz damping both x faces, both y faces and the bottom
 * (cerjan.f:4-247 evaluated on every rank), ghost values outside the global
 * domain are zero (memory.f:163-213), media ghosts are the padmedia replicas
 * (structure.f:563-676, including its (nx+1,0,0) corner), qp/qs ghosts are
 * zero (they are never padded; mediaswp only fills interior rank boundaries).
 *
 * Index convention: Fortran's (i,j,k) with i=1..nx, j=1..ny, k=1..nz, where
 * k = nz is the free surface (k = nz+1-z for file depth index z,
 * structure.f:488, source.f:188, station.f:50).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef REAL
#define REAL float
#endif

#if defined(REAL_IS_DOUBLE)
#define SFX(name) name##_f64
#define EXP exp
#define LOG log
#define SQRT sqrt
#define ATAN atan
#elif defined(MIXED)
#define SFX(name) name##_mix
#define EXP expf
#define LOG logf
#define SQRT sqrtf
#define ATAN atanf
#else
#define SFX(name) name##_f32
#define EXP expf
#define LOG logf
#define SQRT sqrtf
#define ATAN atanf
#endif

typedef REAL real;

/*
 * MIXED: the mixed-storage variant of the device path, restated for
 * checking it.  The reference has no such mode; this is the FP32
 * restatement with (a) the memory variables r1..r6 stored as FP16 (round to
 * nearest even) of r * 2^(15-e), e shared by each 64-point x row segment
 * (i = 64s+1 .. 64s+64) at one (j, k): the exponent of the segment's
 * largest |r| (f * 2^e, f in [0.5, 1)) clamped to [-80, 120]; the new values
 * enter the same step's stresses unrounded, as on the device, (b) the attenuation weights
 * qp, qs stored as FP16 of value * 2^e, e per array putting its largest
 * magnitude in [2^13, 2^14), and the kernels' exact multiplications by
 * 2^e / 2^-e where the stored values are used, (c) the density and the
 * compliances d1, mu, lam stored as FP16 deviations from an FP32 profile per
 * level (quantize_media).  Arithmetic stays FP32 in the same order.
 */
#if defined(MIXED)
static inline float as_float(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t as_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
/* value of the FP16 nearest (ties to even) to f; |f| < 65504 */
static inline float round_fp16(float f)
{
    uint32_t x = as_bits(f);
    const uint32_t sign = x & 0x80000000u, ax = x & 0x7FFFFFFFu;
    if (ax < 0x38800000u) {          /* below 2^-14: FP16 subnormal grid 2^-24 */
        float q = rintf(fabsf(f) * 16777216.0f) / 16777216.0f;
        return sign ? -q : q;
    }
    uint32_t r = ax + 0xFFFu + ((ax >> 13) & 1u);
    return as_float(sign | (r & 0xFFFFE000u));
}
#define MISCALE(v, a) ((v) = (v) * o->misc[a])
#define RSTORE(v) (v)   /* rounded per row segment after the row, quantize_row */
#else
#define MISCALE(v, a) ((void)0)
#define RSTORE(v) (v)
#endif

/* Parameters the time step needs, as parsed from IN3D (io.f:327-418).  Real
 * fields are carried as double and rounded to `real` on entry. */
typedef struct {
    int nx, ny, nz;          /* NX NY NZ */
    int npx, npy, npz;       /* NPX NPY NPZ: only the SGT ownership rule uses them */
    double dh, dt;           /* DH DT */
    int nt;                  /* nint(TMAX/DT)+1, pmcl3d.f:287 */
    int nd;                  /* ND */
    double arbc;             /* ARBC */
    int socalq;              /* SOCALQ */
    double fl, fh, fp;       /* FL FH FP */
    int igreen;              /* 1..6 */
    int nst;                 /* NST */
    int ntiskp_sgt;          /* NTISKP_SGT */
    int write_step;          /* WRITE_STEP */
    int nvar;                /* 3 or 5 (8: 3 skipped + 5) */
} awp_oracle_params;

typedef struct {
    awp_oracle_params p;
    int nx, ny, nz;
    long NXP, NYP, NZP;
    size_t nvol;
    real dh, dt;
    /* wavefield (halo 2) */
    real *u1, *v1, *w1, *xx, *yy, *zz, *xy, *xz, *yz;
    /* memory variables (interior only, but stored on the same padded grid) */
    real *r1, *r2, *r3, *r4, *r5, *r6;
    /* media: d1 (density), mu, lam (compliances 1/mu, 1/lambda), qp, qs */
    real *d1, *mu, *lam, *qp, *qs;
    /* separable sponge profiles, index 1..n */
    real *fx, *fy, *fz;
    real msc[5], misc[5];     /* media storage scales d1 mu lam qp qs (MIXED) */
    real tau[2][2][2];       /* tau(idx,idy,idz), relax.f:15-34 */
    /* sources (global indices) */
    int npsrc;
    int *psrc;               /* [npsrc][3] */
    real *stf;               /* [npsrc][nst][6]: axx ayy azz axz ayz axy */
    /* SGT points */
    int npts;                /* points in the cordfile */
    int nkept;
    int *kidx;               /* [nkept] global output index (cordfile line - 1) */
    int *kpos;               /* [nkept][3] global (i,j,k) */
    real *sg;                /* [nkept][3] sg1 sg2 sg3 */
    int nsamp_max;
    int nsamp;
    real *sgt;               /* [npts][6][nsamp_max], -1 where never sampled */
    int step;                /* last completed step (0 before the loop) */
} oracle_t;

#define IX(o, i, j, k) (((size_t)((k) + 2) * (o)->NYP + (size_t)((j) + 2)) * (o)->NXP + (size_t)((i) + 2))

static real *alloc_vol(oracle_t *o)
{
    real *a = (real *)calloc(o->nvol, sizeof(real));
    return a;
}

/* Fortran MOD for the non-negative/negative operands used here. */
static int fmod_i(int a, int b) { return a % b; }

/* --- media set-up: inimesh case 1 transform, structure.f:449-524 -------- */
static void setup_media(oracle_t *o, const float *media, int nvar)
{
    const int nx = o->nx, ny = o->ny, nz = o->nz;
    const int voff = (nvar == 8) ? 3 : 0;
    const real pi = (real)4. * ATAN((real)1.);
    const real fp = (real)o->p.fp, fl = (real)o->p.fl, fh = (real)o->p.fh;
    const real w0 = (real)2 * pi * fp;
    const real ww1 = (real)2 * pi * fl;
    const real w2 = (real)2 * pi * fh;
    const real taumax = (real)1. / ww1;
    const real taumin = (real)1. / w2;
    const real tmp1 = (real)2. / pi * (LOG(taumax / taumin));
    const real tmp2 = (real)2. / pi * LOG(w0 * taumin);
    const real lw = LOG(w2 / w0);

    for (int zf = 0; zf < nz; ++zf) {
        const int k = nz - zf;               /* structure.f:488 */
        for (int j = 1; j <= ny; ++j) {
            for (int i = 1; i <= nx; ++i) {
                const float *rec = media + (((size_t)zf * ny + (j - 1)) * nx + (i - 1)) * nvar;
                real vp = (real)rec[voff + 0];
                real vs = (real)rec[voff + 1];
                real dd = (real)rec[voff + 2];
                real pq, sq;
                if (nvar == 3) {               /* structure.f:452-455 */
                    sq = (real)0.05f * vs;
#if defined(REAL_IS_DOUBLE)
                    sq = 0.05 * vs;
#endif
                    pq = (real)2.0 * sq;
                } else {
                    pq = (real)rec[voff + 3];
                    sq = (real)rec[voff + 4];
                }
                if (o->p.socalq == 1) {        /* structure.f:482-487 */
                    vs = vs * ((real)1 + (lw) / (pi * sq));
                    vp = vp * ((real)1 + (lw) / (pi * pq));
                    real vpvs = vp / vs;
#if defined(REAL_IS_DOUBLE)
                    if (vpvs < 1.45) vs = vp / 1.45;
#else
                    if (vpvs < 1.45f) vs = vp / 1.45f;
#endif
                }
                const size_t id = IX(o, i, j, k);
                o->mu[id] = (real)1. / (dd * vs * vs);
                o->lam[id] = (real)1. / (dd * (vp * vp - (real)2. * vs * vs));
                o->d1[id] = dd;
                real qpinv, qsinv;             /* structure.f:506-520 */
                if (pq <= (real)0.0) {
                    qpinv = 0; qsinv = 0;
                } else {
                    qpinv = (real)1. / pq;
                    qsinv = (real)1. / sq;
                }
                o->qp[id] = tmp1 * qpinv / ((real)1.0 - tmp2 * qpinv);
                o->qs[id] = tmp1 * qsinv / ((real)1.0 - tmp2 * qsinv);
            }
        }
    }

    /* padmedia, structure.f:563-676, on the global domain */
    real *A[3] = {o->lam, o->mu, o->d1};
    for (int a = 0; a < 3; ++a) {
        real *m = A[a];
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j) {
                m[IX(o, 0, j, k)] = m[IX(o, 1, j, k)];
                m[IX(o, nx + 1, j, k)] = m[IX(o, nx, j, k)];
            }
        for (int k = 1; k <= nz; ++k)
            for (int i = 1; i <= nx; ++i) {
                m[IX(o, i, 0, k)] = m[IX(o, i, 1, k)];
                m[IX(o, i, ny + 1, k)] = m[IX(o, i, ny, k)];
            }
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                m[IX(o, i, j, 0)] = m[IX(o, i, j, 1)];
        for (int i = 1; i <= nx; ++i) {
            m[IX(o, i, 0, 0)] = m[IX(o, i, 1, 1)];
            m[IX(o, i, ny + 1, 0)] = m[IX(o, i, ny, 1)];
            m[IX(o, i, 0, nz + 1)] = m[IX(o, i, 1, nz)];
            m[IX(o, i, ny + 1, nz + 1)] = m[IX(o, i, ny, nz)];
        }
        for (int j = 1; j <= ny; ++j) {
            m[IX(o, 0, j, 0)] = m[IX(o, 1, j, 1)];
            m[IX(o, nx + 1, j, 0)] = m[IX(o, nx, j, 1)];
            m[IX(o, 0, j, nz + 1)] = m[IX(o, 1, j, nz)];
            m[IX(o, nx + 1, j, nz + 1)] = m[IX(o, nx, j, nz)];
        }
        for (int k = 1; k <= nz; ++k) {
            m[IX(o, 0, 0, k)] = m[IX(o, 1, 1, k)];
            m[IX(o, nx + 1, 0, k)] = m[IX(o, nx, 1, k)];
            m[IX(o, 0, ny + 1, k)] = m[IX(o, 1, ny, k)];
            m[IX(o, nx + 1, ny + 1, k)] = m[IX(o, nx, ny, k)];
        }
        m[IX(o, 0, 0, 0)] = m[IX(o, 1, 1, 1)];
        m[IX(o, 0, ny + 1, 0)] = m[IX(o, 1, ny, 1)];
        m[IX(o, 0, 0, nz + 1)] = m[IX(o, 1, 1, nz)];
        m[IX(o, nx + 1, 0, nz + 1)] = m[IX(o, nx, 1, nz)];
        m[IX(o, nx + 1, ny + 1, 0)] = m[IX(o, nx, ny, 1)];
        m[IX(o, 0, ny + 1, nz + 1)] = m[IX(o, 1, ny, nz)];
        m[IX(o, nx + 1, ny + 1, nz + 1)] = m[IX(o, nx, ny, nz)];
    }
    /* the (nxt+1,0,0) corner takes lam(nxt,1,1) for all three arrays
     * (structure.f:653-655) */
    o->lam[IX(o, nx + 1, 0, 0)] = o->lam[IX(o, nx, 1, 1)];
    o->mu[IX(o, nx + 1, 0, 0)] = o->lam[IX(o, nx, 1, 1)];
    o->d1[IX(o, nx + 1, 0, 0)] = o->lam[IX(o, nx, 1, 1)];

    /* tausub, relax.f:4-39 */
    static const int tautem[2][2][2] = {{{1, 8}, {7, 2}}, {{6, 3}, {4, 5}}}; /* [idx][idy][idz] */
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) {
                real t = (real)tautem[a][b][c];
                t = (t - (real)0.5) / 8;
                t = (real)2.0 * t - (real)1.0;
                o->tau[a][b][c] = EXP((real)0.5 * (LOG(taumax * taumin) + LOG(taumax / taumin) * t));
            }
}

#if defined(MIXED)
/* d1, mu, lam (structure.f:488-490 after padmedia) to their mixed storage:
 * per level k a profile value bar(k), the midrange (min + max) / 2 in double
 * rounded to float of the level's nonzero interior values (i = 1..nx,
 * j = 1..ny; levels below 1 and above nz take level 1's and level nz's, 0 for
 * a level without nonzero values), and per cell the FP16 (round to nearest
 * even) of h = (float)(x / bar - 1) in double, used as fmaf(bar, h, bar).
 * On a layered medium h = 0 and the values are exact; zeros stay zero. */
static void quantize_media(oracle_t *o)
{
    real *M[3] = {o->d1, o->mu, o->lam};
    const int nx = o->nx, ny = o->ny, nz = o->nz;
    float *bar = (float *)malloc(sizeof(float) * (size_t)nz);
    for (int a = 0; a < 3; ++a) {
        for (int k = 1; k <= nz; ++k) {
            float lo = 0.0f, hi = 0.0f;
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    const float v = M[a][IX(o, i, j, k)];
                    if (v == 0.0f) continue;
                    if (lo == 0.0f || v < lo) lo = v;
                    if (v > hi) hi = v;
                }
            bar[k - 1] = (float)(((double)lo + (double)hi) * 0.5);
        }
        for (int k = -2; k <= nz + 2; ++k) {
            const float b = bar[(k < 1 ? 1 : (k > nz ? nz : k)) - 1];
            for (int j = -2; j <= ny + 2; ++j)
                for (int i = -2; i <= nx + 2; ++i) {
                    real *v = &M[a][IX(o, i, j, k)];
                    const float h = b != 0.0f ? round_fp16((float)((double)*v / (double)b - 1.0)) : 0.0f;
                    *v = fmaf(b, h, b);
                }
        }
    }
    free(bar);
}
#endif

/* --- Cerjan profiles, cerjan.f:4-247 on >= 2 ranks per axis ------------- */
static void setup_cerjan(oracle_t *o)
{
    const int nd = o->p.nd, nx = o->nx, ny = o->ny, nz = o->nz;
    const real alpha = SQRT(-LOG((real)o->p.arbc)) / (real)nd;
    for (int i = 1; i <= nx; ++i) {
        real f = 1;
        if (i <= nd) { real a = alpha * (real)(nd - i + 1); f = EXP(-(a * a)); }
        else if (i >= nx - nd + 1) { real a = alpha * (real)(nd - (nx - i)); f = EXP(-(a * a)); }
        o->fx[i] = f;
    }
    for (int j = 1; j <= ny; ++j) {
        real f = 1;
        if (j <= nd) { real a = alpha * (real)(nd - j + 1); f = EXP(-(a * a)); }
        else if (j >= ny - nd + 1) { real a = alpha * (real)(nd - (ny - j)); f = EXP(-(a * a)); }
        o->fy[j] = f;
    }
    for (int k = 1; k <= nz; ++k) {
        real f = 1;
        if (k <= nd) { real a = alpha * (real)(nd - k + 1); f = EXP(-(a * a)); }
        o->fz[k] = f;
    }
}

static inline real damp(const oracle_t *o, int i, int j, int k)
{
    return (o->fx[i] * o->fy[j]) * o->fz[k];
}

/* --- velocity: uxx1/vyy1/wzz1 (operator.f:92-244) + addcrjvel ----------- */
static void dvelc(oracle_t *o)
{
    const int nx = o->nx, ny = o->ny, nz = o->nz;
    const real dth = o->dt / o->dh;
    const real c1 = (real)9. / (real)8.;
    const real c2 = (real)-1. / (real)24.;
    const long sx = 1, sy = o->NXP, sz = o->NXP * o->NYP;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const size_t p = IX(o, i, j, k);
                const real *d1 = o->d1, *xx = o->xx, *yy = o->yy, *zz = o->zz;
                const real *xy = o->xy, *xz = o->xz, *yz = o->yz;
                real d, q;
                d = (real)0.25 * ((d1[p] + d1[p - sy]) + (d1[p - sz] + d1[p - sy - sz]));
                q = dth / d;
                o->u1[p] = o->u1[p] + q * (
                    c1 * (xx[p] - xx[p - sx]) + c2 * (xx[p + sx] - xx[p - 2 * sx]) +
                    c1 * (xy[p] - xy[p - sy]) + c2 * (xy[p + sy] - xy[p - 2 * sy]) +
                    c1 * (xz[p] - xz[p - sz]) + c2 * (xz[p + sz] - xz[p - 2 * sz]));
                d = (real)0.25 * ((d1[p] + d1[p + sx]) + (d1[p - sz] + d1[p + sx - sz]));
                q = dth / d;
                o->v1[p] = o->v1[p] + q * (
                    c1 * (xy[p + sx] - xy[p]) + c2 * (xy[p + 2 * sx] - xy[p - sx]) +
                    c1 * (yy[p + sy] - yy[p]) + c2 * (yy[p + 2 * sy] - yy[p - sy]) +
                    c1 * (yz[p] - yz[p - sz]) + c2 * (yz[p + sz] - yz[p - 2 * sz]));
                d = (real)0.25 * ((d1[p] + d1[p + sx]) + (d1[p - sy] + d1[p + sx - sy]));
                q = dth / d;
                o->w1[p] = o->w1[p] + q * (
                    c1 * (xz[p + sx] - xz[p]) + c2 * (xz[p + 2 * sx] - xz[p - sx]) +
                    c1 * (yz[p] - yz[p - sy]) + c2 * (yz[p + sy] - yz[p - 2 * sy]) +
                    c1 * (zz[p + sz] - zz[p]) + c2 * (zz[p + 2 * sz] - zz[p - sz]));
                /* addcrjvel, cerjan.f:292-328 (D == 1 outside the sponge) */
                const real D = damp(o, i, j, k);
                o->u1[p] = o->u1[p] * D;
                o->v1[p] = o->v1[p] * D;
                o->w1[p] = o->w1[p] * D;
            }
}

/* --- free surface velocity images: fvelxy, fvelz (operator.f:246-306) --- */
static void fvel(oracle_t *o)
{
    const int nx = o->nx, ny = o->ny, nz = o->nz;
    const long sx = 1, sy = o->NXP, sz = o->NXP * o->NYP;
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            size_t p = IX(o, i, j, nz);
            o->u1[p + sz] = o->u1[p] - (o->w1[p] - o->w1[p - sx]);
        }
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            size_t p = IX(o, i, j, nz);
            o->v1[p + sz] = o->v1[p] - (o->w1[p + sy] - o->w1[p]);
        }
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            size_t p = IX(o, i, j, nz);
            size_t q = p + sz; /* nz+1 */
            real xl = (real)1. / o->lam[p];
            real xm = (real)2. / o->mu[p];
            real xl2m = xm + xl;
            o->w1[q] = o->w1[p - sz] - (xl / xl2m) * (
                (o->u1[q + sx] - o->u1[q]) +
                (o->u1[p + sx] - o->u1[p]) +
                (o->v1[q] - o->v1[q - sy]) +
                (o->v1[p] - o->v1[p - sy]));
        }
}

/* --- stress + memory variables: allq (viscoop.f:49-191) + addcrjstr ----- */
#if defined(MIXED)
/* Memory variables of row (j, k) to their stored values: per 64-point row
 * segment, FP16 (round to nearest even) of r * 2^(15-e) times 2^(e-15). */
static void quantize_row(oracle_t *o, int j, int k)
{
    real *rr[6] = {o->r1, o->r2, o->r3, o->r4, o->r5, o->r6};
    for (int a = 0; a < 6; ++a)
        for (int s0 = 1; s0 <= o->nx; s0 += 64) {
            const int s1 = s0 + 63 < o->nx ? s0 + 63 : o->nx;
            float mx = 0.0f;
            for (int i = s0; i <= s1; ++i) mx = fmaxf(mx, fabsf(rr[a][IX(o, i, j, k)]));
            int e = 0;
            if (mx > 0.0f) frexpf(mx, &e);
            e = e < -80 ? -80 : (e > 120 ? 120 : e);
            for (int i = s0; i <= s1; ++i) {
                const size_t p = IX(o, i, j, k);
                rr[a][p] = ldexpf(round_fp16(ldexpf(rr[a][p], 15 - e)), e - 15);
            }
        }
}
#endif

static void dstrqc(oracle_t *o)
{
    const int nx = o->nx, ny = o->ny, nz = o->nz;
    const real dh = o->dh, dt = o->dt;
    const real dh1 = (real)1. / dh;
    const real dth = dt / dh;
    const real c1 = (real)9. / (real)8.;
    const real c2 = (real)-1. / (real)24.;
    const real dt1 = (real)1. / dt;
    real tau1[2][2][2], tau2[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) {
                real tauu = o->tau[a][b][c];
                tau1[a][b][c] = (real)1 / ((tauu * dt1) + ((real)1. / (real)2.));
                tau2[a][b][c] = (tauu * dt1) - ((real)1. / (real)2.);
            }
    const long sx = 1, sy = o->NXP, sz = o->NXP * o->NYP;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k) {
        const int itz = fmod_i((nz + 1) - k, 2);
        for (int j = 1; j <= ny; ++j) {
            const int ity = fmod_i(j, 2);
            for (int i = 1; i <= nx; ++i) {
                const int itx = fmod_i(i, 2);
                const real vx1 = tau1[itx][ity][itz];
                const real vx2 = tau2[itx][ity][itz];
                const size_t p = IX(o, i, j, k);
                const real *lam = o->lam, *mu = o->mu, *qp = o->qp, *qs = o->qs;
                const real *u1 = o->u1, *v1 = o->v1, *w1 = o->w1;
                real vxl = (real)8. / (lam[p] + lam[p + sx] + lam[p - sy] + lam[p + sx - sy] +
                                       lam[p - sz] + lam[p + sx - sz] + lam[p - sy - sz] + lam[p + sx - sy - sz]);
                real vxm = (real)8. / (mu[p] + mu[p + sx] + mu[p - sy] + mu[p + sx - sy] +
                                       mu[p - sz] + mu[p + sx - sz] + mu[p - sy - sz] + mu[p + sx - sy - sz]);
                vxl = vxl + (real)2. * vxm;
                real vqpa = (real)0.125 * (qp[p] + qp[p + sx] + qp[p - sy] + qp[p + sx - sy] +
                                           qp[p - sz] + qp[p + sx - sz] + qp[p - sy - sz] + qp[p + sx - sy - sz]);
                real vqsa = (real)0.125 * (qs[p] + qs[p + sx] + qs[p - sy] + qs[p + sx - sy] +
                                           qs[p - sz] + qs[p + sx - sz] + qs[p - sy - sz] + qs[p + sx - sy - sz]);
                MISCALE(vqpa, 3);
                MISCALE(vqsa, 4);
                real vt1 = c1 * (u1[p + sx] - u1[p]) + c2 * (u1[p + 2 * sx] - u1[p - sx]);
                real vt2 = c1 * (v1[p] - v1[p - sy]) + c2 * (v1[p + sy] - v1[p - 2 * sy]);
                real vt3 = c1 * (w1[p] - w1[p - sz]) + c2 * (w1[p + sz] - w1[p - 2 * sz]);
                real vt4 = c1 * (u1[p + sy] - u1[p]) + c2 * (u1[p + 2 * sy] - u1[p - sy]) +
                           c1 * (v1[p] - v1[p - sx]) + c2 * (v1[p + sx] - v1[p - 2 * sx]);
                real vt5 = c1 * (u1[p + sz] - u1[p]) + c2 * (u1[p + 2 * sz] - u1[p - sz]) +
                           c1 * (w1[p] - w1[p - sx]) + c2 * (w1[p + sx] - w1[p - 2 * sx]);
                real vt6 = c1 * (v1[p + sz] - v1[p]) + c2 * (v1[p + 2 * sz] - v1[p - sz]);
                real vt7 = c1 * (w1[p + sy] - w1[p]) + c2 * (w1[p + 2 * sy] - w1[p - sy]);

                real vh1 = -vxm * vqsa;
                real vtmp = vxl * (vt1 + vt2 + vt3);
                real va1 = -vqpa * vxl * dh1 * (vt1 + vt2 + vt3) / (real)2.;

                real *xx = o->xx, *yy = o->yy, *zz = o->zz, *xy = o->xy, *xz = o->xz, *yz = o->yz;
                real *r1 = o->r1, *r2 = o->r2, *r3 = o->r3, *r4 = o->r4, *r5 = o->r5, *r6 = o->r6;
                xx[p] = xx[p] + dth * (vtmp - (real)2. * vxm * (vt2 + vt3)) + dt * r1[p];
                yy[p] = yy[p] + dth * (vtmp - (real)2. * vxm * (vt1 + vt3)) + dt * r2[p];
                zz[p] = zz[p] + dth * (vtmp - (real)2. * vxm * (vt1 + vt2)) + dt * r3[p];

                real vxmu = (real)2. / (mu[p] + mu[p - sz]);
                real vaas = (real)0.5 * (qs[p] + qs[p - sz]);
                MISCALE(vaas, 4);
                real vh2_r4 = -vaas * vxmu / (real)2.;
                xy[p] = xy[p] + vxmu * dth * vt4 + dt * r4[p];
                vxmu = (real)2. / (mu[p] + mu[p - sy]);
                vaas = (real)0.5 * (qs[p] + qs[p - sy]);
                MISCALE(vaas, 4);
                real vh2_r5 = -vaas * vxmu / (real)2.;
                xz[p] = xz[p] + vxmu * dth * vt5 + dt * r5[p];
                vxmu = (real)2. / (mu[p] + mu[p + sx]);
                vaas = (real)0.5 * (qs[p] + qs[p + sx]);
                MISCALE(vaas, 4);
                real vh2_r6 = -vaas * vxmu / (real)2.;
                yz[p] = yz[p] + vxmu * (dth * vt6 + dth * vt7) + dt * r6[p];

                const real n1 = (vx2 * r1[p] - vh1 * (vt2 + vt3) * dh1 + va1) * vx1;
                const real n2 = (vx2 * r2[p] - vh1 * (vt1 + vt3) * dh1 + va1) * vx1;
                const real n3 = (vx2 * r3[p] - vh1 * (vt1 + vt2) * dh1 + va1) * vx1;
                const real n4 = (vx2 * r4[p] + vh2_r4 * dh1 * vt4) * vx1;
                const real n5 = (vx2 * r5[p] + vh2_r5 * dh1 * vt5) * vx1;
                const real n6 = (vx2 * r6[p] + vh2_r6 * dh1 * (vt6 + vt7)) * vx1;
                r1[p] = RSTORE(n1); r2[p] = RSTORE(n2); r3[p] = RSTORE(n3);
                r4[p] = RSTORE(n4); r5[p] = RSTORE(n5); r6[p] = RSTORE(n6);

                xx[p] = xx[p] + dt * n1;
                yy[p] = yy[p] + dt * n2;
                zz[p] = zz[p] + dt * n3;
                xy[p] = xy[p] + dt * n4;
                xz[p] = xz[p] + dt * n5;
                yz[p] = yz[p] + dt * n6;

                /* addcrjstr, cerjan.f:249-290 (r1..r6 are not damped) */
                const real D = damp(o, i, j, k);
                xx[p] = xx[p] * D;
                yy[p] = yy[p] * D;
                zz[p] = zz[p] * D;
                xz[p] = xz[p] * D;
                yz[p] = yz[p] * D;
                xy[p] = xy[p] * D;
            }
#if defined(MIXED)
            quantize_row(o, j, k);
#endif
        }
    }
}

/* --- fstr, operator.f:575-612 ------------------------------------------ */
static void fstr(oracle_t *o)
{
    const int nx = o->nx, ny = o->ny, nz = o->nz;
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
            o->zz[IX(o, i, j, nz + 1)] = -o->zz[IX(o, i, j, nz)];
            o->zz[IX(o, i, j, nz + 2)] = -o->zz[IX(o, i, j, nz - 1)];
            o->xz[IX(o, i, j, nz + 1)] = -o->xz[IX(o, i, j, nz - 1)];
            o->xz[IX(o, i, j, nz + 2)] = -o->xz[IX(o, i, j, nz - 2)];
            o->yz[IX(o, i, j, nz + 1)] = -o->yz[IX(o, i, j, nz - 1)];
            o->yz[IX(o, i, j, nz + 2)] = -o->yz[IX(o, i, j, nz - 2)];
            o->xz[IX(o, i, j, nz)] = 0;
            o->yz[IX(o, i, j, nz)] = 0;
        }
}

/* --- addsrc, source.f:812-913, igreen 1..6 ------------------------------- */
static void addsrc(oracle_t *o, int istep)
{
    if (istep > o->p.nst) return;
    const int nz = o->nz;
    const real dx = o->dh, dt = o->dt;
    const real vtst = dt / (dx * dx * dx);
    const real vtst1 = (real)1. / (dx * dx);
    for (int s = 0; s < o->npsrc; ++s) {
        const int *ps = o->psrc + 3 * s;
        const real *a = o->stf + ((size_t)s * o->p.nst + (istep - 1)) * 6;
        const real axx = a[0], ayy = a[1], azz = a[2];
        const size_t p = IX(o, ps[0], ps[1], ps[2]);
        switch (o->p.igreen) {
        case 1: { real a1 = axx * vtst / o->d1[p]; o->u1[p] = o->u1[p] + a1; } break;
        case 2: { real a1 = ayy * vtst / o->d1[p]; o->v1[p] = o->v1[p] + a1; } break;
        case 3: { real a1 = azz * vtst / o->d1[p]; o->w1[p] = o->w1[p] + a1; } break;
        case 4:
            o->xz[IX(o, ps[0], ps[1], nz)] = axx * vtst1;
            o->xz[IX(o, ps[0], ps[1], nz + 1)] = (real)2. * axx * vtst1 - o->xz[IX(o, ps[0], ps[1], nz - 1)];
            o->xz[IX(o, ps[0], ps[1], nz + 2)] = (real)2. * axx * vtst1 - o->xz[IX(o, ps[0], ps[1], nz - 2)];
            break;
        case 5:
            o->yz[IX(o, ps[0], ps[1], nz)] = ayy * vtst1;
            o->yz[IX(o, ps[0], ps[1], nz + 1)] = (real)2. * ayy * vtst1 - o->yz[IX(o, ps[0], ps[1], nz - 1)];
            o->yz[IX(o, ps[0], ps[1], nz + 2)] = (real)2. * ayy * vtst1 - o->yz[IX(o, ps[0], ps[1], nz - 2)];
            break;
        case 6:
            o->zz[IX(o, ps[0], ps[1], nz + 1)] = (real)2. * azz * vtst1 - o->zz[IX(o, ps[0], ps[1], nz)];
            o->zz[IX(o, ps[0], ps[1], nz + 2)] = (real)2. * azz * vtst1 - o->zz[IX(o, ps[0], ps[1], nz - 1)];
            break;
        default: break;
        }
    }
}

/* --- SGT gather, pmcl3d.f:1536-1557 ------------------------------------- */
static void sgt_gather(oracle_t *o)
{
    const int s = o->nsamp;
    if (s >= o->nsamp_max) return;
    for (int n = 0; n < o->nkept; ++n) {
        const int *q = o->kpos + 3 * n;
        const size_t p = IX(o, q[0], q[1], q[2]);
        const real sg1 = o->sg[3 * n], sg2 = o->sg[3 * n + 1], sg3 = o->sg[3 * n + 2];
        const real xx = o->xx[p], yy = o->yy[p], zz = o->zz[p];
        real *dst = o->sgt + (size_t)o->kidx[n] * 6 * o->nsamp_max;
        dst[0 * o->nsamp_max + s] = sg1 * xx + sg2 * yy + sg2 * zz;
        dst[1 * o->nsamp_max + s] = sg2 * xx + sg1 * yy + sg2 * zz;
        dst[2 * o->nsamp_max + s] = sg2 * xx + sg2 * yy + sg1 * zz;
        dst[3 * o->nsamp_max + s] = sg3 * o->xy[p];
        dst[4 * o->nsamp_max + s] = sg3 * o->xz[p];
        dst[5 * o->nsamp_max + s] = sg3 * o->yz[p];
    }
    o->nsamp++;
}

/* ======================================================================= */

void SFX(awp_oracle_destroy)(void *h)
{
    oracle_t *o = (oracle_t *)h;
    if (!o) return;
    real *arrs[] = {o->u1, o->v1, o->w1, o->xx, o->yy, o->zz, o->xy, o->xz, o->yz,
                    o->r1, o->r2, o->r3, o->r4, o->r5, o->r6,
                    o->d1, o->mu, o->lam, o->qp, o->qs, o->fx, o->fy, o->fz,
                    o->stf, o->sg, o->sgt};
    for (size_t a = 0; a < sizeof(arrs) / sizeof(arrs[0]); ++a) free(arrs[a]);
    free(o->psrc); free(o->kidx); free(o->kpos);
    free(o);
}

/*
 * Create a run.  media: float32 [nz][ny][nx][nvar] in file order.
 * src: [npsrc][3] file coordinates (x, y, z from the surface, 1-based);
 * stf: [npsrc][nst][6].  pts: [npts][3] cordfile coordinates.
 * Returns NULL on invalid input.
 */
void *SFX(awp_oracle_create)(const awp_oracle_params *prm, const float *media,
                             int npsrc, const int *src, const float *stf,
                             int npts, const int *pts)
{
    if (!prm || prm->nx < 4 || prm->ny < 4 || prm->nz < 4) return NULL;
    if (prm->igreen < 1 || prm->igreen > 6) return NULL;
    if (prm->nvar != 3 && prm->nvar != 5 && prm->nvar != 8) return NULL;
    if (prm->npx < 1 || prm->npy < 1 || prm->npz < 1) return NULL;
    if (prm->nx % prm->npx || prm->ny % prm->npy || prm->nz % prm->npz) return NULL;
    oracle_t *o = (oracle_t *)calloc(1, sizeof(oracle_t));
    o->p = *prm;
    o->nx = prm->nx; o->ny = prm->ny; o->nz = prm->nz;
    o->NXP = o->nx + 5; o->NYP = o->ny + 5; o->NZP = o->nz + 5;
    o->nvol = (size_t)o->NXP * o->NYP * o->NZP;
    o->dh = (real)prm->dh; o->dt = (real)prm->dt;
    real **vols[] = {&o->u1, &o->v1, &o->w1, &o->xx, &o->yy, &o->zz, &o->xy, &o->xz, &o->yz,
                     &o->r1, &o->r2, &o->r3, &o->r4, &o->r5, &o->r6,
                     &o->d1, &o->mu, &o->lam, &o->qp, &o->qs};
    for (size_t a = 0; a < sizeof(vols) / sizeof(vols[0]); ++a) {
        *vols[a] = alloc_vol(o);
        if (!*vols[a]) { SFX(awp_oracle_destroy)(o); return NULL; }
    }
    o->fx = (real *)calloc(o->nx + 2, sizeof(real));
    o->fy = (real *)calloc(o->ny + 2, sizeof(real));
    o->fz = (real *)calloc(o->nz + 2, sizeof(real));

    setup_media(o, media, prm->nvar);
#if defined(MIXED)
    quantize_media(o);
#endif
    setup_cerjan(o);

    /* sources: rdsrcpos (source.f:170-198) + getsrc (source.f:378-402) */
    o->npsrc = npsrc;
    o->psrc = (int *)calloc(3 * (size_t)(npsrc > 0 ? npsrc : 1), sizeof(int));
    o->stf = (real *)calloc((size_t)(npsrc > 0 ? npsrc : 1) * prm->nst * 6, sizeof(real));
    for (int s = 0; s < npsrc; ++s) {
        const int x = src[3 * s], y = src[3 * s + 1], z = src[3 * s + 2];
        if (x < 1 || x > o->nx || y < 1 || y > o->ny || z < 1 || z > o->nz) {
            SFX(awp_oracle_destroy)(o); return NULL;
        }
        o->psrc[3 * s] = x; o->psrc[3 * s + 1] = y; o->psrc[3 * s + 2] = o->nz + 1 - z;
        for (size_t q = 0; q < (size_t)prm->nst * 6; ++q)
            o->stf[(size_t)s * prm->nst * 6 + q] = (real)stf[(size_t)s * prm->nst * 6 + q];
    }

    /* SGT ownership, station.f:31-80, evaluated for every rank of the
     * NPX x NPY x NPZ decomposition; kept points are re-expressed globally. */
    const int nxt = o->nx / prm->npx, nyt = o->ny / prm->npy, nzt = o->nz / prm->npz;
    o->npts = npts;
    o->kidx = (int *)calloc((size_t)(npts > 0 ? npts : 1), sizeof(int));
    o->kpos = (int *)calloc(3 * (size_t)(npts > 0 ? npts : 1), sizeof(int));
    o->sg = (real *)calloc(3 * (size_t)(npts > 0 ? npts : 1), sizeof(real));
    o->nkept = 0;
    for (int n = 0; n < npts; ++n) {
        const int x = pts[3 * n], y = pts[3 * n + 1], z = pts[3 * n + 2];
        const int nzp = (o->nz - z) / nzt;     /* Fortran int(): truncation */
        const int nyp = (y - 1) / nyt;
        const int nxp = (x - 1) / nxt;
        if (nzp < 0 || nzp >= prm->npz || nyp < 0 || nyp >= prm->npy || nxp < 0 || nxp >= prm->npx)
            continue;
        int a;
        a = x % nxt; const int lx = a ? a : nxt;
        a = y % nyt; const int ly = a ? a : nyt;
        a = (o->nz - z + 1) % nzt; const int lz = a ? a : nzt;
        const int gi = nxp * nxt + lx, gj = nyp * nyt + ly, gk = nzp * nzt + lz;
        o->kidx[o->nkept] = n;
        o->kpos[3 * o->nkept] = gi; o->kpos[3 * o->nkept + 1] = gj; o->kpos[3 * o->nkept + 2] = gk;
        /* sg1/sg2/sg3 at the node, pmcl3d.f:1158-1170 */
        const size_t p = IX(o, gi, gj, gk);
        const real lam = o->lam[p], mu = o->mu[p];
        o->sg[3 * o->nkept + 0] = ((real)1. / lam + (real)1. / mu) /
                                  ((real)1. / mu * ((real)3.0 / lam + (real)2.0 / mu));
        o->sg[3 * o->nkept + 1] = -(((real)1. / lam) /
                                    ((real)2.0 / mu * ((real)3.0 / lam + (real)2.0 / mu)));
        o->sg[3 * o->nkept + 2] = mu;
        o->nkept++;
    }
    for (int a = 0; a < 5; ++a) { o->msc[a] = 1; o->misc[a] = 1; }
#if defined(MIXED)
    /* qp, qs to FP16 storage of value * 2^e (d1, mu, lam were stored as
     * deviations from their level profiles by quantize_media, before the
     * SGT coefficients above were taken from them, as the device does) */
    {
        real *M[5] = {o->d1, o->mu, o->lam, o->qp, o->qs};
        for (int a = 3; a < 5; ++a) {
            float mx = 0.0f;
            for (size_t q = 0; q < o->nvol; ++q) mx = fmaxf(mx, fabsf(M[a][q]));
            int ex = 0;
            if (mx > 0.0f) frexpf(mx, &ex);
            o->msc[a] = ldexpf(1.0f, 14 - ex);
            o->misc[a] = ldexpf(1.0f, ex - 14);
            for (size_t q = 0; q < o->nvol; ++q) M[a][q] = round_fp16(M[a][q] * o->msc[a]);
        }
    }
#endif
    o->nsamp_max = prm->nt / prm->ntiskp_sgt;
    o->sgt = (real *)malloc((size_t)(npts > 0 ? npts : 1) * 6 * (o->nsamp_max > 0 ? o->nsamp_max : 1) * sizeof(real));
    for (size_t q = 0; q < (size_t)(npts > 0 ? npts : 1) * 6 * (o->nsamp_max > 0 ? o->nsamp_max : 1); ++q)
        o->sgt[q] = -1;
    o->nsamp = 0;
    o->step = 0;
    /* first source sample before the loop, pmcl3d.f:988 */
    if (prm->igreen >= 1 && prm->igreen <= 3) addsrc(o, 1);
    else if (prm->igreen >= 4) addsrc(o, 1);
    return o;
}

/* One reference time step i = ++step: pmcl3d.f:1183-1557 in order. */
static void one_step(oracle_t *o)
{
    const int i = ++o->step;
    dvelc(o);          /* dvelc + addcrjvel */
    fvel(o);           /* fvelxy, fvelz */
    dstrqc(o);         /* dstrqc + addcrjstr */
    if (o->p.igreen >= 4) {                     /* pmcl3d.f:1380-1385 */
        fstr(o);
        addsrc(o, i + 1);
    } else {
        addsrc(o, i + 1);
        fstr(o);
    }
    if (i % o->p.ntiskp_sgt == 0) sgt_gather(o);
}

int SFX(awp_oracle_run)(void *h, int nsteps, int nthreads)
{
    oracle_t *o = (oracle_t *)h;
    if (!o) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    for (int s = 0; s < nsteps && o->step < o->p.nt; ++s) one_step(o);
    return o->step;
}

/* Copy an interior field out as [k][j][i] (k = 1 bottom ... nz surface),
 * x fastest.  which: 0 u1 1 v1 2 w1 3 xx 4 yy 5 zz 6 xy 7 xz 8 yz 9..14 r1..r6
 * 15 d1 16 mu 17 lam 18 qp 19 qs.  kext = 0 copies k = 1..nz; kext = 2 copies
 * k = 1..nz+2 (the free-surface image planes too). */
int SFX(awp_oracle_field)(void *h, int which, int kext, real *out)
{
    oracle_t *o = (oracle_t *)h;
    if (!o) return -1;
    real *src[] = {o->u1, o->v1, o->w1, o->xx, o->yy, o->zz, o->xy, o->xz, o->yz,
                   o->r1, o->r2, o->r3, o->r4, o->r5, o->r6,
                   o->d1, o->mu, o->lam, o->qp, o->qs};
    if (which < 0 || which >= (int)(sizeof(src) / sizeof(src[0]))) return -2;
    const real *a = src[which];
    const real f = which >= 15 ? o->misc[which - 15] : (real)1;
    size_t q = 0;
    for (int k = 1; k <= o->nz + kext; ++k)
        for (int j = 1; j <= o->ny; ++j)
            for (int i = 1; i <= o->nx; ++i) out[q++] = a[IX(o, i, j, k)] * f;
    return 0;
}

/* SGT samples so far: [npts][6][nsamp_max] (unsampled slots stay -1).
 * Returns the number of samples taken. */
int SFX(awp_oracle_sgt)(void *h, real *out)
{
    oracle_t *o = (oracle_t *)h;
    if (!o) return -1;
    memcpy(out, o->sgt, (size_t)o->npts * 6 * o->nsamp_max * sizeof(real));
    return o->nsamp;
}

int SFX(awp_oracle_info)(void *h, int *out4)
{
    oracle_t *o = (oracle_t *)h;
    if (!o) return -1;
    out4[0] = o->step; out4[1] = o->nsamp; out4[2] = o->nsamp_max; out4[3] = o->nkept;
    return 0;
}

/* Global (i,j,k) of each kept SGT point and its cordfile index. */
int SFX(awp_oracle_points)(void *h, int *kidx, int *kpos)
{
    oracle_t *o = (oracle_t *)h;
    if (!o) return -1;
    memcpy(kidx, o->kidx, sizeof(int) * o->nkept);
    memcpy(kpos, o->kpos, sizeof(int) * 3 * o->nkept);
    return o->nkept;
}
