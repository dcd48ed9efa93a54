// awp_ctx.hip -- the C ABI of include/awpsgt.h: one context per SGT run,
// owning its device buffers and HIP stream, driving the kernels of
// awp_kernels.hip in the reference's step order (pmcl3d.f:1183-1557).

#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "awp_ctx.hpp"
#include "awp_kernels.hpp"
#include "awpsgt.h"

using namespace awp;

// ---------------------------------------------------------------------------

extern "C" const char *awp_last_error(const awp_ctx *ctx) { return ctx ? ctx->err : "null context"; }

namespace {
// the device address of the readback scratch, set up once per device
// (defined with awp_get_field_noalloc below)
int scratch_setup(int device, char *err, size_t errlen);
}  // namespace

// 32-bit byte offsets where every array is below 4 GiB
static bool use_o32(const Geo &g) { return offsets32(g); }

// Placement of a run's volume arrays.  One allocation holds, in order: the
// media d1 mu lam qp qs (4 bytes FP32, 2 bytes mixed; a 2-byte array starts
// 64 bytes into its slot so that its interior rows start on 128-byte lines),
// the mixed mode's row-segment exponents, the 9 FP32 arrays (velocities,
// stresses) of wavefield buffer 0, r1..r6 (4 or 2 bytes) and, for a run that
// holds the fused step's second wavefield buffer, its 9 arrays.  awp_init
// zeroes everything after the media.
struct VolLayout {
    size_t m[3], q[2], rexp;   // media, exponent slots
    size_t rexp_n;
    int nseg;
    size_t media_begin, media_end;   // [media_begin, media_end): the media arrays
    size_t zero_end;                 // [media_end, zero_end): exponents and state (zeroed by awp_init)
    size_t wf[2][9], r[6];           // the state arrays
    size_t total;                    // bytes of the allocation
};

static VolLayout vol_layout(const Geo &g, int storage, bool two_bufs)
{
    VolLayout L;
    const size_t nv = (size_t)g.nvol;
    const bool mixed = storage == AWP_STORAGE_MIXED;
    const size_t a4 = nv * 4;
    const size_t eslot = mixed ? nv * 2 + 128 : nv * 4, eoff = mixed ? 64 : 0;
    size_t o = L.media_begin = 0;
    // d1, mu, lam; mixed storage keeps mu and lam as one array of (mu, lam)
    // FP16 pairs, 4 bytes a point, in their two slots (both pointers at its
    // start, a multiple of 128)
    for (int a = 0; a < 3; ++a, o += eslot) L.m[a] = (mixed && a == 2) ? L.m[1] : (mixed && a == 1) ? o : o + eoff;
    // qp, qs: mixed storage keeps them as one array of (qp, qs) FP16 pairs,
    // 4 bytes a point, in the two slots (both pointers at its start)
    const size_t qbase = o;   // a multiple of 128: rows of the pairs start on a line
    for (int a = 0; a < 2; ++a, o += eslot) L.q[a] = mixed ? qbase : o + eoff;
    L.media_end = o;
    L.nseg = (g.nx + TX - 1) / TX;
    L.rexp_n = mixed ? (size_t)g.nz * g.ny * L.nseg : 0;
    L.rexp = o;
    o += (8 * L.rexp_n + 255) / 256 * 256;
    for (int a = 0; a < 9; ++a, o += a4) L.wf[0][a] = o;
    for (int a = 0; a < 6; ++a, o += eslot) L.r[a] = o + eoff;
    for (int a = 0; a < 9; ++a) {
        L.wf[1][a] = two_bufs ? o : L.wf[0][a];
        if (two_bufs) o += a4;
    }
    L.zero_end = L.total = o;
    return L;
}

// Points every volume array of the run at its place in the allocation.
static void place_arrays(awp_ctx *ctx)
{
    const VolLayout L = vol_layout(ctx->g, ctx->storage, ctx->two_bufs);
    char *const base = (char *)ctx->vol;
    for (int b = 0; b < 2; ++b)
        for (int a = 0; a < 9; ++a) ctx->wf[b][a] = (float *)(base + L.wf[b][a]);
    float **f32[9] = {&ctx->s.u1, &ctx->s.v1, &ctx->s.w1, &ctx->s.xx, &ctx->s.yy, &ctx->s.zz,
                      &ctx->s.xy, &ctx->s.xz, &ctx->s.yz};
    for (int a = 0; a < 9; ++a) *f32[a] = ctx->wf[0][a];
    void **rr[6] = {&ctx->s.r1, &ctx->s.r2, &ctx->s.r3, &ctx->s.r4, &ctx->s.r5, &ctx->s.r6};
    for (int a = 0; a < 6; ++a) *rr[a] = base + L.r[a];
    void **mm[3] = {&ctx->s.d1, &ctx->s.mu, &ctx->s.lam};
    for (int a = 0; a < 3; ++a) *mm[a] = base + L.m[a];
    ctx->s.qp = base + L.q[0];
    ctx->s.qs = base + L.q[1];
    ctx->s.rexp = L.rexp_n ? (signed char *)(base + L.rexp) : nullptr;
    ctx->s.rexp_n = (long long)L.rexp_n;
    ctx->s.nseg = L.nseg;
    ctx->zero_begin = L.media_end;
    ctx->zero_end = L.zero_end;
    ctx->cur = 0;
}

// Time-step scheme of a new run.  AWP_SCHEME=split: the two-stage step
// (velocity kernel, then stress kernel, in place); AWP_SCHEME=fused: the
// one-sweep fused step (a second wavefield buffer; every array addressed
// with 32-bit byte offsets from a base rebased per z chunk, so chunks of at
// most fused_kc_max levels).  Unset: fused for FP32 storage, two-stage for
// mixed storage (1500x1000x300, one run's buffers: fused FP32 16.8-16.9 ms
// per step on 64 x 11 tiles against 17.2-17.3 two-stage; fused mixed 23.7
// against 14.7).  awp_tune
// and awp_set_scheme switch a run that holds both buffers between the two.
static bool want_fused(const Geo &g, int storage)
{
    if (fused_kc_max(g) < 16) return false;   // planes of more than ~40M points
    const char *scheme = getenv("AWP_SCHEME");
    if (scheme && !strcmp(scheme, "fused")) return true;
    if (scheme && !strcmp(scheme, "split")) return false;
    return storage == AWP_STORAGE_FP32;
}

// z chunks of `nchunk` requested: the top chunk keeps >= 3 levels (the
// block reaching k = nz writes the free-surface stress images from the two
// levels below it)
static void set_chunks(Launch &l, const Geo &g, int nchunk, bool fused)
{
    nchunk = std::max(1, nchunk);
    int lo = 1;
    if (fused) {
        // the fused step's 32-bit offsets span one chunk (fused_kc_max)
        const int kmax = std::max(1, fused_kc_max(g));
        lo = (g.nz + kmax - 1) / kmax;
        nchunk = std::max(nchunk, lo);
    }
    auto ok = [&g](int n) {
        const int kc = (g.nz + n - 1) / n;
        return n == 1 || (g.nz - (n - 1) * kc >= 3 && (n - 1) * kc < g.nz);
    };
    int n = nchunk;
    while (n > lo && !ok(n)) --n;
    if (!ok(n)) {
        n = nchunk;
        while (n < g.nz && !ok(n)) ++n;
    }
    l.kc = (g.nz + n - 1) / n;
    l.nchunk = (g.nz + l.kc - 1) / l.kc;
}

static Launch choose_launch(const Geo &g, bool fused, bool mixed, int device)
{
    Launch l;
    l.tile_y = 8;
    l.tile_xw = 1;
    l.unroll = 1;
    const char *un = getenv("AWP_UNROLL");
    if (un && atoi(un) == 2) l.unroll = 2;
    // tuning overrides: AWP_TILE_Y (rows per block: 1, 2, 4, 8 or 16 for
    // the two-stage kernels, 4, 7, 8 or 11 for the fused one) and AWP_KC (z
    // levels per block)
    const char *ty = getenv("AWP_TILE_Y");
    // fused defaults measured at 1500x1000x300: FP32 64x11 tiles (16-wave
    // blocks fill a CU's wave slots: 16.2-16.3 ms per step against 17.7-18.1
    // for 64x8), mixed 64x7 tiles, both without load prefetch (awp_tune and
    // awp_batch_tune search the rest)
    if (fused) l.tile_y = mixed ? 7 : 11;
    if (ty && (atoi(ty) == 1 || atoi(ty) == 2 || atoi(ty) == 4 || atoi(ty) == 7 || atoi(ty) == 8 || atoi(ty) == 11 ||
               atoi(ty) == 16))
        l.tile_y = atoi(ty);
    if (fused && l.tile_y != 4 && l.tile_y != 7 && l.tile_y != 11) l.tile_y = 8;
    if (!fused && l.tile_y == 11) l.tile_y = 8;
    if (!fused && l.tile_y == 7) l.tile_y = 8;
    // tile patches of the fused step (AWP_PATCH=<w>x<h> overrides): 4 x 4
    // tiles of 64 x 8 points (1500x1000x300 FP32, 4 z chunks: 17.21 ms per
    // step against 17.62 for 2 x 16 and 17.75 for 3 x 8)
    l.pw = 4;
    l.ph = 4;
    // tile claim (AWP_CLAIM=0 never, 1 when the launch has at least four
    // tiles per persistent block -- the default --, 2 always; AWP_XROT=0..7
    // rotates the XCD ranges): persistent blocks that claim tiles per XCD.
    // 1500x1000x300: fused 16.81-16.92 ms per step against 17.11-17.20 with
    // one block per tile (same run); 200x200x100 (400 two-stage tiles) one
    // block per tile is faster
    l.claim = 1;
    l.xrot = 0;
    const char *ce = getenv("AWP_CLAIM");
    if (ce && (!strcmp(ce, "0") || !strcmp(ce, "1") || !strcmp(ce, "2"))) l.claim = ce[0] - '0';
    const char *xr = getenv("AWP_XROT");
    if (xr) l.xrot = atoi(xr) & 7;
    const char *pe = getenv("AWP_PATCH");
    if (pe) {
        int w = 0, h = 0;
        if (sscanf(pe, "%dx%d", &w, &h) == 2 && w >= 1 && h >= 1 && w * h <= 4096) {
            l.pw = w;
            l.ph = h;
        }
    }
    l.o32 = use_o32(g);
    // Form and blocks per CU of the two-stage stress kernel, measured at
    // 1500x1000x300 with the tile patches and 25-level z chunks: the 32-bit
    // byte-offset form, as many blocks as registers allow, in both storage
    // modes (FP32 stress 12.38 ms against 13.03 in the element-indexed form;
    // mixed 10.41 against 10.40 with a cap of 2 blocks per CU).  AWP_O32=0/1
    // (stress form) and AWP_BPC=<velocity>x<stress> (0 = as registers allow)
    // override.
    l.o32_s = true;
    l.bpc_v = 0;
    l.bpc_s = 0;
    const char *o32e = getenv("AWP_O32");
    if (o32e && (!strcmp(o32e, "0") || !strcmp(o32e, "1"))) l.o32_s = o32e[0] == '1';
    // velocity kernel form (AWP_O32V=0/1): the element-indexed form, measured
    // at 1500x1000x300 on the same run and buffers against the 32-bit
    // byte-offset form: 5.04 against 5.28 ms mixed, 5.19 against 5.45 ms FP32
    l.o32_v = false;
    const char *o32v = getenv("AWP_O32V");
    if (o32v && (!strcmp(o32v, "0") || !strcmp(o32v, "1"))) l.o32_v = o32v[0] == '1';
    const char *bpc = getenv("AWP_BPC");
    if (bpc) {
        int v = 0, s2 = 0;
        if (sscanf(bpc, "%dx%d", &v, &s2) == 2 && v >= 0 && v <= 16 && s2 >= 0 && s2 <= 16) {
            l.bpc_v = v;
            l.bpc_s = s2;
        }
    }
    set_block_caps(l, device);
    // tile patches of the two-stage kernels (AWP_SPATCH=<w>x<h>, 0x0 for
    // whole rows of tiles).  Measured at 1500x1000x300 against rows of
    // tiles: 3 x 8 patches, FP32 stress 13.13 -> 12.80 ms; 3 x 2, mixed
    // stress 12.12 -> 11.63 ms (an XCD's resident tiles then share more of
    // their halo rows in its L2).
    l.spw = 3;
    l.sph = mixed ? 2 : 8;
    const char *sp = getenv("AWP_SPATCH");
    if (sp) {
        int w = 0, h = 0;
        if (sscanf(sp, "%dx%d", &w, &h) == 2 && w >= 0 && h >= 0 && w * h <= 4096) {
            l.spw = w;
            l.sph = h;
        }
    }
    // AWP_TILE_XW=2: two-stage tiles of two waves side by side in x (128
    // points wide).  Blocks of more than 512 threads (64 x 16, 128 x 8
    // tiles) exist only in the byte-offset stress form without z unroll.
    const char *xw = getenv("AWP_TILE_XW");
    if (!fused && xw && atoi(xw) == 2 && l.tile_y <= 8) l.tile_xw = 2;
    if (!fused && l.tile_y * l.tile_xw > 8 && (!l.o32 || !l.o32_s || l.unroll != 1)) {
        l.tile_y = 8;
        l.tile_xw = 1;
    }
    if (fused) {
        // the fused step keeps its load prefetch depth (0, 1, 2) in `unroll`
        const char *pf = getenv("AWP_PF");
        // (3: no prefetch, two levels per barrier, on 64 x 11 tiles only --
        // their default: 1500x1000x300 16.17-16.27 ms per step against
        // 16.48-16.58 with one level per barrier on the same buffers)
        l.unroll = (pf && atoi(pf) >= 0 && atoi(pf) <= 3) ? atoi(pf) : (l.tile_y == 11 ? 3 : 0);
        if (l.tile_y == 11) {
            if (l.unroll != 0 && l.unroll != 3) l.unroll = 3;
        } else if (l.tile_y == 8 || l.unroll == 3) {
            l.unroll = 0;
        }
    }
    const long long bxy = (long long)((g.nx + TX - 1) / TX) * ((g.ny + l.tile_y - 1) / l.tile_y);
    int nchunk;
    if (fused && l.tile_y == 11) {
        // chunks of about 100 levels: at 1500x1000x300 with 4 x 4 patches,
        // 3 chunks ran 16.21-16.23 ms per step on two boxes; 4 chunks
        // 16.34 on one and 17.71 on the other, 2 chunks 17.92 (each chunk
        // recomputes 4 velocity levels)
        nchunk = std::max(1, (g.nz + 50) / 100);
        nchunk = std::min(nchunk, std::max(1, g.nz / 16));
        (void)bxy;
    } else if (fused) {
        // 64 x 8, 64 x 7, 64 x 4 tiles: an even number of chunks of about
        // 75 levels (1500x1000x300, 64 x 8: 17.2-17.7 ms per step with 2 or
        // 4 chunks, 19.0-19.6 ms with 3 or 5)
        nchunk = 2 * std::max(1, (g.nz + 75) / 150);
        nchunk = std::min(nchunk, std::max(1, g.nz / 16));
    } else {
        // z chunks of about SPLIT_KC levels: with the tile patches, the
        // tiles an XCD holds at once then span few levels and share their
        // halo lines in L2 (1500x1000x300: chunks of 150 levels 18.64 ms
        // per step FP32, 17.46 mixed; of 25 levels 18.29 and 16.25; of 12,
        // 18.57 and 16.37 -- the z queues' start-up loads take over)
        nchunk = split_nchunk(g.nz, SPLIT_KC);
    }
    const char *kc_env = getenv("AWP_KC");
    if (kc_env && atoi(kc_env) >= 3) nchunk = std::max(1, (g.nz + atoi(kc_env) - 1) / atoi(kc_env));
    set_chunks(l, g, nchunk, fused);
    return l;
}

extern "C" int awp_create(awp_ctx **out, const awp_in3d *p, int device, int storage, char *err, size_t errlen)
{
    if (!out || !p) {
        copy_err(err, errlen, "null argument");
        return AWP_EINVAL;
    }
    *out = nullptr;
    int rc = awp_in3d_validate(p, err, errlen);
    if (rc) return rc;
    if (storage != AWP_STORAGE_FP32 && storage != AWP_STORAGE_MIXED) {
        copy_err(err, errlen, "unknown storage mode");
        return AWP_EINVAL;
    }
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        snprintf(err, errlen, "hipSetDevice(%d): %s", device, hipGetErrorString(e));
        return AWP_EHIP;
    }
    rc = scratch_setup(device, err, errlen);
    if (rc) return rc;
    awp_ctx *ctx = new awp_ctx();
    memset(ctx->err, 0, sizeof(ctx->err));
    ctx->p = *p;
    ctx->device = device;
    ctx->storage = storage;
    ctx->g = make_geo(p->nx, p->ny, p->nz);
    ctx->fused = want_fused(ctx->g, storage);
    ctx->two_bufs = ctx->fused;
    ctx->l = choose_launch(ctx->g, ctx->fused, storage == AWP_STORAGE_MIXED, device);
    ctx->vol = nullptr;
    ctx->prof = nullptr;
    ctx->d_sjob = nullptr;
    ctx->d_claim = nullptr;
    ctx->d_fjob = nullptr;
    ctx->fgrid = ctx->ftotal = 0;
    ctx->split_total = ctx->split_fs = 0;
    ctx->sgrid_v = ctx->sgrid_s = 0;
    ctx->npsrc = 0;
    ctx->d_psrc = ctx->d_ptop = nullptr;
    ctx->d_stf = nullptr;
    ctx->npts = ctx->nkept = 0;
    ctx->nsamp_max = p->nt / p->ntiskp_sgt;
    ctx->d_kpos = nullptr;
    ctx->d_kidx = nullptr;
    ctx->d_sg = nullptr;
    ctx->d_sgt = nullptr;
    ctx->media_set = ctx->source_set = ctx->points_set = ctx->initialised = false;
    ctx->step = ctx->nsamp = 0;
    ctx->timing = false;
    ctx->t_vel = ctx->t_str = ctx->t_all = 0;
    ctx->bytes = 0;
    ctx->vol_bytes = ctx->zero_begin = ctx->zero_end = 0;
    ctx->sg_done = false;
    ctx->d_trace = nullptr;
    ctx->trace_pos = 0;

    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        snprintf(err, errlen, "hipStreamCreate: %s", hipGetErrorString(e));
        delete ctx;
        return AWP_EHIP;
    }
    // the volume arrays and their placement (vol_layout)
    const VolLayout L = vol_layout(ctx->g, storage, ctx->two_bufs);
    const size_t nbytes = L.total;
    e = hipMalloc((void **)&ctx->vol, nbytes);
    if (e != hipSuccess) {
        snprintf(err, errlen, "hipMalloc(%zu bytes) for the wavefield and media: %s", nbytes, hipGetErrorString(e));
        hipStreamDestroy(ctx->stream);
        delete ctx;
        return AWP_ENOMEM;
    }
    ctx->bytes += nbytes;
    ctx->vol_bytes = nbytes;
    place_arrays(ctx);
    MScale one = {1, 1, 1, 1};
    ctx->s.ms = one;
    // sponge profiles fx fy fz, then (mixed) the [3][pz] level profiles of
    // d1, mu, lam
    const size_t np = (size_t)(p->nx + 2) + (p->ny + 2) + (p->nz + 2);
    const size_t nmb = storage == AWP_STORAGE_MIXED ? 3 * (size_t)ctx->g.pz : 0;
    e = hipMalloc((void **)&ctx->prof, (np + nmb) * sizeof(float));
    if (e != hipSuccess) {
        snprintf(err, errlen, "hipMalloc profiles: %s", hipGetErrorString(e));
        awp_destroy(ctx);
        return AWP_ENOMEM;
    }
    ctx->bytes += (np + nmb) * sizeof(float);
    e = hipMalloc((void **)&ctx->d_claim, awp::CLAIM_WORDS * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(ctx->d_claim, 0, awp::CLAIM_WORDS * sizeof(unsigned long long));
    if (e != hipSuccess) {
        snprintf(err, errlen, "hipMalloc tile claim counters: %s", hipGetErrorString(e));
        awp_destroy(ctx);
        return AWP_ENOMEM;
    }
    ctx->s.mbar = nmb ? ctx->prof + np : nullptr;
    ctx->s.fx = ctx->prof;
    ctx->s.fy = ctx->prof + (p->nx + 2);
    ctx->s.fz = ctx->prof + (p->nx + 2) + (p->ny + 2);
    e = hipMemsetAsync(ctx->vol, 0, nbytes, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        snprintf(err, errlen, "hipMemset: %s", hipGetErrorString(e));
        awp_destroy(ctx);
        return AWP_EHIP;
    }

    // coefficients of the time step (viscoop.f:67-88), single precision
    const float dt = p->dt, dh = p->dh;
    ctx->c.dth = dt / dh;
    ctx->c.dt = dt;
    ctx->c.dh1 = 1.0f / dh;
    ctx->c.c1 = 9.0f / 8.0f;
    ctx->c.c2 = -1.0f / 24.0f;
    // sponge profiles, cerjan.f:4-247 (both faces in x and y, bottom in z)
    {
        std::vector<float> h(np, 1.0f);
        const int nd = p->nd;
        const float alpha = sqrtf(-logf(p->arbc)) / (float)nd;
        float *fx = h.data(), *fy = fx + (p->nx + 2), *fz = fy + (p->ny + 2);
        for (int i = 1; i <= p->nx; ++i) {
            float f = 1.0f;
            if (i <= nd) { float a = alpha * (float)(nd - i + 1); f = expf(-(a * a)); }
            else if (i >= p->nx - nd + 1) { float a = alpha * (float)(nd - (p->nx - i)); f = expf(-(a * a)); }
            fx[i] = f;
        }
        for (int j = 1; j <= p->ny; ++j) {
            float f = 1.0f;
            if (j <= nd) { float a = alpha * (float)(nd - j + 1); f = expf(-(a * a)); }
            else if (j >= p->ny - nd + 1) { float a = alpha * (float)(nd - (p->ny - j)); f = expf(-(a * a)); }
            fy[j] = f;
        }
        for (int k = 1; k <= p->nz; ++k) {
            float f = 1.0f;
            if (k <= nd) { float a = alpha * (float)(nd - k + 1); f = expf(-(a * a)); }
            fz[k] = f;
        }
        e = hipMemcpy(ctx->prof, h.data(), np * sizeof(float), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            snprintf(err, errlen, "hipMemcpy profiles: %s", hipGetErrorString(e));
            awp_destroy(ctx);
            return AWP_EHIP;
        }
    }
    ctx->vtst = dt / (dh * dh * dh);
    ctx->vtst1 = 1.0f / (dh * dh);
    // the reference's master rank traces its local node (ND+2, ND+2, ND+2),
    // global (ND+2, ND+2, ND+2) counted from the bottom (pmcl3d.f:1476-1479;
    // rank 0 holds the first block of every axis); clamped into the grid
    ctx->trace_ijk[0] = std::min(p->nd + 2, p->nx);
    ctx->trace_ijk[1] = std::min(p->nd + 2, p->ny);
    ctx->trace_ijk[2] = std::min(p->nd + 2, p->nz);
    ctx->trace_pos = gidx(ctx->g, ctx->trace_ijk[0], ctx->trace_ijk[1], ctx->trace_ijk[2]);
    e = hipMalloc((void **)&ctx->d_trace, (size_t)std::max(p->nt, 1) * 3 * sizeof(float));
    if (e != hipSuccess) {
        snprintf(err, errlen, "hipMalloc trace: %s", hipGetErrorString(e));
        awp_destroy(ctx);
        return AWP_ENOMEM;
    }
    ctx->bytes += (size_t)std::max(p->nt, 1) * 3 * sizeof(float);
    *out = ctx;
    return AWP_OK;
}

extern "C" int awp_destroy(awp_ctx *ctx)
{
    if (!ctx) return AWP_OK;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    for (auto &e : ctx->ev) hipEventDestroy(e);
    hipFree(ctx->vol);
    hipFree(ctx->prof);
    hipFree(ctx->d_trace);
    hipFree(ctx->d_sjob);
    hipFree(ctx->d_claim);
    hipFree(ctx->d_fjob);
    hipFree(ctx->d_psrc);
    hipFree(ctx->d_ptop);
    hipFree(ctx->d_stf);
    hipFree(ctx->d_kpos);
    hipFree(ctx->d_kidx);
    hipFree(ctx->d_sg);
    hipFree(ctx->d_sgt);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return AWP_OK;
}

// ---------------------------------------------------------------------------
// media
// ---------------------------------------------------------------------------
static int media_from_device(awp_ctx *ctx, const float *dmedia)
{
    const awp_in3d &p = ctx->p;
    // constants of structure.f:461-468 and relax.f, single precision
    const float pi = 4.0f * atanf(1.0f);
    const float w0 = 2.0f * pi * p.fp;
    const float ww1 = 2.0f * pi * p.fl;
    const float w2 = 2.0f * pi * p.fh;
    const float taumax = 1.0f / ww1;
    const float taumin = 1.0f / w2;
    MediaConst mc;
    mc.pi = pi;
    mc.tmp1 = 2.0f / pi * (logf(taumax / taumin));
    mc.tmp2 = 2.0f / pi * logf(w0 * taumin);
    mc.lw = logf(w2 / w0);
    mc.socalq = p.socalq;
    mc.nvar = p.nvar;
    // relaxation times (relax.f:4-39) and tau1/tau2 (viscoop.f:79-88)
    static const int tautem[2][2][2] = {{{1, 8}, {7, 2}}, {{6, 3}, {4, 5}}};
    const float dt1 = 1.0f / p.dt;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) {
                float t = (float)tautem[a][b][c];
                t = (t - 0.5f) / 8;
                t = 2.0f * t - 1.0f;
                const float tau = expf(0.5f * (logf(taumax * taumin) + logf(taumax / taumin) * t));
                ctx->c.tau1[(a * 2 + b) * 2 + c] = 1.0f / ((tau * dt1) + (1.0f / 2.0f));
                ctx->c.tau2[(a * 2 + b) * 2 + c] = (tau * dt1) - (1.0f / 2.0f);
            }
    if (!ctx->mixed()) {
        const Media32 m = ctx->media32();
        HIPTRY(ctx, launch_media_setup(ctx->g, m, dmedia, mc, ctx->stream));
        HIPTRY(ctx, launch_media_ghosts(ctx->g, m, ctx->stream));
        HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    } else {
        // The FP32 media are set up in five arrays of wavefield buffer 0
        // (state that awp_init zeroes before any step), then stored:
        // d1, mu, lam as FP16 deviations from a profile value per level,
        // qp and qs as FP16 of value * 2^e.
        const size_t nv = (size_t)ctx->g.nvol;
        const int nz = ctx->g.nz, pz = ctx->g.pz;
        float *tmp[5];
        for (int a = 0; a < 5; ++a) tmp[a] = ctx->wf[0][a];
        Media32 m = {tmp[0], tmp[1], tmp[2], tmp[3], tmp[4]};
        hipError_t e = hipSuccess;
        for (int a = 0; a < 5 && e == hipSuccess; ++a) e = hipMemsetAsync(tmp[a], 0, nv * sizeof(float), ctx->stream);
        if (e == hipSuccess) e = launch_media_setup(ctx->g, m, dmedia, mc, ctx->stream);
        if (e == hipSuccess) e = launch_media_ghosts(ctx->g, m, ctx->stream);
        // level profiles: (smallest + largest nonzero interior value) / 2 of
        // each level 1..nz; levels outside take level 1's and level nz's
        const int nblk = 1024;
        const size_t nrange = 3 * 2 * (size_t)nz;
        float *part = nullptr;
        if (e == hipSuccess) e = hipMalloc((void **)&part, 2 * nblk * sizeof(float) + nrange * sizeof(unsigned));
        unsigned *range = part ? (unsigned *)(part + 2 * nblk) : nullptr;
        std::vector<unsigned> hr(nrange);
        for (size_t q = 0; q < nrange; ++q) hr[q] = (q & 1) ? 0u : 0xFFFFFFFFu;
        if (e == hipSuccess) e = hipMemcpyAsync(range, hr.data(), nrange * sizeof(unsigned), hipMemcpyHostToDevice, ctx->stream);
        for (int a = 0; a < 3 && e == hipSuccess; ++a) e = launch_level_range(ctx->g, tmp[a], range + 2 * (size_t)nz * a, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(hr.data(), range, nrange * sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        auto key_value = [](unsigned key) {
            const unsigned b = (key & 0x80000000u) ? (key & 0x7FFFFFFFu) : ~key;
            float f;
            memcpy(&f, &b, 4);
            return f;
        };
        std::vector<float> bar(3 * (size_t)pz, 0.0f);
        for (int a = 0; a < 3; ++a) {
            std::vector<float> lv(nz, 0.0f);
            for (int k = 1; k <= nz; ++k) {
                const unsigned lo = hr[2 * ((size_t)nz * a + k - 1)], hi = hr[2 * ((size_t)nz * a + k - 1) + 1];
                if (lo != 0xFFFFFFFFu)
                    lv[k - 1] = (float)(((double)key_value(lo) + (double)key_value(hi)) * 0.5);
            }
            for (int k = -1; k <= nz + 2; ++k) bar[mbar_idx(ctx->g, a, k)] = lv[std::min(std::max(k, 1), nz) - 1];
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(ctx->s.mbar, bar.data(), bar.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess)
            e = launch_media_encode(ctx->g, tmp[0], (_Float16 *)ctx->s.d1, ctx->s.mbar, ctx->stream);
        if (e == hipSuccess)
            e = launch_media_encode2(ctx->g, tmp[1], tmp[2], (unsigned *)ctx->s.mu, ctx->s.mbar + (size_t)pz,
                                     ctx->s.mbar + 2 * (size_t)pz, ctx->stream);
        // qp, qs: per-array power of two putting the largest magnitude in [2^13, 2^14)
        const float *arrs[2] = {m.qp, m.qs};
        for (int a = 0; a < 2 && e == hipSuccess; ++a) e = launch_absmax(arrs[a], (long long)nv, part + a * nblk, nblk, ctx->stream);
        std::vector<float> hp(2 * nblk);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) e = hipMemcpy(hp.data(), part, hp.size() * sizeof(float), hipMemcpyDeviceToHost);
        float sc[2], isc[2];
        for (int a = 0; a < 2; ++a) {
            float mx = 0.0f;
            for (int b = 0; b < nblk; ++b) mx = std::max(mx, hp[a * nblk + b]);
            int ex = 0;
            if (mx > 0.0f) frexpf(mx, &ex);   // mx = f * 2^ex, f in [0.5, 1)
            sc[a] = ldexpf(1.0f, 14 - ex);    // mx * sc in [2^13, 2^14)
            isc[a] = ldexpf(1.0f, ex - 14);
        }
        if (e == hipSuccess)
            e = launch_to_qpair(m.qp, m.qs, (unsigned *)ctx->s.qp, sc[0], sc[1], (long long)nv, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        hipFree(part);
        if (e != hipSuccess) return fail(ctx, AWP_EHIP, "mixed media set-up: %s", hipGetErrorString(e));
        MScale ms = {sc[0], sc[1], isc[0], isc[1]};
        ctx->s.ms = ms;
    }
    ctx->sg_done = false;  // computed at awp_init from the stored mu, lam
    ctx->media_set = true;
    ctx->initialised = false;
    return AWP_OK;
}

extern "C" int awp_set_media(awp_ctx *ctx, const float *media_host)
{
    if (!ctx || !media_host) return fail(ctx, AWP_EINVAL, "null argument");
    HIPTRY(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)ctx->p.nx * ctx->p.ny * ctx->p.nz * ctx->p.nvar;
    float *d = nullptr;
    hipError_t e = hipMalloc((void **)&d, n * sizeof(float));
    if (e != hipSuccess) return fail(ctx, AWP_ENOMEM, "hipMalloc(%zu) for the media file: %s", n * 4, hipGetErrorString(e));
    e = hipMemcpy(d, media_host, n * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(d);
        return fail(ctx, AWP_EHIP, "hipMemcpy media: %s", hipGetErrorString(e));
    }
    int rc = media_from_device(ctx, d);
    hipFree(d);
    return rc;
}

extern "C" int awp_set_media_rwg(awp_ctx *ctx, const float *vp_km, const float *vs_km, const float *rho_g)
{
    if (!ctx || !vp_km || !vs_km || !rho_g) return fail(ctx, AWP_EINVAL, "null argument");
    if (ctx->p.nvar != 3) return fail(ctx, AWP_EINVAL, "RWG meshes carry vp, vs, rho: IN3D needs NVAR=3");
    HIPTRY(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)ctx->p.nx * ctx->p.ny * ctx->p.nz;
    float *buf = nullptr;
    hipError_t e = hipMalloc((void **)&buf, 6 * n * sizeof(float));
    if (e != hipSuccess)
        return fail(ctx, AWP_ENOMEM, "hipMalloc(%zu) for the RWG mesh: %s", 24 * n, hipGetErrorString(e));
    const float *src[3] = {vp_km, vs_km, rho_g};
    for (int a = 0; a < 3 && e == hipSuccess; ++a)
        e = hipMemcpy(buf + a * n, src[a], n * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = launch_rwg_media(buf, buf + n, buf + 2 * n, ctx->p.nx, ctx->p.ny, ctx->p.nz, buf + 3 * n, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        hipFree(buf);
        return fail(ctx, AWP_EHIP, "RWG mesh: %s", hipGetErrorString(e));
    }
    int rc = media_from_device(ctx, buf + 3 * n);
    hipFree(buf);
    return rc;
}

extern "C" int awp_set_media_device(awp_ctx *ctx, const float *media_dev)
{
    if (!ctx || !media_dev) return fail(ctx, AWP_EINVAL, "null argument");
    HIPTRY(ctx, hipSetDevice(ctx->device));
    return media_from_device(ctx, media_dev);
}

// ---------------------------------------------------------------------------
// sources (source.f:170-198, 378-402)
// ---------------------------------------------------------------------------
extern "C" int awp_set_source(awp_ctx *ctx, int nsrc, const int *xyz, const float *stf)
{
    if (!ctx || nsrc < 1 || !xyz || !stf) return fail(ctx, AWP_EINVAL, "bad source arguments");
    const awp_in3d &p = ctx->p;
    std::vector<long long> idx(nsrc), top(nsrc);
    for (int s = 0; s < nsrc; ++s) {
        const int x = xyz[3 * s], y = xyz[3 * s + 1], z = xyz[3 * s + 2];
        // rdsrcpos stops on coordinates outside the grid (source.f:178-193)
        if (x < 1 || x > p.nx) return fail(ctx, AWP_EINVAL, "X source coordinate %d is outside range 1-%d", x, p.nx);
        if (y < 1 || y > p.ny) return fail(ctx, AWP_EINVAL, "Y source coordinate %d is outside range 1-%d", y, p.ny);
        if (z < 1 || z > p.nz) return fail(ctx, AWP_EINVAL, "Z source coordinate %d is outside range 1-%d", z, p.nz);
        // surface sources write the top plane of the rank that owns them
        // (source.f:870-902 use nzt): only a source at z = 1 makes that the
        // free surface of the whole model
        if (p.igreen >= 4 && z != 1)
            return fail(ctx, AWP_EINVAL, "igreen=%d is a surface source; its depth index must be 1, got %d", p.igreen, z);
        idx[s] = gidx(ctx->g, x, y, p.nz + 1 - z);
        top[s] = gidx(ctx->g, x, y, p.nz);
    }
    HIPTRY(ctx, hipSetDevice(ctx->device));
    hipFree(ctx->d_psrc);
    hipFree(ctx->d_ptop);
    hipFree(ctx->d_stf);
    ctx->d_psrc = ctx->d_ptop = nullptr;
    ctx->d_stf = nullptr;
    HIPTRY(ctx, hipMalloc((void **)&ctx->d_psrc, nsrc * sizeof(long long)));
    HIPTRY(ctx, hipMalloc((void **)&ctx->d_ptop, nsrc * sizeof(long long)));
    const size_t ns = (size_t)nsrc * p.nst * 6;
    HIPTRY(ctx, hipMalloc((void **)&ctx->d_stf, ns * sizeof(float)));
    HIPTRY(ctx, hipMemcpy(ctx->d_psrc, idx.data(), nsrc * sizeof(long long), hipMemcpyHostToDevice));
    HIPTRY(ctx, hipMemcpy(ctx->d_ptop, top.data(), nsrc * sizeof(long long), hipMemcpyHostToDevice));
    HIPTRY(ctx, hipMemcpy(ctx->d_stf, stf, ns * sizeof(float), hipMemcpyHostToDevice));
    ctx->npsrc = nsrc;
    ctx->source_set = true;
    ctx->initialised = false;
    return AWP_OK;
}

// ---------------------------------------------------------------------------
// SGT points (station.f:31-80)
// ---------------------------------------------------------------------------
__global__ void k_fill(float *a, size_t n, float v)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; t < n; t += stride) a[t] = v;
}

extern "C" int awp_set_sgt_points(awp_ctx *ctx, int npts, const int *xyz, int *nkept)
{
    if (!ctx || npts < 0 || (npts > 0 && !xyz)) return fail(ctx, AWP_EINVAL, "bad point arguments");
    std::vector<int> keep(npts > 0 ? npts : 1), ijk(3 * (size_t)(npts > 0 ? npts : 1));
    int kept = awp_sgt_select(&ctx->p, npts, xyz, keep.data(), ijk.data(), ctx->err, sizeof(ctx->err));
    if (kept < 0) return kept;
    std::vector<long long> kpos;
    std::vector<int> kidx;
    kpos.reserve(kept);
    kidx.reserve(kept);
    for (int n = 0; n < npts; ++n)
        if (keep[n]) {
            kpos.push_back(gidx(ctx->g, ijk[3 * n], ijk[3 * n + 1], ijk[3 * n + 2]));
            kidx.push_back(n);
        }
    HIPTRY(ctx, hipSetDevice(ctx->device));
    hipFree(ctx->d_kpos);
    hipFree(ctx->d_kidx);
    hipFree(ctx->d_sg);
    hipFree(ctx->d_sgt);
    ctx->d_kpos = nullptr;
    ctx->d_kidx = nullptr;
    ctx->d_sg = nullptr;
    ctx->d_sgt = nullptr;
    const size_t kk = (size_t)std::max(kept, 1);
    HIPTRY(ctx, hipMalloc((void **)&ctx->d_kpos, kk * sizeof(long long)));
    HIPTRY(ctx, hipMalloc((void **)&ctx->d_kidx, kk * sizeof(int)));
    HIPTRY(ctx, hipMalloc((void **)&ctx->d_sg, kk * 3 * sizeof(float)));
    const size_t nout = (size_t)std::max(npts, 1) * 6 * std::max(ctx->nsamp_max, 1);
    hipError_t e = hipMalloc((void **)&ctx->d_sgt, nout * sizeof(float));
    if (e != hipSuccess)
        return fail(ctx, AWP_ENOMEM, "hipMalloc(%zu bytes) for the SGT buffer: %s", nout * 4, hipGetErrorString(e));
    if (kept > 0) {
        HIPTRY(ctx, hipMemcpy(ctx->d_kpos, kpos.data(), kept * sizeof(long long), hipMemcpyHostToDevice));
        HIPTRY(ctx, hipMemcpy(ctx->d_kidx, kidx.data(), kept * sizeof(int), hipMemcpyHostToDevice));
    }
    ctx->keep.assign(keep.begin(), keep.begin() + npts);
    ctx->npts = npts;
    ctx->nkept = kept;
    ctx->points_set = true;
    ctx->initialised = false;
    ctx->sg_done = false;   // d_sg is new: awp_init fills it
    if (nkept) *nkept = kept;
    return AWP_OK;
}

// ---------------------------------------------------------------------------
// run
// ---------------------------------------------------------------------------
// Uploads the two-stage step's one-entry job table for the run's current
// launch shape and wavefield buffer (after the stream has drained: its
// kernels read the table).
static int upload_sjob(awp_ctx *ctx)
{
    HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    if (!ctx->d_sjob) HIPTRY(ctx, hipMalloc((void **)&ctx->d_sjob, sizeof(SJob)));
    const SJob q = split_job(ctx->g, ctx->view(ctx->cur), ctx->c, ctx->l, 0, 0);
    ctx->split_total = split_tiles(q);
    ctx->split_fs = split_fs_blocks(q);
    awp::split_batch_grids(ctx->mixed(), ctx->l, ctx->device, &ctx->sgrid_v, &ctx->sgrid_s);
    HIPTRY(ctx, hipMemcpy(ctx->d_sjob, &q, sizeof(SJob), hipMemcpyHostToDevice));
    return AWP_OK;
}

// Uploads the claimed fused step's one-entry job table (buffer 0 and 1 of
// the run; a launch's parity picks which is read) for the run's launch
// shape, and sizes its persistent grid.
static int upload_fjob(awp_ctx *ctx)
{
    HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    if (!ctx->d_fjob) HIPTRY(ctx, hipMalloc((void **)&ctx->d_fjob, sizeof(awp::FJob)));
    awp::FJob f;
    f.g = ctx->g;
    f.s[0] = ctx->view(0);
    f.s[1] = ctx->view(1);
    f.c = ctx->c;
    f.kc = ctx->l.kc;
    f.t = awp::fused_tiles(ctx->g, ctx->l);
    f.tile0 = 0;
    ctx->ftotal = f.t.nbx * f.t.nby * f.t.nbz;
    ctx->fgrid = awp::fused_batch_grid(ctx->mixed(), ctx->l, ctx->device);
    HIPTRY(ctx, hipMemcpy(ctx->d_fjob, &f, sizeof(f), hipMemcpyHostToDevice));
    return AWP_OK;
}

// the two-stage step's kernels in buffer 0: claimed (persistent blocks) or
// one block per tile; e: the timing events of the step, or null
static int split_launch(awp_ctx *ctx, hipStream_t st, hipEvent_t *e)
{
    const bool cl = ctx->l.claim == 2 ||
                    (ctx->l.claim == 1 && ctx->split_total >= 4 * std::max(ctx->sgrid_v, ctx->sgrid_s));
    const int gv = cl ? ctx->sgrid_v : 1 << 30, gs = cl ? ctx->sgrid_s : 1 << 30;
    unsigned long long *claim = cl ? ctx->d_claim : nullptr;
    HIPTRY(ctx, awp::launch_split_velocity(ctx->mixed(), ctx->l, ctx->l.o32, ctx->d_sjob, 1, ctx->split_total, gv,
                                           claim, st));
    if (e) HIPTRY(ctx, hipEventRecord(e[1], st));
    HIPTRY(ctx, awp::launch_split_fs(ctx->mixed(), ctx->d_sjob, 1, ctx->split_fs, st));
    if (e) HIPTRY(ctx, hipEventRecord(e[2], st));
    HIPTRY(ctx, awp::launch_split_stress(ctx->mixed(), ctx->l, ctx->l.o32, ctx->d_sjob, 1, ctx->split_total, gs,
                                         claim, st));
    return AWP_OK;
}

// the fused step from buffer cur to cur ^ 1: claimed (persistent blocks over
// the one-entry job table) or static (one block per tile)
static hipError_t fused_launch(awp_ctx *ctx, hipStream_t st)
{
    if (ctx->l.claim == 2 || (ctx->l.claim == 1 && ctx->ftotal >= 4 * ctx->fgrid))
        return awp::launch_step_fused_batch(ctx->mixed(), ctx->l, ctx->d_fjob, 1, ctx->ftotal, ctx->fgrid, ctx->cur,
                                            ctx->d_claim, st);
    return awp::launch_step_fused(ctx->mixed(), ctx->g, ctx->view(ctx->cur), ctx->view(ctx->cur ^ 1), ctx->c,
                                  ctx->l, st);
}

// After ctx->fused or ctx->l changed: the two-stage step works in buffer 0,
// so a state the fused step left in buffer 1 moves there first (the whole
// buffer: every value the next step reads), and its job table follows the
// launch shape.
static int scheme_changed(awp_ctx *ctx)
{
    if (!ctx->fused && ctx->cur == 1) {
        for (int a = 0; a < 9; ++a)
            HIPTRY(ctx, hipMemcpyAsync(ctx->wf[0][a], ctx->wf[1][a], (size_t)ctx->g.nvol * sizeof(float),
                                       hipMemcpyDeviceToDevice, ctx->stream));
        ctx->cur = 0;
    }
    if (!ctx->initialised) return AWP_OK;
    return ctx->fused ? upload_fjob(ctx) : upload_sjob(ctx);
}

extern "C" int awp_init(awp_ctx *ctx)
{
    if (!ctx) return AWP_EINVAL;
    if (!ctx->media_set || !ctx->source_set || !ctx->points_set)
        return fail(ctx, AWP_ESTATE, "awp_init needs media, source and SGT points first");
    HIPTRY(ctx, hipSetDevice(ctx->device));
    // the exponent slots and the state room (wavefield buffers, memory variables)
    HIPTRY(ctx, hipMemsetAsync((char *)ctx->vol + ctx->zero_begin, 0, ctx->zero_end - ctx->zero_begin, ctx->stream));
    const size_t nout = (size_t)std::max(ctx->npts, 1) * 6 * std::max(ctx->nsamp_max, 1);
    hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, ctx->stream, ctx->d_sgt, nout, -1.0f);
    HIPTRY(ctx, hipGetLastError());
    if (!ctx->sg_done) {
        HIPTRY(ctx, launch_sgt_coef(ctx->g, ctx->s, ctx->nkept, ctx->d_kpos, ctx->d_sg, ctx->stream));
        ctx->sg_done = true;
    }
    // the two-stage step's one-run job table (state, coefficients, tiles)
    ctx->cur = 0;
    {
        int rc = ctx->fused ? upload_fjob(ctx) : upload_sjob(ctx);
        if (rc) return rc;
    }
    // first source sample before the loop (pmcl3d.f:988)
    HIPTRY(ctx, launch_source(ctx->g, ctx->view(0), ctx->p.igreen, ctx->npsrc, ctx->d_psrc, ctx->d_ptop,
                              ctx->d_stf, ctx->p.nst, 1, ctx->vtst, ctx->vtst1, true, nullptr, 0, ctx->stream));
    HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->step = 0;
    ctx->nsamp = 0;
    ctx->initialised = true;
    return AWP_OK;
}

static int ensure_events(awp_ctx *ctx, size_t n)
{
    while (ctx->ev.size() < n) {
        hipEvent_t e;
        HIPTRY(ctx, hipEventCreate(&e));
        ctx->ev.push_back(e);
    }
    return AWP_OK;
}

extern "C" int awp_step(awp_ctx *ctx, int nsteps)
{
    if (!ctx) return AWP_EINVAL;
    if (!ctx->initialised) return fail(ctx, AWP_ESTATE, "awp_step before awp_init");
    if (nsteps < 0) return fail(ctx, AWP_EINVAL, "nsteps < 0");
    HIPTRY(ctx, hipSetDevice(ctx->device));
    const awp_in3d &p = ctx->p;
    const int todo = std::min(nsteps, p.nt - ctx->step);
    if (ctx->timing) {
        int rc = ensure_events(ctx, (size_t)6 * std::max(todo, 1));
        if (rc) return rc;
    }
    hipStream_t st = ctx->stream;
    if (!ctx->fused && ctx->cur != 0 && todo > 0) {
        // a fused batch stepped this two-stage run and left its state in
        // buffer 1: the two-stage step works in buffer 0 (its job table
        // points there)
        int rc = scheme_changed(ctx);
        if (rc) return rc;
    }
    for (int n = 0; n < todo; ++n) {
        const int i = ++ctx->step;
        hipEvent_t *e = ctx->timing ? &ctx->ev[(size_t)6 * n] : nullptr;
        if (e) HIPTRY(ctx, hipEventRecord(e[0], st));
        const bool mx = ctx->mixed();
        State sn;   // the state after this step
        if (ctx->fused) {
            const State sb = ctx->view(ctx->cur ^ 1);
            if (e) HIPTRY(ctx, hipEventRecord(e[1], st));
            if (e) HIPTRY(ctx, hipEventRecord(e[2], st));
            HIPTRY(ctx, fused_launch(ctx, st));
            ctx->cur ^= 1;
            sn = sb;
        } else {
            // the stage kernels over this run's one-entry job table (in
            // buffer 0)
            sn = ctx->view(ctx->cur);
            int rc = split_launch(ctx, st, e);
            if (rc) return rc;
        }
        if (e) HIPTRY(ctx, hipEventRecord(e[3], st));
        // source sample i + 1 (while the source lasts), then the trace row
        HIPTRY(ctx, launch_source(ctx->g, sn, p.igreen, ctx->npsrc, ctx->d_psrc, ctx->d_ptop, ctx->d_stf, p.nst,
                                  i + 1, ctx->vtst, ctx->vtst1, i + 1 <= p.nst, ctx->d_trace, ctx->trace_pos, st));
        if (i % p.ntiskp_sgt == 0 && ctx->nsamp < ctx->nsamp_max) {
            HIPTRY(ctx, launch_sgt_gather(sn, ctx->nkept, ctx->d_kpos, ctx->d_kidx, ctx->d_sg, ctx->d_sgt,
                                          ctx->nsamp_max, ctx->nsamp, st));
            ctx->nsamp++;
        }
        if (e) HIPTRY(ctx, hipEventRecord(e[4], st));
    }
    if (ctx->timing && todo > 0) {
        HIPTRY(ctx, hipStreamSynchronize(st));
        double tv = 0, ts = 0, ta = 0;
        for (int n = 0; n < todo; ++n) {
            hipEvent_t *e = &ctx->ev[(size_t)6 * n];
            float a, b, c;
            HIPTRY(ctx, hipEventElapsedTime(&a, e[0], e[1]));
            HIPTRY(ctx, hipEventElapsedTime(&b, e[2], e[3]));
            HIPTRY(ctx, hipEventElapsedTime(&c, e[0], e[4]));
            tv += a; ts += b; ta += c;
        }
        ctx->t_vel = tv;
        ctx->t_str = ts;
        ctx->t_all = ta;
    }
    return AWP_OK;
}

extern "C" int awp_set_scheme(awp_ctx *ctx, int fused)
{
    if (!ctx) return AWP_EINVAL;
    if (fused && !ctx->two_bufs)
        return fail(ctx, AWP_EINVAL, "the fused step needs the second wavefield buffer (a run created with the "
                                     "fused scheme)");
    HIPTRY(ctx, hipSetDevice(ctx->device));
    HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->fused = fused != 0;
    ctx->l = choose_launch(ctx->g, ctx->fused, ctx->mixed(), ctx->device);
    return scheme_changed(ctx);
}

extern "C" int awp_get_scheme(const awp_ctx *ctx, int *fused, int *tile_y, int *nchunk)
{
    if (!ctx) return AWP_EINVAL;
    if (fused) *fused = ctx->fused ? 1 : 0;
    if (tile_y) *tile_y = ctx->l.tile_y;
    if (nchunk) *nchunk = ctx->l.nchunk;
    return AWP_OK;
}

namespace {

// `n` steps of the step kernels alone (no source, no SGT sample), for timing
int trial_steps(awp_ctx *ctx, int n)
{
    for (int k = 0; k < n; ++k) {
        if (ctx->fused) {
            HIPTRY(ctx, fused_launch(ctx, ctx->stream));
            ctx->cur ^= 1;
        } else {
            int rc = split_launch(ctx, ctx->stream, nullptr);
            if (rc) return rc;
        }
    }
    return AWP_OK;
}

// one warm-up step, then `reps` timed ones (milliseconds in *ms)
int time_trial(awp_ctx *ctx, int reps, float *ms)
{
    int r = trial_steps(ctx, 1);
    if (r) return r;
    HIPTRY(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    r = trial_steps(ctx, reps);
    if (r) return r;
    HIPTRY(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    HIPTRY(ctx, hipEventSynchronize(ctx->ev[1]));
    HIPTRY(ctx, hipEventElapsedTime(ms, ctx->ev[0], ctx->ev[1]));
    return AWP_OK;
}

}  // namespace

extern "C" int awp_tune(awp_ctx *ctx, int reps)
{
    if (!ctx) return AWP_EINVAL;
    if (!ctx->initialised || ctx->step != 0)
        return fail(ctx, AWP_ESTATE, "awp_tune needs an initialised run that has not stepped");
    if (reps < 1) reps = 3;
    HIPTRY(ctx, hipSetDevice(ctx->device));
    int rc = ensure_events(ctx, 2);
    if (rc) return rc;
    // candidates: the two-stage step with its defaults, and on grids of
    // fewer tiles than the GPU holds blocks (200x200x100: 400 tiles) with
    // shorter z chunks and 64 x 4 tiles for more of them; with both buffers,
    // the fused step on 64 x 11 tiles (FP32; z chunks of about 100, 75, 60
    // levels x tile patches 4x4, 2x2, 8x2, and larger patches) and 64 x 8
    // tiles, or on 64 x 7 tiles (mixed).  The fastest is kept.
    struct Cand {
        bool fused;
        int ty, kc, pw, ph;
    };
    std::vector<Cand> cands;
    cands.push_back({false, 0, 0, 0, 0});
    {
        const long long tiles = (long long)((ctx->g.nx + TX - 1) / TX) * ((ctx->g.ny + 7) / 8) *
                                split_nchunk(ctx->g.nz, SPLIT_KC);
        if (tiles < 4096)
            for (int ty : {8, 4})
                for (int kc : {12, 8}) cands.push_back({false, ty, kc, 0, 0});
    }
    if (ctx->two_bufs) {
        if (!ctx->mixed()) {
            const int pat[3][2] = {{4, 4}, {2, 2}, {8, 2}};
            for (int kc : {100, 75, 60})
                for (const auto &pp : pat) cands.push_back({true, 11, kc, pp[0], pp[1]});
            // larger patches: more of an XCD's 32 resident tiles are
            // neighbours that share their halo lines
            const int big[3][2] = {{4, 8}, {8, 4}, {6, 5}};
            for (const auto &pp : big) cands.push_back({true, 11, 100, pp[0], pp[1]});
            cands.push_back({true, 11, 75, 3, 3});
            // two chunks: less velocity recomputed at chunk edges
            cands.push_back({true, 11, 150, 4, 4});
            cands.push_back({true, 11, 150, 2, 2});
            cands.push_back({true, 8, 75, 4, 4});
            // small grids: more z chunks, or 64 x 4 tiles, for more tiles
            const long long fxy = (long long)((ctx->g.nx + TX - 1) / TX) * ((ctx->g.ny + 10) / 11);
            if (fxy * std::max(1, ctx->g.nz / 100) < 4096) {
                for (int kc : {25, 16}) cands.push_back({true, 11, kc, 2, 2});
                cands.push_back({true, 4, 16, 4, 4});
            }
        } else {
            cands.push_back({true, 7, 75, 4, 4});
        }
    }
    auto apply = [ctx](const Cand &cd) -> int {
        HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
        ctx->fused = cd.fused;
        Launch l = choose_launch(ctx->g, cd.fused, ctx->mixed(), ctx->device);
        if (!cd.fused && cd.ty) {
            l.tile_y = cd.ty;
            l.tile_xw = 1;
            set_chunks(l, ctx->g, split_nchunk(ctx->g.nz, cd.kc), false);
        }
        if (cd.fused) {
            l.tile_y = cd.ty;
            l.unroll = cd.ty == 11 ? 3 : 0;   // 64 x 11: two levels per barrier
            l.pw = cd.pw;
            l.ph = cd.ph;
            set_chunks(l, ctx->g, std::min((ctx->g.nz + cd.kc / 2) / cd.kc, std::max(1, ctx->g.nz / 16)), true);
        }
        ctx->l = l;
        return scheme_changed(ctx);
    };
    // one warm-up step, then `reps` timed ones (ms in *ms)
    auto time_steps = [ctx, reps](float *ms) -> int { return time_trial(ctx, reps, ms); };
    // the fastest of the candidates
    size_t pick = 0;
    // AWP_TUNE_LOG=1: every timing on stderr
    const char *tl = getenv("AWP_TUNE_LOG");
    const bool log = tl && atoi(tl) == 1;
    // every candidate once, then the three fastest again (their first
    // timings lie within a percent of each other at 1500x1000x300)
    auto search = [&]() -> int {
        std::vector<std::pair<float, size_t>> first;
        for (size_t n = 0; n < cands.size(); ++n) {
            float ms = 0;
            int r = apply(cands[n]);
            if (!r) r = time_steps(&ms);
            if (r) return r;
            if (log)
                fprintf(stderr, "awp_tune: %s ty %d kc %d patch %dx%d: %.3f ms per step\n",
                        cands[n].fused ? "fused" : "two-stage", cands[n].ty, cands[n].kc, cands[n].pw, cands[n].ph,
                        ms / reps);
            first.push_back({ms, n});
        }
        std::sort(first.begin(), first.end());
        float best = 1e30f;
        pick = first[0].second;
        for (size_t q = 0; q < std::min<size_t>(3, first.size()); ++q) {
            float ms = 0, ms2 = 0;
            const size_t n = first[q].second;
            int r = apply(cands[n]);
            if (!r) r = time_steps(&ms);
            if (!r) r = time_steps(&ms2);
            if (r) return r;
            if (log)
                fprintf(stderr, "awp_tune again: %s ty %d kc %d patch %dx%d: %.3f ms per step\n",
                        cands[n].fused ? "fused" : "two-stage", cands[n].ty, cands[n].kc, cands[n].pw, cands[n].ph,
                        (ms + ms2) / (2 * reps));
            if (ms + ms2 < best) {
                best = ms + ms2;
                pick = n;
            }
        }
        return apply(cands[pick]);
    };
    rc = search();
    if (rc) return rc;
    // the trial steps changed the state: start the run again
    return awp_init(ctx);
}

extern "C" int awp_synchronize(awp_ctx *ctx)
{
    if (!ctx) return AWP_EINVAL;
    HIPTRY(ctx, hipSetDevice(ctx->device));
    HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    return AWP_OK;
}

extern "C" int awp_progress(const awp_ctx *ctx, int *step, int *nsamples)
{
    if (!ctx) return AWP_EINVAL;
    if (step) *step = ctx->step;
    if (nsamples) *nsamples = ctx->nsamp;
    return AWP_OK;
}

extern "C" int awp_set_timing(awp_ctx *ctx, int on)
{
    if (!ctx) return AWP_EINVAL;
    ctx->timing = on != 0;
    return AWP_OK;
}

extern "C" int awp_get_timing(awp_ctx *ctx, double *out3)
{
    if (!ctx || !out3) return AWP_EINVAL;
    out3[0] = ctx->t_vel;
    out3[1] = ctx->t_str;
    out3[2] = ctx->t_all;
    return AWP_OK;
}

extern "C" int awp_estimate_bytes(const awp_in3d *p, int storage, int npts, size_t *steady, size_t *peak)
{
    if (!p || (storage != AWP_STORAGE_FP32 && storage != AWP_STORAGE_MIXED) || npts < 0) return AWP_EINVAL;
    const Geo g = make_geo(p->nx, p->ny, p->nz);
    const bool fused = want_fused(g, storage);
    const size_t nv = (size_t)g.nvol;
    const VolLayout L = vol_layout(g, storage, fused);
    const size_t vol = L.total;
    const size_t prof = ((size_t)p->nx + p->ny + p->nz + 6 + (storage == AWP_STORAGE_MIXED ? 3 * (size_t)g.pz : 0)) *
                        sizeof(float);
    const size_t nsamp_max = (size_t)std::max(p->nt / std::max(p->ntiskp_sgt, 1), 1);
    const size_t sgt = (size_t)std::max(npts, 1) * 6 * nsamp_max * sizeof(float);
    const size_t pts = (size_t)std::max(npts, 1) * (sizeof(long long) + sizeof(int) + 3 * sizeof(float));
    const size_t src = (size_t)p->nsrc * (2 * sizeof(long long) + (size_t)p->nst * 6 * sizeof(float));
    const size_t trace = (size_t)std::max(p->nt, 1) * 3 * sizeof(float);
    const size_t st = vol + prof + sgt + pts + src + trace;
    // awp_set_media: the media file, plus (mixed mode) the reductions' small
    // buffers; the FP32 media of the mixed set-up live in wavefield buffer 0
    size_t tr = (size_t)p->nx * p->ny * p->nz * p->nvar * sizeof(float);
    if (storage == AWP_STORAGE_MIXED) tr += 2 * 1024 * sizeof(float) + 6 * (size_t)p->nz * sizeof(unsigned);
    (void)nv;
    if (steady) *steady = st;
    if (peak) *peak = st + tr;
    return AWP_OK;
}

extern "C" int awp_device_memory(int device, size_t *free_bytes, size_t *total_bytes)
{
    size_t f = 0, t = 0;
    if (hipSetDevice(device) != hipSuccess || hipMemGetInfo(&f, &t) != hipSuccess) return AWP_EHIP;
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return AWP_OK;
}

extern "C" int awp_device_bytes(const awp_ctx *ctx, size_t *bytes)
{
    if (!ctx || !bytes) return AWP_EINVAL;
    size_t b = ctx->bytes;
    b += (size_t)std::max(ctx->npts, 1) * 6 * std::max(ctx->nsamp_max, 1) * sizeof(float);
    b += (size_t)std::max(ctx->nkept, 1) * (sizeof(long long) + sizeof(int) + 3 * sizeof(float));
    b += (size_t)ctx->npsrc * (2 * sizeof(long long) + (size_t)ctx->p.nst * 6 * sizeof(float));
    *bytes = b;
    return AWP_OK;
}

// ---------------------------------------------------------------------------
// outputs
// ---------------------------------------------------------------------------
extern "C" int awp_get_sgt(awp_ctx *ctx, float *host, size_t count)
{
    if (!ctx || !host) return AWP_EINVAL;
    if (!ctx->points_set) return fail(ctx, AWP_ESTATE, "no SGT points");
    const size_t n = (size_t)ctx->npts * 6 * ctx->nsamp_max;
    if (count < n) return fail(ctx, AWP_EINVAL, "buffer holds %zu floats, need %zu", count, n);
    HIPTRY(ctx, hipSetDevice(ctx->device));
    HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    if (n) HIPTRY(ctx, hipMemcpy(host, ctx->d_sgt, n * sizeof(float), hipMemcpyDeviceToHost));
    return AWP_OK;
}

extern "C" int awp_get_trace(awp_ctx *ctx, float *host, size_t count, int *ijk)
{
    if (!ctx) return AWP_EINVAL;
    if (ijk) {
        for (int a = 0; a < 3; ++a) ijk[a] = ctx->trace_ijk[a];
    }
    if (!host) return AWP_OK;
    const size_t n = (size_t)ctx->step * 3;
    if (count < n) return fail(ctx, AWP_EINVAL, "buffer holds %zu floats, need %zu", count, n);
    HIPTRY(ctx, hipSetDevice(ctx->device));
    HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    if (n) HIPTRY(ctx, hipMemcpy(host, ctx->d_trace, n * sizeof(float), hipMemcpyDeviceToHost));
    return AWP_OK;
}

extern "C" int awp_write_sgt_file(awp_ctx *ctx, const char *path)
{
    if (!ctx || !path) return AWP_EINVAL;
    if (!ctx->points_set) return fail(ctx, AWP_ESTATE, "no SGT points");
    const awp_in3d &p = ctx->p;
    int nt_out = 0, nflush = 0;
    awp_sgt_file_geometry(&p, &nt_out, &nflush);
    // buffers that filled during the steps run so far
    const int avail = std::min(nflush, ctx->nsamp / p.write_step);
    if (avail <= 0) return AWP_OK;   // the reference opens the file at the first flush
    const int ws = p.write_step, ms = ctx->nsamp_max;
    const size_t pts_per_chunk = 16384;
    std::vector<float> host(pts_per_chunk * 6 * (size_t)ms);
    int fd = open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (fd < 0) return fail(ctx, AWP_EIO, "cannot create '%s'", path);
    HIPTRY(ctx, hipSetDevice(ctx->device));
    HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    const size_t slot = (size_t)4 * nt_out;          // bytes per (point, component)
    std::vector<char> img;
    for (size_t g0 = 0; g0 < (size_t)ctx->npts; g0 += pts_per_chunk) {
        const size_t g1 = std::min((size_t)ctx->npts, g0 + pts_per_chunk);
        size_t last = g1;   // one past the last kept point of the chunk
        while (last > g0 && !ctx->keep[last - 1]) --last;
        size_t first = g0;
        while (first < last && !ctx->keep[first]) ++first;
        if (first >= last) continue;
        const size_t n = (size_t)(last - g0) * 6 * ms;
        hipError_t e = hipMemcpy(host.data(), ctx->d_sgt + g0 * 6 * (size_t)ms, n * sizeof(float), hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            close(fd);
            return fail(ctx, AWP_EHIP, "hipMemcpy SGT: %s", hipGetErrorString(e));
        }
        // image of the file bytes [first*24*nt_out, end of last point's last flush)
        const size_t off0 = first * 6 * slot;
        const size_t end = (last - 1) * 6 * slot + 5 * slot + (size_t)(avail - 1) * 4 * ws + 4 * (size_t)ws;
        img.assign(end - off0, 0);
        for (size_t g = first; g < last; ++g) {
            if (!ctx->keep[g]) continue;
            for (int c = 0; c < 6; ++c)
                for (int f = 1; f <= avail; ++f) {
                    const size_t off = g * 6 * slot + c * slot + (size_t)(f - 1) * 4 * ws;
                    memcpy(&img[off - off0], &host[((g - g0) * 6 + c) * (size_t)ms + (size_t)(f - 1) * ws],
                           4 * (size_t)ws);
                }
        }
        size_t done = 0;
        while (done < img.size()) {
            ssize_t w = pwrite(fd, img.data() + done, img.size() - done, (off_t)(off0 + done));
            if (w <= 0) {
                close(fd);
                return fail(ctx, AWP_EIO, "write '%s' failed", path);
            }
            done += (size_t)w;
        }
    }
    if (close(fd)) return fail(ctx, AWP_EIO, "close '%s' failed", path);
    return AWP_OK;
}

extern "C" int awp_write_sgt_file_rwg(awp_ctx *ctx, const char *path, int zsource)
{
    if (!ctx || !path) return AWP_EINVAL;
    if (!ctx->points_set) return fail(ctx, AWP_ESTATE, "no SGT points");
    const awp_in3d &p = ctx->p;
    int nt_out = 0, nflush = 0;
    awp_sgt_file_geometry(&p, &nt_out, &nflush);
    const int avail = std::min(nflush, ctx->nsamp / p.write_step);
    if (avail <= 0) return AWP_OK;   // no raw file, nothing to reformat
    // the raw file ends with the last kept point's last full buffer; the
    // reformatter takes the complete 6 * NT_OUT records of it
    long long last = -1;
    for (long long g = (long long)ctx->npts - 1; g >= 0; --g)
        if (ctx->keep[g]) {
            last = g;
            break;
        }
    const size_t slot = (size_t)4 * nt_out;
    const size_t raw_size = last < 0 ? 0 : (size_t)last * 6 * slot + 5 * slot + (size_t)avail * 4 * p.write_step;
    const size_t nrec = raw_size / (6 * slot);
    const int navail = avail * p.write_step;
    FILE *f = fopen(path, "wb");
    if (!f) return fail(ctx, AWP_EIO, "cannot create '%s'", path);
    HIPTRY(ctx, hipSetDevice(ctx->device));
    HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    const size_t chunk = 8192;
    const size_t per = 6 * (size_t)nt_out;
    float *d_out = nullptr;
    int *d_keep = nullptr;
    std::vector<float> host(chunk * per);
    std::vector<int> kk(chunk);
    hipError_t e = hipMalloc((void **)&d_out, std::max<size_t>(chunk * per, 1) * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&d_keep, chunk * sizeof(int));
    bool ok = e == hipSuccess;
    for (size_t g0 = 0; ok && g0 < nrec; g0 += chunk) {
        const size_t n = std::min(chunk, nrec - g0);
        for (size_t k = 0; k < n; ++k) kk[k] = ctx->keep[g0 + k];
        e = hipMemcpy(d_keep, kk.data(), n * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = launch_sgt_rwg(ctx->d_sgt, d_keep, (int)g0, (int)n, ctx->nsamp_max, nt_out, navail, zsource ? 1 : 0,
                               d_out, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) e = hipMemcpy(host.data(), d_out, n * per * sizeof(float), hipMemcpyDeviceToHost);
        ok = e == hipSuccess && fwrite(host.data(), sizeof(float), n * per, f) == n * per;
    }
    hipFree(d_out);
    hipFree(d_keep);
    ok = (fclose(f) == 0) && ok;
    if (e != hipSuccess) return fail(ctx, AWP_EHIP, "RWG SGT: %s", hipGetErrorString(e));
    return ok ? AWP_OK : fail(ctx, AWP_EIO, "write '%s' failed", path);
}

extern "C" int awp_write_sgt_file_fp16(awp_ctx *ctx, const char *path)
{
    if (!ctx || !path) return AWP_EINVAL;
    if (!ctx->points_set) return fail(ctx, AWP_ESTATE, "no SGT points");
    const awp_in3d &p = ctx->p;
    int nt_out = 0, nflush = 0;
    awp_sgt_file_geometry(&p, &nt_out, &nflush);
    const int avail = std::min(nflush, ctx->nsamp / p.write_step);
    const int nsamp = avail * p.write_step;   // the samples the raw file holds
    const size_t npts = (size_t)ctx->npts;
    HIPTRY(ctx, hipSetDevice(ctx->device));
    float *d_scale = nullptr;
    _Float16 *d_data = nullptr;
    const size_t ndata = npts * 6 * (size_t)std::max(nsamp, 1);
    HIPTRY(ctx, hipMalloc((void **)&d_scale, std::max<size_t>(npts, 1) * 6 * sizeof(float)));
    hipError_t e = hipMalloc((void **)&d_data, ndata * sizeof(_Float16));
    if (e != hipSuccess) {
        hipFree(d_scale);
        return fail(ctx, AWP_ENOMEM, "hipMalloc FP16 SGT: %s", hipGetErrorString(e));
    }
    e = hipMemsetAsync(d_scale, 0, std::max<size_t>(npts, 1) * 6 * sizeof(float), ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d_data, 0, ndata * sizeof(_Float16), ctx->stream);
    if (e == hipSuccess && nsamp > 0)
        e = launch_sgt_fp16(ctx->d_sgt, ctx->d_kidx, ctx->nkept, ctx->nsamp_max, nsamp, d_scale, d_data, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    std::vector<float> scale(npts * 6);
    std::vector<uint16_t> data(npts * 6 * (size_t)nsamp);
    if (e == hipSuccess && npts) e = hipMemcpy(scale.data(), d_scale, npts * 6 * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess && !data.empty())
        e = hipMemcpy(data.data(), d_data, data.size() * sizeof(uint16_t), hipMemcpyDeviceToHost);
    hipFree(d_scale);
    hipFree(d_data);
    if (e != hipSuccess) return fail(ctx, AWP_EHIP, "FP16 SGT: %s", hipGetErrorString(e));
    unsigned char hdr[64];
    memset(hdr, 0, sizeof(hdr));
    memcpy(hdr, "AWPSGT16", 8);
    const uint32_t h32[4] = {1u, (uint32_t)npts, 6u, (uint32_t)nsamp};
    memcpy(hdr + 8, h32, sizeof(h32));
    const uint32_t ntk = (uint32_t)p.ntiskp_sgt;
    memcpy(hdr + 24, &ntk, 4);
    memcpy(hdr + 28, &p.dt, 4);
    const uint32_t nto = (uint32_t)nt_out;   // the raw file's series length
    memcpy(hdr + 32, &nto, 4);
    FILE *f = fopen(path, "wb");
    if (!f) return fail(ctx, AWP_EIO, "cannot create '%s'", path);
    bool ok = fwrite(hdr, 1, 64, f) == 64;
    ok = ok && fwrite(scale.data(), sizeof(float), scale.size(), f) == scale.size();
    ok = ok && fwrite(data.data(), sizeof(uint16_t), data.size(), f) == data.size();
    ok = (fclose(f) == 0) && ok;
    return ok ? AWP_OK : fail(ctx, AWP_EIO, "write '%s' failed", path);
}

extern "C" int awp_get_field(awp_ctx *ctx, int which, int kext, float *host)
{
    if (!ctx || !host || which < 0 || which >= 20 || (kext != 0 && kext != 2)) return fail(ctx, AWP_EINVAL, "bad field request");
    HIPTRY(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)ctx->p.nx * ctx->p.ny * (ctx->p.nz + kext);
    float *d = nullptr;
    hipError_t e = hipMalloc((void **)&d, n * sizeof(float));
    if (e != hipSuccess) return fail(ctx, AWP_ENOMEM, "hipMalloc field: %s", hipGetErrorString(e));
    e = launch_pack_field(ctx->mixed(), ctx->g, ctx->view(ctx->cur), which, kext, d, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(host, d, n * sizeof(float), hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess) return fail(ctx, AWP_EHIP, "field copy: %s", hipGetErrorString(e));
    return AWP_OK;
}

// ---------------------------------------------------------------------------
// Field readback without allocation (awp_get_field_noalloc).  The staging
// buffer is a file-scope device array, `scratch`, of SCRATCH_FLOATS floats
// (32 MiB; every GPU the library runs on holds its own copy until the process
// ends).  Its device address is taken once per device with that device
// current (scratch_setup, called by awp_create) and kept in scratch_dev[];
// host code uses only that address, never the array's name.  Every call on
// one GPU shares the array: calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t SCRATCH_FLOATS = (size_t)8 << 20;
constexpr int SCRATCH_MAX_DEVICES = 64;
__device__ __align__(256) float scratch[SCRATCH_FLOATS];
float *scratch_dev[SCRATCH_MAX_DEVICES];
std::mutex scratch_mu;

// the address of `scratch` on `device` (the current device), once per device
int scratch_setup(int device, char *err, size_t errlen)
{
    if (device < 0 || device >= SCRATCH_MAX_DEVICES) {
        snprintf(err, errlen, "device %d beyond the %d the readback scratch serves", device, SCRATCH_MAX_DEVICES);
        return AWP_EINVAL;
    }
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (scratch_dev[device]) return AWP_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        snprintf(err, errlen, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        return AWP_EHIP;
    }
    scratch_dev[device] = (float *)ptr;
    return AWP_OK;
}
}  // namespace

extern "C" int awp_get_field_noalloc(awp_ctx *ctx, int which, int kext, float *host)
{
    if (!ctx || !host || which < 0 || which >= 20 || (kext != 0 && kext != 2)) return fail(ctx, AWP_EINVAL, "bad field request");
    const size_t plane = (size_t)ctx->p.nx * ctx->p.ny;
    if (plane > SCRATCH_FLOATS)
        return fail(ctx, AWP_EINVAL, "awp_get_field_noalloc: a level of %zu points exceeds the %zu-point scratch", plane,
                    SCRATCH_FLOATS);
    float *d = nullptr;
    {
        std::lock_guard<std::mutex> lk(scratch_mu);
        d = scratch_dev[ctx->device];
    }
    if (!d) return fail(ctx, AWP_ESTATE, "awp_get_field_noalloc: no scratch address for device %d", ctx->device);
    HIPTRY(ctx, hipSetDevice(ctx->device));
    const int nk = ctx->p.nz + kext;
    const int per = (int)std::min<size_t>((size_t)nk, SCRATCH_FLOATS / plane);
    const State st = ctx->view(ctx->cur);
    // chunks of levels through the scratch; stream order keeps each copy
    // ahead of the next chunk's pack
    for (int k0 = 0; k0 < nk; k0 += per) {
        const int n = std::min(per, nk - k0);
        HIPTRY(ctx, launch_pack_levels(ctx->mixed(), ctx->g, st, which, kext, k0, n, d, ctx->stream));
        HIPTRY(ctx, hipMemcpyAsync(host + (size_t)k0 * plane, d, (size_t)n * plane * sizeof(float),
                                   hipMemcpyDeviceToHost, ctx->stream));
    }
    HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    return AWP_OK;
}

// ---------------------------------------------------------------------------
// checkpoint / restart (the reference's restart files: pmcl3d.f:1016-1097,
// 1964-2045).  File: a 72-byte header, then the memory variables'
// exponent slots (mixed), the 9 arrays of wavefield buffer 0, r1..r6, the 9
// arrays of the second wavefield buffer where the run holds one, the SGT
// sample buffer and the per-step trace -- each array's padded volume as it
// sits in HBM, array by array -- streamed through a pinned staging buffer.
// ---------------------------------------------------------------------------
namespace {

struct CkpHead {
    char magic[8];                  // "AWPCKP04"
    int32_t nx, ny, nz, storage;
    int32_t fused, cur, step, nsamp;
    int32_t npts, nsamp_max, nt, reserved;
    uint64_t state_bytes, sgt_bytes;
};
static_assert(sizeof(CkpHead) == 72, "checkpoint header is 72 bytes");

size_t sgt_buffer_bytes(const awp_ctx *ctx)
{
    return (size_t)std::max(ctx->npts, 1) * 6 * std::max(ctx->nsamp_max, 1) * sizeof(float);
}

// the state a checkpoint holds: exponent slots, wavefield buffer 0, r1..r6,
// and buffer 1 where the run holds one (each array's elements, no pads)
size_t state_payload_bytes(const awp_ctx *ctx)
{
    const size_t nv = (size_t)ctx->g.nvol;
    const size_t rb = ctx->mixed() ? 2 : 4;
    return 8 * (size_t)ctx->s.rexp_n + 9 * nv * 4 + 6 * nv * rb + (ctx->two_bufs ? 9 * nv * 4 : 0);
}

CkpHead ckp_head_of(const awp_ctx *ctx)
{
    CkpHead h;
    memset(&h, 0, sizeof h);
    memcpy(h.magic, "AWPCKP04", 8);
    h.nx = ctx->p.nx;
    h.ny = ctx->p.ny;
    h.nz = ctx->p.nz;
    h.storage = ctx->storage;
    h.fused = ctx->two_bufs ? 1 : 0;   // the buffer layout (the step in use may differ)
    h.cur = ctx->cur;
    h.step = ctx->step;
    h.nsamp = ctx->nsamp;
    h.npts = ctx->npts;
    h.nsamp_max = ctx->nsamp_max;
    h.nt = ctx->p.nt;
    h.state_bytes = state_payload_bytes(ctx);
    h.sgt_bytes = sgt_buffer_bytes(ctx);
    return h;
}

// copies `bytes` between device memory and the file in chunks
int ckp_stream(awp_ctx *ctx, FILE *f, char *dev, size_t bytes, char *stage, size_t chunk, bool save)
{
    for (size_t off = 0; off < bytes; off += chunk) {
        const size_t n = std::min(chunk, bytes - off);
        if (save) {
            HIPTRY(ctx, hipMemcpyAsync(stage, dev + off, n, hipMemcpyDeviceToHost, ctx->stream));
            HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
            if (fwrite(stage, 1, n, f) != n) return fail(ctx, AWP_EIO, "checkpoint write failed");
        } else {
            if (fread(stage, 1, n, f) != n) return fail(ctx, AWP_EIO, "checkpoint file is short");
            HIPTRY(ctx, hipMemcpyAsync(dev + off, stage, n, hipMemcpyHostToDevice, ctx->stream));
            HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
        }
    }
    return AWP_OK;
}

int ckp_run(awp_ctx *ctx, FILE *f, bool save)
{
    const size_t chunk = (size_t)64 << 20;
    char *stage = nullptr;
    HIPTRY(ctx, hipHostMalloc((void **)&stage, chunk, hipHostMallocDefault));
    const size_t nv = (size_t)ctx->g.nvol;
    int rc = ckp_stream(ctx, f, (char *)ctx->s.rexp, 8 * (size_t)ctx->s.rexp_n, stage, chunk, save);
    for (int a = 0; a < 9 && !rc; ++a) rc = ckp_stream(ctx, f, (char *)ctx->wf[0][a], nv * 4, stage, chunk, save);
    void *rr[6] = {ctx->s.r1, ctx->s.r2, ctx->s.r3, ctx->s.r4, ctx->s.r5, ctx->s.r6};
    for (int a = 0; a < 6 && !rc; ++a)
        rc = ckp_stream(ctx, f, (char *)rr[a], nv * (ctx->mixed() ? 2 : 4), stage, chunk, save);
    for (int a = 0; a < 9 && !rc && ctx->two_bufs; ++a)
        rc = ckp_stream(ctx, f, (char *)ctx->wf[1][a], nv * 4, stage, chunk, save);
    if (!rc) rc = ckp_stream(ctx, f, (char *)ctx->d_sgt, sgt_buffer_bytes(ctx), stage, chunk, save);
    if (!rc)
        rc = ckp_stream(ctx, f, (char *)ctx->d_trace, (size_t)std::max(ctx->p.nt, 1) * 3 * sizeof(float), stage,
                        chunk, save);
    hipHostFree(stage);
    return rc;
}

}  // namespace

extern "C" int awp_checkpoint_save(awp_ctx *ctx, const char *path)
{
    if (!ctx || !path) return fail(ctx, AWP_EINVAL, "awp_checkpoint_save: null argument");
    if (!ctx->initialised) return fail(ctx, AWP_ESTATE, "awp_checkpoint_save before awp_init");
    HIPTRY(ctx, hipSetDevice(ctx->device));
    HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return fail(ctx, AWP_EIO, "cannot create '%s'", tmp.c_str());
    const CkpHead h = ckp_head_of(ctx);
    int rc = fwrite(&h, sizeof h, 1, f) == 1 ? AWP_OK : fail(ctx, AWP_EIO, "checkpoint write failed");
    if (!rc) rc = ckp_run(ctx, f, true);
    if (fclose(f) != 0 && !rc) rc = fail(ctx, AWP_EIO, "checkpoint write failed");
    if (!rc && rename(tmp.c_str(), path) != 0) rc = fail(ctx, AWP_EIO, "cannot rename '%s' to '%s'", tmp.c_str(), path);
    if (rc) remove(tmp.c_str());
    return rc;
}

extern "C" int awp_checkpoint_load(awp_ctx *ctx, const char *path)
{
    if (!ctx || !path) return fail(ctx, AWP_EINVAL, "awp_checkpoint_load: null argument");
    if (!ctx->initialised) return fail(ctx, AWP_ESTATE, "awp_checkpoint_load needs an initialised run (awp_init)");
    HIPTRY(ctx, hipSetDevice(ctx->device));
    FILE *f = fopen(path, "rb");
    if (!f) return fail(ctx, AWP_EIO, "cannot open '%s'", path);
    CkpHead h;
    const CkpHead want = ckp_head_of(ctx);
    int rc = AWP_OK;
    if (fread(&h, sizeof h, 1, f) != 1 || memcmp(h.magic, want.magic, 8)) {
        rc = fail(ctx, AWP_EINVAL, "'%s' is not a checkpoint file", path);
    } else if (h.nx != want.nx || h.ny != want.ny || h.nz != want.nz || h.storage != want.storage ||
               h.fused != want.fused || h.npts != want.npts || h.nsamp_max != want.nsamp_max || h.nt != want.nt ||
               h.state_bytes != want.state_bytes || h.sgt_bytes != want.sgt_bytes) {
        rc = fail(ctx, AWP_EINVAL,
                  "checkpoint '%s' is of another run (grid, storage, wavefield buffers, points or steps differ)", path);
    } else if (h.step < 0 || h.step > h.nt || h.nsamp < 0 || h.nsamp > h.nsamp_max || (h.cur != 0 && h.cur != 1) ||
               (h.cur == 1 && !ctx->two_bufs)) {
        rc = fail(ctx, AWP_EINVAL, "checkpoint '%s' has an invalid step count", path);
    }
    if (!rc) {
        rc = ckp_run(ctx, f, false);
        // a partly restored state is no state: the run needs awp_init again
        if (rc) ctx->initialised = false;
    }
    fclose(f);
    if (rc) return rc;
    ctx->step = h.step;
    ctx->nsamp = h.nsamp;
    ctx->cur = h.cur;
    return scheme_changed(ctx);
}
