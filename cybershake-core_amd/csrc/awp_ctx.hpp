// awp_ctx.hpp -- the run context behind the C ABI (one SGT component run on
// one GPU): its device buffers, stream, step counters and error text.  Shared
// by awp_ctx.hip (single runs) and awp_batch.hip (runs advanced together).
#pragma once

#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "awp_kernels.hpp"
#include "awpsgt.h"

using awp::Coef;
using awp::Geo;
using awp::Launch;
using awp::Media32;
using awp::State;

struct awp_ctx {
    awp_in3d p;
    int device;
    int storage;
    Geo g;
    State s;
    Coef c;
    Launch l;
    hipStream_t stream;
    float *vol;              // one allocation for the volume arrays (media, exponents, state)
    // wavefield buffers (u1 v1 w1 xx yy zz xy xz yz): the fused step reads
    // wf[cur] and writes wf[cur ^ 1]; the two-stage step uses wf[0] only
    float *wf[2][9];
    int cur;
    bool fused;              // the step in use: fused one-sweep (true) or two-stage
    bool two_bufs;           // the allocation holds wf[1] (the fused step can run)
    // the two-stage step's job table (one entry: this run), uploaded by
    // awp_init, and its tile and free-surface block counts
    awp::SJob *d_sjob;
    // the claimed fused step's one-entry job table (both buffers), its
    // persistent grid and tile claim counters (CLAIM_WORDS, zero between
    // launches)
    awp::FJob *d_fjob;
    int fgrid, ftotal;
    unsigned long long *d_claim;
    int split_total, split_fs;
    int sgrid_v, sgrid_s;    // the two-stage kernels' persistent grids (claimed launches)
    float *prof;             // fx | fy | fz
    // the per-step job-value trace (CHKJ, pmcl3d.f:1476-1479): u1 v1 w1 at
    // node trace_ijk after every step, [nt][3]
    float *d_trace;
    long long trace_pos;
    int trace_ijk[3];
    // sources
    int npsrc;
    long long *d_psrc;       // [npsrc] node index
    long long *d_ptop;       // [npsrc] index of (x, y, nz)
    float *d_stf;            // [npsrc][nst][6]
    float vtst, vtst1;
    // SGT
    int npts, nkept, nsamp_max;
    long long *d_kpos;
    int *d_kidx;
    float *d_sg;
    float *d_sgt;            // [npts][6][nsamp_max]
    std::vector<int> keep;   // per cordfile point
    // progress
    bool media_set, source_set, points_set, initialised;
    int step, nsamp;
    // timing
    bool timing;
    std::vector<hipEvent_t> ev;  // pairs per timed launch: vel, stress, step
    double t_vel, t_str, t_all;
    size_t bytes;
    size_t vol_bytes;        // the volume allocation
    size_t zero_begin, zero_end;   // in vol: exponents and state, zeroed by awp_init
    bool sg_done;            // SGT coefficients computed from FP32 media
    char err[512];

    bool mixed() const { return storage == AWP_STORAGE_MIXED; }
    // the state with wavefield buffer `which` (r1..r6, media, profiles shared)
    State view(int which) const
    {
        State x = s;
        float **f[9] = {&x.u1, &x.v1, &x.w1, &x.xx, &x.yy, &x.zz, &x.xy, &x.xz, &x.yz};
        for (int a = 0; a < 9; ++a) *f[a] = wf[which][a];
        return x;
    }
    // the media arrays of an FP32-storage run
    Media32 media32() const
    {
        Media32 m = {(float *)s.d1, (float *)s.mu, (float *)s.lam, (float *)s.qp, (float *)s.qs};
        return m;
    }
};

inline int fail(awp_ctx *ctx, int code, const char *fmt, ...)
{
    if (ctx) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(ctx->err, sizeof(ctx->err), fmt, ap);
        va_end(ap);
    }
    return code;
}

#define HIPTRY(ctx, call)                                                                       \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail((ctx), AWP_EHIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                    \
    } while (0)

inline void copy_err(char *dst, size_t n, const char *src)
{
    if (dst && n) {
        strncpy(dst, src, n - 1);
        dst[n - 1] = 0;
    }
}

