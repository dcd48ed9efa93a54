// awp_batch.hip -- several SGT runs (components, sites) advanced together on
// one GPU.  Each stage of a time step is one persistent launch over the
// tiles of every run (a grid walking the concatenated tile list; the fused
// kernel's blocks claim its tiles per XCD, claim_tile): in the
// two-stage scheme the velocity kernel, the free-surface images and the
// stress kernel, in the fused scheme the one-sweep kernel; then one launch
// for every run's source and SGT sample.  Two steps (both buffer parities of
// the fused scheme) are captured in a HIP graph once and replayed.  The runs
// keep their own contexts (buffers, counters, outputs): the batch only
// drives them, so awp_get_sgt / awp_write_sgt_file work on each run
// afterwards.
//
// A batch uses the fused scheme when every run was created with it (two
// wavefield buffers), else the two-stage one.  The tuner times the step
// kernels of the batch for each candidate scheme (the fused one only where
// the runs hold two buffers), tile height, prefetch depth or z unroll and z
// chunking on the runs' own buffers and keeps the fastest; it needs runs
// that have not stepped yet and re-initialises them afterwards.
//
// The reference runs one SGT component per MPI job (awp_odc_wrapper.sh:14);
// there is no counterpart of a batch in it.

#include <hip/hip_runtime.h>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "awp_ctx.hpp"
#include "awp_kernels.hpp"
#include "awpsgt.h"

using namespace awp;

struct awp_batch {
    std::vector<awp_ctx *> runs;
    int device;
    bool mixed;
    bool split;                  // two-stage scheme (else fused)
    bool can_fuse;               // every run holds the fused scheme's two buffers
    bool persist;                // persistent grids (else one block per tile)
    bool o32;                    // 32-bit byte offsets reach every run's arrays
    Launch l;                    // tile_y and prefetch depth / z unroll (unroll) of the batch
    std::vector<int> nchunk;     // fused scheme: z chunks per run
    int kc_split;                // two-stage scheme: z chunk length (levels)
    hipStream_t stream;
    FJob *d_jobs;
    SJob *d_sjobs;
    AuxJob *d_aux;
    int *d_counter;
    unsigned long long *d_claim;   // the fused kernel's tile claim counters (CLAIM_WORDS)
    int total, fs_total, max_kept;
    int grid, grid_v, grid_s;    // persistent blocks: fused kernel; velocity, stress kernels
    hipGraphExec_t graph;
    bool graph_ok;
    bool use_graph;
    hipEvent_t e0, e1;
    double t_last;
    char err[512];
};

static int bfail(awp_batch *b, int code, const char *fmt, ...)
{
    if (b) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(b->err, sizeof(b->err), fmt, ap);
        va_end(ap);
    }
    return code;
}

#define BTRY(b, call)                                                                             \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return bfail((b), AWP_EHIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),    \
                         __FILE__, __LINE__);                                                     \
    } while (0)

extern "C" const char *awp_batch_last_error(const awp_batch *b) { return b ? b->err : "null batch"; }

// z chunks of a run with `nchunk` requested (chunks of >= 16 levels fused,
// >= 8 two-stage): the top chunk keeps >= 3 levels (the free-surface stress
// images need the two levels below the surface)
static Launch run_launch(const Geo &g, const Launch &bl, int nchunk, bool split)
{
    Launch l = bl;
    nchunk = std::max(1, std::min(nchunk, std::max(1, g.nz / (split ? 8 : 16))));
    int lo = 1;
    if (!split) {
        // the fused step's 32-bit offsets span one chunk
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
    return l;
}

// Builds the device job tables from the runs' current buffer parity: in the
// fused scheme run r reads s[0] = its current state and writes s[1]; the
// two-stage scheme updates the current state in place.
static int upload_tables(awp_batch *b)
{
    const int n = (int)b->runs.size();
    std::vector<FJob> fj(n);
    std::vector<SJob> sj(n);
    std::vector<AuxJob> aj(n);
    int tile0 = 0, fs0 = 0;
    b->max_kept = 0;
    for (int r = 0; r < n; ++r) {
        awp_ctx *c = b->runs[r];
        const Launch l = run_launch(c->g, b->l, b->split ? split_nchunk(c->g.nz, b->kc_split) : b->nchunk[r],
                                    b->split);
        AuxJob &x = aj[r];
        if (b->split) {
            SJob &q = sj[r];
            q = split_job(c->g, c->view(c->cur), c->c, l, tile0, fs0);
            tile0 += split_tiles(q);
            fs0 += split_fs_blocks(q);
            x.s[0] = x.s[1] = q.s;
        } else {
            FJob &f = fj[r];
            f.g = c->g;
            f.s[0] = c->view(c->cur);
            f.s[1] = c->view(c->cur ^ 1);
            f.c = c->c;
            f.kc = l.kc;
            f.t = fused_tiles(c->g, l);
            f.tile0 = tile0;
            tile0 += f.t.nbx * f.t.nby * f.t.nbz;
            x.s[0] = f.s[0];
            x.s[1] = f.s[1];
        }
        x.g = c->g;
        x.igreen = c->p.igreen;
        x.npsrc = c->npsrc;
        x.nst = c->p.nst;
        x.ntiskp = c->p.ntiskp_sgt;
        x.nsamp_max = c->nsamp_max;
        x.nkept = c->nkept;
        x.psrc = c->d_psrc;
        x.ptop = c->d_ptop;
        x.stf = c->d_stf;
        x.vtst = c->vtst;
        x.vtst1 = c->vtst1;
        x.kpos = c->d_kpos;
        x.kidx = c->d_kidx;
        x.sg = c->d_sg;
        x.sgt = c->d_sgt;
        x.counter = b->d_counter + r;
        x.trace = c->d_trace;
        x.tpos = c->trace_pos;
        b->max_kept = std::max(b->max_kept, c->nkept);
    }
    b->total = tile0;
    b->fs_total = fs0;
    if (b->split)
        BTRY(b, hipMemcpyAsync(b->d_sjobs, sj.data(), n * sizeof(SJob), hipMemcpyHostToDevice, b->stream));
    else
        BTRY(b, hipMemcpyAsync(b->d_jobs, fj.data(), n * sizeof(FJob), hipMemcpyHostToDevice, b->stream));
    BTRY(b, hipMemcpyAsync(b->d_aux, aj.data(), n * sizeof(AuxJob), hipMemcpyHostToDevice, b->stream));
    BTRY(b, hipStreamSynchronize(b->stream));   // fj / sj / aj are freed on return
    return AWP_OK;
}

static int set_counters(awp_batch *b)
{
    const int n = (int)b->runs.size();
    std::vector<int> cnt(n);
    for (int r = 0; r < n; ++r) cnt[r] = b->runs[r]->step;
    BTRY(b, hipMemcpyAsync(b->d_counter, cnt.data(), n * sizeof(int), hipMemcpyHostToDevice, b->stream));
    // the host vector must outlive the copy
    BTRY(b, hipStreamSynchronize(b->stream));
    return AWP_OK;
}

// the step kernels of every run (no source / SGT work)
static hipError_t launch_kernels(awp_batch *b, int par)
{
    const int n = (int)b->runs.size();
    if (b->split)
        return launch_step_split_batch(b->mixed, b->l, b->o32, b->d_sjobs, n, b->total, b->fs_total,
                                       b->grid_v, b->grid_s, b->persist && b->l.claim != 0 ? b->d_claim : nullptr,
                                       b->stream);
    return launch_step_fused_batch(b->mixed, b->l, b->d_jobs, n, b->total, b->grid, par, b->d_claim, b->stream);
}

// grid sizes of the current scheme and launch shape: the blocks that fit on
// the device at once (persistent), or one block per tile (the launchers
// clamp the grid to the tile count)
static void set_grids(awp_batch *b)
{
    if (!b->persist) {
        b->grid = b->grid_v = b->grid_s = 1 << 30;
        return;
    }
    if (b->split)
        split_batch_grids(b->mixed, b->l, b->device, &b->grid_v, &b->grid_s);
    else
        b->grid = fused_batch_grid(b->mixed, b->l, b->device);
}

// one time step of every run at relative parity `par` (fused scheme)
static hipError_t launch_one(awp_batch *b, int par)
{
    hipError_t e = launch_kernels(b, par);
    if (e != hipSuccess) return e;
    return launch_aux_batch(b->d_aux, (int)b->runs.size(), b->max_kept, b->split ? 0 : par, b->stream);
}

static void drop_graph(awp_batch *b)
{
    if (b->graph_ok) hipGraphExecDestroy(b->graph);
    b->graph_ok = false;
}

static int build_graph(awp_batch *b)
{
    drop_graph(b);
    hipGraph_t g = nullptr;
    BTRY(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = launch_one(b, 0);
    if (e == hipSuccess) e = launch_one(b, 1);
    hipError_t e2 = hipStreamEndCapture(b->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return bfail(b, AWP_EHIP, "graph capture: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    }
    e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return bfail(b, AWP_EHIP, "graph instantiate: %s", hipGetErrorString(e));
    b->graph_ok = true;
    return AWP_OK;
}

extern "C" int awp_batch_create(awp_batch **out, awp_ctx *const *runs, int nruns, char *err, size_t errlen)
{
    if (!out || !runs || nruns < 1) {
        copy_err(err, errlen, "awp_batch_create: no runs");
        return AWP_EINVAL;
    }
    *out = nullptr;
    for (int r = 0; r < nruns; ++r) {
        const awp_ctx *c = runs[r];
        char msg[160];
        if (!c) {
            snprintf(msg, sizeof msg, "run %d is null", r);
        } else if (!c->initialised) {
            snprintf(msg, sizeof msg, "run %d is not initialised (awp_init)", r);
        } else if (c->device != runs[0]->device || c->storage != runs[0]->storage) {
            snprintf(msg, sizeof msg, "run %d differs from run 0 in device or storage mode", r);
        } else {
            bool dup = false;
            for (int q = 0; q < r; ++q) dup = dup || runs[q] == c;
            if (!dup) continue;
            snprintf(msg, sizeof msg, "run %d appears twice", r);
        }
        copy_err(err, errlen, msg);
        return AWP_EINVAL;
    }
    awp_batch *b = new awp_batch();
    b->runs.assign(runs, runs + nruns);
    b->device = runs[0]->device;
    b->mixed = runs[0]->mixed();
    // the fused scheme when every run holds both buffers and steps fused
    b->can_fuse = true;
    bool all_fused = true;
    for (int r = 0; r < nruns; ++r) {
        b->can_fuse = b->can_fuse && runs[r]->two_bufs;
        all_fused = all_fused && runs[r]->fused;
    }
    b->split = !(b->can_fuse && all_fused);
    // AWP_BATCH_SCHEME=fused / split overrides (fused only where every run
    // holds both buffers; each run's own step is unchanged)
    const char *bs = getenv("AWP_BATCH_SCHEME");
    if (bs && !strcmp(bs, "fused") && b->can_fuse) b->split = false;
    if (bs && !strcmp(bs, "split")) b->split = true;
    b->persist = true;
    b->o32 = true;
    for (int r = 0; r < nruns; ++r) b->o32 = b->o32 && runs[r]->l.o32;
    b->l = runs[0]->l;
    if (b->split) {
        // two-stage kernels: {64, 128} x {1, 2, 4, 8} tiles, z unroll 1 or 2;
        // blocks of more than 512 threads (64 x 16, 128 x 8) in the
        // byte-offset stress form without z unroll
        if (b->l.unroll != 2) b->l.unroll = 1;
        if (b->l.tile_xw != 2) b->l.tile_xw = 1;
        const bool tall_ok = b->o32 && b->l.o32_s && b->l.unroll == 1;
        if (b->l.tile_y != 1 && b->l.tile_y != 2 && b->l.tile_y != 4 && b->l.tile_y != 16) b->l.tile_y = 8;
        if (b->l.tile_y * b->l.tile_xw > 16 || (b->l.tile_y * b->l.tile_xw > 8 && !tall_ok)) {
            b->l.tile_y = 8;
            b->l.tile_xw = 1;
        }
    } else {
        b->l.tile_y = (b->l.tile_y == 4 || b->l.tile_y == 7 || b->l.tile_y == 8 || b->l.tile_y == 11) ? b->l.tile_y
                                                                                                     : 8;
        // 64 x 8: no prefetch; 64 x 11: none, with one or two levels per barrier
        if (b->l.tile_y == 8 || (b->l.tile_y == 11 && b->l.unroll != 3)) b->l.unroll = 0;
        b->l.tile_xw = 1;
    }
    b->nchunk.assign(nruns, 1);
    b->kc_split = SPLIT_KC;
    b->d_jobs = nullptr;
    b->d_sjobs = nullptr;
    b->d_aux = nullptr;
    b->d_counter = nullptr;
    b->d_claim = nullptr;
    b->graph_ok = false;
    const char *g_env = getenv("AWP_GRAPH");
    b->use_graph = !(g_env && !strcmp(g_env, "0"));
    b->t_last = 0;
    b->err[0] = 0;
    hipError_t e = hipSetDevice(b->device);
    if (e == hipSuccess) e = hipStreamCreate(&b->stream);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_jobs, nruns * sizeof(FJob));
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_sjobs, nruns * sizeof(SJob));
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_aux, nruns * sizeof(AuxJob));
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_counter, nruns * sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_claim, awp::CLAIM_WORDS * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(b->d_claim, 0, awp::CLAIM_WORDS * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipEventCreate(&b->e0);
    if (e == hipSuccess) e = hipEventCreate(&b->e1);
    if (e != hipSuccess) {
        snprintf(b->err, sizeof b->err, "batch set-up: %s", hipGetErrorString(e));
        copy_err(err, errlen, b->err);
        awp_batch_destroy(b);
        return AWP_EHIP;
    }
    // the runs' own streams may still hold work (awp_init); the batch
    // stream starts after it
    for (awp_ctx *c : b->runs) hipStreamSynchronize(c->stream);
    set_grids(b);
    *out = b;
    return AWP_OK;
}

extern "C" int awp_batch_destroy(awp_batch *b)
{
    if (!b) return AWP_EINVAL;
    hipSetDevice(b->device);
    drop_graph(b);
    if (b->stream) hipStreamSynchronize(b->stream);
    hipFree(b->d_jobs);
    hipFree(b->d_sjobs);
    hipFree(b->d_aux);
    hipFree(b->d_counter);
    hipFree(b->d_claim);
    hipEventDestroy(b->e0);
    hipEventDestroy(b->e1);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
    return AWP_OK;
}

extern "C" int awp_batch_step(awp_batch *b, int nsteps)
{
    if (!b) return AWP_EINVAL;
    if (nsteps < 0) return bfail(b, AWP_EINVAL, "nsteps < 0");
    BTRY(b, hipSetDevice(b->device));
    int todo = nsteps;
    for (awp_ctx *c : b->runs) {
        if (!c->initialised) return bfail(b, AWP_ESTATE, "a run was re-created or not initialised");
        todo = std::min(todo, c->p.nt - c->step);
    }
    b->t_last = 0;
    if (todo <= 0) return AWP_OK;
    for (awp_ctx *c : b->runs) BTRY(b, hipStreamSynchronize(c->stream));
    int rc = upload_tables(b);
    if (rc) return rc;
    rc = set_counters(b);
    if (rc) return rc;
    if (b->use_graph && todo >= 2 && !b->graph_ok) {
        rc = build_graph(b);
        if (rc) return rc;
    }
    BTRY(b, hipEventRecord(b->e0, b->stream));
    int done = 0;
    if (b->use_graph) {
        for (; done + 2 <= todo; done += 2) BTRY(b, hipGraphLaunch(b->graph, b->stream));
    }
    for (int par = 0; done < todo; ++done, par ^= 1) BTRY(b, launch_one(b, par));
    BTRY(b, hipEventRecord(b->e1, b->stream));
    BTRY(b, hipEventSynchronize(b->e1));
    float ms = 0;
    BTRY(b, hipEventElapsedTime(&ms, b->e0, b->e1));
    b->t_last = ms;
    for (awp_ctx *c : b->runs) {
        c->step += todo;
        c->nsamp = std::min(c->nsamp_max, c->step / c->p.ntiskp_sgt);
        if (!b->split) c->cur ^= (todo & 1);
    }
    return AWP_OK;
}

extern "C" int awp_batch_synchronize(awp_batch *b)
{
    if (!b) return AWP_EINVAL;
    BTRY(b, hipSetDevice(b->device));
    BTRY(b, hipStreamSynchronize(b->stream));
    return AWP_OK;
}

extern "C" int awp_batch_time(const awp_batch *b, double *ms)
{
    if (!b || !ms) return AWP_EINVAL;
    *ms = b->t_last;
    return AWP_OK;
}

extern "C" int awp_batch_get_scheme(const awp_batch *b, int *fused, int *persistent)
{
    if (!b) return AWP_EINVAL;
    if (fused) *fused = b->split ? 0 : 1;
    if (persistent) *persistent = b->persist ? 1 : 0;
    return AWP_OK;
}

extern "C" int awp_batch_get_launch(const awp_batch *b, int *tile_y, int *prefetch, int *nchunk_run0)
{
    if (!b) return AWP_EINVAL;
    if (tile_y) *tile_y = b->l.tile_y;
    if (prefetch) *prefetch = b->l.unroll;
    if (nchunk_run0)
        *nchunk_run0 = b->split ? run_launch(b->runs[0]->g, b->l, split_nchunk(b->runs[0]->g.nz, b->kc_split), true).nchunk
                                : b->nchunk[0];
    return AWP_OK;
}

extern "C" int awp_batch_tune(awp_batch *b, int reps)
{
    if (!b) return AWP_EINVAL;
    for (awp_ctx *c : b->runs)
        if (c->step != 0) return bfail(b, AWP_ESTATE, "tuning needs runs that have not stepped");
    if (reps < 1) reps = 2;
    BTRY(b, hipSetDevice(b->device));
    struct Cand {
        bool split, persist;
        int ty, pf, nchunk;   // nchunk: z chunks per run (fused), chunk length in levels (two-stage)
        int xw;      // two-stage: waves per tile row
        int sform;   // two-stage stress kernel: 0 = 64-bit offsets, blocks as registers allow;
                     // 1 = 32-bit offsets, 2 blocks per CU
        bool patch;  // two-stage: the runs' tile patches (else rows of tiles)
    };
    std::vector<Cand> cands;
    for (bool ps : {true, false}) {
        // two-stage: tile rows x z unroll x z chunk length (levels) x stress
        // kernel form x tile order
        for (int ty : {4, 8})
            for (int un : {1, 2})
                for (int kc : {SPLIT_KC, 2 * SPLIT_KC})
                    for (int sf : {0, 1})
                        for (bool pt : {true, false}) cands.push_back({true, ps, ty, un, kc, 1, sf, pt});
        // 128 x 4 tiles; 64 x 16 and 128 x 8 tiles (1024-thread blocks:
        // byte-offset stress form, no z unroll)
        for (int kc : {SPLIT_KC, 2 * SPLIT_KC})
            for (bool pt : {true, false}) {
                for (int sf : {0, 1}) cands.push_back({true, ps, 4, 1, kc, 2, sf, pt});
                if (b->o32) {
                    cands.push_back({true, ps, 16, 1, kc, 1, 1, pt});
                    cands.push_back({true, ps, 8, 1, kc, 2, 1, pt});
                }
            }
        // fused: tile rows x prefetch depth (64x8 and 64x11 tiles only
        // without prefetch: their 13- and 16-wave blocks have 128 registers
        // per lane) x z chunks
        const int tp[8][2] = {{11, 3}, {11, 0}, {8, 0}, {7, 0}, {7, 1}, {7, 2}, {4, 0}, {4, 2}};
        if (b->can_fuse)
            for (const auto &v : tp)
                for (int nc : {2, 4}) cands.push_back({false, ps, v[0], v[1], nc, 1, 0, true});
    }
    const int spw0 = b->l.spw, sph0 = b->l.sph;
    auto apply_form = [b, spw0, sph0](int sf, bool patch) {
        b->l.o32_s = sf == 1;
        b->l.bpc_s = sf == 1 ? 2 : 0;
        b->l.spw = patch ? spw0 : 0;
        b->l.sph = patch ? sph0 : 0;
        set_block_caps(b->l, b->device);
    };
    // times the step kernels of candidate cd over n launches after one
    // warm-up (ms in *out)
    auto time_cand = [&](const Cand &cd, int n, double *out) -> int {
        b->split = cd.split;
        b->persist = cd.persist;
        apply_form(cd.sform, cd.patch);
        b->l.tile_y = cd.ty;
        b->l.tile_xw = cd.xw;
        b->l.unroll = cd.pf;
        if (cd.split)
            b->kc_split = cd.nchunk;
        else
            std::fill(b->nchunk.begin(), b->nchunk.end(), cd.nchunk);
        int rc = upload_tables(b);
        if (rc) return rc;
        set_grids(b);
        BTRY(b, launch_kernels(b, 0));
        BTRY(b, hipEventRecord(b->e0, b->stream));
        for (int k = 0; k < n; ++k) BTRY(b, launch_kernels(b, k & 1));
        BTRY(b, hipEventRecord(b->e1, b->stream));
        BTRY(b, hipEventSynchronize(b->e1));
        float ms = 0;
        BTRY(b, hipEventElapsedTime(&ms, b->e0, b->e1));
        *out = ms / n;
        return AWP_OK;
    };
    // every candidate once, then the three fastest again over twice the
    // launches (one pass over ~100 candidates leaves picks within timing
    // noise of each other)
    // AWP_TUNE_LOG=1: every timing on stderr
    const char *tl = getenv("AWP_TUNE_LOG");
    const bool log = tl && atoi(tl) == 1;
    auto show = [&](const char *pass, const Cand &cd, double ms) {
        if (log)
            fprintf(stderr, "awp_batch_tune %s: %s %s ty %d xw %d pf/unroll %d chunks %d form %d patch %d: %.3f ms\n",
                    pass, cd.split ? "two-stage" : "fused", cd.persist ? "persistent" : "per-tile", cd.ty, cd.xw,
                    cd.pf, cd.nchunk, cd.sform, cd.patch ? 1 : 0, ms);
    };
    std::vector<std::pair<double, size_t>> first;
    for (size_t n = 0; n < cands.size(); ++n) {
        double ms = 0;
        int rc = time_cand(cands[n], reps, &ms);
        if (rc) return rc;
        show("first", cands[n], ms);
        first.push_back({ms, n});
    }
    std::sort(first.begin(), first.end());
    double best = 1e30;
    Cand pick = cands[first[0].second];
    for (size_t q = 0; q < std::min<size_t>(3, first.size()); ++q) {
        double ms = 0;
        int rc = time_cand(cands[first[q].second], 2 * reps, &ms);
        if (rc) return rc;
        show("again", cands[first[q].second], ms);
        if (ms < best) {
            best = ms;
            pick = cands[first[q].second];
        }
    }
    b->split = pick.split;
    b->persist = pick.persist;
    apply_form(pick.sform, pick.patch);
    b->l.tile_y = pick.ty;
    b->l.tile_xw = pick.xw;
    b->l.unroll = pick.pf;
    if (pick.split)
        b->kc_split = pick.nchunk;
    else
        std::fill(b->nchunk.begin(), b->nchunk.end(), pick.nchunk);
    set_grids(b);
    drop_graph(b);
    // the trial steps changed the runs' state: start them again
    for (awp_ctx *c : b->runs) {
        int rc = awp_init(c);
        if (rc) return bfail(b, rc, "re-initialising a run after tuning: %s", awp_last_error(c));
    }
    b->t_last = best;
    return AWP_OK;
}
