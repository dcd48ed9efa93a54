/*
 * awpsgt.h -- C ABI of the MI355X AWP-ODC-SGT strain-Green-tensor solver.
 *
 * One context is one SGT component run (one IN3D): it owns its device
 * buffers and its HIP stream.  All functions return 0 on success and a
 * negative AWP_E* code on failure; the text of the last error of a context is
 * available from awp_last_error().  Plain pointers and sizes only.
 *
 * The reference's host<->device seam is the set of extern "C" launch wrappers
 * the AWP-GPU-SGT C host calls (AWP-GPU-SGT/src/kernel_kp1_fix_havg_ifault5.cu):
 * SetDeviceConstValue (:24), dvelcx_H (:73), dstrqc_H (:124), addsrc_H (:145),
 * ComputeSGT (:380), over buffers the host allocates (pmcl3d.c:537-823).  This
 * ABI replaces that seam with a context (no __constant__ globals, so several
 * runs can live in one process) and keeps the CPU solver's semantics
 * (AWP-ODC-SGT/src/pmcl3d.f:1183-1557).  Each entry point names the reference
 * code it replaces.
 */
#ifndef AWPSGT_H
#define AWPSGT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AWPSGT_ABI_VERSION 1

enum {
    AWP_OK = 0,
    AWP_EINVAL = -1,   /* invalid argument or unsupported IN3D setting */
    AWP_EHIP = -2,     /* HIP runtime error */
    AWP_ENOMEM = -3,   /* allocation failed */
    AWP_EIO = -4,      /* file I/O error */
    AWP_ESTATE = -5    /* call out of order (e.g. step before media) */
};

/* Storage modes of the state and media arrays (arithmetic is FP32 in all). */
enum {
    AWP_STORAGE_FP32 = 0,   /* everything FP32: the reference's storage */
    AWP_STORAGE_MIXED = 1   /* r1..r6 FP16 with per-row exponents, qp/qs FP16 (scaled), d1/lam/mu FP16
                               deviations from an FP32 profile per level; qp with qs and mu with lam
                               share one 4-byte word a point */
};

/* The IN3D parameters the SGT path uses, in the reference's units.  Filled by
 * awp_in3d_read() (io.f:327-418) or by the caller. */
typedef struct {
    int igreen;          /* 1,2,3 body force x/y/z; 4,5,6 surface xz/yz/zz */
    float tmax, dh, dt;  /* TMAX (s), DH (m), DT (s) */
    int npc;             /* must be 0 (Cerjan) */
    int nd;              /* ND sponge width (cells) */
    float arbc;          /* ARBC */
    float pht;
    int nsrc, nst;       /* NSRC, NST */
    int nx, ny, nz;      /* grid */
    int npx, npy, npz;   /* reference MPI decomposition (SGT point rule only) */
    int ifault;          /* must be 0 (text source) */
    int checkpoint, isfcvlm, imd5, ivelocity;
    int mediarestart;    /* must be 1 (media file) */
    int nvar;            /* 3, 5 or 8 */
    int iost, partdeg, io_opt, perf_meas, idyna;
    int socalq;          /* SOCALQ */
    int nve;             /* must be 1 (visco-elastic) */
    float mu_s, mu_d, fl, fh, fp;
    int read_step, write_step, write_step2;
    int nbgx, nedx, nskpx, nbgy, nedy, nskpy, nbgz, nedz, nskpz, ntiskp;
    int nbgx2, nedx2, nskpx2, nbgy2, nedy2, nskpy2, nbgz2, nedz2, nskpz2, ntiskp2;
    int ntiskp_sgt;      /* NTISKP_SGT */
    char chkp[256], chkj[256], insrc[256], invel[256], insgt[256];
    char sxgro[256], sygro[256], szgro[256], sxgro2[256], sygro2[256], szgro2[256];
    char sgtgro[256], sgsn[256];
    int nt;              /* derived: nint(TMAX/DT)+1 (pmcl3d.f:287) */
} awp_in3d;

typedef struct awp_ctx awp_ctx;

/* ---------------------------------------------------------------------- */
/* Host-side formats (C, no device work)                                   */
/* ---------------------------------------------------------------------- */

/* Parse an IN3D file: 72 positional list-directed records (io.f:327-418,
 * set_names.f:29-47 for the file-name overrides, which the caller applies).
 * Computes nt.  Replaces rddt3d (io.f:231). */
int awp_in3d_read(const char *path, awp_in3d *out, char *err, size_t errlen);

/* Check the IN3D settings against what the SGT path supports (igreen 1..6,
 * NPC=0, NVE=1, IFAULT=0, MEDIARESTART=1, NVAR 3/5/8, sponge fits, ...).
 * Unsupported values are errors, never silently ignored. */
int awp_in3d_validate(const awp_in3d *p, char *err, size_t errlen);

/* Read an IFAULT=0 text source (source.f:170-198): per source an "x y z"
 * line then NST lines of axx ayy azz axz ayz axy.  xyz [nsrc][3] (file
 * coordinates, z=1 at the surface), stf [nsrc][nst][6]. */
int awp_source_read(const char *path, int nsrc, int nst, int *xyz, float *stf,
                    char *err, size_t errlen);

/* Count of points in a cordfile (io.f:419-442: leading '#' lines skipped). */
int awp_cordfile_count(const char *path, int *npts, char *err, size_t errlen);
/* Read the points, [npts][3] (x, y, z; 1-based, z=1 at the surface). */
int awp_cordfile_read(const char *path, int npts, int *xyz, char *err, size_t errlen);

/* SGT point selection (station.f:31-80), as the reference evaluates it on its
 * NPX x NPY x NPZ ranks.  For each cordfile point: keep[n] = 1 and
 * ijk[n] = global node (i, j, k; k = nz is the surface) when some rank owns
 * it, else keep[n] = 0 (the reference silently drops it).  Points with a
 * negative x or y, or z > nz+1, make the reference index outside its arrays;
 * they are rejected with AWP_EINVAL.  Returns the number kept (>= 0). */
int awp_sgt_select(const awp_in3d *p, int npts, const int *xyz, int *keep, int *ijk,
                   char *err, size_t errlen);

/* Geometry of the raw SGT file (pmcl3d.f:1582-1601): NT_OUT samples per
 * (point, component) slot and the number of WRITE_STEP buffers flushed. */
int awp_sgt_file_geometry(const awp_in3d *p, int *nt_out, int *nflush);

/* ---- acceptance checks of a raw SGT file (host only) --------------------- */

/* NaN / zero-run scan of a float32 file (SgtTest/src/check_for_nans.c:19-56,
 * which a CyberShake SGT must pass): status 0 clean, 2 a NaN at byte
 * `offset`, 3 `zero_run` consecutive values with |v| < 1e-20 ending at byte
 * `offset` (the reference uses 24000); nvalues = floats read. */
typedef struct {
    int status;
    long long offset;
    long long nvalues;
} awp_sgt_scan_result;
int awp_sgt_scan(const char *path, long long zero_run, awp_sgt_scan_result *out, char *err, size_t errlen);

/* MD5 (RFC 1321) of a file as 32 lower-case hex digits, and the check
 * CheckSgt/CheckSgt.py:44-78 makes: the file's digest against the first
 * word of the first line of an md5 file (0 equal, 1 different, < 0 error). */
int awp_md5_file(const char *path, char hex[33], char *err, size_t errlen);
int awp_md5_bytes(const void *data, size_t n, char hex[33]);
int awp_md5_check(const char *path, const char *md5path, char *err, size_t errlen);

/* Comparison of two raw SGT files of npts points x 6 components x nt
 * samples (AWP-HIP-SGT/utils/compare_sgts.c:7-119): its statistics -- mean
 * difference (test - ref), largest signed difference and its float index,
 * percent differences over the values whose reference value is > 1e-10
 * (mean, mean absolute, largest absolute and its index, histogram over
 * |pct| < 1e-4, 1e-3, 1e-2, 0.1, 0.3, 1, 3, 10, 100, 1000, >= 1000) -- and
 * the relative L2 error of every (point, component) time series whose
 * reference norm is at least floor_frac x the largest series norm: the
 * worst of them and its series index (point * 6 + component), the counts
 * checked and below the floor, and the largest absolute L2 difference
 * among the series below the floor. */
#define AWP_SGT_CMP_BINS 11
typedef struct {
    long long nvalues;
    double avg_diff;
    float max_diff;            /* largest test - ref (signed, as the reference tool) */
    long long max_diff_index;
    float max_abs_diff;
    long long pct_count;
    double avg_pct, avg_abs_pct;
    float max_pct;
    long long max_pct_index;
    long long bins[AWP_SGT_CMP_BINS];
    double max_series_norm;
    long long series_checked, series_below_floor;
    double worst_rel_l2;
    long long worst_series;
    double worst_abs_below_floor;
} awp_sgt_cmp;
int awp_sgt_compare(const char *ref_path, const char *test_path, long long npts, int nt, double floor_frac,
                    awp_sgt_cmp *out, char *err, size_t errlen);

/* ---------------------------------------------------------------------- */
/* Device run (HIP)                                                        */
/* ---------------------------------------------------------------------- */

/* Create a context on HIP device `device` (allocates every device buffer;
 * replaces the host's cudaMalloc block, pmcl3d.c:537-823, and
 * SetDeviceConstValue, kernel...cu:24).  Two time-step schemes give the same
 * bits: the two-stage step (velocity kernel, free-surface images, stress
 * kernel, in place) and the one-sweep fused step, which reads one wavefield
 * buffer and writes a second one (36 more bytes per grid point).  FP32 runs
 * take the fused step, mixed-storage runs the two-stage one; AWP_SCHEME=split
 * or AWP_SCHEME=fused in the environment overrides (grids whose arrays reach
 * 4 GiB always take the two-stage step). */
int awp_create(awp_ctx **out, const awp_in3d *p, int device, int storage,
               char *err, size_t errlen);
int awp_destroy(awp_ctx *ctx);
const char *awp_last_error(const awp_ctx *ctx);

/* Media in file layout, float32 [nz][ny][nx][nvar] (structure.f:130-170).
 * The transform to compliances, density and Q weights (structure.f:449-524),
 * the ghost replicas (padmedia/afsmedia, structure.f:563-707), the
 * relaxation times (relax.f:4-39) and the sponge (cerjan.f:4-247) are set up
 * on the device.  _device takes a device pointer. */
int awp_set_media(awp_ctx *ctx, const float *media_host);
/* The same from an RWG velocity mesh (the .p/.s/.d files, each float32,
 * x fastest, then z, then y, km/s and g/cm^3, on the RWG grid whose x is
 * the AWP y): the conversion SgtHead's reformat_velocity makes into the AWP
 * media file (SgtHead/src/reformat_velocity.c:62-72, x1000 to m/s, kg/m^3)
 * runs on the device, no media file is written.  Needs NVAR = 3. */
int awp_set_media_rwg(awp_ctx *ctx, const float *vp_km, const float *vs_km, const float *rho_g);
int awp_set_media_device(awp_ctx *ctx, const float *media_dev);

/* Point sources: xyz [nsrc][3] file coordinates, stf [nsrc][nst][6]
 * (inisource/getsrc, source.f:3-402). */
int awp_set_source(awp_ctx *ctx, int nsrc, const int *xyz, const float *stf);

/* SGT output points in cordfile order (getrec SGT branch, station.f:31-80);
 * *nkept receives the number the reference rule keeps. */
int awp_set_sgt_points(awp_ctx *ctx, int npts, const int *xyz, int *nkept);

/* Zero the wavefield and memory variables, inject source sample 1
 * (pmcl3d.f:988) and reset the step counter. */
int awp_init(awp_ctx *ctx);

/* Time-step scheme of a run; every scheme and launch shape gives the same
 * bits.  awp_tune times `reps` steps (default 3) of each candidate -- the
 * two-stage step (on grids of few tiles also with 64x4 tiles and shorter z
 * chunks) and, in a run that holds the second wavefield buffer, the fused
 * step on several tile, z-chunk and tile-patch shapes --, times the three
 * fastest again over twice the steps, keeps the fastest and starts the run
 * again (awp_init): call it on an initialised run before its first step.  awp_set_scheme(ctx, 1 / 0) switches to the fused /
 * two-stage step with its default launch shape at any point between steps
 * (the fused step needs the second buffer).  awp_get_scheme reports the step
 * in use (1 fused, 0 two-stage), its tile rows and its z chunks; any pointer
 * may be NULL.  No counterpart in the reference, which has one kernel
 * sequence (AWP-GPU-SGT/src/pmcl3d.c:841-917). */
int awp_tune(awp_ctx *ctx, int reps);
int awp_set_scheme(awp_ctx *ctx, int fused);
int awp_get_scheme(const awp_ctx *ctx, int *fused, int *tile_y, int *nchunk);

/* Advance nsteps full time steps (pmcl3d.f:1183-1557: velocity, sponge,
 * free-surface velocity, stress + memory variables, sponge, source sample
 * i+1, free-surface stress, SGT sample when i % NTISKP_SGT == 0).
 * Asynchronous on the context's stream.  Replaces the dvelcx_H / dstrqc_H /
 * addsrc_H / ComputeSGT call sequence of AWP-GPU-SGT/src/pmcl3d.c:841-917. */
int awp_step(awp_ctx *ctx, int nsteps);
int awp_synchronize(awp_ctx *ctx);
/* Completed steps and SGT samples so far. */
int awp_progress(const awp_ctx *ctx, int *step, int *nsamples);

/* SGT samples so far, [npts][6][nsamp_max] float32 with nsamp_max =
 * nt / NTISKP_SGT; slots of points the rule drops and samples not yet taken
 * are -1.  (Strains exx eyy ezz exy exz eyz, pmcl3d.f:1546-1557.) */
int awp_get_sgt(awp_ctx *ctx, float *host, size_t count);

/* The job-value trace the reference's master rank writes to the CHKJ file
 * after every step (pmcl3d.f:1476-1479, format (I6,3E10.3)): u1, v1, w1 at
 * the node (ND+2, ND+2, ND+2) counted from the bottom (the reference's rank 0
 * holds the first block of every axis; clamped into the grid), [step][3]
 * float32 for the steps completed so far.  ijk (may be NULL) receives the
 * node; host may be NULL to query it only. */
int awp_get_trace(awp_ctx *ctx, float *host, size_t count, int *ijk);

/* Write the raw SGT file exactly as the reference does (pmcl3d.f:1582-1601):
 * only full WRITE_STEP buffers, each at byte offset
 * g*24*NT_OUT + c*4*NT_OUT + (f-1)*4*WRITE_STEP; dropped points leave holes.
 * No file is created when no buffer fills.  The file is truncated first. */
int awp_write_sgt_file(awp_ctx *ctx, const char *path);

/* Write, in one pass from the device buffer, the file SgtHead's reformatter
 * makes of the raw file (`reformat_awp <raw> NT_OUT <out> [-z]`,
 * SgtHead/src/reformat_awp.c:7-66): RWG component order XX YY ZZ XY XZ YZ
 * with XX <- YY, YY <- XX, XZ <- -YZ, YZ <- -XZ, every value first times -2
 * when zsource != 0 (the Z-force component), complete records of the raw
 * file only, its holes as zeros.  Replaces awp_write_sgt_file + reformat_awp
 * (reformat_awp_mpi.c:91-122) in a CyberShake SGT workflow. */
int awp_write_sgt_file_rwg(awp_ctx *ctx, const char *path, int zsource);

/* Write the SGT in the compact FP16 format (no counterpart in the
 * reference; a storage option for the raw SGT, decoded by awp_sgt16_read):
 *   bytes 0-7 "AWPSGT16"; u32 version=1, npts, ncomp=6, nsamp, NTISKP_SGT;
 *   f32 DT; u32 NT_OUT (the raw file's series length, >= nsamp); zero
 *   padding to byte 64; f32 scale[npts][6]; then
 *   FP16 data[npts][6][nsamp], value = data * scale.
 * nsamp is the number of samples the raw file holds (full WRITE_STEP
 * buffers); scale = 2^(e-15) per series, where the series peak is
 * f * 2^e' with f in [0.5, 1) and e = max(e', -134) (so scale >= 2^-149,
 * the smallest FP32), and data = fp16(value * 2^(15-e)), round to nearest
 * even; points the reference drops and all-zero series have scale 0.  Error
 * per value <= 2^-11 of the series peak (+ 2^-149 for peaks below 2^-134). */
int awp_write_sgt_file_fp16(awp_ctx *ctx, const char *path);

/* Dimensions of an FP16 SGT file, and its values decoded to float32 as
 * [npts][6][nsamp] (out may be NULL to query the dimensions only). */
int awp_sgt16_read(const char *path, int *npts, int *nsamp, float *out, size_t count,
                   char *err, size_t errlen);
/* Points first .. first+count-1 of an FP16 SGT file decoded to float32
 * [count][6][NT_OUT] (samples the raw file leaves empty are zero) in the
 * layout a consumer of the FP32 files reads: layout 0 the raw AWP file's
 * records; 1 the RWG records SgtHead's reformat_awp writes
 * (SgtHead/src/reformat_awp.c:38-62), 2 the same for the Z source (-z).
 * Layouts 1/2 give DirectSynth's reader the slice readSGT_MPI takes from the
 * reformatted file (DirectSynth/src/sgt_handler.c:282-333), to FP16
 * rounding. */
int awp_sgt16_read_points(const char *path, int first, int count, int layout, float *out, size_t n,
                          char *err, size_t errlen);

/* Copy one interior field to the host as [k][j][i], x fastest, k = 1 at the
 * bottom; kext = 2 also copies the two free-surface image planes.  which:
 * 0 u1 1 v1 2 w1 3 xx 4 yy 5 zz 6 xy 7 xz 8 yz 9..14 r1..r6 15 d1 16 mu
 * 17 lam 18 qp 19 qs.  For tests and checkpoints. */
int awp_get_field(awp_ctx *ctx, int which, int kext, float *host);

/* awp_get_field without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * field is packed, a chunk of levels at a time, into a file-scope device
 * scratch array of 8M floats (32 MiB, held on every GPU the library runs on
 * until the process ends) and copied to `host`.  Rejects (AWP_EINVAL) grids
 * whose nx*ny exceeds 8M points.  Calls that overlap on one GPU share the
 * scratch and race; calls on different GPUs do not. */
int awp_get_field_noalloc(awp_ctx *ctx, int which, int kext, float *host);

/* Device time of the last awp_step() call per kernel family, from HIP events
 * on the context's stream (milliseconds): out[0] velocity, out[1] stress,
 * out[2] whole step sequence.  Enabled by awp_set_timing(ctx, 1). */
int awp_set_timing(awp_ctx *ctx, int on);
int awp_get_timing(awp_ctx *ctx, double *out3);

/* Bytes of device memory the context holds. */
int awp_device_bytes(const awp_ctx *ctx, size_t *bytes);

/* Device memory a run of these parameters with npts SGT points needs:
 * *steady while it steps, *peak during awp_set_media (the media file and
 * set-up temporaries on the device as well).  For admitting runs to a GPU
 * under a memory budget.  No counterpart in the reference. */
int awp_estimate_bytes(const awp_in3d *p, int storage, int npts, size_t *steady, size_t *peak);

/* Free and total device memory of a GPU (for admission decisions). */
int awp_device_memory(int device, size_t *free_bytes, size_t *total_bytes);

/* ---- checkpoint / restart ----------------------------------------------
 * awp_checkpoint_save writes the run's evolving state -- the wavefield and
 * memory variables in their storage format, the SGT samples so far, the
 * step count -- to `path` (through path.tmp, renamed when complete);
 * awp_checkpoint_load restores it into an initialised context of the same
 * grid, storage mode, wavefield buffers (one or two; the step in use may
 * differ) and SGT points (media and source are the
 * caller's inputs as at the start), after which awp_step continues bit for
 * bit as the uninterrupted run (a load that fails after it has started
 * copying leaves the run uninitialised).  Replaces the reference's per-rank restart
 * files (pmcl3d.f:1016-1097 read at start-up, :1964-2045 written every
 * CHECKPOINT steps), which hold the same state for one MPI rank. */
int awp_checkpoint_save(awp_ctx *ctx, const char *path);
int awp_checkpoint_load(awp_ctx *ctx, const char *path);

/* ---- SGT header (SgtHead write_head) -------------------------------------
 * The separate header file SgtHead's write_head writes beside an AWP SGT
 * run (SgtHead/src/write_head.c:51-469): one sgtmaster record, then one
 * sgtindex and one sgtheader record per point of the RWG coordfile, in its
 * order, with the site and point coordinates projected on RWG's great-circle
 * model frame and lambda, mu (dyne/cm^2) and rho (g/cm^3) read from the
 * AWP-format media file ([z][x][y][3] Vp, Vs m/s, rho kg/m^3, as
 * reformat_velocity writes it).  Records and byte layout are the reference's
 * (structure.h: 36-byte master, 24-byte index, 128-byte header). */
typedef struct {
    const char *modelbox;   /* model box file (5th line: "mlon= ... mlat= ... mrot= ...") */
    const char *coordfile;  /* RWG coordfile: count, then "x y z indx lon lat dep" lines */
    const char *fdloc;      /* site grid location "x y" */
    const char *gridout;    /* RWG gridout (nx=, ny=, nz= blocks) */
    const char *media;      /* AWP-format media file */
    const char *out;        /* header file to write */
    float spacing;          /* grid spacing, km */
    int nt;                 /* time steps of the run */
    float dt;               /* time step, s */
    int sgt_tinc;           /* time decimation of the SGT (NTISKP_SGT) */
    char component;         /* 'x', 'y' or 'z' force */
    float moment;           /* force moment */
    float flo;              /* frequency the source was filtered at (tst = -3/flo) */
    int mu_corr;            /* 1: corrected mu with Qs = 0.05 Vs (write_head -c) */
} awp_head_params;
int awp_write_sgt_head(const awp_head_params *hp, char *err, size_t errlen);

/* ---- batches ------------------------------------------------------------
 * Several initialised runs (same device and storage mode) advanced
 * together: each stage of a time step is one persistent launch over the
 * tiles of every run (two-stage scheme: velocity, free-surface images,
 * stress; fused scheme: the one-sweep kernel), then one launch for every
 * run's source and SGT sample, replayed from a HIP graph.  The fused scheme
 * is used when every run holds two wavefield buffers and steps fused (the
 * FP32 default, or AWP_SCHEME=fused), else the two-stage one.  Each run keeps its own
 * outputs (awp_get_sgt, awp_write_sgt_file).  The reference runs one SGT
 * component per MPI job (AWP-ODC-SGT/utils/awp_odc_wrapper.sh:14); these
 * entry points have no counterpart in it. */
typedef struct awp_batch awp_batch;
int awp_batch_create(awp_batch **out, awp_ctx *const *runs, int nruns, char *err, size_t errlen);
/* Times the batch's step kernels for each scheme (fused only when the runs
 * hold two buffers) / tile height / prefetch depth or z unroll / z chunking
 * / persistent or one-block-per-tile grid candidate (reps launches each),
 * times the three fastest again over twice the launches and keeps the
 * fastest.  Needs runs that have not stepped; re-initialises them
 * (awp_init) afterwards. */
int awp_batch_tune(awp_batch *b, int reps);
/* Advances every run min(nsteps, steps it has left) steps; returns when
 * they are done. */
int awp_batch_step(awp_batch *b, int nsteps);
int awp_batch_synchronize(awp_batch *b);
/* Device milliseconds of the last awp_batch_step (the last tuning launch
 * after awp_batch_tune). */
int awp_batch_time(const awp_batch *b, double *ms);
/* The kernel variant in use: tile rows, prefetch depth (fused) or z unroll
 * (two-stage), z chunks of run 0. */
int awp_batch_get_launch(const awp_batch *b, int *tile_y, int *prefetch, int *nchunk);
/* *fused = 1 for the fused scheme, 0 for the two-stage one; *persistent = 1
 * when the stage kernels run persistent grids (the blocks that fit on the
 * GPU at once claim the tiles of the list per XCD), 0 for one block per tile
 * (the tuner picks).  Either pointer may be NULL. */
int awp_batch_get_scheme(const awp_batch *b, int *fused, int *persistent);
const char *awp_batch_last_error(const awp_batch *b);
int awp_batch_destroy(awp_batch *b);

#ifdef __cplusplus
}
#endif

#endif /* AWPSGT_H */
