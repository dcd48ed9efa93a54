/*
 * awp_pmcl3d.c -- C host of the MI355X SGT solver, a drop-in for the
 * reference's `pmcl3d <IN3D> [src] [media] [chkp] [chkj] ...` run in SGT mode
 * (AWP-ODC-SGT/src/pmcl3d.f, argument handling set_names.f:29-47 and
 * io.f:444-479).  It reads the same IN3D, media, source and cordfile, runs
 * the time loop on one GPU through libawpsgt, and writes the raw SGT file
 * the reference writes (SGTGRO), plus the CHKP parameter summary
 * (io.f:481-537).
 *
 *   awp_pmcl3d [IN3D] [src] [media] [chkp] [chkj]
 *   awp_pmcl3d --batch IN3D_1 [IN3D_2 ...]
 *
 * The batch form runs several SGT components (e.g. a site's X and Y runs,
 * each with its own IN3D as the reference's workflow writes them) in one
 * process: runs are admitted in queue order while their device memory
 * (awp_estimate_bytes) fits the budget, each admitted wave is advanced as
 * one batch (awp_batch_*), and every run writes its own outputs.  The
 * reference runs one component per MPI job (utils/awp_odc_wrapper.sh:14).
 *
 * Environment: AWP_DEVICE selects the HIP device (default 0);
 * AWP_STORAGE=mixed selects mixed storage (FP16 memory variables, FP16
 * attenuation weights); AWP_SGT_FP16=<path> also writes the SGT in the
 * compact FP16 format to <path>; AWP_SGT_RWG=<path> also writes the file
 * SgtHead's reformat_awp would make of the raw file (RWG component order,
 * with its -z scaling for a Z-force run, igreen 3), in one pass.  In batch
 * form the two paths get ".<n>" appended (n = position in the list);
 * AWP_HBM_BUDGET_GB caps the device memory a wave may use (default: the free
 * memory less 2 GiB) and AWP_TUNE=1 runs the batch tuner first.  A single
 * run picks its fastest time-step scheme and launch shape first (awp_tune;
 * AWP_TUNE=0 skips that); batch runs take the two-stage step unless
 * AWP_SCHEME says otherwise (the fused step's second wavefield buffer would
 * admit fewer runs per wave).
 */
#define _POSIX_C_SOURCE 200809L
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "awpsgt.h"

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static int die(const char *what, const char *msg)
{
    fprintf(stderr, "awp_pmcl3d: %s: %s\n", what, msg);
    return 1;
}

static void set_name(char *dst, const char *src)
{
    strncpy(dst, src, 255);
    dst[255] = 0;
}

/* Media file, float32 [nz][ny][nx][nvar] (structure.f:130-170). */
static float *read_media(const awp_in3d *p, char *err, size_t errlen)
{
    const size_t n = (size_t)p->nx * p->ny * p->nz * p->nvar;
    FILE *f = fopen(p->invel, "rb");
    if (!f) {
        snprintf(err, errlen, "cannot open media '%s': %s", p->invel, strerror(errno));
        return NULL;
    }
    float *m = (float *)malloc(n * sizeof(float));
    if (!m) {
        fclose(f);
        snprintf(err, errlen, "out of host memory for %zu media values", n);
        return NULL;
    }
    size_t got = fread(m, sizeof(float), n, f);
    fclose(f);
    if (got != n) {
        free(m);
        snprintf(err, errlen, "media '%s' holds %zu values, IN3D needs %zu", p->invel, got, n);
        return NULL;
    }
    return m;
}

/* The reference's CHKP summary (wrpm3d, io.f:481-537), single-domain form. */
static void write_chkp(const awp_in3d *p, const float *media)
{
    if (!p->chkp[0]) return;
    FILE *f = fopen(p->chkp, "w");
    if (!f) return;
    float vpmin = 1e30f, vpmax = 0, vsmin = 1e30f, vsmax = 0, dmin = 1e30f, dmax = 0;
    const size_t n = (size_t)p->nx * p->ny * p->nz;
    const int off = (p->nvar == 8) ? 3 : 0;
    for (size_t q = 0; q < n; ++q) {
        const float *r = media + q * p->nvar + off;
        if (r[0] < vpmin) vpmin = r[0];
        if (r[0] > vpmax) vpmax = r[0];
        if (r[1] < vsmin) vsmin = r[1];
        if (r[1] > vsmax) vsmax = r[1];
        if (r[2] < dmin) dmin = r[2];
        if (r[2] > dmax) dmax = r[2];
    }
    fprintf(f, " igreen = %d\n", p->igreen);
    fprintf(f, "STABILITY CRITERIA .5 > CMAX*DT/DX = %15.7E\n", vpmax * p->dt / p->dh);
    fprintf(f, "# OF X NODES PER PROC%5d\n", p->nx / p->npx);
    fprintf(f, "# OF Y NODES PER PROC%5d\n", p->ny / p->npy);
    fprintf(f, "# OF Z NODES PER PROC%5d\n", p->nz / p->npz);
    fprintf(f, "# OF TIME STEPS%5d\n", p->nt);
    fprintf(f, "DISCRETIZATION IN SPACE%10.4f\n", p->dh);
    fprintf(f, "DISCRETIZATION IN TIME%10.4f\n", p->dt);
    fprintf(f, "PML REFLECTION COEFFICIENT%13.4G\n", p->arbc);
    fprintf(f, "HIGHEST P-VELOCITY ENCOUNTERED%10.2f\n", vpmax);
    fprintf(f, "LOWEST P-VELOCITY ENCOUNTERED%10.2f\n", vpmin);
    fprintf(f, "HIGHEST S-VELOCITY ENCOUNTERED%10.2f\n", vsmax);
    fprintf(f, "LOWEST S-VELOCITY ENCOUNTERED%10.2f\n", vsmin);
    fprintf(f, "HIGHEST DENSITY ENCOUNTERED%15.2f\n", dmax);
    fprintf(f, "LOWEST  DENSITY ENCOUNTERED%15.2f\n", dmin);
    fprintf(f, "SKIP OF SEISMOGRAMS IN TIME (LOOP COUNTER)%5d\n", p->ntiskp);
    fprintf(f, "ABC CONDITION, PML=1 OR CERJAN=0:%5d\n", p->npc);
    fprintf(f, "FD SCHEME, VISCO=1 OR ELASTIC=0:%5d\n", p->nve);
    fprintf(f, "Q, FL,FP,FH:%14.5G%14.5G%14.5G\n", p->fl, p->fp, p->fh);
    fclose(f);
}

typedef struct {
    awp_in3d p;
    int npts;
    size_t steady, peak;
    awp_ctx *ctx;
} job_t;

/* Reads one run's inputs and sets up its context (media, source, points,
 * init).  The IN3D is read and validated by the caller. */
static int setup_run(job_t *j, int device, int storage, char *err, size_t errlen)
{
    awp_in3d *p = &j->p;
    int rc = 1;
    int *sxyz = (int *)malloc(sizeof(int) * 3 * (size_t)p->nsrc);
    float *stf = (float *)malloc(sizeof(float) * 6 * (size_t)p->nsrc * p->nst);
    int *pts = (int *)malloc(sizeof(int) * 3 * (size_t)(j->npts > 0 ? j->npts : 1));
    float *media = NULL;
    if (!sxyz || !stf || !pts) {
        snprintf(err, errlen, "out of host memory");
        goto done;
    }
    if (awp_source_read(p->insrc, p->nsrc, p->nst, sxyz, stf, err, errlen)) goto done;
    if (awp_cordfile_read(p->insgt, j->npts, pts, err, errlen)) goto done;
    media = read_media(p, err, errlen);
    if (!media) goto done;
    if (awp_create(&j->ctx, p, device, storage, err, errlen)) goto done;
    int nkept = 0;
    if (awp_set_sgt_points(j->ctx, j->npts, pts, &nkept) || awp_set_media(j->ctx, media) ||
        awp_set_source(j->ctx, p->nsrc, sxyz, stf) || awp_init(j->ctx)) {
        snprintf(err, errlen, "%s", awp_last_error(j->ctx));
        goto done;
    }
    write_chkp(p, media);
    printf(" %s: SGT points kept %d of %d; %d time steps\n", p->sgtgro, nkept, j->npts, p->nt);
    rc = 0;
done:
    free(media);
    free(sxyz);
    free(stf);
    free(pts);
    return rc;
}

/* A value in Fortran's E10.3 edit descriptor: 0.ddd mantissa, two-digit
 * exponent ("-0.542E-42", " 0.000E+00"); a three-digit exponent drops the E. */
static void fort_e10_3(float v, char out[16])
{
    if (isnan(v)) {
        snprintf(out, 16, "%10s", "NaN");
        return;
    }
    if (isinf(v)) {
        snprintf(out, 16, "%10s", v < 0 ? "-Infinity" : "Infinity");
        return;
    }
    if (v == 0.0f) {
        snprintf(out, 16, "%10s", signbit(v) ? "-0.000E+00" : "0.000E+00");
        return;
    }
    /* three significant digits d.dd x 10^e  ->  0.ddd x 10^(e+1) */
    char t[32];
    snprintf(t, sizeof t, "%.2e", fabs((double)v));
    const int e = atoi(strchr(t, 'e') + 1) + 1;
    const char sign = v < 0 ? '-' : ' ';
    char body[24];
    if (e >= -99 && e <= 99)
        snprintf(body, sizeof body, "0.%c%c%cE%c%02d", t[0], t[2], t[3], e < 0 ? '-' : '+', e < 0 ? -e : e);
    else
        snprintf(body, sizeof body, "0.%c%c%c%c%03d", t[0], t[2], t[3], e < 0 ? '-' : '+', e < 0 ? -e : e);
    snprintf(out, 16, "%c%s", sign, body);
}

/* The CHKJ trace (pmcl3d.f:1476-1479: write(24,'(I6,3E10.3)') i, u1, v1, w1
 * at the master rank's node (ND+2, ND+2, ND+2) after every step). */
static int write_chkj(job_t *j)
{
    if (!j->p.chkj[0]) return 0;
    int steps = 0;
    if (awp_progress(j->ctx, &steps, NULL)) return die("CHKJ", awp_last_error(j->ctx));
    float *tr = (float *)malloc(sizeof(float) * 3 * (size_t)(steps > 0 ? steps : 1));
    if (!tr) return die("CHKJ", "out of host memory");
    if (awp_get_trace(j->ctx, tr, (size_t)3 * steps, NULL)) {
        free(tr);
        return die("CHKJ", awp_last_error(j->ctx));
    }
    FILE *f = fopen(j->p.chkj, "w");
    if (!f) {
        free(tr);
        return die("CHKJ", "cannot create the file");
    }
    for (int i = 0; i < steps; ++i) {
        char a[16], b[16], c[16];
        fort_e10_3(tr[3 * i], a);
        fort_e10_3(tr[3 * i + 1], b);
        fort_e10_3(tr[3 * i + 2], c);
        fprintf(f, "%6d%s%s%s\n", i + 1, a, b, c);
    }
    free(tr);
    if (fclose(f)) return die("CHKJ", "write failed");
    return 0;
}

static int write_outputs(job_t *j, const char *fp16_path, const char *rwg_path)
{
    if (awp_write_sgt_file(j->ctx, j->p.sgtgro)) return die("SGT output", awp_last_error(j->ctx));
    if (write_chkj(j)) return 1;
    if (fp16_path && fp16_path[0] && awp_write_sgt_file_fp16(j->ctx, fp16_path))
        return die("FP16 SGT output", awp_last_error(j->ctx));
    if (rwg_path && rwg_path[0] && awp_write_sgt_file_rwg(j->ctx, rwg_path, j->p.igreen == 3))
        return die("RWG SGT output", awp_last_error(j->ctx));
    return 0;
}

static int run_batch(int n, char **in3d, int device, int storage)
{
    char err[512];
    job_t *jobs = (job_t *)calloc((size_t)n, sizeof(job_t));
    if (!jobs) return die("batch", "out of memory");
    for (int k = 0; k < n; ++k) {
        if (awp_in3d_read(in3d[k], &jobs[k].p, err, sizeof(err)) ||
            awp_in3d_validate(&jobs[k].p, err, sizeof(err)))
            return die(in3d[k], err);
        if (awp_cordfile_count(jobs[k].p.insgt, &jobs[k].npts, err, sizeof(err))) return die(in3d[k], err);
        if (awp_estimate_bytes(&jobs[k].p, storage, jobs[k].npts, &jobs[k].steady, &jobs[k].peak))
            return die(in3d[k], "bad parameters");
    }
    size_t free_b = 0, total_b = 0;
    if (awp_device_memory(device, &free_b, &total_b)) return die("device", "cannot query memory");
    size_t budget = free_b > ((size_t)2 << 30) ? free_b - ((size_t)2 << 30) : free_b;
    const char *bud = getenv("AWP_HBM_BUDGET_GB");
    if (bud && atof(bud) > 0) budget = (size_t)(atof(bud) * 1e9);
    const char *fp16 = getenv("AWP_SGT_FP16"), *rwg = getenv("AWP_SGT_RWG");
    const int tune = getenv("AWP_TUNE") && atoi(getenv("AWP_TUNE")) == 1;
    enum { MAXW = 256 };
    awp_ctx *ctxs[MAXW];
    int k0 = 0, wave = 0;
    while (k0 < n) {
        /* admission in queue order: steady bytes plus the largest set-up
         * transient of the wave within the budget */
        size_t st = 0, tr = 0;
        int k1 = k0;
        while (k1 < n && k1 - k0 < MAXW) {
            const size_t t = jobs[k1].peak - jobs[k1].steady;
            if (k1 == k0 && jobs[k1].peak > budget) return die(in3d[k1], "needs more device memory than the budget");
            if (k1 > k0 && st + jobs[k1].steady + (t > tr ? t : tr) > budget) break;
            st += jobs[k1].steady;
            if (t > tr) tr = t;
            ++k1;
        }
        printf(" wave %d: runs %d..%d, %.1f GB of a %.1f GB budget\n", wave, k0, k1 - 1, st / 1e9, budget / 1e9);
        int minnt = 0x7fffffff;
        for (int k = k0; k < k1; ++k) {
            if (setup_run(&jobs[k], device, storage, err, sizeof(err))) return die(in3d[k], err);
            ctxs[k - k0] = jobs[k].ctx;
            if (jobs[k].p.nt < minnt) minnt = jobs[k].p.nt;
        }
        awp_batch *b = NULL;
        if (awp_batch_create(&b, ctxs, k1 - k0, err, sizeof(err))) return die("batch", err);
        if (tune && awp_batch_tune(b, 2)) return die("batch tune", awp_batch_last_error(b));
        const double t0 = now_s();
        for (int done = 0; done < minnt;) {
            const int m = minnt - done < 100 ? minnt - done : 100;
            if (awp_batch_step(b, m)) return die("batch step", awp_batch_last_error(b));
            done += m;
            printf(" TIME STEP %d OF %d\n", done, minnt);
            fflush(stdout);
        }
        const double t1 = now_s();
        awp_batch_destroy(b);
        double pts = 0;
        for (int k = k0; k < k1; ++k) {
            /* runs longer than the shortest finish on their own */
            if (jobs[k].p.nt > minnt && awp_step(jobs[k].ctx, jobs[k].p.nt - minnt))
                return die("awp_step", awp_last_error(jobs[k].ctx));
            if (awp_synchronize(jobs[k].ctx)) return die("awp_step", awp_last_error(jobs[k].ctx));
            char f16[300] = "", rw[300] = "";
            if (fp16 && fp16[0]) snprintf(f16, sizeof f16, "%s.%d", fp16, k);
            if (rwg && rwg[0]) snprintf(rw, sizeof rw, "%s.%d", rwg, k);
            if (write_outputs(&jobs[k], f16, rw)) return 1;
            pts += (double)jobs[k].p.nx * jobs[k].p.ny * jobs[k].p.nz;
            awp_destroy(jobs[k].ctx);
            jobs[k].ctx = NULL;
        }
        if (minnt > 0) {
            printf(" computing time per time step = %12.6f sec (wave of %d runs)\n", (t1 - t0) / minnt, k1 - k0);
            printf(" point-updates per second = %.4g\n", pts * minnt / (t1 - t0));
        }
        k0 = k1;
        ++wave;
    }
    free(jobs);
    return 0;
}

int main(int argc, char **argv)
{
    char err[512];
    const char *dev_env = getenv("AWP_DEVICE");
    const int device = dev_env ? atoi(dev_env) : 0;
    const char *st_env = getenv("AWP_STORAGE");
    int storage = AWP_STORAGE_FP32;
    if (st_env && st_env[0]) {
        if (!strcmp(st_env, "mixed")) storage = AWP_STORAGE_MIXED;
        else if (strcmp(st_env, "fp32")) return die("AWP_STORAGE", "expected 'fp32' or 'mixed'");
    }
    if (argc >= 2 && !strcmp(argv[1], "--batch")) {
        if (argc < 3) return die("--batch", "no IN3D files given");
        setenv("AWP_SCHEME", "split", 0);
        return run_batch(argc - 2, argv + 2, device, storage);
    }

    const char *in3d_path = argc >= 2 ? argv[1] : "IN3D";
    job_t j;
    memset(&j, 0, sizeof j);
    awp_in3d *p = &j.p;
    if (awp_in3d_read(in3d_path, p, err, sizeof(err))) return die("IN3D", err);
    /* command-line overrides of the file names (io.f:444-479) */
    if (argc >= 3) set_name(p->insrc, argv[2]);
    if (argc >= 4) set_name(p->invel, argv[3]);
    if (argc >= 5) set_name(p->chkp, argv[4]);
    if (argc >= 6) set_name(p->chkj, argv[5]);
    if (awp_in3d_validate(p, err, sizeof(err))) return die("IN3D", err);
    const char *fp16_path = getenv("AWP_SGT_FP16");
    const char *rwg_path = getenv("AWP_SGT_RWG");

    double t0 = now_s();
    if (awp_cordfile_count(p->insgt, &j.npts, err, sizeof(err)))
        return die("cordfile", err);  /* no station list: SGT mode needs one */
    printf(" Use SGT output mode, sgt_numsta= %d\n", j.npts);
    if (setup_run(&j, device, storage, err, sizeof(err))) return die("set-up", err);
    const char *tune_env = getenv("AWP_TUNE");
    if (!(tune_env && !strcmp(tune_env, "0"))) {
        int fused = 0, ty = 0, nc = 0;
        if (awp_tune(j.ctx, 3) || awp_get_scheme(j.ctx, &fused, &ty, &nc))
            return die("tune", awp_last_error(j.ctx));
        printf(" time step: %s, 64 x %d tiles, %d z chunks\n", fused ? "fused one-sweep" : "two-stage", ty, nc);
    }
    double t2 = now_s();

    /* CHECKPOINT > 0: resume from output_ckp/awp_ckp.bin when it exists and
       write it every CHECKPOINT steps (pmcl3d.f:1016-1097, 1964-2045; the
       reference keeps one file set per rank in the same directory) */
    const char *ckp = "output_ckp/awp_ckp.bin";
    int done = 0;
    if (p->checkpoint > 0 && access(ckp, R_OK) == 0) {
        if (awp_checkpoint_load(j.ctx, ckp)) return die("checkpoint", awp_last_error(j.ctx));
        if (awp_progress(j.ctx, &done, NULL)) return die("checkpoint", awp_last_error(j.ctx));
        printf(" Checkpoint found, starting from step %d\n", done);
    }
    const int chunk = 100;
    double tl = now_s();
    const int first = done;
    while (done < p->nt) {
        int n = p->nt - done < chunk ? p->nt - done : chunk;
        if (p->checkpoint > 0) {
            const int next_ckp = (done / p->checkpoint + 1) * p->checkpoint;
            if (done + n > next_ckp) n = next_ckp - done;
        }
        if (awp_step(j.ctx, n)) return die("awp_step", awp_last_error(j.ctx));
        done += n;
        if (awp_synchronize(j.ctx)) return die("awp_step", awp_last_error(j.ctx));
        if (p->checkpoint > 0 && done % p->checkpoint == 0 && done < p->nt) {
            if (awp_checkpoint_save(j.ctx, ckp)) return die("checkpoint", awp_last_error(j.ctx));
            printf(" Writing checkpoint file at step %d\n", done);
        }
        printf(" TIME STEP %d OF %d\n", done, p->nt);
        fflush(stdout);
    }
    double t3 = now_s();
    if (write_outputs(&j, fp16_path, rwg_path)) return 1;
    double t4 = now_s();
    const double per_step = (t3 - tl) / (p->nt > first ? p->nt - first : 1);
    printf(" inialization time = %10.3f sec\n", t2 - t0);
    printf(" computing time per time step = %12.6f sec\n", per_step);
    printf(" point-updates per second = %.4g\n", (double)p->nx * p->ny * p->nz / per_step);
    printf(" SGT write time = %10.3f sec\n", t4 - t3);
    awp_destroy(j.ctx);
    return 0;
}
