/*
 * awp_pmcl3d.c -- C host of the MI355X SGT solver, a drop-in for the
 * reference's `pmcl3d <IN3D> [src] [media] [chkp] [chkj] ...` run in SGT mode
 * (AWP-ODC-SGT/src/pmcl3d.f, argument handling set_names.f:29-47 and
 * io.f:444-479).  It reads the same IN3D, media, source and cordfile, runs
 * the time loop on one GPU through libawpsgt, and writes the raw SGT file
 * the reference writes (SGTGRO), plus the CHKP parameter summary
 * (io.f:481-537).
 *
 * Environment: AWP_DEVICE selects the HIP device (default 0);
 * AWP_STORAGE=mixed selects mixed storage (BF16 memory variables, FP16
 * attenuation weights); AWP_SGT_FP16=<path> also writes the SGT in the
 * compact FP16 format to <path>; AWP_SGT_RWG=<path> also writes the file
 * SgtHead's reformat_awp would make of the raw file (RWG component order,
 * with its -z scaling for a Z-force run, igreen 3), in one pass.
 */
#define _POSIX_C_SOURCE 200809L
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

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

int main(int argc, char **argv)
{
    char err[512];
    const char *in3d_path = argc >= 2 ? argv[1] : "IN3D";
    awp_in3d p;
    if (awp_in3d_read(in3d_path, &p, err, sizeof(err))) return die("IN3D", err);
    /* command-line overrides of the file names (io.f:444-479) */
    if (argc >= 3) set_name(p.insrc, argv[2]);
    if (argc >= 4) set_name(p.invel, argv[3]);
    if (argc >= 5) set_name(p.chkp, argv[4]);
    if (argc >= 6) set_name(p.chkj, argv[5]);
    if (awp_in3d_validate(&p, err, sizeof(err))) return die("IN3D", err);

    const char *dev_env = getenv("AWP_DEVICE");
    const int device = dev_env ? atoi(dev_env) : 0;
    const char *st_env = getenv("AWP_STORAGE");
    int storage = AWP_STORAGE_FP32;
    if (st_env && st_env[0]) {
        if (!strcmp(st_env, "mixed")) storage = AWP_STORAGE_MIXED;
        else if (strcmp(st_env, "fp32")) return die("AWP_STORAGE", "expected 'fp32' or 'mixed'");
    }
    const char *fp16_path = getenv("AWP_SGT_FP16");
    const char *rwg_path = getenv("AWP_SGT_RWG");
    double t0 = now_s();

    int *sxyz = (int *)malloc(sizeof(int) * 3 * (size_t)p.nsrc);
    float *stf = (float *)malloc(sizeof(float) * 6 * (size_t)p.nsrc * p.nst);
    if (!sxyz || !stf) return die("source", "out of memory");
    if (awp_source_read(p.insrc, p.nsrc, p.nst, sxyz, stf, err, sizeof(err))) return die("source", err);

    int npts = 0;
    if (awp_cordfile_count(p.insgt, &npts, err, sizeof(err)))
        return die("cordfile", err);  /* no station list: SGT mode needs one */
    int *pts = (int *)malloc(sizeof(int) * 3 * (size_t)(npts > 0 ? npts : 1));
    if (!pts) return die("cordfile", "out of memory");
    if (awp_cordfile_read(p.insgt, npts, pts, err, sizeof(err))) return die("cordfile", err);
    printf(" Use SGT output mode, sgt_numsta= %d\n", npts);

    float *media = read_media(&p, err, sizeof(err));
    if (!media) return die("media", err);
    double t1 = now_s();

    awp_ctx *ctx = NULL;
    if (awp_create(&ctx, &p, device, storage, err, sizeof(err))) return die("awp_create", err);
    int nkept = 0;
    if (awp_set_sgt_points(ctx, npts, pts, &nkept) || awp_set_media(ctx, media) ||
        awp_set_source(ctx, p.nsrc, sxyz, stf) || awp_init(ctx))
        return die("set-up", awp_last_error(ctx));
    write_chkp(&p, media);
    free(media);
    double t2 = now_s();
    printf(" SGT points kept %d of %d; %d time steps\n", nkept, npts, p.nt);

    const int chunk = 100;
    double tl = now_s();
    for (int done = 0; done < p.nt;) {
        int n = p.nt - done < chunk ? p.nt - done : chunk;
        if (awp_step(ctx, n)) return die("awp_step", awp_last_error(ctx));
        done += n;
        if (awp_synchronize(ctx)) return die("awp_step", awp_last_error(ctx));
        printf(" TIME STEP %d OF %d\n", done, p.nt);
        fflush(stdout);
    }
    double t3 = now_s();
    if (awp_write_sgt_file(ctx, p.sgtgro)) return die("SGT output", awp_last_error(ctx));
    if (fp16_path && fp16_path[0] && awp_write_sgt_file_fp16(ctx, fp16_path))
        return die("FP16 SGT output", awp_last_error(ctx));
    if (rwg_path && rwg_path[0] && awp_write_sgt_file_rwg(ctx, rwg_path, p.igreen == 3))
        return die("RWG SGT output", awp_last_error(ctx));
    double t4 = now_s();
    const double per_step = (t3 - tl) / p.nt;
    printf(" inialization time = %10.3f sec\n", t2 - t0);
    printf("    read input time = %10.3f sec\n", t1 - t0);
    printf(" computing time per time step = %12.6f sec\n", per_step);
    printf(" point-updates per second = %.4g\n", (double)p.nx * p.ny * p.nz / per_step);
    printf(" SGT write time = %10.3f sec\n", t4 - t3);
    awp_destroy(ctx);
    free(sxyz);
    free(stf);
    free(pts);
    return 0;
}
