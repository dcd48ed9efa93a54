/*
 * harness.c -- drives the host C parsers and writers of libawpsgt
 * (host/awp_io.c, host/awp_head.c, host/awp_sgtcheck.c) for a CPU build
 * under AddressSanitizer and UndefinedBehaviorSanitizer (tests/test_sanitize.py).
 * Every call's return code is printed; a parser fed a malformed file must
 * return an error, never read or write out of bounds.
 *
 *   harness run <IN3D>                  IN3D, source, cordfile, point rule, geometry
 *   harness head <modelbox> <coordfile> <fdloc> <gridout> <media> <out>
 *   harness sgt16 <file> <npts> <nt_out>
 *   harness check <file>                scan, md5, compare with itself
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "awpsgt.h"

static char err[512];

static int run(const char *in3d)
{
    awp_in3d p;
    int rc = awp_in3d_read(in3d, &p, err, sizeof err);
    printf("in3d_read %d %s\n", rc, rc ? err : "");
    if (rc) return 0;
    rc = awp_in3d_validate(&p, err, sizeof err);
    printf("in3d_validate %d %s\n", rc, rc ? err : "");
    if (rc) return 0;
    int *xyz = malloc(sizeof(int) * 3 * (size_t)p.nsrc);
    float *stf = malloc(sizeof(float) * 6 * (size_t)p.nsrc * (size_t)p.nst);
    rc = awp_source_read(p.insrc, p.nsrc, p.nst, xyz, stf, err, sizeof err);
    printf("source_read %d %s\n", rc, rc ? err : "");
    free(xyz);
    free(stf);
    int npts = 0;
    rc = awp_cordfile_count(p.insgt, &npts, err, sizeof err);
    printf("cordfile_count %d %d %s\n", rc, npts, rc ? err : "");
    if (rc || npts < 0 || npts > 100000000) return 0;
    int *pts = malloc(sizeof(int) * 3 * (size_t)(npts ? npts : 1));
    rc = awp_cordfile_read(p.insgt, npts, pts, err, sizeof err);
    printf("cordfile_read %d %s\n", rc, rc ? err : "");
    if (!rc) {
        int *keep = malloc(sizeof(int) * (size_t)(npts ? npts : 1));
        int *ijk = malloc(sizeof(int) * 3 * (size_t)(npts ? npts : 1));
        rc = awp_sgt_select(&p, npts, pts, keep, ijk, err, sizeof err);
        printf("sgt_select %d %s\n", rc, rc < 0 ? err : "");
        free(keep);
        free(ijk);
    }
    free(pts);
    int nt_out = 0, nflush = 0;
    rc = awp_sgt_file_geometry(&p, &nt_out, &nflush);
    printf("geometry %d %d %d\n", rc, nt_out, nflush);
    return 0;
}

static int head(char **a)
{
    awp_head_params hp;
    memset(&hp, 0, sizeof hp);
    hp.modelbox = a[0];
    hp.coordfile = a[1];
    hp.fdloc = a[2];
    hp.gridout = a[3];
    hp.media = a[4];
    hp.out = a[5];
    hp.spacing = 0.1f;
    hp.nt = 2000;
    hp.dt = 0.005f;
    hp.sgt_tinc = 10;
    hp.component = 'x';
    hp.moment = 1e20f;
    hp.flo = 1.0f;
    hp.mu_corr = 1;
    int rc = awp_write_sgt_head(&hp, err, sizeof err);
    printf("write_sgt_head %d %s\n", rc, rc ? err : "");
    return 0;
}

static int sgt16(const char *path, int npts, int nt_out)
{
    int n = 0, ns = 0;
    int rc = awp_sgt16_read(path, &n, &ns, NULL, 0, err, sizeof err);
    printf("sgt16_read %d %d %d %s\n", rc, n, ns, rc ? err : "");
    if (rc || n <= 0 || ns <= 0 || n > 1000000 || ns > 100000) return 0;
    float *out = malloc(sizeof(float) * (size_t)n * 6 * (size_t)ns);
    rc = awp_sgt16_read(path, &n, &ns, out, (size_t)n * 6 * ns, err, sizeof err);
    printf("sgt16_read_all %d %s\n", rc, rc ? err : "");
    free(out);
    const int cnt = npts < n ? npts : n;
    float *o2 = malloc(sizeof(float) * (size_t)(cnt ? cnt : 1) * 6 * (size_t)nt_out);
    for (int layout = 0; layout < 3; ++layout) {
        rc = awp_sgt16_read_points(path, 0, cnt, layout, o2, (size_t)cnt * 6 * nt_out, err, sizeof err);
        printf("sgt16_read_points %d %d %s\n", layout, rc, rc ? err : "");
    }
    free(o2);
    return 0;
}

static int check(const char *path)
{
    awp_sgt_scan_result r;
    int rc = awp_sgt_scan(path, 24000, &r, err, sizeof err);
    printf("scan %d %d %lld\n", rc, r.status, r.nvalues);
    char hex[33];
    rc = awp_md5_file(path, hex, err, sizeof err);
    printf("md5 %d %s\n", rc, rc ? err : hex);
    awp_sgt_cmp c;
    rc = awp_sgt_compare(path, path, 1, 10, 1e-5, &c, err, sizeof err);
    printf("compare %d %s\n", rc, rc ? err : "");
    return 0;
}

int main(int argc, char **argv)
{
    if (argc >= 3 && !strcmp(argv[1], "run")) return run(argv[2]);
    if (argc >= 8 && !strcmp(argv[1], "head")) return head(argv + 2);
    if (argc >= 5 && !strcmp(argv[1], "sgt16")) return sgt16(argv[2], atoi(argv[3]), atoi(argv[4]));
    if (argc >= 3 && !strcmp(argv[1], "check")) return check(argv[2]);
    fprintf(stderr, "usage: harness run|head|sgt16|check ...\n");
    return 1;
}
