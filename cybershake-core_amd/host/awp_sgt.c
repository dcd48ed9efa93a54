/*
 * awp_sgt.c -- command-line acceptance checks of raw SGT files, over the
 * library calls of host/awp_sgtcheck.c:
 *
 *   awp_sgt scan <sgt file> [zero_run]          NaN / zero-run scan
 *        (SgtTest/src/check_for_nans.c: exit 0 clean, 2 NaN, 3 zero run)
 *   awp_sgt md5 <sgt file> [md5 file]           print the digest, or check
 *        it against an md5 file (CheckSgt/CheckSgt.py: exit 1 on mismatch)
 *   awp_sgt compare <ref> <test> <npts> <nt> [floor]
 *        statistics of AWP-HIP-SGT/utils/compare_sgts.c plus the worst
 *        per-series relative L2 error (floor default 1e-6)
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "awpsgt.h"

static int usage(const char *me)
{
    fprintf(stderr,
            "usage: %s scan <sgt file> [zero_run]\n"
            "       %s md5 <sgt file> [md5 file]\n"
            "       %s compare <reference file> <test file> <num pts> <timesteps> [floor]\n",
            me, me, me);
    return 1;
}

int main(int argc, char **argv)
{
    char err[512] = "";
    if (argc < 3) return usage(argv[0]);
    if (!strcmp(argv[1], "scan")) {
        const long long zr = argc > 3 ? atoll(argv[3]) : 24000;
        awp_sgt_scan_result r;
        if (awp_sgt_scan(argv[2], zr, &r, err, sizeof err)) {
            fprintf(stderr, "%s\n", err);
            return 1;
        }
        if (r.status == 2) {
            printf("NaN in %s at byte offset %lld\n", argv[2], r.offset);
            return 2;
        }
        if (r.status == 3) {
            printf("Found %lld consecutive values with an abs less than 1e-20, at byte offset %lld, index %lld.\n",
                   zr, r.offset, r.offset / 4);
            return 3;
        }
        printf("Read %lld floats.\n", r.nvalues);
        return 0;
    }
    if (!strcmp(argv[1], "md5")) {
        char hex[33];
        if (awp_md5_file(argv[2], hex, err, sizeof err)) {
            fprintf(stderr, "%s\n", err);
            return 1;
        }
        if (argc < 4) {
            printf("%s  %s\n", hex, argv[2]);
            return 0;
        }
        printf("New md5sum: %s\n", hex);
        const int rc = awp_md5_check(argv[2], argv[3], err, sizeof err);
        if (rc) {
            printf("%s\n", err);
            return 1;
        }
        return 0;
    }
    if (!strcmp(argv[1], "compare")) {
        if (argc < 6) return usage(argv[0]);
        const double fl = argc > 6 ? atof(argv[6]) : 1e-6;
        awp_sgt_cmp c;
        if (awp_sgt_compare(argv[2], argv[3], atoll(argv[4]), atoi(argv[5]), fl, &c, err, sizeof err)) {
            fprintf(stderr, "%s\n", err);
            return 1;
        }
        static const char *cut[AWP_SGT_CMP_BINS - 1] = {"0.0001", "0.001", "0.01", "0.1", "0.3",
                                                        "1", "3", "10", "100", "1000"};
        printf("Average diff = %e.\n", c.avg_diff);
        printf("Using percent cutoff of 1e-10, average percent diff = %f%%, average absolute percent diff = %f%%\n",
               c.avg_pct, c.avg_abs_pct);
        printf("Largest diff of %e at float index %lld.\n", c.max_diff, c.max_diff_index);
        printf("Largest percent diff of %f%% at float index %lld.\n", c.max_pct, c.max_pct_index);
        printf("Absolute percentage difference histogram (upper bin edges, %%):\n");
        for (int q = 0; q < AWP_SGT_CMP_BINS; ++q)
            printf("  < %-8s %12lld  %6.2f%%\n", q < AWP_SGT_CMP_BINS - 1 ? cut[q] : "inf", c.bins[q],
                   c.pct_count ? 100.0 * (double)c.bins[q] / (double)c.pct_count : 0.0);
        printf("Per-series relative L2: worst %.3e (series %lld = point %lld component %lld) over %lld series; "
               "%lld series below %.1e x the largest series norm (largest absolute L2 difference among them %.3e)\n",
               c.worst_rel_l2, c.worst_series, c.worst_series >= 0 ? c.worst_series / 6 : -1,
               c.worst_series >= 0 ? c.worst_series % 6 : -1, c.series_checked, c.series_below_floor, fl,
               c.worst_abs_below_floor);
        return 0;
    }
    return usage(argv[0]);
}
