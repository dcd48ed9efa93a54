/*
 * awp_sgtcheck.c -- acceptance checks of a raw SGT file, host only.
 *
 * The CyberShake workflow accepts an SGT file after three checks, restated
 * here as library calls (and as the awp_sgt command, host/awp_sgt.c):
 *
 *  - NaN / zero-run scan (SgtTest/src/check_for_nans.c:19-56): the file is
 *    read as float32 in order; the first NaN fails the file (result 2, with
 *    its byte offset); a run of `zero_run` consecutive values with
 *    |v| < 1e-20 fails it (result 3; 24000 in the reference, "for 1 Hz").
 *  - md5 (CheckSgt/CheckSgt.py:44-78): the MD5 of the whole file against the
 *    first 32 characters of the first line of an md5 file.  The digest is
 *    RFC 1321's (the same bytes hashlib.md5 gives), computed here.
 *  - comparison of two raw files (AWP-HIP-SGT/utils/compare_sgts.c:7-119):
 *    the reference tool's statistics (mean difference, largest signed
 *    difference, percent differences of the values whose reference value is
 *    above 1e-10 and their histogram over the bins 1e-4 .. 1000 percent),
 *    plus, per (point, component) time series, the relative L2 error with a
 *    stated floor: series whose reference norm is below floor x the largest
 *    series norm are counted apart (their relative error is noise).
 *
 * No device work; every function returns 0 or a negative AWP_E* code.
 */
#define _FILE_OFFSET_BITS 64
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "awpsgt.h"

static void seterr(char *err, size_t n, const char *fmt, const char *arg)
{
    if (err && n) snprintf(err, n, fmt, arg);
}

/* ---------------------------------------------------------------------- */
/* NaN / zero-run scan                                                      */
/* ---------------------------------------------------------------------- */
int awp_sgt_scan(const char *path, long long zero_run, awp_sgt_scan_result *out, char *err, size_t errlen)
{
    if (!path || !out || zero_run < 1) {
        seterr(err, errlen, "%s", "awp_sgt_scan: bad argument");
        return AWP_EINVAL;
    }
    memset(out, 0, sizeof(*out));
    out->offset = -1;
    FILE *f = fopen(path, "rb");
    if (!f) {
        seterr(err, errlen, "cannot open '%s'", path);
        return AWP_EIO;
    }
    const size_t chunk = (size_t)1 << 22;   /* floats per read */
    float *buf = (float *)malloc(chunk * sizeof(float));
    if (!buf) {
        fclose(f);
        seterr(err, errlen, "%s", "out of host memory");
        return AWP_ENOMEM;
    }
    long long run = 0, pos = 0;
    size_t n;
    while (out->status == 0 && (n = fread(buf, sizeof(float), chunk, f)) > 0) {
        for (size_t i = 0; i < n; ++i) {
            const float v = buf[i];
            if (isnan(v)) {
                out->status = 2;
                out->offset = (pos + (long long)i) * 4;
                break;
            }
            run = fabsf(v) < 1e-20f ? run + 1 : 0;
            if (run >= zero_run) {
                out->status = 3;
                out->offset = (pos + (long long)i) * 4;
                break;
            }
        }
        pos += (long long)n;
    }
    const int bad = ferror(f);
    free(buf);
    fclose(f);
    if (bad) {
        seterr(err, errlen, "read error in '%s'", path);
        return AWP_EIO;
    }
    out->nvalues = out->status ? out->offset / 4 + 1 : pos;
    return AWP_OK;
}

/* ---------------------------------------------------------------------- */
/* MD5 (RFC 1321)                                                           */
/* ---------------------------------------------------------------------- */
typedef struct {
    uint32_t h[4];
    uint64_t len;          /* bytes hashed */
    unsigned char blk[64];
    size_t nblk;
} md5_state;

static uint32_t md5_t[64];
static int md5_t_ready;

/* T[i] = floor(2^32 * |sin(i + 1)|) (RFC 1321 section 3.4) */
static void md5_table(void)
{
    if (md5_t_ready) return;
    for (int i = 0; i < 64; ++i) md5_t[i] = (uint32_t)(fabs(sin((double)(i + 1))) * 4294967296.0);
    md5_t_ready = 1;
}

static uint32_t rol(uint32_t x, int s) { return (x << s) | (x >> (32 - s)); }

static void md5_block(md5_state *st, const unsigned char *p)
{
    static const int sh[4][4] = {{7, 12, 17, 22}, {5, 9, 14, 20}, {4, 11, 16, 23}, {6, 10, 15, 21}};
    uint32_t x[16];
    for (int i = 0; i < 16; ++i)
        x[i] = (uint32_t)p[4 * i] | ((uint32_t)p[4 * i + 1] << 8) | ((uint32_t)p[4 * i + 2] << 16) |
               ((uint32_t)p[4 * i + 3] << 24);
    uint32_t a = st->h[0], b = st->h[1], c = st->h[2], d = st->h[3];
    for (int i = 0; i < 64; ++i) {
        const int r = i / 16;
        uint32_t fn;
        int k;
        switch (r) {
        case 0: fn = (b & c) | (~b & d); k = i; break;
        case 1: fn = (b & d) | (c & ~d); k = (5 * i + 1) & 15; break;
        case 2: fn = b ^ c ^ d; k = (3 * i + 5) & 15; break;
        default: fn = c ^ (b | ~d); k = (7 * i) & 15; break;
        }
        const uint32_t t = d;
        d = c;
        c = b;
        b = b + rol(a + fn + md5_t[i] + x[k], sh[r][i & 3]);
        a = t;
    }
    st->h[0] += a;
    st->h[1] += b;
    st->h[2] += c;
    st->h[3] += d;
}

static void md5_init(md5_state *st)
{
    md5_table();
    st->h[0] = 0x67452301u;
    st->h[1] = 0xefcdab89u;
    st->h[2] = 0x98badcfeu;
    st->h[3] = 0x10325476u;
    st->len = 0;
    st->nblk = 0;
}

static void md5_update(md5_state *st, const unsigned char *p, size_t n)
{
    st->len += n;
    if (st->nblk) {
        const size_t take = n < 64 - st->nblk ? n : 64 - st->nblk;
        memcpy(st->blk + st->nblk, p, take);
        st->nblk += take;
        p += take;
        n -= take;
        if (st->nblk == 64) {
            md5_block(st, st->blk);
            st->nblk = 0;
        }
    }
    for (; n >= 64; p += 64, n -= 64) md5_block(st, p);
    if (n) {
        memcpy(st->blk, p, n);
        st->nblk = n;
    }
}

static void md5_final(md5_state *st, char hex[33])
{
    const uint64_t bits = st->len * 8;
    unsigned char pad[72];
    const size_t np = (st->nblk < 56 ? 56 : 120) - st->nblk;
    memset(pad, 0, sizeof pad);
    pad[0] = 0x80;
    unsigned char lenb[8];
    for (int i = 0; i < 8; ++i) lenb[i] = (unsigned char)(bits >> (8 * i));
    md5_update(st, pad, np);
    md5_update(st, lenb, 8);
    for (int w = 0; w < 4; ++w)
        for (int i = 0; i < 4; ++i) sprintf(hex + 8 * w + 2 * i, "%02x", (unsigned)((st->h[w] >> (8 * i)) & 0xff));
    hex[32] = 0;
}

int awp_md5_bytes(const void *data, size_t n, char hex[33])
{
    if ((!data && n) || !hex) return AWP_EINVAL;
    md5_state st;
    md5_init(&st);
    md5_update(&st, (const unsigned char *)data, n);
    md5_final(&st, hex);
    return AWP_OK;
}

int awp_md5_file(const char *path, char hex[33], char *err, size_t errlen)
{
    if (!path || !hex) return AWP_EINVAL;
    FILE *f = fopen(path, "rb");
    if (!f) {
        seterr(err, errlen, "cannot open '%s'", path);
        return AWP_EIO;
    }
    const size_t chunk = (size_t)1 << 24;
    unsigned char *buf = (unsigned char *)malloc(chunk);
    if (!buf) {
        fclose(f);
        return AWP_ENOMEM;
    }
    md5_state st;
    md5_init(&st);
    size_t n;
    while ((n = fread(buf, 1, chunk, f)) > 0) md5_update(&st, buf, n);
    const int bad = ferror(f);
    free(buf);
    fclose(f);
    if (bad) {
        seterr(err, errlen, "read error in '%s'", path);
        return AWP_EIO;
    }
    md5_final(&st, hex);
    return AWP_OK;
}

int awp_md5_check(const char *path, const char *md5path, char *err, size_t errlen)
{
    if (!path || !md5path) return AWP_EINVAL;
    FILE *f = fopen(md5path, "r");
    if (!f) {
        seterr(err, errlen, "cannot open '%s'", md5path);
        return AWP_EIO;
    }
    char line[256];
    const int got = fgets(line, sizeof line, f) != NULL;
    fclose(f);
    /* the first word of the first line ("<md5>  <name>" as md5sum writes) */
    size_t w = 0;
    while (got && line[w] && line[w] != ' ' && line[w] != '\t' && line[w] != '\n' && line[w] != '\r') ++w;
    if (!got || w != 32) {
        seterr(err, errlen, "invalid md5sum in '%s'", md5path);
        return AWP_EINVAL;
    }
    line[32] = 0;
    char hex[33];
    const int rc = awp_md5_file(path, hex, err, errlen);
    if (rc) return rc;
    for (int i = 0; i < 32; ++i)
        if ((line[i] | 0x20) != hex[i]) {
            seterr(err, errlen, "md5 checksums do not match for '%s'", path);
            return 1;
        }
    return AWP_OK;
}

/* ---------------------------------------------------------------------- */
/* comparison of two raw files                                              */
/* ---------------------------------------------------------------------- */
static const float pct_cut[AWP_SGT_CMP_BINS - 1] = {0.0001f, 0.001f, 0.01f, 0.1f, 0.3f, 1.0f, 3.0f, 10.0f, 100.0f,
                                                    1000.0f};

int awp_sgt_compare(const char *ref_path, const char *test_path, long long npts, int nt, double floor_frac,
                    awp_sgt_cmp *out, char *err, size_t errlen)
{
    if (!ref_path || !test_path || !out || npts < 0 || nt < 1 || floor_frac < 0) {
        seterr(err, errlen, "%s", "awp_sgt_compare: bad argument");
        return AWP_EINVAL;
    }
    memset(out, 0, sizeof(*out));
    out->max_diff_index = -1;
    out->max_pct_index = -1;
    out->worst_series = -1;
    FILE *fr = fopen(ref_path, "rb");
    if (!fr) {
        seterr(err, errlen, "cannot open '%s'", ref_path);
        return AWP_EIO;
    }
    FILE *ft = fopen(test_path, "rb");
    if (!ft) {
        fclose(fr);
        seterr(err, errlen, "cannot open '%s'", test_path);
        return AWP_EIO;
    }
    const long long nser = npts * 6;
    double *nref = (double *)calloc((size_t)(nser > 0 ? nser : 1), sizeof(double));
    double *ndif = (double *)calloc((size_t)(nser > 0 ? nser : 1), sizeof(double));
    float *a = (float *)malloc((size_t)nt * sizeof(float));
    float *b = (float *)malloc((size_t)nt * sizeof(float));
    int rc = AWP_OK;
    if (!nref || !ndif || !a || !b) rc = AWP_ENOMEM;
    double sum_diff = 0, sum_pct = 0, sum_abs_pct = 0;
    for (long long s = 0; !rc && s < nser; ++s) {
        if (fread(a, sizeof(float), (size_t)nt, fr) != (size_t)nt || fread(b, sizeof(float), (size_t)nt, ft) != (size_t)nt) {
            seterr(err, errlen, "%s", "a file is shorter than npts x 6 x nt floats");
            rc = AWP_EIO;
            break;
        }
        double sr = 0, sd = 0;
        for (int t = 0; t < nt; ++t) {
            const long long idx = s * nt + t;
            const float diff = b[t] - a[t];
            sum_diff += diff;
            if (diff > out->max_diff) {
                out->max_diff = diff;
                out->max_diff_index = idx;
            }
            if (fabsf(diff) > out->max_abs_diff) out->max_abs_diff = fabsf(diff);
            if (a[t] > 1e-10f) {
                const float pct = diff / a[t] * 100.0f;
                const float apct = fabsf(pct);
                sum_pct += pct;
                sum_abs_pct += apct;
                out->pct_count++;
                if (apct > out->max_pct) {
                    out->max_pct = apct;
                    out->max_pct_index = idx;
                }
                int bin = AWP_SGT_CMP_BINS - 1;
                for (int q = 0; q < AWP_SGT_CMP_BINS - 1; ++q)
                    if (apct < pct_cut[q]) {
                        bin = q;
                        break;
                    }
                out->bins[bin]++;
            }
            sr += (double)a[t] * a[t];
            sd += (double)diff * diff;
        }
        nref[s] = sqrt(sr);
        ndif[s] = sqrt(sd);
    }
    if (!rc) {
        const double nval = (double)nser * nt;
        out->nvalues = (long long)nval;
        out->avg_diff = nval > 0 ? sum_diff / nval : 0;
        out->avg_pct = out->pct_count ? sum_pct / (double)out->pct_count : 0;
        out->avg_abs_pct = out->pct_count ? sum_abs_pct / (double)out->pct_count : 0;
        double mx = 0;
        for (long long s = 0; s < nser; ++s) mx = nref[s] > mx ? nref[s] : mx;
        out->max_series_norm = mx;
        const double fl = floor_frac * mx;
        for (long long s = 0; s < nser; ++s) {
            if (nref[s] <= 0 || nref[s] < fl) {
                out->series_below_floor++;
                if (ndif[s] > out->worst_abs_below_floor) out->worst_abs_below_floor = ndif[s];
                continue;
            }
            out->series_checked++;
            const double e = ndif[s] / nref[s];
            if (e > out->worst_rel_l2 || out->worst_series < 0) {
                out->worst_rel_l2 = e;
                out->worst_series = s;
            }
        }
    }
    free(nref);
    free(ndif);
    free(a);
    free(b);
    fclose(fr);
    fclose(ft);
    return rc;
}
