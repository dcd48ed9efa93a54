/*
 * awp_io.c -- host-side formats of the AWP-ODC-SGT drop-in surface: the IN3D
 * parameter file, the IFAULT=0 text source, the SGT cordfile, the SGT point
 * selection rule and the raw SGT file geometry.  Plain C, no device code.
 */
#define _XOPEN_SOURCE 700
#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "awpsgt.h"

static int seterr(char *err, size_t errlen, int code, const char *fmt, ...)
{
    if (err && errlen) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(err, errlen, fmt, ap);
        va_end(ap);
    }
    return code;
}

/* ------------------------------------------------------------------------ */
/* IN3D (io.f:327-418)                                                       */
/* ------------------------------------------------------------------------ */

enum { K_INT, K_REAL, K_STR };

typedef struct {
    const char *label;
    int kind;
    size_t off;
} in3d_field;

#define FI(name, lab) {lab, K_INT, offsetof(awp_in3d, name)}
#define FR(name, lab) {lab, K_REAL, offsetof(awp_in3d, name)}
#define FS(name, lab) {lab, K_STR, offsetof(awp_in3d, name)}

/* The order of the READ statements of rddt3d, io.f:327-418. */
static const in3d_field IN3D_FIELDS[] = {
    FI(igreen, "IGREEN"), FR(tmax, "TMAX"), FR(dh, "DH"), FR(dt, "DT"), FI(npc, "NPC"),
    FI(nd, "ND"), FR(arbc, "ARBC"), FR(pht, "PHT"), FI(nsrc, "NSRC"), FI(nst, "NST"),
    FI(nx, "NX"), FI(ny, "NY"), FI(nz, "NZ"), FI(npx, "NPX"), FI(npy, "NPY"), FI(npz, "NPZ"),
    FI(ifault, "IFAULT"), FI(checkpoint, "CHECKPOINT"), FI(isfcvlm, "ISFCVLM"),
    FI(imd5, "IMD5"), FI(ivelocity, "IVELOCITY"), FI(mediarestart, "MEDIARESTART"),
    FI(nvar, "NVAR"), FI(iost, "IOST"), FI(partdeg, "PARTDEG"), FI(io_opt, "IO_OPT"),
    FI(perf_meas, "PERF_MEAS"), FI(idyna, "IDYNA"), FI(socalq, "SOCALQ"), FI(nve, "NVE"),
    FR(mu_s, "MU_S"), FR(mu_d, "MU_D"), FR(fl, "FL"), FR(fh, "FH"), FR(fp, "FP"),
    FI(read_step, "READ_STEP"), FI(write_step, "WRITE_STEP"), FI(write_step2, "WRITE_STEP2"),
    FI(nbgx, "NBGX"), FI(nedx, "NEDX"), FI(nskpx, "NSKPX"), FI(nbgy, "NBGY"),
    FI(nedy, "NEDY"), FI(nskpy, "NSKPY"), FI(nbgz, "NBGZ"), FI(nedz, "NEDZ"),
    FI(nskpz, "NSKPZ"), FI(ntiskp, "NTISKP"),
    FI(nbgx2, "NBGX2"), FI(nedx2, "NEDX2"), FI(nskpx2, "NSKPX2"), FI(nbgy2, "NBGY2"),
    FI(nedy2, "NEDY2"), FI(nskpy2, "NSKPY2"), FI(nbgz2, "NBGZ2"), FI(nedz2, "NEDZ2"),
    FI(nskpz2, "NSKPZ2"), FI(ntiskp2, "NTISKP2"), FI(ntiskp_sgt, "NTISKP_SGT"),
    FS(chkp, "CHKP"), FS(chkj, "CHKJ"), FS(insrc, "INSRC"), FS(invel, "INVEL"),
    FS(insgt, "INSGT"), FS(sxgro, "SXRGO"), FS(sygro, "SYRGO"), FS(szgro, "SZRGO"),
    FS(sxgro2, "SXRGO2"), FS(sygro2, "SYRGO2"), FS(szgro2, "SZRGO2"), FS(sgtgro, "SGTGRO"),
    FS(sgsn, "SGSN"),
};
#define N_IN3D ((int)(sizeof(IN3D_FIELDS) / sizeof(IN3D_FIELDS[0])))

/* First list-directed item of a record: a quoted string (doubled quotes
 * stand for one), or the characters up to a blank, comma or slash.  Returns
 * 0 for a record holding no item (blank line). */
static int first_item(const char *line, char *out, size_t outlen)
{
    const char *s = line;
    while (*s == ' ' || *s == '\t' || *s == '\r' || *s == '\n') s++;
    if (!*s) return 0;
    size_t n = 0;
    if (*s == '\'' || *s == '"') {
        char q = *s++;
        while (*s) {
            if (*s == q) {
                if (s[1] == q) { if (n + 1 < outlen) out[n++] = q; s += 2; continue; }
                break;
            }
            if (*s == '\n' || *s == '\r') break;
            if (n + 1 < outlen) out[n++] = *s;
            s++;
        }
    } else {
        while (*s && *s != ' ' && *s != '\t' && *s != ',' && *s != '/' && *s != '\r' && *s != '\n') {
            if (n + 1 < outlen) out[n++] = *s;
            s++;
        }
    }
    out[n] = 0;
    return 1;
}

static int parse_int(const char *tok, int *v)
{
    char *end;
    errno = 0;
    long x = strtol(tok, &end, 10);
    if (end == tok || *end || errno) return -1;
    *v = (int)x;
    return 0;
}

static int parse_real(const char *tok, float *v)
{
    char buf[128];
    size_t n = strlen(tok);
    if (n == 0 || n >= sizeof(buf)) return -1;
    for (size_t i = 0; i <= n; ++i) buf[i] = (tok[i] == 'd' || tok[i] == 'D') ? 'e' : tok[i];
    char *end;
    errno = 0;
    float x = strtof(buf, &end); /* correctly rounded, like Fortran's read */
    if (end == buf || *end) return -1;
    *v = x;
    return 0;
}

int awp_in3d_read(const char *path, awp_in3d *out, char *err, size_t errlen)
{
    if (!path || !out) return seterr(err, errlen, AWP_EINVAL, "null argument");
    FILE *f = fopen(path, "r");
    if (!f) return seterr(err, errlen, AWP_EIO, "cannot open IN3D '%s': %s", path, strerror(errno));
    memset(out, 0, sizeof(*out));
    char line[4096], tok[1024];
    int field = 0, lineno = 0;
    while (field < N_IN3D && fgets(line, sizeof(line), f)) {
        lineno++;
        if (!first_item(line, tok, sizeof(tok))) continue;
        const in3d_field *fd = &IN3D_FIELDS[field];
        char *dst = (char *)out + fd->off;
        if (fd->kind == K_INT) {
            if (parse_int(tok, (int *)dst)) {
                fclose(f);
                return seterr(err, errlen, AWP_EINVAL, "IN3D line %d: '%s' is not an integer (%s)", lineno, tok, fd->label);
            }
        } else if (fd->kind == K_REAL) {
            if (parse_real(tok, (float *)dst)) {
                fclose(f);
                return seterr(err, errlen, AWP_EINVAL, "IN3D line %d: '%s' is not a real (%s)", lineno, tok, fd->label);
            }
        } else {
            strncpy(dst, tok, 255);
            dst[255] = 0;
        }
        field++;
    }
    fclose(f);
    if (field < N_IN3D)
        return seterr(err, errlen, AWP_EINVAL, "IN3D '%s': %d records, expected %d (missing %s)",
                      path, field, N_IN3D, IN3D_FIELDS[field].label);
    /* nt = nint(tmax/dt) + 1, single precision (pmcl3d.f:287) */
    float q = out->tmax / out->dt;
    out->nt = (int)lroundf(q) + 1;
    return AWP_OK;
}

int awp_in3d_validate(const awp_in3d *p, char *err, size_t errlen)
{
    if (!p) return seterr(err, errlen, AWP_EINVAL, "null argument");
    if (p->igreen < 1 || p->igreen > 6)
        return seterr(err, errlen, AWP_EINVAL, "igreen=%d unsupported (SGT modes 1..6)", p->igreen);
    if (p->npc != 0) return seterr(err, errlen, AWP_EINVAL, "NPC=%d unsupported (only Cerjan, NPC=0)", p->npc);
    if (p->nve != 1) return seterr(err, errlen, AWP_EINVAL, "NVE=%d unsupported (only visco-elastic, NVE=1)", p->nve);
    if (p->ifault != 0) return seterr(err, errlen, AWP_EINVAL, "IFAULT=%d unsupported (only text point source, IFAULT=0)", p->ifault);
    if (p->mediarestart != 1 && p->mediarestart != 2)
        return seterr(err, errlen, AWP_EINVAL, "MEDIARESTART=%d unsupported (media file, 1 or 2)", p->mediarestart);
    if (p->nvar != 3 && p->nvar != 5 && p->nvar != 8)
        return seterr(err, errlen, AWP_EINVAL, "NVAR=%d unsupported (3, 5 or 8)", p->nvar);
    if (p->nx < 8 || p->ny < 8 || p->nz < 8)
        return seterr(err, errlen, AWP_EINVAL, "grid %dx%dx%d too small", p->nx, p->ny, p->nz);
    if (p->npx < 1 || p->npy < 1 || p->npz < 1 || p->nx % p->npx || p->ny % p->npy || p->nz % p->npz)
        return seterr(err, errlen, AWP_EINVAL, "NPX/NPY/NPZ %d/%d/%d must divide NX/NY/NZ", p->npx, p->npy, p->npz);
    if (p->nd < 1 || 2 * p->nd > p->nx || 2 * p->nd > p->ny || p->nd > p->nz)
        return seterr(err, errlen, AWP_EINVAL, "sponge ND=%d does not fit the grid", p->nd);
    if (!(p->arbc > 0.f && p->arbc < 1.f)) return seterr(err, errlen, AWP_EINVAL, "ARBC=%g outside (0,1)", p->arbc);
    if (!(p->dh > 0.f) || !(p->dt > 0.f)) return seterr(err, errlen, AWP_EINVAL, "DH and DT must be positive");
    if (!(p->fl > 0.f && p->fh > 0.f && p->fp > 0.f)) return seterr(err, errlen, AWP_EINVAL, "FL/FH/FP must be positive");
    if (p->nsrc < 1 || p->nst < 1) return seterr(err, errlen, AWP_EINVAL, "NSRC and NST must be >= 1");
    if (p->nt < 1) return seterr(err, errlen, AWP_EINVAL, "TMAX/DT gives no time step");
    if (p->ntiskp_sgt < 1 || p->write_step < 1)
        return seterr(err, errlen, AWP_EINVAL, "NTISKP_SGT and WRITE_STEP must be >= 1");
    return AWP_OK;
}

/* ------------------------------------------------------------------------ */
/* source (source.f:170-198, IFAULT=0)                                       */
/* ------------------------------------------------------------------------ */

static int read_numbers(FILE *f, int want, double *vals, float *fvals, int *lineno)
{
    char line[4096];
    while (fgets(line, sizeof(line), f)) {
        (*lineno)++;
        char *s = line;
        int got = 0;
        while (got < want) {
            while (*s == ' ' || *s == '\t' || *s == ',' || *s == '\r' || *s == '\n') s++;
            if (!*s) break;
            char tok[128];
            size_t n = 0;
            while (*s && *s != ' ' && *s != '\t' && *s != ',' && *s != '\r' && *s != '\n') {
                if (n + 1 < sizeof(tok)) tok[n++] = (*s == 'd' || *s == 'D') ? 'e' : *s;
                s++;
            }
            tok[n] = 0;
            char *end;
            double v = strtod(tok, &end);
            if (end == tok || *end) return -2;
            /* strtof rounds the decimal once, as Fortran's single-precision
             * list-directed read does */
            if (fvals) fvals[got] = strtof(tok, NULL);
            vals[got++] = v;
        }
        if (got == 0) continue;      /* blank record */
        if (got < want) return -2;
        return 0;
    }
    return -1;
}

int awp_source_read(const char *path, int nsrc, int nst, int *xyz, float *stf, char *err, size_t errlen)
{
    if (!path || nsrc < 1 || nst < 1 || !xyz || !stf) return seterr(err, errlen, AWP_EINVAL, "bad arguments");
    FILE *f = fopen(path, "r");
    if (!f) return seterr(err, errlen, AWP_EIO, "cannot open source '%s': %s", path, strerror(errno));
    int lineno = 0;
    double v[6];
    float fv[6];
    for (int s = 0; s < nsrc; ++s) {
        int rc = read_numbers(f, 3, v, NULL, &lineno);
        if (rc) { fclose(f); return seterr(err, errlen, AWP_EINVAL, "source '%s' line %d: expected 'x y z'", path, lineno + (rc == -1)); }
        for (int c = 0; c < 3; ++c) {
            if (v[c] != floor(v[c])) { fclose(f); return seterr(err, errlen, AWP_EINVAL, "source line %d: non-integer coordinate", lineno); }
            xyz[3 * s + c] = (int)v[c];
        }
        for (int t = 0; t < nst; ++t) {
            rc = read_numbers(f, 6, v, fv, &lineno);
            if (rc) { fclose(f); return seterr(err, errlen, AWP_EINVAL, "source '%s': sample %d of source %d missing or malformed (line %d)", path, t + 1, s + 1, lineno); }
            for (int c = 0; c < 6; ++c) stf[((size_t)s * nst + t) * 6 + c] = fv[c];
        }
    }
    fclose(f);
    return AWP_OK;
}

/* ------------------------------------------------------------------------ */
/* cordfile (io.f:419-442)                                                   */
/* ------------------------------------------------------------------------ */

static FILE *cord_open(const char *path, int *n, char *err, size_t errlen, int *rc)
{
    FILE *f = fopen(path, "r");
    if (!f) { *rc = seterr(err, errlen, AWP_EIO, "cannot open cordfile '%s': %s", path, strerror(errno)); return NULL; }
    char line[4096];
    while (fgets(line, sizeof(line), f)) {
        const char *s = line;
        while (*s == ' ' || *s == '\t') s++;
        if (*s == '#') continue;          /* io.f:425-429 */
        if (*s == '\n' || *s == '\r' || !*s) continue;
        char *end;
        long v = strtol(s, &end, 10);
        if (end == s || v < 0) { fclose(f); *rc = seterr(err, errlen, AWP_EINVAL, "cordfile '%s': bad point count", path); return NULL; }
        *n = (int)v;
        *rc = 0;
        return f;
    }
    fclose(f);
    *rc = seterr(err, errlen, AWP_EINVAL, "cordfile '%s': no point count", path);
    return NULL;
}

int awp_cordfile_count(const char *path, int *npts, char *err, size_t errlen)
{
    int rc;
    FILE *f = cord_open(path, npts, err, errlen, &rc);
    if (!f) return rc;
    fclose(f);
    return AWP_OK;
}

int awp_cordfile_read(const char *path, int npts, int *xyz, char *err, size_t errlen)
{
    int n, rc;
    FILE *f = cord_open(path, &n, err, errlen, &rc);
    if (!f) return rc;
    if (n != npts) { fclose(f); return seterr(err, errlen, AWP_EINVAL, "cordfile count %d != %d", n, npts); }
    int lineno = 0;
    double v[3];
    for (int i = 0; i < n; ++i) {
        rc = read_numbers(f, 3, v, NULL, &lineno);
        if (rc) { fclose(f); return seterr(err, errlen, AWP_EINVAL, "cordfile '%s': point %d missing or malformed", path, i + 1); }
        for (int c = 0; c < 3; ++c) xyz[3 * i + c] = (int)v[c];
    }
    fclose(f);
    return AWP_OK;
}

/* ------------------------------------------------------------------------ */
/* SGT point selection (station.f:31-80)                                     */
/* ------------------------------------------------------------------------ */

/* Local index rule of station.f:44-54: a = mod(v, n); a = 0 maps to n. */
static int local_index(int v, int n)
{
    int a = v % n;
    return a ? a : n;
}

int awp_sgt_select(const awp_in3d *p, int npts, const int *xyz, int *keep, int *ijk, char *err, size_t errlen)
{
    if (!p || npts < 0 || (npts && (!xyz || !keep || !ijk))) return seterr(err, errlen, AWP_EINVAL, "bad arguments");
    const int nxt = p->nx / p->npx, nyt = p->ny / p->npy, nzt = p->nz / p->npz;
    int kept = 0;
    for (int n = 0; n < npts; ++n) {
        const int x = xyz[3 * n], y = xyz[3 * n + 1], z = xyz[3 * n + 2];
        if (x < 0 || y < 0 || z > p->nz + 1)
            return seterr(err, errlen, AWP_EINVAL,
                          "cordfile point %d (%d %d %d) would index outside the reference's arrays", n + 1, x, y, z);
        /* Fortran int() truncates toward zero, as C division does */
        const int nzp = (p->nz - z) / nzt;
        const int nyp = (y - 1) / nyt;
        const int nxp = (x - 1) / nxt;
        keep[n] = 0;
        ijk[3 * n] = ijk[3 * n + 1] = ijk[3 * n + 2] = 0;
        if (nzp < 0 || nzp >= p->npz || nyp < 0 || nyp >= p->npy || nxp < 0 || nxp >= p->npx) continue;
        keep[n] = 1;
        ijk[3 * n] = nxp * nxt + local_index(x, nxt);
        ijk[3 * n + 1] = nyp * nyt + local_index(y, nyt);
        ijk[3 * n + 2] = nzp * nzt + local_index(p->nz - z + 1, nzt);
        kept++;
    }
    return kept;
}

/* ------------------------------------------------------------------------ */
/* compact FP16 SGT file (format in awpsgt.h)                                */
/* ------------------------------------------------------------------------ */

static float half_to_float(uint16_t h)
{
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t ex = (h >> 10) & 0x1Fu, man = h & 0x3FFu;
    float v;
    if (ex == 0)
        v = ldexpf((float)man, -24);               /* zero and subnormals */
    else if (ex == 31)
        v = man ? NAN : INFINITY;
    else
        v = ldexpf((float)(man | 0x400u), (int)ex - 25);
    return sign ? -v : v;
}

int awp_sgt16_read(const char *path, int *npts, int *nsamp, float *out, size_t count, char *err, size_t errlen)
{
    FILE *f = fopen(path, "rb");
    if (!f) return seterr(err, errlen, AWP_EIO, "cannot open '%s': %s", path, strerror(errno));
    unsigned char hdr[64];
    uint32_t h32[4];
    if (fread(hdr, 1, 64, f) != 64 || memcmp(hdr, "AWPSGT16", 8) != 0) {
        fclose(f);
        return seterr(err, errlen, AWP_EINVAL, "'%s' is not an FP16 SGT file", path);
    }
    memcpy(h32, hdr + 8, sizeof(h32));
    if (h32[0] != 1u || h32[2] != 6u) {
        fclose(f);
        return seterr(err, errlen, AWP_EINVAL, "'%s': unsupported version or component count", path);
    }
    const size_t np = h32[1], ns = h32[3];
    if (npts) *npts = (int)np;
    if (nsamp) *nsamp = (int)ns;
    if (!out) {
        fclose(f);
        return AWP_OK;
    }
    if (count < np * 6 * ns) {
        fclose(f);
        return seterr(err, errlen, AWP_EINVAL, "buffer holds %zu values, file has %zu", count, np * 6 * ns);
    }
    float *scale = (float *)malloc(sizeof(float) * 6 * (np ? np : 1));
    uint16_t *row = (uint16_t *)malloc(sizeof(uint16_t) * (ns ? ns : 1));
    int rc = AWP_OK;
    if (!scale || !row || fread(scale, sizeof(float), 6 * np, f) != 6 * np) {
        rc = seterr(err, errlen, AWP_EIO, "'%s': truncated", path);
    } else {
        for (size_t s = 0; s < 6 * np && rc == AWP_OK; ++s) {
            if (ns && fread(row, sizeof(uint16_t), ns, f) != ns) {
                rc = seterr(err, errlen, AWP_EIO, "'%s': truncated", path);
                break;
            }
            for (size_t t = 0; t < ns; ++t) out[s * ns + t] = half_to_float(row[t]) * scale[s];
        }
    }
    free(scale);
    free(row);
    fclose(f);
    return rc;
}

int awp_sgt16_read_points(const char *path, int first, int count, int layout, float *out, size_t n, char *err,
                          size_t errlen)
{
    if (!path || first < 0 || count < 0 || layout < 0 || layout > 2 || (count > 0 && !out))
        return seterr(err, errlen, AWP_EINVAL, "bad arguments");
    FILE *f = fopen(path, "rb");
    if (!f) return seterr(err, errlen, AWP_EIO, "cannot open '%s': %s", path, strerror(errno));
    unsigned char hdr[64];
    uint32_t h32[4], nto = 0;
    if (fread(hdr, 1, 64, f) != 64 || memcmp(hdr, "AWPSGT16", 8) != 0) {
        fclose(f);
        return seterr(err, errlen, AWP_EINVAL, "'%s' is not an FP16 SGT file", path);
    }
    memcpy(h32, hdr + 8, sizeof(h32));
    memcpy(&nto, hdr + 32, 4);
    const size_t np = h32[1], ns = h32[3];
    const size_t nt = nto ? nto : ns;   /* series length of the raw file */
    if (h32[0] != 1u || h32[2] != 6u || nt < ns) {
        fclose(f);
        return seterr(err, errlen, AWP_EINVAL, "'%s': unsupported version or geometry", path);
    }
    if ((size_t)first + (size_t)count > np) {
        fclose(f);
        return seterr(err, errlen, AWP_EINVAL, "points %d..%d beyond the file's %zu", first, first + count - 1, np);
    }
    if (n < (size_t)count * 6 * nt) {
        fclose(f);
        return seterr(err, errlen, AWP_EINVAL, "buffer holds %zu values, need %zu", n, (size_t)count * 6 * nt);
    }
    float sc[6];
    uint16_t *row = (uint16_t *)malloc(sizeof(uint16_t) * 6 * (ns ? ns : 1));
    float *v = (float *)malloc(sizeof(float) * 6 * nt);
    int rc = row && v ? AWP_OK : seterr(err, errlen, AWP_ENOMEM, "out of host memory");
    /* input component of each output component: AWP order as is, or the
     * RWG order of SgtHead/src/reformat_awp.c:46-61 */
    static const int from_rwg[6] = {1, 0, 2, 3, 5, 4};
    for (int k = 0; k < count && rc == AWP_OK; ++k) {
        const size_t g = (size_t)first + k;
        if (fseeko(f, (off_t)(64 + g * 6 * 4), SEEK_SET) || fread(sc, sizeof(float), 6, f) != 6 ||
            fseeko(f, (off_t)(64 + np * 6 * 4 + g * 6 * ns * 2), SEEK_SET) ||
            (ns && fread(row, sizeof(uint16_t), 6 * ns, f) != 6 * ns)) {
            rc = seterr(err, errlen, AWP_EIO, "'%s': truncated", path);
            break;
        }
        for (int c = 0; c < 6; ++c)
            for (size_t t = 0; t < nt; ++t) {
                float x = t < ns ? half_to_float(row[c * ns + t]) * sc[c] : 0.0f;
                if (layout == 2) x = (float)(-2.0 * x);   /* reformat_awp.c:40-44 */
                v[c * nt + t] = x;
            }
        float *o = out + (size_t)k * 6 * nt;
        for (int c = 0; c < 6; ++c) {
            const int ci = layout ? from_rwg[c] : c;
            const int neg = layout && c >= 4;
            for (size_t t = 0; t < nt; ++t) o[c * nt + t] = neg ? (float)(-1.0 * v[ci * nt + t]) : v[ci * nt + t];
        }
    }
    free(row);
    free(v);
    fclose(f);
    return rc;
}

int awp_sgt_file_geometry(const awp_in3d *p, int *nt_out, int *nflush)
{
    if (!p || p->write_step < 1 || p->ntiskp_sgt < 1) return AWP_EINVAL;
    /* NT_OUT = int(int(nt/WS)*WS/NTISKP_SGT) (pmcl3d.f:1591-1592) */
    if (nt_out) *nt_out = (int)((long long)(p->nt / p->write_step) * p->write_step / p->ntiskp_sgt);
    /* buffers fill every WS samples; a partial last buffer is never written */
    if (nflush) *nflush = (p->nt / p->ntiskp_sgt) / p->write_step;
    return AWP_OK;
}
