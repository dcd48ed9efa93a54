/*
 * awp_head.c -- the SGT header file of an AWP run, as SgtHead's write_head
 * makes it (SgtHead/src/write_head.c:51-469): a master record, then an index
 * record and a header record per SGT point, with the site and point
 * positions on RWG's great-circle model projection (head_subs.c:3-126 and
 * the geoproj = 1 branch of init_modelc, :230-256) and the elastic moduli
 * and density at each point from the AWP-format media file.  Plain C, no
 * device code.
 *
 * The downstream tools compare headers byte for byte, so every expression
 * keeps the types the reference evaluates it in: a float / float is a
 * single-precision division even where its result meets a double, and a
 * value the reference stores in a float is rounded to float before its next
 * use.
 */
#define _XOPEN_SOURCE 700
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/types.h>

#include "awpsgt.h"

/* Records of the header file (SgtHead/src/structure.h:164-221): the layout
 * is the file format. */
typedef struct {
    int geoproj;
    float modellon, modellat, modelrot, xshift, yshift;
    int globnp, localnp, nt;
} head_master;
typedef struct {
    long long indx;
    int xsgt, ysgt, zsgt;
    float h;
} head_index;
typedef struct {
    long long indx;
    int geoproj;
    float modellon, modellat, modelrot, xshift, yshift;
    int nt;
    float xazim, dt, tst, h, src_lat, src_lon, src_dep;
    int xsrc, ysrc, zsrc;
    float sgt_lat, sgt_lon, sgt_dep;
    int xsgt, ysgt, zsgt;
    float cdist, lam, mu, rho, xmom, ymom, zmom;
} head_record;
_Static_assert(sizeof(head_master) == 36, "sgtmaster is 36 bytes");
_Static_assert(sizeof(head_index) == 24, "sgtindex is 24 bytes");
_Static_assert(sizeof(head_record) == 128, "sgtheader is 128 bytes");

/* RWG's constants (SgtHead/src/defs.h:59-63) */
#define RAD_PER_DEG 0.01745329
#define EARTH_RADIUS_KM 6378.139f

static int fail(char *err, size_t errlen, int code, const char *fmt, ...)
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
/* Great-circle model frame                                                 */
/* ------------------------------------------------------------------------ */

/* The model box is a patch of the sphere: model x and y are arc lengths
 * along two orthogonal great circles through the box centre.  `rot` turns a
 * point of the model frame (unit sphere, pole at the box centre) into
 * geographic axes: rotate by the model angle about the local vertical, tilt
 * by the colatitude of the centre, rotate by its longitude. */
typedef struct {
    double rot[9];
    double g0, b0;      /* half the box extents in y and x, as angles */
    float erad;
    float xshift, yshift;
} gc_frame;

static void gc_frame_init(gc_frame *f, float lon, float lat, float angle, int nx, int ny, float h)
{
    const float xlen = nx * h, ylen = ny * h;
    const double a = (double)angle * RAD_PER_DEG;
    const double t = (double)(90.0 - lat) * RAD_PER_DEG;
    const double p = (double)lon * RAD_PER_DEG;
    const double ca = cos(a), sa = sin(a), ct = cos(t), st = sin(t), cp = cos(p), sp = sin(p);
    double *r = f->rot;
    r[0] = ca * ct * cp + sa * sp;
    r[1] = sa * ct * cp - ca * sp;
    r[2] = st * cp;
    r[3] = ca * ct * sp - sa * cp;
    r[4] = sa * ct * sp + ca * cp;
    r[5] = st * sp;
    r[6] = -ca * st;
    r[7] = -sa * st;
    r[8] = ct;
    f->erad = EARTH_RADIUS_KM;
    f->xshift = -0.5 * xlen;
    f->yshift = -0.5 * ylen;
    f->g0 = (double)(0.5 * ylen) / (double)f->erad;
    f->b0 = (double)(0.5 * xlen) / (double)f->erad;
}

/* model (x, y) in km from the box corner -> longitude, latitude (degrees) */
static void gc_to_lonlat(const gc_frame *f, float x, float y, float *lon, float *lat)
{
    const double b = x / f->erad - f->b0;   /* angle along x from the centre */
    const double g = y / f->erad - f->g0;   /* and along y */
    const double cb = cos(b), sb = sin(b), cg = cos(g), sg = sin(g);
    /* the point on the unit sphere in the model frame */
    const double s = sqrt(1.0 + sb * sb * sg * sg);
    const double xp = sg * cb * s;
    const double yp = sb * cg * s;
    const double zp = sqrt(1.0 - xp * xp - yp * yp);
    const double *r = f->rot;
    const double xg = xp * r[0] + yp * r[1] + zp * r[2];
    const double yg = xp * r[3] + yp * r[4] + zp * r[5];
    const double zg = xp * r[6] + yp * r[7] + zp * r[8];
    *lat = 90.0 - atan(sqrt(xg * xg + yg * yg) / zg) / RAD_PER_DEG;
    float lo = 0.0f;
    if (xg != 0.0) lo = atan(yg / xg) / RAD_PER_DEG;
    if (xg < 0.0) lo = lo - 180.0;
    while (lo < -180.0) lo = lo + 360.0;
    *lon = lo;
}

/* ------------------------------------------------------------------------ */
/* Input files                                                              */
/* ------------------------------------------------------------------------ */

/* modelbox: the 5th line holds "<name>= lon <name>= lat <name>= rot" */
static int read_modelbox(const char *path, float *lon, float *lat, float *rot, char *err, size_t errlen)
{
    FILE *f = fopen(path, "r");
    if (!f) return fail(err, errlen, AWP_EIO, "cannot open model box file %s", path);
    char line[1024] = {0};
    int ok = 1;
    for (int n = 0; n < 5 && ok; ++n) ok = fgets(line, sizeof line, f) != NULL;
    fclose(f);
    if (!ok) return fail(err, errlen, AWP_EIO, "%s: fewer than 5 lines", path);
    const char *tok[6];
    int nt = 0;
    for (char *c = strtok(line, "= "); c && nt < 6; c = strtok(NULL, "= ")) tok[nt++] = c;
    if (nt < 6) return fail(err, errlen, AWP_EINVAL, "%s: line 5 is not 'mlon= .. mlat= .. mrot= ..'", path);
    *lon = (float)atof(tok[1]);
    *lat = (float)atof(tok[3]);
    *rot = (float)atof(tok[5]);
    return AWP_OK;
}

/* gridout: blocks "nx=<n>" + n lines + a length line, then ny, then nz */
static int read_gridout(const char *path, int *nx, int *ny, int *nz, char *err, size_t errlen)
{
    FILE *f = fopen(path, "r");
    if (!f) return fail(err, errlen, AWP_EIO, "cannot open gridout file %s", path);
    char line[1024];
    int *dims[3] = {nx, ny, nz};
    const char *key[3] = {"nx=", "ny=", "nz="};
    int rc = AWP_OK;
    if (!fgets(line, sizeof line, f)) rc = AWP_EIO;
    for (int d = 0; d < 3 && rc == AWP_OK; ++d) {
        if (!fgets(line, sizeof line, f) || strncmp(line, key[d], 3) || sscanf(line + 3, "%d", dims[d]) != 1 ||
            *dims[d] < 1) {
            rc = AWP_EINVAL;
            break;
        }
        if (d < 2)
            for (int n = 0; n < *dims[d] + 1 && rc == AWP_OK; ++n)
                if (!fgets(line, sizeof line, f)) rc = AWP_EIO;
    }
    fclose(f);
    if (rc) return fail(err, errlen, rc, "%s: not an RWG gridout file (nx=, ny=, nz= blocks)", path);
    return AWP_OK;
}

typedef struct {
    int x, y, z;
    long long indx;
    float dep;
} rwg_point;

/* RWG coordfile: '#' lines, the count, then "x y z indx lon lat dep" */
static int read_rwg_coordfile(const char *path, rwg_point **out, int *np, char *err, size_t errlen)
{
    FILE *f = fopen(path, "r");
    if (!f) return fail(err, errlen, AWP_EIO, "cannot open coordfile %s", path);
    char line[1024];
    int n = -1;
    while (fgets(line, sizeof line, f))
        if (line[0] != '#') {
            if (sscanf(line, "%d", &n) != 1) n = -1;
            break;
        }
    if (n < 0) {
        fclose(f);
        return fail(err, errlen, AWP_EINVAL, "%s: no point count", path);
    }
    rwg_point *p = (rwg_point *)malloc(sizeof(rwg_point) * (size_t)(n > 0 ? n : 1));
    if (!p) {
        fclose(f);
        return fail(err, errlen, AWP_ENOMEM, "coordfile: out of memory");
    }
    for (int k = 0; k < n; ++k) {
        float lon, lat;
        if (!fgets(line, sizeof line, f) ||
            sscanf(line, "%d %d %d %lld %f %f %f", &p[k].x, &p[k].y, &p[k].z, &p[k].indx, &lon, &lat, &p[k].dep) !=
                7) {
            free(p);
            fclose(f);
            return fail(err, errlen, AWP_EINVAL, "%s: point %d is not 'x y z indx lon lat dep'", path, k);
        }
    }
    fclose(f);
    *out = p;
    *np = n;
    return AWP_OK;
}

/* ------------------------------------------------------------------------ */

/* Moduli and density of one point from the two media planes it needs: `a`
 * is the plane above the point (the point's own plane at the surface), `b`
 * the next one down; each [x][y][3] Vp, Vs (m/s), rho (kg/m^3).  At y = 0
 * the density comes from y = 1 (the mesh generator copies that row). */
static void point_media(const float *a, const float *b, int x, int y, int z, int nx, int ny, int nz, int mu_corr,
                        head_record *h)
{
    const float km = 0.001f;             /* m/s -> km/s, kg/m^3 -> g/cm^3 */
    const float to_dyne = 1.0e+10f;      /* (km/s)^2 g/cm^3 -> dyne/cm^2 */
    const float fh = 25.0f, fp = 1.0f;   /* reference frequencies of the correction */
    const long vp = 3L * ((long)x * ny + y), vs = vp + 1;
    const int yr = y ? y : 1;
    const long r0 = 3L * ((long)x * ny + yr) + 2;
    const long rx = 3L * ((long)(x + 1) * ny + yr) + 2;
    const long ry = 3L * ((long)x * ny + y + 1) + 2;
    float mu;
    if (mu_corr) {
        const float qs = 0.05 * a[vs];
        const float vsc = a[vs] * km * (1 + (log(fh / fp) / (qs * 3.14159)));
        mu = a[r0] * km * vsc * vsc;
    } else {
        mu = a[vs] * km * a[vs] * km * a[r0] * km;
    }
    const float lam = a[vp] * km * a[vp] * km * a[r0] * km - 2 * mu;
    h->mu = to_dyne * mu;
    h->lam = to_dyne * lam;
    const float rho0 = a[r0] * km;
    if (x == nx - 1 || z == nz - 1) {
        h->rho = rho0;
    } else {
        /* harmonic mean over the point and its +x, +y, +z neighbours */
        const float bx = 2.0 / (rho0 + a[rx] * km);
        const float by = 2.0 / (rho0 + a[ry] * km);
        const float bz = 2.0 / (rho0 + b[r0] * km);
        h->rho = 3.0 / (bx + by + bz);
    }
}

int awp_write_sgt_head(const awp_head_params *hp, char *err, size_t errlen)
{
    if (!hp || !hp->modelbox || !hp->coordfile || !hp->fdloc || !hp->gridout || !hp->media || !hp->out)
        return fail(err, errlen, AWP_EINVAL, "awp_write_sgt_head: missing file name");
    if (hp->sgt_tinc < 1 || hp->spacing <= 0.0f || hp->flo == 0.0f)
        return fail(err, errlen, AWP_EINVAL, "awp_write_sgt_head: bad spacing, time decimation or frequency");
    if (hp->component != 'x' && hp->component != 'y' && hp->component != 'z')
        return fail(err, errlen, AWP_EINVAL, "awp_write_sgt_head: component must be x, y or z");
    float mlon, mlat, mrot;
    int rc = read_modelbox(hp->modelbox, &mlon, &mlat, &mrot, err, errlen);
    if (rc) return rc;
    int nx, ny, nz;
    rc = read_gridout(hp->gridout, &nx, &ny, &nz, err, errlen);
    if (rc) return rc;
    int xsrc, ysrc;
    FILE *f = fopen(hp->fdloc, "r");
    if (!f) return fail(err, errlen, AWP_EIO, "cannot open fdloc file %s", hp->fdloc);
    rc = fscanf(f, "%d %d", &xsrc, &ysrc) == 2 ? AWP_OK : AWP_EINVAL;
    fclose(f);
    if (rc) return fail(err, errlen, rc, "%s: not 'x y'", hp->fdloc);
    rwg_point *pt = NULL;
    int np = 0;
    rc = read_rwg_coordfile(hp->coordfile, &pt, &np, err, errlen);
    if (rc) return rc;
    for (int k = 0; k < np; ++k)
        if (pt[k].x < 0 || pt[k].x >= nx || pt[k].y < 0 || pt[k].y >= ny || pt[k].z < 0 || pt[k].z >= nz) {
            free(pt);
            return fail(err, errlen, AWP_EINVAL, "coordfile point %d (%d %d %d) outside the %dx%dx%d grid", k,
                        pt[k].x, pt[k].y, pt[k].z, nx, ny, nz);
        }

    gc_frame fr;
    gc_frame_init(&fr, mlon, mlat, mrot, nx, ny, hp->spacing);
    const float h = hp->spacing;
    head_master m;
    memset(&m, 0, sizeof m);
    m.geoproj = 1;
    m.modellon = mlon;
    m.modellat = mlat;
    m.modelrot = mrot;
    m.xshift = fr.xshift;
    m.yshift = fr.yshift;
    m.globnp = m.localnp = np;
    m.nt = hp->nt / hp->sgt_tinc;

    head_index *ix = (head_index *)calloc((size_t)(np > 0 ? np : 1), sizeof(head_index));
    head_record *hd = (head_record *)calloc((size_t)(np > 0 ? np : 1), sizeof(head_record));
    const size_t plane = (size_t)3 * nx * ny;
    float *pa = (float *)malloc(plane * sizeof(float)), *pb = (float *)malloc(plane * sizeof(float));
    int *first = (int *)malloc(sizeof(int) * (size_t)nz), *next = (int *)malloc(sizeof(int) * (size_t)(np + 1));
    FILE *mf = fopen(hp->media, "rb");
    if (!ix || !hd || !pa || !pb || !first || !next || !mf) {
        rc = mf ? fail(err, errlen, AWP_ENOMEM, "awp_write_sgt_head: out of memory")
                : fail(err, errlen, AWP_EIO, "cannot open media file %s", hp->media);
        goto done;
    }

    /* the site: same for every record */
    float src_lon, src_lat;
    gc_to_lonlat(&fr, xsrc * h, ysrc * h, &src_lon, &src_lat);

    /* points by depth, so each pair of media planes is read once */
    for (int z = 0; z < nz; ++z) first[z] = -1;
    for (int k = np - 1; k >= 0; --k) {
        next[k] = first[pt[k].z];
        first[pt[k].z] = k;
    }
    for (int z = 0; z < nz; ++z) {
        if (first[z] < 0) continue;
        const int p0 = z > 0 ? z - 1 : 0;
        const int nplanes = p0 + 1 < nz ? 2 : 1;
        if (fseeko(mf, (off_t)p0 * (off_t)plane * (off_t)sizeof(float), SEEK_SET) ||
            fread(pa, sizeof(float), plane, mf) != plane ||
            (nplanes == 2 && fread(pb, sizeof(float), plane, mf) != plane)) {
            rc = fail(err, errlen, AWP_EIO, "%s: short media file (planes %d..%d of %dx%dx%d)", hp->media, p0,
                      p0 + nplanes - 1, nx, ny, nz);
            goto done;
        }
        for (int k = first[z]; k >= 0; k = next[k]) {
            const rwg_point *q = &pt[k];
            head_index *si = &ix[k];
            head_record *r = &hd[k];
            si->indx = q->indx;
            si->xsgt = q->x;
            si->ysgt = q->y;
            si->zsgt = q->z;
            si->h = h;
            r->indx = q->indx;
            r->geoproj = 1;
            r->modellon = mlon;
            r->modellat = mlat;
            r->modelrot = mrot;
            r->xshift = fr.xshift;
            r->yshift = fr.yshift;
            r->xazim = 90.0 + mrot;
            r->nt = m.nt;
            r->dt = hp->dt * hp->sgt_tinc;
            r->h = h;
            r->xsrc = xsrc;
            r->ysrc = ysrc;
            r->zsrc = 1;
            r->xsgt = q->x;
            r->ysgt = q->y;
            r->zsgt = q->z;
            r->xmom = hp->component == 'x' ? hp->moment : 0.0f;
            r->ymom = hp->component == 'y' ? hp->moment : 0.0f;
            r->zmom = hp->component == 'z' ? hp->moment : 0.0f;
            r->tst = -3.0 / hp->flo;
            r->src_lon = src_lon;
            r->src_lat = src_lat;
            r->src_dep = (r->zsrc - 1) * h;
            gc_to_lonlat(&fr, q->x * h, q->y * h, &r->sgt_lon, &r->sgt_lat);
            r->sgt_dep = q->dep;
            const float dx = r->xsgt - r->xsrc, dy = r->ysgt - r->ysrc, dz = r->zsgt - r->zsrc;
            r->cdist = h * sqrt(dx * dx + dy * dy + dz * dz);
            point_media(pa, pb, q->x, q->y, q->z, nx, ny, nz, hp->mu_corr, r);
        }
    }

    {
        FILE *o = fopen(hp->out, "wb");
        if (!o) {
            rc = fail(err, errlen, AWP_EIO, "cannot create %s", hp->out);
            goto done;
        }
        const int ok = fwrite(&m, sizeof m, 1, o) == 1 &&
                       fwrite(ix, sizeof(head_index), (size_t)np, o) == (size_t)np &&
                       fwrite(hd, sizeof(head_record), (size_t)np, o) == (size_t)np;
        if (fclose(o) || !ok) rc = fail(err, errlen, AWP_EIO, "writing %s failed", hp->out);
    }

done:
    if (mf) fclose(mf);
    free(pt);
    free(ix);
    free(hd);
    free(pa);
    free(pb);
    free(first);
    free(next);
    return rc;
}
