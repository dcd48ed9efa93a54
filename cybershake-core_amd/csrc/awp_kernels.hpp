// awp_kernels.hpp -- device data layout and kernel argument blocks of the
// AWP-ODC-SGT time step on MI355X (gfx950).
//
// Layout in HBM: every volume array is x-fastest with a padded pitch.  The
// interior node (i, j, k), 1-based with k = nz at the free surface (the
// reference's internal convention, structure.f:488), lives at element
//     base + k*sz + j*sy + i
// with base chosen so that (1,1,1) sits at (XO, YO, ZO) = (32, 2, 2).  XO = 32
// puts the start of every interior row on a 128-byte line, so a 64-lane wave
// reading 64 consecutive x values touches exactly two lines.  Two ghost planes
// on every side hold the reference's halo values (zero outside the domain,
// free-surface images at k = nz+1, nz+2).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace awp {

constexpr int XO = 32;   // x offset of i = 1
constexpr int YO = 2;    // y offset of j = 1
constexpr int ZO = 2;    // z offset of k = 1
constexpr int TX = 64;   // threads per block in x: one wave64 row

struct Geo {
    int nx, ny, nz;
    int px, py, pz;        // padded extents
    long long sy, sz;      // strides of j and k (elements)
    long long base;        // element index of (i,j,k) = (0,0,0)
    long long nvol;        // px*py*pz
};

static inline Geo make_geo(int nx, int ny, int nz)
{
    Geo g;
    g.nx = nx; g.ny = ny; g.nz = nz;
    int need = XO + nx + 2;
    g.px = (need + 31) / 32 * 32;
    g.py = ny + 2 * YO;
    g.pz = nz + 2 * ZO;
    g.sy = g.px;
    g.sz = (long long)g.px * g.py;
    g.base = (long long)(ZO - 1) * g.sz + (long long)(YO - 1) * g.sy + (XO - 1);
    g.nvol = g.sz * g.pz;
    return g;
}

__host__ __device__ inline long long gidx(const Geo &g, int i, int j, int k)
{
    return g.base + (long long)k * g.sz + (long long)j * g.sy + i;
}

// Pointers of one run's state.  Media and memory variables are FP32 here;
// the mixed-storage variants reinterpret them (see awp_kernels.hip).
struct State {
    float *u1, *v1, *w1;
    float *xx, *yy, *zz, *xy, *xz, *yz;
    float *r1, *r2, *r3, *r4, *r5, *r6;
    float *d1, *mu, *lam, *qp, *qs;
    float *fx, *fy, *fz;       // sponge profiles, index 1..n (fx[0] unused)
};

struct Coef {
    float dth;       // dt/dh
    float dt;
    float dh1;       // 1/dh
    float c1, c2;    // 9/8, -1/24
    float tau1[8];   // [(itx*2+ity)*2+itz]
    float tau2[8];
};

struct Launch {
    int tile_y;      // threads per block in y
    int kc;          // z levels per block (chunk)
    int nchunk;      // number of z chunks
    int unroll;      // z-loop unroll factor (1 or 2)
};

// host launchers (awp_kernels.hip)
hipError_t launch_velocity(const Geo &g, const State &s, const Coef &c, const Launch &l, hipStream_t st);
hipError_t launch_fs_velocity(const Geo &g, const State &s, hipStream_t st);
hipError_t launch_stress(const Geo &g, const State &s, const Coef &c, const Launch &l, hipStream_t st);
hipError_t launch_source(const Geo &g, const State &s, int igreen, int npsrc, const long long *psrc_idx,
                         const long long *psrc_top, const float *stf, int nst, int istep, float vtst,
                         float vtst1, hipStream_t st);
hipError_t launch_sgt_coef(const Geo &g, const State &s, int nkept, const long long *kpos, float *sg, hipStream_t st);
hipError_t launch_sgt_gather(const Geo &g, const State &s, int nkept, const long long *kpos, const int *kidx,
                             const float *sg, float *out, int nsamp_max, int sample, hipStream_t st);
struct MediaConst {
    float pi, tmp1, tmp2, lw;   // 4*atan(1), Q-fit constants, log(w2/w0)
    int socalq, nvar;
};
hipError_t launch_media_setup(const Geo &g, const State &s, const float *media, const MediaConst &mc, hipStream_t st);
hipError_t launch_media_ghosts(const Geo &g, const State &s, hipStream_t st);
hipError_t launch_pack_field(const Geo &g, const float *src, int kext, float *dst, hipStream_t st);

}  // namespace awp
