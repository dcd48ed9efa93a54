// awp_kernels.hpp -- device data layout and kernel argument blocks of the
// AWP-ODC-SGT time step on MI355X (gfx950).
//
// Layout in HBM: every volume array is x-fastest with a padded pitch.  The
// interior node (i, j, k), 1-based with k = nz at the free surface (the
// reference's internal convention, structure.f:488), lives at element
//     base + k*sz + j*sy + i
// with base chosen so that (1,1,1) sits at (XO, YO, ZO) = (32, 2, 2).  XO = 32
// puts the start of every interior row of a 4-byte array on a 128-byte line,
// so a 64-lane wave reading 64 consecutive x values touches exactly two
// lines; 2-byte arrays start 64 bytes into their allocation so that their
// rows start on a line too (one line per wave access).  Two ghost planes
// on every side hold the reference's halo values (zero outside the domain,
// free-surface images at k = nz+1, nz+2).
//
// Storage: velocities and stresses are FP32 always.  The memory variables
// r1..r6, the attenuation coefficients qp/qs, the density d1 and the
// compliances mu/lam are FP32 (StoreF32) or, in the mixed mode (StoreMixed),
// FP16 (see StoreMixed).  Each FP16 qp/qs array holds value * 2^e, a
// per-array power of two that puts its largest magnitude in [2^13, 2^14);
// the kernels undo it with an exact multiply, so FP16's only effect is the
// 11-bit rounding of the stored value.  d1/mu/lam are stored as FP16
// deviations h from an FP32 profile value b(k) per level, x = fma(b, h, b)
// (mdec): the error is FP16's rounding of x/b - 1, zero on a layered medium
// (a plain FP16 rounding of the values themselves is constant across a layer,
// shifts the wave speeds coherently and took the SGT error past 1e-3).
// Arithmetic is FP32 in both modes.
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
    // pitch a multiple of 64 elements: every row of a 4-byte array starts on
    // a 256-byte boundary and (with the 64-byte shifted base the 2-byte
    // arrays get) every row of a 2-byte array on a 128-byte line
    int need = XO + nx + 2;
    g.px = (need + 63) / 64 * 64;
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

// exponent slot of row segment seg at (j, k) (interior j, k); each slot is
// 8 bytes: the exponents of r1..r6 in bytes 0..5 (one load, one store)
__host__ __device__ inline long long rexp_idx(const Geo &g, int nseg, int seg, int j, int k)
{
    return ((long long)(k - 1) * g.ny + (j - 1)) * nseg + seg;
}
__device__ inline int rexp_of(unsigned long long w, int q) { return (int)(signed char)(w >> (8 * q)); }
__device__ inline unsigned long long rexp_pack(const int e[6])
{
    unsigned long long w = 0;
    for (int q = 0; q < 6; ++q) w |= (unsigned long long)(unsigned char)(signed char)e[q] << (8 * q);
    return w;
}

// ---- storage element types ------------------------------------------------
struct Bf16 {
    unsigned short u;
};

__host__ __device__ inline float ld(float x) { return x; }
__device__ inline float ld(Bf16 x) { return __uint_as_float(((unsigned)x.u) << 16); }
__device__ inline float ld(_Float16 x) { return (float)x; }

template <class T> __device__ inline T st(float f);
template <> __device__ inline float st<float>(float f) { return f; }
// round to nearest even (finite values)
template <> __device__ inline Bf16 st<Bf16>(float f)
{
    unsigned b = __float_as_uint(f);
    b += 0x7FFFu + ((b >> 16) & 1u);
    Bf16 r;
    r.u = (unsigned short)(b >> 16);
    return r;
}
template <> __device__ inline _Float16 st<_Float16>(float f) { return (_Float16)f; }

struct StoreF32 {
    using R = float;   // r1..r6
    using Q = float;   // qp, qs
    using M = float;   // d1, mu, lam
    static constexpr bool mixed = false;
};
struct StoreMixed {
    using R = _Float16;   // r * 2^(15-e), e per row segment (see below)
    using Q = _Float16;
    using M = _Float16;   // x / b(k) - 1, b(k) the level's profile value
    static constexpr bool mixed = true;
};

// The density and the compliances from their storage: FP32 as stored, or
// (mixed) the FP16 deviation h from the level's profile value b,
// x = fma(b, h, b) -- one rounding, the same on the host's fmaf.
__host__ __device__ __forceinline__ float mdec(float x, float) { return x; }
__device__ __forceinline__ float mdec(_Float16 h, float b) { return fmaf(b, (float)h, b); }
// Profile entry of array a (0 d1, 1 mu, 2 lam) at level k (-1 <= k <= nz+2)
// in the [3][pz] profile table of a mixed-storage run.
__host__ __device__ __forceinline__ long long mbar_idx(const Geo &g, int a, int k)
{
    return (long long)a * g.pz + k + 1;
}

// Memory variables in mixed storage: each stored value is the FP16 (round to
// nearest even) of r * 2^(15-e), where e is shared by the 64 points of one x
// row segment (i = 64s+1 .. 64s+64, the lanes of one wave) at one (j, k):
// the exponent of the segment's largest |r| (f * 2^e, f in [0.5, 1)),
// clamped to [REXP_MIN, REXP_MAX], kept in one byte per segment and array
// (the six bytes of a segment side by side, rexp_idx).  Every
// value keeps FP16's 11-bit relative precision over 30 binades below the
// segment's largest; decoding is the exact h * 2^(e-15).  (BF16 storage of r,
// 8-bit precision, let the SGT error grow past 1e-3 over 1500 steps.)
constexpr int REXP_MIN = -80, REXP_MAX = 120;
__device__ inline int row_exponent(float amax)
{
    int e = 0;
    if (amax > 0.0f) frexpf(amax, &e);
    return min(max(e, REXP_MIN), REXP_MAX);
}
// Maximum of v >= 0 over the 64 lanes of the wave (all lanes active), the
// same in every lane.  Non-negative floats order as their bit patterns, so
// the reduction runs on the integers: four DPP steps (quad permutes [1,0,3,2]
// and [2,3,0,1], half-row and row mirrors) leave each 16-lane row's maximum
// in every lane of the row, and the four rows meet through readlane in
// scalar registers.  (A shuffle per step would cost an LDS-path permute.)
__device__ inline float wave_max(float v)
{
    int x = __float_as_int(v);
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, false));
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x140, 0xF, 0xF, false));
    const int a = max(__builtin_amdgcn_readlane(x, 0), __builtin_amdgcn_readlane(x, 16));
    const int b = max(__builtin_amdgcn_readlane(x, 32), __builtin_amdgcn_readlane(x, 48));
    return __int_as_float(max(a, b));
}
// The exponents of the six memory variables of one row segment after an
// update, packed as rexp_pack does: row_exponent(wave_max(act ? |v[q]| : 0))
// for q = 0..5, all 64 lanes active, the same in every lane.  ew_old holds
// the row's exponents before the update.  A row's largest value usually
// stays in its binade from one step to the next, which two ballots per array
// establish (every |v| < 2^e_old, some |v| >= 2^(e_old-1) -- or, at the clamp
// REXP_MIN, some v != 0): the exponent is then e_old; for an array that
// fails them a third ballot finds an all-zero row (exponent 0).  Only when an
// array is neither does the wave take the DPP reductions, for all six in one
// uniform branch (so the six values stay in registers).  The result equals
// the reductions' in every case.
__device__ inline unsigned long long row_exponents6(const float (&v)[6], const bool act,
                                                    const unsigned long long ew_old)
{
    int e[6];
    bool kept[6];
    bool all_kept = true;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const int eo = rexp_of(ew_old, q);
        const float a = fabsf(v[q]);
        const float hi = __int_as_float((eo + 127) << 23);                                     // 2^eo
        const float lo = eo == REXP_MIN ? __int_as_float(1) : __int_as_float((eo + 126) << 23);
        const bool below = __builtin_amdgcn_ballot_w64(!act || a < hi) == ~0ull;
        const bool some = __builtin_amdgcn_ballot_w64(act && a >= lo) != 0ull;
        kept[q] = below && some;
        e[q] = eo;
        all_kept = all_kept && kept[q];
    }
    if (!all_kept) {
        // the arrays whose largest value left its binade: an all-zero row
        // takes exponent 0, any other one the reductions
        bool slow = false;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            if (!kept[q]) {
                const bool zero = __builtin_amdgcn_ballot_w64(act && v[q] != 0.0f) == 0ull;
                e[q] = 0;
                slow = slow || !zero;
            }
        }
        if (slow) {
#pragma unroll
            for (int q = 0; q < 6; ++q) e[q] = row_exponent(wave_max(act ? fabsf(v[q]) : 0.0f));
        }
    }
    return rexp_pack(e);
}
// The exponent slot `slot` of a row segment (the same in every active lane
// of the wave), read through the scalar data cache: the exponents then live
// in scalar registers and their unpacking, the bounds of the binade test and
// the decode offsets are scalar work.  Each slot is read once per launch, by
// the wave that rewrites it afterwards (with a vector store of its new value,
// which depends on this read), and a launch starts with the scalar cache
// invalidated, so the read sees the previous launch's value.
template <class T>
__device__ __forceinline__ unsigned long long rexp_load(const T *rexp, long long slot)
{
    using CP = const __attribute__((address_space(4))) unsigned long long *;
    const int u = __builtin_amdgcn_readfirstlane((int)slot);
    return ((CP)(const void *)rexp)[u];
}
__device__ inline _Float16 r_encode(float r, int e) { return (_Float16)ldexpf(r, 15 - e); }
// Mixed storage keeps a point's qp and qs side by side in one 4-byte word
// (FP16 qp in the low half, FP16 qs in the high half), so one load brings
// both; these unpack the halves.
__device__ __forceinline__ float q_lo(unsigned w) { return (float)__builtin_bit_cast(_Float16, (unsigned short)(w & 0xFFFFu)); }
__device__ __forceinline__ float q_hi(unsigned w) { return (float)__builtin_bit_cast(_Float16, (unsigned short)(w >> 16)); }
// The same for mu (low half) and lam (high half), whose FP16 deviations
// mdec decodes.
__device__ __forceinline__ _Float16 h_lo(unsigned w) { return __builtin_bit_cast(_Float16, (unsigned short)(w & 0xFFFFu)); }
__device__ __forceinline__ _Float16 h_hi(unsigned w) { return __builtin_bit_cast(_Float16, (unsigned short)(w >> 16)); }
// h * 2^(e-15) as one multiply by the power of two (exact scaling, the
// same rounding as ldexpf; e - 15 lies in [REXP_MIN - 15, REXP_MAX - 15],
// inside the normal exponent range)
__device__ inline float r_decode(_Float16 h, int e) { return (float)h * __int_as_float((e - 15 + 127) << 23); }

// Power-of-two storage scales of the FP16 qp/qs arrays (1 in FP32 mode) and
// their inverses.
struct MScale {
    float qp, qs;
    float iqp, iqs;
};

// Pointers of one run's state.  r1..r6 and qp/qs point to R / Q arrays of
// the storage mode.
struct State {
    float *u1, *v1, *w1;
    float *xx, *yy, *zz, *xy, *xz, *yz;
    void *d1, *mu, *lam;       // M arrays of the storage mode
    float *mbar;               // mixed: [3][pz] level profiles of d1, mu, lam (mbar_idx)
    void *r1, *r2, *r3, *r4, *r5, *r6;
    signed char *rexp;         // mixed: rexp_n 8-byte slots of row-segment exponents
    long long rexp_n;          // nz * ny * nseg
    int nseg;                  // (nx + 63) / 64
    void *qp, *qs;
    float *fx, *fy, *fz;       // sponge profiles, index 1..n (fx[0] unused)
    MScale ms;
};

// Pointers a kernel reads from a job table in device memory are generic
// (flat) pointers to the compiler, and loads through them become flat loads,
// which wait on both the vector-memory and the LDS counters (kernel-argument
// pointers are promoted to global ones by the compiler; table pointers are
// not).  The volume kernels therefore access the state through GState, the
// same pointers typed in the global address space.
template <class T>
using G = __attribute__((address_space(1))) T;

struct GState {
    G<float> *u1, *v1, *w1;
    G<float> *xx, *yy, *zz, *xy, *xz, *yz;
    G<void> *d1, *mu, *lam;
    G<float> *mbar;
    G<void> *r1, *r2, *r3, *r4, *r5, *r6;
    G<signed char> *rexp;
    long long rexp_n;
    int nseg;
    G<void> *qp, *qs;
    G<float> *fx, *fy, *fz;
    MScale ms;
};

__device__ __forceinline__ GState as_global(const State &a)
{
    GState s;
    s.u1 = (G<float> *)a.u1; s.v1 = (G<float> *)a.v1; s.w1 = (G<float> *)a.w1;
    s.xx = (G<float> *)a.xx; s.yy = (G<float> *)a.yy; s.zz = (G<float> *)a.zz;
    s.xy = (G<float> *)a.xy; s.xz = (G<float> *)a.xz; s.yz = (G<float> *)a.yz;
    s.d1 = (G<void> *)a.d1; s.mu = (G<void> *)a.mu; s.lam = (G<void> *)a.lam;
    s.mbar = (G<float> *)a.mbar;
    s.r1 = (G<void> *)a.r1; s.r2 = (G<void> *)a.r2; s.r3 = (G<void> *)a.r3;
    s.r4 = (G<void> *)a.r4; s.r5 = (G<void> *)a.r5; s.r6 = (G<void> *)a.r6;
    s.rexp = (G<signed char> *)a.rexp;
    s.rexp_n = a.rexp_n;
    s.nseg = a.nseg;
    s.qp = (G<void> *)a.qp; s.qs = (G<void> *)a.qs;
    s.fx = (G<float> *)a.fx; s.fy = (G<float> *)a.fy; s.fz = (G<float> *)a.fz;
    s.ms = a.ms;
    return s;
}

// Loads and stores at a byte offset from an array base.  With a 32-bit
// offset (O = unsigned) and a uniform base (scalar registers), an access
// takes the scalar-base + 32-bit vector-offset form, one offset register
// serves every array of an element size, and +-1, +-2 x neighbours become
// immediate offsets; this needs every array below 4 GiB.  A 64-bit offset
// (O = unsigned long long) addresses any size.
template <class T, class O>
__device__ __forceinline__ T gl(const G<T> *base, O off, int imm = 0)
{
    return *reinterpret_cast<const G<T> *>(reinterpret_cast<const G<char> *>(base) + off + imm);
}
template <class T, class O>
__device__ __forceinline__ void gs(G<T> *base, O off, T v, int imm = 0)
{
    *reinterpret_cast<G<T> *>(reinterpret_cast<G<char> *>(base) + off + imm) = v;
}
// Streamed access.  A value of a pointwise array (one read and one write per
// step at the point itself, never a neighbour's halo) goes through the
// caches with the non-temporal hint, leaving L2 to the halo lines of the
// stencil arrays that neighbouring tiles read again.  AWP_NT=0 at compile
// time turns the hint off.
#ifndef AWP_NT
#define AWP_NT 1
#endif
template <class T>
__device__ __forceinline__ T nt_ld(const G<T> *a)
{
    if constexpr (AWP_NT != 0)
        return __builtin_nontemporal_load(a);
    else
        return *a;
}
template <class T>
__device__ __forceinline__ void nt_st(G<T> *a, T v)
{
    if constexpr (AWP_NT != 0)
        __builtin_nontemporal_store(v, a);
    else
        *a = v;
}
template <class T, class O>
__device__ __forceinline__ T gln(const G<T> *base, O off)
{
    return nt_ld(reinterpret_cast<const G<T> *>(reinterpret_cast<const G<char> *>(base) + off));
}
template <class T, class O>
__device__ __forceinline__ void gsn(G<T> *base, O off, T v)
{
    nt_st(reinterpret_cast<G<T> *>(reinterpret_cast<G<char> *>(base) + off), v);
}
// Position of one thread's point in the volume arrays while it marches up in
// z, as a 32-bit byte offset into a 4-byte array (2-byte arrays take it
// halved), and access to its neighbours (di, dj, dk): every access is a
// uniform base plus a shared offset register plus an immediate.  Needs every
// array below 4 GiB.
struct Cursor32 {
    unsigned o;        // the point
    unsigned sy, sz;   // row and level strides
    __device__ Cursor32(const Geo &g, long long e)
        : o((unsigned)(e * 4)), sy((unsigned)(g.sy * 4)), sz((unsigned)(g.sz * 4)) {}
    // (unsigned arithmetic: exact modulo 2^32 for offsets inside the array)
    __device__ __forceinline__ unsigned at(int dj, int dk) const { return o + (unsigned)dj * sy + (unsigned)dk * sz; }
    template <class T>
    __device__ __forceinline__ T ld(const G<T> *p, int di = 0, int dj = 0, int dk = 0) const
    {
        return gl(p, at(dj, dk) >> (sizeof(T) == 2 ? 1 : 0), di * (int)sizeof(T));
    }
    template <class T>
    __device__ __forceinline__ void st(G<T> *p, T v, int di = 0, int dj = 0, int dk = 0) const
    {
        gs(p, at(dj, dk) >> (sizeof(T) == 2 ? 1 : 0), v, di * (int)sizeof(T));
    }
    // the point itself, streamed (nt_ld / nt_st)
    template <class T>
    __device__ __forceinline__ T ldn(const G<T> *p) const
    {
        return gln(p, o >> (sizeof(T) == 2 ? 1 : 0));
    }
    template <class T>
    __device__ __forceinline__ void stn(G<T> *p, T v) const
    {
        gsn(p, o >> (sizeof(T) == 2 ? 1 : 0), v);
    }
    __device__ __forceinline__ void up() { o += sz; }
};

// z chunks of the two-stage kernels: about `kc` levels each
constexpr int SPLIT_KC = 25;
static inline int split_nchunk(int nz, int kc)
{
    const int n = (nz + kc / 2) / kc;
    return n < 1 ? 1 : n;
}

// The fused step addresses every volume array with 32-bit byte offsets
// from a base rebased per z chunk (the chunk's lowest level, kb - 4, or -1):
// the offsets then span at most kc + 10 levels of the padded plane, whatever
// the size of the arrays.  Largest chunk for which they fit:
static inline int fused_kc_max(const Geo &g)
{
    const long long lv = (long long)(((1ull << 32) - 1) / (unsigned long long)(g.sz * 4));
    return (int)(lv - 10 > 1 << 20 ? 1 << 20 : lv - 10);
}

// whether 32-bit byte offsets reach every element of a run's arrays (the
// 32-bit forms need fewer registers and address instructions, but the
// compiler then schedules the two-stage FP32 kernels with fewer loads in
// flight; which form is faster is measured per kernel and storage mode)
static inline bool offsets32(const Geo &g) { return (unsigned long long)g.nvol * 4ull < (1ull << 32); }

// FP32 media (the storage in FP32 mode; in mixed mode qp/qs point to set-up
// temporaries)
struct Media32 {
    float *d1, *mu, *lam, *qp, *qs;
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
    int tile_xw;     // two-stage kernels: waves per tile row (1 or 2: tiles 64 or 128 points wide)
    int kc;          // z levels per block (chunk)
    int nchunk;      // number of z chunks
    int unroll;      // z-loop unroll factor (1 or 2); prefetch depth of the fused step
    int pw, ph;      // fused step: tile patch width / height (tiles)
    int claim;       // persistent blocks claiming tiles per XCD (claim_tile): 0 never, 1 with >= 4 tiles a block, 2 always
    int xrot;        // claimed fused launches: XCD x starts on range (x + xrot) % 8
    int spw, sph;    // two-stage kernels: tile patch width / height (0 = rows of tiles)
    bool o32;        // 32-bit byte offsets reach every array (offsets32() of the run's geometry)
    bool o32_s;      // two-stage stress kernel: its 32-bit byte-offset form (only where o32)
    bool o32_v;      // two-stage velocity kernel: its 32-bit byte-offset form (only where o32)
    int bpc_v, bpc_s;   // two-stage kernels: most blocks per CU (0 = as registers allow)
    unsigned lds_v, lds_s;   // the dynamic LDS bytes per block that enforce them (set_block_caps)
};


// Dynamic LDS a block reserves so that at most `bpc` blocks share a CU:
// capping the blocks in flight keeps the tiles an XCD works on at once, and
// with them the halo lines they share, within its L2.  0 (no cap) when bpc
// <= 0 or when the per-block LDS limit cannot enforce it.
static inline unsigned lds_cap(int bpc, int lds_cu, int lds_block)
{
    if (bpc <= 0 || lds_cu <= 0) return 0u;
    const int want = lds_cu / bpc - 512;
    return (want > 0 && want <= lds_block) ? (unsigned)want : 0u;
}
// fills lds_v / lds_s of `l` from bpc_v / bpc_s for device `device`
void set_block_caps(Launch &l, int device);

// (bx, by, bz) of local tile number lt in the patch order of FTiles
__host__ __device__ inline void patch_tile(int lt, int nbx, int nby, int pw, int ph, int &bx, int &by, int &bz)
{
    const int nxy = nbx * nby;
    bz = lt / nxy;
    int q = lt - bz * nxy;
    const int rowt = nbx * ph;                       // tiles of one full patch row
    const int npy = (nby + ph - 1) / ph;
    int pr = q / rowt;
    if (pr > npy - 1) pr = npy - 1;
    q -= pr * rowt;
    const int h = nby - pr * ph < ph ? nby - pr * ph : ph;
    const int npx = (nbx + pw - 1) / pw;
    int pc = q / (pw * h);
    if (pc > npx - 1) pc = npx - 1;
    q -= pc * pw * h;
    const int w = nbx - pc * pw < pw ? nbx - pc * pw : pw;
    bx = pc * pw + q % w;
    by = pr * ph + q / w;
}

struct MediaConst {
    float pi, tmp1, tmp2, lw;   // 4*atan(1), Q-fit constants, log(w2/w0)
    int socalq, nvar;
};

// host launchers (awp_kernels.hip); `mixed` selects the storage mode
// source sample istep (when apply); with trace set, then row istep - 2 of
// the per-step trace (u1 v1 w1 at element tpos)
hipError_t launch_source(const Geo &g, const State &s, int igreen, int npsrc, const long long *psrc_idx,
                         const long long *psrc_top, const float *stf, int nst, int istep, float vtst,
                         float vtst1, bool apply, float *trace, long long tpos, hipStream_t st);
// sg1..sg3 at the SGT nodes from the stored mu, lam (decoded in mixed storage)
hipError_t launch_sgt_coef(const Geo &g, const State &s, int nkept, const long long *kpos, float *sg, hipStream_t st);
hipError_t launch_sgt_gather(const State &s, int nkept, const long long *kpos, const int *kidx,
                             const float *sg, float *out, int nsamp_max, int sample, hipStream_t st);
hipError_t launch_media_setup(const Geo &g, const Media32 &m, const float *media, const MediaConst &mc, hipStream_t st);
hipError_t launch_media_ghosts(const Geo &g, const Media32 &m, hipStream_t st);
// per-block maxima of |a| over n elements into part[nblk] (nblk <= 4096)
// per level k = 1..nz of a (FP32 media array): order-preserving keys of the
// smallest and largest nonzero interior value, atomically into range[2(k-1)],
// range[2(k-1)+1] (which start at 0xFFFFFFFF and 0)
hipError_t launch_level_range(const Geo &g, const float *a, unsigned *range, hipStream_t st);
// dst = FP16 deviations of src from the level profile bar[0..pz-1] (mixed storage)
hipError_t launch_media_encode(const Geo &g, const float *src, _Float16 *dst, const float *bar, hipStream_t st);
hipError_t launch_absmax(const float *a, long long n, float *part, int nblk, hipStream_t st);
hipError_t launch_to_half(const float *src, _Float16 *dst, float scale, long long n, hipStream_t st);
// mixed storage: mu, lam (FP32) to (mu, lam) pairs of FP16 deviations from
// their level profiles (k_media_encode's rounding)
hipError_t launch_media_encode2(const Geo &g, const float *mu, const float *lam, unsigned *dst, const float *bar_mu,
                                const float *bar_lam, hipStream_t st);
// mixed storage: qp, qs (FP32, n points) to (qp, qs) FP16 pairs scaled by sp, ss
hipError_t launch_to_qpair(const float *qp, const float *qs, unsigned *dst, float sp, float ss, long long n,
                           hipStream_t st);
hipError_t launch_sgt_fp16(const float *sgt, const int *kidx, int nkept, int nsamp_max, int nsamp, float *scales,
                           _Float16 *out, hipStream_t st);
// the fused one-sweep step (awp_fused.hip): reads state a, writes state b,
// one block per tile (the static form; the claimed form of a run is
// launch_step_fused_batch over a one-entry table)
hipError_t launch_step_fused(bool mixed, const Geo &g, const State &a, const State &b, const Coef &c,
                             const Launch &l, hipStream_t st);

// tiles of one run for the fused step: nbx * nby * nbz column tiles
// Tile order: within each z chunk, tiles are grouped in patches of pw x ph
// tiles (x fastest inside a patch, patches x fastest, then y); each XCD takes
// a contiguous range of that order, so the blocks an XCD runs at one time
// form a compact patch whose tiles read each other's halo lines from L2
// while they march through z together.
struct FTiles {
    int nbx, nby, nbz;
    int per;   // tiles per XCD (single-run launch)
    int pw, ph;
};
FTiles fused_tiles(const Geo &g, const Launch &l);

// Tile claim of a persistent launch.  The tiles [0, total) form 8 ranges of
// `per` consecutive tiles (the last one shorter).  A block on XCD x takes the
// next unclaimed tile from the front of range (x + rot) % 8, the XCD read
// from the hardware register, so that an XCD sweeps one compact part of the
// domain whatever order the dispatcher hands blocks to XCDs in.  When that
// range is spent it takes tiles from the back of the following ranges: a
// faster XCD finishes a slower one's work on tiles away from the ones the
// owner is working on.  Counter q
// (one 64-bit word per 128-byte line, line q): tiles taken from the front
// of range q (low half) and from its back (high half); a claim is valid
// while the two sum to less than the range's tiles, so no tile is taken
// twice.  Line 8: blocks done -- the last block to finish clears all nine
// for the next launch.
static constexpr int CLAIM_LINE = 16;                  // 64-bit words per 128-byte line
static constexpr int CLAIM_WORDS = 9 * CLAIM_LINE;     // 64-bit words of the claim buffer

// the XCD the calling wave runs on
__device__ inline int xcc_id()
{
    unsigned v;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
    return (int)(v & 7);
}

// the next tile for a block on this XCD, or -1 when every tile is taken
// (called by one lane of the block)
__device__ inline int claim_tile(unsigned long long *claim, int per, int total, int rot)
{
    const int x = (xcc_id() + rot) & 7;
    for (int d = 0; d < 8; ++d) {
        const int q = (x + d) & 7;
        const int n = min(per, total - q * per);   // tiles of range q
        if (n <= 0) continue;
        const unsigned long long o = atomicAdd(&claim[q * CLAIM_LINE], d == 0 ? 1ull : 1ull << 32);
        const unsigned f = (unsigned)o, b = (unsigned)(o >> 32);
        if (f + b < (unsigned)n) return q * per + (d == 0 ? (int)f : n - 1 - (int)b);
    }
    return -1;
}

// a block has taken its last (failed) claim; the last of the grid clears the
// counters (called by one lane of the block)
__device__ inline void claim_done(unsigned long long *claim, int nblocks)
{
    __threadfence();
    if (atomicAdd(&claim[8 * CLAIM_LINE], 1ull) == (unsigned long long)(nblocks - 1)) {
        __threadfence();
        for (int q = 0; q < 9; ++q) atomicExch(&claim[q * CLAIM_LINE], 0ull);
    }
}

// one run of a batched fused step; s[p] is read and s[p ^ 1] written when
// the launch's parity is p
struct FJob {
    Geo g;
    State s[2];
    Coef c;
    int kc;
    FTiles t;
    int tile0;   // first global tile of the run
};
// one launch for all runs (d_jobs in device memory, tile0 ascending);
// tile_y / unroll / xrot of `l` select the kernel variant and the XCD range
// rotation, grid = persistent blocks claiming tiles through `claim`
// (CLAIM_WORDS words of device memory, zero before the first launch; each
// launch leaves them zero; launches that share them must not overlap)
hipError_t launch_step_fused_batch(bool mixed, const Launch &l, const FJob *d_jobs, int njobs, int total,
                                   int grid, int par, unsigned long long *claim, hipStream_t st);
// blocks of the batched kernel that fit on the device at once
int fused_batch_grid(bool mixed, const Launch &l, int device);

// one run of a batched two-stage step (state updated in place); tiles of
// 64 x tile_y points in patch order (patch_tile) within each z chunk
struct SJob {
    Geo g;
    State s;
    Coef c;
    int kc;
    int nbx, nby, nbz;
    int tile0;   // first global tile of the run (velocity and stress stages)
    int fs0;     // first block of the run in the free-surface launch
    int pw, ph;  // tile patches (patch_tile); pw = nbx, ph = 1 is row order
};
// The job of one run, its tiles numbered from tile0 and its free-surface
// blocks from fs0.
static inline SJob split_job(const Geo &g, const State &s, const Coef &c, const Launch &l, int tile0, int fs0)
{
    SJob q;
    q.g = g;
    q.s = s;
    q.c = c;
    q.kc = l.kc;
    q.nbx = (g.nx + TX * l.tile_xw - 1) / (TX * l.tile_xw);
    q.nby = (g.ny + l.tile_y - 1) / l.tile_y;
    q.nbz = l.nchunk;
    q.tile0 = tile0;
    q.fs0 = fs0;
    q.pw = l.spw > 0 ? (l.spw < q.nbx ? l.spw : q.nbx) : q.nbx;
    q.ph = l.sph > 0 ? l.sph : 1;
    return q;
}
static inline int split_tiles(const SJob &q) { return q.nbx * q.nby * q.nbz; }
static inline int split_fs_blocks(const SJob &q) { return ((q.g.nx + TX - 1) / TX) * ((q.g.ny + 3) / 4); }

// velocity, free-surface and stress stages of every run: three launches,
// the first and last over the tile list (grid_v / grid_s blocks, from
// split_batch_grids, or one block per tile when larger); fs_total = sum of
// the runs' free-surface blocks; o32 = offsets32() holds for every run.
// With `claim` (CLAIM_WORDS words, zero before the first launch, left zero
// by each; launches sharing it must not overlap) the blocks claim tiles per
// XCD (claim_tile, ranges rotated by l.xrot); without, the static walk.
hipError_t launch_step_split_batch(bool mixed, const Launch &l, bool o32, const SJob *d_jobs, int njobs,
                                   int total, int fs_total, int grid_v, int grid_s, unsigned long long *claim,
                                   hipStream_t st);
hipError_t launch_split_velocity(bool mixed, const Launch &l, bool o32, const SJob *d_jobs, int njobs, int total,
                                 int grid_v, unsigned long long *claim, hipStream_t st);
hipError_t launch_split_fs(bool mixed, const SJob *d_jobs, int njobs, int fs_total, hipStream_t st);
hipError_t launch_split_stress(bool mixed, const Launch &l, bool o32, const SJob *d_jobs, int njobs, int total,
                               int grid_s, unsigned long long *claim, hipStream_t st);
void split_batch_grids(bool mixed, const Launch &l, int device, int *grid_v, int *grid_s);

// per-run data of the batched source / SGT stage
struct AuxJob {
    State s[2];                 // as FJob: s[p ^ 1] is the state after the step
    Geo g;
    int igreen, npsrc, nst, ntiskp, nsamp_max, nkept;
    const long long *psrc, *ptop;
    const float *stf;
    float vtst, vtst1;
    const long long *kpos;
    const int *kidx;
    const float *sg;
    float *sgt;
    int *counter;               // steps completed before the current one
    float *trace;               // [nt][3] u1 v1 w1 at tpos after each step
    long long tpos;
};
// the source and SGT work of one step for every run, then the counters
hipError_t launch_aux_batch(const AuxJob *d_jobs, int njobs, int max_kept, int par, hipStream_t st);
// RWG mesh arrays (AWP [X][Z][Y], y fastest) -> AWP media file [Z][Y][X][3] x 1000
hipError_t launch_rwg_media(const float *p, const float *s, const float *d, int NX, int NY, int NZ, float *out,
                            hipStream_t st);
// raw-file records g0 .. g0+nrec-1 in SgtHead's RWG order and signs
// (keep[nrec] marks the points the reference writes; navail samples filled)
hipError_t launch_sgt_rwg(const float *sgt, const int *keep, int g0, int nrec, int nsamp_max, int nt_out,
                          int navail, int zsrc, float *out, hipStream_t st);
// copy field `which` (0..19) to dst as [k][j][i] in physical units
hipError_t launch_pack_field(bool mixed, const Geo &g, const State &s, int which, int kext, float *dst,
                             hipStream_t st);
// levels k0+1 .. k0+nk of field `which` to dst as [k - 1 - k0][j][i]
hipError_t launch_pack_levels(bool mixed, const Geo &g, const State &s, int which, int kext, int k0, int nk,
                              float *dst, hipStream_t st);

}  // namespace awp
