// device.hpp -- storage types, device descriptors and reduction helpers shared
// by the gfx950 kernels of the pressure projection path.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace phg {

constexpr int TB = 256;          // threads per block (4 wave64)
constexpr int MAXL = 12;         // multigrid levels (pressure_mg.cpp:141)
constexpr int NPART = 4;         // FP64 partials per block: dq, rr, max|r|, rz

// ---- storage types: FP32, FP16 and BF16 (BF16 kept as raw bits, RNE) ----
struct bf16_t {
    uint16_t u;
};
// coefficient tag: level-0 face coefficients read from the 1-byte face mask
struct Mask01 {};

__device__ __forceinline__ float to_f(float v) { return v; }
__device__ __forceinline__ float to_f(_Float16 v) { return (float)v; }
__device__ __forceinline__ float to_f(bf16_t v) { return __uint_as_float((uint32_t)v.u << 16); }

template <class T> __device__ __forceinline__ T from_f(float v);
template <> __device__ __forceinline__ float from_f<float>(float v) { return v; }
template <> __device__ __forceinline__ _Float16 from_f<_Float16>(float v) { return (_Float16)v; }
template <> __device__ __forceinline__ bf16_t from_f<bf16_t>(float v) {
    uint32_t b = __float_as_uint(v);
    b += 0x7FFFu + ((b >> 16) & 1u);   // round to nearest even
    bf16_t r;
    r.u = (uint16_t)(b >> 16);
    return r;
}

// a value as it reads back after a store in T (FP32: itself)
template <class T> __device__ __forceinline__ float zq(float v) { return to_f(from_f<T>(v)); }

// Global-memory access.  Pointers read from a descriptor are generic, and a
// generic access compiles to a flat instruction, which counts against both the
// vector-memory and the LDS wait counters: every LDS wait would then also wait
// for outstanding loads.  These casts give the global address space, so the
// compiler emits global_load / global_store.
// Indices are 32-bit unsigned (every array here has fewer than 2^31
// elements): with the base pointer uniform across the wave, the compiler then
// forms the address as a scalar base plus a 32-bit vector offset instead of
// 64-bit vector arithmetic for every access.
// The byte offset is formed in 32 bits, so the address is base + zero-extended
// offset, which the scalar-base form of the load/store encodes directly.
#define PHG_GLOBAL __attribute__((address_space(1)))
template <class T> __device__ __forceinline__ const PHG_GLOBAL T* gptr(const T* p, unsigned i) {
    return (const PHG_GLOBAL T*)((const PHG_GLOBAL char*)p + i * (unsigned)sizeof(T));
}
template <class T> __device__ __forceinline__ PHG_GLOBAL T* gptr(T* p, unsigned i) {
    return (PHG_GLOBAL T*)((PHG_GLOBAL char*)p + i * (unsigned)sizeof(T));
}
template <class T> __device__ __forceinline__ T gld(const T* p, unsigned i) { return *gptr(p, i); }
template <class T> __device__ __forceinline__ void gst(T* p, unsigned i, T v) { *gptr(p, i) = v; }
// two adjacent floats in one 8-byte access; two halves in one 4-byte access
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
template <class T> struct pair_of;
template <> struct pair_of<float> { using type = f32x2; };
template <> struct pair_of<_Float16> { using type = f16x2; };
// BF16 is a struct of raw bits: access the bits
template <> __device__ __forceinline__ bf16_t gld(const bf16_t* p, unsigned i) {
    bf16_t r;
    r.u = *gptr(reinterpret_cast<const uint16_t*>(p), i);
    return r;
}
template <> __device__ __forceinline__ void gst(bf16_t* p, unsigned i, bf16_t v) {
    *gptr(reinterpret_cast<uint16_t*>(p), i) = v.u;
}

template <class T> __device__ __forceinline__ float ld(const void* p, unsigned i) {
    return to_f(gld(static_cast<const T*>(p), i));
}
template <class T> __device__ __forceinline__ void st(void* p, unsigned i, float v) {
    gst(static_cast<T*>(p), i, from_f<T>(v));
}
// Streaming stores for the level-0 vectors: each pass writes them once and
// the next kernel reads them back from HBM (a batch's level-0 vectors far
// exceed L2 and the last-level cache), so they are marked evict-first.
template <class T> __device__ __forceinline__ void gst_nt(T* p, unsigned i, T v) {
    __builtin_nontemporal_store(v, gptr(p, i));
}
template <> __device__ __forceinline__ void gst_nt(bf16_t* p, unsigned i, bf16_t v) {
    __builtin_nontemporal_store(v.u, gptr(reinterpret_cast<uint16_t*>(p), i));
}
template <class T> __device__ __forceinline__ void st_nt(void* p, unsigned i, float v) {
    gst_nt(static_cast<T*>(p), i, from_f<T>(v));
}

// ELL column and value of a coarse level in one load: (column << 16 | FP16
// value bits) for the FP16 coefficient type, (column, FP32 value bits) for FP32
template <class CT> struct CVPack;
template <> struct CVPack<_Float16> {
    using T = uint32_t;
    static __device__ __forceinline__ int col(T p) { return (int)(p >> 16); }
    static __device__ __forceinline__ float val(T p) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(p & 0xFFFFu)); }
};
template <> struct CVPack<float> {
    using T = uint64_t;   // column in the low word, value bits in the high word
    static __device__ __forceinline__ int col(T p) { return (int)(uint32_t)p; }
    static __device__ __forceinline__ float val(T p) { return __uint_as_float((uint32_t)(p >> 32)); }
};

// ---- descriptors -----------------------------------------------------------
// One multigrid level of one system.  Operator arrays belong to the grid and
// are shared by every system on it; x, b, r, t are the system's own.
struct DLevel {
    int n, nx, ny, nz;            // nx.. level 0 only
    int ellk, pad;
    const float* ax32;            // level 0 masked face coefficients, FP32 (outer matvec)
    const float* ay32;
    const float* az32;
    const void* axc;              // level 0 coefficients in the V-cycle storage type
    const void* ayc;
    const void* azc;
    const int* col;               // coarse: ELL columns, column-major [ellk][n]
    const void* val;              // coarse: ELL values (V-cycle coefficient type)
    const void* diag;             // coarse: diagonal (V-cycle coefficient type)
    const uint8_t* mask;          // level 0, grids whose coefficients are all 0 or 1: bit d set when
                                  // face d (-x,+x,-y,+y,-z,+z) has coefficient 1; else null
    const int* agg;               // unknown -> next-level unknown (-1: none)
    const uint32_t* aggm;         // level 0 of a 0/1 grid: agg (0xFFFFFF: none) | face mask << 24
    const int* cptr;              // children of next-level unknowns into this level
    const int* cidx;
    const void* cv;               // coarse levels under 65536 rows: ELL column and value in one word
                                  // (CVPack), or null
    const uint32_t* cblk;         // level 0: per level-1 aggregate, its 2x2x2 block's corner cell |
                                  // child mask << 24 (bit dx | dy << 1 | dz << 2), or null
    void* x;                      // work vectors in the V-cycle vector type
    void* b;
    void* r;
    void* t;
};

// scalar slots (FP64) and state slots (int) of one system
// S_ZS: power-of-two scale applied to the V-cycle input (its residual made
// O(1) so FP16/BF16 storage stays in range; exact in FP32); S_ZSN: the next one
// S_RPREV: ||r|| at the start of the current round (restarts must make progress)
// S_PNEXT: the ||r||^2 / ||b||^2 below which the residual's closed-zone means
// are removed next (see k_proj_r)
// S_DS: the power-of-two scale the last direction d is stored with (the S_ZS of
// its z; in FP32 d is stored unscaled); S_ALPHA2, S_DS2: alpha and the scale of
// the direction before it
// S_PW: the factor every prolongation applies to the coarse level's correction:
// the level scale times the solve's coarse-correction weight (4 for weight 1)
enum { S_RZ = 0, S_DQ, S_ALPHA, S_BETA, S_RR, S_BB, S_TOL, S_MAXR, S_INV, S_ZS, S_ZSN, S_RPREV, S_PNEXT, S_DS, S_ALPHA2, S_DS2,
       S_PW, S_NSCAL = 18 };
// ST_XPEND: x += alpha d updates still owed by a stopped system (0, 1 or 2:
// k_dir_apply applies them in pairs, on even iterations)
// ST_FLEX: flexible beta (mixed precision); ST_VC: a CG iteration has run (q is defined)
// ST_ITOT: iterations of earlier rounds (restarts from the true residual)
// ST_RESTART: this round restarts from the recomputed residual (see k_fin_init)
// ST_LAST: how the round before a restart ended (converged or broke down)
// ST_PROJ: removal of the residual's closed-zone means (k_proj): 1 sums due, 2 subtraction due
enum {
    ST_ACTIVE = 0, ST_ITERS, ST_MAXIT, ST_STATUS, ST_RULE, ST_ZEROX, ST_XPEND, ST_FLEX, ST_VC, ST_ITOT, ST_RESTART,
    ST_LAST, ST_PROJ, ST_NSTATE = 16
};

// One system of a batch.  Level 0 is swept by columns: a block takes TB
// consecutive (i, j) columns of one k-chunk of `kchunk` planes, so that
// tiles0 = ceil(nx*ny/TB) * ceil(nz/kchunk) blocks cover the grid.
struct DSys {
    int nl;                       // levels; the coarsest is nl-1
    int nc;                       // coarsest size
    int tiles0;                   // level-0 blocks (stride of the partials)
    int kchunk;                   // planes per level-0 block
    int nu0;                      // level-0 red-black sweeps before and after the coarse correction
    const double* cinv;           // nc*nc FP64 inverse of the grounded coarsest operator (nc <= CP_MIN)
    const double* cinvp;          // nc > CP_MIN: its upper triangle as 64x64 tiles (k_vc_coarse_p)
    int cnt;                      // 64-row strips of a packed inverse (0: cinv is read whole)
    double* cpart;                // [cnt][cnt][64]: block (R, S) of the inverse times strip S of b
    const float* b;               // outer CG vectors, FP32 on level-0 cells
    float* x;
    float* r;                     // residual, double-buffered: r_k lives in r (k even) or r1 (k odd),
    float* r1;                    // so a pass that updates it never overwrites what a neighbour reads
    float* q;
    void* d0;                     // direction, double-buffered: d_k lives in d[k & 1]; z's type, in
    void* d1;                     // the mixed precisions scaled by S_DS
    void* z;                      // preconditioned residual, V-cycle vector type
    void* z2;                     // the pre-smoothed iterate of the fused passes (z's type): the fused
                                  // post-smoothing reads it in its halo rows while it writes z
    const float* invd;            // 1 / diag of level 0 (0 where diag <= 0), FP32
    double* scal;
    int* state;
    double* part;                 // [tiles0][NPART]
    // closed-zone mean removal (src/sim/gasphase.cpp:628-639)
    int nclosed, nztask;
    const int4* ztask;            // (zone, first, end) chunks of the zone cell list
    const int* ztptr;             // tasks of each zone: ztptr[z] .. ztptr[z+1]
    const int* zcells;            // cells sorted by closed zone
    const int* zclosed;           // per cell: closed-zone index or -1
    double* zpart;                // per task
    double* zmean;                // per closed zone
    // the coarsest level by the reference's 20 forward/backward Gauss-Seidel
    // sweeps instead of the dense inverse (k_vc_coarse_gs): the level schedule
    // of each sweep (HostHierarchy::gs_ord / gs_lptr)
    const int* gs_ord;
    const int* gs_lptr;
    int gs_nf, gs_nb;
    DLevel L[MAXL];
};

// ---- small helpers -------------------------------------------------------
__device__ __forceinline__ void cell_ijk(int c, int nx, int ny, int& i, int& j, int& k) {
    const int sxy = nx * ny;
    k = c / sxy;
    const int r = c - k * sxy;
    j = r / nx;
    i = r - j * nx;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Fixed-tree block sum of NV doubles; thread q < NV writes out[q].
template <int NV>
__device__ __forceinline__ void block_sums(double (&v)[NV], double* out) {
    __shared__ double sh[NV][TB / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        v[q] = wave_sum(v[q]);
        if (lane == 0) sh[q][w] = v[q];
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0;
#pragma unroll
        for (int ww = 0; ww < TB / 64; ++ww) s += sh[threadIdx.x][ww];
        out[threadIdx.x] = s;
    }
}

__device__ __forceinline__ float block_max(float v) {
    __shared__ float sh[TB / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    v = wave_max(v);
    if (lane == 0) sh[w] = v;
    __syncthreads();
    float m = sh[0];
#pragma unroll
    for (int ww = 1; ww < TB / 64; ++ww) m = fmaxf(m, sh[ww]);
    return m;
}

// max of non-negative floats through their bit patterns (order independent)
// Max of non-negative floats (their bit patterns order as unsigned ints).  A
// block whose value does not exceed the one already stored skips the atomic:
// thousands of blocks updating one address would otherwise queue at the L2
// (a stale read only costs an unneeded atomic).
__device__ __forceinline__ void atomic_max_nonneg(float* addr, float v) {
    unsigned int* a = reinterpret_cast<unsigned int*>(addr);
    const unsigned int bits = __float_as_uint(v);
    if (bits > __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(a, bits);
}

}  // namespace phg
