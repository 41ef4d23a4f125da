// kernels_pcg.hpp -- batched MG-PCG kernels for gfx950.
//
// Every kernel works on a batch of independent systems.  The (system, tile)
// pair of a block comes from a tile table built at batch creation and
// ordered so that consecutive tiles of a system land on the same XCD (blocks
// are dispatched round-robin over the 8 XCDs, each with its own L2).
// Systems that have stopped leave at once.
//
// Level 0 (the structured grid, where nearly all the bytes are) is swept by
// columns: a thread owns one (i, j) column and walks the planes of its block's
// k-chunk, so consecutive lanes read consecutive cells of a plane and the
// index arithmetic is done once per column.  One CG iteration is three
// level-0 passes on grids with 0/1 coefficients (reference: the CG loop
// src/sim/gasphase.cpp:603-627 and the V-cycle src/sim/pressure_mg.cpp:270-300):
//   k_dir_apply   d = z + beta d_old, q = A d, fused d.q, and the deferred
//                 x += alpha_prev d_old of the previous iteration
//   k_pre_fused   r -= alpha q with |r|^2 and max|r|, the red and black
//                 pre-smoothing half sweeps from zero, and the level-0
//                 residual into the V-cycle scratch
//   ... restriction, coarse levels, coarsest solve, prolongation ...
//   k_post_fused  prolongation, the black and red post-smoothing half sweeps,
//                 and the r.z (and q.z) partials
// Other grids run the half sweeps as separate passes with the same results:
//   k_pre_red     r -= alpha q, fused |r|^2 and max|r|; first pre-smoothing
//                 half sweep from zero (red: r/diag, black: 0)
//   k_rb_black    black half sweep                         (gsColor(1))
//   k_resid0      level-0 residual into the V-cycle scratch (residualGrid)
//   k_post_black  prolongation fused into the black half sweep: a GS update
//                 replaces the cell's own value, so only the red neighbours'
//                 prolonged values are needed, and they are formed on the fly
//   k_post_red    red half sweep fused with the r.z partial
// The V-cycle's right-hand side on level 0 is the CG residual r and its
// iterate is z, so no copy-in/copy-out (pressure_mg.cpp:273,299) is needed.
#pragma once
#include <type_traits>

#include "device.hpp"

namespace phg {

// ---------------------------------------------------------------------------
// level 0 stencil access
// ---------------------------------------------------------------------------
struct Cell {
    int c, i, j, k;
};

// the six face coefficients of a cell (-x, +x, -y, +y, -z, +z): from face
// arrays in the reference's layout (FP32 or FP16), or, for grids whose
// coefficients are all 0 or 1, from the 1-byte face mask (1 B/cell instead of
// 12 B/cell of FP32 faces, exact)
template <class CT>
__device__ __forceinline__ void coefs6(const DLevel& L, const void* ax, const void* ay, const void* az, const Cell& q,
                                       float a[6]) {
    if constexpr (std::is_same<CT, Mask01>::value) {
        const unsigned b = gld(L.mask, q.c);
#pragma unroll
        for (int d = 0; d < 6; ++d) a[d] = (float)((b >> d) & 1u);
    } else {
        const int nx = L.nx, ny = L.ny;
        const int xf = q.c + q.j + ny * q.k;   // i + (nx+1)(j + ny k)
        const int yf = q.c + nx * q.k;         // i + nx (j + (ny+1) k)
        a[0] = ld<CT>(ax, xf);
        a[1] = ld<CT>(ax, xf + 1);
        a[2] = ld<CT>(ay, yf);
        a[3] = ld<CT>(ay, yf + nx);
        a[4] = ld<CT>(az, q.c);
        a[5] = ld<CT>(az, q.c + nx * ny);
    }
}

__device__ __forceinline__ float diag6(const float a[6]) { return a[0] + a[1] + a[2] + a[3] + a[4] + a[5]; }

// a cell's coefficients held in registers between their load and their use
// (the 1-byte mask, or the six values)
template <class CT>
struct CoefRegs {
    float a[6];
    __device__ __forceinline__ void load(const DLevel& L, const void* ax, const void* ay, const void* az,
                                         const Cell& q) {
        coefs6<CT>(L, ax, ay, az, q, a);
    }
    __device__ __forceinline__ void get(float out[6]) const {
#pragma unroll
        for (int d = 0; d < 6; ++d) out[d] = a[d];
    }
};
template <>
struct CoefRegs<Mask01> {
    unsigned m;
    __device__ __forceinline__ void load(const DLevel& L, const void*, const void*, const void*, const Cell& q) {
        m = gld(L.mask, q.c);
    }
    __device__ __forceinline__ void get(float out[6]) const {
#pragma unroll
        for (int d = 0; d < 6; ++d) out[d] = (float)((m >> d) & 1u);
    }
};

// sum over existing neighbours of a_f * v(nb) (a missing neighbour is the
// Dirichlet boundary, value 0); v is a functor of the neighbour index.
// Branch-free: a missing neighbour reads the cell itself with weight 0, so
// all six loads are issued together.
template <class F>
__device__ __forceinline__ float nb_sum_f(const Cell& q, int nx, int ny, int nz, const float a[6], F v) {
    const int sxy = nx * ny;
    const bool e0 = q.i > 0, e1 = q.i < nx - 1, e2 = q.j > 0, e3 = q.j < ny - 1, e4 = q.k > 0, e5 = q.k < nz - 1;
    const float v0 = v(e0 ? q.c - 1 : q.c), v1 = v(e1 ? q.c + 1 : q.c), v2 = v(e2 ? q.c - nx : q.c),
                v3 = v(e3 ? q.c + nx : q.c), v4 = v(e4 ? q.c - sxy : q.c), v5 = v(e5 ? q.c + sxy : q.c);
    return (e0 ? a[0] : 0.0f) * v0 + (e1 ? a[1] : 0.0f) * v1 + (e2 ? a[2] : 0.0f) * v2 + (e3 ? a[3] : 0.0f) * v3 +
           (e4 ? a[4] : 0.0f) * v4 + (e5 ? a[5] : 0.0f) * v5;
}

// CG residual r_k: buffer k & 1.  The pre-smoothing of iteration k reads r_k
// and writes r_{k+1}; the rest of that V-cycle reads r_{k+1}; the V-cycle
// before the first iteration reads r_0 (ST_VC is still 0 then).
// A system's state is the same for the whole block: reading it through
// readfirstlane keeps it (and the vector pointers chosen by it) in scalar
// registers, so accesses through those pointers take the scalar-base form.
__device__ __forceinline__ int st_uni(const DSys& s, int slot) { return __builtin_amdgcn_readfirstlane(s.state[slot]); }
__device__ __forceinline__ float* rbuf(const DSys& s, int k) { return (k & 1) ? s.r1 : s.r; }
__device__ __forceinline__ const float* r_vc(const DSys& s) {
    return st_uni(s, ST_VC) ? rbuf(s, st_uni(s, ST_ITERS) + 1) : s.r;
}

// the column and k-range of a level-0 block
struct ColTile {
    int i, j, c2, k0, k1;
    bool valid;
};
__device__ __forceinline__ ColTile col_tile(const DSys& s, int tile) {
    const DLevel& L = s.L[0];
    const int sxy = L.nx * L.ny;
    const int nct = (sxy + TB - 1) / TB;
    const int ct = tile % nct, kc = tile / nct;
    ColTile t;
    t.c2 = ct * TB + threadIdx.x;
    t.valid = t.c2 < sxy;
    t.j = t.c2 / L.nx;
    t.i = t.c2 - t.j * L.nx;
    t.k0 = kc * s.kchunk;
    t.k1 = min(L.nz, t.k0 + s.kchunk);
    return t;
}

// The partial row of each TB-column tile of a block of NT = TB or 2 TB
// threads: lane 0 of the tile's first wave adds the tile's wave sums in wave
// order, as a TB-thread block adds them, so the partials are bitwise those
// of TB-thread blocks.
template <class F>
__device__ __forceinline__ void tile_rows(const DSys& s, const int2& e, int c0, F write) {
    const int tid = (int)threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (lane != 0 || (w % (TB / 64)) != 0) return;
    const int h = w / (TB / 64);
    if (h > 0 && c0 + h * TB >= s.L[0].nx * s.L[0].ny) return;   // no such column tile
    write(w, s.part + (long)(e.y + h) * NPART);
}

// Coarse right-hand sides are restricted by summing children, so each level's
// is up to 8 times the one above (a point source) while the Galerkin operator
// grows 4 times: 2-byte storage would overflow a few levels down.  Level l+1
// therefore holds its vectors scaled by 1/4 (restriction multiplies by 1/4,
// prolongation by 4).  The scale is a power of two, so FP32 results are
// bitwise those of the unscaled cycle.
// A solve may weight every coarse-grid correction by w (0 < w < 2): the
// prolongation then adds 4 w times the coarse value (S_PW).  Piecewise-constant
// aggregation under-corrects smooth errors, and a weight of 1.8 takes the
// house workload from 29.5 CG iterations to 19; w = 1 is the reference's cycle,
// bit for bit.  The symmetric cycle stays a symmetric positive definite
// preconditioner for 0 < w < 2.
constexpr float VC_LEVEL_SCALE = 4.0f;

// ---------------------------------------------------------------------------
// plane march with LDS-staged lateral neighbours
// ---------------------------------------------------------------------------
// A block owns the TB columns [c0, c0 + TB) of one k-chunk.  For every plane
// k it stages the value of its columns, and of one grid row on either side
// (the -nx / +nx neighbours), in LDS; the four lateral neighbours are then
// read from LDS and the two vertical ones come from a register pipeline along
// k.  Each staged value is loaded from global memory once per plane instead of
// up to seven times.  lds: 2 * (TB + 2 nx) floats, double-buffered so that
// one barrier per plane suffices.
//
// The kernel supplies
//   raw(n)  -> R   the global loads behind the value of cell n (no arithmetic),
//   mid(R&)        loads that depend on raw's (issued a plane before use),
//   fin(R)  -> float  the value from them,
//   pre(q)  -> P   the cell's own operands (coefficients, r, ...),
//   body(k, v(k-1), v(k), v(k+1), &plane[own column], P).
// Raw planes are loaded three planes ahead and operands two ahead, with
// clamped in-bounds indices and no branch around any load, so the compiler's
// waits let each plane's loads overlap the work of the planes before it.
// Halo values beyond the first MARCH_HALO * TB (nx > MARCH_HALO * TB / 2) are
// loaded in the plane itself.
constexpr int MARCH_HALO = 2;
constexpr int MARCH_MAX_NX = 2048;   // LDS plane buffers: 2 (TB + 2 nx) floats

// NT: threads (own columns) of the block; HS: halo slots per thread, enough
// for 2 nx <= HS NT (wider rows take the loop in `step`)
template <int NT = TB, int HS = MARCH_HALO, class RAW, class MID, class FIN, class PRE, class BODY>
__device__ __forceinline__ void march(const DSys& s, const ColTile& t, float* lds, RAW raw, MID mid, FIN fin, PRE pre,
                                      BODY body) {
    const DLevel& L = s.L[0];
    const int nx = L.nx, nz = L.nz, sxy = nx * L.ny;
    const int W = NT + 2 * nx;
    const int c0 = t.c2 - (int)threadIdx.x;
    using R = decltype(raw(0));
    using P = decltype(pre(Cell{}));
    struct Plane {
        R own;
        R h[HS];
    };
    const int cown = t.valid ? t.c2 : sxy - 1;   // any in-bounds column for idle lanes
    int hcol[HS];
    bool hin[HS];
#pragma unroll
    for (int m = 0; m < HS; ++m) {
        const int hh = (int)threadIdx.x + m * NT;
        const int col = hh < nx ? c0 - nx + hh : c0 + NT + (hh - nx);
        hin[m] = hh < 2 * nx && col >= 0 && col < sxy;
        hcol[m] = hin[m] ? col : cown;
    }
    auto kcl = [&](int k) { return min(max(k, 0), nz - 1); };
    auto load = [&](int k, Plane& pl) {
        const unsigned base = (unsigned)sxy * (unsigned)kcl(k);
        pl.own = raw(cown + base);
#pragma unroll
        for (int m = 0; m < HS; ++m) pl.h[m] = raw(hcol[m] + base);
    };
    auto mid_plane = [&](Plane& pl) {
        mid(pl.own);
#pragma unroll
        for (int m = 0; m < HS; ++m) mid(pl.h[m]);
    };
    auto fin_own = [&](int k, const R& r) {
        const float v = fin(r);
        return (t.valid && k >= 0 && k < nz) ? v : 0.0f;
    };
    auto fin_halo = [&](int k, const Plane& pl, float (&h)[HS]) {
#pragma unroll
        for (int m = 0; m < HS; ++m) {
            const float v = fin(pl.h[m]);
            h[m] = (hin[m] && k < nz) ? v : 0.0f;
        }
    };
    // idle lanes (columns past the grid) load the last column's operands
    const int oi = t.valid ? t.i : nx - 1, oj = t.valid ? t.j : L.ny - 1;
    auto opnd = [&](int k) { return pre(Cell{cown + sxy * kcl(k), oi, oj, kcl(k)}); };

    // Register rings, unrolled by three so that no register that a load is
    // still filling is ever copied (a copy would make the wave wait for it):
    // raw planes X0..X2 hold k+1, k+2 and receive k+3; operands P0..P2 hold k,
    // k+1 and receive k+2.
    Plane X0, X1, X2;
    float vprev, vcur, hfin[HS];
    {
        Plane Z;
        load(t.k0 - 1, Z);
        mid_plane(Z);
        vprev = fin_own(t.k0 - 1, Z.own);
        load(t.k0, Z);
        mid_plane(Z);
        vcur = fin_own(t.k0, Z.own);
        fin_halo(t.k0, Z, hfin);
    }
    load(t.k0 + 1, X0);
    load(t.k0 + 2, X1);
    mid_plane(X0);
    P P0 = opnd(t.k0), P1 = opnd(t.k0 + 1), P2{};
    auto step = [&](int k, const Plane& A, Plane& B, Plane& Cdst, const P& pcur, P& pdst) {
        float* buf = lds + (k & 1) * W;
        buf[nx + threadIdx.x] = vcur;   // 0 for columns past the grid
#pragma unroll
        for (int m = 0; m < HS; ++m) {
            const int hh = (int)threadIdx.x + m * NT;
            if (hh < 2 * nx) buf[hh < nx ? hh : NT + hh] = hfin[m];
        }
        for (int hh = (int)threadIdx.x + HS * NT; hh < 2 * nx; hh += NT) {
            const int col = hh < nx ? c0 - nx + hh : c0 + NT + (hh - nx);
            float v = 0.0f;
            if (col >= 0 && col < sxy) {
                R rv = raw((unsigned)col + (unsigned)sxy * (unsigned)k);
                mid(rv);
                v = fin(rv);
            }
            buf[hh < nx ? hh : NT + hh] = v;
        }
        __syncthreads();
        const float vnext = fin_own(k + 1, A.own);
        float hnext[HS];
        fin_halo(k + 1, A, hnext);
        // no branch around loads (a branch would merge the wait counters and
        // force an early wait): past the chunk they repeat the last plane
        // needed, which the caches serve
        load(min(k + 3, t.k1), Cdst);
        pdst = opnd(min(k + 2, t.k1 - 1));
        mid_plane(B);   // plane k + 2
        if (t.valid) body(k, vprev, vcur, vnext, (const float*)(buf + nx + threadIdx.x), pcur);
        vprev = vcur;
        vcur = vnext;
#pragma unroll
        for (int m = 0; m < HS; ++m) hfin[m] = hnext[m];
    };
    // whole groups of three steps in the loop, the rest after it: no step
    // inside the loop is conditional, so the rings keep their registers
    // around the loop and no load is waited on before its plane is used
    int k = t.k0;
    for (; k + 2 < t.k1; k += 3) {
        step(k, X0, X1, X2, P0, P2);
        step(k + 1, X1, X2, X0, P1, P0);
        step(k + 2, X2, X0, X1, P2, P1);
    }
    if (k < t.k1) step(k, X0, X1, X2, P0, P2);
    if (k + 1 < t.k1) step(k + 1, X1, X2, X0, P1, P0);
}

// a_f-weighted sum of the six neighbours: lateral ones at p[-1], p[+1],
// p[-nx], p[+nx] (the staged plane), vertical ones vp (k-1) and vn (k+1);
// missing neighbours (Dirichlet boundary) weigh 0
__device__ __forceinline__ float lat_sum(const Cell& q, int nx, int ny, int nz, const float a[6], const float* p,
                                         float vp, float vn) {
    const bool e0 = q.i > 0, e1 = q.i < nx - 1, e2 = q.j > 0, e3 = q.j < ny - 1, e4 = q.k > 0, e5 = q.k < nz - 1;
    return (e0 ? a[0] : 0.0f) * p[-1] + (e1 ? a[1] : 0.0f) * p[1] + (e2 ? a[2] : 0.0f) * p[-nx] +
           (e3 ? a[3] : 0.0f) * p[nx] + (e4 ? a[4] : 0.0f) * vp + (e5 ? a[5] : 0.0f) * vn;
}

// bit f set when face f of the cell at (i, j, k) has a neighbour cell
__device__ __forceinline__ unsigned edge_bits(int i, int j, int k, int nx, int ny, int nz) {
    return (i > 0 ? 1u : 0u) | (i < nx - 1 ? 2u : 0u) | (j > 0 ? 4u : 0u) | (j < ny - 1 ? 8u : 0u) |
           (k > 0 ? 16u : 0u) | (k < nz - 1 ? 32u : 0u);
}

// lat_sum for 0/1 coefficients given as bits (face mask and edge bits
// already combined).  The products are exact and the sum runs in the same
// face order, so the result is bit-identical to lat_sum's.
__device__ __forceinline__ float lat_bits(unsigned b, const float* p, int nx, float vp, float vn) {
    return (float)(b & 1u) * p[-1] + (float)((b >> 1) & 1u) * p[1] + (float)((b >> 2) & 1u) * p[-nx] +
           (float)((b >> 3) & 1u) * p[nx] + (float)((b >> 4) & 1u) * vp + (float)((b >> 5) & 1u) * vn;
}

// x / d for a face count d in 1..6, correctly rounded without a divide
// sequence: the reciprocal of d rounded to nearest by one Newton step from
// the hardware reciprocal (any start within two units in the last place
// lands on it), then the quotient x (1/d) corrected by its exact FMA
// residual (Markstein: a faithful quotient q plus (1/d)(x - q d) rounds to
// x / d).  Equal to x / d for every float x whose quotient is normal
// (checked exhaustively for d = 1..6).
__device__ __forceinline__ float div_count(float x, float d) {
    const float r0 = __builtin_amdgcn_rcpf(d);
    const float r = fmaf(r0, fmaf(-d, r0, 1.0f), r0);
    const float q0 = x * r;
    return fmaf(fmaf(-q0, d, x), r, q0);
}

// Lookup tables of the fused passes, built in LDS by each block.  For a
// six-bit face set (face mask and edge bits combined), the 0/1 weights of
// lat_bits as bytes, faces 0-3 in .x and 4-5 in .y, which convert to floats
// without a bit extract each; for a face-mask byte, its face count d and
// RN(1/d) exactly as div_count forms them.  The weights, d and 1/d are the
// same floats, and the sums run in the same order, so results are bitwise
// those of lat_bits and div_count.
struct FuseLut {
    uint2 w[64];
    float2 d[256];
};
__device__ __forceinline__ void fuse_lut_init(FuseLut& T, int tid, int nt) {
    for (int q = tid; q < 64; q += nt) {
        unsigned lo = 0, hi = 0;
#pragma unroll
        for (int f = 0; f < 4; ++f) lo |= ((q >> f) & 1u) << (8 * f);
        hi = ((q >> 4) & 1u) | (((q >> 5) & 1u) << 8);
        T.w[q] = make_uint2(lo, hi);
    }
    for (int q = tid; q < 256; q += nt) {
        const float d = (float)__popc(q);
        float r = 0.0f;
        if (q != 0) {
            const float r0 = __builtin_amdgcn_rcpf(d);
            r = fmaf(r0, fmaf(-d, r0, 1.0f), r0);
        }
        T.d[q] = make_float2(d, r);
    }
}
__device__ __forceinline__ float lat_lut(const FuseLut& T, unsigned b, const float* p, int nx, float vp, float vn) {
    const uint2 w = T.w[b];
    return (float)(w.x & 255u) * p[-1] + (float)((w.x >> 8) & 255u) * p[1] + (float)((w.x >> 16) & 255u) * p[-nx] +
           (float)(w.x >> 24) * p[nx] + (float)(w.y & 255u) * vp + (float)(w.y >> 8) * vn;
}
// div_count(x, d) given r = RN(1/d)
__device__ __forceinline__ float div_rn(float x, float d, float r) {
    const float q0 = x * r;
    return fmaf(fmaf(-q0, d, x), r, q0);
}

// a cell's diagonal, neighbour sum and division by the diagonal, from its
// coefficient registers: the six values in general, the mask bits for a
// 0/1 grid (same results, fewer instructions).  The division stays
// correctly rounded: the reference's gsColor divides, and a plain product
// with the reciprocal face count, one unit in the last place away, made the
// FP32 iteration diverge on point-source problems with several closed zones.
template <class CT>
__device__ __forceinline__ float cell_diag(const CoefRegs<CT>& cf) {
    if constexpr (std::is_same<CT, Mask01>::value)
        return (float)__popc(cf.m);
    else
        return diag6(cf.a);
}
template <class CT>
__device__ __forceinline__ float cell_lat(const CoefRegs<CT>& cf, const Cell& q, int nx, int ny, int nz,
                                          const float* p, float vp, float vn) {
    if constexpr (std::is_same<CT, Mask01>::value)
        return lat_bits(cf.m & edge_bits(q.i, q.j, q.k, nx, ny, nz), p, nx, vp, vn);
    else
        return lat_sum(q, nx, ny, nz, cf.a, p, vp, vn);
}
template <class CT>
__device__ __forceinline__ float cell_div(float x, float dg) {
    if constexpr (std::is_same<CT, Mask01>::value)
        return div_count(x, dg);
    else
        return x / dg;
}

// a cell's V-cycle coefficients and CG residual, loaded ahead by the march
template <class CT>
struct CoefR {
    CoefRegs<CT> cf;
    float r;
};
template <class CT>
__device__ __forceinline__ auto cr_pre(const DSys& s, const DLevel& L) {
    const float* rv = r_vc(s);
    return [&L, rv](const Cell& q) {
        CoefR<CT> o;
        o.cf.load(L, L.axc, L.ayc, L.azc, q);
        o.r = gld(rv, q.c);
        return o;
    };
}

#define PHG_SYS_PROLOGUE                  \
    const int2 e = tab[blockIdx.x];       \
    const DSys& s = sys[e.x];             \
    if (s.state[ST_ACTIVE] == 0) return;

// ---------------------------------------------------------------------------
// outer CG kernels (FP32 vectors, FP64 reductions)
// ---------------------------------------------------------------------------
// r = b - A x on gas cells (0 elsewhere), or x = 0 and r = b when starting
// from zero; partials |b|^2 (slot 0), |r|^2 (1), max|r| (2)
template <class OT>
// hi: the residual of a restart, with every product of the FP32 x accumulated
// in FP64 and rounded once (iterative refinement: a residual formed in FP32
// is rounding noise once it falls to eps |A| |x|)
__global__ void __launch_bounds__(TB) k_resid_init(const DSys* __restrict__ sys, const int2* __restrict__ tab, int hi) {
    PHG_SYS_PROLOGUE
    const DLevel& L = s.L[0];
    const ColTile t = col_tile(s, e.y);
    const int sxy = L.nx * L.ny;
    double v[2] = {0.0, 0.0};
    double mag2[1] = {0.0};   // hi: sum of (|b| + |dg x| + sum |a x_nb|)^2, the size of the terms
    float m = 0.0f;
    if (t.valid) {
        const bool zx = s.state[ST_ZEROX] != 0;
        for (int k = t.k0; k < t.k1; ++k) {
            const Cell q{t.c2 + sxy * k, t.i, t.j, k};
            float a[6];
            coefs6<OT>(L, L.ax32, L.ay32, L.az32, q, a);
            const float dg = diag6(a);
            const float b = s.b[q.c];
            float r = 0.0f;
            if (zx) {
                s.x[q.c] = 0.0f;
                if (dg > 0) r = b;
            } else if (dg > 0 && hi) {
                const int sx = L.nx * L.ny;
                const bool e0 = q.i > 0, e1 = q.i < L.nx - 1, e2 = q.j > 0, e3 = q.j < L.ny - 1, e4 = q.k > 0,
                           e5 = q.k < L.nz - 1;
                const double xc = (double)dg * (double)gld(s.x, q.c);
                double acc = (double)b - xc, mg = fabs((double)b) + fabs(xc), tm;
                if (e0) { tm = (double)a[0] * (double)gld(s.x, q.c - 1); acc += tm; mg += fabs(tm); }
                if (e1) { tm = (double)a[1] * (double)gld(s.x, q.c + 1); acc += tm; mg += fabs(tm); }
                if (e2) { tm = (double)a[2] * (double)gld(s.x, q.c - L.nx); acc += tm; mg += fabs(tm); }
                if (e3) { tm = (double)a[3] * (double)gld(s.x, q.c + L.nx); acc += tm; mg += fabs(tm); }
                if (e4) { tm = (double)a[4] * (double)gld(s.x, q.c - sx); acc += tm; mg += fabs(tm); }
                if (e5) { tm = (double)a[5] * (double)gld(s.x, q.c + sx); acc += tm; mg += fabs(tm); }
                r = (float)acc;
                mag2[0] += mg * mg;
            } else if (dg > 0) {
                r = b - (dg * s.x[q.c] - nb_sum_f(q, L.nx, L.ny, L.nz, a, [&](int n) { return s.x[n]; }));
            }
            s.r[q.c] = r;
            v[0] += (double)b * b;
            v[1] += (double)r * r;
            m = fmaxf(m, fabsf(r));
        }
    }
    double* out = s.part + (long)e.y * NPART;
    block_sums<2>(v, out);
    m = block_max(m);
    if (threadIdx.x == 0) out[2] = m;
    if (hi) {
        __syncthreads();
        block_sums<1>(mag2, out + 3);
    }
}

// d_k = z + beta d_{k-1} (d_0 = z), q = A d_k with the partial d.q (slot 0)
// (gasphase.cpp:620-622 and the apply of :606), and on even k >= 2 the two
// previous iterations' x += alpha d (gasphase.cpp:612), deferred to this pass:
// x = (x + alpha_{k-2} d_{k-2}) + alpha_{k-1} d_{k-1}, the same two FMAs in the
// same order as one update per iteration, with x read and written every
// second iteration only.  d is double-buffered because neighbours read
// d_{k-1} while d_k is written; d_{k-2}, in the buffer d_k goes to, is read
// at the cell before d_k is stored there.
// d is stored in z's type: FP32, or in the mixed precisions FP16 scaled by
// the power of two of its z (S_DS once stored), so its values stay near an
// RMS of 1.  q = A d and x += alpha d use d as stored, so x and r stay
// consistent; only the direction itself is rounded, which the flexible beta
// of the mixed precisions absorbs like the rounding of z.
template <class OT, class ZT, int NT = TB, int HS = MARCH_HALO>
__global__ void __launch_bounds__(NT) k_dir_apply(const DSys* __restrict__ sys, const int2* __restrict__ tab) {
    PHG_SYS_PROLOGUE
    const DLevel& L = s.L[0];
    const ColTile t = col_tile(s, e.y);
    const int sxy = L.nx * L.ny;
    const int it = st_uni(s, ST_ITERS);
    const bool first = it == 0;
    constexpr bool scaled = !std::is_same<ZT, float>::value;
    ZT* __restrict__ dnew = static_cast<ZT*>((it & 1) ? s.d1 : s.d0);
    const ZT* __restrict__ dold = static_cast<const ZT*>((it & 1) ? s.d0 : s.d1);
    const bool xup = it >= 2 && !(it & 1);   // the two owed x updates are applied here
    const float beta = (float)s.scal[S_BETA], alpha = (float)s.scal[S_ALPHA], alpha2 = (float)s.scal[S_ALPHA2];
    const void* z = s.z;
    const float izs = (float)(1.0 / s.scal[S_ZS]);   // z is stored scaled by S_ZS (a power of two)
    const float ds = scaled ? (float)s.scal[S_ZS] : 1.0f, ids = scaled ? izs : 1.0f;   // d_k's scale
    const float idso = scaled && !first ? (float)(1.0 / s.scal[S_DS]) : 1.0f;         // d_{k-1}'s
    const float idso2 = scaled && xup ? (float)(1.0 / s.scal[S_DS2]) : 1.0f;          // d_{k-2}'s
    // d = z + beta d_old (d = z on the first iteration, where d_old is not read),
    // as it reads back once stored
    auto raw = [&](unsigned n) { return float2{ld<ZT>(z, n), ld<ZT>(dold, n)}; };
    auto fin = [&](const float2& v) {
        const float d = first ? v.x * izs : fmaf(beta, v.y * idso, v.x * izs);
        return scaled ? zq<ZT>(d * ds) * ids : d;
    };
    extern __shared__ float lds[];
    double acc = 0.0;
    struct Op {
        CoefRegs<OT> cf;
        float x, dold, dold2;
    };
    auto pre = [&](const Cell& q) {
        Op o;
        o.cf.load(L, L.ax32, L.ay32, L.az32, q);
        if (xup) {
            o.x = gld(s.x, q.c);
            o.dold = ld<ZT>(dold, q.c) * idso;
            o.dold2 = ld<ZT>(dnew, q.c) * idso2;
        }
        return o;
    };
    march<NT, HS>(s, t, lds, raw, [](auto&) {}, fin, pre, [&](int k, float vp, float dc, float vn, const float* p, const Op& o) {
        const Cell q{t.c2 + sxy * k, t.i, t.j, k};
        const float dg = cell_diag(o.cf);
        float y = 0.0f;
        if (dg > 0) y = dg * dc - cell_lat(o.cf, q, L.nx, L.ny, L.nz, p, vp, vn);
        if (xup) gst_nt(s.x, q.c, fmaf(alpha, o.dold, fmaf(alpha2, o.dold2, o.x)));
        st_nt<ZT>(dnew, q.c, dc * ds);   // exact: dc * ds is a ZT value
        gst_nt(s.q, q.c, y);
        acc += (double)y * dc;
    });
    if constexpr (NT == TB) {
        double v[1] = {acc};
        block_sums<1>(v, s.part + (long)e.y * NPART);
    } else {
        __shared__ double sh[NT / 64];
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        const double v = wave_sum(acc);
        if (lane == 0) sh[w] = v;
        __syncthreads();
        tile_rows(s, e, t.c2 - (int)threadIdx.x, [&](int w0, double* out) {
            double a = 0.0;
#pragma unroll
            for (int ww = 0; ww < TB / 64; ++ww) a += sh[w0 + ww];
            out[0] = a;
        });
    }
}

// k_dir_apply for 0/1 grids with an even plane size, on column pairs: a
// block of TB threads owns the 2 TB columns [c0, c0 + 2 TB) of one k-chunk
// (the even column tiles, like the 2 TB-thread form) and each lane the two
// adjacent columns c0 + 2 tid and c0 + 2 tid + 1.  A pair starts at an even
// cell index, so z, d_old and d_new are one 4-byte access per lane in FP16
// (8 bytes in FP32), q one 8-byte access and the face masks one 2-byte
// access: half the memory instructions of one column per lane, each covering
// twice the bytes.  A march over FP16 vectors with one 2-byte access per lane
// moves bytes at about half the rate of 4-byte accesses, so the pairing is
// what lets the FP16 direction pay.
// The staged plane holds the own columns and nxe = nx rounded up to even
// columns on either side (pairs stay aligned for an odd nx); the halo needs
// nxe <= TB.  The d.q terms of each column are summed along k as in
// k_dir_apply, then handed through LDS to the lane that owns the column in
// the 2 TB-thread form, whose wave and tile sums they follow, so the partials,
// and with them every result, are bitwise those of k_dir_apply.  x is read and
// written as a pair when the caller's vector is 8-byte aligned, else one cell
// at a time (two 4-byte stores per lane write each line in two halves).
template <class ZT>
__global__ void __launch_bounds__(TB) k_dir_apply_pair(const DSys* __restrict__ sys, const int2* __restrict__ tab) {
    PHG_SYS_PROLOGUE
    using P2 = typename pair_of<ZT>::type;
    const DLevel& L = s.L[0];
    const int nx = L.nx, ny = L.ny, nz = L.nz, sxy = nx * ny;
    const int nct = (sxy + TB - 1) / TB;
    const int tid = (int)threadIdx.x;
    const int c0 = (e.y % nct) * TB;   // an even column tile: a multiple of 2 TB
    const int kc = e.y / nct;
    const int k0 = kc * s.kchunk, k1 = min(nz, k0 + s.kchunk);
    const int nxe = (nx + 1) & ~1;
    const int W = 2 * TB + 2 * nxe;
    const int ca = c0 + 2 * tid;                 // the lane's first column
    const bool valid = ca < sxy;                 // sxy is even: both columns or neither
    const int cown = valid ? ca : sxy - 2;
    // the lane's halo pair: left of the own columns for hh < nxe / 2, right after
    const int hh = tid;
    const int hu = hh < nxe / 2 ? hh : hh + TB;
    const int hcol = c0 - nxe + 2 * hu;
    const bool hin = hh < nxe && hcol >= 0 && hcol < sxy;
    const int hcl = hin ? hcol : cown;
    const int hslot = 2 * hu;                    // its W index
    // (i, j) edge bits of the two own cells (the second may start the next row)
    unsigned ijb[2];
    int ci[2], cj[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int c = cown + h;
        cj[h] = c / nx;
        ci[h] = c - cj[h] * nx;
        ijb[h] = (ci[h] > 0 ? 1u : 0u) | (ci[h] < nx - 1 ? 2u : 0u) | (cj[h] > 0 ? 4u : 0u) | (cj[h] < ny - 1 ? 8u : 0u);
    }

    const int it = st_uni(s, ST_ITERS);
    const bool first = it == 0;
    constexpr bool scaled = !std::is_same<ZT, float>::value;
    ZT* __restrict__ dnew = static_cast<ZT*>((it & 1) ? s.d1 : s.d0);
    const ZT* __restrict__ dold = static_cast<const ZT*>((it & 1) ? s.d0 : s.d1);
    const bool xup = it >= 2 && !(it & 1);
    const bool xpair = (reinterpret_cast<uintptr_t>(s.x) & 7u) == 0;   // the same for the whole block
    const float beta = (float)s.scal[S_BETA], alpha = (float)s.scal[S_ALPHA], alpha2 = (float)s.scal[S_ALPHA2];
    const ZT* __restrict__ z = static_cast<const ZT*>(s.z);
    const float izs = (float)(1.0 / s.scal[S_ZS]);
    const float ds = scaled ? (float)s.scal[S_ZS] : 1.0f, ids = scaled ? izs : 1.0f;
    const float idso = scaled && !first ? (float)(1.0 / s.scal[S_DS]) : 1.0f;
    const float idso2 = scaled && xup ? (float)(1.0 / s.scal[S_DS2]) : 1.0f;
    auto fin1 = [&](float zv, float dv) {
        const float d = first ? zv * izs : fmaf(beta, dv * idso, zv * izs);
        return scaled ? zq<ZT>(d * ds) * ids : d;
    };
    struct Raw {
        P2 z, d, hz, hd;
    };
    auto load = [&](int k, Raw& X) {
        const unsigned base = (unsigned)sxy * (unsigned)min(max(k, 0), nz - 1);
        const unsigned io = (cown + base) >> 1, ih = (hcl + base) >> 1;
        X.z = gld(reinterpret_cast<const P2*>(z), io);
        X.d = gld(reinterpret_cast<const P2*>(dold), io);
        X.hz = gld(reinterpret_cast<const P2*>(z), ih);
        X.hd = gld(reinterpret_cast<const P2*>(dold), ih);
    };
    auto fin_own = [&](int k, const Raw& X) {
        const bool in = valid && k >= 0 && k < nz;
        return float2{in ? fin1((float)X.z.x, (float)X.d.x) : 0.0f, in ? fin1((float)X.z.y, (float)X.d.y) : 0.0f};
    };
    auto fin_halo = [&](int k, const Raw& X) {
        const bool in = hin && k < nz;
        return float2{in ? fin1((float)X.hz.x, (float)X.hd.x) : 0.0f, in ? fin1((float)X.hz.y, (float)X.hd.y) : 0.0f};
    };
    struct Op {
        uint16_t m;   // face masks of the pair, first cell in bits 0-7
        float x[2], d2[2];
    };
    auto opnd = [&](int k) {
        Op o;
        const unsigned c = cown + (unsigned)sxy * (unsigned)min(max(k, 0), nz - 1);
        o.m = gld(reinterpret_cast<const uint16_t*>(L.mask), c >> 1);
        if (xup) {
            const P2 d2 = gld(reinterpret_cast<const P2*>(dnew), c >> 1);
            if (xpair) {
                const f32x2 xv = gld(reinterpret_cast<const f32x2*>(s.x), c >> 1);
                o.x[0] = xv.x;
                o.x[1] = xv.y;
            } else {
                o.x[0] = gld(s.x, c);
                o.x[1] = gld(s.x, c + 1);
            }
            o.d2[0] = (float)d2.x * idso2;
            o.d2[1] = (float)d2.y * idso2;
        }
        return o;
    };
    extern __shared__ float lds[];
    double acc[2] = {0.0, 0.0};
    Raw X0, X1, X2;
    float2 vprev, vcur, hcur;
    P2 dcur;   // d_old of the own pair at plane k, as stored
    {
        Raw Z;
        load(k0 - 1, Z);
        vprev = fin_own(k0 - 1, Z);
        load(k0, Z);
        vcur = fin_own(k0, Z);
        hcur = fin_halo(k0, Z);
        dcur = Z.d;
    }
    load(k0 + 1, X0);
    load(k0 + 2, X1);
    Op P0 = opnd(k0), P1 = opnd(k0 + 1), P2r{};
    auto step = [&](int k, const Raw& A, Raw& Cdst, const Op& o, Op& pdst) {
        float* buf = lds + (k & 1) * W;
        *reinterpret_cast<float2*>(buf + nxe + 2 * tid) = vcur;
        if (hh < nxe) *reinterpret_cast<float2*>(buf + hslot) = hcur;
        __syncthreads();
        const float2 vnext = fin_own(k + 1, A), hnext = fin_halo(k + 1, A);
        const P2 dnext = A.d;
        load(min(k + 3, k1), Cdst);
        pdst = opnd(min(k + 2, k1 - 1));
        if (valid) {
            const unsigned kb = (k > 0 ? 16u : 0u) | (k < nz - 1 ? 32u : 0u);
            const float* p = buf + nxe + 2 * tid;
            const float vc[2] = {vcur.x, vcur.y}, vp[2] = {vprev.x, vprev.y}, vn[2] = {vnext.x, vnext.y};
            const float dd[2] = {(float)dcur.x * idso, (float)dcur.y * idso};
            const unsigned c = (unsigned)cown + (unsigned)sxy * (unsigned)k;
            float y[2], xn[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const unsigned m = (h == 0 ? o.m : (unsigned)o.m >> 8) & 255u;
                const float dg = (float)__popc(m);
                y[h] = 0.0f;
                if (dg > 0) y[h] = dg * vc[h] - lat_bits(m & (ijb[h] | kb), p + h, nx, vp[h], vn[h]);
                xn[h] = fmaf(alpha, dd[h], fmaf(alpha2, o.d2[h], o.x[h]));
                acc[h] += (double)y[h] * vc[h];
            }
            if (xup) {
                if (xpair) {
                    gst_nt(reinterpret_cast<f32x2*>(s.x), c >> 1, f32x2{xn[0], xn[1]});
                } else {
                    gst_nt(s.x, c, xn[0]);
                    gst_nt(s.x, c + 1, xn[1]);
                }
            }
            P2 dst;
            dst.x = (ZT)(vc[0] * ds);   // exact: vc * ds is a ZT value
            dst.y = (ZT)(vc[1] * ds);
            gst_nt(reinterpret_cast<P2*>(dnew), c >> 1, dst);
            gst_nt(reinterpret_cast<f32x2*>(s.q), c >> 1, f32x2{y[0], y[1]});
        }
        vprev = vcur;
        vcur = vnext;
        hcur = hnext;
        dcur = dnext;
    };
    int k = k0;
    for (; k + 2 < k1; k += 3) {
        step(k, X0, X2, P0, P2r);
        step(k + 1, X1, X0, P1, P0);
        step(k + 2, X2, X1, P2r, P1);
    }
    if (k < k1) step(k, X0, X2, P0, P2r);
    if (k + 1 < k1) step(k + 1, X1, X0, P1, P0);
    // partials: column c0 + m is thread m of the 2 TB-thread form; its wave
    // sums are taken over the same lanes and added per TB-column tile in
    // wave order
    __shared__ double colacc[2 * TB];
    __shared__ double sh[2 * TB / 64];
    colacc[2 * tid] = acc[0];
    colacc[2 * tid + 1] = acc[1];
    __syncthreads();
    const int lane = tid & 63, w = tid >> 6;
    const double v0 = wave_sum(colacc[tid]), v1 = wave_sum(colacc[tid + TB]);
    if (lane == 0) {
        sh[w] = v0;
        sh[TB / 64 + w] = v1;
    }
    __syncthreads();
    if (lane == 0 && w < 2 && (w == 0 || c0 + TB < sxy)) {
        double a = 0.0;
#pragma unroll
        for (int ww = 0; ww < TB / 64; ++ww) a += sh[w * (TB / 64) + ww];
        s.part[(long)(e.y + w) * NPART] = a;
    }
}

// after the loop: the x += alpha d updates systems that stopped still owe
// (ST_XPEND: the last direction, d in buffer (it - 1) & 1 with S_ALPHA and
// S_DS, and for 2 first the one before it, with S_ALPHA2 and S_DS2)
template <class ZT>
__global__ void __launch_bounds__(TB) k_x_flush(const DSys* __restrict__ sys, const int2* __restrict__ tab) {
    const int2 e = tab[blockIdx.x];
    const DSys& s = sys[e.x];
    const int owed = s.state[ST_XPEND];
    if (owed == 0) return;
    const ColTile t = col_tile(s, e.y);
    if (!t.valid) return;
    const int sxy = s.L[0].nx * s.L[0].ny;
    const int it = st_uni(s, ST_ITERS);
    const void* dk = ((it - 1) & 1) ? s.d1 : s.d0;
    const void* dk2 = (it & 1) ? s.d1 : s.d0;
    constexpr bool scaled = !std::is_same<ZT, float>::value;
    const float ids = scaled ? (float)(1.0 / s.scal[S_DS]) : 1.0f;
    const float ids2 = scaled && owed == 2 ? (float)(1.0 / s.scal[S_DS2]) : 1.0f;
    const float alpha = (float)s.scal[S_ALPHA], alpha2 = (float)s.scal[S_ALPHA2];
    for (int k = t.k0; k < t.k1; ++k) {
        const int c = t.c2 + sxy * k;
        float x = s.x[c];
        if (owed == 2) x = fmaf(alpha2, ld<ZT>(dk2, c) * ids2, x);
        s.x[c] = fmaf(alpha, ld<ZT>(dk, c) * ids, x);
    }
}

// ---- per-system finalisers: fixed-order sums of the block partials --------
__device__ __forceinline__ void sum_parts(const DSys& s, int slot, double* out) {
    double v[1] = {0.0};
    for (int t = threadIdx.x; t < s.tiles0; t += TB) v[0] += s.part[(long)t * NPART + slot];
    block_sums<1>(v, out);
}
__device__ __forceinline__ float max_parts(const DSys& s) {
    float m = 0.0f;
    for (int t = threadIdx.x; t < s.tiles0; t += TB) m = fmaxf(m, (float)s.part[(long)t * NPART + 2]);
    return block_max(m);
}

// power-of-two scale that brings a residual of squared norm rr over n cells to rms ~1
__device__ __forceinline__ double vcycle_scale(double rr, int n) {
    if (!(rr > 0)) return 1.0;
    double e = rint(0.5 * log2((double)n / rr));
    e = fmin(fmax(e, -60.0), 60.0);
    return exp2(e);
}

__device__ __forceinline__ bool stop_reached(const DSys& s, double rr, float maxr) {
    // the reference's test, in FP32: res = max|r| / inv against pressureTol (gasphase.cpp:601-604)
    if (s.state[ST_RULE] == 1) return maxr / (float)s.scal[S_INV] <= (float)s.scal[S_TOL];
    const double bb = s.scal[S_BB];
    const double rel = bb > 0 ? sqrt(rr) / sqrt(bb) : 0.0;
    return rel <= s.scal[S_TOL];
}

// after k_resid_init: |b|^2, |r|^2, max|r|; stop at iteration 0 if already converged
__global__ void __launch_bounds__(TB) k_fin_init(const DSys* __restrict__ sys, const int2* __restrict__ tab) {
    PHG_SYS_PROLOGUE
    __shared__ double res[3];
    sum_parts(s, 0, &res[0]);
    __syncthreads();
    sum_parts(s, 1, &res[1]);
    __syncthreads();
    const bool restart = s.state[ST_RESTART] != 0;
    if (restart) sum_parts(s, 3, &res[2]);
    const float m = max_parts(s);
    __syncthreads();
    if (threadIdx.x == 0) {
        // A restart only pays while the true residual stands clear of the noise
        // floor of an FP32 solution (eps_32 times the size of the terms of
        // b - A x) and has fallen well below the residual the last round
        // started from: otherwise it is rounding noise, and iterating on it can
        // only damage x.  Then a system whose last round converged counts as
        // converged; one whose last round broke down does so only if its
        // residual is at that noise floor, and keeps its breakdown status
        // otherwise.
        bool floor_hit = false, at_noise = false;
        const double rn = sqrt(res[1]);
        if (restart) {
            s.state[ST_RESTART] = 0;
            // the term sizes add linearly here, the rounding errors about in
            // quadrature: a quarter of eps times their norm marks the noise
            at_noise = s.state[ST_RULE] == 0 && rn <= 0.25 * 5.9604644775390625e-8 * sqrt(res[2]);
            floor_hit = s.state[ST_RULE] == 0 && (at_noise || rn > 0.25 * s.scal[S_RPREV]);
        }
        if (floor_hit) {
            // no restart: the solve ends with the previous round's figures
            const int last = s.state[ST_LAST];
            s.state[ST_ITERS] = 0;
            s.state[ST_XPEND] = 0;
            s.state[ST_STATUS] = last == 1 || at_noise ? 1 : last;
            s.state[ST_ACTIVE] = 0;
            return;
        }
        s.scal[S_RPREV] = rn;
        s.scal[S_BB] = res[0];
        s.scal[S_RR] = res[1];
        s.scal[S_MAXR] = m;
        s.scal[S_ZS] = s.scal[S_ZSN] = vcycle_scale(res[1], s.L[0].n);
        s.scal[S_PNEXT] = 1e-6;
        s.state[ST_PROJ] = 0;
        s.state[ST_ITERS] = 0;
        s.state[ST_XPEND] = 0;
        // a right-hand side or warm start that is not finite leaves nothing to
        // iterate on: the system stops at once with its own status
        if (!isfinite(res[0]) || !isfinite(res[1])) {
            s.state[ST_STATUS] = 5;
            s.state[ST_ACTIVE] = 0;
            return;
        }
        const bool done = floor_hit || stop_reached(s, res[1], m);
        if (done || s.state[ST_MAXIT] <= 0) {
            s.state[ST_STATUS] = done ? 1 : 2;
            s.state[ST_ACTIVE] = 0;
        }
    }
}

// after the first V-cycle: rz; stop if rz <= 0 (the reference's loop test)
__global__ void __launch_bounds__(TB) k_fin_rz0(const DSys* __restrict__ sys, const int2* __restrict__ tab) {
    PHG_SYS_PROLOGUE
    __shared__ double res[1];
    sum_parts(s, 3, &res[0]);
    __syncthreads();
    if (threadIdx.x == 0) {
        s.scal[S_RZ] = res[0] / s.scal[S_ZS];
        if (!(res[0] > 0)) {
            s.state[ST_STATUS] = 4;
            s.state[ST_ACTIVE] = 0;
        }
    }
}

// after k_dir_apply: alpha = rz / dq, or stop on dq <= 0 (gasphase.cpp:606-608)
__global__ void __launch_bounds__(TB) k_fin_alpha(const DSys* __restrict__ sys, const int2* __restrict__ tab) {
    PHG_SYS_PROLOGUE
    __shared__ double res[1];
    sum_parts(s, 0, &res[0]);
    __syncthreads();
    if (threadIdx.x == 0) {
        const double dq = res[0];
        s.scal[S_DQ] = dq;
        if (!(dq > 0)) {
            s.state[ST_STATUS] = 3;
            s.state[ST_ACTIVE] = 0;
            // on odd iterations k_dir_apply left alpha_{k-1} d_{k-1} owed
            s.state[ST_XPEND] = s.state[ST_ITERS] & 1;
        } else {
            s.scal[S_ALPHA2] = s.scal[S_ALPHA];
            s.scal[S_DS2] = s.scal[S_DS];
            s.scal[S_DS] = s.scal[S_ZS];   // the scale k_dir_apply stored d with
            s.scal[S_ALPHA] = s.scal[S_RZ] / dq;
            s.scal[S_ZS] = s.scal[S_ZSN];   // z of the previous V-cycle has been consumed
            // k_proj: sums taken -> subtract this iteration; subtracted -> done
            const int pj = s.state[ST_PROJ];
            if (pj) s.state[ST_PROJ] = pj == 1 ? 2 : 0;
            s.state[ST_VC] = 1;             // q is defined from here on
        }
    }
}

// after k_post_red: beta, iteration count and the stop test (gasphase.cpp:617-625).
// beta = rz_new / rz (the reference's); with a mixed-precision V-cycle, whose
// rounding makes the preconditioner vary slightly between applications,
// the flexible (Polak-Ribiere) form beta = z_new.(r_new - r_old) / rz
// = -alpha q.z_new / rz, which equals it for a fixed preconditioner
__global__ void __launch_bounds__(TB) k_fin_beta(const DSys* __restrict__ sys, const int2* __restrict__ tab) {
    PHG_SYS_PROLOGUE
    __shared__ double res[3];
    sum_parts(s, 3, &res[0]);
    __syncthreads();
    sum_parts(s, 1, &res[1]);
    __syncthreads();
    sum_parts(s, 0, &res[2]);
    const float m = max_parts(s);
    __syncthreads();
    if (threadIdx.x == 0) {
        const double rz_new = res[0] / s.scal[S_ZS];
        s.scal[S_ZSN] = vcycle_scale(res[1], s.L[0].n);
        if (s.state[ST_FLEX]) s.scal[S_BETA] = -s.scal[S_ALPHA] * (res[2] / s.scal[S_ZS]) / s.scal[S_RZ];
        else s.scal[S_BETA] = rz_new / s.scal[S_RZ];
        s.scal[S_RZ] = rz_new;
        s.scal[S_RR] = res[1];
        s.scal[S_MAXR] = m;
        const int it = s.state[ST_ITERS] + 1;
        s.state[ST_ITERS] = it;
        int status = 0;
        if (!isfinite(res[1])) status = 5;   // the reference's `res > tol` ends here too, unflagged
        else if (stop_reached(s, res[1], m)) status = 1;
        else if (it >= s.state[ST_MAXIT]) status = 2;
        else if (!(rz_new > 0)) status = 4;
        // each further fall of ||r|| / ||b|| by 1e3 (from 1e-3 on) has k_proj_r
        // clear the residual's closed-zone means once
        if (!status && s.nclosed > 0 && s.state[ST_PROJ] == 0 && res[1] <= s.scal[S_PNEXT] * s.scal[S_BB]) {
            s.state[ST_PROJ] = 1;
            s.scal[S_PNEXT] *= 1e-6;
        }
        if (status) {
            s.state[ST_STATUS] = status;
            s.state[ST_ACTIVE] = 0;
            // this iteration's x += alpha d is owed, and on odd iterations the
            // one before it (k_dir_apply applies them in pairs)
            s.state[ST_XPEND] = (it & 1) ? 1 : 2;
        }
    }
}

// between rounds of a solve: the owed x += alpha d has been applied (x_flush);
// a system that converged or broke down (d.Ad <= 0 or r.z <= 0, which the
// rounding of an FP16/BF16 V-cycle can bring about once the recursive
// residual nears the accuracy an FP32 x can hold) is started again from its
// x, with its true residual b - A x (k_resid_init, then k_fin_init stops it
// at once if that already meets the tolerance); one that hit the cap stays done
__global__ void k_refine_reset(const DSys* __restrict__ sys, int nsys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsys) return;
    int* st = sys[i].state;
    st[ST_XPEND] = 0;
    const int last = st[ST_STATUS];
    if (last != 1 && last != 3 && last != 4) return;
    st[ST_LAST] = last;
    st[ST_ITOT] += st[ST_ITERS];
    st[ST_ITERS] = 0;
    st[ST_STATUS] = 0;
    st[ST_ZEROX] = 0;
    st[ST_VC] = 0;
    st[ST_RESTART] = 1;
    st[ST_ACTIVE] = 1;
}

__global__ void k_count_active(const DSys* __restrict__ sys, int nsys, int* out) {
    int n = 0;
    for (int i = threadIdx.x; i < nsys; i += blockDim.x) n += sys[i].state[ST_ACTIVE] != 0;
    __shared__ int sh[TB];
    sh[threadIdx.x] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int i = 0; i < blockDim.x; ++i) t += sh[i];
        *out = t;
    }
}

// ---------------------------------------------------------------------------
// V-cycle, level 0 (in place on the outer vectors: b = r, x = z)
// ---------------------------------------------------------------------------
// r -= alpha q (gasphase.cpp:613; skipped on the first V-cycle) with the
// partials |r|^2 (slot 1) and max|r| (slot 2), then the first pre-smoothing
// half sweep from z = 0: red cells z = r / diag, black cells z = 0
// (gsColor(0) on a zero iterate, pressure_mg.cpp:276,282)
template <class CT, class ZT>
__global__ void __launch_bounds__(TB) k_pre_red(const DSys* __restrict__ sys, const int2* __restrict__ tab, int first) {
    PHG_SYS_PROLOGUE
    const DLevel& L = s.L[0];
    const ColTile t = col_tile(s, e.y);
    const int sxy = L.nx * L.ny;
    const float alpha = (float)s.scal[S_ALPHA];
    const float zs = (float)s.scal[S_ZS];
    const int it = st_uni(s, ST_ITERS);
    const float* __restrict__ rin = rbuf(s, it);   // r_it (r_0 on the first V-cycle)
    float* __restrict__ rout = rbuf(s, it + 1);
    double v[1] = {0.0};
    float m = 0.0f;
    if (t.valid) {
        for (int k = t.k0; k < t.k1; ++k) {
            const int c = t.c2 + sxy * k;
            float r = gld(rin, c);
            if (!first) {
                r = fmaf(-alpha, gld(s.q, c), r);
                gst_nt(rout, c, r);
                v[0] += (double)r * r;
                m = fmaxf(m, fabsf(r));
            }
            float z = 0.0f;
            if (((t.i + t.j + k) & 1) == 0) {
                if constexpr (std::is_same<CT, Mask01>::value) {
                    const int dg = __popc((unsigned)gld(L.mask, c));
                    if (dg > 0) z = div_count(r * zs, (float)dg);
                } else {
                    z = (r * zs) * gld(s.invd, c);
                }
            }
            st_nt<ZT>(s.z, c, z);
        }
    }
    if (first) return;
    double* out = s.part + (long)e.y * NPART;
    __shared__ double sh[TB / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    v[0] = wave_sum(v[0]);
    if (lane == 0) sh[w] = v[0];
    __syncthreads();
    m = block_max(m);
    if (threadIdx.x == 0) {
        double acc = 0.0;
#pragma unroll
        for (int ww = 0; ww < TB / 64; ++ww) acc += sh[ww];
        out[1] = acc;
        out[2] = m;
    }
}

// black half sweep z = (r + sum a z_nb) / diag on cells with i+j+k odd (gsColor(1))
template <class CT, class ZT>
__global__ void __launch_bounds__(TB) k_rb_black(const DSys* __restrict__ sys, const int2* __restrict__ tab, int sweep) {
    PHG_SYS_PROLOGUE
    if (sweep >= s.nu0) return;   // sweep of a system that smooths fewer times
    const DLevel& L = s.L[0];
    const ColTile t = col_tile(s, e.y);
    const int sxy = L.nx * L.ny;
    const void* z = s.z;
    const float zs = (float)s.scal[S_ZS];
    extern __shared__ float lds[];
    auto raw = [&](unsigned n) { return ld<ZT>(z, n); };
    march(s, t, lds, raw, [](auto&) {}, [](float v) { return v; }, cr_pre<CT>(s, L), [&](int k, float vp, float vc, float vn, const float* p,
                                                                  const CoefR<CT>& o) {
        if (((t.i + t.j + k) & 1) == 0) return;
        const Cell q{t.c2 + sxy * k, t.i, t.j, k};
        const float dg = cell_diag(o.cf);
        float zc = 0.0f;
        if (dg > 0) zc = cell_div<CT>(o.r * zs + cell_lat(o.cf, q, L.nx, L.ny, L.nz, p, vp, vn), dg);
        st_nt<ZT>(s.z, q.c, zc);
    });
}

// level-0 residual r0 = r - A z into the V-cycle scratch (residualGrid, pressure_mg.cpp:217-235)
template <class CT, class ZT, class VT>
__global__ void __launch_bounds__(TB) k_resid0(const DSys* __restrict__ sys, const int2* __restrict__ tab) {
    PHG_SYS_PROLOGUE
    const DLevel& L = s.L[0];
    const ColTile t = col_tile(s, e.y);
    const int sxy = L.nx * L.ny;
    const void* z = s.z;
    const float zs = (float)s.scal[S_ZS];
    extern __shared__ float lds[];
    auto raw = [&](unsigned n) { return ld<ZT>(z, n); };
    march(s, t, lds, raw, [](auto&) {}, [](float v) { return v; }, cr_pre<CT>(s, L), [&](int k, float vp, float vc, float vn, const float* p,
                                                                  const CoefR<CT>& o) {
        const Cell q{t.c2 + sxy * k, t.i, t.j, k};
        const float dg = cell_diag(o.cf);
        float r = 0.0f;
        if (dg > 0) r = o.r * zs - dg * vc + cell_lat(o.cf, q, L.nx, L.ny, L.nz, p, vp, vn);
        st_nt<VT>(L.r, q.c, r);
    });
}

// prolongation fused into the black post-smoothing half sweep: black cells
// z = (r + sum a (z_nb + t1[agg_nb])) / diag, where the red neighbours' values
// are formed as prolonged on the fly (pressure_mg.cpp:293-296).  The black
// cells' own prolonged values are never read: the sweep replaces them.
template <class CT, class ZT, class VT>
__global__ void __launch_bounds__(TB) k_post_black(const DSys* __restrict__ sys, const int2* __restrict__ tab) {
    PHG_SYS_PROLOGUE
    const DLevel& L = s.L[0];
    const ColTile t = col_tile(s, e.y);
    const int sxy = L.nx * L.ny;
    const void* z = s.z;
    const void* t1 = s.L[1].t;
    const int* __restrict__ agg = L.agg;
    // prolonged value z + t1[agg]: the aggregate index is loaded with z, the
    // level-1 value (a small, cache-resident vector) when the value is formed
    struct ZG {
        float z;
        int g;
        float c;
    };
    auto raw = [&](unsigned n) { return ZG{ld<ZT>(z, n), gld(agg, n), 0.0f}; };
    auto mid = [&](ZG& v) { v.c = ld<VT>(t1, max(v.g, 0)); };
    const float pw = (float)s.scal[S_PW];
    auto fin = [&](const ZG& v) { return v.z + (v.g >= 0 ? pw * v.c : 0.0f); };
    const float zs = (float)s.scal[S_ZS];
    extern __shared__ float lds[];
    march(s, t, lds, raw, mid, fin, cr_pre<CT>(s, L), [&](int k, float vp, float vc, float vn, const float* p,
                                                     const CoefR<CT>& o) {
        if (((t.i + t.j + k) & 1) == 0) return;
        const Cell q{t.c2 + sxy * k, t.i, t.j, k};
        const float dg = cell_diag(o.cf);
        float zc = 0.0f;
        if (dg > 0) zc = cell_div<CT>(o.r * zs + cell_lat(o.cf, q, L.nx, L.ny, L.nz, p, vp, vn), dg);
        st_nt<ZT>(s.z, q.c, zc);
    });
}

// red post-smoothing half sweep (gsColor(0)) fused with the partial r.z
// (slot 3) and, for flexible CG, q.z (slot 0)
template <class CT, class ZT>
__global__ void __launch_bounds__(TB) k_post_red(const DSys* __restrict__ sys, const int2* __restrict__ tab, int sweep) {
    PHG_SYS_PROLOGUE
    if (sweep >= s.nu0) return;   // sweep of a system that smooths fewer times
    const DLevel& L = s.L[0];
    const ColTile t = col_tile(s, e.y);
    const int sxy = L.nx * L.ny;
    const void* z = s.z;
    const float zs = (float)s.scal[S_ZS];
    const float pw = (float)s.scal[S_PW];   // level scale x coarse-correction weight
    const bool with_q = st_uni(s, ST_FLEX) != 0 && st_uni(s, ST_VC) != 0;   // q exists once an iteration has run
    const float* rv = r_vc(s);
    double v[1] = {0.0}, vq = 0.0;
    extern __shared__ float lds[];
    struct Op {
        CoefRegs<CT> cf;
        float r, q;
    };
    auto pre = [&](const Cell& q) {
        Op o;
        o.cf.load(L, L.axc, L.ayc, L.azc, q);
        o.r = gld(rv, q.c);
        o.q = with_q ? gld(s.q, q.c) : 0.0f;
        return o;
    };
    auto raw = [&](unsigned n) { return ld<ZT>(z, n); };
    march(s, t, lds, raw, [](auto&) {}, [](float v) { return v; }, pre, [&](int k, float vp, float vc, float vn, const float* p, const Op& o) {
        const Cell q{t.c2 + sxy * k, t.i, t.j, k};
        const float r = o.r;
        float zc = vc;   // black cells: final from the black sweep
        if (((t.i + t.j + k) & 1) == 0) {
            const float dg = cell_diag(o.cf);
            zc = 0.0f;
            if (dg > 0) zc = zq<ZT>(cell_div<CT>(r * zs + cell_lat(o.cf, q, L.nx, L.ny, L.nz, p, vp, vn), dg));
            st_nt<ZT>(s.z, q.c, zc);   // the dots below use the value as stored
        }
        v[0] += (double)r * zc;
        if (with_q) vq += (double)o.q * zc;
    });
    double* out = s.part + (long)e.y * NPART;
    __shared__ double sh[2][TB / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    v[0] = wave_sum(v[0]);
    if (with_q) vq = wave_sum(vq);
    if (lane == 0) {
        sh[0][w] = v[0];
        sh[1][w] = vq;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double acc = 0.0, accq = 0.0;
#pragma unroll
        for (int ww = 0; ww < TB / 64; ++ww) {
            acc += sh[0][ww];
            accq += sh[1][ww];
        }
        out[3] = acc;
        if (with_q) out[0] = accq;
    }
}

// ---------------------------------------------------------------------------
// fused pre-smoothing: k_pre_red + k_rb_black + k_resid0 in one pass
// ---------------------------------------------------------------------------
// For grids with 0/1 coefficients.  The black sweep at a cell needs the red
// values of its neighbours, and the residual at a cell the smoothed values of
// its neighbours, so the residual of plane k needs red values two cells away.
// A block owns TB columns and marches its k-chunk with a lead of two planes:
//   stage plane k+2   the new residual rn = r - alpha q and, over the block's
//                     columns +- 2 grid rows (W2 = TB + 4 nx), the red value
//                     rn zs / diag (black cells: rn zs) and the face mask
//   smooth plane k+1  over the columns +- 1 row (W1 = TB + 2 nx): red cells
//                     keep their value, black cells take the Gauss-Seidel update
//   emit plane k      z, the updated r and the residual r0 = rn zs - A z of
//                     the block's own columns, with |r|^2 and max|r|.
// Every value is formed by the same operations as in the three separate
// kernels, so the results are bitwise the same; the pass reads r and q and
// writes r, z and r0 once (18 B/cell in place of 36).  z goes to the system's
// z2, from where the fused post-smoothing reads it.
// Staging and smoothing work on pairs of adjacent columns (2u and 2u+1 of
// the W2 range, which starts at an even column).  The two cells of a pair
// have opposite colours (along a row i changes by one; across the end of a
// row, which only a pair of an odd nx straddles, i + j changes by nx - 2),
// so a lane does one red and one black cell and no lane idles through the
// other colour's work: one division per pair when staging, one stencil per
// pair when smoothing.
// LDS rings of three planes: staged values (after a two-float pad), smoothed
// values, the own columns' new residual, face masks.  Three slots suffice: a
// plane's staged values are last read by the smoothing of the previous step
// and its smoothed values by the previous emit, both behind a barrier of the
// step that overwrites them; the own-column values are read before the
// step's second barrier for the same reason.
constexpr int FUSE_MAX_NX = 192;   // staging pairs NT/2 + 2 nx <= 2 NT

// A fused block has NT = TB, 2 TB or 4 TB threads and owns NT columns: the
// column tile of its table entry and the NT / TB - 1 after it (the table then
// lists every second or fourth column tile only).  A wider block stages fewer
// halo rows per own column (TB + 4 nx against 2 TB + 4 nx for twice the
// columns); 4 TB takes one staging and one smoothing pair per thread up to
// nx = FUSE_MAX_NX, where 2 TB needs two staging pairs past nx = 128.
// The passes keep rings of three planes in LDS.  Their strides are those of
// the configuration's widest grid, so every access is a per-thread base plus a
// constant, and the ring slot of each step is known at compile time (the march
// is unrolled by three).
template <int V>
using ic = std::integral_constant<int, V>;
__host__ __device__ constexpr int fuse_nxmax(int nt, int p2, int p1) {   // widest nx of a configuration
    return nt == 4 * TB   ? FUSE_MAX_NX
           : nt == 2 * TB ? (p2 == 1 ? 128 : FUSE_MAX_NX)
                          : (p2 == 1 ? 64 : (p1 == 1 ? 127 : FUSE_MAX_NX));
}
__host__ __device__ constexpr int fuse_sr(int nt, int nxm) { return nt + 4 * nxm; }       // staged values, masks
__host__ __device__ constexpr int fuse_sz(int nt, int nxm) { return nt + 2 * nxm + 2; }   // smoothed values
__host__ __device__ constexpr size_t fuse_lds(int nt, int p2, int p1, bool post) {
    const int m = fuse_nxmax(nt, p2, p1), sr = fuse_sr(nt, m), sz = fuse_sz(nt, m);
    return (size_t)(2 + 3 * sr) * 4 + (size_t)3 * sz * 4 + (size_t)3 * nt * 4 + (size_t)3 * sr +
           (post ? (size_t)3 * nt * 4 : 0);
}

// P2, P1: staging and smoothing pairs per thread, for NT/2 + 2 nx and at
// most NT/2 + nx + 1 pairs
template <class VT, class ZT, bool FIRST, int NT, int P2, int P1>
__global__ void __launch_bounds__(NT) k_pre_fused(const DSys* __restrict__ sys, const int2* __restrict__ tab) {
    PHG_SYS_PROLOGUE
    const DLevel& L = s.L[0];
    const ColTile t = col_tile(s, e.y);
    const int nx = L.nx, ny = L.ny, nz = L.nz, sxy = nx * ny;
    const int W1 = NT + 2 * nx, W2 = NT + 4 * nx;
    const int tid = (int)threadIdx.x;
    const int c0 = t.c2 - tid;     // a multiple of TB, so even (NT = 2 TB: of 2 TB)
    const int zo = nx & ~1;        // Zb holds W2 index h at h - zo, so pairs stay 8-byte aligned
    const float alpha = (float)s.scal[S_ALPHA];
    const float zs = (float)s.scal[S_ZS];
    const int it = st_uni(s, ST_ITERS);
    const float* __restrict__ rin = rbuf(s, it);   // r_it: read here, also in the halo rows of other blocks
    float* __restrict__ rout = rbuf(s, it + 1);    // r_{it+1}: written for the own columns only
    constexpr int NXM = fuse_nxmax(NT, P2, P1), SR = fuse_sr(NT, NXM), SZ = fuse_sz(NT, NXM);
    extern __shared__ float lds[];
    float* Rb = lds + 2;                 // [3][SR]; Rb[-1] is read (and ignored) by an edge pair of an odd nx
    float* Zb = Rb + 3 * SR;             // [3][SZ]
    float* Ob = Zb + 3 * SZ;             // [3][NT]: rn of the own columns
    uint8_t* Mb = reinterpret_cast<uint8_t*>(Ob + 3 * NT);   // [3][SR]
    auto ij_bits = [&](int c) {                              // the (i, j) part of edge_bits
        const int j = c / nx, i = c - j * nx;
        return (i > 0 ? 1u : 0u) | (i < nx - 1 ? 2u : 0u) | (j > 0 ? 4u : 0u) | (j < ny - 1 ? 8u : 0u);
    };
    __shared__ FuseLut lut;
    fuse_lut_init(lut, tid, NT);
    __syncthreads();

    // staging pairs (W2 indices 2u, 2u+1): clamped column of the first
    // cell, in-grid flags, colour (parity of i + j) of the first cell.  The
    // range starts at an even column, so a pair is either wholly left of
    // column 0 or starts inside the grid.
    const int NP2 = W2 / 2;
    int ca[P2], par2[P2];
    bool ina[P2], inb[P2];
#pragma unroll
    for (int m = 0; m < P2; ++m) {
        const int u = tid + m * NT;
        const int col = c0 - 2 * nx + 2 * u;
        ina[m] = u < NP2 && col >= 0 && col < sxy;
        inb[m] = u < NP2 && col + 1 >= 0 && col + 1 < sxy;
        ca[m] = min(max(col, 0), sxy - 1);
        par2[m] = (ca[m] % nx + ca[m] / nx) & 1;
    }
    // smoothing pairs: W2 indices ha, ha + 1 over the W1 range [nx, nx + W1)
    // (for an odd nx the first and last pair stick out by one cell)
    const int u1 = nx >> 1, NP1 = ((NT + 3 * nx - 1) >> 1) - u1 + 1;
    int ha1[P1], par1[P1];
    unsigned ija[P1], ijb[P1];
    bool ina1[P1], inb1[P1];
#pragma unroll
    for (int m = 0; m < P1; ++m) {
        const bool use = tid + m * NT < NP1;
        const int h = 2 * (u1 + tid + m * NT), col = c0 - 2 * nx + h;
        ha1[m] = h;
        ina1[m] = use && h >= nx && h < nx + W1 && col >= 0 && col < sxy;
        inb1[m] = use && h + 1 >= nx && h + 1 < nx + W1 && col + 1 >= 0 && col + 1 < sxy;
        const int cca = min(max(col, 0), sxy - 1), ccb = min(max(col + 1, 0), sxy - 1);
        ija[m] = ij_bits(cca);
        ijb[m] = ij_bits(ccb);
        par1[m] = (cca % nx + cca / nx) & 1;
    }
    // Every vector is 8-byte aligned and the plane size is even (the host
    // runs this kernel only then), so a pair, which starts at an even index
    // inside the grid, is one 8-byte load and its two mask bytes one 2-byte
    // load: each memory instruction covers whole consecutive lines.  The
    // loaded registers are not touched until the plane is staged three steps
    // later, so the loads stay in flight.  Pairs outside the grid, and the
    // unused slots of the last lanes, load a clamped pair that is not used.
    struct Raw {
        f32x2 r[P2], q[P2];
        uint16_t m[P2];   // mask bytes of the pair: first cell in bits 0-7
    };
    auto load = [&](int p, Raw& X) {
        const unsigned base = (unsigned)sxy * (unsigned)min(max(p, 0), nz - 1);
#pragma unroll
        for (int m = 0; m < P2; ++m) {
            const unsigned i2 = (ca[m] + base) >> 1;
            X.r[m] = gld(reinterpret_cast<const f32x2*>(rin), i2);
            if (!FIRST) X.q[m] = gld(reinterpret_cast<const f32x2*>(s.q), i2);
            X.m[m] = gld(reinterpret_cast<const uint16_t*>(L.mask), i2);
        }
    };
    // stage plane p from its raw loads: red cells rn zs / diag (0 without
    // open faces), black cells rn zs, out-of-grid cells 0
    auto stage = [&](int p, const Raw& X, auto S) {   // S: the ring slot of plane p
        constexpr int sp = decltype(S)::value;
        const bool pin = p >= 0 && p < nz;
        float* R = Rb + sp * SR;
        uint8_t* M = Mb + sp * SR;
        float* O = Ob + sp * NT;
#pragma unroll
        for (int m = 0; m < P2; ++m) {
            const int u = tid + m * NT;
            if (u >= NP2) break;
            float rna = X.r[m].x, rnb = X.r[m].y;
            if (!FIRST) {
                rna = fmaf(-alpha, X.q[m].x, rna);
                rnb = fmaf(-alpha, X.q[m].y, rnb);
            }
            const unsigned ma = X.m[m] & 255u, mb = (unsigned)X.m[m] >> 8;
            const bool oka = pin && ina[m], okb = pin && inb[m];
            const bool ared = ((par2[m] + p) & 1) == 0;   // the first cell is red
            const float rsa = rna * zs, rsb = rnb * zs;
            const float2 D = lut.d[ared ? ma : mb];   // face count, its reciprocal
            // the red value from zero, as the separate red pass stores it in z
            const float vr = D.x > 0 ? zq<ZT>(div_rn(ared ? rsa : rsb, D.x, D.y)) : 0.0f;
            *reinterpret_cast<float2*>(R + 2 * u) =
                float2{oka ? (ared ? vr : rsa) : 0.0f, okb ? (ared ? rsb : vr) : 0.0f};
            *reinterpret_cast<uint16_t*>(M + 2 * u) = (uint16_t)((oka ? ma : 0u) | ((okb ? mb : 0u) << 8));
            const int o = 2 * u - 2 * nx;
            if (o >= 0 && o < NT) *reinterpret_cast<float2*>(O + o) = float2{oka ? rna : 0.0f, okb ? rnb : 0.0f};
        }
    };
    // smoothed values of plane p over W1: red cells keep their value, black
    // cells take the Gauss-Seidel update from their red neighbours
    auto smooth = [&](int p, auto S) {   // S: the ring slot of plane p
        constexpr int sp = decltype(S)::value;
        const bool pin = p >= 0 && p < nz;
        const unsigned kb = (p > 0 ? 16u : 0u) | (p < nz - 1 ? 32u : 0u);
        const float* R = Rb + sp * SR;
        const float* Rm = Rb + ((sp + 2) % 3) * SR;
        const float* Rp = Rb + ((sp + 1) % 3) * SR;
        const uint8_t* M = Mb + sp * SR;
        float* Z = Zb + sp * SZ - zo;
#pragma unroll
        for (int m = 0; m < P1; ++m) {
            if (tid + m * NT >= NP1) break;
            const bool ared = ((par1[m] + p) & 1) == 0;
            const int hb = ha1[m] + (ared ? 1 : 0);    // the black cell of the pair
            const float xm = R[hb - 1], xp = R[hb + 1];
            const unsigned mk = M[hb];
            const float2 D = lut.d[mk];
            float zb = 0.0f;
            if (pin && (ared ? inb1[m] : ina1[m]) && D.x > 0)
                zb = zq<ZT>(div_rn(R[hb] + lat_lut(lut, mk & ((ared ? ijb[m] : ija[m]) | kb), R + hb, nx, Rm[hb],
                                                   Rp[hb]),
                                   D.x, D.y));
            const float zr = pin && (ared ? ina1[m] : inb1[m]) ? (ared ? xm : xp) : 0.0f;
            *reinterpret_cast<float2*>(Z + ha1[m]) = ared ? float2{zr, zb} : float2{zb, zr};
        }
    };
    double acc = 0.0;
    float mx = 0.0f;
    // emit plane k of the own column: its rn and mask are read before the
    // step's second barrier (the next step's staging reuses their slot)
    const int hz = 2 * nx + tid - zo;   // the own column in Zb
    struct Own {
        float rn;
        unsigned m;
    };
    auto own = [&](auto S) {
        constexpr int sp = decltype(S)::value;
        return Own{Ob[sp * NT + tid], (unsigned)Mb[sp * SR + 2 * nx + tid]};
    };
    auto emit = [&](int k, auto S, const Own& o) {
        constexpr int sp = decltype(S)::value;
        if (!t.valid) return;
        const int c = t.c2 + sxy * k;
        const float rn = o.rn;
        const float* Z = Zb + sp * SZ + hz;
        const float zc = *Z;
        const float dg = lut.d[o.m].x;
        float r0 = 0.0f;
        if (dg > 0)
            r0 = rn * zs - dg * zc +
                 lat_lut(lut, o.m & edge_bits(t.i, t.j, k, nx, ny, nz), Z, nx, Zb[((sp + 2) % 3) * SZ + hz],
                         Zb[((sp + 1) % 3) * SZ + hz]);
        st_nt<ZT>(s.z2, c, zc);   // read by the fused post-smoothing (exact: zc is a ZT value)
        st_nt<VT>(L.r, c, r0);
        if (!FIRST) {
            gst_nt(rout, c, rn);
            acc += (double)rn * rn;
            mx = fmaxf(mx, fabsf(rn));
        }
    };
    // one plane of the march: stage k+2, smooth k+1, emit k (for k >= k0).
    // X holds the raw loads of plane k+2 and is refilled with plane k+5.
    const int ks = t.k0 - 2;
    // J: the ring slot of plane k, (k - ks) mod 3
    auto step = [&](int k, Raw& X, auto J) {
        constexpr int j = decltype(J)::value;
        stage(k + 2, X, ic<(j + 2) % 3>());
        load(min(k + 5, t.k1 + 1), X);   // past the chunk: repeats the last plane needed
        __syncthreads();
        smooth(k + 1, ic<(j + 1) % 3>());
        const Own o = own(J);
        __syncthreads();
        if (k >= t.k0) emit(k, J, o);
    };
    Raw X0, X1, X2;
    {
        Raw Y;
        load(ks, Y);
        stage(ks, Y, ic<0>());
        load(ks + 1, Y);
        stage(ks + 1, Y, ic<1>());
    }
    load(ks + 2, X0);
    load(ks + 3, X1);
    load(ks + 4, X2);
    // whole groups of three steps in the loop, the rest after it: no step
    // inside the loop is conditional, so the three raw buffers keep their
    // registers around the loop and their loads are never waited on early
    int k = ks;
    for (; k + 2 < t.k1; k += 3) {
        step(k, X0, ic<0>());
        step(k + 1, X1, ic<1>());
        step(k + 2, X2, ic<2>());
    }
    if (k < t.k1) step(k, X0, ic<0>());
    if (k + 1 < t.k1) step(k + 1, X1, ic<1>());
    if (FIRST) return;
    __shared__ double sh[NT / 64];
    __shared__ float shm[NT / 64];
    const int lane = tid & 63, w = tid >> 6;
    const double v = wave_sum(acc);
    const float wm = wave_max(mx);
    if (lane == 0) {
        sh[w] = v;
        shm[w] = wm;
    }
    __syncthreads();
    tile_rows(s, e, c0, [&](int w0, double* out) {
        double tot = 0.0;
        float m = 0.0f;
#pragma unroll
        for (int ww = 0; ww < TB / 64; ++ww) {
            tot += sh[w0 + ww];
            m = fmaxf(m, shm[w0 + ww]);
        }
        out[1] = tot;
        out[2] = m;
    });
}

// ---------------------------------------------------------------------------
// fused post-smoothing: k_post_black + k_post_red in one pass
// ---------------------------------------------------------------------------
// For the grids of k_pre_fused.  The black update of a cell needs the
// prolonged values z + t1[agg] of its red neighbours, and the red update of a
// cell the new values of its black neighbours, so the red update of plane k
// needs prolonged values two cells away.  The march mirrors k_pre_fused:
//   stage plane k+2   over the columns +- 2 rows, on column pairs: red cells
//                     hold their prolonged value, black cells r zs; the
//                     pair's one level-1 gather was issued a step earlier
//   smooth plane k+1  over the columns +- 1 row: black cells take the
//                     Gauss-Seidel update (the red cells' slots are unused)
//   emit plane k      red cells of the own columns take their update from the
//                     new black values; z of every own cell is written once,
//                     with the partial sums r.z and (flexible CG) q.z.
// Every value is formed by the same operations as in the two separate
// kernels, and each thread sums its column's terms in the same order, so the
// results are bitwise the same; the pass reads z and r once and writes z once
// (12 B/cell, 16 with q, in place of about 27).

template <class VT, class ZT, int NT, int P2, int P1>
__global__ void __launch_bounds__(NT) k_post_fused(const DSys* __restrict__ sys, const int2* __restrict__ tab) {
    using ZP = typename pair_of<ZT>::type;
    PHG_SYS_PROLOGUE
    const DLevel& L = s.L[0];
    const ColTile t = col_tile(s, e.y);
    const int nx = L.nx, ny = L.ny, nz = L.nz, sxy = nx * ny;
    const int W1 = NT + 2 * nx, W2 = NT + 4 * nx;
    const int tid = (int)threadIdx.x;
    const int c0 = t.c2 - tid;     // a multiple of TB, so even (NT = 2 TB: of 2 TB)
    const int zo = nx & ~1;        // Zb holds W2 index h at h - zo, so pairs stay 8-byte aligned
    const float zs = (float)s.scal[S_ZS];
    const float pw = (float)s.scal[S_PW];   // level scale x coarse-correction weight
    const bool with_q = st_uni(s, ST_FLEX) != 0 && st_uni(s, ST_VC) != 0;   // q exists once an iteration has run
    const float* __restrict__ rv = r_vc(s);
    const ZT* __restrict__ z = static_cast<const ZT*>(s.z2);   // the pre-smoothed iterate, never written here
    const VT* __restrict__ t1 = reinterpret_cast<const VT*>(s.L[1].t);
    constexpr int NXM = fuse_nxmax(NT, P2, P1), SR = fuse_sr(NT, NXM), SZ = fuse_sz(NT, NXM);
    extern __shared__ float lds[];
    float* Rb = lds + 2;                 // [3][SR]; Rb[-1] is read (and ignored) by an edge pair of an odd nx
    float* Zb = Rb + 3 * SR;             // [3][SZ]
    float* Ob = Zb + 3 * SZ;             // [3][NT]: r of the own columns
    uint8_t* Mb = reinterpret_cast<uint8_t*>(Ob + 3 * NT);   // [3][SR]
    float* Qb = reinterpret_cast<float*>(Mb + 3 * SR);       // [3][NT]: q of the own columns
    auto ij_bits = [&](int c) {                              // the (i, j) part of edge_bits
        const int j = c / nx, i = c - j * nx;
        return (i > 0 ? 1u : 0u) | (i < nx - 1 ? 2u : 0u) | (j > 0 ? 4u : 0u) | (j < ny - 1 ? 8u : 0u);
    };
    __shared__ FuseLut lut;
    fuse_lut_init(lut, tid, NT);
    __syncthreads();
    // staging pairs (as in k_pre_fused)
    const int NP2 = W2 / 2;
    int ca[P2], par2[P2];
    bool ina[P2], inb[P2];
#pragma unroll
    for (int m = 0; m < P2; ++m) {
        const int u = tid + m * NT;
        const int col = c0 - 2 * nx + 2 * u;
        ina[m] = u < NP2 && col >= 0 && col < sxy;
        inb[m] = u < NP2 && col + 1 >= 0 && col + 1 < sxy;
        ca[m] = min(max(col, 0), sxy - 1);
        par2[m] = (ca[m] % nx + ca[m] / nx) & 1;
    }
    // smoothing pairs (as in k_pre_fused)
    const int u1 = nx >> 1, NP1 = ((NT + 3 * nx - 1) >> 1) - u1 + 1;
    int ha1[P1], par1[P1];
    unsigned ija[P1], ijb[P1];
    bool ina1[P1], inb1[P1];
#pragma unroll
    for (int m = 0; m < P1; ++m) {
        const bool use = tid + m * NT < NP1;
        const int h = 2 * (u1 + tid + m * NT), col = c0 - 2 * nx + h;
        ha1[m] = h;
        ina1[m] = use && h >= nx && h < nx + W1 && col >= 0 && col < sxy;
        inb1[m] = use && h + 1 >= nx && h + 1 < nx + W1 && col + 1 >= 0 && col + 1 < sxy;
        const int cca = min(max(col, 0), sxy - 1), ccb = min(max(col + 1, 0), sxy - 1);
        ija[m] = ij_bits(cca);
        ijb[m] = ij_bits(ccb);
        par1[m] = (cca % nx + cca / nx) & 1;
    }
    const unsigned cown = t.valid ? (unsigned)t.c2 : (unsigned)(sxy - 1);   // idle lanes read a valid column
    // Raw loads of one plane, issued three steps before the plane is staged;
    // the level-1 value of each pair's red cell is gathered one step before.
    // Nothing touches a loaded register before then, so the loads stay in
    // flight (pairs outside the grid load a clamped pair that is not used).
    typedef int i32x2 __attribute__((ext_vector_type(2)));
    struct Raw {
        ZP z[P2];
        f32x2 r[P2];
        i32x2 a[P2];      // the pair's level-1 aggregates (0xFFFFFF: none) | face masks << 24
        int g[P2];        // the red cell's aggregate, once gathered
        VT c[P2];         // level-1 value of the red cell's aggregate
        float q;          // q of the own column
    };
    auto load = [&](int p, Raw& X) {
        const unsigned base = (unsigned)sxy * (unsigned)min(max(p, 0), nz - 1);
#pragma unroll
        for (int m = 0; m < P2; ++m) {
            const unsigned i2 = (ca[m] + base) >> 1;
            X.z[m] = gld(reinterpret_cast<const ZP*>(z), i2);
            X.r[m] = gld(reinterpret_cast<const f32x2*>(rv), i2);
            X.a[m] = gld(reinterpret_cast<const i32x2*>(L.aggm), i2);
        }
        X.q = gld(s.q, cown + base);
    };
    auto gather = [&](int p, Raw& X) {
#pragma unroll
        for (int m = 0; m < P2; ++m) {
            const bool ared = ((par2[m] + p) & 1) == 0;
            const int g = (ared ? X.a[m].x : X.a[m].y) & 0xFFFFFF;
            X.g[m] = g;
            X.c[m] = gld(t1, g == 0xFFFFFF ? 0u : (unsigned)g);
        }
    };
    // stage plane p: red cells z + t1[agg] (the prolongation), black cells r zs
    auto stage = [&](int p, const Raw& X, auto S) {   // S: the ring slot of plane p
        constexpr int sp = decltype(S)::value;
        const bool pin = p >= 0 && p < nz;
        float* R = Rb + sp * SR;
        uint8_t* M = Mb + sp * SR;
        float* O = Ob + sp * NT;
#pragma unroll
        for (int m = 0; m < P2; ++m) {
            const int u = tid + m * NT;
            if (u >= NP2) break;
            const bool oka = pin && ina[m], okb = pin && inb[m];
            const bool ared = ((par2[m] + p) & 1) == 0;   // the first cell is red
            const float zr = (float)(ared ? X.z[m].x : X.z[m].y) + (X.g[m] != 0xFFFFFF ? pw * to_f(X.c[m]) : 0.0f);
            const float va = ared ? zr : X.r[m].x * zs, vb = ared ? X.r[m].y * zs : zr;
            *reinterpret_cast<float2*>(R + 2 * u) = float2{oka ? va : 0.0f, okb ? vb : 0.0f};
            *reinterpret_cast<uint16_t*>(M + 2 * u) =
                (uint16_t)((oka ? ((unsigned)X.a[m].x >> 24) : 0u) | ((okb ? ((unsigned)X.a[m].y >> 24) : 0u) << 8));
            const int o = 2 * u - 2 * nx;
            if (o >= 0 && o < NT) *reinterpret_cast<float2*>(O + o) = float2{X.r[m].x, X.r[m].y};
        }
        Qb[sp * NT + tid] = X.q;
    };
    // black cells of plane p over W1 take the Gauss-Seidel update from the
    // prolonged values of their red neighbours.  The red cells' slots are
    // written too (as 0): a stencil reads every neighbour slot, weighting a
    // missing neighbour by 0, and for an even nx the slot past a row's end
    // belongs to a red cell of the next row; a stale value there could be a
    // NaN, which a weight of 0 does not remove.
    auto smooth = [&](int p, auto S) {   // S: the ring slot of plane p
        constexpr int sp = decltype(S)::value;
        const bool pin = p >= 0 && p < nz;
        const unsigned kb = (p > 0 ? 16u : 0u) | (p < nz - 1 ? 32u : 0u);
        const float* R = Rb + sp * SR;
        const float* Rm = Rb + ((sp + 2) % 3) * SR;
        const float* Rp = Rb + ((sp + 1) % 3) * SR;
        const uint8_t* M = Mb + sp * SR;
        float* Z = Zb + sp * SZ - zo;
#pragma unroll
        for (int m = 0; m < P1; ++m) {
            if (tid + m * NT >= NP1) break;
            const bool ared = ((par1[m] + p) & 1) == 0;
            const int hb = ha1[m] + (ared ? 1 : 0);    // the black cell of the pair
            const unsigned mk = M[hb];
            const float2 D = lut.d[mk];   // face count, its reciprocal
            float zb = 0.0f;
            if (pin && (ared ? inb1[m] : ina1[m]) && D.x > 0)   // as the separate black pass stores it in z
                zb = zq<ZT>(div_rn(R[hb] + lat_lut(lut, mk & ((ared ? ijb[m] : ija[m]) | kb), R + hb, nx, Rm[hb],
                                                   Rp[hb]),
                                   D.x, D.y));
            *reinterpret_cast<float2*>(Z + ha1[m]) = ared ? float2{0.0f, zb} : float2{zb, 0.0f};
        }
    };
    double acc = 0.0, accq = 0.0;
    // emit plane k of the own column: its r, mask and q are read before the
    // step's second barrier (the next step's staging reuses their slot)
    const int hz = 2 * nx + tid - zo;   // the own column in Zb
    struct Own {
        float r, q;
        unsigned m;
    };
    auto own = [&](auto S) {
        constexpr int sp = decltype(S)::value;
        return Own{Ob[sp * NT + tid], Qb[sp * NT + tid], (unsigned)Mb[sp * SR + 2 * nx + tid]};
    };
    auto emit = [&](int k, auto S, const Own& o) {
        constexpr int sp = decltype(S)::value;
        if (!t.valid) return;
        const int c = t.c2 + sxy * k;
        const float r = o.r;
        const float* Z = Zb + sp * SZ + hz;
        float zc = *Z;   // black cells: the new value
        if (((t.i + t.j + k) & 1) == 0) {
            const float2 D = lut.d[o.m];
            zc = 0.0f;
            if (D.x > 0)
                zc = zq<ZT>(div_rn(r * zs + lat_lut(lut, o.m & edge_bits(t.i, t.j, k, nx, ny, nz), Z, nx,
                                                    Zb[((sp + 2) % 3) * SZ + hz], Zb[((sp + 1) % 3) * SZ + hz]),
                                   D.x, D.y));
        }
        st_nt<ZT>(s.z, c, zc);   // exact; the dots use the value as stored
        acc += (double)r * zc;
        if (with_q) accq += (double)o.q * zc;
    };
    const int ks = t.k0 - 2;
    // stage k+2 (gathered a step ago), gather for plane k+3, refill the
    // staged buffer with plane k+5, smooth k+1, emit k.  The gathers go out
    // before the plane loads: memory operations complete in order for the
    // wave's counter, so waiting for a gather at the next step then leaves
    // the plane loads in flight.
    // J: the ring slot of plane k, (k - ks) mod 3
    auto step = [&](int k, Raw& X, Raw& Xn, auto J) {
        constexpr int j = decltype(J)::value;
        stage(k + 2, X, ic<(j + 2) % 3>());
        gather(min(k + 3, t.k1 + 1), Xn);
        load(min(k + 5, t.k1 + 1), X);   // past the chunk: repeats the last plane needed
        __syncthreads();
        smooth(k + 1, ic<(j + 1) % 3>());
        const Own o = own(J);
        __syncthreads();
        if (k >= t.k0) emit(k, J, o);
    };
    Raw X0, X1, X2;
    {
        Raw Y;
        load(ks, Y);
        gather(ks, Y);
        stage(ks, Y, ic<0>());
        load(ks + 1, Y);
        gather(ks + 1, Y);
        stage(ks + 1, Y, ic<1>());
    }
    load(ks + 2, X0);
    load(ks + 3, X1);
    load(ks + 4, X2);
    gather(ks + 2, X0);
    int k = ks;
    for (; k + 2 < t.k1; k += 3) {
        step(k, X0, X1, ic<0>());
        step(k + 1, X1, X2, ic<1>());
        step(k + 2, X2, X0, ic<2>());
    }
    if (k < t.k1) step(k, X0, X1, ic<0>());
    if (k + 1 < t.k1) step(k + 1, X1, X2, ic<1>());
    __shared__ double sh[2][NT / 64];
    const int lane = tid & 63, w = tid >> 6;
    acc = wave_sum(acc);
    if (with_q) accq = wave_sum(accq);
    if (lane == 0) {
        sh[0][w] = acc;
        sh[1][w] = accq;
    }
    __syncthreads();
    tile_rows(s, e, c0, [&](int w0, double* out) {
        double a = 0.0, aq = 0.0;
#pragma unroll
        for (int ww = 0; ww < TB / 64; ++ww) {
            a += sh[0][w0 + ww];
            aq += sh[1][w0 + ww];
        }
        out[3] = a;
        if (with_q) out[0] = aq;
    });
}

// z (V-cycle type) -> FP32 copy, for phg_precondition
template <class VT>
__global__ void __launch_bounds__(TB) k_z_out(const DSys* __restrict__ sys, const int2* __restrict__ tab, float* out) {
    const int2 e = tab[blockIdx.x];
    const DSys& s = sys[e.x];
    const ColTile t = col_tile(s, e.y);
    if (!t.valid) return;
    const int sxy = s.L[0].nx * s.L[0].ny;
    const float izs = (float)(1.0 / s.scal[S_ZS]);
    for (int k = t.k0; k < t.k1; ++k) out[t.c2 + sxy * k] = ld<VT>(s.z, t.c2 + sxy * k) * izs;
}

// ---------------------------------------------------------------------------
// V-cycle, coarse levels (ELL operators)
// ---------------------------------------------------------------------------
// Coarse levels 1 .. L-2 (ELL operators).  A block takes up to CV_GRP systems
// that share one grid and 256 unknowns: each thread reads its row of the
// shared operator once and applies it to every system of the group, whose
// own vectors it gathers.  The arithmetic of each system is the same as if it
// were alone.  tab: (first system, count, row tile, L | R << 8) with L the
// levels of the systems' grid and R the rows of the tile, so a block tests
// both before it reads any descriptor.
constexpr int CV_GRP = 4;

struct Grp {
    int first, count;
    unsigned active;   // bit g: system first + g still iterates
};
// The batch's state rows are one array with ST_NSTATE ints per system, so the
// flags are read at an address known from the table entry alone, at the same
// time as the system's descriptor rather than after it.  Each of these kernels
// is a short chain of dependent loads per block, so every load of the CV_GRP
// systems is issued without a test of the group's size: a slot past the group
// reads its last system again (in bounds) and its result is never stored.  A
// load under a per-slot test waits before the next slot's load is issued.
// The flags are read by lanes 0..count-1 of every wave with one vector load
// and combined by a ballot (scalar loads of them were issued one at a time),
// so the whole block calls live_of before any of its threads has left.
__device__ __forceinline__ int slot(const int4& e, int g) { return e.x + min(g, e.y - 1); }
// bit g: system first + g (g < count <= 64) still iterates
__device__ __forceinline__ unsigned long long live_of(const int* __restrict__ states, int first, int count) {
    const int g = (int)(threadIdx.x & 63u);
    const int a = states[(first + min(g, count - 1)) * ST_NSTATE + ST_ACTIVE];
    return __ballot(g < count && a != 0);
}
__device__ __forceinline__ Grp grp_of(const int* __restrict__ states, const int4& e) {
    return Grp{e.x, e.y, (unsigned)live_of(states, e.x, e.y)};   // e.y <= CV_GRP
}

// acc[g] += the sum over the row's ELL entries, in entry order, of
// val * src_g[col], for the systems of a group.  Entries are taken eight at a
// time with every load of the eight issued before the first is used: a loop
// over single entries waits twice per entry (for the column, then for the
// gather through it).  Past the row's width the loads repeat its last entry
// and are not used (the width is the level's, the same for every thread).
constexpr int ELL_UNROLL = 8;
template <class CT, class VT>
__device__ __forceinline__ void ell_accum(const DLevel& L, unsigned I, const void* const (&src)[CV_GRP], int count,
                                          float (&acc)[CV_GRP]) {
    const int K = L.ellk;   // >= 1
    int p0 = 0;
    do {
        int c[ELL_UNROLL];
        CT v[ELL_UNROLL];
#pragma unroll
        for (int q = 0; q < ELL_UNROLL; ++q) {
            const unsigned o = (unsigned)min(p0 + q, K - 1) * (unsigned)L.n + I;
            c[q] = gld(L.col, o);
            v[q] = gld(static_cast<const CT*>(L.val), o);
        }
        VT x[CV_GRP][ELL_UNROLL];
#pragma unroll
        for (int g = 0; g < CV_GRP; ++g)
            if (g < count)
#pragma unroll
                for (int q = 0; q < ELL_UNROLL; ++q) x[g][q] = gld(static_cast<const VT*>(src[g]), (unsigned)c[q]);
#pragma unroll
        for (int q = 0; q < ELL_UNROLL; ++q) {
            if (p0 + q >= K) break;
            const float vq = to_f(v[q]);
#pragma unroll
            for (int g = 0; g < CV_GRP; ++g)
                if (g < count) acc[g] += vq * to_f(x[g][q]);
        }
        p0 += ELL_UNROLL;
    } while (p0 < K);
}

// gather restriction: b_{l+1}[I] = sum over children of r_l, ascending
// (the serial sum of pressure_mg.cpp:285-286 in the same order).  A row has
// at most 8 children at level 0 (a piece of a 2x2x2 block) and rarely more
// further down; the first eight are gathered together.  Where level l+1 is
// not the system's coarsest, the same thread then takes the first Jacobi
// sweep from zero there, x = omega b / diag (jacobi, pressure_mg.cpp:237-247,
// from x = 0: pointwise), which saves a launch and a read of b.
template <class CT, class VT>
__global__ void __launch_bounds__(TB) k_vc_restrict(const DSys* __restrict__ sys, const int* __restrict__ states,
                                                    const int4* __restrict__ tab, int l) {
    const int4 e = tab[blockIdx.x];
    if (l + 1 >= (e.w & 255)) return;
    const DSys& s0 = sys[e.x];
    const Grp G = grp_of(states, e);
    if (!G.active || (int)threadIdx.x >= e.w >> 8) return;
    const DLevel& F = s0.L[l];
    const int I = e.z * TB + threadIdx.x;
    const void* r[CV_GRP];
    float acc[CV_GRP];
#pragma unroll
    for (int g = 0; g < CV_GRP; ++g) {
        r[g] = sys[slot(e, g)].L[l].r;
        acc[g] = 0.0f;
    }
    const bool jac = l + 2 < (e.w & 255);
    const float dg = ld<CT>(s0.L[l + 1].diag, I);   // level l+1 exists: read without a branch
    const int p0 = gld(F.cptr, I), p1 = gld(F.cptr, I + 1);   // p1 > p0: an aggregate has a child
    int pb = p0;
    do {
        int c[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) c[q] = gld(F.cidx, (unsigned)min(pb + q, p1 - 1));
        VT x[CV_GRP][8];
#pragma unroll
        for (int g = 0; g < CV_GRP; ++g)
            if (g < G.count)
#pragma unroll
                for (int q = 0; q < 8; ++q) x[g][q] = gld(static_cast<const VT*>(r[g]), (unsigned)c[q]);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (pb + q >= p1) break;
#pragma unroll
            for (int g = 0; g < CV_GRP; ++g)
                if (g < G.count) acc[g] += to_f(x[g][q]);
        }
        pb += 8;
    } while (pb < p1);
#pragma unroll
    for (int g = 0; g < CV_GRP; ++g) {
        if (!((G.active >> g) & 1u)) continue;
        const DLevel& Lg = sys[G.first + g].L[l + 1];
        // b as stored: its value in VT, which the sweep below uses
        const float b = to_f(from_f<VT>(acc[g] * (1.0f / VC_LEVEL_SCALE)));
        st<VT>(Lg.b, I, b);
        if (jac) st<VT>(Lg.x, I, dg > 0 ? (0.7f * b) / dg : 0.0f);
    }
}

// damped Jacobi sweep src -> dst, omega = 0.7
template <class CT, class VT>
__global__ void __launch_bounds__(TB) k_vc_jac(const DSys* __restrict__ sys, const int* __restrict__ states,
                                               const int4* __restrict__ tab, int l,
                                               int src_is_x) {
    const int4 e = tab[blockIdx.x];
    if (l >= (e.w & 255) - 1) return;
    const DSys& s0 = sys[e.x];
    const Grp G = grp_of(states, e);
    if (!G.active || (int)threadIdx.x >= e.w >> 8) return;
    const DLevel& L = s0.L[l];
    const int I = e.z * TB + threadIdx.x;
    const void* src[CV_GRP];
    float acc[CV_GRP], self[CV_GRP];
#pragma unroll
    for (int g = 0; g < CV_GRP; ++g) {
        const DLevel& Lg = sys[slot(e, g)].L[l];
        src[g] = src_is_x ? Lg.x : Lg.t;
        acc[g] = ld<VT>(Lg.b, I);
        self[g] = ld<VT>(src[g], I);
    }
    const float dg = ld<CT>(L.diag, I);
    ell_accum<CT, VT>(L, (unsigned)I, src, G.count, acc);   // padded rows too: their loads are in bounds
#pragma unroll
    for (int g = 0; g < CV_GRP; ++g) {
        if (!((G.active >> g) & 1u)) continue;
        const DLevel& Lg = sys[G.first + g].L[l];
        float out = 0.0f;
        if (dg > 0) out = (1.0f - 0.7f) * self[g] + (0.7f * acc[g]) / dg;
        st<VT>(src_is_x ? Lg.t : Lg.x, I, out);
    }
}

// coarse residual of the current iterate t (residualSparse, pressure_mg.cpp:248-257)
template <class CT, class VT>
__global__ void __launch_bounds__(TB) k_vc_resid(const DSys* __restrict__ sys, const int* __restrict__ states,
                                                 const int4* __restrict__ tab, int l) {
    const int4 e = tab[blockIdx.x];
    if (l >= (e.w & 255) - 1) return;
    const DSys& s0 = sys[e.x];
    const Grp G = grp_of(states, e);
    if (!G.active || (int)threadIdx.x >= e.w >> 8) return;
    const DLevel& L = s0.L[l];
    const int I = e.z * TB + threadIdx.x;
    const float dg = ld<CT>(L.diag, I);
    const void* t[CV_GRP];
    float r[CV_GRP];
#pragma unroll
    for (int g = 0; g < CV_GRP; ++g) {
        const DLevel& Lg = sys[slot(e, g)].L[l];
        t[g] = Lg.t;
        r[g] = ld<VT>(Lg.b, I) - dg * ld<VT>(Lg.t, I);
    }
    ell_accum<CT, VT>(L, (unsigned)I, t, G.count, r);
#pragma unroll
    for (int g = 0; g < CV_GRP; ++g)
        if ((G.active >> g) & 1u) st<VT>(sys[G.first + g].L[l].r, I, dg > 0 ? r[g] : 0.0f);
}

// t_l += w t_{l+1}[agg]
template <class VT>
__global__ void __launch_bounds__(TB) k_vc_prolong(const DSys* __restrict__ sys, const int* __restrict__ states,
                                                   const int4* __restrict__ tab, int l) {
    const int4 e = tab[blockIdx.x];
    if (l >= (e.w & 255) - 1) return;
    const DSys& s0 = sys[e.x];
    const Grp G = grp_of(states, e);
    if (!G.active || (int)threadIdx.x >= e.w >> 8) return;
    const DLevel& L = s0.L[l];
    const int I = e.z * TB + threadIdx.x;
    const int a = gld(L.agg, I);
    float t[CV_GRP];
#pragma unroll
    for (int g = 0; g < CV_GRP; ++g) {
        const DSys& sg = sys[slot(e, g)];
        t[g] = ld<VT>(sg.L[l].t, I) + (float)sg.scal[S_PW] * ld<VT>(sg.L[l + 1].t, a);
    }
#pragma unroll
    for (int g = 0; g < CV_GRP; ++g)
        if ((G.active >> g) & 1u) st<VT>(sys[G.first + g].L[l].t, I, t[g]);
}

// ---------------------------------------------------------------------------
// V-cycle, the coarse levels that fit in one block's LDS
// ---------------------------------------------------------------------------
// From its first coarse level s whose vectors fit in one block's LDS, a
// system's levels above the coarsest are swept by one block of CT_NT threads
// in two launches instead of five or six per level:
//   k_vc_down  the restriction into s, then on each level s .. nl-2 the first
//              Jacobi sweep from zero, one more sweep, the residual, and the
//              restriction into the next level, ending with the coarsest
//              right-hand side;
//   k_vc_up    after the coarsest solve, on each level nl-2 .. s the
//              prolongation and two sweeps.
// A level's two vectors (iterate and second Jacobi buffer) live in LDS and a
// barrier separates the passes.  Rows are dealt to the threads round-robin,
// CT_U at a time: every load of a thread's CT_U rows is issued before the
// first is used, since a block's time is its chain of dependent loads.
// LDS holds two regions A and B of n_s values each.  Level s keeps its
// iterate in A and its second buffer in B; a level l > s keeps both in A for
// even l - s and in B for odd (2 n_l <= n_s), so a restriction or
// prolongation reads one level while it writes the other.  Each value is
// formed by the same operations as in k_vc_restrict, k_vc_jac, k_vc_resid and
// k_vc_prolong and rounded to VT where they store it, so the results are
// theirs bit for bit; what other launches read (b and t of levels s .. nl-2,
// the coarsest b, t of level s) is stored to the level's vectors as they
// store it.
constexpr int CT_NT = 1024;
constexpr int CT_U = 4;
constexpr int CT_EC = 6;   // ELL entries loaded together (a coarse level's usual row width)

// LDS bytes of the down and up blocks of a system whose LDS levels start at s
__host__ __device__ constexpr size_t tail_lds(int n_s, size_t vts) { return 2 * (size_t)n_s * vts; }

// The CT_U rows I0 + u CT_NT of a pass; rows past n repeat row n - 1 (their
// loads stay in bounds; nothing is stored for them)
template <int U = CT_U>
struct TailRows {
    int I[U];
    bool ok[U];
    __device__ __forceinline__ TailRows(int I0, int n) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            ok[u] = I0 + u * CT_NT < n;
            I[u] = ok[u] ? I0 + u * CT_NT : n - 1;
        }
    }
};

// acc[u] + the sum over row I[u]'s ELL entries, in entry order, of
// val * src[col] (src in LDS): ell_accum for one system and CT_U rows, each
// entry's column and value in one load (the level's packed ELL)
template <class CT, class VT>
__device__ __forceinline__ void ell_rows(const DLevel& L, const TailRows<>& R, const VT* src, float (&acc)[CT_U]) {
    using P = CVPack<CT>;
    const typename P::T* cv = static_cast<const typename P::T*>(L.cv);
    const int K = L.ellk;   // >= 1
    int p0 = 0;
    do {
        typename P::T e[CT_U][CT_EC];
#pragma unroll
        for (int q = 0; q < CT_EC; ++q)
#pragma unroll
            for (int u = 0; u < CT_U; ++u)
                e[u][q] = gld(cv, (unsigned)min(p0 + q, K - 1) * (unsigned)L.n + (unsigned)R.I[u]);
        float x[CT_U][CT_EC];
#pragma unroll
        for (int q = 0; q < CT_EC; ++q)
#pragma unroll
            for (int u = 0; u < CT_U; ++u) x[u][q] = to_f(src[P::col(e[u][q])]);
#pragma unroll
        for (int q = 0; q < CT_EC; ++q) {
            if (p0 + q >= K) break;
#pragma unroll
            for (int u = 0; u < CT_U; ++u) {
                const float vq = P::val(e[u][q]);
                acc[u] += vq * x[u][q];
            }
        }
        p0 += CT_EC;
    } while (p0 < K);
}

// two adjacent cells 2 i2, 2 i2 + 1 of a V-cycle vector in one load
template <class VT> __device__ __forceinline__ float2 ld_pair(const void* p, unsigned i2);
template <> __device__ __forceinline__ float2 ld_pair<float>(const void* p, unsigned i2) {
    const f32x2 v = gld(static_cast<const f32x2*>(p), i2);
    return float2{v.x, v.y};
}
template <> __device__ __forceinline__ float2 ld_pair<_Float16>(const void* p, unsigned i2) {
    const f16x2 v = gld(static_cast<const f16x2*>(p), i2);
    return float2{(float)v.x, (float)v.y};
}
template <> __device__ __forceinline__ float2 ld_pair<bf16_t>(const void* p, unsigned i2) {
    const uint32_t v = gld(static_cast<const uint32_t*>(p), i2);
    return float2{__uint_as_float(v << 16), __uint_as_float(v & 0xFFFF0000u)};
}

// restriction from level 0 into level 1 (k_vc_restrict) by the level-1
// aggregates' 2x2x2 blocks: b of rows R, the children's r0 summed in
// ascending cell order from 0 and scaled (the children list's order).  The
// block words exist for grids of even nx only, where a block's corner cell is
// even, so each of its four x-pairs is one aligned load (a pair without
// children loads the corner's, in bounds).
template <class VT>
__device__ __forceinline__ void block_rows(const DLevel& L0, const TailRows<>& R, float (&b)[CT_U]) {
    const unsigned nx = (unsigned)L0.nx, sxy = nx * (unsigned)L0.ny;
    const unsigned off[4] = {0u, nx, sxy, sxy + nx};
    uint32_t w[CT_U];
#pragma unroll
    for (int u = 0; u < CT_U; ++u) w[u] = gld(L0.cblk, (unsigned)R.I[u]);
    float x[CT_U][8];
#pragma unroll
    for (int u = 0; u < CT_U; ++u) {
        const unsigned base = w[u] & 0xFFFFFFu;
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const float2 v = ld_pair<VT>(L0.r, ((w[u] >> (24 + 2 * h)) & 3u ? base + off[h] : base) >> 1);
            x[u][2 * h] = v.x;
            x[u][2 * h + 1] = v.y;
        }
    }
#pragma unroll
    for (int u = 0; u < CT_U; ++u) {
        float acc = 0.0f;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if ((w[u] >> (24 + q)) & 1u) acc += x[u][q];
        b[u] = to_f(from_f<VT>(acc * (1.0f / VC_LEVEL_SCALE)));
    }
}

// restriction (k_vc_restrict): b of rows R of the next level, the children of
// each in level `from` summed in list order from 0 and scaled; val(c) reads
// child c
template <class VT, int U, class F>
__device__ __forceinline__ void child_rows(const DLevel& from, const TailRows<U>& R, F val, float (&b)[U]) {
    int p0[U], p1[U], m = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        p0[u] = gld(from.cptr, (unsigned)R.I[u]);
        p1[u] = gld(from.cptr, (unsigned)R.I[u] + 1u);   // p1 > p0: an aggregate has a child
    }
    float acc[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        acc[u] = 0.0f;
        m = max(m, p1[u] - p0[u]);
    }
    for (int j = 0; j < m; j += 8) {
        float x[U][8];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            int c[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) c[q] = gld(from.cidx, (unsigned)min(p0[u] + j + q, p1[u] - 1));
#pragma unroll
            for (int q = 0; q < 8; ++q) x[u][q] = val(c[q]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (p0[u] + j + q < p1[u]) acc[u] += x[u][q];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) b[u] = to_f(from_f<VT>(acc[u] * (1.0f / VC_LEVEL_SCALE)));
}

// one damped Jacobi sweep src -> dst (k_vc_jac); the new values also go to
// `out` (a level vector, or null)
template <class CT, class VT>
__device__ __forceinline__ void tail_jac(const DLevel& L, const VT* src, VT* dst, void* out) {
    for (int I0 = (int)threadIdx.x; I0 < L.n; I0 += CT_U * CT_NT) {
        const TailRows<> R(I0, L.n);
        float dg[CT_U], acc[CT_U], self[CT_U];
#pragma unroll
        for (int u = 0; u < CT_U; ++u) {
            dg[u] = ld<CT>(L.diag, (unsigned)R.I[u]);
            acc[u] = ld<VT>(L.b, (unsigned)R.I[u]);
            self[u] = to_f(src[R.I[u]]);
        }
        ell_rows<CT, VT>(L, R, src, acc);
#pragma unroll
        for (int u = 0; u < CT_U; ++u) {
            if (!R.ok[u]) continue;
            float o = 0.0f;
            if (dg[u] > 0) o = (1.0f - 0.7f) * self[u] + (0.7f * acc[u]) / dg[u];
            dst[R.I[u]] = from_f<VT>(o);
            if (out) st<VT>(out, (unsigned)R.I[u], o);
        }
    }
}

// a level's two LDS vectors: level s in A and B, a level l > s in A (even
// l - s) or B (odd)
template <class VT>
__device__ __forceinline__ VT* tail_region(VT* lds, int ns, int l, int s) {
    return lds + (((l - s) & 1) ? ns : 0);
}

template <class CT, class VT>
__global__ void __launch_bounds__(CT_NT) __attribute__((amdgpu_waves_per_eu(8))) k_vc_down(const DSys* __restrict__ sys, const int* __restrict__ states,
                                                   const int2* __restrict__ tab) {
    const int2 e = tab[blockIdx.x];   // (system, first LDS level)
    if (!states[e.x * ST_NSTATE + ST_ACTIVE]) return;
    const DSys& S = sys[e.x];
    const int nl = S.nl, s = e.y, ns = S.L[s].n, tid = (int)threadIdx.x;
    extern __shared__ float tail_sh[];
    VT* const lds = reinterpret_cast<VT*>(tail_sh);
    for (int l = s; l < nl; ++l) {
        const DLevel& F = S.L[l - 1];
        const DLevel& L = S.L[l];
        const int n = L.n;
        const VT* Rl = tail_region(lds, ns, l - 1, s);   // level l-1's residual (l > s)
        VT* X = tail_region(lds, ns, l, s);
        VT* T = l == s ? lds + ns : X + n;
        const bool top = l == nl - 1;
        // restriction into level l, and below the coarsest the first Jacobi
        // sweep from zero into X
        auto restrict_rows = [&](auto rows_tag, auto sum) {
            constexpr int U = decltype(rows_tag)::value;
            for (int I0 = tid; I0 < n; I0 += U * CT_NT) {
                const TailRows<U> R(I0, n);
                float b[U], dg[U];
                sum(R, b);
#pragma unroll
                for (int u = 0; u < U; ++u) dg[u] = top ? 0.0f : ld<CT>(L.diag, (unsigned)R.I[u]);
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (!R.ok[u]) continue;
                    st<VT>(L.b, (unsigned)R.I[u], b[u]);
                    if (!top) X[R.I[u]] = from_f<VT>(dg[u] > 0 ? (0.7f * b[u]) / dg[u] : 0.0f);
                }
            }
        };
        using U2 = std::integral_constant<int, 2>;
        if (l > s)
            restrict_rows(U2{}, [&](const TailRows<2>& R, float (&b)[2]) {
                child_rows<VT>(F, R, [&](int c) { return to_f(Rl[c]); }, b);
            });
        else if (F.cblk)
            restrict_rows(std::integral_constant<int, CT_U>{},
                          [&](const TailRows<>& R, float (&b)[CT_U]) { block_rows<VT>(F, R, b); });
        else
            restrict_rows(U2{}, [&](const TailRows<2>& R, float (&b)[2]) {
                child_rows<VT>(F, R, [&](int c) { return ld<VT>(F.r, (unsigned)c); }, b);
            });
        if (top) break;
        __syncthreads();
        // one sweep x -> t; t is the iterate the up leg corrects
        tail_jac<CT, VT>(L, X, T, L.t);
        __syncthreads();
        // the residual of t into X (k_vc_resid)
        for (int I0 = tid; I0 < n; I0 += CT_U * CT_NT) {
            const TailRows<> R(I0, n);
            float dg[CT_U], r[CT_U];
#pragma unroll
            for (int u = 0; u < CT_U; ++u) {
                dg[u] = ld<CT>(L.diag, (unsigned)R.I[u]);
                r[u] = ld<VT>(L.b, (unsigned)R.I[u]) - dg[u] * to_f(T[R.I[u]]);
            }
            ell_rows<CT, VT>(L, R, T, r);
#pragma unroll
            for (int u = 0; u < CT_U; ++u)
                if (R.ok[u]) X[R.I[u]] = from_f<VT>(dg[u] > 0 ? r[u] : 0.0f);
        }
        __syncthreads();
    }
}

template <class CT, class VT>
__global__ void __launch_bounds__(CT_NT) __attribute__((amdgpu_waves_per_eu(8))) k_vc_up(const DSys* __restrict__ sys, const int* __restrict__ states,
                                                 const int2* __restrict__ tab) {
    const int2 e = tab[blockIdx.x];
    if (!states[e.x * ST_NSTATE + ST_ACTIVE]) return;
    const DSys& S = sys[e.x];
    const int nl = S.nl, s = e.y, ns = S.L[s].n, tid = (int)threadIdx.x;
    extern __shared__ float tail_sh[];
    VT* const lds = reinterpret_cast<VT*>(tail_sh);
    const float pw = (float)S.scal[S_PW];
    for (int l = nl - 2; l >= s; --l) {
        const DLevel& L = S.L[l];
        const DLevel& Cl = S.L[l + 1];
        const int n = L.n;
        // the prolonged iterate P and the middle buffer Q: level s prolongs
        // into A while it reads level s+1 from B
        VT* P = l == s ? lds : tail_region(lds, ns, l, s) + n;
        VT* Q = l == s ? lds + ns : tail_region(lds, ns, l, s);
        const VT* Ct = tail_region(lds, ns, l + 1, s) + Cl.n;   // t of level l+1 below the coarsest
        const bool top = l + 1 == nl - 1;
        // t_l += w t_{l+1}[agg] (k_vc_prolong)
        for (int I0 = tid; I0 < n; I0 += CT_U * CT_NT) {
            const TailRows<> R(I0, n);
            int a[CT_U];
            float t[CT_U], tc[CT_U];
#pragma unroll
            for (int u = 0; u < CT_U; ++u) {
                a[u] = gld(L.agg, (unsigned)R.I[u]);
                t[u] = ld<VT>(L.t, (unsigned)R.I[u]);
            }
#pragma unroll
            for (int u = 0; u < CT_U; ++u) tc[u] = top ? ld<VT>(Cl.t, (unsigned)a[u]) : to_f(Ct[a[u]]);
#pragma unroll
            for (int u = 0; u < CT_U; ++u)
                if (R.ok[u]) P[R.I[u]] = from_f<VT>(t[u] + pw * tc[u]);
        }
        __syncthreads();
        tail_jac<CT, VT>(L, P, Q, nullptr);
        __syncthreads();
        tail_jac<CT, VT>(L, Q, P, l == s ? L.t : nullptr);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// coarsest level: t = B b in FP64, B the projected inverse of the grounded
// coarsest operator (symmetric, so column i is read as row i).  A block takes
// up to CGS systems that share one grid, so each entry of B is read once for
// all of them, and CG_ROWS rows: lane = row, B read coalesced along a column,
// the CGS right-hand sides broadcast from LDS ([n][CGS] floats: the V-cycle
// vectors hold at most FP32, so the copy is exact).  The four waves split the
// column range and their partials are added in a fixed order.  A wave issues
// the loads of CG_UNROLL columns before their products, which it adds in
// column order, so the sums do not depend on CG_UNROLL or CGS.
// The reference approximates this solve with 20 forward/backward Gauss-Seidel
// sweeps (pressure_mg.cpp:277-281), a chain of dependent row updates that
// leaves a GPU mostly idle; k_vc_coarse_gs runs them where a grid asks for
// the reference's iterates.
constexpr int CG_ROWS = 64;
constexpr int CG_SYS_MAX = 16;
constexpr int CG_UNROLL = 8;

// LDS bytes of k_vc_coarse for `cgs` systems of up to `n` unknowns
__host__ __device__ constexpr size_t coarse_lds(int cgs, int n) {
    return (size_t)n * cgs * 4 > (size_t)3 * CG_ROWS * cgs * 8 ? (size_t)n * cgs * 4 : (size_t)3 * CG_ROWS * cgs * 8;
}

// acc[g] += a b[g] for the CGS right-hand sides of one column
template <int CGS>
__device__ __forceinline__ void coarse_acc(double (&acc)[CGS], double a, const float* b) {
    if constexpr (CGS % 4 == 0) {
        const float4* bv = reinterpret_cast<const float4*>(b);
#pragma unroll
        for (int q = 0; q < CGS / 4; ++q) {
            const float4 b4 = bv[q];
            acc[4 * q + 0] += a * (double)b4.x;
            acc[4 * q + 1] += a * (double)b4.y;
            acc[4 * q + 2] += a * (double)b4.z;
            acc[4 * q + 3] += a * (double)b4.w;
        }
    } else {
#pragma unroll
        for (int g = 0; g < CGS; ++g) acc[g] += a * (double)b[g];
    }
}

// b of the group's systems into LDS, [rows][CGS] from row r0 (rows past n
// and systems past the group read as 0), and their coarsest t (or, for
// `tsel` 1, their product buffers) into tp.  Thread k takes system k % CGS
// throughout (TB is a multiple of CGS), so it reads that system's pointer
// once, and issues eight loads before it stores any: a loop of load, store
// per row and system waits once per element.
template <class VT, int CGS>
__device__ __forceinline__ void coarse_stage(const DSys* __restrict__ sys, const int4& e, int top, int r0, int rows,
                                             int n, float* bsh, void** tp, int tsel = 0) {
    constexpr int JS = TB / CGS, SU = 8;
    static_assert(TB % CGS == 0, "a thread keeps one system");
    const int g = (int)threadIdx.x % CGS;
    const DSys& sg = sys[e.x + min(g, e.y - 1)];
    const void* b = sg.L[top].b;
    if ((int)threadIdx.x < CGS) tp[g] = tsel ? (void*)sg.cpart : sg.L[top].t;
    for (int j0 = (int)threadIdx.x / CGS; j0 < rows; j0 += SU * JS) {
        float v[SU];
#pragma unroll
        for (int u = 0; u < SU; ++u) v[u] = ld<VT>(b, (unsigned)min(r0 + j0 + u * JS, n - 1));
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            const int j = j0 + u * JS;
            if (j < rows) bsh[j * CGS + g] = g < e.y && r0 + j < n ? v[u] : 0.0f;
        }
    }
}

template <class VT, int CGS>
__global__ void __launch_bounds__(TB) k_vc_coarse(const DSys* __restrict__ sys, const int* __restrict__ states,
                                                  const int4* __restrict__ tab) {
    const int4 e = tab[blockIdx.x];   // first system, count, row tile
    const unsigned long long live = live_of(states, e.x, e.y);
    if (!live) return;
    const DSys& s0 = sys[e.x];
    const int n = s0.nc;
    const double* __restrict__ A = s0.cinv;
    extern __shared__ double dsh[];
    __shared__ void* tsh[CGS];
    float* bsh = reinterpret_cast<float*>(dsh);   // [n][CGS]
    coarse_stage<VT, CGS>(sys, e, s0.nl - 1, 0, n, n, bsh, tsh);
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = e.z * CG_ROWS + lane;
    const int ic = i < n ? i : n - 1;
    const int jper = (n + 3) / 4, j0 = w * jper, j1 = min(n, j0 + jper);
    double acc[CGS];
#pragma unroll
    for (int g = 0; g < CGS; ++g) acc[g] = 0.0;
    int j = j0;
    for (; j + CG_UNROLL <= j1; j += CG_UNROLL) {
        double a[CG_UNROLL];
#pragma unroll
        for (int u = 0; u < CG_UNROLL; ++u) a[u] = gld(A, (unsigned)(j + u) * (unsigned)n + (unsigned)ic);
#pragma unroll
        for (int u = 0; u < CG_UNROLL; ++u) coarse_acc<CGS>(acc, a[u], bsh + (j + u) * CGS);
    }
    for (; j < j1; ++j) coarse_acc<CGS>(acc, gld(A, (unsigned)j * (unsigned)n + (unsigned)ic), bsh + j * CGS);
    __syncthreads();   // bsh is reused for the partials of waves 1..3
    double* red = dsh;   // [3][CG_ROWS][CGS]
    if (w > 0) {
#pragma unroll
        for (int g = 0; g < CGS; ++g) red[((w - 1) * CG_ROWS + lane) * CGS + g] = acc[g];
    }
    __syncthreads();
    if (w == 0 && i < n) {
#pragma unroll
        for (int g = 0; g < CGS; ++g) {
            if (g >= e.y) break;
            const double v = acc[g] + red[(0 * CG_ROWS + lane) * CGS + g] + red[(1 * CG_ROWS + lane) * CGS + g] +
                             red[(2 * CG_ROWS + lane) * CGS + g];
            if ((live >> g) & 1u) st<VT>(tsh[g], i, (float)v);
        }
    }
}

// Coarsest level on the FP64 matrix cores for grids that many systems of the
// batch share: t_j = B b_j for all of them is one GEMM (nc x nc times
// nc x m).  A block takes CM_ROWS = 16 rows and 16 ST systems of one grid;
// its four waves split the columns, wave w taking the w-th quarter of every
// chunk of CM_KC.  For each group of four columns a wave issues ST
// v_mfma_f64_16x16x4f64: lane l holds B[row l % 16][col l / 16] (B is
// symmetric, so that is column row of the stored matrix, read coalesced) and
// b_{l % 16}[col l / 16], and accumulates t of rows l / 16 + 4 r and system
// l % 16 of the tile.  Wave 0 then adds the other waves' partial sums in wave
// order.  The right-hand sides go through LDS CM_KC columns at a time in their
// storage type, zero past nc, in two buffers: while a chunk's MFMAs run, the
// next chunk's right-hand sides and matrix columns are already on their way
// into registers, and each chunk costs one load round trip, so chunks are
// long.  Small tiles give the chip several
// waves per SIMD, which the MFMA chains need to issue back to back.  Each
// output comes from its own row of B and column of b only, so a system's
// result does not depend on the other systems of its block; it differs from
// the vector kernel's column-ordered sums by FP64 rounding only.  The inverse
// is read once per 16 ST systems instead of once per 16.
constexpr int CM_ROWS = 16, CM_KC = 128;
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <class VT, int ST>
__global__ void __launch_bounds__(TB) k_vc_coarse_mfma(const DSys* __restrict__ sys, const int* __restrict__ states,
                                                       const int4* __restrict__ tab) {
    constexpr int NS = 16 * ST;
    constexpr int BL = CM_KC * NS / TB;     // right-hand-side values per thread per chunk
    constexpr int AW = CM_KC / 4 / 4;      // MFMA column groups per wave per chunk (a quarter of it)
    static_assert(TB == 4 * 64, "four waves");
    static_assert(CM_KC * NS % TB == 0 && TB % CM_KC == 0, "whole staging rounds");
    const int4 e = tab[blockIdx.x];   // first system, count (<= NS), row tile
    const unsigned long long live = live_of(states, e.x, e.y);
    if (!live) return;
    const DSys& s0 = sys[e.x];
    const int n = s0.nc, top = s0.nl - 1;
    const double* __restrict__ A = s0.cinv;
    // two staging buffers, padded so the transposing store hits distinct
    // banks; after the last chunk they hold the partial sums of waves 1..3
    constexpr int SB = 2 * CM_KC * (NS + 1);
    static_assert(SB * sizeof(VT) >= 3 * 64 * 4 * ST * sizeof(double), "partials fit the staging buffers");
    __shared__ __align__(16) VT bsh[SB];
    __shared__ const void* bp[NS];
    __shared__ void* tp[NS];
    const int tid = (int)threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid < NS) {
        const DSys& sg = sys[e.x + min(tid, e.y - 1)];
        bp[tid] = sg.L[top].b;
        tp[tid] = sg.L[top].t;
    }
    __syncthreads();
    const unsigned arow = (unsigned)min(e.z * CM_ROWS + (lane & 15), n - 1);
    const int kq = lane >> 4;
    // staging: thread tid takes column k0 + tid % CM_KC of systems
    // tid / CM_KC + u TB / CM_KC, consecutive threads on consecutive columns
    const int skk = tid % CM_KC, sj = tid / CM_KC;
    // the raw values stay in their storage type until stash: converting them
    // where they are loaded would wait for the loads before the MFMAs
    VT bn[BL];
    double an[AW];
    auto fetch = [&](int k0) {
        const unsigned k = (unsigned)min(k0 + skk, n - 1);
#pragma unroll
        for (int u = 0; u < BL; ++u) bn[u] = gld(static_cast<const VT*>(bp[sj + u * (TB / CM_KC)]), k);
#pragma unroll
        for (int u = 0; u < AW; ++u)
            an[u] = gld(A, (unsigned)min(k0 + CM_KC / 4 * w + 4 * u + kq, n - 1) * (unsigned)n + arow);
    };
    const VT zero = from_f<VT>(0.0f);
    auto stash = [&](int buf, int k0) {
        const bool kin = k0 + skk < n;
#pragma unroll
        for (int u = 0; u < BL; ++u) {
            const int j = sj + u * (TB / CM_KC);
            bsh[(buf * CM_KC + skk) * (NS + 1) + j] = kin && j < e.y ? bn[u] : zero;
        }
    };
    f64x4 acc[ST];
#pragma unroll
    for (int t = 0; t < ST; ++t) acc[t] = f64x4{0.0, 0.0, 0.0, 0.0};
    fetch(0);
    stash(0, 0);
    __syncthreads();
    for (int k0 = 0, buf = 0; k0 < n; k0 += CM_KC, buf ^= 1) {
        double a[AW];
#pragma unroll
        for (int u = 0; u < AW; ++u) a[u] = an[u];
        const bool more = k0 + CM_KC < n;
        if (more) fetch(k0 + CM_KC);
#pragma unroll
        for (int u = 0; u < AW; ++u) {
            const VT* brow = bsh + (buf * CM_KC + CM_KC / 4 * w + 4 * u + kq) * (NS + 1) + (lane & 15);
#pragma unroll
            for (int t = 0; t < ST; ++t)
                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], (double)to_f(brow[16 * t]), acc[t], 0, 0, 0);
        }
        if (more) stash(buf ^ 1, k0 + CM_KC);
        __syncthreads();
    }
    double* part = reinterpret_cast<double*>(bsh);   // [3][4 ST][64]
    if (w > 0) {
#pragma unroll
        for (int t = 0; t < ST; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) part[((w - 1) * 4 * ST + 4 * t + r) * 64 + lane] = acc[t][r];
    }
    __syncthreads();
    if (w > 0) return;
#pragma unroll
    for (int t = 0; t < ST; ++t) {
        const int j = 16 * t + (lane & 15);
        if (j >= e.y || !((live >> j) & 1u)) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = e.z * CM_ROWS + kq + 4 * r;
            const double v = acc[t][r] + part[(4 * t + r) * 64 + lane] + part[((4 * ST) + 4 * t + r) * 64 + lane] +
                             part[((8 * ST) + 4 * t + r) * 64 + lane];
            if (i < n) st<VT>(tp[j], (unsigned)i, (float)v);
        }
    }
}

// Coarsest level by the reference's own method (pressure_mg.cpp:258-281):
// from x = 0, 20 times a forward then a backward Gauss-Seidel sweep,
// x_I = (b_I + sum val x_col) / diag_I over the row's entries in order,
// x_I = 0 where diag_I <= 0.  One block per system keeps x and b in LDS and
// runs each sweep level by level (DSys::gs_ord): the rows of a level read
// only rows of earlier levels (updated in this sweep) or of later ones (not
// yet), exactly as the serial sweep does, so every x_I is formed by the same
// operations on the same values.  No contraction into FMA, as in the
// reference's arithmetic.  Used for grids asked to follow the reference's
// iterates (phg_grid_set_coarse) and for coarsest levels too large for the
// dense inverse.
template <class CC, class VT>
__global__ void __launch_bounds__(TB) k_vc_coarse_gs(const DSys* __restrict__ sys, const int* __restrict__ states,
                                                     const int2* __restrict__ tab) {
#pragma clang fp contract(off)
    const int si = tab[blockIdx.x].x;
    if (states[si * ST_NSTATE + ST_ACTIVE] == 0) return;
    const DSys& s = sys[si];
    const DLevel& L = s.L[s.nl - 1];
    const int n = L.n, K = L.ellk;
    extern __shared__ float gsh[];
    float* x = gsh;        // [n]
    float* b = gsh + n;    // [n]
    for (int i = threadIdx.x; i < n; i += TB) {
        x[i] = 0.0f;
        b[i] = ld<VT>(L.b, (unsigned)i);
    }
    __syncthreads();
    for (int it = 0; it < 20; ++it) {
        for (int dir = 0; dir < 2; ++dir) {
            const int* __restrict__ ord = s.gs_ord + (size_t)dir * n;
            const int* __restrict__ lp = s.gs_lptr + (dir ? s.gs_nf + 1 : 0);
            const int nlev = dir ? s.gs_nb : s.gs_nf;
            for (int lv = 0; lv < nlev; ++lv) {
                for (int t = lp[lv] + (int)threadIdx.x; t < lp[lv + 1]; t += TB) {
                    const int I = ord[t];
                    const float dg = ld<CC>(L.diag, (unsigned)I);
                    float v = 0.0f;
                    if (dg > 0) {
                        // padded entries (col = I, val = 0) add an exact zero
                        float acc = b[I];
                        for (int q = 0; q < K; ++q) {
                            const unsigned e = (unsigned)q * (unsigned)n + (unsigned)I;
                            acc = acc + ld<CC>(L.val, e) * x[gld(L.col, e)];
                        }
                        v = acc / dg;
                    }
                    x[I] = v;
                }
                __syncthreads();
            }
        }
    }
    for (int i = threadIdx.x; i < n; i += TB) st<VT>(L.t, (unsigned)i, x[i]);
}

// Coarsest level of a grid with more than CP_MIN unknowns: the inverse is
// exactly symmetric, so only its upper triangle is stored, as 64x64 tiles
// (I <= J, tile-major in row order, each column-major: element (r, c) of
// tile (I, J) is B[64 I + r][64 J + c], zero past nc).  A block takes one tile
// and up to CGS systems of its grid: it reads the tile once (lane = row,
// coalesced along a column) and forms both of its products, B_IJ b_J for the
// rows of strip I and, for I < J, B_IJ^T b_I for the rows of strip J, through
// a transposed copy in LDS.  Each product sums the block's 64 terms in wave
// order; k_vc_coarse_r then adds the 64-row products of a strip over all
// strips S in order, so t_R = sum_S B_RS b_S.  Every element of the inverse
// is read once per group instead of twice.
constexpr int CP_MIN = 1024;

__host__ __device__ constexpr size_t coarse_p_lds(int cgs) {
    return (size_t)64 * 65 * 8 + (size_t)2 * 64 * cgs * 4 + (size_t)2 * 3 * 64 * cgs * 8;
}

template <class VT, int CGS>
__global__ void __launch_bounds__(TB) k_vc_coarse_p(const DSys* __restrict__ sys, const int* __restrict__ states,
                                                    const int4* __restrict__ tab) {
    const int4 e = tab[blockIdx.x];   // first system, count, strip I, strip J >= I
    if (!live_of(states, e.x, e.y)) return;
    const DSys& s0 = sys[e.x];
    const int n = s0.nc, nt = s0.cnt, I = e.z, J = e.w;
    extern __shared__ double dsh[];
    __shared__ void* psh[CGS];
    double* tl = dsh;                                           // [c][65]: column c of the tile
    float* bI = reinterpret_cast<float*>(dsh + 64 * 65);       // [64][CGS]
    float* bJ = bI + 64 * CGS;                                  // [64][CGS]
    double* redr = reinterpret_cast<double*>(bJ + 64 * CGS);   // [3][64][CGS]
    double* redc = redr + 3 * 64 * CGS;                         // [3][64][CGS]
    coarse_stage<VT, CGS>(sys, e, s0.nl - 1, 64 * I, 64, n, bI, psh, 1);
    coarse_stage<VT, CGS>(sys, e, s0.nl - 1, 64 * J, 64, n, bJ, psh, 1);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const size_t tix = (size_t)I * nt - (size_t)I * (I - 1) / 2 + (J - I);
    const double* __restrict__ T = s0.cinvp + tix * 4096;
    double a[16];   // row lane, columns 16 w .. 16 w + 15
#pragma unroll
    for (int i = 0; i < 16; ++i) a[i] = gld(T, (unsigned)((16 * w + i) * 64 + lane));
#pragma unroll
    for (int i = 0; i < 16; ++i) tl[(16 * w + i) * 65 + lane] = a[i];
    __syncthreads();
    double ar[CGS], ac[CGS];
#pragma unroll
    for (int g = 0; g < CGS; ++g) ar[g] = ac[g] = 0.0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const float* bj = bJ + (16 * w + i) * CGS;
#pragma unroll
        for (int g = 0; g < CGS; ++g) ar[g] += a[i] * (double)bj[g];
    }
    if (I < J) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const double t = tl[lane * 65 + 16 * w + i];   // element (16 w + i, lane)
            const float* bi = bI + (16 * w + i) * CGS;
#pragma unroll
            for (int g = 0; g < CGS; ++g) ac[g] += t * (double)bi[g];
        }
    }
    if (w > 0) {
#pragma unroll
        for (int g = 0; g < CGS; ++g) {
            redr[((w - 1) * 64 + lane) * CGS + g] = ar[g];
            redc[((w - 1) * 64 + lane) * CGS + g] = ac[g];
        }
    }
    __syncthreads();
    if (w == 0) {
#pragma unroll
        for (int g = 0; g < CGS; ++g) {
            if (g >= e.y) break;
            double* P = static_cast<double*>(psh[g]);
            P[((size_t)I * nt + J) * 64 + lane] = ar[g] + redr[(0 * 64 + lane) * CGS + g] +
                                                  redr[(1 * 64 + lane) * CGS + g] + redr[(2 * 64 + lane) * CGS + g];
            if (I < J)
                P[((size_t)J * nt + I) * 64 + lane] = ac[g] + redc[(0 * 64 + lane) * CGS + g] +
                                                      redc[(1 * 64 + lane) * CGS + g] + redc[(2 * 64 + lane) * CGS + g];
        }
    }
}

// t_R = sum over strips S, in order, of the products k_vc_coarse_p left;
// a wave per system of the group, lane = row
template <class VT>
__global__ void __launch_bounds__(TB) k_vc_coarse_r(const DSys* __restrict__ sys, const int* __restrict__ states,
                                                    const int4* __restrict__ tab) {
    const int4 e = tab[blockIdx.x];   // first system, count, strip R
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long live = live_of(states, e.x, e.y);
    for (int g = w; g < e.y; g += TB / 64) {
        const DSys& s = sys[e.x + g];
        if (!((live >> g) & 1u)) continue;
        const int nt = s.cnt, i = 64 * e.z + lane;
        const double* __restrict__ P = s.cpart + (size_t)e.z * nt * 64 + lane;
        double y = 0.0;
        int S = 0;
        for (; S + 8 <= nt; S += 8) {
            double p[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) p[u] = P[(size_t)(S + u) * 64];
#pragma unroll
            for (int u = 0; u < 8; ++u) y += p[u];
        }
        for (; S < nt; ++S) y += P[(size_t)S * 64];
        if (i < s.nc) st<VT>(s.L[s.nl - 1].t, i, (float)y);
    }
}

// ---------------------------------------------------------------------------
// the residual kept in range(A).  On a closed zone A's null space holds the
// constant vector, and b is free of it, so r is in exact arithmetic; but the
// FP32 rounding of r -= alpha q adds a small constant per zone, taken on
// mostly in the first iterations, when alpha q is large.  CG cannot reduce
// it, and once the rest of r has fallen to its size the recursion loses
// conjugacy: ||r|| grows by orders of magnitude until d.Ad <= 0 (a point
// source in a closed room shows it near 1e-6).  The reference keeps r as it
// comes (gasphase.cpp:603-627).  Here, when ||r|| / ||b|| first falls below
// 1e-3, 1e-6, ..., r's mean on each closed zone is removed, over two
// iterations with one launch of k_proj each (ST_PROJ: k_fin_beta sets 1 at
// the crossing, k_fin_alpha turns 1 into 2 and 2 into 0):
// - 1: fixed-order FP64 partial sums of r_k over the chunks of each zone's
//   cell list (the tasks of the final mean removal);
// - 2: each zone's mean formed from them in task order and subtracted in FP32
//   from r_{k+1} = r_k - alpha q, as k_zone_sub does.  The mean of alpha q on
//   a closed zone is rounding only, so this leaves one iteration's rounding.
// A block takes PROJ_TASKS tasks, so the launch of an iteration without a
// projection stays small.
// ---------------------------------------------------------------------------
constexpr int PROJ_TASKS = 8;

// Most iterations have no projection due: lane q of every wave reads task q's
// flags in the batch's state array, and a ballot tells the block which tasks
// to take, so an idle block leaves after two dependent loads.
__global__ void __launch_bounds__(TB) k_proj(const DSys* __restrict__ sys, const int* __restrict__ states,
                                             const int2* __restrict__ tab, int ntab) {
    __shared__ double out[1];
    const int lane = (int)(threadIdx.x & 63u), base = blockIdx.x * PROJ_TASKS;
    const int* sq = states + tab[min(base + min(lane, PROJ_TASKS - 1), ntab - 1)].x * ST_NSTATE;
    const unsigned long long due =
        __ballot(lane < PROJ_TASKS && base + lane < ntab && sq[ST_PROJ] != 0 && sq[ST_ACTIVE] != 0);
    for (int q = 0; q < PROJ_TASKS; ++q) {
        if (!((due >> q) & 1u)) continue;
        const int idx = base + q;
        const int2 e = tab[idx];
        const DSys& s = sys[e.x];
        const int pj = st_uni(s, ST_PROJ);
        float* __restrict__ r = rbuf(s, st_uni(s, ST_ITERS));   // r of the iteration just finished
        const int4 tk = s.ztask[e.y];
        int p = tk.y + threadIdx.x;
        if (pj == 1) {
            // a thread adds its cells in increasing order, eight loads issued before their use
            double v[1] = {0.0};
            for (; p + 7 * TB < tk.z; p += 8 * TB) {
                int c[8];
                float x[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) c[u] = s.zcells[p + u * TB];
#pragma unroll
                for (int u = 0; u < 8; ++u) x[u] = r[c[u]];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[0] += (double)x[u];
            }
            for (; p < tk.z; p += TB) v[0] += (double)r[s.zcells[p]];
            block_sums<1>(v, out);
            __syncthreads();
            if (threadIdx.x == 0) s.zpart[e.y] = out[0];
            __syncthreads();
        } else {
            const int t0 = s.ztptr[tk.x], t1 = s.ztptr[tk.x + 1];
            double acc = 0.0;
            for (int t = t0; t < t1; ++t) acc += s.zpart[t];
            const float mean = (float)(acc / (double)(s.ztask[t1 - 1].z - s.ztask[t0].y));
            for (; p + 7 * TB < tk.z; p += 8 * TB) {
                int c[8];
                float x[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) c[u] = s.zcells[p + u * TB];
#pragma unroll
                for (int u = 0; u < 8; ++u) x[u] = r[c[u]];
#pragma unroll
                for (int u = 0; u < 8; ++u) r[c[u]] = x[u] - mean;
            }
            for (; p < tk.z; p += TB) r[s.zcells[p]] -= mean;
        }
    }
}

// ---------------------------------------------------------------------------
// closed-zone mean removal for every system of a batch (gasphase.cpp:628-639):
// fixed-order FP64 sums over chunks of each zone's cell list, then per zone
// ---------------------------------------------------------------------------
// `which`: 0 the solution x (after the solve), 1 the residual r_0 of a system
// restarted from its true residual, whose rounding can leave a component along
// the null space (a constant per closed zone) that CG cannot reduce
// the three steps shared by the batched kernels and the workspace kernels
// of phg_zone_mean_noalloc (same sums in the same order)
__device__ __forceinline__ void zone_task_sum(const int4 tk, const int* __restrict__ zcells,
                                              const float* __restrict__ x, double* __restrict__ slot) {
    double v[1] = {0.0};
    for (int p = tk.y + threadIdx.x; p < tk.z; p += TB) v[0] += (double)x[zcells[p]];
    __shared__ double out[1];
    block_sums<1>(v, out);
    __syncthreads();
    if (threadIdx.x == 0) *slot = out[0];
}

__device__ __forceinline__ void zone_means(int nclosed, const int* __restrict__ ztptr, const int4* __restrict__ ztask,
                                           const double* __restrict__ zpart, double* __restrict__ zmean) {
    for (int z = threadIdx.x; z < nclosed; z += TB) {
        double acc = 0.0;
        int cnt = 0;
        for (int t = ztptr[z]; t < ztptr[z + 1]; ++t) {
            acc += zpart[t];
            cnt += ztask[t].z - ztask[t].y;
        }
        zmean[z] = cnt > 0 ? acc / (double)cnt : 0.0;
    }
}

__global__ void __launch_bounds__(TB) k_zone_part(const DSys* __restrict__ sys, const int2* __restrict__ tab,
                                                  int which) {
    const int2 e = tab[blockIdx.x];
    const DSys& s = sys[e.x];
    if (which == 1 && s.state[ST_ACTIVE] == 0) return;
    zone_task_sum(s.ztask[e.y], s.zcells, which == 1 ? s.r : s.x, &s.zpart[e.y]);
}

__global__ void __launch_bounds__(TB) k_zone_fin(const DSys* __restrict__ sys, const int2* __restrict__ tab,
                                                 int which) {
    const DSys& s = sys[tab[blockIdx.x].x];
    if (which == 1 && s.state[ST_ACTIVE] == 0) return;
    zone_means(s.nclosed, s.ztptr, s.ztask, s.zpart, s.zmean);
}

__global__ void __launch_bounds__(TB) k_zone_sub(const DSys* __restrict__ sys, const int2* __restrict__ tab,
                                                 int which) {
    const int2 e = tab[blockIdx.x];
    const DSys& s = sys[e.x];
    if (s.nclosed == 0) return;
    if (which == 1 && s.state[ST_ACTIVE] == 0) return;
    float* x = which == 1 ? s.r : s.x;
    const ColTile t = col_tile(s, e.y);
    if (!t.valid) return;
    const int sxy = s.L[0].nx * s.L[0].ny;
    for (int k = t.k0; k < t.k1; ++k) {
        const int c = t.c2 + sxy * k;
        const int z = s.zclosed[c];
        if (z >= 0) x[c] -= (float)s.zmean[z];
    }
}

// Workspace forms for one grid: the partials and the means live in a buffer
// the caller names (ws: ntask partials, then nclosed means), so no batch and
// no allocation is involved.  One block per task; one block for the means;
// one thread per cell for the subtraction.
__global__ void __launch_bounds__(TB) k_zone_part_ws(const int4* __restrict__ ztask, const int* __restrict__ zcells,
                                                     const float* __restrict__ x, double* __restrict__ ws) {
    zone_task_sum(ztask[blockIdx.x], zcells, x, &ws[blockIdx.x]);
}

__global__ void __launch_bounds__(TB) k_zone_fin_ws(int nclosed, int ntask, const int* __restrict__ ztptr,
                                                    const int4* __restrict__ ztask, double* __restrict__ ws) {
    zone_means(nclosed, ztptr, ztask, ws, ws + ntask);
}

__global__ void __launch_bounds__(TB) k_zone_sub_ws(int n, const int* __restrict__ zclosed,
                                                    const double* __restrict__ zmean, float* __restrict__ x) {
    const int c = blockIdx.x * TB + threadIdx.x;
    if (c >= n) return;
    const int z = zclosed[c];
    if (z >= 0) x[c] -= (float)zmean[z];
}

}  // namespace phg

namespace phg {
// y = A x on level 0 for phg_apply (FP32 operator), partial x.y in slot 0
template <class OT>
__global__ void __launch_bounds__(TB) k_apply_plain(const DSys* __restrict__ sys, const int2* __restrict__ tab,
                                                    const float* __restrict__ x, float* __restrict__ y) {
    const int2 e = tab[blockIdx.x];
    const DSys& s = sys[e.x];
    const DLevel& L = s.L[0];
    const ColTile t = col_tile(s, e.y);
    const int sxy = L.nx * L.ny;
    double v[1] = {0.0};
    if (t.valid) {
        for (int k = t.k0; k < t.k1; ++k) {
            const Cell q{t.c2 + sxy * k, t.i, t.j, k};
            float a[6];
            coefs6<OT>(L, L.ax32, L.ay32, L.az32, q, a);
            const float dg = diag6(a);
            float yc = 0.0f;
            if (dg > 0) yc = dg * x[q.c] - nb_sum_f(q, L.nx, L.ny, L.nz, a, [&](int n) { return x[n]; });
            y[q.c] = yc;
            v[0] += (double)yc * x[q.c];
        }
    }
    block_sums<1>(v, s.part + (long)e.y * NPART);
}

// plain sum of partial slot 0 into scal[S_DQ] (no state change)
__global__ void __launch_bounds__(TB) k_fin_dot(const DSys* __restrict__ sys, const int2* __restrict__ tab) {
    const DSys& s = sys[tab[blockIdx.x].x];
    __shared__ double res[1];
    sum_parts(s, 0, &res[0]);
    __syncthreads();
    if (threadIdx.x == 0) s.scal[S_DQ] = res[0];
}
}  // namespace phg
