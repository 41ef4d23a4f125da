// kernels_pcg.hpp -- batched MG-PCG kernels for gfx950.
//
// Every kernel works on a batch of independent systems.  A block handles one
// 256-unknown tile of one system; the (system, tile) pair comes from a tile
// table built at batch creation and ordered so that consecutive tiles of a
// system land on the same XCD (blocks are dispatched round-robin over the
// 8 XCDs, each with its own L2).  Systems that have stopped leave at once.
//
// Reference: the CG loop src/sim/gasphase.cpp:571-627 (apply :551-570, FP64
// dot products, FP64 alpha/beta, FP32 vector updates) and the V-cycle
// src/sim/pressure_mg.cpp:195-300.  Level 0 of the V-cycle works in place on
// the outer vectors: its right-hand side is the CG residual r and its
// iterate is z, so no copy-in/copy-out (pressure_mg.cpp:273,299) is needed.
#pragma once
#include "device.hpp"

namespace phg {

// ---------------------------------------------------------------------------
// level 0 stencil access
// ---------------------------------------------------------------------------
struct Cell {
    int c, i, j, k;
};

__device__ __forceinline__ Cell tile_cell(int tile, int nx, int ny) {
    Cell q;
    q.c = tile * TB + threadIdx.x;
    cell_ijk(q.c, nx, ny, q.i, q.j, q.k);
    return q;
}

// the six face coefficients of a cell: -x, +x, -y, +y, -z, +z
template <class CT>
__device__ __forceinline__ void coefs6(const void* ax, const void* ay, const void* az, const Cell& q, int nx, int ny,
                                       float a[6]) {
    const int xf = q.c + q.j + ny * q.k;   // i + (nx+1)(j + ny k)
    const int yf = q.c + nx * q.k;         // i + nx (j + (ny+1) k)
    a[0] = ld<CT>(ax, xf);
    a[1] = ld<CT>(ax, xf + 1);
    a[2] = ld<CT>(ay, yf);
    a[3] = ld<CT>(ay, yf + nx);
    a[4] = ld<CT>(az, q.c);
    a[5] = ld<CT>(az, q.c + nx * ny);
}

// sum over existing neighbours of a_f * x_nb (a missing neighbour is the
// Dirichlet boundary, value 0)
template <class XT>
__device__ __forceinline__ float nb_sum(const void* x, const Cell& q, int nx, int ny, int nz, const float a[6]) {
    const int sxy = nx * ny;
    float s = 0.0f;
    if (q.i > 0) s += a[0] * ld<XT>(x, q.c - 1);
    if (q.i < nx - 1) s += a[1] * ld<XT>(x, q.c + 1);
    if (q.j > 0) s += a[2] * ld<XT>(x, q.c - nx);
    if (q.j < ny - 1) s += a[3] * ld<XT>(x, q.c + nx);
    if (q.k > 0) s += a[4] * ld<XT>(x, q.c - sxy);
    if (q.k < nz - 1) s += a[5] * ld<XT>(x, q.c + sxy);
    return s;
}

__device__ __forceinline__ float diag6(const float a[6]) { return a[0] + a[1] + a[2] + a[3] + a[4] + a[5]; }

#define PHG_SYS_PROLOGUE                  \
    const int2 e = tab[blockIdx.x];       \
    const DSys& s = sys[e.x];             \
    if (s.state[ST_ACTIVE] == 0) return;

// ---------------------------------------------------------------------------
// outer CG kernels (FP32 vectors, FP64 reductions)
// ---------------------------------------------------------------------------
// r = b - A x on gas cells (0 elsewhere); partials |b|^2 (slot 0), |r|^2 (1), max|r| (2)
__global__ void __launch_bounds__(TB) k_resid_init(const DSys* __restrict__ sys, const int2* __restrict__ tab) {
    PHG_SYS_PROLOGUE
    const DLevel& L = s.L[0];
    const Cell q = tile_cell(e.y, L.nx, L.ny);
    double v[2] = {0.0, 0.0};
    float m = 0.0f;
    if (q.c < L.n) {
        float a[6];
        coefs6<float>(L.ax32, L.ay32, L.az32, q, L.nx, L.ny, a);
        const float dg = diag6(a);
        const float b = s.b[q.c];
        float r = 0.0f;
        if (s.state[ST_ZEROX]) {
            s.x[q.c] = 0.0f;
            if (dg > 0) r = b;
        } else if (dg > 0) {
            r = b - (dg * s.x[q.c] - nb_sum<float>(s.x, q, L.nx, L.ny, L.nz, a));
        }
        s.r[q.c] = r;
        v[0] = (double)b * b;
        v[1] = (double)r * r;
        m = fabsf(r);
    }
    double* out = s.part + (long)e.y * NPART;
    block_sums<2>(v, out);
    m = block_max(m);
    if (threadIdx.x == 0) out[2] = m;
}

// q = A d and the partial d.q (slot 0)  (apply, gasphase.cpp:551-570)
__global__ void __launch_bounds__(TB) k_apply(const DSys* __restrict__ sys, const int2* __restrict__ tab) {
    PHG_SYS_PROLOGUE
    const DLevel& L = s.L[0];
    const Cell q = tile_cell(e.y, L.nx, L.ny);
    double v[1] = {0.0};
    if (q.c < L.n) {
        float a[6];
        coefs6<float>(L.ax32, L.ay32, L.az32, q, L.nx, L.ny, a);
        const float dg = diag6(a);
        float y = 0.0f;
        const float dc = s.d[q.c];
        if (dg > 0) y = dg * dc - nb_sum<float>(s.d, q, L.nx, L.ny, L.nz, a);
        s.q[q.c] = y;
        v[0] = (double)y * dc;
    }
    block_sums<1>(v, s.part + (long)e.y * NPART);
}

// x += alpha d, r -= alpha q (gasphase.cpp:609-615); partials |r|^2 (slot 1), max|r| (2).
// Cells with diag <= 0 hold d = q = r = 0, so they need no test.
__global__ void __launch_bounds__(TB) k_axpy(const DSys* __restrict__ sys, const int2* __restrict__ tab) {
    PHG_SYS_PROLOGUE
    const int n = s.L[0].n;
    const int c = e.y * TB + threadIdx.x;
    const double alpha = s.scal[S_ALPHA];
    double v[1] = {0.0};
    float m = 0.0f;
    if (c < n) {
        s.x[c] += (float)(alpha * s.d[c]);
        const float r = s.r[c] - (float)(alpha * s.q[c]);
        s.r[c] = r;
        v[0] = (double)r * r;
        m = fabsf(r);
    }
    double* out = s.part + (long)e.y * NPART;
    // slot 1 <- |r|^2, slot 2 <- max|r|
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

// partial r.z (slot 3)
__global__ void __launch_bounds__(TB) k_dot_rz(const DSys* __restrict__ sys, const int2* __restrict__ tab) {
    PHG_SYS_PROLOGUE
    const int c = e.y * TB + threadIdx.x;
    double v[1] = {0.0};
    if (c < s.L[0].n) v[0] = (double)s.r[c] * s.z[c];
    __shared__ double sh[TB / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    v[0] = wave_sum(v[0]);
    if (lane == 0) sh[w] = v[0];
    __syncthreads();
    if (threadIdx.x == 0) {
        double acc = 0.0;
#pragma unroll
        for (int ww = 0; ww < TB / 64; ++ww) acc += sh[ww];
        s.part[(long)e.y * NPART + 3] = acc;
    }
}

// d = z + beta d (gasphase.cpp:620-622); first = 1 copies d = z
__global__ void __launch_bounds__(TB) k_update_d(const DSys* __restrict__ sys, const int2* __restrict__ tab, int first) {
    PHG_SYS_PROLOGUE
    const int c = e.y * TB + threadIdx.x;
    if (c >= s.L[0].n) return;
    if (first) s.d[c] = s.z[c];
    else s.d[c] = s.z[c] + (float)(s.scal[S_BETA] * s.d[c]);
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

__device__ __forceinline__ bool stop_reached(const DSys& s, double rr, float maxr) {
    if (s.state[ST_RULE] == 1) return (double)(maxr / (float)s.scal[S_INV]) <= s.scal[S_TOL];
    const double bb = s.scal[S_BB];
    const double rel = bb > 0 ? sqrt(rr) / sqrt(bb) : 0.0;
    return rel <= s.scal[S_TOL];
}

// after k_resid_init: |b|^2, |r|^2, max|r|; stop at iteration 0 if already converged
__global__ void __launch_bounds__(TB) k_fin_init(const DSys* __restrict__ sys, const int2* __restrict__ tab) {
    PHG_SYS_PROLOGUE
    __shared__ double res[2];
    sum_parts(s, 0, &res[0]);
    __syncthreads();
    sum_parts(s, 1, &res[1]);
    const float m = max_parts(s);
    __syncthreads();
    if (threadIdx.x == 0) {
        s.scal[S_BB] = res[0];
        s.scal[S_RR] = res[1];
        s.scal[S_MAXR] = m;
        s.state[ST_ITERS] = 0;
        if (stop_reached(s, res[1], m) || s.state[ST_MAXIT] <= 0) {
            s.state[ST_STATUS] = stop_reached(s, res[1], m) ? 1 : 2;
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
        s.scal[S_RZ] = res[0];
        if (!(res[0] > 0)) {
            s.state[ST_STATUS] = 3;
            s.state[ST_ACTIVE] = 0;
        }
    }
}

// after k_apply: alpha = rz / dq, or stop on dq <= 0 (gasphase.cpp:606-608)
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
        } else {
            s.scal[S_ALPHA] = s.scal[S_RZ] / dq;
        }
    }
}

// after k_dot_rz: beta, iteration count and the stop test (gasphase.cpp:617-625)
__global__ void __launch_bounds__(TB) k_fin_beta(const DSys* __restrict__ sys, const int2* __restrict__ tab) {
    PHG_SYS_PROLOGUE
    __shared__ double res[2];
    sum_parts(s, 3, &res[0]);
    __syncthreads();
    sum_parts(s, 1, &res[1]);
    const float m = max_parts(s);
    __syncthreads();
    if (threadIdx.x == 0) {
        const double rz_new = res[0];
        s.scal[S_BETA] = rz_new / s.scal[S_RZ];
        s.scal[S_RZ] = rz_new;
        s.scal[S_RR] = res[1];
        s.scal[S_MAXR] = m;
        const int it = s.state[ST_ITERS] + 1;
        s.state[ST_ITERS] = it;
        if (stop_reached(s, res[1], m)) {
            s.state[ST_STATUS] = 1;
            s.state[ST_ACTIVE] = 0;
        } else if (it >= s.state[ST_MAXIT]) {
            s.state[ST_STATUS] = 2;
            s.state[ST_ACTIVE] = 0;
        } else if (!(rz_new > 0)) {
            s.state[ST_STATUS] = 3;
            s.state[ST_ACTIVE] = 0;
        }
    }
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
// first pre-smoothing half sweep from x = 0: colour 0 cells x = b / diag,
// colour 1 cells x = 0 (gsColor(0) on a zero iterate, pressure_mg.cpp:276,282)
template <class CT>
__global__ void __launch_bounds__(TB) k_vc_smooth0(const DSys* __restrict__ sys, const int2* __restrict__ tab) {
    PHG_SYS_PROLOGUE
    const DLevel& L = s.L[0];
    const Cell q = tile_cell(e.y, L.nx, L.ny);
    if (q.c >= L.n) return;
    float z = 0.0f;
    if (((q.i + q.j + q.k) & 1) == 0) {
        float a[6];
        coefs6<CT>(L.axc, L.ayc, L.azc, q, L.nx, L.ny, a);
        const float dg = diag6(a);
        if (dg > 0) z = s.r[q.c] / dg;
    }
    s.z[q.c] = z;
}

// red-black Gauss-Seidel half sweep on cells of one colour (gsColor, pressure_mg.cpp:195-216)
template <class CT>
__global__ void __launch_bounds__(TB) k_vc_rb(const DSys* __restrict__ sys, const int2* __restrict__ tab, int color) {
    PHG_SYS_PROLOGUE
    const DLevel& L = s.L[0];
    const Cell q = tile_cell(e.y, L.nx, L.ny);
    if (q.c >= L.n || ((q.i + q.j + q.k + color) & 1) != 0) return;
    float a[6];
    coefs6<CT>(L.axc, L.ayc, L.azc, q, L.nx, L.ny, a);
    const float dg = diag6(a);
    float z = 0.0f;
    if (dg > 0) z = (s.r[q.c] + nb_sum<float>(s.z, q, L.nx, L.ny, L.nz, a)) / dg;
    s.z[q.c] = z;
}

// level-0 residual r0 = b - A z into the V-cycle scratch (residualGrid, pressure_mg.cpp:217-235)
template <class CT, class VT>
__global__ void __launch_bounds__(TB) k_vc_resid0(const DSys* __restrict__ sys, const int2* __restrict__ tab) {
    PHG_SYS_PROLOGUE
    const DLevel& L = s.L[0];
    const Cell q = tile_cell(e.y, L.nx, L.ny);
    if (q.c >= L.n) return;
    float a[6];
    coefs6<CT>(L.axc, L.ayc, L.azc, q, L.nx, L.ny, a);
    const float dg = diag6(a);
    float r = 0.0f;
    if (dg > 0) r = s.r[q.c] - dg * s.z[q.c] + nb_sum<float>(s.z, q, L.nx, L.ny, L.nz, a);
    st<VT>(L.r, q.c, r);
}

// prolongation into level 0 and nothing else: z += t1[agg] (pressure_mg.cpp:293-295)
template <class VT>
__global__ void __launch_bounds__(TB) k_vc_prolong0(const DSys* __restrict__ sys, const int2* __restrict__ tab) {
    PHG_SYS_PROLOGUE
    const DLevel& L = s.L[0];
    const int c = e.y * TB + threadIdx.x;
    if (c >= L.n) return;
    const int a = L.agg[c];
    if (a >= 0) s.z[c] += ld<VT>(s.L[1].t, a);
}

// ---------------------------------------------------------------------------
// V-cycle, coarse levels (ELL operators)
// ---------------------------------------------------------------------------
// gather restriction: b_{l+1}[I] = sum over children of r_l, ascending
// (the serial sum of pressure_mg.cpp:285-286 in the same order)
template <class VT>
__global__ void __launch_bounds__(TB) k_vc_restrict(const DSys* __restrict__ sys, const int2* __restrict__ tab, int l) {
    PHG_SYS_PROLOGUE
    if (l + 1 >= s.nl) return;
    const DLevel& F = s.L[l];
    const DLevel& C = s.L[l + 1];
    const int I = e.y * TB + threadIdx.x;
    if (I >= C.n) return;
    float acc = 0.0f;
    for (int p = F.cptr[I]; p < F.cptr[I + 1]; ++p) acc += ld<VT>(F.r, F.cidx[p]);
    st<VT>(C.b, I, acc);
}

// first Jacobi sweep from zero: x = omega b / diag (jacobi, pressure_mg.cpp:237-247)
template <class CT, class VT>
__global__ void __launch_bounds__(TB) k_vc_jac0(const DSys* __restrict__ sys, const int2* __restrict__ tab, int l) {
    PHG_SYS_PROLOGUE
    if (l >= s.nl - 1) return;
    const DLevel& L = s.L[l];
    const int I = e.y * TB + threadIdx.x;
    if (I >= L.n) return;
    const float dg = ld<CT>(L.diag, I);
    float x = 0.0f;
    if (dg > 0) x = (0.7f * ld<VT>(L.b, I)) / dg;
    st<VT>(L.x, I, x);
}

// damped Jacobi sweep src -> dst, omega = 0.7
template <class CT, class VT>
__global__ void __launch_bounds__(TB) k_vc_jac(const DSys* __restrict__ sys, const int2* __restrict__ tab, int l,
                                               int src_is_x) {
    PHG_SYS_PROLOGUE
    if (l >= s.nl - 1) return;
    const DLevel& L = s.L[l];
    const int I = e.y * TB + threadIdx.x;
    if (I >= L.n) return;
    const void* src = src_is_x ? L.x : L.t;
    void* dst = src_is_x ? L.t : L.x;
    const float dg = ld<CT>(L.diag, I);
    float out = 0.0f;
    if (dg > 0) {
        float acc = ld<VT>(L.b, I);
        for (int p = 0; p < L.ellk; ++p) {
            const long o = (long)p * L.n + I;
            acc += ld<CT>(L.val, o) * ld<VT>(src, L.col[o]);
        }
        out = (1.0f - 0.7f) * ld<VT>(src, I) + (0.7f * acc) / dg;
    }
    st<VT>(dst, I, out);
}

// coarse residual of the current iterate t (residualSparse, pressure_mg.cpp:248-257)
template <class CT, class VT>
__global__ void __launch_bounds__(TB) k_vc_resid(const DSys* __restrict__ sys, const int2* __restrict__ tab, int l) {
    PHG_SYS_PROLOGUE
    if (l >= s.nl - 1) return;
    const DLevel& L = s.L[l];
    const int I = e.y * TB + threadIdx.x;
    if (I >= L.n) return;
    const float dg = ld<CT>(L.diag, I);
    float r = 0.0f;
    if (dg > 0) {
        r = ld<VT>(L.b, I) - dg * ld<VT>(L.t, I);
        for (int p = 0; p < L.ellk; ++p) {
            const long o = (long)p * L.n + I;
            r += ld<CT>(L.val, o) * ld<VT>(L.t, L.col[o]);
        }
    }
    st<VT>(L.r, I, r);
}

// t_l += t_{l+1}[agg]
template <class VT>
__global__ void __launch_bounds__(TB) k_vc_prolong(const DSys* __restrict__ sys, const int2* __restrict__ tab, int l) {
    PHG_SYS_PROLOGUE
    if (l >= s.nl - 1) return;
    const DLevel& L = s.L[l];
    const int I = e.y * TB + threadIdx.x;
    if (I >= L.n) return;
    st<VT>(L.t, I, ld<VT>(L.t, I) + ld<VT>(s.L[l + 1].t, L.agg[I]));
}

// coarsest level: t = (grounded A_c)^-1 b in FP64.  A block takes up to
// CG_SYS systems that share one grid (so each row of the inverse is read once
// for all of them) and CG_ROWS rows; each wave handles CG_ROWS/4 rows with its
// 64 lanes striding the row, so the inverse is read coalesced.
// The reference approximates this solve with 20 forward/backward Gauss-Seidel
// sweeps (pressure_mg.cpp:277-281); a serial sweep has no place on a GPU.
constexpr int CG_SYS = 8;
constexpr int CG_ROWS = 32;

template <class VT>
__global__ void __launch_bounds__(TB) k_vc_coarse(const DSys* __restrict__ sys, const int4* __restrict__ tab) {
    const int4 e = tab[blockIdx.x];   // first system, count, row tile
    const DSys& s0 = sys[e.x];
    const int n = s0.nc;
    const double* __restrict__ A = s0.cinv;
    int live = 0;
    for (int g = 0; g < e.y; ++g) live |= sys[e.x + g].state[ST_ACTIVE];
    if (!live) return;
    extern __shared__ double bsh[];   // [count][n]
    for (int g = 0; g < e.y; ++g) {
        const DSys& s = sys[e.x + g];
        const void* b = s.L[s.nl - 1].b;
        for (int j = threadIdx.x; j < n; j += TB) bsh[g * n + j] = (double)ld<VT>(b, j);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int rr = 0; rr < CG_ROWS / 4; ++rr) {
        const int i = e.z * CG_ROWS + w * (CG_ROWS / 4) + rr;
        if (i >= n) break;
        double acc[CG_SYS];
#pragma unroll
        for (int g = 0; g < CG_SYS; ++g) acc[g] = 0.0;
        for (int j = lane; j < n; j += 64) {
            const double a = A[(long)i * n + j];
#pragma unroll
            for (int g = 0; g < CG_SYS; ++g)
                if (g < e.y) acc[g] += a * bsh[g * n + j];
        }
#pragma unroll
        for (int g = 0; g < CG_SYS; ++g) {
            if (g >= e.y) break;
            const double v = wave_sum(acc[g]);
            const DSys& s = sys[e.x + g];
            if (lane == 0 && s.state[ST_ACTIVE]) st<VT>(s.L[s.nl - 1].t, i, (float)v);
        }
    }
}

}  // namespace phg
