// hierarchy_dev.hip -- levels 0 and 1 of the pressure multigrid formed on the
// device, for rebuilds after face changes (the reference's mgDirty_,
// src/sim/pressure_mg.cpp:28-136).  The results equal build_hierarchy's bit
// for bit:
//   level 0   diagonal (the six coefficients summed left to right), the
//             open-boundary flags;
//   level 1   connected pieces of 2x2x2 cell blocks (one thread per block:
//             closure of the at most eight cells over open internal faces),
//             numbered in the order of their lowest cell (a prefix sum over
//             "lowest cell of its piece" flags), as the reference's serial
//             numbering does;
//   Galerkin  each coupling between two pieces is an event (cell, axis) in
//             the reference's loop order; its two row entries are sorted
//             stably by (row, column), so repeated couplings are summed in
//             FP32 in event order, and each row's entries then stably by
//             their first event, the reference's first-insertion order; each
//             piece's diagonal sums its cells' diagonals, then removes its
//             internal faces twice, in FP64 in the reference's order.
// The coarser levels (at most a seventh of the cells) continue on the host
// from the downloaded level 1 with the reference's own code.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdio>
#include <vector>

#include "hierarchy.hpp"

namespace phg {
namespace {

constexpr int BT = 256;
constexpr int ROW_BITS = 21;   // level-1 rows and columns below 2^21
constexpr int EV_BITS = 26;    // events (3 per cell) below 2^26

struct Dims {
    int nx, ny, nz;
    __device__ int xf(int i, int j, int k) const { return i + (nx + 1) * (j + ny * k); }
    __device__ int yf(int i, int j, int k) const { return i + nx * (j + (ny + 1) * k); }
    __device__ int zf(int i, int j, int k) const { return i + nx * (j + ny * k); }
};

// diagonal and open-boundary flag of every cell (pressure_mg.cpp:47-57)
__global__ void __launch_bounds__(BT) k_b_cells(Dims D, const float* __restrict__ ax, const float* __restrict__ ay,
                                                const float* __restrict__ az, const uint8_t* __restrict__ blocked,
                                                float* __restrict__ diag, uint8_t* __restrict__ leak) {
#pragma clang fp contract(off)
    const int n = D.nx * D.ny * D.nz;
    const int c = blockIdx.x * BT + threadIdx.x;
    if (c >= n) return;
    const int i = c % D.nx, j = (c / D.nx) % D.ny, k = c / (D.nx * D.ny);
    if (blocked[c]) {
        diag[c] = 0.0f;
        leak[c] = 0;
        return;
    }
    const float a0 = ax[D.xf(i, j, k)], a1 = ax[D.xf(i + 1, j, k)], a2 = ay[D.yf(i, j, k)], a3 = ay[D.yf(i, j + 1, k)],
                a4 = az[D.zf(i, j, k)], a5 = az[D.zf(i, j, k + 1)];
    const float d = a0 + a1 + a2 + a3 + a4 + a5;
    diag[c] = d;
    const bool bnd = (i == 0 && a0 > 0) || (i == D.nx - 1 && a1 > 0) || (j == 0 && a2 > 0) ||
                     (j == D.ny - 1 && a3 > 0) || (k == 0 && a4 > 0) || (k == D.nz - 1 && a5 > 0);
    leak[c] = (d > 0 && bnd) ? 1 : 0;
}

// the cells of a 2x2x2 block, local index l = dx + 2 dy + 4 dz (ascending
// cell index); -1 outside the grid
__device__ __forceinline__ void block_cells(const Dims& D, int i0, int j0, int k0, int (&cell)[8]) {
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        const int i = i0 + (l & 1), j = j0 + ((l >> 1) & 1), k = k0 + (l >> 2);
        cell[l] = (i < D.nx && j < D.ny && k < D.nz) ? i + D.nx * (j + D.ny * k) : -1;
    }
}

// one thread per 2x2x2 block: each cell's piece, as its lowest cell (-1 for
// cells with diag <= 0), joined over internal faces with a > 0 between cells
// with diag > 0 (pressure_mg.cpp:61-72)
__global__ void __launch_bounds__(BT) k_b_roots(Dims D, const float* __restrict__ ax, const float* __restrict__ ay,
                                                const float* __restrict__ az, const float* __restrict__ diag,
                                                int* __restrict__ root, int* __restrict__ first) {
    const int bx = (D.nx + 1) / 2, by = (D.ny + 1) / 2, bz = (D.nz + 1) / 2;
    const int b = blockIdx.x * BT + threadIdx.x;
    if (b >= bx * by * bz) return;
    const int i0 = 2 * (b % bx), j0 = 2 * ((b / bx) % by), k0 = 2 * (b / (bx * by));
    int cell[8];
    block_cells(D, i0, j0, k0, cell);
    unsigned live = 0, adj[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        adj[l] = 0;
        if (cell[l] >= 0 && diag[cell[l]] > 0) live |= 1u << l;
    }
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        if (!(live >> l & 1)) continue;
        const int i = i0 + (l & 1), j = j0 + ((l >> 1) & 1), k = k0 + (l >> 2);
        if (!(l & 1) && (live >> (l + 1) & 1) && ax[D.xf(i + 1, j, k)] > 0) {
            adj[l] |= 1u << (l + 1);
            adj[l + 1] |= 1u << l;
        }
        if (!(l & 2) && (live >> (l + 2) & 1) && ay[D.yf(i, j + 1, k)] > 0) {
            adj[l] |= 1u << (l + 2);
            adj[l + 2] |= 1u << l;
        }
        if (!(l & 4) && (live >> (l + 4) & 1) && az[D.zf(i, j, k + 1)] > 0) {
            adj[l] |= 1u << (l + 4);
            adj[l + 4] |= 1u << l;
        }
    }
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        if (cell[l] < 0) continue;
        if (!(live >> l & 1)) {
            root[cell[l]] = -1;
            first[cell[l]] = 0;
            continue;
        }
        unsigned reach = 1u << l;
        for (int it = 0; it < 7; ++it) {
            unsigned nxt = reach;
#pragma unroll
            for (int m = 0; m < 8; ++m)
                if (reach >> m & 1) nxt |= adj[m];
            reach = nxt;
        }
        const int low = __ffs(reach) - 1;   // the piece's lowest cell
        root[cell[l]] = cell[low];
        first[cell[l]] = low == l ? 1 : 0;
    }
}

// aggregate numbers from the prefix sum over the lowest-cell flags; the
// block coordinates and lowest cell of each piece
__global__ void __launch_bounds__(BT) k_b_agg(Dims D, const int* __restrict__ root, const int* __restrict__ pos,
                                              int* __restrict__ agg, int* __restrict__ afirst, int* __restrict__ bi,
                                              int* __restrict__ bj, int* __restrict__ bk) {
    const int n = D.nx * D.ny * D.nz;
    const int c = blockIdx.x * BT + threadIdx.x;
    if (c >= n) return;
    const int r = root[c];
    agg[c] = r >= 0 ? pos[r] : -1;
    if (r == c) {
        const int A = pos[c];
        afirst[A] = c;
        bi[A] = (c % D.nx) >> 1;
        bj[A] = ((c / D.nx) % D.ny) >> 1;
        bk[A] = (c / (D.nx * D.ny)) >> 1;
    }
}

// per piece: its cells in ascending order (count, or the child list), the
// FP64 diagonal (pressure_mg.cpp:111-126) and the open-boundary flag
__global__ void __launch_bounds__(BT) k_b_pieces(Dims D, int n1, const int* __restrict__ afirst,
                                                 const int* __restrict__ agg, const float* __restrict__ ax,
                                                 const float* __restrict__ ay, const float* __restrict__ az,
                                                 const float* __restrict__ diag, const uint8_t* __restrict__ leak0,
                                                 int* __restrict__ cnt, const int* __restrict__ cptr,
                                                 int* __restrict__ cidx, float* __restrict__ diag1,
                                                 uint8_t* __restrict__ leak1) {
    const int A = blockIdx.x * BT + threadIdx.x;
    if (A >= n1) return;
    const int c0 = afirst[A];
    const int sxy = D.nx * D.ny;
    const int i0 = (c0 % D.nx) & ~1, j0 = ((c0 / D.nx) % D.ny) & ~1, k0 = (c0 / sxy) & ~1;
    int cell[8];
    block_cells(D, i0, j0, k0, cell);
    if (cptr == nullptr) {   // first pass: the piece's size
        int m = 0;
        for (int l = 0; l < 8; ++l)
            if (cell[l] >= 0 && agg[cell[l]] == A) ++m;
        cnt[A] = m;
        return;
    }
    double dg = 0.0;
    uint8_t lk = 0;
    int m = cptr[A];
    for (int l = 0; l < 8; ++l) {
        const int c = cell[l];
        if (c < 0 || agg[c] != A) continue;
        cidx[m++] = c;
        dg += diag[c];
        lk |= leak0[c];
    }
    for (int l = 0; l < 8; ++l) {   // internal faces, from both cells
        const int c = cell[l];
        if (c < 0 || agg[c] != A) continue;
        const int i = c % D.nx, j = (c / D.nx) % D.ny, k = c / sxy;
        const int nb[3] = {i + 1 < D.nx ? c + 1 : -1, j + 1 < D.ny ? c + D.nx : -1, k + 1 < D.nz ? c + sxy : -1};
        const float w[3] = {i + 1 < D.nx ? ax[D.xf(i + 1, j, k)] : 0.0f, j + 1 < D.ny ? ay[D.yf(i, j + 1, k)] : 0.0f,
                            k + 1 < D.nz ? az[D.zf(i, j, k + 1)] : 0.0f};
        for (int d = 0; d < 3; ++d)
            if (nb[d] >= 0 && diag[nb[d]] > 0 && w[d] > 0 && agg[nb[d]] == A) dg -= 2.0 * w[d];
    }
    diag1[A] = (float)(dg < 0.0 ? 0.0 : dg);   // std::max(dg, 0.0)
    leak1[A] = lk;
}

// couplings between pieces: event e = 3 c + axis in the reference's loop order
// (pressure_mg.cpp:98-108); flags first, then the two row entries of each
__global__ void __launch_bounds__(BT) k_b_events(Dims D, const int* __restrict__ agg, const float* __restrict__ ax,
                                                 const float* __restrict__ ay, const float* __restrict__ az,
                                                 const float* __restrict__ diag, const int* __restrict__ epos,
                                                 int* __restrict__ flag, unsigned long long* __restrict__ key,
                                                 unsigned long long* __restrict__ val) {
    const int n = D.nx * D.ny * D.nz, sxy = D.nx * D.ny;
    const int e = blockIdx.x * BT + threadIdx.x;
    if (e >= 3 * n) return;
    const int c = e / 3, d = e - 3 * c;
    const int i = c % D.nx, j = (c / D.nx) % D.ny, k = c / sxy;
    int nb = -1;
    float w = 0.0f;
    if (d == 0 && i + 1 < D.nx) nb = c + 1, w = ax[D.xf(i + 1, j, k)];
    if (d == 1 && j + 1 < D.ny) nb = c + D.nx, w = ay[D.yf(i, j + 1, k)];
    if (d == 2 && k + 1 < D.nz) nb = c + sxy, w = az[D.zf(i, j, k + 1)];
    bool ok = diag[c] > 0 && nb >= 0 && w > 0;
    int A = -1, B = -1;
    if (ok) {
        A = agg[c];
        B = agg[nb];
        ok = A >= 0 && B >= 0 && A != B;
    }
    if (epos == nullptr) {
        flag[e] = ok ? 1 : 0;
        return;
    }
    if (!ok) return;
    const int p = 2 * epos[e];
    const unsigned long long payload = ((unsigned long long)(unsigned)e << 32) | (unsigned long long)__float_as_uint(w);
    key[p] = ((unsigned long long)A << ROW_BITS) | (unsigned long long)B;
    val[p] = payload;
    key[p + 1] = ((unsigned long long)B << ROW_BITS) | (unsigned long long)A;
    val[p + 1] = payload;
}

// one (row, column) run of the key-sorted entries: its FP32 sum in event
// order and its first event, keyed for the row order
__global__ void __launch_bounds__(BT) k_b_runs(int m, const unsigned long long* __restrict__ key,
                                               const unsigned long long* __restrict__ val, const int* __restrict__ rpos,
                                               unsigned long long* __restrict__ key2,
                                               unsigned long long* __restrict__ val2, int* __restrict__ rowcnt) {
#pragma clang fp contract(off)
    const int p = blockIdx.x * BT + threadIdx.x;
    if (p >= m || (p > 0 && key[p] == key[p - 1])) return;
    const unsigned long long kk = key[p];
    const unsigned e0 = (unsigned)(val[p] >> 32);
    float v = __uint_as_float((unsigned)val[p]);
    for (int q = p + 1; q < m && key[q] == kk; ++q) v = v + __uint_as_float((unsigned)val[q]);
    const unsigned row = (unsigned)(kk >> ROW_BITS), col = (unsigned)(kk & ((1ull << ROW_BITS) - 1));
    const int r = rpos[p];
    key2[r] = ((unsigned long long)row << EV_BITS) | e0;
    val2[r] = ((unsigned long long)col << 32) | __float_as_uint(v);
    atomicAdd(&rowcnt[row], 1);
}

__global__ void __launch_bounds__(BT) k_b_run_flags(int m, const unsigned long long* __restrict__ key,
                                                    int* __restrict__ f) {
    const int p = blockIdx.x * BT + threadIdx.x;
    if (p < m) f[p] = (p == 0 || key[p] != key[p - 1]) ? 1 : 0;
}

__global__ void __launch_bounds__(BT) k_b_csr(int nnz, const unsigned long long* __restrict__ val2,
                                              int* __restrict__ col, float* __restrict__ v) {
    const int p = blockIdx.x * BT + threadIdx.x;
    if (p >= nnz) return;
    col[p] = (int)(val2[p] >> 32);
    v[p] = __uint_as_float((unsigned)val2[p]);
}

// device buffers of one build, freed on every path out
struct Pool {
    std::vector<void*> p;
    ~Pool() {
        for (void* q : p) (void)hipFree(q);
    }
    template <class T>
    T* get(size_t n) {
        void* q = nullptr;
        if (hipMalloc(&q, (n ? n : 1) * sizeof(T)) != hipSuccess) return nullptr;
        p.push_back(q);
        return static_cast<T*>(q);
    }
};

#define B_OK(call)                                                               \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess) return std::string(#call ": ") + hipGetErrorString(e_); \
    } while (0)

int blocks(long n) { return (int)((n + BT - 1) / BT); }

// exclusive prefix sum of n ints into out; returns the total
std::string scan(hipStream_t s, Pool& pool, const int* in, int* out, int n, int* total) {
    size_t tb = 0;
    B_OK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, n, s));
    void* tmp = pool.get<char>(tb);
    if (!tmp) return "out of device memory";
    B_OK(hipcub::DeviceScan::ExclusiveSum(tmp, tb, in, out, n, s));
    int a = 0, b = 0;
    B_OK(hipMemcpyAsync(&a, out + n - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    B_OK(hipMemcpyAsync(&b, in + n - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    B_OK(hipStreamSynchronize(s));
    *total = a + b;
    return "";
}

// stable sort of (key, val) pairs on the key's low `bits` bits
std::string sort_pairs(hipStream_t s, Pool& pool, unsigned long long* k_in, unsigned long long* v_in,
                       unsigned long long* k_out, unsigned long long* v_out, int n, int bits) {
    size_t tb = 0;
    B_OK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k_in, k_out, v_in, v_out, n, 0, bits, s));
    void* tmp = pool.get<char>(tb);
    if (!tmp) return "out of device memory";
    B_OK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, k_in, k_out, v_in, v_out, n, 0, bits, s));
    return "";
}

template <class T>
std::string down(hipStream_t s, std::vector<T>& v, const T* d, size_t n) {
    v.resize(n);
    if (n) B_OK(hipMemcpyAsync(v.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, s));
    return "";
}

}  // namespace

std::string device_levels01(void* stream, int nx, int ny, int nz, const float* d_ax, const float* d_ay,
                            const float* d_az, const uint8_t* d_blocked, const float* ax, const float* ay,
                            const float* az, HostHierarchy& H) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long n = (long)nx * ny * nz;
    if (3 * n >= (1L << EV_BITS)) return "grid too large for the device build's event keys";
    const Dims D{nx, ny, nz};
    Pool pool;
    float* diag = pool.get<float>(n);
    uint8_t* leak = pool.get<uint8_t>(n);
    int* root = pool.get<int>(n);
    int* first = pool.get<int>(n);
    int* pos = pool.get<int>(n);
    int* agg = pool.get<int>(n);
    if (!diag || !leak || !root || !first || !pos || !agg) return "out of device memory";
    hipLaunchKernelGGL(k_b_cells, dim3(blocks(n)), dim3(BT), 0, s, D, d_ax, d_ay, d_az, d_blocked, diag, leak);
    const long nb = (long)((nx + 1) / 2) * ((ny + 1) / 2) * ((nz + 1) / 2);
    hipLaunchKernelGGL(k_b_roots, dim3(blocks(nb)), dim3(BT), 0, s, D, d_ax, d_ay, d_az, (const float*)diag, root,
                       first);
    B_OK(hipGetLastError());
    int n1 = 0;
    std::string err = scan(s, pool, first, pos, (int)n, &n1);
    if (!err.empty()) return err;
    if (n1 >= (1 << ROW_BITS)) return "too many level-1 unknowns for the device build's keys";
    int* afirst = pool.get<int>(n1);
    int* bi = pool.get<int>(n1);
    int* bj = pool.get<int>(n1);
    int* bk = pool.get<int>(n1);
    int* cnt = pool.get<int>(n1);
    int* cptr = pool.get<int>(n1 + 1);
    float* diag1 = pool.get<float>(n1);
    uint8_t* leak1 = pool.get<uint8_t>(n1);
    if (!afirst || !bi || !bj || !bk || !cnt || !cptr || !diag1 || !leak1) return "out of device memory";
    hipLaunchKernelGGL(k_b_agg, dim3(blocks(n)), dim3(BT), 0, s, D, (const int*)root, (const int*)pos, agg, afirst, bi,
                       bj, bk);
    int members = 0;
    int* cidx = nullptr;
    if (n1 > 0) {
        hipLaunchKernelGGL(k_b_pieces, dim3(blocks(n1)), dim3(BT), 0, s, D, n1, (const int*)afirst, (const int*)agg,
                           d_ax, d_ay, d_az, (const float*)diag, (const uint8_t*)leak, cnt, (const int*)nullptr,
                           (int*)nullptr, diag1, leak1);
        err = scan(s, pool, cnt, cptr, n1, &members);
        if (!err.empty()) return err;
        B_OK(hipMemcpyAsync(cptr + n1, &members, sizeof(int), hipMemcpyHostToDevice, s));
        cidx = pool.get<int>(members);
        if (!cidx) return "out of device memory";
        hipLaunchKernelGGL(k_b_pieces, dim3(blocks(n1)), dim3(BT), 0, s, D, n1, (const int*)afirst, (const int*)agg,
                           d_ax, d_ay, d_az, (const float*)diag, (const uint8_t*)leak, cnt, (const int*)cptr, cidx,
                           diag1, leak1);
    } else {
        B_OK(hipMemsetAsync(cptr, 0, sizeof(int), s));
    }
    // the couplings between pieces, then the level-1 rows
    int* flag = pool.get<int>(3 * n);
    int* epos = pool.get<int>(3 * n);
    if (!flag || !epos) return "out of device memory";
    hipLaunchKernelGGL(k_b_events, dim3(blocks(3 * n)), dim3(BT), 0, s, D, (const int*)agg, d_ax, d_ay, d_az,
                       (const float*)diag, (const int*)nullptr, flag, (unsigned long long*)nullptr,
                       (unsigned long long*)nullptr);
    int nev = 0;
    err = scan(s, pool, flag, epos, (int)(3 * n), &nev);
    if (!err.empty()) return err;
    const int m = 2 * nev;
    std::vector<int> rowStart(n1 + 1, 0), col;
    std::vector<float> val;
    if (m > 0) {
        auto* key = pool.get<unsigned long long>(m);
        auto* v = pool.get<unsigned long long>(m);
        auto* key_s = pool.get<unsigned long long>(m);
        auto* v_s = pool.get<unsigned long long>(m);
        int* rf = pool.get<int>(m);
        int* rpos = pool.get<int>(m);
        int* rowcnt = pool.get<int>(n1 + 1);
        if (!key || !v || !key_s || !v_s || !rf || !rpos || !rowcnt) return "out of device memory";
        hipLaunchKernelGGL(k_b_events, dim3(blocks(3 * n)), dim3(BT), 0, s, D, (const int*)agg, d_ax, d_ay, d_az,
                           (const float*)diag, (const int*)epos, flag, key, v);
        err = sort_pairs(s, pool, key, v, key_s, v_s, m, 2 * ROW_BITS);
        if (!err.empty()) return err;
        hipLaunchKernelGGL(k_b_run_flags, dim3(blocks(m)), dim3(BT), 0, s, m, (const unsigned long long*)key_s, rf);
        int nnz = 0;
        err = scan(s, pool, rf, rpos, m, &nnz);
        if (!err.empty()) return err;
        auto* key2 = pool.get<unsigned long long>(nnz);
        auto* v2 = pool.get<unsigned long long>(nnz);
        auto* key2_s = pool.get<unsigned long long>(nnz);
        auto* v2_s = pool.get<unsigned long long>(nnz);
        int* dcol = pool.get<int>(nnz);
        float* dval = pool.get<float>(nnz);
        if (!key2 || !v2 || !key2_s || !v2_s || !dcol || !dval) return "out of device memory";
        B_OK(hipMemsetAsync(rowcnt, 0, (n1 + 1) * sizeof(int), s));
        hipLaunchKernelGGL(k_b_runs, dim3(blocks(m)), dim3(BT), 0, s, m, (const unsigned long long*)key_s,
                           (const unsigned long long*)v_s, (const int*)rpos, key2, v2, rowcnt);
        err = sort_pairs(s, pool, key2, v2, key2_s, v2_s, nnz, ROW_BITS + EV_BITS);
        if (!err.empty()) return err;
        hipLaunchKernelGGL(k_b_csr, dim3(blocks(nnz)), dim3(BT), 0, s, nnz, (const unsigned long long*)v2_s, dcol,
                           dval);
        B_OK(hipGetLastError());
        std::vector<int> rc;
        err = down(s, rc, (const int*)rowcnt, (size_t)n1);
        if (err.empty()) err = down(s, col, (const int*)dcol, (size_t)nnz);
        if (err.empty()) err = down(s, val, (const float*)dval, (size_t)nnz);
        if (!err.empty()) return err;
        B_OK(hipStreamSynchronize(s));
        for (int I = 0; I < n1; ++I) rowStart[I + 1] = rowStart[I] + rc[I];
    }
    // levels 0 and 1 into the host hierarchy
    H = HostHierarchy();
    H.L.resize(2);
    HostLevel& L0 = H.L[0];
    L0.n = (int)n;
    L0.nx = nx;
    L0.ny = ny;
    L0.nz = nz;
    L0.ax.assign(ax, ax + (size_t)(nx + 1) * ny * nz);
    L0.ay.assign(ay, ay + (size_t)nx * (ny + 1) * nz);
    L0.az.assign(az, az + (size_t)nx * ny * (nz + 1));
    HostLevel& C = H.L[1];
    C.n = n1;
    err = down(s, L0.diag, (const float*)diag, (size_t)n);
    if (err.empty()) err = down(s, L0.leak, (const uint8_t*)leak, (size_t)n);
    if (err.empty()) err = down(s, L0.agg, (const int*)agg, (size_t)n);
    if (err.empty()) err = down(s, L0.cptr, (const int*)cptr, (size_t)n1 + 1);
    if (err.empty() && cidx) err = down(s, L0.cidx, (const int*)cidx, (size_t)members);
    if (err.empty()) err = down(s, C.diag, (const float*)diag1, (size_t)n1);
    if (err.empty()) err = down(s, C.leak, (const uint8_t*)leak1, (size_t)n1);
    if (err.empty()) err = down(s, C.bi, (const int*)bi, (size_t)n1);
    if (err.empty()) err = down(s, C.bj, (const int*)bj, (size_t)n1);
    if (err.empty()) err = down(s, C.bk, (const int*)bk, (size_t)n1);
    if (!err.empty()) return err;
    B_OK(hipStreamSynchronize(s));
    C.rowStart = std::move(rowStart);
    C.col = std::move(col);
    C.val = std::move(val);
    return "";
}

}  // namespace phg
