// kernels_transport.hpp -- scalar transport of the gas step: Superbee-limited
// upwind advection of T and nine species with species diffusion and a
// neighbour-bounded result (Solver::advectScalars, src/sim/gasphase.cpp:685-784).
//
// Design: a face-centred plane march.  A block owns a TR_TX x TR_TY tile of
// (i, j) columns and a chunk of k planes, one column per thread.  For each
// plane it
//   1. stages the plane's cell data (10 fields, background pressure, eddy
//      viscosity, a cell status) with two halo cells on every side in LDS;
//   2. evaluates every x and y face of the tile once, for all ten fields:
//      the limited upwind face value, the face's upwind density (from the
//      limited face temperature), and the species-diffusion weights;
//   3. lets each thread assemble its own cell from its six faces: the four
//      x/y faces from LDS, the two z faces from a register window of its
//      column (planes k-1 .. k+2), the upper one evaluated now and carried up
//      as the next plane's lower face.
// Each face is thus evaluated once instead of once from each side.  Its
// value is the one the reference forms from either side (the limiter sees
// the same upwind, downwind and far-upwind cells in the same roles), and each
// cell sums its faces in the reference's order (-x, +x, -y, +y, -z, +z) with
// the reference's operations, without FMA contraction, so results are the
// reference's bit for bit.
//
// Bytes per cell (one pass each): 10 fields + Pb + nu + u, v, w + fx, fy, fz
// read (18 x 4 B), blocked (1 B), 10 fields written (40 B): 113 B.
#pragma once
#include "device.hpp"

namespace phg {

constexpr int TR_NF = 10;                 // T, Yo, Yf, Ys, Ya, Yp, Yc, Yw, Yco, Yai
constexpr int TR_PB = TR_NF;              // slot of the background pressure
constexpr int TR_NU = TR_NF + 1;          // slot of the eddy viscosity
constexpr int TR_NV = TR_NF + 2;          // values staged per cell
constexpr int TR_TX = 32, TR_TY = 8;      // tile of columns: one per thread
constexpr int TR_THREADS = TR_TX * TR_TY;
constexpr int TR_LX = TR_TX + 4, TR_LY = TR_TY + 4;   // plane with two halo cells on each side
constexpr int TR_NXF = (TR_TX + 1) * TR_TY;          // x faces of a tile plane
constexpr int TR_NYF = TR_TX * (TR_TY + 1);          // y faces
constexpr int TR_NHALO = 4 * TR_TY + 4 * TR_TX;      // halo cells of a plane
static_assert(TR_NHALO <= TR_THREADS, "one halo cell per thread");

// cell status: a gas cell, a solid cell, or a position outside the domain
// (whose values are the ambient ones: the transport sees ambient air there)
enum : int { TR_GAS = 0, TR_SOLID = 1, TR_OUT = 2 };

struct TransportArgs {
    int nx, ny, nz, kchunk, tiles_x, tiles_y;
    float dx, dt;
    const float* u;
    const float* v;
    const float* w;
    const float* fx;
    const float* fy;
    const float* fz;
    const uint8_t* blocked;
    const float* Pb;
    const float* nu;
    const float* in[TR_NF];
    float* out[TR_NF];
};

// ambient values outside the open boundary (gasphase.cpp:688) and P0
__device__ __forceinline__ float tr_ambient(int q) {
    switch (q) {
        case 0: return 293.15f;
        case 1: return 0.232f;
        case 6: return 0.0006f;
        case 7: return 0.0073f;
        case TR_PB: return 101325.0f;
        default: return 0.0f;
    }
}

// std::min / std::max / std::clamp semantics (the first argument wins ties,
// so signed zeros come out as the reference's)
__device__ __forceinline__ float tr_min(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float tr_max(float a, float b) { return a < b ? b : a; }

// Superbee limiter: the largest of 0, min(2r, 1), min(r, 2) (gasphase.cpp:35)
__device__ __forceinline__ float tr_superbee(float r) {
#pragma clang fp contract(off)
    const float m1 = tr_min(2.0f * r, 1.0f), m2 = tr_min(r, 2.0f);
    float x = 0.0f;
    if (x < m1) x = m1;
    if (x < m2) x = m2;
    return x;
}

// one face's record (its open flag is kept apart, as a byte, so that the
// face records of a tile plane and the plane itself fit three blocks per CU)
struct TrFace {
    float v[TR_NF];   // limited upwind face value of each field
    float flux;       // f u, positive along +axis
    float rup;        // upwind density at the face
    float dw, dr;     // f D and f rho_f D of the species diffusion
};

// One face between cell a (low side) and b (high side); aa lies beyond a and
// bb beyond b along the axis.  Cells outside the domain hold ambient values.
// Returns whether the face is open (f > 0 and neither cell solid); the
// record is written only then.
__device__ __forceinline__ bool tr_face(float fo, float vel, const float (&va)[TR_NV], const float (&vb)[TR_NV],
                                        const float (&vaa)[TR_NV], const float (&vbb)[TR_NV], int sa, int sb, int saa,
                                        int sbb, TrFace& f) {
#pragma clang fp contract(off)
    const float RS = 287.0f, T_MIN = 200.0f, SC_T = 0.5f;
    if (!(fo > 0.0f && sa != TR_SOLID && sb != TR_SOLID)) return false;
    const float flux = fo * vel;
    const bool fwd = flux > 0.0f;                     // a is upwind (with no flow, b: unused then)
    const bool far_solid = (fwd ? saa : sbb) == TR_SOLID;
    {   // upwind density from the limited face temperature (gasphase.cpp:726-742)
        const float tu = fwd ? va[0] : vb[0], tn = fwd ? vb[0] : va[0];
        const float tfar = far_solid ? tu : (fwd ? vaa[0] : vbb[0]);
        const float den = tn - tu;
        const float tf = tu + (fabsf(den) > 1e-6f ? 0.5f * tr_superbee((tu - tfar) / den) * den : 0.0f);
        const float pu = fwd ? va[TR_PB] : vb[TR_PB];
        f.rup = pu / (RS * tr_max(tf, T_MIN));
    }
#pragma unroll
    for (int q = 0; q < TR_NF; ++q) {   // limited upwind value (gasphase.cpp:754-764)
        const float fu = fwd ? va[q] : vb[q], fn = fwd ? vb[q] : va[q];
        const float ffar = far_solid ? fu : (fwd ? vaa[q] : vbb[q]);
        const float den = fn - fu;
        float face = fu;
        if (fabsf(den) > 1e-12f) face += 0.5f * tr_superbee((fu - ffar) / den) * den;
        f.v[q] = face;
    }
    // species diffusion weights (gasphase.cpp:767-773): across the boundary
    // the inside cell's viscosity and density stand for both sides
    const bool aout = sa == TR_OUT, bout = sb == TR_OUT;
    const float nua = aout ? vb[TR_NU] : va[TR_NU], nub = bout ? va[TR_NU] : vb[TR_NU];
    const float D = 0.5f * (nua + nub) / SC_T + 2.0e-5f;
    const float ra = va[TR_PB] / (RS * va[0]), rb = vb[TR_PB] / (RS * vb[0]);
    const float rf = aout ? rb : (bout ? ra : 0.5f * (ra + rb));
    f.flux = flux;
    f.dw = fo * D;
    f.dr = fo * rf * D;
    return true;
}

__device__ __forceinline__ void tr_load_cell(const TransportArgs& A, int i, int j, int k, float (&v)[TR_NV], int& st) {
    if (i < 0 || i >= A.nx || j < 0 || j >= A.ny || k < 0 || k >= A.nz) {
#pragma unroll
        for (int q = 0; q < TR_NV; ++q) v[q] = tr_ambient(q);
        st = TR_OUT;
        return;
    }
    const unsigned c = (unsigned)(i + A.nx * (j + A.ny * k));
#pragma unroll
    for (int q = 0; q < TR_NF; ++q) v[q] = A.in[q][c];
    v[TR_PB] = A.Pb[c];
    v[TR_NU] = A.nu[c];
    st = A.blocked[c] ? TR_SOLID : TR_GAS;
}

__global__ void __launch_bounds__(TR_THREADS) k_transport(TransportArgs A) {
#pragma clang fp contract(off)
    __shared__ float pl[TR_NV][TR_LY][TR_LX];
    __shared__ uint8_t pst[TR_LY][TR_LX];
    __shared__ TrFace xf[TR_NXF];
    __shared__ TrFace yf[TR_NYF];
    __shared__ uint8_t xo[TR_NXF], yo[TR_NYF];   // open flags of the face records
    const float RS = 287.0f;
    const int tx = (int)threadIdx.x % TR_TX, ty = (int)threadIdx.x / TR_TX;
    const int ntile = A.tiles_x * A.tiles_y;
    const int tile = (int)blockIdx.x % ntile, chunk = (int)blockIdx.x / ntile;
    const int i0 = (tile % A.tiles_x) * TR_TX, j0 = (tile / A.tiles_x) * TR_TY;
    const int k0 = chunk * A.kchunk, k1 = min(A.nz, k0 + A.kchunk);
    const int i = i0 + tx, j = j0 + ty;
    const bool own = i < A.nx && j < A.ny;
    const int sxy = A.nx * A.ny;
    const float s = A.dt / A.dx;
    // halo cell of this thread (if any): x halo (4 per row), then y halo (4 rows)
    int hlx = -1, hly = -1;
    {
        const int t = (int)threadIdx.x;
        if (t < 4 * TR_TY) {
            const int e = t % 4;
            hlx = e < 2 ? e : TR_TX + e;
            hly = 2 + t / 4;
        } else if (t < TR_NHALO) {
            const int e = (t - 4 * TR_TY) / TR_TX;
            hly = e < 2 ? e : TR_TY + e;
            hlx = 2 + (t - 4 * TR_TY) % TR_TX;
        }
    }
    // register window of the own column: planes k-1, k, k+1, k+2
    float W[4][TR_NV];
    int WS[4];
    TrFace zlo;   // the face below plane k
    bool zlo_o;
    {
        float P[TR_NV];
        int PS;
        tr_load_cell(A, i, j, k0 - 2, P, PS);
#pragma unroll
        for (int p = 0; p < 3; ++p) tr_load_cell(A, i, j, k0 - 1 + p, W[p], WS[p]);
        float fo = 0.0f, wv = 0.0f;
        if (own) {
            const unsigned f = (unsigned)(i + A.nx * j + sxy * k0);
            fo = A.fz[f];
            wv = A.w[f];
        }
        zlo_o = tr_face(fo, wv, W[0], W[1], P, W[2], WS[0], WS[1], PS, WS[2], zlo);
        tr_load_cell(A, i, j, k0 + 2, W[3], WS[3]);
    }
    for (int k = k0; k < k1; ++k) {
        // 1. the plane into LDS: own columns from the window, halo from memory
#pragma unroll
        for (int q = 0; q < TR_NV; ++q) pl[q][ty + 2][tx + 2] = W[1][q];
        pst[ty + 2][tx + 2] = (uint8_t)WS[1];
        if (hlx >= 0) {
            float h[TR_NV];
            int hs;
            tr_load_cell(A, i0 + hlx - 2, j0 + hly - 2, k, h, hs);
#pragma unroll
            for (int q = 0; q < TR_NV; ++q) pl[q][hly][hlx] = h[q];
            pst[hly][hlx] = (uint8_t)hs;
        }
        __syncthreads();
        // 2. every x and y face of the tile plane, once
        for (int e = (int)threadIdx.x; e < TR_NXF + TR_NYF; e += TR_THREADS) {
            const bool isx = e < TR_NXF;
            const int r = isx ? e : e - TR_NXF;
            // x face (fi, fj) lies between cells fi - 1 and fi of row fj; y face (fi, fj)
            // between rows fj - 1 and fj of column fi
            const int fi = isx ? r % (TR_TX + 1) : r % TR_TX, fj = isx ? r / (TR_TX + 1) : r / TR_TX;
            const int gi = i0 + fi, gj = j0 + fj;
            float fo = 0.0f, vel = 0.0f;
            if (isx ? (gi <= A.nx && gj < A.ny) : (gi < A.nx && gj <= A.ny)) {
                const unsigned f = isx ? (unsigned)(gi + (A.nx + 1) * (gj + A.ny * k))
                                       : (unsigned)(gi + A.nx * (gj + (A.ny + 1) * k));
                fo = isx ? A.fx[f] : A.fy[f];
                vel = isx ? A.u[f] : A.v[f];
            }
            // LDS coordinates of aa, a, b, bb along the face's axis
            const int bx = fi + 2, by = fj + 2;
            const int dx_ = isx ? 1 : 0, dy_ = isx ? 0 : 1;
            float vaa[TR_NV], va[TR_NV], vb[TR_NV], vbb[TR_NV];
#pragma unroll
            for (int q = 0; q < TR_NV; ++q) {
                vaa[q] = pl[q][by - 2 * dy_][bx - 2 * dx_];
                va[q] = pl[q][by - dy_][bx - dx_];
                vb[q] = pl[q][by][bx];
                vbb[q] = pl[q][by + dy_][bx + dx_];
            }
            TrFace& out = isx ? xf[r] : yf[r];
            const bool o = tr_face(fo, vel, va, vb, vaa, vbb, pst[by - dy_][bx - dx_], pst[by][bx],
                                   pst[by - 2 * dy_][bx - 2 * dx_], pst[by + dy_][bx + dx_], out);
            (isx ? xo[r] : yo[r]) = o ? 1 : 0;
        }
        // the z face above the own cell, from the column window
        TrFace zhi;
        bool zhi_o;
        {
            float fo = 0.0f, wv = 0.0f;
            if (own) {
                const unsigned f = (unsigned)(i + A.nx * j + sxy * (k + 1));
                fo = A.fz[f];
                wv = A.w[f];
            }
            zhi_o = tr_face(fo, wv, W[1], W[2], W[0], W[3], WS[1], WS[2], WS[0], WS[3], zhi);
        }
        __syncthreads();
        // 3. the own cell from its six faces, in the reference's order
        if (own) {
            const unsigned c = (unsigned)(i + A.nx * j + sxy * k);
            if (WS[1] == TR_SOLID) {
#pragma unroll
                for (int q = 0; q < TR_NF; ++q) A.out[q][c] = W[1][q];
            } else {
                // the six faces, -x +x -y +y -z +z: four LDS records and the two
                // z faces of the column
                const TrFace& X0 = xf[ty * (TR_TX + 1) + tx];
                const TrFace& X1 = xf[ty * (TR_TX + 1) + tx + 1];
                const TrFace& Y0 = yf[ty * TR_TX + tx];
                const TrFace& Y1 = yf[(ty + 1) * TR_TX + tx];
                auto pick = [&](int d, auto get) {
                    return d == 0 ? get(X0) : d == 1 ? get(X1) : d == 2 ? get(Y0) : d == 3 ? get(Y1) : d == 4 ? get(zlo)
                                                                                                       : get(zhi);
                };
                const float rc = W[1][TR_PB] / (RS * W[1][0]);
                float mden = rc;
                float un[6], ru[6], dw[6], dr[6];
                int op[6];
#pragma unroll
                for (int d = 0; d < 6; ++d) {
                    op[d] = d == 0   ? xo[ty * (TR_TX + 1) + tx]
                            : d == 1 ? xo[ty * (TR_TX + 1) + tx + 1]
                            : d == 2 ? yo[ty * TR_TX + tx]
                            : d == 3 ? yo[(ty + 1) * TR_TX + tx]
                            : d == 4 ? (int)zlo_o
                                     : (int)zhi_o;
                    un[d] = 0.0f;
                    ru[d] = dw[d] = dr[d] = 0.0f;
                    if (!op[d]) continue;
                    const float fl = pick(d, [](const TrFace& f) { return f.flux; });
                    un[d] = (d & 1) ? fl : -fl;   // outward from the cell
                    ru[d] = pick(d, [](const TrFace& f) { return f.rup; });
                    dw[d] = pick(d, [](const TrFace& f) { return f.dw; });
                    dr[d] = pick(d, [](const TrFace& f) { return f.dr; });
                    mden -= s * ru[d] * un[d];
                }
                const bool conservative = mden > 0.25f * rc;
#pragma unroll
                for (int q = 0; q < TR_NF; ++q) {
                    const float fc = W[1][q];
                    // neighbour values: x/y from the LDS plane, z from the window
                    const float fnb[6] = {pl[q][ty + 2][tx + 1], pl[q][ty + 2][tx + 3], pl[q][ty + 1][tx + 2],
                                          pl[q][ty + 3][tx + 2], W[0][q], W[2][q]};
                    float lo = fc, hi = fc, acc = 0.0f, macc = 0.0f;
#pragma unroll
                    for (int d = 0; d < 6; ++d) {
                        if (!op[d]) continue;
                        const float fn = fnb[d];
                        lo = tr_min(lo, fn);
                        hi = tr_max(hi, fn);
                        const float fv = pick(d, [q](const TrFace& f) { return f.v[q]; });
                        const float face = un[d] != 0.0f ? fv : fc;
                        acc -= un[d] * (face - fc);
                        macc -= ru[d] * un[d] * face;
                        if (q > 0) {
                            acc += dw[d] * (fn - fc) / A.dx;
                            macc += dr[d] * (fn - fc) / A.dx;
                        }
                    }
                    const float val = conservative ? (rc * fc + s * macc) / mden : fc + s * acc;
                    A.out[q][c] = val < lo ? lo : (hi < val ? hi : val);
                }
            }
        }
        // advance the window and carry the face; the LDS plane and face
        // records are rewritten only after every thread has read them
        zlo = zhi;
        zlo_o = zhi_o;
#pragma unroll
        for (int p = 0; p < 3; ++p) {
#pragma unroll
            for (int q = 0; q < TR_NV; ++q) W[p][q] = W[p + 1][q];
            WS[p] = WS[p + 1];
        }
        if (k + 1 < k1) tr_load_cell(A, i, j, k + 3, W[3], WS[3]);
        __syncthreads();
    }
}

}  // namespace phg
