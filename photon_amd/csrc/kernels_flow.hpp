// kernels_flow.hpp -- the rest of the gas step on the device: eddy viscosity
// (Solver::turbulence, src/sim/gasphase.cpp:163-218), heat conduction
// (heatDiffusion, :221-248), momentum (:275-399), burner gas
// (injectBurners / addGasSpecies, :92-105, src/sim/solver.cpp:213-224),
// combustion (solver.cpp:284-344), the start-of-step clears and the zone
// background pressure (solver.cpp:231-234,258-263).
//
// Every float expression keeps the reference's operation order and is
// compiled without FMA contraction; the host computes the constants that
// come from the C library (exp of the mixing time, the wind profile), so
// results match the reference's except where a device pow (molecular
// viscosity and conductivity) rounds differently from the host's.
#pragma once
#include "device.hpp"

namespace phg {

namespace gasc {   // constants of the gas model (solver.h:159-181, gasphase.cpp:25-33)
constexpr float RS = 287.0f, T0 = 293.15f, P0 = 101325.0f, CP = 1100.0f, GRAV = 9.81f;
constexpr float VMAX = 40.0f, PR_T = 0.5f, C_NU = 0.1f;
constexpr float DHC_FUEL = 25.0e6f, S_O2 = 1.95f, Y_H2O = 0.761f, C_FRAC = 0.597f, CO_YIELD = 0.02f;
constexpr float T_CRIT = 1600.0f, T_AUTO = 800.0f;
}  // namespace gasc

__device__ __forceinline__ float fl_min(float a, float b) { return b < a ? b : a; }   // std::min
__device__ __forceinline__ float fl_max(float a, float b) { return a < b ? b : a; }   // std::max
__device__ __forceinline__ float fl_clamp(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }

__device__ __forceinline__ float nu_mol(float T) {   // gasphase.cpp:37
#pragma clang fp contract(off)
    return 1.5e-5f * powf(T / 293.15f, 1.7f);
}
__device__ __forceinline__ float k_mol(float T) {    // gasphase.cpp:38
#pragma clang fp contract(off)
    return 0.0258f * powf(T / 293.15f, 0.8f);
}

// ---------------------------------------------------------------------------
// Eddy viscosity: a plane march over a TU_TX x TU_TY tile of columns.  For
// each plane the cell-centred velocity of the tile and a one-cell halo is
// formed in LDS from the face velocities, filtered along x (the tile's rows
// and the halo rows) and along y (the tile); each column keeps the xy-filtered
// velocity of planes k-1, k, k+1 in registers for the z pass.  Solid cells
// and the domain edge mirror the cell (gasphase.cpp:183-199).
// Bytes per cell: u, v, w (one face each), T, blocked read, nu written: 21 B.
// ---------------------------------------------------------------------------
constexpr int TU_TX = 32, TU_TY = 8, TU_THREADS = TU_TX * TU_TY;
enum : int { TU_GAS = 0, TU_SOLID = 1, TU_OUT = 2 };

struct TurbArgs {
    int nx, ny, nz, kchunk, tiles_x, tiles_y;
    float dx;
    const float* u;
    const float* v;
    const float* w;
    const float* T;
    const uint8_t* blocked;
    float* nu;
    float* max_nu;
};

struct TurbPlane {   // one column's values on one plane
    float g[3];      // xy-filtered cell velocity
    float uc[3];     // cell velocity
    int st;          // TU_GAS / TU_SOLID / TU_OUT
};

__device__ __forceinline__ TurbPlane turb_plane(const TurbArgs& A, int i0, int j0, int kp, float (*ucs)[TU_TY + 2][TU_TX + 2],
                                                int (*sts)[TU_TX + 2], float (*fxs)[TU_TY + 2][TU_TX]) {
#pragma clang fp contract(off)
    const int tx = (int)threadIdx.x % TU_TX, ty = (int)threadIdx.x / TU_TX;
    TurbPlane out;
    __syncthreads();   // the previous plane's LDS has been read
    // cell velocities of the tile and its one-cell halo (domain edge: outside)
    for (int e = (int)threadIdx.x; e < (TU_TY + 2) * (TU_TX + 2); e += TU_THREADS) {
        const int ly = e / (TU_TX + 2), lx = e % (TU_TX + 2);
        const int i = i0 + lx - 1, j = j0 + ly - 1;
        int st = TU_OUT;
        float a = 0.0f, b = 0.0f, c3 = 0.0f;
        if (kp >= 0 && kp < A.nz && i >= 0 && i < A.nx && j >= 0 && j < A.ny) {
            const unsigned c = (unsigned)(i + A.nx * (j + A.ny * kp));
            st = A.blocked[c] ? TU_SOLID : TU_GAS;
            if (st == TU_GAS) {
                const unsigned xf = (unsigned)(i + (A.nx + 1) * (j + A.ny * kp));
                const unsigned yf = (unsigned)(i + A.nx * (j + (A.ny + 1) * kp));
                a = 0.5f * (A.u[xf] + A.u[xf + 1]);
                b = 0.5f * (A.v[yf] + A.v[yf + A.nx]);
                c3 = 0.5f * (A.w[c] + A.w[c + (unsigned)(A.nx * A.ny)]);
            }
        }
        ucs[0][ly][lx] = a;
        ucs[1][ly][lx] = b;
        ucs[2][ly][lx] = c3;
        sts[ly][lx] = st;
    }
    __syncthreads();
    // x pass over the tile's columns on every staged row
    for (int e = (int)threadIdx.x; e < (TU_TY + 2) * TU_TX; e += TU_THREADS) {
        const int ly = e / TU_TX, lx = e % TU_TX + 1;
        const bool gas = sts[ly][lx] == TU_GAS;
        const bool l = sts[ly][lx - 1] == TU_GAS, r = sts[ly][lx + 1] == TU_GAS;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const float in = ucs[q][ly][lx];
            const float lo = l ? ucs[q][ly][lx - 1] : in, hi = r ? ucs[q][ly][lx + 1] : in;
            fxs[q][ly][lx - 1] = gas ? 0.25f * lo + 0.5f * in + 0.25f * hi : 0.0f;
        }
    }
    __syncthreads();
    // y pass for the own column
    const int ly = ty + 1, lx = tx + 1;
    out.st = sts[ly][lx];
    const bool d = sts[ly - 1][lx] == TU_GAS, u = sts[ly + 1][lx] == TU_GAS;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        out.uc[q] = ucs[q][ly][lx];
        const float in = fxs[q][ly][tx];
        const float lo = d ? fxs[q][ly - 1][tx] : in, hi = u ? fxs[q][ly + 1][tx] : in;
        out.g[q] = out.st == TU_GAS ? 0.25f * lo + 0.5f * in + 0.25f * hi : 0.0f;
    }
    return out;
}

__global__ void __launch_bounds__(TU_THREADS) k_turbulence(TurbArgs A) {
#pragma clang fp contract(off)
    __shared__ float ucs[3][TU_TY + 2][TU_TX + 2];
    __shared__ int sts[TU_TY + 2][TU_TX + 2];
    __shared__ float fxs[3][TU_TY + 2][TU_TX];
    const int ntile = A.tiles_x * A.tiles_y;
    const int tile = (int)blockIdx.x % ntile, chunk = (int)blockIdx.x / ntile;
    const int i0 = (tile % A.tiles_x) * TU_TX, j0 = (tile / A.tiles_x) * TU_TY;
    const int k0 = chunk * A.kchunk, k1 = min(A.nz, k0 + A.kchunk);
    const int i = i0 + (int)threadIdx.x % TU_TX, j = j0 + (int)threadIdx.x / TU_TX;
    const bool own = i < A.nx && j < A.ny;
    TurbPlane lo = turb_plane(A, i0, j0, k0 - 1, ucs, sts, fxs);
    TurbPlane mid = turb_plane(A, i0, j0, k0, ucs, sts, fxs);
    float mx = 0.0f;
    for (int k = k0; k < k1; ++k) {
        const TurbPlane hi = turb_plane(A, i0, j0, k + 1, ucs, sts, fxs);
        if (own) {
            const unsigned c = (unsigned)(i + A.nx * (j + A.ny * k));
            float nu = 0.0f;
            if (mid.st == TU_GAS) {
                float f[3];
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const float in = mid.g[q];
                    const float a = lo.st == TU_GAS ? lo.g[q] : in, b = hi.st == TU_GAS ? hi.g[q] : in;
                    f[q] = 0.25f * a + 0.5f * in + 0.25f * b;
                }
                const float du = mid.uc[0] - f[0], dv = mid.uc[1] - f[1], dw = mid.uc[2] - f[2];
                const float ks = 0.5f * (du * du + dv * dv + dw * dw);
                nu = fl_min(1.0f, gasc::C_NU * A.dx * sqrtf(ks) + nu_mol(A.T[c]));
            }
            A.nu[c] = nu;
            mx = fl_max(mx, nu);
        }
        lo = mid;
        mid = hi;
    }
    mx = block_max(mx);
    if (threadIdx.x == 0) atomic_max_nonneg(A.max_nu, mx);
}

// ---------------------------------------------------------------------------
// Heat conduction into Q, one thread per cell (gasphase.cpp:221-248):
// molecular + turbulent conductivity on each open face, ambient air outside
// open boundary faces.  Bytes per cell: T, nu, Pb, Q (rw), fx fy fz, blocked:
// 29 B (neighbour values come from cache).
// ---------------------------------------------------------------------------
struct CellGrid {
    int nx, ny, nz;
    float dx;
    const float* fx;
    const float* fy;
    const float* fz;
    const uint8_t* blocked;
};

__global__ void __launch_bounds__(TB) k_heat_diffusion(CellGrid g, float dt, const float* __restrict__ T,
                                                       const float* __restrict__ nu, const float* __restrict__ Pb,
                                                       float* __restrict__ Q) {
#pragma clang fp contract(off)
    using namespace gasc;
    const int n = g.nx * g.ny * g.nz;
    const int c = blockIdx.x * TB + threadIdx.x;
    if (c >= n || g.blocked[c]) return;
    const int sxy = g.nx * g.ny;
    const int k = c / sxy, j = (c - k * sxy) / g.nx, i = c - k * sxy - j * g.nx;
    const int xf = i + (g.nx + 1) * (j + g.ny * k), yf = i + g.nx * (j + (g.ny + 1) * k);
    const float fo[6] = {g.fx[xf], g.fx[xf + 1], g.fy[yf], g.fy[yf + g.nx], g.fz[c], g.fz[c + sxy]};
    const int nb[6] = {i > 0 ? c - 1 : -1,         i < g.nx - 1 ? c + 1 : -1, j > 0 ? c - g.nx : -1,
                       j < g.ny - 1 ? c + g.nx : -1, k > 0 ? c - sxy : -1,      k < g.nz - 1 ? c + sxy : -1};
    const float Tc = T[c], nuc = nu[c];
    const float rc = Pb[c] / (RS * Tc);
    float q = 0.0f;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        if (fo[d] <= 0.0f) continue;
        const int m = nb[d];
        if (m >= 0 && g.blocked[m]) continue;
        const float Tn = m >= 0 ? T[m] : T0;
        const float nun = m >= 0 ? nu[m] : nuc;
        const float rn = m >= 0 ? Pb[m] / (RS * T[m]) : P0 / (RS * T0);
        const float Tf = 0.5f * (Tc + Tn);
        const float kf = 0.5f * (rc + rn) * CP * 0.5f * (nuc + nun) / PR_T + k_mol(Tf);
        q += fo[d] * kf * g.dx * (Tn - Tc);
    }
    Q[c] += q * dt;
}

// ---------------------------------------------------------------------------
// Momentum of one velocity component C on its faces, one thread per face
// (gasphase.cpp:275-381).  The face's five-face lines along each axis give
// the van Leer limited upwind fluxes and the Laplacian; closed or leak
// neighbours mirror the face (free slip).  Boundary faces take the wind
// table's inflow (host windAt, per side and level) or keep their value for
// the zero-gradient pass.  Bytes per face: vel (5 reads from cache), the
// other two components (4 faces each, cached), fo, two cells' T Pb nu p,
// out written: about 40 B.
// ---------------------------------------------------------------------------
struct FlowArgs {
    int nx, ny, nz;
    float dx, dt;
    const float* u;
    const float* v;
    const float* w;
    const float* fx;
    const float* fy;
    const float* fz;
    const uint8_t* lx;       // leak faces or null
    const uint8_t* ly;
    const uint8_t* lz;
    const float* T;
    const float* Pb;
    const float* nu;
    const float* p;
    const float* wind;       // [axis 0..1][side 0..1][k]: inflow on that boundary face, NaN if none; or null
    float* u2;
    float* v2;
    float* w2;
};

__device__ __forceinline__ float van_leer(float a, float b) {   // gasphase.cpp:34
#pragma clang fp contract(off)
    return a * b > 0 ? 2 * a * b / (a + b) : 0.0f;
}

template <int C>
__global__ void __launch_bounds__(TB) k_momentum(FlowArgs A) {
#pragma clang fp contract(off)
    using namespace gasc;
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const int dX = nx + (C == 0), dY = ny + (C == 1), dZ = nz + (C == 2);
    const int f = blockIdx.x * TB + threadIdx.x;
    if (f >= dX * dY * dZ) return;
    const float* vel = C == 0 ? A.u : C == 1 ? A.v : A.w;
    const float* fo = C == 0 ? A.fx : C == 1 ? A.fy : A.fz;
    const uint8_t* lk = C == 0 ? A.lx : C == 1 ? A.ly : A.lz;
    float* out = C == 0 ? A.u2 : C == 1 ? A.v2 : A.w2;
    if (fo[f] <= 0.0f) {
        out[f] = 0.0f;
        return;
    }
    if (lk && lk[f]) {   // set by the crack flow
        out[f] = vel[f];
        return;
    }
    const int i = f % dX, j = (f / dX) % dY, k = f / (dX * dY);
    const int id[3] = {i, j, k}, dim[3] = {dX, dY, dZ}, st[3] = {1, dX, dX * dY};
    const int ncomp = C == 0 ? nx : C == 1 ? ny : nz;
    if (id[C] == 0 || id[C] >= ncomp) {
        // boundary face: wind inflow, or left for the zero-gradient pass
        const float wv = A.wind && C < 2 ? A.wind[(C * 2 + (id[C] == 0 ? 0 : 1)) * nz + k] : __builtin_nanf("");
        out[f] = wv == wv ? wv : vel[f];
        return;
    }
    const float v0 = vel[f];
    const float idx = 1.0f / A.dx, idx2 = 1.0f / (A.dx * A.dx);
    // the face's line along each axis: m2 m1 [v0] p1 p2, closed/leak neighbours mirrored
    float m1[3], p1[3], m2[3], p2[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int s = -1; s <= 1; s += 2) {
            float n1 = v0, n2;
            bool ok1 = false;
            if (id[a] + s >= 0 && id[a] + s < dim[a]) {
                const int g1 = f + s * st[a];
                ok1 = fo[g1] > 0.0f && !(lk && lk[g1]);
                n1 = ok1 ? vel[g1] : v0;
            }
            n2 = n1;
            if (id[a] + 2 * s >= 0 && id[a] + 2 * s < dim[a]) {
                const int g1 = f + s * st[a], g2 = f + 2 * s * st[a];
                n2 = fo[g2] > 0.0f && fo[g1] > 0.0f && !(lk && (lk[g2] || lk[g1])) ? vel[g2] : n1;
            }
            if (s < 0) {
                m1[a] = n1;
                m2[a] = n2;
            } else {
                p1[a] = n1;
                p2[a] = n2;
            }
        }
    }
    // advecting velocity: this component, and the four faces of each other
    // component around the face (gasphase.cpp:330-345)
    float adv[3];
    adv[C] = v0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (a == C) continue;
        const float* va = a == 0 ? A.u : a == 1 ? A.v : A.w;
        const int aX = nx + (a == 0), aY = ny + (a == 1), na = a == 0 ? nx : a == 1 ? ny : nz;
        float s = 0.0f;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int pp[3] = {i, j, k};
            if (q & 1) pp[a] += 1;
            if (q & 2) pp[C] -= 1;
            pp[a] = min(pp[a], na);
            pp[C] = max(pp[C], 0);
            s += va[pp[0] + aX * (pp[1] + aY * pp[2])];
        }
        adv[a] = 0.25f * s;
    }
    float acc = 0.0f, lap = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float fp, fm;
        if (adv[a] >= 0.0f) {
            fp = v0 + 0.5f * van_leer(v0 - m1[a], p1[a] - v0);
            fm = m1[a] + 0.5f * van_leer(m1[a] - m2[a], v0 - m1[a]);
        } else {
            fp = p1[a] - 0.5f * van_leer(p1[a] - v0, p2[a] - p1[a]);
            fm = v0 - 0.5f * van_leer(v0 - m1[a], p1[a] - v0);
        }
        acc -= adv[a] * (fp - fm) * idx;
        lap += (m1[a] - 2 * v0 + p1[a]) * idx2;
    }
    // the two cells the face separates
    const int cb = i + nx * (j + ny * k), ca = cb - (C == 0 ? 1 : C == 1 ? nx : nx * ny);
    const float ra = A.Pb[ca] / (RS * A.T[ca]), rb = A.Pb[cb] / (RS * A.T[cb]);
    const float rf = 0.5f * (ra + rb);
    const float nuf = 0.5f * (A.nu[ca] + A.nu[cb]);
    acc += nuf * lap;
    if (C == 2) {
        const float r0 = 0.5f * (A.Pb[ca] + A.Pb[cb]) / (RS * T0);
        acc += GRAV * (r0 - rf) / rf;   // buoyancy of the full density difference
    }
    const float pt = rf * 0.5f * (A.p[ca] + A.p[cb]);
    acc += pt * (1.0f / rb - 1.0f / ra) * idx;   // baroclinic correction
    out[f] = fl_clamp(v0 + A.dt * acc, -VMAX, VMAX);
}

// open boundary faces without wind inflow take the new value of the face one
// cell inside (zero gradient; gasphase.cpp:382-397); the ground is not one
template <int C>
__global__ void __launch_bounds__(TB) k_momentum_edge(FlowArgs A) {
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    const int dX = nx + (C == 0), dY = ny + (C == 1);
    const int na = C == 0 ? ny : C == 1 ? nz : nx;   // the two axes across the boundary plane
    const int nb = C == 0 ? nz : C == 1 ? nx : ny;
    const int t = blockIdx.x * TB + threadIdx.x;
    if (t >= 2 * na * nb) return;
    const int side = t / (na * nb), r = t % (na * nb), a2 = r % na, b2 = r / na;
    if (C == 2 && side == 0) return;
    const int ncomp = C == 0 ? nx : C == 1 ? ny : nz;
    int id[3];
    id[C] = side == 0 ? 0 : ncomp;
    id[(C + 1) % 3] = a2;
    id[(C + 2) % 3] = b2;
    const int f = id[0] + dX * (id[1] + dY * id[2]);
    const float* fo = C == 0 ? A.fx : C == 1 ? A.fy : A.fz;
    const uint8_t* lk = C == 0 ? A.lx : C == 1 ? A.ly : A.lz;
    float* out = C == 0 ? A.u2 : C == 1 ? A.v2 : A.w2;
    if (fo[f] <= 0.0f || (lk && lk[f])) return;
    if (A.wind && C < 2) {
        const float wv = A.wind[(C * 2 + side) * nz + id[2]];
        if (wv == wv) return;
    }
    const int st = C == 0 ? 1 : C == 1 ? dX : dX * dY;
    const int g = f + (side == 0 ? st : -st);
    out[f] = fo[g] > 0.0f && !(lk && lk[g]) ? out[g] : 0.0f;
}

// ---------------------------------------------------------------------------
// Gas added to listed cells (addGasSpecies, solver.cpp:213-224), with the
// burner's pilot flag (injectBurners, gasphase.cpp:92-105)
// ---------------------------------------------------------------------------
struct GasFields {
    float* T;
    float* Yo;
    float* Yf;
    float* Ys;
    float* Ya;
    float* Yp;
    float* Yc;
    float* Yw;
    float* Yco;
    float* Yai;
    float* Pb;
    float* massSrc;
    float* Q;
    float* Qr;
    float* hrr;
    uint8_t* pilot;
    uint8_t* lit;
};

__global__ void __launch_bounds__(TB) k_add_gas(int ncell, const int* __restrict__ cells, const uint8_t* blocked,
                                                float dx, float fuel_gen, float aerosol, float soot_pot, float ait,
                                                int set_pilot, GasFields F) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * TB + threadIdx.x;
    if (t >= ncell) return;
    const int c = cells[t];
    if (blocked[c]) return;
    if (set_pilot) F.pilot[c] = 1;
    const float m = F.Pb[c] / (gasc::RS * F.T[c]) * dx * dx * dx;
    const float madd = fuel_gen + aerosol;
    if (madd <= 0.0f) return;
    const float mt = m + madd, sc = m / mt;
    F.Yo[c] *= sc;
    F.Ys[c] *= sc;
    F.Yc[c] *= sc;
    F.Yw[c] *= sc;
    F.Yco[c] *= sc;
    F.Yai[c] = F.Yai[c] * sc + fuel_gen * ait / mt;
    F.Yf[c] = F.Yf[c] * sc + fuel_gen / mt;
    F.Ya[c] = F.Ya[c] * sc + aerosol / mt;
    F.Yp[c] = F.Yp[c] * sc + soot_pot / mt;
    F.massSrc[c] += madd;
}

// ---------------------------------------------------------------------------
// Combustion, one thread per cell (solver.cpp:284-344).  lit (last step) is
// read; the new flags go to lit_new, copied back by k_copy_u8 afterwards.
// frac = 1 - exp(-dt / tau_mix) comes from the host.  Each block leaves the
// FP64 sum of its cells' heat release (W) in part[blockIdx.x].
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(TB) k_combustion(int nx, int ny, int nz, float dx, float dt, float frac, float chiR,
                                                   const uint8_t* __restrict__ blocked, GasFields F,
                                                   uint8_t* __restrict__ lit_new, double* __restrict__ part) {
#pragma clang fp contract(off)
    using namespace gasc;
    const int n = nx * ny * nz, sz = nx * ny;
    const int c = blockIdx.x * TB + threadIdx.x;
    const float V = dx * dx * dx;
    double h = 0.0;
    if (c < n) {
        uint8_t lnew = 0;
        float hr = 0.0f;
        const float yf = F.Yf[c];
        bool go = !blocked[c] && !(yf < 1e-6f);
        float Tc = 0.0f, ait = 0.0f;
        bool pil = false;
        if (go) {
            Tc = F.T[c];
            ait = yf > 1e-9f ? F.Yai[c] / yf : T_AUTO;
            pil = F.pilot[c] != 0;
            bool ign = pil || Tc > ait;
            if (!ign) {
                const int i = c % nx, j = (c / nx) % ny, k = c / sz;
                ign = (i > 0 && F.lit[c - 1]) || (i < nx - 1 && F.lit[c + 1]) || (j > 0 && F.lit[c - nx]) ||
                      (j < ny - 1 && F.lit[c + nx]) || (k > 0 && F.lit[c - sz]) || (k < nz - 1 && F.lit[c + sz]);
            }
            go = ign;
        }
        float burnable = 0.0f;
        if (go) {
            burnable = fl_min(yf, F.Yo[c] / S_O2);
            go = !(burnable <= 1e-7f);
        }
        if (go) {
            const float tAd = Tc + DHC_FUEL / (CP * (1.0f + S_O2 / fl_max(F.Yo[c], 1e-4f)));
            go = pil || !(tAd < T_CRIT);
        }
        if (go) {
            const float m = F.Pb[c] / (RS * F.T[c]) * V;
            const float dmf = burnable * m * frac;
            const float q = dmf * DHC_FUEL;
            const float sootMade = yf > 0 ? F.Yp[c] * (dmf / m) / yf : 0;
            const float coMade = CO_YIELD * dmf / m;
            const float co2Made = fl_max(0.0f, (C_FRAC * dmf / m - sootMade - (12.0f / 28.0f) * coMade) * (44.0f / 12.0f));
            const float o2Used = S_O2 * dmf / m - (32.0f / 12.0f) * sootMade - (16.0f / 28.0f) * coMade;
            const float yfn = fl_max(0.0f, yf - dmf / m);
            F.Yf[c] = yfn;
            F.Yai[c] = yfn * ait;
            F.Yo[c] = fl_max(0.0f, F.Yo[c] - o2Used);
            F.Yp[c] = fl_max(0.0f, F.Yp[c] - sootMade);
            F.Ys[c] += sootMade;
            F.Yco[c] += coMade;
            F.Yc[c] += co2Made;
            F.Yw[c] += Y_H2O * dmf / m;
            F.Ya[c] *= (1.0f - frac);
            F.Q[c] += q * (1.0f - chiR);
            F.Qr[c] = q * chiR;
            hr = q / (dt * V);
            lnew = q > 1e-3f ? 1 : 0;
        }
        F.hrr[c] = hr;
        lit_new[c] = lnew;
        if (!blocked[c]) h = (double)(hr * V);
    }
    double v[1] = {h};
    block_sums<1>(v, part + blockIdx.x);
}

__global__ void k_copy_u8(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// the start of a step: the per-step sources cleared (solver.cpp:231-234)
__global__ void __launch_bounds__(TB) k_step_clear(int n, float* __restrict__ Q, float* __restrict__ Qr,
                                                   uint8_t* __restrict__ pilot, float* __restrict__ massSrc) {
    const int c = blockIdx.x * TB + threadIdx.x;
    if (c >= n) return;
    Q[c] = 0.0f;
    Qr[c] = 0.0f;
    pilot[c] = 0;
    massSrc[c] = 0.0f;
}

// closed zones: per task the FP64 sum of S V (the zone integral of the
// divergence source, gasphase.cpp:405-409), summed per zone by k_zone_fin_sum
__global__ void __launch_bounds__(TB) k_zone_sv(const int4* __restrict__ ztask, const int* __restrict__ zcells,
                                                const float* __restrict__ S, float V, double* __restrict__ out) {
    const int4 tk = ztask[blockIdx.x];
    double v[1] = {0.0};
    for (int p = tk.y + threadIdx.x; p < tk.z; p += TB) v[0] += (double)(S[zcells[p]] * V);
    __shared__ double r[1];
    block_sums<1>(v, r);
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = r[0];
}

__global__ void __launch_bounds__(TB) k_zone_fin_sum(int nclosed, const int* __restrict__ ztptr,
                                                     const double* __restrict__ part, double* __restrict__ out) {
    for (int z = threadIdx.x; z < nclosed; z += TB) {
        double acc = 0.0;
        for (int t = ztptr[z]; t < ztptr[z + 1]; ++t) acc += part[t];
        out[z] = acc;
    }
}

// per closed zone value (shift of S, or background pressure) applied to its cells
__global__ void __launch_bounds__(TB) k_zone_apply(int n, const int* __restrict__ zclosed, const float* __restrict__ val,
                                                   float* __restrict__ x, int subtract) {
    const int c = blockIdx.x * TB + threadIdx.x;
    if (c >= n) return;
    const int z = zclosed[c];
    if (z < 0) return;
    if (subtract) x[c] -= val[z];
    else x[c] = val[z];
}

}  // namespace phg
