// kernels_step.hpp -- the kernels around the pressure solve: RHS assembly,
// velocity correction (Solver::project, src/sim/gasphase.cpp:505-677) and
// the divergence source (energyAndExpansion, :251-271).  One thread per cell
// or face; all of them are HBM-bound elementwise / 7-point passes.  Scalar
// transport (advectScalars) is in kernels_transport.hpp.
#pragma once
#include "device.hpp"

namespace phg {

struct DGrid {
    int nx, ny, nz, n;
    const float* ax;      // masked matrix coefficients (the reference's coefs(), gasphase.cpp:518-527)
    const float* ay;
    const float* az;
    const float* fx;      // face open fractions
    const float* fy;
    const float* fz;
    const uint8_t* blocked;
};

__device__ __forceinline__ float h_out(float vel, bool inflow) { return inflow ? -0.5f * vel * vel : 0.0f; }

// std::clamp's semantics (a NaN passes through, as in the reference)
__device__ __forceinline__ float clamp_ref(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }

// b = (S - div(f u*)) dx^2/dt plus the known outside H on open boundary
// faces, diag = sum of the masked coefficients, p = 0 in solids (gasphase.cpp:528-549)
__global__ void __launch_bounds__(TB) k_rhs(DGrid g, float dx, float inv, const float* __restrict__ u,
                                            const float* __restrict__ v, const float* __restrict__ w,
                                            const float* __restrict__ S, float* __restrict__ b,
                                            float* __restrict__ diag, float* __restrict__ p) {
    const int c = blockIdx.x * TB + threadIdx.x;
    if (c >= g.n) return;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    int i, j, k;
    cell_ijk(c, nx, ny, i, j, k);
    if (g.blocked[c]) {
        b[c] = 0.0f;
        if (diag) diag[c] = 0.0f;
        p[c] = 0.0f;
        return;
    }
    const int a0 = c + j + ny * k, a1 = a0 + 1, b0 = c + nx * k, b1 = b0 + nx, c0 = c, c1 = c + nx * ny;
    const float div = (g.fx[a1] * u[a1] - g.fx[a0] * u[a0] + g.fy[b1] * v[b1] - g.fy[b0] * v[b0] +
                       g.fz[c1] * w[c1] - g.fz[c0] * w[c0]) / dx;
    float bb = (S[c] - div) * inv;
    const float co[6] = {g.ax[a0], g.ax[a1], g.ay[b0], g.ay[b1], g.az[c0], g.az[c1]};
    if (i == 0 && co[0] > 0) bb += co[0] * h_out(u[a0], u[a0] > 0);
    if (i == nx - 1 && co[1] > 0) bb += co[1] * h_out(u[a1], u[a1] < 0);
    if (j == 0 && co[2] > 0) bb += co[2] * h_out(v[b0], v[b0] > 0);
    if (j == ny - 1 && co[3] > 0) bb += co[3] * h_out(v[b1], v[b1] < 0);
    if (k == nz - 1 && co[5] > 0) bb += co[5] * h_out(w[c1], w[c1] < 0);
    b[c] = bb;
    if (diag) diag[c] = co[0] + co[1] + co[2] + co[3] + co[4] + co[5];
}

__global__ void k_fill_f32(float* p, int n, float v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// u -= dt/dx grad H on matrix faces, clamp +-40 m/s, max |u| f
// (gasphase.cpp:640-676).  Faces are numbered x faces, then y faces, then the
// z faces above each cell (k = 1 .. nz; the ground face is not touched).
__global__ void __launch_bounds__(TB) k_correct(DGrid g, float sc, const float* __restrict__ p, float* __restrict__ u,
                                                float* __restrict__ v, float* __restrict__ w, float* __restrict__ vmax) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int nfx = (nx + 1) * ny * nz, nfy = nx * (ny + 1) * nz, nfz = nx * ny * nz;
    const int t = blockIdx.x * TB + threadIdx.x;
    const float VMAX = 40.0f;
    float vm = 0.0f;
    if (t < nfx) {
        const int f = t, i = f % (nx + 1), jk = f / (nx + 1), j = jk % ny, k = jk / ny;
        if (g.fx[f] <= 0) {
            u[f] = 0.0f;
        } else {
            float uf = u[f];
            if (g.ax[f] > 0) {
                const int c = i + nx * (j + ny * k);
                const float hr = i < nx ? p[c] : h_out(uf, uf < 0);
                const float hl = i > 0 ? p[c - 1] : h_out(uf, uf > 0);
                uf = clamp_ref(uf - sc * (hr - hl), -VMAX, VMAX);
                u[f] = uf;
            }
            vm = fabsf(uf) * g.fx[f];
        }
    } else if (t < nfx + nfy) {
        const int f = t - nfx, i = f % nx, jk = f / nx, j = jk % (ny + 1), k = jk / (ny + 1);
        if (g.fy[f] <= 0) {
            v[f] = 0.0f;
        } else {
            float vf = v[f];
            if (g.ay[f] > 0) {
                const int c = i + nx * (j + ny * k);
                const float hr = j < ny ? p[c] : h_out(vf, vf < 0);
                const float hl = j > 0 ? p[c - nx] : h_out(vf, vf > 0);
                vf = clamp_ref(vf - sc * (hr - hl), -VMAX, VMAX);
                v[f] = vf;
            }
            vm = fabsf(vf) * g.fy[f];
        }
    } else if (t < nfx + nfy + nfz) {
        const int c = t - nfx - nfy;          // the cell below the face
        const int f = c + nx * ny;            // zface(i, j, k + 1)
        const int k = c / (nx * ny);
        if (g.fz[f] <= 0) {
            w[f] = 0.0f;
        } else {
            float wf = w[f];
            if (g.az[f] > 0) {
                const float hu = k + 1 < nz ? p[f] : h_out(wf, wf < 0);
                wf = clamp_ref(wf - sc * (hu - p[c]), -VMAX, VMAX);
                w[f] = wf;
            }
            vm = fabsf(wf) * g.fz[f];
        }
    }
    vm = block_max(vm);
    if (threadIdx.x == 0) atomic_max_nonneg(vmax, vm);
}

// T from the heat budget, divergence source S (energyAndExpansion, gasphase.cpp:251-271)
__global__ void __launch_bounds__(TB) k_energy(int n, float dx, float dt, const uint8_t* __restrict__ blocked,
                                               const float* __restrict__ Pb, const float* __restrict__ Q,
                                               const float* __restrict__ massSrc, float* __restrict__ T,
                                               float* __restrict__ S, float* __restrict__ maxdiv) {
    const float RS = 287.0f, CP = 1100.0f, T_MIN = 200.0f, T_MAX = 2300.0f, D_MAX = 50.0f;
    const int c = blockIdx.x * TB + threadIdx.x;
    float md = 0.0f;
    if (c < n) {
        if (blocked[c]) {
            S[c] = 0.0f;
        } else {
            const float V = dx * dx * dx;
            const float t = T[c], pb = Pb[c], q = Q[c];
            const float rho = pb / (RS * t);
            const float m = rho * V;
            const float tn = clamp_ref(t + q / (m * CP), T_MIN, T_MAX);
            const float rhon = pb / (RS * tn);
            const float D = q / (m * CP * t * dt) + massSrc[c] / (rhon * V * dt);
            T[c] = tn;
            const float sv = clamp_ref(D, -D_MAX, D_MAX);
            S[c] = sv;
            md = fabsf(sv);
        }
    }
    md = block_max(md);
    if (threadIdx.x == 0) atomic_max_nonneg(maxdiv, md);
}

}  // namespace phg

namespace phg {
// plain 16-byte-per-lane copy (calibration of the HBM byte counters)
__global__ void __launch_bounds__(TB) k_copy16(const float4* __restrict__ src, float4* __restrict__ dst, long n) {
    const long i = (long)blockIdx.x * TB + threadIdx.x;   // one 16-byte element per thread
    if (i < n) dst[i] = src[i];
}
}  // namespace phg
