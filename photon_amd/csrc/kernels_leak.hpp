// kernels_leak.hpp -- crack flow and the background pressure of closed zones
// (Solver::leakage, src/sim/gasphase.cpp:402-498) on the device.
//
// A crack is a face of an intact barrier whose open fraction is below 0.05:
// it is left out of the pressure matrix and carries the orifice flow
// cd * sqrt(2 |dp| / rho_upwind), capped at 40 m/s.  Per step:
//   k_crack_prep      per (closed zone, crack to another zone or outside):
//                     the other side's zone pressure and both densities,
//                     fixed during the zone's root search;
//   k_crack_bisect    one workgroup per closed zone: the end-of-step
//                     background pressure Pt of
//                       Pt - P = dt gamma P / Vz (integral of S dV - net outflow(Pt))
//                     by 40 bisection steps on [P - 20 kPa, P + 20 kPa]
//                     (P kept when the bracket does not hold); each
//                     evaluation sums the zone's crack flows in FP64 across
//                     the workgroup;
//   k_crack_velocity  per crack: relax the face velocity halfway to the
//                     orifice velocity of the new zone pressures plus the
//                     dynamic-pressure difference, and its volume flow;
//   k_crack_zone_out  one workgroup per closed zone: the net crack outflow
//                     (FP64), the rise rate zoneDP and the shift of S that
//                     keeps the projection compatible.
// Float expressions keep the reference's order and are not contracted.  The
// reference sums in FP64 sequentially; here the FP64 sums are trees, so a
// float formed from one can differ by a rounding where the sum falls within
// a few 2^-53 of a float boundary.
#pragma once
#include "device.hpp"
#include "kernels_flow.hpp"

namespace phg {

constexpr float GAS_GAMMA = gasc::CP / (gasc::CP - gasc::RS);   // solver.h:175

struct CrackDev {
    int ncr = 0;                      // crack faces, x then y then z, each in face order
    const int4* face = nullptr;       // (axis, face, cell a (low side), cell b (high side)); -1 outside
    const int2* fzone = nullptr;      // zone of a and of b (-1 outside)
    const float* oa = nullptr;        // open fraction * dx^2
    const int* bptr = nullptr;        // closed zone k: bisection entries [bptr[k], bptr[k+1])
    const int* bent = nullptr;        //   crack << 1 | (zone is side a), cracks to another zone or outside
    const int* optr = nullptr;        // closed zone k: outflow entries [optr[k], optr[k+1])
    const int* oent = nullptr;        //   crack << 1 | (zone is side b: subtract), every crack touching it
};

__device__ __forceinline__ float crack_speed(float dp, float rho_up) {   // gasphase.cpp:431-434
#pragma clang fp contract(off)
    const float s = dp >= 0 ? 1.0f : -1.0f;
    return s * fl_min(40.0f, 1.0f * sqrtf(2.0f * fabsf(dp) / fl_max(rho_up, 0.05f)));
}

__device__ __forceinline__ float cell_rho(int c, const float* __restrict__ Pb, const float* __restrict__ T) {
#pragma clang fp contract(off)
    return c >= 0 ? Pb[c] / (gasc::RS * T[c]) : gasc::P0 / (gasc::RS * gasc::T0);   // Solver::gasDensity
}

// (pressure of the other side, density of this side, density of the other side, open area)
__global__ void __launch_bounds__(TB) k_crack_prep(CrackDev cd, int nent, const float* __restrict__ zoneP,
                                                   const float* __restrict__ Pb, const float* __restrict__ T,
                                                   float4* __restrict__ pre) {
    const int e = blockIdx.x * TB + threadIdx.x;
    if (e >= nent) return;
    const int ent = cd.bent[e], q = ent >> 1;
    const bool mine_a = ent & 1;
    const int4 fc = cd.face[q];
    const int2 zz = cd.fzone[q];
    const int self = mine_a ? fc.z : fc.w, other = mine_a ? fc.w : fc.z;
    const int zo = mine_a ? zz.y : zz.x;
    pre[e] = make_float4(zo >= 0 ? zoneP[zo] : gasc::P0, cell_rho(self, Pb, T), cell_rho(other, Pb, T), cd.oa[q]);
}

// Uniform across the workgroup: every thread forms the same total in the
// same order, so all take the same branch.
__device__ __forceinline__ float crack_resid(float Pt, float P, float c, double dint, const float4* __restrict__ pre,
                                             int e0, int e1, double* red) {
#pragma clang fp contract(off)
    double acc = 0.0;
    for (int e = e0 + (int)threadIdx.x; e < e1; e += TB) {
        const float4 t = pre[e];
        const float dp = Pt - t.x;
        const float rho = dp >= 0 ? t.y : t.z;
        acc += (double)(t.w * fabsf(crack_speed(dp, rho)) * (dp >= 0 ? 1.0f : -1.0f));
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    double out = 0.0;
#pragma unroll
    for (int w = 0; w < TB / 64; ++w) out += red[w];
    __syncthreads();   // red is rewritten by the next evaluation
    return Pt - P - c * (float)(dint - out);
}

__global__ void __launch_bounds__(TB) k_crack_bisect(CrackDev cd, const int* __restrict__ czone,
                                                     const float* __restrict__ zvol, const double* __restrict__ dint,
                                                     const float* __restrict__ zoneP, const float4* __restrict__ pre,
                                                     float dt, float* __restrict__ Pnew) {
#pragma clang fp contract(off)
    __shared__ double red[TB / 64];
    const int k = blockIdx.x, z = czone[k];
    const int e0 = cd.bptr[k], e1 = cd.bptr[k + 1];
    const float P = zoneP[z], Vz = zvol[k];
    const double dI = dint[k];
    const float c = dt * GAS_GAMMA * P / Vz;
    float lo = P - 20000.0f, hi = P + 20000.0f;
    const float rl = crack_resid(lo, P, c, dI, pre, e0, e1, red);
    const float rh = crack_resid(hi, P, c, dI, pre, e0, e1, red);
    float Pn = P;
    if (!(rl > 0 || rh < 0)) {
        for (int it = 0; it < 40; ++it) {
            const float mid = 0.5f * (lo + hi);
            if (crack_resid(mid, P, c, dI, pre, e0, e1, red) > 0) hi = mid;
            else lo = mid;
        }
        Pn = 0.5f * (lo + hi);
    }
    if (threadIdx.x == 0) Pnew[z] = Pn;
}

__global__ void __launch_bounds__(TB) k_crack_velocity(CrackDev cd, const float* __restrict__ Pnew,
                                                       const float* __restrict__ Pb, const float* __restrict__ T,
                                                       const float* __restrict__ p, float* __restrict__ u,
                                                       float* __restrict__ v, float* __restrict__ w,
                                                       float* __restrict__ qflow) {
#pragma clang fp contract(off)
    const int q = blockIdx.x * TB + threadIdx.x;
    if (q >= cd.ncr) return;
    const int4 fc = cd.face[q];
    const int2 zz = cd.fzone[q];
    const float pa = zz.x >= 0 ? Pnew[zz.x] : gasc::P0, pb = zz.y >= 0 ? Pnew[zz.y] : gasc::P0;
    const float ra = cell_rho(fc.z, Pb, T), rb = cell_rho(fc.w, Pb, T);
    const float Ha = fc.z >= 0 ? p[fc.z] : 0.0f, Hb = fc.w >= 0 ? p[fc.w] : 0.0f;
    const float dp = (pa - pb) + 0.5f * (ra + rb) * (Ha - Hb);
    const float un = crack_speed(dp, dp >= 0 ? ra : rb);
    float* vel = fc.x == 0 ? u : fc.x == 1 ? v : w;
    const float nv = 0.5f * (vel[fc.y] + un);
    vel[fc.y] = nv;
    qflow[q] = cd.oa[q] * nv;
}

// zdp[k]: the closed zone's rise rate; shift[k]: what S loses in it
__global__ void __launch_bounds__(TB) k_crack_zone_out(CrackDev cd, const int* __restrict__ czone,
                                                       const float* __restrict__ zvol, const double* __restrict__ dint,
                                                       const float* __restrict__ zoneP, const float* __restrict__ qflow,
                                                       float* __restrict__ zdp, float* __restrict__ shift) {
#pragma clang fp contract(off)
    __shared__ double red[TB / 64];
    const int k = blockIdx.x;
    const int e0 = cd.optr ? cd.optr[k] : 0, e1 = cd.optr ? cd.optr[k + 1] : 0;   // no crack faces: no outflow
    double acc = 0.0;
    for (int e = e0 + (int)threadIdx.x; e < e1; e += TB) {
        const int ent = cd.oent[e];
        const double qf = (double)qflow[ent >> 1];
        acc += (ent & 1) ? -qf : qf;
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double out = 0.0;
#pragma unroll
        for (int w = 0; w < TB / 64; ++w) out += red[w];
        const float Vz = zvol[k];
        const float d = (float)(dint[k] - out);
        zdp[k] = GAS_GAMMA * zoneP[czone[k]] * d / Vz;
        shift[k] = d / Vz;
    }
}

// the closed zones' background pressure follows its rise rate
// (solver.cpp:258-263): zoneP[z] = clamp(zoneP[z] + dt zdp[k], P0 - 20 kPa,
// P0 + 50 kPa), in the reference's float operations; val[k] takes the new
// pressure for k_zone_apply
__global__ void __launch_bounds__(TB) k_zone_pressure(int nclosed, const int* __restrict__ czone,
                                                      const float* __restrict__ zdp, float dt,
                                                      float* __restrict__ zoneP, float* __restrict__ val) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * TB + threadIdx.x;
    if (k >= nclosed) return;
    const int z = czone[k];
    const float v = fl_clamp(zoneP[z] + dt * zdp[k], 101325.0f - 20000.0f, 101325.0f + 50000.0f);
    zoneP[z] = v;
    val[k] = v;
}

// face velocities set to zero where a face update closes the face
// (Solver::updateFace, solver.cpp:92); list: (axis, face)
__global__ void __launch_bounds__(TB) k_face_zero(const int2* __restrict__ list, int n, float* __restrict__ u,
                                                  float* __restrict__ v, float* __restrict__ w) {
    const int q = blockIdx.x * TB + threadIdx.x;
    if (q >= n) return;
    const int2 e = list[q];
    (e.x == 0 ? u : e.x == 1 ? v : w)[e.y] = 0.0f;
}

}  // namespace phg
