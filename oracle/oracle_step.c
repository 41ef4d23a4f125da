/*
 * oracle_step.c -- CPU restatement of the rest of Photon's gas-phase step:
 * eddy viscosity, heat conduction, momentum with wind inflow, burner gas,
 * combustion and the stable time step (src/sim/gasphase.cpp,
 * src/sim/solver.cpp; each function cites the lines it restates).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never shipped or
 * called by it.  Plain C11 in ISO mode, so no FMA contraction: every float
 * expression is evaluated in the reference's order, and the results match the
 * reference's own solver built without contraction bit for bit (pow, sin and
 * cos come from the same C library).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define RS 287.0f
#define T0 293.15f
#define P0 101325.0f
#define CP 1100.0f
#define GRAV 9.81f
#define VMAX 40.0f
#define PR_T 0.5f
#define C_NU 0.1f

static float fminr(float a, float b) { return b < a ? b : a; }   /* std::min */
static float fmaxr(float a, float b) { return a < b ? b : a; }   /* std::max */
static float clampr(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }   /* std::clamp */

static float nu_mol(float T) { return 1.5e-5f * powf(T / 293.15f, 1.7f); }   /* gasphase.cpp:37 */
static float k_mol(float T) { return 0.0258f * powf(T / 293.15f, 0.8f); }    /* gasphase.cpp:38 */
static float van_leer(float a, float b) { return a * b > 0 ? 2 * a * b / (a + b) : 0.0f; }   /* gasphase.cpp:34 */

/* windAt (gasphase.cpp:52-61): inflow velocity on a boundary face, NaN if none */
float orc_wind_at(const orc_wind *wd, int axis, int side, float z) {
    if (!wd || wd->speed <= 0 || axis == 2) return NAN;
    float th = wd->from_deg * 3.14159265f / 180.0f;
    float U = wd->speed * powf(fmaxr(z, 0.05f) / 10.0f, wd->exponent);
    float comp = axis == 0 ? -sinf(th) * U : -cosf(th) * U;
    if (side == 0 && comp > 0) return comp;
    if (side == 1 && comp < 0) return comp;
    return NAN;
}

/* stableDt (gasphase.cpp:63-70) */
float orc_stable_dt(float dt_max, float profile_dt_max, float cfl, float dx, float max_vel, float max_nu,
                    float max_div) {
    float dt = fminr(dt_max, profile_dt_max);
    dt = fminr(dt, cfl * dx / fmaxr(max_vel, 0.1f));
    float alpha = max_nu * (1.0f + 1.0f / PR_T);
    if (alpha > 0) dt = fminr(dt, 0.9f * dx * dx / (6.0f * alpha));
    if (max_div > 0) dt = fminr(dt, 0.25f / max_div);
    return fmaxr(dt, 1e-3f);
}

#define CELL(i, j, k) ((i) + g->nx * ((j) + g->ny * (k)))
#define XF(i, j, k) ((i) + (g->nx + 1) * ((j) + g->ny * (k)))
#define YF(i, j, k) ((i) + g->nx * ((j) + (g->ny + 1) * (k)))
#define ZF(i, j, k) ((i) + g->nx * ((j) + g->ny * (k)))

/* Deardorff eddy viscosity (turbulence, gasphase.cpp:163-218): cell-centred
 * velocity, a separable 1-2-1 test filter (x, then y, then z; solid
 * neighbours and the domain edge mirror the cell), nu = min(1, C dx sqrt(k) +
 * nu_mol(T)).  Returns maxNu. */
float orc_turbulence(const orc_grid *g, const float *u, const float *v, const float *w, const float *T, float *nu) {
    const int nx = g->nx, ny = g->ny, nz = g->nz, N = nx * ny * nz;
    float *buf = (float *)calloc((size_t)9 * N, sizeof(float));
    float *uc[3] = {buf, buf + N, buf + 2 * N};
    float *f[3] = {buf + 3 * N, buf + 4 * N, buf + 5 * N};
    float *gg[3] = {buf + 6 * N, buf + 7 * N, buf + 8 * N};
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int c = CELL(i, j, k);
                if (g->blocked[c]) { uc[0][c] = uc[1][c] = uc[2][c] = 0; continue; }
                uc[0][c] = 0.5f * (u[XF(i, j, k)] + u[XF(i + 1, j, k)]);
                uc[1][c] = 0.5f * (v[YF(i, j, k)] + v[YF(i, j + 1, k)]);
                uc[2][c] = 0.5f * (w[ZF(i, j, k)] + w[ZF(i, j, k + 1)]);
            }
    const int str[3] = {1, nx, nx * ny};
    for (int q = 0; q < 3; ++q)
        for (int axis = 0; axis < 3; ++axis) {
            const float *in = axis == 0 ? uc[q] : axis == 1 ? f[q] : gg[q];
            float *out = axis == 1 ? gg[q] : f[q];
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i) {
                        int c = CELL(i, j, k);
                        if (g->blocked[c]) { out[c] = 0; continue; }
                        int id = axis == 0 ? i : axis == 1 ? j : k, n = axis == 0 ? nx : axis == 1 ? ny : nz;
                        float lo = id > 0 && !g->blocked[c - str[axis]] ? in[c - str[axis]] : in[c];
                        float hi = id < n - 1 && !g->blocked[c + str[axis]] ? in[c + str[axis]] : in[c];
                        out[c] = 0.25f * lo + 0.5f * in[c] + 0.25f * hi;
                    }
        }
    float m = 0;
    for (int c = 0; c < N; ++c) {
        if (g->blocked[c]) { nu[c] = 0; continue; }
        float du = uc[0][c] - f[0][c], dv = uc[1][c] - f[1][c], dw = uc[2][c] - f[2][c];
        float ks = 0.5f * (du * du + dv * dv + dw * dw);
        nu[c] = fminr(1.0f, C_NU * g->dx * sqrtf(ks) + nu_mol(T[c]));
        m = fmaxr(m, nu[c]);
    }
    free(buf);
    return m;
}

/* heatDiffusion (gasphase.cpp:221-248): molecular + turbulent conduction
 * between gas cells and with the ambient outside open boundary faces, into Q */
void orc_heat_diffusion(const orc_grid *g, float dt, const float *T, const float *nu, const float *Pb, float *Q) {
    const int nx = g->nx, ny = g->ny, nz = g->nz, sxy = nx * ny;
    const float dx = g->dx;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int c = CELL(i, j, k);
                if (g->blocked[c]) continue;
                float fo[6] = {g->fx[XF(i, j, k)], g->fx[XF(i + 1, j, k)], g->fy[YF(i, j, k)],
                               g->fy[YF(i, j + 1, k)], g->fz[ZF(i, j, k)],  g->fz[ZF(i, j, k + 1)]};
                int nb[6] = {i > 0 ? c - 1 : -1,       i < nx - 1 ? c + 1 : -1,  j > 0 ? c - nx : -1,
                             j < ny - 1 ? c + nx : -1, k > 0 ? c - sxy : -1,     k < nz - 1 ? c + sxy : -1};
                float rc = Pb[c] / (RS * T[c]), q = 0;
                for (int d = 0; d < 6; ++d) {
                    if (fo[d] <= 0) continue;
                    int n = nb[d];
                    if (n >= 0 && g->blocked[n]) continue;
                    float Tn = n >= 0 ? T[n] : T0;
                    float nun = n >= 0 ? nu[n] : nu[c];
                    float rn = n >= 0 ? Pb[n] / (RS * T[n]) : P0 / (RS * T0);
                    float Tf = 0.5f * (T[c] + Tn);
                    float kf = 0.5f * (rc + rn) * CP * 0.5f * (nu[c] + nun) / PR_T + k_mol(Tf);
                    q += fo[d] * kf * dx * (Tn - T[c]);
                }
                Q[c] += q * dt;
            }
}

/* momentum (gasphase.cpp:275-399): each velocity component on its faces:
 * limited (van Leer) upwind advection by the face-averaged velocity, eddy
 * viscosity, buoyancy of the full density difference (w), the baroclinic
 * term; wind inflow on boundary faces it enters through, zero gradient on the
 * other open boundary faces.  u,v,w in; u2,v2,w2 out. */
void orc_momentum(const orc_grid *g, const orc_wind *wd, float dt, const float *u, const float *v, const float *w,
                  const float *T, const float *Pb, const float *nu, const float *p, float *u2, float *v2, float *w2) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const float dx = g->dx, idx = 1.0f / dx, idx2 = 1.0f / (dx * dx);
    const float oz = wd ? wd->oz : 0.0f;
    for (int comp = 0; comp < 3; ++comp) {
        const float *vel = comp == 0 ? u : comp == 1 ? v : w;
        const float *fo = comp == 0 ? g->fx : comp == 1 ? g->fy : g->fz;
        const uint8_t *lk = comp == 0 ? g->lx : comp == 1 ? g->ly : g->lz;
        float *out = comp == 0 ? u2 : comp == 1 ? v2 : w2;
        const int dX = nx + (comp == 0), dY = ny + (comp == 1), dZ = nz + (comp == 2);
        const int st[3] = {1, dX, dX * dY}, dim[3] = {dX, dY, dZ};
        const int ncomp = comp == 0 ? nx : comp == 1 ? ny : nz;
#define LK(f) (lk && lk[f])
        for (int k = 0; k < dZ; ++k)
            for (int j = 0; j < dY; ++j)
                for (int i = 0; i < dX; ++i) {
                    int f = i + dX * (j + dY * k);
                    if (fo[f] <= 0) { out[f] = 0; continue; }
                    if (LK(f)) { out[f] = vel[f]; continue; }
                    int id[3] = {i, j, k};
                    int lo = id[comp] - 1;
                    int bnd = lo < 0 || id[comp] >= ncomp;
                    int ca = -1, cb = -1;
                    {
                        int a[3] = {i, j, k};
                        a[comp] -= 1;
                        if (a[comp] >= 0) ca = CELL(a[0], a[1], a[2]);
                        if (id[comp] < ncomp) cb = CELL(i, j, k);
                    }
                    float zc = oz + (comp == 2 ? k : k + 0.5f) * dx;
                    if (bnd) {
                        float wv = orc_wind_at(wd, comp, lo < 0 ? 0 : 1, zc);
                        out[f] = !isnan(wv) ? wv : vel[f];
                        continue;
                    }
                    float v0 = vel[f];
                    float m1v[3], p1v[3], m2v[3], p2v[3];
                    for (int a = 0; a < 3; ++a) {
                        for (int s = -1; s <= 1; s += 2) {
                            int ni = id[a] + s;
                            float n1 = v0;
                            if (!(ni < 0 || ni >= dim[a])) {
                                int gq = f + s * st[a];
                                n1 = fo[gq] > 0 && !LK(gq) ? vel[gq] : v0;
                            }
                            int ni2 = id[a] + 2 * s;
                            float n2 = n1;
                            if (!(ni2 < 0 || ni2 >= dim[a])) {
                                int g2 = f + 2 * s * st[a], g1 = f + s * st[a];
                                n2 = fo[g2] > 0 && fo[g1] > 0 && !LK(g2) && !LK(g1) ? vel[g2] : n1;
                            }
                            if (s < 0) { m1v[a] = n1; m2v[a] = n2; } else { p1v[a] = n1; p2v[a] = n2; }
                        }
                    }
                    float adv[3];
                    adv[comp] = v0;
                    for (int a = 0; a < 3; ++a) {
                        if (a == comp) continue;
                        const float *va = a == 0 ? u : a == 1 ? v : w;
                        int aX = nx + (a == 0), aY = ny + (a == 1);
                        int na = a == 0 ? nx : a == 1 ? ny : nz;
                        float s = 0;
                        for (int q = 0; q < 4; ++q) {
                            int pp[3] = {i, j, k};
                            if (q & 1) pp[a] += 1;
                            if (q & 2) pp[comp] -= 1;
                            if (pp[a] > na) pp[a] = na;
                            if (pp[comp] < 0) pp[comp] = 0;
                            s += va[pp[0] + aX * (pp[1] + aY * pp[2])];
                        }
                        adv[a] = 0.25f * s;
                    }
                    float acc = 0, lap = 0;
                    for (int a = 0; a < 3; ++a) {
                        float m1 = m1v[a], p1 = p1v[a], m2 = m2v[a], p2 = p2v[a];
                        float fp, fm;
                        if (adv[a] >= 0) {
                            fp = v0 + 0.5f * van_leer(v0 - m1, p1 - v0);
                            fm = m1 + 0.5f * van_leer(m1 - m2, v0 - m1);
                        } else {
                            fp = p1 - 0.5f * van_leer(p1 - v0, p2 - p1);
                            fm = v0 - 0.5f * van_leer(v0 - m1, p1 - v0);
                        }
                        acc -= adv[a] * (fp - fm) * idx;
                        lap += (m1 - 2 * v0 + p1) * idx2;
                    }
                    float ra = Pb[ca] / (RS * T[ca]), rb = Pb[cb] / (RS * T[cb]);
                    float rf = 0.5f * (ra + rb);
                    float nuf = 0.5f * (nu[ca] + nu[cb]);
                    acc += nuf * lap;
                    if (comp == 2) {
                        float r0 = 0.5f * (Pb[ca] + Pb[cb]) / (RS * T0);
                        acc += GRAV * (r0 - rf) / rf;
                    }
                    float pt = rf * 0.5f * (p[ca] + p[cb]);
                    acc += pt * (1.0f / rb - 1.0f / ra) * idx;
                    out[f] = clampr(v0 + dt * acc, -VMAX, VMAX);
                }
        /* open boundary faces without wind inflow: the new value one face inside */
        int nA = dim[(comp + 1) % 3], nB = dim[(comp + 2) % 3];
        for (int b2 = 0; b2 < nB; ++b2)
            for (int a2 = 0; a2 < nA; ++a2)
                for (int side = (comp == 2 ? 1 : 0); side < 2; ++side) {
                    int id[3];
                    id[comp] = side == 0 ? 0 : ncomp;
                    id[(comp + 1) % 3] = a2;
                    id[(comp + 2) % 3] = b2;
                    int f = id[0] + dX * (id[1] + dY * id[2]);
                    if (fo[f] <= 0 || LK(f)) continue;
                    float zc = oz + (comp == 2 ? id[2] : id[2] + 0.5f) * dx;
                    if (!isnan(orc_wind_at(wd, comp, side, zc))) continue;
                    int gq = f + (side == 0 ? 1 : -1) * st[comp];
                    out[f] = fo[gq] > 0 && !LK(gq) ? out[gq] : 0.0f;
                }
#undef LK
    }
}

/* addGasSpecies (solver.cpp:213-224) for each listed gas cell, and the
 * burner's pilot flag (injectBurners, gasphase.cpp:92-105) */
void orc_add_gas(int ncell, const int *cells, float dx, float fuel_gen, float aerosol, float soot_pot, float ait,
                 int set_pilot, const uint8_t *blocked, const orc_species *sp, uint8_t *pilot) {
    for (int q = 0; q < ncell; ++q) {
        int c = cells[q];
        if (blocked[c]) continue;
        if (set_pilot) pilot[c] = 1;
        float m = sp->Pb[c] / (RS * sp->T[c]) * dx * dx * dx;
        float madd = fuel_gen + aerosol;
        if (madd <= 0) continue;
        float mt = m + madd, sc = m / mt;
        sp->Yo[c] *= sc; sp->Ys[c] *= sc; sp->Yc[c] *= sc; sp->Yw[c] *= sc; sp->Yco[c] *= sc;
        sp->Yai[c] = sp->Yai[c] * sc + fuel_gen * ait / mt;
        sp->Yf[c] = sp->Yf[c] * sc + fuel_gen / mt;
        sp->Ya[c] = sp->Ya[c] * sc + aerosol / mt;
        sp->Yp[c] = sp->Yp[c] * sc + soot_pot / mt;
        sp->massSrc[c] += madd;
    }
}

/* combustion (solver.cpp:284-344): mixing-limited burning with ignition by a
 * pilot, a burning neighbour or the mixture's auto-ignition temperature, and
 * the critical-flame-temperature extinction test.  lit is read (last step)
 * and rewritten.  Returns the total heat release rate (W). */
double orc_combustion(const orc_grid *g, float dt, int radiation_on, const orc_species *sp, float *Q, float *Qr,
                      float *hrr, const uint8_t *pilot, uint8_t *lit) {
    const float DHC_FUEL = 25.0e6f, S_O2 = 1.95f, Y_H2O = 0.761f, C_FRAC = 0.597f, CO_YIELD = 0.02f;
    const float CHI_R = 0.35f, T_CRIT = 1600.0f, T_AUTO = 800.0f, TAU_MIX = 0.1f;
    const int nx = g->nx, ny = g->ny, nz = g->nz, N = nx * ny * nz, sz = nx * ny;
    const float V = g->dx * g->dx * g->dx;
    const float s = S_O2;
    const float frac = 1.0f - expf(-dt / TAU_MIX);
    const float chiR = radiation_on ? CHI_R : 0.0f;
    uint8_t *litNew = (uint8_t *)calloc((size_t)N, 1);
    double total = 0;
    for (int c = 0; c < N; ++c) {
        hrr[c] = 0;
        if (g->blocked[c]) continue;
        float yf = sp->Yf[c];
        if (yf < 1e-6f) continue;
        float Tc = sp->T[c];
        float ait = yf > 1e-9f ? sp->Yai[c] / yf : T_AUTO;
        int ign = pilot[c] || Tc > ait;
        if (!ign) {
            int i = c % nx, j = (c / nx) % ny, k = c / sz;
            ign = (i > 0 && lit[c - 1]) || (i < nx - 1 && lit[c + 1]) || (j > 0 && lit[c - nx]) ||
                  (j < ny - 1 && lit[c + nx]) || (k > 0 && lit[c - sz]) || (k < nz - 1 && lit[c + sz]);
        }
        if (!ign) continue;
        float burnable = fminr(yf, sp->Yo[c] / s);
        if (burnable <= 1e-7f) continue;
        float tAd = Tc + DHC_FUEL / (CP * (1.0f + S_O2 / fmaxr(sp->Yo[c], 1e-4f)));
        if (!pilot[c] && tAd < T_CRIT) continue;
        float m = sp->Pb[c] / (RS * sp->T[c]) * V;
        float dmf = burnable * m * frac;
        float q = dmf * DHC_FUEL;
        float sootMade = yf > 0 ? sp->Yp[c] * (dmf / m) / yf : 0;
        float coMade = CO_YIELD * dmf / m;
        float co2Made = fmaxr(0.0f, (C_FRAC * dmf / m - sootMade - (12.0f / 28.0f) * coMade) * (44.0f / 12.0f));
        float o2Used = s * dmf / m - (32.0f / 12.0f) * sootMade - (16.0f / 28.0f) * coMade;
        sp->Yf[c] = fmaxr(0.0f, yf - dmf / m);
        sp->Yai[c] = sp->Yf[c] * ait;
        sp->Yo[c] = fmaxr(0.0f, sp->Yo[c] - o2Used);
        sp->Yp[c] = fmaxr(0.0f, sp->Yp[c] - sootMade);
        sp->Ys[c] += sootMade;
        sp->Yco[c] += coMade;
        sp->Yc[c] += co2Made;
        sp->Yw[c] += Y_H2O * dmf / m;
        sp->Ya[c] *= (1.0f - frac);
        Q[c] += q * (1.0f - chiR);
        Qr[c] = q * chiR;
        hrr[c] = q / (dt * V);
        litNew[c] = q > 1e-3f ? 1 : 0;
    }
    for (int c = 0; c < N; ++c) {
        lit[c] = litNew[c];
        if (!g->blocked[c]) total += hrr[c] * V;
    }
    free(litNew);
    return total;
}

/* ---- crack flow and closed zones' background pressure ---------------------
 * leakage (gasphase.cpp:402-498).  Crack faces (lx/ly/lz set, neither side
 * solid) carry the orifice flow cd*sqrt(2|dp|/rho_up), capped at 40 m/s.  For
 * each closed zone the end-of-step background pressure Pt solves
 *   Pt - P = dt*gamma*P/Vz * (integral of S dV - net crack outflow(Pt))
 * by 40 bisection steps on [P - 20 kPa, P + 20 kPa] (left at P when the
 * bracket does not hold).  Each crack velocity is then relaxed halfway to the
 * orifice velocity of the zone pressures plus the dynamic-pressure term, and
 * every closed zone gets its rise rate zoneDP and the matching shift of S.
 * All sums are FP64, in the reference's order: cells ascending, crack faces
 * x then y then z, each in face order.  zone_vol: each zone's volume, m^3, as
 * a float.  Returns the number of crack faces. */
typedef struct { int axis, f, a, b; } orc_crack;

static float crack_speed(float dp, float rho_up) {   /* gasphase.cpp:431-434 */
    const float sgn = dp >= 0 ? 1.0f : -1.0f;
    return sgn * fminr(40.0f, 1.0f * sqrtf(2.0f * fabsf(dp) / fmaxr(rho_up, 0.05f)));
}

int orc_leakage(const orc_grid *g, float dt, const int *zone, int nzones, const uint8_t *zone_open,
                const float *zone_vol, const float *zoneP, const float *T, const float *Pb, const float *p, float *u,
                float *v, float *w, float *S, float *zoneDP) {
    const int nx = g->nx, ny = g->ny, nz = g->nz, N = nx * ny * nz;
    const float dx = g->dx, V = dx * dx * dx, A = dx * dx;
    const float gamma = CP / (CP - RS), rho_amb = P0 / (RS * T0);
    double *dint = (double *)calloc((size_t)nzones + 1, sizeof(double));
    for (int c = 0; c < N; ++c)
        if (zone[c] >= 0) dint[zone[c]] += S[c] * V;
    /* the crack faces */
    const int nf[3] = {(nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)};
    const uint8_t *lk[3] = {g->lx, g->ly, g->lz};
    const float *fo[3] = {g->fx, g->fy, g->fz};
    float *vel[3] = {u, v, w};
    int ncr = 0;
    for (int ax = 0; ax < 3; ++ax)
        if (lk[ax])
            for (int f = 0; f < nf[ax]; ++f) ncr += lk[ax][f] != 0;
    orc_crack *cr = (orc_crack *)malloc(sizeof(orc_crack) * (size_t)(ncr + 1));
    int m = 0;
    for (int ax = 0; ax < 3; ++ax) {
        if (!lk[ax]) continue;
        const int ex = nx + (ax == 0), ey = ny + (ax == 1);
        const int lim = ax == 0 ? nx : ax == 1 ? ny : nz;
        for (int f = 0; f < nf[ax]; ++f) {
            if (!lk[ax][f]) continue;
            const int i = f % ex, j = (f / ex) % ey, k = f / (ex * ey);
            const int pos = ax == 0 ? i : ax == 1 ? j : k;
            const int step = ax == 0 ? 1 : ax == 1 ? nx : nx * ny;
            const int cb = pos < lim ? i + nx * (j + ny * k) : -1;
            const int ca = pos > 0 ? i + nx * (j + ny * k) - step : -1;
            if ((ca >= 0 && g->blocked[ca]) || (cb >= 0 && g->blocked[cb])) continue;
            cr[m].axis = ax; cr[m].f = f; cr[m].a = ca; cr[m].b = cb;
            ++m;
        }
    }
    ncr = m;
#define ZP(c, P) ((c) >= 0 && zone[c] >= 0 ? (P)[zone[c]] : P0)
#define RHO(c) ((c) >= 0 ? Pb[c] / (RS * T[c]) : rho_amb)
    /* each zone's cracks to another zone or outside, in crack order */
    int *cnt = (int *)calloc((size_t)nzones + 1, sizeof(int));
    for (int q = 0; q < ncr; ++q) {
        const int za = cr[q].a >= 0 ? zone[cr[q].a] : -1, zb = cr[q].b >= 0 ? zone[cr[q].b] : -1;
        if (za == zb) continue;
        if (za >= 0) cnt[za]++;
        if (zb >= 0) cnt[zb]++;
    }
    int *ptr = (int *)calloc((size_t)nzones + 2, sizeof(int));
    for (int z = 0; z < nzones; ++z) ptr[z + 1] = ptr[z] + cnt[z];
    int *lst = (int *)malloc(sizeof(int) * (size_t)(ptr[nzones] + 1));
    memset(cnt, 0, sizeof(int) * (size_t)(nzones + 1));
    for (int q = 0; q < ncr; ++q) {
        const int za = cr[q].a >= 0 ? zone[cr[q].a] : -1, zb = cr[q].b >= 0 ? zone[cr[q].b] : -1;
        if (za == zb) continue;
        if (za >= 0) lst[ptr[za] + cnt[za]++] = q;
        if (zb >= 0) lst[ptr[zb] + cnt[zb]++] = q;
    }
    float *Pnew = (float *)malloc(sizeof(float) * (size_t)(nzones + 1));
    for (int z = 0; z < nzones; ++z) Pnew[z] = zoneP[z];
    for (int z = 0; z < nzones; ++z) {
        if (zone_open[z]) continue;
        const float P = zoneP[z], Vz = zone_vol[z];
        float lo = P - 20000.0f, hi = P + 20000.0f;
        float rl = 0, rh = 0;
        for (int pass = 0; pass < 42; ++pass) {
            const float Pt = pass == 0 ? lo : pass == 1 ? hi : 0.5f * (lo + hi);
            double outflow = 0;
            for (int e = ptr[z]; e < ptr[z + 1]; ++e) {
                const orc_crack *c = &cr[lst[e]];
                const int mine_a = c->a >= 0 && zone[c->a] == z;
                const int other = mine_a ? c->b : c->a;
                const float dp = Pt - ZP(other, zoneP);
                const int up = dp >= 0 ? (mine_a ? c->a : c->b) : other;
                outflow += fo[c->axis][c->f] * A * fabsf(crack_speed(dp, RHO(up))) * (dp >= 0 ? 1 : -1);
            }
            const float r = Pt - P - dt * gamma * P / Vz * (float)(dint[z] - outflow);
            if (pass == 0) rl = r;
            else if (pass == 1) {
                rh = r;
                if (rl > 0 || rh < 0) break;
            } else if (r > 0) hi = Pt;
            else lo = Pt;
        }
        Pnew[z] = (rl > 0 || rh < 0) ? P : 0.5f * (lo + hi);
    }
    /* relaxed crack velocities and each zone's net outflow */
    double *out = (double *)calloc((size_t)nzones + 1, sizeof(double));
    for (int q = 0; q < ncr; ++q) {
        const orc_crack *c = &cr[q];
        const float pa = ZP(c->a, Pnew), pb = ZP(c->b, Pnew);
        const float ra = RHO(c->a), rb = RHO(c->b);
        const float Ha = c->a >= 0 ? p[c->a] : 0, Hb = c->b >= 0 ? p[c->b] : 0;
        const float dp = (pa - pb) + 0.5f * (ra + rb) * (Ha - Hb);
        const float un = crack_speed(dp, dp >= 0 ? ra : rb);
        float *vv = vel[c->axis];
        vv[c->f] = 0.5f * (vv[c->f] + un);
        const float flow = fo[c->axis][c->f] * A * vv[c->f];
        if (c->a >= 0 && zone[c->a] >= 0) out[zone[c->a]] += flow;
        if (c->b >= 0 && zone[c->b] >= 0) out[zone[c->b]] -= flow;
    }
#undef ZP
#undef RHO
    float *shift = (float *)calloc((size_t)nzones + 1, sizeof(float));
    for (int z = 0; z < nzones; ++z) {
        zoneDP[z] = 0;
        if (zone_open[z]) continue;
        const float Vz = zone_vol[z];
        zoneDP[z] = gamma * zoneP[z] * (float)(dint[z] - out[z]) / Vz;
        shift[z] = (float)(dint[z] - out[z]) / Vz;
    }
    for (int c = 0; c < N; ++c)
        if (zone[c] >= 0 && !zone_open[zone[c]]) S[c] -= shift[zone[c]];
    free(shift); free(out); free(Pnew); free(lst); free(ptr); free(cnt); free(cr); free(dint);
    return ncr;
}
