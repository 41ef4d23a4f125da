// ref_driver.cpp -- C entry points over the reference's own photon::Solver,
// used to pin the oracle and as the CPU baseline.  TEST INFRASTRUCTURE ONLY.
//
// Built by oracle/Makefile together with the reference's unmodified sources
// (src/sim/*.cpp, src/scene/*.cpp under the reference checkout) into
// oracle/_ref/libphoton_ref.so.  Only this file sees the Solver's private
// members (it is compiled with private mapped to public): the synthetic
// grids of the tests cannot be made through Scene + discretizer
// (src/sim/discretize.cpp:18-20,282 always add margins and a roof), so the
// driver writes the solver's arrays directly and calls the private phases.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>
#define private public
#include "sim/solver.h"
#undef private

using photon::Solver;

namespace {
std::vector<float>* field(Solver* s, const char* name) {
    std::string n(name);
    if (n == "u") return &s->u;
    if (n == "v") return &s->v;
    if (n == "w") return &s->w;
    if (n == "fx") return &s->fx;
    if (n == "fy") return &s->fy;
    if (n == "fz") return &s->fz;
    if (n == "ax") return &s->ax_;
    if (n == "ay") return &s->ay_;
    if (n == "az") return &s->az_;
    if (n == "p") return &s->p;
    if (n == "S") return &s->S;
    if (n == "T") return &s->T;
    if (n == "Yo") return &s->Yo;
    if (n == "Yf") return &s->Yf;
    if (n == "Ys") return &s->Ys;
    if (n == "Ya") return &s->Ya;
    if (n == "Yp") return &s->Yp;
    if (n == "Yc") return &s->Yc;
    if (n == "Yw") return &s->Yw;
    if (n == "Yco") return &s->Yco;
    if (n == "Yai") return &s->Yai;
    if (n == "nu") return &s->nu;
    if (n == "Pb") return &s->Pb;
    if (n == "Q") return &s->Q;
    if (n == "massSrc") return &s->massSrc;
    if (n == "rhs") return &s->rhs_;
    if (n == "diag") return &s->tmp_[9];
    if (n == "Qr") return &s->Qr;
    if (n == "hrr") return &s->hrr;
    return nullptr;
}
std::vector<uint8_t>* bytes(Solver* s, const char* name) {
    std::string n(name);
    if (n == "lit") return &s->lit;
    if (n == "pilot") return &s->pilot;
    if (n == "lx") return &s->lx_;
    if (n == "ly") return &s->ly_;
    if (n == "lz") return &s->lz_;
    return nullptr;
}
}  // namespace

extern "C" {

void* ref_create(int threads) { return new Solver(threads); }
void ref_destroy(void* h) { delete static_cast<Solver*>(h); }

// Geometry straight into the solver's arrays; zones from the reference's own
// computeZones (gasphase.cpp:111-158).  Returns the zone count.
int ref_setup(void* h, int nx, int ny, int nz, float dx, const float* fx, const float* fy, const float* fz,
              const float* ax, const float* ay, const float* az, const uint8_t* blocked) {
    Solver* s = static_cast<Solver*>(h);
    s->nx = nx; s->ny = ny; s->nz = nz; s->N = nx * ny * nz;
    s->dx = dx; s->ox = s->oy = s->oz = 0;
    s->wind.speed = 0;
    const size_t N = (size_t)s->N, nfx = (size_t)(nx + 1) * ny * nz, nfy = (size_t)nx * (ny + 1) * nz,
                 nfz = (size_t)nx * ny * (nz + 1);
    s->fx.assign(fx, fx + nfx); s->fy.assign(fy, fy + nfy); s->fz.assign(fz, fz + nfz);
    s->ax_.assign(ax, ax + nfx); s->ay_.assign(ay, ay + nfy); s->az_.assign(az, az + nfz);
    s->lx_.assign(nfx, 0); s->ly_.assign(nfy, 0); s->lz_.assign(nfz, 0);
    s->u.assign(nfx, 0.0f); s->v.assign(nfy, 0.0f); s->w.assign(nfz, 0.0f);
    s->blocked.assign(blocked, blocked + N);
    s->p.assign(N, 0.0f); s->S.assign(N, 0.0f);
    s->T.assign(N, Solver::T0);
    s->Yo.assign(N, Solver::YO2_AMB); s->Yf.assign(N, 0.0f); s->Ys.assign(N, 0.0f); s->Ya.assign(N, 0.0f);
    s->Yp.assign(N, 0.0f); s->Yc.assign(N, Solver::YCO2_AMB); s->Yw.assign(N, Solver::YH2O_AMB);
    s->Yco.assign(N, 0.0f); s->Yai.assign(N, 0.0f);
    s->nu.assign(N, 0.0f); s->Q.assign(N, 0.0f); s->massSrc.assign(N, 0.0f);
    s->Pb.assign(N, Solver::P0);
    s->rhs_.assign(N, 0.0f);
    for (auto& t : s->tmp_) t.assign(N, 0.0f);
    s->cgR_.assign(N, 0.0f); s->cgZ_.assign(N, 0.0f); s->cgP_.assign(N, 0.0f); s->cgQ_.assign(N, 0.0f);
    s->Qr.assign(N, 0.0f); s->hrr.assign(N, 0.0f); s->lit.assign(N, 0); s->pilot.assign(N, 0);
    s->u2_.assign(nfx, 0.0f); s->v2_.assign(nfy, 0.0f); s->w2_.assign(nfz, 0.0f);
    // a gas-only world: no radiation, no surface exchange, no solids or barriers
    // (the Phase 3 comparison switches, as buildRoom in test_main.cpp:420-433)
    s->radiationOn = false;
    s->adiabaticSurfaces = true;
    s->zone.assign(N, -1);
    s->computeZones();
    s->mgDirty_ = true;
    return (int)s->zoneP.size();
}

int ref_set_bytes(void* h, const char* name, const uint8_t* data) {
    std::vector<uint8_t>* f = bytes(static_cast<Solver*>(h), name);
    if (!f) return -1;
    std::copy(data, data + f->size(), f->begin());
    return (int)f->size();
}

int ref_get_bytes(void* h, const char* name, uint8_t* out) {
    std::vector<uint8_t>* f = bytes(static_cast<Solver*>(h), name);
    if (!f) return -1;
    std::copy(f->begin(), f->end(), out);
    return (int)f->size();
}

// the private phases of Solver::step (solver.cpp:226-269), one at a time
float ref_turbulence(void* h) {
    Solver* s = static_cast<Solver*>(h);
    s->turbulence();
    return s->maxNu;
}
void ref_heat_diffusion(void* h, float dt) { static_cast<Solver*>(h)->heatDiffusion(dt); }
void ref_momentum(void* h, float dt) { static_cast<Solver*>(h)->momentum(dt); }
void ref_inject_burners(void* h, float dt) { static_cast<Solver*>(h)->injectBurners(dt); }
void ref_combustion(void* h, float dt) { static_cast<Solver*>(h)->combustion(dt); }
void ref_leakage(void* h, float dt) { static_cast<Solver*>(h)->leakage(dt); }

// stableDt (gasphase.cpp:63-70) from the given step statistics and profile
float ref_stable_dt(void* h, float dtMax, float maxVel, float maxNu, float maxDiv, float cfl, float profileDtMax) {
    Solver* s = static_cast<Solver*>(h);
    s->maxVel = maxVel;
    s->maxNu = maxNu;
    s->maxDiv = maxDiv;
    s->profile.cfl = cfl;
    s->profile.dtMax = profileDtMax;
    return s->stableDt(dtMax);
}

// a burner over a floor rectangle (Solver::addBurner, gasphase.cpp:72-90)
int ref_add_burner(void* h, float x0, float y0, float x1, float y1, float z, float peakW, float rampS) {
    photon::Burner b;
    b.x0 = x0; b.y0 = y0; b.x1 = x1; b.y1 = y1; b.z = z; b.peakW = peakW; b.rampS = rampS;
    b.label = "burner";
    return static_cast<Solver*>(h)->addBurner(b);
}
int ref_burner_cells(void* h, int b, int* cells) {
    const std::vector<int>& c = static_cast<Solver*>(h)->burners[b].cells;
    if (cells) std::copy(c.begin(), c.end(), cells);
    return (int)c.size();
}

// the whole step with the solver's stop rule (profile); returns dt
float ref_step(void* h, float dtMax, float tol, int maxIters) {
    Solver* s = static_cast<Solver*>(h);
    s->profile.pressureTol = tol;
    s->profile.pressureMaxIters = maxIters;
    return s->step(dtMax);
}

// scalars of the last step: time, maxVel, maxNu, maxDiv, pressureResidual,
// hrr; pressureIters as the return value
int ref_scalars(void* h, float* out6) {
    Solver* s = static_cast<Solver*>(h);
    out6[0] = s->time; out6[1] = s->maxVel; out6[2] = s->maxNu; out6[3] = s->maxDiv;
    out6[4] = s->pressureResidual; out6[5] = s->stats.hrr;
    return s->pressureIters;
}

// zone background pressures and their rise rates of the last step
int ref_zone_pressure(void* h, float* zoneP, float* zoneDP) {
    Solver* s = static_cast<Solver*>(h);
    if (zoneP) std::copy(s->zoneP.begin(), s->zoneP.end(), zoneP);
    if (zoneDP) {
        for (size_t z = 0; z < s->zoneP.size(); ++z) zoneDP[z] = z < s->zoneDP_.size() ? s->zoneDP_[z] : 0.0f;
    }
    return (int)s->zoneP.size();
}

// set the zone background pressures (and each zone's Pb), as a step leaves
// them (solver.cpp:258-263); returns the zone count
int ref_set_zone_pressure(void* h, const float* zoneP) {
    Solver* s = static_cast<Solver*>(h);
    std::copy(zoneP, zoneP + s->zoneP.size(), s->zoneP.begin());
    for (int c = 0; c < s->N; ++c)
        if (s->zone[c] >= 0) s->Pb[c] = s->zoneP[s->zone[c]];
    return (int)s->zoneP.size();
}

// Face changes as Solver::updateFace leaves them (solver.cpp:88-100): open
// fraction, matrix coefficient and crack flag written, the velocity of a
// closed face zeroed, zonesDirty_ set when the coefficient or the flag
// changed (the next step recomputes the zones, which marks mgDirty_).
void ref_update_faces(void* h, int n, const int* axis, const int* face, const float* f, const float* a,
                      const uint8_t* crack) {
    Solver* s = static_cast<Solver*>(h);
    for (int q = 0; q < n; ++q) {
        const int d = axis[q], e = face[q];
        std::vector<float>& fo = d == 0 ? s->fx : d == 1 ? s->fy : s->fz;
        std::vector<float>& av = d == 0 ? s->ax_ : d == 1 ? s->ay_ : s->az_;
        std::vector<uint8_t>& lv = d == 0 ? s->lx_ : d == 1 ? s->ly_ : s->lz_;
        std::vector<float>& vel = d == 0 ? s->u : d == 1 ? s->v : s->w;
        fo[e] = f[q];
        if (f[q] == 0) vel[e] = 0;
        if (av[e] != a[q] || lv[e] != crack[q]) s->zonesDirty_ = true;
        av[e] = a[q];
        lv[e] = crack[q];
    }
}

// crack faces the last leakage handled (Solver::leakFaces, solver.h:279)
int ref_leak_faces(void* h) { return static_cast<Solver*>(h)->leakFaces; }

// Wind at the open boundaries (Solver::wind, solver.h:106-111); the
// hierarchy is rebuilt on the next project (its level 0 masks wind sides,
// pressure_mg.cpp:31-46).
void ref_set_wind(void* h, float speed, float fromDeg, float exponent) {
    Solver* s = static_cast<Solver*>(h);
    s->wind.speed = speed;
    s->wind.fromDeg = fromDeg;
    s->wind.exponent = exponent;
    s->mgDirty_ = true;
}

// windAt (gasphase.cpp:52-61): the inflow velocity on a boundary face, NaN if none
float ref_wind_at(void* h, int axis, int side, float z) { return static_cast<Solver*>(h)->windAt(axis, side, z); }

void ref_get_zones(void* h, int* zone, uint8_t* zoneOpen) {
    Solver* s = static_cast<Solver*>(h);
    std::copy(s->zone.begin(), s->zone.end(), zone);
    std::copy(s->zoneOpen.begin(), s->zoneOpen.end(), zoneOpen);
}

int ref_set_field(void* h, const char* name, const float* data) {
    std::vector<float>* f = field(static_cast<Solver*>(h), name);
    if (!f) return -1;
    std::copy(data, data + f->size(), f->begin());
    return (int)f->size();
}

int ref_get_field(void* h, const char* name, float* out) {
    std::vector<float>* f = field(static_cast<Solver*>(h), name);
    if (!f) return -1;
    std::copy(f->begin(), f->end(), out);
    return (int)f->size();
}

// project(dt) with the given stop rule parameters (solver.h:89-90): returns iterations.
int ref_project(void* h, float dt, float tol, int maxIters, float* residual, float* maxVel) {
    Solver* s = static_cast<Solver*>(h);
    s->profile.pressureTol = tol;
    s->profile.pressureMaxIters = maxIters;
    s->project(dt);
    if (residual) *residual = s->pressureResidual;
    if (maxVel) *maxVel = s->maxVel;
    return s->pressureIters;
}

void ref_precondition(void* h, const float* r, float* z) {
    Solver* s = static_cast<Solver*>(h);
    std::vector<float> rv(r, r + s->N), zv(s->N, 0.0f);
    s->pressurePrecondition(rv, zv);
    std::copy(zv.begin(), zv.end(), z);
}

static void ensure_levels(Solver* s) {
    if (s->mgDirty_ || s->mg_.empty()) s->buildPressureLevels();
}

int ref_mg_nlevels(void* h) {
    Solver* s = static_cast<Solver*>(h);
    ensure_levels(s);
    return (int)s->mg_.size();
}
int ref_mg_level_n(void* h, int l) { return static_cast<Solver*>(h)->mg_[l].n; }
int ref_mg_level_nnz(void* h, int l) { return (int)static_cast<Solver*>(h)->mg_[l].col.size(); }
int ref_mg_level_has_agg(void* h, int l) { return static_cast<Solver*>(h)->mg_[l].agg.empty() ? 0 : 1; }
void ref_mg_level_get(void* h, int l, int* rowStart, int* col, float* val, float* diag, int* agg) {
    const photon::PressureLevel& L = static_cast<Solver*>(h)->mg_[l];
    if (rowStart && !L.rowStart.empty()) std::copy(L.rowStart.begin(), L.rowStart.end(), rowStart);
    if (col) std::copy(L.col.begin(), L.col.end(), col);
    if (val) std::copy(L.val.begin(), L.val.end(), val);
    if (diag) std::copy(L.diag.begin(), L.diag.end(), diag);
    if (agg) {
        if (L.agg.empty()) std::fill(agg, agg + L.n, -1);
        else std::copy(L.agg.begin(), L.agg.end(), agg);
    }
}

float ref_energy(void* h, float dt) {
    Solver* s = static_cast<Solver*>(h);
    s->energyAndExpansion(dt);
    return s->maxDiv;
}

void ref_advect(void* h, float dt) { static_cast<Solver*>(h)->advectScalars(dt); }

int ref_threads(void* h) { return static_cast<Solver*>(h)->pool_->threadCount(); }

}  // extern "C"
