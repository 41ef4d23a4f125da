// gas_phase.cpp -- photon_gpu::GasPhase over the C ABI of libphoton_gas.so.
// Each member names the reference function it stands for.
#include "gas_phase.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>

namespace photon_gpu {

namespace {
constexpr float T0 = 293.15f, P0 = 101325.0f, RS = 287.0f, CP = 1100.0f;
constexpr float DHC_FUEL = 25.0e6f;       // solver.h:165
constexpr float BURNER_SOOT = 0.01f;      // gasphase.cpp:93
constexpr float T_AUTO = 800.0f;          // solver.h:179
}  // namespace

float Burner::powerAt(float t) const {   // solver.h:121-125
    if (t < start || t > start + duration) return 0;
    float r = rampS > 0 ? std::min(1.0f, (t - start) / rampS) : 1.0f;
    return peakW * r * r;
}

struct GasPhase::Copy {
    void* dev;
    void* host;
    size_t bytes;
};

GasPhase::~GasPhase() {
    if (ctx_) {
        for (void* p : dmem_) phg_free(ctx_, p);
        if (grid_ >= 0) phg_grid_destroy(ctx_, grid_);
        phg_destroy(ctx_);
    }
}

bool GasPhase::check(int rc, const char* what) {
    if (rc == PHG_OK) return true;
    if (err_.empty()) err_ = std::string(what) + ": " + (ctx_ ? phg_last_error(ctx_) : "no context");
    return false;
}

template <class T>
T* GasPhase::alloc(size_t n) {
    void* p = nullptr;
    if (!check(phg_malloc(ctx_, &p, std::max<size_t>(n, 1) * sizeof(T)), "phg_malloc")) return nullptr;
    dmem_.push_back(p);
    return static_cast<T*>(p);
}

phg_gas GasPhase::gas() const {
    phg_gas g;
    g.T = dF_[0];
    g.Yo = dF_[1];
    g.Yf = dF_[2];
    g.Ys = dF_[3];
    g.Ya = dF_[4];
    g.Yp = dF_[5];
    g.Yc = dF_[6];
    g.Yw = dF_[7];
    g.Yco = dF_[8];
    g.Yai = dF_[9];
    g.Pb = dPb_;
    g.massSrc = dMs_;
    g.Q = dQ_;
    g.Qr = dQr_;
    g.hrr = dHrr_;
    g.pilot = dPilot_;
    g.lit = dLit_;
    return g;
}

// the host vectors and their device copies, for syncToHost / syncToDevice
std::vector<GasPhase::Copy> GasPhase::copies() const {
    auto* self = const_cast<GasPhase*>(this);
    std::vector<float>* F[10] = {&self->T, &self->Yo, &self->Yf, &self->Ys, &self->Ya,
                                 &self->Yp, &self->Yc, &self->Yw, &self->Yco, &self->Yai};
    std::vector<Copy> c = {{dU_, self->u.data(), u.size() * 4},   {dV_, self->v.data(), v.size() * 4},
                           {dW_, self->w.data(), w.size() * 4},   {dP_, self->p.data(), p.size() * 4},
                           {dS_, self->S.data(), S.size() * 4},   {dNu_, self->nu.data(), nu.size() * 4},
                           {dPb_, self->Pb.data(), Pb.size() * 4}, {dQ_, self->Q.data(), Q.size() * 4},
                           {dQr_, self->Qr.data(), Qr.size() * 4}, {dMs_, self->massSrc.data(), massSrc.size() * 4},
                           {dHrr_, self->hrr.data(), hrr.size() * 4}, {dPilot_, self->pilot.data(), pilot.size()},
                           {dLit_, self->lit.data(), lit.size()}};
    for (int q = 0; q < 10; ++q) c.push_back({dF_[q], F[q]->data(), F[q]->size() * 4});
    return c;
}

void GasPhase::syncToHost() {
    for (const Copy& c : copies()) check(phg_memcpy_d2h(ctx_, c.host, c.dev, c.bytes), "syncToHost");
}
void GasPhase::syncToDevice() {
    for (const Copy& c : copies()) check(phg_memcpy_h2d(ctx_, c.dev, c.host, c.bytes), "syncToDevice");
}

namespace {
constexpr char kStateMagic[8] = {'P', 'H', 'G', 'S', 'T', 'A', 'T', '1'};

// FNV-1a over the geometry a state belongs to
uint64_t geometry_hash(const std::vector<float>& fx, const std::vector<float>& fy, const std::vector<float>& fz,
                       const std::vector<uint8_t>& blocked, const std::vector<int>& zone) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    };
    mix(fx.data(), fx.size() * 4);
    mix(fy.data(), fy.size() * 4);
    mix(fz.data(), fz.size() * 4);
    mix(blocked.data(), blocked.size());
    mix(zone.data(), zone.size() * 4);
    return h;
}

struct Writer {
    std::ofstream out;
    template <class T>
    void pod(const T& v) { out.write(reinterpret_cast<const char*>(&v), sizeof(T)); }
    template <class T>
    void vec(const std::vector<T>& v) {
        const uint64_t n = v.size();
        pod(n);
        out.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(n * sizeof(T)));
    }
};

struct Reader {
    std::ifstream in;
    template <class T>
    bool pod(T& v) { return (bool)in.read(reinterpret_cast<char*>(&v), sizeof(T)); }
    template <class T>
    bool vec(std::vector<T>& v, size_t expect) {
        uint64_t n = 0;
        if (!pod(n) || (expect != (size_t)-1 && n != expect) || n > (1ull << 34)) return false;
        v.resize(n);
        return (bool)in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(T)));
    }
};
}  // namespace

bool GasPhase::saveState(const std::string& path, std::string& err) {
    syncToHost();
    if (!ok()) {
        err = err_;
        return false;
    }
    Writer wr;
    wr.out.open(path, std::ios::binary);
    if (!wr.out) {
        err = "cannot write " + path;
        return false;
    }
    wr.out.write(kStateMagic, 8);
    wr.pod(nx);
    wr.pod(ny);
    wr.pod(nz);
    wr.pod(dx);
    wr.pod(geometry_hash(fx, fy, fz, blocked, zone));
    wr.pod(time);
    wr.pod(maxVel);
    wr.pod(maxNu);
    wr.pod(maxDiv);
    wr.pod(pressureIters);
    wr.pod(pressureResidual);
    wr.vec(zoneP);
    wr.vec(zoneDP);
    for (const std::vector<float>* f : {&u, &v, &w, &p, &S, &nu, &Pb, &Q, &Qr, &massSrc, &hrr, &T, &Yo, &Yf, &Ys, &Ya,
                                        &Yp, &Yc, &Yw, &Yco, &Yai})
        wr.vec(*f);
    wr.vec(pilot);
    wr.vec(lit);
    wr.pod((uint64_t)burners.size());
    for (const Burner& b : burners) {
        for (float x : {b.x0, b.y0, b.x1, b.y1, b.z, b.peakW, b.rampS, b.start, b.duration}) wr.pod(x);
        wr.vec(b.cells);
    }
    if (!wr.out) {
        err = "write failed: " + path;
        return false;
    }
    return true;
}

bool GasPhase::loadState(const std::string& path, std::string& err) {
    Reader r;
    r.in.open(path, std::ios::binary);
    char magic[8];
    if (!r.in || !r.in.read(magic, 8) || std::memcmp(magic, kStateMagic, 8) != 0) {
        err = "not a gas state file: " + path;
        return false;
    }
    int sx = 0, sy = 0, sz = 0;
    float sdx = 0;
    uint64_t hash = 0;
    r.pod(sx);
    r.pod(sy);
    r.pod(sz);
    r.pod(sdx);
    r.pod(hash);
    if (sx != nx || sy != ny || sz != nz || sdx != dx || hash != geometry_hash(fx, fy, fz, blocked, zone)) {
        err = "the state was saved from another geometry";
        return false;
    }
    const size_t nfx = u.size(), nfy = v.size(), nfz = w.size(), n = (size_t)N, nzn = zoneP.size();
    bool good = r.pod(time) && r.pod(maxVel) && r.pod(maxNu) && r.pod(maxDiv) && r.pod(pressureIters) &&
                r.pod(pressureResidual) && r.vec(zoneP, nzn) && r.vec(zoneDP, nzn) && r.vec(u, nfx) && r.vec(v, nfy) &&
                r.vec(w, nfz);
    for (std::vector<float>* f : {&p, &S, &nu, &Pb, &Q, &Qr, &massSrc, &hrr, &T, &Yo, &Yf, &Ys, &Ya, &Yp, &Yc, &Yw,
                                  &Yco, &Yai})
        good = good && r.vec(*f, n);
    good = good && r.vec(pilot, n) && r.vec(lit, n);
    uint64_t nb = 0;
    good = good && r.pod(nb) && nb < (1u << 20);
    if (!good) {
        err = "truncated or malformed state file: " + path;
        return false;
    }
    for (int* d : dBurnerCells_) phg_free(ctx_, d), dmem_.erase(std::find(dmem_.begin(), dmem_.end(), (void*)d));
    dBurnerCells_.clear();
    burners.clear();
    for (uint64_t i = 0; i < nb; ++i) {
        Burner b;
        float* fields[9] = {&b.x0, &b.y0, &b.x1, &b.y1, &b.z, &b.peakW, &b.rampS, &b.start, &b.duration};
        for (float* x : fields) good = good && r.pod(*x);
        good = good && r.vec(b.cells, (size_t)-1);
        for (int c : b.cells) good = good && c >= 0 && c < N;
        if (!good) {
            err = "malformed burner in " + path;
            return false;
        }
        int* d = alloc<int>(b.cells.size());
        if (!d || !check(phg_memcpy_h2d(ctx_, d, b.cells.data(), b.cells.size() * sizeof(int)), "burner cells")) break;
        dBurnerCells_.push_back(d);
        burners.push_back(b);
    }
    syncToDevice();
    err = err_;
    return ok();
}

bool GasPhase::build(int nx_, int ny_, int nz_, float dx_, const std::vector<float>& fx_,
                     const std::vector<float>& fy_, const std::vector<float>& fz_, const std::vector<float>& ax,
                     const std::vector<float>& ay, const std::vector<float>& az, const std::vector<uint8_t>& blocked_,
                     std::string& err, int device, bool refCoarse, int precision) {
    nx = nx_;
    ny = ny_;
    nz = nz_;
    N = nx * ny * nz;
    dx = dx_;
    const size_t nfx = (size_t)(nx + 1) * ny * nz, nfy = (size_t)nx * (ny + 1) * nz, nfz = (size_t)nx * ny * (nz + 1);
    if (fx_.size() != nfx || fy_.size() != nfy || fz_.size() != nfz || ax.size() != nfx || ay.size() != nfy ||
        az.size() != nfz || blocked_.size() != (size_t)N) {
        err = "geometry arrays do not match the grid";
        return false;
    }
    fx = fx_;
    fy = fy_;
    fz = fz_;
    blocked = blocked_;
    if (phg_create(&ctx_, device) != PHG_OK) {
        err = "phg_create failed: no usable GPU";
        return false;
    }
    // sides where windAt prescribes inflow close for the pressure (gasphase.cpp:510)
    uint8_t wc[4] = {0, 0, 0, 0};
    if (wind.speed > 0) {
        const float th = wind.fromDeg * 3.14159265f / 180.0f;
        const float sx = -std::sin(th), sy = -std::cos(th);
        wc[0] = sx > 0;
        wc[1] = sx < 0;
        wc[2] = sy > 0;
        wc[3] = sy < 0;
    }
    if (!check(phg_grid_create(ctx_, nx, ny, nz, dx, fx.data(), fy.data(), fz.data(), ax.data(), ay.data(), az.data(),
                               blocked.data(), wind.speed > 0 ? wc : nullptr, precision, &grid_),
               "phg_grid_create") ||
        (refCoarse && !check(phg_grid_set_coarse(ctx_, grid_, PHG_COARSE_GS20), "phg_grid_set_coarse"))) {
        err = err_;
        return false;
    }
    // zones (computeZones, gasphase.cpp:111-158)
    int nl = 0, nzn = 0, rl = 0;
    int lvn[16];
    check(phg_grid_info(ctx_, grid_, &nl, lvn, &nzn, &rl), "phg_grid_info");
    zone.assign(N, -1);
    zoneOpen.assign(std::max(nzn, 1), 0);
    check(phg_grid_zones(ctx_, grid_, zone.data(), zoneOpen.data()), "phg_grid_zones");
    zoneOpen.resize(nzn);
    zoneP.assign(nzn, P0);
    zoneDP.assign(nzn, 0.0f);
    // state at ambient, still air (as Solver::build leaves the gas)
    u.assign(nfx, 0.0f);
    v.assign(nfy, 0.0f);
    w.assign(nfz, 0.0f);
    for (auto* f : {&p, &S, &nu, &Q, &Qr, &massSrc, &hrr, &Yf, &Ys, &Ya, &Yp, &Yco, &Yai}) f->assign(N, 0.0f);
    T.assign(N, T0);
    Yo.assign(N, 0.232f);
    Yc.assign(N, 0.0006f);
    Yw.assign(N, 0.0073f);
    Pb.assign(N, P0);
    pilot.assign(N, 0);
    lit.assign(N, 0);
    dU_ = alloc<float>(nfx);
    dV_ = alloc<float>(nfy);
    dW_ = alloc<float>(nfz);
    dU2_ = alloc<float>(nfx);
    dV2_ = alloc<float>(nfy);
    dW2_ = alloc<float>(nfz);
    for (float** d : {&dP_, &dS_, &dNu_, &dPb_, &dQ_, &dQr_, &dMs_, &dHrr_}) *d = alloc<float>(N);
    for (int q = 0; q < 10; ++q) {
        dF_[q] = alloc<float>(N);
        dFt_[q] = alloc<float>(N);
    }
    dPilot_ = alloc<uint8_t>(N);
    dLit_ = alloc<uint8_t>(N);
    if (!err_.empty()) {
        err = err_;
        return false;
    }
    syncToDevice();
    err = err_;
    return err_.empty();
}

int GasPhase::addBurner(const Burner& b0) {   // gasphase.cpp:72-90
    Burner b = b0;
    b.cells.clear();
    auto cellAt = [&](float x, float y, float z) {
        int i = (int)std::floor((x - ox) / dx), j = (int)std::floor((y - oy) / dx), k = (int)std::floor((z - oz) / dx);
        if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) return -1;
        return i + nx * (j + ny * k);
    };
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            float x = ox + (i + 0.5f) * dx, y = oy + (j + 0.5f) * dx;
            if (x < b.x0 || x > b.x1 || y < b.y0 || y > b.y1) continue;
            int c = cellAt(x, y, b.z + 0.5f * dx);
            if (c >= 0 && !blocked[c]) b.cells.push_back(c);
        }
    if (b.cells.empty()) {
        int c = cellAt(0.5f * (b.x0 + b.x1), 0.5f * (b.y0 + b.y1), b.z + 0.5f * dx);
        if (c < 0 || blocked[c]) return -1;
        b.cells.push_back(c);
    }
    int* d = alloc<int>(b.cells.size());
    if (!d || !check(phg_memcpy_h2d(ctx_, d, b.cells.data(), b.cells.size() * sizeof(int)), "burner cells")) return -1;
    dBurnerCells_.push_back(d);
    burners.push_back(b);
    return (int)burners.size() - 1;
}

float GasPhase::stableDt(float dtMax) const {
    float dt = dtMax;
    phg_stable_dt(dtMax, profile.dtMax, profile.cfl, dx, maxVel, maxNu, maxDiv, &dt);
    return dt;
}

void GasPhase::stepClear() {
    const phg_gas g = gas();
    check(phg_step_clear(ctx_, grid_, &g), "phg_step_clear");
}

void GasPhase::injectBurners(float dt) {   // gasphase.cpp:92-105
    const phg_gas g = gas();
    for (size_t bi = 0; bi < burners.size(); ++bi) {
        const Burner& b = burners[bi];
        float P = b.powerAt(time);
        if (P <= 0) continue;
        float mf = P * dt / DHC_FUEL / b.cells.size();
        check(phg_add_gas(ctx_, grid_, (int)b.cells.size(), dBurnerCells_[bi], mf, 0.0f, mf * BURNER_SOOT, T_AUTO, 1,
                          &g),
              "phg_add_gas");
    }
}

void GasPhase::combustion(float dt) {
    const phg_gas g = gas();
    check(phg_combustion(ctx_, grid_, dt, radiationOn ? 1 : 0, &g, &hrrTotal), "phg_combustion");
}

void GasPhase::turbulence() {
    check(phg_turbulence(ctx_, grid_, dU_, dV_, dW_, dF_[0], dNu_, deferStats_ ? nullptr : &maxNu), "phg_turbulence");
}

void GasPhase::heatDiffusion(float dt) {
    check(phg_heat_diffusion(ctx_, grid_, dt, dF_[0], dNu_, dPb_, dQ_), "phg_heat_diffusion");
}

void GasPhase::energyAndExpansion(float dt) {
    check(phg_energy_expansion(ctx_, grid_, dt, dF_[0], dS_, dPb_, dQ_, dMs_, deferStats_ ? nullptr : &maxDiv),
          "phg_energy_expansion");
}

void GasPhase::momentum(float dt) {
    const phg_wind wd{wind.speed, wind.fromDeg, wind.exponent, oz};
    if (check(phg_momentum(ctx_, grid_, dt, &wd, dU_, dV_, dW_, dF_[0], dPb_, dNu_, dP_, dU2_, dV2_, dW2_),
              "phg_momentum")) {
        std::swap(dU_, dU2_);
        std::swap(dV_, dV2_);
        std::swap(dW_, dW2_);
    }
}

bool GasPhase::setCracks(const std::vector<uint8_t>& lx, const std::vector<uint8_t>& ly,
                         const std::vector<uint8_t>& lz) {
    const size_t nfx = (size_t)(nx + 1) * ny * nz, nfy = (size_t)nx * (ny + 1) * nz, nfz = (size_t)nx * ny * (nz + 1);
    if ((!lx.empty() && lx.size() != nfx) || (!ly.empty() && ly.size() != nfy) || (!lz.empty() && lz.size() != nfz)) {
        err_ = "crack arrays do not match the grid";
        return false;
    }
    return check(phg_grid_set_cracks(ctx_, grid_, lx.empty() ? nullptr : lx.data(), ly.empty() ? nullptr : ly.data(),
                                     lz.empty() ? nullptr : lz.data()),
                 "phg_grid_set_cracks");
}

bool GasPhase::updateFaces(const std::vector<FaceUpdate>& faces) {
    const int n = (int)faces.size();
    std::vector<int> axis(n), face(n);
    std::vector<float> fo(n), co(n);
    std::vector<uint8_t> cr(n);
    for (int q = 0; q < n; ++q) {
        axis[q] = faces[q].axis;
        face[q] = faces[q].face;
        fo[q] = faces[q].open;
        co[q] = faces[q].coef;
        cr[q] = faces[q].crack ? 1 : 0;
    }
    int rebuilt = 0;
    if (!check(phg_grid_update_faces(ctx_, grid_, n, axis.data(), face.data(), fo.data(), co.data(), cr.data(), dU_,
                                     dV_, dW_, &rebuilt),
               "phg_grid_update_faces"))
        return false;
    for (int q = 0; q < n; ++q) {
        std::vector<float>& f = axis[q] == 0 ? fx : axis[q] == 1 ? fy : fz;
        f[face[q]] = fo[q];
    }
    if (!rebuilt) return true;
    // computeZones: closed zones at their cells' mean background pressure
    int nl = 0, nzn = 0, rl = 0;
    int lvn[16];
    if (!check(phg_grid_info(ctx_, grid_, &nl, lvn, &nzn, &rl), "phg_grid_info")) return false;
    zoneOpen.assign(std::max(nzn, 1), 0);
    if (!check(phg_grid_zones(ctx_, grid_, zone.data(), zoneOpen.data()), "phg_grid_zones")) return false;
    zoneOpen.resize(nzn);
    std::vector<float> pb(N);
    if (!check(phg_memcpy_d2h(ctx_, pb.data(), dPb_, N * sizeof(float)), "phg_memcpy_d2h")) return false;
    std::vector<double> pSum(nzn, 0.0), vSum(nzn, 0.0);
    for (int c = 0; c < N; ++c)
        if (zone[c] >= 0) {
            pSum[zone[c]] += pb[c];
            vSum[zone[c]] += 1;
        }
    zoneP.assign(nzn, P0);
    zoneDP.assign(nzn, 0.0f);
    for (int z = 0; z < nzn; ++z)
        if (!zoneOpen[z]) zoneP[z] = (float)(pSum[z] / std::max(1.0, vSum[z]));
    for (int c = 0; c < N; ++c) pb[c] = zone[c] >= 0 ? zoneP[zone[c]] : P0;
    Pb = pb;
    return check(phg_memcpy_h2d(ctx_, dPb_, pb.data(), N * sizeof(float)), "phg_memcpy_h2d");
}

void GasPhase::leakage(float dt) {   // gasphase.cpp:402-498
    zoneDP.assign(zoneP.size(), 0.0f);
    if (zoneP.empty()) return;
    check(phg_zone_leakage(ctx_, grid_, dt, zoneP.data(), dF_[0], dPb_, dP_, dU_, dV_, dW_, dS_, zoneDP.data(),
                           &leakFaces),
          "phg_zone_leakage");
}

void GasPhase::project(float dt) {   // gasphase.cpp:505-677
    phg_solve_params prm{};
    prm.tol = profile.pressureTol;
    prm.max_iters = profile.pressureMaxIters;
    prm.stop_rule = PHG_STOP_MAXNORM;
    prm.demean = 1;
    prm.coarse_weight = profile.pressureCoarseWeight;
    // the set-up and iterations as one graph sized from the last step's count
    prm.use_graph = 2;
    prm.check_every = pressureIters > 0 ? (pressureIters + 4) / 4 * 4 : 8;
    phg_solve_stats st{};
    float vmax = 0;
    if (!check(phg_project(ctx_, grid_, dt, dU_, dV_, dW_, dP_, dS_, &prm, &st, &vmax), "phg_project")) {
        pressureIters = -1;   // step() has no error path: flag it
        pressureResidual = NAN;
        return;
    }
    pressureIters = st.iters;
    pressureResidual = (float)st.maxres;
    maxVel = vmax;
}

void GasPhase::advectScalars(float dt) {   // gasphase.cpp:685-784; tmp_ swapped in
    if (check(phg_advect_scalars(ctx_, grid_, dt, dU_, dV_, dW_, dPb_, dNu_, dF_, dFt_), "phg_advect_scalars"))
        for (int q = 0; q < 10; ++q) std::swap(dF_[q], dFt_[q]);
}

void GasPhase::zonePressure(float dt) {   // solver.cpp:258-263
    bool any = false;
    for (size_t z = 0; z < zoneP.size(); ++z)
        if (!zoneOpen[z]) {
            zoneP[z] = std::clamp(zoneP[z] + dt * zoneDP[z], P0 - 20000.0f, P0 + 50000.0f);
            any = true;
        }
    if (any) check(phg_zone_values(ctx_, grid_, zoneP.data(), dPb_, 0), "phg_zone_values");
}

float GasPhase::step(float dtMax) {   // solver.cpp:226-269, gas-only
    float dt = stableDt(dtMax);
    stepClear();
    injectBurners(dt);
    combustion(dt);
    // maxNu and maxDiv stay on the device until the solve has waited
    deferStats_ = true;
    turbulence();
    heatDiffusion(dt);
    energyAndExpansion(dt);
    deferStats_ = false;
    momentum(dt);
    leakage(dt);
    project(dt);
    check(phg_grid_stats(ctx_, grid_, nullptr, &maxNu, &maxDiv), "phg_grid_stats");
    advectScalars(dt);
    zonePressure(dt);
    time += dt;
    return dt;
}

}  // namespace photon_gpu
