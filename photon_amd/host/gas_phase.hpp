// gas_phase.hpp -- a C++20 host class with the gas-phase members of the
// reference's photon::Solver (src/sim/solver.h:156-347), whose phases run on
// an MI355X through the C ABI of include/photon_gas.h.
//
// It is what a Solver maintainer binds: the state lives in device memory for
// the whole run, the host vectors of the same names (u, v, w, p, T, Yo, ...)
// are filled on syncToHost() and pushed with syncToDevice(), and step()
// follows Solver::step (src/sim/solver.cpp:226-269) for a gas-only world (no
// radiation, solids or barriers: the reference's Phase 3 comparison
// switches).  No HIP types appear here; the class links libphoton_gas.so.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/photon_gas.h"

namespace photon_gpu {

// Solver::wind (solver.h:106-111)
struct Wind {
    float speed = 0;
    float fromDeg = 0;
    float exponent = 0.25f;
};

// One face's state as Solver::updateFace leaves it (solver.cpp:66-101)
struct FaceUpdate {
    int axis = 0, face = 0;
    float open = 1.0f;        // open fraction (fx/fy/fz)
    float coef = 1.0f;        // pressure-matrix coefficient (ax_/ay_/az_)
    bool crack = false;       // lx_/ly_/lz_
};

// Solver's Burner (solver.h:113-127): fuel gas over a floor rectangle
struct Burner {
    float x0 = 0, y0 = 0, x1 = 0, y1 = 0, z = 0;
    float peakW = 0, rampS = 0, start = 0, duration = 1e9f;
    std::vector<int> cells;
    float powerAt(float t) const;
};

// the profile fields the gas step reads (FidelityProfile, solver.h:82-104)
struct Profile {
    float cfl = 0.8f;
    float dtMax = 0.05f;
    int pressureMaxIters = 80;
    float pressureTol = 0.02f;   // 1/s, max-norm stop rule
    // weight of the V-cycle's coarse-grid corrections (phg_solve_params.coarse_weight):
    // 0 keeps the reference's cycle and its pressureIters; 1.8 takes about a third fewer
    // iterations to the same pressureTol (no counterpart in the reference's profile)
    float pressureCoarseWeight = 0.0f;
};

class GasPhase {
public:
    GasPhase() = default;
    ~GasPhase();
    GasPhase(const GasPhase&) = delete;
    GasPhase& operator=(const GasPhase&) = delete;

    // Geometry in the reference's layout (solver.h:216-220): face open
    // fractions, matrix coefficients (ax_/ay_/az_), solid mask.  `refCoarse`
    // selects the reference's coarsest Gauss-Seidel sweeps, so the pressure
    // iterates follow the reference's.  Returns false with a message.
    bool build(int nx, int ny, int nz, float dx, const std::vector<float>& fx, const std::vector<float>& fy,
               const std::vector<float>& fz, const std::vector<float>& ax, const std::vector<float>& ay,
               const std::vector<float>& az, const std::vector<uint8_t>& blocked, std::string& err, int device = 0,
               bool refCoarse = true, int precision = PHG_PREC_FP32);

    // Crack faces (Solver::lx_, ly_, lz_, solver.h:297): one byte per face of
    // each axis, empty for none on that axis.  Returns false on error.
    bool setCracks(const std::vector<uint8_t>& lx, const std::vector<uint8_t>& ly, const std::vector<uint8_t>& lz);

    // Face changes between steps (barrier failure, door toggles): the grid's
    // faces and cracks are updated; when the pressure matrix changes the
    // zones are computed again (closed zones take their cells' mean Pb,
    // computeZones, gasphase.cpp:111-158) and the hierarchy is rebuilt.
    // Returns false on error.
    bool updateFaces(const std::vector<FaceUpdate>& faces);

    float step(float dtMax);                 // Solver::step, returns dt
    int addBurner(const Burner& b);          // Solver::addBurner, -1 if it covers no gas cell

    // the phases (private members of the reference Solver)
    float stableDt(float dtMax) const;
    void stepClear();
    void injectBurners(float dt);
    void combustion(float dt);
    void turbulence();
    void heatDiffusion(float dt);
    void energyAndExpansion(float dt);
    void momentum(float dt);
    void leakage(float dt);
    void project(float dt);
    void advectScalars(float dt);
    void zonePressure(float dt);

    // Checkpoint and resume (the role of Solver::saveState / loadState,
    // src/sim/state.cpp:62-195) for the gas state: every field a later step
    // reads, zone pressures, the statistics stableDt uses, time and burners.
    // A state belongs to the geometry it was saved from (checked by a hash
    // of the open fractions, coefficients' zones and solid mask); resuming
    // continues bit for bit as if uninterrupted.
    bool saveState(const std::string& path, std::string& err);
    bool loadState(const std::string& path, std::string& err);

    void syncToHost();     // device state -> the public vectors
    void syncToDevice();   // the public vectors -> device state
    const std::string& error() const { return err_; }
    bool ok() const { return err_.empty(); }

    // --- state, as Solver's public members ---------------------------------
    int nx = 0, ny = 0, nz = 0, N = 0;
    float dx = 0.2f, ox = 0, oy = 0, oz = 0;
    float time = 0;
    std::vector<float> u, v, w, fx, fy, fz, p, S, T, Yo, Yf, Ys, Ya, Yp, Yc, Yw, Yco, Yai, nu, Pb, Q, Qr, massSrc, hrr;
    std::vector<uint8_t> blocked, pilot, lit;
    std::vector<int> zone;
    std::vector<uint8_t> zoneOpen;
    std::vector<float> zoneP, zoneDP;
    std::vector<Burner> burners;
    Wind wind;
    Profile profile;
    bool radiationOn = false;
    float maxVel = 0, maxNu = 0, maxDiv = 0;
    int leakFaces = 0;                       // crack faces handled by the last leakage
    int pressureIters = 0;
    float pressureResidual = 0;
    double hrrTotal = 0;

private:
    phg_ctx* ctx_ = nullptr;
    int grid_ = -1;
    std::string err_;
    bool deferStats_ = false;   // step(): maxNu and maxDiv read after the solve, in one wait
    std::vector<void*> dmem_;
    float *dU_ = nullptr, *dV_ = nullptr, *dW_ = nullptr, *dU2_ = nullptr, *dV2_ = nullptr, *dW2_ = nullptr;
    float *dP_ = nullptr, *dS_ = nullptr, *dNu_ = nullptr;
    float* dF_[10] = {};      // T, Yo, Yf, Ys, Ya, Yp, Yc, Yw, Yco, Yai
    float* dFt_[10] = {};     // transport output, swapped in
    float *dPb_ = nullptr, *dQ_ = nullptr, *dQr_ = nullptr, *dMs_ = nullptr, *dHrr_ = nullptr;
    uint8_t *dPilot_ = nullptr, *dLit_ = nullptr;
    std::vector<int*> dBurnerCells_;
    bool check(int rc, const char* what);
    template <class T>
    T* alloc(size_t n);
    phg_gas gas() const;
    struct Copy;
    std::vector<Copy> copies() const;
};

}  // namespace photon_gpu
