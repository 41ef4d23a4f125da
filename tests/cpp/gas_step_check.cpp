// gas_step_check.cpp -- a C++ caller of the MI355X gas phase through the
// compiled drop-in class photon_gpu::GasPhase (photon_amd/host/gas_phase.hpp)
// and the C ABI beneath it.  TEST PROGRAM: tests/test_gpu_dropin_cpp.py
// writes a case directory (meta.txt plus raw little-endian arrays), runs this
// program on it and compares what it writes with the reference's fixtures.
//
//   gas_step_check phases <dir>   project (fixed iterations, max-norm rule),
//                                 energyAndExpansion, advectScalars
//   gas_step_check step <dir>     Solver::step from still air with a burner
//   gas_step_check faces <dir>    the same through face changes between steps
//   gas_step_check resume <dir>   saveState after some steps, loadState in a
//                                 fresh object, compared with an uninterrupted run
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../photon_amd/host/gas_phase.hpp"

namespace {
std::map<std::string, std::string> read_meta(const std::string& dir) {
    std::map<std::string, std::string> m;
    std::ifstream in(dir + "/meta.txt");
    std::string line;
    while (std::getline(in, line)) {
        const size_t eq = line.find('=');
        if (eq != std::string::npos) m[line.substr(0, eq)] = line.substr(eq + 1);
    }
    return m;
}

template <class T>
std::vector<T> read_raw(const std::string& dir, const std::string& name) {
    std::ifstream in(dir + "/" + name + ".bin", std::ios::binary | std::ios::ate);
    if (!in) {
        std::fprintf(stderr, "missing %s/%s.bin\n", dir.c_str(), name.c_str());
        std::exit(2);
    }
    const size_t bytes = (size_t)in.tellg();
    std::vector<T> v(bytes / sizeof(T));
    in.seekg(0);
    in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)bytes);
    return v;
}

bool exists(const std::string& dir, const std::string& name) {
    std::ifstream in(dir + "/" + name + ".bin", std::ios::binary);
    return (bool)in;
}

template <class T>
void write_raw(const std::string& dir, const std::string& name, const std::vector<T>& v) {
    std::ofstream out(dir + "/out_" + name + ".bin", std::ios::binary);
    out.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

float num(const std::map<std::string, std::string>& m, const char* key, float dflt = 0.0f) {
    auto it = m.find(key);
    return it == m.end() ? dflt : std::strtof(it->second.c_str(), nullptr);
}
}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s phases|step <dir>\n", argv[0]);
        return 2;
    }
    const std::string mode = argv[1], dir = argv[2];
    const auto meta = read_meta(dir);
    // a GasPhase on the case's geometry, with its cracks and profile
    auto make = [&](photon_gpu::GasPhase& g) {
        g.wind.speed = num(meta, "wind_speed");
        g.wind.fromDeg = num(meta, "wind_from");
        g.wind.exponent = num(meta, "wind_exponent", 0.25f);
        std::string err;
        if (!g.build((int)num(meta, "nx"), (int)num(meta, "ny"), (int)num(meta, "nz"), num(meta, "dx"),
                     read_raw<float>(dir, "fx"), read_raw<float>(dir, "fy"), read_raw<float>(dir, "fz"),
                     read_raw<float>(dir, "ax"), read_raw<float>(dir, "ay"), read_raw<float>(dir, "az"),
                     read_raw<uint8_t>(dir, "blocked"), err)) {
            std::fprintf(stderr, "build: %s\n", err.c_str());
            return false;
        }
        if (exists(dir, "lx") &&
            !g.setCracks(read_raw<uint8_t>(dir, "lx"), read_raw<uint8_t>(dir, "ly"), read_raw<uint8_t>(dir, "lz"))) {
            std::fprintf(stderr, "setCracks: %s\n", g.error().c_str());
            return false;
        }
        g.profile.pressureTol = num(meta, "tol", 0.02f);
        g.profile.pressureMaxIters = (int)num(meta, "max_iters", 80);
        return true;
    };
    auto burner = [&](photon_gpu::GasPhase& g) {
        photon_gpu::Burner b;
        b.x0 = num(meta, "bx0");
        b.y0 = num(meta, "by0");
        b.x1 = num(meta, "bx1");
        b.y1 = num(meta, "by1");
        b.peakW = num(meta, "peak");
        b.rampS = num(meta, "ramp");
        return g.addBurner(b) >= 0;
    };
    photon_gpu::GasPhase gp;
    if (!make(gp)) return 1;
    std::FILE* log = std::fopen((dir + "/out_log.txt").c_str(), "w");
    if (mode == "phases") {
        const float dt = num(meta, "dt");
        // project after a fixed number of iterations (pressureTol 0)
        gp.u = read_raw<float>(dir, "pj_u0");
        gp.v = read_raw<float>(dir, "pj_v0");
        gp.w = read_raw<float>(dir, "pj_w0");
        gp.S = read_raw<float>(dir, "pj_S");
        gp.syncToDevice();
        gp.profile.pressureTol = 0.0f;
        gp.profile.pressureMaxIters = (int)num(meta, "pj_K");
        gp.profile.pressureCoarseWeight = num(meta, "coarse_weight", 0.0f);
        gp.project(dt);
        gp.syncToHost();
        std::fprintf(log, "project %d %.9g %.9g\n", gp.pressureIters, gp.pressureResidual, gp.maxVel);
        write_raw(dir, "p", gp.p);
        write_raw(dir, "u", gp.u);
        write_raw(dir, "v", gp.v);
        write_raw(dir, "w", gp.w);
        // energyAndExpansion
        gp.T = read_raw<float>(dir, "en_T0");
        gp.Q = read_raw<float>(dir, "en_Q");
        gp.massSrc = read_raw<float>(dir, "en_ms");
        gp.syncToDevice();
        gp.energyAndExpansion(dt);
        gp.syncToHost();
        std::fprintf(log, "energy %.9g\n", gp.maxDiv);
        write_raw(dir, "en_T", gp.T);
        write_raw(dir, "en_S", gp.S);
        // advectScalars on the corrected velocities
        gp.T = read_raw<float>(dir, "en_T0");
        std::vector<float>* F[9] = {&gp.Yo, &gp.Yf, &gp.Ys, &gp.Ya, &gp.Yp, &gp.Yc, &gp.Yw, &gp.Yco, &gp.Yai};
        const char* names[9] = {"Yo", "Yf", "Ys", "Ya", "Yp", "Yc", "Yw", "Yco", "Yai"};
        for (int q = 0; q < 9; ++q) *F[q] = read_raw<float>(dir, std::string("ad_in_") + names[q]);
        gp.nu = read_raw<float>(dir, "ad_nu");
        gp.u = read_raw<float>(dir, "pj_u");
        gp.v = read_raw<float>(dir, "pj_v");
        gp.w = read_raw<float>(dir, "pj_w");
        gp.syncToDevice();
        gp.advectScalars(dt);
        gp.syncToHost();
        write_raw(dir, "ad_T", gp.T);
        for (int q = 0; q < 9; ++q) write_raw(dir, std::string("ad_") + names[q], *F[q]);
    } else if (mode == "step") {
        photon_gpu::Burner b;
        b.x0 = num(meta, "bx0");
        b.y0 = num(meta, "by0");
        b.x1 = num(meta, "bx1");
        b.y1 = num(meta, "by1");
        b.peakW = num(meta, "peak");
        if (gp.addBurner(b) < 0) {
            std::fprintf(stderr, "burner covers no gas cell\n");
            return 1;
        }
        const int nsteps = (int)num(meta, "nsteps");
        gp.profile.pressureTol = num(meta, "tol");
        gp.profile.pressureMaxIters = (int)num(meta, "max_iters");
        for (int k = 0; k < nsteps; ++k) {
            const float dt = gp.step(num(meta, "dt_max"));
            std::fprintf(log, "step %d %.9g %d %.9g %.9g\n", k, dt, gp.pressureIters, gp.pressureResidual, gp.maxVel);
        }
        gp.syncToHost();
        write_raw(dir, "u", gp.u);
        write_raw(dir, "v", gp.v);
        write_raw(dir, "w", gp.w);
        write_raw(dir, "p", gp.p);
        write_raw(dir, "T", gp.T);
        write_raw(dir, "Yo", gp.Yo);
        write_raw(dir, "zoneP", gp.zoneP);
    } else if (mode == "resume") {
        // Solver's resume test (test_main.cpp:519-546): nsteps uninterrupted
        // (out_a_*), and ksave steps, saveState, a fresh object, loadState,
        // the remaining steps (out_b_*)
        const int nsteps = (int)num(meta, "nsteps"), ksave = (int)num(meta, "ksave");
        const float dtMax = num(meta, "dt_max");
        const std::string state = dir + "/state.bin";
        if (!burner(gp)) return 1;
        for (int k = 0; k < nsteps; ++k) gp.step(dtMax);
        gp.syncToHost();
        {
            photon_gpu::GasPhase g1;
            if (!make(g1) || !burner(g1)) return 1;
            for (int k = 0; k < ksave; ++k) g1.step(dtMax);
            std::string err;
            if (!g1.saveState(state, err)) {
                std::fprintf(stderr, "saveState: %s\n", err.c_str());
                return 1;
            }
        }
        photon_gpu::GasPhase g2;
        if (!make(g2)) return 1;
        std::string err;
        if (!g2.loadState(state, err)) {
            std::fprintf(stderr, "loadState: %s\n", err.c_str());
            return 1;
        }
        for (int k = ksave; k < nsteps; ++k) g2.step(dtMax);
        g2.syncToHost();
        std::fprintf(log, "time %.9g %.9g iters %d %d\n", gp.time, g2.time, gp.pressureIters, g2.pressureIters);
        for (auto* pr : {&gp, &g2}) {
            const std::string tag = pr == &gp ? "a_" : "b_";
            write_raw(dir, tag + "u", pr->u);
            write_raw(dir, tag + "w", pr->w);
            write_raw(dir, tag + "p", pr->p);
            write_raw(dir, tag + "T", pr->T);
            write_raw(dir, tag + "Yf", pr->Yf);
            write_raw(dir, tag + "lit", pr->lit);
            write_raw(dir, tag + "zoneP", pr->zoneP);
        }
        // a state from another geometry is refused
        photon_gpu::GasPhase g3;
        std::vector<float> fx3 = read_raw<float>(dir, "fx");
        fx3[fx3.size() / 2] = fx3[fx3.size() / 2] > 0 ? 0.0f : 1.0f;
        if (g3.build((int)num(meta, "nx"), (int)num(meta, "ny"), (int)num(meta, "nz"), num(meta, "dx"), fx3,
                     read_raw<float>(dir, "fy"), read_raw<float>(dir, "fz"), read_raw<float>(dir, "ax"),
                     read_raw<float>(dir, "ay"), read_raw<float>(dir, "az"), read_raw<uint8_t>(dir, "blocked"), err)) {
            const bool refused = !g3.loadState(state, err);
            std::fprintf(log, "other_geometry_refused %d\n", refused ? 1 : 0);
        }
    } else if (mode == "faces") {
        // the step with face changes: event e's arrays evE_{axis,face,f,a,crack}
        // are applied before step evE_at
        photon_gpu::Burner b;
        b.x0 = num(meta, "bx0");
        b.y0 = num(meta, "by0");
        b.x1 = num(meta, "bx1");
        b.y1 = num(meta, "by1");
        b.peakW = num(meta, "peak");
        if (gp.addBurner(b) < 0) {
            std::fprintf(stderr, "burner covers no gas cell\n");
            return 1;
        }
        gp.profile.pressureTol = num(meta, "tol");
        gp.profile.pressureMaxIters = (int)num(meta, "max_iters");
        const int nsteps = (int)num(meta, "nsteps"), nev = (int)num(meta, "nevents");
        for (int k = 0; k < nsteps; ++k) {
            for (int e = 0; e < nev; ++e) {
                const std::string pre = "ev" + std::to_string(e) + "_";
                if ((int)num(meta, (pre + "at").c_str()) != k) continue;
                const auto ax = read_raw<int>(dir, pre + "axis"), fc = read_raw<int>(dir, pre + "face");
                const auto fo = read_raw<float>(dir, pre + "f"), co = read_raw<float>(dir, pre + "a");
                const auto cr = read_raw<uint8_t>(dir, pre + "crack");
                std::vector<photon_gpu::FaceUpdate> up(ax.size());
                for (size_t q = 0; q < ax.size(); ++q) {
                    up[q].axis = ax[q];
                    up[q].face = fc[q];
                    up[q].open = fo[q];
                    up[q].coef = co[q];
                    up[q].crack = cr[q] != 0;
                }
                if (!gp.updateFaces(up)) {
                    std::fprintf(stderr, "updateFaces: %s\n", gp.error().c_str());
                    return 1;
                }
            }
            const float dt = gp.step(num(meta, "dt_max"));
            std::fprintf(log, "step %d %.9g %d %d\n", k, dt, gp.pressureIters, (int)gp.zoneP.size());
        }
        gp.syncToHost();
        write_raw(dir, "u", gp.u);
        write_raw(dir, "T", gp.T);
        write_raw(dir, "p", gp.p);
        write_raw(dir, "zoneP", gp.zoneP);
    } else {
        std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
        return 2;
    }
    std::fclose(log);
    if (!gp.ok()) {
        std::fprintf(stderr, "%s\n", gp.error().c_str());
        return 1;
    }
    return 0;
}
