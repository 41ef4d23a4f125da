"""Write the golden vectors of the pressure projection path from the
reference's own photon::Solver, compiled from its sources into
oracle/_ref/libphoton_ref_pin.so (oracle/Makefile, -ffp-contract=off).

Run in a container that has the reference checkout:
    make -C oracle && python tests/golden/make_golden.py
Each fixture is an .npz of inputs and the reference's outputs on a seeded
grid: the multigrid hierarchy (level sizes, level-1 operator), one V-cycle
z = M^-1 r (pressurePrecondition), project(dt) after a fixed number of CG
iterations (pressure, face velocities, pressureIters, pressureResidual,
maxVel), energyAndExpansion and advectScalars on seeded fields.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle as po  # noqa: E402
from photon_amd import scenarios as sc  # noqa: E402

DT = 0.01
CASES = {
    "room": lambda: sc.room(),
    "walls": lambda: sc.house(7, 24, 16, 12),
    # the room with a window in the y-high wall, wind from the north: the
    # window side takes inflow, so windAt closes it for the pressure matrix
    "wind": lambda: sc.room(window=True),
}
# Solver::wind (speed m/s at 10 m, degrees it blows from, profile exponent)
WIND = {"wind": (3.0, 0.0, 0.25)}
FIELDS = ("T", "Yo", "Yf", "Ys", "Ya", "Yp", "Yc", "Yw", "Yco", "Yai")


def flow_inputs(g, seed):
    rng = np.random.default_rng(seed)
    u = (0.5 * rng.standard_normal(len(g.fx))).astype(np.float32) * g.fx
    v = (0.5 * rng.standard_normal(len(g.fy))).astype(np.float32) * g.fy
    w = (0.5 * rng.standard_normal(len(g.fz))).astype(np.float32) * g.fz
    S = sc.divergence_field(g, seed + 3)
    return u.astype(np.float32), v.astype(np.float32), w.astype(np.float32), S


def scalar_inputs(g, seed):
    rng = np.random.default_rng(seed)
    T = (293.15 + 300 * rng.random(g.n)).astype(np.float32)
    ys = {}
    for name in FIELDS[1:]:
        base = {"Yo": 0.232, "Yc": 0.0006, "Yw": 0.0073}.get(name, 0.0)
        ys[name] = (base + 0.05 * rng.random(g.n)).astype(np.float32)
    nu = (1e-3 * rng.random(g.n)).astype(np.float32)
    Q = (50.0 * rng.random(g.n)).astype(np.float32)
    Q[rng.random(g.n) < 0.7] = 0
    ms = (1e-5 * rng.random(g.n)).astype(np.float32)
    return T, ys, nu, Q, ms


def make(name, g, K=6):
    out = {}
    R = po.RefSolver(g, threads=1, pin=True)
    if name in WIND:
        R.set_wind(*WIND[name])
        out["wind"] = np.array(WIND[name], np.float32)
        # sides (x-low, x-high, y-low, y-high) where windAt prescribes inflow;
        # the power law never changes the sign, so one height decides a side
        out["wind_closed"] = np.array([0 if np.isnan(R.wind_at(a, sd, 1.0)) else 1 for a in (0, 1) for sd in (0, 1)],
                                      np.uint8)
    zone, zo = R.zones()
    out["zone"], out["zone_open"] = zone, zo
    lv = R.levels()
    out["level_n"] = np.array([l["n"] for l in lv], np.int32)
    out["l1_rowStart"], out["l1_col"], out["l1_val"] = lv[1]["rowStart"], lv[1]["col"], lv[1]["val"]
    out["l1_diag"], out["l0_agg"] = lv[1]["diag"], lv[0]["agg"]
    r = sc.rhs_from_divergence(g, sc.divergence_field(g, 11), DT)
    out["vc_r"], out["vc_z"] = r, R.precondition(r)
    u, v, w, S = flow_inputs(g, 41)
    out["pj_u0"], out["pj_v0"], out["pj_w0"], out["pj_S"] = u, v, w, S
    for f, a in (("u", u), ("v", v), ("w", w), ("S", S)):
        R.set(f, a)
    R.set("p", np.zeros(g.n, np.float32))
    it, res, vm = R.project(DT, 0.0, K)
    out["pj_K"] = np.int32(K)
    out["pj_iters"], out["pj_residual"], out["pj_maxvel"] = np.int32(it), np.float32(res), np.float32(vm)
    out["pj_p"] = R.get("p", g.n)
    out["pj_u"], out["pj_v"], out["pj_w"] = R.get("u", len(g.fx)), R.get("v", len(g.fy)), R.get("w", len(g.fz))
    out["pj_rhs"] = R.get("rhs", g.n)
    # energyAndExpansion
    T, ys, nu, Q, ms = scalar_inputs(g, 77)
    R.set("T", T); R.set("Q", Q); R.set("massSrc", ms); R.set("Pb", np.full(g.n, 101325.0, np.float32))
    out["en_T0"], out["en_Q"], out["en_ms"] = T, Q, ms
    out["en_maxdiv"] = np.float32(R.energy(DT))
    out["en_T"], out["en_S"] = R.get("T", g.n), R.get("S", g.n)
    # advectScalars, with the corrected velocities of the project step
    R.set("T", T)
    for f in FIELDS[1:]:
        R.set(f, ys[f])
    R.set("nu", nu)
    out["ad_nu"] = nu
    for f in FIELDS[1:]:
        out["ad_in_" + f] = ys[f]
    R.advect(DT)
    for f in FIELDS:
        out["ad_out_" + f] = R.get(f, g.n)
    np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), **out)
    print(name, g.shape, "levels", out["level_n"], "project iters", it)


if __name__ == "__main__":
    if not po.ref_available(pin=True):
        sys.exit("oracle/_ref/libphoton_ref_pin.so not built (needs the reference checkout): make -C oracle")
    names = sys.argv[1:] or list(CASES)
    for name in names:
        make(name, CASES[name]())
