"""GPU parity of the kernels around the solve (RHS, zone mean, velocity
correction, divergence source, scalar transport) and of project(dt) as a
whole, against the CPU oracle (which matches the reference bit for bit).

Elementwise kernels: FP32 with FMA contraction on the GPU, so within a few
ulp (rtol 1e-5).  project(): pressure and corrected velocities within 1e-5
relative L2 of the oracle solved with the same stop rule.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from photon_amd import gas
from photon_amd import scenarios as sc

pytestmark = pytest.mark.gpu

DT = 0.01


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def flow_state(g, seed):
    """Seeded predicted velocities (zero on closed faces) and a divergence source."""
    u = (0.5 * sc.normal_field(seed, len(g.fx))).astype(np.float32) * g.fx
    v = (0.5 * sc.normal_field(seed + 1, len(g.fy))).astype(np.float32) * g.fy
    w = (0.5 * sc.normal_field(seed + 2, len(g.fz))).astype(np.float32) * g.fz
    S = sc.divergence_field(g, seed + 3)
    return u.astype(np.float32), v.astype(np.float32), w.astype(np.float32), S


@pytest.fixture(scope="module", params=["room", "house_small"])
def scen(request):
    return sc.room() if request.param == "room" else sc.house(3, 30, 20, 14)


def test_rhs(ctx, scen):
    g = scen
    u, v, w, S = flow_state(g, 21)
    b_ref, d_ref = po.rhs(g, DT, u, v, w, S)
    G = gas.Grid(ctx, g)
    du, dv, dw, dS = (ctx.to_device(a) for a in (u, v, w, S))
    db, dd, dp = ctx.array(g.n), ctx.array(g.n), ctx.to_device(np.ones(g.n, np.float32))
    ctx._check(ctx.L.phg_rhs(ctx.h, G.id, DT, du.ptr, dv.ptr, dw.ptr, dS.ptr, db.ptr, dd.ptr, dp.ptr))
    np.testing.assert_allclose(db.download(), b_ref, rtol=1e-5, atol=1e-5)
    np.testing.assert_array_equal(dd.download(), d_ref)
    p = dp.download()
    assert np.all(p[g.blocked != 0] == 0) and np.all(p[g.blocked == 0] == 1)
    G.destroy()


def test_leakage_makes_closed_zones_compatible(ctx, scen):
    """leakage() without cracks (gasphase.cpp:487-498): on every closed zone S
    loses its mean (the zone integral over the zone volume); open zones keep
    S.  Pinned by the FP64 zone means and by the compatibility of the shifted
    source (zero integral per closed zone)."""
    g = scen
    S = (sc.normal_field(41, g.n) + 3.0).astype(np.float32)   # a net source in every zone
    S[g.blocked != 0] = 0
    solver = gas.GasSolver(ctx, g)
    solver.set("S", S)
    shift = solver.leakage(report=True)
    out = solver.get("S")
    for z in range(len(g.zone_open)):
        sel = g.zone == z
        if g.zone_open[z]:
            assert shift[z] == 0 and np.array_equal(out[sel], S[sel])
        else:
            mean = S[sel].astype(np.float64).mean()
            assert abs(shift[z] - mean) <= 1e-6 * abs(mean)
            np.testing.assert_allclose(out[sel], S[sel] - np.float32(mean), rtol=0, atol=2e-6)
            assert abs(out[sel].astype(np.float64).sum()) <= 1e-5 * np.abs(S[sel]).sum()
    assert np.array_equal(out[g.blocked != 0], S[g.blocked != 0])


def test_zone_mean_and_correct(ctx, scen):
    g = scen
    u, v, w, _ = flow_state(g, 31)
    p = sc.normal_field(35, g.n).astype(np.float32)
    p[g.blocked != 0] = 0
    p_ref = po.zone_mean(g, p.copy())
    G = gas.Grid(ctx, g)
    dp = ctx.to_device(p)
    ctx._check(ctx.L.phg_zone_mean(ctx.h, G.id, dp.ptr))
    np.testing.assert_allclose(dp.download(), p_ref, rtol=1e-6, atol=1e-6)
    uu, vv, ww = u.copy(), v.copy(), w.copy()
    vm_ref = po.correct(g, DT, p_ref, uu, vv, ww)
    du, dv, dw = ctx.to_device(u), ctx.to_device(v), ctx.to_device(w)
    dp.upload(p_ref)
    import ctypes
    vm = ctypes.c_float()
    ctx._check(ctx.L.phg_correct(ctx.h, G.id, DT, dp.ptr, du.ptr, dv.ptr, dw.ptr, ctypes.byref(vm)))
    np.testing.assert_allclose(du.download(), uu, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(dv.download(), vv, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(dw.download(), ww, rtol=1e-5, atol=1e-6)
    assert abs(vm.value - vm_ref) <= 1e-5 * vm_ref
    G.destroy()


def test_project_matches_oracle(ctx, scen):
    g = scen
    u, v, w, S = flow_state(g, 41)
    H = po.Hierarchy(g)
    uo, vo, wo, po_ = u.copy(), v.copy(), w.copy(), np.zeros(g.n, np.float32)
    o = po.project(H, g, DT, uo, vo, wo, po_, S, tol=1e-6, max_iters=400, stop=po.STOP_REL_L2,
                   coarse=po.COARSE_EXACT)
    s = gas.GasSolver(ctx, g, tol=1e-6, max_iters=400)
    for name, arr in (("u", u), ("v", v), ("w", w), ("S", S)):
        s.set(name, arr)
    it = s.project(DT)
    p = s.get("p")
    print(f"iters gpu={it} oracle={o['iters']} p err={rel(p, po_):.2e} u err={rel(s.get('u'), uo):.2e}")
    assert abs(it - o["iters"]) <= 2
    assert rel(p, po_) < 1e-5
    for name, ref in (("u", uo), ("v", vo), ("w", wo)):
        assert rel(s.get(name), ref) < 1e-5, name
    assert abs(s.maxVel - o["max_vel"]) <= 1e-4 * max(o["max_vel"], 1e-6)


def test_energy_expansion(ctx, scen):
    g = scen
    rng = np.random.default_rng(5)
    T = (293.15 + 200 * rng.random(g.n)).astype(np.float32)
    Pb = np.full(g.n, 101325.0, np.float32)
    Q = (rng.random(g.n) * 50.0).astype(np.float32)
    Q[rng.random(g.n) < 0.7] = 0
    ms = (rng.random(g.n) * 1e-5).astype(np.float32)
    S = np.zeros(g.n, np.float32)
    T_ref, S_ref = T.copy(), S.copy()
    md_ref = po.energy_expansion(g, DT, Pb, Q, ms, T_ref, S_ref)
    s = gas.GasSolver(ctx, g)
    for name, arr in (("T", T), ("Pb", Pb), ("Q", Q), ("massSrc", ms)):
        s.set(name, arr)
    md = s.energy_and_expansion(DT)
    np.testing.assert_allclose(s.get("T"), T_ref, rtol=1e-6)
    np.testing.assert_allclose(s.get("S"), S_ref, rtol=1e-5, atol=1e-6)
    assert abs(md - md_ref) <= 1e-5 * md_ref


def test_advect_scalars(ctx, scen):
    g = scen
    u, v, w, _ = flow_state(g, 51)
    rng = np.random.default_rng(9)
    fields = []
    for q, name in enumerate(gas.FIELD_NAMES):
        base = {"T": 293.15, "Yo": 0.232, "Yc": 0.0006, "Yw": 0.0073}.get(name, 0.0)
        scale = 300.0 if name == "T" else 0.05
        fields.append((base + scale * rng.random(g.n)).astype(np.float32))
    Pb = np.full(g.n, 101325.0, np.float32)
    nu = (1e-3 * rng.random(g.n)).astype(np.float32)
    ref = po.advect_scalars(g, DT, u, v, w, Pb, nu, fields)
    s = gas.GasSolver(ctx, g)
    for name, arr in (("u", u), ("v", v), ("w", w), ("Pb", Pb), ("nu", nu)):
        s.set(name, arr)
    for name, arr in zip(gas.FIELD_NAMES, fields):
        s.set(name, arr)
    s.advect_scalars(DT)
    for name, r in zip(gas.FIELD_NAMES, ref):
        np.testing.assert_allclose(s.get(name), r, rtol=2e-5, atol=1e-7, err_msg=name)


def test_multistep_path_room_burner(ctx):
    """The room case (20x15x12 at 0.2 m, door open to ambient) with a 100 kW
    2x2-cell burner at the floor centre, 30 steps of dt = 0.01 s through the
    phases this library implements, in the order of Solver::step
    (src/sim/solver.cpp:252-257): energyAndExpansion, the closed-zone source
    shift of leakage, project (||r||/||b|| <= 1e-6), advectScalars.  The
    same sequence through the CPU oracle (the reference's method, 20-sweep
    coarsest smoother).  Stated per-field tolerances after 30 steps: face
    velocities 1e-5 relative L2, temperature and species 1e-6; pressure 1e-5
    relative L2 after the first step and, on every step, max |difference|
    within 1e-5 of the run's pressure scale.  Per-solve differences (the GPU's
    exact coarsest solve stops at a different iterate of the same tolerance)
    carry from step to step."""
    g = sc.room()
    dt, nsteps, hrr = 0.01, 30, 100e3
    n = g.n
    Q = np.zeros(n, np.float32)
    i0, j0 = g.nx // 2 - 1, g.ny // 2 - 1
    for j in (j0, j0 + 1):
        for i in (i0, i0 + 1):
            assert g.blocked[i + g.nx * j] == 0
            Q[i + g.nx * j] = hrr * dt / 4.0
    # GPU
    s = gas.GasSolver(ctx, g, tol=1e-6, max_iters=400)
    s.set("Q", Q)
    # oracle
    H = po.Hierarchy(g)
    T = np.full(n, 293.15, np.float32)
    fields = [T, np.full(n, 0.232, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32),
              np.zeros(n, np.float32), np.zeros(n, np.float32), np.full(n, 0.0006, np.float32),
              np.full(n, 0.0073, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32)]
    Pb = np.full(n, 101325.0, np.float32)
    nu = np.zeros(n, np.float32)
    zero = np.zeros(n, np.float32)
    u, v, w = (np.zeros(len(a), np.float32) for a in (g.fx, g.fy, g.fz))
    p = np.zeros(n, np.float32)
    S = np.zeros(n, np.float32)
    closed = [z for z in range(len(g.zone_open)) if not g.zone_open[z]]
    iters, p_err, p_scale = [], [], 0.0
    for step in range(nsteps):
        s.energy_and_expansion(dt)
        s.leakage()
        it_gpu = s.project(dt)
        s.advect_scalars(dt)
        Tcur = fields[0]
        po.energy_expansion(g, dt, Pb, Q, zero, Tcur, S)
        for z in closed:
            sel = g.zone == z
            S[sel] -= np.float32(S[sel].astype(np.float64).mean())
        out = po.project(H, g, dt, u, v, w, p, S, tol=1e-6)
        fields = po.advect_scalars(g, dt, u, v, w, Pb, nu, fields)
        iters.append((it_gpu, out["iters"]))
        pg = s.get("p")
        p_scale = max(p_scale, float(np.abs(p).max()))
        p_err.append((rel(pg, p), float(np.abs(pg - p).max())))
    errs = {"u": rel(s.get("u"), u), "v": rel(s.get("v"), v), "w": rel(s.get("w"), w),
            "T": rel(s.get("T"), fields[0]), "Yo": rel(s.get("Yo"), fields[1]), "Yc": rel(s.get("Yc"), fields[6])}
    worst_p = max(e[1] for e in p_err) / p_scale
    print("room burner 30 steps: iters (gpu, cpu)", iters[:3], "...", iters[-3:], "errors",
          {k: f"{v:.2e}" for k, v in errs.items()}, f"p step 1 rel {p_err[0][0]:.2e}",
          f"p max|diff| / pressure scale over the run {worst_p:.2e}")
    assert fields[0].max() > 300.0          # the burner heated the room
    # Without the momentum phase the velocity only changes by the projection,
    # so after the first step each step's pressure corrects a residual change
    # and is itself near zero: it is compared against the run's pressure scale
    assert p_err[0][0] < 1e-5
    assert worst_p < 1e-5
    for k in ("u", "v", "w"):
        assert errs[k] < 1e-5, (k, errs[k])
    for k in ("T", "Yo", "Yc"):
        assert errs[k] < 1e-6, (k, errs[k])


def test_graph_sized_steps_match_direct_launches(ctx):
    """GasSolver.project runs each solve as one captured graph sized from the
    previous step's iteration count (re-captured when that count moves to
    another multiple of 4); over a burner's first steps the pressure and the
    corrected velocity equal those of direct launches bit for bit."""
    g = sc.room()
    dt = 0.01
    Q = np.zeros(g.n, np.float32)
    Q[g.nx // 2 + g.nx * (g.ny // 2)] = 250.0
    out = []
    for graph_steps in (True, False):
        s = gas.GasSolver(ctx, g, tol=1e-6, max_iters=400, graph_steps=graph_steps)
        s.set("Q", Q)
        its = []
        for _ in range(8):
            s.energy_and_expansion(dt, report=False)
            s.leakage()
            its.append(s.project(dt))
            s.advect_scalars(dt)
        out.append((its, s.get("p"), s.get("u"), s.get("T")))
        for a in list(s.d.values()) + s._tmp:
            a.free()
        s.grid.destroy()
    assert out[0][0] == out[1][0]
    for k in (1, 2, 3):
        assert np.array_equal(out[0][k], out[1][k]), k
