"""GPU parity of the drop-in path: phg_project under the reference's own stop
rule (max|r| dt/dx^2 <= pressureTol, PHG_STOP_MAXNORM), against the golden
vectors the reference's Solver::project wrote (tests/golden/make_golden.py)
and against the oracle run in the same mode.

With the grid set to the reference's coarsest solve (PHG_COARSE_GS20: 20
forward/backward Gauss-Seidel sweeps) the library follows the reference's
iterates, so after the fixture's K iterations (tol = 0) the pressure, the
corrected face velocities, pressureIters, pressureResidual and maxVel agree
with the reference to FP32 rounding (bar: 1e-5 relative L2 for fields, 1e-4
relative for the two scalars).  The wind fixture has its y-high side closed
by windAt (gasphase.cpp:52-61), the wind_closed argument of the grid.

Also here: the V-cycle with the Gauss-Seidel coarsest level against the
reference's pressurePrecondition output; grids whose coarsening stalls above
the dense inverse's 2048 unknowns; the refusals and statuses the C ABI gives
for a bad MAXNORM call and a non-finite right-hand side; the whole-solve
graph re-captured when `refine` changes.
"""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from photon_amd import gas
from photon_amd import scenarios as sc

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DT = 0.01
CASES = {"room": sc.room, "walls": lambda: sc.house(7, 24, 16, 12), "wind": lambda: sc.room(window=True)}


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def golden(name):
    return np.load(os.path.join(HERE, "golden", f"ref_{name}.npz"))


def wind_closed(G):
    return G["wind_closed"] if "wind_closed" in G.files else None


def run_project(ctx, g, G, tol, max_iters, coarse, graph=True):
    s = gas.GasSolver(ctx, g, wind_closed=wind_closed(G), tol=tol, max_iters=max_iters,
                      stop_rule=gas.STOP_MAXNORM, coarse=coarse, graph_steps=graph)
    for name in ("u", "v", "w"):
        s.set(name, G[f"pj_{name}0"])
    s.set("S", G["pj_S"])
    s.set("p", np.zeros(g.n, np.float32))
    it = s.project(DT)
    out = dict(iters=it, residual=s.pressureResidual, max_vel=s.maxVel, p=s.get("p"), u=s.get("u"), v=s.get("v"),
               w=s.get("w"))
    s.grid.destroy()
    return out


@pytest.mark.parametrize("name", list(CASES))
def test_gs20_vcycle_matches_reference_precondition(ctx, name):
    g, G = CASES[name](), golden(name)
    grid = gas.Grid(ctx, g, wind_closed=wind_closed(G))
    grid.coarse = gas.COARSE_GS20
    assert grid.coarse == gas.COARSE_GS20
    dr, dz = ctx.to_device(G["vc_r"]), ctx.array(g.n)
    grid.precondition(dr, dz)
    z = dz.download()
    e = rel(z, G["vc_z"])
    print(f"{name}: V-cycle with the reference's coarsest sweeps vs pressurePrecondition: {e:.2e}")
    assert e < 1e-5
    grid.destroy()


@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("name", list(CASES))
def test_project_follows_reference_iterates(ctx, name, graph):
    g, G = CASES[name](), golden(name)
    K = int(G["pj_K"])
    o = run_project(ctx, g, G, tol=0.0, max_iters=K, coarse=gas.COARSE_GS20, graph=graph)
    errs = {f: rel(o[f], G[f"pj_{f}"]) for f in ("p", "u", "v", "w")}
    r_err = abs(o["residual"] - float(G["pj_residual"])) / float(G["pj_residual"])
    v_err = abs(o["max_vel"] - float(G["pj_maxvel"])) / float(G["pj_maxvel"])
    print(f"{name} graph={graph}: iters {o['iters']}/{int(G['pj_iters'])} residual {o['residual']:.6g} "
          f"({float(G['pj_residual']):.6g}, {r_err:.1e}) maxVel {v_err:.1e} fields {errs}")
    assert o["iters"] == int(G["pj_iters"]) == K
    assert r_err < 1e-4 and v_err < 1e-4
    for f, e in errs.items():
        assert e < 1e-5, (f, e)


@pytest.mark.parametrize("tol", [0.02, 0.002])   # realtime / reference profiles (solver.h:89-90,100-101)
@pytest.mark.parametrize("coarse", [gas.COARSE_GS20, gas.COARSE_EXACT])
@pytest.mark.parametrize("name", list(CASES))
def test_project_maxnorm_matches_oracle(ctx, name, coarse, tol):
    g, G = CASES[name](), golden(name)
    wc = wind_closed(G)
    H = po.Hierarchy(g, wind_closed=wc)
    u, v, w = G["pj_u0"].copy(), G["pj_v0"].copy(), G["pj_w0"].copy()
    p = np.zeros(g.n, np.float32)
    oc = po.COARSE_GS20 if coarse == gas.COARSE_GS20 else po.COARSE_EXACT
    ref = po.project(H, g, DT, u, v, w, p, G["pj_S"], tol=tol, max_iters=80, stop=po.STOP_MAXNORM, coarse=oc,
                     wind_closed=wc)
    o = run_project(ctx, g, G, tol=tol, max_iters=80, coarse=coarse)
    print(f"{name} coarse={coarse} tol={tol}: iters gpu {o['iters']} oracle {ref['iters']}, residual "
          f"{o['residual']:.4g} / {ref['residual']:.4g}, p {rel(o['p'], p):.2e}, u {rel(o['u'], u):.2e}")
    assert o["iters"] == ref["iters"]
    assert o["residual"] <= tol and abs(o["residual"] - ref["residual"]) <= 1e-4 * ref["residual"]
    assert rel(o["p"], p) < 1e-5
    for f, a in (("u", u), ("v", v), ("w", w)):
        assert rel(o[f], a) < 1e-5, f
    assert abs(o["max_vel"] - ref["max_vel"]) <= 1e-5 * ref["max_vel"]


def pockets(n):
    """n^3 fluid cells in sealed 2x2x2 pockets (every face between pockets
    closed): the coarsening stalls at one unknown per pocket."""
    nfx = (n + 1) * n * n
    fx = np.ones(nfx, np.float32).reshape(n, n, n + 1)
    fy = np.ones(nfx, np.float32).reshape(n, n + 1, n)
    fz = np.ones(nfx, np.float32).reshape(n + 1, n, n)
    fx[:, :, 0::2] = 0.0
    fy[:, 0::2, :] = 0.0
    fz[0::2, :, :] = 0.0
    return sc.Grid(n, n, n, 0.2, np.zeros(n ** 3, np.uint8), fx.ravel(), fy.ravel(), fz.ravel(),
                   fx.ravel().copy(), fy.ravel().copy(), fz.ravel().copy())


def test_stalled_coarsening_takes_the_reference_sweeps(ctx):
    # 32^3 cells in 4096 pockets: too many coarsest unknowns for the dense
    # inverse (2048); the grid is built, takes the Gauss-Seidel sweeps, and
    # refuses the dense mode with a message
    g = pockets(32)
    grid = gas.Grid(ctx, g)
    assert grid.info()["levels"] == [32 ** 3, 16 ** 3]
    assert grid.coarse == gas.COARSE_GS20
    with pytest.raises(gas.PhgError, match="too many for the dense inverse"):
        grid.coarse = gas.COARSE_EXACT
    b = po.closed_zone_demean(g, sc.normal_field(4, g.n)).astype(np.float32)
    H = po.Hierarchy(g)
    x64, _, _ = H.pcg64(b)
    db, dx = ctx.to_device(b), ctx.array(g.n)
    dx.zero()
    st = grid.solve(db, dx, gas.solve_params(tol=1e-6, max_iters=100))
    x = dx.download()
    print(f"pockets 32^3: {st}, err vs FP64 oracle {rel(x, x64):.2e}")
    assert st["status"] == gas.CONVERGED and rel(x, x64) < 1e-5
    grid.destroy()


def test_maxnorm_needs_positive_inv_and_nan_rhs_is_flagged(ctx):
    g = sc.house(7, 24, 16, 12)
    grid = gas.Grid(ctx, g)
    b = sc.rhs_from_divergence(g, sc.divergence_field(g, 8), DT)
    db, dx = ctx.to_device(b), ctx.array(g.n)
    dx.zero()
    for inv in (0.0, -1.0, float("nan")):
        with pytest.raises(gas.PhgError, match="PHG_STOP_MAXNORM needs inv"):
            grid.solve(db, dx, gas.solve_params(tol=0.02, stop_rule=gas.STOP_MAXNORM, inv=inv))
    st = grid.solve(db, dx, gas.solve_params(tol=0.02, stop_rule=gas.STOP_MAXNORM, inv=4.0, max_iters=80))
    assert st["status"] == gas.CONVERGED and st["maxres"] <= 0.02
    bn = b.copy()
    bn[123] = np.nan
    dbn = ctx.to_device(bn)
    dx.zero()
    st = grid.solve(dbn, dx, gas.solve_params(tol=1e-6, max_iters=50, refine=2))
    assert st["status"] == gas.NONFINITE and st["iters"] == 0
    grid.destroy()


def test_whole_solve_graph_recaptured_when_refine_changes(ctx):
    # the set-up's residual precision depends on refine: a graph captured
    # with refine = 0 must not be replayed for refine = 2
    g = sc.house(7, 24, 16, 12)
    b = sc.rhs_from_divergence(g, sc.divergence_field(g, 9), DT)
    x0 = (0.01 * sc.normal_field(10, g.n)).astype(np.float32)
    res = {}
    for graph in (2, 0):
        grid = gas.Grid(ctx, g)
        db, dx = ctx.to_device(b), ctx.to_device(x0)
        grid.solve(db, dx, gas.solve_params(tol=1e-6, max_iters=60, use_graph=graph, check_every=16))
        dx.upload(x0)
        grid.solve(db, dx, gas.solve_params(tol=1e-6, max_iters=60, use_graph=graph, check_every=16, refine=2))
        res[graph] = dx.download()
        grid.destroy()
    assert np.array_equal(res[2].view(np.uint32), res[0].view(np.uint32))
