"""The compiled C++ drop-in on the GPU: tests/cpp/gas_step_check (built by
__graft_entry__.build() with g++) drives photon_gpu::GasPhase, the
Solver-shaped host class over the C ABI, and its results are compared with
the fixtures the reference's own Solver wrote.

phases: project after the fixture's K iterations under the max-norm rule
with the reference's coarsest sweeps (pressureIters, pressureResidual,
maxVel, p, u, v, w), energyAndExpansion (T, S, maxDiv) and advectScalars (ten
fields, bit for bit).  step: 20 steps of Solver::step from still air with a
burner: every step's dt and pressureIters, then the fields."""
import os
import subprocess

import numpy as np
import pytest

from photon_amd import gas
from photon_amd import scenarios as sc

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "gas_step_check")
CASES = {"room": sc.room, "walls": lambda: sc.house(7, 24, 16, 12), "wind": lambda: sc.room(window=True),
         "cracked": sc.cracked}
SPECIES = ("Yo", "Yf", "Ys", "Ya", "Yp", "Yc", "Yw", "Yco", "Yai")


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def write_case(d, g, meta, arrays):
    os.makedirs(d, exist_ok=True)
    base = dict(nx=g.nx, ny=g.ny, nz=g.nz, dx=g.dx)
    base.update(meta)
    with open(os.path.join(d, "meta.txt"), "w") as f:
        for k, v in base.items():
            f.write(f"{k}={v}\n")
    geo = dict(fx=g.fx, fy=g.fy, fz=g.fz, ax=g.ax, ay=g.ay, az=g.az)
    for k, a in geo.items():
        np.ascontiguousarray(a, np.float32).tofile(os.path.join(d, k + ".bin"))
    np.ascontiguousarray(g.blocked, np.uint8).tofile(os.path.join(d, "blocked.bin"))
    if g.leaks is not None:
        for k, a in zip(("lx", "ly", "lz"), g.leaks):
            np.ascontiguousarray(a, np.uint8).tofile(os.path.join(d, k + ".bin"))
    for k, a in arrays.items():
        np.ascontiguousarray(a).tofile(os.path.join(d, k + ".bin"))


def run(mode, d):
    assert os.path.exists(EXE), "tests/cpp/gas_step_check not built: run __graft_entry__.build()"
    print(f"gas_step_check {mode} {d}", flush=True)
    r = subprocess.run([EXE, mode, d], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    with open(os.path.join(d, "out_log.txt")) as f:
        return f.read().split("\n")


def out(d, name, dtype=np.float32):
    return np.fromfile(os.path.join(d, f"out_{name}.bin"), dtype)


def wind_meta(G):
    if "wind" not in G.files:
        return {}
    s, fr, ex = (float(x) for x in G["wind"])
    return dict(wind_speed=s, wind_from=fr, wind_exponent=ex)


@pytest.mark.parametrize("name", list(CASES))
def test_cpp_dropin_phases_match_reference(tmp_path, name):
    g = CASES[name]()
    G = np.load(os.path.join(HERE, "golden", f"ref_{name}.npz"))
    d = str(tmp_path / name)
    arrays = {k: G[k] for k in ("pj_u0", "pj_v0", "pj_w0", "pj_S", "en_T0", "en_Q", "en_ms", "ad_nu", "pj_u", "pj_v",
                                "pj_w")}
    arrays.update({"ad_in_" + n: G["ad_in_" + n] for n in SPECIES})
    write_case(d, g, dict(dt=0.01, pj_K=int(G["pj_K"]), **wind_meta(G)), arrays)
    log = run("phases", d)
    it, res, vmax = log[0].split()[1:]
    assert int(it) == int(G["pj_iters"])
    assert abs(float(res) - float(G["pj_residual"])) <= 1e-4 * float(G["pj_residual"])
    assert abs(float(vmax) - float(G["pj_maxvel"])) <= 1e-4 * max(float(G["pj_maxvel"]), 1e-30)
    for f in ("p", "u", "v", "w"):
        assert rel(out(d, f), G["pj_" + f]) < 1e-5, f
    assert np.array_equal(out(d, "en_T"), G["en_T"]) and np.array_equal(out(d, "en_S"), G["en_S"])
    assert np.float32(float(log[1].split()[1])) == G["en_maxdiv"]
    assert np.array_equal(out(d, "ad_T").view(np.uint32), G["ad_out_T"].view(np.uint32))
    for n in SPECIES:
        assert np.array_equal(out(d, "ad_" + n).view(np.uint32), G["ad_out_" + n].view(np.uint32)), n


@pytest.mark.parametrize("name", ["room", "walls"])
def test_cpp_dropin_coarse_weight_reaches_the_solve(ctx, tmp_path, name):
    """Profile::pressureCoarseWeight goes into the solve: GasPhase::project
    with weight 1.8 (the reference's coarsest sweeps, K iterations) equals the
    Python GasSolver's projection with the same weight bit for bit, and
    differs from the weight-0 projection (the reference's cycle, which the
    golden p follows)."""
    g = CASES[name]()
    G = np.load(os.path.join(HERE, "golden", f"ref_{name}.npz"))
    K = int(G["pj_K"])
    arrays = {k: G[k] for k in ("pj_u0", "pj_v0", "pj_w0", "pj_S", "en_T0", "en_Q", "en_ms", "ad_nu", "pj_u", "pj_v",
                                "pj_w")}
    arrays.update({"ad_in_" + n: G["ad_in_" + n] for n in SPECIES})
    outs = {}
    for cw in (0.0, 1.8):
        d = str(tmp_path / f"{name}_{cw}")
        write_case(d, g, dict(dt=0.01, pj_K=K, coarse_weight=cw), arrays)
        run("phases", d)
        outs[cw] = {f: out(d, f) for f in ("p", "u", "v", "w")}
    s = gas.GasSolver(ctx, g, tol=0.0, max_iters=K, stop_rule=gas.STOP_MAXNORM, coarse=gas.COARSE_GS20,
                      coarse_weight=1.8)
    for f in ("u", "v", "w"):
        s.set(f, G["pj_" + f + "0"])
    s.set("S", G["pj_S"])
    s.project(0.01)
    for f in ("p", "u", "v", "w"):
        assert np.array_equal(outs[1.8][f], s.get(f)), f
    assert not np.array_equal(outs[1.8]["p"], outs[0.0]["p"])
    assert rel(outs[0.0]["p"], G["pj_p"]) < 1e-5


@pytest.mark.parametrize("name", list(CASES))
def test_cpp_dropin_step_follows_reference(tmp_path, name):
    g = CASES[name]()
    G = np.load(os.path.join(HERE, "golden", f"ref_{name}_step.npz"))
    x0, y0, x1, y1, peak = (float(x) for x in G["burner"])
    rows = G["st_rows"]
    d = str(tmp_path / name)
    write_case(d, g, dict(bx0=x0, by0=y0, bx1=x1, by1=y1, peak=peak, nsteps=len(rows), dt_max=0.01, tol=0.02,
                          max_iters=80, **wind_meta(G)), {})
    log = [line.split() for line in run("step", d) if line.startswith("step")]
    assert len(log) == len(rows)
    for k, ln in enumerate(log):
        assert np.float32(float(ln[2])) == np.float32(rows[k, 0]), (k, ln)
        assert int(ln[3]) == int(rows[k, 1]), (k, ln, rows[k])
    for f in ("u", "v", "w"):
        assert rel(out(d, f), G["st_" + f]) < 1e-5, f
    assert rel(out(d, "T"), G["st_T"]) < 1e-6 and rel(out(d, "Yo"), G["st_Yo"]) < 1e-6
    p = out(d, "p")
    assert np.abs(p - G["st_p"]).max() <= 1e-4 * np.abs(G["st_p"]).max()
    assert np.abs(out(d, "zoneP") - G["st_zoneP"]).max() <= 1e-4 * 101325


def test_cpp_dropin_face_changes_follow_reference(tmp_path):
    """GasPhase::updateFaces between steps (the partition door fails after
    10 steps, the window shuts to a crack and the panel turns solid after
    20): every step's dt, pressureIters and zone count, then the fields, as
    the reference's own run (tests/golden/ref_cracked_faces.npz)."""
    g = sc.cracked()
    G = np.load(os.path.join(HERE, "golden", "ref_cracked_faces.npz"))
    x0, y0, x1, y1, peak = (float(x) for x in G["burner"])
    rows = G["st_rows"]
    d = str(tmp_path / "faces")
    arrays, meta = {}, dict(bx0=x0, by0=y0, bx1=x1, by1=y1, peak=peak, nsteps=len(rows), dt_max=0.01, tol=0.02,
                           max_iters=80, nevents=2)
    for e, (at, upd) in enumerate(zip((10, 20), sc.cracked_events(g))):
        meta[f"ev{e}_at"] = at
        for k, a in zip(("axis", "face", "f", "a", "crack"), upd):
            arrays[f"ev{e}_{k}"] = a
    write_case(d, g, meta, arrays)
    log = [line.split() for line in run("faces", d) if line.startswith("step")]
    assert len(log) == len(rows)
    for k, ln in enumerate(log):
        assert np.float32(float(ln[2])) == np.float32(rows[k, 0]), (k, ln)
        assert int(ln[3]) == int(rows[k, 1]) and int(ln[4]) == int(rows[k, 4]), (k, ln, rows[k])
    assert rel(out(d, "u"), G["st_u"]) < 1e-5
    assert rel(out(d, "T"), G["st_T"]) < 1e-6
    assert np.abs(out(d, "zoneP") - G["st_zoneP"]).max() <= 1e-4 * 101325


def test_cpp_dropin_resume_matches_uninterrupted(tmp_path):
    """GasPhase::saveState after 15 steps of the room with a ramped burner,
    loadState in a fresh object, 15 more steps: bit for bit the uninterrupted
    30 steps (the reference's resume test, test_main.cpp:519-546); a state
    saved from another geometry is refused."""
    g = sc.room()
    d = str(tmp_path / "resume")
    write_case(d, g, dict(bx0=1.8, by0=1.2, bx1=2.2, by1=1.6, peak=100e3, ramp=0.2, nsteps=30, ksave=15, dt_max=0.01,
                          tol=0.02, max_iters=80), {})
    log = run("resume", d)
    assert "other_geometry_refused 1" in log
    t = log[0].split()
    assert t[1] == t[2] and t[4] == t[5], log[0]
    for f, dt in (("u", np.float32), ("w", np.float32), ("p", np.float32), ("T", np.float32), ("Yf", np.float32),
                  ("lit", np.uint8), ("zoneP", np.float32)):
        a, b = out(d, "a_" + f, dt), out(d, "b_" + f, dt)
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), f
    assert out(d, "a_T").max() > 300.0
