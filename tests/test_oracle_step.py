"""CPU tests of the oracle's restatement of the rest of the gas step
(oracle/oracle_step.c): eddy viscosity, heat conduction, momentum with wind,
burner gas and combustion, stableDt.

Pins: bit-exact agreement with the golden vectors written by the reference's
own Solver (tests/golden/ref_<case>_phases.npz, room / walled house / windy
room), and properties the physics fixes independently of any restatement:
still air stays still, a uniform flow has no sub-grid energy, conduction in a
closed box conserves heat, burning conserves element mass shares; crack
flow (leakage) against the reference's own leakage on the cracked two-room
grid, and its direction and balance from the orifice law."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from photon_amd import scenarios as sc

HERE = os.path.dirname(os.path.abspath(__file__))
DT = 0.01
CASES = {"room": sc.room, "walls": lambda: sc.house(7, 24, 16, 12), "wind": lambda: sc.room(window=True),
         "cracked": sc.cracked}
SPECIES = ("Yo", "Yf", "Ys", "Ya", "Yp", "Yc", "Yw", "Yco", "Yai")


def golden(name):
    return np.load(os.path.join(HERE, "golden", f"ref_{name}_phases.npz"))


def bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype == np.float32:
        return np.array_equal(a.view(np.uint32), np.asarray(b, np.float32).view(np.uint32))
    return np.array_equal(a, b)


@pytest.fixture(scope="module", params=list(CASES))
def case(request):
    return request.param, CASES[request.param](), golden(request.param)


def wind_of(G):
    return po.wind(*G["wind"].tolist()) if "wind" in G.files else None


def test_turbulence_matches_golden(case):
    _, g, G = case
    nu, m = po.turbulence(g, G["in_u"], G["in_v"], G["in_w"], G["in_T"])
    assert bits(nu, G["tu_nu"]) and np.float32(m) == G["tu_maxnu"]


def test_heat_diffusion_matches_golden(case):
    _, g, G = case
    Q = G["in_Q"].copy()
    po.heat_diffusion(g, DT, G["in_T"], G["tu_nu"], G["in_Pb"], Q)
    assert bits(Q, G["hd_Q"])


def test_momentum_matches_golden(case):
    _, g, G = case
    u, v, w = po.momentum(g, DT, G["in_u"], G["in_v"], G["in_w"], G["in_T"], G["in_Pb"], G["tu_nu"], G["in_p"],
                          wind_of(G), leaks=g.leaks)
    assert bits(u, G["mo_u"]) and bits(v, G["mo_v"]) and bits(w, G["mo_w"])


def test_burner_and_combustion_match_golden(case):
    _, g, G = case
    f = {n: G["in_" + n].copy() for n in ("T", "Pb") + SPECIES}
    f["massSrc"] = np.zeros(g.n, np.float32)
    Q, Qr, hrr = (np.zeros(g.n, np.float32) for _ in range(3))
    lit, pilot = G["in_lit"].copy(), G["in_pilot"].copy()
    cells = G["cb_cells"]
    f32 = np.float32
    mf = f32(50e3) * f32(DT) / f32(25e6) / f32(len(cells))
    po.add_gas(g, cells, mf, 0.0, mf * f32(0.01), 800.0, True, f, pilot)
    po.combustion(g, DT, False, f, Q, Qr, hrr, pilot, lit)
    for n in ("T",) + SPECIES:
        assert bits(f[n], G["cb_" + n]), n
    for n, a in (("Q", Q), ("Qr", Qr), ("hrr", hrr), ("massSrc", f["massSrc"])):
        assert bits(a, G["cb_" + n]), n
    assert bits(lit, G["cb_lit"]) and bits(pilot, G["cb_pilot"])
    assert (G["cb_hrr"] > 0).sum() > 10   # the fixture burns


def test_stable_dt_matches_golden(case):
    _, g, G = case
    dt_max, max_vel, max_nu, max_div, cfl, pdt = (float(x) for x in G["sd_in"])
    assert np.float32(po.stable_dt(dt_max, pdt, cfl, g.dx, max_vel, max_nu, max_div)) == G["sd_dt"]


# ----------------------------------------------------------- physics pins --
def test_still_air_stays_still():
    # uniform T and Pb, zero velocity and pressure: no force anywhere (the
    # reference's still_air_stays_still, test_main.cpp:437-445, for one step)
    g = sc.house(3, 20, 14, 10, n_walls=4, n_open=2)
    z = [np.zeros(len(a), np.float32) for a in (g.fx, g.fy, g.fz)]
    T = np.full(g.n, 293.15, np.float32)
    Pb = np.full(g.n, 101325.0, np.float32)
    nu, m = po.turbulence(g, *z, T)
    u, v, w = po.momentum(g, DT, *z, T, Pb, nu, np.zeros(g.n, np.float32))
    assert not u.any() and not v.any() and not w.any()
    # nu is the molecular viscosity of air at T0 in every gas cell
    assert np.allclose(nu[g.blocked == 0], 1.5e-5, rtol=1e-6) and np.all(nu[g.blocked == 1] == 0)


def test_uniform_flow_has_no_subgrid_energy():
    g = sc.room()
    u = np.full(len(g.fx), 0.7, np.float32)
    v = np.full(len(g.fy), -0.3, np.float32)
    w = np.full(len(g.fz), 0.2, np.float32)
    nu, m = po.turbulence(g, u, v, w, np.full(g.n, 400.0, np.float32))
    assert np.allclose(nu, 1.5e-5 * (400.0 / 293.15) ** 1.7, rtol=1e-5)


def test_conduction_in_a_closed_box_conserves_heat():
    g = sc.house(5, 16, 12, 8, n_walls=3, n_open=1)    # closed domain
    rng = np.random.default_rng(3)
    T = (293.15 + 500 * rng.random(g.n)).astype(np.float32)
    nu = (1e-3 * rng.random(g.n)).astype(np.float32)
    Pb = np.full(g.n, 101325.0, np.float32)
    Q = np.zeros(g.n, np.float32)
    po.heat_diffusion(g, DT, T, nu, Pb, Q)
    assert abs(Q.astype(np.float64).sum()) <= 1e-5 * np.abs(Q).astype(np.float64).sum()
    assert np.abs(Q).sum() > 0


def test_burning_conserves_element_shares():
    # one cell's burn: fuel lost equals CO2/soot/CO carbon plus O2 balance
    g = sc.room()
    f = {n: np.zeros(g.n, np.float32) for n in ("T", "Pb") + SPECIES + ("massSrc",)}
    f["T"][:] = 1200.0
    f["Pb"][:] = 101325.0
    f["Yo"][:] = 0.2
    f["Yf"][:] = 0.05
    f["Yai"][:] = 0.05 * 800.0
    Q, Qr, hrr = (np.zeros(g.n, np.float32) for _ in range(3))
    lit = np.zeros(g.n, np.uint8)
    pilot = np.zeros(g.n, np.uint8)
    y0 = {n: f[n].astype(np.float64).copy() for n in SPECIES}
    tot = po.combustion(g, DT, False, f, Q, Qr, hrr, pilot, lit)
    dfuel = y0["Yf"] - f["Yf"]
    frac = 1 - np.exp(-DT / 0.1)
    assert np.allclose(dfuel, 0.05 * frac, rtol=1e-5)             # mixing-limited share of the fuel
    carbon = (f["Yc"] - y0["Yc"]) * 12 / 44 + (f["Yco"] - y0["Yco"]) * 12 / 28 + (f["Ys"] - y0["Ys"])
    assert np.allclose(carbon, 0.597 * dfuel, rtol=1e-4)
    assert np.allclose(Q, dfuel * 25e6 * (101325.0 / (287.0 * 1200.0)) * 0.2 ** 3, rtol=1e-4)
    assert tot > 0 and lit.all()


# ---- crack flow and closed-zone pressure (leakage, gasphase.cpp:402-498) ----
LEAK = np.load(os.path.join(HERE, "golden", "ref_cracked_leakage.npz"))


def leak_run(g, G, key):
    ins = {f: G[f"{key}_in_{f}"].copy() for f in ("u", "v", "w", "S")}
    zdp, nf = po.leakage(g, DT, g.leaks, G[f"{key}_in_zoneP"], G[f"{key}_in_T"], G[f"{key}_in_Pb"],
                         G[f"{key}_in_p"], ins["u"], ins["v"], ins["w"], ins["S"])
    return ins, zdp, nf


@pytest.mark.parametrize("key", ["lk", "lk2"])
def test_leakage_matches_golden(key):
    g = sc.cracked()
    out, zdp, nf = leak_run(g, LEAK, key)
    assert nf == int(LEAK[f"{key}_faces"]) == sum(int(a.sum()) for a in g.leaks)
    for f in ("u", "v", "w", "S"):
        assert bits(out[f], LEAK[f"{key}_{f}"]), f
    assert bits(zdp, LEAK[f"{key}_zoneDP"])
    # only crack faces move
    for f, lk in zip("uvw", g.leaks):
        moved = out[f] != LEAK[f"{key}_in_{f}"]
        assert not (moved & (lk == 0)).any()


def face_cells(g, axis):
    """(low-side cell, high-side cell) of every face of an axis, -1 outside."""
    nx, ny, nz = g.shape
    ex, ey, ez = nx + (axis == 0), ny + (axis == 1), nz + (axis == 2)
    k, j, i = np.meshgrid(np.arange(ez), np.arange(ey), np.arange(ex), indexing="ij")
    pos = (i, j, k)[axis].ravel()
    lim = (nx, ny, nz)[axis]
    step = (1, nx, nx * ny)[axis]
    c = (i + nx * (j + ny * k)).ravel()
    cb = np.where(pos < lim, c, -1)
    ca = np.where(pos > 0, c - step, -1)
    return ca, cb


def test_leakage_pressurised_zone_vents_outward():
    """A closed zone above ambient with still air, no source and no dynamic
    pressure: every crack blows out of it (positive velocity is low side to
    high side), its rise rate is negative, S in it rises uniformly by the
    vented volume per unit volume, and the open zone is untouched."""
    g = sc.cracked()
    zone, zo = g.zone, g.zone_open
    z = int(np.flatnonzero(zo == 0)[0])
    zp = np.where(zo == 0, np.float32(101325.0 + 40.0), np.float32(101325.0)).astype(np.float32)
    T = np.full(g.n, 293.15, np.float32)
    Pb = np.where(zone >= 0, zp[np.maximum(zone, 0)], np.float32(101325.0)).astype(np.float32)
    u, v, w = (np.zeros(len(a), np.float32) for a in (g.fx, g.fy, g.fz))
    S = np.zeros(g.n, np.float32)
    zdp, nf = po.leakage(g, DT, g.leaks, zp, T, Pb, np.zeros(g.n, np.float32), u, v, w, S)
    assert nf > 0 and zdp[z] < 0 and all(zdp[k] == 0 for k in np.flatnonzero(zo))
    for axis, (vel, lk) in enumerate(zip((u, v, w), g.leaks)):
        ca, cb = face_cells(g, axis)
        za = np.where(ca >= 0, zone[np.maximum(ca, 0)], -1)
        zb = np.where(cb >= 0, zone[np.maximum(cb, 0)], -1)
        crack = lk.astype(bool)
        assert (vel[crack & (za == z) & (zb != z)] > 0).all()
        assert (vel[crack & (zb == z) & (za != z)] < 0).all()
        assert (vel[crack & (za != z) & (zb != z)] == 0).all()
        assert (vel[~crack] == 0).all()
    assert np.ptp(S[zone == z]) == 0 and S[zone == z][0] > 0
    assert (S[zone != z] == 0).all()


def test_leakage_without_cracks_is_the_mean_shift():
    """With no crack faces the closed zones' S loses its volume mean and
    zoneDP = gamma P0 * integral(S dV) / Vz (the no-crack source shift)."""
    g = sc.room(door=False)
    zo = g.zone_open
    rng = np.random.default_rng(3)
    S = rng.random(g.n).astype(np.float32) * (g.blocked == 0)
    S0 = S.copy()
    zp = np.full(len(zo), 101325.0, np.float32)
    u, v, w = (np.zeros(len(a), np.float32) for a in (g.fx, g.fy, g.fz))
    zdp, nf = po.leakage(g, DT, None, zp, np.full(g.n, 293.15, np.float32), np.full(g.n, 101325.0, np.float32),
                         np.zeros(g.n, np.float32), u, v, w, S)
    assert nf == 0 and not u.any() and not v.any() and not w.any()
    z = g.zone == 0
    Vz = po.zone_volumes(g)[0]
    dint = np.sum(S0[z].astype(np.float64) * np.float64(np.float32(g.dx) ** 3))
    assert abs(S[z].astype(np.float64).sum() * float(np.float32(g.dx) ** 3)) < 1e-6 * dint
    gamma = np.float32(1100.0) / (np.float32(1100.0) - np.float32(287.0))
    assert abs(zdp[0] - gamma * 101325.0 * dint / Vz) < 1e-5 * abs(zdp[0])


def test_face_change_zones_match_reference():
    """The cracked grid's two face changes (door fails, window shuts to a
    crack and the panel turns solid) applied as updateFace does: the zones
    and their openness after each equal the reference's own recomputed
    zones (computeZones after zonesDirty_), and the oracle's hierarchy of the
    final geometry has the level sizes the reference rebuilds."""
    G = np.load(os.path.join(HERE, "golden", "ref_cracked_faces.npz"))
    g = sc.cracked()
    for k, upd in zip((10, 20), sc.cracked_events(g)):
        g = sc.apply_face_updates(g, upd)
        zone, zo = po.zones(g)
        assert np.array_equal(zone, G[f"zone_{k}"]) and np.array_equal(zo, G[f"zone_open_{k}"]), k
        assert np.array_equal(g.zone, zone) and np.array_equal(g.zone_open, zo)
    assert len(g.zone_open) == 1 and g.zone_open[0] == 0
