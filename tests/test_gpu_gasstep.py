"""GPU parity of the gas-step kernels around the projection, through the C ABI.

Scalar transport (advectScalars, src/sim/gasphase.cpp:685-784) runs as a
face-centred plane march that evaluates each face once; it forms every value
with the reference's operations in the reference's order and without FMA
contraction, so the bar is bit-exact: against the golden vectors written by
the reference's own Solver (room, walled house, windy room) and against the
oracle on larger walled grids, including grids whose sizes leave partial
column tiles and k-chunks.
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
GOLDEN = {"room": sc.room, "walls": lambda: sc.house(7, 24, 16, 12), "wind": lambda: sc.room(window=True)}


def golden(name):
    return np.load(os.path.join(HERE, "golden", f"ref_{name}.npz"))


def bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def run_transport(ctx, g, u, v, w, Pb, nu, fields):
    s = gas.GasSolver(ctx, g)
    for name, arr in (("u", u), ("v", v), ("w", w), ("Pb", Pb), ("nu", nu)):
        s.set(name, arr)
    for name, arr in zip(gas.FIELD_NAMES, fields):
        s.set(name, arr)
    s.advect_scalars(DT)
    out = [s.get(name) for name in gas.FIELD_NAMES]
    s.grid.destroy()
    return out


@pytest.mark.parametrize("name", list(GOLDEN))
def test_transport_matches_reference_bitwise(ctx, name):
    g, G = GOLDEN[name](), golden(name)
    fields = [G["en_T0"]] + [G["ad_in_" + f] for f in gas.FIELD_NAMES[1:]]
    out = run_transport(ctx, g, G["pj_u"], G["pj_v"], G["pj_w"], np.full(g.n, 101325.0, np.float32), G["ad_nu"],
                        fields)
    for f, o in zip(gas.FIELD_NAMES, out):
        ref = G["ad_out_" + f]
        assert bits_equal(o, ref), (f, int(np.sum(o != ref)), float(np.max(np.abs(o - ref))))


ORACLE_GRIDS = {
    "house37x21x13": lambda: sc.house(5, 37, 21, 13, n_walls=6, n_open=3),
    "house80x50x40": lambda: sc.house(5, 80, 50, 40, n_walls=12, n_open=6),
    "room40x30x24_window": lambda: sc.room(0.1, window=True),   # open door and window: boundary faces
}


@pytest.mark.parametrize("name", list(ORACLE_GRIDS))
def test_transport_matches_oracle_bitwise(ctx, name):
    g = ORACLE_GRIDS[name]()
    shape = g.shape
    rng = np.random.default_rng(g.nx)
    u = (0.8 * rng.standard_normal(len(g.fx))).astype(np.float32) * g.fx
    v = (0.8 * rng.standard_normal(len(g.fy))).astype(np.float32) * g.fy
    w = (0.8 * rng.standard_normal(len(g.fz))).astype(np.float32) * g.fz
    u[rng.random(len(u)) < 0.1] = 0.0                      # faces without flow
    fields = []
    for name in gas.FIELD_NAMES:
        base = {"T": 293.15, "Yo": 0.232, "Yc": 0.0006, "Yw": 0.0073}.get(name, 0.0)
        scale = 600.0 if name == "T" else 0.05
        f = (base + scale * rng.random(g.n)).astype(np.float32)
        f[rng.random(g.n) < 0.2] = base                     # flat patches: zero limiter denominators
        fields.append(f)
    Pb = (101325.0 + 50.0 * rng.standard_normal(g.n)).astype(np.float32)
    nu = (2e-3 * rng.random(g.n)).astype(np.float32)
    ref = po.advect_scalars(g, DT, u, v, w, Pb, nu, fields)
    out = run_transport(ctx, g, u, v, w, Pb, nu, fields)
    for f, o, r in zip(gas.FIELD_NAMES, out, ref):
        assert bits_equal(o, r), (shape, f, int(np.sum(o != r)), float(np.max(np.abs(o - r))))
