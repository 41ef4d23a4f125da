"""Face changes rebuilt in place with levels 0 and 1 of the hierarchy formed
on the device (phg_grid_update_faces): the result must be the grid a fresh
host set-up makes of the changed geometry, bit for bit.

For each case a grid is created on the initial geometry and updated with a
set of face changes; a second grid is created directly on the changed
geometry by the host set-up (buildPressureLevels and computeZones restated,
themselves pinned to the reference's golden hierarchies).  Level sizes,
zones, one V-cycle z = M^-1 r and a whole solve must agree bit for bit: any
difference in aggregates, Galerkin values, summation order, child lists or
the coarsest solve would change those bits.  The cases cover walls with
openings, cracks, wind-closed sides, fractional (vented) coefficients, a
door failure that merges zones and faces that close."""
import numpy as np
import pytest

from photon_amd import gas
from photon_amd import scenarios as sc

pytestmark = pytest.mark.gpu


def toggles(g, seed, nface=60):
    """Face changes on faces between two gas cells: open faces close
    (open fraction 0) and closed ones open fully, chosen by a seeded draw."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = g.shape
    gas_ = (g.blocked.reshape(nz, ny, nx) == 0)
    out = []
    for axis in range(3):
        if axis == 0:
            both = np.zeros((nz, ny, nx + 1), bool)
            both[:, :, 1:nx] = gas_[:, :, 1:] & gas_[:, :, :-1]
            f = g.fx
        elif axis == 1:
            both = np.zeros((nz, ny + 1, nx), bool)
            both[:, 1:ny, :] = gas_[:, 1:, :] & gas_[:, :-1, :]
            f = g.fy
        else:
            both = np.zeros((nz + 1, ny, nx), bool)
            both[1:nz] = gas_[1:] & gas_[:-1]
            f = g.fz
        idx = np.flatnonzero(both.ravel())
        pick = rng.choice(idx, size=min(nface // 3, len(idx)), replace=False)
        newf = np.where(f[pick] > 0, 0.0, 1.0).astype(np.float32)
        out.append((np.full(len(pick), axis, np.int32), pick.astype(np.int32), newf, newf.copy(),
                    np.zeros(len(pick), np.uint8)))
    return tuple(np.concatenate([o[q] for o in out]) for q in range(5))


def vented(g, seed):
    """Open faces given fractional coefficients (vents), as a face change."""
    rng = np.random.default_rng(seed)
    f = g.fx
    idx = np.flatnonzero(f > 0)
    pick = rng.choice(idx, size=min(200, len(idx)), replace=False)
    val = (0.05 + 0.9 * rng.random(len(pick))).astype(np.float32)
    return (np.zeros(len(pick), np.int32), pick.astype(np.int32), val, val.copy(), np.zeros(len(pick), np.uint8))


CASES = {
    "walls24_toggles": (lambda: sc.house(7, 24, 16, 12), lambda g: toggles(g, 1), None),
    "house80_toggles": (lambda: sc.house(5, 80, 50, 40, n_walls=12, n_open=6), lambda g: toggles(g, 2, 300), None),
    "cracked_door_fails": (sc.cracked, lambda g: sc.cracked_events(g)[0], None),
    "cracked_window_shuts": (sc.cracked, lambda g: sc.cracked_events(g)[1], None),
    "room_wind_closed_side": (lambda: sc.room(window=True), lambda g: toggles(g, 3), np.array([0, 0, 0, 1], np.uint8)),
    "odd_grid_vents": (lambda: sc.house(3, 37, 21, 13, n_walls=6, n_open=3), lambda g: vented(g, 4), None),
}


@pytest.mark.parametrize("precision", [gas.PREC_FP32, gas.PREC_MIXED_F16])
@pytest.mark.parametrize("name", list(CASES))
def test_device_rebuild_equals_host_setup(ctx, name, precision):
    make, change, wc = CASES[name]
    g0 = make()
    upd = change(g0)
    g1 = sc.apply_face_updates(g0, upd)
    fresh = gas.Grid(ctx, g1, precision=precision, wind_closed=wc)
    moved = gas.Grid(ctx, g0, precision=precision, wind_closed=wc)
    rb = moved.update_faces(upd)
    print(name, "rebuilt", rb, "levels", fresh.info()["levels"], flush=True)
    assert rb == 2                                  # levels 0 and 1 came from the device
    assert moved.info() == fresh.info()
    za, zoa = fresh.zones()
    zb, zob = moved.zones()
    assert np.array_equal(za, zb) and np.array_equal(zoa, zob)
    n = g1.n
    r = sc.normal_field(11, n).astype(np.float32) * (g1.blocked == 0)
    dr = ctx.to_device(r)
    za_, zb_ = ctx.array(n), ctx.array(n)
    fresh.precondition(dr, za_)
    moved.precondition(dr, zb_)
    assert np.array_equal(za_.download().view(np.uint32), zb_.download().view(np.uint32))
    b = ctx.to_device(sc.rhs_from_divergence(g1, sc.divergence_field(g1, 7), 0.01))
    xa, xb = ctx.array(n), ctx.array(n)
    prm = gas.solve_params(tol=1e-6, max_iters=400, demean=True, zero_x=True)
    sa = fresh.solve(b, xa, prm)
    sb = moved.solve(b, xb, prm)
    assert sa == sb
    assert np.array_equal(xa.download().view(np.uint32), xb.download().view(np.uint32))
    fresh.destroy()
    moved.destroy()
