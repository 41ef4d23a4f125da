"""phg_zone_mean_noalloc: the closed-zone mean removal without any device
allocation, against phg_zone_mean (the existing entry point) bit for bit.

The workspace is one fixed per-device array, shared by every call: two calls
in a row (the second on a smaller grid) must each equal phg_zone_mean, and a
grid whose tasks and zones exceed the workspace must be refused before any
work is queued."""
import ctypes as C

import numpy as np
import pytest

from photon_amd import gas
from photon_amd import scenarios as sc

pytestmark = pytest.mark.gpu


def _noalloc(ctx, grid, d):
    ctx._check(ctx.L.phg_zone_mean_noalloc(ctx.h, grid.id, d.ptr))
    ctx.synchronize()


def _with_existing(ctx, grid, x):
    d = ctx.to_device(x)
    ctx._check(ctx.L.phg_zone_mean(ctx.h, grid.id, d.ptr))
    out = d.download()
    d.free()
    return out


def test_noalloc_zone_mean_equals_zone_mean_twice(ctx):
    big = sc.house(3)                    # 80x50x40, closed domain: closed zones of up to 160k cells
    small = sc.house(7, 24, 16, 12)      # 24x16x12
    for g, seed in ((big, 5), (small, 6)):
        grid = gas.Grid(ctx, g)
        print(f"noalloc zone mean: grid {g.shape}, seed {seed}", flush=True)
        x = (sc.normal_field(seed, g.n) + 3.0).astype(np.float32)
        x[g.blocked.astype(bool)] = 0.0
        ref = _with_existing(ctx, grid, x)
        d = ctx.to_device(x)
        _noalloc(ctx, grid, d)
        got = d.download()
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
        # it did something: every closed zone now has mean ~0
        zone, zo = g.zone, g.zone_open
        for z in np.flatnonzero(zo == 0)[:4]:
            m = float(np.mean(got[zone == z].astype(np.float64)))
            assert abs(m) < 1e-5, (z, m)
        d.free()
        grid.destroy()


def test_noalloc_zone_mean_refuses_a_grid_beyond_its_workspace(ctx):
    # 24^3 cells in sealed 2x2x2 pockets (every face between pockets closed):
    # 1728 closed zones of one task each = 3456 workspace entries > 2048
    n = 24
    nfx = (n + 1) * n * n
    fx = np.ones(nfx, np.float32).reshape(n, n, n + 1)
    fy = np.ones(nfx, np.float32).reshape(n, n + 1, n)
    fz = np.ones(nfx, np.float32).reshape(n + 1, n, n)
    fx[:, :, 0::2] = 0.0
    fy[:, 0::2, :] = 0.0
    fz[0::2, :, :] = 0.0
    g = sc.Grid(n, n, n, 0.2, np.zeros(n ** 3, np.uint8), fx.ravel(), fy.ravel(), fz.ravel(),
                fx.ravel().copy(), fy.ravel().copy(), fz.ravel().copy())
    grid = gas.Grid(ctx, g)
    assert grid.info()["nzones"] == (n // 2) ** 3
    x = sc.normal_field(9, g.n).astype(np.float32)
    d = ctx.to_device(x)
    rc = ctx.L.phg_zone_mean_noalloc(ctx.h, grid.id, d.ptr)
    assert rc == -1
    msg = ctx.L.phg_last_error(ctx.h).decode()
    assert "limit 2048" in msg and "1728 closed zones" in msg, msg
    ctx.synchronize()
    assert np.array_equal(d.download(), x)     # nothing was queued
    # the existing entry point still serves this grid
    ref = _with_existing(ctx, grid, x)
    for z in range(3):
        assert abs(float(np.mean(ref[g.zone == z].astype(np.float64)))) < 1e-6
    d.free()
    grid.destroy()
