"""Autotuner of the batched pressure solve: per size class of grid, the
level-0 tile depth (planes per block, `kchunk`) and the level-0 smoother
sweeps (`sweeps0`, red-black pairs on each side of the V-cycle).

The reference has no counterpart (it runs one scenario on the CPU with a fixed
V-cycle, src/sim/pressure_mg.cpp:270-300).  Here a batch mixes grids of very
different sizes; the best block shape depends on the grid (short columns want
whole-column blocks, tall ones a split so that the launch fills the chip),
and one more smoothing sweep trades level-0 passes per iteration against
iterations.  The tuner times each candidate on the systems of one class as
their own batch, with the same right-hand sides and stop rule as the real
solve, and keeps the fastest time to solution.  Every candidate solves the
same linear systems to the same tolerance, so the choice changes speed only.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import gas

# (upper bound on cells, name); the last class is open-ended
SIZE_CLASSES = ((16_000, "small"), (64_000, "medium"), (None, "large"))
# kchunk 0 stands for whole columns (kchunk = nz of the grid)
KCHUNKS = (4, 8, 16, 0)
SWEEPS = (1, 2)


def size_class(n_cells: int) -> str:
    for bound, name in SIZE_CLASSES:
        if bound is None or n_cells < bound:
            return name
    raise AssertionError("unreachable")


@dataclass
class Choice:
    kchunk: int      # 0 = whole columns
    sweeps0: int
    seconds: float   # time to solution of the class's systems as one batch


def _kc(kchunk: int, nz: int) -> int:
    """Planes per block for the library: whole columns for 0, the library's
    own choice (0 there) for a negative value."""
    if kchunk < 0:
        return 0
    return nz if kchunk == 0 else min(kchunk, nz)


def settings(grids, table: dict) -> tuple[list[int], list[int]]:
    """Per-system kchunk and sweeps0 for `grids` (gas.Grid objects) from a
    table of class name -> Choice."""
    kcs, sws = [], []
    for G in grids:
        c = table[size_class(G.scen.n)]
        kcs.append(_kc(c.kchunk, G.scen.nz))
        sws.append(c.sweeps0)
    return kcs, sws


def _time_batch(ctx, grids, bs, xs, params, kcs, sws, repeats):
    B = gas.Batch(ctx, grids, sys_kchunk=kcs, sys_sweeps0=sws)
    try:
        B.solve(bs, xs, params)   # warm-up (and graph capture)
        ctx.synchronize()
        best = float("inf")
        stats = None
        for _ in range(repeats):
            t0 = time.perf_counter()
            stats = B.solve(bs, xs, params)
            ctx.synchronize()
            best = min(best, time.perf_counter() - t0)
        ok = all(s["status"] == gas.CONVERGED for s in stats)
        return best, ok
    finally:
        B.destroy()


def tune(ctx, grids, bs, xs, params, repeats: int = 2, kchunks=KCHUNKS, sweeps=SWEEPS) -> dict:
    """Time every (kchunk, sweeps0) candidate on each size class present in
    `grids` (gas.Grid objects; bs, xs: device pointers or arrays of the
    right-hand sides and solutions, one per grid).  Returns class name ->
    Choice.  Candidates that fail to converge are never chosen."""
    table = {}
    for _, name in SIZE_CLASSES:
        idx = [k for k, G in enumerate(grids) if size_class(G.scen.n) == name]
        if not idx:
            continue
        g = [grids[k] for k in idx]
        b = [bs[k] for k in idx]
        x = [xs[k] for k in idx]
        best = None
        for kc in kchunks:
            for nu in sweeps:
                t, ok = _time_batch(ctx, g, b, x, params, [_kc(kc, G.scen.nz) for G in g], [nu] * len(g), repeats)
                if ok and (best is None or t < best.seconds):
                    best = Choice(kc, nu, t)
        if best is None:
            best = Choice(0, 1, float("nan"))
        table[name] = best
    return table


def default_table() -> dict:
    """The settings without tuning: the library's own block depth, one sweep."""
    return {name: Choice(-1, 1, float("nan")) for _, name in SIZE_CLASSES}


def table_to_json(table: dict) -> dict:
    return {k: {"kchunk": ("whole column" if v.kchunk == 0 else v.kchunk), "sweeps0": v.sweeps0,
                "class_seconds": None if np.isnan(v.seconds) else round(v.seconds, 6)}
            for k, v in table.items()}
