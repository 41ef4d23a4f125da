"""Autotuner of the batched pressure solve: per size class of grid, the
level-0 tile depth (planes per block, `kchunk`) and the level-0 smoother
sweeps (`sweeps0`, red-black pairs on each side of the V-cycle).

The reference has no counterpart (it runs one scenario on the CPU with a fixed
V-cycle, src/sim/pressure_mg.cpp:270-300).  Here a batch mixes grids of very
different sizes; the best block shape depends on the grid (short columns want
whole-column blocks, tall ones a split so that the launch fills the chip),
and one more smoothing sweep trades level-0 passes per iteration against
iterations.  A batch's time is set by the mix (its slowest systems and the
launches they keep alive), so a candidate is judged inside the whole batch,
with the same right-hand sides and stop rule as the real solve: coordinate
descent from the library's defaults, one class at a time, largest first.
Every candidate solves the same linear systems to the same tolerance, so the
choice changes speed only.
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
    kchunk: int      # 0 = whole columns, -1 = the library's own choice
    sweeps0: int
    seconds: float   # the whole batch's time once this class was tuned (classes are tuned largest first)


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


def tune(ctx, grids, bs, xs, params, repeats: int = 2, kchunks=KCHUNKS, sweeps=SWEEPS, margin: float = 0.01) -> dict:
    """Per size class present in `grids` (gas.Grid objects; bs, xs: device
    pointers or arrays of the right-hand sides and solutions, one per grid),
    the (kchunk, sweeps0) that makes the whole batch fastest.  Starting from
    the library's defaults, each class in turn, largest first, tries every
    candidate with the other classes held at their current choice, timing the
    whole batch, and keeps one that converges and is faster by more than
    `margin`.  Returns class name -> Choice (seconds: the batch's time after
    that class's turn), and under the key "_default" the batch's time with the
    library's defaults (Choice(-1, 1, seconds))."""
    present = [name for _, name in SIZE_CLASSES if any(size_class(G.scen.n) == name for G in grids)]
    table = {name: Choice(-1, 1, float("nan")) for name in present}

    def run(tab):
        kcs, sws = settings(grids, tab)
        return _time_batch(ctx, grids, bs, xs, params, kcs, sws, repeats)

    best, ok = run(table)
    if not ok:
        best = float("inf")
    base = best
    after = {}
    for name in reversed(present):
        for kc in kchunks:
            for nu in sweeps:
                trial = dict(table)
                trial[name] = Choice(kc, nu, float("nan"))
                t, ok = run(trial)
                if ok and t < best * (1.0 - margin):
                    best, table = t, trial
        after[name] = best
    out = {name: Choice(c.kchunk, c.sweeps0, after[name]) for name, c in table.items()}
    out["_default"] = Choice(-1, 1, base)
    return out


def default_table() -> dict:
    """The settings without tuning: the library's own block depth, one sweep."""
    return {name: Choice(-1, 1, float("nan")) for _, name in SIZE_CLASSES}


def table_to_json(table: dict) -> dict:
    def kc(v):
        return "whole column" if v == 0 else "library default" if v < 0 else v
    out = {k: {"kchunk": kc(v.kchunk), "sweeps0": v.sweeps0,
               "batch_seconds_after_its_turn": None if np.isnan(v.seconds) else round(v.seconds, 6)}
           for k, v in table.items() if not k.startswith("_")}
    if "_default" in table:
        out["batch_seconds_with_defaults"] = round(table["_default"].seconds, 6)
    return out
