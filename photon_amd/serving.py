"""Device-resident request loop with continuous batching of fire scenarios.

The reference runs background jobs one at a time: a runner polls a job
directory, builds one Solver and steps it to the end before taking the next
(src/app/cli.cpp:23-101).  Here a single GPU context serves a stream of
scenarios:

- requests wait in a FIFO queue and are admitted when their estimated HBM
  footprint fits the budget and a slot is free, at any step boundary (so work
  that arrives while others run joins them at once);
- a request's host set-up (geometry, multigrid hierarchy with its dense
  coarsest inverse, zones: gas.GridPlan) starts on a worker thread when it
  arrives, while the device steps the admitted scenarios; admission then only
  installs the grid and fields on the device (it waits for the set-up only
  when nothing else is running);
- every scenario keeps its fields resident on the device; a step advances all
  admitted scenarios together: energyAndExpansion, the closed-zone source
  shift of leakage (sealed rooms), the right-hand side, ONE
  batched MG-PCG launch sequence for all their pressure solves, the velocity
  correction and advectScalars (src/sim/solver.cpp:252-257, the phases of
  Solver::step this library implements);
- a scenario that finishes releases its memory and slot at the end of the step.

Per-system block depth and sweeps are fixed by the grid alone (not by the mix),
and a system's arithmetic does not depend on the other systems of a batch, so a
scenario's results are the same bit for bit alone and in any mix.
"""
from __future__ import annotations

import collections
import concurrent.futures
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import gas
from . import scenarios as sc

# device bytes per cell: fields (u v w, 16 cell fields, 10 advection
# temporaries, b, diag), geometry and hierarchy, and the solver's work vectors
# (the step batch and the grid's own one-system batch)
BYTES_PER_CELL = 268
# the dense coarsest inverse is at most 2048^2 FP64
COARSE_BYTES_MAX = 2048 * 2048 * 8


@dataclass
class JobSpec:
    job_id: int
    seed: int
    cells: int            # requested size; the grid has aspect 4 : 3 : 2.4
    steps: int
    hrr_kw: float
    arrival_s: float = 0.0


def grid_for(spec: JobSpec) -> sc.Grid:
    """A scenario grid of about `spec.cells` cells with the room's aspect
    (4 : 3 : 2.4 m, as the 20x15x12 room), walls and openings scaled to its
    floor area, 0.2 m cells."""
    nx = max(8, int(round((spec.cells / 0.45) ** (1.0 / 3.0))))
    ny = max(6, int(round(0.75 * nx)))
    nz = max(5, int(round(0.6 * nx)))
    scale = (nx * ny) / (80 * 50)
    nw = max(2, int(round(40 * scale)))
    no = max(1, int(round(10 * scale)))
    return sc.house(spec.seed, nx, ny, nz, 0.2, nw, no)


def footprint_bytes(g: sc.Grid) -> int:
    return BYTES_PER_CELL * g.n + COARSE_BYTES_MAX


def kchunk_for(g: sc.Grid) -> int:
    """Planes per level-0 block from the grid alone: whole columns for short
    grids, 16-plane chunks for tall ones."""
    return g.nz if g.nz <= 40 else 16


@dataclass
class Job:
    spec: JobSpec
    grid: sc.Grid
    solver: gas.GasSolver
    bytes: int
    admitted_s: float
    steps_done: int = 0
    iters: list = field(default_factory=list)


class Server:
    """Continuous-batching server on one GPU context."""

    def __init__(self, ctx: gas.Context, hbm_budget_bytes: float = 64e9, max_active: int = 64,
                 precision: int = gas.PREC_FP32, dt: float = 0.01, tol: float = 1e-6, max_iters: int = 400,
                 refine: int = 1, warm_start: bool = False, prep_workers: int = 2):
        self.ctx = ctx
        # host set-up of queued requests (0: done at admission, on this thread)
        self._pool = concurrent.futures.ThreadPoolExecutor(prep_workers) if prep_workers > 0 else None
        self._prep: dict = {}
        self.budget = hbm_budget_bytes
        self.max_active = max_active
        self.precision = precision
        self.dt = dt
        # Without the momentum phase (outside this library) the velocity only
        # changes by the projection itself, so a step's right-hand side is
        # small, and a warm start from the last pressure sits far from its
        # solution in relative terms; each step therefore solves from zero
        # unless asked otherwise (the reference warm-starts, gasphase.cpp:576).
        self.params = gas.solve_params(tol=tol, max_iters=max_iters, demean=True, check_every=8,
                                       zero_x=not warm_start, refine=refine)
        self.queue: collections.deque = collections.deque()
        self.active: list[Job] = []
        self.finished: list[Job] = []
        self.used = 0
        self._batch = None
        self._batch_ids = None
        self.step_latencies: list[float] = []   # one entry per job-step
        self.steps_run = 0
        # host seconds of the loop's thread spent admitting (device grid and
        # fields, plus geometry and hierarchy when prep_workers is 0, or the
        # wait for them when nothing else runs), rebuilding the batch after
        # the active set changed, and stepping
        self.t_admit = self.t_rebuild = self.t_step = self.t_solve = 0.0
        self.steps = 0

    # ---- requests --------------------------------------------------------------
    def submit(self, spec: JobSpec):
        self.queue.append(spec)
        if self._pool is not None:
            self._prep[spec.job_id] = self._pool.submit(self._prepare, spec)

    def _prepare(self, spec: JobSpec):
        """Host work of a request: its geometry, grid plan and burner field."""
        g = grid_for(spec)
        return g, gas.GridPlan(g), self._burner_heat(g, spec)

    def _admit(self, now_s: float):
        t0 = time.perf_counter()
        try:
            self._admit_inner(now_s)
        finally:
            self.t_admit += time.perf_counter() - t0

    def _admit_inner(self, now_s: float):
        while self.queue and len(self.active) < self.max_active:
            spec = self.queue[0]
            if self._pool is not None:
                fut = self._prep[spec.job_id]
                if not fut.done() and self.active:
                    break   # set-up still running: admit at a later step boundary
                g, plan, q = fut.result()
            else:
                g, plan, q = self._prepare_inline(spec)
            need = footprint_bytes(g)
            if self.used + need > self.budget and self.active:
                break   # wait for memory to free up (a lone job is always admitted)
            self.queue.popleft()
            self._prep.pop(spec.job_id, None)
            solver = gas.GasSolver(self.ctx, g, precision=self.precision, plan=plan)
            solver.set("Q", q)
            self.active.append(Job(spec, g, solver, need, now_s))
            self.used += need

    def _prepare_inline(self, spec: JobSpec):
        g = grid_for(spec)
        return g, None, self._burner_heat(g, spec)

    def _burner_heat(self, g: sc.Grid, spec: JobSpec) -> np.ndarray:
        """Q (J per step) of a 2x2-cell floor burner of spec.hrr_kw at a seeded
        gas floor cell (the burner of the room case)."""
        rng = sc.Rng(spec.seed ^ 0xB0B0)
        blk = g.blocked.reshape(g.nz, g.ny, g.nx)
        floor = [(i, j) for j in range(g.ny - 1) for i in range(g.nx - 1) if not blk[0, j:j + 2, i:i + 2].any()]
        i0, j0 = floor[rng.randint(0, len(floor) - 1)]
        q = np.zeros(g.n, np.float32)
        for j in (j0, j0 + 1):
            for i in (i0, i0 + 1):
                q[i + g.nx * j] = spec.hrr_kw * 1e3 * self.dt / 4.0
        return q

    # ---- one step of every admitted scenario ---------------------------------------
    def _batch_for_active(self):
        ids = tuple(j.spec.job_id for j in self.active)
        if ids != self._batch_ids:
            t0 = time.perf_counter()
            if self._batch is not None:
                self._batch.destroy()
            Gs = [j.solver.grid for j in self.active]
            self._batch = gas.Batch(self.ctx, Gs, sys_kchunk=[kchunk_for(j.grid) for j in self.active])
            self._batch_ids = ids
            self.t_rebuild += time.perf_counter() - t0
        return self._batch

    def step(self, now_s: float = 0.0) -> int:
        """Admit what fits, advance every active scenario by one step, retire
        the finished ones.  Returns the number of scenarios stepped."""
        self._admit(now_s)
        if not self.active:
            return 0
        t0 = time.perf_counter()
        dt = self.dt
        for j in self.active:
            j.solver.energy_and_expansion(dt, report=False)   # queued: no wait for maxDiv
            j.solver.leakage()                                # sealed rooms: a solvable pressure equation
            j.solver.rhs(dt)
            # the divergence of the corrected velocity sums to zero over a sealed
            # room only up to FP32 rounding; with no momentum phase that
            # rounding is a large part of b, so b is projected the same way
            j.solver.leakage(field="b")
        B = self._batch_for_active()
        ts = time.perf_counter()
        stats = B.solve([j.solver.d["b"] for j in self.active], [j.solver.d["p"] for j in self.active], self.params)
        self.t_solve += time.perf_counter() - ts
        for j, st in zip(self.active, stats):
            j.solver.correct(dt, report=False)
            j.solver.advect_scalars(dt)
            j.iters.append(st["iters"])
            j.steps_done += 1
        self.ctx.synchronize()
        lat = time.perf_counter() - t0
        self.t_step += lat
        self.steps += 1
        n = len(self.active)
        self.step_latencies.extend([lat] * n)
        self.steps_run += n
        done = [j for j in self.active if j.steps_done >= j.spec.steps]
        if done:
            self.active = [j for j in self.active if j.steps_done < j.spec.steps]
            for j in done:
                self.used -= j.bytes
                self.finished.append(j)
        return n

    def release(self, job: Job):
        """Free a finished job's device state."""
        s = job.solver
        for a in list(s.d.values()) + s._tmp:
            a.free()
        s.grid.destroy()

    def close(self):
        if self._pool is not None:
            for fut in self._prep.values():
                fut.cancel()
            self._pool.shutdown(wait=True)
            self._pool = None
            self._prep.clear()
        if self._batch is not None:
            self._batch.destroy()
            self._batch = None
        for j in self.active + self.finished:
            self.release(j)
        self.active, self.finished = [], []


def poisson_stream(seed: int, rate_per_s: float, window_s: float, steps: int, cells_lo: int = 3600,
                   cells_hi: int = 1_280_000) -> list[JobSpec]:
    """Poisson arrivals over `window_s` seconds; sizes log-uniform in
    [cells_lo, cells_hi]; burners uniform in [10, 500] kW."""
    rng = sc.Rng(seed)
    t, out, k = 0.0, [], 0
    while True:
        t += -math.log(max(rng.uniform(), 1e-300)) / rate_per_s
        if t > window_s:
            break
        cells = int(round(math.exp(math.log(cells_lo) + (math.log(cells_hi) - math.log(cells_lo)) * rng.uniform())))
        out.append(JobSpec(k, 1000 + k, cells, steps, 10.0 + 490.0 * rng.uniform(), t))
        k += 1
    return out


def serve(ctx, specs: list[JobSpec], **server_kw) -> dict:
    """Run the stream to completion against the wall clock: a request joins
    the queue when its arrival time has passed.  Reports throughput (job-steps
    per second) and per-step latency percentiles."""
    srv = Server(ctx, **server_kw)
    pending = sorted(specs, key=lambda s: s.arrival_s)
    t0 = time.perf_counter()
    i = 0
    waits = []
    iters_all = []
    while i < len(pending) or srv.queue or srv.active:
        now = time.perf_counter() - t0
        while i < len(pending) and pending[i].arrival_s <= now:
            srv.submit(pending[i])
            i += 1
        before = {j.spec.job_id for j in srv.active}
        if srv.step(now) == 0 and i < len(pending):
            time.sleep(max(0.0, min(0.01, pending[i].arrival_s - now)))
        for j in srv.active:
            if j.spec.job_id not in before:
                waits.append(j.admitted_s - j.spec.arrival_s)
        for j in list(srv.finished):
            iters_all.append(j.iters)
            srv.release(j)
            srv.finished.remove(j)
    wall = time.perf_counter() - t0
    its = np.concatenate([np.asarray(v, np.int64) for v in iters_all]) if iters_all else np.zeros(1, np.int64)
    lat = np.array(srv.step_latencies) if srv.step_latencies else np.zeros(1)
    out = {"jobs": len(specs), "job_steps": srv.steps_run, "wall_s": round(wall, 3),
           "steps_per_s": round(srv.steps_run / wall, 1),
           "step_latency_ms_p50": round(1e3 * float(np.percentile(lat, 50)), 3),
           "step_latency_ms_p99": round(1e3 * float(np.percentile(lat, 99)), 3),
           "admission_wait_s_mean": round(float(np.mean(waits)), 3) if waits else 0.0,
           "loop_steps": srv.steps,
           "pcg_iters_mean": round(float(its.mean()), 2), "pcg_iters_p99": int(np.percentile(its, 99)),
           "pcg_iters_max": int(its.max()),
           "host_s": {"admit": round(srv.t_admit, 3), "batch_rebuild": round(srv.t_rebuild, 3),
                      "steps": round(srv.t_step, 3), "of_which_solve": round(srv.t_solve, 3)}}
    srv.close()
    return out
