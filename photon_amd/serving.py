"""Device-resident request loop with continuous batching of fire scenarios.

The reference runs background jobs one at a time: a runner reads a job file,
builds one Solver, steps it to the end, reports through a status file and
obeys "pause", "resume" and "cancel" written to a control file
(src/app/cli.cpp:23-101).  Here a single GPU context serves a stream of
scenarios:

- requests wait in a FIFO queue and are admitted when their estimated HBM
  footprint fits the budget and a slot is free, at any step boundary (work
  that arrives while others run joins them at once);
- a request's host set-up (geometry, multigrid hierarchy with its dense
  coarsest inverse, zones: gas.GridPlan) starts on a worker thread when it
  arrives, while the device steps the admitted scenarios;
- every scenario keeps its whole gas state resident on the device, its zone
  background pressures included, so a step queues its work without waiting
  for the device until the solve.  A loop step advances every running
  scenario by one Solver::step
  (src/sim/solver.cpp:226-269) at its own stable dt: burners, combustion,
  eddy viscosity, conduction, energy and expansion, momentum, the closed
  zones' pressure rise and source shift, the right-hand side, then ONE batched
  MG-PCG launch sequence for all their pressure solves, the velocity
  correction, scalar transport and the zone background pressure;
- control as the reference's runner: pause (stays resident, not stepped),
  resume, cancel, and evict (a paused scenario's state moved to host memory,
  its HBM freed; resume re-admits it);
- a scenario whose pressure solve does not converge (iteration cap,
  breakdown, or a NaN / infinite residual) is retired as failed at that step;
  the batch's systems are independent, so its neighbours' results are the
  same bit for bit with or without it;
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

# device bytes per cell: fields (u v w and their momentum copies, 21 cell
# fields, 10 transport temporaries, b, diag, flags), geometry and hierarchy,
# and the solver's work vectors (the step batch and the grid's one-system batch)
BYTES_PER_CELL = 320
# the dense coarsest inverse is at most 2048^2 FP64
COARSE_BYTES_MAX = 2048 * 2048 * 8

RUNNING, PAUSED, EVICTED, DONE, CANCELLED, FAILED = "running", "paused", "evicted", "done", "cancelled", "failed"


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


def burner_rect(g: sc.Grid, spec: JobSpec):
    """A 0.4 x 0.4 m floor burner over a seeded 2x2 patch of gas cells: the
    rectangle (x0, y0, x1, y1) that holds exactly those cell centres."""
    rng = sc.Rng(spec.seed ^ 0xB0B0)
    blk = g.blocked.reshape(g.nz, g.ny, g.nx)
    floor = [(i, j) for j in range(g.ny - 1) for i in range(g.nx - 1) if not blk[0, j:j + 2, i:i + 2].any()]
    i0, j0 = floor[rng.randint(0, len(floor) - 1)]
    return ((i0 + 0.25) * g.dx, (j0 + 0.25) * g.dx, (i0 + 1.75) * g.dx, (j0 + 1.75) * g.dx)


@dataclass
class Job:
    spec: JobSpec
    grid: sc.Grid
    solver: gas.GasSolver | None
    bytes: int
    admitted_s: float
    steps_done: int = 0
    iters: list = field(default_factory=list)
    state: str = RUNNING
    reason: str = ""
    snap: dict | None = None      # host copy of an evicted job's state
    dt: float = 0.0
    result: dict | None = None    # fields copied to the host when it finished (keep_fields)


def solve_failed(st) -> str:
    """Why a pressure solve's statistics mark its scenario as failed ("" if not)."""
    if st["status"] == gas.NONFINITE or not np.isfinite(st["relres"]):
        return "non-finite residual"
    if st["status"] == gas.MAXITER:
        return "iteration cap"
    if st["status"] in (gas.BREAKDOWN, gas.BREAKDOWN_RZ):
        return "breakdown"
    return ""


class Server:
    """Continuous-batching server on one GPU context."""

    def __init__(self, ctx: gas.Context, hbm_budget_bytes: float = 64e9, max_active: int = 64,
                 precision: int = gas.PREC_MIXED_F16, dt_max: float = 0.01, tol: float = 1e-6,
                 max_iters: int = 400, refine: int = 1, prep_workers: int = 2, keep_fields=(),
                 coarse_weight: float = gas.COARSE_WEIGHT):
        self.ctx = ctx
        self.keep_fields = tuple(keep_fields)   # fields copied to the host when a job ends
        # host set-up of queued requests (0: done at admission, on this thread)
        self._pool = concurrent.futures.ThreadPoolExecutor(prep_workers) if prep_workers > 0 else None
        self._prep: dict = {}
        self.budget = hbm_budget_bytes
        self.max_active = max_active
        self.precision = precision
        self.dt_max = dt_max
        # warm start from the last pressure (gasphase.cpp:576)
        # the over-corrected V-cycle: a third fewer iterations to the same stop rule
        self.params = gas.solve_params(tol=tol, max_iters=max_iters, demean=True, check_every=8, refine=refine,
                                       coarse_weight=coarse_weight)
        self.queue: collections.deque = collections.deque()
        self.active: list[Job] = []          # resident: running or paused
        self.evicted: dict[int, Job] = {}
        self.finished: list[Job] = []
        self.jobs: dict[int, Job] = {}
        self.used = 0
        self._batch = None
        self._batch_ids = None
        self.step_latencies: list[float] = []   # one entry per job-step
        self.steps_run = 0
        self.t_admit = self.t_rebuild = self.t_step = self.t_solve = 0.0
        self.steps = 0

    # ---- requests and control ------------------------------------------------------
    def submit(self, spec: JobSpec):
        self.queue.append(spec)
        if self._pool is not None:
            self._prep[spec.job_id] = self._pool.submit(self._prepare, spec)

    def control(self, job_id: int, cmd: str) -> str:
        """"pause", "resume", "cancel" (as the reference's control file) or
        "evict" (a paused job's state to host memory, its HBM freed).  Returns
        the job's state afterwards."""
        j = self.jobs.get(job_id)
        if j is None:   # still queued
            if cmd == "cancel":
                for spec in list(self.queue):
                    if spec.job_id == job_id:
                        self.queue.remove(spec)
                        fut = self._prep.pop(job_id, None)
                        if fut is not None:
                            fut.cancel()
                        return CANCELLED
            return "queued"
        if j.state in (DONE, CANCELLED, FAILED):
            return j.state
        if cmd == "pause" and j.state == RUNNING:
            j.state = PAUSED
        elif cmd == "resume" and j.state == PAUSED:
            j.state = RUNNING
        elif cmd == "resume" and j.state == EVICTED:
            self._readmit(j)
        elif cmd == "evict" and j.state in (RUNNING, PAUSED):
            j.snap = j.solver.snapshot()
            self._retire(j, EVICTED)
            self.evicted[job_id] = j
        elif cmd == "cancel":
            if j.state == EVICTED:
                self.evicted.pop(job_id, None)
                j.state, j.snap = CANCELLED, None
                self.finished.append(j)
            else:
                self._retire(j, CANCELLED)
        return j.state

    def status(self, job_id: int) -> dict:
        """The reference runner's status record (cli.cpp:27-37): state,
        simulated time, steps, the last pressure iterations, hottest gas."""
        j = self.jobs.get(job_id)
        if j is None:
            return {"state": "queued" if any(s.job_id == job_id for s in self.queue) else "unknown"}
        out = {"state": j.state, "steps": j.steps_done, "reason": j.reason,
               "pressure_iters": j.iters[-1] if j.iters else 0}
        if j.solver is not None:
            out["sim_time"] = j.solver.time
            out["max_c"] = float(j.solver.get("T").max()) - 273.15
        elif j.snap is not None:
            out["sim_time"] = float(j.snap["time"])
        return out

    def _prepare(self, spec: JobSpec):
        """Host work of a request: its geometry and grid plan."""
        g = grid_for(spec)
        return g, gas.GridPlan(g)

    def _new_solver(self, g, plan, spec):
        s = gas.GasSolver(self.ctx, g, precision=self.precision, plan=plan, graph_steps=False, resident_zones=True)
        x0, y0, x1, y1 = burner_rect(g, spec)
        s.add_burner(x0, y0, x1, y1, 0.0, spec.hrr_kw * 1e3)
        return s

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
                g, plan = fut.result()
            else:
                g, plan = grid_for(spec), None
            need = footprint_bytes(g)
            if self.used + need > self.budget and self.active:
                break   # wait for memory to free up (a lone job is always admitted)
            self.queue.popleft()
            self._prep.pop(spec.job_id, None)
            j = Job(spec, g, self._new_solver(g, plan, spec), need, now_s)
            self.active.append(j)
            self.jobs[spec.job_id] = j
            self.used += need

    def _readmit(self, j: Job):
        if self.used + j.bytes > self.budget and self.active:
            return   # stays evicted until memory frees up
        j.solver = self._new_solver(j.grid, None, j.spec)
        j.solver.restore(j.snap)
        j.snap = None
        j.state = RUNNING
        self.evicted.pop(j.spec.job_id, None)
        self.active.append(j)
        self.used += j.bytes

    def _retire(self, j: Job, state: str, reason: str = ""):
        j.state, j.reason = state, reason
        if j in self.active:
            self.active.remove(j)
            self.used -= j.bytes
        if j.solver is not None:
            if self.keep_fields and state in (DONE, FAILED, CANCELLED):
                j.result = {f: j.solver.get(f) for f in self.keep_fields}
            self._release_solver(j.solver)
            j.solver = None
        if state != EVICTED:
            self.finished.append(j)

    # ---- one step of every running scenario ----------------------------------------
    def _batch_for(self, jobs):
        ids = tuple(j.spec.job_id for j in jobs)
        if ids != self._batch_ids:
            t0 = time.perf_counter()
            if self._batch is not None:
                self._batch.destroy()
            self._batch = gas.Batch(self.ctx, [j.solver.grid for j in jobs],
                                    sys_kchunk=[kchunk_for(j.grid) for j in jobs])
            self._batch_ids = ids
            self.t_rebuild += time.perf_counter() - t0
        return self._batch

    def step(self, now_s: float = 0.0) -> int:
        """Admit what fits, advance every running scenario by one step, retire
        the finished and failed ones.  Returns the number of scenarios stepped."""
        self._admit(now_s)
        run = [j for j in self.active if j.state == RUNNING]
        if not run:
            return 0
        t0 = time.perf_counter()
        for j in run:   # queued on the stream, no waits
            s = j.solver
            j.dt = dt = s.stable_dt(self.dt_max)
            s.step_clear()
            s.inject_burners(dt)
            s.combustion(dt, report=False)
            s.turbulence(report=False)
            s.heat_diffusion(dt)
            s.energy_and_expansion(dt, report=False)
            s.momentum(dt)
        for j in run:   # closed zones' integrals: the first call waits for the phases above
            j.solver.leakage_zones(j.dt)
            j.solver.rhs(j.dt)
        B = self._batch_for(run)
        ts = time.perf_counter()
        stats = B.solve([j.solver.d["b"] for j in run], [j.solver.d["p"] for j in run], self.params)
        self.t_solve += time.perf_counter() - ts
        ok = []
        for j, st in zip(run, stats):
            j.iters.append(st["iters"])
            why = solve_failed(st)
            if why:
                self._retire(j, FAILED, why)
                continue
            s = j.solver
            s.correct(j.dt, report=False)
            s.advect_scalars(j.dt)
            s.zone_pressure(j.dt)
            s.time = float(np.float32(s.time) + np.float32(j.dt))
            s.pressureIters = st["iters"]
            j.steps_done += 1
            ok.append(j)
        for j in ok:
            j.solver.read_stats()   # maxVel, maxNu, maxDiv for the next stableDt
        lat = time.perf_counter() - t0
        self.t_step += lat
        self.steps += 1
        self.step_latencies.extend([lat] * len(run))
        self.steps_run += len(run)
        for j in [j for j in ok if j.steps_done >= j.spec.steps]:
            self._retire(j, DONE)
        return len(run)

    def _release_solver(self, s: gas.GasSolver):
        for a in list(s.d.values()) + s._tmp:
            a.free()
        for b in s.burners:
            b["dcells"].free()
        s.grid.destroy()

    def release(self, job: Job):
        """Forget a finished job (its device state is already freed)."""
        if job in self.finished:
            self.finished.remove(job)

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
        for j in list(self.active):
            self._retire(j, CANCELLED)
        self.active, self.finished, self.evicted = [], [], {}


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


def serve(ctx, specs: list[JobSpec], progress=None, **server_kw) -> dict:
    """Run the stream to completion against the wall clock: a request joins
    the queue when its arrival time has passed.  Reports throughput (job-steps
    per second) and per-step latency percentiles.  progress(loop steps,
    job-steps, jobs finished), if given, is called every 500 loop steps."""
    srv = Server(ctx, **server_kw)
    pending = sorted(specs, key=lambda s: s.arrival_s)
    t0 = time.perf_counter()
    i = 0
    waits = []
    iters_all = []
    states = collections.Counter()
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
            states[j.state] += 1
            srv.release(j)
        if progress is not None and srv.steps and srv.steps % 500 == 0:
            progress(srv.steps, srv.steps_run, sum(states.values()))
    wall = time.perf_counter() - t0
    its = np.concatenate([np.asarray(v, np.int64) for v in iters_all]) if iters_all else np.zeros(1, np.int64)
    lat = np.array(srv.step_latencies) if srv.step_latencies else np.zeros(1)
    out = {"jobs": len(specs), "job_steps": srv.steps_run, "wall_s": round(wall, 3),
           "steps_per_s": round(srv.steps_run / wall, 1),
           "step_latency_ms_p50": round(1e3 * float(np.percentile(lat, 50)), 3),
           "step_latency_ms_p99": round(1e3 * float(np.percentile(lat, 99)), 3),
           "admission_wait_s_mean": round(float(np.mean(waits)), 3) if waits else 0.0,
           "loop_steps": srv.steps, "job_states": dict(states),
           "pcg_iters_mean": round(float(its.mean()), 2), "pcg_iters_p99": int(np.percentile(its, 99)),
           "pcg_iters_max": int(its.max()),
           "host_s": {"admit": round(srv.t_admit, 3), "batch_rebuild": round(srv.t_rebuild, 3),
                      "steps": round(srv.t_step, 3), "of_which_solve": round(srv.t_solve, 3)}}
    srv.close()
    return out
