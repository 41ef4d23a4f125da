"""The reference's gas step (Solver::step, src/sim/solver.cpp:226-269) for a
gas-only world, run through the oracle's restatements (oracle.c,
oracle_step.c) on the host.

TEST INFRASTRUCTURE ONLY: replays scenarios the GPU ran (tests/), never used
by the photon_amd package.  The pressure solve is the oracle's FP32 MG-PCG
with its exact coarsest solve and the relative-L2 stop rule (the serving
loop's), warm-started from the last pressure.
"""
from __future__ import annotations

import numpy as np

from . import pyoracle as po

F32 = np.float32
P0, T0 = F32(101325.0), F32(293.15)
GAMMA = F32(1100.0) / (F32(1100.0) - F32(287.0))
SPECIES = ("Yo", "Yf", "Ys", "Ya", "Yp", "Yc", "Yw", "Yco", "Yai")
FIELDS = ("T",) + SPECIES


class OracleGas:
    def __init__(self, grid, tol=1e-6, dt_max_profile=0.05, cfl=0.8):
        g = self.g = grid
        n = g.n
        self.H = po.Hierarchy(g)
        self.tol = tol
        self.dt_max_profile, self.cfl = dt_max_profile, cfl
        self.u, self.v, self.w = (np.zeros(len(a), np.float32) for a in (g.fx, g.fy, g.fz))
        self.f = {n_: np.zeros(n, np.float32) for n_ in FIELDS + ("Pb", "massSrc")}
        self.f["T"][:] = T0
        self.f["Yo"][:] = 0.232
        self.f["Yc"][:] = 0.0006
        self.f["Yw"][:] = 0.0073
        self.f["Pb"][:] = P0
        self.p, self.S, self.nu = (np.zeros(n, np.float32) for _ in range(3))
        self.Q, self.Qr, self.hrr = (np.zeros(n, np.float32) for _ in range(3))
        self.pilot, self.lit = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
        self.zone, self.zone_open = g.zone, g.zone_open
        self.zoneP = np.full(len(self.zone_open), P0, np.float32)
        cnt = np.bincount(self.zone[self.zone >= 0], minlength=len(self.zone_open)).astype(np.float64)
        dxd = np.float64(F32(g.dx))
        self.zone_vol = (cnt * dxd * dxd * dxd).astype(np.float32)
        self.time = F32(0)
        self.max_vel = self.max_nu = self.max_div = 0.0
        self.burners = []
        self.iters = []

    def add_burner_cells(self, cells, peak_w):
        self.burners.append((np.asarray(cells, np.int32), F32(peak_w)))

    def step(self, dt_max):
        g, f = self.g, self.f
        dt = F32(po.stable_dt(dt_max, self.dt_max_profile, self.cfl, g.dx, self.max_vel, self.max_nu, self.max_div))
        self.Q[:] = 0
        self.Qr[:] = 0
        self.pilot[:] = 0
        f["massSrc"][:] = 0
        for cells, peak in self.burners:      # injectBurners, constant power
            mf = peak * dt / F32(25e6) / F32(len(cells))
            po.add_gas(g, cells, mf, 0.0, mf * F32(0.01), 800.0, True, f, self.pilot)
        po.combustion(g, dt, False, f, self.Q, self.Qr, self.hrr, self.pilot, self.lit)
        self.nu, self.max_nu = po.turbulence(g, self.u, self.v, self.w, f["T"])
        po.heat_diffusion(g, dt, f["T"], self.nu, f["Pb"], self.Q)
        self.max_div = po.energy_expansion(g, dt, f["Pb"], self.Q, f["massSrc"], f["T"], self.S)
        self.u, self.v, self.w = po.momentum(g, dt, self.u, self.v, self.w, f["T"], f["Pb"], self.nu, self.p)
        # leakage without cracks: closed zones' rise rate and source shift
        V = F32(g.dx) * F32(g.dx) * F32(g.dx)
        zdp = np.zeros(len(self.zone_open), np.float32)
        for z in np.flatnonzero(self.zone_open == 0):
            sel = self.zone == z
            dI = F32(np.sum((self.S[sel] * V).astype(np.float64)))
            Vz = self.zone_vol[z]
            zdp[z] = GAMMA * self.zoneP[z] * dI / Vz
            self.S[sel] -= dI / Vz
        o = po.project(self.H, g, dt, self.u, self.v, self.w, self.p, self.S, tol=self.tol, max_iters=400,
                       stop=po.STOP_REL_L2, coarse=po.COARSE_EXACT)
        self.iters.append(o["iters"])
        self.max_vel = o["max_vel"]
        out = po.advect_scalars(g, dt, self.u, self.v, self.w, f["Pb"], self.nu, [f[n] for n in FIELDS])
        for n, a in zip(FIELDS, out):
            f[n] = a
        for z in np.flatnonzero(self.zone_open == 0):
            self.zoneP[z] = min(max(self.zoneP[z] + dt * zdp[z], P0 - F32(20000)), P0 + F32(50000))
            f["Pb"][self.zone == z] = self.zoneP[z]
        self.time = self.time + dt
        return float(dt)
