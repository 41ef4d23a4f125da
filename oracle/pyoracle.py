"""ctypes bindings of the CPU checker: the C restatement (liboracle.so) and the
reference's own solver (_ref/libphoton_ref.so, when it has been built).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the photon_amd package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(HERE, "liboracle.so")
_REF = os.path.join(HERE, "_ref", "libphoton_ref.so")          # the reference's GCC flags
_REF_PIN = os.path.join(HERE, "_ref", "libphoton_ref_pin.so")  # same, -ffp-contract=off

COARSE_GS20, COARSE_EXACT = 0, 1
STOP_MAXNORM, STOP_REL_L2 = 0, 1

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


class ProjectArgs(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dx", C.c_float), ("dt", C.c_float),
                ("fx", C.c_void_p), ("fy", C.c_void_p), ("fz", C.c_void_p),
                ("ax", C.c_void_p), ("ay", C.c_void_p), ("az", C.c_void_p),
                ("blocked", C.c_void_p), ("zone", C.c_void_p), ("zone_open", C.c_void_p),
                ("wind_closed", C.c_void_p), ("tol", C.c_float), ("max_iters", C.c_int),
                ("stop_mode", C.c_int), ("coarse_mode", C.c_int)]


class ProjectOut(C.Structure):
    _fields_ = [("iters", C.c_int), ("residual", C.c_float), ("max_vel", C.c_float), ("relres", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError(f"oracle not built: {_LIB} (run make -C oracle)")
        L = C.CDLL(_LIB)
        L.orc_mg_build.restype = C.c_void_p
        L.orc_mg_build.argtypes = [C.c_int, C.c_int, C.c_int, _f32p, _f32p, _f32p, _u8p, C.c_void_p]
        L.orc_mg_free.argtypes = [C.c_void_p]
        L.orc_mg_nlevels.argtypes = [C.c_void_p]
        L.orc_mg_level_n.argtypes = [C.c_void_p, C.c_int]
        L.orc_mg_level_nnz.argtypes = [C.c_void_p, C.c_int]
        L.orc_mg_level_get.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 9
        L.orc_precondition_f32.argtypes = [C.c_void_p, _f32p, _f32p, C.c_int]
        L.orc_precondition_f64.argtypes = [C.c_void_p, _f64p, _f64p, C.c_int]
        L.orc_apply_f32.restype = C.c_double
        L.orc_apply_f32.argtypes = [C.c_void_p, _f32p, _f32p]
        L.orc_coarse_solve_f64.argtypes = [C.c_void_p, _f64p, _f64p]
        L.orc_pcg_f32.argtypes = [C.c_void_p, _f32p, _f32p, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int,
                                  C.POINTER(C.c_double)]
        L.orc_pcg_f64.argtypes = [C.c_void_p, _f64p, _f64p, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.orc_true_relres_f32.restype = C.c_double
        L.orc_true_relres_f32.argtypes = [C.c_void_p, _f32p, _f32p]
        L.orc_true_relres_f64.restype = C.c_double
        L.orc_true_relres_f64.argtypes = [C.c_void_p, _f64p, _f64p]
        L.orc_compute_zones.argtypes = [C.c_int, C.c_int, C.c_int, _f32p, _f32p, _f32p, _u8p, _i32p, _u8p, C.c_int]
        L.orc_project.argtypes = [C.c_void_p, C.POINTER(ProjectArgs), _f32p, _f32p, _f32p, _f32p, _f32p,
                                  C.c_void_p, C.c_void_p, C.POINTER(ProjectOut)]
        L.orc_rhs.argtypes = [C.POINTER(ProjectArgs), _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p]
        L.orc_zone_mean.argtypes = [C.POINTER(ProjectArgs), C.c_int, _f32p]
        L.orc_correct.restype = C.c_float
        L.orc_correct.argtypes = [C.POINTER(ProjectArgs), _f32p, _f32p, _f32p, _f32p]
        L.orc_energy_expansion.restype = C.c_float
        L.orc_energy_expansion.argtypes = [C.c_int, C.c_float, C.c_float, _u8p, _f32p, _f32p, _f32p, _f32p, _f32p]
        L.orc_advect_scalars.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, _f32p, _f32p, _f32p,
                                         _f32p, _f32p, _f32p, _u8p, _f32p, _f32p, C.c_void_p, C.c_void_p]
        L.orc_wind_at.restype = C.c_float
        L.orc_wind_at.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_float]
        L.orc_stable_dt.restype = C.c_float
        L.orc_stable_dt.argtypes = [C.c_float] * 7
        L.orc_turbulence.restype = C.c_float
        L.orc_turbulence.argtypes = [C.POINTER(OrcGrid), _f32p, _f32p, _f32p, _f32p, _f32p]
        L.orc_heat_diffusion.argtypes = [C.POINTER(OrcGrid), C.c_float, _f32p, _f32p, _f32p, _f32p]
        L.orc_momentum.argtypes = [C.POINTER(OrcGrid), C.c_void_p, C.c_float] + [_f32p] * 10
        L.orc_add_gas.argtypes = [C.c_int, _i32p, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int,
                                  _u8p, C.POINTER(OrcSpecies), _u8p]
        L.orc_combustion.restype = C.c_double
        L.orc_combustion.argtypes = [C.POINTER(OrcGrid), C.c_float, C.c_int, C.POINTER(OrcSpecies), _f32p, _f32p,
                                     _f32p, _u8p, _u8p]
        L.orc_leakage.argtypes = [C.POINTER(OrcGrid), C.c_float, _i32p, C.c_int, _u8p, _f32p, _f32p, _f32p, _f32p,
                                  _f32p, _f32p, _f32p, _f32p, _f32p, _f32p]
        _lib = L
    return _lib


class OrcGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dx", C.c_float),
                ("fx", C.c_void_p), ("fy", C.c_void_p), ("fz", C.c_void_p), ("blocked", C.c_void_p),
                ("lx", C.c_void_p), ("ly", C.c_void_p), ("lz", C.c_void_p)]


class OrcWind(C.Structure):
    _fields_ = [("speed", C.c_float), ("from_deg", C.c_float), ("exponent", C.c_float), ("oz", C.c_float)]


SPECIES = ("T", "Pb", "Yo", "Yf", "Ys", "Ya", "Yp", "Yc", "Yw", "Yco", "Yai", "massSrc")


class OrcSpecies(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in SPECIES]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Hierarchy:
    """The oracle's multigrid hierarchy (buildPressureLevels restated)."""

    def __init__(self, grid, wind_closed=None):
        L = lib()
        self.grid = grid
        self._wc = None if wind_closed is None else np.ascontiguousarray(wind_closed, np.uint8)
        self.h = L.orc_mg_build(grid.nx, grid.ny, grid.nz, grid.ax, grid.ay, grid.az, grid.blocked, _ptr(self._wc))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_mg_free(self.h)
            self.h = None

    @property
    def nlevels(self):
        return lib().orc_mg_nlevels(self.h)

    def level(self, l):
        L = lib()
        n = L.orc_mg_level_n(self.h, l)
        nnz = L.orc_mg_level_nnz(self.h, l)
        out = dict(n=n, diag=np.zeros(n, np.float32), agg=np.zeros(n, np.int32), leak=np.zeros(n, np.uint8))
        if l > 0:
            out.update(rowStart=np.zeros(n + 1, np.int32), col=np.zeros(nnz, np.int32), val=np.zeros(nnz, np.float32),
                       bi=np.zeros(n, np.int32), bj=np.zeros(n, np.int32), bk=np.zeros(n, np.int32))
        g = out.get
        L.orc_mg_level_get(self.h, l, _ptr(g("rowStart")), _ptr(g("col")), _ptr(g("val")), _ptr(out["diag"]),
                           _ptr(out["agg"]), _ptr(g("bi")), _ptr(g("bj")), _ptr(g("bk")), _ptr(out["leak"]))
        return out

    def precondition(self, r, coarse=COARSE_GS20, fp64=False):
        if fp64:
            r = np.ascontiguousarray(r, np.float64)
            z = np.zeros_like(r)
            lib().orc_precondition_f64(self.h, r, z, coarse)
        else:
            r = np.ascontiguousarray(r, np.float32)
            z = np.zeros_like(r)
            lib().orc_precondition_f32(self.h, r, z, coarse)
        return z

    def coarse_solve(self, b):
        """The exact coarsest solve alone (FP64) on a coarsest-level vector."""
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros_like(b)
        lib().orc_coarse_solve_f64(self.h, b, x)
        return x

    def apply(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros_like(x)
        dot = lib().orc_apply_f32(self.h, x, y)
        return y, dot

    def pcg(self, b, x0=None, tol=1e-6, max_iters=1000, stop=STOP_REL_L2, coarse=COARSE_GS20, inv=1.0):
        b = np.ascontiguousarray(b, np.float32)
        x = np.zeros_like(b) if x0 is None else np.array(x0, np.float32)
        res = C.c_double(0)
        it = lib().orc_pcg_f32(self.h, b, x, inv, tol, max_iters, stop, coarse, C.byref(res))
        return x, it, res.value

    def pcg64(self, b, tol=1e-12, max_iters=2000, coarse=COARSE_EXACT):
        """The FP64 oracle: b is first projected onto range(A) (its mean is
        removed on every closed zone, whose constant vector spans A's null
        space there), solved to ||r||/||b|| <= tol in FP64, and the solution's
        closed-zone means are set to zero (the reference's convention,
        src/sim/gasphase.cpp:628-639)."""
        b = closed_zone_demean(self.grid, np.array(b, np.float64))
        x = np.zeros_like(b)
        res = C.c_double(0)
        it = lib().orc_pcg_f64(self.h, b, x, tol, max_iters, coarse, C.byref(res))
        return closed_zone_demean(self.grid, x), it, res.value

    def true_relres(self, b, x):
        if x.dtype == np.float64:
            return lib().orc_true_relres_f64(self.h, np.ascontiguousarray(b, np.float64), x)
        return lib().orc_true_relres_f32(self.h, np.ascontiguousarray(b, np.float32),
                                         np.ascontiguousarray(x, np.float32))


def closed_zone_demean(grid, x):
    """Remove the mean of x (FP64) on each closed pressure zone; zero solids."""
    x = np.array(x, np.float64)
    z = grid.zone
    sel = z >= 0
    nzn = len(grid.zone_open)
    sums = np.bincount(z[sel], weights=x[sel], minlength=nzn)
    cnt = np.bincount(z[sel], minlength=nzn)
    mean = np.where((grid.zone_open == 0) & (cnt > 0), sums / np.maximum(cnt, 1), 0.0)
    x[sel] -= mean[z[sel]]
    x[~sel] = 0.0
    return x


def project_args(grid, dt, tol=1e-6, max_iters=1000, stop=STOP_REL_L2, coarse=COARSE_GS20, wind_closed=None):
    a = ProjectArgs()
    a.nx, a.ny, a.nz, a.dx, a.dt = grid.nx, grid.ny, grid.nz, grid.dx, dt
    a._keep = [grid.fx, grid.fy, grid.fz, grid.ax, grid.ay, grid.az, grid.blocked, grid.zone, grid.zone_open]
    a.fx, a.fy, a.fz = _ptr(grid.fx), _ptr(grid.fy), _ptr(grid.fz)
    a.ax, a.ay, a.az = _ptr(grid.ax), _ptr(grid.ay), _ptr(grid.az)
    a.blocked, a.zone, a.zone_open = _ptr(grid.blocked), _ptr(grid.zone), _ptr(grid.zone_open)
    if wind_closed is not None:
        wc = np.ascontiguousarray(wind_closed, np.uint8)
        a._keep.append(wc)
        a.wind_closed = _ptr(wc)
    a.tol, a.max_iters, a.stop_mode, a.coarse_mode = tol, max_iters, stop, coarse
    return a


def project(hier, grid, dt, u, v, w, p, S, **kw):
    """project(dt) restated; u, v, w, p are updated in place."""
    a = project_args(grid, dt, **kw)
    out = ProjectOut()
    b = np.zeros(grid.n, np.float32)
    d = np.zeros(grid.n, np.float32)
    lib().orc_project(hier.h, C.byref(a), u, v, w, p, np.ascontiguousarray(S, np.float32), _ptr(b), _ptr(d),
                      C.byref(out))
    return dict(iters=out.iters, residual=out.residual, max_vel=out.max_vel, relres=out.relres, b=b, diag=d)


def orc_grid(grid, leaks=None):
    """The oracle's view of a scenario grid (kept alive with the grid)."""
    og = OrcGrid(grid.nx, grid.ny, grid.nz, grid.dx, _ptr(grid.fx), _ptr(grid.fy), _ptr(grid.fz),
                 _ptr(grid.blocked), None, None, None)
    if leaks is not None:
        og._leaks = [np.ascontiguousarray(a, np.uint8) for a in leaks]
        og.lx, og.ly, og.lz = (_ptr(a) for a in og._leaks)
    return og


def wind(speed, from_deg, exponent=0.25, oz=0.0):
    return OrcWind(speed, from_deg, exponent, oz)


def wind_at(wd, axis, side, z):
    return lib().orc_wind_at(C.byref(wd) if wd is not None else None, axis, side, z)


def stable_dt(dt_max, profile_dt_max, cfl, dx, max_vel, max_nu, max_div):
    return lib().orc_stable_dt(dt_max, profile_dt_max, cfl, dx, max_vel, max_nu, max_div)


def turbulence(grid, u, v, w, T):
    nu = np.zeros(grid.n, np.float32)
    m = lib().orc_turbulence(C.byref(orc_grid(grid)), *[np.ascontiguousarray(a, np.float32) for a in (u, v, w, T)],
                             nu)
    return nu, m


def heat_diffusion(grid, dt, T, nu, Pb, Q):
    """Q is updated in place."""
    lib().orc_heat_diffusion(C.byref(orc_grid(grid)), dt, *[np.ascontiguousarray(a, np.float32) for a in (T, nu, Pb)],
                             Q)


def momentum(grid, dt, u, v, w, T, Pb, nu, p, wd=None, leaks=None):
    """Returns the new (u, v, w)."""
    outs = [np.zeros(len(a), np.float32) for a in (u, v, w)]
    ins = [np.ascontiguousarray(a, np.float32) for a in (u, v, w, T, Pb, nu, p)]
    lib().orc_momentum(C.byref(orc_grid(grid, leaks)), C.byref(wd) if wd is not None else None, dt, *ins, *outs)
    return outs


def zone_volumes(grid):
    """Each zone's volume as the reference keeps it (gasphase.cpp:155-156):
    the cell count times dx^3 in FP64, then read as a float."""
    zone = grid.zone
    cnt = np.bincount(zone[zone >= 0], minlength=len(grid.zone_open)).astype(np.float64)
    dxd = np.float64(np.float32(grid.dx))
    return (cnt * dxd * dxd * dxd).astype(np.float32)


def leakage(grid, dt, leaks, zoneP, T, Pb, p, u, v, w, S):
    """leakage(dt) restated; u, v, w and S are updated in place.  Returns
    (zoneDP, number of crack faces)."""
    zoneP = np.ascontiguousarray(zoneP, np.float32)
    zdp = np.zeros(max(len(zoneP), 1), np.float32)
    zo = np.ascontiguousarray(grid.zone_open, np.uint8)
    n = lib().orc_leakage(C.byref(orc_grid(grid, leaks)), dt, np.ascontiguousarray(grid.zone, np.int32),
                          len(zoneP), zo, zone_volumes(grid), zoneP,
                          *[np.ascontiguousarray(a, np.float32) for a in (T, Pb, p)], u, v, w, S, zdp)
    return zdp[:len(zoneP)], n


def species(f):
    """An OrcSpecies over a dict of float32 cell arrays (updated in place)."""
    sp = OrcSpecies(*[_ptr(f[n]) for n in SPECIES])
    sp._keep = f
    return sp


def add_gas(grid, cells, fuel_gen, aerosol, soot_pot, ait, set_pilot, f, pilot):
    cells = np.ascontiguousarray(cells, np.int32)
    lib().orc_add_gas(len(cells), cells, grid.dx, fuel_gen, aerosol, soot_pot, ait, 1 if set_pilot else 0,
                      grid.blocked, C.byref(species(f)), pilot)


def combustion(grid, dt, radiation_on, f, Q, Qr, hrr, pilot, lit):
    return lib().orc_combustion(C.byref(orc_grid(grid)), dt, 1 if radiation_on else 0, C.byref(species(f)), Q, Qr,
                                hrr, pilot, lit)


def rhs(grid, dt, u, v, w, S):
    a = project_args(grid, dt)
    b = np.zeros(grid.n, np.float32)
    d = np.zeros(grid.n, np.float32)
    p = np.zeros(grid.n, np.float32)
    lib().orc_rhs(C.byref(a), u, v, w, np.ascontiguousarray(S, np.float32), b, d, p)
    return b, d


def zone_mean(grid, p):
    a = project_args(grid, 1.0)
    lib().orc_zone_mean(C.byref(a), len(grid.zone_open), p)
    return p


def correct(grid, dt, p, u, v, w):
    a = project_args(grid, dt)
    return lib().orc_correct(C.byref(a), np.ascontiguousarray(p, np.float32), u, v, w)


def energy_expansion(grid, dt, Pb, Q, massSrc, T, S):
    return lib().orc_energy_expansion(grid.n, grid.dx, dt, grid.blocked, Pb, Q, massSrc, T, S)


FIELD_NAMES = ("T", "Yo", "Yf", "Ys", "Ya", "Yp", "Yc", "Yw", "Yco", "Yai")


def advect_scalars(grid, dt, u, v, w, Pb, nu, fields):
    ins = [np.ascontiguousarray(f, np.float32) for f in fields]
    outs = [np.zeros(grid.n, np.float32) for _ in range(10)]
    ip = (C.c_void_p * 10)(*[f.ctypes.data for f in ins])
    op = (C.c_void_p * 10)(*[f.ctypes.data for f in outs])
    lib().orc_advect_scalars(grid.nx, grid.ny, grid.nz, grid.dx, dt, u, v, w, grid.fx, grid.fy, grid.fz,
                             grid.blocked, Pb, nu, C.cast(ip, C.c_void_p), C.cast(op, C.c_void_p))
    return outs


def zones(grid):
    zone = np.zeros(grid.n, np.int32)
    zopen = np.zeros(grid.n + 1, np.uint8)
    nzn = lib().orc_compute_zones(grid.nx, grid.ny, grid.nz, grid.ax, grid.ay, grid.az, grid.blocked, zone, zopen,
                                  grid.n + 1)
    return zone, zopen[:nzn]


# ---------------------------------------------------------------------------
# the reference's own solver (oracle/_ref), optional
# ---------------------------------------------------------------------------
def ref_available(pin=True):
    return os.path.exists(_REF_PIN if pin else _REF)


class RefSolver:
    """The reference photon::Solver driven through ref_driver.cpp."""

    _libs = {}

    @classmethod
    def _lib(cls, pin):
        path = _REF_PIN if pin else _REF
        if path not in cls._libs:
            L = C.CDLL(path)
            L.ref_create.restype = C.c_void_p
            L.ref_create.argtypes = [C.c_int]
            L.ref_destroy.argtypes = [C.c_void_p]
            L.ref_setup.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_float] + [_f32p] * 6 + [_u8p]
            L.ref_get_zones.argtypes = [C.c_void_p, _i32p, _u8p]
            L.ref_set_field.argtypes = [C.c_void_p, C.c_char_p, _f32p]
            L.ref_get_field.argtypes = [C.c_void_p, C.c_char_p, _f32p]
            L.ref_project.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_int, C.POINTER(C.c_float),
                                      C.POINTER(C.c_float)]
            L.ref_precondition.argtypes = [C.c_void_p, _f32p, _f32p]
            L.ref_mg_nlevels.argtypes = [C.c_void_p]
            L.ref_mg_level_n.argtypes = [C.c_void_p, C.c_int]
            L.ref_mg_level_nnz.argtypes = [C.c_void_p, C.c_int]
            L.ref_mg_level_has_agg.argtypes = [C.c_void_p, C.c_int]
            L.ref_mg_level_get.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 5
            L.ref_energy.restype = C.c_float
            L.ref_energy.argtypes = [C.c_void_p, C.c_float]
            L.ref_advect.argtypes = [C.c_void_p, C.c_float]
            L.ref_threads.argtypes = [C.c_void_p]
            L.ref_set_wind.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_float]
            L.ref_set_bytes.argtypes = [C.c_void_p, C.c_char_p, _u8p]
            L.ref_get_bytes.argtypes = [C.c_void_p, C.c_char_p, _u8p]
            L.ref_turbulence.restype = C.c_float
            L.ref_turbulence.argtypes = [C.c_void_p]
            for nm in ("ref_heat_diffusion", "ref_momentum", "ref_inject_burners", "ref_combustion", "ref_leakage"):
                getattr(L, nm).argtypes = [C.c_void_p, C.c_float]
            L.ref_stable_dt.restype = C.c_float
            L.ref_stable_dt.argtypes = [C.c_void_p] + [C.c_float] * 6
            L.ref_add_burner.argtypes = [C.c_void_p] + [C.c_float] * 7
            L.ref_burner_cells.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
            L.ref_step.restype = C.c_float
            L.ref_step.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_int]
            L.ref_scalars.argtypes = [C.c_void_p, _f32p]
            L.ref_zone_pressure.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
            L.ref_set_zone_pressure.argtypes = [C.c_void_p, _f32p]
            L.ref_leak_faces.argtypes = [C.c_void_p]
            L.ref_update_faces.argtypes = [C.c_void_p, C.c_int, _i32p, _i32p, _f32p, _f32p, _u8p]
            L.ref_wind_at.restype = C.c_float
            L.ref_wind_at.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_float]
            cls._libs[path] = L
        return cls._libs[path]

    def __init__(self, grid, threads=1, pin=True):
        """pin=True: the build without FMA contraction (bitwise-comparable with
        the oracle); pin=False: the reference's own flags (CPU baseline)."""
        L = self._L = self._lib(pin)
        self.grid = grid
        self.h = L.ref_create(threads)
        self.nzones = L.ref_setup(self.h, grid.nx, grid.ny, grid.nz, grid.dx, grid.fx, grid.fy, grid.fz,
                                  grid.ax, grid.ay, grid.az, grid.blocked)

    def __del__(self):
        if getattr(self, "h", None) and self._L is not None:
            self._L.ref_destroy(self.h)
            self.h = None

    def zones(self):
        self.nzones = self._L.ref_zone_pressure(self.h, None, None)   # the current zone count
        zone = np.zeros(self.grid.n, np.int32)
        zo = np.zeros(max(self.nzones, 1), np.uint8)
        self._L.ref_get_zones(self.h, zone, zo)
        return zone, zo[:self.nzones]

    def set(self, name, arr):
        n = self._L.ref_set_field(self.h, name.encode(), np.ascontiguousarray(arr, np.float32))
        if n < 0:
            raise KeyError(name)

    def get(self, name, size):
        out = np.zeros(size, np.float32)
        if self._L.ref_get_field(self.h, name.encode(), out) < 0:
            raise KeyError(name)
        return out

    def project(self, dt, tol, max_iters):
        res, vm = C.c_float(0), C.c_float(0)
        it = self._L.ref_project(self.h, dt, tol, max_iters, C.byref(res), C.byref(vm))
        return it, res.value, vm.value

    def precondition(self, r):
        z = np.zeros(self.grid.n, np.float32)
        self._L.ref_precondition(self.h, np.ascontiguousarray(r, np.float32), z)
        return z

    def levels(self):
        L = self._L
        nl = L.ref_mg_nlevels(self.h)
        out = []
        for l in range(nl):
            n = L.ref_mg_level_n(self.h, l)
            nnz = L.ref_mg_level_nnz(self.h, l)
            d = dict(n=n, diag=np.zeros(n, np.float32), agg=np.zeros(n, np.int32))
            if l > 0:
                d.update(rowStart=np.zeros(n + 1, np.int32), col=np.zeros(nnz, np.int32),
                         val=np.zeros(nnz, np.float32))
            L.ref_mg_level_get(self.h, l, _ptr(d.get("rowStart")), _ptr(d.get("col")), _ptr(d.get("val")),
                               _ptr(d["diag"]), _ptr(d["agg"]))
            out.append(d)
        return out

    def set_bytes(self, name, arr):
        if self._L.ref_set_bytes(self.h, name.encode(), np.ascontiguousarray(arr, np.uint8)) < 0:
            raise KeyError(name)

    def get_bytes(self, name, size):
        out = np.zeros(size, np.uint8)
        if self._L.ref_get_bytes(self.h, name.encode(), out) < 0:
            raise KeyError(name)
        return out

    def turbulence(self):
        return self._L.ref_turbulence(self.h)

    def heat_diffusion(self, dt):
        self._L.ref_heat_diffusion(self.h, dt)

    def momentum(self, dt):
        self._L.ref_momentum(self.h, dt)

    def inject_burners(self, dt):
        self._L.ref_inject_burners(self.h, dt)

    def combustion(self, dt):
        self._L.ref_combustion(self.h, dt)

    def leakage(self, dt):
        self._L.ref_leakage(self.h, dt)

    def stable_dt(self, dt_max, max_vel, max_nu, max_div, cfl=0.8, profile_dt_max=0.05):
        return self._L.ref_stable_dt(self.h, dt_max, max_vel, max_nu, max_div, cfl, profile_dt_max)

    def add_burner(self, x0, y0, x1, y1, z, peak_w, ramp_s=0.0):
        b = self._L.ref_add_burner(self.h, x0, y0, x1, y1, z, peak_w, ramp_s)
        n = self._L.ref_burner_cells(self.h, b, None)
        cells = np.zeros(n, np.int32)
        self._L.ref_burner_cells(self.h, b, cells.ctypes.data)
        return b, cells

    def step(self, dt_max, tol, max_iters):
        return self._L.ref_step(self.h, dt_max, tol, max_iters)

    def scalars(self):
        out = np.zeros(6, np.float32)
        it = self._L.ref_scalars(self.h, out)
        return dict(time=out[0], max_vel=out[1], max_nu=out[2], max_div=out[3], residual=out[4], hrr=out[5],
                    iters=it)

    def zone_pressure(self):
        n = self._L.ref_zone_pressure(self.h, None, None)
        zp, zdp = np.zeros(max(n, 1), np.float32), np.zeros(max(n, 1), np.float32)
        self._L.ref_zone_pressure(self.h, zp.ctypes.data, zdp.ctypes.data)
        return zp[:n], zdp[:n]

    def set_zone_pressure(self, zoneP):
        self._L.ref_set_zone_pressure(self.h, np.ascontiguousarray(zoneP, np.float32))

    def update_faces(self, axis, face, f, a, crack):
        self._L.ref_update_faces(self.h, len(axis), np.ascontiguousarray(axis, np.int32),
                                 np.ascontiguousarray(face, np.int32), np.ascontiguousarray(f, np.float32),
                                 np.ascontiguousarray(a, np.float32), np.ascontiguousarray(crack, np.uint8))

    def leak_faces(self):
        return self._L.ref_leak_faces(self.h)

    def set_wind(self, speed, from_deg, exponent=0.25):
        self._L.ref_set_wind(self.h, speed, from_deg, exponent)

    def wind_at(self, axis, side, z):
        return self._L.ref_wind_at(self.h, axis, side, z)

    def energy(self, dt):
        return self._L.ref_energy(self.h, dt)

    def advect(self, dt):
        self._L.ref_advect(self.h, dt)

    @property
    def threads(self):
        return self._L.ref_threads(self.h)
