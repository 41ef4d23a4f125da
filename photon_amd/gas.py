"""Python host side of the MI355X gas-phase path, over the C ABI of
libphoton_gas.so (include/photon_gas.h).

It mirrors the pieces of the reference's photon::Solver that the path
replaces (src/sim/solver.h:156-347): a GasSolver owns one grid's geometry and
its device-resident fields and exposes project(dt), advect_scalars(dt) and
energy_and_expansion(dt) with the reference's names and meaning, plus the
batched solve used for many scenarios at once.  No CPU fallback exists: if
the HIP library is missing or no GPU is present the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libphoton_gas.so")

PREC_FP32, PREC_MIXED_F16, PREC_MIXED_BF16 = 0, 1, 2
STOP_REL_L2, STOP_MAXNORM = 0, 1
CONVERGED, MAXITER, BREAKDOWN, BREAKDOWN_RZ, NONFINITE = 1, 2, 3, 4, 5
COARSE_EXACT, COARSE_GS20 = 0, 1
# coarse-grid correction weight of the over-corrected V-cycle (solve_params)
COARSE_WEIGHT = 1.8

EXPORTS = (
    "phg_create", "phg_destroy", "phg_last_error", "phg_synchronize", "phg_stream", "phg_malloc", "phg_free",
    "phg_memcpy_h2d", "phg_memcpy_d2h", "phg_memset", "phg_grid_create", "phg_grid_destroy", "phg_grid_info",
    "phg_grid_zones", "phg_apply", "phg_precondition", "phg_batch_create", "phg_batch_destroy",
    "phg_batch_solve", "phg_solve", "phg_project", "phg_rhs", "phg_zone_mean", "phg_correct",
    "phg_energy_expansion", "phg_advect_scalars", "phg_profile_enable", "phg_profile_select", "phg_profile_get",
    "phg_profile_reset",
    "phg_hierarchy_shape", "phg_batch_create_ex", "phg_copy", "phg_leakage",
    "phg_grid_plan_build", "phg_grid_plan_free", "phg_grid_create_planned", "phg_zone_mean_noalloc",
    "phg_grid_set_coarse", "phg_grid_get_coarse",
    "phg_stable_dt", "phg_turbulence", "phg_heat_diffusion", "phg_momentum", "phg_step_clear", "phg_add_gas",
    "phg_combustion", "phg_zone_integrals", "phg_zone_values", "phg_grid_stats", "phg_batch_graph_check",
    "phg_grid_set_cracks", "phg_zone_leakage", "phg_grid_update_faces", "phg_zone_pressures_set",
    "phg_zone_pressures_get", "phg_zone_leakage_resident", "phg_zone_pressure_resident",
)


class PhgError(RuntimeError):
    pass


class SolveParams(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iters", C.c_int), ("stop_rule", C.c_int), ("inv", C.c_float),
                ("demean", C.c_int), ("check_every", C.c_int), ("use_graph", C.c_int), ("zero_x", C.c_int),
                ("refine", C.c_int), ("coarse_weight", C.c_float)]


class BatchOpts(C.Structure):
    _fields_ = [("kchunk", C.c_int), ("level0_blocks", C.c_int), ("unfused", C.c_int), ("sweeps0", C.c_int),
                ("sys_kchunk", C.c_void_p), ("sys_sweeps0", C.c_void_p), ("fused_threads", C.c_int),
                ("coarse_lds", C.c_int)]


class Wind(C.Structure):
    """Solver::wind (src/sim/solver.h:106-111) and the grid's floor height."""
    _fields_ = [("speed", C.c_float), ("from_deg", C.c_float), ("exponent", C.c_float), ("oz", C.c_float)]


GAS_FIELDS = ("T", "Yo", "Yf", "Ys", "Ya", "Yp", "Yc", "Yw", "Yco", "Yai", "Pb", "massSrc", "Q", "Qr", "hrr",
              "pilot", "lit")


class GasPtrs(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in GAS_FIELDS]


class SolveStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("status", C.c_int), ("relres", C.c_double), ("maxres", C.c_double)]


_lib = None


def load_library():
    """Load libphoton_gas.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PhgError(f"{LIB_PATH} is missing: build it with `python -m photon_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, i32, f32, f64, sz = C.c_void_p, C.c_int, C.c_float, C.c_double, C.c_size_t
    for name in EXPORTS:
        getattr(L, name).restype = C.c_int
    L.phg_create.argtypes = [C.POINTER(vp), i32]
    L.phg_destroy.argtypes = [vp]
    L.phg_destroy.restype = None
    L.phg_last_error.argtypes = [vp]
    L.phg_last_error.restype = C.c_char_p
    L.phg_synchronize.argtypes = [vp]
    L.phg_stream.argtypes = [vp]
    L.phg_stream.restype = vp
    L.phg_malloc.argtypes = [vp, C.POINTER(vp), sz]
    L.phg_free.argtypes = [vp, vp]
    L.phg_memcpy_h2d.argtypes = [vp, vp, vp, sz]
    L.phg_memcpy_d2h.argtypes = [vp, vp, vp, sz]
    L.phg_memset.argtypes = [vp, vp, i32, sz]
    L.phg_grid_create.argtypes = [vp, i32, i32, i32, f32] + [vp] * 8 + [i32, C.POINTER(i32)]
    L.phg_grid_destroy.argtypes = [vp, i32]
    L.phg_grid_info.argtypes = [vp, i32, vp, vp, vp, vp]
    L.phg_grid_zones.argtypes = [vp, i32, vp, vp]
    L.phg_apply.argtypes = [vp, i32, vp, vp, C.POINTER(f64)]
    L.phg_precondition.argtypes = [vp, i32, vp, vp]
    L.phg_batch_create.argtypes = [vp, i32, vp, C.POINTER(i32)]
    L.phg_copy.argtypes = [vp, vp, vp, sz]
    L.phg_batch_create_ex.argtypes = [vp, i32, vp, C.POINTER(BatchOpts), C.POINTER(i32)]
    L.phg_batch_destroy.argtypes = [vp, i32]
    L.phg_batch_solve.argtypes = [vp, i32, vp, vp, C.POINTER(SolveParams), vp]
    L.phg_solve.argtypes = [vp, i32, vp, vp, C.POINTER(SolveParams), vp]
    L.phg_project.argtypes = [vp, i32, f32, vp, vp, vp, vp, vp, C.POINTER(SolveParams), vp, C.POINTER(f32)]
    L.phg_rhs.argtypes = [vp, i32, f32, vp, vp, vp, vp, vp, vp, vp]
    L.phg_zone_mean.argtypes = [vp, i32, vp]
    L.phg_zone_mean_noalloc.argtypes = [vp, i32, vp]
    L.phg_grid_set_coarse.argtypes = [vp, i32, i32]
    L.phg_grid_get_coarse.argtypes = [vp, i32, C.POINTER(i32)]
    L.phg_stable_dt.argtypes = [f32] * 7 + [C.POINTER(f32)]
    L.phg_turbulence.argtypes = [vp, i32, vp, vp, vp, vp, vp, C.POINTER(f32)]
    L.phg_heat_diffusion.argtypes = [vp, i32, f32, vp, vp, vp, vp]
    L.phg_momentum.argtypes = [vp, i32, f32, C.POINTER(Wind)] + [vp] * 10
    L.phg_step_clear.argtypes = [vp, i32, C.POINTER(GasPtrs)]
    L.phg_add_gas.argtypes = [vp, i32, i32, vp, f32, f32, f32, f32, i32, C.POINTER(GasPtrs)]
    L.phg_combustion.argtypes = [vp, i32, f32, i32, C.POINTER(GasPtrs), C.POINTER(f64)]
    L.phg_zone_integrals.argtypes = [vp, i32, vp, vp]
    L.phg_zone_values.argtypes = [vp, i32, vp, vp, i32]
    L.phg_grid_stats.argtypes = [vp, i32, C.POINTER(f32), C.POINTER(f32), C.POINTER(f32)]
    L.phg_batch_graph_check.argtypes = [vp, i32, i32, i32, C.POINTER(i32), C.c_char_p, sz]
    L.phg_leakage.argtypes = [vp, i32, vp, vp]
    L.phg_grid_set_cracks.argtypes = [vp, i32, vp, vp, vp]
    L.phg_zone_leakage.argtypes = [vp, i32, f32] + [vp] * 9 + [vp]
    L.phg_zone_pressures_set.argtypes = [vp, i32, vp]
    L.phg_zone_pressures_get.argtypes = [vp, i32, vp, vp]
    L.phg_zone_leakage_resident.argtypes = [vp, i32, f32] + [vp] * 7
    L.phg_zone_pressure_resident.argtypes = [vp, i32, f32, vp]
    L.phg_grid_update_faces.argtypes = [vp, i32, i32] + [vp] * 9
    L.phg_correct.argtypes = [vp, i32, f32, vp, vp, vp, vp, C.POINTER(f32)]
    L.phg_energy_expansion.argtypes = [vp, i32, f32, vp, vp, vp, vp, vp, C.POINTER(f32)]
    L.phg_advect_scalars.argtypes = [vp, i32, f32, vp, vp, vp, vp, vp, vp, vp]
    L.phg_profile_enable.argtypes = [vp, i32]
    L.phg_profile_select.argtypes = [vp, C.c_ulonglong]
    L.phg_profile_get.argtypes = [vp, i32, C.POINTER(C.c_char_p), C.POINTER(f64), C.POINTER(C.c_longlong), C.POINTER(f64)]
    L.phg_profile_reset.argtypes = [vp]
    L.phg_hierarchy_shape.argtypes = [i32, i32, i32] + [vp] * 5 + [C.POINTER(i32), vp, vp]
    L.phg_grid_plan_build.argtypes = [i32, i32, i32] + [vp] * 5 + [C.POINTER(vp)]
    L.phg_grid_plan_free.argtypes = [vp]
    L.phg_grid_plan_free.restype = None
    L.phg_grid_create_planned.argtypes = [vp, i32, i32, i32, f32] + [vp] * 8 + [i32, vp, C.POINTER(i32)]
    _lib = L
    return L


class DeviceArray:
    """A device allocation of the context (float32 / int32 / uint8 ...)."""

    def __init__(self, ctx: "Context", shape, dtype=np.float32):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        self.shape = (shape,) if isinstance(shape, (int, np.integer)) else tuple(shape)
        self.size = int(np.prod(self.shape))
        self.nbytes = self.size * self.dtype.itemsize
        p = C.c_void_p()
        ctx._check(ctx.L.phg_malloc(ctx.h, C.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, ctx, arr, dtype=None):
        a = np.ascontiguousarray(arr, dtype=dtype or arr.dtype)
        d = cls(ctx, a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, arr):
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        assert a.size == self.size, (a.size, self.size)
        self.ctx._check(self.ctx.L.phg_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data, self.nbytes))

    def download(self):
        out = np.empty(self.shape, self.dtype)
        self.ctx._check(self.ctx.L.phg_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def zero(self):
        self.ctx._check(self.ctx.L.phg_memset(self.ctx.h, self.ptr, 0, self.nbytes))

    def free(self):
        if self.ptr is not None and self.ctx.h:
            self.ctx.L.phg_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One device and one HIP stream (phg_create)."""

    def __init__(self, device: int = 0):
        self.L = load_library()
        h = C.c_void_p()
        rc = self.L.phg_create(C.byref(h), device)
        if rc != 0:
            raise PhgError(f"phg_create(device={device}) failed with {rc}: no usable GPU")
        self.h = h.value

    def _check(self, rc):
        if rc != 0:
            raise PhgError(f"photon_gas error {rc}: {self.L.phg_last_error(self.h).decode()}")
        return rc

    def close(self):
        if self.h:
            self.L.phg_destroy(self.h)
            self.h = None

    def synchronize(self):
        self._check(self.L.phg_synchronize(self.h))

    def array(self, shape, dtype=np.float32):
        return DeviceArray(self, shape, dtype)

    def to_device(self, arr, dtype=None):
        return DeviceArray.from_host(self, arr, dtype)

    def copy(self, dst, src, nbytes):
        """Device-to-device copy by the library's streaming copy kernel (16 B
        per lane); the reference point for bandwidth and counter calibration."""
        self._check(self.L.phg_copy(self.h, dst.ptr, src.ptr, nbytes))

    # profiling
    def profile(self, on=True, classes=None):
        """Time kernel launches by class with HIP events; `classes` (names)
        limits the timing to those classes (None: every class)."""
        mask = (1 << 64) - 1
        if classes is not None:
            names = self.profile_names()
            mask = 0
            for c in classes:
                if c in names:
                    mask |= 1 << names.index(c)
        self._check(self.L.phg_profile_select(self.h, mask))
        self._check(self.L.phg_profile_enable(self.h, 1 if on else 0))

    def profile_names(self):
        out = []
        while True:
            name = C.c_char_p()
            if self.L.phg_profile_get(self.h, len(out), C.byref(name), None, None, None) != 0:
                return out
            out.append(name.value.decode())

    def profile_reset(self):
        self._check(self.L.phg_profile_reset(self.h))

    def profile_read(self):
        out = {}
        i = 0
        while True:
            name, ms, n, by = C.c_char_p(), C.c_double(), C.c_longlong(), C.c_double()
            if self.L.phg_profile_get(self.h, i, C.byref(name), C.byref(ms), C.byref(n), C.byref(by)) != 0:
                break
            out[name.value.decode()] = (ms.value, n.value)
            i += 1
        return out


def solve_params(tol=1e-6, max_iters=200, stop_rule=STOP_REL_L2, inv=1.0, demean=True, check_every=4,
                 use_graph=False, zero_x=False, refine=0, coarse_weight=0.0):
    """use_graph: False/0 direct launches, True/1 one captured iteration
    replayed, 2 the whole solve (set-up + check_every iterations) as one graph.
    coarse_weight: weight of the V-cycle's coarse-grid corrections (0 or 1:
    the reference's cycle; COARSE_WEIGHT over-corrects, taking about a third
    fewer iterations)."""
    g = int(use_graph) if not isinstance(use_graph, bool) else (1 if use_graph else 0)
    return SolveParams(tol, max_iters, stop_rule, inv, 1 if demean else 0, check_every, g, 1 if zero_x else 0,
                       refine, coarse_weight)


class GridPlan:
    """The host half of a grid's set-up (phg_grid_plan_build: the hierarchy,
    its dense coarsest inverse and the zones), built without a device; safe
    to build on a worker thread (ctypes releases the GIL for the call)."""

    def __init__(self, scen, wind_closed=None):
        L = load_library()
        self.scen = scen
        self.wind_closed = wind_closed
        arrs = [np.ascontiguousarray(a, np.float32) for a in (scen.ax, scen.ay, scen.az)]
        blk = np.ascontiguousarray(scen.blocked, np.uint8)
        wc = None if wind_closed is None else np.ascontiguousarray(wind_closed, np.uint8)
        h = C.c_void_p()
        rc = L.phg_grid_plan_build(scen.nx, scen.ny, scen.nz, *[a.ctypes.data for a in arrs], blk.ctypes.data,
                                   None if wc is None else wc.ctypes.data, C.byref(h))
        if rc != 0:
            raise PhgError(f"phg_grid_plan_build failed with {rc}")
        self.h = h

    def free(self):
        if self.h:
            load_library().phg_grid_plan_free(self.h)
            self.h = None

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.phg_grid_plan_free(self.h)
            self.h = None


class Grid:
    """Geometry of one scenario on the device with its multigrid hierarchy
    (phg_grid_create: buildPressureLevels + computeZones).  `plan`: a GridPlan
    of the same geometry whose host work is taken instead of redone."""

    def __init__(self, ctx: Context, scen, precision=PREC_FP32, wind_closed=None, plan: GridPlan | None = None):
        self.ctx = ctx
        self.scen = scen
        self.precision = precision
        self.nx, self.ny, self.nz, self.dx = scen.nx, scen.ny, scen.nz, scen.dx
        self.n = scen.nx * scen.ny * scen.nz
        arrs = [np.ascontiguousarray(a, np.float32) for a in (scen.fx, scen.fy, scen.fz, scen.ax, scen.ay, scen.az)]
        blk = np.ascontiguousarray(scen.blocked, np.uint8)
        wc = None if wind_closed is None else np.ascontiguousarray(wind_closed, np.uint8)
        gid = C.c_int()
        if plan is None:
            ctx._check(ctx.L.phg_grid_create(ctx.h, scen.nx, scen.ny, scen.nz, scen.dx,
                                             *[a.ctypes.data for a in arrs], blk.ctypes.data,
                                             None if wc is None else wc.ctypes.data, precision, C.byref(gid)))
        else:
            same_wc = (plan.wind_closed is None and wind_closed is None) or (
                plan.wind_closed is not None and wind_closed is not None
                and np.array_equal(np.asarray(plan.wind_closed, np.uint8), np.asarray(wind_closed, np.uint8)))
            if plan.scen is not scen or not same_wc or not plan.h:
                raise PhgError("the plan was built for another geometry or is used")
            ctx._check(ctx.L.phg_grid_create_planned(ctx.h, scen.nx, scen.ny, scen.nz, scen.dx,
                                                     *[a.ctypes.data for a in arrs], blk.ctypes.data,
                                                     None if wc is None else wc.ctypes.data, precision, plan.h,
                                                     C.byref(gid)))
            plan.free()
        self.id = gid.value
        self.leak_faces = 0
        if getattr(scen, "leaks", None) is not None:
            self.set_cracks(scen.leaks)

    def update_faces(self, upd, u=None, v=None, w=None):
        """Face changes (axis, face, open fraction, coefficient, crack flag
        arrays; phg_grid_update_faces); the velocities of closing faces are
        zeroed in u, v, w (device arrays) when given.  Returns 0 when the set-up
        stood, 2 when the zones and hierarchy were set up again with levels 0
        and 1 formed on the device, 1 when wholly on the host."""
        axis, face, f, a, c = upd
        arrs = [np.ascontiguousarray(axis, np.int32), np.ascontiguousarray(face, np.int32),
                np.ascontiguousarray(f, np.float32), np.ascontiguousarray(a, np.float32),
                np.ascontiguousarray(c, np.uint8)]
        reb = C.c_int()
        self.ctx._check(self.ctx.L.phg_grid_update_faces(
            self.ctx.h, self.id, len(arrs[0]), *[x.ctypes.data for x in arrs],
            *[None if x is None else x.ptr for x in (u, v, w)], C.byref(reb)))
        return reb.value

    def set_cracks(self, leaks):
        """Crack faces (lx, ly, lz: one byte per face, None for none) of the
        grid (phg_grid_set_cracks); None clears them."""
        lk = [None, None, None] if leaks is None else [None if a is None else np.ascontiguousarray(a, np.uint8)
                                                        for a in leaks]
        self._leaks = lk
        self.ctx._check(self.ctx.L.phg_grid_set_cracks(self.ctx.h, self.id,
                                                       *[None if a is None else a.ctypes.data for a in lk]))

    def info(self):
        nl, nz, rl = C.c_int(), C.c_int(), C.c_int()
        sizes = np.zeros(16, np.int32)
        self.ctx._check(self.ctx.L.phg_grid_info(self.ctx.h, self.id, C.addressof(nl), sizes.ctypes.data,
                                                 C.addressof(nz), C.addressof(rl)))
        return dict(levels=sizes[:nl.value].tolist(), nzones=nz.value, ref_levels=rl.value)

    @property
    def coarse(self):
        """How the V-cycle solves the coarsest level: COARSE_EXACT (dense
        inverse) or COARSE_GS20 (the reference's 20 Gauss-Seidel sweeps)."""
        m = C.c_int()
        self.ctx._check(self.ctx.L.phg_grid_get_coarse(self.ctx.h, self.id, C.byref(m)))
        return m.value

    @coarse.setter
    def coarse(self, mode):
        self.ctx._check(self.ctx.L.phg_grid_set_coarse(self.ctx.h, self.id, int(mode)))

    def zones(self):
        nz = self.info()["nzones"]
        zone = np.zeros(self.n, np.int32)
        zo = np.zeros(max(nz, 1), np.uint8)
        self.ctx._check(self.ctx.L.phg_grid_zones(self.ctx.h, self.id, zone.ctypes.data, zo.ctypes.data))
        return zone, zo[:nz]

    def apply(self, x: DeviceArray, y: DeviceArray):
        dot = C.c_double()
        self.ctx._check(self.ctx.L.phg_apply(self.ctx.h, self.id, x.ptr, y.ptr, C.byref(dot)))
        return dot.value

    def precondition(self, r: DeviceArray, z: DeviceArray):
        self.ctx._check(self.ctx.L.phg_precondition(self.ctx.h, self.id, r.ptr, z.ptr))

    def solve(self, b: DeviceArray, x: DeviceArray, params: SolveParams):
        st = SolveStats()
        self.ctx._check(self.ctx.L.phg_solve(self.ctx.h, self.id, b.ptr, x.ptr, C.byref(params), C.addressof(st)))
        return dict(iters=st.iters, status=st.status, relres=st.relres, maxres=st.maxres)

    def destroy(self):
        if self.id is not None and self.ctx.h:
            self.ctx.L.phg_grid_destroy(self.ctx.h, self.id)
            self.id = None


class Batch:
    """A fixed set of systems solved together (phg_batch_create / _solve)."""

    def __init__(self, ctx: Context, grids, kchunk=0, level0_blocks=0, unfused=False, sweeps0=0, sys_kchunk=None,
                 sys_sweeps0=None, fused_threads=0, coarse_lds=0):
        self.ctx = ctx
        self.grids = list(grids)
        ids = np.array([g.id for g in self.grids], np.int32)
        bid = C.c_int()
        kc = None if sys_kchunk is None else np.ascontiguousarray(sys_kchunk, np.int32)
        sw = None if sys_sweeps0 is None else np.ascontiguousarray(sys_sweeps0, np.int32)
        for a in (kc, sw):
            assert a is None or len(a) == len(ids)
        opts = BatchOpts(kchunk, level0_blocks, 1 if unfused else 0, sweeps0,
                         None if kc is None else kc.ctypes.data, None if sw is None else sw.ctypes.data,
                         fused_threads, coarse_lds)
        ctx._check(ctx.L.phg_batch_create_ex(ctx.h, len(ids), ids.ctypes.data, C.byref(opts), C.byref(bid)))
        self.id = bid.value

    def solve(self, bs, xs, params: SolveParams):
        n = len(self.grids)
        assert len(bs) == n and len(xs) == n
        bp = (C.c_void_p * n)(*[b.ptr if isinstance(b, DeviceArray) else b for b in bs])
        xp = (C.c_void_p * n)(*[x.ptr if isinstance(x, DeviceArray) else x for x in xs])
        st = (SolveStats * n)()
        self.ctx._check(self.ctx.L.phg_batch_solve(self.ctx.h, self.id, bp, xp, C.byref(params), st))
        return [dict(iters=s.iters, status=s.status, relres=s.relres, maxres=s.maxres) for s in st]

    def graph_check(self, iterations, refine=0):
        """Capture the whole-solve graph and check its nodes against the
        device's launch limits (phg_batch_graph_check): (nodes, summary)."""
        n = C.c_int()
        buf = C.create_string_buffer(512)
        self.ctx._check(self.ctx.L.phg_batch_graph_check(self.ctx.h, self.id, iterations, refine, C.byref(n), buf,
                                                         len(buf)))
        return n.value, buf.value.decode()

    def destroy(self):
        if self.id is not None and self.ctx.h:
            self.ctx.L.phg_batch_destroy(self.ctx.h, self.id)
            self.id = None


FIELD_NAMES = ("T", "Yo", "Yf", "Ys", "Ya", "Yp", "Yc", "Yw", "Yco", "Yai")


def burner_cells(g, x0, y0, x1, y1, z, ox=0.0, oy=0.0, oz=0.0):
    """The gas cells a burner over the floor rectangle (x0, y0)-(x1, y1) at
    height z releases into, by Solver::addBurner's rule (gasphase.cpp:72-90:
    cell centres inside the rectangle, else the cell under its centre); None
    when it covers no gas cell."""
    f32 = np.float32
    dx = f32(g.dx)

    def cell_at(x, y, zz):
        i, j, k = (int(np.floor((f32(a) - f32(o)) / dx)) for a, o in ((x, ox), (y, oy), (zz, oz)))
        if i < 0 or j < 0 or k < 0 or i >= g.nx or j >= g.ny or k >= g.nz:
            return -1
        return i + g.nx * (j + g.ny * k)

    cells = []
    zc = f32(z) + f32(0.5) * dx
    for j in range(g.ny):
        for i in range(g.nx):
            x, y = f32(ox) + (f32(i) + f32(0.5)) * dx, f32(oy) + (f32(j) + f32(0.5)) * dx
            if x < f32(x0) or x > f32(x1) or y < f32(y0) or y > f32(y1):
                continue
            c = cell_at(x, y, zc)
            if c >= 0 and not g.blocked[c]:
                cells.append(c)
    if not cells:
        c = cell_at(f32(0.5) * (f32(x0) + f32(x1)), f32(0.5) * (f32(y0) + f32(y1)), zc)
        if c < 0 or g.blocked[c]:
            return None
        cells.append(c)
    return np.array(cells, np.int32)


def hierarchy_shape(scen, wind_closed=None):
    """Level sizes and level-0 aggregates the library builds for a geometry
    (host only, no device needed)."""
    L = load_library()
    arrs = [np.ascontiguousarray(a, np.float32) for a in (scen.ax, scen.ay, scen.az)]
    blk = np.ascontiguousarray(scen.blocked, np.uint8)
    wc = None if wind_closed is None else np.ascontiguousarray(wind_closed, np.uint8)
    nl = C.c_int()
    sizes = np.zeros(16, np.int32)
    agg = np.zeros(scen.nx * scen.ny * scen.nz, np.int32)
    rc = L.phg_hierarchy_shape(scen.nx, scen.ny, scen.nz, *[a.ctypes.data for a in arrs], blk.ctypes.data,
                               None if wc is None else wc.ctypes.data, C.byref(nl), sizes.ctypes.data,
                               agg.ctypes.data)
    if rc != 0:
        raise PhgError(f"phg_hierarchy_shape failed with {rc}")
    return sizes[:nl.value].tolist(), agg


class GasSolver:
    """Device-resident gas-phase fields of one grid with the reference's
    phase names: project(dt), advect_scalars(dt), energy_and_expansion(dt)
    (src/sim/gasphase.cpp:251-271, :505-677, :685-784).  Host numpy copies
    are read with get(name) and written with set(name, array).

    With `graph_steps` (the default), project() runs each solve's set-up and
    iterations as one captured graph covering the previous step's iteration
    count rounded up to a multiple of 4: consecutive steps' solves need
    similar counts, so the host waits about once per solve.  Replays give
    the same results bit for bit as direct launches.

    With `resident_zones`, the zones' background pressures and rise rates stay
    on the device (phg_zone_*_resident), so a step queues its work without
    waiting for the device; zoneP and zoneDP on the host are refreshed by
    zones_from_device() (snapshots do so themselves).

    `coarse_weight` (0: the reference's V-cycle) weights the V-cycle's
    coarse-grid corrections; COARSE_WEIGHT takes about a third fewer
    iterations to the same stop rule, with iteration counts that then differ
    from the reference's."""

    CELL_FIELDS = ("p", "S", "Pb", "Q", "Qr", "hrr", "massSrc", "nu") + FIELD_NAMES
    BYTE_FIELDS = ("pilot", "lit")

    def __init__(self, ctx: Context, scen, precision=PREC_FP32, wind_closed=None, tol=1e-6, max_iters=400,
                 stop_rule=STOP_REL_L2, plan: GridPlan | None = None, graph_steps=True, coarse=None, wind=None,
                 radiation_on=False, dt_max_profile=0.05, cfl=0.8, resident_zones=False, coarse_weight=0.0):
        self.ctx = ctx
        self.scen = scen
        self.graph_steps = graph_steps
        self.grid = Grid(ctx, scen, precision, wind_closed, plan=plan)
        if coarse is not None:
            self.grid.coarse = coarse
        n = self.grid.n
        self.d = {}
        for name in ("u", "v", "w"):
            size = {"u": len(scen.fx), "v": len(scen.fy), "w": len(scen.fz)}[name]
            self.d[name] = ctx.array(size)
            self.d[name].zero()
            self.d[name + "2"] = ctx.array(size)      # momentum writes here, then the two swap
            self.d[name + "2"].zero()
        init = dict(T=293.15, Yo=0.232, Yc=0.0006, Yw=0.0073, Pb=101325.0)
        for name in self.CELL_FIELDS:
            self.d[name] = ctx.to_device(np.full(n, init.get(name, 0.0), np.float32))
        for name in self.BYTE_FIELDS:
            self.d[name] = ctx.to_device(np.zeros(n, np.uint8))
        # Solver::wind, radiationOn and the profile's dtMax and cfl (solver.h:82-111,184-187)
        self.wind = Wind(*(wind if wind is not None else (0.0, 0.0, 0.25, 0.0)))
        self.radiation_on = radiation_on
        self.dt_max_profile, self.cfl = dt_max_profile, cfl
        self.time = 0.0
        self.burners = []
        self.maxNu = 0.0
        self.hrr_total = 0.0
        # zones: background pressure per zone and volumes (computeZones, gasphase.cpp:111-158)
        zone, zopen = self.grid.zones()
        self.zone_open = zopen
        self.zoneP = np.full(len(zopen), 101325.0, np.float32)
        self.zoneDP = np.zeros(len(zopen), np.float32)
        cnt = np.bincount(zone[zone >= 0], minlength=len(zopen)).astype(np.float64)
        dxd = np.float64(np.float32(scen.dx))
        self.zone_vol = (cnt * dxd * dxd * dxd).astype(np.float32)
        self.resident_zones = resident_zones
        self._zones_to_device()
        self._tmp = [ctx.array(n) for _ in range(10)]
        self.d["b"] = ctx.array(n)       # right-hand side of the split projection (rhs / correct)
        self.d["diag"] = ctx.array(n)
        self.params = solve_params(tol=tol, max_iters=max_iters, stop_rule=stop_rule, coarse_weight=coarse_weight)
        self.pressureIters = 0
        self.pressureResidual = 0.0
        self.pressureRelRes = 0.0
        self.maxVel = 0.0
        self.maxDiv = 0.0

    def set(self, name, arr):
        self.d[name].upload(np.asarray(arr, self.d[name].dtype))

    def get(self, name):
        return self.d[name].download()

    def project(self, dt):
        if self.graph_steps:
            self.params.use_graph = 2
            self.params.check_every = max(4, -(-(self.pressureIters + 1) // 4) * 4) if self.pressureIters > 0 else 8
        st = SolveStats()
        vm = C.c_float()
        d = self.d
        self.ctx._check(self.ctx.L.phg_project(self.ctx.h, self.grid.id, dt, d["u"].ptr, d["v"].ptr, d["w"].ptr,
                                               d["p"].ptr, d["S"].ptr, C.byref(self.params), C.addressof(st),
                                               C.byref(vm)))
        self.pressureIters = st.iters
        self.pressureResidual = st.maxres
        self.pressureRelRes = st.relres
        self.maxVel = vm.value
        return st.iters

    def leakage(self, report=False, field="S"):
        """The compatible source of Solver::leakage for grids without cracks
        (src/sim/gasphase.cpp:487-498): S -= its mean on every closed zone.
        With `report`, returns the per-zone shift (1/s; 0 on open zones).
        `field` = "b" applies the same projection to the right-hand side."""
        nz = len(self.scen.zone_open) if self.scen.zone_open is not None else 0
        out = np.zeros(max(nz, 1), np.float32)
        self.ctx._check(self.ctx.L.phg_leakage(self.ctx.h, self.grid.id, self.d[field].ptr,
                                               out.ctypes.data if report else None))
        return out[:nz] if report else None

    def rhs(self, dt):
        """The right-hand side and diagonal of project() into d["b"], d["diag"]
        (src/sim/gasphase.cpp:528-549); p is zeroed in solids."""
        d = self.d
        self.ctx._check(self.ctx.L.phg_rhs(self.ctx.h, self.grid.id, dt, d["u"].ptr, d["v"].ptr, d["w"].ptr,
                                           d["S"].ptr, d["b"].ptr, d["diag"].ptr, d["p"].ptr))

    def correct(self, dt, report=True):
        """Face-velocity correction from p (src/sim/gasphase.cpp:640-676); sets
        maxVel when `report` (which waits for the device)."""
        vm = C.c_float()
        d = self.d
        self.ctx._check(self.ctx.L.phg_correct(self.ctx.h, self.grid.id, dt, d["p"].ptr, d["u"].ptr, d["v"].ptr,
                                               d["w"].ptr, C.byref(vm) if report else None))
        if report:
            self.maxVel = vm.value
        return self.maxVel

    def energy_and_expansion(self, dt, report=True):
        """Sets maxDiv when `report` (which waits for the device)."""
        md = C.c_float()
        d = self.d
        self.ctx._check(self.ctx.L.phg_energy_expansion(self.ctx.h, self.grid.id, dt, d["T"].ptr, d["S"].ptr,
                                                        d["Pb"].ptr, d["Q"].ptr, d["massSrc"].ptr,
                                                        C.byref(md) if report else None))
        if report:
            self.maxDiv = md.value
        return self.maxDiv

    # ---- the rest of the gas step (Solver::step, src/sim/solver.cpp:226-269)
    def _gas(self):
        return GasPtrs(*[self.d[n].ptr for n in GAS_FIELDS])

    def turbulence(self, report=True):
        """Solver::turbulence: nu and maxNu (gasphase.cpp:163-218); maxNu is
        read back when `report` (which waits for the device)."""
        m = C.c_float()
        d = self.d
        self.ctx._check(self.ctx.L.phg_turbulence(self.ctx.h, self.grid.id, d["u"].ptr, d["v"].ptr, d["w"].ptr,
                                                  d["T"].ptr, d["nu"].ptr, C.byref(m) if report else None))
        if report:
            self.maxNu = m.value
        return self.maxNu

    def read_stats(self):
        """maxVel, maxNu and maxDiv of the last queued correction, turbulence
        and energy calls, in one wait (phg_grid_stats)."""
        a, b, c = C.c_float(), C.c_float(), C.c_float()
        self.ctx._check(self.ctx.L.phg_grid_stats(self.ctx.h, self.grid.id, C.byref(a), C.byref(b), C.byref(c)))
        self.maxVel, self.maxNu, self.maxDiv = a.value, b.value, c.value
        return self.maxVel, self.maxNu, self.maxDiv

    def heat_diffusion(self, dt):
        d = self.d
        self.ctx._check(self.ctx.L.phg_heat_diffusion(self.ctx.h, self.grid.id, dt, d["T"].ptr, d["nu"].ptr,
                                                      d["Pb"].ptr, d["Q"].ptr))

    def momentum(self, dt):
        """Solver::momentum (gasphase.cpp:275-399): new face velocities, swapped in."""
        d = self.d
        self.ctx._check(self.ctx.L.phg_momentum(self.ctx.h, self.grid.id, dt, C.byref(self.wind), d["u"].ptr,
                                                d["v"].ptr, d["w"].ptr, d["T"].ptr, d["Pb"].ptr, d["nu"].ptr,
                                                d["p"].ptr, d["u2"].ptr, d["v2"].ptr, d["w2"].ptr))
        for f in ("u", "v", "w"):
            d[f], d[f + "2"] = d[f + "2"], d[f]

    def add_burner(self, x0, y0, x1, y1, z, peak_w, ramp_s=0.0, start=0.0, duration=1e9, ox=0.0, oy=0.0, oz=0.0):
        """Solver::addBurner (gasphase.cpp:72-90): returns its index, or -1 if
        it covers no gas cell."""
        arr = burner_cells(self.scen, x0, y0, x1, y1, z, ox, oy, oz)
        if arr is None:
            return -1
        f32 = np.float32
        self.burners.append(dict(cells=arr, dcells=self.ctx.to_device(arr), peak=f32(peak_w), ramp=f32(ramp_s),
                                 start=f32(start), duration=f32(duration)))
        return len(self.burners) - 1

    def inject_burners(self, dt):
        """Solver::injectBurners (gasphase.cpp:92-105)."""
        f32 = np.float32
        t = f32(self.time)
        for b in self.burners:
            if t < b["start"] or t > b["start"] + b["duration"]:
                continue
            r = min(f32(1.0), (t - b["start"]) / b["ramp"]) if b["ramp"] > 0 else f32(1.0)
            P = b["peak"] * r * r
            if P <= 0:
                continue
            mf = P * f32(dt) / f32(25.0e6) / f32(len(b["cells"]))
            self.ctx._check(self.ctx.L.phg_add_gas(self.ctx.h, self.grid.id, len(b["cells"]), b["dcells"].ptr,
                                                   mf, 0.0, mf * f32(0.01), 800.0, 1, C.byref(self._gas())))

    def combustion(self, dt, report=True):
        tot = C.c_double()
        self.ctx._check(self.ctx.L.phg_combustion(self.ctx.h, self.grid.id, dt, 1 if self.radiation_on else 0,
                                                  C.byref(self._gas()), C.byref(tot) if report else None))
        if report:
            self.hrr_total = tot.value
        return self.hrr_total

    def step_clear(self):
        self.ctx._check(self.ctx.L.phg_step_clear(self.ctx.h, self.grid.id, C.byref(self._gas())))

    def update_faces(self, upd):
        """Face changes between steps (barrier failure, a door opened or
        shut): the results of Solver::updateFace (solver.cpp:66-101) for a
        list of faces, as (axis, face, open fraction, coefficient, crack
        flag) arrays.  When the pressure matrix changes, the zones are
        computed again and each closed zone's background pressure is its
        cells' mean Pb, every cell's Pb its zone's (computeZones,
        gasphase.cpp:111-158); the hierarchy is rebuilt on the device side."""
        from . import scenarios as sc
        d = self.d
        rebuilt = self.grid.update_faces(upd, d["u"], d["v"], d["w"])
        self.scen = sc.apply_face_updates(self.scen, upd)
        self.grid.scen = self.scen
        if rebuilt:
            zone, zopen = self.grid.zones()
            Pb = self.get("Pb")
            nzn = len(zopen)
            g = zone >= 0
            cnt = np.bincount(zone[g], minlength=nzn).astype(np.float64)
            psum = np.bincount(zone[g], weights=Pb[g].astype(np.float64), minlength=nzn)
            zp = np.where(zopen != 0, np.float32(101325.0),
                          (psum / np.maximum(cnt, 1.0)).astype(np.float32)).astype(np.float32)
            self.zone_open = zopen
            self.zoneP = zp
            self.zoneDP = np.zeros(nzn, np.float32)
            dxd = np.float64(np.float32(self.scen.dx))
            self.zone_vol = (cnt * dxd * dxd * dxd).astype(np.float32)
            self.set("Pb", np.where(g, zp[np.maximum(zone, 0)], np.float32(101325.0)).astype(np.float32))
            self._zones_to_device()
        return rebuilt

    def _zones_to_device(self):
        if self.resident_zones and len(self.zone_open):
            zp = np.ascontiguousarray(self.zoneP, np.float32)
            self.ctx._check(self.ctx.L.phg_zone_pressures_set(self.ctx.h, self.grid.id, zp.ctypes.data))

    def zones_from_device(self):
        """zoneP and zoneDP from the device (resident zones; waits)."""
        if self.resident_zones and len(self.zone_open):
            zp = np.zeros(len(self.zone_open), np.float32)
            zd = np.zeros(len(self.zone_open), np.float32)
            self.ctx._check(self.ctx.L.phg_zone_pressures_get(self.ctx.h, self.grid.id, zp.ctypes.data,
                                                              zd.ctypes.data))
            self.zoneP[:] = zp
            self.zoneDP[:] = zd
        return self.zoneP

    def leakage_zones(self, dt):
        """Solver::leakage (gasphase.cpp:402-498) on the device
        (phg_zone_leakage): crack flow, each closed zone's background pressure
        root, the relaxed crack velocities, zoneDP and the shift of S.  With
        resident zones nothing comes back to the host (returns None)."""
        d = self.d
        if self.resident_zones:
            self.ctx._check(self.ctx.L.phg_zone_leakage_resident(self.ctx.h, self.grid.id, dt, d["T"].ptr,
                                                                 d["Pb"].ptr, d["p"].ptr, d["u"].ptr, d["v"].ptr,
                                                                 d["w"].ptr, d["S"].ptr))
            return None
        nz = len(self.zone_open)
        zdp = np.zeros(max(nz, 1), np.float32)
        zp = np.ascontiguousarray(self.zoneP, np.float32)
        nf = C.c_int()
        self.ctx._check(self.ctx.L.phg_zone_leakage(self.ctx.h, self.grid.id, dt, zp.ctypes.data, d["T"].ptr,
                                                    d["Pb"].ptr, d["p"].ptr, d["u"].ptr, d["v"].ptr, d["w"].ptr,
                                                    d["S"].ptr, zdp.ctypes.data, C.byref(nf)))
        self.zoneDP[:] = zdp[:nz]
        self.leakFaces = nf.value
        return self.zoneDP

    def zone_pressure(self, dt):
        """Closed zones' background pressure follows its rise rate (solver.cpp:258-263)."""
        f32 = np.float32
        if not len(self.zone_open) or self.zone_open.all():
            return
        if self.resident_zones:
            self.ctx._check(self.ctx.L.phg_zone_pressure_resident(self.ctx.h, self.grid.id, dt, self.d["Pb"].ptr))
            return
        for z in range(len(self.zone_open)):
            if not self.zone_open[z]:
                v = self.zoneP[z] + f32(dt) * self.zoneDP[z]
                self.zoneP[z] = min(max(v, f32(101325.0 - 20000.0)), f32(101325.0 + 50000.0))
        self.ctx._check(self.ctx.L.phg_zone_values(self.ctx.h, self.grid.id, self.zoneP.ctypes.data,
                                                   self.d["Pb"].ptr, 0))

    # ---- checkpoint and resume (the role of Solver::saveState / loadState,
    # src/sim/state.cpp:62-195): every field a later step reads, the zone
    # pressures, the step statistics stableDt uses, time and burners
    STATE_MAGIC = "PHGGAS01"
    STATE_FIELDS = ("u", "v", "w", "p", "S", "nu", "Pb", "hrr", "lit") + FIELD_NAMES

    def fingerprint(self):
        """Hash of the geometry and wind a state belongs to."""
        import hashlib
        h = hashlib.sha256()
        g = self.scen
        h.update(np.array([g.nx, g.ny, g.nz], np.int64).tobytes())
        h.update(np.float32(g.dx).tobytes())
        for a in (g.fx, g.fy, g.fz, g.ax, g.ay, g.az):
            h.update(np.ascontiguousarray(a, np.float32).tobytes())
        h.update(np.ascontiguousarray(g.blocked, np.uint8).tobytes())
        h.update(np.array([self.wind.speed, self.wind.from_deg, self.wind.exponent, self.wind.oz], np.float32).tobytes())
        return h.hexdigest()

    def snapshot(self):
        """The full state as host arrays (a dict)."""
        self.zones_from_device()
        st = {"magic": self.STATE_MAGIC, "fingerprint": self.fingerprint(), "time": np.float32(self.time),
              "zoneP": self.zoneP.copy(),
              "stats": np.array([self.maxVel, self.maxNu, self.maxDiv], np.float32),
              "pressureIters": self.pressureIters,
              "burners": [{k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in b.items() if k != "dcells"}
                          for b in self.burners]}
        for f in self.STATE_FIELDS:
            st["f_" + f] = self.get(f)
        return st

    def restore(self, st):
        """Load a snapshot taken from a solver of the same geometry and wind
        (refused otherwise, with a message)."""
        if st.get("magic") != self.STATE_MAGIC:
            raise PhgError("not a gas state")
        if st.get("fingerprint") != self.fingerprint():
            raise PhgError("the state belongs to another geometry or wind")
        for f in self.STATE_FIELDS:
            self.set(f, st["f_" + f])
        self.time = float(st["time"])
        self.zoneP[:] = st["zoneP"]
        self._zones_to_device()
        self.maxVel, self.maxNu, self.maxDiv = (float(x) for x in st["stats"])
        self.pressureIters = int(st["pressureIters"])
        self.burners = []
        for b in st["burners"]:
            nb = dict(b)
            nb["dcells"] = self.ctx.to_device(np.asarray(b["cells"], np.int32))
            self.burners.append(nb)

    def save_state(self, path):
        """Write the snapshot as an .npz of plain arrays (no pickled objects)."""
        st = self.snapshot()
        flat = {k: v for k, v in st.items() if k != "burners"}
        flat["magic"] = np.array(st["magic"])
        flat["fingerprint"] = np.array(st["fingerprint"])
        flat["pressureIters"] = np.int32(st["pressureIters"])
        for i, b in enumerate(st["burners"]):
            flat[f"burner{i}_cells"] = b["cells"]
            flat[f"burner{i}_params"] = np.array([b["peak"], b["ramp"], b["start"], b["duration"]], np.float32)
        with open(path, "wb") as f:
            np.savez(f, **flat)

    def load_state(self, path):
        with np.load(path, allow_pickle=False) as z:
            st = {k: z[k] for k in z.files}
        st["magic"], st["fingerprint"] = str(st["magic"]), str(st["fingerprint"])
        burners, i = [], 0
        while f"burner{i}_cells" in st:
            peak, ramp, start, dur = (np.float32(x) for x in st.pop(f"burner{i}_params"))
            burners.append(dict(cells=st.pop(f"burner{i}_cells"), peak=peak, ramp=ramp, start=start, duration=dur))
            i += 1
        st["burners"] = burners
        self.restore(st)

    def stable_dt(self, dt_max):
        out = C.c_float()
        self.ctx._check(self.ctx.L.phg_stable_dt(dt_max, self.dt_max_profile, self.cfl, self.scen.dx, self.maxVel,
                                                 self.maxNu, self.maxDiv, C.byref(out)))
        return out.value

    def step(self, dt_max):
        """Solver::step (src/sim/solver.cpp:226-269) for a gas-only world (no
        radiation, solids or barriers): returns dt."""
        dt = self.stable_dt(dt_max)
        self.step_clear()
        self.inject_burners(dt)
        self.combustion(dt)
        # maxNu and maxDiv stay on the device until the solve has waited
        self.turbulence(report=False)
        self.heat_diffusion(dt)
        self.energy_and_expansion(dt, report=False)
        self.momentum(dt)
        self.leakage_zones(dt)
        self.project(dt)
        self.read_stats()
        self.advect_scalars(dt)
        self.zone_pressure(dt)
        self.time = float(np.float32(self.time) + np.float32(dt))
        return dt

    def advect_scalars(self, dt):
        d = self.d
        ins = (C.c_void_p * 10)(*[d[f].ptr for f in FIELD_NAMES])
        outs = (C.c_void_p * 10)(*[t.ptr for t in self._tmp])
        self.ctx._check(self.ctx.L.phg_advect_scalars(self.ctx.h, self.grid.id, dt, d["u"].ptr, d["v"].ptr,
                                                      d["w"].ptr, d["Pb"].ptr, d["nu"].ptr, ins, outs))
        # swap, as the reference swaps tmp_ into the fields (gasphase.cpp:783)
        for k, f in enumerate(FIELD_NAMES):
            d[f], self._tmp[k] = self._tmp[k], d[f]
