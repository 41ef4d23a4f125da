"""CPU tests of the host-side logic: the scenario generator, the library's
host-built hierarchy against the oracle, and the C ABI's exported symbols."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import pyoracle as po
from photon_amd import gas
from photon_amd import scenarios as sc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_splitmix64_known_answer():
    # SplitMix64 from seed 0: first outputs of the published reference generator
    out = sc.splitmix64(0, 3)
    assert [int(v) for v in out] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_normal_field_moments_and_determinism():
    a = sc.normal_field(42, 200001)
    assert abs(a.mean()) < 0.01 and abs(a.std() - 1) < 0.01
    assert np.array_equal(a, sc.normal_field(42, 200001))


def test_generator_is_seeded_and_refines():
    g1, g2 = sc.house(1), sc.house(1)
    assert np.array_equal(g1.blocked, g2.blocked) and np.array_equal(g1.ax, g2.ax)
    assert not np.array_equal(sc.house(2).blocked, g1.blocked)
    fine = sc.house(1, 160, 100, 80, 0.1)
    # a 0.2 m cell is solid iff its 8 children at 0.1 m are
    c = g1.blocked.reshape(40, 50, 80)
    f = fine.blocked.reshape(80, 100, 160).reshape(40, 2, 50, 2, 80, 2).min(axis=(1, 3, 5))
    assert np.array_equal(c, f)


def test_zero_mean_rhs_per_closed_zone():
    g = sc.house(1)
    S = sc.divergence_field(g, 5).astype(np.float64)
    for z in range(len(g.zone_open)):
        m = g.zone == z
        assert abs(S[m].mean()) < 1e-6
    assert np.all(S[g.blocked != 0] == 0)


def test_zones_agree_with_oracle():
    for g in (sc.room(), sc.house(3), sc.batch_grid(2)):
        zone, zo = po.zones(g)
        assert np.array_equal(zone, g.zone) and np.array_equal(zo, g.zone_open)


def test_library_hierarchy_matches_oracle():
    for g in (sc.room(), sc.house(7, 24, 16, 12), sc.house(1), sc.batch_grid(5)):
        sizes, agg0 = gas.hierarchy_shape(g)
        H = po.Hierarchy(g)
        assert sizes == [H.level(l)["n"] for l in range(H.nlevels)]
        assert np.array_equal(agg0, H.level(0)["agg"])


def test_grid_plan_builds_without_a_device():
    """The host half of grid set-up runs with no context, on worker threads
    at once, and refuses shapes phg_grid_create refuses."""
    import concurrent.futures
    gs = [sc.room(), sc.house(7, 24, 16, 12), sc.batch_grid(5)]
    with concurrent.futures.ThreadPoolExecutor(3) as pool:
        plans = list(pool.map(gas.GridPlan, gs))
    for p in plans:
        assert p.h
        p.free()
        assert not p.h
    g = sc.room()
    L = gas.load_library()
    h = ctypes.c_void_p()
    arrs = [np.ascontiguousarray(a, np.float32) for a in (g.ax, g.ay, g.az)]
    blk = np.ascontiguousarray(g.blocked, np.uint8)
    assert L.phg_grid_plan_build(1, g.ny, g.nz, *[a.ctypes.data for a in arrs], blk.ctypes.data, None,
                                 ctypes.byref(h)) != 0
    assert not h.value


def pocket_grid(n):
    """n^3 cells cut into sealed 2x2x2 pockets: open faces only inside a
    pocket, so coarsening stalls at level 1 with one unknown per pocket."""
    fx = np.zeros((n, n, n + 1), np.float32)
    fx[:, :, 1:n:2] = 1.0
    fy = np.zeros((n, n + 1, n), np.float32)
    fy[:, 1:n:2, :] = 1.0
    fz = np.zeros((n + 1, n, n), np.float32)
    fz[1:n:2] = 1.0
    return sc.Grid(n, n, n, 0.2, np.zeros(n ** 3, np.uint8), fx.ravel(), fy.ravel(), fz.ravel(), fx.ravel().copy(),
                   fy.ravel().copy(), fz.ravel().copy())


def test_stalled_coarsening_plans_promptly_or_refuses_cleanly():
    """Thousands of sealed pockets stall the coarsening (C.n > 0.8 F.n,
    pressure_mg.cpp:176): with 13,824 coarsest unknowns the plan takes the
    reference's Gauss-Seidel sweeps (no dense inverse, built at once); with
    32,768, beyond what the sweeps' level schedule keeps, the host set-up
    refuses with PHG_EINVAL, promptly, instead of an O(n^3) inverse or an
    exception through the C ABI."""
    import time
    t = time.perf_counter()
    p = gas.GridPlan(pocket_grid(48))
    p.free()
    with pytest.raises(gas.PhgError, match="-1"):
        gas.GridPlan(pocket_grid(64))
    assert time.perf_counter() - t < 10.0


def test_abi_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "photon_gas.h")).read()
    declared = set(re.findall(r"\b(phg_\w+)\s*\(", header))
    assert declared, "no declarations parsed"
    lib = ctypes.CDLL(gas.LIB_PATH)
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert declared <= set(gas.EXPORTS)


def test_abi_builds_from_cpp_and_c(tmp_path):
    """The header compiles as C and as C++, and a C++ caller (as the reference's
    Solver would be) links the library and runs its host-only half of grid
    set-up; without a GPU, phg_create refuses with an error code."""
    import shutil
    import subprocess
    if not shutil.which("g++") or not shutil.which("gcc"):
        pytest.skip("no host compiler")
    inc = os.path.join(ROOT, "include")
    src = tmp_path / "caller.cpp"
    src.write_text(r'''
#include "photon_gas.h"
#include <cstdio>
#include <vector>
int main() {
    const int nx = 6, ny = 5, nz = 4, n = nx * ny * nz;
    std::vector<float> ax((nx + 1) * ny * nz, 1.0f), ay(nx * (ny + 1) * nz, 1.0f), az(nx * ny * (nz + 1), 1.0f);
    std::vector<uint8_t> blocked(n, 0);
    phg_grid_plan* plan = nullptr;
    const int rc = phg_grid_plan_build(nx, ny, nz, ax.data(), ay.data(), az.data(), blocked.data(), nullptr, &plan);
    std::printf("plan %d %d\n", rc, plan != nullptr);
    phg_grid_plan_free(plan);
    phg_ctx* ctx = nullptr;
    std::printf("create %d\n", phg_create(&ctx, 0));
    if (ctx) phg_destroy(ctx);
    return 0;
}
''')
    exe = tmp_path / "caller"
    libdir = os.path.dirname(gas.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-I", inc, str(src), "-L", libdir, "-lphoton_gas",
                    f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True, timeout=120)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120).stdout.split("\n")
    assert out[0] == "plan 0 1"
    assert out[1].startswith("create ")
    subprocess.run(["gcc", "-x", "c", "-std=c99", "-fsyntax-only", "-I", inc, "-"], input='#include "photon_gas.h"\n',
                   text=True, check=True, timeout=60)


def test_no_gpu_gives_clear_error_not_fallback():
    # in a container without a GPU the context must refuse, never compute on the CPU
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from photon_amd import gas\n"
            "try:\n    gas.Context(0)\nexcept gas.PhgError as e:\n    print('refused', e)\n"
            "else:\n    print('context')\n") % ROOT
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120).stdout
    assert "refused" in out or "context" in out


def test_serving_stream_and_sizes():
    from photon_amd import serving as sv
    specs = sv.poisson_stream(5, rate_per_s=20.0, window_s=30.0, steps=10)
    # Poisson arrivals: about rate * window requests, sorted arrival times
    assert 480 < len(specs) < 720
    t = np.array([s.arrival_s for s in specs])
    assert np.all(np.diff(t) > 0) and t[-1] <= 30.0
    # sizes log-uniform in [3.6k, 1.28M]: the log-size median sits mid-range
    logc = np.log([s.cells for s in specs])
    assert np.log(3600) - 0.01 <= logc.min() and logc.max() <= np.log(1_280_000) + 0.01
    assert abs(np.median(logc) - 0.5 * (np.log(3600) + np.log(1_280_000))) < 0.6
    assert all(10.0 <= s.hrr_kw <= 500.0 for s in specs)
    # the grid keeps the room's 4 : 3 : 2.4 aspect and the requested size
    g = sv.grid_for(sv.JobSpec(0, 1, 3600, 1, 100.0))
    assert (g.nx, g.ny, g.nz) == (20, 15, 12)
    g = sv.grid_for(sv.JobSpec(0, 1, 200_000, 1, 100.0))
    assert abs(g.n / 200_000 - 1) < 0.1
    assert sv.footprint_bytes(g) > sv.footprint_bytes(sv.grid_for(sv.JobSpec(0, 1, 3600, 1, 1.0)))
