"""GPU parity of the MG-PCG path against the CPU oracle, through the C ABI.

Bars: the level-0 operator and dot product match the oracle to FP32 rounding;
the V-cycle matches the oracle's V-cycle with the same (exact, FP64) coarsest
solve to 1e-5; solutions at ||r||/||b|| <= 1e-6 match the FP64 oracle and the
reference-method FP32 oracle within 1e-5 relative L2;
batched results equal single-system results bit for bit (runs of 32+ systems
on one grid, whose coarsest product runs on the FP64 matrix cores, to FP64
rounding of it); runs repeat bit for bit.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from photon_amd import gas
from photon_amd import scenarios as sc

pytestmark = pytest.mark.gpu

DT = 0.01


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def small():
    g = sc.house(7, 24, 16, 12)
    return g, po.Hierarchy(g)


@pytest.fixture(scope="module")
def house():
    g = sc.house(1)
    return g, po.Hierarchy(g)


def test_hierarchy_matches_oracle(ctx, small, house):
    for g, H in (small, house):
        G = gas.Grid(ctx, g)
        info = G.info()
        ref = [H.level(l)["n"] for l in range(H.nlevels)]
        assert info["levels"][:len(ref)] == ref
        assert info["ref_levels"] == len(ref)
        zone, zo = G.zones()
        assert np.array_equal(zone, g.zone) and np.array_equal(zo, g.zone_open)
        G.destroy()


def test_apply_matches_oracle(ctx, small):
    g, H = small
    G = gas.Grid(ctx, g)
    x = sc.normal_field(5, g.n).astype(np.float32)
    y_ref, dot_ref = H.apply(x)
    dx, dy = ctx.to_device(x), ctx.array(g.n)
    dot = G.apply(dx, dy)
    y = dy.download()
    np.testing.assert_allclose(y, y_ref, rtol=1e-6, atol=1e-5)
    assert abs(dot - dot_ref) <= 1e-6 * abs(dot_ref)
    G.destroy()


@pytest.mark.parametrize("prec,tol", [(gas.PREC_FP32, 1e-5), (gas.PREC_MIXED_F16, 3e-3), (gas.PREC_MIXED_BF16, 3e-2)])
def test_vcycle_matches_oracle(ctx, small, prec, tol):
    g, H = small
    G = gas.Grid(ctx, g, precision=prec)
    r = sc.rhs_from_divergence(g, sc.divergence_field(g, 11), DT)
    z_ref = H.precondition(r, coarse=po.COARSE_EXACT)
    dr, dz = ctx.to_device(r), ctx.array(g.n)
    G.precondition(dr, dz)
    z = dz.download()
    assert rel(z, z_ref) < tol
    G.destroy()


@pytest.mark.parametrize("prec", [gas.PREC_FP32, gas.PREC_MIXED_F16, gas.PREC_MIXED_BF16])
def test_solve_matches_fp64_oracle(ctx, house, prec):
    g, H = house
    b = sc.rhs_from_divergence(g, sc.divergence_field(g, 123), DT)
    x64, _, _ = H.pcg64(b)
    x_cpu, it_cpu, _ = H.pcg(b, tol=1e-6)           # reference method, FP32
    x_cpu = po.closed_zone_demean(g, x_cpu)
    G = gas.Grid(ctx, g, precision=prec)
    db, dx = ctx.to_device(b), ctx.array(g.n)
    dx.zero()
    st = G.solve(db, dx, gas.solve_params(tol=1e-6, max_iters=200))
    x = dx.download()
    print(f"prec={prec} iters gpu={st['iters']} cpu={it_cpu} relres={st['relres']:.3e} "
          f"err_fp64={rel(x, x64):.3e} diff_cpu={rel(x, x_cpu):.3e} true={H.true_relres(b, x):.3e}")
    assert st["status"] == gas.CONVERGED and st["relres"] <= 1e-6
    assert rel(x, x64) < 1e-5
    assert rel(x, x_cpu) < 1e-5
    assert st["iters"] <= it_cpu + 10
    G.destroy()


def test_batch_equals_single_and_repeats(ctx, small):
    g, H = small
    G = gas.Grid(ctx, g)
    bs = [sc.rhs_from_divergence(g, sc.divergence_field(g, 100 + s), DT) for s in range(5)]
    p = gas.solve_params(tol=1e-6, max_iters=200)
    singles = []
    for b in bs:
        db, dx = ctx.to_device(b), ctx.array(g.n)
        dx.zero()
        G.solve(db, dx, p)
        singles.append(dx.download())
    B = gas.Batch(ctx, [G] * len(bs))
    for rep in range(2):
        dbs = [ctx.to_device(b) for b in bs]
        dxs = [ctx.array(g.n) for _ in bs]
        for d in dxs:
            d.zero()
        stats = B.solve(dbs, dxs, p)
        for k, d in enumerate(dxs):
            assert stats[k]["status"] == gas.CONVERGED
            assert np.array_equal(d.download(), singles[k]), f"system {k} rep {rep}"
    B.destroy()
    G.destroy()


# walled grids with an odd nx and an even plane size: the fused pass's column
# pairs then straddle row ends; one per pair-slot configuration (nx <= 64,
# <= 127, <= 192)
ODD_NX = {"odd37": (11, 37, 22, 10), "odd101": (12, 101, 14, 8), "odd131": (13, 131, 10, 6)}


@pytest.mark.parametrize("prec", [gas.PREC_FP32, gas.PREC_MIXED_F16])
@pytest.mark.parametrize("grid,kchunk,width", [("small", 3, 0), ("small", 0, 256), ("house", 7, 0), ("house", 40, 0),
                                               ("house", 40, 256), ("house", 40, 1024), ("odd37", 3, 0),
                                               ("odd101", 0, 256), ("odd101", 0, 1024), ("odd131", 4, 0),
                                               ("odd131", 4, 256), ("odd131", 4, 512)])
def test_fused_presmoothing_is_bitwise_the_separate_passes(ctx, small, house, prec, grid, kchunk, width):
    """The fused smoothing passes form every value with the same operations
    as the separate half sweeps, so whole solves agree bit for bit, for
    k-chunks that cut the grid unevenly and for whole columns, on grids of
    many column tiles (their halo rows belong to other blocks), on grids of
    odd nx, where a column pair can hold the last cell of one row and the
    first of the next, and for blocks of 256, 512 and 1024 threads (the
    default: 512, and 1024 for grids wider than 128 cells)."""
    if grid in ODD_NX:
        g = sc.house(*ODD_NX[grid])
        assert g.nx % 2 == 1 and (g.nx * g.ny) % 2 == 0
    else:
        g, _ = small if grid == "small" else house
    G = gas.Grid(ctx, g, precision=prec)
    bs = [sc.rhs_from_divergence(g, sc.divergence_field(g, 300 + s), DT) for s in range(8)]
    p = gas.solve_params(tol=1e-6, max_iters=200, zero_x=True)
    res = []
    for unfused in (False, True):
        B = gas.Batch(ctx, [G] * len(bs), kchunk=kchunk, unfused=unfused, fused_threads=width)
        dbs = [ctx.to_device(b) for b in bs]
        dxs = [ctx.array(g.n) for _ in bs]
        stats = B.solve(dbs, dxs, p)
        res.append(([d.download() for d in dxs], [s["iters"] for s in stats]))
        assert all(s["status"] == gas.CONVERGED for s in stats)
        B.destroy()
    assert res[0][1] == res[1][1]
    for a, b in zip(res[0][0], res[1][0]):
        assert np.array_equal(a, b)
    G.destroy()


@pytest.mark.parametrize("prec", [gas.PREC_MIXED_F16, gas.PREC_MIXED_BF16])
@pytest.mark.parametrize("grid,kchunk,warm,xoff", [("small", 3, False, 0), ("house", 7, True, 0),
                                                   ("house", 40, False, 1), ("odd37", 3, True, 1),
                                                   ("odd131", 4, False, 0), ("wide", 5, False, 0)])
def test_direction_update_on_column_pairs_is_bitwise_one_column_per_thread(ctx, small, house, prec, grid, kchunk,
                                                                           warm, xoff):
    """In the mixed precisions (FP16 direction) the direction update runs on
    column pairs by default for 0/1 grids with an even plane size up to 256
    cells wide (k_dir_apply_pair: 256 threads, two adjacent cells per lane);
    a batch that asks for 512-thread blocks keeps one column per thread.  The pair form sums each column's
    d.q terms along k and reduces them over the lanes and tiles the
    one-column form uses, so whole solves agree bit for bit: iteration
    counts and solutions, with x read and written on every second iteration
    (warm starts included; as pairs, or one cell at a time when x is not
    8-byte aligned), for pairs that straddle a row end (odd nx), a grid 256
    cells wide (the widest the pair form takes) and k-chunks that cut the
    grid unevenly."""
    if grid in ODD_NX:
        g = sc.house(*ODD_NX[grid])
    elif grid == "wide":
        g = sc.house(14, 256, 6, 9)
    else:
        g, _ = small if grid == "small" else house
    assert (g.nx * g.ny) % 2 == 0 and g.nx <= 256
    G = gas.Grid(ctx, g, precision=prec)
    bs = [sc.rhs_from_divergence(g, sc.divergence_field(g, 700 + s), DT) for s in range(4)]
    x0 = [(0.01 * sc.divergence_field(g, 900 + s)).astype(np.float32) for s in range(4)]
    p = gas.solve_params(tol=1e-6, max_iters=200, zero_x=not warm)
    res = []
    for width in (0, 512):
        B = gas.Batch(ctx, [G] * len(bs), kchunk=kchunk, fused_threads=width)
        dbs = [ctx.to_device(b) for b in bs]
        # x at an offset of xoff floats inside a larger buffer (xoff = 1: 4-byte aligned only)
        bufs = [ctx.to_device(np.concatenate([np.zeros(xoff, np.float32), x if warm else np.zeros_like(x)]))
                for x in x0]
        stats = B.solve(dbs, [d.ptr + 4 * xoff for d in bufs], p)
        assert all(s["status"] == gas.CONVERGED for s in stats)
        res.append(([d.download()[xoff:] for d in bufs], [s["iters"] for s in stats]))
        B.destroy()
    assert res[0][1] == res[1][1]
    for k, (a, b) in enumerate(zip(res[0][0], res[1][0])):
        assert np.array_equal(a, b), f"system {k}"
    G.destroy()


@pytest.mark.parametrize("prec", [gas.PREC_FP32, gas.PREC_MIXED_F16])
def test_coarse_weight_one_is_the_reference_cycle(ctx, house, prec):
    """coarse_weight 1 and the default 0 both give the reference's V-cycle:
    the prolongation then adds 4 t (the level scale), as before the weight
    existed, so the solves agree bit for bit."""
    g, _ = house
    G = gas.Grid(ctx, g, precision=prec)
    bs = [sc.rhs_from_divergence(g, sc.divergence_field(g, 40 + s), DT) for s in range(3)]
    res = []
    for cw in (0.0, 1.0):
        B = gas.Batch(ctx, [G] * len(bs))
        dbs = [ctx.to_device(b) for b in bs]
        dxs = [ctx.array(g.n) for _ in bs]
        st = B.solve(dbs, dxs, gas.solve_params(tol=1e-6, max_iters=200, zero_x=True, coarse_weight=cw))
        res.append(([d.download() for d in dxs], [s["iters"] for s in st]))
        B.destroy()
    assert res[0][1] == res[1][1]
    for a, b in zip(res[0][0], res[1][0]):
        assert np.array_equal(a, b)
    G.destroy()


@pytest.mark.parametrize("prec", [gas.PREC_FP32, gas.PREC_MIXED_F16, gas.PREC_MIXED_BF16])
def test_over_corrected_vcycle_meets_the_fp64_oracle_in_fewer_iterations(ctx, house, prec):
    """With the coarse-grid corrections weighted by COARSE_WEIGHT the V-cycle
    is no longer the reference's, so the bar is the solution: within 1e-5 of
    the FP64 oracle (a direct-solve-pinned FP64 PCG, independent of the
    preconditioner) at ||r||/||b|| <= 1e-6, with the true residual formed in
    FP64 from the stored x at its FP32 floor, and in at most three quarters of
    the reference cycle's iterations on the walled house."""
    g, H = house
    G = gas.Grid(ctx, g, precision=prec)
    bs = [sc.rhs_from_divergence(g, sc.divergence_field(g, 123 + s), DT) for s in range(2)]
    its = {}
    for cw in (0.0, gas.COARSE_WEIGHT):
        B = gas.Batch(ctx, [G] * len(bs))
        dbs = [ctx.to_device(b) for b in bs]
        dxs = [ctx.array(g.n) for _ in bs]
        st = B.solve(dbs, dxs, gas.solve_params(tol=1e-6, max_iters=200, zero_x=True, coarse_weight=cw))
        its[cw] = [s["iters"] for s in st]
        assert all(s["status"] == gas.CONVERGED and s["relres"] <= 1e-6 for s in st)
        if cw > 0:
            for k, b in enumerate(bs):
                x = dxs[k].download()
                x64, _, _ = H.pcg64(b)
                err, true = rel(x, x64), H.true_relres(b, x)
                print(f"prec={prec} system {k}: iters {its[0.0][k]} -> {its[cw][k]}, err_fp64={err:.2e}, "
                      f"true relres {true:.2e}")
                assert err < 1e-5
                assert true < 1e-4   # an FP32 x itself leaves ~1e-5 (so does the reference cycle)
        B.destroy()
    for a, b in zip(its[0.0], its[gas.COARSE_WEIGHT]):
        assert b <= 0.75 * a
    G.destroy()


@pytest.mark.parametrize("prec", [gas.PREC_FP32, gas.PREC_MIXED_F16])
def test_weighted_prolongations_agree_across_kernel_forms(ctx, small, house, prec):
    """Every prolongation applies the weight: the fused post pass and the
    separate black sweep at level 0, the LDS up pass and the per-level
    launches on the coarse levels.  With coarse_weight = COARSE_WEIGHT the
    fused and separate passes, and the LDS and per-level coarse forms, give
    the same solves bit for bit, and a system alone equals it in a batch."""
    for g in (small[0], house[0]):
        G = gas.Grid(ctx, g, precision=prec)
        bs = [sc.rhs_from_divergence(g, sc.divergence_field(g, 60 + s), DT) for s in range(3)]
        p = gas.solve_params(tol=1e-6, max_iters=200, zero_x=True, coarse_weight=gas.COARSE_WEIGHT)
        res = []
        for opts in ({}, {"unfused": True}, {"coarse_lds": 2}, {"coarse_lds": 1}):
            B = gas.Batch(ctx, [G] * len(bs), **opts)
            dbs = [ctx.to_device(b) for b in bs]
            dxs = [ctx.array(g.n) for _ in bs]
            st = B.solve(dbs, dxs, p)
            assert all(s["status"] == gas.CONVERGED for s in st)
            res.append(([d.download() for d in dxs], [s["iters"] for s in st]))
            B.destroy()
        B = gas.Batch(ctx, [G])
        db, dx = ctx.to_device(bs[1]), ctx.array(g.n)
        st = B.solve([db], [dx], p)
        res.append(([None, dx.download(), None], [None, st[0]["iters"], None]))
        B.destroy()
        for r in res[1:]:
            for k in range(len(bs)):
                if r[0][k] is None:
                    continue
                assert r[1][k] == res[0][1][k]
                assert np.array_equal(r[0][k], res[0][0][k]), (g.nx, k)
        G.destroy()


@pytest.mark.parametrize("cw", [2.0, -0.5, float("nan")])
def test_coarse_weight_outside_its_range_is_refused(ctx, small, cw):
    g, _ = small
    G = gas.Grid(ctx, g)
    db, dx = ctx.to_device(np.ones(g.n, np.float32)), ctx.array(g.n)
    with pytest.raises(gas.PhgError):
        G.solve(db, dx, gas.solve_params(tol=1e-6, max_iters=10, coarse_weight=cw))
    G.destroy()


@pytest.mark.parametrize("prec", [gas.PREC_FP32, gas.PREC_MIXED_F16])
@pytest.mark.parametrize("width", [0, 256])
def test_batch_splits_smoothing_per_system(ctx, prec, width):
    """Fused passes are chosen per system: in one batch the grids of even
    plane size and nx <= 192 smoothed once take them (by default on 512-thread
    blocks up to 128 cells wide and 1024-thread blocks above, each group in
    the configuration its widest grid needs), while a grid of odd plane size,
    one wider than 192 cells and a system given two sweeps take the separate
    half sweeps.
    Every system's solve is bit for bit the all-separate batch's."""
    gs = [sc.house(21, 24, 18, 10), sc.house(22, 25, 19, 8), sc.house(23, 200, 12, 6),
          sc.house(24, 131, 10, 6), sc.house(25, 24, 18, 10)]
    assert (gs[1].nx * gs[1].ny) % 2 == 1 and gs[2].nx > 192
    Gs = [gas.Grid(ctx, g, precision=prec) for g in gs]
    sweeps = [1, 1, 1, 1, 2]
    bs = [sc.rhs_from_divergence(g, sc.divergence_field(g, 500 + s), DT) for s, g in enumerate(gs)]
    p = gas.solve_params(tol=1e-6, max_iters=200, zero_x=True)
    res = []
    for unfused in (False, True):
        B = gas.Batch(ctx, Gs, kchunk=4, unfused=unfused, sys_sweeps0=sweeps, fused_threads=width)
        dbs = [ctx.to_device(b) for b in bs]
        dxs = [ctx.array(g.n) for g in gs]
        stats = B.solve(dbs, dxs, p)
        assert all(s["status"] == gas.CONVERGED for s in stats)
        res.append(([d.download() for d in dxs], [s["iters"] for s in stats]))
        B.destroy()
    assert res[0][1] == res[1][1]
    for k, (a, b) in enumerate(zip(res[0][0], res[1][0])):
        assert np.array_equal(a, b), f"system {k}"
    for G in Gs:
        G.destroy()


@pytest.mark.parametrize("prec", [gas.PREC_FP32, gas.PREC_MIXED_F16, gas.PREC_MIXED_BF16])
def test_coarse_levels_in_lds_equal_split_launches(ctx, prec):
    """The coarse levels that fit in one block's LDS run as one down and one
    up launch per system (k_vc_down / k_vc_up; forced here, by default only for
    batches of at least two such blocks per CU, so not for this one); every
    system's solve is bit for bit the one whose coarse levels each take their
    own launches.  The
    mix holds the 80x50x40 house twice (its LDS levels start at level 1 in
    FP16 and BF16, at level 2 in FP32), batch scenarios of four sizes (three
    or four levels) and, in FP16, the 160x100x80 house (LDS levels from
    level 2, level 1 on the separate launches)."""
    house = sc.house(1)
    gs = [house, house] + [sc.batch_grid(s) for s in (0, 3, 5, 10)]
    if prec == gas.PREC_MIXED_F16:
        gs.append(sc.house(1, 160, 100, 80, 0.1))
    Gs = [gas.Grid(ctx, g, precision=prec) for g in gs[1:]]
    Gs.insert(0, Gs[0])   # the house twice on one grid
    bs = [sc.rhs_from_divergence(g, sc.divergence_field(g, 700 + s), DT) for s, g in enumerate(gs)]
    p = gas.solve_params(tol=1e-6, max_iters=200, zero_x=True)
    res = []
    # 1: per-level launches; 2: LDS levels; 0 (default): per-level launches for
    # a batch this small
    for mode in (1, 2, 0):
        B = gas.Batch(ctx, Gs, coarse_lds=mode)
        dbs = [ctx.to_device(b) for b in bs]
        dxs = [ctx.array(g.n) for g in gs]
        ctx.profile(True, classes=["vc_down", "vc_up"])
        ctx.profile_reset()
        stats = B.solve(dbs, dxs, p)
        prof = ctx.profile_read()
        ctx.profile(False)
        assert all(s["status"] == gas.CONVERGED for s in stats), stats
        res.append(([d.download() for d in dxs], [s["iters"] for s in stats]))
        # the LDS launches run exactly when the batch asks for them
        assert (prof["vc_down"][1] > 0) == (mode == 2) and (prof["vc_up"][1] > 0) == (mode == 2)
        B.destroy()
    for other in res[1:]:
        assert other[1] == res[0][1]
        for k, (a, b) in enumerate(zip(res[0][0], other[0])):
            assert np.array_equal(a, b), f"system {k}"
    for G in Gs[1:]:
        G.destroy()


def test_graph_replay_matches_direct(ctx, small):
    g, H = small
    G = gas.Grid(ctx, g)
    b = sc.rhs_from_divergence(g, sc.divergence_field(g, 77), DT)
    out = []
    for use_graph in (False, True, True):
        db, dx = ctx.to_device(b), ctx.array(g.n)
        dx.zero()
        st = G.solve(db, dx, gas.solve_params(tol=1e-6, max_iters=200, use_graph=use_graph))
        out.append((st["iters"], dx.download()))
    assert out[0][0] == out[1][0] == out[2][0]
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[1][1], out[2][1])
    G.destroy()


def test_mixed_grids_batch(ctx):
    grids = [sc.batch_grid(s) for s in range(4)]
    Gs = [gas.Grid(ctx, g) for g in grids]
    B = gas.Batch(ctx, Gs)
    bs = [sc.rhs_from_divergence(g, sc.divergence_field(g, 9 + k), DT) for k, g in enumerate(grids)]
    dbs = [ctx.to_device(b) for b in bs]
    dxs = [ctx.array(g.n) for g in grids]
    for d in dxs:
        d.zero()
    stats = B.solve(dbs, dxs, gas.solve_params(tol=1e-6, max_iters=200))
    for k, g in enumerate(grids):
        H = po.Hierarchy(g)
        x64, _, _ = H.pcg64(bs[k])
        x = dxs[k].download()
        assert stats[k]["status"] == gas.CONVERGED
        assert rel(x, x64) < 1e-5, (k, g.shape, rel(x, x64))
    B.destroy()
    for G in Gs:
        G.destroy()


@pytest.mark.parametrize("prec", [gas.PREC_FP32, gas.PREC_MIXED_F16])
def test_fractional_coefficients(ctx, prec):
    """Partial openings give fractional matrix coefficients: the face-array
    path of the level-0 kernels (the 0/1 face mask does not apply)."""
    g = sc.vented(sc.house(5, 40, 30, 20), 3)
    assert np.any((g.ax > 0) & (g.ax < 1))
    H = po.Hierarchy(g)
    b = sc.rhs_from_divergence(g, sc.divergence_field(g, 8), DT)
    x64, _, _ = H.pcg64(b)
    G = gas.Grid(ctx, g, precision=prec)
    db, dx = ctx.to_device(b), ctx.array(g.n)
    st = G.solve(db, dx, gas.solve_params(tol=1e-6, max_iters=300, zero_x=True))
    x = dx.download()
    print(f"vented prec={prec} iters={st['iters']} err_fp64={rel(x, x64):.2e}")
    assert st["status"] == gas.CONVERGED and rel(x, x64) < 1e-5
    G.destroy()


def test_precision_config_1p28M_cells(ctx):
    """The precision workload: the house generator refined to 0.1 m
    (160x100x80 = 1.28M cells, 5 levels).  FP16 and BF16 V-cycle storage
    against FP32 storage: iterations and error against the FP64 oracle
    (solved to 1e-12), per right-hand side."""
    g = sc.house(1, 160, 100, 80, 0.1)
    H = po.Hierarchy(g)
    rows = []
    for seed in (5, 6):
        b = sc.rhs_from_divergence(g, sc.divergence_field(g, seed), DT)
        x64, it64, _ = H.pcg64(b)
        for prec in (gas.PREC_FP32, gas.PREC_MIXED_F16, gas.PREC_MIXED_BF16):
            G = gas.Grid(ctx, g, precision=prec)
            db, dx = ctx.to_device(b), ctx.array(g.n)
            st = G.solve(db, dx, gas.solve_params(tol=1e-6, max_iters=300, zero_x=True))
            x = dx.download()
            err = rel(x, x64)
            rows.append((seed, prec, st["iters"], st["relres"], err))
            print(f"1.28M seed={seed} prec={prec} iters={st['iters']} relres={st['relres']:.3e} err_fp64={err:.3e}")
            assert st["status"] == gas.CONVERGED and st["relres"] <= 1e-6
            assert err < 1e-5
            G.destroy()
    for seed in (5, 6):
        its = {p: it for s, p, it, _, _ in rows if s == seed}
        assert its[gas.PREC_MIXED_F16] <= its[gas.PREC_FP32] + 2
        assert its[gas.PREC_MIXED_BF16] <= its[gas.PREC_FP32] + 3


@pytest.mark.parametrize("grid,prec", [("small", gas.PREC_FP32), ("small", gas.PREC_MIXED_F16),
                                       ("small", gas.PREC_MIXED_BF16), ("house", gas.PREC_MIXED_F16)])
def test_mfma_coarsest_product_matches_vector_path(ctx, small, house, grid, prec):
    """A batch of 40 systems on one grid takes the coarsest product on the FP64
    matrix cores (k_vc_coarse_mfma); each system solved alone takes the vector
    kernel.  The two sum the product in different orders, so they agree to
    FP64 rounding of the coarsest solve, not bit for bit: the same iteration
    counts within one, solutions within 1e-6 of each other, and the FP64
    oracle within the 1e-5 bar.  The small house's coarsest level has 455
    unknowns, the 80x50x40 house's 712: a partial last row tile and column
    chunk in both."""
    g, H = small if grid == "small" else house
    G = gas.Grid(ctx, g, precision=prec)
    nsys = 40
    bs = [sc.rhs_from_divergence(g, sc.divergence_field(g, 900 + s), DT) for s in range(nsys)]
    p = gas.solve_params(tol=1e-6, max_iters=200, zero_x=True)
    B = gas.Batch(ctx, [G] * nsys)
    dbs = [ctx.to_device(b) for b in bs]
    dxs = [ctx.array(g.n) for _ in bs]
    ctx.profile(True, classes=["vc_coarse", "vc_coarse_mfma"])
    ctx.profile_reset()
    stats = B.solve(dbs, dxs, p)
    prof = ctx.profile_read()
    assert prof["vc_coarse_mfma"][1] > 0 and prof["vc_coarse"][1] == 0   # the whole batch on the matrix cores
    assert all(s["status"] == gas.CONVERGED for s in stats)
    xb = [d.download() for d in dxs]
    B.destroy()
    for k in range(0, nsys, 8):
        db, dx = ctx.to_device(bs[k]), ctx.array(g.n)
        ctx.profile_reset()
        st = G.solve(db, dx, p)
        prof = ctx.profile_read()
        assert prof["vc_coarse_mfma"][1] == 0 and prof["vc_coarse"][1] > 0   # alone: the vector kernel
        xa = dx.download()
        assert abs(st["iters"] - stats[k]["iters"]) <= 1, (k, st["iters"], stats[k]["iters"])
        assert rel(xb[k], xa) < 1e-6, k
    ctx.profile(False)
    for k in (0, nsys - 1):
        x64, _, _ = H.pcg64(bs[k])
        assert rel(xb[k], x64) < 1e-5, k
    G.destroy()
