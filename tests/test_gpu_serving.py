"""GPU tests of the serving loop: scenarios admitted while others run share
each step's batched pressure solve, and each scenario's results are the same
bit for bit alone and in the mix; admission waits for HBM to free up."""
import numpy as np
import pytest

from photon_amd import gas
from photon_amd import serving as sv

pytestmark = pytest.mark.gpu

FIELDS = ("p", "u", "v", "w", "T", "Yo", "S")


def snapshot(job):
    return {f: job.solver.get(f) for f in FIELDS}


def run_alone(ctx, spec):
    srv = sv.Server(ctx, hbm_budget_bytes=64e9)
    srv.submit(spec)
    while srv.active or srv.queue:
        srv.step()
    job = srv.finished[0]
    out = snapshot(job), list(job.iters)
    srv.close()
    return out


def test_results_same_alone_and_in_mix(ctx):
    a = sv.JobSpec(0, 11, 6000, 6, 200.0)
    b = sv.JobSpec(1, 12, 20000, 9, 450.0)
    c = sv.JobSpec(2, 13, 3600, 4, 50.0)
    ref_a, it_a = run_alone(ctx, a)
    ref_c, it_c = run_alone(ctx, c)
    srv = sv.Server(ctx, hbm_budget_bytes=64e9)
    srv.submit(b)
    srv.step()
    srv.submit(a)       # arrives while b is on the GPU
    srv.step()
    srv.step()
    srv.submit(c)
    got = {}
    while srv.active or srv.queue:
        srv.step()
        for j in srv.finished:
            if j.spec.job_id not in got:
                got[j.spec.job_id] = (snapshot(j), list(j.iters))
    assert set(got) == {0, 1, 2}
    for ref, it, jid in ((ref_a, it_a, 0), (ref_c, it_c, 2)):
        snap, its = got[jid]
        assert its == it, (jid, its, it)
        for f in FIELDS:
            assert np.array_equal(snap[f], ref[f]), (jid, f)
    # the scenarios did something: a burner heats its room and drives flow
    snap_a = got[0][0]
    assert snap_a["T"].max() > 300.0 and np.abs(snap_a["u"]).max() > 0.0
    srv.close()


def test_admission_waits_for_memory(ctx):
    big = sv.JobSpec(0, 21, 40000, 3, 300.0)
    small = sv.JobSpec(1, 22, 5000, 2, 100.0)
    budget = sv.footprint_bytes(sv.grid_for(big)) + 1000   # room for the big job only
    srv = sv.Server(ctx, hbm_budget_bytes=budget)
    srv.submit(big)
    srv.submit(small)
    srv.step()
    assert [j.spec.job_id for j in srv.active] == [0] and len(srv.queue) == 1
    while srv.active and srv.active[0].spec.job_id == 0:
        srv.step()
    assert srv.finished[0].spec.job_id == 0
    while srv.active or srv.queue:
        srv.step()
    assert [j.spec.job_id for j in srv.finished] == [0, 1]
    assert srv.used == 0
    srv.close()


def test_poisson_stream_serves_every_job(ctx):
    specs = sv.poisson_stream(3, rate_per_s=20.0, window_s=0.5, steps=3, cells_hi=60000)
    out = sv.serve(ctx, specs, hbm_budget_bytes=8e9, max_active=8)
    assert out["jobs"] == len(specs) and out["job_steps"] == 3 * len(specs)
    assert out["step_latency_ms_p99"] >= out["step_latency_ms_p50"] > 0


def test_planned_grid_is_the_same_grid(ctx):
    """A grid installed from a plan built on a worker thread is the grid
    phg_grid_create builds: same hierarchy and zones, same solve bit for bit."""
    import concurrent.futures

    from photon_amd import scenarios as sc
    g = sc.house(5, 40, 30, 20)
    b = sc.rhs_from_divergence(g, sc.burner_field(g, 5), 0.01)
    p = gas.solve_params(tol=1e-6, max_iters=200, zero_x=True)
    with concurrent.futures.ThreadPoolExecutor(1) as pool:
        plan = pool.submit(gas.GridPlan, g).result()
    out = []
    for pl in (None, plan):
        G = gas.Grid(ctx, g, precision=gas.PREC_MIXED_F16, plan=pl)
        db, dx = ctx.to_device(b), ctx.array(g.n)
        st = G.solve(db, dx, p)
        zone, zo = G.zones()
        out.append((st["iters"], dx.download(), G.info(), zone, zo))
        G.destroy()
    assert not plan.h   # consumed
    assert out[0][0] == out[1][0] and out[0][2] == out[1][2]
    for k in (1, 3, 4):
        assert np.array_equal(out[0][k], out[1][k])
    with pytest.raises(gas.PhgError):
        gas.Grid(ctx, g, plan=plan)   # a plan serves one grid
