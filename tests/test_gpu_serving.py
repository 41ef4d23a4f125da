"""GPU tests of the serving loop: scenarios admitted while others run share
each step's batched pressure solve and run the whole gas step; each
scenario's results are the same bit for bit alone and in the mix; admission
waits for HBM to free up; pause / resume / evict / cancel as the reference's
job runner (src/app/cli.cpp:77-99); a scenario whose solve fails is retired
without changing its neighbours; served scenarios match the oracle's replay
of the reference's step."""
import numpy as np
import pytest

from oracle import pystep
from photon_amd import gas
from photon_amd import serving as sv

pytestmark = pytest.mark.gpu

FIELDS = ("p", "u", "v", "w", "T", "Yo", "Yf", "S", "nu", "Pb")


def run_alone(ctx, spec, **kw):
    srv = sv.Server(ctx, hbm_budget_bytes=64e9, keep_fields=FIELDS, **kw)
    srv.submit(spec)
    while srv.active or srv.queue:
        srv.step()
    job = srv.finished[0]
    out = job.result, list(job.iters)
    srv.close()
    return out


def same(a, b):
    return all(np.array_equal(a[f].view(np.uint32), b[f].view(np.uint32)) for f in FIELDS)


def test_results_same_alone_and_in_mix(ctx):
    a = sv.JobSpec(0, 11, 6000, 6, 200.0)
    b = sv.JobSpec(1, 12, 20000, 9, 450.0)
    c = sv.JobSpec(2, 13, 3600, 4, 50.0)
    ref_a, it_a = run_alone(ctx, a)
    ref_c, it_c = run_alone(ctx, c)
    srv = sv.Server(ctx, hbm_budget_bytes=64e9, keep_fields=FIELDS)
    srv.submit(b)
    srv.step()
    srv.submit(a)       # arrives while b is on the GPU
    srv.step()
    srv.step()
    srv.submit(c)
    while srv.active or srv.queue:
        srv.step()
    got = {j.spec.job_id: (j.result, list(j.iters)) for j in srv.finished}
    assert set(got) == {0, 1, 2} and all(j.state == sv.DONE for j in srv.finished)
    for ref, it, jid in ((ref_a, it_a, 0), (ref_c, it_c, 2)):
        res, its = got[jid]
        assert its == it, (jid, its, it)
        assert same(res, ref), jid
    # the scenarios did something: a burner heats its room and drives flow
    res_a = got[0][0]
    assert res_a["T"].max() > 300.0 and np.abs(res_a["u"]).max() > 0.0 and res_a["Yf"].max() > 0
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
    assert out["job_states"] == {"done": len(specs)}
    assert out["step_latency_ms_p99"] >= out["step_latency_ms_p50"] > 0


def test_failed_job_is_retired_and_neighbours_unchanged(ctx):
    a = sv.JobSpec(0, 31, 5000, 5, 150.0)
    bad = sv.JobSpec(1, 32, 8000, 5, 250.0)
    ref_a, it_a = run_alone(ctx, a)
    srv = sv.Server(ctx, hbm_budget_bytes=64e9, keep_fields=FIELDS, prep_workers=0)
    srv.submit(a)
    srv.submit(bad)
    srv.step()      # both admitted (set-up inline), both stepped
    srv.step()
    jb = srv.jobs[1]
    T = jb.solver.get("T")
    gas_cells = np.flatnonzero(jb.grid.blocked == 0)
    T[gas_cells[len(gas_cells) // 2]] = np.nan      # a corrupted gas cell in one scenario
    jb.solver.set("T", T)
    while srv.active or srv.queue:
        srv.step()
    assert jb.state == sv.FAILED and jb.reason == "non-finite residual" and jb.steps_done == 2
    st = srv.status(1)
    assert st["state"] == "failed" and st["steps"] == 2
    ja = srv.jobs[0]
    assert ja.state == sv.DONE and ja.iters == it_a and same(ja.result, ref_a)
    assert srv.used == 0
    srv.close()


def test_pause_resume_evict_cancel(ctx):
    a = sv.JobSpec(0, 41, 6000, 6, 200.0)
    ref_a, it_a = run_alone(ctx, a)
    srv = sv.Server(ctx, hbm_budget_bytes=64e9, keep_fields=FIELDS)
    srv.submit(a)
    srv.submit(sv.JobSpec(1, 42, 4000, 50, 80.0))
    srv.step()
    srv.step()
    assert srv.control(0, "pause") == sv.PAUSED
    srv.step()
    srv.step()
    assert srv.jobs[0].steps_done == 2 and srv.status(0)["state"] == "paused"
    used = srv.used
    assert srv.control(0, "evict") == sv.EVICTED and srv.used < used and srv.jobs[0].solver is None
    srv.step()
    assert srv.control(0, "resume") == sv.RUNNING
    assert srv.control(1, "cancel") == sv.CANCELLED
    while srv.active or srv.queue:
        srv.step()
    ja = srv.jobs[0]
    assert ja.state == sv.DONE and ja.iters == it_a and same(ja.result, ref_a)   # bitwise resume
    assert srv.jobs[1].state == sv.CANCELLED and srv.used == 0
    assert srv.control(7, "cancel") == "queued"      # unknown job: nothing to do
    srv.close()


@pytest.mark.parametrize("cells,steps", [(3600, 4), (40000, 3), (200000, 2)])
def test_served_job_matches_oracle_step(ctx, cells, steps):
    """One served scenario per size class, replayed through the oracle's
    restatement of Solver::step (oracle/pystep.py; FP32 MG-PCG, exact coarsest
    solve, ||r||/||b|| <= 1e-6): face velocities within 1e-5 relative L2,
    temperature and species within 1e-6, pressure within 5e-5 of its scale
    (two solutions that each meet a 1e-6 relative residual differ by up to
    the operator's condition number times it; 1.2e-5 at 45x34x27)."""
    spec = sv.JobSpec(0, 77 + cells, cells, steps, 300.0)
    res, its = run_alone(ctx, spec)
    g = sv.grid_for(spec)
    orc = pystep.OracleGas(g)
    x0, y0, x1, y1 = sv.burner_rect(g, spec)
    orc.add_burner_cells(gas.burner_cells(g, x0, y0, x1, y1, 0.0), spec.hrr_kw * 1e3)
    for _ in range(steps):
        orc.step(0.01)

    def rel(a, b):
        return float(np.linalg.norm(a.astype(np.float64) - b) / max(np.linalg.norm(b.astype(np.float64)), 1e-30))

    print(f"{g.shape}: iters gpu {its} oracle {orc.iters}, u {rel(res['u'], orc.u):.2e}, "
          f"T {rel(res['T'], orc.f['T']):.2e}, p {np.abs(res['p'] - orc.p).max() / np.abs(orc.p).max():.2e}")
    for f, o in (("u", orc.u), ("v", orc.v), ("w", orc.w)):
        assert rel(res[f], o) < 1e-5, f
    assert rel(res["T"], orc.f["T"]) < 1e-6 and rel(res["Yf"], orc.f["Yf"]) < 1e-6
    assert np.abs(res["p"] - orc.p).max() <= 5e-5 * np.abs(orc.p).max()
