"""Continuous-batching server: every job's result is the same alone and in the
mix, bit for bit in every storage (the batched two-step kernel and the single
lattice's one-step kernel compute the same step), with admission by HBM
footprint forcing jobs to wait and the active set to change between chunks;
and every served result against the CPU oracle: FP64 bit for bit, FP32
within the stated tolerance of the FP64 oracle."""

import numpy as np
import pytest

from tests import oracle_jobs as J

pytestmark = pytest.mark.gpu

# FP32 results (probe rows, u_x profiles) against the FP64 oracle over these
# budgets (at most 500 steps): max absolute difference.
F32_TOL = 2e-6
_ORACLE = {}


def oracle_result(spec):
    key = (spec.kind, spec.lx, spec.ly, spec.steps, spec.plate, spec.reynolds, spec.u_in, spec.omega, spec.delta_p)
    if key not in _ORACLE:
        _ORACLE[key] = J.spec_result(spec, J.run_spec(spec))
    return _ORACLE[key]


def check_against_oracle(rec):
    want = oracle_result(rec.spec)
    for key, val in want.items():
        got = rec.result[key]
        assert got.shape == val.shape, (rec.spec, key)
        if rec.spec.storage == "f64":
            assert np.array_equal(got, val), (rec.spec, key)
        elif rec.spec.storage == "f32":
            assert np.abs(got - val).max() <= F32_TOL, (rec.spec, key, float(np.abs(got - val).max()))


def jobs(storage):
    from lattice_boltzmann_parallel_solver_amd.serving import JobSpec

    return [JobSpec("von_karman", 96, 40, 333, storage, plate=8, reynolds=100.0),
            JobSpec("couette", 48, 48, 150, storage, u_in=0.05, omega=1.0),
            JobSpec("poiseuille", 64, 32, 500, storage, omega=1.5),
            JobSpec("von_karman", 120, 52, 77, storage, plate=12, reynolds=60.0),
            JobSpec("couette", 33, 70, 260, storage, u_in=0.05, omega=0.8),
            JobSpec("von_karman", 80, 36, 410, storage, plate=10, reynolds=150.0)]


def alone(spec):
    lat = spec.start()
    lat.run(spec.relaxation, spec.steps)
    res = spec.result(lat)
    lat.close()
    return res


@pytest.mark.parametrize("policy", ["fifo", "srpt"])
@pytest.mark.parametrize("storage,atol", [("f64", 0.0), ("f32", 0.0), ("f16", 0.0)])
def test_results_same_alone_and_in_the_mix(storage, atol, policy):
    from lattice_boltzmann_parallel_solver_amd.serving import Server, summarize

    specs = jobs(storage)
    arrivals = [(0.0, s) for s in specs]
    budget = sum(sorted(s.footprint() for s in specs)[-3:])  # about three jobs at a time
    server = Server(hbm_budget=budget, chunk_steps=64, policy=policy)
    recs = server.run(arrivals, realtime=False)
    assert all(np.isfinite(r.latency) for r in recs)
    assert any(r.admitted > 0 for r in recs)  # some jobs had to wait for memory
    for rec in recs:
        ref = alone(rec.spec)
        assert rec.result["finite"] and ref["finite"]
        key = "samples" if rec.spec.kind == "von_karman" else "ux_mid"
        a, b = rec.result[key], ref[key]
        assert a.shape == b.shape
        if atol == 0.0:
            assert np.array_equal(a, b), rec.spec
        else:
            np.testing.assert_allclose(a, b, rtol=0, atol=atol)
        check_against_oracle(rec)
    s = summarize(recs, server.makespan)
    assert s["p99_s"] >= s["p50_s"] > 0 and s["unstable"] == 0


def test_realtime_arrivals_are_served():
    from lattice_boltzmann_parallel_solver_amd.serving import Server, c5_jobs, summarize

    arr = c5_jobs(12, seed=1, rate=200.0, vk_size=(256, 64), square=(32, 96), budget=(200, 2000))
    server = Server(chunk_steps=250)
    recs = server.run(arr)
    s = summarize(recs, server.makespan)
    assert s["jobs"] == 12 and s["unstable"] == 0
    assert all(r.finished >= r.arrival for r in recs)


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_srpt_policy_same_results_least_work_first(storage):
    """Shortest-remaining-work scheduling: a chunk advances only the jobs with
    the least remaining node steps (here about two lattices' worth), so the
    run set changes from chunk to chunk; results stay those of each job run
    alone, and the first job to finish is one of the two with the least work
    (the first run set)."""
    from lattice_boltzmann_parallel_solver_amd.serving import Server

    specs = jobs(storage)
    arrivals = [(0.0, s) for s in specs]
    server = Server(chunk_steps=64, policy="srpt", fill_nodes=6000)
    recs = server.run(arrivals, realtime=False)
    for rec in recs:
        ref = alone(rec.spec)
        key = "samples" if rec.spec.kind == "von_karman" else "ux_mid"
        assert np.array_equal(rec.result[key], ref[key]), rec.spec
        check_against_oracle(rec)
    first = min(recs, key=lambda r: r.finished)
    least = sorted(recs, key=lambda r: r.spec.nodes * r.spec.steps)[:2]
    assert first.job_id in [r.job_id for r in least]


def test_footprint_bounds_the_device_memory_of_a_job():
    """Admission by HBM footprint: the estimate is at least what a job's
    context allocates on the device, and at most 32 MiB more."""
    import torch

    from lattice_boltzmann_parallel_solver_amd.serving import JobSpec

    warm = JobSpec("couette", 64, 64, 10, "f32").start()
    warm.close()
    for spec in (JobSpec("von_karman", 1200, 514, 20000, "f32", plate=120),
                 JobSpec("couette", 300, 300, 5000, "f64"),
                 JobSpec("poiseuille", 257, 129, 5000, "f16", omega=1.5)):
        torch.cuda.synchronize()
        free0, _ = torch.cuda.mem_get_info()
        lat = spec.start()
        lat.synchronize()
        free1, _ = torch.cuda.mem_get_info()
        used = free0 - free1
        lat.close()
        assert used <= spec.footprint() <= used + (32 << 20), (spec.kind, used, spec.footprint())


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_unstable_job_fails_early_and_frees_its_slot(storage):
    """A channel at omega ~ 1.99 (Re 955 on a plate of 8: the oracle's state
    grows without bound within about 300 steps) served beside stable jobs, with
    room for only two jobs at a time: the server marks it failed after the
    chunk in which its sums leave the bounds, long before its 5000-step
    budget, frees its slot for the waiting job, and every other job's result
    still equals the oracle's."""
    from lattice_boltzmann_parallel_solver_amd.serving import JobSpec, Server, summarize

    bad = JobSpec("von_karman", 96, 40, 5000, storage, plate=8, reynolds=955.0)
    assert 1.985 < bad.relaxation < 1.995
    good = [JobSpec("von_karman", 120, 52, 700, storage, plate=12, reynolds=60.0),
            JobSpec("couette", 48, 48, 600, storage, u_in=0.05, omega=1.0),
            JobSpec("poiseuille", 64, 32, 500, storage, omega=1.5)]
    specs = [bad] + good
    budget = bad.footprint() + max(s.footprint() for s in good)
    print(f"case: serving with an injected unstable job, {storage}", flush=True)
    server = Server(hbm_budget=budget, chunk_steps=100, policy="fifo")
    recs = server.run([(0.0, s) for s in specs], realtime=False)
    r_bad = recs[0]
    assert r_bad.result["failed"] and not r_bad.result["finite"]
    assert 0 < r_bad.result["failed_after_steps"] <= 1000, r_bad.result
    assert r_bad.finished < max(r.finished for r in recs[1:])
    for rec in recs[1:]:
        assert not rec.result.get("failed"), rec.result
        check_against_oracle(rec)
    s = summarize(recs, server.makespan)
    assert s["failed"] == 1 and s["completed"] == 3


def test_jobs_arriving_during_a_chunk_are_admitted_before_it_ends():
    """A long chunk of one large job, and jobs that arrive while it runs: they
    are admitted (allocated and initialised) during the chunk, and all results
    still equal the oracle's."""
    from lattice_boltzmann_parallel_solver_amd.serving import JobSpec, Server

    # about 3 s of GPU work in one chunk (8192x2048 FP64, one step per launch)
    long_job = JobSpec("couette", 8192, 2048, 8000, "f64", u_in=0.05, omega=1.0)
    late = [JobSpec("couette", 40, 30, 300, "f64", u_in=0.05, omega=0.9),
            JobSpec("von_karman", 96, 40, 333, "f64", plate=8, reynolds=100.0)]
    print("case: serving admission while a chunk runs", flush=True)
    server = Server(chunk_steps=8000, policy="fifo")
    recs = server.run([(0.0, long_job), (1.0, late[0]), (1.1, late[1])], realtime=True)
    assert server.admitted_during_chunks >= 2
    assert all(r.admitted < recs[0].finished for r in recs[1:])
    for rec in recs[1:]:
        check_against_oracle(rec)
