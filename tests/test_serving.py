"""Job server (nudim_server_*): continuous batching under an HBM budget.

Each job runs on its own stream with its own buffers and fixed-order
reductions, so its result must be bit-identical alone and in a mix."""
import numpy as np
import pytest

from tests.util import blob_phantom

pytestmark = pytest.mark.gpu


def job_inputs(hip, n, nz, seed, interval=5):
    rng = np.random.default_rng(seed)
    dens = blob_phantom((nz, n, n), 8, rng, sig=(1.5, 3.0), slab=n // 2)
    return hip.project_2d1d(dens, n, -60, 60, interval)


def job_params(hip, n, nz, prec, denoise, method, interval=5):
    return hip.make_params(n, nz, n, -60, 60, interval, padangle=1.0, weight=True, nhio=2, ner=2,
                           denoise=denoise, method=method, height=n // 2, center=n // 2, precision=prec)


# (n = np, nz, precision, denoise, method); DIFFUSION needs nz == np
CASES = [(32, 32, 1, True, 1), (48, 48, 0, True, 0), (64, 64, 1, False, 1), (32, 48, 0, True, 1),
         (64, 64, 0, False, 1), (48, 32, 1, True, 1)]


def alone(hip, p, proj):
    with hip.Reconstruction(p) as job:
        job.load_projections(proj)
        job.initial()
        r = job.iterate(p.nhio + p.ner)
        return job.volume(), r


def test_alone_and_in_mix_identical(hip):
    jobs = []
    for k, (n, nz, prec, den, meth) in enumerate(CASES):
        p = job_params(hip, n, nz, prec, den, meth)
        proj = job_inputs(hip, n, nz, 100 + k)
        jobs.append((p, proj, alone(hip, p, proj)))
    with hip.Server(max_active=4) as srv:
        ids = {srv.submit(p, proj): k for k, (p, proj, _) in enumerate(jobs)}
        seen = set()
        for _ in jobs:
            rep, vol, rf = srv.next(timeout_ms=120000)
            assert rep["status"] == 0, rep["error"]
            k = ids[rep["job_id"]]
            seen.add(k)
            v_alone, r_alone = jobs[k][2]
            assert np.array_equal(vol, v_alone), "job %d differs in the mix" % k
            assert np.array_equal(rf, r_alone)
            assert rep["t_submit"] <= rep["t_start"] <= rep["t_end"]
        assert seen == set(range(len(jobs)))


def test_budget_serialises_admission(hip):
    p = job_params(hip, 32, 32, 1, False, 1)
    proj = job_inputs(hip, 32, 32, 7)
    need = hip.estimate_bytes(p)
    # room for one job at a time: every job starts after the previous ended
    with hip.Server(max_active=4, hbm_budget=int(need * 1.5)) as srv:
        for _ in range(3):
            srv.submit(p, proj)
        reps = [srv.next(timeout_ms=120000)[0] for _ in range(3)]
    reps.sort(key=lambda r: r["t_start"])
    for a, b in zip(reps, reps[1:]):
        assert b["t_start"] >= a["t_end"] - 1e-6
    # a job larger than the whole budget is refused at submission
    with hip.Server(max_active=1, hbm_budget=int(need // 2)) as srv:
        with pytest.raises(hip.NudimError):
            srv.submit(p, proj)


def test_next_times_out_when_idle(hip):
    with hip.Server(max_active=1) as srv:
        assert srv.next(timeout_ms=10) is None
        assert srv.stats()["pending"] == 0


@pytest.mark.parametrize("policy", ["shortest", "fifo"])
def test_admission_policy_order(hip, policy):
    """One slot: while the first job runs, a larger and then a smaller job
    arrive; least-work-first runs the smaller one next, FIFO the larger."""
    pol = hip.SERVE_SHORTEST if policy == "shortest" else hip.SERVE_FIFO
    big = job_params(hip, 64, 64, 1, False, 1)
    small = job_params(hip, 32, 32, 1, False, 1)
    pb, ps = job_inputs(hip, 64, 64, 1), job_inputs(hip, 32, 32, 2)
    with hip.Server(max_active=1, policy=pol) as srv:
        first = srv.submit(big, pb)
        later_big = srv.submit(big, pb)
        later_small = srv.submit(small, ps)
        reps = {}
        for _ in range(3):
            rep = srv.next(timeout_ms=120000)[0]
            reps[rep["job_id"]] = rep
    assert reps[first]["t_start"] <= min(reps[later_big]["t_start"], reps[later_small]["t_start"])
    if policy == "shortest":
        assert reps[later_small]["t_start"] < reps[later_big]["t_start"]
    else:
        assert reps[later_big]["t_start"] < reps[later_small]["t_start"]


def test_cu_masked_jobs_identical(hip):
    """A job restricted to part of the device (cu_mask), and jobs placed by
    the server's small-job partition, compute exactly what an unrestricted
    job computes: every sum has a fixed order whichever CUs run the blocks."""
    p = job_params(hip, 48, 48, 1, True, 1)
    proj = job_inputs(hip, 48, 48, 11)
    v0, r0 = alone(hip, p, proj)
    pm = job_params(hip, 48, 48, 1, True, 1)
    for i in range(2):
        pm.cu_mask[i] = 0xFFFFFFFF  # the first 64 CUs
    v1, r1 = alone(hip, pm, proj)
    assert np.array_equal(v0, v1) and np.array_equal(r0, r1)
    big = job_params(hip, 64, 64, 1, False, 1)
    pb = job_inputs(hip, 64, 64, 12)
    vb, rb = alone(hip, big, pb)
    small_work = 48 * 48 * 48 * (p.nhio + p.ner)
    with hip.Server(max_active=3, small_cus=32, small_work=small_work) as srv:
        ids = {srv.submit(p, proj): "small", srv.submit(big, pb): "big", srv.submit(p, proj): "small"}
        for _ in range(3):
            rep, vol, rf = srv.next(timeout_ms=120000)
            assert rep["status"] == 0, rep["error"]
            ref = (v0, r0) if ids[rep["job_id"]] == "small" else (vb, rb)
            assert np.array_equal(vol, ref[0]) and np.array_equal(rf, ref[1])
    with pytest.raises(hip.NudimError):
        hip.Server(max_active=1, small_cus=100000)


def test_graph_jobs_in_a_mix_run_directly(hip):
    """A job submitted with use_graph runs with direct launches in the
    server (capturing one job's pass would break the other slots'
    synchronous copies): it succeeds beside other jobs and equals its alone
    run without graphs."""
    p = job_params(hip, 48, 48, 1, True, 1)
    proj = job_inputs(hip, 48, 48, 21)
    v0, r0 = alone(hip, p, proj)
    pg = job_params(hip, 48, 48, 1, True, 1)
    pg.use_graph = 1
    with hip.Server(max_active=4) as srv:
        ids = {srv.submit(pg, proj): "graph"}
        for k in range(3):
            ids[srv.submit(job_params(hip, 32, 32, 1, True, 1), job_inputs(hip, 32, 32, 30 + k))] = "other"
        for _ in range(4):
            rep, vol, rf = srv.next(timeout_ms=120000)
            assert rep["status"] == 0, rep["error"]
            if ids[rep["job_id"]] == "graph":
                assert np.array_equal(vol, v0) and np.array_equal(rf, r0)


def blocker(hip):
    """a job that keeps a one-slot server busy while the test queues others"""
    p = hip.make_params(128, 128, 128, -60, 60, 3, padangle=1.0, weight=True, nhio=0, ner=10)
    return p, job_inputs(hip, 128, 128, 99, interval=3)


@pytest.mark.parametrize("nz,njobs,max_stack,expect", [(64, 5, 0, [5]), (32, 4, 0, [3, 1]), (32, 5, 0, [5]),
                                                        (64, 5, 1, [1] * 5), (64, 5, 2, [2, 2, 1])])
def test_same_parameter_jobs_run_as_one_batch(hip, nz, njobs, max_stack, expect):
    """Jobs with the same parameters queued behind a running one run
    together as one batch context (reported in 'stacked'): all of them at
    nz = 64; at nz = 32 an odd count (four queued jobs run as three and then
    one, since an even count at nz = 32 would group the reductions' planes
    differently); max_stack bounds the batch.  Every job's volume and R
    trace equal its own job's bit for bit."""
    p = job_params(hip, 64, nz, 1, True, 1)
    projs = [job_inputs(hip, 64, nz, 300 + k) for k in range(njobs)]
    ref = [alone(hip, p, pr) for pr in projs]
    bp, bproj = blocker(hip)
    with hip.Server(max_active=1, max_stack=max_stack) as srv:
        srv.submit(bp, bproj)
        ids = {srv.submit(p, pr): k for k, pr in enumerate(projs)}
        got = {}
        for _ in range(njobs + 1):
            rep, vol, rf = srv.next(timeout_ms=120000)
            assert rep["status"] == 0, rep["error"]
            if rep["job_id"] in ids:
                got[ids[rep["job_id"]]] = (rep, vol, rf)
    groups = {}
    for k, (rep, vol, rf) in got.items():
        assert np.array_equal(vol, ref[k][0]), "job %d differs in its batch" % k
        assert np.array_equal(rf, ref[k][1])
        groups.setdefault(rep["t_start"], []).append(rep["stacked"])
    sizes = sorted((len(v) for v in groups.values()), reverse=True)
    assert sizes == expect, groups
    for v in groups.values():
        assert all(s == len(v) for s in v)
