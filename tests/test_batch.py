"""The uneven batch and the serving mix at their configured shapes (needs a
GPU): volumes of the 48-instance batch (nx = nz in {64 ... 256}, np = nx
{1, 1.25, 1.5}, tilts +-{50, 60, 70} at step 2 or 3, SNR 1 or 4, ner = 30
with TV over the slab) and of the serving classes (128^3 at np 192, 256^3 at
np 384, 512^3 at np 768), each run alone and through the job server.

A job's kernels and reductions do not depend on what else runs, so every
result must be bit-identical alone and in the mix.  Without the denoiser
(ER passes, a contracting map) the FP32 form must stay within 1e-4 of the
FP64 plain path at every one of these shapes."""
import numpy as np
import pytest

from nudim_amd import phantom
from tests.util import rel

pytestmark = pytest.mark.gpu


def params(hip, prm, precision, **over):
    q = dict(prm)
    q.update(over)
    return hip.make_params(q["nx"], q["nz"], q["np_"], q["gam_min"], q["gam_max"], q["interval"],
                           padangle=q["padangle"], weight=q["weight"], nhio=q["nhio"], ner=q["ner"],
                           denoise=q["denoise"], method=hip.TV, height=q["height"], center=q["center"],
                           precision=precision)


@pytest.fixture(scope="module")
def subset():
    """eight batch instances spread over the sizes: the distinct (np, nz)
    shapes ordered by work, every other one from the smallest up, plus the
    largest"""
    inst = phantom.batch_instances(48)
    shapes = {}
    for k, (prm, _) in enumerate(inst):
        shapes.setdefault((prm["np_"], prm["nz"]), k)
    order = sorted(shapes.items(), key=lambda kv: kv[0][0] ** 2 * kv[0][1])
    pick = [k for _, k in order[::2]][:7] + [order[-1][1]]
    return [inst[k] for k in sorted(set(pick))]


def alone(hip, p, proj):
    with hip.Reconstruction(p) as job:
        job.load_projections(proj)
        job.initial()
        r = job.iterate(p.nhio + p.ner)
        return job.volume(), r


def served(hip, jobs, slots=4):
    out = {}
    with hip.Server(max_active=slots) as srv:
        ids = {srv.submit(p, proj): k for k, (p, proj) in enumerate(jobs)}
        for _ in jobs:
            rep, vol, rf = srv.next(timeout_ms=600000)
            assert rep["status"] == 0, rep["error"]
            out[ids[rep["job_id"]]] = (vol, rf)
    return [out[k] for k in range(len(jobs))]


def test_batch_subset_alone_equals_served(hip, subset):
    """The configured batch jobs (FP32, 30 ER passes with TV) in a mix of
    eight through the server equal their alone runs bit for bit."""
    jobs = [(params(hip, prm, hip.FP32), proj) for prm, proj in subset]
    shapes = [(p.np, p.nz) for p, _ in jobs]
    print("batch subset shapes (np, nz):", shapes, flush=True)
    assert len(set(shapes)) >= 6
    ref = [alone(hip, p, proj) for p, proj in jobs]
    for k, (vol, rf) in enumerate(served(hip, jobs)):
        assert np.array_equal(vol, ref[k][0]), "job %s differs in the mix" % (shapes[k],)
        assert np.array_equal(rf, ref[k][1])
        assert np.isfinite(vol).all()


def test_batch_subset_er_fp32_vs_fp64(hip, subset):
    """The same volumes, 3 ER passes without the denoiser: FP32 within 1e-4
    relative L2 of the FP64 plain path, R within 1e-5."""
    for prm, proj in subset:
        v64, r64 = alone(hip, params(hip, prm, hip.FP64_PLAIN, ner=3, denoise=False), proj)
        v32, r32 = alone(hip, params(hip, prm, hip.FP32, ner=3, denoise=False), proj)
        print("np %d nz %d: rel %.2e, max |dR| %.2e" % (prm["np_"], prm["nz"], rel(v32, v64),
                                                        np.max(np.abs(r32 - r64))), flush=True)
        assert rel(v32, v64) < 1e-4
        assert np.max(np.abs(r32 - r64)) < 1e-5


def test_serving_mix_with_512(hip):
    """One job of each serving class (128^3 np 192, 256^3 np 384 and 512^3
    np 768 with 121 tilts), four passes each (2 HIO + 2 ER with TV), served
    together: each equals its alone run bit for bit."""
    jobs = []
    for name in ("s128", "s256", "s512"):
        prm, proj = phantom.config_workload(name)
        jobs.append((params(hip, prm, hip.FP32, nhio=2, ner=2), proj))
    assert (jobs[2][0].np, jobs[2][0].nz) == (768, 512)
    ref = [alone(hip, p, proj) for p, proj in jobs]
    for k, (vol, rf) in enumerate(served(hip, jobs, slots=3)):
        assert np.array_equal(rf, ref[k][1])
        assert np.array_equal(vol, ref[k][0])
