"""Batches of volumes of one geometry (nudim_batch_create; needs a GPU).

A batch runs every kernel of a pass over all its volumes at once: the
Fourier side as nvol nz planes, the real side as nvol consecutive volumes.
Every operator works plane by plane, so where the plane grouping of the
per-plane reductions is the same as in a single job (nz a multiple of 64,
or nvol odd) each volume's initial volume, R trace and final volume equal
its own job's bit for bit; otherwise they agree to rounding."""
import numpy as np
import pytest

from tests.util import blob_phantom, rel

pytestmark = pytest.mark.gpu


def cases(hip, np_, nz, nx, K, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(K):
        dens = blob_phantom((nz, np_, np_), 10, rng, sig=(1.5, 3.0), slab=np_ // 2)
        out.append(hip.project_2d1d(dens, nx, -60, 60, 5))
    return out


def alone(hip, p, proj, npass):
    with hip.Reconstruction(p) as job:
        job.load_projections(proj)
        job.initial()
        v0 = job.volume()
        r = job.iterate(npass)
        return v0, r, job.volume()


def batched(hip, p, projs, npass):
    K = len(projs)
    with hip.Batch(p, K) as b:
        for k, pr in enumerate(projs):
            b.load_projections(k, pr)
        b.initial()
        v0 = [b.volume(k) for k in range(K)]
        r = b.iterate(npass)
        assert r.shape == (npass, K)
        return v0, r, [b.volume(k) for k in range(K)]


CONFIGS = {
    # TV over a height band (the rows either side copied per volume)
    "tv_height": dict(np_=64, nz=64, nx=64, kw=dict(denoise=True, height=32, center=32, nhio=2, ner=1)),
    # DIFFUSION with --random padding (each volume draws the same stream)
    "diffusion_random": dict(np_=32, nz=32, nx=32, kw=dict(denoise=True, nhio=1, ner=1, random=True)),
    # no denoiser, support narrower than the grid
    "er_support": dict(np_=64, nz=32, nx=48, kw=dict(nhio=1, ner=2)),
}


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", list(CONFIGS))
def test_batch_volumes_equal_their_own_jobs(hip, name, precision):
    c = CONFIGS[name]
    kw = dict(c["kw"])
    if name.startswith("diffusion"):
        kw["method"] = hip.DIFFUSION
    p = hip.make_params(c["nx"], c["nz"], c["np_"], -60, 60, 5, padangle=1.0, weight=1, precision=precision, **kw)
    npass = kw["nhio"] + kw["ner"]
    projs = cases(hip, c["np_"], c["nz"], c["nx"], 3, 7)
    print("batch case", name, "precision", precision, flush=True)
    bv0, br, bv = batched(hip, p, projs, npass)
    for k, pr in enumerate(projs):
        v0, r, v = alone(hip, p, pr, npass)
        assert np.array_equal(bv0[k], v0), (name, k)
        assert np.array_equal(br[:, k], r), (name, k, br[:, k], r)
        assert np.array_equal(bv[k], v), (name, k)
    # the volumes are distinct (each volume reads its own samples)
    assert not np.array_equal(bv[0], bv[1])


def test_batch_graph_replay_equals_direct(hip):
    """use_graph on a batch: the captured passes (R of every volume copied
    out of the replay) equal the direct launches."""
    projs = cases(hip, 64, 64, 64, 3, 11)
    out = []
    for g in (False, True):
        p = hip.make_params(64, 64, 64, -60, 60, 5, padangle=1.0, weight=1, precision=hip.FP32, denoise=True,
                            height=32, center=32, nhio=2, ner=2, use_graph=g)
        out.append(batched(hip, p, projs, 4))
    assert np.array_equal(out[0][1], out[1][1])
    for k in range(3):
        assert np.array_equal(out[0][2][k], out[1][2][k])


def test_batch_other_plane_grouping_agrees_to_rounding(hip):
    """nvol = 2 at nz = 32: the per-plane reductions group 64 planes, not 32,
    and the interpolation takes its patch form, so sums are added in another
    order; on ER passes without the denoiser (a contracting map: with HIO
    and TV such differences grow ~3e3 times a pass, as any rounding
    difference does there) FP64 volumes and R agree with the single jobs to
    rounding."""
    projs = cases(hip, 32, 32, 32, 2, 13)
    p = hip.make_params(32, 32, 32, -60, 60, 5, padangle=1.0, weight=1, nhio=0, ner=3)
    bv0, br, bv = batched(hip, p, projs, 3)
    for k, pr in enumerate(projs):
        v0, r, v = alone(hip, p, pr, 3)
        assert rel(bv0[k], v0) < 1e-12
        assert np.max(np.abs(br[:, k] - r)) < 1e-12
        assert rel(bv[k], v) < 1e-10


def test_batch_state_errors(hip):
    p = hip.make_params(32, 32, 32, -60, 60, 5, padangle=1.0, weight=1, nhio=0, ner=1, precision=hip.FP32)
    projs = cases(hip, 32, 32, 32, 2, 17)
    L = hip.lib()
    with hip.Batch(p, 2) as b:
        with pytest.raises(hip.NudimError, match="nudim_batch_load_projections"):
            hip._check(L.nudim_load_projections(b._h, hip._ptr(np.ascontiguousarray(projs[0]))))
        b.load_projections(0, projs[0])
        with pytest.raises(hip.NudimError, match="every volume"):
            b.initial()
        with pytest.raises(hip.NudimError, match="out of range"):
            b.load_projections(2, projs[1])
        b.load_projections(1, projs[1])
        b.initial()
        b.iterate(1)
        with pytest.raises(hip.NudimError, match="out of range"):
            b.volume(-1)
        with pytest.raises(hip.NudimError, match="single-volume"):
            hip._check(L.nudim_set_volume(b._h, hip._ptr(np.zeros((32, 32, 32))), 0))
        with pytest.raises(hip.NudimError, match="nudim_batch_get_volume"):
            hip._check(L.nudim_get_volume(b._h, hip._ptr(np.zeros((32, 32, 32)))))
