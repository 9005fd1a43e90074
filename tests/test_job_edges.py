"""Edge cases of whole jobs on the HIP path against the oracle (needs a GPU):
Toeplitz passes over plane chunks inside a job, R near a consistent volume,
degenerate inputs (a constant volume in TV, all-zero projections: the
reference's 0/0 gives NaN, denoise.c:127-142 and the CGNR step size), the
--random padding (src/reconstruction.c:434-469) and the finite-kappa
diffusion conductance."""
import numpy as np
import pytest

from tests.util import blob_phantom, rel

pytestmark = pytest.mark.gpu


def phantom_case(orc, np_, nz, seed, t=5, nblobs=12):
    rng = np.random.default_rng(seed)
    dens = blob_phantom((nz, np_, np_), nblobs, rng, sig=(1.5, 3.0), slab=np_ // 2)
    return dens, orc.project_2d1d(dens, np_, -60, 60, t)


@pytest.mark.parametrize("keep", ["1", "0"])
def test_plane_chunked_job(hip, orc, monkeypatch, keep):
    """A job whose Toeplitz intermediate is split in plane chunks (a 4 KB row
    budget gives two chunks of 64 planes at np = 64, nz = 128) runs the CG
    start's R quadratic-form partials and the kept directions per chunk: its
    volume and R trace equal the single-chunk job bit for bit over HIO and
    ER passes, and R is within 1e-5 of the FP64 plain path."""
    np_, nz = 64, 128
    _, proj = phantom_case(orc, np_, nz, 21, t=3)
    monkeypatch.setenv("NUDIM_KEEP_P", keep)

    def run(prec):
        p = hip.make_params(np_, nz, np_, -60, 60, 3, padangle=1.0, weight=1, nhio=2, ner=2, precision=prec)
        with hip.Reconstruction(p) as job:
            job.load_projections(proj)
            job.initial()
            r = job.iterate(4)
            return r, job.volume()

    r_one, v_one = run(hip.FP32)
    monkeypatch.setenv("NUDIM_TOEP_ROW_BYTES", "4096")
    r_two, v_two = run(hip.FP32)
    monkeypatch.delenv("NUDIM_TOEP_ROW_BYTES")
    assert np.array_equal(r_one, r_two)
    assert np.array_equal(v_one, v_two)
    r64, v64 = run(hip.FP64_PLAIN)
    assert np.max(np.abs(r_two - r64)) < 1e-5
    assert np.quantile(np.abs(v_two - v64), 0.999) < 1e-4 * np.abs(v64).max()


def test_small_r_at_consistent_volume(hip, orc, monkeypatch):
    """From the phantom itself (consistent with its projections up to the
    simulator's own discretisation) R of the next pass is 4e-5 in the
    oracle, far below where passes run (0.3-0.7).  The FP32 R from the known-node
    quadratic form (a cancellation of |m|^2-sized terms) stays finite and
    non-negative, within 2e-3 of the oracle's; the FP32 R through the
    forward NFFT (NUDIM_R_FORM=nfft) within 1e-4."""
    np_ = nz = 32
    dens, proj = phantom_case(orc, np_, nz, 31)
    kw = dict(padangle=1.0, weight=1, nhio=0, ner=1)
    op = orc.params(np_, nz, np_, -60, 60, 5, **kw)
    slices, kind = orc.assemble(op, proj)
    _, r_ref = orc.iterate(op, dens, slices, kind, 0)
    out = {}
    for name, prec, form in (("fp64", hip.FP64_PLAIN, None), ("fp32", hip.FP32, None), ("fp32nfft", hip.FP32, "nfft")):
        if form:
            monkeypatch.setenv("NUDIM_R_FORM", form)
        with hip.Reconstruction(hip.make_params(np_, nz, np_, -60, 60, 5, precision=prec, **kw)) as job:
            job.load_projections(proj)
            job.set_volume(dens)
            out[name] = job.iterate(1)[0]
        monkeypatch.delenv("NUDIM_R_FORM", raising=False)
    print("R at the phantom: oracle %.3e fp64 %.3e fp32 %.3e fp32-nfft %.3e" %
          (r_ref, out["fp64"], out["fp32"], out["fp32nfft"]), flush=True)
    assert r_ref < 1e-4
    assert abs(out["fp64"] - r_ref) < 1e-10
    assert np.isfinite(out["fp32"]) and out["fp32"] >= 0 and abs(out["fp32"] - r_ref) < 2e-3
    assert abs(out["fp32nfft"] - r_ref) < 1e-4


@pytest.mark.parametrize("precision", [0, 1])
def test_tv_constant_volume_is_nan(hip, orc, precision):
    """A constant positive volume: dA = 0 and sum grad^2 = 0, so the update
    lam sqrt(dA) grad / sqrt(sum grad^2) is 0/0: the reference's volume is
    NaN wherever TV updates (the whole volume, or the height band)."""
    u = np.full((8, 12, 12), 0.5)
    ref = orc.tv(u, niter=2)
    got = hip.tv(u, niter=2, precision=precision)
    assert np.array_equal(np.isnan(got), np.isnan(ref)) and np.isnan(ref).all()
    ref = orc.tv_height(u, 4, 6, niter=2)
    got = hip.tv_height(u, 4, 6, niter=2, precision=precision)
    assert np.isnan(ref).sum() == 8 * 5 * 12
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    assert np.array_equal(got[~np.isnan(ref)], ref[~np.isnan(ref)])


@pytest.mark.parametrize("precision", [0, 1])
def test_zero_projections_give_nan(hip, orc, precision):
    """All-zero projections: every CGNR step size is 0/0, so the reference's
    initial volume, every pass's volume and R are NaN."""
    np_ = nz = 16
    kw = dict(padangle=1.0, weight=1, nhio=1, ner=1)
    proj = np.zeros((25, nz, np_))
    img0, img, rt = orc.reconstruct(orc.params(np_, nz, np_, -60, 60, 5, **kw), proj)
    assert np.isnan(img0).all() and np.isnan(img).all() and np.isnan(rt).all()
    with hip.Reconstruction(hip.make_params(np_, nz, np_, -60, 60, 5, precision=precision, **kw)) as job:
        job.load_projections(proj)
        job.initial()
        v0 = job.volume()
        r = job.iterate(2)
        v = job.volume()
    assert np.isnan(v0).all() and np.isnan(v).all() and np.isnan(r).all()


@pytest.mark.parametrize("precision,tol", [(0, 1e-11), (1, 2e-5)])
def test_random_padding_job(hip, orc, precision, tol):
    """--random: the unknown and held-out samples hold gsl_rng_uniform * 10
    from MT19937 with GSL's default seed (4357), which only the initial INFFT
    sees (the projector replaces just the known samples).  The oracle fills
    them from numpy's MT19937; the initial volume and two passes match."""
    np_ = nz = 32
    _, proj = phantom_case(orc, np_, nz, 41)
    kw = dict(padangle=1.0, weight=1, nhio=1, ner=1)
    op = orc.params(np_, nz, np_, -60, 60, 5, **kw)
    slices, kind = orc.assemble(op, proj)
    filled = orc.random_fill(slices, kind)
    img0 = orc.initial(op, filled)
    img1, r1 = orc.iterate(op, img0, slices, kind, 0)
    img2, r2 = orc.iterate(op, img1, slices, kind, 1)
    zero0 = orc.initial(op, slices)
    assert rel(zero0, img0) > 1e-2  # the padding changes the start
    with hip.Reconstruction(hip.make_params(np_, nz, np_, -60, 60, 5, precision=precision, random=True,
                                            **kw)) as job:
        job.load_projections(proj)
        job.initial()
        v0 = job.volume()
        r = job.iterate(2)
        v2 = job.volume()
    assert rel(v0, img0) < tol
    if precision == 0:
        assert rel(v2, img2) < 1e-10 and np.max(np.abs(r - [r1, r2])) < 1e-12
    else:
        # the first pass is HIO: a voxel within rounding of zero may flip
        assert np.quantile(np.abs(v2 - img2), 0.999) < 1e-4 * np.abs(img2).max()
        assert np.max(np.abs(r - [r1, r2])) < 1e-4


@pytest.mark.parametrize("precision", [0, 1])
def test_diffusion_conductance_closed_form(hip, orc, precision):
    """The finite-kappa conductance on the HIP path: a spike of height
    h = kappa moves its two earlier-visited neighbours by lam h / 2 in one
    sweep, as in the oracle (the exponential variant would give lam h / e)."""
    lam, kappa, h = 0.25, 0.5, 0.5
    u = np.zeros((6, 6, 6))
    u[2, 3, 3] = h
    got = hip.diffusion2d(u, lam=lam, kappa=kappa, niter=1, precision=precision)
    ref = orc.diffusion2d(u, lam=lam, kappa=kappa, niter=1)
    assert got[2, 2, 3] == lam * h / 2 and got[2, 3, 2] == lam * h / 2
    assert np.max(np.abs(got - ref)) < (1e-16 if precision == 0 else 1e-7)
    # several sweeps at the reference's kappa on a smooth field
    rng = np.random.default_rng(3)
    v = blob_phantom((12, 12, 12), 4, rng, sig=(1.5, 3.0)) * 0.01
    tol = 1e-12 if precision == 0 else 1e-5
    assert rel(hip.diffusion2d(v, niter=5, precision=precision), orc.diffusion2d(v, niter=5)) < tol


@pytest.mark.parametrize("precision", [0, 1])
def test_diffusion_plane_in_lds_equals_hyperplane_launches(hip, monkeypatch, precision):
    """The diffusion with each z plane held in LDS (one launch for all sweeps)
    equals the one-launch-per-hyperplane form bit for bit, on the parity
    entry point and inside a DIFFUSION job (np = nz = 64, HIO and ER)."""
    rng = np.random.default_rng(17)
    v = blob_phantom((24, 40, 24), 6, rng, sig=(1.5, 3.0)) * 0.01 + 1e-3 * rng.standard_normal((24, 40, 24))
    a = hip.diffusion2d(v, niter=12, precision=precision)
    monkeypatch.setenv("NUDIM_DIFFUSION_LDS", "0")
    b = hip.diffusion2d(v, niter=12, precision=precision)
    monkeypatch.delenv("NUDIM_DIFFUSION_LDS")
    assert np.array_equal(a, b) and not np.array_equal(a, v)
    np_ = nz = 64
    dens = blob_phantom((nz, np_, np_), 10, rng, sig=(1.5, 3.0), slab=np_ // 2)
    proj = hip.project_2d1d(dens, np_, -60, 60, 5)
    out = []
    for env in (None, "0"):
        if env:
            monkeypatch.setenv("NUDIM_DIFFUSION_LDS", env)
        p = hip.make_params(np_, nz, np_, -60, 60, 5, padangle=1.0, weight=1, nhio=1, ner=1, denoise=True,
                            method=hip.DIFFUSION, precision=precision)
        with hip.Reconstruction(p) as job:
            job.load_projections(proj)
            job.initial()
            out.append((job.iterate(2), job.volume()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("prec,np_,nz,denoise", [(0, 64, 64, 0), (1, 64, 64, 1), (0, 96, 48, 1), (1, 128, 128, 1),
                                                 (0, 128, 128, 0), (1, 96, 96, 0)])
def test_estimate_bounds_device_bytes(hip, orc, prec, np_, nz, denoise):
    """The admission estimate (what the job server budgets with) is at least
    what a job holds on the device after set-up and passes: the patch-form
    spreading tables and partial sums (plane counts that are multiples of
    64), the Toeplitz operators and the per-cell maps otherwise."""
    rng = np.random.default_rng(np_ + nz)
    proj = rng.standard_normal((41, nz, np_))
    p = hip.make_params(np_, nz, np_, -60, 60, 3, padangle=1.0, weight=1, nhio=1, ner=1, denoise=bool(denoise),
                        method=hip.TV, height=np_ // 2, center=np_ // 2, precision=prec)
    est = hip.estimate_bytes(p)
    with hip.Reconstruction(p) as job:
        job.load_projections(proj)
        job.initial()
        job.iterate(2)
        held = job.device_bytes()
    assert 0 < held <= est
