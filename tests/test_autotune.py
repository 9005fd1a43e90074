"""FP32 NFFT sizing by measurement (nudim_autotune): for a job whose passes
run the NFFT the chosen (n, W) is a fixed-length FFT size no larger than the
default, its NFFT is within the stated 2e-6 of the oracle's, and a job
created without explicit sizes runs exactly as one given the tuned sizes.
A job whose passes run no NFFT takes the default sizes whatever was
measured before in the process."""
import numpy as np
import pytest

from tests.util import blob_phantom, rel

pytestmark = pytest.mark.gpu

FIXED_N = (128, 256, 384, 512, 768, 1024, 1536)


def params(hip, np_, nz, **kw):
    return hip.make_params(np_, nz, np_, -60, 60, 5, padangle=1.0, weight=True, nhio=1, ner=2,
                           precision=hip.FP32, **kw)


@pytest.mark.parametrize("np_", [64, 192])
def test_autotune_choice_and_accuracy(hip, orc, np_):
    p = params(hip, np_, 64, cg_form=hip.CG_NFFT)
    n, W = hip.autotune(p)
    assert n in FIXED_N and np_ + 16 <= n <= 2 * np_ and W in (4, 6, 8, 10, 12)
    assert hip.autotune(p) == (n, W)  # cached
    # its trafo and adjoint against the oracle's NFFT on this node set
    pmin, pmax = orc.pad_range(-60, 5)
    x = orc.polar_nodes(pmin, pmax, 1.0, np_)
    rng = np.random.default_rng(3)
    P = 64
    fhat = rng.standard_normal((P, np_, np_)) + 1j * rng.standard_normal((P, np_, np_))
    f = rng.standard_normal((P, x.shape[0])) + 1j * rng.standard_normal((P, x.shape[0]))
    ref_f = hip.nfft_trafo(x, fhat, n=2 * 2 ** int(np.ceil(np.log2(np_))), W=14)
    ref_fh = hip.nfft_adjoint(x, f, np_, n=2 * 2 ** int(np.ceil(np.log2(np_))), W=14)
    # the tuner holds its own random planes to 2e-6; other data within 1.5x
    assert rel(hip.nfft_trafo(x, fhat, n=n, W=W, precision=hip.FP32), ref_f) < 3e-6
    assert rel(hip.nfft_adjoint(x, f, np_, n=n, W=W, precision=hip.FP32), ref_fh) < 3e-6


def test_autotuned_job_equals_explicit_sizes(hip, orc, monkeypatch):
    np_ = nz = 64
    rng = np.random.default_rng(5)
    dens = blob_phantom((nz, np_, np_), 10, rng, sig=(1.5, 3.0), slab=np_ // 2)
    proj = orc.project_2d1d(dens, np_, -60, 60, 5)
    n, W = hip.autotune(params(hip, np_, nz, cg_form=hip.CG_NFFT))
    out = []
    for kw in ({}, {"nfft_n": n, "nfft_w": W}):
        with hip.Reconstruction(params(hip, np_, nz, cg_form=hip.CG_NFFT, **kw)) as job:
            job.load_projections(proj)
            job.initial()
            out.append((job.volume(), job.iterate(3), job.volume()))
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
    # switched off: the fixed default sizing
    monkeypatch.setenv("NUDIM_AUTOTUNE", "0")
    assert hip.autotune(params(hip, np_, nz, cg_form=hip.CG_NFFT)) == (128, 8)


def test_no_nfft_job_ignores_earlier_measurement(hip, orc):
    """A Toeplitz-form FP32 job (projected CG start, R from the quadratic
    form: no NFFT in its passes) gives the same result before and after the
    process measured sizes for its node set, and equal to the default sizes."""
    np_, nz = 48, 32
    rng = np.random.default_rng(11)
    dens = blob_phantom((nz, np_, np_), 8, rng, sig=(1.5, 3.0), slab=np_ // 2)
    proj = orc.project_2d1d(dens, np_, -60, 60, 5)

    def run(**kw):
        with hip.Reconstruction(params(hip, np_, nz, **kw)) as job:
            job.load_projections(proj)
            job.initial()
            r = job.iterate(2)
            return job.volume(), r

    before = run()
    assert hip.autotune(params(hip, np_, nz)) == (96, 8)  # defaults, not measured
    hip.autotune(params(hip, np_, nz, cg_form=hip.CG_NFFT))  # measures this node set
    after = run()
    explicit = run(nfft_n=96, nfft_w=8)
    for a, b, c in zip(before, after, explicit):
        assert np.array_equal(a, b) and np.array_equal(a, c)
