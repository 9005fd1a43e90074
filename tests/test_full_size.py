"""Parity at the full workloads.

128^3 (np = nz = 128, 41 tilts -60..60 step 3, weighted): the FP64 plain
path and the FP32 form against the oracle, which takes a few seconds per
pass on the host cores -- the noise-free ER schedule (20 passes) in full,
and the noisy HIO+denoise schedule (SNR 1, nhio 20 + ner 20, TV or
DIFFUSION over the height band) for its first passes against the oracle,
then in full against the FP64 plain path.

256^3 (np = 384, nz = 256, 71 tilts -70..70 step 2, SNR 2, weighted, TV
with height 128 centre 192): the FP32 form against the FP64 plain path over
the same passes from each path's own initial INFFT; at this size the oracle
would take hours.

Without TV the iteration is stable and FP32 is held to 1e-4 relative L2 of
the volume (measured 1.5e-6).  With TV (40 sweeps of t / sqrt(1e-8 + |t|^2)
per pass) the reference's own iteration amplifies rounding-size differences
to a plateau of a few 1e-2 within 4 passes: the FP64 path started from its
initial volume times (1 + 1e-7 noise) ends 2.6e-2 away from itself.  There
FP32 is held to a small multiple of that floor, measured in the test."""
import numpy as np
import pytest

import nudim_amd as nd
from nudim_amd import phantom
from tests.util import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c3():
    prm, proj = phantom.config_workload("c3")
    assert (prm["np_"], prm["nz"], prm["nhio"], prm["height"]) == (384, 256, 30, 128)
    return prm, proj


def make(prm, precision, **over):
    q = dict(prm)
    q.update(over)
    return nd.make_params(q["nx"], q["nz"], q["np_"], q["gam_min"], q["gam_max"], q["interval"],
                          padangle=q["padangle"], weight=q["weight"], nhio=q["nhio"], ner=q["ner"],
                          denoise=q["denoise"], method=nd.TV, height=q["height"], center=q["center"],
                          precision=precision)


def run(prm, proj, precision, npass, perturb=0.0, **over):
    """(initial volume, R trace, volume after npass passes) of one job; with
    perturb, also the R trace and volume from the initial volume times
    (1 + perturb * noise) and a bit-for-bit repeat check."""
    job = nd.Reconstruction(make(prm, precision, **over))
    try:
        job.load_projections(proj)
        job.initial()
        v0 = job.volume()
        r = job.iterate(npass)
        v = job.volume()
        extra = None
        if perturb:
            rng = np.random.default_rng(1)
            job.set_volume(v0 * (1 + perturb * rng.standard_normal(v0.shape)))
            rp = job.iterate(npass)
            extra = (rp, job.volume())
            job.set_volume(v0)
            rr = job.iterate(npass)
            assert np.array_equal(rr, r) and np.array_equal(job.volume(), v), "passes are not repeatable"
        return v0, r, v, extra
    finally:
        job.close()


# ---- 128^3 against the oracle ------------------------------------------------

def oracle_params(orc, prm, **over):
    q = dict(prm)
    q.update(over)
    return orc.params(q["nx"], q["nz"], q["np_"], q["gam_min"], q["gam_max"], q["interval"],
                      padangle=q["padangle"], weight=int(q["weight"]), denoise=int(q.get("denoise", False)),
                      method=q.get("method", 1), height=q.get("height"), center=q.get("center"),
                      nhio=q["nhio"], ner=q["ner"])


def make128(prm, precision, **over):
    q = dict(prm)
    q.update(over)
    return nd.make_params(q["nx"], q["nz"], q["np_"], q["gam_min"], q["gam_max"], q["interval"],
                          padangle=q["padangle"], weight=q["weight"], nhio=q["nhio"], ner=q["ner"],
                          denoise=q.get("denoise", False), method=q.get("method", nd.TV), height=q.get("height"),
                          center=q.get("center"), precision=precision)


def oracle_run(orc, op, img, slices, kind, n, nn0=0):
    rs = []
    for nn in range(nn0, nn0 + n):
        img, r = orc.iterate(op, img, slices, kind, nn)
        rs.append(r)
    return img, np.array(rs)


def hip_run(params, proj, npass, start=None):
    """(initial volume, R trace, final volume) of one job, optionally from
    the given start volume instead of the job's own initial INFFT"""
    with nd.Reconstruction(params) as job:
        job.load_projections(proj)
        job.initial()
        v0 = job.volume()
        if start is not None:
            job.set_volume(start)
        r = job.iterate(npass)
        return v0, r, job.volume()


def flip_tolerant(a, b):
    """the 99.9th percentile of |a - b| over max |b|: HIO (reconstruction.c:682)
    maps a voxel within rounding of zero to either img or img_last - 0.9 img"""
    return np.quantile(np.abs(a - b), 0.999) / np.abs(b).max()


def test_c1_er_against_oracle(orc):
    """The 128^3 reference configuration: initial INFFT and 20 ER passes
    (noise-free, no denoiser) on the FP64 plain path and the FP32 form,
    each against the oracle over the whole R trace."""
    prm, proj = phantom.config_workload("c1")
    assert (prm["np_"], prm["nz"], prm["nhio"], prm["ner"]) == (128, 128, 0, 20)
    op = oracle_params(orc, prm)
    slices, kind = orc.assemble(op, proj)
    img0 = orc.initial(op, slices)
    img, r_ref = oracle_run(orc, op, img0, slices, kind, 20)
    v0_64, r64, v64 = hip_run(make128(prm, nd.FP64_PLAIN), proj, 20)
    v0_32, r32, v32 = hip_run(make128(prm, nd.FP32), proj, 20)
    print("c1 fp64: initial %.2e final %.2e max|dR| %.2e; fp32: initial %.2e final %.2e max|dR| %.2e" %
          (rel(v0_64, img0), rel(v64, img), np.max(np.abs(r64 - r_ref)), rel(v0_32, img0), rel(v32, img),
           np.max(np.abs(r32 - r_ref))), flush=True)
    print("c1 R drift per pass (fp32 - oracle):", np.array2string(r32 - r_ref, precision=2), flush=True)
    assert rel(v0_64, img0) < 1e-10
    assert rel(v64, img) < 1e-10
    assert np.max(np.abs(r64 - r_ref)) < 1e-12
    assert rel(v0_32, img0) < 1e-4
    assert rel(v32, img) < 1e-4
    assert np.max(np.abs(r32 - r_ref)) < 1e-5


@pytest.fixture(scope="module")
def c2(orc):
    prm, proj = phantom.config_workload("c2")
    assert (prm["nhio"], prm["ner"], prm["height"], prm["center"]) == (20, 20, 64, 64)
    op = oracle_params(orc, prm)
    slices, kind = orc.assemble(op, proj)
    return prm, proj, slices, kind, orc.initial(op, slices)


@pytest.mark.parametrize("method", [nd.TV, nd.DIFFUSION])
def test_c2_first_passes_against_oracle(orc, c2, method):
    """Noisy 128^3 (SNR 1) with HIO and the denoiser: the initial volume and
    the first three HIO passes against the oracle.  TV amplifies any
    rounding-size difference (reconstruction.c:630-632; denoise.c:97-142),
    so after the first pass the bar is the oracle's own distance from itself
    when its volume is perturbed by the path's rounding size (1e-13 FP64 at
    the start; 1e-6 FP32 before each pass, as FP32 rounds in every pass)."""
    prm, proj, slices, kind, img0 = c2
    op = oracle_params(orc, prm, method=method)
    img3, r_ref = oracle_run(orc, op, img0, slices, kind, 3)
    rng = np.random.default_rng(2)
    f64, _ = oracle_run(orc, op, img0 * (1 + 1e-13 * rng.standard_normal(img0.shape)), slices, kind, 3)
    # FP32 rounds in every pass: perturb before each of the three passes
    f32 = img0
    for nn in range(3):
        f32, _ = orc.iterate(op, f32 * (1 + 1e-6 * rng.standard_normal(img0.shape)), slices, kind, nn)
    floor64, floor32 = rel(f64, img3), rel(f32, img3)
    out = {}
    for prec in (nd.FP64_PLAIN, nd.FP32):
        v0, r, v3 = hip_run(make128(prm, prec, method=method), proj, 3)
        out[prec] = (v0, r, v3)
    v0, r, v3 = out[nd.FP64_PLAIN]
    w0, q, w3 = out[nd.FP32]
    print("c2 method %d: oracle floors %.2e / %.2e; fp64 initial %.2e pass3 %.2e dR %s; fp32 initial %.2e pass3 %.2e "
          "dR %s" % (method, floor64, floor32, rel(v0, img0), rel(v3, img3), np.abs(r - r_ref), rel(w0, img0),
                     rel(w3, img3), np.abs(q - r_ref)), flush=True)
    assert rel(v0, img0) < 1e-10
    assert abs(r[0] - r_ref[0]) < 1e-12
    assert rel(v3, img3) < 10 * floor64 + 1e-9
    assert rel(w0, img0) < 1e-4
    assert abs(q[0] - r_ref[0]) < 1e-5
    if method == nd.TV:
        assert rel(w3, img3) < 4 * floor32 + 1e-4
    else:
        # DIFFUSION: FP32 rounds in each of the 80 in-place sweeps of every
        # pass (differences near kappa = 0.005 carry ~1e-5 relative error)
        # and the HIO passes amplify that with a heavy tail: the oracle's
        # volume-level perturbations (floor32, ~2e-6) do not model it.
        # Measured here 5e-5 .. 1.5e-4 for FP32 rounding variants of the
        # conductance; the same 5e-4 bar as the 64^3 DIFFUSION runs.
        assert rel(w3, img3) < 5e-4


def hip_run_perturbed(params, proj, npass, eps, seed):
    """R trace and final volume of a job whose volume is multiplied by
    (1 + eps * noise) before every pass: the path's response to rounding-size
    errors made in every pass, as an FP32 run makes them"""
    rng = np.random.default_rng(seed)
    rs = []
    with nd.Reconstruction(params) as job:
        job.load_projections(proj)
        job.initial()
        for nn in range(npass):
            v = job.volume()
            job.set_volume(v * (1 + eps * rng.standard_normal(v.shape)), nn)
            rs.append(job.iterate(1)[0])
        return np.array(rs), job.volume()


@pytest.mark.parametrize("method", [nd.TV, nd.DIFFUSION])
def test_c2_full_schedule_fp32_within_fp64_floor(c2, method):
    """The whole noisy 128^3 schedule (20 HIO + 20 ER passes with the
    denoiser): FP32 against FP64 plain from each path's own initial volume,
    held to twice the FP64 path's own distance from itself when its volume
    is perturbed by 1e-6 (relative, FP32 rounding size) before every pass;
    the per-pass R drift is reported.

    Both denoisers make the schedule sensitive: measured on the FP64 path,
    per-pass perturbations of 1e-7 or 1e-6 end 4e-3 .. 1e-2 away from the
    unperturbed run after 40 passes with DIFFUSION (2.5e-3 .. 2.9e-3 with TV),
    while one perturbation of 1e-7 of the initial volume alone decays to
    2e-8 with DIFFUSION.  FP32 ends 9.6e-3 (DIFFUSION) and 2.7e-3 (TV) away."""
    prm, proj = c2[0], c2[1]
    p64 = make128(prm, nd.FP64_PLAIN, method=method)
    _, r64, v64 = hip_run(p64, proj, 40)
    rp, vp = hip_run_perturbed(p64, proj, 40, 1e-6, 3)
    _, r32, v32 = hip_run(make128(prm, nd.FP32, method=method), proj, 40)
    floor, rfloor = rel(vp, v64), np.max(np.abs(rp - r64))
    drift = r32 - r64
    print("c2 full method %d: fp64 per-pass floor %.2e (R %.2e); fp32 vs fp64 %.2e, max R drift %.2e" %
          (method, floor, rfloor, rel(v32, v64), np.max(np.abs(drift))), flush=True)
    print("c2 full method %d R drift per pass:" % method, np.array2string(drift, precision=2), flush=True)
    assert np.isfinite(v32).all() and np.isfinite(r32).all()
    assert 1e-4 < floor < 0.1  # the response, not a broken perturbation
    assert rel(v32, v64) < 2 * floor
    assert np.max(np.abs(drift)) < 10 * rfloor


# ---- 256^3 -------------------------------------------------------------------

def test_c3_er_fp32_matches_fp64_plain(c3):
    prm, proj = c3
    v0_64, r64, v64, _ = run(prm, proj, nd.FP64_PLAIN, 3, nhio=0, denoise=False)
    v0_32, r32, v32, _ = run(prm, proj, nd.FP32, 3, nhio=0, denoise=False)
    assert v64.shape == (256, 384, 384)
    assert rel(v0_32, v0_64) < 1e-5
    assert rel(v32, v64) < 1e-4
    assert np.max(np.abs(r32 - r64)) < 1e-6


def test_c3_hio_tv_fp32_within_fp64_floor(c3):
    prm, proj = c3
    _, r64, v64, (rp, vp) = run(prm, proj, nd.FP64_PLAIN, 4, perturb=1e-7)
    _, r32, v32, _ = run(prm, proj, nd.FP32, 4)
    floor = rel(vp, v64)
    rfloor = np.max(np.abs(rp - r64))
    assert 1e-3 < floor < 0.1  # the plateau, not a broken perturbation
    assert rel(v32, v64) < 2 * floor
    assert np.max(np.abs(r32 - r64)) < 10 * rfloor
