"""HIP path vs the FP64 oracle on the same seeded inputs (needs a GPU).

Tolerances: FP64 plain path vs oracle, relative L2 <= 1e-10 (same algorithm
and window, only rounding order differs); FP32 form, relative L2 <= 1e-4 for
end-to-end volumes (the stated target) and tighter per kernel.
"""
import numpy as np
import pytest

from tests.util import blob_phantom, rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def polar32(orc):
    pmin, pmax = orc.pad_range(-60, 3)
    return orc.polar_nodes(pmin, pmax, 1.0, 32)


@pytest.mark.parametrize("length,howmany", [(8, 3), (12, 5), (30, 2), (64, 16), (96, 8), (256, 4), (40, 1),
                                            (128, 16), (384, 8), (512, 8), (768, 16), (1024, 8), (1536, 4),
                                            (2048, 4), (768, 3)])
def test_zfft(hip, orc, length, howmany):
    rng = np.random.default_rng(length)
    a = rng.standard_normal((length, howmany)) + 1j * rng.standard_normal((length, howmany))
    for sign in (-1, +1):
        ref = np.fft.fft(a, axis=0) if sign < 0 else np.fft.ifft(a, axis=0) * length
        assert rel(hip.zfft(a, sign), ref) < 1e-13
        assert rel(hip.zfft(a, sign, precision=hip.FP32), ref) < 2e-6
        assert rel(orc.fft_strided(a, sign), ref) < 1e-13


def test_nfft_trafo_adjoint_fp64(hip, orc, polar32):
    rng = np.random.default_rng(1)
    N, P = 32, 4
    fh = rng.standard_normal((P, N, N)) + 1j * rng.standard_normal((P, N, N))
    plan = orc.NFFT(polar32, N)
    f = hip.nfft_trafo(polar32, fh)
    for p in range(P):
        assert rel(f[p], plan.trafo(fh[p])) < 1e-12
        assert rel(f[p], orc.ndft_trafo(polar32, fh[p])) < 1e-12
    g = rng.standard_normal((P, polar32.shape[0])) + 1j * rng.standard_normal((P, polar32.shape[0]))
    a = hip.nfft_adjoint(polar32, g, N)
    for p in range(P):
        assert rel(a[p], plan.adjoint(g[p])) < 1e-12
        assert rel(a[p], orc.ndft_adjoint(polar32, g[p], N)) < 1e-12


@pytest.mark.parametrize("n,W,tol", [(64, 8, 2e-6), (48, 8, 5e-6), (64, 6, 2e-5)])
def test_nfft_fp32(hip, orc, polar32, n, W, tol):
    rng = np.random.default_rng(2)
    N, P = 32, 8
    fh = rng.standard_normal((P, N, N)) + 1j * rng.standard_normal((P, N, N))
    f = hip.nfft_trafo(polar32, fh, n=n, W=W, precision=hip.FP32)
    g = rng.standard_normal((P, polar32.shape[0])) + 1j * rng.standard_normal((P, polar32.shape[0]))
    a = hip.nfft_adjoint(polar32, g, N, n=n, W=W, precision=hip.FP32)
    for p in (0, P - 1):
        assert rel(f[p], orc.ndft_trafo(polar32, fh[p])) < tol
        assert rel(a[p], orc.ndft_adjoint(polar32, g[p], N)) < tol


@pytest.mark.parametrize("fp64,n,W,tol", [(True, 128, 14, 1e-12), (False, 128, 8, 2e-6), (False, 96, 6, 1e-4),
                                          (True, 120, 14, 1e-12)])
def test_nfft_patch_forms(hip, orc, fp64, n, W, tol):
    """64 planes on a grid whose edge is a multiple of 16 take the patch-form
    kernels (8 x 4-cell patches in registers: the gather writes each cell
    once, the interpolation adds per-block-patch partial sums per node); 63
    planes take the per-cell gather and the per-node interpolation.  The two
    forms agree to rounding and both match the NDFT (the dense centre
    patches hold hundreds of nodes, several record chunks; n = 120 takes the
    patch gather and the per-node interpolation)."""
    N = 64
    pmin, pmax = orc.pad_range(-60, 3)
    x = orc.polar_nodes(pmin, pmax, 1.0, N)
    M = x.shape[0]
    rng = np.random.default_rng(7)
    g = rng.standard_normal((64, M)) + 1j * rng.standard_normal((64, M))
    fh = rng.standard_normal((64, N, N)) + 1j * rng.standard_normal((64, N, N))
    prec = hip.FP64_PLAIN if fp64 else hip.FP32
    a64 = hip.nfft_adjoint(x, g, N, n=n, W=W, precision=prec)
    a63 = hip.nfft_adjoint(x, g[:63], N, n=n, W=W, precision=prec)
    assert rel(a64[:63], a63) < (1e-14 if fp64 else 1e-6)
    f64 = hip.nfft_trafo(x, fh, n=n, W=W, precision=prec)
    f63 = hip.nfft_trafo(x, fh[:63], n=n, W=W, precision=prec)
    assert rel(f64[:63], f63) < (1e-14 if fp64 else 1e-6)
    # tol: FP64 rounding, FP32 rounding, window error of sigma 1.5 and W 6 (5.3e-5 in FP64)
    for p in (0, 37, 63):
        assert rel(a64[p], orc.ndft_adjoint(x, g[p], N)) < tol
        assert rel(f64[p], orc.ndft_trafo(x, fh[p])) < tol


@pytest.mark.parametrize("fp64", [True, False])
def test_nfft_patch_gather_crowded_nodes(hip, orc, fp64):
    """Nodes crowded into one corner of the frequency square (one angular
    sector of the grid): the gather keeps its global heavy-first order
    instead of dealing sectors to the XCDs, and equals the per-cell form
    and the NDFT; uniformly spread nodes take the sector order and agree
    the same way."""
    N = 64
    rng = np.random.default_rng(23)
    for x in (0.05 + 0.3 * rng.random((3000, 2)), rng.random((3000, 2)) - 0.5):
        M = x.shape[0]
        g = rng.standard_normal((64, M)) + 1j * rng.standard_normal((64, M))
        prec = hip.FP64_PLAIN if fp64 else hip.FP32
        a64 = hip.nfft_adjoint(x, g, N, precision=prec)
        a63 = hip.nfft_adjoint(x, g[:63], N, precision=prec)
        assert rel(a64[:63], a63) < (1e-14 if fp64 else 1e-6)
        for p in (0, 63):
            assert rel(a64[p], orc.ndft_adjoint(x, g[p], N)) < (1e-10 if fp64 else 2e-5)


@pytest.mark.parametrize("N,n", [(96, 384), (160, 512), (256, 768), (384, 1024), (600, 1536)])
def test_nfft_fp64_lane_pair_lengths(hip, orc, N, n):
    """FP64 grids of 24x16, 32x16, 32x24, 32x32 and 48x32 points (n > 2N:
    the general line form on lane pairs, k_fft2p): trafo and adjoint of two
    planes at 400 random nodes against the NDFT."""
    rng = np.random.default_rng(N)
    x = rng.random((400, 2)) - 0.5
    P = 2
    fh = rng.standard_normal((P, N, N)) + 1j * rng.standard_normal((P, N, N))
    g = rng.standard_normal((P, 400)) + 1j * rng.standard_normal((P, 400))
    f = hip.nfft_trafo(x, fh, n=n, W=14, precision=hip.FP64_PLAIN)
    a = hip.nfft_adjoint(x, g, N, n=n, W=14, precision=hip.FP64_PLAIN)
    for p in range(P):
        assert rel(f[p], orc.ndft_trafo(x, fh[p])) < 1e-11
        assert rel(a[p], orc.ndft_adjoint(x, g[p], N)) < 1e-11


@pytest.mark.parametrize("N", [64, 192])
def test_nfft_centred_half_lengths(hip, orc, N):
    """n = 2N on the fixed-length FFT kernels (16 x 8, 24 x 16): the row and
    column passes of trafo read, and those of the adjoint write, only the N
    modes centred on zero (half-size DFTs).  Reference: the oracle's NFFT at
    its own sizes (n = 2 nextpow2(N), 14 taps)."""
    pmin, pmax = orc.pad_range(-60, 3)
    x = orc.polar_nodes(pmin, pmax, 2.0, N)
    M, P = x.shape[0], 4
    rng = np.random.default_rng(N + 1)
    fh = rng.standard_normal((P, N, N)) + 1j * rng.standard_normal((P, N, N))
    g = rng.standard_normal((P, M)) + 1j * rng.standard_normal((P, M))
    plan = orc.NFFT(x, N)
    for prec, W, tol in ((hip.FP64_PLAIN, 14, 1e-11), (hip.FP32, 8, 3e-6)):
        f = hip.nfft_trafo(x, fh, n=2 * N, W=W, precision=prec)
        a = hip.nfft_adjoint(x, g, N, n=2 * N, W=W, precision=prec)
        for p in (0, P - 1):
            assert rel(f[p], plan.trafo(fh[p])) < tol
            assert rel(a[p], plan.adjoint(g[p])) < tol


@pytest.mark.parametrize("weighted", [True, False])
def test_cgnr(hip, orc, polar32, weighted):
    rng = np.random.default_rng(3)
    N, P, M = 32, 3, polar32.shape[0]
    w = orc.weights(N, M // N, 1.0) if weighted else None
    y = rng.standard_normal((P, M)) + 1j * rng.standard_normal((P, M))
    f0 = rng.standard_normal((P, N, N))
    plan = orc.NFFT(polar32, N)
    ref = np.stack([plan.cgnr(y[p], f0[p], w, 10) for p in range(P)])
    # unweighted normal equations are worse conditioned: rounding grows more
    tol = 1e-10 if weighted else 2e-9
    got = hip.cgnr(polar32, y, f0, w)
    assert rel(got, ref) < tol
    got_t = hip.cgnr(polar32, y, f0, w, form=hip.CG_TOEPLITZ)
    assert rel(got_t, ref) < 10 * tol
    got32 = hip.cgnr(polar32, y, f0, w, n=64, W=8, precision=hip.FP32)
    # Without density compensation the normal matrix of the polar grid is
    # badly scaled (node density ~ 1/r) and FP32 recurrences lose ~3 digits;
    # every configuration runs with --weight.
    assert rel(got32, ref) < (1e-4 if weighted else 5e-2)


@pytest.mark.parametrize("N", [64, 128])
def test_cgnr_register_fft_sizes(hip, orc, N):
    """Toeplitz column pass and NFFT grids on the fixed-length FFT kernels
    (L = 2N = 128 / 256, reference grid n = 128 / 256)."""
    rng = np.random.default_rng(N)
    pmin, pmax = orc.pad_range(-60, 3)
    x = orc.polar_nodes(pmin, pmax, 2.0, N)
    M, P = x.shape[0], 8
    w = orc.weights(N, M // N, 2.0)
    y = rng.standard_normal((P, M)) + 1j * rng.standard_normal((P, M))
    f0 = rng.standard_normal((P, N, N))
    plan = orc.NFFT(x, N)
    ref = np.stack([plan.cgnr(y[p], f0[p], w, 10) for p in (0, P - 1)])
    got = hip.cgnr(x, y, f0, w)[[0, P - 1]]
    assert rel(got, ref) < 1e-10
    got_t = hip.cgnr(x, y, f0, w, form=hip.CG_TOEPLITZ)[[0, P - 1]]
    assert rel(got_t, ref) < 1e-9
    got32 = hip.cgnr(x, y, f0, w, n=2 * N, W=8, precision=hip.FP32)[[0, P - 1]]
    assert rel(got32, ref) < 1e-4


@pytest.mark.parametrize("N", [80, 96, 120, 144, 160, 192, 240, 256, 288, 320, 384, 512, 768])
def test_cgnr_toeplitz_lengths(hip, orc, N):
    """The fused Toeplitz passes at the register-FFT lengths of the full
    workloads, L = 2N = 384 (24 x 16), 512 (32 x 16), 768 (32 x 24, the
    256^3 workload at np = 384), 1024 (32 x 32) and 1536 (48 x 32, np = 768),
    and of the uneven batch's grids: 160 (20 x 8), 192 (16 x 12), 240
    (20 x 12), 288 (24 x 12), 320 (20 x 16), 480 (24 x 20), 576 (24 x 24) and
    640 (32 x 20), with the FP32 NFFT grid FFT at n = 2N on the same lengths,
    against the FP64 NFFT CGNR at the reference's sizes (n = 2 nextpow2(N),
    W = 14), whose parity with the oracle is pinned at N = 32-128 above and
    on one plane here for N <= 256: FP64 Toeplitz to 1e-9, FP32 to 1e-4."""
    rng = np.random.default_rng(N + 1)
    pmin, pmax = orc.pad_range(-70, 2)
    x = orc.polar_nodes(pmin, pmax, 1.0, N)
    M, P = x.shape[0], 16
    w = orc.weights(N, M // N, 1.0)
    y = rng.standard_normal((P, M)) + 1j * rng.standard_normal((P, M))
    f0 = rng.standard_normal((P, N, N))
    ref = hip.cgnr(x, y, f0, w)
    if N <= 256:
        plan = orc.NFFT(x, N)
        assert rel(ref[P - 1], plan.cgnr(y[P - 1], f0[P - 1], w, 10)) < 1e-10
    got_t = hip.cgnr(x, y, f0, w, form=hip.CG_TOEPLITZ)
    assert rel(got_t, ref) < 1e-9
    got32 = hip.cgnr(x, y, f0, w, n=2 * N, W=8, precision=hip.FP32)
    assert rel(got32, ref) < 1e-4
    for p in (0, 7, P - 1):
        assert rel(got32[p], ref[p]) < 1e-4


def test_toeplitz_lane_pairs_job(hip, monkeypatch):
    """A whole job at np = 768 (n = 1536: the 48 x 32 passes, the projected
    CG start's two-input column pass and row inverse with the added vector):
    the passes with two lanes per 48-point line equal the single-lane passes
    to FP32 rounding and the FP64 plain path to the FP32 bar (ER without TV,
    random projections: the parity does not depend on the data's content)."""
    np_, nz, nx = 768, 16, 768
    rng = np.random.default_rng(5)
    nmeas = 41
    proj = rng.standard_normal((nmeas, nz, nx))

    def run(prec):
        p = hip.make_params(nx, nz, np_, -60, 60, 3, padangle=1.0, weight=1, nhio=0, ner=2, precision=prec)
        with hip.Reconstruction(p) as job:
            job.load_projections(proj)
            job.initial()
            r = job.iterate(2)
            return r, job.volume()

    r_pairs, v_pairs = run(hip.FP32)
    monkeypatch.setenv("NUDIM_TOEP_PAIRS", "0")
    r_one, v_one = run(hip.FP32)
    monkeypatch.delenv("NUDIM_TOEP_PAIRS")
    r64, v64 = run(hip.FP64_PLAIN)
    assert rel(v_pairs, v_one) < 1e-5
    assert rel(v_pairs, v64) < 1e-4
    assert np.max(np.abs(r_pairs - r64)) < 1e-5


@pytest.mark.parametrize("nz", [64, 48])
def test_kept_directions_job(hip, monkeypatch, nz):
    """The projected-start solve with its ten search directions kept apart
    (f + sum_k alpha_k p_k applied once, in the z transform's loads; nz = 64
    on the register z transform, nz = 48 on the generic one with separate
    updates) equals the solve that updates f in every row pass to FP32
    rounding, and the FP64 plain path to the FP32 bar: over ER passes
    without the denoiser in relative L2; over HIO and ER passes by the
    flip-tolerant bar, since a voxel near zero may take the other HIO branch
    under the FP32 form's rounding (reconstruction.c:680-695) and the
    Fourier projection then spreads the difference: at most 1 % of the
    voxels beyond 1e-4 of the volume's maximum (0.5 % at nz = 48, 0.005 % at
    nz = 64 measured) and relative L2 within 1e-3 (the FP64 path itself
    moves linearly: 1e-9 on the projections gives 1e-9 on the volume, no
    flips).  The admission estimate counts the ten buffers."""
    np_, nx = 128, 128
    rng = np.random.default_rng(11)
    proj = rng.standard_normal((41, nz, nx))

    def run(prec, nhio):
        p = hip.make_params(nx, nz, np_, -60, 60, 3, padangle=1.0, weight=1, nhio=nhio, ner=4 - nhio,
                            precision=prec)
        with hip.Reconstruction(p) as job:
            job.load_projections(proj)
            job.initial()
            r = job.iterate(4)
            return r, job.volume(), hip.estimate_bytes(p)

    for nhio in (0, 2):
        r_keep, v_keep, b_keep = run(hip.FP32, nhio)
        monkeypatch.setenv("NUDIM_KEEP_P", "0")
        r_one, v_one, b_one = run(hip.FP32, nhio)
        monkeypatch.delenv("NUDIM_KEEP_P")
        r64, v64, _ = run(hip.FP64_PLAIN, nhio)
        assert b_keep - b_one >= 10 * np_ * np_ * nz * 8
        assert rel(v_keep, v_one) < 1e-5
        if nhio == 0:
            assert rel(v_keep, v64) < 1e-4
            assert np.max(np.abs(r_keep - r64)) < 1e-5
        else:
            assert (np.abs(v_keep - v64) > 1e-4 * np.abs(v64).max()).mean() < 0.01
            assert rel(v_keep, v64) < 1e-3
            assert np.max(np.abs(r_keep - r64)) < 1e-4


def test_toeplitz_plane_chunks(hip, orc, monkeypatch):
    """The fused Toeplitz passes over plane chunks (long rows of the 2N x 2N
    intermediate, forced here by a small row budget) equal the single-chunk
    passes bit for bit, and the oracle's CGNR to the FP32 bar."""
    N, P = 64, 128
    pmin, pmax = orc.pad_range(-60, 3)
    x = orc.polar_nodes(pmin, pmax, 2.0, N)
    M = x.shape[0]
    w = orc.weights(N, M // N, 2.0)
    rng = np.random.default_rng(17)
    y = rng.standard_normal((P, M)) + 1j * rng.standard_normal((P, M))
    f0 = rng.standard_normal((P, N, N))
    one = hip.cgnr(x, y, f0, w, n=2 * N, W=8, precision=hip.FP32, form=hip.CG_TOEPLITZ)
    monkeypatch.setenv("NUDIM_TOEP_ROW_BYTES", "4096")  # 2N * 64 planes * 8 B > 4 KB: two chunks of 64
    two = hip.cgnr(x, y, f0, w, n=2 * N, W=8, precision=hip.FP32, form=hip.CG_TOEPLITZ)
    assert np.array_equal(one, two)
    plan = orc.NFFT(x, N)
    for p in (0, 77, P - 1):
        assert rel(two[p], plan.cgnr(y[p], f0[p], w, 10)) < 1e-4


@pytest.mark.parametrize("n,nz,np_,t", [(32, 16, 48, 5), (24, 32, 32, 3), (16, 8, 16, 30)])
def test_project_2d1d(hip, orc, n, nz, np_, t):
    """GPU 2D1D simulator against the oracle's (same algorithm, FP64), and
    against the reference's own brute-force DFT mode (which uses PI =
    3.1415926 in its exponent: agreement to ~1e-7) on the smallest case."""
    rng = np.random.default_rng(n + nz)
    dens = blob_phantom((nz, n, n), 6, rng, sig=(1.5, 3.0), slab=n // 2)
    got = hip.project_2d1d(dens, np_, -60, 60, t)
    ref = orc.project_2d1d(dens, np_, -60, 60, t)
    assert got.shape == ref.shape
    assert rel(got, ref) < 1e-12
    if n == 16:
        assert rel(got, orc.project_dft(dens, np_, -60, 60, t)) < 1e-6
        # 0 degrees: the line integral along y, centred in the row
        zero = got[2]
        assert np.max(np.abs(zero - dens.sum(axis=1))) < 1e-10


def test_projection_ft(hip, orc):
    rng = np.random.default_rng(4)
    proj = rng.standard_normal((5, 16, 24))
    ref = orc.projection_ft(proj, 32)
    assert rel(hip.projection_ft(proj, 32), ref) < 1e-13


def smooth_volume(shape, seed):
    rng = np.random.default_rng(seed)
    v = blob_phantom(shape, 6, rng, sig=(1.5, 3.0), slab=None)
    return v - 0.15 * v.max() + 0.01 * rng.standard_normal(shape)


def perturbed(a, eps, seed):
    rng = np.random.default_rng(seed)
    return a * (1 + eps * rng.standard_normal(a.shape))


def test_tv_fp64(hip, orc):
    u = smooth_volume((10, 16, 16), 5)
    assert rel(hip.tv(u), orc.tv(u)) < 1e-11
    assert rel(hip.tv_height(u, 6, 8), orc.tv_height(u, 6, 8)) < 1e-11
    assert rel(hip.tv(u, niter=1), orc.tv(u, niter=1)) < 1e-14


def test_tv_fp32(hip, orc):
    # 40 sweeps of t/sqrt(1e-8 + |t|^2) amplify input perturbations ~1e3-1e4x
    # (the reference map itself, measured on the oracle): FP32 must stay
    # within a small multiple of the oracle's own response to an input
    # perturbation of FP32 rounding size
    u = smooth_volume((10, 16, 16), 5)
    ref = orc.tv(u)
    floor = rel(orc.tv(perturbed(u, 6e-8, 1)), ref)
    assert rel(hip.tv(u, precision=1), ref) < 4 * floor + 1e-6
    ref_h = orc.tv_height(u, 6, 8)
    floor_h = rel(orc.tv_height(perturbed(u, 6e-8, 2), 6, 8), ref_h)
    assert rel(hip.tv_height(u, 6, 8, precision=1), ref_h) < 4 * floor_h + 1e-6
    # one sweep is well conditioned
    assert rel(hip.tv(u, niter=1, precision=1), orc.tv(u, niter=1)) < 1e-5


# The sweep kernels by contiguous extent nz (the first axis here): nz % 4 == 0
# takes runs of 4 z per lane, FP32 with nz >= 512 and nz % 8 == 0 runs of 8;
# a wave of runs may cover part of a line (1040: edge lanes read their
# neighbours directly), exactly one line (256 / 512) or several (36, 16).
@pytest.mark.parametrize("shape", [(16, 7, 9), (36, 5, 7), (256, 4, 5), (512, 5, 6), (1040, 3, 4)])
def test_tv_runs(hip, orc, shape):
    u = smooth_volume(shape, 11)
    # three sweeps (two through the run kernel) move u by 8-16% and stay well
    # conditioned: the oracle's own response to FP32-sized input noise is
    # below 5e-7 on these shapes
    ref3 = orc.tv(u, niter=3)
    assert rel(hip.tv(u, niter=3), ref3) < 1e-13
    assert rel(hip.tv(u, niter=3, precision=1), ref3) < 1e-5
    assert rel(hip.tv(u, niter=1), orc.tv(u, niter=1)) < 1e-14
    # 20 sweeps: FP32 within a small multiple of the oracle's response to an
    # input perturbation of FP32 rounding size.  Past ~20 sweeps that response
    # varies 10-40x between perturbation samples (t / sqrt(1e-8 + |t|^2) near
    # flat regions), so one sample does not bound an FP32 result there.
    ref20 = orc.tv(u, niter=20)
    floor20 = rel(orc.tv(perturbed(u, 6e-8, 3), niter=20), ref20)
    assert rel(hip.tv(u, niter=20, precision=1), ref20) < 4 * floor20 + 1e-6
    # 40 sweeps in FP64: the same amplification (floor / 6e-8, up to ~1e5)
    # applied to FP64 rounding, 4e-7 * floor ~ 200 ulp
    ref = orc.tv(u)
    floor = rel(orc.tv(perturbed(u, 6e-8, 3)), ref)
    assert rel(hip.tv(u), ref) < 1e-11 + 4e-7 * floor
    # a band of rows (the other rows untouched)
    ny = shape[1]
    ref_h = orc.tv_height(u, 2, ny // 2)
    floor_h = rel(orc.tv_height(perturbed(u, 6e-8, 4), 2, ny // 2), ref_h)
    assert rel(hip.tv_height(u, 2, ny // 2), ref_h) < 1e-11 + 4e-7 * floor_h
    ref_h3 = orc.tv_height(u, 2, ny // 2, niter=3)
    assert rel(hip.tv_height(u, 2, ny // 2, niter=3, precision=1), ref_h3) < 1e-5


@pytest.mark.parametrize("prec,tol", [(0, 1e-11), (1, 2e-5)])
def test_diffusion(hip, orc, prec, tol):
    u = smooth_volume((12, 10, 12), 6)
    assert rel(hip.diffusion2d(u, precision=prec), orc.diffusion2d(u)) < tol
    u2 = smooth_volume((16, 16, 16), 7)
    assert rel(hip.diffusion2d(u2, niter=5, precision=prec), orc.diffusion2d(u2, niter=5)) < tol


@pytest.mark.parametrize("prec", [0, 1])
def test_constraint(hip, orc, prec):
    rng = np.random.default_rng(8)
    img = rng.standard_normal((6, 16, 16))
    last = rng.standard_normal((6, 16, 16))
    tol = 1e-15 if prec == 0 else 1e-6
    for hio in (0, 1):
        assert rel(hip.constraint(img, last, 12, hio, precision=prec), orc.constraint(img, last, 12, hio)) < tol
        assert rel(hip.constraint(img, last, 16, hio, height=6, center=7, precision=prec),
                   orc.constraint(img, last, 16, hio, 1, 6, 7)) < tol


def small_case(orc, np_=32, nz=32, seed=9, noise=0.0):
    rng = np.random.default_rng(seed)
    dens = blob_phantom((nz, np_, np_), 12, rng, sig=(1.5, 3.0), slab=np_ // 2)
    proj = orc.project_2d1d(dens, np_, -60, 60, 5)
    if noise:
        proj = proj + noise * proj.std() * rng.standard_normal(proj.shape)
    return proj


def oracle_passes(orc, op, img, slices, kind, n):
    rs = []
    for nn in range(n):
        img, r = orc.iterate(op, img, slices, kind, nn)
        rs.append(r)
    return img, np.array(rs)


@pytest.mark.parametrize("denoise,method,height", [(0, 1, None), (0, 1, 16), (1, 1, None), (1, 1, 16), (1, 0, None)])
def test_end_to_end_small(hip, orc, denoise, method, height):
    """initial INFFT + 2 HIO + 2 ER passes vs the oracle.

    Without TV the iteration contracts perturbations, so the FP64 path must
    agree to rounding and the FP32 form to 1e-4.  TV amplifies any
    perturbation ~1e3x per pass until a ~5e-3 plateau (measured on the oracle
    against itself), so after the first pass the bar is the oracle's own
    response to an input perturbation of the path's rounding size."""
    np_ = nz = 32
    proj = small_case(orc, np_, nz)
    center = None if height is None else 16
    kw = dict(padangle=1.0, weight=1, denoise=denoise, method=method, nhio=2, ner=2)
    op = orc.params(np_, nz, np_, -60, 60, 5, height=height, center=center, **kw)
    slices, kind = orc.assemble(op, proj)
    img0 = orc.initial(op, slices)
    img1, r1 = orc.iterate(op, img0, slices, kind, 0)
    img_ref, r_ref = oracle_passes(orc, op, img0, slices, kind, 4)
    chaotic = denoise and method == 1
    results = {}
    for prec in (hip.FP64_PLAIN, hip.FP32):
        hp = hip.make_params(np_, nz, np_, -60, 60, 5, height=height, center=center, precision=prec, **kw)
        with hip.Reconstruction(hp) as job:
            job.load_projections(proj)
            job.initial()
            v0 = job.volume()
            ra = job.iterate(1)
            v1 = job.volume()
            rb = job.iterate(3)
            results[prec] = (v0, v1, job.volume(), np.concatenate([ra, rb]))
    v0, v1, v4, r = results[hip.FP64_PLAIN]
    assert rel(v0, img0) < 1e-11
    assert rel(v1, img1) < (1e-8 if chaotic else 1e-10)
    assert abs(r[0] - r_ref[0]) < 1e-12
    if not chaotic:
        assert rel(v4, img_ref) < 1e-9
        assert np.max(np.abs(r - r_ref)) < 1e-10
    else:
        f64, _ = oracle_passes(orc, op, perturbed(img0, 1e-13, 3), slices, kind, 4)
        assert rel(v4, img_ref) < 10 * rel(f64, img_ref) + 1e-9
    # FP32 form.  HIO (reconstruction.c:682) jumps from img to img_last - 0.9 img
    # where img crosses zero, so a voxel within rounding of zero may flip: the
    # HIO passes are held to a flip-tolerant bar (99.9th percentile of the
    # voxel error), the ER passes (a continuous map) to relative L2.
    v0, v1, v4_32, r32 = results[hip.FP32]
    assert rel(v0, img0) < 2e-5
    assert abs(r32[0] - r_ref[0]) < 1e-6
    scale = np.abs(img_ref).max()
    if not denoise:
        assert np.quantile(np.abs(v1 - img1), 0.999) < 1e-4 * np.abs(img1).max()
        assert np.quantile(np.abs(v4_32 - img_ref), 0.999) < 1e-4 * scale
        assert np.max(np.abs(r32 - r_ref)) < 1e-4
    else:
        # the denoiser spreads an HIO flip over its neighbourhood: compare with
        # the oracle's own response to an input perturbation of FP32 size
        f32, _ = oracle_passes(orc, op, perturbed(img0, 1e-6, 4), slices, kind, 4)
        assert rel(v4_32, img_ref) < 4 * rel(f32, img_ref) + 1e-4


@pytest.mark.parametrize("np_,nz", [(80, 96), (48, 192)])
def test_end_to_end_z_lengths(hip, orc, np_, nz):
    """Jobs on the register z transforms of length 96 (16 x 6) and 192
    (16 x 12), the first with the fused Toeplitz passes at L = 160: initial
    volume and two ER passes with TV over the height band against the oracle
    (FP64 plain to 1e-10 before the denoiser's sensitivity sets in, i.e. the
    initial volume and first pass; FP32 within 1e-4 initially)."""
    rng = np.random.default_rng(np_ + nz)
    dens = blob_phantom((nz, np_, np_), 10, rng, sig=(1.5, 3.0), slab=np_ // 2)
    proj = orc.project_2d1d(dens, np_, -60, 60, 3)
    kw = dict(padangle=1.0, weight=1, nhio=0, ner=2, denoise=1, method=1, height=np_ // 2, center=np_ // 2)
    op = orc.params(np_, nz, np_, -60, 60, 3, **kw)
    slices, kind = orc.assemble(op, proj)
    img0 = orc.initial(op, slices)
    img1, r1 = orc.iterate(op, img0, slices, kind, 0)
    for prec, tol0, tol1 in ((hip.FP64_PLAIN, 1e-11, 1e-8), (hip.FP32, 2e-5, None)):
        with hip.Reconstruction(hip.make_params(np_, nz, np_, -60, 60, 3, precision=prec, **kw)) as job:
            job.load_projections(proj)
            job.initial()
            v0 = job.volume()
            r = job.iterate(1)
            v1 = job.volume()
        assert rel(v0, img0) < tol0
        if tol1:
            assert rel(v1, img1) < tol1 and abs(r[0] - r1) < 1e-12
        else:
            assert abs(r[0] - r1) < 1e-5
            f32, _ = oracle_passes(orc, op, perturbed(img0, 1e-6, 4), slices, kind, 1)
            assert rel(v1, img1) < 4 * rel(f32, img1) + 1e-4


@pytest.mark.parametrize("denoise,method,np_", [(0, 1, 32), (1, 0, 32), (0, 1, 64), (1, 0, 64)])
def test_end_to_end_er_fp32(hip, orc, denoise, method, np_):
    """ER passes only (continuous): FP32 form within 1e-4 relative L2.  At
    np = nz = 64 the Toeplitz CGNR runs on the fused fixed-length passes
    (2N = 128) and the NFFT of the initial volume on the patch-form kernels
    (64 planes).

    With DIFFUSION at 64^3 the denoiser amplifies the ~2e-6 per-pass FP32
    rounding with a heavy tail: the oracle fed relative perturbations of 2e-6
    before each pass ends 1.9e-5 .. 5.5e-5 away from itself over six seeds,
    and the FP32 form (either FFT path) ends 2.3e-5 .. 2.1e-4 away over four
    phantoms, so that case is held to 5e-4 (a wrong operator is off by 1e-2
    or more)."""
    nz = np_
    proj = small_case(orc, np_, nz, seed=12)
    kw = dict(padangle=1.0, weight=1, denoise=denoise, method=method, nhio=0, ner=3, height=np_ // 2,
              center=np_ // 2)
    op = orc.params(np_, nz, np_, -60, 60, 5, **kw)
    img0_ref, img_ref, r_ref = orc.reconstruct(op, proj)
    with hip.Reconstruction(hip.make_params(np_, nz, np_, -60, 60, 5, precision=hip.FP32, **kw)) as job:
        job.load_projections(proj)
        job.initial()
        r = job.iterate(3)
        v = job.volume()
    assert rel(v, img_ref) < (5e-4 if (denoise and np_ == 64) else 1e-4)
    assert np.max(np.abs(r - r_ref)) < 1e-5


@pytest.mark.parametrize("np_,nz,weight,nhio", [(64, 64, 1, 0), (64, 48, 0, 0), (128, 32, 1, 0), (64, 48, 1, 2),
                                                (128, 32, 0, 1)])
def test_projected_cg_start(hip, orc, monkeypatch, np_, nz, weight, nhio):
    """Each pass's Toeplitz CGNR starts from z = b_K + T_U fh - T f0 (two
    Toeplitz products, b_K = A^H W K m once per job) instead of the adjoint
    NFFT of W (y - A f0).  The two are the same map:
      - in FP64 (Toeplitz form) the projected start matches the oracle to
        1e-9 (R to 1e-10), HIO or ER; without weights the conditioning
        lifts FP64 rounding to ~4e-9, held to 1e-7;
      - in FP32 with density weights and ER passes the two starts agree to
        FP32 rounding (2e-5) and both stay within 1e-4 of the oracle.
    Without weights the 10-step CGNR is ill-conditioned and FP32 rounding
    moves its iterate at the 1e-2 level whichever start is used, so there the
    projected start is held to no worse than twice the adjoint start's
    distance to the oracle.  HIO's sign rule is discontinuous (a voxel near
    zero flips between img and img_last - 0.9 img), so FP32 runs with HIO
    passes compare the 0.999 quantile of the voxel differences."""
    proj = small_case(orc, np_, nz, seed=31)
    kw = dict(padangle=1.0, weight=weight, nhio=nhio, ner=4 - nhio, height=np_ // 2, center=np_ // 2)
    op = orc.params(np_, nz, np_, -60, 60, 5, **kw)
    _, img_ref, r_ref = orc.reconstruct(op, proj)
    sc = np.abs(img_ref).max()

    def run(prec, start, form=hip.CG_AUTO):
        if start == "nfft":
            monkeypatch.setenv("NUDIM_CG_START", "nfft")
        else:
            monkeypatch.delenv("NUDIM_CG_START", raising=False)
        p = hip.make_params(np_, nz, np_, -60, 60, 5, precision=prec, cg_form=form, **kw)
        with hip.Reconstruction(p) as job:
            job.load_projections(proj)
            job.initial()
            r = job.iterate(p.nhio + p.ner)
            return job.volume(), r, job.device_bytes()

    v64, r64, _ = run(hip.FP64_PLAIN, "projected", hip.CG_TOEPLITZ)
    e64 = rel(v64, img_ref) if nhio == 0 else np.max(np.abs(v64 - img_ref)) / sc
    assert e64 < (1e-9 if weight else 1e-7), e64
    assert np.max(np.abs(r64 - r_ref)) < (1e-10 if weight else 1e-8)
    (va, ra, ba), (vb, rb, bb) = run(hip.FP32, "nfft"), run(hip.FP32, "projected")
    assert bb > ba  # the second kernel, intermediate and b_K are held
    dr, dro = np.max(np.abs(rb - ra)), np.max(np.abs(rb - r_ref))
    if nhio == 0:
        e, eo, ea = rel(vb, va), rel(vb, img_ref), rel(va, img_ref)
    else:
        e = np.quantile(np.abs(vb - va), 0.999) / sc
        eo = np.quantile(np.abs(vb - img_ref), 0.999) / sc
        ea = np.quantile(np.abs(va - img_ref), 0.999) / sc
    if weight:
        assert e < (2e-5 if nhio == 0 else 1e-4), e
        assert eo < 1e-4, (eo, ea)
        assert dr < (1e-7 if nhio == 0 else 1e-5) and dro < 1e-4, (dr, dro)
    else:
        assert eo < 2 * ea + 1e-4, (eo, ea)


@pytest.mark.parametrize("np_,nz", [(64, 128), (64, 256), (32, 512)])
def test_z_transform_lengths(hip, orc, np_, nz):
    """The z transforms at the lengths the fixed-length z kernels cover
    (with the z shift of reconstruction.c:536-621 folded into their index
    maps, and in the FP32 form the CG's last f += alpha p folded into the
    inverse's loads): FP64 initial volume within 1e-11 of the oracle and one
    ER pass within 1e-10; FP32 within 1e-4 after two ER passes (np = 64
    runs the FP32 form on the fused Toeplitz passes, where the fold applies)."""
    rng = np.random.default_rng(nz)
    dens = blob_phantom((nz, np_, np_), 12, rng, sig=(1.5, 3.0), slab=np_ // 2)
    proj = orc.project_2d1d(dens, np_, -60, 60, 5)
    kw = dict(padangle=1.0, weight=1, nhio=0, ner=2)
    op = orc.params(np_, nz, np_, -60, 60, 5, **kw)
    slices, kind = orc.assemble(op, proj)
    img0 = orc.initial(op, slices)
    img1, r1 = orc.iterate(op, img0, slices, kind, 0)
    img2, r2 = orc.iterate(op, img1, slices, kind, 1)
    with hip.Reconstruction(hip.make_params(np_, nz, np_, -60, 60, 5, precision=hip.FP64_PLAIN, **kw)) as job:
        job.load_projections(proj)
        job.initial()
        assert rel(job.volume(), img0) < 1e-11
        r = job.iterate(1)
        assert rel(job.volume(), img1) < 1e-10
        assert abs(r[0] - r1) < 1e-12
    with hip.Reconstruction(hip.make_params(np_, nz, np_, -60, 60, 5, precision=hip.FP32, **kw)) as job:
        job.load_projections(proj)
        job.initial()
        r = job.iterate(2)
        assert rel(job.volume(), img2) < 1e-4
        assert abs(r[1] - r2) < 1e-5


@pytest.mark.parametrize("prec", [0, 1])
def test_interp_spatial_order_same_values(hip, orc, monkeypatch, prec):
    """The per-node interpolation (plane counts that are not multiples of 64)
    visits the nodes sorted by 8 x 8-cell tile (for L2 reuse of the grid
    rows); every node's sum is the same as in node order, bit for bit."""
    N, P = 64, 48
    pmin, pmax = orc.pad_range(-60, 3)
    x = orc.polar_nodes(pmin, pmax, 1.0, N)
    rng = np.random.default_rng(31)
    fh = rng.standard_normal((P, N, N)) + 1j * rng.standard_normal((P, N, N))
    W = 14 if prec == 0 else 8
    a = hip.nfft_trafo(x, fh, n=2 * N, W=W, precision=prec)
    monkeypatch.setenv("NUDIM_INTERP_ORDER", "node")
    b = hip.nfft_trafo(x, fh, n=2 * N, W=W, precision=prec)
    assert np.array_equal(a, b)
