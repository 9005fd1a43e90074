"""Pins of the FP64 oracle (CPU only).

The reference ships no tests or golden vectors (its sample inputs are
missing), so every oracle function is pinned here by something other than
its own formula: a textbook library routine, a closed form, brute force on
tiny inputs, or an invariant that a plausible mistake (a sign, a shift, a
dropped term, a transposed operand) would break.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from tests.util import rel

REF_PI = 3.1415926


# ---- geometry (grid.c, reconstruction.c:372-426) -----------------------------

def test_pad_range(orc):
    assert orc.pad_range(-60, 3) == (-90.0, 90.0)
    assert orc.pad_range(-70, 2) == (-90.0, 90.0)
    a, b = orc.pad_range(-55, 4)        # trunc toward zero: -55 + 8*(-4) / -55 + 36*4
    assert (a, b) == (-87.0, 89.0)


def test_polar_nodes_closed_form(orc):
    np_ = 16
    x = orc.polar_nodes(-90, 90, 1.0, np_)
    assert x.shape == (181 * np_, 2)
    ang = np.repeat(-90.0 + np.arange(181), np_)
    r = np.tile(-0.5 + np.arange(np_) / np_, 181)
    # accumulated sums differ from i*step by a few ulps only
    assert np.max(np.abs(x[:, 0] - r * np.sin(ang / 180 * REF_PI))) < 1e-13
    assert np.max(np.abs(x[:, 1] - r * np.cos(ang / 180 * REF_PI))) < 1e-13
    # the truncated pi is used (grid.h:11): sin(90 deg) is not exactly 1
    assert x[0, 0] != 0.5 and abs(x[0, 0] - 0.5) < 1e-14 + 0.5 * (1 - np.sin(REF_PI / 2)) + 1e-15
    # -90 and +90 rays hold the same line with mirrored radius
    assert np.max(np.abs(x[:np_, 0] + x[-np_:, 0])) < 1e-13


def test_uniform_nodes(orc):
    x = orc.uniform_nodes(6, 4)
    assert np.allclose(x[:6, 0], -0.5 + np.arange(6) / 6, atol=1e-15)
    assert np.allclose(x[::6, 1], -0.5 + np.arange(4) / 4, atol=1e-15)


def test_weights_closed_form(orc):
    np_, nang, pad = 16, 181, 1.0
    w = orc.weights(np_, nang, pad).reshape(nang, np_)
    i = np.arange(np_)
    expect = pad / 180 * REF_PI * np.abs(i - np_ // 2) / (np_ // 2) ** 2
    expect[np_ // 2] = REF_PI / (4 * nang * (np_ // 2) ** 2)
    assert np.allclose(w, expect[None, :], rtol=1e-15, atol=0)
    # odd np: integer division in abs(i - np/2) and pow(np/2, 2)
    w = orc.weights(7, 3, 1.0).reshape(3, 7)
    assert w[0, 3] == REF_PI / (4 * 3 * 9.0) and w[0, 0] == 1.0 / 180 * REF_PI * 3 / 9.0


def test_sample_classes(orc):
    # -60..60 step 3, padangle 1: 41 tilts, 40 enforced (0 deg held out)
    pmin, pmax = orc.pad_range(-60, 3)
    kind, order = orc.mask(8, 181, pmin, 1.0, -60, 60, 3)
    per_ray = kind.reshape(181, 8)[:, 0]
    assert (per_ray == 1).sum() == 40 and (per_ray == 2).sum() == 1
    ang = -90 + np.arange(181)
    assert per_ray[ang == 0][0] == 2
    known = ang[per_ray == 1]
    assert np.all(known % 3 == 0) and known.min() == -60 and known.max() == 60
    # tilt file index of a known ray
    o = order.reshape(181, 8)[:, 0]
    assert o[ang == -60][0] == 0 and o[ang == 60][0] == 40
    # padangle equal to the interval: every ray in range is known
    pmin, pmax = orc.pad_range(-70, 2)
    kind, _ = orc.mask(4, 91, pmin, 2.0, -70, 70, 2)
    assert (kind.reshape(91, 4)[:, 0] == 1).sum() == 70


# ---- FFT and NDFT / NFFT ------------------------------------------------------

@pytest.mark.parametrize("n", [2, 6, 15, 16, 60, 96, 100, 128, 7])
def test_fft_vs_numpy(orc, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, 5)) + 1j * rng.standard_normal((n, 5))
    assert rel(orc.fft_strided(a, -1), np.fft.fft(a, axis=0)) < 1e-14
    assert rel(orc.fft_strided(a, +1), np.fft.ifft(a, axis=0) * n) < 1e-14


def test_ndft_delta_closed_form(orc):
    N = 8
    x = orc.polar_nodes(-90, 90, 15.0, N)
    for k0, k1 in [(0, 0), (3, 6), (N // 2, N // 2)]:
        fh = np.zeros((N, N), np.complex128)
        fh[k0, k1] = 1
        expect = np.exp(-2j * np.pi * ((k0 - N / 2) * x[:, 0] + (k1 - N / 2) * x[:, 1]))
        assert rel(orc.ndft_trafo(x, fh), expect) < 1e-14


def test_ndft_on_uniform_grid_is_centred_fft(orc):
    # nodes (j - N/2)/N on both axes: NDFT == fftshift(fft2(ifftshift(.)))
    N = 12
    g = -0.5 + np.arange(N) / N
    x = np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2)
    rng = np.random.default_rng(0)
    fh = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    expect = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(fh)))
    assert rel(orc.ndft_trafo(x, fh).reshape(N, N), expect) < 1e-13


@pytest.mark.parametrize("N", [16, 24, 32])
def test_nfft_vs_ndft(orc, N):
    pmin, pmax = orc.pad_range(-60, 3)
    x = orc.polar_nodes(pmin, pmax, 1.0, N)
    rng = np.random.default_rng(N)
    fh = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    g = rng.standard_normal(x.shape[0]) + 1j * rng.standard_normal(x.shape[0])
    P = orc.NFFT(x, N)
    assert rel(P.trafo(fh), orc.ndft_trafo(x, fh)) < 1e-12
    assert rel(P.adjoint(g), orc.ndft_adjoint(x, g, N)) < 1e-12
    # adjointness <A f, g> = <f, A^H g>
    lhs = np.vdot(g, P.trafo(fh))
    rhs = np.vdot(P.adjoint(g), fh)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_nfft_non_square(orc):
    x = orc.polar_nodes(-60, 60, 10, 20)
    rng = np.random.default_rng(5)
    fh = rng.standard_normal((12, 16)) + 1j * rng.standard_normal((12, 16))
    P = orc.NFFT(x, 12, 16)
    k0 = np.arange(12)[:, None] - 6
    k1 = np.arange(16)[None, :] - 8
    E = np.exp(-2j * np.pi * (k0[None] * x[:, 0, None, None] + k1[None] * x[:, 1, None, None]))
    assert rel(P.trafo(fh), (E * fh[None]).sum((1, 2))) < 1e-12


def test_projection_ft_is_centred_padded_dft(orc):
    rng = np.random.default_rng(2)
    nz, nx, np_ = 10, 12, 16
    proj = rng.standard_normal((3, nz, nx))
    tomo = orc.projection_ft(proj, np_)
    for t in range(3):
        # centred coordinates: column c -> c - nx/2, row r -> r - nz/2;
        # frequency i -> (i - np/2)/np, j -> (j - nz/2)/nz
        pad = np.zeros((nz, np_))
        pad[:, (np_ - nx) // 2:(np_ - nx) // 2 + nx] = proj[t]
        # shift so that element (nz/2, np/2) sits at index 0 of the DFT
        expect = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(pad)))
        assert rel(tomo[t], expect) < 1e-12


# ---- CGNR --------------------------------------------------------------------

def test_cgnr_is_krylov_least_squares(orc):
    """After k steps from f0, weighted CGNR minimises ||y - A f||_W over
    f0 + span{z0, (A^H W A) z0, ..., (A^H W A)^{k-1} z0}, z0 = A^H W (y - A f0)."""
    N = 8
    pmin, pmax = orc.pad_range(-60, 15)
    x = orc.polar_nodes(pmin, pmax, 15.0, N)
    M = x.shape[0]
    w = orc.weights(N, M // N, 15.0)
    k0 = (np.arange(N) - N / 2)
    A = np.exp(-2j * np.pi * (x[:, 0, None, None] * k0[None, :, None] + x[:, 1, None, None] * k0[None, None, :]))
    A = A.reshape(M, N * N)
    rng = np.random.default_rng(7)
    y = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    f0 = rng.standard_normal(N * N)
    P = orc.NFFT(x, N)
    Wh = np.sqrt(w)
    B = A.conj().T @ (w[:, None] * A)
    for k in (1, 3, 6):
        got = P.cgnr(y, f0.reshape(N, N), w, k).ravel()
        z0 = A.conj().T @ (w * (y - A @ f0))
        K = [z0]
        for _ in range(k - 1):
            K.append(B @ K[-1])
        K = np.stack(K, 1)
        Q, _ = np.linalg.qr(K)
        c, *_ = np.linalg.lstsq(Wh[:, None] * (A @ Q), Wh * (y - A @ f0), rcond=None)
        expect = f0 + Q @ c
        assert rel(got, expect) < 1e-9, k
    # zero steps: the warm start is returned unchanged
    assert rel(P.cgnr(y, f0.reshape(N, N), w, 0).ravel(), f0) == 0


# ---- TV, diffusion, constraint -----------------------------------------------

def tv_functional(u, ylo=0, yhi=None, eps=1e-8):
    """sum over voxels of sqrt(eps + sum_d (u - u[clamped -1 along d])^2)"""
    nz, ny, nx = u.shape
    yhi = ny - 1 if yhi is None else yhi
    dz = u - u[np.r_[0, 0:nz - 1]]
    dy = u - u[:, np.r_[0, 0:ny - 1]]
    dx = u - u[:, :, np.r_[0, 0:nx - 1]]
    return np.sqrt(eps + dz ** 2 + dy ** 2 + dx ** 2)


def test_tv_grad_is_gradient_of_tv_functional(orc):
    rng = np.random.default_rng(3)
    u = rng.standard_normal((5, 6, 7))
    g = orc.tv_grad(u)
    h = 1e-6
    fd = np.zeros_like(u)
    for idx in np.ndindex(u.shape):
        up, um = u.copy(), u.copy()
        up[idx] += h
        um[idx] -= h
        fd[idx] = (tv_functional(up).sum() - tv_functional(um).sum()) / (2 * h)
    assert rel(g, fd) < 1e-7


def test_tv_single_sweep_formula(orc):
    rng = np.random.default_rng(4)
    u = rng.standard_normal((4, 5, 6))
    g = orc.tv_grad(u)
    dA = np.sum(u[u < 0] ** 2)
    lam = float(np.float32(0.2))          # float parameter promoted to double
    expect = u - lam * np.sqrt(dA) * (g / np.sqrt(np.sum(g * g)))
    assert rel(orc.tv(u, niter=1), expect) < 1e-14


def test_tv_invariants(orc):
    rng = np.random.default_rng(5)
    u = np.abs(rng.standard_normal((4, 5, 6))) + 0.1
    assert np.array_equal(orc.tv(u), u)              # no negative voxel: dA = 0
    c = np.full((3, 4, 5), -1.0)
    assert np.all(np.isnan(orc.tv(c, niter=1)))      # constant volume: 0/0


def test_tv_height_band(orc):
    rng = np.random.default_rng(6)
    u = rng.standard_normal((4, 10, 5))
    nheight, ncenter = 4, 5
    lo, hi = (2 * ncenter - nheight) // 2, (2 * ncenter + nheight) // 2   # 3..7 inclusive
    v = orc.tv_height(u, nheight, ncenter, niter=1)
    outside = np.r_[0:lo, hi + 1:10]
    assert np.array_equal(v[:, outside], u[:, outside])
    g = orc.tv_grad(u)[:, lo:hi + 1]
    band = u[:, lo:hi + 1]
    dA = np.sum(band[band < 0] ** 2)
    expect = band - float(np.float32(0.2)) * np.sqrt(dA) * g / np.sqrt(np.sum(g * g))
    assert rel(v[:, lo:hi + 1], expect) < 1e-14


def test_diffusion_linear_limit_is_gauss_seidel(orc):
    """For kappa -> infinity the conductances are 1 and one sweep is the
    lexicographic Gauss-Seidel step (L + D) u_new = (D' + U) u_old of the
    clamped 4-neighbour Laplacian, solvable by a sparse triangular solve."""
    rng = np.random.default_rng(8)
    ny, nx, lam = 6, 5, 0.25
    nz = nx
    u = rng.standard_normal((nz, ny, nx))
    got = orc.diffusion2d(u, lam=lam, kappa=1e12, niter=1)
    n = ny * nx
    L = sp.lil_matrix((n, n))
    R = sp.lil_matrix((n, n))
    for y in range(ny):
        for x in range(nx):
            i = y * nx + x
            diag = 1.0
            for (yy, xx, new) in [(y - 1, x, True), (y, x - 1, True), (y + 1, x, False), (y, x + 1, False)]:
                if 0 <= yy < ny and 0 <= xx < nx:
                    diag -= lam
                    j = yy * nx + xx
                    if new:
                        L[i, j] = -lam
                    else:
                        R[i, j] = lam
            L[i, i] = 1.0
            R[i, i] = diag
    L, R = L.tocsr(), R.tocsr()
    for k in range(nz):
        expect = spla.spsolve_triangular(L, R @ u[k].ravel(), lower=True)
        assert rel(got[k].ravel(), expect) < 1e-9


def test_diffusion_invariants(orc):
    c = np.full((4, 5, 4), 0.3)
    assert np.array_equal(orc.diffusion2d(c), c)
    rng = np.random.default_rng(9)
    u = rng.standard_normal((4, 5, 4))
    # huge gradients and tiny kappa: conductance ~ kappa^2/g^2, flux ~ kappa^2/g
    assert rel(orc.diffusion2d(u, kappa=1e-9, niter=3), u) < 1e-12
    with pytest.raises(ValueError):
        orc.diffusion2d(np.zeros((3, 4, 4)))


@pytest.mark.parametrize("h,c", [(0.5, 0.5), (1.5, 0.1)])
def test_diffusion_conductance_closed_form(orc, h, c):
    """Finite kappa: a single spike of height h in a zero plane.  In one
    lexicographic sweep the voxel above the spike (row y-1) and the one to
    its left (column x-1) are visited before it, each sees the spike as its
    only non-zero neighbour difference g = h, and moves by lam c(g) g with the
    Perona-Malik rational conductance c(g) = 1 / (1 + g^2 / kappa^2): 1/2 at
    g = kappa, 1/10 at g = 3 kappa.  The exponential conductance
    exp(-g^2 / kappa^2) (e^-1, e^-9) or a sign or neighbour mistake fails."""
    lam, kappa = 0.25, 0.5  # exact in float, as the reference's float arguments
    u = np.zeros((6, 6, 6))
    u[2, 3, 3] = h
    got = orc.diffusion2d(u, lam=lam, kappa=kappa, niter=1)
    a = lam * c * h
    assert got[2, 2, 3] == a
    assert got[2, 3, 2] == a
    # voxels before them in the sweep saw only zeros
    assert np.all(got[2, :2] == 0) and np.all(got[2, 2, :3] == 0)
    # other planes are independent
    assert np.all(got[[0, 1, 3, 4, 5]] == 0)


def test_mt19937_stream_known_answers():
    """The --random stream: numpy's MT19937 with init_genrand seeding gives
    the generator's published known answers (seed 5489: first output
    3499211612, 10000th output 4123659995, the value C++ [rand.predef]
    requires of std::mt19937)."""
    g = np.random.MT19937()
    g._legacy_seeding(5489)
    r = g.random_raw(10000)
    assert r[0] == 3499211612 and r[-1] == 4123659995


def test_random_fill(orc):
    """--random fills exactly the unknown and held-out samples, plane by
    plane in node order, with values in [0, 10) from one stream; the known
    samples keep the measured values."""
    np_ = nz = 16
    p = orc.params(np_, nz, np_, -60, 60, 5)
    rng = np.random.default_rng(1)
    proj = rng.standard_normal((25, nz, np_))
    slices, kind = orc.assemble(p, proj)
    filled = orc.random_fill(slices, kind)
    known = kind == 1
    assert np.array_equal(filled[:, known], slices[:, known])
    v = filled[:, ~known]
    assert np.all(v.imag == 0) and np.all((v.real >= 0) & (v.real < 10))
    g = np.random.MT19937()
    g._legacy_seeding(4357)
    first = g.random_raw(3) / 4294967296.0 * 10
    assert np.array_equal(v[0, :3].real, first)
    # the second plane continues the stream
    g._legacy_seeding(4357)
    assert np.array_equal(v[1, :2].real, g.random_raw(v.shape[1] + 2)[-2:] / 4294967296.0 * 10)


def test_constraint_closed_form(orc):
    rng = np.random.default_rng(10)
    img = rng.standard_normal((3, 8, 8))
    last = rng.standard_normal((3, 8, 8))
    nx = 4
    sup = np.zeros((8, 8), bool)
    sup[2:6, 2:6] = True
    er = orc.constraint(img, last, nx, 0)
    expect = np.where(sup[None] & (img >= 0), img, 0.0)
    assert np.array_equal(er, expect)
    hio = orc.constraint(img, last, nx, 1)
    expect = np.where(sup[None] & (img >= 0), img, last - 0.9 * img)
    assert np.array_equal(hio, expect)
    # height band: rows (2c-h)/2 .. (2c+h)/2 inclusive, all columns
    band = orc.constraint(img, last, nx, 0, 1, 3, 4)
    rows = np.zeros(8, bool)
    rows[2:6] = True
    expect = np.where(rows[None, :, None] & (img >= 0), img, 0.0)
    assert np.array_equal(band, expect)


# ---- simulator ---------------------------------------------------------------

def test_simulator_modes_agree_and_zero_degree_is_a_sum(orc):
    from tests.util import blob_phantom
    rng = np.random.default_rng(11)
    d = blob_phantom((12, 12, 12), 4, rng, sig=(1.5, 2.5))
    p1 = orc.project_2d1d(d, 16, -60, 60, 30)
    p2 = orc.project_dft(d, 16, -60, 60, 30)
    # the DFT mode's exponent uses PI = 3.1415926: agreement to ~1e-7
    assert rel(p1, p2) < 1e-6
    # 0 degrees: the line integral along y, centred in the padded row
    zero = p1[2]
    assert np.max(np.abs(zero[:, 2:14] - d.sum(axis=1))) < 1e-10
    assert np.max(np.abs(zero[:, :2])) < 1e-10
