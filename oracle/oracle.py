"""ctypes wrapper of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the parity checker for the HIP path and the CPU
baseline of bench.py.  The product (nudim_amd) never imports this module.

Arrays follow the reference's layouts: volumes are float64 [nz][np][np]
(img[z][y][x], src/reconstruction.c:527), complex arrays are complex128.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    """liboracle.so, and _ref/mrcread where the reference tree is present."""
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_cp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class OracleParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("nz", C.c_int), ("np", C.c_int),
                ("gam_min", C.c_double), ("gam_max", C.c_double),
                ("interval", C.c_double), ("padangle", C.c_double),
                ("weight", C.c_int), ("denoise", C.c_int), ("method", C.c_int),
                ("config_height", C.c_int), ("nheight", C.c_int), ("ncenter", C.c_int),
                ("nhio", C.c_int), ("ner", C.c_int), ("nthreads", C.c_int),
                ("tv_full", C.c_int), ("diffusion_lambda", C.c_float)]


def _declare(L):
    i, d, f, v = C.c_int, C.c_double, C.c_float, C.c_void_p
    sig = {
        "oracle_polar_nodes": (i, [d, d, d, i, v]),
        "oracle_uniform_nodes": (i, [i, i, _dp]),
        "oracle_pad_range": (None, [d, d, C.POINTER(d), C.POINTER(d)]),
        "oracle_weights": (None, [i, i, d, _dp]),
        "oracle_mask": (None, [i, i, d, d, d, d, d, _ip, _ip]),
        "oracle_projection_ft": (None, [i, i, i, i, _dp, _cp]),
        "oracle_ndft_trafo": (None, [i, i, _dp, _cp, _cp]),
        "oracle_ndft_adjoint": (None, [i, i, _dp, _cp, _cp]),
        "oracle_nfft_create": (v, [i, i, i, i, i, i, _dp]),
        "oracle_nfft_destroy": (None, [v]),
        "oracle_nfft_trafo": (None, [v, _cp, _cp]),
        "oracle_nfft_adjoint": (None, [v, _cp, _cp]),
        "oracle_kb_beta": (d, [i, d]),
        "oracle_fft_strided": (None, [_cp, i, i, i]),
        "oracle_fft": (None, [_cp, i, i]),
        "oracle_cgnr": (None, [v, v, _cp, _cp, i]),
        "oracle_tv_grad": (None, [_dp, i, i, i, i, i, _dp]),
        "oracle_tv": (None, [_dp, f, i, i, i, i]),
        "oracle_tv_height": (None, [_dp, f, i, i, i, i, i, i]),
        "oracle_diffusion2d": (i, [_dp, f, f, i, i, i, i]),
        "oracle_constraint": (None, [_dp, _dp, i, i, i, i, i, i, i]),
        "oracle_reconstruct": (i, [C.POINTER(OracleParams), _dp, v, v, v]),
        "oracle_assemble": (i, [C.POINTER(OracleParams), _dp, _cp, _ip]),
        "oracle_initial": (i, [C.POINTER(OracleParams), _cp, _dp]),
        "oracle_iterate": (i, [C.POINTER(OracleParams), _dp, v, _cp, _ip, i, C.POINTER(d)]),
        "oracle_time_fourier_planes": (d, [C.POINTER(OracleParams), _dp, _cp, _ip, i, i]),
        "oracle_project_2d1d": (i, [i, i, i, i, d, d, d, _dp, _dp]),
        "oracle_project_dft": (i, [i, i, i, i, d, d, d, _dp, _dp]),
        "oracle_set_threads": (None, [i]),
        "oracle_wtime": (d, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def _vp(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---- geometry --------------------------------------------------------------

def polar_nodes(amin, amax, ainterval, np_):
    M = lib().oracle_polar_nodes(amin, amax, ainterval, np_, None)
    x = np.zeros((M, 2))
    lib().oracle_polar_nodes(amin, amax, ainterval, np_, _vp(x))
    return x


def uniform_nodes(nx, nz):
    x = np.zeros((nz * nx, 2))
    lib().oracle_uniform_nodes(nx, nz, x)
    return x


def pad_range(gam_min, interval):
    a, b = C.c_double(), C.c_double()
    lib().oracle_pad_range(gam_min, interval, C.byref(a), C.byref(b))
    return a.value, b.value


def weights(np_, nang, padangle):
    w = np.zeros(nang * np_)
    lib().oracle_weights(np_, nang, padangle, w)
    return w


def mask(np_, nang, pad_min, padangle, gam_min, gam_max, interval):
    kind = np.zeros(nang * np_, np.int32)
    order = np.zeros(nang * np_, np.int32)
    lib().oracle_mask(np_, nang, pad_min, padangle, gam_min, gam_max, interval, kind, order)
    return kind, order


# ---- transforms ------------------------------------------------------------

def projection_ft(proj, np_):
    nmeas, nz, nx = proj.shape
    tomo = np.zeros((nmeas, nz, np_), np.complex128)
    lib().oracle_projection_ft(nx, nz, np_, nmeas, np.ascontiguousarray(proj, np.float64), tomo)
    return tomo


def ndft_trafo(x, fhat):
    N = fhat.shape[0]
    f = np.zeros(x.shape[0], np.complex128)
    lib().oracle_ndft_trafo(N, x.shape[0], np.ascontiguousarray(x), np.ascontiguousarray(fhat, np.complex128), f)
    return f


def ndft_adjoint(x, f, N):
    fh = np.zeros((N, N), np.complex128)
    lib().oracle_ndft_adjoint(N, x.shape[0], np.ascontiguousarray(x), np.ascontiguousarray(f, np.complex128), fh)
    return fh


class NFFT:
    """Kaiser-Bessel NFFT plan; the reference's parameters are n = 2*nextpow2(N), W = 14."""

    def __init__(self, x, N0, N1=None, n0=None, n1=None, W=14):
        N1 = N0 if N1 is None else N1
        n0 = n0 or 2 * _np2(N0)
        n1 = n1 or 2 * _np2(N1)
        self.x = np.ascontiguousarray(x, np.float64)
        self.N0, self.N1, self.M = N0, N1, x.shape[0]
        self._p = lib().oracle_nfft_create(N0, N1, n0, n1, W, self.M, self.x)

    def trafo(self, fhat):
        f = np.zeros(self.M, np.complex128)
        lib().oracle_nfft_trafo(self._p, np.ascontiguousarray(fhat, np.complex128), f)
        return f

    def adjoint(self, f):
        fh = np.zeros((self.N0, self.N1), np.complex128)
        lib().oracle_nfft_adjoint(self._p, np.ascontiguousarray(f, np.complex128), fh)
        return fh

    def cgnr(self, y, fhat0, w=None, nsteps=10):
        fh = np.array(fhat0, np.complex128, copy=True, order="C")
        wv = None if w is None else np.ascontiguousarray(w, np.float64)
        lib().oracle_cgnr(self._p, _vp(wv), np.ascontiguousarray(y, np.complex128), fh, nsteps)
        return fh

    def __del__(self):
        if getattr(self, "_p", None):
            lib().oracle_nfft_destroy(self._p)
            self._p = None


def _np2(n):
    p = 1
    while p < n:
        p <<= 1
    return p


def fft_strided(data, sign):
    """unnormalised DFT along axis 0 (FFTW sign convention), in a copy"""
    d = np.array(data, np.complex128, copy=True, order="C")
    lib().oracle_fft_strided(d, d.shape[0], d.size // d.shape[0], sign)
    return d


# ---- regularisers ----------------------------------------------------------

def tv_grad(u, ylo=None, yhi=None):
    nz, ny, nx = u.shape
    g = np.zeros_like(u)
    lib().oracle_tv_grad(np.ascontiguousarray(u), nz, ny, nx,
                         0 if ylo is None else ylo, ny - 1 if yhi is None else yhi, g)
    return g


def tv(u, lam=0.2, niter=40):
    v = np.array(u, np.float64, copy=True, order="C")
    lib().oracle_tv(v, lam, niter, *v.shape)
    return v


def tv_height(u, nheight, ncenter, lam=0.2, niter=40):
    v = np.array(u, np.float64, copy=True, order="C")
    lib().oracle_tv_height(v, lam, niter, *v.shape, nheight, ncenter)
    return v


def diffusion2d(u, lam=0.25, kappa=0.005, niter=80):
    v = np.array(u, np.float64, copy=True, order="C")
    rc = lib().oracle_diffusion2d(v, lam, kappa, niter, *v.shape)
    if rc:
        raise ValueError("diffusion2d needs nz == nx (reference clamps x with the z extent)")
    return v


def constraint(img, img_last, nx, is_hio, config_height=0, nheight=-1, ncenter=-1):
    v = np.array(img, np.float64, copy=True, order="C")
    nz, np_, _ = v.shape
    lib().oracle_constraint(v, np.ascontiguousarray(img_last, np.float64), nz, np_, nx,
                            int(is_hio), int(config_height), nheight, ncenter)
    return v


# ---- end to end ------------------------------------------------------------

def params(nx, nz, np_, gam_min, gam_max, interval, padangle=1.0, weight=1, denoise=0,
           method=1, height=None, center=None, nhio=0, ner=1, nthreads=0, tv_full=False,
           diffusion_lambda=0.0):
    ch = height is not None or center is not None
    return OracleParams(nx, nz, np_, gam_min, gam_max, interval, padangle, int(weight),
                        int(denoise), int(method), int(ch), height or -1, center or -1,
                        nhio, ner, nthreads, int(tv_full), diffusion_lambda)


def reconstruct(p, proj):
    V = p.np * p.np * p.nz
    img0, img = np.zeros(V), np.zeros(V)
    rt = np.zeros(max(1, p.nhio + p.ner))
    rc = lib().oracle_reconstruct(C.byref(p), np.ascontiguousarray(proj, np.float64),
                                  _vp(img0), _vp(img), _vp(rt))
    if rc:
        raise RuntimeError("oracle_reconstruct failed (%d)" % rc)
    sh = (p.nz, p.np, p.np)
    return img0.reshape(sh), img.reshape(sh), rt[: p.nhio + p.ner]


def assemble(p, proj):
    pmin, pmax = pad_range(p.gam_min, p.interval)
    M = polar_nodes(pmin, pmax, p.padangle, p.np).shape[0]
    slices = np.zeros((p.nz, M), np.complex128)
    kind = np.zeros(M, np.int32)
    rc = lib().oracle_assemble(C.byref(p), np.ascontiguousarray(proj, np.float64), slices, kind)
    if rc:
        raise RuntimeError("oracle_assemble failed")
    return slices, kind


def random_fill(slices, kind, seed=0):
    """--random padding (src/reconstruction.c:434-469): every sample that is
    not known (unknown rays and the held-out 0 degree ray) gets
    gsl_rng_uniform(rng) * 10 as its real part, planes outer and nodes inner,
    from gsl_rng_default = gsl_rng_mt19937 with the default seed 0, which
    GSL's mt19937 seeding maps to 4357.  The stream is numpy's MT19937 with
    the init_genrand seeding (`_legacy_seeding`), an implementation of the
    same published generator (Matsumoto and Nishimura 1998) independent of
    the product's; gsl_rng_uniform for mt19937 is the 32-bit output / 2^32.
    Returns a new array."""
    nz, M = slices.shape
    unk = np.flatnonzero(np.asarray(kind) != 1)
    g = np.random.MT19937()
    g._legacy_seeding(4357 if seed == 0 else seed)
    draws = g.random_raw(nz * unk.size).astype(np.float64) / 4294967296.0 * 10
    out = np.array(slices, np.complex128, copy=True)
    out[:, unk] = draws.reshape(nz, unk.size)
    return out


def initial(p, slices):
    img0 = np.zeros((p.nz, p.np, p.np))
    rc = lib().oracle_initial(C.byref(p), np.ascontiguousarray(slices), img0)
    if rc:
        raise RuntimeError("oracle_initial failed")
    return img0


def iterate(p, img, slices, kind, nn):
    v = np.array(img, np.float64, copy=True, order="C")
    last = np.zeros_like(v)
    r = C.c_double()
    rc = lib().oracle_iterate(C.byref(p), v, _vp(last), np.ascontiguousarray(slices), kind, nn, C.byref(r))
    if rc:
        raise RuntimeError("oracle_iterate failed (%d)" % rc)
    return v, r.value


def time_fourier_planes(p, img, slices, kind, j0, j1):
    return lib().oracle_time_fourier_planes(C.byref(p), np.ascontiguousarray(img, np.float64),
                                            np.ascontiguousarray(slices), kind, j0, j1)


def project_2d1d(density, np_, gam_min, gam_max, interval):
    nz, ny, nx = density.shape
    ns = int((gam_max - gam_min) / interval + 0.5) + 1
    out = np.zeros((ns, nz, np_))
    rc = lib().oracle_project_2d1d(nx, ny, nz, np_, gam_min, gam_max, interval,
                                   np.ascontiguousarray(density, np.float64), out)
    if rc:
        raise RuntimeError("oracle_project_2d1d failed")
    return out


def project_dft(density, np_, gam_min, gam_max, interval):
    nz, ny, nx = density.shape
    ns = int((gam_max - gam_min) / interval + 0.5) + 1
    out = np.zeros((ns, nz, np_))
    rc = lib().oracle_project_dft(nx, ny, nz, np_, gam_min, gam_max, interval,
                                  np.ascontiguousarray(density, np.float64), out)
    if rc:
        raise RuntimeError("oracle_project_dft failed")
    return out


def set_threads(n):
    lib().oracle_set_threads(n)


# ---- MRC converter (examples/exp) ------------------------------------------

def mrc_convert(stack, shift_x=61, shift_y=1):
    """examples/exp/mrcread.c:38-80 on stack float32 [nz][ny][nx] (the
    sections as mrc_read_slice converts them, mrcfile.c:182-207): section 0
    dropped; for k >= 1 value (j, i) moves to column (i - 61) mod nx and, for
    i < 61, to row (j - 1) mod ny; output max - value + min in float64 with
    max/min the section's extremes started at -100000 / 100000.  Returns
    {output index nz-1-k: (values [ny*nx] float64, max, min)}."""
    nz, ny, nx = stack.shape
    out = {}
    i = np.arange(nx)
    for k in range(1, nz):
        sec = stack[k].astype(np.float64)
        mx = max(-100000.0, float(sec.max()))
        mn = min(100000.0, float(sec.min()))
        moved = np.empty_like(sec)
        wide = i >= shift_x
        moved[:, (i[wide] - shift_x) % nx] = sec[:, wide]
        rows = (np.arange(ny) - shift_y) % ny
        moved[rows[:, None], ((i[~wide] - shift_x) % nx)[None, :]] = sec[:, ~wide]
        out[nz - 1 - k] = ((mx - moved) + mn).ravel(), mx, mn
    return out
