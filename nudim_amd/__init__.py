"""nudim_amd -- MI355X (gfx950) implementation of NUDIM's HIO/ER iteration.

Python binding of libnudim_hip.so (C ABI in include/nudim.h).  The binding
only marshals host arrays; every computation runs in the HIP library.  If
the library is missing the import of this module still works, but every
call raises: there is no CPU fallback.

Arrays follow the reference's layouts: volumes float64 [nz][np][np]
(img[z][y][x], src/reconstruction.c:527), projections float64
[nmeas][nz][nx] (./files/<prefix>-<j>.txt, src/reconstruction.c:341-358),
complex arrays complex128.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnudim_hip.so")

FP64_PLAIN, FP32 = 0, 1
DIFFUSION, TV = 0, 1
CG_AUTO, CG_NFFT, CG_TOEPLITZ = 0, 1, 2
KC_TOEPLITZ_COL, KC_INTERP, KC_GATHER, KC_LINE_FFT, KC_TV, KC_ZFFT = 1, 2, 3, 4, 5, 6
KC_TOEPLITZ_ROWF, KC_TOEPLITZ_ROWI, KC_TOEPLITZ_COL2, KC_CG_UPDATE, KC_DIFFUSION = 7, 8, 9, 10, 11

_LIB = None


class NudimError(RuntimeError):
    pass


class Params(C.Structure):
    """nudim_params (include/nudim.h)"""
    _fields_ = [("nx", C.c_int), ("nz", C.c_int), ("np", C.c_int),
                ("gam_min", C.c_double), ("gam_max", C.c_double), ("interval", C.c_double),
                ("padangle", C.c_double), ("weight", C.c_int), ("random", C.c_int),
                ("seed", C.c_ulong), ("config_height", C.c_int), ("nheight", C.c_int),
                ("ncenter", C.c_int), ("denoise", C.c_int), ("method", C.c_int),
                ("nhio", C.c_int), ("ner", C.c_int), ("precision", C.c_int), ("device", C.c_int),
                ("nfft_n", C.c_int), ("nfft_w", C.c_int), ("cg_form", C.c_int),
                ("use_graph", C.c_int), ("tv_full", C.c_int), ("diffusion_lambda", C.c_float),
                ("cu_mask", C.c_uint * 8)]


EXPORTS = [
    "nudim_last_error", "nudim_default_params", "nudim_create", "nudim_destroy", "nudim_sizes",
    "nudim_load_projections", "nudim_load_samples", "nudim_initial", "nudim_iterate",
    "nudim_get_volume", "nudim_set_volume", "nudim_set_dump", "nudim_device_bytes",
    "nudim_estimate_bytes", "nudim_autotune", "nudim_synchronize", "nudim_stream_probe", "nudim_profile", "nudim_profile_read", "nudim_nfft_trafo", "nudim_nfft_adjoint",
    "nudim_cgnr", "nudim_zfft", "nudim_tv", "nudim_tv_height", "nudim_tv_height_noalloc",
    "nudim_diffusion2d",
    "nudim_constraint", "nudim_projection_ft", "nudim_project_2d1d",
    "nudim_server_create", "nudim_server_submit", "nudim_server_next", "nudim_server_stats",
    "nudim_server_destroy", "nudim_batch_create", "nudim_batch_load_projections", "nudim_batch_get_volume",
]
EAGAIN = -6


SERVE_FIFO, SERVE_SHORTEST = 0, 1


class ServerOpts(C.Structure):
    """nudim_server_opts (include/nudim.h)"""
    _fields_ = [("hbm_budget", C.c_longlong), ("max_active", C.c_int), ("device", C.c_int),
                ("policy", C.c_int), ("max_wait_s", C.c_double), ("small_cus", C.c_int),
                ("small_work", C.c_double), ("max_stack", C.c_int)]


class JobReport(C.Structure):
    """nudim_job_report (include/nudim.h)"""
    _fields_ = [("job_id", C.c_longlong), ("status", C.c_int), ("t_submit", C.c_double),
                ("t_start", C.c_double), ("t_end", C.c_double), ("device_bytes", C.c_longlong),
                ("error", C.c_char * 256), ("stacked", C.c_int)]


def lib():
    """Load libnudim_hip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise NudimError("libnudim_hip.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _LIB = L
    return _LIB


def _declare(L):
    i, d, f, v, dp, ip = C.c_int, C.c_double, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p
    sig = {
        "nudim_last_error": (C.c_char_p, []),
        "nudim_default_params": (None, [C.POINTER(Params)]),
        "nudim_create": (i, [C.POINTER(Params), C.POINTER(v)]),
        "nudim_destroy": (None, [v]),
        "nudim_sizes": (i, [v, C.POINTER(i), C.POINTER(i), C.POINTER(i)]),
        "nudim_load_projections": (i, [v, dp]),
        "nudim_load_samples": (i, [v, dp, ip]),
        "nudim_initial": (i, [v]),
        "nudim_iterate": (i, [v, i, dp]),
        "nudim_get_volume": (i, [v, dp]),
        "nudim_set_volume": (i, [v, dp, i]),
        "nudim_set_dump": (i, [v, i, dp, dp]),
        "nudim_device_bytes": (C.c_size_t, [v]),
        "nudim_estimate_bytes": (i, [C.POINTER(Params), C.POINTER(C.c_size_t)]),
        "nudim_autotune": (i, [C.POINTER(Params), C.POINTER(i), C.POINTER(i)]),
        "nudim_synchronize": (i, [v]),
        "nudim_stream_probe": (i, [i, C.c_size_t, i, i, C.POINTER(d)]),
        "nudim_profile": (i, [v, i]),
        "nudim_profile_read": (i, [v, C.POINTER(d), C.POINTER(C.c_long), C.POINTER(d)]),
        "nudim_nfft_trafo": (i, [i, i, i, dp, i, i, i, dp, dp]),
        "nudim_nfft_adjoint": (i, [i, i, i, dp, i, i, i, dp, dp]),
        "nudim_cgnr": (i, [i, i, i, i, dp, dp, i, i, i, dp, dp, i]),
        "nudim_zfft": (i, [i, i, i, dp, i]),
        "nudim_tv": (i, [i, dp, i, i, i, f, i]),
        "nudim_tv_height": (i, [i, dp, i, i, i, f, i, i, i]),
        "nudim_tv_height_noalloc": (i, [i, v, i, i, i, f, i, i, i, v]),
        "nudim_diffusion2d": (i, [i, dp, i, i, i, f, f, i]),
        "nudim_constraint": (i, [i, dp, dp, i, i, i, i, i, i, i]),
        "nudim_projection_ft": (i, [i, i, i, i, i, dp, dp]),
        "nudim_project_2d1d": (i, [i, i, i, i, d, d, d, dp, dp]),
        "nudim_server_create": (i, [C.POINTER(ServerOpts), C.POINTER(v)]),
        "nudim_server_submit": (i, [v, C.POINTER(Params), dp, dp, dp, C.POINTER(C.c_longlong)]),
        "nudim_server_next": (i, [v, i, C.POINTER(JobReport)]),
        "nudim_server_stats": (i, [v, C.POINTER(i), C.POINTER(i), C.POINTER(C.c_longlong)]),
        "nudim_server_destroy": (None, [v]),
        "nudim_batch_create": (i, [C.POINTER(Params), i, C.POINTER(v)]),
        "nudim_batch_load_projections": (i, [v, i, dp]),
        "nudim_batch_get_volume": (i, [v, i, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc != 0:
        raise NudimError("%s (code %d)" % (lib().nudim_last_error().decode(errors="replace"), rc))


def _c128(a):
    return np.ascontiguousarray(a, np.complex128)


def _f64(a):
    return np.ascontiguousarray(a, np.float64)


def make_params(nx, nz, np_, gam_min, gam_max, interval, padangle=1.0, weight=True, random=False,
                seed=0, height=None, center=None, denoise=False, method=TV, nhio=0, ner=20,
                precision=FP64_PLAIN, device=0, nfft_n=0, nfft_w=0, cg_form=CG_AUTO,
                use_graph=False, tv_full=False, diffusion_lambda=0.0, cu_mask=None):
    """Parameters with the meaning of reconstruction's flags (src/reconstruction.c:91-112)."""
    p = Params()
    lib().nudim_default_params(C.byref(p))
    p.nx, p.nz, p.np = nx, nz, np_
    p.gam_min, p.gam_max, p.interval, p.padangle = gam_min, gam_max, interval, padangle
    p.weight, p.random, p.seed = int(weight), int(random), seed
    ch = height is not None or center is not None
    p.config_height = int(ch)
    p.nheight = -1 if height is None else height
    p.ncenter = -1 if center is None else center
    p.denoise, p.method = int(denoise), method
    p.nhio, p.ner = nhio, ner
    p.precision, p.device = precision, device
    p.nfft_n, p.nfft_w, p.cg_form, p.use_graph = nfft_n, nfft_w, cg_form, int(use_graph)
    p.tv_full, p.diffusion_lambda = int(tv_full), diffusion_lambda
    if cu_mask:
        for i, w in enumerate(cu_mask):
            p.cu_mask[i] = w
    return p


def estimate_bytes(p):
    b = C.c_size_t()
    _check(lib().nudim_estimate_bytes(C.byref(p), C.byref(b)))
    return b.value


def autotune(p):
    """(n, W) of the job's NFFT: measured on the device for FP32 jobs without
    explicit sizes (once per process and node set), else the fixed choice."""
    n, w = C.c_int(), C.c_int()
    _check(lib().nudim_autotune(C.byref(p), C.byref(n), C.byref(w)))
    return n.value, w.value


class Reconstruction:
    """One reconstruction job on the device (the numerics of reconstruction.c's main)."""

    nvol = 1

    def __init__(self, params):
        self.p = params
        h = C.c_void_p()
        self._create(params, h)
        self._h = h
        nm, M, na = C.c_int(), C.c_int(), C.c_int()
        _check(lib().nudim_sizes(h, C.byref(nm), C.byref(M), C.byref(na)))
        self.nmeas, self.M, self.nang = nm.value, M.value, na.value
        self.shape = (params.nz, params.np, params.np)
        self.nn = 0

    def _create(self, params, h):
        _check(lib().nudim_create(C.byref(params), C.byref(h)))

    def close(self):
        if getattr(self, "_h", None):
            lib().nudim_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def load_projections(self, proj):
        proj = _f64(proj)
        assert proj.shape == (self.nmeas, self.p.nz, self.p.nx), proj.shape
        _check(lib().nudim_load_projections(self._h, _ptr(proj)))

    def load_samples(self, slices, kind=None):
        s = _c128(slices)
        assert s.shape == (self.p.nz, self.M), s.shape
        k = None if kind is None else np.ascontiguousarray(kind, np.int32)
        _check(lib().nudim_load_samples(self._h, _ptr(s), _ptr(k)))

    def initial(self):
        _check(lib().nudim_initial(self._h))
        self.nn = 0

    def iterate(self, count=1):
        r = np.zeros(max(count, 1))
        _check(lib().nudim_iterate(self._h, count, _ptr(r)))
        self.nn += count
        return r[:count]

    def volume(self):
        out = np.zeros(self.shape)
        _check(lib().nudim_get_volume(self._h, _ptr(out)))
        return out

    def set_volume(self, img, nn=0):
        img = _f64(img)
        assert img.shape == self.shape
        _check(lib().nudim_set_volume(self._h, _ptr(img), nn))
        self.nn = nn

    def set_dump(self, nn, before, after):
        self._dump = (before, after)
        _check(lib().nudim_set_dump(self._h, nn, _ptr(before), _ptr(after)))

    def device_bytes(self):
        return lib().nudim_device_bytes(self._h)

    def profile(self, kernel_class):
        """time one kernel class with HIP events during iterate() (0 = off)"""
        _check(lib().nudim_profile(self._h, kernel_class))

    def profile_read(self):
        ms, n, b = C.c_double(), C.c_long(), C.c_double()
        _check(lib().nudim_profile_read(self._h, C.byref(ms), C.byref(n), C.byref(b)))
        return ms.value, n.value, b.value

    def synchronize(self):
        _check(lib().nudim_synchronize(self._h))


class Batch(Reconstruction):
    """nvol volumes of one geometry reconstructed together
    (nudim_batch_create): load each volume's projections, then initial() and
    iterate() run the whole batch; iterate returns R as [count][nvol]."""

    def __init__(self, params, nvol):
        self.nvol = int(nvol)
        super().__init__(params)

    def _create(self, params, h):
        _check(lib().nudim_batch_create(C.byref(params), self.nvol, C.byref(h)))

    def load_projections(self, vol, proj):
        proj = _f64(proj)
        assert proj.shape == (self.nmeas, self.p.nz, self.p.nx), proj.shape
        _check(lib().nudim_batch_load_projections(self._h, vol, _ptr(proj)))

    def iterate(self, count=1):
        r = np.zeros((max(count, 1), self.nvol))
        _check(lib().nudim_iterate(self._h, count, _ptr(r)))
        self.nn += count
        return r[:count]

    def volume(self, vol):
        out = np.zeros(self.shape)
        _check(lib().nudim_batch_get_volume(self._h, vol, _ptr(out)))
        return out


def stream_probe(nbytes, kind, reps=5, device=0):
    """GB/s of the library's plain streaming kernel over nbytes of device
    memory: kind 0 read only, 1 copy (bytes read + written)"""
    g = C.c_double()
    _check(lib().nudim_stream_probe(device, nbytes, kind, reps, C.byref(g)))
    return g.value


# ---- stateless parity entry points ------------------------------------------

def nfft_trafo(x, fhat, n=0, W=14, precision=FP64_PLAIN):
    """fhat [P][N][N] -> f [P][M] (nfft_trafo, NFFT3 sign)."""
    fhat = _c128(fhat)
    if fhat.ndim == 2:
        fhat = fhat[None]
    P, N, _ = fhat.shape
    x = _f64(x)
    M = x.shape[0]
    n = n or 2 * _np2(N)
    f = np.zeros((P, M), np.complex128)
    _check(lib().nudim_nfft_trafo(precision, N, M, _ptr(x), n, W, P, _ptr(fhat), _ptr(f)))
    return f


def nfft_adjoint(x, f, N, n=0, W=14, precision=FP64_PLAIN):
    """f [P][M] -> fhat [P][N][N] (nfft_adjoint)."""
    f = _c128(f)
    if f.ndim == 1:
        f = f[None]
    P, M = f.shape
    x = _f64(x)
    n = n or 2 * _np2(N)
    fh = np.zeros((P, N, N), np.complex128)
    _check(lib().nudim_nfft_adjoint(precision, N, M, _ptr(x), n, W, P, _ptr(f), _ptr(fh)))
    return fh


def cgnr(x, y, fhat0, w=None, nsteps=10, n=0, W=14, precision=FP64_PLAIN, form=CG_AUTO):
    """NFFT3 CGNR: y [P][M], fhat0 [P][N][N] -> fhat [P][N][N]."""
    y = _c128(y)
    fh = np.array(fhat0, np.complex128, copy=True, order="C")
    if y.ndim == 1:
        y, fh = y[None], fh[None]
    P, N, _ = fh.shape
    x = _f64(x)
    M = x.shape[0]
    n = n or 2 * _np2(N)
    wv = None if w is None else _f64(w)
    _check(lib().nudim_cgnr(precision, form, N, M, _ptr(x), _ptr(wv), n, W, P, _ptr(y), _ptr(fh), nsteps))
    return fh


def zfft(data, sign, precision=FP64_PLAIN):
    """unnormalised DFT along axis 0 (FFTW sign), returns a new array"""
    d = np.array(data, np.complex128, copy=True, order="C")
    _check(lib().nudim_zfft(precision, d.shape[0], d.size // d.shape[0], _ptr(d), sign))
    return d


def tv(u, lam=0.2, niter=40, precision=FP64_PLAIN):
    v = np.array(u, np.float64, copy=True, order="C")
    _check(lib().nudim_tv(precision, _ptr(v), *v.shape, lam, niter))
    return v


def tv_height(u, nheight, ncenter, lam=0.2, niter=40, precision=FP64_PLAIN):
    v = np.array(u, np.float64, copy=True, order="C")
    _check(lib().nudim_tv_height(precision, _ptr(v), *v.shape, lam, niter, nheight, ncenter))
    return v


def tv_height_noalloc(dev_ptr, shape, nheight, ncenter, lam=0.2, niter=40, precision=FP64_PLAIN, stream=None):
    """nudim_tv_height_noalloc: TV-height in place on a device array of
    float64 [nz][ny][nx] at address dev_ptr, enqueued on `stream` (a
    hipStream_t address; None: the thread's default stream).  The caller
    synchronises the stream before reading the result."""
    nz, ny, nx = shape
    _check(lib().nudim_tv_height_noalloc(precision, C.c_void_p(dev_ptr), nz, ny, nx, lam, niter, nheight, ncenter,
                                         None if stream is None else C.c_void_p(stream)))


def diffusion2d(u, lam=0.25, kappa=0.005, niter=80, precision=FP64_PLAIN):
    v = np.array(u, np.float64, copy=True, order="C")
    _check(lib().nudim_diffusion2d(precision, _ptr(v), *v.shape, lam, kappa, niter))
    return v


def constraint(img, img_last, nx, is_hio, height=None, center=None, precision=FP64_PLAIN):
    v = np.array(img, np.float64, copy=True, order="C")
    last = _f64(img_last)
    nz, np_, _ = v.shape
    ch = height is not None or center is not None
    _check(lib().nudim_constraint(precision, _ptr(v), _ptr(last), nz, np_, nx, int(is_hio), int(ch),
                                  -1 if height is None else height, -1 if center is None else center))
    return v


def projection_ft(proj, np_):
    proj = _f64(proj)
    nmeas, nz, nx = proj.shape
    tomo = np.zeros((nmeas, nz, np_), np.complex128)
    _check(lib().nudim_projection_ft(FP64_PLAIN, nx, nz, np_, nmeas, _ptr(proj), _ptr(tomo)))
    return tomo


def project_2d1d(density, np_, gam_min, gam_max, interval):
    density = _f64(density)
    nz, ny, nx = density.shape
    ns = int((gam_max - gam_min) / interval + 0.5) + 1
    out = np.zeros((ns, nz, np_))
    _check(lib().nudim_project_2d1d(nx, ny, nz, np_, gam_min, gam_max, interval, _ptr(density), _ptr(out)))
    return out


def _np2(n):
    p = 1
    while p < n:
        p <<= 1
    return p


class Server:
    """Continuous-batching job server (nudim_server_*): jobs of mixed size are
    admitted under an HBM budget (least work first with an aging bound, or
    arrival order) and run concurrently, one stream and host slot thread
    each; pending jobs with the same parameters run together as one batch of
    up to max_stack volumes (1: every job alone)."""

    def __init__(self, max_active=4, hbm_budget=0, device=0, policy=SERVE_SHORTEST, max_wait_s=30.0,
                 small_cus=0, small_work=0.0, max_stack=0):
        o = ServerOpts(hbm_budget, max_active, device, policy, max_wait_s, small_cus, small_work, max_stack)
        h = C.c_void_p()
        _check(lib().nudim_server_create(C.byref(o), C.byref(h)))
        self._h = h
        self._jobs = {}

    def submit(self, params, proj):
        """Queue a job; returns its id.  The volume and R trace are returned
        by next() when it finishes."""
        n = params.nz * params.np * params.np
        vol = np.zeros(n)
        rf = np.zeros(max(params.nhio + params.ner, 1))
        proj = _f64(proj)
        jid = C.c_longlong()
        _check(lib().nudim_server_submit(self._h, C.byref(params), _ptr(proj), _ptr(vol), _ptr(rf), C.byref(jid)))
        self._jobs[jid.value] = (vol, rf, params)
        return jid.value

    def next(self, timeout_ms=-1):
        """(report dict, volume [nz][np][np], R trace) of the next finished
        job, or None after timeout_ms without one."""
        r = JobReport()
        rc = lib().nudim_server_next(self._h, timeout_ms, C.byref(r))
        if rc == EAGAIN:
            return None
        _check(rc)
        vol, rf, p = self._jobs.pop(r.job_id)
        rep = {"job_id": r.job_id, "status": r.status, "t_submit": r.t_submit, "t_start": r.t_start,
               "t_end": r.t_end, "device_bytes": r.device_bytes, "error": r.error.decode(errors="replace"),
               "stacked": r.stacked}
        return rep, vol.reshape(p.nz, p.np, p.np), rf[: p.nhio + p.ner]

    def stats(self):
        pend, act, used = C.c_int(), C.c_int(), C.c_longlong()
        _check(lib().nudim_server_stats(self._h, C.byref(pend), C.byref(act), C.byref(used)))
        return {"pending": pend.value, "active": act.value, "bytes_in_use": used.value}

    def close(self):
        if self._h:
            lib().nudim_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
