"""nudim_tv_height_noalloc (TV-height on caller device memory, scratch in a
file-scope device array) against nudim_tv_height on the same volumes.

Both run the same sweep kernels with the same partial-sum layout, so the
results must be equal bit for bit.  Two calls in a row on each precision,
the second on a smaller volume, so stale scratch from the first call would
show; the first call runs on a stream the test creates, the second on the
thread's default stream.  Device memory comes from the HIP runtime the
library itself links (libamdhip64.so.7), through ctypes."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    return h


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed: %d" % (what, rc))


def _volume(shape, seed):
    rng = np.random.default_rng(seed)
    # mostly positive with some negative voxels, so dA > 0 and the sweeps move u
    return rng.standard_normal(shape) * 0.3 + 0.2


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_matches_tv_height(hip, precision):
    rt = _hip()
    stream = C.c_void_p()
    _ok(rt.hipStreamCreate(C.byref(stream)), "hipStreamCreate")
    cases = [((24, 48, 40), 20, 26, 7, 40, stream.value), ((14, 32, 24), 12, 15, 8, 12, None)]
    try:
        for shape, nheight, ncenter, seed, niter, st in cases:
            print("noalloc precision=%d shape=%s band=%d/%d stream=%s" % (precision, shape, nheight, ncenter,
                                                                         "own" if st else "default"), flush=True)
            u = _volume(shape, seed)
            ref = hip.tv_height(u, nheight, ncenter, niter=niter, precision=precision)
            d = C.c_void_p()
            _ok(rt.hipMalloc(C.byref(d), u.nbytes), "hipMalloc")
            try:
                _ok(rt.hipMemcpy(d, u.ctypes.data, u.nbytes, 1), "hipMemcpy H2D")
                hip.tv_height_noalloc(d.value, shape, nheight, ncenter, niter=niter, precision=precision, stream=st)
                _ok(rt.hipStreamSynchronize(C.c_void_p(st)) if st else rt.hipDeviceSynchronize(), "synchronize")
                got = np.empty_like(u)
                _ok(rt.hipMemcpy(got.ctypes.data, d, u.nbytes, 2), "hipMemcpy D2H")
            finally:
                rt.hipFree(d)
            assert np.isfinite(got).all()
            assert not np.array_equal(got, u)  # the sweeps changed the band
            assert np.array_equal(got, ref), np.abs(got - ref).max()
    finally:
        rt.hipStreamDestroy(stream)
