"""Shared helpers of the tests (no GPU, no oracle)."""
import numpy as np


def rel(a, b):
    """relative L2 difference ||a - b|| / ||b||"""
    a = np.asarray(a)
    b = np.asarray(b)
    d = np.linalg.norm((a - b).ravel())
    n = np.linalg.norm(b.ravel())
    return d / n if n > 0 else d


def blob_phantom(shape, nblobs, rng, sig=(2.0, 6.0), amp=(0.5, 1.0), slab=None):
    """Sum of isotropic Gaussian blobs on a [nz][ny][nx] grid; with `slab` the
    centres lie in a band of that height along y (centred at ny/2) and the
    density outside the band is zeroed."""
    nz, ny, nx = shape
    z = np.arange(nz)[:, None, None]
    y = np.arange(ny)[None, :, None]
    x = np.arange(nx)[None, None, :]
    out = np.zeros(shape)
    for _ in range(nblobs):
        s = rng.uniform(*sig)
        a = rng.uniform(*amp)
        m = min(2 * s, min(shape) / 4)
        cz = rng.uniform(m, nz - m)
        cx = rng.uniform(m, nx - m)
        if slab is None:
            cy = rng.uniform(m, ny - m)
        else:
            lo, hi = ny / 2 - slab / 2, ny / 2 + slab / 2
            cy = rng.uniform(lo + min(m, slab / 4), hi - min(m, slab / 4))
        out += a * np.exp(-((z - cz) ** 2 + (y - cy) ** 2 + (x - cx) ** 2) / (2 * s * s))
    if slab is not None:
        lo, hi = int(np.ceil(ny / 2 - slab / 2)), int(np.floor(ny / 2 + slab / 2))
        out[:, :lo, :] = 0
        out[:, hi + 1:, :] = 0
    return out
