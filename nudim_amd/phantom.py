"""Synthetic blob phantoms and their tilt series (bench and serving workloads).

A phantom is a sum of isotropic 3D Gaussian blobs (sigma ~ U(2,6) voxels,
amplitude ~ U(0.5,1)) whose centres lie in a slab of height h along y,
centred at ny/2.  Its tilt series is computed analytically: the line
integral of A exp(-|p-c|^2 / (2 s^2)) along a ray is A sqrt(2 pi) s times a 2D
Gaussian in the detector plane.  The detector geometry is that of
projection_sim -m 2D1D (src/projection_sim.c:282-405): rotation about z, the
tilt at angle g projects the (y, x) plane onto the detector column
c = np/2 + (x - nx/2) cos g + (y - ny/2) sin g, with g in the reference's
degrees-to-radians convention (PI = 3.1415926, src/grid.h:11).

Noise: Gaussian per projection with sigma_noise = std(projection) / sqrt(SNR)
(SNR as a variance ratio; the reference's --snr is a standard-deviation
ratio, src/projection_sim.c:58-62).
"""
import numpy as np

REF_PI = 3.1415926


def blob_params(nblobs, shape, slab, seed, sig=(2.0, 6.0), amp=(0.5, 1.0)):
    nz, ny, nx = shape
    rng = np.random.Generator(np.random.PCG64(seed))
    s = rng.uniform(*sig, nblobs)
    a = rng.uniform(*amp, nblobs)
    m = np.minimum(2 * s, min(shape) / 4)
    cz = rng.uniform(m, nz - m)
    cx = rng.uniform(m, nx - m)
    lo, hi = ny / 2 - slab / 2, ny / 2 + slab / 2
    mm = np.minimum(m, slab / 4)
    cy = rng.uniform(lo + mm, hi - mm)
    return dict(s=s, a=a, cz=cz, cy=cy, cx=cx, shape=shape, slab=slab)


def density(bp):
    """voxelised phantom [nz][ny][nx] (density outside the slab zeroed)"""
    nz, ny, nx = bp["shape"]
    z = np.arange(nz)[:, None, None]
    y = np.arange(ny)[None, :, None]
    x = np.arange(nx)[None, None, :]
    out = np.zeros((nz, ny, nx))
    for s, a, cz, cy, cx in zip(bp["s"], bp["a"], bp["cz"], bp["cy"], bp["cx"]):
        out += a * np.exp(-((z - cz) ** 2 + (y - cy) ** 2 + (x - cx) ** 2) / (2 * s * s))
    slab = bp["slab"]
    lo, hi = int(np.ceil(ny / 2 - slab / 2)), int(np.floor(ny / 2 + slab / 2))
    out[:, :lo, :] = 0
    out[:, hi + 1:, :] = 0
    return out


def tilt_angles(gam_min, gam_max, interval):
    n = int((gam_max - gam_min) / interval + 0.5) + 1
    ang = gam_min - interval
    out = []
    for _ in range(n):
        ang += interval
        out.append(ang)
    return np.array(out)


def projections(bp, np_, gam_min, gam_max, interval):
    """analytic tilt series [nmeas][nz][np] of the blob phantom"""
    nz, ny, nx = bp["shape"]
    z = np.arange(nz)[None, :]
    c = np.arange(np_)[None, :]
    s, a = bp["s"][:, None], bp["a"][:, None]
    Z = a * np.sqrt(2 * np.pi) * s * np.exp(-((z - bp["cz"][:, None]) ** 2) / (2 * s * s))  # [b][nz]
    out = []
    for g in tilt_angles(gam_min, gam_max, interval):
        t = g / 180 * REF_PI
        sc = np_ / 2 + (bp["cx"] - nx / 2) * np.cos(t) + (bp["cy"] - ny / 2) * np.sin(t)
        S = np.exp(-((c - sc[:, None]) ** 2) / (2 * s * s))  # [b][np]
        out.append(Z.T @ S)
    return np.stack(out)


def add_noise(proj, snr, seed):
    """Gaussian noise per projection at a variance signal-to-noise ratio `snr`."""
    if not snr:
        return proj
    rng = np.random.Generator(np.random.PCG64(seed))
    out = np.empty_like(proj)
    for j in range(proj.shape[0]):
        sd = proj[j].std()
        out[j] = proj[j] + rng.standard_normal(proj[j].shape) * (sd / np.sqrt(snr))
    return out


def config_workload(cfg, seed=None):
    """(params dict, projections) of the named reconstruction configs.

    c1: 128^3, 200 blobs, slab 64, np=128, -60..60 step 3, no noise, ner=20, no denoise.
    c2: c1 phantom, SNR=1, nhio=20 + ner=20, TV with height 64 (centre 64).
    c3: 256^3, 800 blobs, slab 128, np=384, -70..70 step 2, SNR=2, nhio=30 + ner=50,
        TV with height 128 (centre 192); reconstruction grid 384 x 384 x 256.
    s128, s256, s512: the serving mix (SERVE_CLASSES).
    """
    if cfg == "c1":
        bp = blob_params(200, (128, 128, 128), 64, 1000 if seed is None else seed)
        proj = projections(bp, 128, -60, 60, 3)
        prm = dict(nx=128, nz=128, np_=128, gam_min=-60, gam_max=60, interval=3, padangle=1.0, weight=True,
                   nhio=0, ner=20)
    elif cfg == "c2":
        bp = blob_params(200, (128, 128, 128), 64, 2000 if seed is None else seed)
        proj = add_noise(projections(bp, 128, -60, 60, 3), 1.0, 2001)
        prm = dict(nx=128, nz=128, np_=128, gam_min=-60, gam_max=60, interval=3, padangle=1.0, weight=True,
                   nhio=20, ner=20, denoise=True, height=64, center=64)
    elif cfg == "c3":
        bp = blob_params(800, (256, 256, 256), 128, 3000 if seed is None else seed)
        proj = add_noise(projections(bp, 384, -70, 70, 2), 2.0, 3001)
        prm = dict(nx=384, nz=256, np_=384, gam_min=-70, gam_max=70, interval=2, padangle=1.0, weight=True,
                   nhio=30, ner=50, denoise=True, height=128, center=192)
    elif cfg in SERVE_CLASSES:
        n, np_, nblobs, g, t, nmax = SERVE_CLASSES[cfg]
        bp = blob_params(nblobs, (n, n, n), n // 2, 5000 + n if seed is None else seed)
        proj = add_noise(projections(bp, np_, -g, g, t), 2.0, (5001 + n if seed is None else seed + 1))
        prm = dict(nx=np_, nz=n, np_=np_, gam_min=-g, gam_max=g, interval=t, padangle=1.0, weight=True,
                   nhio=20, ner=40, denoise=True, height=n // 2, center=np_ // 2)
    else:
        raise ValueError(cfg)
    return prm, proj


# Job classes of the serving mix (128^3, 256^3, 512^3 blob phantoms, np = 1.5 n,
# SNR 2, nhio=20 + ner=40 with TV over the slab): name -> (n, np, blobs,
# tilt range +-g, step, -)
SERVE_CLASSES = {
    "s128": (128, 192, 200, 60, 3, 0),
    "s256": (256, 384, 800, 70, 2, 0),
    "s512": (512, 768, 3200, 60, 1, 0),
}
SERVE_MIX = (("s128", 0.50), ("s256", 0.35), ("s512", 0.15))


def batch_instances(count=48, seed=4000):
    """The uneven batch: instance i (seed 4000 + i) has nx = nz from
    {64, 96, 128, 192, 256}, np = nx * {1, 1.25, 1.5}, tilts +-{50, 60, 70} at
    step 2 or 3, a blob phantom (slab nx/2, blobs ~ 200 (nx/128)^3), noise at
    SNR 1 or 4, ner = 30 with TV over the slab.  Returns [(params dict,
    projections)]."""
    out = []
    for i in range(count):
        rng = np.random.Generator(np.random.PCG64(seed + i))
        nx = int(rng.choice([64, 96, 128, 192, 256]))
        np_ = int(round(nx * rng.choice([1.0, 1.25, 1.5]) / 2)) * 2
        g = int(rng.choice([50, 60, 70]))
        t = int(rng.choice([2, 3]))
        snr = float(rng.choice([1.0, 4.0]))
        nb = max(8, int(200 * (nx / 128) ** 3))
        bp = blob_params(nb, (nx, nx, nx), nx // 2, seed + i)
        proj = add_noise(projections(bp, np_, -g, g, t), snr, seed + 1000 + i)
        prm = dict(nx=np_, nz=nx, np_=np_, gam_min=-g, gam_max=g, interval=t, padangle=1.0, weight=True,
                   nhio=0, ner=30, denoise=True, height=nx // 2, center=np_ // 2)
        out.append((prm, proj))
    return out
