#!/usr/bin/env python
"""Throughput of NUDIM's HIO/ER iteration on one MI355X.

Workload (default): the 256^3 blob phantom of the reconstruction config
"c3" -- 800 blobs in a 128-voxel slab, tilt series of 71 projections at
-70..70 step 2 with np=384, SNR 2 -- reconstructed on the 384x384x256 grid
with --padangle=1 --weight --denoise --method=TV --height=128 --center=192.
One step is one pass of the HIO/ER loop body over the whole volume
(z FFT, per-plane projector + 10-step CGNR for all 256 planes, z IFFT,
40 TV sweeps, density constraint), inputs resident in HBM.

value  = passes/s of the FP32 batched form (NUDIM_FP32).
Beside it: passes/s of the plain FP64 path (NUDIM_FP64_PLAIN: literal
NFFT3-order CGNR with the reference's NFFT sizes), their ratio, the relative
L2 difference between the two after the same passes, the dominant kernel's
roofline figure (HIP events over the timed region) and a CPU baseline (the
FP64 C restatement on the host cores, bounded sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import nudim_amd as nd  # noqa: E402
from nudim_amd import phantom  # noqa: E402

HBM_PEAK_GBS = 8000.0
KC_NAMES = {1: "k_fft_lines<float> (Toeplitz column pass)", 2: "k_interp<float>", 3: "k_gather<float>",
            4: "k_fft_lines<float> (row passes)", 5: "k_tv_grad/k_tv_update<float>", 6: "k_zfft<float>"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def make_job(prm, proj, precision, **kw):
    p = nd.make_params(prm["nx"], prm["nz"], prm["np_"], prm["gam_min"], prm["gam_max"], prm["interval"],
                       padangle=prm["padangle"], weight=prm["weight"], nhio=prm["nhio"], ner=prm["ner"],
                       denoise=prm.get("denoise", False), method=nd.TV, height=prm.get("height"),
                       center=prm.get("center"), precision=precision, **kw)
    t0 = time.perf_counter()
    job = nd.Reconstruction(p)
    job.load_projections(proj)
    job.initial()
    job.synchronize()
    log("  set-up + initial INFFT %.2f s, device memory %.2f GB" % (time.perf_counter() - t0,
                                                                   job.device_bytes() / 1e9))
    return job


def timed(job, steps, kclass=0):
    if kclass:
        job.profile(kclass)
    job.synchronize()
    t0 = time.perf_counter()
    r = job.iterate(steps)
    job.synchronize()
    dt = time.perf_counter() - t0
    prof = job.profile_read() if kclass else None
    if kclass:
        job.profile(0)
    return dt, r, prof


def cpu_baseline(prm, proj, seconds_budget=30.0):
    """FP64 C restatement (oracle) on the host: Fourier step of a few planes,
    full-volume TV and z FFTs, scaled to one pass of the whole volume."""
    from oracle import oracle
    threads = min(os.cpu_count() or 1, 16)
    oracle.set_threads(threads)
    op = oracle.params(prm["nx"], prm["nz"], prm["np_"], prm["gam_min"], prm["gam_max"], prm["interval"],
                       padangle=prm["padangle"], weight=prm["weight"], denoise=1, method=1,
                       height=prm.get("height"), center=prm.get("center"), nhio=prm["nhio"], ner=prm["ner"],
                       nthreads=threads)
    nz, np_ = prm["nz"], prm["np_"]
    t0 = time.perf_counter()
    slices, kind = oracle.assemble(op, proj)
    t_setup = time.perf_counter() - t0
    rng = np.random.default_rng(0)
    img = np.abs(rng.standard_normal((nz, np_, np_))) * 0.1
    nplanes = threads
    t_planes = oracle.time_fourier_planes(op, img, slices, kind, 0, nplanes)
    t_fourier = t_planes * nz / nplanes
    data = img.reshape(nz, -1).astype(np.complex128)
    t1 = time.perf_counter()
    oracle.fft_strided(data, -1)
    t_zfft = time.perf_counter() - t1
    t1 = time.perf_counter()
    sweeps = 4
    oracle.tv_height(img, prm.get("height", np_), prm.get("center", np_ // 2), niter=sweeps)
    t_tv = (time.perf_counter() - t1) * 40 / sweeps
    t_iter = t_fourier + 2 * t_zfft + t_tv
    log("  cpu: setup %.1fs, %d planes %.2fs -> Fourier step %.1fs, z-FFT %.2fs, TV(40) %.1fs -> %.1f s/pass"
        % (t_setup, nplanes, t_planes, t_fourier, t_zfft, t_tv, t_iter))
    return {"value": 1.0 / t_iter, "unit": "volume-iter/s", "cores": threads, "kind": "port",
            "sample": "FP64 C restatement (-O3, OpenMP over planes, %d threads): per-plane projector + "
                      "10-step NFFT CGNR of %d of %d planes (scaled x%d), 4 TV sweeps (scaled x10), "
                      "one z FFT (x2); sample %.1f s" % (threads, nplanes, nz, nz // nplanes,
                                                          t_planes + t_zfft + t_tv / 10)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--kclass", type=int, default=nd.KC_TOEPLITZ_COL)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    t0 = time.perf_counter()
    prm, proj = phantom.config_workload(args.config)
    log("workload %s: projections %s generated in %.1f s" % (args.config, proj.shape, time.perf_counter() - t0))

    log("FP32 batched form:")
    j32 = make_job(prm, proj, nd.FP32)
    j32.iterate(args.warmup)
    dt32, r32, _ = timed(j32, args.steps)
    rate32 = args.steps / dt32
    log("  %d passes in %.3f s -> %.2f passes/s" % (args.steps, dt32, rate32))
    dtk, _, prof = timed(j32, args.steps, args.kclass)
    kms, klaunch, kbytes = prof
    avg_us = 1e3 * kms / max(klaunch, 1)
    achieved = kbytes / max(klaunch, 1) / (avg_us * 1e-6) / 1e9 if klaunch else 0.0
    log("  %s: %d launches, avg %.1f us, %.1f GB/s algorithmic" % (KC_NAMES.get(args.kclass), klaunch, avg_us,
                                                                 achieved))

    rate64 = None
    acc = {}
    if not args.no_plain:
        log("FP64 plain path:")
        j64 = make_job(prm, proj, nd.FP64_PLAIN)
        r64w = j64.iterate(1)
        dt64, r64, _ = timed(j64, args.plain_steps)
        r64 = np.concatenate([r64w, r64])
        rate64 = args.plain_steps / dt64
        log("  %d passes in %.3f s -> %.3f passes/s" % (args.plain_steps, dt64, rate64))
        if not args.no_accuracy:
            # same number of passes from each path's own initial INFFT
            npass = 1 + args.plain_steps
            v64 = j64.volume()
            j32.initial()
            v0 = j32.volume()
            r32a = j32.iterate(npass)
            v32 = j32.volume()
            j64.initial()
            v0_64 = j64.volume()
            acc = {"rel_l2_initial": rel(v0, v0_64), "rel_l2_after_passes": rel(v32, v64),
                   "passes_compared": npass,
                   "max_abs_rfactor_drift": float(np.max(np.abs(r32a - r64))),
                   "rfactor_fp64": [float(v) for v in r64], "rfactor_fp32": [float(v) for v in r32a]}
            log("  rel L2 FP32 vs FP64: initial %.2e, after %d passes %.2e" % (acc["rel_l2_initial"], npass,
                                                                              acc["rel_l2_after_passes"]))
        j64.close()
    j32.close()

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(prm, proj)
        except Exception as e:  # the baseline is reported, never required
            log("cpu baseline skipped: %s" % e)

    out = {
        "metric": "ER/HIO passes/s at 256^3 (c3), FP32 batched form",
        "value": rate32,
        "unit": "volume-iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 / rate32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "c3: 256^3 blob phantom, np=384, nz=256, 71 tilts -70..70 step 2, SNR 2, "
                               "padangle 1, weight, TV height 128 centre 192; one HIO/ER pass",
                   "grid": [prm["nz"], prm["np_"], prm["np_"]]},
        "fp64_plain_passes_per_s": rate64,
        "speedup_vs_fp64_plain": (rate32 / rate64) if rate64 else None,
        "accuracy": acc or None,
        "roofline": {"kernel": KC_NAMES.get(args.kclass), "bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "avg_us": avg_us, "launches": klaunch},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
