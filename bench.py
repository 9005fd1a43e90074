#!/usr/bin/env python
"""Throughput of NUDIM's HIO/ER iteration on one MI355X.

Workload (default): the 256^3 blob phantom of the reconstruction config
"c3" -- 800 blobs in a 128-voxel slab, tilt series of 71 projections at
-70..70 step 2 with np=384, SNR 2 -- reconstructed on the 384x384x256 grid
with --padangle=1 --weight --denoise --method=TV --height=128 --center=192.
One step is one pass of the HIO/ER loop body over the whole volume
(z FFT, per-plane projector + 10-step CGNR for all 256 planes, z IFFT,
40 TV sweeps, density constraint), inputs resident in HBM.

value  = passes/s of the FP32 batched form (NUDIM_FP32), direct launches.
Beside it: the same passes replayed as hipGraphs, passes/s of the plain FP64
path (NUDIM_FP64_PLAIN: literal NFFT3-order CGNR with the reference's NFFT
sizes), their ratio, the relative L2 difference between the two after the
same passes, the dominant kernel's roofline figure (HIP events over the
timed region), a CPU baseline (the FP64 C restatement on the host cores and
on one core, bounded samples), the uneven batch (48 volumes through the job
server) and the serving traces (100 Poisson arrivals of the mixed-size mix
at 4 jobs/s, a below-capacity trace, and a short trace of the 128^3 and
256^3 classes on both the FP32 form and the FP64 plain path).
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
# profiler classes of the library (nudim_profile) and the kernels they time;
# the template arguments follow the workload: the register FFT split of each
# length (fft2.hip), the NFFT window W, the z length
FFT2_SPLIT = {128: (16, 8), 160: (20, 8), 192: (16, 12), 240: (20, 12), 256: (16, 16), 288: (24, 12),
              320: (20, 16), 384: (24, 16), 480: (24, 20), 512: (32, 16), 576: (24, 24), 640: (32, 20),
              768: (32, 24), 1024: (32, 32), 1536: (48, 32)}
ZFFT2_SPLIT = {64: (8, 8), 96: (16, 6), 128: (16, 8), 192: (16, 12), 256: (16, 16), 512: (32, 16)}


def kc_names(np_, nz, n, W):
    a, b = FFT2_SPLIT.get(2 * np_, ("N1", "N2"))

    def toep(k, what):
        return "k_toep<float,%s,%s,%d> (%s)" % (a, b, k, what)

    grid = ("k_fft2<float,%d,%d> (NFFT grid passes, n=%d)" % (FFT2_SPLIT[n] + (n,)) if n in FFT2_SPLIT
            else "k_fft_lines<float> (NFFT grid passes, n=%d)" % n)
    zf = "k_zfft2<float,%d,%d> (z transforms)" % ZFFT2_SPLIT[nz] if nz in ZFFT2_SPLIT else "k_zfft_fwd/inv<float>"
    gather = ("k_gather_chunk<float,%d> (NFFT adjoint)" % W if nz % 64 == 0
              else "k_gather_w<float,%d> (NFFT adjoint)" % W)
    return {1: toep(1, "Toeplitz column pass, <p,Tp>"), 2: "k_interp_rsums<float,%d> (R sums at the known nodes)" % W, 3: gather, 4: grid,
            5: "k_tv_march<float>" if nz % 4 == 0 else "k_tv_fused<float>", 6: zf,
            7: toep(0, "Toeplitz row forward, p = z + beta p"), 8: toep(2, "Toeplitz row inverse, z -= alpha q"),
            9: toep(3, "two-input column pass, CG start"), 10: "k_cg_zupdate<float>"}


KC_CANDIDATES = (7, 1, 8, 5, 10, 9, 4, 2, 6, 3)
PROFILES = os.path.join(ROOT, "profiles")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


NFFT_CHOICE = {}


def make_job(prm, proj, precision, **kw):
    p = nd.make_params(prm["nx"], prm["nz"], prm["np_"], prm["gam_min"], prm["gam_max"], prm["interval"],
                       padangle=prm["padangle"], weight=prm["weight"], nhio=prm["nhio"], ner=prm["ner"],
                       denoise=prm.get("denoise", False), method=nd.TV, height=prm.get("height"),
                       center=prm.get("center"), precision=precision, **kw)
    t0 = time.perf_counter()
    n, w = nd.autotune(p)
    key = "f32" if precision == nd.FP32 else "f64"
    if key not in NFFT_CHOICE:
        NFFT_CHOICE[key] = {"n": n, "W": w, "choice_s": round(time.perf_counter() - t0, 3)}
        log("  NFFT n=%d, W=%d (%s, %.2f s)" % (n, w, "measured" if precision == nd.FP32 else "reference sizes",
                                                time.perf_counter() - t0))
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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(prm, proj, threads):
    """FP64 C restatement (the oracle) on `threads` host threads: Fourier
    step of a bounded sample of planes, TV sweeps and z FFTs, scaled to one
    pass of the whole volume (about 10-20 s of host work)."""
    from oracle import oracle
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
    nplanes = min(nz, 4 * threads)
    t_planes = oracle.time_fourier_planes(op, img, slices, kind, 0, nplanes)
    t_fourier = t_planes * nz / nplanes
    data = img.reshape(nz, -1).astype(np.complex128)
    t1 = time.perf_counter()
    oracle.fft_strided(data, -1)
    t_zfft = time.perf_counter() - t1
    t1 = time.perf_counter()
    sweeps = 8 if threads > 1 else 2
    oracle.tv_height(img, prm.get("height", np_), prm.get("center", np_ // 2), niter=sweeps)
    t_tv = (time.perf_counter() - t1) * 40 / sweeps
    t_iter = t_fourier + 2 * t_zfft + t_tv
    sample = t_planes + t_zfft + t_tv * sweeps / 40
    log("  cpu, %d threads: setup %.1fs, %d planes %.2fs -> Fourier step %.1fs, z-FFT %.2fs, TV(40) %.1fs -> "
        "%.1f s/pass" % (threads, t_setup, nplanes, t_planes, t_fourier, t_zfft, t_tv, t_iter))
    return {"value": 1.0 / t_iter, "unit": "volume-iter/s", "cores": threads, "kind": "port",
            "sample": "FP64 C restatement of reconstruction.c (-O3, OpenMP over planes, %d threads on %s): "
                      "projector + 10-step NFFT CGNR of %d of %d planes (scaled x%.0f), %d TV sweeps (scaled x%.0f), "
                      "one z FFT (x2); %.1f s of host work" % (threads, cpu_model(), nplanes, nz, nz / nplanes,
                                                               sweeps, 40 / sweeps, sample)}


def workload_name(cfg, prm):
    if cfg == "c3":
        return ("c3: 256^3 blob phantom, np=384, nz=256, 71 tilts -70..70 step 2, SNR 2, padangle 1, weight, "
                "TV height 128 centre 192; one HIO/ER pass")
    return ("%s: np=%d, nz=%d, tilts %g..%g step %g, padangle %g, TV height %s centre %s; one HIO/ER pass"
            % (cfg, prm["np_"], prm["nz"], prm["gam_min"], prm["gam_max"], prm["interval"], prm["padangle"],
               prm.get("height"), prm.get("center")))


def traffic_from_profiles(kclass):
    """HBM bytes per launch of the kernel from the committed counter summary
    (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs, scaled by a copy kernel
    of known size); None when no summary covers the kernel."""
    path = os.path.join(PROFILES, "pmc_traffic_c3.txt")
    try:
        with open(path) as fh:
            for line in fh:
                f = line.split()
                if len(f) >= 2 and f[0] == "class%d" % kclass:
                    return float(f[1])
    except OSError:
        pass
    return None


def cli_leg(args):
    """The drop-in `reconstruction` binary on the 256^3 workload's files:
    71 projection files of 384 x 256 values (%le) written into a scratch
    ./files, then the reference's command line (nhio 30 + ner 50, TV height
    128 centre 192) in FP32, with its own phase times (NUDIM_CLI_TIMES) at
    16 text-I/O threads and at 1.  The outputs are recovered-0.txt (%le, full
    384 x 384 x 256 grid) and recovered.txt (%f, the central 384^2 x 256
    crop): 37.7M lines each."""
    import shutil
    import subprocess
    import tempfile
    _, proj = phantom.config_workload("c3")
    exe = os.path.join(ROOT, "nudim_amd", "bin", "reconstruction")
    flags = ["--prefix=prj", "--pad=384", "--nx=384", "--nz=256", "--max=70", "--min=-70", "-t", "2",
             "--padangle=1", "--weight", "--nhio=30", "--ner=50", "--denoise", "--method=TV", "--height=128",
             "--center=192"]
    d = tempfile.mkdtemp(prefix="nudim_cli_")
    try:
        os.makedirs(os.path.join(d, "files"))
        t0 = time.perf_counter()
        for j in range(proj.shape[0]):
            v = proj[j].ravel()
            with open(os.path.join(d, "files", "prj-%d.txt" % j), "w") as fh:
                fh.write(("%le\n" * v.size) % tuple(v))
        gen_s = time.perf_counter() - t0
        runs = {}
        for threads in (16, 1):
            env = dict(os.environ, NUDIM_PRECISION="fp32", NUDIM_CLI_TIMES="1", NUDIM_IO_THREADS=str(threads))
            t0 = time.perf_counter()
            r = subprocess.run([exe] + flags, cwd=d, env=env, capture_output=True, text=True, timeout=900)
            wall = time.perf_counter() - t0
            if r.returncode != 0:
                raise RuntimeError("reconstruction failed: %s" % r.stderr[-500:])
            line = [ln for ln in r.stderr.splitlines() if ln.startswith("nudim-cli-times")][-1].split()
            ph = {line[i]: float(line[i + 1]) for i in range(1, len(line) - 1, 2) if line[i] != "io-threads"}
            ph["wall"] = wall
            runs["io_threads_%d" % threads] = ph
            log("  cli, %d I/O threads: %s" % (threads, " ".join("%s %.2f s" % kv for kv in ph.items())))
            for f in ("recovered-0.txt", "recovered.txt"):
                os.remove(os.path.join(d, f))
    finally:
        shutil.rmtree(d, ignore_errors=True)
    # the simulator (projection_sim -m 2D1D) at the same sizes: a 256^3
    # density (random: the cost does not depend on the values) to 71 tilts at
    # np = 384, FP64 as the reference
    dens = np.random.default_rng(0).random((256, 256, 256))
    nd.project_2d1d(dens, 384, -70, 70, 2)  # first call: device set-up
    t0 = time.perf_counter()
    nd.project_2d1d(dens, 384, -70, 70, 2)
    sim_s = time.perf_counter() - t0
    log("  projection_sim 2D1D, 256^3 -> 71 x 256 x 384 (FP64): %.3f s" % sim_s)
    return {"files": "71 x 98304 values (%le)", "generated_s": gen_s, "precision": "f32", "runs": runs,
            "projection_sim_2d1d_256cube_71tilts_s": sim_s}


def device_copy_gbs(mib=1024, reps=5):
    """Bandwidth of the runtime's device-to-device copy (hipMemcpy: its blit
    kernel, plain loads and stores) on this GPU now: bytes read + written
    over the best of reps copies of mib MiB -- the practical ceiling for a
    streaming kernel, beside the datasheet peak."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    nbytes = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), nbytes) or hip.hipMalloc(ctypes.byref(b), nbytes):
        raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, nbytes) or hip.hipDeviceSynchronize():
            raise RuntimeError("hipMemset failed")
        best = None
        for _ in range(reps):
            t0 = time.perf_counter()
            if hip.hipMemcpy(b, a, nbytes, 3) or hip.hipDeviceSynchronize():  # device to device
                raise RuntimeError("hipMemcpy failed")
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    return 2.0 * nbytes / best / 1e9


def probe(prm, proj, copy_mib):
    """One FP32 pass, one FP64 plain pass, one device-to-device hipMemcpy of
    copy_mib MiB (the runtime's blit kernel) and one pass each of the
    library's read-only and copy streaming kernels over copy_mib MiB (plain
    16-byte loads and stores): the program the counter runs profile, the
    three known-size kernels being the scale of FETCH_SIZE / WRITE_SIZE."""
    import ctypes
    for prec in (nd.FP32, nd.FP64_PLAIN):
        job = make_job(prm, proj, prec)
        job.iterate(1)
        job.synchronize()
        job.close()
    nbytes = copy_mib << 20
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), nbytes) or hip.hipMalloc(ctypes.byref(b), nbytes):
        raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, nbytes) or hip.hipDeviceSynchronize():
            raise RuntimeError("hipMemset failed")
        if hip.hipMemcpy(b, a, nbytes, 3) or hip.hipDeviceSynchronize():  # device to device
            raise RuntimeError("hipMemcpy failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    rd = nd.stream_probe(nbytes, 0, reps=1)
    cp = nd.stream_probe(nbytes, 1, reps=1)
    log("probe: one FP32 pass, one FP64 plain pass; hipMemcpy device to device of %d MiB; k_stream_read %.0f GB/s, "
        "k_stream_copy %.0f GB/s over %d MiB" % (copy_mib, rd, cp, copy_mib))


def serve(args, precision, jobs=None, rate=None, seed=None, classes=None):
    """Serving leg: a Poisson trace of `jobs` jobs at `rate` arrivals per
    second drawn from the serving mix (128^3 50 %, 256^3 35 %, 512^3 15 %;
    nhio=20 + ner=40 with TV; `classes` restricts the mix to those classes,
    renormalised), submitted at their arrival times to the
    continuous-batching server (admission by HBM budget, args.serve_slots
    jobs in flight).  Latency = finish - arrival."""
    jobs = args.serve_jobs if jobs is None else jobs
    rate = args.serve_rate if rate is None else rate
    seed = args.serve_seed if seed is None else seed
    rng = np.random.default_rng(seed)
    mix = [m for m in phantom.SERVE_MIX if classes is None or m[0] in classes]
    names = [m[0] for m in mix]
    frac = np.array([m[1] for m in mix])
    cls = rng.choice(len(names), jobs, p=frac / frac.sum())
    gaps = rng.exponential(1.0 / rate, jobs)
    arrivals = np.cumsum(gaps) - gaps[0]
    # a few distinct phantoms per class, cycled over the trace
    inputs = {}
    t0 = time.perf_counter()
    for c in sorted(set(cls)):
        inputs[c] = [phantom.config_workload(names[c], seed=7000 + 10 * c + k)
                     for k in range(1 if names[c] == "s512" else 2)]
    log("  serving: %d jobs (%s), inputs generated in %.1f s" % (
        jobs, ", ".join("%s x%d" % (names[c], int((cls == c).sum())) for c in sorted(set(cls))),
        time.perf_counter() - t0))
    # start-up calibration of each job class's NFFT sizing (nudim_autotune),
    # as a server does before it takes requests
    t0 = time.perf_counter()
    for c in sorted(set(cls)):
        prm = inputs[c][0][0]
        nd.autotune(nd.make_params(prm["nx"], prm["nz"], prm["np_"], prm["gam_min"], prm["gam_max"],
                                   prm["interval"], padangle=prm["padangle"], weight=prm["weight"],
                                   precision=precision))
    calib_s = time.perf_counter() - t0
    log("  serving: NFFT sizing measured for %d job classes in %.1f s" % (len(set(cls)), calib_s))
    reps = {}
    policy = nd.SERVE_FIFO if args.serve_fifo else nd.SERVE_SHORTEST
    # small jobs: the 128^3 class (np 192, 60 passes: 2.8e8 work units)
    with nd.Server(max_active=args.serve_slots, policy=policy, small_cus=args.serve_small_cus,
                   small_work=5e8, max_stack=args.serve_stack) as srv:
        meta = {}
        start = time.perf_counter()
        k = 0
        last_log = 0.0
        while k < jobs or len(reps) < jobs:
            now = time.perf_counter() - start
            if now - last_log >= 20.0:
                last_log = now
                log("    %.0f s: %d of %d jobs submitted, %d done" % (now, k, jobs, len(reps)))
            while k < jobs and arrivals[k] <= now:
                prm, proj = inputs[cls[k]][k % len(inputs[cls[k]])]
                p = nd.make_params(prm["nx"], prm["nz"], prm["np_"], prm["gam_min"], prm["gam_max"],
                                   prm["interval"], padangle=prm["padangle"], weight=prm["weight"],
                                   nhio=prm["nhio"], ner=prm["ner"], denoise=True, method=nd.TV,
                                   height=prm["height"], center=prm["center"], precision=precision)
                jid = srv.submit(p, proj)
                meta[jid] = (names[cls[k]], prm["nhio"] + prm["ner"])
                k += 1
            wait_ms = 50 if k >= jobs else max(1, int(1e3 * (arrivals[k] - now)))
            got = srv.next(timeout_ms=min(wait_ms, 50))
            if got is not None:
                rep = got[0]
                if rep["status"] != 0:
                    raise RuntimeError("serving job failed: %s" % rep["error"])
                reps[rep["job_id"]] = rep
    lat = np.array([r["t_end"] - r["t_submit"] for r in reps.values()])
    t_first = min(r["t_submit"] for r in reps.values())
    t_last = max(r["t_end"] for r in reps.values())
    span = t_last - t_first
    passes = sum(meta[j][1] for j in reps)
    by = {}
    for j, r in reps.items():
        by.setdefault(meta[j][0], []).append(r["t_end"] - r["t_submit"])
    by_n = {}
    for j in reps:
        by_n[meta[j][0]] = by_n.get(meta[j][0], 0) + 1
    out = {"jobs": len(reps), "jobs_by_class": by_n, "arrival_rate_jobs_per_s": rate, "slots": args.serve_slots,
           "offered_jobs_per_s": jobs / max(arrivals[-1], 1e-9) if jobs > 1 else None,
           "achieved_jobs_per_s": len(reps) / span, "volume_iters_per_s": passes / span,
           "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
           "mean_latency_s": float(lat.mean()), "makespan_s": span,
           "p50_latency_s_by_class": {k: float(np.percentile(v, 50)) for k, v in sorted(by.items())},
           "p99_latency_s_by_class": {k: float(np.percentile(v, 99)) for k, v in sorted(by.items())},
           "calibration_s": calib_s, "small_job_cus": args.serve_small_cus,
           "max_stack": args.serve_stack if args.serve_stack > 0 else 8,
           "jobs_run_in_batches": int(sum(1 for r in reps.values() if r["stacked"] > 1)),
           "largest_batch": int(max(r["stacked"] for r in reps.values())),
           "precision": "f32" if precision == nd.FP32 else "f64",
           "policy": "fifo" if args.serve_fifo else "shortest work first, 30 s aging",
           "mix": "%s; seed %d" % (", ".join("%s %.0f%%" % (n, 100 * f / frac.sum()) for n, f in mix), seed)}
    log("  serving %s: %d jobs in %.1f s: p50 %.2f s, p99 %.2f s, %.1f volume-iter/s, achieved %.2f jobs/s "
        "(offered %.2f)" % (out["precision"], len(reps), span, out["p50_latency_s"], out["p99_latency_s"],
                            out["volume_iters_per_s"], out["achieved_jobs_per_s"], out["offered_jobs_per_s"] or 0))
    return out


def small_configs():
    """The 128^3 configurations: c1 (noise-free ER passes without the
    denoiser) and c2 (SNR 1, HIO + ER with TV or DIFFUSION over the height
    band), passes/s of the FP32 form and the FP64 plain path, and the
    denoiser's device time per pass (HIP events: TV sweeps, or the diffusion
    call with each z plane held in LDS)."""
    out = {}
    for cfg, method, kc in (("c1", None, 0), ("c2", nd.TV, nd.KC_TV), ("c2", nd.DIFFUSION, nd.KC_DIFFUSION)):
        prm, proj = phantom.config_workload(cfg)
        name = cfg if method is None else "%s_%s" % (cfg, "tv" if method == nd.TV else "diffusion")
        res = {}
        for prec in (nd.FP32, nd.FP64_PLAIN):
            p = nd.make_params(prm["nx"], prm["nz"], prm["np_"], prm["gam_min"], prm["gam_max"], prm["interval"],
                               padangle=prm["padangle"], weight=prm["weight"], nhio=prm["nhio"], ner=prm["ner"],
                               denoise=method is not None, method=method if method is not None else nd.TV,
                               height=prm.get("height"), center=prm.get("center"), precision=prec)
            with nd.Reconstruction(p) as job:
                job.load_projections(proj)
                job.initial()
                job.iterate(2)
                steps = 20 if prec == nd.FP32 else 5
                dt, _, _ = timed(job, steps)
                r = {"passes_per_s": steps / dt}
                if kc:
                    _, _, (kms, kl, kb) = timed(job, steps, kc)
                    if kl:
                        r["denoiser_ms_per_pass"] = kms / steps
                        r["denoiser_algorithmic_GBs"] = kb / (kms * 1e-3) / 1e9
                res["f32" if prec == nd.FP32 else "f64"] = r
        res["speedup"] = res["f32"]["passes_per_s"] / res["f64"]["passes_per_s"]
        out[name] = res
        log("  %s: FP32 %.1f passes/s, FP64 plain %.1f passes/s (%.1fx)%s" % (
            name, res["f32"]["passes_per_s"], res["f64"]["passes_per_s"], res["speedup"],
            "; denoiser %.3f ms/pass FP32" % res["f32"]["denoiser_ms_per_pass"] if kc else ""))
    return out


def batch(args):
    """Uneven batch: 48 independent volumes (phantom.batch_instances: 15+
    sizes from 64^3 to 384 x 384 x 256, 36 distinct node geometries)
    submitted together to the job server, which co-schedules them
    (args.batch_slots in flight, least work first).  Timed three ways: the
    first run in the process (each geometry's tables -- window and
    deconvolution tables, gather maps, the three Toeplitz kernel spectra --
    built on first use), the same again with the geometries resident, and
    the resident jobs one at a time."""
    t0 = time.perf_counter()
    inst = phantom.batch_instances(args.batch_jobs)
    log("  batch: %d instances generated in %.1f s" % (len(inst), time.perf_counter() - t0))

    def params(prm):
        return nd.make_params(prm["nx"], prm["nz"], prm["np_"], prm["gam_min"], prm["gam_max"], prm["interval"],
                              padangle=prm["padangle"], weight=prm["weight"], nhio=prm["nhio"], ner=prm["ner"],
                              denoise=True, method=nd.TV, height=prm["height"], center=prm["center"],
                              precision=nd.FP32)

    passes = sum(prm["nhio"] + prm["ner"] for prm, _ in inst)
    res = {}
    for name, slots in (("cold", args.batch_slots), ("resident", args.batch_slots), ("one_at_a_time", 1)):
        with nd.Server(max_active=slots) as srv:
            t = time.perf_counter()
            for prm, proj in inst:
                srv.submit(params(prm), proj)
            for _ in inst:
                rep = srv.next(timeout_ms=-1)[0]
                if rep["status"] != 0:
                    raise RuntimeError("batch job failed: %s" % rep["error"])
            dt = time.perf_counter() - t
        res[name] = dt
        log("  batch, %s, %d in flight: %d volumes (%d volume-iters) in %.2f s -> %.1f volume-iter/s"
            % (name, slots, len(inst), passes, dt, passes / dt))
    sizes = {}
    for prm, _ in inst:
        k = "%d^3 np %d" % (prm["nz"], prm["np_"])
        sizes[k] = sizes.get(k, 0) + 1
    geoms = len({(p["np_"], p["gam_max"], p["interval"]) for p, _ in inst})
    return {"volumes": len(inst), "volume_iters": passes, "slots": args.batch_slots,
            "seconds": res["resident"], "volume_iters_per_s": passes / res["resident"],
            "seconds_cold": res["cold"], "volume_iters_per_s_cold": passes / res["cold"],
            "seconds_one_at_a_time": res["one_at_a_time"],
            "speedup_over_one_at_a_time": res["one_at_a_time"] / res["resident"],
            "node_geometries": geoms, "sizes": sizes,
            "schedule": "FP32, ner=30 with TV over the slab; every volume its own job, stream and buffers; "
                        "least work first, %d in flight" % args.batch_slots}


def stacked(args):
    """Many small volumes of one geometry (64^3, np 64, +-60 at step 2, SNR 4,
    ner = 30 with TV over the slab), FP32, end to end (create, load, initial
    INFFT, passes, volumes back), four ways: one batch context
    (nudim_batch_create: every kernel of a pass covers all volumes), one job
    after another, and the job server with 8 slots, every job alone
    (max_stack 1) or queued jobs batched (its default, up to 8 a batch)."""
    K, ner = args.stack_vols, 30
    vols = []
    for i in range(K):
        bp = phantom.blob_params(16, (64, 64, 64), 32, 6000 + i)
        vols.append(phantom.add_noise(phantom.projections(bp, 64, -60, 60, 2), 4.0, 7000 + i))
    p = nd.make_params(64, 64, 64, -60, 60, 2, padangle=1.0, weight=True, nhio=0, ner=ner, denoise=True,
                       method=nd.TV, height=32, center=32, precision=nd.FP32)

    def run_batch(projs):
        with nd.Batch(p, len(projs)) as b:
            for k, pr in enumerate(projs):
                b.load_projections(k, pr)
            b.initial()
            b.iterate(ner)
            return [b.volume(k) for k in range(len(projs))]

    def run_one(pr):
        with nd.Reconstruction(p) as job:
            job.load_projections(pr)
            job.initial()
            job.iterate(ner)
            return job.volume()

    def run_server(projs, max_stack, slots=8):
        out = [None] * len(projs)
        with nd.Server(max_active=slots, max_stack=max_stack) as srv:
            ids = {srv.submit(p, pr): k for k, pr in enumerate(projs)}
            for _ in projs:
                rep, v, _ = srv.next(timeout_ms=-1)
                if rep["status"] != 0:
                    raise RuntimeError("stacked leg job failed: %s" % rep["error"])
                out[ids[rep["job_id"]]] = v
        return out

    run_batch(vols[:3])  # the geometry's tables and the code paths warm
    run_one(vols[0])
    res = {}
    for name, fn in (("batch", lambda: run_batch(vols)), ("one_at_a_time", lambda: [run_one(v) for v in vols]),
                     ("server_8_slots_alone", lambda: run_server(vols, 1)),
                     ("server_8_slots_batched", lambda: run_server(vols, 0))):
        t = time.perf_counter()
        out = fn()
        res[name] = (time.perf_counter() - t, out)
        log("  %d volumes of 64^3, %s: %.2f s -> %.0f volume-iter/s"
            % (K, name, res[name][0], K * ner / res[name][0]))
    same = all(np.array_equal(a, b) for a, b in zip(res["batch"][1], res["one_at_a_time"][1]))
    return {"volumes": K, "passes": ner, "shape": [64, 64, 64],
            "volume_iters_per_s": {k: K * ner / v[0] for k, v in res.items()},
            "seconds": {k: v[0] for k, v in res.items()},
            "speedup_batch_over_one_at_a_time": res["one_at_a_time"][0] / res["batch"][0],
            "speedup_batch_over_server_alone": res["server_8_slots_alone"][0] / res["batch"][0],
            "batch_volumes_bit_equal_to_own_jobs": bool(same),
            "schedule": "FP32, ner=30 with TV over the slab, end to end per volume (create, load, initial, passes, "
                        "volume back)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--kclass", type=int, default=0, help="profile one kernel class only (0: all candidates)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=100, help="jobs in the serving trace (0: no serving legs)")
    ap.add_argument("--below-jobs", type=int, default=10, help="jobs in the below-capacity serving trace")
    ap.add_argument("--below-rate", type=float, default=0.3, help="its arrivals per second")
    ap.add_argument("--fp64-jobs", type=int, default=8,
                    help="jobs of the 128^3 and 256^3 classes served on both paths (0: none)")
    ap.add_argument("--no-batch", action="store_true", help="skip the uneven-batch leg")
    ap.add_argument("--no-small", action="store_true", help="skip the 128^3 configurations")
    ap.add_argument("--serve-rate", type=float, default=4.0, help="mean arrivals per second (Poisson)")
    ap.add_argument("--serve-slots", type=int, default=4, help="jobs in flight at once")
    ap.add_argument("--serve-seed", type=int, default=5)
    ap.add_argument("--serve-stack", type=int, default=0,
                    help="most same-parameter jobs the server runs as one batch (0: its default, 8; 1: none)")
    ap.add_argument("--serve-fifo", action="store_true", help="admit in arrival order (default: least work first)")
    ap.add_argument("--serve-small-cus", type=int, default=0,
                    help="CUs reserved for small jobs (128^3 class) in the serving leg (0: none)")
    ap.add_argument("--serve-fp64", action="store_true", help="also serve the trace on the plain FP64 path")
    ap.add_argument("--serve-only", action="store_true", help="run only the serving leg")
    ap.add_argument("--batch", action="store_true", help="run only the uneven-batch leg (48 volumes)")
    ap.add_argument("--cli", action="store_true", help="run only the drop-in CLI on the 256^3 workload's files")
    ap.add_argument("--batch-jobs", type=int, default=48)
    ap.add_argument("--batch-slots", type=int, default=8)
    ap.add_argument("--stack-vols", type=int, default=32, help="volumes of the stacked 64^3 leg (0: skip it)")
    ap.add_argument("--stacked", action="store_true", help="run only the stacked 64^3 leg")
    ap.add_argument("--probe", type=int, default=0, metavar="MIB",
                    help="run one FP32 pass and a MIB-MiB device copy, then exit (for counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    t0 = time.perf_counter()
    prm, proj = phantom.config_workload(args.config)
    log("workload %s: projections %s generated in %.1f s" % (args.config, proj.shape, time.perf_counter() - t0))
    if args.probe:
        probe(prm, proj, args.probe)
        return
    if args.batch:
        print(json.dumps({"batch": batch(args)}), flush=True)
        return
    if args.stacked:
        print(json.dumps({"stacked_64": stacked(args)}), flush=True)
        return
    if args.cli:
        print(json.dumps({"cli": cli_leg(args)}), flush=True)
        return
    if args.serve_only:
        out = {"serving": serve(args, nd.FP32)}
        if args.serve_fp64:
            out["serving_fp64"] = serve(args, nd.FP64_PLAIN)
        print(json.dumps(out), flush=True)
        return

    log("FP32 batched form:")
    j32 = make_job(prm, proj, nd.FP32)
    j32.iterate(args.warmup)
    dt32, r32, _ = timed(j32, args.steps)
    rate32 = args.steps / dt32
    log("  %d passes in %.3f s -> %.2f passes/s" % (args.steps, dt32, rate32))
    # the same passes replayed as captured hipGraphs (one graph per starting
    # buffer state and HIO/ER switch: at most six states, so twelve warm-up
    # passes capture them all)
    jg = make_job(prm, proj, nd.FP32, use_graph=True)
    gwarm = max(args.warmup, 12)
    jg.iterate(gwarm)
    dtg, rg, _ = timed(jg, args.steps)
    jg.close()
    graph = {"passes_per_s": args.steps / dtg, "steps": args.steps, "warmup": gwarm}
    log("  hipGraph replay: %d passes in %.3f s -> %.2f passes/s" % (args.steps, dtg, args.steps / dtg))
    # per-class HIP-event timing over separate timed runs; the dominant
    # kernel is the class with the most device time per pass
    classes = [args.kclass] if args.kclass else list(KC_CANDIDATES)
    KC_NAMES = kc_names(prm["np_"], prm["nz"], NFFT_CHOICE["f32"]["n"], NFFT_CHOICE["f32"]["W"])
    kern = {}
    for kc in classes:
        nsteps = max(2, args.steps // 2)
        _, _, prof = timed(j32, nsteps, kc)
        kms, klaunch, kbytes = prof
        if not klaunch:
            continue
        avg_us = 1e3 * kms / klaunch
        kern[kc] = {"name": KC_NAMES.get(kc), "ms_per_pass": kms / nsteps, "launches_per_pass": klaunch / nsteps,
                    "avg_us": avg_us, "bytes_per_launch": kbytes / klaunch,
                    "achieved_GBs": kbytes / klaunch / (avg_us * 1e-6) / 1e9}
        log("  %s: %.2f ms/pass, %d launches, avg %.1f us, %.1f GB/s algorithmic"
            % (KC_NAMES.get(kc), kms / nsteps, klaunch, avg_us, kern[kc]["achieved_GBs"]))
    dom = max(kern, key=lambda k: kern[k]["ms_per_pass"])
    kd = kern[dom]

    rate64 = None
    acc = {}
    spreading = {}
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
            # the FP64 path's own distance after the same passes from its
            # initial volume times (1 + 1e-7 noise): the chaotic floor of
            # the HIO + TV schedule
            rng = np.random.default_rng(1)
            j64.set_volume(v0_64 * (1 + 1e-7 * rng.standard_normal(v0_64.shape)))
            rp64 = j64.iterate(npass)
            floor = rel(j64.volume(), v64)
            acc = {"criterion": "rel L2 <= 1e-4 of the volume on ER passes without TV (er_no_tv); the configured "
                                "HIO + TV passes are chaotic in the reference's own iteration, so there FP32 is "
                                "compared with the FP64 path's distance from itself after a 1e-7 perturbation",
                   "rel_l2_initial": rel(v0, v0_64),
                   "hio_tv": {"passes": npass, "rel_l2_fp32_vs_fp64": rel(v32, v64),
                              "rel_l2_fp64_vs_fp64_perturbed_1e-7": floor,
                              "max_abs_rfactor_drift": float(np.max(np.abs(r32a - r64))),
                              "max_abs_rfactor_drift_fp64_perturbed": float(np.max(np.abs(rp64 - r64))),
                              "rfactor_fp64": [float(v) for v in r64], "rfactor_fp32": [float(v) for v in r32a]}}
            log("  rel L2 FP32 vs FP64: initial %.2e; after %d HIO+TV passes %.2e (FP64 vs itself after a 1e-7 "
                "perturbation: %.2e)" % (acc["rel_l2_initial"], npass, acc["hio_tv"]["rel_l2_fp32_vs_fp64"], floor))
        # the FP64 plain pass's spreading kernels (12 interpolations and 11
        # gathers per pass): HIP events on the job's stream, one pass each
        n64, w64 = NFFT_CHOICE["f64"]["n"], NFFT_CHOICE["f64"]["W"]
        for kc, what in ((2, "interpolation: k_interp_block<double> + k_interp_sum<double>, W=%d, n=%d"),
                         (3, "gather adjoint: k_gather_patch<double,8,4>, W=%d, n=%d")):
            _, _, (kms, klaunch, kbytes) = timed(j64, 1, kc)
            if not klaunch:
                continue
            avg_us = 1e3 * kms / klaunch
            gbs = kbytes / klaunch / (avg_us * 1e-6) / 1e9
            spreading[("interp", "gather")[kc - 2]] = {
                "kernels": what % (w64, n64), "launches_per_pass": klaunch, "avg_us": avg_us,
                "bytes_per_launch": kbytes / klaunch, "achieved_GBs": gbs, "frac_of_peak": gbs / HBM_PEAK_GBS,
                "ms_per_pass": kms}
            log("  FP64 %s: %d launches, avg %.1f us, %.0f GB/s algorithmic (%.2f of peak)"
                % (what % (w64, n64), klaunch, avg_us, gbs, gbs / HBM_PEAK_GBS))
        j64.close()
    j32.close()
    if not args.no_plain and not args.no_accuracy:
        # The HIO + TV schedule above amplifies any rounding difference
        # (TV ~1e3x per pass, HIO flips voxels near zero), so the FP32 error
        # bar is checked on the continuous part of the map: ER passes without
        # the denoiser, each form from its own initial INFFT.
        prm_er = dict(prm, nhio=0, denoise=False)
        npass = 3
        vols, rs = [], []
        for prec in (nd.FP32, nd.FP64_PLAIN):
            j = make_job(prm_er, proj, prec)
            rs.append(j.iterate(npass))
            vols.append(j.volume())
            j.close()
        e = rel(vols[0], vols[1])
        acc["er_no_tv"] = {"passes": npass, "rel_l2": e, "within_1e-4": bool(e <= 1e-4),
                           "max_abs_rfactor_drift": float(np.max(np.abs(rs[0] - rs[1])))}
        log("  ER without TV, %d passes: rel L2 FP32 vs FP64 %.2e, R drift %.2e"
            % (npass, e, acc["er_no_tv"]["max_abs_rfactor_drift"]))

    configs_128 = None if args.no_small else small_configs()
    batch_out = None if args.no_batch else batch(args)
    stacked_out = stacked(args) if args.stack_vols > 0 else None
    serving = serve(args, nd.FP32) if args.serve_jobs > 0 else None
    below = None
    if args.serve_jobs > 0 and args.below_jobs > 0:
        below = serve(args, nd.FP32, jobs=args.below_jobs, rate=args.below_rate, seed=args.serve_seed + 1)
    fp64_serving = None
    if args.serve_jobs > 0 and args.fp64_jobs > 0:
        # the same short trace on both paths; the 512^3 class is left out of
        # it because one such job takes about 90 s on the FP64 plain path
        small = ("s128", "s256")
        fp64_serving = {"classes": list(small),
                        "f32": serve(args, nd.FP32, jobs=args.fp64_jobs, rate=1.0, seed=args.serve_seed + 2,
                                     classes=small),
                        "f64": serve(args, nd.FP64_PLAIN, jobs=args.fp64_jobs, rate=1.0, seed=args.serve_seed + 2,
                                     classes=small)}
    try:
        copy_gbs = device_copy_gbs()
        log("  device copy (hipMemcpy, 1 GiB): %.0f GB/s read + write" % copy_gbs)
    except Exception as e:  # reported context, never required
        copy_gbs = None
        log("device copy measurement skipped: %s" % e)
    stream = {"read_GBs": nd.stream_probe(1 << 30, 0), "copy_GBs": nd.stream_probe(1 << 30, 1)}
    log("  streaming kernels over 1 GiB: read only %.0f GB/s, copy %.0f GB/s read + write"
        % (stream["read_GBs"], stream["copy_GBs"]))

    cpu = None
    if not args.no_cpu:
        try:
            # the host's cores as the process sees them, at most the 16 that a
            # one-GPU share of the host holds; and one core
            nproc = os.cpu_count() or 1
            cpu = cpu_baseline(prm, proj, min(nproc, 16))
            cpu["nproc"] = nproc
            one = cpu_baseline(prm, proj, 1)
            cpu["one_core"] = {k: one[k] for k in ("value", "unit", "cores", "sample")}
        except Exception as e:  # the baseline is reported, never required
            log("cpu baseline skipped: %s" % e)

    out = {
        "metric": "ER iters/s at 256\u00b3 FP32-batched vs FP64 plain (\u22641e-4 rel L2); p50/p99 job latency",
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
        "config": {"workload": workload_name(args.config, prm),
                   "grid": [prm["nz"], prm["np_"], prm["np_"]], "nfft": NFFT_CHOICE or None},
        "graph": graph,
        "fp64_plain_passes_per_s": rate64,
        "fp64_spreading": spreading or None,
        "speedup_vs_fp64_plain": (rate32 / rate64) if rate64 else None,
        "accuracy": acc or None,
        "roofline": {"kernel": kd["name"], "bound": "hbm", "achieved": kd["achieved_GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": kd["achieved_GBs"] / HBM_PEAK_GBS,
                     "traffic": traffic_from_profiles(dom) if args.config == "c3" else None,
                     "algorithmic_bytes": kd["bytes_per_launch"],
                     "avg_us": kd["avg_us"], "device_copy_GBs": copy_gbs, "stream_kernels_GBs": stream,
                     "frac_of_device_copy": (kd["achieved_GBs"] / copy_gbs) if copy_gbs else None},
        "kernels": {KC_NAMES.get(k): {kk: round(vv, 4) for kk, vv in v.items() if kk != "name"}
                    for k, v in kern.items()},
        "configs_128": configs_128,
        "batch": batch_out,
        "stacked_64": stacked_out,
        "serving": serving,
        "serving_below_capacity": below,
        "serving_fp32_vs_fp64": fp64_serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
