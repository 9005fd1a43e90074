"""Builds libnudim_hip.so (gfx950) and the drop-in C host binaries in-tree.

    python -m nudim_amd.build            # library + binaries
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libnudim_hip.so")
BIN = os.path.join(HERE, "bin")

HIP_SOURCES = ["ops.hip", "api.hip", "fft2.hip", "fft2_more.hip", "serve.hip", "tune.hip"]
CPP_SOURCES = ["geom.cpp"]
ARCH = os.environ.get("NUDIM_OFFLOAD_ARCH", "gfx950")
# Packed FP32 instructions are switched off for all device code.
# Per-thread default streams: the synchronous copies of one job's set-up do
# not wait for other jobs' blocking (CU-masked) streams.
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-fgpu-default-stream=per-thread",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wshadow", "-Wno-unused-function", "-Wno-unused-variable"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def _stale(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def build_lib(force=False, jobs=4):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")]
    headers.append(os.path.join(INCLUDE, "nudim.h"))
    objs, procs = [], []
    for src in HIP_SOURCES + CPP_SOURCES:
        s = os.path.join(CSRC, src)
        o = os.path.join(objdir, src + ".o")
        objs.append(o)
        if force or _stale(o, [s] + headers):
            if src.endswith(".hip"):
                cmd = [hipcc, "-c", "-x", "hip", s, "-o", o, "-I", INCLUDE, "-I", CSRC] + HIPFLAGS
            else:
                cmd = ["g++", "-c", s, "-o", o, "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wshadow", "-I", INCLUDE, "-I", CSRC]
            print("+", " ".join(cmd), flush=True)
            procs.append(subprocess.Popen(cmd))
            if len(procs) >= jobs:
                _wait(procs)
    _wait(procs)
    if force or _stale(LIB, objs):
        tmp = LIB + ".tmp"
        _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
        os.replace(tmp, LIB)
    return LIB


def _wait(procs):
    while procs:
        p = procs.pop(0)
        if p.wait() != 0:
            raise subprocess.CalledProcessError(p.returncode, p.args)


# host programs, their sources (nudim_amd/host) and whether they link libnudim_hip
HOST_PROGRAMS = {
    "reconstruction": (["reconstruction.c", "recon_cli.c", "textio.c"], True),
    "reconstruction_parallel": (["reconstruction_parallel.c", "recon_cli.c", "textio.c"], True),
    "projection_sim": (["projection_sim.c", "textio.c"], True),
    "mrcread": (["mrcread.c", "textio.c"], False),
}


def build_host(force=False):
    os.makedirs(BIN, exist_ok=True)
    out = []
    headers = [os.path.join(HOST, h) for h in sorted(os.listdir(HOST)) if h.endswith(".h")]
    for prog, (srcs, uses_lib) in HOST_PROGRAMS.items():
        exe = os.path.join(BIN, prog)
        paths = [os.path.join(HOST, s) for s in srcs]
        libs = ["-L", HERE, "-lnudim_hip", "-Wl,-rpath,$ORIGIN/.."] if uses_lib else []
        if force or _stale(exe, paths + headers + ([LIB] if uses_lib else []) + [os.path.join(INCLUDE, "nudim.h")]):
            _run(["gcc", "-O2", "-std=gnu99", "-Wall", "-o", exe] + paths + ["-I", INCLUDE, "-I", HOST] + libs +
                 ["-lm", "-lpthread"])
        out.append(exe)
    return out


def build(force=False):
    build_lib(force=force)
    build_host(force=force)


if __name__ == "__main__":
    build(force="--force" in sys.argv)
