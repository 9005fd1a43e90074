"""The drop-in `reconstruction` binary: command line, files, output formats."""
import os
import subprocess

import numpy as np
import pytest

from tests.util import blob_phantom, rel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "nudim_amd", "bin", "reconstruction")


def run(args, cwd, env=None):
    e = dict(os.environ)
    if env:
        e.update(env)
    return subprocess.run([BIN] + args, cwd=cwd, capture_output=True, text=True, env=e, timeout=600)


@pytest.fixture(scope="module")
def binary(hip):
    if not os.path.exists(BIN):
        from nudim_amd import build
        build.build_host()
    return BIN


# (arguments, message of src/reconstruction.c:245-293)
BAD = [
    (["--nx=8", "--nz=8", "-p", "8"], "You need to provide prefix name of tomogram files with --prefix."),
    (["--prefix=prj", "--nz=8"], "You need to provide valid projection size with --nx, --nz."),
    (["--prefix=prj", "--nx=8", "--nz=8"], "You need to provide valid projection dimension in XY plane with --pad."),
    (["--prefix=prj", "--nx=8", "--nz=8", "-p", "8"], "You need to provide valid nhio with --nhio."),
    (["--prefix=prj", "--nx=8", "--nz=8", "-p", "8", "--nhio=0"], "You need to provide valid ner with --ner."),
    (["--prefix=prj", "--nx=8", "--nz=8", "-p", "8", "--nhio=0", "--ner=1"],
     "You need to provide valid angular sampling interval with --interval."),
    (["--prefix=prj", "--nx=8", "--nz=8", "-p", "8", "--nhio=0", "--ner=1", "-t", "3"],
     "You need to provide valid padded angular sampling interval with --padangle."),
    (["--prefix=prj", "--nx=8", "--nz=8", "-p", "8", "--nhio=0", "--ner=1", "-t", "3", "--padangle=2"],
     "You need to provide valid padded angular sampling interval with --padangle."),
    (["--prefix=prj", "--nx=8", "--nz=8", "-p", "8", "--nhio=0", "--ner=1", "-t", "3", "--padangle=1"],
     "You need to provide valid scope of tomogram angles (-90,90] with --max, --min."),
    (["--prefix=prj", "--nx=8", "--nz=8", "-p", "8", "--nhio=0", "--ner=1", "-t", "3", "--padangle=1",
      "--max=60", "--min=-60", "--denoise"], "You must provide valid denoising method or delete --denoise."),
    (["--prefix=prj", "--nx=8", "--nz=8", "-p", "8", "--nhio=0", "--ner=1", "-t", "3", "--padangle=1",
      "--max=60", "--min=-60", "--height=4"],
     "You must provide valid both sample height and center with --height and --center."),
    (["--prefix=prj", "--nx=8x"], "Invalid X dimension!"),
    (["--prefix=prj", "-t", "3a"], "Invalid angular sampling interval!"),
]


@pytest.mark.parametrize("args,msg", BAD)
def test_validation_messages(binary, tmp_path, args, msg):
    r = run(args, tmp_path)
    assert r.returncode == 1
    assert msg in r.stderr


def test_help_and_unknown_option(binary, tmp_path):
    r = run(["-h"], tmp_path)
    assert r.returncode == 0 and "--padangle" in r.stdout
    r = run(["--bogus"], tmp_path)
    assert r.returncode == 1


def test_fp32_without_weight_warns(binary, tmp_path):
    """The FP32 form without --weight is accurate to ~5e-2 only: the CLI
    says so on stderr before it starts (whatever happens after: here there
    are no projection files)."""
    args = ["--prefix=prj", "--nx=16", "--nz=16", "-p", "16", "--nhio=0", "--ner=1", "-t", "30",
            "--padangle=10", "--max=60", "--min=-60"]
    r = run(args, tmp_path, env={"NUDIM_PRECISION": "fp32"})
    assert "without --weight is accurate to about 5e-2" in r.stderr
    r = run(args + ["--weight"], tmp_path, env={"NUDIM_PRECISION": "fp32"})
    assert "without --weight" not in r.stderr


def test_missing_projection_file(binary, tmp_path):
    (tmp_path / "files").mkdir()
    r = run(["--prefix=prj", "--nx=16", "--nz=16", "-p", "16", "--nhio=0", "--ner=1", "-t", "30",
             "--padangle=10", "--max=60", "--min=-60"], tmp_path)
    assert r.returncode != 0


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_cli_end_to_end(binary, orc, tmp_path, prec):
    np_ = nz = 32
    rng = np.random.default_rng(21)
    dens = blob_phantom((nz, np_, np_), 10, rng, sig=(1.5, 3.0), slab=16)
    proj = orc.project_2d1d(dens, np_, -60, 60, 5)
    files = tmp_path / "files"
    files.mkdir()
    for j in range(proj.shape[0]):
        # projection_sim writes "%f\n" (src/projection_sim.c:401-402)
        np.savetxt(files / ("prj-%d.txt" % j), proj[j].ravel(), fmt="%f")
    args = ["--prefix=prj", "--pad=32", "--nx=24", "--nz=32", "--max=60", "--min=-60", "-t", "5",
            "--padangle=1", "--nhio=1", "--ner=2", "--weight", "--print"]
    r = run(args, tmp_path, env={"NUDIM_PRECISION": prec})
    assert r.returncode == 0, r.stderr
    lines = [l for l in r.stdout.splitlines() if l.startswith("Iteration:")]
    assert len(lines) == 3 and lines[0].startswith("Iteration: 0 Measured Residual: ")
    # the oracle on the same (text-rounded) files; with --nx=24 the reader
    # takes the first 24*32 values of each file (src/reconstruction.c:345-356)
    proj_in = np.stack([np.loadtxt(files / ("prj-%d.txt" % j))[: 24 * 32].reshape(32, 24)
                        for j in range(proj.shape[0])])
    op = orc.params(24, nz, np_, -60, 60, 5, padangle=1.0, weight=1, nhio=1, ner=2)
    img0, img, rt = orc.reconstruct(op, proj_in)
    r0 = np.loadtxt(tmp_path / "recovered-0.txt").reshape(nz, np_, np_)
    rec = np.loadtxt(tmp_path / "recovered.txt").reshape(nz, 24, 24)
    crop = img[:, 4:28, 4:28]
    # recovered-0.txt is "%le": 7 significant digits
    tol0 = 1e-6 if prec == "fp64" else 2e-5
    assert rel(r0, img0) < tol0
    # recovered.txt is "%f": 6 decimals
    assert np.max(np.abs(rec - crop)) < (1e-6 if prec == "fp64" else 1e-3 * np.abs(crop).max())
    printed = np.array([float(l.split()[-1]) for l in lines])
    assert np.max(np.abs(printed - rt)) < (1e-6 if prec == "fp64" else 1e-4)


@pytest.mark.gpu
def test_cli_dumps_at_pass_100(binary, hip, orc, tmp_path):
    """--denoise past pass 100: the serial program writes grad_before.txt and
    grad_after.txt ("%f", [z][y][x]) with the volume before and after the
    denoiser of pass 100 (src/reconstruction.c:637-658).  The files are the
    library's own dump of that pass, byte for byte; the dumped volumes are
    pinned to the oracle: the ER constraint of "before" is the oracle's
    pass 100 without the denoiser from the same volume, and "after" is the
    oracle's TV-height of "before"."""
    np_ = nz = 16
    rng = np.random.default_rng(23)
    dens = blob_phantom((nz, np_, np_), 6, rng, sig=(1.5, 3.0), slab=8)
    proj = orc.project_2d1d(dens, np_, -60, 60, 5)
    files = tmp_path / "files"
    files.mkdir()
    for j in range(proj.shape[0]):
        np.savetxt(files / ("prj-%d.txt" % j), proj[j].ravel(), fmt="%f")
    args = ["--prefix=prj", "--pad=16", "--nx=16", "--nz=16", "--max=60", "--min=-60", "-t", "5",
            "--padangle=1", "--nhio=2", "--ner=99", "--weight", "--denoise", "--method=TV", "--height=8",
            "--center=8"]
    r = run(args, tmp_path, env={"NUDIM_PRECISION": "fp64"})
    assert r.returncode == 0, r.stderr
    before_txt = (tmp_path / "grad_before.txt").read_bytes()
    after_txt = (tmp_path / "grad_after.txt").read_bytes()
    # the same job through the library, with full-precision dumps at pass 100
    proj_in = np.stack([np.loadtxt(files / ("prj-%d.txt" % j)).reshape(nz, np_) for j in range(proj.shape[0])])
    hp = hip.make_params(np_, nz, np_, -60, 60, 5, padangle=1.0, weight=1, nhio=2, ner=99, denoise=True,
                         method=hip.TV, height=8, center=8)
    before = np.zeros((nz, np_, np_))
    after = np.zeros((nz, np_, np_))
    with hip.Reconstruction(hp) as job:
        job.load_projections(proj_in)
        job.initial()
        job.set_dump(100, before, after)
        job.iterate(100)
        v100 = job.volume()
        job.iterate(1)
    fmt = lambda a: "".join("%f\n" % v for v in a.ravel()).encode()
    assert before_txt == fmt(before)
    assert after_txt == fmt(after)
    assert not np.array_equal(before, after)
    op = orc.params(np_, nz, np_, -60, 60, 5, padangle=1.0, weight=1, nhio=2, ner=99, height=8, center=8)
    slices, kind = orc.assemble(op, proj_in)
    er, _ = orc.iterate(op, v100, slices, kind, 100)  # pass 100 without the denoiser
    assert rel(orc.constraint(before, v100, np_, 0, 1, 8, 8), er) < 1e-10
    assert rel(after, orc.tv_height(before, 8, 8)) < 1e-11


# ---- projection_sim ---------------------------------------------------------

SIM = os.path.join(ROOT, "nudim_amd", "bin", "projection_sim")


def run_sim(args, cwd):
    return subprocess.run([SIM] + args, cwd=cwd, capture_output=True, text=True, timeout=600)


@pytest.fixture(scope="module")
def sim_binary(hip):
    if not os.path.exists(SIM):
        from nudim_amd import build
        build.build_host()
    return SIM


# (arguments, message of src/projection_sim.c:204-245)
SIM_BAD = [
    (["--nx=8", "--ny=8", "--nz=8"], "You need to provide 3D density with --infile."),
    (["-i", "d.txt", "--nx=8", "--nz=8"], "You need to provide valid dimensions with --nx, --ny, --nz."),
    (["-i", "d.txt", "--nx=8", "--ny=8", "--nz=8"],
     "You need to provide valid projection dimension in XY plane with --pad."),
    (["-i", "d.txt", "--nx=8", "--ny=8", "--nz=8", "-p", "8"],
     "You need to provide valid angular sampling interval with --interval."),
    (["-i", "d.txt", "--nx=8", "--ny=8", "--nz=8", "-p", "8", "-t", "3"],
     "You need to provide valid scope of tomogram angles (-90,90] with --max, --min."),
    (["-i", "d.txt", "--nx=8", "--ny=8", "--nz=8", "-p", "8", "-t", "3", "--max=60", "--min=-60", "--noise"],
     "You must provide valid signal to noise ratio with --snr or delete --noise."),
    (["-i", "d.txt", "--nx=8", "--ny=8", "--nz=8", "-p", "8", "-t", "3", "--max=60", "--min=-60"],
     "Error : Data file d.txt doesn't exist!"),
    (["-i", "d.txt", "--nx=8", "--ny=8", "--nz=8", "-p", "8", "-t", "3", "--max=60", "--min=-60", "-m", "DFT"],
     "is not built here"),
    (["-i", "d.txt", "--nx=8", "--ny=6", "--nz=8", "-p", "8", "-t", "3", "--max=60", "--min=-60", "-m", "2D1D"],
     "square even x-y planes"),
    (["--ny=8y"], "Invalid Y dimension!"),
    (["--snr=x"], "Invalid signal to noise ratio!"),
]


@pytest.mark.parametrize("args,msg", SIM_BAD)
def test_sim_validation_messages(sim_binary, tmp_path, args, msg):
    r = run_sim(args, tmp_path)
    assert r.returncode == 1
    assert msg in r.stderr


def test_sim_help_and_default_method(sim_binary, tmp_path):
    r = run_sim(["-h"], tmp_path)
    assert r.returncode == 0 and "--snr" in r.stdout
    # no --method: the reference's notice, then the density check
    r = run_sim(["-i", "d.txt", "--nx=8", "--ny=8", "--nz=8", "-p", "8", "-t", "3", "--max=60", "--min=-60"],
                tmp_path)
    assert "using the 2D1D method, which is fastest." in r.stderr.replace("  ", " ")


def test_sim_short_density_file(sim_binary, tmp_path):
    np.savetxt(tmp_path / "d.txt", np.zeros(100), fmt="%le")
    r = run_sim(["-i", "d.txt", "--nx=8", "--ny=8", "--nz=8", "-p", "8", "-t", "30", "--max=60", "--min=-60"],
                tmp_path)
    assert r.returncode == 1 and "holds 100 of the 512 values" in r.stderr


@pytest.mark.gpu
def test_sim_end_to_end(sim_binary, orc, tmp_path):
    """density text map -> ./files/prj-<j>.txt ("%f"), against the oracle's
    2D1D simulator on the same (text-rounded) density."""
    n, nz, np_ = 24, 16, 32
    rng = np.random.default_rng(33)
    dens = blob_phantom((nz, n, n), 6, rng, sig=(1.5, 3.0), slab=12)
    np.savetxt(tmp_path / "d.txt", dens.ravel(), fmt="%.17e")
    (tmp_path / "files").mkdir()
    r = run_sim(["-i", "d.txt", "--nx=%d" % n, "--ny=%d" % n, "--nz=%d" % nz, "-p", str(np_), "-t", "10",
                 "--max=60", "--min=-60", "-m", "2D1D"], tmp_path)
    assert r.returncode == 0, r.stderr
    ref = orc.project_2d1d(np.loadtxt(tmp_path / "d.txt").reshape(nz, n, n), np_, -60, 60, 10)
    assert ref.shape[0] == 13
    for j in range(ref.shape[0]):
        got = np.loadtxt(tmp_path / "files" / ("prj-%d.txt" % j)).reshape(nz, np_)
        # "%f": 6 decimals
        assert np.max(np.abs(got - ref[j])) <= 5.1e-7
    # with noise: the same files plus Gaussian noise of the stated level
    (tmp_path / "noisy").mkdir()
    (tmp_path / "noisy" / "files").mkdir()
    np.savetxt(tmp_path / "noisy" / "d.txt", dens.ravel(), fmt="%.17e")
    r = run_sim(["-i", "d.txt", "--nx=%d" % n, "--ny=%d" % n, "--nz=%d" % nz, "-p", str(np_), "-t", "10",
                 "--max=60", "--min=-60", "--noise", "--snr=4"], tmp_path / "noisy")
    assert r.returncode == 0, r.stderr
    for j in (0, 6, 12):
        clean = ref[j].ravel()
        noisy = np.loadtxt(tmp_path / "noisy" / "files" / ("prj-%d.txt" % j))
        ratio = np.std(noisy - clean) / np.std(clean)
        assert 0.15 < ratio < 0.35  # sigma_noise = sigma_signal / snr (src/projection_sim.c:58-62)


# ---- reconstruction_parallel -----------------------------------------------

PAR = os.path.join(ROOT, "nudim_amd", "bin", "reconstruction_parallel")


def run_par(args, cwd, env=None):
    e = dict(os.environ)
    if env:
        e.update(env)
    return subprocess.run([PAR] + args, cwd=cwd, capture_output=True, text=True, env=e, timeout=600)


def grad_mag(u):
    """show_grad_mag of denoise.c:381-402 on u[z][y][x]: the density on the
    border, the norm of the three backward differences inside."""
    g = u.copy()
    c = u[1:-1, 1:-1, 1:-1]
    g[1:-1, 1:-1, 1:-1] = np.sqrt((c - u[:-2, 1:-1, 1:-1]) ** 2 + (c - u[1:-1, :-2, 1:-1]) ** 2 +
                                  (c - u[1:-1, 1:-1, :-2]) ** 2)
    return g


@pytest.fixture(scope="module")
def par_binary(hip):
    if not os.path.exists(PAR):
        from nudim_amd import build
        build.build_host()
    return PAR


def test_par_validation(par_binary, tmp_path):
    base = ["--prefix=prj", "--nx=8", "--nz=8", "-p", "8", "--nhio=0", "--ner=1", "-t", "3", "--padangle=1",
            "--max=60", "--min=-60"]
    r = run_par(base + ["--height=0"], tmp_path)
    assert r.returncode == 1 and "You must provide valid sample height --height." in r.stderr
    # no --center in the parallel program (reconstruction-parallel.c:108-127)
    assert run_par(base + ["--height=4", "--center=4"], tmp_path).returncode == 1
    r = run_par(["-h"], tmp_path)
    assert r.returncode == 0 and "--height.                 Sample height for constraint." in r.stdout
    assert "--center" not in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("method,passes", [("DIFFUSION", 2), ("TV", 1)])
def test_par_end_to_end(par_binary, hip, orc, tmp_path, method, passes):
    """FP64: the parallel program's numerics (band centred at np/2, TV over
    the whole volume, DIFFUSION with lambda 0.4f) against the oracle given the
    same switches; one TV pass keeps TV's rounding amplification ~1e-9."""
    np_ = nz = 32
    rng = np.random.default_rng(23)
    dens = blob_phantom((nz, np_, np_), 10, rng, sig=(1.5, 3.0), slab=16)
    proj = orc.project_2d1d(dens, np_, -60, 60, 5)
    files = tmp_path / "files"
    files.mkdir()
    for j in range(proj.shape[0]):
        np.savetxt(files / ("prj-%d.txt" % j), proj[j].ravel(), fmt="%f")
    args = ["--prefix=prj", "--pad=32", "--nx=32", "--nz=32", "--max=60", "--min=-60", "-t", "5",
            "--padangle=1", "--nhio=0", "--ner=%d" % passes, "--weight", "--height=14", "--denoise",
            "--method=%s" % method, "--print"]
    r = run_par(args, tmp_path, env={"NUDIM_PRECISION": "fp64"})
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip().splitlines()[-1].startswith("Total time used : ")
    proj_in = np.stack([np.loadtxt(files / ("prj-%d.txt" % j)).reshape(32, 32) for j in range(proj.shape[0])])
    op = orc.params(32, nz, np_, -60, 60, 5, padangle=1.0, weight=1, nhio=0, ner=passes, denoise=1,
                    method=0 if method == "DIFFUSION" else 1, height=14, center=16, tv_full=True,
                    diffusion_lambda=0.4 if method == "DIFFUSION" else 0.0)
    img0, img, rt = orc.reconstruct(op, proj_in)
    r0 = np.loadtxt(tmp_path / "recovered-0-parallel.txt").reshape(nz, np_, np_)
    rec = np.loadtxt(tmp_path / "recovered-parallel.txt").reshape(nz, 32, 32)
    assert rel(r0, img0) < 1e-6
    assert np.max(np.abs(rec - img)) < 1e-5
    # gradient-magnitude dumps of the density before / after the denoiser at
    # pass 0: the library's dumps of the same job, the denoiser step against
    # the oracle, grad_mag restated from denoise.c:381-402
    hp = hip.make_params(32, nz, np_, -60, 60, 5, padangle=1.0, weight=True, nhio=0, ner=passes, denoise=True,
                         method=hip.DIFFUSION if method == "DIFFUSION" else hip.TV, height=14, center=16,
                         tv_full=True, diffusion_lambda=0.4 if method == "DIFFUSION" else 0.0)
    before, after = np.zeros((nz, np_, np_)), np.zeros((nz, np_, np_))
    with hip.Reconstruction(hp) as job:
        job.load_projections(proj_in)
        job.initial()
        job.set_dump(0, before, after)
        job.iterate(passes)
    ref_after = orc.diffusion2d(before, lam=0.4) if method == "DIFFUSION" else orc.tv(before)
    assert rel(after, ref_after) < 1e-9
    for name, v in (("grad_before.txt", before), ("grad_after.txt", after)):
        g = np.loadtxt(tmp_path / name)
        assert np.max(np.abs(g - grad_mag(v).ravel())) < 1e-5
    # the band is centred at np/2: rows outside (32-14)/2 .. (32+14)/2 are zero after ER
    assert np.all(rec[:, :9, :] == 0) and np.all(rec[:, 24:, :] == 0)
