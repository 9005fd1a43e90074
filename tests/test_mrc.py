"""The drop-in MRC converter (examples/exp/mrcread.c:3-90, mrcfile.c:5-210):
stdout and every ./files/prj-<j>.txt byte for byte those of the reference's
own converter (compiled from its sources into oracle/_ref where they are
present), values against the oracle's restatement."""
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "nudim_amd", "bin", "mrcread")
REF = os.path.join(ROOT, "oracle", "_ref", "mrcread")


@pytest.fixture(scope="module")
def binary():
    if not os.path.exists(BIN):
        from nudim_amd import build
        build.build_host()
    return BIN


def mrc_bytes(data, mode, rng, tag=b"MAP "):
    """An MRC file: 1024-byte header (MRC2014 word layout) + data [nz][ny][nx]."""
    nz, ny, nx = data.shape
    h = bytearray(1024)
    struct.pack_into("<10i", h, 0, nx, ny, nz, mode, 3, -4, 5, nx, ny, nz)
    struct.pack_into("<6f", h, 40, *rng.uniform(-50, 50, 6).astype(np.float32))
    struct.pack_into("<3i", h, 64, 1, 2, 3)
    struct.pack_into("<3f", h, 76, *rng.uniform(-5, 5, 3).astype(np.float32))
    struct.pack_into("<2h", h, 88, 7, -2)
    h[208:212] = tag
    dt = {0: np.uint8, 1: np.int16, 2: np.float32}[mode]
    return bytes(h) + np.ascontiguousarray(data, dt).tobytes()


def run_in(exe, d, args=()):
    (d / "files").mkdir(exist_ok=True)
    return subprocess.run([exe] + list(args), cwd=d, capture_output=True, text=True, timeout=120)


def stack_for(mode, shape, rng):
    if mode == 0:
        return rng.integers(0, 256, shape).astype(np.uint8)
    if mode == 1:
        return rng.integers(-3000, 3000, shape).astype(np.int16)
    return (rng.standard_normal(shape) * 40).astype(np.float32)


# (mode, nz, ny, nx): nx = 61 leaves the unshifted block empty; ny = 1 wraps
# the row shift onto itself
CASES = [(2, 5, 10, 72), (1, 4, 7, 64), (0, 3, 6, 90), (2, 3, 1, 61), (2, 2, 9, 128)]


@pytest.mark.parametrize("mode,nz,ny,nx", CASES)
def test_mrcread_matches_reference(binary, tmp_path, mode, nz, ny, nx):
    rng = np.random.default_rng(mode * 100 + nx)
    data = stack_for(mode, (nz, ny, nx), rng)
    blob = mrc_bytes(data, mode, rng)
    mine, ref = tmp_path / "mine", tmp_path / "ref"
    for d in (mine, ref):
        d.mkdir()
        (d / "BBa.ali").write_bytes(blob)
    r = run_in(binary, mine)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("Data Dimension: %d %d %d\n" % (nx, ny, nz))
    files = sorted(os.listdir(mine / "files"))
    assert files == sorted("prj-%d.txt" % j for j in range(nz - 1))
    # values: the oracle's restatement (printed with %f: 6 decimals)
    exp = oracle.mrc_convert(data.astype(np.float32))
    for j, (vals, mx, mn) in exp.items():
        got = np.loadtxt(mine / "files" / ("prj-%d.txt" % j))
        assert np.max(np.abs(got - vals)) <= 5e-7 * max(1.0, np.abs(vals).max())
    # bytes: the reference's converter on the same file
    if not os.path.exists(REF):
        pytest.skip("reference converter not built (reference sources absent)")
    rr = run_in(REF, ref)
    assert r.stdout == rr.stdout
    for f in files:
        assert (mine / "files" / f).read_bytes() == (ref / "files" / f).read_bytes(), f


def test_oracle_mrc_convert_pins():
    """The restatement against what the shift fixes: a constant section maps
    to itself (max - c + min = c), the output is a permutation of the
    contrast-inverted section, and single marked entries land where
    mrcread.c:49-63 sends them."""
    nz, ny, nx = 2, 4, 70
    st = np.full((nz, ny, nx), 3.5, np.float32)
    vals, mx, mn = oracle.mrc_convert(st)[0]
    assert mx == mn == 3.5 and np.all(vals == 3.5)
    rng = np.random.default_rng(1)
    st = rng.standard_normal((nz, ny, nx)).astype(np.float32)
    vals, mx, mn = oracle.mrc_convert(st)[0]
    assert np.array_equal(np.sort(vals), np.sort((mx - st[1].astype(np.float64) + mn).ravel()))
    for (j, i, row, col) in [(0, 0, ny - 1, nx - 61), (2, 60, 1, nx - 1), (2, 61, 2, 0), (3, 69, 3, 8)]:
        st = np.zeros((nz, ny, nx), np.float32)
        st[1, j, i] = 1.0
        vals, mx, mn = oracle.mrc_convert(st)[0]
        v = vals.reshape(ny, nx)
        assert v[row, col] == 0.0 and np.count_nonzero(v == 0.0) == 1


def test_mrcread_rejects(binary, tmp_path):
    rng = np.random.default_rng(2)
    data = stack_for(2, (3, 4, 64), rng)
    (tmp_path / "BBa.ali").write_bytes(mrc_bytes(data, 2, rng, tag=b"XXXX"))
    r = run_in(binary, tmp_path)
    assert r.returncode == 1 and "Warning: Not MRC format!" in r.stdout
    (tmp_path / "BBa.ali").write_bytes(mrc_bytes(data, 2, rng)[:-4])
    r = run_in(binary, tmp_path)
    assert r.returncode == 1 and "fewer than" in r.stderr
    blob = bytearray(mrc_bytes(data, 2, rng))
    struct.pack_into("<i", blob, 12, 6)
    (tmp_path / "BBa.ali").write_bytes(bytes(blob))
    r = run_in(binary, tmp_path)
    assert r.returncode == 1 and "mode 6" in r.stderr
    # an explicit input path
    (tmp_path / "stack.mrc").write_bytes(mrc_bytes(data, 2, rng))
    r = run_in(binary, tmp_path, ["stack.mrc"])
    assert r.returncode == 0 and sorted(os.listdir(tmp_path / "files")) == ["prj-0.txt", "prj-1.txt"]
