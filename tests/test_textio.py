"""Threaded text I/O of the command-line tools (nudim_amd/host/textio.c)
against sequential formatting and parsing: printf's %f / %le and strtod are
exact, as are Python's % formatting and float(), so files and values must
match byte for byte and bit for bit."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "nudim_amd", "host")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    d = tmp_path_factory.mktemp("textio")
    exe = str(d / "textio_check")
    subprocess.run(["gcc", "-O2", "-std=gnu99", "-o", exe, os.path.join(ROOT, "tests", "textio_check.c"),
                    os.path.join(HOST, "textio.c"), "-I", HOST, "-lpthread"], check=True)
    return exe


def values(seed, n):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 8, n)
    v[:6] = [0.0, -0.0, 0.5, 1e-300, 123456789.123456789, -2.5e-7]
    return v


@pytest.mark.parametrize("threads", [1, 3, 8])
@pytest.mark.parametrize("fmt", ["%f", "%le"])
def test_write_matches_printf(driver, tmp_path, threads, fmt):
    v = values(threads, 20011)
    src = tmp_path / "in.bin"
    v.tofile(src)
    out = tmp_path / "out.txt"
    subprocess.run([driver, "write", str(src), str(out), fmt + "\n", str(threads)], check=True)
    pyfmt = fmt.replace("l", "")
    assert out.read_bytes() == "".join(pyfmt % x + "\n" for x in v).encode()


@pytest.mark.parametrize("threads", [1, 5, 16])
def test_read_matches_strtod(driver, tmp_path, threads):
    v = values(10 + threads, 30007)
    txt = tmp_path / "in.txt"
    # mixed separators and formats, as fscanf("%le\n") accepts them
    parts = []
    for i, x in enumerate(v):
        s = repr(float(x)) if i % 3 else "%.17e" % x
        parts.append(s + ("\n" if i % 5 else "  \n\t"))
    txt.write_text("".join(parts))
    out = tmp_path / "out.bin"
    r = subprocess.run([driver, "read", str(txt), str(out), str(len(v)), str(threads)], capture_output=True,
                       text=True, check=True)
    assert int(r.stdout) == len(v)
    got = np.fromfile(out, dtype=np.float64)
    assert np.array_equal(got.view(np.uint64), np.array([float(p.split()[0]) for p in parts]).view(np.uint64))


def test_short_and_missing(driver, tmp_path):
    txt = tmp_path / "in.txt"
    txt.write_text("1.0\n2.0\n3.0\n")
    out = tmp_path / "out.bin"
    r = subprocess.run([driver, "read", str(txt), str(out), "10", "4"], capture_output=True, text=True, check=True)
    assert int(r.stdout) == 3
    r = subprocess.run([driver, "read", str(tmp_path / "none.txt"), str(out), "10", "4"], capture_output=True,
                       text=True, check=True)
    assert int(r.stdout) == -1
