"""The correlator record file (include/plegma/PLEGMA_h5twin.h) written by a
host-only C++ program and read back by plegma_amd.h5twin: byte layout, path
normalisation, append on reopen, the reference's skip-if-present and
not-a-group rules (include/io/PLEGMA_hdf5.h:272-279, 665-667) and the
format deduction of writeFile (include/PLEGMA_io.h:8-34)."""
import os
import struct
import subprocess

import numpy as np
import pytest

from plegma_amd import h5twin

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    d = tmp_path_factory.mktemp("h5twin")
    exe = str(d / "selftest")
    src = os.path.join(ROOT, "tests", "cpp", "h5twin_selftest.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I" + os.path.join(ROOT, "include"),
                        "-I/opt/rocm/include", src, "-o", exe], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr)
    fn = str(d / "t.h5.twin")
    r = subprocess.run([exe, fn], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = dict(ln.split(" ", 1) for ln in r.stdout.splitlines())
    return fn, lines, r.stderr


def test_datasets(run):
    fn, _, _ = run
    ds = h5twin.read(fn)
    assert list(ds) == ["/sx01sy02sz03st05/g/x", "/sx01sy02sz03st05/g/y", "/sx01sy02sz03st05/g/mvec",
                        "/sx01sy02sz03st05/h/c", "/empty"]
    x = ds["/sx01sy02sz03st05/g/x"]
    np.testing.assert_array_equal(x.data, (0.5 * np.arange(6) - 1.0).reshape(2, 3))
    assert x.data.dtype == np.float64
    assert x.attrs == {"description": "shape: /a/b", "k": "v"}
    y = ds["/sx01sy02sz03st05/g/y"].data
    assert y.dtype == np.float32 and y.tolist() == [1.5, -2.25, 3.0, np.float32(1e-30)]
    m = ds["/sx01sy02sz03st05/g/mvec"].data
    assert m.dtype == np.int32 and m.tolist() == [[0, 0, 0], [1, -1, 0]]
    np.testing.assert_array_equal(ds["/sx01sy02sz03st05/h/c"].data, [[7, 8], [9, 10]])
    assert ds["/empty"].data.shape == (0, 2)


def test_rules(run):
    _, lines, err = run
    assert lines["dup"] == "0"                 # skipped, old data kept (checked above)
    assert "already exists" in err and "Skipping" in err
    assert lines["exists"] == "1 0"            # groups are implicit, not datasets
    assert lines["errors"] == "2"              # a dataset used as a group and vice versa
    assert lines["bad"] == "1"                 # a file without the magic is refused


def test_formats(run):
    # ASCII_FORMAT 0, HDF5_FORMAT 1, LIME_FORMAT 2, DEFAULT_FORMAT 3
    assert run[1]["formats"] == "1 1 1 0 0 2 3"


def test_byte_layout(run):
    """the first record, decoded by hand from the documented layout"""
    fn, _, _ = run
    b = open(fn, "rb").read()
    assert b[:8] == b"PLGH5TW1"
    tag, n = struct.unpack_from("<II", b, 8)
    assert tag == 0x54455344
    path = b[16:16 + n].decode()
    assert path == "/sx01sy02sz03st05/g/x"
    o = 16 + n
    dtype, ndim = struct.unpack_from("<II", b, o)
    dims = struct.unpack_from("<2Q", b, o + 8)
    assert (dtype, ndim, dims) == (1, 2, (2, 3))


def test_truncated(tmp_path, run):
    fn, _, _ = run
    b = open(fn, "rb").read()
    p = tmp_path / "cut.twin"
    p.write_bytes(b[:-3])
    with pytest.raises(ValueError):
        h5twin.read(str(p))
