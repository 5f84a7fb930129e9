"""MATLAB level-4 and CSV I/O of the C++ host side (physicell-x_amd/host),
driven by tests/cpp/io_demo (no GPU work): load_initial_conditions_from_matlab
from a file written by scipy, load_initial_conditions_from_csv, write_to_matlab
read back by scipy.  Checks the
reference's file layout (BioFVM/BioFVM_microenvironment.cpp:1222-1250,
2341-2387): one column per voxel [x, y, z, dV, densities] in voxel-index
order, voxel centres start + (i + 1/2) dx, and the rejections of a malformed
initial condition."""
import os
import subprocess

import numpy as np
import pytest
import scipy.io

from helpers import ROOT

DEMO = os.path.join(ROOT, "tests", "cpp", "io_demo")


def centres(n, dx):
    return -0.5 * dx * n + (np.arange(n) + 0.5) * dx


def ic_matrix(nx, ny, nz, ns, dx, field, order):
    x, y, z = centres(nx, dx), centres(ny, dx), centres(nz, dx)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    rows = np.column_stack([x[i.ravel()], y[j.ravel()], z[k.ravel()], field.reshape(-1, ns)])
    return rows[order]


def run(args):
    return subprocess.run([DEMO] + [str(a) for a in args], capture_output=True, text=True, timeout=60)


def test_initial_condition_and_matlab_output_round_trip(tmp_path):
    assert os.path.exists(DEMO), "build the C++ host side first (make -C physicell-x_amd/host)"
    nx, ny, nz, ns, dx = 7, 5, 4, 3, 12.5
    rng = np.random.default_rng(3)
    field = rng.uniform(0, 40, nx * ny * nz * ns)
    order = rng.permutation(nx * ny * nz)  # rows in any voxel order
    ic = tmp_path / "ic.mat"
    scipy.io.savemat(str(ic), {"M": ic_matrix(nx, ny, nz, ns, dx, field, order)}, format="4")
    out = tmp_path / "out.mat"
    r = run([nx, ny, nz, ns, dx, ic, out])
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(out, "rb").read()
    hdr = np.frombuffer(raw[:20], np.uint32)
    assert list(hdr) == [0, 4 + ns, nx * ny * nz, 0, len("multiscale_microenvironment")]
    assert raw[20:20 + hdr[4]] == b"multiscale_microenvironment"
    m = scipy.io.loadmat(str(out))["multiscale_microenvironment"]
    assert m.shape == (4 + ns, nx * ny * nz)
    x, y, z = centres(nx, dx), centres(ny, dx), centres(nz, dx)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    assert np.array_equal(m[0], x[i.ravel()]) and np.array_equal(m[1], y[j.ravel()])
    assert np.array_equal(m[2], z[k.ravel()])
    assert np.all(m[3] == dx ** 3)
    assert np.array_equal(m[4:].T.ravel(), field)


def test_malformed_initial_conditions_are_rejected(tmp_path):
    nx, ny, nz, ns, dx = 4, 3, 2, 2, 10.0
    field = np.arange(nx * ny * nz * ns, dtype=float)
    order = np.arange(nx * ny * nz)
    good = ic_matrix(nx, ny, nz, ns, dx, field, order)
    cases = {"rows": good[:-1], "cols": good[:, :-1], "repeat": np.vstack([good[:-1], good[:1]])}
    for name, mat in cases.items():
        p = tmp_path / f"{name}.mat"
        scipy.io.savemat(str(p), {"M": mat}, format="4")
        r = run([nx, ny, nz, ns, dx, p, tmp_path / "o.mat"])
        assert r.returncode == 1, (name, r.stdout)
        assert "ERROR" in r.stdout


def write_csv(path, rows, header=None):
    with open(path, "w") as f:
        if header is not None:
            f.write(header + "\n")
        for r in rows:
            f.write(",".join(repr(float(v)) for v in r) + "\n")


def test_csv_initial_conditions(tmp_path):
    """load_initial_conditions_from_csv (BioFVM/BioFVM_microenvironment.cpp:
    2391-2480): rows in any voxel order, with no header (columns are substrates
    0..ns-1) or with a header naming a subset of the substrates in any order
    (the others keep their values, 0 here)."""
    nx, ny, nz, ns, dx = 5, 4, 3, 3, 7.5
    rng = np.random.default_rng(11)
    field = rng.uniform(-5, 40, nx * ny * nz * ns)
    order = rng.permutation(nx * ny * nz)
    rows = ic_matrix(nx, ny, nz, ns, dx, field, order)
    # no header
    p = tmp_path / "ic.csv"
    write_csv(p, rows)
    out = tmp_path / "o.mat"
    r = run([nx, ny, nz, ns, dx, p, out])
    assert r.returncode == 0, r.stdout + r.stderr
    m = scipy.io.loadmat(str(out))["multiscale_microenvironment"]
    assert np.array_equal(m[4:].T.ravel(), field)
    # header naming substrates 2 and 0 (in that order): substrate 1 stays 0
    p2 = tmp_path / "ic_named.csv"
    write_csv(p2, rows[:, [0, 1, 2, 5, 3]], header="x,y,z,s2,s0")
    r = run([nx, ny, nz, ns, dx, p2, out])
    assert r.returncode == 0, r.stdout + r.stderr
    m = scipy.io.loadmat(str(out))["multiscale_microenvironment"]
    want = field.reshape(-1, ns).copy()
    want[:, 1] = 0.0
    assert np.array_equal(m[4:].T, want)


def test_malformed_csv_initial_conditions_are_rejected(tmp_path):
    nx, ny, nz, ns, dx = 4, 3, 2, 2, 10.0
    field = np.arange(nx * ny * nz * ns, dtype=float)
    good = ic_matrix(nx, ny, nz, ns, dx, field, np.arange(nx * ny * nz))
    cases = {"rows": (good[:-1], None), "repeat": (np.vstack([good[:-1], good[:1]]), None),
             "name": (good, "x,y,z,s0,nope"), "head": (good, "x,q,z,s0,s1"),
             "short": (good[:, :4], "x,y,z,s0,s1")}
    for name, (rows, header) in cases.items():
        p = tmp_path / f"{name}.csv"
        write_csv(p, rows, header)
        r = run([nx, ny, nz, ns, dx, p, tmp_path / "o.mat"])
        assert r.returncode == 1, (name, r.stdout)
        assert "ERROR" in r.stdout, name
    r = run([nx, ny, nz, ns, dx, tmp_path / "missing.csv", tmp_path / "o.mat"])
    assert r.returncode == 1 and "not found" in r.stdout


def test_read_matlab_integer_types(tmp_path):
    # level-4 files of float32 / int32 / int16 data load through the same reader
    nx, ny, nz, ns, dx = 3, 2, 2, 1, 4.0
    field = np.arange(nx * ny * nz * ns, dtype=float) + 1
    mat = ic_matrix(nx, ny, nz, ns, dx, field, np.arange(nx * ny * nz))
    for dt in (np.float32, np.int32, np.int16):
        conv = mat.astype(dt)
        p = tmp_path / f"ic_{np.dtype(dt).name}.mat"
        scipy.io.savemat(str(p), {"M": conv}, format="4")
        out = tmp_path / "o.mat"
        r = run([nx, ny, nz, ns, dx, p, out])
        assert r.returncode == 0, (dt, r.stdout + r.stderr)
        m = scipy.io.loadmat(str(out))["multiscale_microenvironment"]
        assert np.array_equal(m[4:].T.ravel(), conv[:, 3:].astype(float).ravel())
