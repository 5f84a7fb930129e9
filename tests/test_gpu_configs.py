"""GPU parity at the full sizes of the benchmark workloads, against the CPU oracle.

* 400^3 x 4 substrates (the headline workload): Dirichlet on all six faces,
  field = Dirichlet value + N(0, 1), 100,000 agents with secretion and uptake;
* 800^3 x 2 substrates: the oxygen-like and signal substrates with 1,000,000
  agents in a rim-weighted spheroid.

Both are the exact inputs bench.py times (its generators, its seeds).  After
every step the FP64 path must agree with the oracle to 1e-12 and the FP32
scaled-deviation path to 1e-5 (relative max-norm per substrate); the FP32 total
mass must agree to 1e-5 as well.  The oracle runs with all host threads.
"""
import numpy as np
import pytest

import bench
import oracle
from helpers import load_pkg

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-5


@pytest.fixture(scope="module")
def pkg():
    return load_pkg()


def relmax_chunked(a, b, chunk=1 << 24):
    """max|a - b| / max|b| without full-size temporaries."""
    num = 0.0
    den = 0.0
    for s in range(0, a.size, chunk):
        x, y = a[s:s + chunk], b[s:s + chunk]
        num = max(num, float(np.max(np.abs(x - y))))
        den = max(den, float(np.max(np.abs(y))))
    return num / max(den, 1e-300)


def per_substrate_errors(gpu_ref, orc_ref, ns):
    g = gpu_ref.reshape(-1, ns)
    u = orc_ref.reshape(-1, ns)
    return [relmax_chunked(np.ascontiguousarray(g[:, d]), np.ascontiguousarray(u[:, d])) for d in range(ns)]


def run_workload(pkg, n, ns, D, decay, dval, field_ref, voxel, c1, c2, e2, steps, label):
    orc = oracle.Oracle(n, n, n, ns, 20.0, D, decay, np.ones(ns, np.uint8), np.ones(6, np.uint8),
                        np.tile(dval, (6, 1)))
    ag = oracle.AgentSet(voxel, c1, c2, np.zeros_like(c1), e2)
    envs = {}
    try:
        for prec in ("fp64", "fp32"):
            e = pkg.Microenvironment(n, n, n, ns, 20.0, prec)
            e.set_coefficients(D, decay)
            e.set_dirichlet(np.ones(ns, np.uint8), np.ones(6, np.uint8), np.tile(dval, (6, 1)))
            e.upload(field_ref, "ref")
            e.set_agents(voxel, c1, c2, e2)
            envs[prec] = e
        u = field_ref.copy()
        buf = np.empty_like(u)
        for s in range(steps):
            orc.diffusion_step(u, 0.01)
            orc.secretion_uptake(u, ag)
            m_orc = orc.total_mass(u)
            for prec, tol in (("fp64", FP64_TOL), ("fp32", FP32_TOL)):
                e = envs[prec]
                e.step(0.01, 1)
                e.download("ref", out=buf)
                errs = per_substrate_errors(buf, u, ns)
                mass = e.total_mass()
                drift = [abs(mass[d] - m_orc[d]) / abs(m_orc[d]) if m_orc[d] else abs(mass[d]) for d in range(ns)]
                print(f"{label} step {s} {prec}: max rel err per substrate "
                      f"{[f'{x:.2e}' for x in errs]}, mass drift {[f'{x:.2e}' for x in drift]}", flush=True)
                assert max(errs) <= tol, f"{label} {prec} step {s}: {max(errs):.3e} > {tol:.0e}"
                assert max(drift) <= max(tol, 1e-12), f"{label} {prec} step {s}: mass drift {max(drift):.3e}"
    finally:
        for e in envs.values():
            e.close()


def test_config2_400cube_four_substrates_100k_agents(pkg):
    n, ns = 400, 4
    w = bench.make_workload(n, ns, 100_000, 2)
    # bench keeps the field in device (substrate-major) order; the oracle and
    # this upload use the reference order ((i*ny + j)*nz + k)*ns + d
    field_ref = np.ascontiguousarray(w["field"].reshape(ns, n, n, n).transpose(3, 2, 1, 0)).ravel()
    run_workload(pkg, n, ns, w["D"], w["decay"], w["dval"], field_ref, w["voxel"], w["c1"], w["c2"],
                 w["e2"], steps=3, label="400^3x4")


def test_config3_800cube_two_substrates_1m_spheroid_agents(pkg):
    w = bench.large_workload()
    n, ns = w["n"], w["ns"]
    field_ref = np.empty(n ** 3 * ns)
    r = field_ref.reshape(-1, ns)
    for d in range(ns):
        r[:, d] = w["dval"][d]
    run_workload(pkg, n, ns, w["D"], w["decay"], w["dval"], field_ref, w["voxel"], w["c1"], w["c2"],
                 w["e2"], steps=2, label="800^3x2")
