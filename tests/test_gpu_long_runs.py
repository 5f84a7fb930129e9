"""FP32 scaled-deviation form over long runs, against the FP64 oracle.

Near a steady state an FP32 solve whose rounding perturbs the step operator
moves the fixed point by about eps32 / (decay * dt) relative; the sweeps solve
for the increment from the residual (delta form) so that the fixed point is
set by the residual alone.  These runs take the served workload's two smallest
requests (the bench's FP64 spot check) to 1,000 steps, the most a served
request asks for, on both FP32 paths (a microenvironment of its own, and a
batch member), and hold them to 1e-5 relative max-norm per substrate, with
total mass to 1e-5, at every checkpoint.
"""
import numpy as np
import pytest

import bench
import oracle
from helpers import load_pkg, relmax

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-5
CHECKPOINTS = (250, 500, 1000)


@pytest.fixture(scope="module")
def pkg():
    return load_pkg()


def request(i):
    reqs, _ = bench.serving_requests(40, 5)
    return reqs[i]


def oracle_run(r):
    nx, ny, nz, ns = r["nx"], r["ny"], r["nz"], r["ns"]
    orc = oracle.Oracle(nx, ny, nz, ns, 20.0, r["D"], r["decay"], np.ones(ns, np.uint8), np.ones(6, np.uint8),
                        np.tile(r["dval"], (6, 1)))
    ag = oracle.AgentSet(r["voxel"], r["c1"], r["c2"], np.zeros_like(r["c1"]), r["e2"])
    u = np.tile(r["dval"], nx * ny * nz).astype(np.float64)
    out = {}
    for s in range(1, CHECKPOINTS[-1] + 1):
        orc.diffusion_step(u, 0.01)
        orc.secretion_uptake(u, ag)
        if s in CHECKPOINTS:
            out[s] = (u.copy(), orc.total_mass(u))
    return out


def check(r, got, mass, ref, label):
    ns = r["ns"]
    u, m = ref
    g, uu = got.reshape(-1, ns), u.reshape(-1, ns)
    errs = [relmax(g[:, d], uu[:, d]) for d in range(ns)]
    drift = [abs(mass[d] - m[d]) / abs(m[d]) for d in range(ns)]
    print(f"{label}: max rel err {max(errs):.2e}, mass drift {max(drift):.2e}", flush=True)
    assert max(errs) <= FP32_TOL, f"{label}: {max(errs):.3e}"
    assert max(drift) <= FP32_TOL, f"{label}: mass drift {max(drift):.3e}"


@pytest.mark.parametrize("req", [31, 17])
def test_fp32_holds_1e5_over_1000_steps(pkg, req):
    r = request(req)
    ns = r["ns"]
    refs = oracle_run(r)
    env = pkg.Microenvironment(r["nx"], r["ny"], r["nz"], ns, 20.0, "fp32")
    b = pkg.Batch("fp32")
    try:
        env.set_coefficients(r["D"], r["decay"])
        env.set_dirichlet(np.ones(ns, np.uint8), np.ones(6, np.uint8), np.tile(r["dval"], (6, 1)))
        env.fill(r["dval"])
        env.set_agents(r["voxel"], r["c1"], r["c2"], r["e2"])
        mid = b.add(r["nx"], r["ny"], r["nz"], ns, 20.0, r["D"], r["decay"], np.ones(ns, np.uint8),
                    np.ones(6, np.uint8), np.tile(r["dval"], (6, 1)))
        b.fill(mid, r["dval"])
        b.set_agents(mid, r["voxel"], r["c1"], r["c2"], r["e2"])
        done = 0
        for cp in CHECKPOINTS:
            env.step(0.01, cp - done)
            b.step(0.01, cp - done)
            done = cp
            check(r, env.download(), env.total_mass(), refs[cp], f"request {req} solo, {cp} steps")
            check(r, b.download(mid), b.total_mass(mid), refs[cp], f"request {req} batch member, {cp} steps")
    finally:
        env.close()
        b.close()
