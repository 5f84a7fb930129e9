"""Shared test helpers: package loader and seeded problem generators.

Problems follow the workloads the benchmark is quoted on (uniform grids with
dx = 20 um, dt = 0.01 min, D log-uniform, decay uniform, Dirichlet values on
all faces, uniformly placed agents), at sizes the CPU oracle finishes quickly.
"""
import importlib.util
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import oracle  # noqa: E402


def load_pkg():
    name = "physicellx_amd"
    if name in sys.modules:
        return sys.modules[name]
    path = os.path.join(ROOT, "physicell-x_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location(name, path,
                                                  submodule_search_locations=[os.path.dirname(path)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def relmax(a, b):
    """max|a-b| / max|b| (relative max-norm)."""
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


class Problem:
    pass


def make_problem(nx, ny, nz, ns, seed, n_agents=0, dirichlet="all", dt=0.01, dx=20.0,
                 noise=1.0, secretion=True, export=0.0, collide=0):
    """Seeded random problem.  dirichlet: "all" (one value per substrate on all
    six faces, all substrates active), "mixed" (some faces registered, per-face
    values, some substrates inactive) or "off"."""
    rng = np.random.default_rng(seed)
    p = Problem()
    p.nx, p.ny, p.nz, p.ns, p.dt, p.dx = nx, ny, nz, ns, dt, dx
    p.D = 10 ** rng.uniform(3, 5, ns)
    p.decay = rng.uniform(0.01, 1.0, ns)
    base = rng.uniform(0, 40, ns)
    if dirichlet == "all":
        p.active = np.ones(ns, np.uint8)
        p.reg = np.ones(6, np.uint8)
        p.fv = np.tile(base, (6, 1))
    elif dirichlet == "mixed":
        p.active = (np.arange(ns) % 2 == 0).astype(np.uint8)
        p.reg = np.array([1, 0, 1, 1, 0, 1], np.uint8)
        p.fv = rng.uniform(0, 40, (6, ns))
    else:
        p.active = np.zeros(ns, np.uint8)
        p.reg = np.zeros(6, np.uint8)
        p.fv = np.ones((6, ns))
    p.u0 = (base + noise * rng.standard_normal((nx, ny, nz, ns))).ravel()
    n = n_agents
    nvox = nx * ny * nz
    p.voxel = rng.integers(0, nvox, n).astype(np.int32)
    if collide and n > 2 * collide:
        # pile extra agents onto a few voxels, interleaved in index order
        hot = rng.integers(0, nvox, collide)
        p.voxel[: 2 * collide : 2] = hot
        p.voxel[1 : 2 * collide : 2] = hot[::-1]
    W = dx ** 3
    S = rng.uniform(0, 10, (n, ns)) if secretion else np.zeros((n, ns))
    U = rng.uniform(0, 10, (n, ns))
    E = np.full((n, ns), export) * rng.uniform(0, 1, (n, ns))
    rs = np.full((n, ns), 40.0)
    V = np.full(n, 2494.0)
    p.c1, p.c2, p.e1, p.e2 = oracle.agent_constants(dt, W, V, S, U, E, rs)
    p.agent_active = np.ones(n, np.uint8)
    if n > 4:
        p.agent_active[3] = 0
    return p


def oracle_for(p):
    return oracle.Oracle(p.nx, p.ny, p.nz, p.ns, p.dx, p.D, p.decay, p.active, p.reg, p.fv)


def agents_for(p):
    return oracle.AgentSet(p.voxel, p.c1, p.c2, p.e1, p.e2, p.agent_active)


def gpu_env(pkg, p, precision="fp64"):
    env = pkg.Microenvironment(p.nx, p.ny, p.nz, p.ns, p.dx, precision)
    env.set_coefficients(p.D, p.decay)
    env.set_dirichlet(p.active, p.reg, p.fv)
    env.upload(p.u0)
    if p.voxel.size:
        env.set_agents(p.voxel, p.c1, p.c2, p.e2, p.agent_active)
    return env
