"""GPU tests of bulk sources and sinks (simulate_bulk_sources_and_sinks) against
the CPU oracle: FP64 bit-exact (same roundings), FP32 deviation form within
1e-6, rate layouts, and the error when no rates are set."""
import numpy as np
import pytest

import oracle
from helpers import gpu_env, load_pkg, make_problem, oracle_for, agents_for, relmax

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pkg():
    return load_pkg()


def rates(p, seed):
    rng = np.random.default_rng(seed)
    n = p.nx * p.ny * p.nz * p.ns
    S = rng.uniform(0, 5, n)
    T = rng.uniform(0, 40, n)
    U = rng.uniform(0, 5, n)
    S[::7] = 0.0
    U[::5] = 0.0
    return S, T, U


def to_soa(a, p):
    return np.ascontiguousarray(a.reshape(p.nx, p.ny, p.nz, p.ns).transpose(3, 2, 1, 0)).ravel()


@pytest.mark.parametrize("shape", [(17, 12, 9, 3), (40, 33, 21, 1), (1, 5, 3, 2)])
def test_bulk_fp64_bit_exact_both_layouts(pkg, shape):
    p = make_problem(*shape, seed=1000 + sum(shape), dirichlet="mixed")
    S, T, U = rates(p, 7)
    want = oracle.bulk_sources(p.u0.copy(), p.ns, 0.02, S, T, U)
    for layout in ("ref", "soa"):
        env = gpu_env(pkg, p, "fp64")
        try:
            if layout == "ref":
                env.set_bulk_rates(S, T, U)
            else:
                env.set_bulk_rates(to_soa(S, p), to_soa(T, p), to_soa(U, p), "soa")
            env.simulate_bulk_sources_and_sinks(0.02)
            assert np.array_equal(env.download(), want), layout
        finally:
            env.close()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_bulk_after_diffusion_steps(pkg, precision, tol):
    p = make_problem(24, 19, 15, 2, seed=1077, n_agents=60, dirichlet="all", export=2.0)
    S, T, U = rates(p, 9)
    orc = oracle_for(p)
    ag = agents_for(p)
    u = p.u0.copy()
    env = gpu_env(pkg, p, precision)
    try:
        env.set_bulk_rates(S, T, U)
        for _ in range(3):
            orc.diffusion_step(u, p.dt)
            orc.secretion_uptake(u, ag)
            oracle.bulk_sources(u, p.ns, p.dt, S, T, U)
            env.step(p.dt, 1)
            env.simulate_bulk_sources_and_sinks(p.dt)
        assert relmax(env.download(), u) <= tol
    finally:
        env.close()


def test_bulk_without_rates_is_an_error(pkg):
    p = make_problem(8, 7, 6, 1, seed=1099)
    env = gpu_env(pkg, p, "fp64")
    try:
        with pytest.raises(pkg.BioFVMError):
            env.simulate_bulk_sources_and_sinks(0.01)
        S, T, U = rates(p, 3)
        env.set_bulk_rates(S, T, U)
        env.simulate_bulk_sources_and_sinks(0.01)
        env.set_bulk_rates(None, None, None)
        with pytest.raises(pkg.BioFVMError):
            env.simulate_bulk_sources_and_sinks(0.01)
        with pytest.raises(pkg.BioFVMError):
            env.set_bulk_rates(S[:-1], T, U)
    finally:
        env.close()
