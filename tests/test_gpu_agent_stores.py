"""GPU tests of the agent step's two store forms (bfa_set_agent_stores):
whole 128-byte lines and single voxels (the default) give bit for bit the
same field and internalised totals, in FP64 and FP32, with agents piled onto
shared voxels and lines, on a grid whose substrate blocks are not a whole
number of lines (so the first and last lines of a substrate are shared with
its neighbours), with agents on the first and last voxels, through bfa_step
and bfa_secretion_uptake, and with each substrate half's agent step on its
own stream."""
import numpy as np
import pytest

from helpers import gpu_env, load_pkg, make_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pkg():
    return load_pkg()


def problem():
    p = make_problem(23, 19, 17, 3, seed=2200, n_agents=1500, dirichlet="mixed", collide=8, export=0.4)
    n = p.nx * p.ny * p.nz
    assert (n * 8) % 128 and (n * 4) % 128  # substrate blocks end inside a line
    p.voxel[20:24] = 0          # the first voxel, shared
    p.voxel[24:27] = n - 1      # the last voxel, shared
    p.voxel[27] = 1             # the first voxel's line neighbour
    p.voxel[28] = n - 2
    return p


def run(pkg, p, precision, lines, eager=False, split=1):
    env = gpu_env(pkg, p, precision)
    try:
        env.set_substrate_streams(split)
        env.set_agent_stores(lines)
        assert env.agent_stores() == lines
        env.track_internalized(np.linspace(0, 1, p.voxel.size * p.ns), p.e1)
        if eager:
            for _ in range(3):
                env.simulate_diffusion_decay(p.dt)
                env.simulate_cell_sources_and_sinks()
        else:
            env.step(p.dt, 3)
        return env.download(), env.internalized()
    finally:
        env.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("eager", [False, True])
def test_whole_line_and_single_voxel_stores_agree(pkg, precision, eager):
    p = problem()
    a, ta = run(pkg, p, precision, True, eager)
    b, tb = run(pkg, p, precision, False, eager)
    assert np.array_equal(a, b)
    assert np.array_equal(ta, tb)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("lines", [False, True])
def test_agents_per_substrate_half_agree(pkg, precision, lines):
    """On two substrate streams each half's agent step follows its own z
    sweep (and overlaps the other half's sweeps); the result equals one
    agent step after all sweeps, with substrate blocks sharing lines."""
    p = problem()
    a, ta = run(pkg, p, precision, lines, split=2)
    b, tb = run(pkg, p, precision, lines, split=1)
    assert np.array_equal(a, b)
    assert np.array_equal(ta, tb)


def test_agent_store_setting_is_checked(pkg):
    p = make_problem(8, 8, 8, 1, seed=2201, n_agents=10)
    env = gpu_env(pkg, p, "fp64")
    try:
        assert not env.agent_stores()  # default: single voxels
        env.set_agent_stores(True)
        assert env.agent_stores()
        assert pkg.lib().bfa_set_agent_stores(env._h, 2) != 0  # only 0 or 1
        assert env.agent_stores()
    finally:
        env.close()
