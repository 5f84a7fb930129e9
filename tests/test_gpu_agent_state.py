"""GPU tests of bfa_set_agents_state: agents given by their raw state
(position, volume, rates) with the voxel mapping, the uptake constants and the
per-voxel ordering computed on the device.  The result must equal the host
path (oracle voxel mapping and constants through bfa_set_agents) bit for bit
in FP64, and the oracle within the plain bars."""
import numpy as np
import pytest

import oracle
from helpers import load_pkg, make_problem, oracle_for, relmax

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pkg():
    return load_pkg()


def agent_state(p, n, seed):
    rng = np.random.default_rng(seed)
    half = 0.5 * p.dx * np.array([p.nx, p.ny, p.nz])
    pos = rng.uniform(-half, half, (n, 3))
    pos[3] = pos[4] = pos[5]          # shared voxels
    pos[7] = half * 1.2               # outside: takes no part
    pos[8] = -half                    # on the low corner: voxel 0
    pos[9] = half                     # on the high corner: last voxel
    act = np.ones(n, np.uint8)
    act[11] = 0
    st = dict(pos=pos, bbox=np.concatenate([-half, half]), V=rng.uniform(1000, 4000, n),
              S=rng.uniform(0, 10, (n, p.ns)), U=rng.uniform(0, 10, (n, p.ns)),
              E=rng.uniform(0, 20, (n, p.ns)), rs=rng.uniform(0, 40, (n, p.ns)), act=act)
    return st


def host_side(p, st):
    vox = oracle.voxel_indices(st["pos"], st["bbox"], p.dx, p.dx, p.dx, p.nx, p.ny, p.nz)
    c1, c2, e1, e2 = oracle.agent_constants(p.dt, p.dx ** 3, st["V"], st["S"], st["U"], st["E"], st["rs"])
    act = (st["act"] & (vox >= 0)).astype(np.uint8)
    return np.where(vox >= 0, vox, 0).astype(np.int32), act, c1, c2, e1, e2


def make_env(pkg, p, precision):
    env = pkg.Microenvironment(p.nx, p.ny, p.nz, p.ns, p.dx, precision)
    env.set_coefficients(p.D, p.decay)
    env.set_dirichlet(p.active, p.reg, p.fv)
    env.upload(p.u0)
    return env


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_agent_state_equals_host_path_and_oracle(pkg, precision, tol):
    p = make_problem(23, 18, 14, 3, seed=1200, dirichlet="mixed")
    st = agent_state(p, 500, 3)
    vox, act, c1, c2, e1, e2 = host_side(p, st)
    dev = make_env(pkg, p, precision)
    host = make_env(pkg, p, precision)
    try:
        dev.set_agents_state(st["pos"], st["bbox"], st["V"], st["S"], st["U"], st["E"], st["rs"], p.dt,
                             active=st["act"])
        host.set_agents(vox, c1, c2, e2, act)
        dev.step(p.dt, 3)
        host.step(p.dt, 3)
        got = dev.download()
        assert np.array_equal(got, host.download())
        orc = oracle_for(p)
        ag = oracle.AgentSet(vox, c1, c2, e1, e2, act)
        u = p.u0.copy()
        for _ in range(3):
            orc.diffusion_step(u, p.dt)
            orc.secretion_uptake(u, ag)
        assert relmax(got, u) <= tol
    finally:
        dev.close()
        host.close()


def test_agent_state_edge_cases(pkg):
    p = make_problem(9, 8, 7, 2, seed=1210, dirichlet="all")
    env = make_env(pkg, p, "fp64")
    ref = make_env(pkg, p, "fp64")
    try:
        st = agent_state(p, 40, 5)
        # nobody inside the box: no agent work at all
        env.set_agents_state(st["pos"] + 1e6, st["bbox"], st["V"], st["S"], st["U"], st["E"], st["rs"], p.dt)
        env.step(p.dt, 2)
        ref.step(p.dt, 2)
        assert np.array_equal(env.download(), ref.download())
        # an empty list clears the agents
        env.set_agents_state(np.zeros((0, 3)), st["bbox"], [], [], [], [], [], p.dt)
        env.step(p.dt, 1)
        ref.step(p.dt, 1)
        assert np.array_equal(env.download(), ref.download())
        with pytest.raises(pkg.BioFVMError):
            env.set_agents_state(st["pos"], st["bbox"], st["V"], st["S"], st["U"], st["E"], st["rs"], 0.0)
    finally:
        env.close()
        ref.close()
