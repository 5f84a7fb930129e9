"""GPU tests of the gradient field (compute_all_gradient_vectors) against the
CPU oracle: FP64 bit-exact (one subtraction and one division per value, as in
the reference), FP32 from the deviation form within 1e-9 of the field's
largest gradient, the voxel-list form equal to the whole-field form."""
import numpy as np
import pytest

import oracle
from helpers import gpu_env, load_pkg, make_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pkg():
    return load_pkg()


SHAPES = [(13, 9, 7, 2), (5, 1, 4, 1), (2, 3, 1, 3), (64, 40, 33, 4), (1, 6, 5, 2)]


@pytest.mark.parametrize("shape", SHAPES)
def test_gradients_fp64_bit_exact(pkg, shape):
    p = make_problem(*shape, seed=900 + sum(shape), n_agents=20, dirichlet="mixed", export=3.0)
    env = gpu_env(pkg, p, "fp64")
    try:
        g0 = env.gradients()
        assert np.array_equal(g0, oracle.gradients(p.u0, p.nx, p.ny, p.nz, p.ns, p.dx))
        env.step(p.dt, 3)
        u = env.download()
        g = env.gradients()
        assert np.array_equal(g, oracle.gradients(u, p.nx, p.ny, p.nz, p.ns, p.dx))
        soa = env.gradients("soa")  # [q][c][(k*ny + j)*nx + i]
        ref_from_soa = soa.reshape(p.ns, 3, p.nz, p.ny, p.nx).transpose(4, 3, 2, 0, 1).reshape(-1, p.ns, 3)
        assert np.array_equal(ref_from_soa, g)
    finally:
        env.close()


@pytest.mark.parametrize("shape", SHAPES[:4])
def test_gradients_fp32(pkg, shape):
    p = make_problem(*shape, seed=950 + sum(shape), n_agents=20, dirichlet="all", export=3.0)
    env = gpu_env(pkg, p, "fp32")
    try:
        env.step(p.dt, 2)
        u = env.download()
        g = env.gradients()
        ref = oracle.gradients(u, p.nx, p.ny, p.nz, p.ns, p.dx)
        scale = max(np.max(np.abs(ref)), 1e-30)
        assert np.max(np.abs(g - ref)) <= 1e-9 * scale
    finally:
        env.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_gradients_at_voxels_match_whole_field(pkg, precision):
    p = make_problem(21, 17, 12, 3, seed=977, n_agents=40, dirichlet="mixed", export=3.0)
    env = gpu_env(pkg, p, precision)
    try:
        env.step(p.dt, 2)
        g = env.gradients()
        rng = np.random.default_rng(5)
        vox = rng.integers(0, p.nx * p.ny * p.nz, 300).astype(np.int32)
        vox[:8] = [0, p.nx * p.ny * p.nz - 1, p.nz - 1, p.ny * p.nz - 1, 1, p.nz, p.ny * p.nz, 5]
        assert np.array_equal(env.gradients_at(vox), g[vox])
        assert env.gradients_at(np.zeros(0, np.int32)).shape == (0, p.ns, 3)
        with pytest.raises(pkg.BioFVMError):
            env.gradients_at([p.nx * p.ny * p.nz])
        with pytest.raises(pkg.BioFVMError):
            env.gradients_at([-1])
    finally:
        env.close()
