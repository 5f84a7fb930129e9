"""The allocation-free entry point bfa_gradients_at_noalloc, called through the
C ABI, against bfa_gradients_at on the same resident field: bit-identical
output on two calls in a row (the second with fewer voxels, so stale scratch
contents past its end would show), in FP64 and FP32; sizes beyond the static
scratch are refused with an error and leave the output untouched."""
import ctypes

import numpy as np
import pytest

from helpers import gpu_env, load_pkg, make_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pkg():
    return load_pkg()


def abi(pkg):
    L = pkg.lib()
    L.bfa_gradients_at_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int32),
                                           ctypes.POINTER(ctypes.c_double)]
    L.bfa_gradients_at.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int32),
                                   ctypes.POINTER(ctypes.c_double)]
    L.bfa_noalloc_scratch_bytes.restype = ctypes.c_size_t
    L.bfa_last_error.restype = ctypes.c_char_p
    return L


def call(L, fn, env, vox, ns):
    out = np.full((vox.size, ns, 3), np.nan)
    rc = fn(env._h, vox.size, vox.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return rc, out


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_gradients_equal_the_allocating_entry_point(pkg, precision):
    L = abi(pkg)
    p = make_problem(37, 29, 41, 3, seed=51, n_agents=300, dirichlet="mixed")
    env = gpu_env(pkg, p, precision)
    try:
        env.step(p.dt, 3)
        rng = np.random.default_rng(52)
        nvox = p.nx * p.ny * p.nz
        for n in (5000, 700):  # the second call is smaller than the first
            vox = np.ascontiguousarray(rng.integers(0, nvox, n), np.int32)
            vox[:3] = [0, nvox - 1, 0]  # corners, repeated entries
            rc_a, ref = call(L, L.bfa_gradients_at, env, vox, p.ns)
            rc_b, got = call(L, L.bfa_gradients_at_noalloc, env, vox, p.ns)
            assert rc_a == 0 and rc_b == 0, L.bfa_last_error(env._h)
            assert np.array_equal(got, ref), f"n={n}"
        # beyond the static scratch: an error, nothing written
        limit = L.bfa_noalloc_scratch_bytes()
        big = int(limit // (p.ns * 3 * 8)) + 1
        vox = np.zeros(big, np.int32)
        rc, out = call(L, L.bfa_gradients_at_noalloc, env, vox, p.ns)
        assert rc == 1 and b"static scratch" in L.bfa_last_error(env._h)
        assert np.all(np.isnan(out))
    finally:
        env.close()
