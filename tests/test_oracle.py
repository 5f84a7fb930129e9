"""Pins the CPU oracle (oracle/biofvm_oracle.c) before anything is compared to it.

Every check here comes from outside the oracle's own arithmetic: exact-rational
fixtures (tests/golden/lod_exact.json), dense linear solves with numpy,
Neumann cosine eigenmodes, the D=0 and decay=0 closed forms, the backward-Euler
residual of the agent update and the internalised-substrate invariant.
"""
import json
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "lod_exact.json")


def unhex(x):
    if isinstance(x, list):
        return [unhex(v) for v in x]
    return float.fromhex(x) if isinstance(x, str) else x


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def axis_matrix(n, k, lp):
    """Dense LOD axis matrix: diag 1+2k+l' (1+k+l' at the ends, 1+l' if n == 1), off -k."""
    if n == 1:
        return np.array([[1.0 + lp]])
    A = np.diag(np.full(n, 1 + 2 * k + lp)) - k * np.eye(n, k=1) - k * np.eye(n, k=-1)
    A[0, 0] = A[-1, -1] = 1 + k + lp
    return A


def dense_lod(u, nx, ny, nz, ns, D, decay, dt, dx, active=None, table=None, final=True):
    """One LOD step by dense solves; u is [nx, ny, nz, ns]; table is the per-voxel
    Dirichlet value array [nx, ny, nz, ns] (only boundary entries are used);
    final=False leaves out the Dirichlet pass after the last axis."""
    u = u.copy()
    bmask = np.zeros((nx, ny, nz), bool)
    bmask[[0, -1], :, :] = True
    bmask[:, [0, -1], :] = True
    bmask[:, :, [0, -1]] = True

    def dirichlet():
        if active is None:
            return
        for d in range(ns):
            if active[d]:
                u[..., d][bmask] = table[..., d][bmask]

    dirichlet()
    for axis, n in enumerate((nx, ny, nz)):
        for d in range(ns):
            k = D[d] * dt / dx ** 2
            lp = decay[d] * dt / 3
            Ainv = np.linalg.inv(axis_matrix(n, k, lp))
            u[..., d] = np.moveaxis(np.tensordot(Ainv, np.moveaxis(u[..., d], axis, 0), 1), 0, axis)
        if axis < 2 or final:
            dirichlet()
    return u


def test_golden_thomas_coefficients():
    doc = json.load(open(GOLDEN))
    for t in doc["thomas"]:
        consts = oracle.thomas_constants([unhex(t["D"])], [unhex(t["decay"])],
                                         unhex(t["dt"]), unhex(t["dx"]))
        denom, cp = oracle.thomas_axis(t["n"], consts)
        assert relmax(denom[:, 0], np.array(unhex(t["denom"]))) < 1e-15
        assert relmax(cp[:, 0], np.array(unhex(t["cp"]))) < 1e-15


@pytest.mark.parametrize("case_index", [0, 1, 2])
def test_golden_lod_and_agent_steps(case_index):
    doc = json.load(open(GOLDEN))
    c = doc["cases"][case_index]
    nx, ny, nz, ns = c["nx"], c["ny"], c["nz"], c["ns"]
    orc = oracle.Oracle(nx, ny, nz, ns, unhex(c["dx"]), unhex(c["D"]), unhex(c["decay"]),
                        c["dirichlet_active"], c["face_registered"], np.array(unhex(c["face_values"])))
    ag = c["agents"]
    agents = None
    if ag["voxel"]:
        dt, W = unhex(c["dt"]), unhex(c["dx"]) ** 3
        c1, c2, e1, e2 = oracle.agent_constants(dt, W, unhex(ag["V"]), unhex(ag["S"]),
                                                unhex(ag["U"]), unhex(ag["E"]),
                                                unhex(ag["rho_star"]))
        agents = oracle.AgentSet(ag["voxel"], c1, c2, e1, e2, ag["active"])
    for step in c["steps"]:
        u = np.array(unhex(step["input"]))
        orc.diffusion_step(u, unhex(c["dt"]))
        if agents is not None:
            orc.secretion_uptake(u, agents)
        assert relmax(u, np.array(unhex(step["expected"]))) < 1e-13


def test_golden_catches_missing_intermediate_dirichlet():
    """A plausible restatement slip (no Dirichlet pass between the x and y sweeps)
    must fail the fixture: run the fixture case with a solver that skips it."""
    doc = json.load(open(GOLDEN))
    c = doc["cases"][0]
    nx, ny, nz, ns = c["nx"], c["ny"], c["nz"], c["ns"]
    dt, dx = unhex(c["dt"]), unhex(c["dx"])
    orc = oracle.Oracle(nx, ny, nz, ns, dx, unhex(c["D"]), unhex(c["decay"]),
                        c["dirichlet_active"], c["face_registered"], np.array(unhex(c["face_values"])))
    u = np.array(unhex(c["steps"][0]["input"]))
    inactive = oracle.Oracle(nx, ny, nz, ns, dx, unhex(c["D"]), unhex(c["decay"]))
    orc.apply_dirichlet(u)
    inactive.diffusion_step(u, dt)  # all three sweeps with no Dirichlet in between
    orc.apply_dirichlet(u)
    exp = np.array(unhex(c["steps"][0]["expected"]))
    # the fixture's expected output also includes the agents; compare away from them
    assert relmax(u, exp) > 1e-3


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("n", [1, 2, 5, 33])
def test_single_axis_against_dense_solve(axis, n):
    dims = [1, 1, 1]
    dims[axis] = n
    nx, ny, nz = dims
    ns = 2
    rng = np.random.default_rng(7 + n + axis)
    D = 10 ** rng.uniform(3, 5, ns)
    decay = rng.uniform(0.01, 1, ns)
    dt, dx = 0.01, 20.0
    u = rng.uniform(0, 40, nx * ny * nz * ns)
    orc = oracle.Oracle(nx, ny, nz, ns, dx, D, decay)
    got = u.copy()
    orc.diffusion_step(got, dt)
    exp = dense_lod(u.reshape(nx, ny, nz, ns), nx, ny, nz, ns, D, decay, dt, dx).ravel()
    assert relmax(got, exp) < 1e-13


def test_full_step_with_dirichlet_against_dense_solve():
    nx, ny, nz, ns = 6, 7, 8, 2
    rng = np.random.default_rng(11)
    D = 10 ** rng.uniform(3, 5, ns)
    decay = rng.uniform(0.01, 1, ns)
    dt, dx = 0.01, 20.0
    reg = np.array([1, 0, 1, 1, 0, 1], np.uint8)
    fv = rng.uniform(0, 40, (6, ns))
    active = np.array([1, 1], np.uint8)
    # per-voxel table built from the face registration order, last face wins
    table = np.ones((nx, ny, nz, ns))
    sl = [np.s_[0, :, :], np.s_[-1, :, :], np.s_[:, 0, :], np.s_[:, -1, :], np.s_[:, :, 0], np.s_[:, :, -1]]
    for f in range(6):
        if reg[f]:
            table[sl[f]] = fv[f]
    u = rng.uniform(0, 40, nx * ny * nz * ns)
    orc = oracle.Oracle(nx, ny, nz, ns, dx, D, decay, active, reg, fv)
    got = u.copy()
    orc.diffusion_step(got, dt)
    exp = dense_lod(u.reshape(nx, ny, nz, ns), nx, ny, nz, ns, D, decay, dt, dx, active, table).ravel()
    assert relmax(got, exp) < 1e-13


def test_explicit_per_voxel_planes_against_dense_solve():
    """Per-voxel boundary values (nodes overridden inside faces, on edges and
    corners) given as planes reproduce a dense step with that value table."""
    nx, ny, nz, ns = 6, 7, 5, 2
    rng = np.random.default_rng(13)
    D = 10 ** rng.uniform(3, 5, ns)
    decay = rng.uniform(0.01, 1, ns)
    dt, dx = 0.01, 20.0
    active = np.array([1, 1], np.uint8)
    table = rng.uniform(0, 40, (nx, ny, nz, ns))
    t = table
    planes = [t[0], t[-1], t[:, 0], t[:, -1], t[:, :, 0], t[:, :, -1]]
    u = rng.uniform(0, 40, nx * ny * nz * ns)
    orc = oracle.Oracle(nx, ny, nz, ns, dx, D, decay, active, np.ones(6, np.uint8), np.ones((6, ns)))
    orc.set_planes(planes)
    got = u.copy()
    orc.diffusion_step(got, dt)
    exp = dense_lod(u.reshape(nx, ny, nz, ns), nx, ny, nz, ns, D, decay, dt, dx, active, table).ravel()
    assert relmax(got, exp) < 1e-13
    # the face-table form is the special case of the same planes
    planes2 = oracle.dirichlet_planes(nx, ny, nz, ns, np.ones(6, np.uint8), rng.uniform(0, 40, (6, ns)))
    orc.set_planes(planes2)
    a = u.copy()
    orc.diffusion_step(a, dt)
    assert relmax(a, got) > 1e-6


def test_avx256d_variant_skips_the_last_dirichlet_pass():
    nx, ny, nz, ns = 6, 5, 7, 2
    rng = np.random.default_rng(14)
    D = 10 ** rng.uniform(3, 5, ns)
    decay = rng.uniform(0.01, 1, ns)
    dt, dx = 0.01, 20.0
    reg = np.array([1, 1, 0, 1, 1, 0], np.uint8)
    fv = rng.uniform(0, 40, (6, ns))
    active = np.array([1, 0], np.uint8)
    table = np.ones((nx, ny, nz, ns))
    sl = [np.s_[0, :, :], np.s_[-1, :, :], np.s_[:, 0, :], np.s_[:, -1, :], np.s_[:, :, 0], np.s_[:, :, -1]]
    for f in range(6):
        if reg[f]:
            table[sl[f]] = fv[f]
    u = rng.uniform(0, 40, nx * ny * nz * ns)
    avx = oracle.Oracle(nx, ny, nz, ns, dx, D, decay, active, reg, fv, solver="avx256d")
    got = u.copy()
    avx.diffusion_step(got, dt)
    exp = dense_lod(u.reshape(nx, ny, nz, ns), nx, ny, nz, ns, D, decay, dt, dx, active, table,
                    final=False).ravel()
    assert relmax(got, exp) < 1e-13
    # and it differs from BLOCKING exactly on the active substrate's faces
    blk = oracle.Oracle(nx, ny, nz, ns, dx, D, decay, active, reg, fv)
    b = u.copy()
    blk.diffusion_step(b, dt)
    diff = np.abs(b - got).reshape(nx, ny, nz, ns)
    assert np.max(diff[1:-1, 1:-1, 1:-1]) == 0.0 and np.max(diff[..., 1]) == 0.0
    assert np.max(diff[..., 0]) > 1e-6


def test_internalised_tracking_per_agent_matches_its_density_change():
    """Each agent's internalised change is minus the substrate it moved:
    -W * (density after its update - density before), the export included
    (e2 = e1 / W), for agents sharing voxels, in index order."""
    ns, dt, dx = 2, 0.01, 20.0
    W = dx ** 3
    rng = np.random.default_rng(15)
    n = 12
    voxel = np.array([3, 3, 5, 3, 7, 5, 1, 3, 9, 9, 2, 3], np.int32)
    c1, c2, e1, e2 = oracle.agent_constants(dt, W, rng.uniform(1000, 4000, n), rng.uniform(0, 10, (n, ns)),
                                            rng.uniform(0, 10, (n, ns)), rng.uniform(0, 20, (n, ns)),
                                            rng.uniform(0, 40, (n, ns)))
    u = rng.uniform(0, 40, 10 * ns)
    internal = np.zeros((n, ns))
    orc = oracle.Oracle(10, 1, 1, ns, dx, [0.0, 0.0], [0.0, 0.0])
    for a in range(n):
        one = oracle.AgentSet(voxel[a:a + 1], c1[a:a + 1], c2[a:a + 1], e1[a:a + 1], e2[a:a + 1])
        before = u[voxel[a] * ns:(voxel[a] + 1) * ns].copy()
        mine = np.zeros((1, ns))
        orc.secretion_uptake(u, one, mine)
        after = u[voxel[a] * ns:(voxel[a] + 1) * ns]
        np.testing.assert_allclose(mine[0], -W * (after - before), rtol=1e-9, atol=1e-9)
        internal[a] = mine[0]
    # the whole set in one call gives the same per-agent totals
    u2 = u.copy()
    u2[:] = rng.uniform(0, 40, 10 * ns)
    v = u2.copy()
    tot = np.zeros((n, ns))
    orc.secretion_uptake(v, oracle.AgentSet(voxel, c1, c2, e1, e2), tot)
    seq = np.zeros((n, ns))
    w = u2.copy()
    for a in range(n):
        one = oracle.AgentSet(voxel[a:a + 1], c1[a:a + 1], c2[a:a + 1], e1[a:a + 1], e2[a:a + 1])
        m = np.zeros((1, ns))
        orc.secretion_uptake(w, one, m)
        seq[a] = m[0]
    assert np.array_equal(tot, seq) and np.array_equal(v, w)


def test_neumann_cosine_eigenmode_decay():
    nx, ny, nz, ns = 24, 20, 16, 2
    D = np.array([1e5, 3e3])
    decay = np.array([0.1, 0.7])
    dt, dx = 0.01, 20.0
    m = (3, 1, 2)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    mode = (np.cos(np.pi * m[0] * (i + 0.5) / nx) * np.cos(np.pi * m[1] * (j + 0.5) / ny)
            * np.cos(np.pi * m[2] * (k + 0.5) / nz))
    u = np.stack([mode * 5.0, mode * -2.0], axis=-1).ravel().copy()
    orc = oracle.Oracle(nx, ny, nz, ns, dx, D, decay)
    steps = 3
    for _ in range(steps):
        orc.diffusion_step(u, dt)
    for d in range(ns):
        kk = D[d] * dt / dx ** 2
        lp = decay[d] * dt / 3
        mu = np.prod([1 + lp + 2 * kk * (1 - np.cos(np.pi * mm / n)) for mm, n in zip(m, (nx, ny, nz))])
        exp = mode * [5.0, -2.0][d] / mu ** steps
        assert relmax(u.reshape(nx, ny, nz, ns)[..., d], exp) < 1e-12


def test_zero_diffusion_is_pure_decay():
    nx, ny, nz, ns = 9, 8, 7, 3
    rng = np.random.default_rng(3)
    decay = rng.uniform(0.01, 1, ns)
    u = rng.uniform(0, 40, nx * ny * nz * ns)
    orc = oracle.Oracle(nx, ny, nz, ns, 20.0, np.zeros(ns), decay)
    got = u.copy()
    orc.diffusion_step(got, 0.01)
    exp = (u.reshape(-1, ns) / (1 + decay * 0.01 / 3) ** 3).ravel()
    assert relmax(got, exp) < 1e-15


def test_zero_decay_conserves_mass():
    nx, ny, nz, ns = 13, 11, 9, 2
    rng = np.random.default_rng(5)
    u = rng.uniform(0, 40, nx * ny * nz * ns)
    orc = oracle.Oracle(nx, ny, nz, ns, 20.0, [1e5, 2e3], [0.0, 0.0])
    m0 = orc.total_mass(u)
    for _ in range(4):
        orc.diffusion_step(u, 0.01)
    assert np.all(np.abs(orc.total_mass(u) - m0) / m0 < 1e-13)


def test_agent_update_is_backward_euler():
    """The uptake/secretion part is one backward-Euler step of
    d(rho)/dt = V/W*(S*(rho* - rho) - U*rho); net export E*dt/W is then added
    explicitly (BioFVM/BioFVM_basic_agent.cpp:536-552)."""
    rng = np.random.default_rng(9)
    ns, dt, W = 3, 0.01, 8000.0
    V = np.array([2494.0])
    S, U, E, rs = (rng.uniform(0, 10, (1, ns)) for _ in range(4))
    c1, c2, e1, e2 = oracle.agent_constants(dt, W, V, S, U, E, rs)
    orc = oracle.Oracle(3, 3, 3, ns, 20.0, np.zeros(ns), np.zeros(ns))
    u = rng.uniform(0, 40, 27 * ns)
    rho0 = u.reshape(27, ns)[13].copy()
    orc.secretion_uptake(u, oracle.AgentSet([13], c1, c2, e1, e2))
    rho1 = u.reshape(27, ns)[13]
    rho_imp = rho1 - E[0] * dt / W
    lhs = (rho_imp - rho0) / dt
    rhs = V[0] / W * (S[0] * (rs[0] - rho_imp) - U[0] * rho_imp)
    assert np.max(np.abs(lhs - rhs)) / np.max(np.abs(lhs)) < 1e-10
    assert np.allclose(e1, E * dt)


def test_agents_apply_in_index_order():
    rng = np.random.default_rng(10)
    ns, dt, W = 2, 0.01, 8000.0
    n = 2
    c1, c2, e1, e2 = oracle.agent_constants(dt, W, np.full(n, 3000.0), rng.uniform(1, 10, (n, ns)),
                                            rng.uniform(1, 10, (n, ns)), rng.uniform(10, 50, (n, ns)),
                                            rng.uniform(10, 40, (n, ns)))
    orc = oracle.Oracle(3, 3, 3, ns, 20.0, np.zeros(ns), np.zeros(ns))
    u0 = rng.uniform(0, 40, 27 * ns)
    both = u0.copy()
    orc.secretion_uptake(both, oracle.AgentSet([4, 4], c1, c2, e1, e2))
    seq = u0.copy()
    orc.secretion_uptake(seq, oracle.AgentSet([4], c1[:1], c2[:1], e1[:1], e2[:1]))
    orc.secretion_uptake(seq, oracle.AgentSet([4], c1[1:], c2[1:], e1[1:], e2[1:]))
    rev = u0.copy()
    orc.secretion_uptake(rev, oracle.AgentSet([4], c1[1:], c2[1:], e1[1:], e2[1:]))
    orc.secretion_uptake(rev, oracle.AgentSet([4], c1[:1], c2[:1], e1[:1], e2[:1]))
    assert np.array_equal(both, seq)
    assert np.max(np.abs(both - rev)) > 1e-6
    # an inactive agent is skipped
    skip = u0.copy()
    orc.secretion_uptake(skip, oracle.AgentSet([4, 4], c1, c2, e1, e2, [0, 1]))
    only2 = u0.copy()
    orc.secretion_uptake(only2, oracle.AgentSet([4], c1[1:], c2[1:], e1[1:], e2[1:]))
    assert np.array_equal(skip, only2)


def test_internalised_tracking_conserves_total_substrate():
    """decay 0, Dirichlet off: field mass + internalised mass is invariant
    (documentation/math/substrate-internalization-equations.tex:133-165)."""
    nx, ny, nz, ns = 10, 9, 8, 2
    rng = np.random.default_rng(12)
    dt, dx = 0.01, 20.0
    W = dx ** 3
    n = 40
    voxel = rng.integers(0, nx * ny * nz, n).astype(np.int32)
    voxel[5] = voxel[6] = voxel[7]
    c1, c2, e1, e2 = oracle.agent_constants(dt, W, rng.uniform(1000, 4000, n), rng.uniform(0, 10, (n, ns)),
                                            rng.uniform(0, 10, (n, ns)), rng.uniform(0, 20, (n, ns)),
                                            rng.uniform(0, 40, (n, ns)))
    agents = oracle.AgentSet(voxel, c1, c2, e1, e2)
    orc = oracle.Oracle(nx, ny, nz, ns, dx, [1e5, 5e3], [0.0, 0.0])
    u = rng.uniform(0, 40, nx * ny * nz * ns)
    internal = np.zeros((n, ns))
    total0 = orc.total_mass(u) + internal.sum(0)
    for _ in range(5):
        orc.diffusion_step(u, dt)
        orc.secretion_uptake(u, agents, internal)
    total = orc.total_mass(u) + internal.sum(0)
    assert np.all(np.abs(total - total0) / np.abs(total0) < 1e-12)
    assert np.max(np.abs(internal)) > 1.0  # the agents did move substrate


def test_voxel_index_mapping():
    nx, ny, nz, dx = 5, 4, 3, 20.0
    bbox = [-50.0, -40.0, -30.0, 50.0, 40.0, 30.0]
    assert oracle.mesh_nodes(-50.0, 50.0, dx) == nx
    assert oracle.mesh_nodes(0.0, 2000.0, 20.0) == 100
    centers = []
    expect = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                centers.append([bbox[0] + (i + 0.5) * dx, bbox[1] + (j + 0.5) * dx, bbox[2] + (k + 0.5) * dx])
                expect.append((i * ny + j) * nz + k)
    got = oracle.voxel_indices(np.array(centers), bbox, dx, dx, dx, nx, ny, nz)
    assert np.array_equal(got, expect)
    edge = oracle.voxel_indices(np.array([[50.0, 40.0, 30.0], [-50.0, -40.0, -30.0], [50.1, 0, 0]]),
                                bbox, dx, dx, dx, nx, ny, nz)
    assert list(edge) == [(4 * ny + 3) * nz + 2, 0, -1]


# ---- gradient field (compute_all_gradient_vectors) ---------------------------------

def _grid_field(nx, ny, nz, ns, f, dx=20.0):
    """Reference-layout field with u(i,j,k,q) = f(x, y, z, q) at node coordinates."""
    i, j, k, q = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), np.arange(ns), indexing="ij")
    return f(i * dx, j * dx, k * dx, q).astype(float).ravel()


def test_gradient_of_linear_field_is_exact_everywhere():
    # central and one-sided differences are exact for linear functions; a
    # per-substrate factor catches a substrate/component mix-up
    nx, ny, nz, ns, dx = 6, 5, 4, 3, 20.0
    a, b, c = 0.75, -1.5, 0.125
    u = _grid_field(nx, ny, nz, ns, lambda x, y, z, q: (q + 1) * (a * x + b * y + c * z) + 3.0, dx)
    g = oracle.gradients(u, nx, ny, nz, ns, dx)
    want = np.array([[(q + 1) * a, (q + 1) * b, (q + 1) * c] for q in range(ns)])
    assert np.max(np.abs(g - want[None])) < 1e-12


def test_gradient_of_quadratic_field_interior_and_ends():
    # u = x^2 + 3y^2 - z^2: the central difference is exact inside (2x); the
    # forward difference at the low end is f'(0) + h, the backward one at the
    # high end f'(L) - h (times the second-derivative factor of each term)
    nx, ny, nz, ns, h = 7, 6, 5, 1, 10.0
    u = _grid_field(nx, ny, nz, ns, lambda x, y, z, q: x * x + 3 * y * y - z * z, h)
    g = oracle.gradients(u, nx, ny, nz, ns, h).reshape(nx, ny, nz, 3)
    for axis, n, coef in ((0, nx, 1.0), (1, ny, 3.0), (2, nz, -1.0)):
        comp = np.moveaxis(g[..., axis], axis, 0)
        pos = np.arange(n) * h
        expect = 2 * coef * pos
        expect[0] = coef * h          # (u(h) - u(0)) / h
        expect[-1] = coef * (2 * pos[-1] - h)  # (u(L) - u(L-h)) / h
        assert np.allclose(comp, expect.reshape(-1, *([1] * (comp.ndim - 1))), rtol=0, atol=1e-9)


def test_gradient_short_axes():
    # an axis of 2 nodes: both ends take the one-sided difference; an axis
    # of 1 node gets 0 (the reference skips z when nz == 1)
    nx, ny, nz, ns, dx = 2, 3, 1, 1, 5.0
    u = _grid_field(nx, ny, nz, ns, lambda x, y, z, q: 4 * x + y * y, dx)
    g = oracle.gradients(u, nx, ny, nz, ns, dx).reshape(nx, ny, 3)
    assert np.allclose(g[:, :, 0], 4.0)
    assert np.allclose(g[:, 1, 1], 2 * 5.0)                    # central: 2y
    assert np.allclose(g[:, 0, 1], 5.0) and np.allclose(g[:, 2, 1], 2 * 10.0 - 5.0)
    assert np.all(g[:, :, 2] == 0.0)


# ---- bulk sources and sinks (simulate_bulk_sources_and_sinks) -----------------------

def test_bulk_sources_fixed_point_and_implicit_residual():
    rng = np.random.default_rng(11)
    nvox, ns, dt = 500, 3, 0.05
    S = rng.uniform(0, 5, nvox * ns)
    T = rng.uniform(0, 40, nvox * ns)
    U = rng.uniform(0, 5, nvox * ns)
    u0 = rng.uniform(0, 40, nvox * ns)
    u = oracle.bulk_sources(u0.copy(), ns, dt, S, T, U)
    # backward Euler of du/dt = S (T - u) - U u
    resid = (u - u0) / dt - (S * (T - u) - U * u)
    assert np.max(np.abs(resid)) < 1e-9
    # the equilibrium S T / (S + U) is a fixed point
    star = S * T / (S + U)
    v = oracle.bulk_sources(star.copy(), ns, dt, S, T, U)
    assert np.max(np.abs(v - star) / star) < 1e-14
    # no rates: nothing changes
    z = np.zeros_like(S)
    w = oracle.bulk_sources(u0.copy(), ns, dt, z, T, z)
    assert np.array_equal(w, u0)
