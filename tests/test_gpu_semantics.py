"""GPU parity of the reference semantics beside the plain step, through the C ABI:

* per-voxel Dirichlet values (dirichlet_value_vectors on the six faces, with
  nodes overridden on faces and edges), the sweeps and the standalone pass;
* internalised-substrate tracking per agent
  (track_internalized_substrates_in_each_agent), from host constants and from
  raw agent state, with agents sharing voxels, up to 2000 in one voxel;
* the LOD_3D_AVX256D variant (no Dirichlet pass after the z sweep).

Bars as in the other parity tests: FP64 1e-12, FP32 1e-5 (relative max-norm).
"""
import numpy as np
import pytest

import oracle
from helpers import agents_for, gpu_env, load_pkg, make_problem, oracle_for, relmax

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-5


@pytest.fixture(scope="module")
def pkg():
    return load_pkg()


def planes_from_table(table):
    """The six face planes of a per-voxel table [nx, ny, nz, ns] in the C ABI
    order: xmin, xmax [j][k][d]; ymin, ymax [i][k][d]; zmin, zmax [i][j][d]."""
    t = table
    return [np.ascontiguousarray(p) for p in
            (t[0], t[-1], t[:, 0], t[:, -1], t[:, :, 0], t[:, :, -1])]


def random_table(p, seed):
    rng = np.random.default_rng(seed)
    table = rng.uniform(0, 40, (p.nx, p.ny, p.nz, p.ns))
    # nodes added after the faces: one inside a face, one on an edge, one on a corner
    table[0, p.ny // 2, p.nz // 3] = 77.0
    table[p.nx - 1, 0, p.nz // 2] = 55.0
    table[0, 0, 0] = 3.5
    return table


@pytest.mark.parametrize("precision,tol", [("fp64", FP64_TOL), ("fp32", FP32_TOL)])
@pytest.mark.parametrize("shape", [(23, 17, 29, 2), (64, 40, 33, 3), (9, 130, 6, 1)])
def test_per_voxel_dirichlet_planes_match_oracle(pkg, precision, tol, shape):
    p = make_problem(*shape, seed=31, n_agents=200, dirichlet="mixed", collide=5)
    p.active[:] = 1
    p.active[-1] = 0 if p.ns > 1 else 1  # one substrate without Dirichlet
    planes = planes_from_table(random_table(p, 32))
    orc = oracle_for(p)
    orc.set_planes(planes)
    ag = agents_for(p)
    env = gpu_env(pkg, p, precision)
    try:
        env.set_dirichlet_planes(planes)
        u = p.u0.copy()
        for s in range(4):
            orc.diffusion_step(u, p.dt)
            orc.secretion_uptake(u, ag)
            env.step(p.dt, 1)
            got = env.download()
            err = relmax(got, u)
            assert err <= tol, f"step {s}: {err:.3e}"
        # the standalone pass writes the planes' values
        env.upload(u)
        v = u.copy()
        v += 1.0
        env.upload(v)
        env.apply_dirichlet_boundaries_conditions()
        orc.apply_dirichlet(v)
        assert relmax(env.download(), v) <= tol
        # back to the face values: the plain face-table result again
        env.set_dirichlet_planes(None)
        orc2 = oracle_for(p)
        u2 = p.u0.copy()
        env.upload(u2)
        orc2.diffusion_step(u2, p.dt)
        env.simulate_diffusion_decay(p.dt)
        assert relmax(env.download(), u2) <= tol
    finally:
        env.close()


def test_inconsistent_planes_are_rejected(pkg):
    p = make_problem(8, 9, 10, 2, seed=33)
    planes = planes_from_table(random_table(p, 34))
    planes[2] = planes[2].copy()
    planes[2][0, 0, 1] += 1.0  # voxel (0, 0, 0) differs between the xmin and ymin planes
    env = gpu_env(pkg, p, "fp64")
    try:
        with pytest.raises(pkg.BioFVMError, match="different values"):
            env.set_dirichlet_planes(planes)
    finally:
        env.close()


@pytest.mark.parametrize("precision,tol", [("fp64", FP64_TOL), ("fp32", FP32_TOL)])
def test_avx256d_variant_matches_oracle(pkg, precision, tol):
    p = make_problem(31, 24, 40, 3, seed=35, n_agents=300, dirichlet="mixed", collide=4)
    orc = oracle.Oracle(p.nx, p.ny, p.nz, p.ns, p.dx, p.D, p.decay, p.active, p.reg, p.fv,
                        solver="avx256d")
    blk = oracle_for(p)
    ag = agents_for(p)
    env = gpu_env(pkg, p, precision)
    try:
        env.set_solver("avx256d")
        assert env.solver() == "avx256d"
        u = p.u0.copy()
        ub = p.u0.copy()
        for s in range(3):
            orc.diffusion_step(u, p.dt)
            orc.secretion_uptake(u, ag)
            blk.diffusion_step(ub, p.dt)
            blk.secretion_uptake(ub, ag)
            env.step(p.dt, 1)
            got = env.download()
            assert relmax(got, u) <= tol, s
        # the variants differ (faces after the z sweep), so the check has teeth
        assert relmax(ub, u) > 1e-6
        env.set_solver("blocking")
        env.upload(p.u0)
        blk2 = oracle_for(p)
        ub = p.u0.copy()
        blk2.diffusion_step(ub, p.dt)
        env.simulate_diffusion_decay(p.dt)
        assert relmax(env.download(), ub) <= tol
    finally:
        env.close()


def export_problem(seed, n_agents=150, shape=(20, 18, 22, 2), decay=True, dirichlet="all"):
    p = make_problem(*shape, seed=seed, n_agents=n_agents, dirichlet=dirichlet, export=15.0,
                     collide=6)
    if not decay:
        p.decay[:] = 0.0
    return p


@pytest.mark.parametrize("precision,tol", [("fp64", FP64_TOL), ("fp32", FP32_TOL)])
def test_internalized_tracking_matches_oracle(pkg, precision, tol):
    p = export_problem(41)
    orc = oracle_for(p)
    ag = agents_for(p)
    rng = np.random.default_rng(42)
    start = rng.uniform(0, 100, (p.voxel.size, p.ns))
    internal = start.copy()
    env = gpu_env(pkg, p, precision)
    try:
        env.track_internalized(start, p.e1)
        u = p.u0.copy()
        for s in range(3):
            orc.diffusion_step(u, p.dt)
            orc.secretion_uptake(u, ag, internal)
            env.step(p.dt, 1)
            got = env.internalized()
            assert got.shape == internal.shape
            assert relmax(got, internal) <= tol, f"step {s}: {relmax(got, internal):.3e}"
            assert relmax(env.download(), u) <= tol
        # the inactive agent kept its start value, and the agents moved substrate
        assert np.array_equal(env.internalized()[3], start[3])
        assert np.max(np.abs(internal - start)) > 1.0
        # a new agent list of the same length keeps the totals ...
        env.set_agents(p.voxel, p.c1, p.c2, p.e2, p.agent_active)
        orc.secretion_uptake(u, ag, internal)
        env.simulate_cell_sources_and_sinks()
        assert relmax(env.internalized(), internal) <= tol
        # ... one of another length switches tracking off
        env.set_agents(p.voxel[:-1], p.c1[:-1], p.c2[:-1], p.e2[:-1])
        with pytest.raises(pkg.BioFVMError, match="tracking is off"):
            env.internalized()
    finally:
        env.close()


def test_internalized_tracking_conserves_substrate_on_gpu(pkg):
    """decay 0, Dirichlet off: field mass + internalised totals stay constant."""
    p = export_problem(43, n_agents=400, decay=False, dirichlet="off")
    env = gpu_env(pkg, p, "fp64")
    try:
        env.track_internalized(None, p.e1)
        total0 = env.total_mass() + env.internalized().sum(0)
        env.step(p.dt, 5)
        total = env.total_mass() + env.internalized().sum(0)
        assert np.all(np.abs(total - total0) <= 1e-12 * np.abs(total0))
        assert np.max(np.abs(env.internalized())) > 1.0
    finally:
        env.close()


def test_internalized_tracking_from_agent_state(pkg):
    """bfa_set_agents_state computes e1 = export_rate * dt itself."""
    nx, ny, nz, ns = 16, 14, 12, 2
    rng = np.random.default_rng(44)
    dx, dt = 20.0, 0.01
    n = 120
    bbox = np.array([0.0, 0.0, 0.0, nx * dx, ny * dx, nz * dx])
    pos = rng.uniform(0, 1, (n, 3)) * bbox[3:]
    pos[7] = pos[8]  # two agents in one voxel
    pos[9] = [-5.0, 10.0, 10.0]  # outside: takes no part
    vol = rng.uniform(1000, 4000, n)
    S, U = rng.uniform(0, 10, (n, ns)), rng.uniform(0, 10, (n, ns))
    E, rs = rng.uniform(0, 20, (n, ns)), np.full((n, ns), 40.0)
    D, decay = np.array([1e4, 5e4]), np.array([0.1, 0.5])
    u0 = rng.uniform(0, 40, nx * ny * nz * ns)
    vox = oracle.voxel_indices(pos, bbox, dx, dx, dx, nx, ny, nz)
    act = (vox >= 0).astype(np.uint8)
    c1, c2, e1, e2 = oracle.agent_constants(dt, dx ** 3, vol, S, U, E, rs)
    orc = oracle.Oracle(nx, ny, nz, ns, dx, D, decay)
    ag = oracle.AgentSet(np.maximum(vox, 0), c1, c2, e1, e2, act)
    internal = np.zeros((n, ns))
    env = pkg.Microenvironment(nx, ny, nz, ns, dx, "fp64")
    try:
        env.set_coefficients(D, decay)
        env.upload(u0)
        env.set_agents_state(pos, bbox, vol, S, U, E, rs, dt)
        env.track_internalized()
        u = u0.copy()
        for _ in range(2):
            orc.diffusion_step(u, dt)
            orc.secretion_uptake(u, ag, internal)
            env.step(dt, 1)
        assert relmax(env.internalized(), internal) <= FP64_TOL
        assert relmax(env.download(), u) <= FP64_TOL
    finally:
        env.close()


@pytest.mark.parametrize("precision,tol", [("fp64", FP64_TOL), ("fp32", FP32_TOL)])
def test_dt_caching_keeps_the_first_coefficients(pkg, precision, tol):
    """With dt caching the coefficients of the first dt serve every later dt,
    as the reference's solver keeps them (diffusion_solver_setup_done); by
    default a new dt recomputes them."""
    p = make_problem(20, 17, 23, 2, seed=71, dirichlet="all")
    env = gpu_env(pkg, p, precision)
    try:
        env.set_dt_caching(True)
        orc = oracle_for(p)
        u = p.u0.copy()
        for dt in (0.01, 0.02, 0.01):
            orc.diffusion_step(u, 0.01)  # the first dt's coefficients every time
            env.simulate_diffusion_decay(dt)
        assert relmax(env.download(), u) <= tol
        env.set_dt_caching(False)
        env.upload(p.u0)
        orc2 = oracle_for(p)
        v = p.u0.copy()
        orc2.diffusion_step(v, 0.02)
        env.simulate_diffusion_decay(0.02)
        assert relmax(env.download(), v) <= tol
    finally:
        env.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_setters_right_after_asynchronous_steps(pkg, precision):
    """bfa_step only enqueues its graph launches.  Calls that drop the captured
    graph (new coefficients, a new dt, new kernel shapes or tile order, new agents) may come
    while those launches still run; the graph is drained before it is
    destroyed, so the result equals a run synchronised before every call.
    The batch's table regrowth (a member added after a step) likewise."""
    p = make_problem(40, 36, 44, 2, seed=91, n_agents=300, dirichlet="mixed", collide=3)

    def run(sync):
        env = gpu_env(pkg, p, precision)
        try:
            calls = [lambda: env.set_coefficients(p.D * 1.5, p.decay),
                     lambda: env.step(p.dt * 2, 1),
                     lambda: env.set_tiles(8, 8, 8, 0, 16, 16),
                     lambda: env.set_col_runs(3, 2),
                     lambda: env.set_agents(p.voxel[::-1].copy(), p.c1[::-1], p.c2[::-1], p.e2[::-1])]
            for call in calls:
                env.step(p.dt, 15)
                if sync:
                    env.synchronize()
                call()
            env.step(p.dt, 5)
            return env.download()
        finally:
            env.close()

    assert np.array_equal(run(False), run(True))

    def run_batch(sync):
        b = pkg.Batch(precision)
        try:
            q = make_problem(30, 28, 26, 2, seed=92, n_agents=100)
            ids = []
            for _ in range(3):
                mid = b.add(q.nx, q.ny, q.nz, q.ns, q.dx, q.D, q.decay, q.active, q.reg, q.fv)
                b.upload(mid, q.u0)
                b.set_agents(mid, q.voxel, q.c1, q.c2, q.e2)
                ids.append(mid)
                b.step(q.dt, 15)
                if sync:
                    b.synchronize()
            b.step(q.dt, 5)
            return [b.download(m) for m in ids]
        finally:
            b.close()

    for a, c in zip(run_batch(False), run_batch(True)):
        assert np.array_equal(a, c)


def crowded_problem(seed):
    """Agents piled onto single voxels: 2000, 300, 65 and 64 agents (the last
    one at the serial walk's limit), interleaved with scattered agents."""
    p = make_problem(20, 18, 22, 2, seed=seed, n_agents=3000, dirichlet="all", export=15.0, collide=6)
    rng = np.random.default_rng(seed)
    hot = rng.choice(p.nx * p.ny * p.nz, 4, replace=False).astype(np.int32)
    idx = rng.permutation(3000)
    start = 0
    for h, cnt in zip(hot, (2000, 300, 65, 64)):
        p.voxel[np.sort(idx[start:start + cnt])] = h
        start += cnt
    p.agent_active[:] = 1
    p.agent_active[idx[5]] = 0  # one inactive agent inside the 2000
    return p


@pytest.mark.parametrize("precision,tol", [("fp64", FP64_TOL), ("fp32", FP32_TOL)])
@pytest.mark.parametrize("lines", [False, True])
def test_crowded_voxels_match_oracle(pkg, precision, tol, lines):
    """Segments of more than 64 agents go through the wavefront segmented
    reduction (k_agents_long); field and internalised totals stay within the
    bars of the oracle's serial walk, with either agent store form."""
    p = crowded_problem(45)
    orc = oracle_for(p)
    ag = agents_for(p)
    rng = np.random.default_rng(46)
    start = rng.uniform(0, 100, (p.voxel.size, p.ns))
    internal = start.copy()
    env = gpu_env(pkg, p, precision)
    try:
        env.set_agent_stores(lines)
        env.track_internalized(start, p.e1)
        u = p.u0.copy()
        for s in range(3):
            orc.diffusion_step(u, p.dt)
            orc.secretion_uptake(u, ag, internal)
            env.step(p.dt, 1)
            assert relmax(env.download(), u) <= tol, f"step {s}"
            assert relmax(env.internalized(), internal) <= tol, f"step {s}"
        # the crowded voxels moved: uptake by 2000 agents in one voxel
        assert np.max(np.abs(internal - start)) > 1.0
    finally:
        env.close()
