"""GPU parity: the HIP library against the CPU oracle on identical seeded inputs.

Bars (relative max-norm, max|gpu - oracle| / max|oracle|):
  * FP64 path: <= 1e-12 after every step;
  * FP32 scaled-deviation path: <= 1e-5 after every step, per substrate;
  * layout round trips and the Thomas coefficients: exact or <= 4 ulp.
"""
import json
import os

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


def run_compare(pkg, p, steps, precision, tol, per_substrate=False, use_step=False):
    orc = oracle_for(p)
    ag = agents_for(p) if p.voxel.size else None
    u = p.u0.copy()
    env = gpu_env(pkg, p, precision)
    worst = 0.0
    try:
        for s in range(steps):
            orc.diffusion_step(u, p.dt)
            if ag is not None:
                orc.secretion_uptake(u, ag)
            if use_step:
                env.step(p.dt, 1)
            else:
                env.simulate_diffusion_decay(p.dt)
                env.simulate_cell_sources_and_sinks(p.dt)
            g = env.download()
            if per_substrate:
                gg = g.reshape(-1, p.ns)
                uu = u.reshape(-1, p.ns)
                err = max(relmax(gg[:, d], uu[:, d]) for d in range(p.ns))
            else:
                err = relmax(g, u)
            worst = max(worst, err)
            assert err <= tol, f"step {s}: relative max-norm error {err:.3e} > {tol:.0e}"
    finally:
        env.close()
    return worst


@pytest.mark.parametrize("shape", [(17, 23, 29, 1), (32, 32, 32, 2), (33, 7, 65, 3),
                                   (1, 5, 9, 2), (64, 1, 3, 1), (5, 130, 2, 4)])
def test_fp64_diffusion_matches_oracle(pkg, shape):
    p = make_problem(*shape, seed=sum(shape), n_agents=0, dirichlet="all")
    run_compare(pkg, p, 3, "fp64", FP64_TOL)


@pytest.mark.parametrize("shape", [(1024, 3, 2, 1), (3, 1024, 2, 1), (2, 3, 1024, 1), (5, 4, 3, 64)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_maximum_edges_and_substrates(pkg, shape, precision):
    """The largest supported edge on each axis (1024 voxels) and the largest
    substrate count (64), with mixed Dirichlet faces and agents, against the
    oracle (edges of 1025 and 65 substrates are rejected at creation, tested
    on the CPU)."""
    p = make_problem(*shape, seed=sum(shape) + 7, n_agents=12, dirichlet="mixed", collide=2)
    run_compare(pkg, p, 2, precision, FP64_TOL if precision == "fp64" else FP32_TOL)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fill_equals_uploading_the_uniform_field(pkg, precision):
    """bfa_fill gives the uniform field an upload of it gives, and the same
    steps (FP64 bit for bit; FP32 within its bar, the basis being exact for
    a fill); non-finite values are rejected."""
    p = make_problem(37, 21, 18, 3, seed=91, n_agents=40, dirichlet="mixed")
    vals = np.array([3.5, 17.25, 0.0])
    nvox = p.nx * p.ny * p.nz
    a = pkg.Microenvironment(p.nx, p.ny, p.nz, p.ns, p.dx, precision)
    b = pkg.Microenvironment(p.nx, p.ny, p.nz, p.ns, p.dx, precision)
    try:
        for e in (a, b):
            e.set_coefficients(p.D, p.decay)
            e.set_dirichlet(p.active, p.reg, p.fv)
            e.set_agents(p.voxel, p.c1, p.c2, p.e2, p.agent_active)
        a.fill(vals)
        b.upload(np.tile(vals, nvox))
        tol = 0.0 if precision == "fp64" else 1e-7
        assert relmax(a.download(), np.tile(vals, nvox)) <= tol
        a.step(p.dt, 3)
        b.step(p.dt, 3)
        got, want = a.download(), b.download()
        if precision == "fp64":
            assert np.array_equal(got, want)
        else:
            assert relmax(got, want) <= FP32_TOL
        with pytest.raises(pkg.BioFVMError):
            a.fill(np.array([1.0, np.nan, 2.0]))
        # download into a caller's buffer, in both layouts
        for layout in ("ref", "soa"):
            buf = np.full(nvox * p.ns, -1.0)
            assert a.download(layout, out=buf) is buf
            assert np.array_equal(buf, a.download(layout))
        with pytest.raises(ValueError):
            a.download(out=np.empty(nvox * p.ns, np.float32))
        with pytest.raises(ValueError):
            a.download(out=np.empty(nvox * p.ns - 1))
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("dirichlet", ["all", "mixed", "off"])
def test_fp64_dirichlet_variants(pkg, dirichlet):
    p = make_problem(24, 19, 21, 3, seed=5, n_agents=50, dirichlet=dirichlet, export=20.0)
    run_compare(pkg, p, 4, "fp64", FP64_TOL)


def test_fp64_config1_workload(pkg):
    """100^3 voxels, one substrate (D=1e5, decay 0.1), Dirichlet 38 on all
    faces, uniform 38 start, 1000 uptake-only agents (U=10, V=2494), 100 steps."""
    nx = ny = nz = 100
    p = make_problem(nx, ny, nz, 1, seed=1, n_agents=0, dirichlet="all", noise=0.0)
    p.D = np.array([1e5])
    p.decay = np.array([0.1])
    p.fv = np.full((6, 1), 38.0)
    p.u0 = np.full(nx * ny * nz, 38.0)
    rng = np.random.default_rng(1000)
    bbox = [-1000.0, -1000.0, -1000.0, 1000.0, 1000.0, 1000.0]
    pos = rng.uniform(-1000.0, 1000.0, (1000, 3))
    p.voxel = oracle.voxel_indices(pos, bbox, 20.0, 20.0, 20.0, nx, ny, nz)
    n = p.voxel.size
    p.c1, p.c2, p.e1, p.e2 = oracle.agent_constants(0.01, 8000.0, np.full(n, 2494.0),
                                                    np.zeros((n, 1)), np.full((n, 1), 10.0),
                                                    np.zeros((n, 1)), np.zeros((n, 1)))
    p.agent_active = np.ones(n, np.uint8)
    orc = oracle_for(p)
    ag = agents_for(p)
    u = p.u0.copy()
    env = gpu_env(pkg, p, "fp64")
    try:
        for s in range(100):
            orc.diffusion_step(u, p.dt)
            orc.secretion_uptake(u, ag)
            env.step(p.dt, 1)
            if s % 10 == 9 or s < 3:
                g = env.download()
                assert relmax(g, u) <= FP64_TOL, f"step {s}"
                gm = env.total_mass()
                om = orc.total_mass(u)
                assert abs(gm[0] - om[0]) / abs(om[0]) <= 1e-12
    finally:
        env.close()


def test_fp64_agents_sharing_voxels_follow_index_order(pkg):
    p = make_problem(20, 18, 16, 2, seed=77, n_agents=400, dirichlet="mixed", export=30.0,
                     collide=60)
    run_compare(pkg, p, 3, "fp64", FP64_TOL)


def test_fp64_agents_on_boundary_voxels(pkg):
    p = make_problem(12, 10, 9, 2, seed=78, n_agents=0, dirichlet="all")
    nx, ny, nz = p.nx, p.ny, p.nz
    vox = [(0 * ny + 3) * nz + 4, ((nx - 1) * ny + 0) * nz + 0, (5 * ny + 5) * nz + (nz - 1),
           (5 * ny + 5) * nz + 4, (5 * ny + 5) * nz + 4]
    p.voxel = np.array(vox, np.int32)
    rng = np.random.default_rng(3)
    n = len(vox)
    p.c1, p.c2, p.e1, p.e2 = oracle.agent_constants(0.01, 8000.0, np.full(n, 3000.0),
                                                    rng.uniform(0, 10, (n, 2)), rng.uniform(0, 10, (n, 2)),
                                                    rng.uniform(0, 50, (n, 2)), np.full((n, 2), 40.0))
    p.agent_active = np.ones(n, np.uint8)
    run_compare(pkg, p, 3, "fp64", FP64_TOL)


def test_graph_step_equals_eager_calls(pkg):
    p = make_problem(40, 36, 30, 2, seed=9, n_agents=300, dirichlet="mixed", export=5.0)
    a = gpu_env(pkg, p, "fp64")
    b = gpu_env(pkg, p, "fp64")
    try:
        a.step(p.dt, 5)
        for _ in range(5):
            b.simulate_diffusion_decay(p.dt)
            b.simulate_cell_sources_and_sinks()
        assert np.array_equal(a.download(), b.download())
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("shape", [(32, 32, 32, 2), (48, 40, 56, 4), (100, 100, 100, 1)])
def test_fp32_deviation_form_matches_oracle(pkg, shape):
    p = make_problem(*shape, seed=11 + shape[0], n_agents=200, dirichlet="all")
    run_compare(pkg, p, 10, "fp32", FP32_TOL, per_substrate=True, use_step=True)


def test_fp32_mixed_dirichlet_and_export(pkg):
    p = make_problem(30, 26, 22, 3, seed=21, n_agents=300, dirichlet="mixed", export=10.0,
                     collide=40)
    run_compare(pkg, p, 8, "fp32", FP32_TOL, per_substrate=True)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_layout_round_trips(pkg, precision):
    p = make_problem(19, 22, 25, 3, seed=4, dirichlet="off")
    env = pkg.Microenvironment(p.nx, p.ny, p.nz, p.ns, p.dx, precision)
    try:
        env.set_coefficients(p.D, p.decay)
        env.upload(p.u0, "ref")
        soa = env.download("soa")
        ref = env.download("ref")
        exp_soa = p.u0.reshape(p.nx, p.ny, p.nz, p.ns).transpose(3, 2, 1, 0).ravel()
        if precision == "fp64":
            assert np.array_equal(ref, p.u0)
            assert np.array_equal(soa, exp_soa)
        else:
            assert relmax(ref, p.u0) < 1e-7
            assert relmax(soa, exp_soa) < 1e-7
        env.upload(exp_soa, "soa")
        back = env.download("ref")
        assert relmax(back, p.u0) < (1e-7 if precision == "fp32" else 1e-300)
    finally:
        env.close()


def deviation_round_trip(u, b, s, ns, ref):
    """u -> v = float32((u - b) * (1/s)) -> b + s*v: what an FP32 upload and
    download must give bit for bit (s a power of two, so s*v is exact and
    the fused multiply-add rounds once, like the plain sum)."""
    d = np.arange(u.size) % ns if ref else np.arange(u.size) // (u.size // ns)
    v = ((u - b[d]) * (1.0 / s[d])).astype(np.float32)
    return v.astype(np.float64) * s[d] + b[d]


@pytest.mark.parametrize("shape", [(19, 22, 25, 3), (211, 197, 163, 4)])
def test_fp32_transfers_round_deviations_exactly(pkg, shape):
    """FP32 downloads move the deviations over PCIe and expand them on the
    host; with a fixed basis uploads form them on the host too.  Both give
    the device conversion's values bit for bit, in both layouts, for a field
    below the staging threshold and one of four 64 MiB chunks."""
    nx, ny, nz, ns = shape
    rng = np.random.default_rng(78)
    u = rng.uniform(1.0, 50.0, nx * ny * nz * ns)
    soa = u.reshape(nx, ny, nz, ns).transpose(3, 2, 1, 0).ravel()
    env = pkg.Microenvironment(nx, ny, nz, ns, 20.0, "fp32")
    try:
        env.set_coefficients(np.full(ns, 1e3), np.full(ns, 0.1))
        env.upload(u, "ref")  # basis from the field, deviations formed on the device
        b, s = env.fp32_basis()
        dev = env.download("soa")
        assert np.array_equal(env.download("ref"), deviation_round_trip(u, b, s, ns, True))
        assert np.array_equal(dev, deviation_round_trip(soa, b, s, ns, False))
        # the same basis fixed: the host forms the deviations, same device field
        env.set_fp32_basis(b, s)
        env.upload(u, "ref")
        assert np.array_equal(env.download("soa"), dev)
        env.upload(soa, "soa")
        assert np.array_equal(env.download("soa"), dev)
        # another fixed basis
        b2 = b + 0.75
        s2 = s * 4.0
        env.set_fp32_basis(b2, s2)
        env.upload(u, "ref")
        assert np.array_equal(env.download("ref"), deviation_round_trip(u, b2, s2, ns, True))
        env.upload(soa, "soa")
        assert np.array_equal(env.download("soa"), deviation_round_trip(soa, b2, s2, ns, False))
        assert np.array_equal(env.fp32_basis()[0], b2)
    finally:
        env.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_staged_transfers_of_large_fields(pkg, precision):
    """Fields of 16 MiB and more move through the pinned chunk pipeline
    (64 MiB chunks): 211 x 197 x 163 x 4 FP64 values = 207 MiB, three full
    chunks and a partial one, in both layouts, for an environment and for a
    batch member of 160^3 x 4 (two chunks)."""
    nx, ny, nz, ns = 211, 197, 163, 4
    rng = np.random.default_rng(77)
    ref = rng.uniform(1.0, 50.0, nx * ny * nz * ns)
    soa = ref.reshape(nx, ny, nz, ns).transpose(3, 2, 1, 0).ravel()
    tol = 0.0 if precision == "fp64" else 1e-7
    env = pkg.Microenvironment(nx, ny, nz, ns, 20.0, precision)
    try:
        env.set_coefficients(np.full(ns, 1e3), np.full(ns, 0.1))
        env.upload(ref, "ref")
        assert relmax(env.download("soa"), soa) <= tol
        buf = np.empty_like(ref)
        assert relmax(env.download("ref", out=buf), ref) <= tol
        env.upload(soa, "soa")
        assert relmax(env.download("ref"), ref) <= tol
        assert relmax(env.download("soa"), soa) <= tol
        # the same upload and step give the same field again
        env.step(0.01, 1)
        stepped = env.download("ref")
        env.upload(soa, "soa")
        env.step(0.01, 1)
        assert np.array_equal(env.download("ref"), stepped)
    finally:
        env.close()
    n = 160
    mref = rng.uniform(1.0, 50.0, n ** 3 * ns)
    b = pkg.Batch(precision)
    try:
        mid = b.add(n, n, n, ns, 20.0, np.full(ns, 1e3), np.full(ns, 0.1))
        b.upload(mid, mref, "ref")
        assert relmax(b.download(mid, "ref"), mref) <= tol
        msoa = mref.reshape(n, n, n, ns).transpose(3, 2, 1, 0).ravel()
        assert relmax(b.download(mid, "soa"), msoa) <= tol
    finally:
        b.close()


def test_total_mass_and_dirichlet_pass(pkg):
    p = make_problem(21, 17, 13, 2, seed=8, dirichlet="mixed")
    orc = oracle_for(p)
    env = gpu_env(pkg, p, "fp64")
    try:
        u = p.u0.copy()
        assert np.allclose(env.total_mass(), orc.total_mass(u), rtol=1e-13, atol=0)
        orc.apply_dirichlet(u)
        env.apply_dirichlet_boundaries_conditions()
        assert np.array_equal(env.download(), u)
    finally:
        env.close()


def test_golden_fixture_cases_on_gpu(pkg):
    doc = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "lod_exact.json")))
    unhex = lambda x: [unhex(v) for v in x] if isinstance(x, list) else (
        float.fromhex(x) if isinstance(x, str) else x)
    for c in doc["cases"]:
        nx, ny, nz, ns = c["nx"], c["ny"], c["nz"], c["ns"]
        dt, dx = unhex(c["dt"]), unhex(c["dx"])
        env = pkg.Microenvironment(nx, ny, nz, ns, dx, "fp64")
        try:
            env.set_coefficients(unhex(c["D"]), unhex(c["decay"]))
            env.set_dirichlet(c["dirichlet_active"], c["face_registered"], np.array(unhex(c["face_values"])))
            ag = c["agents"]
            if ag["voxel"]:
                c1, c2, e1, e2 = oracle.agent_constants(dt, dx ** 3, unhex(ag["V"]), unhex(ag["S"]),
                                                        unhex(ag["U"]), unhex(ag["E"]),
                                                        unhex(ag["rho_star"]))
                env.set_agents(ag["voxel"], c1, c2, e2, ag["active"])
            for step in c["steps"]:
                env.upload(np.array(unhex(step["input"])))
                env.step(dt, 1)
                got = env.download()
                assert relmax(got, np.array(unhex(step["expected"]))) < 1e-12, c["name"]
        finally:
            env.close()


def test_changing_dt_recomputes_coefficients(pkg):
    p = make_problem(16, 16, 16, 1, seed=2, dirichlet="all")
    orc = oracle_for(p)
    u = p.u0.copy()
    env = gpu_env(pkg, p, "fp64")
    try:
        for dt in (0.01, 0.02, 0.01):
            orc.precompute(dt)
            orc.diffusion_step(u, dt)
            env.simulate_diffusion_decay(dt)
        assert relmax(env.download(), u) <= FP64_TOL
    finally:
        env.close()


TILES = [(4, 4, 4, 1, 64, 64), (8, 16, 32, 8, 32, 16), (32, 8, 16, 32, 16, 64),
         (16, 32, 4, 64, 64, 32), (4, 8, 8, 0, 16, 16), (8, 4, 16, 0, 32, 64),
         (16, 16, 4, 0, 64, 16), (32, 32, 32, 0, 32, 32), (8, 25, 25, 0, 32, 16),
         (16, 25, 8, 0, 64, 32)]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("tiles", TILES)
def test_tile_shapes_give_the_same_answer(pkg, tiles, precision):
    p = make_problem(40, 33, 47, 2, seed=31, n_agents=100, dirichlet="mixed")
    base = gpu_env(pkg, p, precision)
    other = gpu_env(pkg, p, precision)
    try:
        other.set_tiles(*tiles)
        assert other.tiles()[:4] == tiles[:4]
        base.step(p.dt, 2)
        other.step(p.dt, 2)
        assert relmax(other.download(), base.download()) <= (1e-14 if precision == "fp64" else 2e-6)
    finally:
        base.close()
        other.close()


@pytest.mark.parametrize("precision,tol", [("fp64", FP64_TOL), ("fp32", FP32_TOL)])
def test_column_chunks_of_25_nodes_against_oracle(pkg, precision, tol):
    """25-node column chunks: a 400-node y line in 16 whole chunks, a 57-node z
    line in two whole chunks and a partial one, with Dirichlet and agents."""
    p = make_problem(12, 400, 57, 2, seed=57, n_agents=300, dirichlet="mixed", collide=4)
    orc = oracle_for(p)
    ag = agents_for(p)
    env = gpu_env(pkg, p, precision)
    try:
        env.set_tiles(8, 25, 25, 0, 32, 16)
        assert env.tiles()[1:3] == (25, 25)
        u = p.u0.copy()
        for s in range(3):
            orc.diffusion_step(u, p.dt)
            orc.secretion_uptake(u, ag)
            env.step(p.dt, 1)
            assert relmax(env.download(), u) <= tol, s
    finally:
        env.close()


@pytest.mark.parametrize("nx", [1, 3, 63, 64, 65, 130, 257, 1000])
def test_x_sweep_forms_at_odd_row_lengths(pkg, nx):
    """Wave-local and LDS-tiled x sweeps against the oracle for row lengths that
    leave partial chunks, single-lane rows and unaligned rows."""
    p = make_problem(nx, 6, 5, 2, seed=nx, n_agents=0, dirichlet="mixed")
    orc = oracle_for(p)
    u = p.u0.copy()
    orc.diffusion_step(u, p.dt)
    for shape in [(0, 0, 0, 0, 0, 0), (4, 0, 0, 0, 0, 0), (8, 0, 0, 0, 0, 0), (16, 0, 0, 0, 0, 0),
                  (20, 0, 0, 0, 0, 0), (28, 0, 0, 0, 0, 0), (32, 0, 0, 0, 0, 0), (0, 0, 0, 8, 0, 0)]:
        env = gpu_env(pkg, p, "fp64")
        try:
            try:
                env.set_tiles(*shape)
            except pkg.BioFVMError:
                continue  # shape not admissible for this row length
            env.simulate_diffusion_decay(p.dt)
            assert relmax(env.download(), u) <= FP64_TOL, shape
        finally:
            env.close()


@pytest.mark.parametrize("nx", [57, 400, 800])
def test_fp32_wave_x_chunk_lengths_match_oracle(pkg, nx):
    """FP32 wave-local x sweep at every chunk length the row admits (whole
    rows of 3 x 20-node chunks at nx = 400, 2 x 28 at 800, partial chunks at
    57) against the oracle, three steps with Dirichlet on some faces."""
    p = make_problem(nx, 6, 5, 2, seed=nx + 1, n_agents=0, dirichlet="mixed")
    orc = oracle_for(p)
    ref = []
    u = p.u0.copy()
    for _ in range(3):
        orc.diffusion_step(u, p.dt)
        ref.append(u.copy())
    ran = 0
    for L in (4, 8, 16, 20, 28, 32):
        env = gpu_env(pkg, p, "fp32")
        try:
            try:
                env.set_tiles(L, 0, 0, 0, 0, 0)
            except pkg.BioFVMError:
                continue  # more than 64 chunks per row
            ran += 1
            for s in range(3):
                env.simulate_diffusion_decay(p.dt)
                g = env.download().reshape(-1, p.ns)
                r = ref[s].reshape(-1, p.ns)
                err = max(relmax(g[:, d], r[:, d]) for d in range(p.ns))
                assert err <= FP32_TOL, (L, s, err)
        finally:
            env.close()
    assert ran >= 3


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("planes", [1, 3, 7, 40])
@pytest.mark.parametrize("lx", [0, 20, 28])
def test_slab_schedule_is_bit_identical(pkg, precision, planes, lx):
    """x/y sweeps in k-slabs on two streams give exactly the whole-volume result
    (every line is solved by the same arithmetic either way), eager and graph,
    at the default x chunk and at the 20- and 28-node chunks the FP64 search
    pairs with slabs."""
    p = make_problem(36, 29, 23, 3, seed=41, n_agents=200, dirichlet="mixed", export=5.0)
    ref = gpu_env(pkg, p, precision)
    sl = gpu_env(pkg, p, precision)
    eager = gpu_env(pkg, p, precision)
    try:
        if lx:
            for e in (ref, sl, eager):
                e.set_tiles(lx, 0, 0, 0, 0, 0)
        sl.set_slab(planes)
        eager.set_slab(planes)
        assert sl.slab() == min(planes, p.nz)
        ref.step(p.dt, 3)
        sl.step(p.dt, 3)
        for _ in range(3):
            eager.simulate_diffusion_decay(p.dt)
            eager.simulate_cell_sources_and_sinks()
        r = ref.download()
        assert np.array_equal(sl.download(), r)
        assert np.array_equal(eager.download(), r)
    finally:
        ref.close()
        sl.close()
        eager.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_x_grid_is_bit_identical(pkg, precision):
    """Any number of blocks per CU for the wave-local x sweep gives exactly the
    default result (the grid only changes which wave solves which row);
    out-of-range values are rejected and leave the setting unchanged."""
    p = make_problem(57, 33, 29, 3, seed=47, n_agents=120, dirichlet="mixed")
    ref = gpu_env(pkg, p, precision)
    try:
        ref.step(p.dt, 2)
        r = ref.download()
        req, resident = ref.x_grid()
        assert req == 0 and resident >= 1
        for g in (1, 2, 3, resident + 2):
            env = gpu_env(pkg, p, precision)
            try:
                env.set_x_grid(g)
                assert env.x_grid()[0] == g
                for bad in (-1, 65):
                    with pytest.raises(pkg.BioFVMError):
                        env.set_x_grid(bad)
                assert env.x_grid()[0] == g
                env.step(p.dt, 2)
                assert np.array_equal(env.download(), r), g
            finally:
                env.close()
    finally:
        ref.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_col_runs_are_bit_identical(pkg, precision):
    """Any number of tile sub-runs per XCD in the y and z sweeps gives exactly
    the default result (the order only changes which block solves which
    column tile), on launches whose tile count is not a multiple of 8 and at
    more runs than an XCD has tiles; out-of-range values are rejected and
    leave the setting unchanged."""
    p = make_problem(57, 33, 29, 3, seed=48, n_agents=120, dirichlet="mixed")
    ref = gpu_env(pkg, p, precision)
    try:
        ref.set_tiles(8, 8, 8, 0, 16, 32)
        ref.step(p.dt, 2)
        r = ref.download()
        assert ref.col_runs() == (1, 1)
        for ry, rz in ((2, 3), (5, 1), (1, 7), (64, 64)):
            env = gpu_env(pkg, p, precision)
            try:
                env.set_tiles(8, 8, 8, 0, 16, 32)
                env.set_col_runs(ry, rz)
                assert env.col_runs() == (ry, rz)
                for bad in ((0, 1), (1, 0), (65, 1), (1, -2)):
                    with pytest.raises(pkg.BioFVMError):
                        env.set_col_runs(*bad)
                assert env.col_runs() == (ry, rz)
                env.step(p.dt, 2)
                assert np.array_equal(env.download(), r), (ry, rz)
            finally:
                env.close()
    finally:
        ref.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("ns", [2, 3])
def test_substrate_stream_schedule_is_bit_identical(pkg, precision, ns):
    """The two substrate halves swept on parallel streams give exactly the
    single-launch result, eager and graph; a bad group count is rejected."""
    p = make_problem(41, 30, 27, ns, seed=43, n_agents=150, dirichlet="mixed", export=5.0)
    ref = gpu_env(pkg, p, precision)
    sp = gpu_env(pkg, p, precision)
    eager = gpu_env(pkg, p, precision)
    try:
        sp.set_substrate_streams(2)
        eager.set_substrate_streams(2)
        assert sp.substrate_streams() == 2
        with pytest.raises(pkg.BioFVMError):
            sp.set_substrate_streams(3)
        assert sp.substrate_streams() == 2
        ref.step(p.dt, 3)
        sp.step(p.dt, 3)
        for _ in range(3):
            eager.simulate_diffusion_decay(p.dt)
            eager.simulate_cell_sources_and_sinks()
        r = ref.download()
        assert np.array_equal(sp.download(), r)
        assert np.array_equal(eager.download(), r)
    finally:
        ref.close()
        sp.close()
        eager.close()


def test_invalid_tiles_are_rejected_and_state_kept(pkg):
    p = make_problem(20, 20, 20, 1, seed=3, dirichlet="off")
    env = gpu_env(pkg, p, "fp64")
    try:
        before = env.tiles()
        for bad in [(3, 0, 0, 0, 0, 0), (0, 0, 0, 3, 0, 0), (0, 0, 0, 0, 48, 0), (64, 0, 0, 0, 0, 0),
                    (0, 0, 0, 0, 0, 8)]:
            with pytest.raises(pkg.BioFVMError):
                env.set_tiles(*bad)
            assert env.tiles() == before
    finally:
        env.close()


def test_full_size_fp64_eigenmode_and_mass(pkg):
    """400^3 x 4 at its real shape: with Dirichlet off and no agents a separable
    Neumann cosine mode decays by prod 1/mu per step, and with zero decay the
    total mass is conserved."""
    nx = ny = nz = 400
    ns = 4
    D = np.array([1e3, 1e4, 5e4, 1e5])
    decay = np.array([0.0, 0.0, 0.3, 1.0])
    dt, dx = 0.01, 20.0
    m = (1, 2, 3)
    ci = np.cos(np.pi * m[0] * (np.arange(nx) + 0.5) / nx)
    cj = np.cos(np.pi * m[1] * (np.arange(ny) + 0.5) / ny)
    ck = np.cos(np.pi * m[2] * (np.arange(nz) + 0.5) / nz)
    mode = (ck[:, None, None] * cj[None, :, None] * ci[None, None, :]).ravel()  # SoA order per substrate
    amp = np.array([1.0, 2.0, -3.0, 0.5])
    field = np.concatenate([20.0 + amp[d] * mode for d in range(ns)])
    env = pkg.Microenvironment(nx, ny, nz, ns, dx, "fp64")
    try:
        env.set_coefficients(D, decay)
        env.upload(field, "soa")
        m0 = env.total_mass()
        steps = 3
        env.step(dt, steps)
        got = env.download("soa").reshape(ns, -1)
        m1 = env.total_mass()
        for d in range(ns):
            k = D[d] * dt / dx ** 2
            lp = decay[d] * dt / 3
            mu = np.prod([1 + lp + 2 * k * (1 - np.cos(np.pi * mm / n)) for mm, n in zip(m, (nx, ny, nz))])
            const = 20.0 / (1 + lp) ** (3 * steps)
            exp = const + amp[d] * mode / mu ** steps
            assert relmax(got[d], exp) < 1e-12, d
        assert abs(m1[0] - m0[0]) / m0[0] < 1e-12 and abs(m1[1] - m0[1]) / m0[1] < 1e-12
    finally:
        env.close()
