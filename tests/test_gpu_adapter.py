"""The C++ host side (physicell-x_amd/host), driven like the reference's main
loop by tests/cpp/adapter_demo.cpp, against the CPU oracle: FP64 <= 1e-12
relative max-norm, total mass per step <= 1e-12 relative."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from helpers import ROOT, relmax

pytestmark = pytest.mark.gpu

DEMO = os.path.join(ROOT, "tests", "cpp", "adapter_demo")


def write_case(path, nx, ny, nz, ns, steps, dx, dt, D, decay, active, reg, fv, field, agents, precision):
    with open(path, "wb") as f:
        np.array([nx, ny, nz, ns, steps, len(agents["V"]), precision], np.int32).tofile(f)
        parts = [np.array([dx, dt]), D, decay, active.astype(float), reg.astype(float), fv.ravel(), field]
        for p in parts:
            np.asarray(p, np.float64).tofile(f)
        for a in range(len(agents["V"])):
            np.concatenate([agents["pos"][a], [agents["V"][a]], agents["S"][a], agents["U"][a],
                            agents["E"][a], agents["rs"][a]]).astype(np.float64).tofile(f)


@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 1e-5)])
def test_cpp_host_side_matches_oracle(tmp_path, precision, tol):
    assert os.path.exists(DEMO), "build the C++ host side first (make -C physicell-x_amd/host)"
    rng = np.random.default_rng(31)
    nx, ny, nz, ns, steps, dx, dt = 30, 26, 34, 2, 6, 20.0, 0.01
    D = 10 ** rng.uniform(3, 5, ns)
    decay = rng.uniform(0.01, 1.0, ns)
    active = np.array([1, 0], np.uint8)
    reg = np.array([1, 1, 1, 0, 1, 1], np.uint8)
    fv = rng.uniform(0, 40, (6, ns))
    field = (fv[0] + rng.standard_normal((nx, ny, nz, ns))).ravel()
    na = 300
    half = 0.5 * dx * np.array([nx, ny, nz])
    pos = rng.uniform(-half, half, (na, 3))
    pos[5] = pos[6] = pos[7]
    pos[9] = half * 1.5  # outside the domain: deactivated
    agents = dict(pos=pos, V=rng.uniform(1000, 4000, na), S=rng.uniform(0, 10, (na, ns)),
                  U=rng.uniform(0, 10, (na, ns)), E=rng.uniform(0, 20, (na, ns)),
                  rs=rng.uniform(0, 40, (na, ns)))
    case = tmp_path / "case.bin"
    out = tmp_path / "out.bin"
    write_case(case, nx, ny, nz, ns, steps, dx, dt, D, decay, active, reg, fv, field, agents, precision)
    r = subprocess.run([DEMO, str(case), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    res = np.fromfile(out, np.float64)
    nf = field.size
    got, masses = res[:nf], res[nf: nf + steps * ns].reshape(steps, ns)
    grad = res[nf + steps * ns: nf + steps * ns + 3 * nf].reshape(-1, ns, 3)
    bulk = res[nf + steps * ns + 3 * nf:]
    assert bulk.size == nf
    # oracle: same mesh, same voxel mapping, same constants
    bbox = [-half[0], -half[1], -half[2], half[0], half[1], half[2]]
    vox = oracle.voxel_indices(pos, bbox, dx, dx, dx, nx, ny, nz)
    c1, c2, e1, e2 = oracle.agent_constants(dt, dx ** 3, agents["V"], agents["S"], agents["U"], agents["E"],
                                            agents["rs"])
    act = (vox >= 0).astype(np.uint8)
    ag = oracle.AgentSet(np.where(vox >= 0, vox, 0), c1, c2, e1, e2, act)
    orc = oracle.Oracle(nx, ny, nz, ns, dx, D, decay, active, reg, fv)
    u = field.copy()
    for s in range(steps):
        orc.diffusion_step(u, dt)
        orc.secretion_uptake(u, ag)
        m = orc.total_mass(u)
        assert np.max(np.abs(masses[s] - m) / np.abs(m)) <= max(tol, 1e-12), s
    g, uu = got.reshape(-1, ns), u.reshape(-1, ns)
    assert max(relmax(g[:, d], uu[:, d]) for d in range(ns)) <= tol
    # gradients of the final field (compute_all_gradient_vectors), and one
    # bulk source/sink call with the demo's rate callbacks
    gref = oracle.gradients(got, nx, ny, nz, ns, dx)
    if precision == 0:
        assert np.array_equal(grad, gref)
    else:
        # reading the field through density_field() marks the host copy as
        # written, so the FP32 device field is re-quantised on the next upload:
        # the gradients then agree to a few FP32 units of the field over dx
        assert np.max(np.abs(grad - gref)) <= 4 * 2.0 ** -24 * np.max(np.abs(got)) / dx
    n = np.repeat(np.arange(nx * ny * nz), ns)
    d = np.tile(np.arange(ns), nx * ny * nz)
    S = np.where(n % 3 == 0, 0.5 * (d + 1), 0.0)
    T = 30.0 + d
    U = np.where(n % 5 == 0, 0.2, 0.0)
    bref = oracle.bulk_sources(got.copy(), ns, dt, S, T.astype(float), U)
    if precision == 0:
        assert np.array_equal(bulk, bref)
    else:
        assert relmax(bulk, bref) <= 1e-6


def write_extra(path, ns, nodes, singles, solver, track):
    with open(path, "ab") as f:
        np.array([len(nodes), len(singles), solver, track], np.int32).tofile(f)
        for v, vals in nodes:
            np.concatenate([[float(v)], vals]).astype(np.float64).tofile(f)
        for v, d, val in singles:
            np.array([v, d, val], np.float64).tofile(f)


def planes_with_nodes(nx, ny, nz, ns, reg, fv, nodes, singles):
    """Face planes after the faces (registration order) and then the nodes,
    each voxel keeping the last value written to it."""
    planes = oracle.dirichlet_planes(nx, ny, nz, ns, reg, fv)

    def put(v, d, val):
        i, j, k = v // (ny * nz), (v // nz) % ny, v % nz
        if i == 0:
            planes[0][j, k, d] = val
        if i == nx - 1:
            planes[1][j, k, d] = val
        if j == 0:
            planes[2][i, k, d] = val
        if j == ny - 1:
            planes[3][i, k, d] = val
        if k == 0:
            planes[4][i, j, d] = val
        if k == nz - 1:
            planes[5][i, j, d] = val

    for v, vals in nodes:
        for d in range(ns):
            put(v, d, vals[d])
    for v, d, val in singles:
        put(v, d, val)
    return planes


@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 1e-5)])
def test_cpp_host_side_dirichlet_nodes_avx256d_and_tracking(tmp_path, precision, tol):
    """add_dirichlet_node / update_dirichlet_node on faces, edges and inside the
    domain, the AVX256D solver pointer and internalised-substrate tracking,
    through the C++ mirror, against the oracle."""
    assert os.path.exists(DEMO), "build the C++ host side first (make -C physicell-x_amd/host)"
    rng = np.random.default_rng(37)
    nx, ny, nz, ns, steps, dx, dt = 21, 18, 25, 2, 4, 20.0, 0.01
    D = 10 ** rng.uniform(3, 5, ns)
    decay = rng.uniform(0.01, 1.0, ns)
    active = np.array([1, 1], np.uint8)
    reg = np.array([1, 1, 1, 1, 1, 1], np.uint8)
    fv = rng.uniform(0, 40, (6, ns))
    field = (fv[0] + rng.standard_normal((nx, ny, nz, ns))).ravel()

    def vid(i, j, k):
        return (i * ny + j) * nz + k

    nodes = [(vid(0, 5, 7), [70.0, 3.0]), (vid(nx - 1, 0, 4), [11.0, 12.0]),   # face, edge
             (vid(4, 6, nz - 1), [22.0, 33.0]), (vid(5, 5, 5), [99.0, 99.0]),  # face, interior (never read)
             (vid(0, 0, 0), [1.5, 2.5])]                                         # corner
    singles = [(vid(3, ny - 1, 9), 1, 44.0), (vid(0, 5, 7), 0, 71.0)]
    na = 200
    half = 0.5 * dx * np.array([nx, ny, nz])
    pos = rng.uniform(-half, half, (na, 3))
    pos[5] = pos[6] = pos[7]
    agents = dict(pos=pos, V=rng.uniform(1000, 4000, na), S=rng.uniform(0, 10, (na, ns)),
                  U=rng.uniform(0, 10, (na, ns)), E=rng.uniform(0, 20, (na, ns)),
                  rs=rng.uniform(0, 40, (na, ns)))
    case = tmp_path / "case.bin"
    out = tmp_path / "out.bin"
    write_case(case, nx, ny, nz, ns, steps, dx, dt, D, decay, active, reg, fv, field, agents, precision)
    write_extra(case, ns, nodes, singles, solver=1, track=1)
    r = subprocess.run([DEMO, str(case), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    res = np.fromfile(out, np.float64)
    nf = field.size
    got = res[:nf]
    internal_got = res[-na * ns:].reshape(na, ns)
    bbox = [-half[0], -half[1], -half[2], half[0], half[1], half[2]]
    vox = oracle.voxel_indices(pos, bbox, dx, dx, dx, nx, ny, nz)
    c1, c2, e1, e2 = oracle.agent_constants(dt, dx ** 3, agents["V"], agents["S"], agents["U"], agents["E"],
                                            agents["rs"])
    act = (vox >= 0).astype(np.uint8)
    ag = oracle.AgentSet(np.where(vox >= 0, vox, 0), c1, c2, e1, e2, act)
    orc = oracle.Oracle(nx, ny, nz, ns, dx, D, decay, active, reg, fv, solver="avx256d")
    orc.set_planes(planes_with_nodes(nx, ny, nz, ns, reg, fv, nodes, singles))
    u = field.copy()
    internal = np.zeros((na, ns))
    for s in range(steps):
        orc.diffusion_step(u, dt)
        orc.secretion_uptake(u, ag, internal)
    g, uu = got.reshape(-1, ns), u.reshape(-1, ns)
    assert max(relmax(g[:, d], uu[:, d]) for d in range(ns)) <= tol
    assert relmax(internal_got, internal) <= tol
    assert np.max(np.abs(internal)) > 1.0


BINDING = os.path.join(ROOT, "tests", "cpp", "binding_demo")


def test_reference_binding_templates_against_oracle(tmp_path):
    """The reference-tree binding (BioFVM_reference_binding.h) instantiated on
    stand-ins of the reference's public types: the MPI-signature solver at
    world size 1, the container-level agent call with internalised tracking,
    per-voxel Dirichlet nodes read through the face pass, and Dirichlet changes
    between steps picked up by the next call."""
    assert os.path.exists(BINDING), "build the C++ host side first (make -C physicell-x_amd/host)"
    rng = np.random.default_rng(61)
    nx, ny, nz, ns, steps, change, dx, dt = 19, 23, 17, 2, 5, 2, 20.0, 0.01
    D = 10 ** rng.uniform(3, 5, ns)
    decay = rng.uniform(0.01, 1.0, ns)
    active = np.array([1, 1], np.uint8)
    field = rng.uniform(0, 40, nx * ny * nz * ns)

    def vid(i, j, k):
        return (i * ny + j) * nz + k

    # every face voxel gets its own value vector, then a few overrides
    nodes = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                if i in (0, nx - 1) or j in (0, ny - 1) or k in (0, nz - 1):
                    nodes.append((vid(i, j, k), rng.uniform(0, 40, ns)))
    nodes.append((vid(0, 3, 4), np.array([77.0, 66.0])))
    nodes.append((vid(5, 5, 5), np.array([99.0, 99.0])))  # interior: never read
    chg = (vid(nx - 1, 7, 8), 0, 55.0, 1)  # (voxel, substrate, value) then substrate 1 switched off
    na = 150
    voxel = rng.integers(0, nx * ny * nz, na).astype(np.int32)
    voxel[4] = voxel[5] = voxel[6]
    act = np.ones(na, np.uint8)
    act[9] = 0
    c1, c2, e1, e2 = oracle.agent_constants(dt, dx ** 3, rng.uniform(1000, 4000, na), rng.uniform(0, 10, (na, ns)),
                                            rng.uniform(0, 10, (na, ns)), rng.uniform(0, 20, (na, ns)),
                                            rng.uniform(0, 40, (na, ns)))
    case, out = tmp_path / "case.bin", tmp_path / "out.bin"
    with open(case, "wb") as f:
        np.array([nx, ny, nz, ns, steps, na, len(nodes), change], np.int32).tofile(f)
        np.concatenate([[dx, dt], D, decay, active.astype(float), field]).astype(np.float64).tofile(f)
        for v, vals in nodes:
            np.concatenate([[float(v)], vals]).astype(np.float64).tofile(f)
        np.array(chg, np.float64).tofile(f)
        for a in range(na):
            np.concatenate([[voxel[a], act[a]], c1[a], c2[a], e1[a], e2[a]]).astype(np.float64).tofile(f)
    r = subprocess.run([BINDING, str(case), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "one GPU" in r.stdout  # the world-size-2 call printed its error
    res = np.fromfile(out, np.float64)
    got, internal_got = res[:field.size], res[field.size:].reshape(na, ns)
    table = np.ones((nx, ny, nz, ns))
    for v, vals in nodes:
        table[v // (ny * nz), (v // nz) % ny, v % nz] = vals
    orc = oracle.Oracle(nx, ny, nz, ns, dx, D, decay, active, np.ones(6, np.uint8), np.ones((6, ns)))
    ag = oracle.AgentSet(voxel, c1, c2, e1, e2, act)
    u = field.copy()
    internal = np.zeros((na, ns))
    t = table
    for s in range(steps):
        if s == change:
            v, d, val, off = chg
            t[v // (ny * nz), (v // nz) % ny, v % nz, d] = val
            orc.active[off] = 0
        orc.set_planes([t[0], t[-1], t[:, 0], t[:, -1], t[:, :, 0], t[:, :, -1]])
        orc.diffusion_step(u, dt)
        orc.secretion_uptake(u, ag, internal)
    assert relmax(got, u) <= 1e-12
    assert relmax(internal_got, internal) <= 1e-12
