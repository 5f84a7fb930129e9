"""Generate exact-rational golden vectors for the BioFVM LOD step.

The reference ships no golden vectors for this path (its diffusion test,
tests/mpi/diffusion/main.cpp:201-215, only times the solver), and it cannot be
built here (it needs MPI).  These fixtures are therefore computed from the
mathematics the reference implements, in exact rational arithmetic
(``fractions.Fraction``), independently of any floating-point restatement:

* one diffusion step is the LOD product  u <- Dir(Tz^-1 Dir(Ty^-1 Dir(Tx^-1 Dir(u))))
  (BioFVM/BioFVM_solvers.cpp:345-778), where each axis matrix T has diagonal
  1+2k+l' (1+k+l' at both ends, 1+l' for a single node) and off-diagonals -k,
  with k = D*dt/dx^2 and l' = decay*dt/3 (BioFVM/BioFVM_solvers.cpp:194-217);
  each line system is solved here by dense Gaussian elimination, not Thomas;
* Dir overwrites all six faces of every active substrate with the voxel's
  Dirichlet value; a voxel's value is 1.0 unless a registered face contains
  it, in which case the last registered face in the order xmin, xmax, ymin,
  ymax, zmin, zmax wins (BioFVM/BioFVM_microenvironment.cpp:218-226, 348-446,
  632, 2649-2786);
* each active agent, in index order, applies the backward-Euler update
  rho <- (rho + c1)/c2 + e2 with c1 = dt*V/W*S*rho*, c2 = 1 + dt*V/W*(S+U),
  e2 = E*dt/W (documentation/math/substrate-internalization-equations.tex:
  104-121; BioFVM/BioFVM_basic_agent.cpp:224-274, 509-559).

Each step's input is the previous step's exact result rounded to float64, so
a checker is compared one step at a time.  Values are stored as float.hex()
strings.  Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import random
from fractions import Fraction as F

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lod_exact.json")


def axis_matrix(n, k, lp):
    A = [[F(0)] * n for _ in range(n)]
    if n == 1:
        A[0][0] = 1 + lp
        return A
    for i in range(n):
        A[i][i] = 1 + (1 if i in (0, n - 1) else 2) * k + lp
        if i > 0:
            A[i][i - 1] = -k
        if i < n - 1:
            A[i][i + 1] = -k
    return A


def solve_dense(A, b):
    n = len(b)
    M = [row[:] + [b[i]] for i, row in enumerate(A)]
    for c in range(n):
        piv = max(range(c, n), key=lambda r: abs(M[r][c]))
        M[c], M[piv] = M[piv], M[c]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c] / M[c][c]
                M[r] = [x - f * y for x, y in zip(M[r], M[c])]
    return [M[i][n] / M[i][i] for i in range(n)]


def dirichlet_table(nx, ny, nz, ns, registered, values):
    table = {}
    order = [
        (0, lambda: [(0, j, k) for k in range(nz) for j in range(ny)]),
        (1, lambda: [(nx - 1, j, k) for k in range(nz) for j in range(ny)]),
        (2, lambda: [(i, 0, k) for k in range(nz) for i in range(nx)]),
        (3, lambda: [(i, ny - 1, k) for k in range(nz) for i in range(nx)]),
        (4, lambda: [(i, j, 0) for j in range(ny) for i in range(nx)]),
        (5, lambda: [(i, j, nz - 1) for j in range(ny) for i in range(nx)]),
    ]
    for f, voxels in order:
        if registered[f]:
            for v in voxels():
                table[v] = [F(x) for x in values[f]]
    return table


def on_boundary(i, j, k, nx, ny, nz):
    return i in (0, nx - 1) or j in (0, ny - 1) or k in (0, nz - 1)


def apply_dir(u, nx, ny, nz, ns, active, table):
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                if on_boundary(i, j, k, nx, ny, nz):
                    vals = table.get((i, j, k), [F(1)] * ns)
                    for d in range(ns):
                        if active[d]:
                            u[(i, j, k, d)] = vals[d]


def sweep(u, axis, n_axis, dims, ns, mats):
    nx, ny, nz = dims
    ranges = [range(nx), range(ny), range(nz)]
    others = [r for a, r in enumerate(ranges) if a != axis]
    for d in range(ns):
        for p in others[0]:
            for q in others[1]:
                def key(t):
                    idx = [0, 0, 0]
                    rest = [p, q]
                    o = 0
                    for a in range(3):
                        if a == axis:
                            idx[a] = t
                        else:
                            idx[a] = rest[o]
                            o += 1
                    return (idx[0], idx[1], idx[2], d)
                rhs = [u[key(t)] for t in range(n_axis)]
                x = solve_dense(mats[d], rhs)
                for t in range(n_axis):
                    u[key(t)] = x[t]


def exact_step(u, case):
    nx, ny, nz, ns = case["nx"], case["ny"], case["nz"], case["ns"]
    dt, dx = F(case["dt"]), F(case["dx"])
    table = dirichlet_table(nx, ny, nz, ns, case["face_registered"], case["face_values"])
    active = case["dirichlet_active"]
    ks = [F(D) * dt / (dx * dx) for D in case["D"]]
    lps = [F(lam) * dt / 3 for lam in case["decay"]]
    apply_dir(u, nx, ny, nz, ns, active, table)
    for axis, n in enumerate((nx, ny, nz)):
        mats = [axis_matrix(n, ks[d], lps[d]) for d in range(ns)]
        sweep(u, axis, n, (nx, ny, nz), ns, mats)
        apply_dir(u, nx, ny, nz, ns, active, table)
    W = dx * dx * dx
    ag = case["agents"]
    for a in range(len(ag["voxel"])):
        if not ag["active"][a]:
            continue
        v = ag["voxel"][a]
        k = v % nz
        j = (v // nz) % ny
        i = v // (nz * ny)
        f = dt * F(ag["V"][a]) / W
        for d in range(ns):
            S, U, E, rs = (F(ag[x][a][d]) for x in ("S", "U", "E", "rho_star"))
            c1 = f * S * rs
            c2 = 1 + f * (S + U)
            e2 = E * dt / W
            u[(i, j, k, d)] = (u[(i, j, k, d)] + c1) / c2 + e2
    return u


def to_flat(u, nx, ny, nz, ns):
    return [float(u[(i, j, k, d)]) for i in range(nx) for j in range(ny)
            for k in range(nz) for d in range(ns)]


def from_flat(vals, nx, ny, nz, ns):
    u = {}
    it = iter(vals)
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                for d in range(ns):
                    u[(i, j, k, d)] = F(next(it))
    return u


def make_case(name, nx, ny, nz, ns, rng, dirichlet_active, face_registered, agents,
              nsteps, D=None, decay=None):
    D = D or [10 ** rng.uniform(3, 5) for _ in range(ns)]
    decay = decay or [rng.uniform(0.01, 1.0) for _ in range(ns)]
    face_values = [[rng.uniform(0, 40) for _ in range(ns)] for _ in range(6)]
    case = dict(name=name, nx=nx, ny=ny, nz=nz, ns=ns, dx=20.0, dt=0.01, D=D, decay=decay,
                dirichlet_active=dirichlet_active, face_registered=face_registered,
                face_values=face_values, agents=agents)
    u0 = [rng.uniform(0, 40) for _ in range(nx * ny * nz * ns)]
    steps = []
    cur = u0
    for _ in range(nsteps):
        u = from_flat(cur, nx, ny, nz, ns)
        out = to_flat(exact_step(u, case), nx, ny, nz, ns)
        steps.append({"input": cur, "expected": out})
        cur = out
    case["steps"] = steps
    return case


def agents_block(rng, nvox, ns, voxels, active):
    n = len(voxels)
    return dict(
        voxel=voxels,
        active=active,
        V=[rng.uniform(1000, 4000) for _ in range(n)],
        S=[[rng.uniform(0, 10) for _ in range(ns)] for _ in range(n)],
        U=[[rng.uniform(0, 10) for _ in range(ns)] for _ in range(n)],
        E=[[rng.uniform(0, 50) for _ in range(ns)] for _ in range(n)],
        rho_star=[[rng.uniform(0, 40) for _ in range(ns)] for _ in range(n)],
    )


def thomas_exact(n, D, lam, dt, dx):
    k = F(D) * F(dt) / (F(dx) * F(dx))
    lp = F(lam) * F(dt) / 3
    A = axis_matrix(n, k, lp)
    denom, cp = [], []
    for i in range(n):
        d = A[i][i] - (k * k / denom[i - 1] if i > 0 else 0)
        denom.append(d)
        cp.append(-k / d)
    return [float(x) for x in denom], [float(x) for x in cp]


def hexify(obj):
    if isinstance(obj, float):
        return obj.hex()
    if isinstance(obj, list):
        return [hexify(x) for x in obj]
    if isinstance(obj, dict):
        return {k: (v if k in ("voxel", "active", "dirichlet_active", "face_registered")
                    else hexify(v)) for k, v in obj.items()}
    return obj


def main():
    rng = random.Random(20261202)
    cases = []
    # A: 4x5x6, two substrates, Dirichlet on substrate 0 only, ymax never
    # registered (its voxels keep 1.0 unless another face claims them).
    nx, ny, nz, ns = 4, 5, 6, 2
    vox = lambda i, j, k: (i * ny + j) * nz + k
    ag = agents_block(rng, nx * ny * nz, ns,
                      [vox(1, 2, 3), vox(2, 2, 2), vox(1, 2, 3), vox(0, 1, 1),
                       vox(2, 3, 4), vox(1, 2, 3)],
                      [1, 1, 1, 1, 0, 1])
    cases.append(make_case("dirichlet_agents_4x5x6x2", nx, ny, nz, ns, rng, [1, 0],
                           [1, 1, 1, 0, 1, 1], ag, 2))
    # B: 5x4x1 single-node z axis (denominator 1+l'), no Dirichlet, no agents.
    nx, ny, nz, ns = 5, 4, 1, 1
    cases.append(make_case("single_z_node_5x4x1x1", nx, ny, nz, ns, rng, [0], [0] * 6,
                           agents_block(rng, 20, 1, [], []), 2))
    # C: 3x3x7, three substrates, all faces registered, all active, agents on
    # a shared interior voxel in non-monotone order.
    nx, ny, nz, ns = 3, 3, 7, 3
    vox = lambda i, j, k: (i * ny + j) * nz + k
    ag = agents_block(rng, nx * ny * nz, ns, [vox(1, 1, 3), vox(1, 1, 2), vox(1, 1, 3)],
                      [1, 1, 1])
    cases.append(make_case("all_faces_3x3x7x3", nx, ny, nz, ns, rng, [1, 1, 1], [1] * 6, ag, 2))
    thomas = []
    for n in (1, 2, 7, 40):
        D, lam = 10 ** rng.uniform(3, 5), rng.uniform(0.01, 1.0)
        denom, cp = thomas_exact(n, D, lam, 0.01, 20.0)
        thomas.append(dict(n=n, D=D, decay=lam, dt=0.01, dx=20.0, denom=denom, cp=cp))
    doc = {"generator": "tests/golden/make_golden.py",
           "cases": [hexify(c) for c in cases],
           "thomas": [hexify(t) for t in thomas]}
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=0)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
