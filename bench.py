"""Benchmark: BioFVM LOD diffusion-decay + agent secretion/uptake on one MI355X.

Workload (default): 400^3 voxels x 4 substrates, dx = 20 um, dt = 0.01 min;
D log-uniform in [1e3, 1e5] um^2/min, decay uniform in [0.01, 1] /min,
Dirichlet value uniform in [0, 40] per substrate on all six faces, initial
field = Dirichlet value + N(0,1); 100,000 agents at uniform random voxels with
secretion and uptake rates uniform in [0, 10] /min, saturation density 40,
volume 2494 um^3.  Synthetic data from a fixed seed.

One "step" is one simulate_diffusion_decay(dt) followed by one
simulate_cell_sources_and_sinks pass.  The headline value is the FP32
scaled-deviation form in voxel*substrate updates per second; the FP64 path is
measured beside it, and the FP32 result is compared with the FP64 result
after the timed steps (the FP64 path itself is held to 1e-12 of the CPU
oracle by the test-suite).

The timed region runs K steps with one HIP event pair around every kernel on
the library's stream; the per-kernel mean durations give the roofline figure
of the dominant kernel.  The CPU baseline times the FP64 oracle (C/OpenMP) on
the host on a bounded number of steps of the same workload.

Usage: python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--no-cpu]
Prints one JSON line.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "physicell-x_amd")
PEAK_HBM_GBS = 8000.0
METRIC = "voxel·substrate updates/s at 400³ (FP32 vs FP64); p50/p99 batched-solve latency"


def load_pkg():
    name = "physicellx_amd"
    if name in sys.modules:
        return sys.modules[name]
    path = os.path.join(PKG_DIR, "__init__.py")
    spec = importlib.util.spec_from_file_location(name, path, submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def agent_constants(dt, W, V, S, U, E, rho_star):
    """Basic_Agent::set_internal_uptake_constants (BioFVM/BioFVM_basic_agent.cpp:224-274)
    evaluated in numpy for the synthetic agents."""
    f = (dt * V / W)[:, None]
    c1 = S * rho_star * f
    c2 = 1.0 + f * S + f * U
    e2 = E * dt / W
    return c1, c2, e2


def make_workload(n, ns, n_agents, seed):
    rng = np.random.default_rng(seed)
    D = 10 ** rng.uniform(3, 5, ns)
    decay = rng.uniform(0.01, 1.0, ns)
    dval = rng.uniform(0, 40, ns)
    # device (SoA) order: substrate-major blocks of n^3
    field = np.empty(ns * n ** 3)
    for d in range(ns):
        blk = field[d * n ** 3:(d + 1) * n ** 3]
        blk[:] = rng.standard_normal(n ** 3)
        blk += dval[d]
    voxel = rng.integers(0, n ** 3, n_agents).astype(np.int32)  # reference numbering
    S = rng.uniform(0, 10, (n_agents, ns))
    U = rng.uniform(0, 10, (n_agents, ns))
    E = np.zeros((n_agents, ns))
    rs = np.full((n_agents, ns), 40.0)
    V = np.full(n_agents, 2494.0)
    c1, c2, e2 = agent_constants(0.01, 20.0 ** 3, V, S, U, E, rs)
    return dict(D=D, decay=decay, dval=dval, field=field, voxel=voxel, c1=c1, c2=c2, e2=e2,
                S=S, U=U, E=E, rs=rs, V=V)


def make_env(pkg, w, n, ns, precision, tiles=None, tune_agents=False):
    """The workload's environment; tune_agents picks the agent step's stores
    by measured step time (autotune.autotune_agents) and then restores the
    initial field, else tiles["agent_lines"] (default single voxels) holds."""
    env = pkg.Microenvironment(n, n, n, ns, 20.0, precision)
    if tiles is not None:
        env.set_tiles(tiles["Lx"], tiles["Ly"], tiles["Lz"], tiles["rows_x"], tiles["cols_y"],
                      tiles["cols_z"])
        env.set_x_grid(tiles.get("x_grid", 0))
        env.set_col_runs(tiles.get("runs_y", 1), tiles.get("runs_z", 1))
        env.set_slab(tiles.get("slab", 0))
        env.set_substrate_streams(tiles.get("substrate_streams", 1))
        env.set_agent_stores(bool(tiles.get("agent_lines", 0)))
    env.set_coefficients(w["D"], w["decay"])
    env.set_dirichlet(np.ones(ns, np.uint8), np.ones(6, np.uint8), np.tile(w["dval"], (6, 1)))
    env.upload(w["field"], "soa")
    env.set_agents(w["voxel"], w["c1"], w["c2"], w["e2"])
    if tune_agents:
        import importlib
        importlib.import_module("physicellx_amd.autotune").autotune_agents(env, 0.01)
        env.upload(w["field"], "soa")
    return env


def time_gpu(env, dt, steps, warmup):
    """K timed steps as the library runs them (bfa_step: one captured graph
    per step, replayed), bracketed by device synchronisation; then K more
    steps launched kernel by kernel with HIP events around every kernel, for
    the per-kernel times."""
    env.step(dt, warmup)
    env.synchronize()
    t0 = time.perf_counter()
    env.step(dt, steps)
    env.synchronize()
    wall = time.perf_counter() - t0
    t1 = time.perf_counter()
    ms = env.time_kernels(dt, steps)
    env.synchronize()
    eager_wall = time.perf_counter() - t1
    return wall, ms, eager_wall


def run_host_boundary(env, dt, updates, prec, reps=3, fix_basis=False):
    """The drop-in solver form, where the caller's field stays in host memory
    (FP64, reference layout, pageable): every step uploads it, runs one step
    and downloads it again.  Median over reps of each part, PCIe included.
    FP32 downloads move 4-byte deviations; FP32 uploads do too once the basis
    is fixed (fix_basis: the current basis is fixed first), otherwise they
    move the FP64 field, from which the device derives the basis."""
    host = env.download("ref")
    fp32 = prec == "fp32"
    if fix_basis:
        env.set_fp32_basis(*env.fp32_basis())
    parts = []
    for _ in range(reps + 1):  # the first round pays first-touch costs
        t0 = time.perf_counter()
        env.upload(host, "ref")
        env.synchronize()
        t1 = time.perf_counter()
        env.step(dt, 1)
        env.synchronize()
        t2 = time.perf_counter()
        env.download("ref", out=host)
        parts.append((t1 - t0, t2 - t1, time.perf_counter() - t2))
    p = np.array(parts[1:]) * 1e3
    total = float(np.median(p.sum(axis=1)))
    return {"upload_ms": float(np.median(p[:, 0])), "step_ms": float(np.median(p[:, 1])),
            "download_ms": float(np.median(p[:, 2])), "ms_per_step": total,
            "value": updates / (total * 1e-3), "field_bytes_each_way": int(host.nbytes),
            "pcie_bytes_up": int(host.size * (4 if fp32 and fix_basis else 8)),
            "pcie_bytes_down": int(host.size * (4 if fp32 else 8)), "reps": reps}


def load_traffic(kernel_key, tiles):
    """(HBM bytes per launch, y and z tile sub-runs of the counter run) of a
    sweep kernel, from the counter runs recorded in profiles/traffic.json, or
    (None, None) unless an entry was measured with this exact kernel shape
    (tiles = Lx, Ly, Lz, rows_x, cols_y, cols_z): another shape moves other
    bytes, so its figure is not reported.  The tile order (sub-runs per XCD)
    only changes which block solves which tile, and the counters read the
    same bytes for every order measured (z at 400^3 and 800^3: 1.000x at one
    run and at the tuned count), so an entry of another order is reported,
    with its order."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None, None
    try:
        ent = json.load(open(path)).get("entries", {}).get(kernel_key)
    except Exception:
        return None, None
    if not ent or [int(t) for t in ent.get("tiles", [])] != [int(t) for t in list(tiles)[:6]]:
        return None, None
    return ent.get("bytes"), [int(r) for r in ent.get("runs", [1, 1])]


def parse_tiles(spec):
    """Lx,Ly,Lz,rows_x,cols_y,cols_z[,slab[,streams[,x_grid[,runs_y,runs_z[,agent_lines]]]]]
    -> shape dict."""
    if not spec:
        return None
    v = [int(x) for x in spec.split(",")]
    return dict(Lx=v[0], Ly=v[1], Lz=v[2], rows_x=v[3], cols_y=v[4], cols_z=v[5],
                slab=v[6] if len(v) > 6 else 0, substrate_streams=v[7] if len(v) > 7 else 1,
                x_grid=v[8] if len(v) > 8 else 0, runs_y=v[9] if len(v) > 9 else 1,
                runs_z=v[10] if len(v) > 10 else 1, agent_lines=v[11] if len(v) > 11 else 0)


def host_cpu():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count()}


def cpu_baseline(w, n, ns, budget_s=12.0, max_steps=40, single_thread_steps=1):
    # thread placement as the reference's own timing runs
    # (tests/mpi/diffusion/launch_test.sh:15-16), read when OpenMP starts
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    os.environ.setdefault("OMP_PLACES", "threads")
    sys.path.insert(0, ROOT)
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    oracle.set_num_threads(threads)
    ref_field = np.ascontiguousarray(w["field"].reshape(ns, n, n, n).transpose(3, 2, 1, 0)).ravel()
    orc = oracle.Oracle(n, n, n, ns, 20.0, w["D"], w["decay"], np.ones(ns, np.uint8),
                        np.ones(6, np.uint8), np.tile(w["dval"], (6, 1)))
    W = 20.0 ** 3
    c1, c2, e1, e2 = oracle.agent_constants(0.01, W, w["V"], w["S"], w["U"], w["E"], w["rs"])
    ag = oracle.AgentSet(w["voxel"], c1, c2, e1, e2)
    orc.precompute(0.01)
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps:
        orc.diffusion_step(ref_field, 0.01)
        orc.secretion_uptake(ref_field, ag)
        steps += 1
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    # one thread, for the per-core rate
    oracle.set_num_threads(1)
    t1 = time.perf_counter()
    for _ in range(single_thread_steps):
        orc.diffusion_step(ref_field, 0.01)
        orc.secretion_uptake(ref_field, ag)
    el1 = time.perf_counter() - t1
    oracle.set_num_threads(threads)
    return dict(value=n ** 3 * ns * steps / el, unit="voxel·substrate updates/s", cores=threads,
                kind="port",
                sample=f"FP64 oracle (C/OpenMP, reference loop order) on the {n}^3 x {ns} workload, "
                       f"{steps} steps after the Thomas precompute, {el:.1f} s, {threads} threads",
                single_thread={"value": n ** 3 * ns * single_thread_steps / el1,
                               "steps": single_thread_steps, "seconds": el1},
                host=dict(host_cpu(), omp_proc_bind=os.environ.get("OMP_PROC_BIND"),
                          omp_places=os.environ.get("OMP_PLACES")))


def ensemble_members(n_env, seed):
    """256-member ensemble: edges i.i.d. U{32..160}, 1-4 substrates, D
    log-uniform [1e3, 1e5], decay U[0.01, 1], Dirichlet U[0, 40] on all faces,
    start at the Dirichlet value, agents = 0.1% of voxels at uniform voxels
    with secretion/uptake rates U[0, 10]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_env):
        nx, ny, nz = (int(v) for v in rng.integers(32, 161, 3))
        ns = int(rng.integers(1, 5))
        D = 10 ** rng.uniform(3, 5, ns)
        decay = rng.uniform(0.01, 1.0, ns)
        dval = rng.uniform(0, 40, ns)
        nvox = nx * ny * nz
        na = max(1, nvox // 1000)
        voxel = rng.integers(0, nvox, na).astype(np.int32)
        c1, c2, e2 = agent_constants(0.01, 20.0 ** 3, np.full(na, 2494.0), rng.uniform(0, 10, (na, ns)),
                                     rng.uniform(0, 10, (na, ns)), np.zeros((na, ns)),
                                     np.full((na, ns), 40.0))
        out.append(dict(nx=nx, ny=ny, nz=nz, ns=ns, D=D, decay=decay, dval=dval, voxel=voxel,
                        c1=c1, c2=c2, e2=e2))
    return out


def run_ensemble(pkg, n_env=256, steps=100, repeats=20, seed=4, sequential=True, tune=True, shapes=None):
    """Batched ensemble vs sequential plain solves of the same members.  The
    batch kernel shapes are autotuned on the batch itself, after which every
    member is reset to its initial state; shapes = {precision: (Lx, Ly, Lz,
    cols_y, cols_z, col_threads, streams)} pins them instead (profiling runs)."""
    import importlib
    AT = importlib.import_module("physicellx_amd.autotune")
    mem = ensemble_members(n_env, seed)
    updates = sum(m["nx"] * m["ny"] * m["nz"] * m["ns"] for m in mem) * steps
    res = {"members": n_env, "steps": steps, "updates_per_run": updates}
    fields = {}
    for prec in ("fp32", "fp64"):
        b = pkg.Batch(prec)
        ids = []
        for m in mem:
            mid = b.add(m["nx"], m["ny"], m["nz"], m["ns"], 20.0, m["D"], m["decay"],
                        np.ones(m["ns"], np.uint8), np.ones(6, np.uint8), np.tile(m["dval"], (6, 1)))
            b.fill(mid, m["dval"])
            b.set_agents(mid, m["voxel"], m["c1"], m["c2"], m["e2"])
            ids.append(mid)
        tuned = None
        shape = (shapes or {}).get(prec)
        if shape is not None:
            b.set_col_threads(0)
            b.set_tiles(*shape[:5])
            b.set_col_threads(shape[5])
            b.set_streams(shape[6])
        elif tune:
            tuned = AT.autotune_batch(b, 0.01)
            for mid, m in zip(ids, mem):
                b.fill(mid, m["dval"])
        b.step(0.01, steps)  # the configured run from the initial state (also the warm-up)
        b.synchronize()
        fields[prec] = [b.download(ids[k], "soa") for k in (0, n_env // 2, n_env - 1)]
        lat = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            b.step(0.01, steps)
            b.synchronize()
            lat.append(time.perf_counter() - t0)
        lat = np.array(lat)
        res[prec] = {"value": updates / float(np.median(lat)),
                     "alg_frac_of_peak": updates / float(np.median(lat)) * (24 if prec == "fp32" else 48)
                     / (PEAK_HBM_GBS * 1e9),
                     "p50_s": float(np.percentile(lat, 50)),
                     "p99_s": float(np.percentile(lat, 99)), "runs": repeats, "tiles": list(b.tiles()),
                     "col_threads": list(b.col_threads()), "streams": b.streams(),
                     "autotune_trials": tuned["trials"] if tuned else 0, "pinned": shape is not None}
        b.close()
    errs = [float(np.max(np.abs(a - c)) / np.max(np.abs(c))) for a, c in zip(fields["fp32"], fields["fp64"])]
    res["fp32_vs_fp64_max_rel_err"] = max(errs)
    res["fp32_vs_fp64_after_steps"] = steps
    if sequential:
        envs = []
        for m in mem:
            e = pkg.Microenvironment(m["nx"], m["ny"], m["nz"], m["ns"], 20.0, "fp32")
            e.set_coefficients(m["D"], m["decay"])
            e.set_dirichlet(np.ones(m["ns"], np.uint8), np.ones(6, np.uint8), np.tile(m["dval"], (6, 1)))
            e.upload(np.tile(m["dval"], m["nx"] * m["ny"] * m["nz"]))
            e.set_agents(m["voxel"], m["c1"], m["c2"], m["e2"])
            e.step(0.01, 1)
            e.synchronize()
            envs.append(e)
        t0 = time.perf_counter()
        for e in envs:
            e.step(0.01, steps)
            e.synchronize()
        seq = time.perf_counter() - t0
        for e in envs:
            e.close()
        res["sequential_fp32"] = {"value": updates / seq, "seconds": seq}
        res["speedup_vs_sequential_fp32"] = res["fp32"]["value"] / (updates / seq)
    return res


# serving: jobs of at least this many voxels run as their own Microenvironment
# beside the batch.  Measured on the serving workload (40 requests, edges
# 48-256): all in the batch 1.31e11 updates/s; from 64^3 up solo 1.58e11;
# from 100^3 1.47e11; from 128^3 1.44e11.
SOLO_VOXELS = 64 ** 3


def serving_requests(n_req, seed, rate=200.0, edge=(48, 256), ns_max=8, steps=(10, 1000)):
    """Poisson arrivals; edges U{48..256}, substrates U{1..8}, steps U{10..1000},
    D log-uniform [1e3, 1e5], decay U[0.01, 1], Dirichlet U[0, 40], agents =
    0.05% of voxels with random rates."""
    rng = np.random.default_rng(seed)
    offs = np.cumsum(rng.exponential(1.0 / rate, n_req))
    reqs = []
    for _ in range(n_req):
        nx, ny, nz = (int(v) for v in rng.integers(edge[0], edge[1] + 1, 3))
        ns = int(rng.integers(1, ns_max + 1))
        st = int(rng.integers(steps[0], steps[1] + 1))
        D = 10 ** rng.uniform(3, 5, ns)
        decay = rng.uniform(0.01, 1.0, ns)
        dval = rng.uniform(0, 40, ns)
        nvox = nx * ny * nz
        na = max(1, nvox // 2000)
        voxel = rng.integers(0, nvox, na).astype(np.int32)
        c1, c2, e2 = agent_constants(0.01, 20.0 ** 3, np.full(na, 2494.0), rng.uniform(0, 10, (na, ns)),
                                     rng.uniform(0, 10, (na, ns)), np.zeros((na, ns)),
                                     np.full((na, ns), 40.0))
        reqs.append(dict(nx=nx, ny=ny, nz=nz, ns=ns, steps=st, D=D, decay=decay, dval=dval,
                         voxel=voxel, c1=c1, c2=c2, e2=e2))
    return reqs, offs


def run_serving(pkg, n_req=200, seed=5, rate=200.0, budget=200e9, k=10, spot=2, tiles=None, sequential=True,
                solo_voxels=SOLO_VOXELS):
    import importlib
    S = importlib.import_module("physicellx_amd.serving")
    reqs, offs = serving_requests(n_req, seed, rate)

    def job(r, want):
        return S.Job(nx=r["nx"], ny=r["ny"], nz=r["nz"], ns=r["ns"], steps=r["steps"], D=r["D"],
                     decay=r["decay"], active=np.ones(r["ns"], np.uint8),
                     face_registered=np.ones(6, np.uint8), face_values=np.tile(r["dval"], (6, 1)),
                     init_values=r["dval"], agent_voxel=r["voxel"], agent_c1=r["c1"],
                     agent_c2=r["c2"], agent_e2=r["e2"], want_field=want)

    sizes = [r["nx"] * r["ny"] * r["nz"] * r["ns"] for r in reqs]
    spot_ids = list(np.argsort(sizes)[:spot])
    jobs = [job(r, i in spot_ids) for i, r in enumerate(reqs)]
    srv = S.Server(pkg, "fp32", hbm_budget=budget, steps_per_iteration=k, max_edge=256, tiles=tiles,
                   solo_voxels=solo_voxels)
    wall, done = S.run_arrivals(srv, jobs, offs)
    srv.close()
    updates = sum(j.updates for j in done)
    res = {"requests": n_req, "rate_per_s": rate, "steps_per_iteration": k,
           "offered_updates_per_s": updates / float(offs[-1]) if offs[-1] > 0 else None,
           "value": updates / wall, "unit": "voxel·substrate updates/s", "wall_s": wall,
           "iterations": srv.iterations, "solo_voxels": solo_voxels,
           "solo_jobs": sum(1 for j in done if srv.is_solo(j))}
    res.update(S.latency_stats(done))
    # sweep traffic of the served FP32 updates (24 B each) against the 8 TB/s peak
    res["alg_frac_of_peak"] = res["value"] * 24 / (PEAK_HBM_GBS * 1e9)
    # FP64 spot check of the smallest requests, run alone
    errs = []
    for i in spot_ids:
        r, j = reqs[i], jobs[i]
        with pkg.Batch("fp64") as b:
            mid = b.add(r["nx"], r["ny"], r["nz"], r["ns"], 20.0, r["D"], r["decay"],
                        np.ones(r["ns"], np.uint8), np.ones(6, np.uint8), np.tile(r["dval"], (6, 1)))
            b.fill(mid, r["dval"])
            b.set_agents(mid, r["voxel"], r["c1"], r["c2"], r["e2"])
            b.step(0.01, r["steps"])
            ref = b.download(mid)
        errs.append(float(np.max(np.abs(j.result - ref)) / np.max(np.abs(ref))))
    res["fp64_spot_check_max_rel_err"] = max(errs) if errs else None
    if not sequential:
        return res
    # the same requests solved one after another with the plain path (step time only)
    sample = reqs
    tot_u, tot_t = 0, 0.0
    for r in sample:
        e = pkg.Microenvironment(r["nx"], r["ny"], r["nz"], r["ns"], 20.0, "fp32")
        e.set_coefficients(r["D"], r["decay"])
        e.set_dirichlet(np.ones(r["ns"], np.uint8), np.ones(6, np.uint8), np.tile(r["dval"], (6, 1)))
        e.upload(np.tile(r["dval"], r["nx"] * r["ny"] * r["nz"]))
        e.set_agents(r["voxel"], r["c1"], r["c2"], r["e2"])
        e.step(0.01, 1)
        e.synchronize()
        t0 = time.perf_counter()
        e.step(0.01, r["steps"])
        e.synchronize()
        tot_t += time.perf_counter() - t0
        tot_u += r["nx"] * r["ny"] * r["nz"] * r["ns"] * r["steps"]
        e.close()
    res["sequential_plain_sample"] = {"value": tot_u / tot_t, "requests": len(sample)}
    res["speedup_vs_sequential"] = res["value"] / (tot_u / tot_t)
    return res


def run_aux(env, w, updates, reps=10):
    """The kernels beside the step, on the same FP32 400^3 x 4 field: the
    whole-field gradient (device kernel time; reads the field, writes 3 FP64
    values per voxel-substrate), the gradient at the agents' voxels through
    the C ABI (host round trip included), and one bulk source/sink update
    (reads the field and two FP64 rate fields, writes the field)."""
    out = {}
    gms = env.time_gradient(reps)
    gbytes = updates * (4 + 3 * 8)
    out["gradient"] = {"kernel_ms": gms, "alg_bytes": gbytes, "GBps": gbytes / (gms * 1e-3) / 1e9}
    vox = np.ascontiguousarray(w["voxel"], np.int32)
    env.gradients_at(vox)
    t0 = time.perf_counter()
    for _ in range(reps):
        env.gradients_at(vox)
    out["gradient_at_agents"] = {"voxels": int(vox.size), "ms_per_call": (time.perf_counter() - t0) / reps * 1e3}
    env.set_bulk_rates(np.full(updates, 0.1), np.full(updates, 30.0), np.full(updates, 0.05))
    env.simulate_bulk_sources_and_sinks(0.01)
    env.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        env.simulate_bulk_sources_and_sinks(0.01)
    env.synchronize()
    bms = (time.perf_counter() - t0) / reps * 1e3
    bbytes = updates * (4 + 16 + 4)
    out["bulk_sources"] = {"ms_per_call": bms, "alg_bytes": bbytes, "GBps": bbytes / (bms * 1e-3) / 1e9}
    env.set_bulk_rates(None, None, None)
    return out


def large_workload(n=800, seed=3, na=1_000_000):
    """Config-3 workload: 800^3 x 2, an oxygen-like substrate (D=1e5, decay
    0.1, Dirichlet 38) and a signal (D=1e3, decay 0.4, Dirichlet 0) on all six
    faces, the field starting at the Dirichlet values; 1e6 agents in a
    radius-6 mm spheroid at the centre, density weighted toward the rim,
    uptake 10/min on substrate 0 and secretion 1/min on substrate 1
    (saturation density 40 on substrate 1), volume 2494 um^3."""
    rng = np.random.default_rng(seed)
    ns = 2
    D = np.array([1e5, 1e3])
    decay = np.array([0.1, 0.4])
    dval = np.array([38.0, 0.0])
    # radius r in [0, R] with density ~ exp(-((R - r)/w)^2) r^2 (rim-weighted)
    R = 6000.0
    r = R * (1 - np.abs(rng.normal(0, 0.15, na)))
    r = np.clip(r, 0, R)
    v = rng.normal(size=(na, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    pos = v * r[:, None]
    L = n * 20.0
    idx = np.clip(np.floor((pos + L / 2) / 20.0).astype(np.int64), 0, n - 1)
    voxel = ((idx[:, 0] * n + idx[:, 1]) * n + idx[:, 2]).astype(np.int32)
    S = np.zeros((na, ns))
    U = np.zeros((na, ns))
    U[:, 0] = 10.0
    S[:, 1] = 1.0
    rs = np.zeros((na, ns))
    rs[:, 1] = 40.0
    V = np.full(na, 2494.0)
    E = np.zeros((na, ns))
    c1, c2, e2 = agent_constants(0.01, 20.0 ** 3, V, S, U, E, rs)
    return dict(n=n, ns=ns, D=D, decay=decay, dval=dval, voxel=voxel, c1=c1, c2=c2, e2=e2,
                S=S, U=U, E=E, rs=rs, V=V)


def run_large(pkg, n=800, steps=20, seed=3, tune=True, fixed=None):
    """800^3 x 2 FP32 scaled-deviation form and FP64 path on the config-3
    workload (large_workload): autotuned shapes, graph-replayed step time,
    per-kernel times with the dominant kernel's roofline, and FP32 vs FP64
    max relative error and mass drift per substrate after the run.  fixed =
    {"fp32": shape, "fp64": shape} (parse_tiles) pins a precision's shapes."""
    import importlib
    at = importlib.import_module("physicellx_amd.autotune")
    w = large_workload(n, seed)
    ns, D, decay, dval = w["ns"], w["D"], w["decay"], w["dval"]
    voxel, c1, c2, e2 = w["voxel"], w["c1"], w["c2"], w["e2"]
    updates = n ** 3 * ns
    out = {"workload": f"{n}^3 voxels x {ns} substrates, {voxel.size} spheroid agents, dt=0.01 min",
           "grid": [n, n, n], "substrates": ns, "agents": int(voxel.size), "steps": steps}
    envs = {}
    for prec in ("fp32", "fp64"):
        tiles = (fixed or {}).get(prec)
        if tiles is None and tune:
            tiles = at.autotune(pkg, n, n, n, ns, prec, 0.01)
        e = pkg.Microenvironment(n, n, n, ns, 20.0, prec)
        if tiles is not None:
            e.set_tiles(tiles["Lx"], tiles["Ly"], tiles["Lz"], tiles["rows_x"], tiles["cols_y"], tiles["cols_z"])
            e.set_x_grid(tiles.get("x_grid", 0))
            e.set_col_runs(tiles.get("runs_y", 1), tiles.get("runs_z", 1))
            e.set_slab(tiles.get("slab", 0))
            e.set_substrate_streams(tiles.get("substrate_streams", 1))
            e.set_agent_stores(bool(tiles.get("agent_lines", 0)))
        e.set_coefficients(D, decay)
        e.set_dirichlet(np.ones(ns, np.uint8), np.ones(6, np.uint8), np.tile(dval, (6, 1)))
        e.fill(dval)
        e.set_agents(voxel, c1, c2, e2)
        if tune and (fixed or {}).get(prec) is None:
            at.autotune_agents(e, 0.01)  # then back to the initial field
            e.fill(dval)
        e.step(0.01, 1)
        e.synchronize()
        t0 = time.perf_counter()
        e.step(0.01, steps - 1)
        e.synchronize()
        dt_ = time.perf_counter() - t0
        ms = e.time_kernels(0.01, 5)
        tsize = 4 if prec == "fp32" else 8
        # with the k-slab schedule the x slot times the overlapped x/y slab pair
        names = ["xy_slab" if e.slab() else "x", "y", "z"]
        dom = int(np.argmax(ms[:3]))
        sweeps = 2 if (e.slab() and dom == 0) else 1
        alg = updates * 2 * tsize * sweeps
        achieved = alg / (ms[dom] * 1e-3) / 1e9
        traf = load_traffic(f"sweep_{names[dom]}_{prec}_{n}", e.tiles())
        out[prec] = {"value": updates * (steps - 1) / dt_, "ms_per_step": 1e3 * dt_ / (steps - 1),
                     "step_alg_frac_of_peak": updates * 6 * tsize / (dt_ / (steps - 1)) / 1e9 / PEAK_HBM_GBS,
                     "kernel_ms": dict(x=ms[0], y=ms[1], z=ms[2], agents=ms[3]),
                     "roofline": {"bound": "hbm", "kernel": f"sweep_{names[dom]}", "sweeps_in_kernel": sweeps,
                                  "alg_bytes_per_launch": alg,
                                  "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                  "frac": achieved / PEAK_HBM_GBS,
                                  "traffic": traf[0], "traffic_col_runs": traf[1]},
                     "tiles": list(e.tiles()), "autotuned": tune and (fixed or {}).get(prec) is None,
                     "schedule": {"slab": e.slab(), "substrate_streams": e.substrate_streams(),
                                  "x_grid": e.x_grid()[0], "runs_y": e.col_runs()[0], "runs_z": e.col_runs()[1]},
                     "agent_stores": "whole lines" if e.agent_stores() else "single voxels"}
        envs[prec] = e
    f32 = envs["fp32"].download("soa").reshape(ns, -1)
    f64 = envs["fp64"].download("soa").reshape(ns, -1)
    m32, m64 = envs["fp32"].total_mass(), envs["fp64"].total_mass()
    out["fp32_vs_fp64"] = {
        "max_rel_err_per_substrate": [float(np.max(np.abs(f32[d] - f64[d])) / np.max(np.abs(f64[d])))
                                      for d in range(ns)],
        "mass_drift_per_substrate": [float(abs(m32[d] - m64[d]) / abs(m64[d])) for d in range(ns)],
        "steps": steps + 5, "norm": "max|f32 - f64| / max|f64| per substrate"}
    for e in envs.values():
        e.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=400)
    ap.add_argument("--ns", type=int, default=4)
    ap.add_argument("--agents", type=int, default=100_000)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--mode", default="default", choices=["default", "single", "ensemble", "serve", "large"])
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--no-seq", action="store_true", help="ensemble mode: skip the sequential plain solves")
    ap.add_argument("--tiles", default="", help="fixed FP32 shapes "
                    "Lx,Ly,Lz,rows_x,cols_y,cols_z[,slab[,streams[,x_grid[,runs_y,runs_z]]]] "
                    "instead of autotuning (profiling runs)")
    ap.add_argument("--tiles64", default="", help="fixed FP64 shapes, same fields as --tiles")
    ap.add_argument("--repeats", type=int, default=20, help="ensemble mode: timed batched solves")
    ap.add_argument("--batch-shape", default="", help="ensemble mode: fixed batch shapes "
                    "Lx,Ly,Lz,cols_y,cols_z,col_threads,streams instead of autotuning (FP32)")
    ap.add_argument("--batch-shape64", default="", help="ensemble mode: fixed FP64 batch shapes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the domain is resident on a single MI355X)", file=sys.stderr)
        sys.exit(2)
    pkg = load_pkg()
    if args.mode == "ensemble":
        shapes = {p: tuple(int(x) for x in v.split(",")) for p, v in
                  (("fp32", args.batch_shape), ("fp64", args.batch_shape64)) if v}
        print(json.dumps({"ensemble": run_ensemble(pkg, sequential=not args.no_seq, repeats=args.repeats,
                                                   shapes=shapes)}))
        return
    if args.mode == "serve":
        print(json.dumps({"serving": run_serving(pkg, n_req=args.requests)}))
        return
    if args.mode == "large":
        fixed = {"fp32": parse_tiles(args.tiles), "fp64": parse_tiles(args.tiles64)}
        print(json.dumps({"large": run_large(pkg, steps=args.steps, fixed=fixed)}))
        return
    n, ns = args.n, args.ns
    dt = 0.01
    w = make_workload(n, ns, args.agents, args.seed)
    updates = n ** 3 * ns

    tiles = {"fp32": parse_tiles(args.tiles), "fp64": parse_tiles(args.tiles64)}
    if not args.no_tune:
        import importlib
        at = importlib.import_module("physicellx_amd.autotune")
        for prec in ("fp32",) if args.no_fp64 else ("fp32", "fp64"):
            if tiles[prec] is not None:
                continue
            r = at.autotune(pkg, n, n, n, ns, prec, dt)
            tiles[prec] = r
            if args.verbose:
                for t in r["trials"]:
                    print(prec, json.dumps(t), file=sys.stderr)
    copy_gbs = 2 * 2 ** 30 / (pkg.copy_kernel_ms(2 ** 30, 10) * 1e-3) / 1e9
    copy_stream_gbs = 2 * 2 ** 30 / (pkg.copy_stream_ms(2 ** 30, 10) * 1e-3) / 1e9

    tune_ag = not args.no_tune
    env32 = make_env(pkg, w, n, ns, "fp32", tiles["fp32"], tune_agents=tune_ag and not args.tiles)
    wall32, ms32, gwall32 = time_gpu(env32, dt, args.steps, args.warmup)
    tiles32 = env32.tiles()
    runs32 = env32.col_runs()
    lines32 = env32.agent_stores()
    split32 = env32.substrate_streams()
    res = {}
    if not args.no_fp64:
        env64 = make_env(pkg, w, n, ns, "fp64", tiles["fp64"], tune_agents=tune_ag and not args.tiles64)
        wall64, ms64, gwall64 = time_gpu(env64, dt, args.steps, args.warmup)
        # both have now run warmup + 2*steps steps from the same start
        f32 = env32.download("soa").reshape(ns, -1)
        f64 = env64.download("soa").reshape(ns, -1)
        m32, m64 = env32.total_mass(), env64.total_mass()
        errs = [float(np.max(np.abs(f32[d] - f64[d])) / np.max(np.abs(f64[d]))) for d in range(ns)]
        drift = [float(abs(m32[d] - m64[d]) / abs(m64[d])) for d in range(ns)]
        res["fp64"] = dict(value=updates * args.steps / wall64, ms_per_step=1e3 * wall64 / args.steps,
                           eager_ms_per_step=1e3 * gwall64 / args.steps,
                           kernel_ms=dict(x=ms64[0], y=ms64[1], z=ms64[2], agents=ms64[3]),
                           tiles=list(env64.tiles()), col_runs=list(env64.col_runs()),
                           substrate_streams=env64.substrate_streams(), slab=env64.slab(),
                           agent_stores="whole lines" if env64.agent_stores() else "single voxels")
        res["fp32_vs_fp64"] = dict(max_rel_err_per_substrate=errs, mass_drift_per_substrate=drift,
                                   steps=args.warmup + 2 * args.steps,
                                   norm="max|f32 - f64| / max|f64| per substrate")
        host_boundary = {"fp64": run_host_boundary(env64, dt, updates, "fp64")}
        env64.close()
    else:
        host_boundary = {}
    aux = run_aux(env32, w, updates)
    host_boundary["fp32"] = run_host_boundary(env32, dt, updates, "fp32")
    host_boundary["fp32_fixed_basis"] = run_host_boundary(env32, dt, updates, "fp32", fix_basis=True)
    env32.close()

    slab32 = tiles["fp32"]["slab"] if tiles["fp32"] else 0
    names = ["xy_slab" if slab32 else "x", "y", "z"]
    dom = int(np.argmax(ms32[:3]))
    sweeps = 2 if (slab32 and dom == 0) else 1
    alg_bytes = updates * 2 * 4 * sweeps  # one FP32 read + one write per voxel-substrate per sweep
    achieved = alg_bytes / (ms32[dom] * 1e-3) / 1e9
    traffic, traffic_runs = load_traffic(f"sweep_{names[dom]}_fp32_{n}", tiles32)
    out = {
        "metric": METRIC,
        "value": updates * args.steps / wall32,
        "unit": "voxel·substrate updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall32 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{n}^3 voxels x {ns} substrates, {args.agents} agents, dt=0.01 min, "
                               "FP32 scaled-deviation LOD step + agents",
                   "grid": [n, n, n], "substrates": ns, "agents": args.agents, "seed": args.seed},
        "roofline": {"bound": "hbm", "kernel": f"sweep_{names[dom]}", "sweeps_in_kernel": sweeps,
                     "alg_bytes_per_launch": alg_bytes, "achieved": achieved,
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                     "traffic": traffic, "traffic_col_runs": traffic_runs},
        "kernel_ms": dict(x=ms32[0], y=ms32[1], z=ms32[2], agents=ms32[3]),
        "step_alg_frac_of_peak": (updates * 6 * 4) / (wall32 / args.steps) / 1e9 / PEAK_HBM_GBS,
        "copy_kernel_GBps": copy_gbs,
        "copy_stream_GBps": copy_stream_gbs,
        "eager_ms_per_step": 1e3 * gwall32 / args.steps,
        "tiles": dict(Lx=tiles32[0], Ly=tiles32[1], Lz=tiles32[2], rows_x=tiles32[3],
                      cols_y=tiles32[4], cols_z=tiles32[5], slab=slab32,
                      substrate_streams=split32,
                      x_grid=(tiles["fp32"] or {}).get("x_grid", 0),
                      runs_y=runs32[0], runs_z=runs32[1],
                      agent_stores="whole lines" if lines32 else "single voxels",
                      autotuned=not args.no_tune and not args.tiles),
        "aux_kernels": aux,
        "host_boundary": dict(host_boundary, note="field in host memory (FP64, reference layout): "
                              "upload + one step + download per step, PCIe included; not the value"),
    }
    out.update(res)
    if args.mode == "default":
        out["large"] = run_large(pkg)
        ens = run_ensemble(pkg, repeats=20)
        srv = run_serving(pkg, n_req=args.requests)
        out["batched_solve"] = {"p50_s": ens["fp32"]["p50_s"], "p99_s": ens["fp32"]["p99_s"],
                                "workload": "256-member ensemble, 100 steps per solve, FP32"}
        out["ensemble"] = ens
        out["serving"] = srv
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, n, ns)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
