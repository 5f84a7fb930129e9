"""Measured kernel-shape selection for the LOD sweeps.

For one grid shape and precision, times every admissible shape of the three
sweep kernels on the GPU (HIP events around each launch, via
``Microenvironment.time_kernels``) and returns the fastest per axis:

* x sweep: the wave-local form with chunk length L in {4, 8, 16, 32}, and
  in FP64 also 20 and 28 (a row's ceil(nx/L) <= 64 chunks share one
  wavefront), or the LDS-tiled
  form with R rows per tile and chunk length L; for the wave-local form, the
  number of resident blocks per compute unit (fewer waves in flight can
  stream faster);
* y and z sweeps: TX in {16, 32, 64} x-columns per block and chunk length L
  in {4, 8, 16, 25, 32}, with TX * ceil(n/L) <= 1024 threads, chosen per axis
  (25-node chunks cut a 400-node line into 16 chunks: 32 columns then fill
  eight whole wavefronts with no padded nodes);
* y and z sweeps at the chosen shapes: how many sub-runs of tiles each XCD
  advances together (``Microenvironment.set_col_runs``), per axis;
* schedule: whole-volume sweeps, the two substrate halves swept on parallel
  streams, or (FP64) x/y sweeps in k-slabs of a few planes, the slabs with
  each of the three fastest wave x chunk lengths (chosen by graph-replayed
  whole-step time; slabs have to win by 2 %).

Results are cached per (nx, ny, nz, ns, precision) in memory and, when a path
is given, in a JSON file.
"""
from __future__ import annotations

import json
import os

import numpy as np

# Wave-local x chunk lengths per precision.  FP64 takes 20 and 28 too: at
# 400^3 x 4 with the k-slab schedule its step went from 2.39 to 2.20 ms with
# 20-node chunks.  FP32 does not: its x sweep timed alone is no faster with
# them (0.393 against 0.394 ms), and in the step with 100k agents it is
# slower (1.27-1.33 against 1.22 ms), which a search without agents misses.
X_WAVE_L = {"fp32": (4, 8, 16, 32), "fp64": (4, 8, 16, 20, 28, 32)}
X_LDS = [(R, L) for R in (8, 16, 32) for L in (8, 16, 32)]
COL_TX = (16, 32, 64)
COL_L = (4, 8, 16, 25, 32)
SLAB_BYTES = (8 << 20, 16 << 20, 32 << 20, 64 << 20, 128 << 20)
X_GRID = (2, 3, 4, 5, 6, 8)
COL_RUNS = (1, 2, 3, 4, 5, 6, 8, 12)
SLAB_X_SHAPES = 3
SLAB_MARGIN = 0.98

_cache: dict = {}


def _key(nx, ny, nz, ns, precision):
    return f"{nx}x{ny}x{nz}x{ns}:{precision}"


def _time(env, dt, reps, shape):
    try:
        env.set_tiles(*shape)
    except Exception:
        return None
    env.time_kernels(dt, 1)  # first launch of a shape pays code-object load
    return env.time_kernels(dt, reps)


def _step_ms(env, dt, reps):
    """Mean graph-replayed step time (ms) over 2*reps + 2 steps."""
    import time

    env.step(dt, 2)  # capture the schedule's graph
    env.synchronize()
    n = 2 * reps + 2
    t0 = time.perf_counter()
    env.step(dt, n)
    env.synchronize()
    return (time.perf_counter() - t0) / n * 1e3


def autotune(pkg, nx, ny, nz, ns, precision="fp32", dt=0.01, reps=6, cache_path=None,
             verbose=False):
    """Return {"Lx", "rows_x", "Ly", "cols_y", "Lz", "cols_z", "x_grid", "runs_y",
    "runs_z", "slab", "substrate_streams", "ms", "trials"}."""
    key = _key(nx, ny, nz, ns, precision)
    if key in _cache:
        return _cache[key]
    if cache_path and os.path.exists(cache_path):
        try:
            doc = json.load(open(cache_path))
            if key in doc:
                _cache[key] = doc[key]
                return doc[key]
        except Exception:
            pass
    rng = np.random.default_rng(0)
    env = pkg.Microenvironment(nx, ny, nz, ns, 20.0, precision)
    trials = []
    try:
        env.set_coefficients(10 ** rng.uniform(3, 5, ns), rng.uniform(0.01, 1, ns))
        env.set_dirichlet(np.ones(ns, np.uint8), np.ones(6, np.uint8), np.full((6, ns), 20.0))
        if nx * ny * nz * ns <= 1 << 27:
            env.upload(20.0 + rng.standard_normal(nx * ny * nz * ns), "soa")
        else:  # kernel times do not depend on the values; skip a host field of this size
            env.fill(np.full(ns, 21.0))
        # wave-local x sweep: chunk length and resident blocks per CU together
        # (the best block count differs between chunk lengths)
        best_x = None
        wave_best = {}  # chunk length -> (ms, blocks per CU) of its best grid
        for L in X_WAVE_L["fp32" if precision in ("fp32", "f32", "float", 1) else "fp64"]:
            env.set_x_grid(0)
            ms = _time(env, dt, reps, (L, 0, 0, 0, 0, 0))
            if ms is None:
                continue
            trials.append({"axis": "x", "form": "wave", "L": L, "blocks_per_cu": 0, "ms": ms[0]})
            wave_best[L] = (ms[0], 0)
            if best_x is None or ms[0] < best_x[0]:
                best_x = (ms[0], L, 0, 0)
            resident = env.x_grid()[1]
            for g in X_GRID:
                if g >= resident:
                    break
                env.set_x_grid(g)
                env.time_kernels(dt, 1)
                ms = env.time_kernels(dt, reps)
                trials.append({"axis": "x", "form": "wave", "L": L, "blocks_per_cu": g, "ms": ms[0]})
                if ms[0] < wave_best[L][0]:
                    wave_best[L] = (ms[0], g)
                if ms[0] < best_x[0]:
                    best_x = (ms[0], L, 0, g)
        env.set_x_grid(0)
        for R, L in X_LDS:
            ms = _time(env, dt, reps, (L, 0, 0, R, 0, 0))
            if ms is not None:
                trials.append({"axis": "x", "form": "lds", "R": R, "L": L, "ms": ms[0]})
                if best_x is None or ms[0] < best_x[0]:
                    best_x = (ms[0], L, R, 0)
        best_y = best_z = None
        env.set_x_grid(best_x[3])
        for TX in COL_TX:
            for L in COL_L:
                ms = _time(env, dt, reps, (best_x[1], L, L, best_x[2], TX, TX))
                if ms is None:
                    continue
                trials.append({"axis": "yz", "TX": TX, "L": L, "ms_y": ms[1], "ms_z": ms[2]})
                if best_y is None or ms[1] < best_y[0]:
                    best_y = (ms[1], L, TX)
                if best_z is None or ms[2] < best_z[0]:
                    best_z = (ms[2], L, TX)
        shape = (best_x[1], best_y[1], best_z[1], best_x[2], best_y[2], best_z[2])
        env.set_tiles(*shape)
        best_grid = (best_x[0], best_x[3])
        # tile sub-runs per XCD of the y and z sweeps, per axis (one launch
        # times both axes at the same run count)
        best_ry = (best_y[0], 1)
        best_rz = (best_z[0], 1)
        for R in COL_RUNS[1:]:
            env.set_col_runs(R, R)
            env.time_kernels(dt, 1)
            ms = env.time_kernels(dt, reps)
            trials.append({"axis": "yz", "runs": R, "ms_y": ms[1], "ms_z": ms[2]})
            if ms[1] < best_ry[0]:
                best_ry = (ms[1], R)
            if ms[2] < best_rz[0]:
                best_rz = (ms[2], R)
        env.set_col_runs(best_ry[1], best_rz[1])
        # schedules, by graph-replayed step time: whole-volume sweeps, the two
        # substrate halves on parallel streams, x/y k-slabs.  The k-slab
        # schedule runs the x sweep of one slab beside the y sweep of the one
        # before, and the x chunk length that is fastest alone is not always
        # the fastest there (FP64 at 800^3 x 2: 28-node chunks in slabs of 25
        # planes 9.29 ms per step, against 9.56 with the 32-node chunks that
        # win alone), so slabs are timed with the SLAB_X_SHAPES fastest chunk
        # lengths, each at its best grid.
        cands = [(0, 1, best_x[1], best_x[3])]
        if ns > 1:
            cands.append((0, 2, best_x[1], best_x[3]))
        # FP32 keeps whole-volume sweeps: on three boxes its k-slab schedule
        # came out between 2 % slower (alternating runs with 100k agents) and
        # 2 % faster (this search, without agents) at 400^3 x 4, never clearly
        # ahead, while it costs the single launch per axis whose counters
        # and roofline the rest of the measurement rests on
        fp32 = precision in ("fp32", "f32", "float", 1)
        if best_x[2] == 0 and not fp32:
            plane = nx * ny * 8
            xs = sorted(wave_best.items(), key=lambda kv: kv[1][0])[:SLAB_X_SHAPES]
            cands += [(max(1, nb // plane), 1, L, g) for nb in SLAB_BYTES if max(1, nb // plane) < nz
                      for L, (_, g) in xs]
        best_whole = best_slab = None
        for slab, split, L, g in cands:
            env.set_tiles(L, *shape[1:])
            env.set_x_grid(g)
            env.set_slab(slab)
            env.set_substrate_streams(split)
            t = _step_ms(env, dt, reps)
            trials.append({"axis": "schedule", "slab": slab, "substrate_streams": split, "Lx": L,
                           "blocks_per_cu": g, "ms_step": t})
            if slab == 0 and (best_whole is None or t < best_whole[0]):
                best_whole = (t, slab, split, L, g)
            if slab > 0 and (best_slab is None or t < best_slab[0]):
                best_slab = (t, slab, split, L, g)
        # a k-slab schedule has to win by SLAB_MARGIN: smaller differences are
        # within the spread of one measurement, and whole-volume sweeps keep
        # one launch per axis (FP32 at 400^3 x 4: slabs of 209 planes with
        # 16-node x chunks 1.229 ms against 1.236 whole-volume on one box,
        # 1.238-1.279 against 1.214-1.257 in alternating runs on another)
        best_sched = best_whole
        if best_slab is not None and best_slab[0] < SLAB_MARGIN * best_whole[0]:
            best_sched = best_slab
        env.set_slab(0)
        env.set_substrate_streams(1)
        res = {"Lx": best_sched[3], "rows_x": best_x[2], "Ly": best_y[1], "cols_y": best_y[2],
               "Lz": best_z[1], "cols_z": best_z[2], "x_grid": best_sched[4],
               "runs_y": best_ry[1], "runs_z": best_rz[1], "slab": best_sched[1],
               "substrate_streams": best_sched[2],
               "ms": {"x": wave_best[best_sched[3]][0] if best_sched[3] != best_x[1] else best_grid[0],
                      "y": best_ry[0], "z": best_rz[0], "step": best_sched[0]},
               "trials": trials}
    finally:
        env.close()
    if verbose:
        for t in trials:
            print(t)
    _cache[key] = res
    if cache_path:
        doc = {}
        if os.path.exists(cache_path):
            try:
                doc = json.load(open(cache_path))
            except Exception:
                doc = {}
        doc[key] = {k: v for k, v in res.items() if k != "trials"}
        with open(cache_path, "w") as f:
            json.dump(doc, f, indent=1)
    return res


def autotune_agents(env, dt=0.01, steps=20, reps=2):
    """Choose, for the environment's current agents and shapes, how the agent
    step stores the field (whole 128-byte lines or single voxels) and, with
    several substrates and no k-slab schedule, whether the two substrate
    halves run on parallel streams (each half's agent step then overlaps the
    other half's sweeps), by graph-replayed step time; the fastest pair is
    left set.  Every trial advances the field (reps * (steps + 4) steps per
    candidate): tune before the run.  Returns {"whole_lines",
    "substrate_streams", "ms_step": {(whole_lines, streams): ms}}."""
    splits = (1, 2) if env.ns > 1 and env.slab() == 0 else (env.substrate_streams(),)
    best = {}
    for rep in range(reps):
        for lines in (True, False):
            for split in splits:
                env.set_agent_stores(lines)
                env.set_substrate_streams(split)
                t = _step_ms(env, dt, steps // 2)
                best[(lines, split)] = min(best.get((lines, split), t), t)
    pick = min(best, key=best.get)
    env.set_agent_stores(pick[0])
    env.set_substrate_streams(pick[1])
    return {"whole_lines": pick[0], "substrate_streams": pick[1], "ms_step": best}


def apply(env, res):
    env.set_tiles(res["Lx"], res["Ly"], res["Lz"], res["rows_x"], res["cols_y"], res["cols_z"])
    env.set_x_grid(res.get("x_grid", 0))
    env.set_col_runs(res.get("runs_y", 1), res.get("runs_z", 1))
    env.set_slab(res.get("slab", 0))
    env.set_substrate_streams(res.get("substrate_streams", 1))


# ---- batches ------------------------------------------------------------------
BATCH_X_L = (8, 16)
BATCH_COL_L = (8, 16, 32)
BATCH_COL_THREADS = (256, 384, 512, 640, 768, 896, 1024)
BATCH_STREAMS = (2, 3, 4)


def autotune_batch(batch, dt=0.01, steps=4, verbose=False, finalists=3):
    """Pick the batch kernel shapes for the batch's current membership.

    Times ``steps`` whole batched steps (one captured graph replay) for every
    admissible x chunk length, y/z chunk length and threads per y/z column
    block (an item whose lines have P chunks runs as tiles of about
    threads / P columns) with one stream group; combinations the batch
    rejects (lines too long for the block) are skipped.  The ``finalists``
    fastest shapes are then timed over 3 * ``steps`` steps with one to four
    stream groups each (the best shape depends on the group count), and the
    fastest pair is kept.  Every trial advances all members: tune before
    loading the state to be solved, or reload it afterwards.  Returns
    {"tiles": (Lx, Ly, Lz, cols_y, cols_z), "col_threads", "streams",
    "ms_per_step", "trials"}.
    """
    import time

    def timed(n):
        batch.step(dt, n)  # capture and first replay
        batch.synchronize()
        t0 = time.perf_counter()
        batch.step(dt, n)
        batch.synchronize()
        return (time.perf_counter() - t0) / n * 1e3

    def apply_shape(shape, thr):
        batch.set_col_threads(0)
        batch.set_tiles(*shape)
        batch.set_col_threads(thr)

    trials = []
    batch.set_streams(1)
    for lx in BATCH_X_L:
        for L in BATCH_COL_L:
            for thr in BATCH_COL_THREADS:
                shape = (lx, L, L, 16, 16)
                try:
                    apply_shape(shape, thr)
                except Exception:
                    continue
                trials.append({"tiles": list(shape), "col_threads": thr, "streams": 1, "ms_per_step": timed(steps)})
                if verbose:
                    print(json.dumps(trials[-1]), flush=True)
    if not trials:
        raise RuntimeError("no admissible batch kernel shape")
    cands = []
    for t in sorted(trials, key=lambda t: t["ms_per_step"]):
        c = (tuple(t["tiles"]), t["col_threads"])
        if c not in cands:
            cands.append(c)
        if len(cands) == finalists:
            break
    best = None
    for shape, thr in cands:
        apply_shape(shape, thr)
        for groups in (1,) + BATCH_STREAMS:
            batch.set_streams(groups)
            ms = timed(3 * steps)
            trials.append({"tiles": list(shape), "col_threads": thr, "streams": groups, "ms_per_step": ms})
            if verbose:
                print(json.dumps(trials[-1]), flush=True)
            if best is None or ms < best[3]:
                best = (shape, thr, groups, ms)
    batch.set_streams(best[2])
    apply_shape(best[0], best[1])
    return {"tiles": best[0], "col_threads": best[1], "streams": best[2], "ms_per_step": best[3],
            "trials": len(trials)}


