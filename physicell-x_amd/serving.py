"""Serving loop: continuous batching of diffusion-decay solve jobs on one GPU.

A job is one microenvironment (grid, substrates, coefficients, Dirichlet
state, initial condition, agents) to be advanced a given number of steps.  The
server keeps every admitted job's field resident in device memory between
steps and advances all of them together: each iteration runs up to
``steps_per_iteration`` steps, every job taking exactly the steps it still
needs.  Jobs below ``solo_voxels`` voxels are members of one ``Batch`` (one
persistent kernel per sweep over all of them; the batch skips a member's work
once its count for the call is used).  Larger jobs, whose grids keep the
single-environment kernels near their roofline, run each as its own
``Microenvironment`` on its own stream, launched next to the batch so that
their kernels overlap on the device.  Jobs that arrive while others are on
the GPU wait in a FIFO queue and are admitted at the next iteration
boundary, as soon as the device-memory budget has room for them; finished
jobs are retired at once (result downloaded, memory freed), which admits the
next ones.

Which path a job takes depends on its own size only, and the batch's kernel
shapes are fixed when the server starts (a solo job's shapes follow from its
grid), so a job's result does not depend on which other jobs shared its
iterations.

There is no counterpart in the reference (it advances one Microenvironment
per process); per job the arithmetic is that of
``Microenvironment.simulate_diffusion_decay`` + ``simulate_cell_sources_and_sinks``.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field
from typing import Optional

import numpy as np


@dataclass
class Job:
    nx: int
    ny: int
    nz: int
    ns: int
    steps: int
    D: np.ndarray
    decay: np.ndarray
    dx: float = 20.0
    dt: float = 0.01
    active: Optional[np.ndarray] = None
    face_registered: Optional[np.ndarray] = None
    face_values: Optional[np.ndarray] = None
    init_values: Optional[np.ndarray] = None   # uniform start per substrate ...
    init_field: Optional[np.ndarray] = None    # ... or a full field (reference layout)
    agent_voxel: Optional[np.ndarray] = None
    agent_c1: Optional[np.ndarray] = None
    agent_c2: Optional[np.ndarray] = None
    agent_e2: Optional[np.ndarray] = None
    want_field: bool = False
    # filled in by the server
    id: int = -1
    arrival: float = 0.0
    admitted: float = 0.0
    finished: float = 0.0
    remaining: int = 0
    member: int = -1
    env: object = None   # the job's own Microenvironment (solo jobs)
    bytes: int = 0
    mass: Optional[np.ndarray] = None
    result: Optional[np.ndarray] = None
    error: Optional[str] = None

    @property
    def updates(self):
        return self.nx * self.ny * self.nz * self.ns * self.steps

    @property
    def latency(self):
        return self.finished - self.arrival


class Server:
    """Continuous-batching solve server over one ``Batch``.

    ``hbm_budget`` bounds the device bytes of admitted jobs (per
    ``member_bytes``); a job larger than the whole budget is rejected at
    submission.  ``max_edge`` fixes the batch kernel shapes for the largest
    grid edge the server accepts.  Jobs of at least ``solo_voxels`` voxels run
    as their own Microenvironment next to the batch (None: all in the batch).
    """

    def __init__(self, pkg, precision="fp32", hbm_budget=200e9, steps_per_iteration=10,
                 max_edge=256, dt=0.01, clock=time.perf_counter, tiles=None, solo_voxels=None):
        self.pkg = pkg
        self.precision = precision
        self.batch = pkg.Batch(precision)
        self.budget = float(hbm_budget)
        self.k = int(steps_per_iteration)
        self.max_edge = int(max_edge)
        self.dt = float(dt)
        self.clock = clock
        # shapes that cover max_edge: x rows in <= 64 lanes; y/z lines in
        # chunks of 16 nodes, as many columns per tile (<= 64) as the block
        # holds (1024 threads, 512 for chunks of 128 bytes or more)
        # (tiles: an explicit (Lx, Ly, Lz, cols_y, cols_z) instead)
        lx = 8 if max_edge <= 512 else 16
        lyz = 16 if max_edge <= 512 else 32
        limit = 512 if lyz * (4 if precision == "fp32" else 8) >= 128 else 1024
        p = -(-max_edge // lyz)
        cols = next((c for c in (64, 32, 16) if c * p <= limit), 16)
        self.batch.set_tiles(*(tiles if tiles is not None else (lx, lyz, lyz, cols, cols)))
        self.solo_voxels = solo_voxels
        self.queue = collections.deque()
        self.active = {}    # batch member id -> job
        self.solo = {}      # job id -> job (own Microenvironment)
        self.done = []
        self.rejected = []
        self.used = 0
        self.iterations = 0
        self._next_id = 0

    def close(self):
        for job in self.solo.values():
            job.env.close()
            job.env = None
        self.solo.clear()
        self.batch.close()

    # -- submission and admission ------------------------------------------------
    def submit(self, job: Job, arrival: Optional[float] = None) -> int:
        job.id = self._next_id
        self._next_id += 1
        job.arrival = self.clock() if arrival is None else arrival
        job.remaining = int(job.steps)
        job.bytes = self.pkg.member_bytes(job.nx, job.ny, job.nz, job.ns, self.precision)
        if max(job.nx, job.ny, job.nz) > self.max_edge:
            job.error = f"edge above the server limit {self.max_edge}"
            self.rejected.append(job)
        elif job.bytes > self.budget:
            job.error = "larger than the device-memory budget"
            self.rejected.append(job)
        elif abs(job.dt - self.dt) > 0.0:
            job.error = f"dt {job.dt} differs from the server's {self.dt}"
            self.rejected.append(job)
        else:
            self.queue.append(job)
        return job.id

    def is_solo(self, job):
        return self.solo_voxels is not None and job.nx * job.ny * job.nz >= self.solo_voxels

    def _admit_solo(self, job, act, reg, fv):
        env = self.pkg.Microenvironment(job.nx, job.ny, job.nz, job.ns, job.dx, self.precision)
        try:
            env.set_coefficients(job.D, job.decay)
            env.set_dirichlet(act, reg, fv)
            if job.init_field is not None:
                env.upload(job.init_field)
            else:
                env.fill(job.init_values if job.init_values is not None else np.zeros(job.ns))
            if job.agent_voxel is not None and len(job.agent_voxel):
                env.set_agents(job.agent_voxel, job.agent_c1, job.agent_c2, job.agent_e2)
        except Exception:
            env.close()
            raise
        job.env = env
        self.solo[job.id] = job

    def _admit(self):
        admitted = 0
        while self.queue and self.used + self.queue[0].bytes <= self.budget:
            job = self.queue.popleft()
            ns = job.ns
            act = job.active if job.active is not None else np.zeros(ns, np.uint8)
            reg = job.face_registered if job.face_registered is not None else np.zeros(6, np.uint8)
            fv = job.face_values if job.face_values is not None else np.ones((6, ns))
            if self.is_solo(job):
                try:
                    self._admit_solo(job, act, reg, fv)
                except self.pkg.BioFVMError:  # device allocation failed: retry later
                    self.queue.appendleft(job)
                    if not self.busy_on_device():
                        raise
                    break
                job.admitted = self.clock()
                self.used += job.bytes
                admitted += 1
                continue
            try:
                mid = self.batch.add(job.nx, job.ny, job.nz, ns, job.dx, job.D, job.decay, act, reg, fv)
            except self.pkg.BioFVMError as exc:  # device allocation failed: retry later
                self.queue.appendleft(job)
                if not self.busy_on_device():
                    raise
                del exc
                break
            try:
                if job.init_field is not None:
                    self.batch.upload(mid, job.init_field)
                else:
                    vals = job.init_values if job.init_values is not None else np.zeros(ns)
                    self.batch.fill(mid, vals)
                if job.agent_voxel is not None and len(job.agent_voxel):
                    self.batch.set_agents(mid, job.agent_voxel, job.agent_c1, job.agent_c2, job.agent_e2)
            except Exception:
                # the member was allocated but not set up: free it so its
                # device memory does not leak outside the budget, and reject
                # the job (its inputs are what failed)
                self.batch.remove(mid)
                raise
            job.member = mid
            job.admitted = self.clock()
            self.used += job.bytes
            self.active[mid] = job
            admitted += 1
        return admitted

    # -- the loop ------------------------------------------------------------------------
    def busy_on_device(self):
        return bool(self.active or self.solo)

    def busy(self):
        return bool(self.queue or self.active or self.solo)

    def iterate(self):
        """Admit what fits, run one iteration of up to k steps, retire finished jobs.
        Returns the jobs finished in this iteration."""
        self._admit()
        if not self.busy_on_device():
            return []
        k = self.k
        # solo jobs first (each asynchronous on its own stream), then the batch;
        # all of them run concurrently until the synchronisations below
        for job in self.solo.values():
            job.env.step(self.dt, min(k, job.remaining))
        if self.active:
            # a fixed run length keeps one captured graph; jobs needing fewer
            # steps stop early inside it
            for mid, job in self.active.items():
                self.batch.set_member_steps(mid, min(k, job.remaining))
            self.batch.step(self.dt, k)
        for job in self.solo.values():
            job.env.synchronize()
        if self.active:
            self.batch.synchronize()
        self.iterations += 1
        finished = []
        for jid in list(self.solo):
            job = self.solo[jid]
            job.remaining -= min(k, job.remaining)
            if job.remaining == 0:
                if job.want_field:
                    job.result = job.env.download()
                job.mass = job.env.total_mass() if job.want_field else None
                job.env.close()
                job.env = None
                self.used -= job.bytes
                job.finished = self.clock()
                del self.solo[jid]
                self.done.append(job)
                finished.append(job)
        for mid in list(self.active):
            job = self.active[mid]
            job.remaining -= min(k, job.remaining)
            if job.remaining == 0:
                if job.want_field:
                    job.result = self.batch.download(mid)
                job.mass = self.batch.total_mass(mid) if job.want_field else None
                self.batch.remove(mid)
                self.used -= job.bytes
                job.finished = self.clock()
                del self.active[mid]
                self.done.append(job)
                finished.append(job)
        return finished

    def drain(self):
        while self.busy():
            self.iterate()


def run_arrivals(server: Server, jobs, offsets, clock=time.perf_counter):
    """Submit jobs at the given arrival offsets (seconds from now) while the
    server iterates; returns (wall seconds, finished jobs)."""
    order = np.argsort(offsets, kind="stable")
    pending = collections.deque((offsets[i], jobs[i]) for i in order)
    t0 = clock()
    while pending or server.busy():
        now = clock() - t0
        while pending and pending[0][0] <= now:
            off, job = pending.popleft()
            server.submit(job, arrival=t0 + off)
        if server.busy():
            server.iterate()
        elif pending:
            time.sleep(max(0.0, min(0.01, pending[0][0] - (clock() - t0))))
    return clock() - t0, list(server.done)


def latency_stats(jobs):
    lat = np.array([j.latency for j in jobs])
    if lat.size == 0:
        return {"p50_s": None, "p99_s": None, "mean_s": None}
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max())}
