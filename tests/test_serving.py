"""CPU tests of the serving loop's host logic (admission by device-memory
budget, FIFO order, exact per-job step counts, retirement, rejection), driven
through a stand-in batch that records what the loop asks of it."""
import importlib.util
import os
import sys

import numpy as np
import pytest

from helpers import ROOT


def load_serving():
    path = os.path.join(ROOT, "physicell-x_amd", "serving.py")
    spec = importlib.util.spec_from_file_location("bfa_serving_under_test", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[spec.name] = mod
    spec.loader.exec_module(mod)
    return mod


class FakeError(RuntimeError):
    pass


class FakeBatch:
    def __init__(self, precision):
        self.members = {}
        self.steps_done = {}
        self.member_steps = {}
        self.calls = []
        self.next_id = 0
        self.fail_add = False
        self.fail_setup = False

    def set_tiles(self, *a):
        self.tiles = a

    def add(self, nx, ny, nz, ns, dx, D, decay, act, reg, fv):
        if self.fail_add:
            raise FakeError("out of memory")
        mid = self.next_id
        self.next_id += 1
        self.members[mid] = (nx, ny, nz, ns)
        self.steps_done[mid] = 0
        self.member_steps[mid] = 1 << 30
        return mid

    def fill(self, mid, vals):
        self.calls.append(("fill", mid))

    def upload(self, mid, f):
        self.calls.append(("upload", mid))

    def set_agents(self, *a):
        if self.fail_setup:
            raise FakeError("bad agent arrays")
        self.calls.append(("agents", a[0]))

    def set_member_steps(self, mid, s):
        self.member_steps[mid] = s

    def step(self, dt, k):
        self.calls.append(("step", k, sorted(self.members)))
        for mid in self.members:
            self.steps_done[mid] += min(k, self.member_steps[mid])

    def synchronize(self):
        pass

    def download(self, mid):
        return np.full(4, float(self.steps_done[mid]))

    def total_mass(self, mid):
        return np.zeros(self.members[mid][3])

    def remove(self, mid):
        del self.members[mid]

    def close(self):
        pass


class FakeEnv:
    """A solo job's Microenvironment: counts the steps it is asked for."""

    def __init__(self, pkg, nx, ny, nz, ns, dx, precision):
        if pkg.fail_env:
            raise FakeError("out of memory")
        self.pkg, self.ns, self.nvox = pkg, ns, nx * ny * nz
        self.steps_done, self.pending, self.closed = 0, 0, False
        self.uploaded = None
        pkg.envs.append(self)

    def set_coefficients(self, D, decay):
        pass

    def set_dirichlet(self, act, reg, fv):
        pass

    def upload(self, field):
        self.uploaded = np.asarray(field)

    def fill(self, values):
        self.uploaded = np.tile(np.asarray(values, dtype=np.float64), self.nvox)

    def set_agents(self, *a):
        pass

    def step(self, dt, k):
        self.pending += k   # asynchronous until synchronize
        self.pkg.calls.append(("env_step", self.nvox, k))

    def synchronize(self):
        self.steps_done += self.pending
        self.pending = 0

    def download(self):
        assert self.pending == 0
        return np.full(4, float(self.steps_done))

    def total_mass(self):
        return np.zeros(self.ns)

    def close(self):
        self.closed = True


class FakePkg:
    BioFVMError = FakeError

    def __init__(self):
        self.batch = None
        self.envs = []
        self.calls = []
        self.fail_env = False

    def Batch(self, precision):
        self.batch = FakeBatch(precision)
        return self.batch

    def Microenvironment(self, *a):
        return FakeEnv(self, *a)

    @staticmethod
    def member_bytes(nx, ny, nz, ns, precision):
        return nx * ny * nz * ns * 4


def job(S, n, ns, steps, **kw):
    return S.Job(nx=n, ny=n, nz=n, ns=ns, steps=steps, D=np.ones(ns), decay=np.ones(ns),
                 want_field=True, **kw)


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 1.0
        return self.t


def test_admission_respects_budget_and_fifo():
    S = load_serving()
    pkg = FakePkg()
    srv = S.Server(pkg, hbm_budget=2 * 10 ** 3 * 4 * 2, steps_per_iteration=4, clock=Clock())
    jobs = [job(S, 10, 2, 6), job(S, 10, 2, 3), job(S, 10, 2, 5)]  # 8000 bytes each, budget 16000
    for j in jobs:
        srv.submit(j)
    srv.iterate()
    assert pkg.batch.calls[-1] == ("step", 4, [0, 1])   # third waits for memory
    assert [j.id for j in srv.queue] == [2]
    assert jobs[1].remaining == 0 and jobs[1] in srv.done  # 3 steps: done inside the first call
    assert jobs[0].remaining == 2
    srv.iterate()                                        # job 2 admitted into the freed room
    assert pkg.batch.calls[-1] == ("step", 4, [0, 2])
    srv.drain()
    assert all(j.remaining == 0 for j in jobs)
    # every job got exactly its own number of steps
    assert [float(j.result[0]) for j in jobs] == [6.0, 3.0, 5.0]
    assert srv.used == 0 and not srv.active
    # FIFO: the third job was admitted only after the second retired
    assert jobs[2].admitted > jobs[1].finished


def test_rejects_oversized_and_mismatched_jobs():
    S = load_serving()
    pkg = FakePkg()
    srv = S.Server(pkg, hbm_budget=1000, max_edge=16, clock=Clock())
    srv.submit(job(S, 10, 1, 3))            # 4000 bytes > budget
    srv.submit(job(S, 20, 1, 3))            # edge above limit
    srv.submit(job(S, 4, 1, 3, dt=0.02))    # other dt
    assert len(srv.rejected) == 3 and not srv.queue
    assert all(j.error for j in srv.rejected)


def test_allocation_failure_requeues_when_others_run():
    S = load_serving()
    pkg = FakePkg()
    srv = S.Server(pkg, hbm_budget=1e9, steps_per_iteration=2, clock=Clock())
    a = job(S, 4, 1, 4)
    srv.submit(a)
    srv.iterate()
    pkg.batch.fail_add = True
    b = job(S, 4, 1, 2)
    srv.submit(b)
    srv.iterate()                      # b cannot be allocated yet: stays queued
    assert list(srv.queue) == [b] and a.remaining == 0
    pkg.batch.fail_add = False
    srv.drain()
    assert b.remaining == 0 and float(b.result[0]) == 2.0
    # with nothing else running, an allocation failure is reported
    pkg.batch.fail_add = True
    srv.submit(job(S, 4, 1, 2))
    with pytest.raises(FakeError):
        srv.iterate()


def test_large_jobs_run_solo_beside_the_batch():
    S = load_serving()
    pkg = FakePkg()
    srv = S.Server(pkg, hbm_budget=1e9, steps_per_iteration=4, clock=Clock(), solo_voxels=1000)
    small = [job(S, 8, 2, 6), job(S, 9, 1, 3)]                       # 512, 729 voxels: batch
    big = [job(S, 10, 2, 5), job(S, 12, 3, 9, init_values=[1.0, 2.0, 3.0])]  # 1000, 1728: solo
    for j in small + big:
        srv.submit(j)
    srv.iterate()
    assert pkg.batch.calls[-1] == ("step", 4, [0, 1]) and len(pkg.envs) == 2
    assert [c for c in pkg.calls if c[0] == "env_step"] == [("env_step", 1000, 4), ("env_step", 1728, 4)]
    # a uniform initial value is uploaded in the reference layout (substrate fastest)
    np.testing.assert_array_equal(pkg.envs[1].uploaded[:6], [1.0, 2.0, 3.0, 1.0, 2.0, 3.0])
    srv.drain()
    assert [float(j.result[0]) for j in small + big] == [6.0, 3.0, 5.0, 9.0]
    assert all(e.closed for e in pkg.envs) and srv.used == 0 and not srv.solo
    # each solo job was asked for exactly its own number of steps
    per_job = {}
    for _, nvox, k in (c for c in pkg.calls if c[0] == "env_step"):
        per_job[nvox] = per_job.get(nvox, 0) + k
    assert per_job == {1000: 5, 1728: 9}


def test_solo_allocation_failure_requeues_when_others_run():
    S = load_serving()
    pkg = FakePkg()
    srv = S.Server(pkg, hbm_budget=1e9, steps_per_iteration=2, clock=Clock(), solo_voxels=1000)
    a = job(S, 4, 1, 6)                 # batch
    srv.submit(a)
    srv.iterate()
    pkg.fail_env = True
    b = job(S, 10, 1, 2)                # solo, cannot be allocated yet
    srv.submit(b)
    srv.iterate()
    assert list(srv.queue) == [b] and srv.used == a.bytes
    pkg.fail_env = False
    srv.drain()
    assert b.remaining == 0 and float(b.result[0]) == 2.0 and a.remaining == 0
    pkg.fail_env = True
    srv.submit(job(S, 10, 1, 2))        # nothing else on the device: reported
    with pytest.raises(FakeError):
        srv.iterate()


def test_latency_stats_and_arrivals():
    S = load_serving()
    pkg = FakePkg()
    srv = S.Server(pkg, hbm_budget=1e9, steps_per_iteration=5)
    jobs = [job(S, 4, 1, s) for s in (3, 7, 11, 2)]
    wall, done = S.run_arrivals(srv, jobs, np.array([0.0, 0.0, 0.001, 0.002]))
    assert len(done) == 4 and wall >= 0.0
    st = S.latency_stats(done)
    assert st["p50_s"] <= st["p99_s"] <= st["max_s"]
    assert sorted(float(j.result[0]) for j in done) == [2.0, 3.0, 7.0, 11.0]


def test_member_set_up_failure_frees_the_member():
    S = load_serving()
    pkg = FakePkg()
    srv = S.Server(pkg, hbm_budget=1e9, steps_per_iteration=2, clock=Clock())
    srv.iterate()
    pkg.batch.fail_setup = True
    bad = job(S, 4, 1, 2, agent_voxel=np.array([0, 1], np.int32), agent_c1=np.zeros(2),
              agent_c2=np.ones(2), agent_e2=np.zeros(2))
    srv.submit(bad)
    with pytest.raises(FakeError):
        srv.iterate()
    # the member added for it is gone and nothing is charged to the budget
    assert not pkg.batch.members and srv.used == 0 and not srv.active
    pkg.batch.fail_setup = False
    ok = job(S, 4, 1, 2)
    srv.submit(ok)
    srv.drain()
    assert float(ok.result[0]) == 2.0 and srv.used == 0
