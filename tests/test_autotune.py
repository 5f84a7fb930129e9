"""CPU tests of the autotuners' searches: every admissible combination of
chunk lengths and column-block size is timed once, rejected ones are
skipped, and the fastest is left applied.  Driven through a stand-in batch
or microenvironment whose times are a known function of the shape (on a
fake clock)."""
import importlib.util
import os
import sys
import time

import pytest

from helpers import ROOT


def load_autotune():
    path = os.path.join(ROOT, "physicell-x_amd", "autotune.py")
    spec = importlib.util.spec_from_file_location("bfa_autotune_under_test", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[spec.name] = mod
    spec.loader.exec_module(mod)
    return mod


class Rejected(RuntimeError):
    pass


class FakeBatch:
    """Longest lines of 160 nodes; FP32 rules: a block holds at most 1024
    threads (512 for chunks of 128 bytes or more) and at least 16 columns of
    the longest lines."""

    def __init__(self, clock):
        self.clock = clock
        self.tiles = None
        self.threads = 0
        self.groups = 1
        self.timed = []

    def _limit(self, L):
        return 512 if L * 4 >= 128 else 1024

    def set_tiles(self, lx, ly, lz, cy, cz):
        P = -(-160 // ly)
        if P * cy > self._limit(ly) or (self.threads and not 16 * P <= self.threads <= self._limit(ly)):
            raise Rejected("shape")
        self.tiles = (lx, ly, lz, cy, cz)

    def set_col_threads(self, thr):
        if thr % 64:
            raise Rejected("threads")
        if thr and self.tiles is not None:
            P = -(-160 // self.tiles[1])
            if not 16 * P <= thr <= self._limit(self.tiles[1]):
                raise Rejected("threads")
        self.threads = thr

    def set_streams(self, g):
        if g not in (1, 2, 3, 4):
            raise Rejected("groups")
        self.groups = g

    def cost(self):
        lx, L = self.tiles[0], self.tiles[1]
        # unique minimum at Lx 8, L 16, 768 threads; two stream groups 5 % faster
        c = 3.0 + 0.2 * (lx != 8) + 0.1 * abs(L - 16) / 8 + abs(self.threads - 768) / 1000.0
        return c * (0.95 if self.groups == 2 else 1.0)

    def step(self, dt, n):
        self.clock.t += self.cost() * n * 1e-3
        self.timed.append((self.tiles, self.threads, self.groups))

    def synchronize(self):
        pass


class Clock:
    t = 0.0

    def __call__(self):
        return self.t


def test_batch_autotune_picks_the_fastest_admissible_shape(monkeypatch):
    at = load_autotune()
    clock = Clock()
    monkeypatch.setattr(time, "perf_counter", clock)
    b = FakeBatch(clock)
    res = at.autotune_batch(b, 0.01, steps=4)
    assert res["tiles"][:3] == (8, 16, 16) and res["col_threads"] == 768 and res["streams"] == 2
    assert res["ms_per_step"] == pytest.approx(3.0 * 0.95)
    assert b.tiles[:3] == (8, 16, 16) and b.threads == 768 and b.groups == 2  # left applied
    tried = {(t[0], t[1], thr) for t, thr, g in b.timed if g == 1}
    # the three fastest single-group shapes are the finalists, each timed
    # with one to four groups: (L 16, 768 threads) 3.0, (L 8, 768) 3.1, then
    # (L 16, 640) 3.128, timed before the equally fast (L 16, 896)
    finals = {(t[0], t[1], thr, g) for t, thr, g in b.timed if g > 1}
    assert finals == {(8, L, thr, g) for L, thr in ((16, 768), (8, 768), (16, 640)) for g in (2, 3, 4)}
    # every admissible combination was timed, no rejected one was
    for lx in at.BATCH_X_L:
        for L in at.BATCH_COL_L:
            for thr in at.BATCH_COL_THREADS:
                P = -(-160 // L)
                ok = 16 * P <= thr <= (512 if L * 4 >= 128 else 1024)
                assert ((lx, L, thr) in tried) == ok, (lx, L, thr)
    assert res["trials"] == len(tried) + 3 * 4


class FakeEnv:
    """Stand-in Microenvironment of a 400-node cube: kernel times are a known
    function of the shape, with unique minima at x chunk 8 with 4 blocks per
    CU, y 25-node chunks x 32 columns in one tile run, z the same shape in
    3 runs, and the two-substrate-group schedule 2 % faster."""

    n = 400
    log = []

    def __init__(self, nx, ny, nz, ns, dx, precision):
        self.shape = (8, 16, 16, 0, 32, 32)
        self.grid = 0
        self.runs = (1, 1)
        self.slab = 0
        self.split = 1
        FakeEnv.log.append(self)
        self.timed = []

    def set_coefficients(self, D, decay):
        pass

    def set_dirichlet(self, *a):
        pass

    def upload(self, *a):
        pass

    def fill(self, *a):
        pass

    def set_tiles(self, lx, ly, lz, rx, cy, cz):
        ly, lz = ly or 16, lz or 16  # 0: the library's default column shape
        cy, cz = cy or 32, cz or 32
        if rx == 0 and -(-self.n // lx) > 64:
            raise Rejected("x chunk")
        for L, TX in ((ly, cy), (lz, cz)):
            if TX * -(-self.n // L) > 1024:
                raise Rejected("column block")
        self.shape = (lx, ly, lz, rx, cy, cz)

    def set_x_grid(self, g):
        self.grid = g

    def x_grid(self):
        return self.grid, 6 if self.shape[0] == 8 else 5

    def set_col_runs(self, ry, rz):
        self.runs = (ry, rz)

    def set_slab(self, s):
        self.slab = s

    def set_substrate_streams(self, g):
        self.split = g

    def kernel_ms(self):
        lx, ly, lz, rx, cy, cz = self.shape
        x = 0.40 + (0.05 if rx else 0.0) + 0.01 * abs(lx - 8) / 4 + 0.002 * abs((self.grid or 6) - 4)
        y = 0.37 + 0.003 * abs(ly - 25) + 0.0004 * abs(cy - 32) + 0.002 * (self.runs[0] - 1)
        z = 0.42 + 0.003 * abs(lz - 25) + 0.0004 * abs(cz - 32) + 0.004 * abs(self.runs[1] - 3)
        return [x, y, z, 0.03]

    def time_kernels(self, dt, reps):
        self.timed.append((self.shape, self.grid, self.runs))
        return self.kernel_ms()

    def step(self, dt, n):
        t = sum(self.kernel_ms()) * (0.98 if self.split == 2 else 1.0) * (1.01 if self.slab else 1.0)
        self.clock.t += t * n * 1e-3

    def synchronize(self):
        pass

    def close(self):
        pass


def test_autotune_picks_the_fastest_shapes_grid_runs_and_schedule(monkeypatch):
    at = load_autotune()
    clock = Clock()
    monkeypatch.setattr(time, "perf_counter", clock)
    FakeEnv.clock = clock
    FakeEnv.log = []

    class Pkg:
        Microenvironment = FakeEnv

    res = at.autotune(Pkg, 400, 400, 400, 4, "fp32", 0.01)
    assert (res["Lx"], res["rows_x"], res["x_grid"]) == (8, 0, 4)
    assert (res["Ly"], res["cols_y"], res["Lz"], res["cols_z"]) == (25, 32, 25, 32)
    assert (res["runs_y"], res["runs_z"]) == (1, 3)
    assert (res["slab"], res["substrate_streams"]) == (0, 2)
    assert not [t for t in res["trials"] if t.get("axis") == "schedule" and t["slab"]]  # FP32: none
    assert res["ms"]["z"] == pytest.approx(0.42) and res["ms"]["y"] == pytest.approx(0.37)
    env = FakeEnv.log[0]
    timed = env.timed
    # every admissible y/z shape was timed at one tile run, every run count at the chosen shape
    yz = {(s[1], s[4]) for s, g, r in timed if s[0] == 8 and s[3] == 0 and r == (1, 1)}
    for TX in at.COL_TX:
        for L in at.COL_L:
            ok = TX * -(-400 // L) <= 1024
            assert ((L, TX) in yz) == ok, (L, TX)
    runs = {r for s, g, r in timed if s == (8, 25, 25, 0, 32, 32)}
    assert runs == {(R, R) for R in at.COL_RUNS}
    # the result applies back onto an environment
    e2 = FakeEnv(400, 400, 400, 4, 20.0, "fp32")
    at.apply(e2, res)
    assert e2.shape == (8, 25, 25, 0, 32, 32) and e2.grid == 4 and e2.runs == (1, 3) and e2.split == 2


class FakeSlabEnv(FakeEnv):
    """FakeEnv whose k-slab schedule runs 10 % faster with x chunks of 16 at
    four blocks per CU, a shape that is not the fastest x sweep alone."""

    log = []
    gain = 0.9

    def __init__(self, *a):
        super().__init__(*a)
        FakeSlabEnv.log.append(self)

    def step(self, dt, n):
        t = sum(self.kernel_ms()) * (0.98 if self.split == 2 else 1.0)
        if self.slab:
            t *= self.gain if (self.shape[0], self.grid) == (16, 4) else 1.01
        self.clock.t += t * n * 1e-3


def test_autotune_times_slabs_with_other_x_chunk_lengths(monkeypatch):
    at = load_autotune()
    clock = Clock()
    monkeypatch.setattr(time, "perf_counter", clock)
    FakeEnv.clock = clock
    FakeSlabEnv.log = []

    class Pkg:
        Microenvironment = FakeSlabEnv

    res = at.autotune(Pkg, 400, 400, 400, 4, "fp64", 0.01)
    # the x sweep alone is fastest at chunk 8 (grid 4); the slab schedule with
    # chunk 16 at its own best grid (4) wins the whole step
    assert res["slab"] > 0 and res["substrate_streams"] == 1
    assert (res["Lx"], res["x_grid"], res["rows_x"]) == (16, 4, 0)
    assert res["ms"]["x"] == pytest.approx(0.42)  # the chosen chunk's own x time
    assert (res["Ly"], res["cols_y"], res["Lz"], res["cols_z"]) == (25, 32, 25, 32)
    e2 = FakeEnv(400, 400, 400, 4, 20.0, "fp32")
    at.apply(e2, res)
    assert e2.shape == (16, 25, 25, 0, 32, 32) and e2.grid == 4 and e2.slab == res["slab"]


def test_autotune_keeps_whole_volume_sweeps_unless_slabs_win_clearly(monkeypatch):
    """A k-slab schedule 1 % faster than the best whole-volume one is not
    taken: it has to win by the margin.  Whole volume: kernels 1.22 ms x 0.98
    (two substrate groups) = 1.1956; slabs with 16-node x chunks: 1.24 x 0.955
    = 1.1842, 0.990 of it."""
    at = load_autotune()
    clock = Clock()
    monkeypatch.setattr(time, "perf_counter", clock)
    FakeEnv.clock = clock

    class Near(FakeSlabEnv):
        gain = 0.955

    class Pkg:
        Microenvironment = Near

    res = at.autotune(Pkg, 400, 400, 400, 4, "fp64", 0.01)
    assert (res["slab"], res["substrate_streams"]) == (0, 2)
    assert (res["Lx"], res["x_grid"]) == (8, 4)
    slabs = [t["ms_step"] for t in res["trials"] if t.get("axis") == "schedule" and t["slab"]]
    assert min(slabs) == pytest.approx(0.990 * res["ms"]["step"], rel=1e-3)


class FakeAgentsEnv:
    """Step time 1.00 ms with whole-line agent stores, 0.97 with single
    voxels; 2 % less on two substrate streams with single voxels only."""

    ns = 4

    def __init__(self, clock):
        self.clock = clock
        self.lines = True
        self.split = 1
        self.steps = 0

    def set_agent_stores(self, lines):
        self.lines = bool(lines)

    def set_substrate_streams(self, g):
        self.split = g

    def substrate_streams(self):
        return self.split

    def slab(self):
        return 0

    def step(self, dt, n):
        t = 1.0 if self.lines else (0.97 * (0.98 if self.split == 2 else 1.0))
        self.clock.t += t * n * 1e-3
        self.steps += n

    def synchronize(self):
        pass


def test_autotune_agents_keeps_the_fastest_store_form_and_schedule(monkeypatch):
    at = load_autotune()
    clock = Clock()
    monkeypatch.setattr(time, "perf_counter", clock)
    env = FakeAgentsEnv(clock)
    res = at.autotune_agents(env, 0.01, steps=20, reps=2)
    assert res["whole_lines"] is False and res["substrate_streams"] == 2
    assert env.lines is False and env.split == 2
    assert res["ms_step"][(True, 1)] == pytest.approx(1.0)
    assert res["ms_step"][(False, 2)] == pytest.approx(0.97 * 0.98)
    assert env.steps == 2 * 4 * (20 + 4)
