"""CPU tests of the batch autotuner's search: every admissible combination of
chunk lengths and column-block size is timed once, rejected ones are
skipped, and the fastest is left applied.  Driven through a stand-in batch
whose step time is a known function of the shape (on a fake clock)."""
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
