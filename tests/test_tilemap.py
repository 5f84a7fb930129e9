"""The workgroup-to-tile maps of the sweep launches, on the CPU: xcd_runs is a
bijection for every launch size and run count tried, equals xcd_block with
one run, gives each XCD the same share of tiles, and advances each sub-run
through consecutive tiles (driven by tests/cpp/tilemap_test, no GPU work)."""
import os
import subprocess

from helpers import ROOT

TEST = os.path.join(ROOT, "tests", "cpp", "tilemap_test")


def test_tile_maps_are_bijections_with_contiguous_runs():
    assert os.path.exists(TEST), "build the C++ host side first (make -C physicell-x_amd/host)"
    r = subprocess.run([TEST], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 failures" in r.stdout
