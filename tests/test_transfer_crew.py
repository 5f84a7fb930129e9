"""The host-thread crew of the pinned transfer pipeline, on the CPU: copies
of many sizes through crews of 1-16 threads are byte-exact and stay inside
the destination, element jobs visit every element once, and the substrate
runs of the FP32 converting transfers follow the field layouts (driven by
tests/cpp/copy_crew_test, no GPU work)."""
import os
import subprocess

from helpers import ROOT

TEST = os.path.join(ROOT, "tests", "cpp", "copy_crew_test")


def test_copy_crew_is_byte_exact():
    assert os.path.exists(TEST), "build the C++ host side first (make -C physicell-x_amd/host)"
    r = subprocess.run([TEST], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
