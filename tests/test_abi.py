"""CPU checks of the C ABI: the library loads, exports every function the
header declares, and rejects bad arguments before touching a device."""
import ctypes
import os
import re

import pytest

from helpers import ROOT, load_pkg

HEADER = os.path.join(ROOT, "include", "biofvm_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bfa_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("bfa_create", "bfa_destroy", "bfa_set_coefficients", "bfa_set_dirichlet",
                     "bfa_upload", "bfa_download", "bfa_set_agents", "bfa_diffusion_decay",
                     "bfa_secretion_uptake", "bfa_step", "bfa_total_mass", "bfa_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    pkg = load_pkg()
    L = pkg.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


@pytest.mark.parametrize("dims", [(0, 4, 4, 1), (4, 4, 2000, 1), (1025, 4, 4, 1), (4, 1025, 4, 1),
                                  (4, 4, 4, 0), (4, 4, 4, 65)])
def test_create_rejects_bad_sizes_without_a_device(dims):
    pkg = load_pkg()
    L = pkg.lib()
    h = ctypes.c_void_p()
    rc = L.bfa_create(ctypes.byref(h), *dims, 20.0, 0)
    assert rc == 1  # BFA_ERR_ARG
    assert not h.value
    assert b"outside" in L.bfa_last_error(None)


def test_create_rejects_bad_precision_and_dx():
    L = load_pkg().lib()
    h = ctypes.c_void_p()
    assert L.bfa_create(ctypes.byref(h), 4, 4, 4, 1, 20.0, 7) == 1
    assert L.bfa_create(ctypes.byref(h), 4, 4, 4, 1, -1.0, 0) == 1
    assert L.bfa_create(ctypes.byref(h), 4, 4, 4, 1, float("nan"), 0) == 1
