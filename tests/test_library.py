"""The HIP library as a drop-in: it loads, exports every entry point the C
header declares, names its backend, and never computes on the host (CPU)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import PRODUCT_SO, ROOT
from solver_cpp_amd.capi import _SIGNATURES, Library, ptr

HEADER = os.path.join(ROOT, "include", "solver_cpp_amd.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:sc|f)_[A-Za-z0-9_]+)\s*\(", text)))


def _exported(path):
    lib = C.CDLL(path, mode=C.RTLD_LOCAL)
    return {n for n in _declared() if hasattr(lib, n)}


def test_header_and_bindings_agree():
    from solver_cpp_amd.capi import BOUNDARY_SIGNATURES, PROFILE_SIGNATURES
    declared = set(_declared())
    bound = set(_SIGNATURES) | set(BOUNDARY_SIGNATURES) | set(PROFILE_SIGNATURES)
    assert bound == declared, (sorted(bound - declared), sorted(declared - bound))


def test_product_exports_every_declared_symbol():
    if not os.path.exists(PRODUCT_SO):
        pytest.fail(f"{PRODUCT_SO} is missing: python -m solver_cpp_amd.build")
    missing = sorted(set(_declared()) - _exported(PRODUCT_SO))
    assert not missing, missing


def test_oracle_exports_the_compute_symbols(oracle):
    missing = sorted(set(_SIGNATURES) - _exported(oracle.path))
    assert not missing, missing


def test_backend_names(oracle):
    assert oracle.backend == "cpu-oracle"
    assert Library(PRODUCT_SO).backend == "hip-gfx950"


def test_product_has_no_host_fallback():
    """Without a GPU every computing entry point fails with a message instead
    of computing on the host.  (On a GPU machine the call succeeds: skip.)"""
    lib = Library(PRODUCT_SO)
    cells = np.array([4, 2, 2], dtype=np.int32)
    x = np.random.default_rng(0).standard_normal(16)
    out = np.zeros(2 * 3 * 2 * 2)
    rc = lib.sc_fft_r2c(ptr(cells, C.POINTER(C.c_int)), 1, ptr(x), ptr(out))
    if rc == 0:
        pytest.skip("a HIP device is present")
    assert lib.sc_last_error().decode()
