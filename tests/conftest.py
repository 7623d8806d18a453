import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
PRODUCT_SO = os.path.join(ROOT, "solver_cpp_amd", "_build", "libsolver_cpp_amd.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle (test infrastructure), built on demand."""
    if not os.path.exists(ORACLE_SO):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "-j8"])
    from solver_cpp_amd.capi import Library
    return Library(ORACLE_SO)


@pytest.fixture(scope="session")
def product():
    """The HIP library.  A missing build is an error, never a skip."""
    if not os.path.exists(PRODUCT_SO):
        raise FileNotFoundError(f"{PRODUCT_SO} is missing: python -m solver_cpp_amd.build")
    from solver_cpp_amd.capi import Library
    return Library(PRODUCT_SO)


@pytest.fixture(scope="session", params=["oracle", pytest.param("product", marks=pytest.mark.gpu)])
def lib(request):
    """Both implementations behind the same C ABI (the HIP one needs a GPU)."""
    return request.getfixturevalue(request.param)


@pytest.fixture(scope="session")
def goldens():
    with open(os.path.join(GOLDEN, "reference_goldens.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def trace():
    return np.load(os.path.join(GOLDEN, "fortran_trace.npz"))


def arr(g):
    """Golden tensor (Fortran order) -> flat float64 numpy array."""
    if "re" in g:
        return np.array(g["re"]) + 1j * np.array(g["im"])
    return np.array(g["data"], dtype=np.float64)


def cplx_interleaved(a):
    a = np.asarray(a, dtype=np.complex128)
    out = np.empty(2 * a.size)
    out[0::2] = a.real
    out[1::2] = a.imag
    return out
