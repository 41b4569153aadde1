import ctypes
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session", autouse=True)
def _built_oracle():
    """The oracle library is test infrastructure: build it if it is missing."""
    lib = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(lib):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "-s"])
    yield


SHIM_SRC = os.path.join(ROOT, "tests", "adjoint_shim.cpp")
SHIM_LIB = os.path.join(ROOT, "tests", "_shim", "adjoint_shim.so")


def build_adjoint_shim():
    """Host-only build of the adjoint's local derivative functions (test
    infrastructure; __graft_entry__.build() makes it too)."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    os.makedirs(os.path.dirname(SHIM_LIB), exist_ok=True)
    tmp = SHIM_LIB + ".tmp"
    subprocess.check_call([hipcc, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", "-ffp-contract=off",
                           "-shared", "-fPIC", SHIM_SRC, "-o", tmp])
    os.replace(tmp, SHIM_LIB)


@pytest.fixture(scope="session")
def shim():
    if not os.path.exists(SHIM_LIB) or os.path.getmtime(SHIM_LIB) < max(
            os.path.getmtime(SHIM_SRC), os.path.getmtime(os.path.join(ROOT, "harm_ramt_amd", "csrc", "harm_adjoint.h")),
            os.path.getmtime(os.path.join(ROOT, "harm_ramt_amd", "csrc", "harm_device.h")),
            os.path.getmtime(os.path.join(ROOT, "harm_ramt_amd", "csrc", "harm_eno.h"))):
        build_adjoint_shim()
    lib = ctypes.CDLL(SHIM_LIB)
    dp = ctypes.POINTER(ctypes.c_double)
    lib.shim_bsq.restype = ctypes.c_double
    lib.shim_slope3.restype = ctypes.c_double
    lib.shim_slope3.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    lib.shim_slope3_coef.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, dp]
    lib.shim_eno_line.restype = ctypes.c_double
    lib.shim_eno_line.argtypes = [ctypes.c_int, dp, ctypes.c_int, ctypes.c_int]
    lib.shim_eno_weight.restype = ctypes.c_double
    lib.shim_eno_weight.argtypes = [ctypes.c_int] * 4
    return lib
