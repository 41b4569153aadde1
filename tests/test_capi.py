"""The C-ABI library: builds, loads without a GPU, exports every entry point
include/harm_amd.h declares, and its host-side generators work on the CPU."""
import ctypes
import os
import re

import numpy as np
import pytest

import harm_ramt_amd as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "harm_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(harm_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = H.load_library()
    declared = header_functions()
    assert len(declared) >= 25
    missing = [f for f in declared if not hasattr(lib, f)]
    assert not missing, missing
    assert sorted(H.EXPORTS) == declared


def test_no_cpu_fallback_when_library_missing(tmp_path):
    with pytest.raises(RuntimeError):
        H._lib_saved = H._lib
        H._lib = None
        try:
            H.load_library(str(tmp_path / "nope.so"))
        finally:
            H._lib = H._lib_saved


def test_parameter_errors_are_reported_before_any_gpu_call():
    lib = H.load_library()
    prm = H.make_params(2, 64, 0.1, 0.1)          # n1 too small
    h = ctypes.c_void_p()
    assert lib.harm_create(ctypes.byref(prm), ctypes.byref(h)) != 0
    assert b"n1,n2" in lib.harm_last_error()
    prm = H.make_params(16, 16, 0.1, 0.1, fluxb=2)   # FLUXCD not on the GPU path
    assert lib.harm_create(ctypes.byref(prm), ctypes.byref(h)) != 0
    assert b"FLUXB" in lib.harm_last_error()
    prm = H.make_params(16, 16, 0.1, 0.1, bc=(H.PERIODIC, H.OUTFLOW, H.POLARAXIS, H.POLARAXIS))
    assert lib.harm_create(ctypes.byref(prm), ctypes.byref(h)) != 0
    assert b"periodic" in lib.harm_last_error()
    # bounds.c rules: an even N2 (bound_disk's equator between two rows)
    prm = H.make_params(16, 15, 0.1, 0.1, boundsfile=H.BOUNDS_DISK, bc=(H.FIXED,) * 4)
    assert lib.harm_create(ctypes.byref(prm), ctypes.byref(h)) != 0
    assert b"even n2" in lib.harm_last_error()
    prm = H.make_params(16, 16, 0.1, 0.1, extrap=3)
    assert lib.harm_create(ctypes.byref(prm), ctypes.byref(h)) != 0
    assert b"extrap" in lib.harm_last_error()
    prm = H.make_params(16, 16, 0.1, 0.1, extrap=2, rescaletype=2)
    assert lib.harm_create(ctypes.byref(prm), ctypes.byref(h)) != 0
    assert b"rescaletype" in lib.harm_last_error()
    # bounds.fishmon.c has no FIXED branch
    prm = H.make_params(16, 16, 0.1, 0.1, bc=(H.FIXED, H.FIXED, H.POLARAXIS, H.POLARAXIS))
    assert lib.harm_create(ctypes.byref(prm), ctypes.byref(h)) != 0
    assert b"FIXED sides need HARM_BOUNDS_DISK" in lib.harm_last_error()


def test_host_generators_run_without_gpu():
    g = H.grid_mks(16, 32)
    assert g.dx2 == 1.0 / 32 and g.Rin < 1.0 + np.sqrt(1.0 - g.a ** 2)   # inner edge inside the horizon
    geom = H.Geometry(g)
    assert np.all(geom.gdet[..., 3] > 0)
    # gcon is the inverse of gcov at every zone and position
    eye = np.einsum("...ij,...jk->...ik", geom.gcov, geom.gcon)
    assert np.max(np.abs(eye - np.eye(4))) < 1e-10
    p = H.init_fishmon(geom)
    assert np.all(np.isfinite(p)) and p[3:-3, 3:-3, 0].max() == 1.0
    go = H.Geometry(H.grid_cartesian(8, 8))
    po = H.init_orszag_tang(go)
    assert np.allclose(po[..., 0], 25.0 / (36.0 * np.pi))
