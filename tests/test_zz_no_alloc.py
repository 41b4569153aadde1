"""harm_utoprimgen_noalloc (staging in the library's fixed device array,
no allocation) against harm_utoprimgen (per-call device allocations), through
the C-ABI: identical solutions and flags on two calls in a row, the second
with fewer zones, and a clear error beyond the staging capacity."""
import ctypes
import sys

import numpy as np
import pytest

import harm_ramt_amd as H
from oracle.oracle import Oracle
from tests.helpers import problem, same

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _states(o, n, n1, n2, seed):
    rng = np.random.default_rng(seed)
    zones = np.stack([rng.integers(-3, n1 + 3, n), rng.integers(-3, n2 + 3, n)], axis=1).astype(np.int32)
    U = np.zeros((n, 8))
    guess = np.zeros((n, 8))
    for t in range(n):
        rho = 10 ** rng.uniform(-4, 0)
        pr = np.array([rho, rho * 10 ** rng.uniform(-3, 1), *rng.normal(0, 1, 3),
                       *(rng.normal(0, 0.3, 3) * np.sqrt(rho))])
        U[t] = o.primtoU(int(zones[t, 0]), int(zones[t, 1]), 3, pr)
        guess[t] = pr * (1 + rng.normal(0, 0.05, 8))
    return zones, U, guess


def _call(h, name, zones, U, guess):
    zones = np.ascontiguousarray(zones, dtype=np.int32)
    U = np.ascontiguousarray(U, dtype=np.float64)
    out = np.ascontiguousarray(guess, dtype=np.float64).copy()
    fl = np.full(len(zones), -7, dtype=np.int32)
    rc = getattr(h.lib, name)(h.h, len(zones), zones.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                              U.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                              fl.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return rc, out, fl


def test_noalloc_inversions_equal_the_allocating_entry_point():
    say("[case] harm_utoprimgen_noalloc vs harm_utoprimgen: 3000 then 700 zones, 24x32 torus geometry")
    n1, n2 = 24, 32
    g, geom, p0, prm = problem("ks", n1, n2)
    h = H.Harm(prm, geom)
    o = Oracle(H.params_dict(prm))
    o.set_geometry(*geom.tables())
    for n, seed in ((3000, 21), (700, 22)):
        zones, U, guess = _states(o, n, n1, n2, seed)
        rc_a, out_a, fl_a = _call(h, "harm_utoprimgen", zones, U, guess)
        rc_b, out_b, fl_b = _call(h, "harm_utoprimgen_noalloc", zones, U, guess)
        assert rc_a == 0 and rc_b == 0, h.lib.harm_last_error()
        assert np.array_equal(fl_a, fl_b), f"{n} zones: flags"
        assert same(out_a, out_b), f"{n} zones: states"
        assert (fl_b == -7).sum() == 0
        say(f"  {n} zones: {int((fl_b != 0).sum())} flagged inversions, identical")
    h.close()


def test_noalloc_refuses_more_zones_than_it_stages():
    say("[case] harm_utoprimgen_noalloc beyond its capacity")
    g, geom, p0, prm = problem("ks", 16, 16)
    h = H.Harm(prm, geom)
    n = 16385
    assert n > 16384
    zones = np.zeros((n, 2), dtype=np.int32)
    U = np.zeros((n, 8))
    guess = np.ones((n, 8))
    rc, out, fl = _call(h, "harm_utoprimgen_noalloc", zones, U, guess)
    assert rc != 0 and b"staging capacity" in h.lib.harm_last_error()
    assert np.all(out == 1.0) and np.all(fl == -7)   # nothing written
    h.close()
