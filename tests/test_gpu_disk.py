"""The reference's default problem through the C-ABI: init.c's Ramesh disk
(PROBLEMTYPE 7, force-free) on the cylindrical-Minkowski grid with bounds.c's
boundaries (every side FIXED from panalytic, bound_disk after every ghost
fill), stepped from the state, panalytic and omegafanalytic the reference's
own init() set up (tests/golden/disk_*.npz):

- every step's t and dt, and the states at the saved steps, equal the
  reference's bit for bit (signs of zeros included), in single steps and in
  graph-replayed steps, before and after t = 15 (bound_disk's switch from
  the analytic field to the outflowed one);
- every step equals the oracle's;
- the standalone bound_prim equals the oracle's (fills, then bound_disk's
  symmetrisation, field and velocity);
- the outflow extrapolation EXTRAP 1 on the torus is covered by the step
  goldens (torus_extrap1_para_hll_rk4) in test_gpu_parity.py.
"""
import sys

import numpy as np
import pytest

import harm_ramt_amd as H
from oracle.oracle import Oracle
from tests.helpers import disk_cases, disk_problem, load_golden

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a).view(np.int64)


def contexts(gd, tuning=None):
    g, geom, p0, pan, om, prm = disk_problem(gd)
    h = H.Harm(prm, geom)
    h.set_analytic(pan, om)
    h.set_array(H.P, p0)
    h.set_time(float(gd["t0"]), float(gd["dt0"]), 1e30)
    if tuning:
        h.set_tuning(*tuning)
    o = Oracle(H.params_dict(prm))
    o.set_geometry(*geom.tables())
    o.set_analytic(pan, om)
    o.set_array(0, p0)
    o.set_time(float(gd["t0"]), float(gd["dt0"]), 1e30)
    return h, o


@pytest.mark.parametrize("name", disk_cases())
def test_gpu_default_problem_matches_reference(name):
    gd = load_golden(name)
    n = int(gd["nsteps"])
    say(f"[case] {name}: Ramesh disk 32x64 CYLMINK, bounds.c, {n} steps from t0={float(gd['t0'])}")
    h, o = contexts(gd)
    saved = set(int(s) for s in gd["saved_steps"])
    for s in range(1, n + 1):
        if s <= n // 2:
            h.step_ch()
        else:
            h.step_n(1)
        o.step()
        t, dt = h.get_time()
        assert (t, dt) == (gd["t"][s - 1], gd["dt"][s - 1]), f"step {s}: t, dt"
        p = h.get_array(H.P)
        assert np.array_equal(bits(p), bits(o.get_array(0))), f"step {s}: state vs oracle"
        if s in saved:
            assert np.array_equal(bits(p), bits(gd[f"p_{s}"])), f"step {s}: state vs reference"
    nfail, failed = h.failures()
    assert failed == o.failed() == 0


@pytest.mark.parametrize("tuning", [(H.FLUX_FACE_64x4, H.BOUND_KERNELS), (H.FLUX_SHARED, H.BOUND_INUPDATE)])
def test_gpu_default_problem_every_schedule(tuning):
    """the general fills run whatever schedule variant is set (a graph of
    10 steps against the oracle)"""
    gd = load_golden("disk_default_mc_laxf_rk2")
    say(f"[case] Ramesh disk, variants {tuning}, 10 graph-replayed steps vs oracle")
    h, o = contexts(gd, tuning)
    h.step_n(10)
    for _ in range(10):
        o.step()
    assert h.get_time() == o.get_time()
    assert np.array_equal(bits(h.get_array(H.P)), bits(o.get_array(0)))


def test_gpu_default_problem_bound_prim():
    gd = load_golden("disk_default_mc_laxf_rk2")
    say("[case] Ramesh disk: standalone bound_prim (FIXED fills + bound_disk) vs oracle")
    h, o = contexts(gd)
    p = np.array(gd["p_10"])
    h.set_array(H.P, p)
    o.set_array(0, p)
    h.bound_prim(H.P)
    o.bound_prim(0)
    assert np.array_equal(bits(h.get_array(H.P)), bits(o.get_array(0)))
