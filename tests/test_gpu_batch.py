"""Batched runs (harm_batch) through the C-ABI: every run of a batch must be
bit-identical to the same context stepped alone, and to the oracle."""
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


def torus(n1, n2, a, rmax, beta, seed):
    g = H.grid_mks(n1, n2, a=a, Rout=40.0, hslope=0.3)
    geom = H.Geometry(g)
    p0 = H.init_fishmon(geom, rin=6.0, rmax=rmax, beta=beta, randfact=4e-2, seed=seed)
    prm = H.make_params(n1, n2, g.dx1, g.dx2, lim=H.PARA, fluxmethod=H.HLLFLUX, timeorder=4, cour=0.9)
    return geom, p0, prm


def rk4_problems():
    """Three RK4 runs of different size, metric and parameters."""
    out = [torus(32, 64, 0.9375, 12.0, 100.0, 7), torus(48, 40, 0.5, 14.0, 50.0, 3)]
    g, geom, p0, prm = problem("mink", 32, 32, lim=H.MC)
    out.append((geom, p0, prm))
    return out


def start(geom, p0, prm, dt0):
    h = H.Harm(prm, geom)
    h.set_array(H.P, p0)
    h.bound_prim(H.P)
    h.set_time(0.0, dt0, 1e30)
    return h


DT0 = [1e-3, 1e-3, 1e-4]


@pytest.mark.parametrize("fv", [H.FLUX_FACE_64x4, H.FLUX_FACE_64x2])
@pytest.mark.parametrize("bv", [H.BOUND_KERNELS, H.BOUND_BLOCK])
def test_batch_runs_equal_separate_runs(fv, bv):
    say(f"[case] batch of 3 RK4 runs (torus 32x64, torus 48x40 a=0.5, OT 32x32), flux={fv} bound={bv}, 20 steps")
    probs = rk4_problems()
    batched = [start(*pp, dt) for pp, dt in zip(probs, DT0)]
    alone = [start(*pp, dt) for pp, dt in zip(probs, DT0)]
    b = H.Batch(batched)
    b.set_tuning(fv, bv)
    assert b.get_tuning() == (fv, bv)
    b.step_n(5)
    b.step_n(15)
    for k, (hb, ha) in enumerate(zip(batched, alone)):
        ha.step_n(20)
        assert hb.get_time() == ha.get_time(), f"run {k}: t, dt"
        assert hb.nstroke() == ha.nstroke(), f"run {k}: Newton iterations"
        assert same(hb.get_array(H.P), ha.get_array(H.P)), f"run {k}: state"
        assert hb.failures() == ha.failures()
    b.close()


def test_batch_matches_oracle_and_autotunes():
    say("[case] autotuned batch of 3 RK4 runs vs the oracle, 12 steps")
    probs = rk4_problems()
    batched = [start(*pp, dt) for pp, dt in zip(probs, DT0)]
    b = H.Batch(batched)
    b.step_n(12)
    fv, bv = b.get_tuning()
    say(f"  chosen flux={fv} bound={bv}")
    for k, ((geom, p0, prm), dt0) in enumerate(zip(probs, DT0)):
        o = Oracle(H.params_dict(prm))
        o.set_geometry(*geom.tables())
        o.set_array(0, p0)
        o.bound_prim(0)
        o.set_time(0.0, dt0, 1e30)
        for _ in range(12):
            o.step()
        assert batched[k].get_time() == o.get_time(), f"run {k}: t, dt"
        assert same(batched[k].get_array(H.P), o.get_array(0)), f"run {k}: state"
        assert batched[k].nstroke() == o.nstroke()
    b.close()


def test_batch_of_force_free_runs():
    say("[case] batch of 2 force-free monopoles (32x64, 48x64), RK2, 30 steps")
    probs = []
    for n1, n2 in ((32, 64), (48, 64)):
        g, geom, p0, prm = problem("ffde", n1, n2, lim=H.MC, fluxmethod=H.LAXFFLUX, timeorder=2)
        probs.append((geom, p0, prm))
    batched = [start(*pp, 1e-3) for pp in probs]
    alone = [start(*pp, 1e-3) for pp in probs]
    b = H.Batch(batched)
    b.step_n(30)
    for k, (hb, ha) in enumerate(zip(batched, alone)):
        ha.step_n(30)
        assert hb.get_time() == ha.get_time()
        assert same(hb.get_array(H.P), ha.get_array(H.P)), f"ffde run {k}"
    b.close()


def test_batch_rejects_mixed_time_orders():
    say("[case] batch argument checks")
    g, geom, p0, prm = problem("mink", 16, 16, lim=H.MC, timeorder=4)
    g2, geom2, p02, prm2 = problem("mink", 16, 16, lim=H.MC, timeorder=2)
    a, c = start(geom, p0, prm, 1e-4), start(geom2, p02, prm2, 1e-4)
    with pytest.raises(H.HarmError, match="TIMEORDER"):
        H.Batch([a, c])


def test_config4_force_free_spin_batch_512():
    """The force-free workload at full size: 16 split monopoles on MKS 512x512
    (Rout 100 M, hslope 0.1) with spins evenly spaced in [0.1, 0.99], MC,
    LAXF, midpoint RK2 (initbase.c:151-157), advanced as one batch for 3 steps.
    Every run is bit-identical to its context stepped alone from the same
    state, and the a=0.99 run to the oracle."""
    from concurrent.futures import ThreadPoolExecutor
    n, nsteps, dt0 = 512, 3, 1e-3
    spins = [float(x) for x in np.linspace(0.1, 0.99, 16)]
    say(f"[case] force-free batch: 16 spins in [0.1, 0.99], {n}x{n}, MC+LAXF+RK2, {nsteps} steps")

    def build(a):
        g = H.grid_mks(n, n, a=a, Rout=100.0, hslope=0.1)
        geom = H.Geometry(g)
        p0 = H.init_monopole(geom)
        prm = H.make_params(n, n, g.dx1, g.dx2, lim=H.MC, fluxmethod=H.LAXFFLUX, timeorder=2, eomtype=H.EOMFFDE,
                            poledeath=0)
        return geom, p0, prm

    runs, alone, keep = [], [], None
    with ThreadPoolExecutor(max_workers=8) as ex:
        for k, (geom, p0, prm) in enumerate(ex.map(build, spins)):
            h = start(geom, p0, prm, dt0)
            h.step_n(nsteps)
            alone.append((h.get_array(H.P), h.get_time(), h.failures()))
            h.set_array(H.P, p0)
            h.bound_prim(H.P)
            h.set_time(0.0, dt0, 1e30)
            runs.append(h)
            if k == len(spins) - 1:
                keep = (geom, p0, prm)
    say("  each run stepped alone; now the batch")
    b = H.Batch(runs)
    b.step_n(nsteps)
    for k, h in enumerate(runs):
        P, t, fl = alone[k]
        assert h.get_time() == t, f"spin {spins[k]}: t, dt"
        assert same(h.get_array(H.P), P), f"spin {spins[k]}: state"
        assert h.failures() == fl and fl[1] == 0
    b.close()
    geom, p0, prm = keep
    o = Oracle(H.params_dict(prm))
    o.set_geometry(*geom.tables())
    o.set_array(0, p0)
    o.bound_prim(0)
    o.set_time(0.0, dt0, 1e30)
    for _ in range(nsteps):
        o.step()
    assert runs[-1].get_time() == o.get_time()
    assert same(runs[-1].get_array(H.P), o.get_array(0)), "a=0.99 run vs the oracle"
    for h in runs:
        h.close()


def test_config2_sixty_four_run_draw():
    """The batched sweep's own 64 runs (seed 2024: grids 64x64 .. 512x256,
    2.81 M zones, spins in [0, 0.99]) advanced as one batch for 3 steps:
    every run bit-identical to its context stepped alone from the same
    state, and the two smallest runs to the oracle."""
    from concurrent.futures import ThreadPoolExecutor
    import bench
    draws = bench.batch_draws(64, 2024)
    nsteps, dt0 = 3, 1e-5
    say(f"[case] the batched sweep's 64-run draw (seed 2024), {sum(d['n1'] * d['n2'] for d in draws)} zones, "
        f"{nsteps} steps")
    small = sorted(range(64), key=lambda k: draws[k]["n1"] * draws[k]["n2"])[:2]
    runs, alone, keep = [], [], {}
    with ThreadPoolExecutor(max_workers=8) as ex:
        for k, (geom, p0, prm) in enumerate(ex.map(bench.build_torus, draws)):
            h = H.Harm(prm, geom)
            h.set_array(H.P, p0)
            h.bound_prim(H.P)
            h.set_time(0.0, dt0, 2000.0)
            h.step_n(nsteps)
            alone.append((h.get_array(H.P), h.get_time(), h.nstroke(), h.failures()))
            h.set_array(H.P, p0)
            h.bound_prim(H.P)
            h.set_time(0.0, dt0, 2000.0)
            runs.append(h)
            if k in small:
                keep[k] = (geom, p0, prm)
    say("  each run stepped alone; now the batch")
    b = H.Batch(runs)
    b.step_n(nsteps)
    for k, h in enumerate(runs):
        P, t, ns, fl = alone[k]
        g = f"run {k} ({draws[k]['n1']}x{draws[k]['n2']}, a={draws[k]['a']:.3f})"
        assert h.get_time() == t, g + ": t, dt"
        assert h.nstroke() - ns == ns, g + ": Newton iterations"
        assert same(h.get_array(H.P), P), g + ": state"
        assert h.failures()[1] == fl[1] == 0
    b.close()
    for k in small:
        geom, p0, prm = keep[k]
        o = Oracle(H.params_dict(prm))
        o.set_geometry(*geom.tables())
        o.set_array(0, p0)
        o.bound_prim(0)
        o.set_time(0.0, dt0, 2000.0)
        for _ in range(nsteps):
            o.step()
        assert runs[k].get_time() == o.get_time(), f"run {k}: t, dt vs the oracle"
        assert same(runs[k].get_array(H.P), o.get_array(0)), f"run {k}: state vs the oracle"
    for h in runs:
        h.close()
