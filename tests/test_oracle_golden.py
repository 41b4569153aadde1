"""The oracle against the golden vectors produced by the compiled reference
(tests/golden/make_golden.py): bit-identical states, dt sequence and Newton
iteration counts."""
import numpy as np
import pytest

import harm_ramt_amd as H
from oracle.oracle import Oracle
from tests.helpers import disk_cases, disk_problem, golden_cases, golden_problem, load_golden, problem, same, sha


@pytest.mark.parametrize("name", golden_cases())
def test_oracle_matches_reference(name):
    gd = load_golden(name)
    kind, n1, n2 = str(gd["kind"]), int(gd["n1"]), int(gd["n2"])
    g, geom, p0, prm = golden_problem(gd)
    assert sha(p0) == str(gd["input_sha"]), "input state generator changed"
    assert sha(np.concatenate([a.ravel() for a in geom.tables()])) == str(gd["geom_sha"])
    o = Oracle(H.params_dict(prm))
    o.set_geometry(*geom.tables())
    if prm.extrap == 2:
        o.set_radius(geom.radius)
    o.set_array(0, p0)
    o.bound_prim(0)
    o.set_time(0.0, float(gd["dt0"]), 1e30)
    saved = set(int(s) for s in gd["saved_steps"])
    for s in range(1, int(gd["nsteps"]) + 1):
        ns0 = o.nstroke()
        o.step()
        t, dt = o.get_time()
        assert t == gd["t"][s - 1] and dt == gd["dt"][s - 1], f"step {s} t/dt"
        assert o.nstroke() - ns0 == int(gd["nstroke"][s - 1]), f"step {s} Newton iterations"
        if s in saved:
            assert same(o.get_array(0), gd[f"p_{s}"]), f"step {s} state differs from the reference"


def test_geometry_matches_reference_set_grid():
    gd = np.load(__import__("tests.helpers", fromlist=["GOLDEN"]).GOLDEN + "/ks_geometry_32x64.npz")
    g = H.grid_mks(32, 64)
    geom = H.Geometry(g)
    allt = np.concatenate([geom.gcov.ravel(), geom.gcon.ravel(), geom.gdet.ravel(), geom.conn.ravel()])
    assert g.Rin == float(gd["Rin"])
    assert g.startx1 == gd["startx"][0] and g.startx2 == gd["startx"][1]
    assert g.dx1 == gd["dx"][0] and g.dx2 == gd["dx"][1]
    assert same(allt[gd["sample_idx"]], gd["sample"])
    import hashlib
    assert hashlib.sha256(allt.tobytes()).hexdigest() == str(gd["sha"])


def bits(a):
    return np.ascontiguousarray(a).view(np.int64)


@pytest.mark.parametrize("name", disk_cases())
def test_oracle_matches_reference_default_problem(name):
    """The reference's default problem (init.c PROBLEMTYPE 7 on the
    cylindrical-Minkowski grid, bounds.c: FIXED from panalytic and
    bound_disk, force-free): the oracle stepped from the state the
    reference's own init() set up equals its steps bit for bit, signs of
    zeros included, with bound_disk's static-field branch (t <= 15) and its
    outflow branch (t > 15)."""
    gd = load_golden(name)
    g, geom, p0, pan, om, prm = disk_problem(gd)
    o = Oracle(H.params_dict(prm))
    o.set_geometry(*geom.tables())
    o.set_analytic(pan, om)
    o.set_array(0, p0)
    o.set_time(float(gd["t0"]), float(gd["dt0"]), 1e30)
    saved = set(int(s) for s in gd["saved_steps"])
    for s in range(1, int(gd["nsteps"]) + 1):
        o.step()
        t, dt = o.get_time()
        assert t == gd["t"][s - 1] and dt == gd["dt"][s - 1], f"step {s} t/dt"
        if s in saved:
            assert np.array_equal(bits(o.get_array(0)), bits(gd[f"p_{s}"])), f"step {s} state"


def test_cylmink_geometry_matches_reference_set_grid():
    """harm_grid_cylmink (defcoord 666, numerical Jacobian, numerical
    inverse and connection) against the reference's set_grid() tables of
    its default problem, by sha256 of all four tables"""
    import hashlib
    for name in disk_cases():
        gd = load_golden(name)
        g, geom, *_ = disk_problem(gd)
        allt = np.concatenate([a.ravel() for a in geom.tables()[:4]])
        assert hashlib.sha256(allt.tobytes()).hexdigest() == str(gd["tables_sha"]), name


def test_extrap1_outflow_against_closed_form():
    """EXTRAP 1 on the outflow faces of a state whose ghosts the inflow
    check and limit_gamma leave alone (static fluid): densities and B1
    scaled by gdet(edge)/gdet(ghost), U2, U3, B2, B3 by (1 + (i - ri) dx1)
    below and (1 - (i - ri) dx1) above, U1 by (1 - (i - ri) dx1) below and
    (1 - 2 (i - ri) dx1) above"""
    n1, n2 = 16, 32
    g = H.grid_mks(n1, n2)
    geom = H.Geometry(g)
    prm = H.make_params(n1, n2, g.dx1, g.dx2, extrap=1, poledeath=0)  # POLEDEATH would zero the pole rows' B2
    rng = np.random.default_rng(5)
    p = np.zeros((n1 + 6, n2 + 6, 8))
    p[..., 0] = 1.0 + rng.random(p.shape[:2])
    p[..., 1] = 0.5 + rng.random(p.shape[:2])
    p[..., 5:8] = 1e-3 * rng.standard_normal(p.shape[:2] + (3,))
    o = Oracle(H.params_dict(prm))
    o.set_geometry(*geom.tables())
    o.set_array(0, p)
    o.bound_prim(0)
    q = o.get_array(0)
    gd = geom.gdet.reshape(n1 + 6, n2 + 6, 4)[..., 3]
    for side, ri, ghosts in ((0, 0, (-1, -2, -3)), (1, n1 - 1, (n1, n1 + 1, n1 + 2))):
        for i in ghosts:
            f = gd[ri + 3, 3:-3] / gd[i + 3, 3:-3]
            src = p[ri + 3, 3:-3]
            exp = np.empty_like(src)
            exp[:, 0:2] = src[:, 0:2] * f[:, None]
            exp[:, 5] = src[:, 5] * f
            lin = (1. - (i - ri) * g.dx1) if side else (1. + (i - ri) * g.dx1)
            exp[:, 2] = src[:, 2] * ((1. - 2 * (i - ri) * g.dx1) if side else (1. - (i - ri) * g.dx1))
            exp[:, [3, 4, 6, 7]] = src[:, [3, 4, 6, 7]] * lin
            # densities and field: the closed form (the velocities, zero
            # here, are what the inflow check then sets at the outer face)
            keep = [0, 1, 5, 6, 7]
            np.testing.assert_allclose(q[i + 3, 3:-3][:, keep], exp[:, keep], rtol=1e-15, atol=0)
