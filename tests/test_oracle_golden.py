"""The oracle against the golden vectors produced by the compiled reference
(tests/golden/make_golden.py): bit-identical states, dt sequence and Newton
iteration counts."""
import numpy as np
import pytest

import harm_ramt_amd as H
from oracle.oracle import Oracle
from tests.helpers import golden_cases, golden_problem, load_golden, problem, same, sha


@pytest.mark.parametrize("name", golden_cases())
def test_oracle_matches_reference(name):
    gd = load_golden(name)
    kind, n1, n2 = str(gd["kind"]), int(gd["n1"]), int(gd["n2"])
    g, geom, p0, prm = golden_problem(gd)
    assert sha(p0) == str(gd["input_sha"]), "input state generator changed"
    assert sha(np.concatenate([a.ravel() for a in geom.tables()])) == str(gd["geom_sha"])
    o = Oracle(H.params_dict(prm))
    o.set_geometry(*geom.tables())
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
