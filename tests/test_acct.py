"""Conservation accounting and dump-time reductions against the compiled
reference (tests/golden/diag_*.npz: the reference's accumulators after its
steps, written by its own diag_flux/diag_source/diag_fixup, and its
divbmaxavg()/constotal() of the final state).

- The energy regions of setflux()/setjetflux() from the host generator are
  the reference's.
- The oracle's accumulators, div B reduction and totals are the reference's
  to the bit (every accumulator, including the floor/failure ones).
- Closed forms: constotal of a uniform state at rest in flat space is
  rho dVF N, the energy total -rho-u per zone; a field with B1 a function of
  x2 only and B2 of x1 only has zero corner divergence.
- The cases cover floors and repaired inversion failures (torus) and
  limit_gamma on face and zone states (Orszag-Tang with GAMMAMAX 1.1),
  CHECKSOLUTION (b^2 flags and checksolution repairs), DOENO 1 and 2, and
  force-free, whose every inversion books its dissipation
  (phys.ffde.c:572-607).  Force-free fixup1zone() books COUNTFLOORACT from
  locals the reference never sets under EOMFFDE (fixup.c:283-313); the
  force-free reference case is the reference built with those locals zero
  (oracle/ref, ffdezi), which the oracle reproduces in full with pr0 = 0;
  with its own choice (pr0 = the incoming state) every accumulator but
  that category and fladd is the reference's.
- The device path (-m gpu): the same accumulators bit for bit on the
  reference's cases, at a larger grid against the oracle, with the state of
  the steps unchanged by the accounting, and the dump-time reductions of the
  final state.
"""
import numpy as np
import pytest

import harm_ramt_amd as H
from oracle.oracle import Oracle
from tests.helpers import golden_problem, load_golden, problem, same, sha

CASES = ["diag_torus_para_hll_rk4", "diag_ot_mc_hll_rk2", "diag_ot_gammamax_mc_hll_rk2",
         "diag_torus_field_checksolution_rk4", "diag_torus_eno_u_para_hll_rk4", "diag_ot_eno_du_mc_hll_rk2"]
FFDE_CASE = "diag_ffde_monopole_mc_laxf_rk2"
STEP_KEYS = ["pdot", "pdotterms", "pcum", "sourceadd", "sourceaddterms", "pdottermsjet2", "frdot", "lumvsr"]


def _oracle_run(gd, nsteps=None, pr0_zero=False):
    g, geom, p0, prm = golden_problem(gd)
    assert sha(p0) == str(gd["input_sha"])
    o = Oracle(H.params_dict(prm))
    o.set_geometry(*geom.tables())
    o.set_array(0, p0)
    o.bound_prim(0)
    o.set_time(0.0, float(gd["dt0"]), 1e30)
    o.acct_setup(gd["enerpos"], gd["doflux"], 2 * np.pi)
    if pr0_zero:
        o.acct_ffde_pr0_zero()
    for _ in range(int(gd["nsteps"]) if nsteps is None else nsteps):
        o.step()
    return o


@pytest.mark.parametrize("name", CASES + [FFDE_CASE])
def test_regions_match_reference(name):
    gd = load_golden(name)
    g, geom, p0, prm = golden_problem(gd)
    ep, df = H.ener_regions(g, float(gd["h_over_r_jet"]))
    assert np.array_equal(ep, gd["enerpos"]) and np.array_equal(df, gd["doflux"])


def test_regions_reject_edge_on_grid_end():
    g = H.grid_mks(16, 32)
    with pytest.raises(H.HarmError):  # jet edge theta = pi: the last x2 face
        H.ener_regions(g, -np.pi / 2)


@pytest.mark.parametrize("name", CASES)
def test_oracle_accounting_matches_reference(name):
    gd = load_golden(name)
    o = _oracle_run(gd)
    assert same(o.get_array(0), gd["p"])
    a = o.acct()
    assert a["dVF"] == gd["dVF"]
    for k in STEP_KEYS + ["fladd", "fladdterms"]:
        assert same(a[k], gd[k]), k
    mx, av, im, jm = o.divbmaxavg()
    assert mx == gd["divbmax"] and av == gd["divbavg"]
    for r in range(3):
        assert same(o.constotal(r), gd["constotal"][r]), r


def test_oracle_force_free_accounting_matches_reference():
    """every accumulator of the force-free reference case, with fixup1zone's
    events measured from pr0 = 0 as the zero-initialised reference books them"""
    gd = load_golden(FFDE_CASE)
    assert int(gd["zero_init_locals"]) == 1
    o = _oracle_run(gd, pr0_zero=True)
    assert same(o.get_array(0), gd["p"])
    a = o.acct()
    for k in STEP_KEYS + ["fladd", "fladdterms"]:
        assert same(a[k], gd[k]), k
    mx, av, im, jm = o.divbmaxavg()
    assert mx == gd["divbmax"] and av == gd["divbavg"]
    for r in range(3):
        assert same(o.constotal(r), gd["constotal"][r]), r


def test_oracle_force_free_accounting_own_choice():
    """with pr0 = the incoming state only fixup1zone's COUNTFLOORACT events
    differ: the dissipation account (categories 0 and 1 of the inversions),
    limit_gamma's and every flux and source accumulator are the reference's;
    the jets' regions hold no fixup1zone event and match in full"""
    gd = load_golden(FFDE_CASE)
    o = _oracle_run(gd)
    a = o.acct()
    for k in STEP_KEYS:
        assert same(a[k], gd[k]), k
    ft, want = a["fladdterms"], gd["fladdterms"]
    for r in range(3):
        for cat in range(ft.shape[1]):
            if (r, cat) != (0, 1):
                assert same(ft[r][cat], want[r][cat]), (r, cat)
    assert same(a["fladd"][1:], gd["fladd"][1:])
    assert not same(ft[0][1], want[0][1])  # limit_gamma acted in region 0
    # the dissipation account: energy and momenta, never the field
    assert np.count_nonzero(want[0][0][1:5]) and not np.count_nonzero(want[0][0][5:])
    assert np.count_nonzero(want[1][1][1:5])


def test_reference_cases_exercise_the_accumulators():
    """the fixtures are not trivially zero where the tests rely on them"""
    t, o = load_golden(CASES[0]), load_golden(CASES[1])
    assert np.count_nonzero(t["fladdterms"][0][1]) and np.count_nonzero(t["fladdterms"][0][0])  # floors, failures
    assert np.count_nonzero(t["sourceadd"][0]) and np.count_nonzero(t["lumvsr"])
    assert np.count_nonzero(o["pdottermsjet2"]) and np.count_nonzero(t["pcum"][1])
    assert t["divbmax"] > 0 and o["divbmax"] > 0
    gm = load_golden(CASES[2])
    assert all(np.count_nonzero(gm["fladdterms"][r][2]) for r in range(3))  # limit_gamma events


def test_oracle_constotal_uniform_flat_state():
    n1, n2 = 8, 12
    g, geom, p0, prm = problem("mink", n1, n2)
    rho, u = 0.7, 0.3
    p = np.zeros_like(p0)
    p[..., 0], p[..., 1] = rho, u
    o = Oracle(H.params_dict(prm))
    o.set_geometry(*geom.tables())
    o.set_array(0, p)
    ep = np.array([[n1 - 1, 0, n2 - 1, 0], [3, 2, 5, 4], [-100] * 4], dtype=np.int32)
    df = np.array([[n1, 0, n2, 0], [-1] * 4, [-1] * 4], dtype=np.int32)
    o.acct_setup(ep, df, 2 * np.pi)
    dVF = (g.dx1 * g.dx2) * 2 * np.pi
    v0, v1, v2 = o.constotal(0), o.constotal(1), o.constotal(2)
    # at rest in Minkowski: rho u^t = rho, T^t_t = -(rho + u) (rest mass kept), no momentum, no field
    assert np.allclose(v0[:2], [rho * dVF * n1 * n2, -(rho + u) * dVF * n1 * n2], rtol=1e-13, atol=0)
    assert np.all(v0[2:] == 0)
    assert np.allclose(v1[:2], [rho * dVF * 4, -(rho + u) * dVF * 4], rtol=1e-13, atol=0)
    assert np.all(v2 == 0)


def test_oracle_divb_zero_for_separable_field():
    n1, n2 = 16, 16
    g, geom, p0, prm = problem("mink", n1, n2)
    p = p0.copy()
    x = np.arange(-3, n1 + 3)[:, None] * np.ones((1, n2 + 6))
    y = np.ones((n1 + 6, 1)) * np.arange(-3, n2 + 3)[None, :]
    p[..., 5] = 1.0 + 0.3 * np.sin(0.7 * y)   # B1(x2)
    p[..., 6] = 0.5 + 0.2 * np.cos(0.4 * x)   # B2(x1)
    o = Oracle(H.params_dict(prm))
    o.set_geometry(*geom.tables())
    o.set_array(0, p)
    mx, av, im, jm = o.divbmaxavg()
    assert mx < 1e-14 and av < 1e-14  # zero up to the rounding of the corner sums
    # one perturbed zone: the largest divergence is at one of its four corners
    p[10, 7, 5] += 1e-3
    o.set_array(0, p)
    mx, av, im, jm = o.divbmaxavg()
    assert mx > 1e-5 and (im - 7, jm - 4) in [(0, 0), (1, 0), (0, 1), (1, 1)]


# ---------------------------------------------------------------- GPU ----

def _gpu_run(gd, acct=True, nsteps=None, graph=False, tuning=None):
    g, geom, p0, prm = golden_problem(gd)
    c = H.Harm(prm, geom)
    if tuning is not None:
        c.set_tuning(*tuning)
    c.set_array(H.P, p0)
    c.bound_prim(H.P)
    c.set_time(0.0, float(gd["dt0"]), 1e30)
    if acct:
        c.acct_setup(gd["enerpos"], gd["doflux"], 2 * np.pi)
    n = int(gd["nsteps"]) if nsteps is None else nsteps
    if graph:
        c.step_n(n)
    else:
        for _ in range(n):
            c.step_ch()
    return c


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("graph", [False, True])
def test_gpu_accounting_matches_reference(name, graph):
    gd = load_golden(name)
    print(f"[case] {name} graph={graph}", flush=True)
    c = _gpu_run(gd, graph=graph)
    assert same(c.get_array(H.P), gd["p"]), "accounting changed the step"
    a = c.acct()
    for k in STEP_KEYS + ["fladd", "fladdterms"]:
        assert same(a[k], gd[k]), k
    mx, av, im, jm = c.divbmaxavg()
    assert mx == gd["divbmax"] and av == gd["divbavg"]
    for r in range(3):
        assert same(c.constotal(r), gd["constotal"][r]), r


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [False, True])
def test_gpu_force_free_accounting_matches_reference(graph):
    """force-free: the dissipation account of every inversion and the rest of
    the accumulators are the reference's; fixup1zone's COUNTFLOORACT events
    (region 0 only here) are the oracle's, measured from the incoming state"""
    gd = load_golden(FFDE_CASE)
    print(f"[case] {FFDE_CASE} graph={graph}", flush=True)
    c = _gpu_run(gd, graph=graph)
    assert same(c.get_array(H.P), gd["p"]), "accounting changed the step"
    a = c.acct()
    for k in STEP_KEYS:
        assert same(a[k], gd[k]), k
    ft, want = a["fladdterms"], gd["fladdterms"]
    for r in range(3):
        for cat in range(ft.shape[1]):
            if (r, cat) != (0, 1):
                assert same(ft[r][cat], want[r][cat]), (r, cat)
    assert same(a["fladd"][1:], gd["fladd"][1:])
    o = _oracle_run(gd).acct()
    assert same(a["fladd"], o["fladd"]) and same(ft, o["fladdterms"])


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [False, True])
def test_gpu_tree_sums_match_ordered_sums(graph):
    """HARM_ACCT_TREE on a 96x128 torus over 6 steps: the state and the
    serial accumulators (surface integrals, frdot, lumvsr) are the ordered
    mode's to the bit; the source and fixup totals agree to 1e-12 of each
    array's largest value; a second tree run gives the same bits"""
    n1, n2, nsteps = 96, 128, 6
    g, geom, p0, prm = problem("ks", n1, n2)
    ep, df = H.ener_regions(g, 0.4)
    out = []
    for order in (H.ACCT_ORDERED, H.ACCT_TREE, H.ACCT_TREE):
        print(f"[case] torus {n1}x{n2} accounting order {order} graph={graph}", flush=True)
        c = H.Harm(prm, geom)
        c.set_array(H.P, p0)
        c.bound_prim(H.P)
        c.set_time(0.0, 1e-3, 1e30)
        c.acct_setup(ep, df, 2 * np.pi)
        c.acct_set_order(order)
        if graph:
            c.step_n(nsteps)
        else:
            for _ in range(nsteps):
                c.step_ch()
        out.append((c.get_array(H.P), c.acct()))
        c.close()
    (p_o, a_o), (p_t, a_t), (p_t2, a_t2) = out
    assert same(p_o, p_t)
    for k in STEP_KEYS:
        if k not in ("sourceadd", "sourceaddterms"):
            assert same(a_t[k], a_o[k]), k
    for k in ("sourceadd", "sourceaddterms", "fladd", "fladdterms"):
        scale = np.abs(a_o[k]).max()
        assert scale > 0, k
        assert np.abs(a_t[k] - a_o[k]).max() <= 1e-12 * scale, k
        assert same(a_t2[k], a_t[k]), k


@pytest.mark.gpu
def test_gpu_accounting_matches_oracle_larger_grid():
    """a 96x128 torus over 6 steps, every flux and boundary variant pair in turn
    gives the oracle's accumulators"""
    n1, n2, nsteps = 96, 128, 6
    g, geom, p0, prm = problem("ks", n1, n2)
    ep, df = H.ener_regions(g, 0.4)
    o = Oracle(H.params_dict(prm))
    o.set_geometry(*geom.tables())
    o.set_array(0, p0)
    o.bound_prim(0)
    o.set_time(0.0, 1e-3, 1e30)
    o.acct_setup(ep, df, 2 * np.pi)
    for _ in range(nsteps):
        o.step()
    want = o.acct()
    wdiv = o.divbmaxavg()
    for tv in [(H.FLUX_FACE_64x4, H.BOUND_KERNELS), (H.FLUX_SHARED, H.BOUND_FUSED)]:
        print(f"[case] torus {n1}x{n2} tuning {tv}", flush=True)
        c = H.Harm(prm, geom)
        c.set_tuning(*tv)
        c.set_array(H.P, p0)
        c.bound_prim(H.P)
        c.set_time(0.0, 1e-3, 1e30)
        c.acct_setup(ep, df, 2 * np.pi)
        for _ in range(nsteps):
            c.step_ch()
        assert same(c.get_array(H.P), o.get_array(0))
        a = c.acct()
        for k in STEP_KEYS + ["fladd", "fladdterms"]:
            assert same(a[k], want[k]), k
        assert c.divbmaxavg() == wdiv
        for r in range(3):
            assert same(c.constotal(r), o.constotal(r)), r


@pytest.mark.gpu
def test_gpu_accounting_off_and_explicit_diag_flux():
    gd = load_golden(CASES[1])
    c = _gpu_run(gd, nsteps=3)
    a3 = c.acct()
    c.acct_off()
    c.step_ch()
    b = c.acct()
    for k in STEP_KEYS:
        assert same(a3[k], b[k]), k  # nothing accumulates while off
    with pytest.raises(H.HarmError):
        c.diag_flux(1e-4)
    # diag_flux on request: pdot of the current fluxes, pcum += pdot Dt
    c.acct_setup(gd["enerpos"], gd["doflux"], 2 * np.pi)
    c.diag_flux(0.5)
    d = c.acct()
    assert np.count_nonzero(d["pdot"][0]) and same(d["pcum"], d["pdot"] * 0.5)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,eno", [("ffde", 0), ("ks_mono", 0), ("ks", 1), ("mink", 2)])
def test_gpu_accounting_matches_oracle_other_modes(kind, eno):
    """force-free, CHECKSOLUTION (b^2 flags, checksolution repairs), DOENO 1
    and 2: the device accumulators are the oracle's"""
    n1, n2, nsteps = (32, 32, 8) if kind == "mink" else (32, 64, 8)
    print(f"[case] {kind} eno={eno}", flush=True)
    g, geom, p0, prm = problem(kind, n1, n2, eno=eno)
    ep, df = H.ener_regions(g, 0.4)
    dt0 = 1e-4 if kind == "mink" else 1e-3
    o = Oracle(H.params_dict(prm))
    o.set_geometry(*geom.tables())
    o.set_array(0, p0)
    o.bound_prim(0)
    o.set_time(0.0, dt0, 1e30)
    o.acct_setup(ep, df, 2 * np.pi)
    c = H.Harm(prm, geom)
    c.set_array(H.P, p0)
    c.bound_prim(H.P)
    c.set_time(0.0, dt0, 1e30)
    c.acct_setup(ep, df, 2 * np.pi)
    for _ in range(nsteps):
        o.step()
        c.step_ch()
    assert same(c.get_array(H.P), o.get_array(0))
    want, got = o.acct(), c.acct()
    for k in STEP_KEYS + ["fladd", "fladdterms"]:
        assert same(got[k], want[k]), k


@pytest.mark.gpu
def test_gpu_standalone_post_fixup_is_not_accounted():
    """a repair outside a step (fixup.c:37 called on its own) adds no fixup
    events to the next step's accumulators"""
    gd = load_golden(CASES[0])
    g, geom, p0, prm = golden_problem(gd)
    p = p0.copy()
    p[10:14, 20:24, 0] = 1e-20  # below the density floor: the repair changes these zones
    print("[case] standalone post_fixup with accounting", flush=True)
    runs = []
    for setup_first in (True, False):
        c = H.Harm(prm, geom)
        c.set_array(H.P, p)
        c.bound_prim(H.P)
        c.set_time(0.0, float(gd["dt0"]), 1e30)
        if setup_first:
            c.acct_setup(gd["enerpos"], gd["doflux"], 2 * np.pi)
        c.post_fixup(H.P)
        if not setup_first:
            c.acct_setup(gd["enerpos"], gd["doflux"], 2 * np.pi)
        c.step_ch()
        runs.append((c.get_array(H.P), c.acct()))
    (sa, aa), (sb, ab) = runs
    assert not same(sa, p)  # the repair and the step ran
    assert same(sa, sb)
    for k in STEP_KEYS + ["fladd", "fladdterms"]:
        assert same(aa[k], ab[k]), k
