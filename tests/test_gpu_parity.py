"""Parity of the HIP path (through the C-ABI) with the oracle and with the
reference's golden vectors.  The GPU step is required to be bit-identical
(IEEE equality) to the CPU step: the library is built with -ffp-contract=off
and keeps the reference's operation order.  The one exception is CSSLOPE,
whose reference arithmetic is x87 long double (interp.c:202); there the bar
is 1e-12 relative.
"""
import sys

import numpy as np
import pytest

import harm_ramt_amd as H
from oracle.oracle import Oracle
from tests.helpers import golden_cases, golden_problem, load_golden, problem, same, sha

pytestmark = pytest.mark.gpu
NG = 3


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_ctx(geom, prm, p0, dt0):
    h = H.Harm(prm, geom)
    h.set_array(H.P, p0)
    h.bound_prim(H.P)
    h.set_time(0.0, dt0, 1e30)
    return h


def oracle_ctx(geom, prm, p0, dt0):
    o = Oracle(H.params_dict(prm))
    o.set_geometry(*geom.tables())
    if prm.extrap == 2:
        o.set_radius(geom.radius)
    o.set_array(0, p0)
    o.bound_prim(0)
    o.set_time(0.0, dt0, 1e30)
    return o


def assert_close_or_same(a, b, exact, what):
    if exact:
        bad = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b))))
        assert len(bad) == 0, f"{what}: {len(bad)} values differ, first {bad[:3].tolist()}"
    else:
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-14 * np.abs(b).max(), err_msg=what)


@pytest.mark.parametrize("name", golden_cases())
def test_gpu_step_matches_reference(name):
    gd = load_golden(name)
    kind, n1, n2 = str(gd["kind"]), int(gd["n1"]), int(gd["n2"])
    say(f"[case] {name} {kind} {n1}x{n2} lim={int(gd['lim'])} steps={int(gd['nsteps'])}")
    g, geom, p0, prm = golden_problem(gd)
    assert sha(p0) == str(gd["input_sha"])
    exact = int(gd["lim"]) != H.CSSLOPE
    h = gpu_ctx(geom, prm, p0, float(gd["dt0"]))
    saved = set(int(s) for s in gd["saved_steps"])
    ns_prev = 0
    for s in range(1, int(gd["nsteps"]) + 1):
        h.step_ch()
        t, dt = h.get_time()
        ns = h.nstroke()
        if exact:
            assert t == gd["t"][s - 1] and dt == gd["dt"][s - 1], f"step {s}: t,dt {t},{dt}"
            assert ns - ns_prev == int(gd["nstroke"][s - 1]), f"step {s}: Newton iterations"
        ns_prev = ns
        if s in saved:
            assert_close_or_same(h.get_array(H.P), gd[f"p_{s}"], exact, f"{name} step {s}")
    assert h.failures()[1] == 0


def _same_fluxes(h, o, n1, n2, what):
    """F1 over fluxcalc's x1 faces (i in [0,N1], j in [-1,N2]) and F2 over its
    x2 faces (i in [-1,N1], j in [0,N2]), GPU against oracle"""
    F1g, F2g, F1o, F2o = h.get_array(H.F1), h.get_array(H.F2), o.get_array(3), o.get_array(4)
    assert same(F1g[NG:NG + n1 + 1, NG - 1:NG + n2 + 1], F1o[NG:NG + n1 + 1, NG - 1:NG + n2 + 1]), what + ": F1"
    assert same(F2g[NG - 1:NG + n1 + 1, NG:NG + n2 + 1], F2o[NG - 1:NG + n1 + 1, NG:NG + n2 + 1]), what + ": F2"


@pytest.mark.parametrize("kind,lim,flux", [("ks", H.PARA, H.HLLFLUX), ("ks", H.MC, H.LAXFFLUX),
                                           ("mink", H.MC, H.HLLFLUX)])
def test_gpu_entry_points_match_oracle(kind, lim, flux):
    say(f"[case] entry points {kind} lim={lim} flux={flux}")
    n1, n2 = 16, 32
    g, geom, p0, prm = problem(kind, n1, n2, lim=lim, fluxmethod=flux)
    h = gpu_ctx(geom, prm, p0, 1e-3)
    o = oracle_ctx(geom, prm, p0, 1e-3)
    assert same(h.get_array(H.P), o.get_array(0)), "bound_prim"
    # fluxcalc: the fluxes before fix_flux, then fix_flux as a call of its own
    nd1, nd2 = h.fluxcalc(H.P, 1), h.fluxcalc(H.P, 2)
    od1, od2 = o.fluxcalc(0, 1), o.fluxcalc(0, 2)
    assert nd1 == od1 and nd2 == od2, "Courant limits"
    _same_fluxes(h, o, n1, n2, "fluxcalc")
    if prm.fixupflux:
        h.fix_flux()
        o.fix_flux()
        _same_fluxes(h, o, n1, n2, "fix_flux")
    # flux_ct
    h.flux_ct()
    o.flux_ct()
    assert same(h.get_array(H.F1)[NG:NG + n1 + 1, NG:NG + n2], o.get_array(3)[NG:NG + n1 + 1, NG:NG + n2])
    assert same(h.get_array(H.F2)[NG:NG + n1, NG:NG + n2 + 1], o.get_array(4)[NG:NG + n1, NG:NG + n2 + 1])
    # advance: one RK4 substage (pi=p, pb=p, pf=pk0) then a second (pb=pk0, pf=pk1)
    cu, cn = [1.0, 0.0, 0.5], [1.0, 1.0 / 6.0, 0.0]
    ng, no = h.advance(H.P, H.P, cu, H.PK0, cn, 0, 4), o.advance(0, 0, cu, 1, cn, 0, 4)
    assert ng == no
    for sel in (H.PK0, H.ULAST, H.UNEW):
        assert same(h.get_array(sel)[NG:-NG, NG:-NG], o.get_array(sel)[NG:-NG, NG:-NG]), f"advance array {sel}"
    h.bound_prim(H.PK0)
    o.bound_prim(1)
    assert same(h.get_array(H.PK0), o.get_array(1)), "bound_prim(pk0)"
    cu, cn = [1.0, 0.0, 0.5], [0.0, 1.0 / 3.0, 0.0]
    assert h.advance(H.P, H.PK0, cu, H.PK1, cn, 1, 4) == o.advance(0, 1, cu, 2, cn, 1, 4)
    for sel in (H.PK1, H.ULAST, H.UNEW):
        assert same(h.get_array(sel)[NG:-NG, NG:-NG], o.get_array(sel)[NG:-NG, NG:-NG]), f"advance array {sel}"


@pytest.mark.parametrize("kind", ["ks", "mink"])
def test_gpu_fix_flux_and_flux_ct_read_the_held_fluxes(kind):
    """fix_flux() and flux_ct() act on the fluxes the context holds: fluxes
    edited through harm_set_array between fluxcalc and fix_flux, and between
    fix_flux and flux_ct, give the oracle's results for the same edits."""
    say(f"[case] fix_flux/flux_ct on edited fluxes, {kind}")
    n1, n2 = 16, 32
    g, geom, p0, prm = problem(kind, n1, n2)
    h = gpu_ctx(geom, prm, p0, 1e-3)
    o = oracle_ctx(geom, prm, p0, 1e-3)
    for d in (1, 2):
        assert h.fluxcalc(H.P, d) == o.fluxcalc(0, d)
    rng = np.random.default_rng(5)
    for sel in (H.F1, H.F2):
        F = o.get_array(sel)
        F *= 1.0 + 0.25 * rng.standard_normal(F.shape)
        F[..., 0] = rng.standard_normal(F.shape[:2])   # both signs of the mass flux at the outflow faces
        h.set_array(sel, F)
        o.set_array(sel, F)
    h.fix_flux()
    o.fix_flux()
    _same_fluxes(h, o, n1, n2, "fix_flux of edited fluxes")
    for sel in (H.F1, H.F2):
        F = o.get_array(sel)
        F[..., 5:8] *= 1.0 + 0.5 * rng.standard_normal(F[..., 5:8].shape)
        h.set_array(sel, F)
        o.set_array(sel, F)
    h.flux_ct()
    o.flux_ct()
    _same_fluxes(h, o, n1, n2, "flux_ct of edited fluxes")


def test_gpu_utoprimgen_matches_oracle():
    say("[case] utoprimgen batch 16x16 torus geometry, 4096 random states")
    g, geom, p0, prm = problem("ks", 16, 16)
    h = H.Harm(prm, geom)
    o = oracle_ctx(geom, prm, p0, 1e-3)
    rng = np.random.default_rng(11)
    n = 4096
    zones = rng.integers(-2, 18, size=(n, 2)).astype(np.int32)
    U = np.zeros((n, 8))
    guess = np.zeros((n, 8))
    for t in range(n):
        i, j = int(zones[t, 0]), int(zones[t, 1])
        rho = 10 ** rng.uniform(-4, 0)
        pr = np.array([rho, rho * 10 ** rng.uniform(-3, 1), *rng.normal(0, 1, 3), *(rng.normal(0, 0.3, 3) * np.sqrt(rho))])
        U[t] = o.primtoU(i, j, 3, pr)
        guess[t] = pr * (1 + rng.normal(0, 0.05, 8))
    out_g, fl_g = h.utoprimgen(zones, U, guess)
    fails = 0
    for t in range(n):
        out_o, fl_o, _ = o.utoprimgen(int(zones[t, 0]), int(zones[t, 1]), U[t], guess[t])
        assert fl_g[t] == fl_o, f"pflag zone {t}"
        assert same(out_g[t], out_o), f"state {t}"
        fails += fl_o != 0
    say(f"  inversions flagged: {fails} of {n} (both paths agree)")


def test_gpu_ffde_utoprimgen_matches_oracle():
    say("[case] force-free utoprimgen batch on the 16x32 monopole geometry, 4096 random states")
    g, geom, p0, prm = problem("ffde", 16, 32, lim=H.MC, fluxmethod=H.LAXFFLUX, timeorder=2)
    h = H.Harm(prm, geom)
    o = oracle_ctx(geom, prm, p0, 1e-3)
    rng = np.random.default_rng(12)
    n = 4096
    zones = np.stack([rng.integers(-2, 18, n), rng.integers(-2, 34, n)], axis=1).astype(np.int32)
    U = np.zeros((n, 8))
    for t in range(n):
        pr = np.array([0.0, 0.0, *rng.normal(0, 2.0, 3), *rng.normal(0, 1.0, 3)])
        U[t] = o.primtoU(int(zones[t, 0]), int(zones[t, 1]), 3, pr)
    guess = np.zeros((n, 8))
    out_g, fl_g = h.utoprimgen(zones, U, guess)
    for t in range(n):
        out_o, fl_o, _ = o.utoprimgen(int(zones[t, 0]), int(zones[t, 1]), U[t], guess[t])
        assert fl_g[t] == fl_o, f"flag zone {t}"
        assert same(out_g[t], out_o), f"state {t}"


def test_gpu_graph_replay_equals_single_steps():
    """harm_step_n replays graphs of 8 steps and of one step: 12 and then
    19 steps (two 8-step replays and three single ones) equal the same
    steps launched one harm_step at a time"""
    say("[case] harm_step_n (HIP graph) vs harm_step, 64x64 torus, 12 + 19 steps")
    g, geom, p0, prm = problem("ks", 64, 64)
    a = gpu_ctx(geom, prm, p0, 1e-3)
    b = gpu_ctx(geom, prm, p0, 1e-3)
    for _ in range(12):
        a.step_ch()
    b.step_n(12)
    assert same(a.get_array(H.P), b.get_array(H.P))
    assert a.get_time() == b.get_time()
    for _ in range(19):
        a.step_ch()
    b.step_n(19)
    assert same(a.get_array(H.P), b.get_array(H.P))
    assert a.get_time() == b.get_time() and a.nstroke() == b.nstroke()


@pytest.mark.parametrize("kind,n", [("ks", 256), ("mink", 256), ("ffde", 512)])
def test_gpu_fullsize_matches_oracle(kind, n):
    say(f"[case] full size {kind} {n}x{n}, 6 steps bitwise vs oracle")
    if kind == "ffde":   # the reference's force-free defaults (initbase.c:151-157)
        g, geom, p0, prm = problem(kind, n, n, lim=H.MC, fluxmethod=H.LAXFFLUX, timeorder=2)
    else:
        g, geom, p0, prm = problem(kind, n, n, lim=H.PARA if kind == "ks" else H.MC)
    dt0 = 1e-4 if kind == "mink" else 1e-3
    h = gpu_ctx(geom, prm, p0, dt0)
    o = oracle_ctx(geom, prm, p0, dt0)
    for s in range(6):
        h.step_ch()
        o.step()
    assert h.get_time() == o.get_time()
    assert same(h.get_array(H.P), o.get_array(0))
    assert h.nstroke() == o.nstroke()


def test_gpu_largest_grid_matches_oracle():
    """a 1024x1024 torus (1.06 M zones with ghosts, 9.5 M fixup-event slots
    when accounting) for 2 steps through the autotuned graph path, bitwise
    against the oracle: index arithmetic at four times the benchmark's size"""
    n = 1024
    say(f"[case] torus {n}x{n}, 2 steps via harm_step_n vs oracle")
    g, geom, p0, prm = problem("ks", n, n)
    h = gpu_ctx(geom, prm, p0, 1e-3)
    o = oracle_ctx(geom, prm, p0, 1e-3)
    h.step_n(2)
    for _ in range(2):
        o.step()
    assert h.get_time() == o.get_time()
    assert same(h.get_array(H.P), o.get_array(0))
    assert h.nstroke() == o.nstroke()


@pytest.mark.parametrize("kind,eno", [("ks", 1), ("ks", 2), ("mink", 1), ("mink", 2)])
def test_gpu_eno_graph_matches_oracle(kind, eno):
    """DOENO through the autotuned graph path at 128x128, bitwise vs the oracle."""
    say(f"[case] DOENO={eno} {kind} 128x128, 6 steps via harm_step_n vs oracle")
    g, geom, p0, prm = problem(kind, 128, 128, lim=H.PARA if kind == "ks" else H.MC, eno=eno)
    dt0 = 1e-4 if kind == "mink" else 1e-3
    h = gpu_ctx(geom, prm, p0, dt0)
    o = oracle_ctx(geom, prm, p0, dt0)
    h.step_n(6)
    for _ in range(6):
        o.step()
    assert h.get_time() == o.get_time()
    assert same(h.get_array(H.P), o.get_array(0))
    assert h.nstroke() == o.nstroke()


def test_gpu_fullsize_ot_invariants():
    say("[case] OT 256x256 on the GPU: 30 steps, conservation and corner div B")
    n = 256
    g, geom, p0, prm = problem("mink", n, n, lim=H.MC)
    h = gpu_ctx(geom, prm, p0, 1e-4)

    def divb(p):
        B1, B2 = p[NG:-NG + 1, NG:-NG + 1, 5], p[NG:-NG + 1, NG:-NG + 1, 6]
        return (0.5 * (B1[1:, 1:] + B1[1:, :-1] - B1[:-1, 1:] - B1[:-1, :-1]) / g.dx1 +
                0.5 * (B2[1:, 1:] + B2[:-1, 1:] - B2[1:, :-1] - B2[:-1, :-1]) / g.dx2)

    d0 = divb(h.get_array(H.P))
    h.step_n(30)
    p = h.get_array(H.P)
    assert np.all(np.isfinite(p))
    assert np.max(np.abs(divb(p) - d0)) < 1e-11 * np.abs(p[..., 5:7]).max() / g.dx1
    n_fail, failed = h.failures()
    assert failed == 0


@pytest.mark.parametrize("fv", [H.FLUX_FACE_64x4, H.FLUX_FACE_64x2, H.FLUX_PAIR, H.FLUX_SHARED])
@pytest.mark.parametrize("bv", [H.BOUND_KERNELS, H.BOUND_BLOCK, H.BOUND_FIXONE, H.BOUND_FUSED, H.BOUND_INUPDATE])
def test_gpu_variants_match_oracle(fv, bv):
    say(f"[case] kernel variants flux={fv} bound={bv}: 32x64 torus, 30 steps (with repaired inversions)")
    g, geom, p0, prm = problem("ks", 32, 64)
    h = gpu_ctx(geom, prm, p0, 1e-3)
    h.set_tuning(fv, bv)
    assert h.get_tuning() == (fv, bv)
    o = oracle_ctx(geom, prm, p0, 1e-3)
    for _ in range(10):
        h.step_ch()
    h.step_n(20)
    for _ in range(30):
        o.step()
    assert h.get_time() == o.get_time()
    assert same(h.get_array(H.P), o.get_array(0))
    assert h.nstroke() == o.nstroke()
    nfail, failed = h.failures()
    assert failed == o.failed() == 0
    assert nfail > 0, "the run is meant to exercise the fixup of failed inversions"


@pytest.mark.parametrize("lim", [H.DONOR, H.VANL, H.MINM, H.MC, H.PARA, H.CSSLOPE, H.NLIMCENT, H.NLIMUP, H.NLIMDOWN])
def test_gpu_shared_flux_every_limiter(lim):
    """The zone-shared reconstruction variant with every limiter, on the torus
    (outflow, polar axis, fix_flux) and on OT (periodic), against the oracle."""
    for kind, n1, n2, dt0 in (("ks", 32, 64, 1e-3), ("mink", 40, 72, 1e-4)):
        say(f"[case] FLUX_SHARED lim={lim} {kind} {n1}x{n2}, 5 steps vs oracle")
        g, geom, p0, prm = problem(kind, n1, n2, lim=lim)
        h = gpu_ctx(geom, prm, p0, dt0)
        h.set_tuning(H.FLUX_SHARED, H.BOUND_FUSED)
        o = oracle_ctx(geom, prm, p0, dt0)
        exact = lim != H.CSSLOPE
        for s in range(5):
            h.step_ch()
            o.step()
            if exact:
                assert h.get_time() == o.get_time(), f"{kind} step {s + 1}"
        assert_close_or_same(h.get_array(H.P), o.get_array(0), exact, f"{kind} lim {lim}")


@pytest.mark.parametrize("case", ["mink", "ffde", "ks_x1periodic_5", "ks_x1periodic_12", "ks_nopoledeath"])
@pytest.mark.parametrize("bv", [H.BOUND_FUSED, H.BOUND_FIXONE, H.BOUND_INUPDATE])
def test_gpu_boundary_variants_other_boundaries(case, bv):
    """The one-launch ghost fill on the other boundary combinations: periodic
    both ways, polar without POLEDEATH, and periodic x1 with polar x2 and
    POLEDEATH (a pole row's x1 ghosts read the far side), including a
    5-zone x1 extent where the two sides' source zones overlap."""
    say(f"[case] boundary variant {bv} on {case}, 8 steps bitwise vs oracle")
    if case == "mink":
        g, geom, p0, prm = problem("mink", 32, 32, lim=H.MC)
    elif case == "ffde":
        g, geom, p0, prm = problem("ffde", 16, 32, lim=H.MC, fluxmethod=H.LAXFFLUX, timeorder=2)
    elif case.startswith("ks_x1periodic"):
        n1 = int(case.rsplit("_", 1)[1])
        g, geom, p0, prm = problem("ks", n1, 32)
        prm.bc[H.X1UP] = prm.bc[H.X1DN] = H.PERIODIC
    else:
        g, geom, p0, prm = problem("ks", 32, 64)
        prm.poledeath = 0
    dt0 = 1e-4 if case == "mink" else 1e-3
    h = gpu_ctx(geom, prm, p0, dt0)
    h.set_tuning(H.FLUX_FACE_64x4, bv)
    o = oracle_ctx(geom, prm, p0, dt0)
    for s in range(8):
        h.step_ch()
        o.step()
        assert h.get_time() == o.get_time(), f"step {s + 1}"
        assert same(h.get_array(H.P), o.get_array(0)), f"step {s + 1}"
    assert h.nstroke() == o.nstroke()


def strong_field_at_edges(p, n1, n2):
    """B raised by ten times the local sqrt(rho) in the three zones next to
    both polar axes (B2) and next to the inner radial edge (B1): inversions
    then fail on the pole rows and the edge columns, so the repair and the
    second bound_prim() act on zones that ghost zones are copied from."""
    s = slice(NG, NG + n1)
    for js in (slice(NG, 2 * NG), slice(n2, n2 + NG)):
        p[s, js, 6] += 10.0 * np.sqrt(p[s, js, 0])
    p[NG:2 * NG, NG:NG + n2, 5] += 10.0 * np.sqrt(p[NG:2 * NG, NG:NG + n2, 0])


def bits(a):
    return np.ascontiguousarray(a).view(np.int64)


@pytest.mark.parametrize("case", ["poledeath", "nopoledeath", "x1periodic_12", "x1periodic_5", "outflow_5",
                                  "outflow_4"])
def test_gpu_fill_in_update_with_failures_at_the_edges(case):
    """HARM_BOUND_INUPDATE (the first bound_prim() written by the update's
    edge zones, the second one as the images of the repaired zones and of
    the pole rows' edge zones) against the oracle and, bit for bit with the
    signs of zeros, against the one-launch fills, on a torus whose
    inversions fail on the pole rows and at the inner edge.  N1 = 5: the up
    edge's workgroup rows beyond the grid write its ghost columns; N1 = 4:
    both edges in one workgroup, each edge zone writes its own columns and
    the repair pass recomputes the pole rows' columns itself."""
    n1 = int(case.rsplit("_", 1)[1]) if case.startswith(("x1periodic", "outflow")) else 32
    n2 = 64
    say(f"[case] fill in the update, {case}: torus {n1}x{n2} with failures at the edges, 6 steps")
    g, geom, p0, prm = problem("ks", n1, n2)
    if case == "nopoledeath":
        prm.poledeath = 0
    if case.startswith("x1periodic"):
        prm.bc[H.X1UP] = prm.bc[H.X1DN] = H.PERIODIC
    p0 = p0.copy()
    strong_field_at_edges(p0, n1, n2)
    a = gpu_ctx(geom, prm, p0, 1e-3)
    a.set_tuning(H.FLUX_SHARED, H.BOUND_INUPDATE)
    b = gpu_ctx(geom, prm, p0, 1e-3)
    b.set_tuning(H.FLUX_SHARED, H.BOUND_FUSED)
    o = oracle_ctx(geom, prm, p0, 1e-3)
    pole = edge = 0
    for s in range(6):
        a.step_ch()
        b.step_ch()
        o.step()
        assert a.get_time() == o.get_time() == b.get_time(), f"step {s + 1}"
        pa = a.get_array(H.P)
        assert same(pa, o.get_array(0)), f"step {s + 1}: state vs oracle"
        assert np.array_equal(bits(pa), bits(b.get_array(H.P))), f"step {s + 1}: bits vs one-launch fills"
        fl = o.pflag()[NG:NG + n1, NG:NG + n2, 0]
        assert np.array_equal(a.get_pflag()[NG:NG + n1, NG:NG + n2, 0], fl), f"step {s + 1}: pflag"
        pole += int(np.count_nonzero(fl[:, [0, n2 - 1]]))
        edge += int(np.count_nonzero(fl[[0, n1 - 1], :]))
    assert a.nstroke() == o.nstroke() == b.nstroke()
    nfail, failed = a.failures()
    assert failed == o.failed() == 0
    say(f"  flagged zones over the run: pole rows {pole}, x1 edge columns {edge}")
    assert pole > 0 and edge > 0, "the run is meant to repair pole-row and edge zones"


def fixed_x1_torus(n1, n2):
    """A torus with FIXED x1 sides filled from panalytic (bounds.c; panalytic
    = the initial state, omegaf = 0) whose field is raised next to both x1
    edges (in the polar quarters, away from bound_disk's equator rows) and
    both poles, so that inversions fail on the edge columns."""
    g, geom, p0, prm = problem("ks", n1, n2)
    prm.bc[H.X1DN] = prm.bc[H.X1UP] = H.FIXED
    prm.boundsfile = H.BOUNDS_DISK
    prm.disk_first = 0
    p0 = p0.copy()
    s = slice(NG, NG + n1)
    for js in (slice(NG, 2 * NG), slice(n2, n2 + NG)):
        p0[s, js, 6] += 10.0 * np.sqrt(p0[s, js, 0])
    for i0 in (NG, n1):
        for js in (slice(NG, NG + n2 // 4), slice(NG + 3 * n2 // 4, NG + n2)):
            p0[i0:i0 + NG, js, 5] += 10.0 * np.sqrt(p0[i0:i0 + NG, js, 0])
    return g, geom, p0, prm


def test_gpu_fixed_x1_sides_repair_counts_ghost_flags_as_failed():
    """bound_pflag does nothing on a FIXED side (boundsint.c:34), so those
    ghosts keep set_arrays()'s initial flag -100 (set_arrays.c:86) and the
    repair of an edge zone counts them as failed neighbours, not as copies
    of the edge zone's flag.  Bitwise against the oracle over 6 steps, with
    failures on both edge columns."""
    n1, n2 = 32, 64
    say(f"[case] FIXED x1 sides: torus {n1}x{n2} with failures at the edges, 6 steps")
    g, geom, p0, prm = fixed_x1_torus(n1, n2)
    om = np.zeros_like(p0)
    h = H.Harm(prm, geom)
    h.set_analytic(p0, om)
    h.set_array(H.P, p0)
    h.bound_prim(H.P)
    h.set_time(0.0, 1e-3, 1e30)
    o = Oracle(H.params_dict(prm))
    o.set_geometry(*geom.tables())
    o.set_analytic(p0, om)
    o.set_array(0, p0)
    o.bound_prim(0)
    o.set_time(0.0, 1e-3, 1e30)
    edge = 0
    for s in range(6):
        h.step_ch()
        o.step()
        assert h.get_time() == o.get_time(), f"step {s + 1}"
        assert same(h.get_array(H.P), o.get_array(0)), f"step {s + 1}: state vs oracle"
        fl = o.pflag()[NG:NG + n1, NG:NG + n2, 0]
        assert np.array_equal(h.get_pflag()[NG:NG + n1, NG:NG + n2, 0], fl), f"step {s + 1}: pflag"
        edge += int(np.count_nonzero(fl[[0, n1 - 1], :]))
    assert h.nstroke() == o.nstroke()
    say(f"  flagged zones on the x1 edge columns over the run: {edge}")
    assert edge > 0, "the run is meant to repair edge zones"


def test_gpu_newton_spread_counts():
    """harm_get_newton_spread: its iteration total is harm_get_nstroke's, and
    the wavefront maxima bound it from above (64 lanes a wavefront)"""
    say("[case] Newton spread counters on the 64x64 torus, 5 steps")
    g, geom, p0, prm = problem("ks", 64, 64)
    h = gpu_ctx(geom, prm, p0, 1e-3)
    h.step_n(5)
    ns, wm = h.newton_spread()
    assert ns == h.nstroke() > 0
    assert 64 * wm >= ns and wm > 0


def test_gpu_autotune_picks_a_variant():
    say("[case] autotuner on the 64x64 torus")
    g, geom, p0, prm = problem("ks", 64, 64)
    h = gpu_ctx(geom, prm, p0, 1e-3)
    o = oracle_ctx(geom, prm, p0, 1e-3)
    h.step_n(3)
    fv, bv = h.get_tuning()
    say(f"  chosen flux={fv} bound={bv}")
    assert fv in (H.FLUX_FACE_64x4, H.FLUX_FACE_64x2, H.FLUX_PAIR, H.FLUX_SHARED)
    assert bv in (H.BOUND_KERNELS, H.BOUND_BLOCK, H.BOUND_FIXONE, H.BOUND_FUSED, H.BOUND_INUPDATE)
    for _ in range(3):
        o.step()
    assert same(h.get_array(H.P), o.get_array(0))
    assert h.get_time() == o.get_time()


@pytest.mark.parametrize("fv", [H.FLUX_FACE_64x4, H.FLUX_PAIR, H.FLUX_SHARED])
@pytest.mark.parametrize("bv", [H.BOUND_KERNELS, H.BOUND_KERNELS_SEQCHECK, H.BOUND_FIXONE, H.BOUND_FUSED,
                                H.BOUND_INUPDATE])
def test_gpu_checksolution_matches_oracle(fv, bv):
    say(f"[case] CHECKSOLUTION: field-threaded torus 32x64, 25 steps, flux={fv} bound={bv}")
    g, geom, p0, prm = problem("ks_mono", 32, 64)
    assert prm.checksolution == 1
    h = gpu_ctx(geom, prm, p0, 1e-3)
    h.set_tuning(fv, bv)
    o = oracle_ctx(geom, prm, p0, 1e-3)
    for s in range(25):
        h.step_ch()
        o.step()
        assert h.get_time() == o.get_time(), f"step {s + 1}"
    assert same(h.get_array(H.P), o.get_array(0))
    assert h.nstroke() == o.nstroke()
    pf_g = h.get_pflag()[3:-3, 3:-3]
    pf_o = o.pflag()[3:-3, 3:-3]
    assert np.array_equal(pf_g[..., 0], pf_o[..., 0]), "FLAGUTOPRIMFAIL"
    assert np.array_equal(pf_g[..., 2:4], pf_o[..., 2:4]), "b^2 flags"
    assert (pf_g[..., 2:4] > 0).sum() > 100, "the problem is meant to flag the magnetised atmosphere"


def test_gpu_checksolution_autotuned_graph():
    say("[case] CHECKSOLUTION with the autotuner and graph replay, 40 steps")
    g, geom, p0, prm = problem("ks_mono", 32, 64)
    h = gpu_ctx(geom, prm, p0, 1e-3)
    o = oracle_ctx(geom, prm, p0, 1e-3)
    h.step_n(40)
    for _ in range(40):
        o.step()
    assert h.get_tuning()[1] != H.BOUND_BLOCK   # the one-workgroup pass has no fixup_checksolution
    assert same(h.get_array(H.P), o.get_array(0))
    assert h.get_time() == o.get_time()


def test_gpu_checksolution_autotuned_region_grows():
    """The autotuner's trial steps leave the persistent gamma table as they
    found it.  With a 5e-3 split-monopole field the b^2-flagged region gains
    zones in the third step (oracle: 342 flagged at the first step, 2 more at
    the third), so trial steps write gamma values where the first real step
    must still see the table's earlier contents.  The autotuned run must equal
    a run forced to the same variants, and the oracle, bit for bit."""
    say("[case] CHECKSOLUTION, flagged region growing, autotuned vs forced vs oracle, 12 steps")
    g, geom, p0, prm = problem("ks", 32, 64)
    p0 = p0.copy()
    p0[..., 5] += H.init_monopole(geom, 5e-3)[..., 5]
    prm = H.make_params(32, 64, g.dx1, g.dx2, lim=H.PARA, fluxmethod=H.HLLFLUX, timeorder=4, checksolution=1)
    a = gpu_ctx(geom, prm, p0, 1e-3)
    a.step_n(12)
    fv, bv = a.get_tuning()
    b = gpu_ctx(geom, prm, p0, 1e-3)
    b.set_tuning(fv, bv)
    b.step_n(12)
    o = oracle_ctx(geom, prm, p0, 1e-3)
    for _ in range(12):
        o.step()
    assert same(a.get_array(H.P), b.get_array(H.P))
    assert same(a.get_array(H.P), o.get_array(0))
    assert a.get_time() == o.get_time()
    assert np.array_equal(a.get_pflag()[3:-3, 3:-3, 0], o.pflag()[3:-3, 3:-3, 0])
