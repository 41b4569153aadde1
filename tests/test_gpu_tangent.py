"""The tangent-linear (forward-mode) step through the C-ABI (harm_tl), on the
force-free split monopole (EOMFFDE, outflow + polar-axis boundaries,
FIXUPFLUX, flux-CT):

- its values are the plain step's to the bit (state, t, dt after every step);
- the derivative of the outgoing energy flux at r ~ 20 M with respect to the
  black-hole spin, from the tangents of the geometry tables and of the initial
  state (fourth-order differences of the host generators), agrees with the
  central finite difference of two full runs at a -+ h with regenerated grids
  and initial states and the base run's dt sequence imposed.
"""
import sys

import numpy as np
import pytest

import harm_ramt_amd as H
from tests.helpers import same

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def monopole(n1, n2, a, lim):
    g = H.grid_mks(n1, n2, a=a, Rout=100.0, hslope=0.1)
    geom = H.Geometry(g)
    p0 = H.init_monopole(geom)
    prm = H.make_params(n1, n2, g.dx1, g.dx2, lim=lim, fluxmethod=H.LAXFFLUX, timeorder=2, eomtype=H.EOMFFDE,
                        poledeath=0)
    return g, geom, p0, prm


def start(geom, p0, prm, dt0=1e-3):
    h = H.Harm(prm, geom)
    h.set_array(H.P, p0)
    h.bound_prim(H.P)
    h.set_time(0.0, dt0, 1e30)
    h.set_tuning(H.FLUX_FACE_64x2, H.BOUND_KERNELS)
    return h


def row_at_radius(g, r0):
    rs = [H.zone_coords(g, i, g.n2 // 2)[2] for i in range(g.n1)]
    return int(np.argmin([abs(r - r0) for r in rs]))


@pytest.mark.parametrize("lim", [H.MC, H.NLIMCENT])
def test_tangent_run_values_equal_the_plain_step(lim):
    n1, n2, ns = 32, 64, 20
    say(f"[case] tangent-linear run vs plain steps, monopole {n1}x{n2} lim={lim}, {ns} steps")
    g, geom, p0, prm = monopole(n1, n2, 0.9, lim)
    a = start(geom, p0, prm)
    b = start(geom, p0, prm)
    t = H.Tangent(b)
    rng = np.random.default_rng(3)
    t.set_state_tangent(rng.standard_normal(p0.shape))
    for s in range(ns):
        a.step_ch()
        t.step_n(1)
        assert a.get_time() == b.get_time(), f"step {s + 1}: t, dt"
        assert same(a.get_array(H.P), b.get_array(H.P)), f"step {s + 1}: state"
    assert np.all(np.isfinite(t.state_tangent()))
    t.close()


@pytest.mark.parametrize("lim,tol", [(H.NLIMCENT, 1e-7), (H.MC, 1e-4)])
def test_spin_derivative_of_the_energy_flux_matches_finite_differences(lim, tol):
    n1, n2, ns, a0 = 32, 64, 30, 0.9
    say(f"[case] dL/da (energy flux at r~20 M) tangent-linear vs central differences, {n1}x{n2} lim={lim}")
    g, geom, p0, prm = monopole(n1, n2, a0, lim)
    i20 = row_at_radius(g, 20.0)
    dgcov, dgcon, dgdet, dconn, dp0, ddx1, ddx2 = H.spin_tangents(
        lambda a: H.grid_mks(n1, n2, a=a, Rout=100.0, hslope=0.1), H.init_monopole, a0)
    base = start(geom, p0, prm)
    t = H.Tangent(base)
    t.set_geometry_tangent(dgcov, dgcon, dgdet, dconn)
    t.set_spacing_tangent(ddx1, ddx2)
    t.set_state_tangent(dp0)
    dts = []
    for _ in range(ns):
        dts.append(base.get_time()[1])
        t.step_n(1)
    L, dL = t.energy_flux(i20)

    def run(a):
        gg, ge, pp, pr = monopole(n1, n2, a, lim)
        h = start(ge, pp, pr)
        tt = 0.0
        for dt in dts:
            h.set_time(tt, dt, 1e30)
            h.step_ch()
            tt = h.get_time()[0]
        tz = H.Tangent(h)
        v = tz.energy_flux(i20)[0]
        tz.close()
        h.close()
        return v

    hstep = 1e-5
    fd = (run(a0 + hstep) - run(a0 - hstep)) / (2 * hstep)
    rel = abs(dL - fd) / abs(fd)
    say(f"  row {i20}: L = {L:.12e}, dL/da tangent {dL:.12e}, central difference {fd:.12e}, rel {rel:.2e}")
    assert np.isfinite(dL) and abs(fd) > 0
    assert rel < tol
    t.close()
