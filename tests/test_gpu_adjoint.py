"""The discrete adjoint (harm_adj_*) on the GPU.

Two independent checks of the gradient of the magnetic energy after n steps
of the (perturbed) Orszag-Tang problem with respect to the initial state:
  - central finite differences of the forward step with the same time-step
    sequence, for schemes whose forward map is differentiable (NLIM* and
    DONOR slopes; HLL/LAXF and the wave-speed max/min are smooth away from
    ties, which a generic state avoids);
  - the CPU reference sweep (tests/adjoint_reference.py) for the limiters with
    kinks (MC, VANL, MINM), where finite differences cross branch switches
    but the active-branch gradient is still defined and must agree to
    round-off.
"""
import sys

import numpy as np
import pytest

import harm_ramt_amd as H

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ot_setup(n1, n2=None, lim=H.MC, flux=H.HLLFLUX, timeorder=4):
    n2 = n2 or n1
    g = H.grid_cartesian(n1, n2)
    geom = H.Geometry(g)
    prm = H.make_params(n1, n2, g.dx1, g.dx2, lim=lim, fluxmethod=flux, timeorder=timeorder, gam=5.0 / 3.0,
                        cour=0.8, bc=(H.PERIODIC,) * 4, fixupflux=0, poledeath=0)
    return g, geom, prm


def ot_state(g, B0=1.0 / np.sqrt(4 * np.pi), vamp=0.5, drho=None, gam=5.0 / 3.0):
    """The relativistic Orszag-Tang state as a function of its parameters."""
    n1, n2 = g.n1, g.n2
    i = np.arange(-3, n1 + 3)[:, None]
    j = np.arange(-3, n2 + 3)[None, :]
    x = g.startx1 + (i + 0.5) * g.dx1 + 0 * j
    y = g.startx2 + (j + 0.5) * g.dx2 + 0 * i
    v1, v2 = -vamp * np.sin(2 * np.pi * y), vamp * np.sin(2 * np.pi * x)
    gamma = 1 / np.sqrt(1 - v1 * v1 - v2 * v2)
    p = np.zeros((n1 + 6, n2 + 6, 8))
    p[..., 0] = 25 / (36 * np.pi)
    if drho is not None:
        p[..., 0] += drho
    p[..., 1] = 5 / (12 * np.pi) / (gam - 1)
    p[..., 2], p[..., 3] = gamma * v1, gamma * v2
    p[..., 5], p[..., 6] = -B0 * np.sin(2 * np.pi * y), B0 * np.sin(4 * np.pi * x)
    return p


def generic(p, seed=11, eps=1e-3):
    """OT with a small random perturbation of every primitive: no exact
    symmetry ties or zero slopes left (those are kinks of the scheme)."""
    rng = np.random.default_rng(seed)
    q = p.copy()
    q[3:-3, 3:-3] *= 1 + rng.normal(0, eps, q[3:-3, 3:-3].shape)
    q[3:-3, 3:-3, 4] += rng.normal(0, eps, q.shape[0:2])[3:-3, 3:-3]
    q[3:-3, 3:-3, 7] += rng.normal(0, eps, q.shape[0:2])[3:-3, 3:-3]
    return q


def magnetic_energy(geom, prm, p0, dts):
    """J after len(dts) steps with the time steps imposed."""
    h = H.Harm(prm, geom)
    h.set_array(H.P, p0)
    h.bound_prim(H.P)
    h.set_time(0.0, dts[0], 1e30)
    a = H.Adjoint(h, len(dts), len(dts))
    a.forward(dts)
    J = a.seed_magnetic_energy()
    a.close()
    h.close()
    return J


def gradient(geom, prm, p0, nsteps, K, dt0=1e-3):
    h = H.Harm(prm, geom)
    h.set_array(H.P, p0)
    h.bound_prim(H.P)
    h.set_time(0.0, dt0, 1e30)
    a = H.Adjoint(h, nsteps, K)
    a.forward()
    dts = a.dts()
    pN = h.get_array(H.P)
    J = a.seed_magnetic_energy()
    lam = a.backward()
    ns = a.nonsmooth()
    after = h.get_array(H.P)
    a.close()
    h.close()
    return J, lam, dts, ns, pN, after


def interior(a):
    return a[3:-3, 3:-3]


def fd_along(geom, prm, p0, dts, v, h=1e-6):
    return (magnetic_energy(geom, prm, p0 + h * v, dts) - magnetic_energy(geom, prm, p0 - h * v, dts)) / (2 * h)


def test_adjoint_matches_finite_differences_on_random_directions():
    n, nsteps = 32, 8
    say(f"[case] adjoint OT {n}x{n} NLIMCENT/HLL/RK4 with flux-CT, {nsteps} steps, checkpoints every 3")
    g, geom, prm = ot_setup(n, lim=H.NLIMCENT)
    p0 = generic(ot_state(g))
    J, lam, dts, ns, pN, after = gradient(geom, prm, p0, nsteps, 3)
    say(f"  J = {J:.15e}, non-smooth points {ns}")
    assert ns == (0, 0)
    assert np.array_equal(after, pN), "the context is left at the end of the forward run"
    assert np.all(lam[:3] == 0) and np.all(lam[-3:] == 0) and np.all(lam[:, :3] == 0) and np.all(lam[:, -3:] == 0)
    rng = np.random.default_rng(3)
    scale = np.abs(interior(p0)).max(axis=(0, 1)) + 1e-3
    for trial in range(4):
        v = np.zeros_like(p0)
        interior(v)[...] = rng.normal(0, 1, interior(v).shape) * scale
        fd = fd_along(geom, prm, p0, dts, v)
        ad = float((lam * v).sum())
        say(f"  direction {trial}: adjoint {ad:.12e}  finite difference {fd:.12e}")
        assert ad == pytest.approx(fd, rel=1e-6)


def test_adjoint_single_zone_components():
    n, nsteps = 16, 4
    say(f"[case] adjoint OT {n}x{n} NLIMCENT, {nsteps} steps: single-zone derivatives of every primitive")
    g, geom, prm = ot_setup(n, lim=H.NLIMCENT)
    p0 = generic(ot_state(g))
    J, lam, dts, ns, pN, after = gradient(geom, prm, p0, nsteps, 2)
    assert ns == (0, 0)
    for (i, j) in ((3 + 5, 3 + 7), (3 + 0, 3 + 15)):
        for k in range(8):
            v = np.zeros_like(p0)
            v[i, j, k] = 1.0
            fd = fd_along(geom, prm, p0, dts, v, h=1e-5)
            assert lam[i, j, k] == pytest.approx(fd, rel=1e-5, abs=1e-10 * abs(J)), f"zone {(i, j)} prim {k}"


@pytest.mark.parametrize("lim,flux,to", [(H.NLIMUP, H.LAXFFLUX, 2), (H.DONOR, H.HLLFLUX, 3), (H.NLIMDOWN, H.HLLFLUX, 1)])
def test_adjoint_other_smooth_schemes(lim, flux, to):
    n, nsteps = 24, 5
    say(f"[case] adjoint OT {n}x{n} lim={lim} flux={flux} RK{to}, {nsteps} steps vs finite differences")
    g, geom, prm = ot_setup(n, lim=lim, flux=flux, timeorder=to)
    p0 = generic(ot_state(g))
    J, lam, dts, ns, pN, after = gradient(geom, prm, p0, nsteps, 2)
    assert ns == (0, 0)
    rng = np.random.default_rng(4)
    v = np.zeros_like(p0)
    interior(v)[...] = rng.normal(0, 1, interior(v).shape) * (np.abs(interior(p0)).max(axis=(0, 1)) + 1e-3)
    assert float((lam * v).sum()) == pytest.approx(fd_along(geom, prm, p0, dts, v), rel=1e-6)


@pytest.mark.parametrize("lim,flux,to", [(H.MC, H.HLLFLUX, 4), (H.VANL, H.LAXFFLUX, 2), (H.MINM, H.HLLFLUX, 3)])
def test_adjoint_matches_cpu_reference_sweep(shim, lim, flux, to):
    from tests.adjoint_reference import ReferenceAdjoint
    n1, n2, nsteps = 8, 10, 2
    say(f"[case] adjoint OT {n1}x{n2} lim={lim} flux={flux} RK{to}, {nsteps} steps vs the CPU reference sweep")
    g, geom, prm = ot_setup(n1, n2, lim=lim, flux=flux, timeorder=to)
    rng = np.random.default_rng(5)
    drho = np.zeros((n1 + 6, n2 + 6))
    drho[3:-3, 3:-3] = rng.normal(0, 1e-3, (n1, n2))
    p0 = ot_state(g, drho=drho)
    J, lam, dts, ns, pN, after = gradient(geom, prm, p0, nsteps, 1)
    ref = ReferenceAdjoint(shim, geom, prm).gradient(p0, dts)
    err = np.abs(lam - ref).max() / np.abs(ref).max()
    say(f"  max |GPU - reference| / max |reference| = {err:.2e}")
    assert err < 1e-11


def test_adjoint_is_independent_of_the_checkpoint_interval():
    n, nsteps = 32, 7
    say(f"[case] adjoint OT {n}x{n} MC, {nsteps} steps: checkpoint intervals 1, 3, 7 give the same gradient")
    g, geom, prm = ot_setup(n)
    p0 = ot_state(g)
    lams = [gradient(geom, prm, p0, nsteps, K)[1] for K in (1, 3, 7)]
    assert np.array_equal(lams[0], lams[1]) and np.array_equal(lams[0], lams[2])


def test_adjoint_parameter_gradients():
    n, nsteps = 32, 10
    say(f"[case] adjoint OT {n}x{n} NLIMCENT, {nsteps} steps: dJ/dB0, dJ/dv_amplitude, dJ/d(rho perturbation)")
    g, geom, prm = ot_setup(n, lim=H.NLIMCENT)
    rng = np.random.default_rng(9)
    drho = np.zeros((n + 6, n + 6))
    drho[3:-3, 3:-3] = rng.normal(0, 1e-3, (n, n))
    B0, vamp = 1.0 / np.sqrt(4 * np.pi), 0.5
    p0 = ot_state(g, B0, vamp, drho)
    J, lam, dts, ns, pN, after = gradient(geom, prm, p0, nsteps, 4)
    assert ns == (0, 0)
    h = 1e-6
    for name, fwd in (("B0", lambda e: ot_state(g, B0 + e, vamp, drho)),
                      ("vamp", lambda e: ot_state(g, B0, vamp + e, drho))):
        dpdq = (fwd(h) - fwd(-h)) / (2 * h)
        ad = float((lam * dpdq).sum())
        fd = (magnetic_energy(geom, prm, fwd(h), dts) - magnetic_energy(geom, prm, fwd(-h), dts)) / (2 * h)
        say(f"  dJ/d{name}: adjoint {ad:.12e}  finite difference {fd:.12e}")
        assert ad == pytest.approx(fd, rel=1e-6)
    w = rng.normal(0, 1, (n, n))
    v = np.zeros_like(p0)
    v[3:-3, 3:-3, 0] = w
    ad = float((lam[3:-3, 3:-3, 0] * w).sum())
    assert ad == pytest.approx(fd_along(geom, prm, p0, dts, v), rel=1e-6)


def test_adjoint_rejects_unsupported_configurations():
    say("[case] adjoint argument checks")
    g = H.grid_mks(16, 32)
    geom = H.Geometry(g)
    prm = H.make_params(16, 32, g.dx1, g.dx2)
    h = H.Harm(prm, geom)
    with pytest.raises(H.HarmError, match="PERIODIC"):
        H.Adjoint(h, 4)
