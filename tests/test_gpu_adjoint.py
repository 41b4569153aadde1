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


def agree(ad, fd, lam, v, rtol=1e-6):
    """A directional derivative is a sum with cancellation: compare on the
    scale of its terms, sum |lambda v|."""
    return abs(ad - fd) <= rtol * float(np.abs(lam * v).sum())


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
        assert agree(ad, fd, lam, v)


def test_adjoint_single_zone_components():
    n, nsteps = 16, 4
    say(f"[case] adjoint OT {n}x{n} NLIMCENT, {nsteps} steps: single-zone derivatives of every primitive")
    g, geom, prm = ot_setup(n, lim=H.NLIMCENT)
    p0 = generic(ot_state(g))
    J, lam, dts, ns, pN, after = gradient(geom, prm, p0, nsteps, 2)
    assert ns == (0, 0)
    # h = 1e-6: at the corner zone a wave-speed max/min switches branch
    # within 1e-5 of rho and u (the CPU reference sweep gives the same
    # lambda there; finite differences converge to it only below that);
    # the absolute floor is the round-off of J's difference quotient
    for (i, j) in ((3 + 5, 3 + 7), (3 + 0, 3 + 15)):
        for k in range(8):
            v = np.zeros_like(p0)
            v[i, j, k] = 1.0
            fd = fd_along(geom, prm, p0, dts, v, h=1e-6)
            assert lam[i, j, k] == pytest.approx(fd, rel=1e-5, abs=3e-10 * abs(J)), f"zone {(i, j)} prim {k}"


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
    assert agree(float((lam * v).sum()), fd_along(geom, prm, p0, dts, v), lam, v)


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


@pytest.mark.parametrize("K", [3, 7])
def test_adjoint_second_backward_gives_the_same_gradient(K):
    """the last segment's states come from the forward run the first time;
    a second backward sweep (re-seeded) recomputes what the first overwrote"""
    n, nsteps = 32, 7
    say(f"[case] adjoint OT {n}x{n}, {nsteps} steps, K={K}: two backward sweeps after one forward")
    g, geom, prm = ot_setup(n)
    p0 = ot_state(g)
    h = H.Harm(prm, geom)
    h.set_array(H.P, p0)
    h.bound_prim(H.P)
    h.set_time(0.0, 1e-3, 1e30)
    a = H.Adjoint(h, nsteps, K)
    a.forward()
    a.seed_magnetic_energy()
    lam1 = a.backward()
    a.seed_magnetic_energy()
    lam2 = a.backward()
    a.close()
    h.close()
    assert np.array_equal(lam1, lam2)


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
        assert agree(ad, fd, lam, dpdq)
    w = rng.normal(0, 1, (n, n))
    v = np.zeros_like(p0)
    v[3:-3, 3:-3, 0] = w
    ad = float((lam[3:-3, 3:-3, 0] * w).sum())
    assert agree(ad, fd_along(geom, prm, p0, dts, v), lam, v)


def test_adjoint_rejects_unsupported_configurations():
    say("[case] adjoint argument checks")
    g = H.grid_mks(16, 32)
    geom = H.Geometry(g)
    prm = H.make_params(16, 32, g.dx1, g.dx2)
    h = H.Harm(prm, geom)
    with pytest.raises(H.HarmError, match="PERIODIC"):
        H.Adjoint(h, 4)


def ot_state_torch(g, B0, vamp, drho, gam=5.0 / 3.0):
    """ot_state() in torch, differentiable in its parameters."""
    import torch
    n1, n2 = g.n1, g.n2
    i = torch.arange(-3, n1 + 3, dtype=torch.float64)[:, None]
    j = torch.arange(-3, n2 + 3, dtype=torch.float64)[None, :]
    x = g.startx1 + (i + 0.5) * g.dx1 + 0 * j
    y = g.startx2 + (j + 0.5) * g.dx2 + 0 * i
    v1, v2 = -vamp * torch.sin(2 * np.pi * y), vamp * torch.sin(2 * np.pi * x)
    gamma = 1 / torch.sqrt(1 - v1 * v1 - v2 * v2)
    z = torch.zeros_like(x)
    return torch.stack([25 / (36 * np.pi) + drho, z + 5 / (12 * np.pi) / (gam - 1), gamma * v1, gamma * v2, z,
                        -B0 * torch.sin(2 * np.pi * y), B0 * torch.sin(4 * np.pi * x), z], -1)


@pytest.mark.parametrize("lim,flux,to", [(H.MC, H.HLLFLUX, 4), (H.NLIMCENT, H.LAXFFLUX, 2)])
def test_adjoint_matches_forward_mode_reference(lim, flux, to):
    import torch
    from torch.func import jvp
    from tests.torch_reference import TorchStep
    n, nsteps = 16, 4
    say(f"[case] adjoint OT {n}x{n} lim={lim} flux={flux} RK{to}, {nsteps} steps vs torch forward mode")
    g, geom, prm = ot_setup(n, lim=lim, flux=flux, timeorder=to)
    p0 = generic(ot_state(g))
    J, lam, dts, ns, pN, after = gradient(geom, prm, p0, nsteps, 2)
    T = TorchStep(n, n, g.dx1, g.dx2, lim=lim, fluxmethod=flux, timeorder=to)
    f = lambda p: T.magnetic_energy(T.run(p, dts))
    rng = np.random.default_rng(6)
    for trial in range(2):
        v = np.zeros_like(p0)
        interior(v)[...] = rng.normal(0, 1, interior(v).shape)
        Jt, dJ = jvp(f, (torch.from_numpy(p0),), (torch.from_numpy(v),))
        ad = float((lam * v).sum())
        say(f"  J gpu {J:.15e} torch {float(Jt):.15e}; dJ[v] adjoint {ad:.13e} forward mode {float(dJ):.13e}")
        assert float(Jt) == pytest.approx(J, rel=1e-12)
        assert agree(ad, float(dJ), lam, v, rtol=1e-10)


def test_torch_autograd_harness_parameter_gradients():
    """Config-3-style gradients through harm_ramt_amd.autograd: p0 built from
    (B0, v amplitude, rho perturbation field) in torch, J from the GPU, dJ/dq
    by J.backward() through the GPU reverse sweep; checked against torch
    forward mode through the reference step."""
    import torch
    from torch.func import jvp
    from harm_ramt_amd.autograd import MagneticEnergyAfterSteps
    from tests.torch_reference import TorchStep
    n, nsteps = 16, 6
    say(f"[case] torch.autograd harness, OT {n}x{n} MC/HLL/RK4, {nsteps} steps: dJ/dB0, dJ/dvamp, dJ/ddrho")
    g, geom, prm = ot_setup(n)
    rng = np.random.default_rng(8)
    drho0 = np.zeros((n + 6, n + 6))
    drho0[3:-3, 3:-3] = rng.normal(0, 1e-3, (n, n))
    B0 = torch.tensor(1.0 / np.sqrt(4 * np.pi), dtype=torch.float64, requires_grad=True)
    vamp = torch.tensor(0.5, dtype=torch.float64, requires_grad=True)
    drho = torch.tensor(drho0, requires_grad=True)
    harm = H.Harm(prm, geom)
    J = MagneticEnergyAfterSteps.apply(ot_state_torch(g, B0, vamp, drho), harm, nsteps, 1e-3, 2, None)
    J.backward()
    # the dt sequence the GPU forward used, for the reference
    h2 = H.Harm(prm, geom)
    h2.set_array(H.P, ot_state_torch(g, B0, vamp, drho).detach().numpy())
    h2.bound_prim(H.P)
    h2.set_time(0.0, 1e-3, 1e30)
    a = H.Adjoint(h2, nsteps, nsteps)
    a.forward()
    dts = a.dts()
    T = TorchStep(n, n, g.dx1, g.dx2, lim=H.MC, fluxmethod=H.HLLFLUX, timeorder=4)
    base = (B0.detach(), vamp.detach(), drho.detach())
    w = torch.from_numpy(rng.normal(0, 1, (n + 6, n + 6)))
    w[:3], w[-3:], w[:, :3], w[:, -3:] = 0, 0, 0, 0
    for name, tangent, grad in (("B0", (1.0, 0.0, 0.0 * w), B0.grad), ("vamp", (0.0, 1.0, 0.0 * w), vamp.grad),
                                ("drho", (0.0, 0.0, w), None)):
        t = (torch.tensor(tangent[0], dtype=torch.float64), torch.tensor(tangent[1], dtype=torch.float64), tangent[2])
        Jt, dJ = jvp(lambda b, va, dr: T.magnetic_energy(T.run(ot_state_torch(g, b, va, dr), dts)), base, t)
        ad = float(grad) if grad is not None else float((drho.grad * w).sum())
        scale = abs(ad) if grad is not None else float((drho.grad * w).abs().sum())
        say(f"  dJ/d{name}: torch.autograd through the GPU adjoint {ad:.13e}, forward mode {float(dJ):.13e}")
        assert float(Jt) == pytest.approx(float(J.detach()), rel=1e-12)
        assert abs(ad - float(dJ)) <= 1e-9 * scale


def _host_branch(lim, yl, yc, yr):
    """slope_branch on numpy arrays (interp.c:106-149 minmod choices)"""
    def mm(a, b):
        code = np.where(np.abs(a) < np.abs(b), 1, 2)
        m = np.where(np.abs(a) < np.abs(b), a, b)
        zero = a * b <= 0
        return np.where(zero, 0, code), np.where(zero, 0.0, m)
    if lim == H.MC:
        cin, m = mm(2.0 * (yc - yl), 2.0 * (yr - yc))
        cout, _ = mm(0.5 * (yr - yl), m)
        return cin | (cout << 2)
    if lim == H.MINM:
        return mm(yc - yl, yr - yc)[0]
    if lim == H.VANL:
        return np.where((yc - yl) * (yr - yc) <= 0.0, 0, 1)
    return np.zeros_like(yc, dtype=int)


def _host_switches(lim, a, b):
    """branch codes of two bounded states compared where the faces use them"""
    n1, n2 = a.shape[0] - 6, a.shape[1] - 6
    n = c = 0
    # x1 slopes of zones i = -1..N1 (rows 2..N1+3), j interior; x2 likewise
    ca = _host_branch(lim, a[1:n1 + 3, 3:n2 + 3], a[2:n1 + 4, 3:n2 + 3], a[3:n1 + 5, 3:n2 + 3])
    cb = _host_branch(lim, b[1:n1 + 3, 3:n2 + 3], b[2:n1 + 4, 3:n2 + 3], b[3:n1 + 5, 3:n2 + 3])
    n += int((ca != cb).sum())
    c += ca.size
    ca = _host_branch(lim, a[3:n1 + 3, 1:n2 + 3], a[3:n1 + 3, 2:n2 + 4], a[3:n1 + 3, 3:n2 + 5])
    cb = _host_branch(lim, b[3:n1 + 3, 1:n2 + 3], b[3:n1 + 3, 2:n2 + 4], b[3:n1 + 3, 3:n2 + 5])
    n += int((ca != cb).sum())
    c += ca.size
    return n, c


@pytest.mark.parametrize("lim", [H.MC, H.MINM, H.VANL])
def test_branch_switches_count_limiter_branches(lim):
    """harm_adj_branch_switches against the branch codes formed on the host
    from the two bounded initial states (one forward-Euler step: its only
    reconstruction input); zero for a run against itself and for a limiter
    without branches"""
    n = 24
    say(f"[case] limiter branch switches, OT {n}x{n}, lim={lim}")
    g, geom, prm = ot_setup(n, lim=lim, timeorder=1)
    base = generic(ot_state(g))
    w = np.random.default_rng(5).normal(0, 1, base.shape)
    runs, states = [], []
    for e in (1e-3, -1e-3, 1e-3):
        h = H.Harm(prm, geom)
        h.set_array(H.P, base + e * w)
        h.bound_prim(H.P)
        states.append(h.get_array(H.P))
        h.set_time(0.0, 1e-4, 1e30)
        a = H.Adjoint(h, 1, 1)
        a.forward([1e-4])
        runs.append((h, a))
    got = runs[0][1].branch_switches(runs[1][1])
    want = _host_switches(lim, states[0], states[1])
    assert got == want and want[0] > 0
    assert runs[0][1].branch_switches(runs[2][1]) == (0, want[1])
    for h, a in runs:
        a.close()
        h.close()
