"""The reverse-mode (vector-Jacobian) functions of the discrete adjoint
(harm_ramt_amd/csrc/harm_adjoint.h) against central finite differences of
the forward functions they reverse (harm_device.h), on the CPU: a host-only
build of the same headers (tests/adjoint_shim.cpp).  For random states u,
directions v and adjoint weights w the check is
    <vjp(w), v>  ==  <w, df(u)[v]>
with df(u)[v] a fourth-order central difference, to its accuracy."""
import ctypes

import numpy as np
import pytest

import harm_ramt_amd as H

dp = ctypes.POINTER(ctypes.c_double)


def d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.ctypes.data_as(dp), a


class Geo:
    """Metric of one zone of a Kerr-Schild (MKS, a=0.9375) grid or Minkowski."""

    def __init__(self, kind="ks", loc=3):
        if kind == "ks":
            g = H.grid_mks(16, 32)
        else:
            g = H.grid_cartesian(8, 8)
        geom = H.Geometry(g)
        i, j = 3 + 8, 3 + 10
        if kind != "ks":
            i, j = 5, 5
        self.gcov = np.ascontiguousarray(geom.gcov[i, j, loc].ravel())
        self.gcon = np.ascontiguousarray(geom.gcon[i, j, loc].ravel())
        self.gdet = float(geom.gdet[i, j, loc])
        self.conn = np.ascontiguousarray(geom.conn[i, j].ravel())

    def args(self):
        return self.gcov.ctypes.data_as(dp), self.gcon.ctypes.data_as(dp), ctypes.c_double(self.gdet)


def params(eomtype=H.EOMGRMHD, fluxmethod=H.HLLFLUX, lim=H.MC, gam=4.0 / 3.0):
    return np.array([gam, eomtype, fluxmethod, 25.0, lim, 0.9, 0.05, 0.03])


def rand_state(rng, ffde=False):
    p = np.array([rng.uniform(0.2, 2.0), rng.uniform(0.1, 1.0), *rng.normal(0, 0.4, 3), *rng.normal(0, 0.6, 3)])
    if ffde:
        p[0] = p[1] = 0.0
    return p


def central(f, x, v, h=1e-4):
    """Fourth-order central difference of f along v."""
    return (8 * (f(x + h * v) - f(x - h * v)) - (f(x + 2 * h * v) - f(x - 2 * h * v))) / (12 * h)


def check_vjp(f, vjp, x, rng, nout, ntrial=6, rtol=1e-7):
    """<vjp(w), v> against the central difference of f along v."""
    for _ in range(ntrial):
        v = rng.normal(0, 1, x.size)
        w = rng.normal(0, 1, nout)
        fd = central(f, x, v)
        lhs = vjp(w) @ v
        rhs = w @ fd
        scale = np.abs(vjp(w)).max() * np.abs(v).max() + np.abs(w).max() * np.abs(fd).max() + 1e-30
        assert abs(lhs - rhs) <= rtol * scale, f"vjp {lhs} vs finite difference {rhs}"


@pytest.mark.parametrize("kind", ["ks", "mink"])
@pytest.mark.parametrize("eom", [H.EOMGRMHD, H.EOMFFDE])
@pytest.mark.parametrize("dirn", [0, 1, 2])
def test_primtoflux_vjp(shim, kind, eom, dirn):
    geo, P = Geo(kind), params(eomtype=eom)
    rng = np.random.default_rng(10 + dirn + 3 * eom)
    Pp, _ = d(P)
    for _ in range(5):
        x = rand_state(rng, ffde=(eom == H.EOMFFDE))

        def f(p):
            F = np.zeros(8)
            pp, p_ = d(p)
            shim.shim_primtoflux(Pp, *geo.args(), pp, dirn, F.ctypes.data_as(dp))
            return F

        def vjp(w):
            pb = np.zeros(8)
            pp, p_ = d(x)
            wp, w_ = d(w)
            shim.shim_primtoflux_vjp(Pp, *geo.args(), pp, dirn, wp, pb.ctypes.data_as(dp))
            return pb

        check_vjp(f, vjp, x, rng, 8)


@pytest.mark.parametrize("kind", ["ks", "mink"])
@pytest.mark.parametrize("dirn", [1, 2])
def test_vchar_vjp(shim, kind, dirn):
    geo, P = Geo(kind, loc=dirn - 1), params()
    rng = np.random.default_rng(20 + dirn)
    Pp, _ = d(P)
    for _ in range(8):
        x = rand_state(rng)

        def f(p):
            v = np.zeros(2)
            pp, p_ = d(p)
            shim.shim_vchar(Pp, *geo.args(), pp, dirn, v.ctypes.data_as(dp))
            return v

        def vjp(w):
            pb = np.zeros(8)
            pp, p_ = d(x)
            wp, w_ = d(w)
            ns = shim.shim_vchar_vjp(Pp, *geo.args(), pp, dirn, wp, pb.ctypes.data_as(dp))
            assert ns == 0
            return pb

        check_vjp(f, vjp, x, rng, 2)


@pytest.mark.parametrize("kind", ["ks", "mink"])
@pytest.mark.parametrize("flux", [H.HLLFLUX, H.LAXFFLUX])
@pytest.mark.parametrize("dirn", [1, 2])
def test_face_flux_vjp(shim, kind, flux, dirn):
    geo, P = Geo(kind, loc=dirn - 1), params(fluxmethod=flux)
    rng = np.random.default_rng(30 + dirn + 5 * flux)
    Pp, _ = d(P)
    for _ in range(6):
        pl = rand_state(rng)
        pr = pl * (1 + rng.normal(0, 0.1, 8))
        x = np.concatenate([pl, pr])

        def f(z):
            F = np.zeros(8)
            a, a_ = d(z[:8])
            b, b_ = d(z[8:])
            shim.shim_face_flux(Pp, *geo.args(), a, b, dirn, F.ctypes.data_as(dp))
            return F

        def vjp(w):
            plb, prb = np.zeros(8), np.zeros(8)
            a, a_ = d(pl)
            b, b_ = d(pr)
            wp, w_ = d(w)
            ns = shim.shim_face_flux_vjp(Pp, *geo.args(), a, b, dirn, wp, plb.ctypes.data_as(dp),
                                         prb.ctypes.data_as(dp))
            assert ns == 0
            return np.concatenate([plb, prb])

        check_vjp(f, vjp, x, rng, 8)


def test_source_vjp(shim):
    geo, P = Geo("ks"), params()
    rng = np.random.default_rng(40)
    Pp, _ = d(P)
    cp = geo.conn.ctypes.data_as(dp)
    assert np.abs(geo.conn).max() > 0
    for _ in range(6):
        x = rand_state(rng)

        def f(p):
            dU = np.zeros(8)
            pp, p_ = d(p)
            shim.shim_source(Pp, *geo.args(), cp, pp, dU.ctypes.data_as(dp))
            return dU

        def vjp(w):
            pb = np.zeros(8)
            pp, p_ = d(x)
            wp, w_ = d(w)
            shim.shim_source_vjp(Pp, *geo.args(), cp, pp, wp, pb.ctypes.data_as(dp))
            return pb

        check_vjp(f, vjp, x, rng, 8)


def test_bsq_vjp(shim):
    geo = Geo("ks")
    rng = np.random.default_rng(50)
    for _ in range(6):
        x = rand_state(rng)

        def f(p):
            pp, p_ = d(p)
            return np.array([shim.shim_bsq(*geo.args(), pp)])

        def vjp(w):
            pb = np.zeros(8)
            pp, p_ = d(x)
            shim.shim_bsq_vjp(*geo.args(), pp, pb.ctypes.data_as(dp))
            return pb * w[0]

        check_vjp(f, vjp, x, rng, 1)


@pytest.mark.parametrize("kind", ["ks", "mink"])
def test_inversion_vjp_is_inverse_transpose_jacobian(shim, kind):
    # p = P(U) with U = primtoU(p): Ub = J^{-T} pbar, i.e. <Ub, J v> = <pbar, v>
    # for every v, with J v the central difference of primtoU along v
    geo, P = Geo(kind), params()
    rng = np.random.default_rng(60)
    Pp, _ = d(P)
    for _ in range(8):
        x = rand_state(rng)
        pbar = rng.normal(0, 1, 8)
        Ub = np.zeros(8)
        xp, x_ = d(x)
        bp, b_ = d(pbar)
        assert shim.shim_inversion_vjp(Pp, *geo.args(), xp, bp, Ub.ctypes.data_as(dp)) == 0
        for _ in range(4):
            v = rng.normal(0, 1, 8)

            def U(p):
                F = np.zeros(8)
                pp, p_ = d(p)
                shim.shim_primtoflux(Pp, *geo.args(), pp, 0, F.ctypes.data_as(dp))
                return F

            Jv = central(U, x, v)
            scale = np.abs(Ub * Jv).sum() + np.abs(pbar * v).sum()   # the sums cancel: compare to their terms
            assert abs(Ub @ Jv - pbar @ v) <= 1e-9 * scale


@pytest.mark.parametrize("lim", [H.MC, H.VANL, H.MINM, H.NLIMCENT, H.NLIMUP, H.NLIMDOWN, H.DONOR])
def test_slope_coefficients_are_the_active_branch_derivative(shim, lim):
    rng = np.random.default_rng(70 + lim)
    c = np.zeros(3)
    n = 0
    for _ in range(300):
        y = rng.normal(0, 1, 3)
        shim.shim_slope3_coef(lim, *y, c.ctypes.data_as(dp))
        h = 1e-7
        fd = np.zeros(3)
        for m in range(3):
            e = np.zeros(3)
            e[m] = h
            fd[m] = (shim.shim_slope3(lim, *(y + e)) - shim.shim_slope3(lim, *(y - e))) / (2 * h)
        # points within h of a branch switch differ; everywhere else exact
        if np.allclose(fd, c, rtol=1e-6, atol=1e-6):
            n += 1
    assert n >= 290
