"""TEST INFRASTRUCTURE: a CPU reference of the discrete adjoint's reverse
sweep, written independently of the GPU kernels (plain Python loops over
zones and faces) on top of the oracle's forward substages and the local
vector-Jacobian products of tests/adjoint_shim.cpp (each checked against
finite differences in tests/test_adjoint_math.py).  It follows the same
mathematics as harm_ramt_amd/csrc/harm_adjoint.inl -- periodic bound_prim
transposed, inversion by the implicit-function theorem, RK combination
reversed, flux-CT and divergence transposed, face fluxes and reconstruction
reversed on their active branches -- so the GPU gradient must agree with it
to round-off, for every limiter including the piecewise-linear ones whose
kinks make finite differences unreliable.  Small grids only (seconds)."""
import ctypes

import numpy as np

import harm_ramt_amd as H
from oracle.oracle import Oracle

NG = 3
dp = ctypes.POINTER(ctypes.c_double)

# setup_rktimestep (step_ch.c:365-460): CUf, Cunew, and the pi/pb/pf selectors
RK = {4: ([[1, 0, .5], [1, 0, .5], [1, 0, 1], [1, 0, 1]], [[1, 1 / 6, 0], [0, 1 / 3, 0], [0, 1 / 3, 0], [0, 1 / 6, 0]],
          [0, 0, 0, 0], [0, 1, 2, 1], [1, 2, 1, 0]),
      3: ([[1, 0, 1], [3 / 4, 1 / 4, 1 / 4], [1 / 3, 2 / 3, 2 / 3]], [[0, 0, 0], [0, 0, 0], [0, 0, 1]],
          [0, 0, 0], [0, 1, 2], [1, 2, 0]),
      2: ([[1, 0, .5], [1, 0, 1]], [[0, 0, 0], [0, 0, 1]], [0, 0], [0, 1], [1, 0]),
      1: ([[1, 0, 1]], [[0, 0, 1]], [0], [0], [0])}


def _p(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.ctypes.data_as(dp), a


class ReferenceAdjoint:
    def __init__(self, shim, geom, prm):
        self.lib, self.geom, self.prm = shim, geom, prm
        self.n1, self.n2 = prm.n1, prm.n2
        self.pd = H.params_dict(prm)
        self.P = np.array([prm.gam, prm.eomtype, prm.fluxmethod, prm.gammamax, prm.lim, prm.cour, prm.dx1, prm.dx2])
        self.to = prm.timeorder
        self.ct = prm.fluxb == H.FLUXCTTOTH

    def _geo(self, i, j, loc):
        a, ka = _p(self.geom.gcov[i + NG, j + NG, loc].ravel())
        b, kb = _p(self.geom.gcon[i + NG, j + NG, loc].ravel())
        return (a, b, ctypes.c_double(self.geom.gdet[i + NG, j + NG, loc])), (ka, kb)

    def oracle(self, p0, dt):
        o = Oracle(self.pd)
        o.set_geometry(*self.geom.tables())
        o.set_array(0, p0)
        o.bound_prim(0)
        o.set_time(0.0, dt, 1e30)
        return o

    def forward_states(self, o, dt):
        """One step with dt imposed; returns [p_n, pf_0, ..., pf_last]."""
        CUf, Cn, pii, pbb, pff = RK[self.to]
        t, _ = o.get_time()
        o.set_time(t, dt, 1e30)
        st = [o.get_array(0).copy()]
        for s in range(len(CUf)):
            o.advance(pii[s], pbb[s], CUf[s], pff[s], Cn[s], s, len(CUf))
            o.bound_prim(pff[s])
            st.append(o.get_array(pff[s]).copy())
        o.set_time(t + dt, dt, 1e30)
        return st

    def boundT(self, lam):
        n1, n2 = self.n1, self.n2
        for i in range(-NG, n1 + NG):
            for j in list(range(-NG, 0)) + list(range(n2, n2 + NG)):
                js = j + n2 if j < 0 else j - n2
                lam[i + NG, js + NG] += lam[i + NG, j + NG]
                lam[i + NG, j + NG] = 0
        for j in range(n2):
            for i in list(range(-NG, 0)) + list(range(n1, n1 + NG)):
                is_ = i + n1 if i < 0 else i - n1
                lam[is_ + NG, j + NG] += lam[i + NG, j + NG]
                lam[i + NG, j + NG] = 0

    def seed_magnetic_energy(self, pN):
        lam = np.zeros_like(pN)
        for i in range(self.n1):
            for j in range(self.n2):
                ga, keep = self._geo(i, j, 3)
                pb = np.zeros(8)
                pp, kp = _p(pN[i + NG, j + NG])
                self.lib.shim_bsq_vjp(ga[0], ga[1], ga[2], pp, pb.ctypes.data_as(dp))
                lam[i + NG, j + NG] = 0.5 * pb * self.geom.gdet[i + NG, j + NG, 3] * self.prm.dx1 * self.prm.dx2
        return lam

    def reverse_step(self, st, lam, dt):
        lib, n1, n2 = self.lib, self.n1, self.n2
        CUf, Cn, pii, pbb, pff = RK[self.to]
        nto = len(CUf)
        P, kP = _p(self.P)
        lun, lul, lUi = np.zeros_like(lam), np.zeros_like(lam), np.zeros_like(lam)
        for s in range(nto - 1, -1, -1):
            a0, a1, a2 = CUf[s]
            b0, b1, b2 = Cn[s]
            last = s == nto - 1
            pf, pb = st[s + 1], st[s]
            self.boundT(lam)
            ldU = np.zeros_like(lam)
            for i in range(n1):
                for j in range(n2):
                    I, J = i + NG, j + NG
                    ga, keep = self._geo(i, j, 3)
                    pp, kp = _p(pf[I, J])
                    lp, kl = _p(lam[I, J])
                    Ub = np.zeros(8)
                    assert lib.shim_inversion_vjp(P, *ga, pp, lp, Ub.ctypes.data_as(dp)) == 0
                    if last:
                        lun[I, J] += Ub
                    lUf = (0 if last else Ub) + b2 * lun[I, J] + lul[I, J]
                    lUi[I, J] += b0 * lun[I, J] + a0 * lUf
                    ldU[I, J] = b1 * dt * lun[I, J] + a2 * dt * lUf
                    lul[I, J] = a1 * lUf

            def ld(i, j, k):
                return ldU[i + NG, j + NG, k] if (0 <= i < n1 and 0 <= j < n2) else 0.

            idx = {1: 1 / self.prm.dx1, 2: 1 / self.prm.dx2}
            lemf = {}
            for i in range(n1 + 1):
                for j in range(n2 + 1):
                    f1 = lambda a, b: (ld(a, b, 6) - ld(a - 1, b, 6)) * idx[1]
                    f2 = lambda a, b: (ld(a, b, 5) - ld(a, b - 1, 5)) * idx[2]
                    lemf[(i, j)] = 0.5 * (f1(i, j) + f1(i, j - 1)) - 0.5 * (f2(i, j) + f2(i - 1, j))

            def le(a, b):
                return self.geom.gdet[a + NG, b + NG, 2] * lemf[(a, b)] if (0 <= a <= n1 and 0 <= b <= n2) else 0.

            lf = {1: np.zeros(lam.shape[:2] + (16,)), 2: np.zeros(lam.shape[:2] + (16,))}
            for d in (1, 2):
                for i in (range(0, n1 + 1) if d == 1 else range(-1, n1 + 1)):
                    for j in (range(-1, n2 + 1) if d == 1 else range(0, n2 + 1)):
                        dn = (lambda k: ld(i - 1, j, k)) if d == 1 else (lambda k: ld(i, j - 1, k))
                        Fob = np.array([(ld(i, j, k) - dn(k)) * idx[d] for k in range(8)])
                        gdf = self.geom.gdet[i + NG, j + NG, d - 1]
                        if self.ct:
                            if d == 1:
                                Fob[5], Fob[6] = 0., 0.25 * (le(i, j) + le(i, j + 1)) / gdf
                            else:
                                Fob[6], Fob[5] = 0., -0.25 * (le(i, j) + le(i + 1, j)) / gdf
                        zz = (lambda m: (i + m, j)) if d == 1 else (lambda m: (i, j + m))
                        pl, pr = np.zeros(8), np.zeros(8)
                        for k in range(8):
                            y = [pb[zz(m)[0] + NG, zz(m)[1] + NG, k] for m in (-2, -1, 0, 1)]
                            pl[k] = y[1] + 0.5 * lib.shim_slope3(self.prm.lim, y[0], y[1], y[2])
                            pr[k] = y[2] - 0.5 * lib.shim_slope3(self.prm.lim, y[1], y[2], y[3])
                        ga, keep = self._geo(i, j, d - 1)
                        plb, prb = np.zeros(8), np.zeros(8)
                        a, ka = _p(pl)
                        b, kb = _p(pr)
                        fb, kf = _p(Fob)
                        lib.shim_face_flux_vjp(P, *ga, a, b, d, fb, plb.ctypes.data_as(dp), prb.ctypes.data_as(dp))
                        lf[d][i + NG, j + NG, :8] = plb
                        lf[d][i + NG, j + NG, 8:] = prb
            new = np.zeros_like(lam)
            c = np.zeros(3)
            for i in range(-NG, n1 + NG):
                for j in range(-NG, n2 + NG):
                    for d in (1, 2):
                        cc, nmax = (i, n1 + NG) if d == 1 else (j, n2 + NG)
                        F = (lambda f: (f, j)) if d == 1 else (lambda f: (i, f))

                        def lpl(f, k):
                            return lf[d][F(f)[0] + NG, F(f)[1] + NG, k] if -NG <= f < nmax else 0.

                        def lpr(f, k):
                            return lf[d][F(f)[0] + NG, F(f)[1] + NG, 8 + k] if -NG <= f < nmax else 0.

                        for k in range(8):
                            acc = lpl(cc + 1, k) + lpr(cc, k)
                            for w in (cc - 1, cc, cc + 1):
                                if w - 1 < -NG or w + 1 >= nmax:
                                    continue
                                ldq = 0.5 * lpl(w + 1, k) - 0.5 * lpr(w, k)
                                if ldq == 0:
                                    continue
                                pv = [pb[F(m)[0] + NG, F(m)[1] + NG, k] for m in (w - 1, w, w + 1)]
                                lib.shim_slope3_coef(self.prm.lim, *pv, c.ctypes.data_as(dp))
                                acc += (c[2] if w == cc - 1 else (c[1] if w == cc else c[0])) * ldq
                            new[i + NG, j + NG, k] += acc
            # the source term (zero connection in Minkowski, kept for generality)
            for i in range(n1):
                for j in range(n2):
                    ga, keep = self._geo(i, j, 3)
                    cn, kc = _p(self.geom.conn[i + NG, j + NG].ravel())
                    pp, kp = _p(pb[i + NG, j + NG])
                    du, kd = _p(ldU[i + NG, j + NG])
                    sb = np.zeros(8)
                    lib.shim_source_vjp(P, *ga, cn, pp, du, sb.ctypes.data_as(dp))
                    new[i + NG, j + NG] += sb
            lam = new
            if s == 0:
                for i in range(n1):
                    for j in range(n2):
                        ga, keep = self._geo(i, j, 3)
                        pb_ = np.zeros(8)
                        pp, kp = _p(pb[i + NG, j + NG])
                        ui, ku = _p(lUi[i + NG, j + NG])
                        lib.shim_primtoflux_vjp(P, *ga, pp, 0, ui, pb_.ctypes.data_as(dp))
                        lam[i + NG, j + NG] += pb_
        return lam

    def gradient(self, p0, dts):
        """lambda_0 of the magnetic energy after len(dts) steps."""
        o = self.oracle(p0, dts[0])
        states = [self.forward_states(o, dt) for dt in dts]
        lam = self.seed_magnetic_energy(o.get_array(0))
        for n in range(len(dts) - 1, -1, -1):
            lam = self.reverse_step(states[n], lam, dts[n])
        self.boundT(lam)
        return lam
