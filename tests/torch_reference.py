"""TEST INFRASTRUCTURE: an independent float64 torch restatement of the
conservative step for the Minkowski, periodic configurations (the
Orszag-Tang problem), differentiable by torch.func, used as the reference
forward-mode gradient for the GPU discrete adjoint: torch.func.jvp of this
step along a direction v must equal <lambda_0, v> from the reverse sweep.

Written from the mathematics of the reference step (step_ch.c fluxcalc /
flux_ct / advance, phys.c state and fluxes, vchar.c wave speeds, interp.c
slopes) in vectorised torch, not from the GPU code: zone-wise formulas on
whole arrays, a periodic wrap for bound_prim, and the inversion of U(p) by a
Newton iteration in the primitive variables (a different algorithm from the
reference's 2D Newton in (W, v^2); both converge to the same root), so its
derivatives follow the implicit function theorem through the iteration.
Only the Minkowski metric (gcov = diag(-1,1,1,1), gdet = 1, no source term),
MC/NLIM*/DONOR slopes, HLL/LAXF, flux-CT and RK1-4 are covered."""
import torch
from torch.func import jacfwd, vmap

NG = 3
RK = {4: ([[1, 0, .5], [1, 0, .5], [1, 0, 1], [1, 0, 1]], [[1, 1 / 6, 0], [0, 1 / 3, 0], [0, 1 / 3, 0], [0, 1 / 6, 0]],
          ["p", "p", "p", "p"], ["p", "k0", "k1", "k0"], ["k0", "k1", "k0", "p"]),
      3: ([[1, 0, 1], [3 / 4, 1 / 4, 1 / 4], [1 / 3, 2 / 3, 2 / 3]], [[0, 0, 0], [0, 0, 0], [0, 0, 1]],
          ["p", "p", "p"], ["p", "k0", "k1"], ["k0", "k1", "p"]),
      2: ([[1, 0, .5], [1, 0, 1]], [[0, 0, 0], [0, 0, 1]], ["p", "p"], ["p", "k0"], ["k0", "p"]),
      1: ([[1, 0, 1]], [[0, 0, 1]], ["p"], ["p"], ["p"])}


class TorchStep:
    def __init__(self, n1, n2, dx1, dx2, gam=5.0 / 3.0, lim=3, fluxmethod=1, timeorder=4, ct=True, newton_iters=12):
        self.n1, self.n2, self.dx1, self.dx2, self.gam = n1, n2, dx1, dx2, gam
        self.lim, self.fluxmethod, self.to, self.ct, self.iters = lim, fluxmethod, timeorder, ct, newton_iters
        self.wi = (torch.arange(n1 + 2 * NG) - NG) % n1
        self.wj = (torch.arange(n2 + 2 * NG) - NG) % n2

    # ---- bound_prim: periodic in both directions ----
    def bound(self, p):
        inner = p[NG:NG + self.n1, NG:NG + self.n2]
        return inner[self.wi][:, self.wj]

    # ---- state and fluxes in Minkowski (u-tilde^i = u^i, alpha = 1) ----
    def state(self, p):
        u = p[..., 2:5]
        B = p[..., 5:8]
        gamma = torch.sqrt(1 + (u * u).sum(-1))
        b0 = (B * u).sum(-1)
        bi = (B + b0[..., None] * u) / gamma[..., None]
        ucon = torch.cat([gamma[..., None], u], -1)
        ucov = torch.cat([-gamma[..., None], u], -1)
        bcon = torch.cat([b0[..., None], bi], -1)
        bcov = torch.cat([-b0[..., None], bi], -1)
        return ucon, ucov, bcon, bcov

    def flux(self, p, d):
        ucon, ucov, bcon, bcov = self.state(p)
        rho, uu = p[..., 0], p[..., 1]
        P = (self.gam - 1) * uu
        bsq = (bcon * bcov).sum(-1)
        eta = P + rho + uu + bsq
        ptot = P + 0.5 * bsq
        delta = torch.zeros(4, dtype=p.dtype)
        delta[d] = 1.0
        T = eta[..., None] * ucon[..., d:d + 1] * ucov + ptot[..., None] * delta - bcon[..., d:d + 1] * bcov
        F0 = rho * ucon[..., d]
        if d == 0:
            FB = p[..., 5:8]
        else:
            FB = (p[..., 5:8] * ucon[..., d:d + 1] - p[..., 4 + d:5 + d] * ucon[..., 1:4]) / ucon[..., 0:1]
        return torch.cat([F0[..., None], (T[..., 0] + F0)[..., None], T[..., 1:4], FB], -1)

    def speeds(self, p, d):
        ucon, ucov, bcon, bcov = self.state(p)
        rho, uu = p[..., 0], p[..., 1]
        g = self.gam
        bsq = (bcon * bcov).sum(-1)
        WW = rho + g * uu
        cs2 = g * (g - 1) * uu / WW
        EF = bsq + WW
        va2 = bsq / EF
        cms2 = cs2 + va2 - cs2 * va2
        Au, Bu = ucon[..., d], ucon[..., 0]
        Asq, Bsq, AB = 1.0, -1.0, 0.0
        A = Bu * Bu - (Bsq + Bu * Bu) * cms2
        Bq = 2 * (Au * Bu - (AB + Au * Bu) * cms2)
        C = Au * Au - (Asq + Au * Au) * cms2
        sd = torch.sqrt(Bq * Bq - 4 * A * C)
        vm = (Bq - sd) / (2 * A)
        vp = (Bq + sd) / (2 * A)
        return torch.maximum(vp, vm), torch.minimum(vp, vm)

    # ---- reconstruction: slope of every zone along dim (edges zero) ----
    def slopes(self, p, dim):
        y = p.movedim(dim, 0)
        yl, yc, yr = y[:-2], y[1:-1], y[2:]

        def minmod(a, b):
            return torch.where(a * b <= 0, torch.zeros_like(a), torch.where(a.abs() < b.abs(), a, b))

        if self.lim == 3:
            dq = minmod(0.5 * (yr - yl), minmod(2 * (yc - yl), 2 * (yr - yc)))
        elif self.lim == 101:
            dq = 0.5 * (yr - yl)
        elif self.lim == 102:
            dq = yc - yl
        elif self.lim == 103:
            dq = yr - yc
        else:
            dq = torch.zeros_like(yc)
        z = torch.zeros_like(y[:1])
        return torch.cat([z, dq, z], 0).movedim(0, dim)

    def face_fluxes(self, p, d):
        """F at the faces of direction d between every pair of neighbours:
        index m along d is the face between zones m-1 and m (m >= 1)."""
        dim = d - 1
        dq = self.slopes(p, dim)
        pm = p.movedim(dim, 0)
        dqm = dq.movedim(dim, 0)
        pl = (pm[:-1] + 0.5 * dqm[:-1])
        pr = (pm[1:] - 0.5 * dqm[1:])
        kb = 4 + d
        bav = 0.5 * (pl[..., kb] + pr[..., kb])
        pl = torch.cat([pl[..., :kb], bav[..., None], pl[..., kb + 1:]], -1)
        pr = torch.cat([pr[..., :kb], bav[..., None], pr[..., kb + 1:]], -1)
        Fl, Fr = self.flux(pl, d), self.flux(pr, d)
        Ul, Ur = self.flux(pl, 0), self.flux(pr, 0)
        cmaxl, cminl = self.speeds(pl, d)
        cmaxr, cminr = self.speeds(pr, d)
        zero = torch.zeros_like(cmaxl)
        cmax = torch.maximum(zero, torch.maximum(cmaxl, cmaxr))
        cmin = torch.maximum(zero, -torch.minimum(cminl, cminr))
        if self.fluxmethod == 1:
            F = (cmax[..., None] * Fl + cmin[..., None] * Fr - (cmax * cmin)[..., None] * (Ur - Ul)) / \
                (cmax + cmin)[..., None]
        else:
            ctop = torch.maximum(cmax, cmin)
            F = 0.5 * (Fl + Fr - ctop[..., None] * (Ur - Ul))
        # prepend a dummy face so that index m is the face between m-1 and m
        F = torch.cat([torch.zeros_like(F[:1]), F], 0)
        return F.movedim(0, dim)

    def dU(self, p):
        n1, n2 = self.n1, self.n2
        F1 = self.face_fluxes(p, 1)   # F1[I, J]: x1-face between zones I-1 and I (array indices)
        F2 = self.face_fluxes(p, 2)
        if self.ct:
            # corner EMF at (i, j): 0.25 (F1[B2](i,j) + F1[B2](i,j-1) - F2[B1](i,j) - F2[B1](i-1,j))
            I = torch.arange(NG, NG + n1 + 1)[:, None]
            J = torch.arange(NG, NG + n2 + 1)[None, :]
            emf = 0.25 * (F1[I, J, 6] + F1[I, J - 1, 6] - F2[I, J, 5] - F2[I - 1, J, 5])
            f1b2 = 0.5 * (emf[:, :-1] + emf[:, 1:])          # x1 faces i in [0,n1], j in [0,n2)
            f2b1 = -0.5 * (emf[:-1, :] + emf[1:, :])         # x2 faces i in [0,n1), j in [0,n2]
            F1i = F1[NG:NG + n1 + 1, NG:NG + n2]
            F1i = torch.cat([F1i[..., :5], torch.zeros_like(f1b2)[..., None], f1b2[..., None], F1i[..., 7:]], -1)
            F2i = F2[NG:NG + n1, NG:NG + n2 + 1]
            F2i = torch.cat([F2i[..., :5], f2b1[..., None], torch.zeros_like(f2b1)[..., None], F2i[..., 7:]], -1)
        else:
            F1i = F1[NG:NG + n1 + 1, NG:NG + n2]
            F2i = F2[NG:NG + n1, NG:NG + n2 + 1]
        return -(F1i[1:] - F1i[:-1]) / self.dx1 - (F2i[:, 1:] - F2i[:, :-1]) / self.dx2

    # ---- inversion: Newton in (rho, u, u-tilde^i) with B = U[B] ----
    def invert(self, U, guess):
        B = U[..., 5:8]

        def U5(x, b):
            return self.flux(torch.cat([x, b]), 0)[:5]

        J5 = vmap(jacfwd(U5, argnums=0))
        x = guess[..., :5].reshape(-1, 5)
        b = B.reshape(-1, 3)
        target = U[..., :5].reshape(-1, 5)
        for _ in range(self.iters):
            f = vmap(U5)(x, b) - target
            x = x - torch.linalg.solve(J5(x, b), f.unsqueeze(-1)).squeeze(-1)
        return torch.cat([x.reshape(U.shape[:-1] + (5,)), B], -1)

    def step(self, p, dt):
        n1, n2 = self.n1, self.n2
        CUf, Cn, pi_, pb_, pf_ = RK[self.to]
        inner = (slice(NG, NG + n1), slice(NG, NG + n2))
        arrs = {"p": p}
        Ui = self.flux(p[inner], 0)
        ulast = torch.zeros_like(Ui)
        unew = torch.zeros_like(Ui)
        nto = len(CUf)
        for s in range(nto):
            pb = arrs[pb_[s]]
            dU = self.dU(pb)
            Uf = CUf[s][0] * Ui + CUf[s][1] * ulast + CUf[s][2] * dt * dU
            unew = unew + Cn[s][0] * Ui + Cn[s][1] * dt * dU + Cn[s][2] * Uf
            ulast = Uf
            X = unew if s == nto - 1 else Uf
            pf_in = self.invert(X, p[inner])
            full = p.clone() * 0
            full = full.index_put((torch.arange(NG, NG + n1)[:, None], torch.arange(NG, NG + n2)[None, :]), pf_in)
            arrs[pf_[s]] = self.bound(full)
        return arrs["p"]

    def magnetic_energy(self, p):
        ucon, ucov, bcon, bcov = self.state(p[NG:NG + self.n1, NG:NG + self.n2])
        return 0.5 * (bcon * bcov).sum() * self.dx1 * self.dx2

    def run(self, p0, dts):
        p = self.bound(p0)
        for dt in dts:
            p = self.step(p, float(dt))
        return p
