// harm_adjoint.h -- reverse-mode (vector-Jacobian) counterparts of the
// per-zone / per-face physics in harm_device.h, for the discrete adjoint of
// the conservative step.  Each *_rev function takes the primal inputs (it
// recomputes what it needs), the adjoints of the primal outputs, and ADDS the
// adjoints of the primal inputs into the given accumulators.  Piecewise
// definitions (limiters, max/min, clamps) are differentiated on the branch the
// primal took.  Host+device, so the same code is checked against finite
// differences on the CPU.
#pragma once
#include "harm_device.h"

namespace harm {

struct StateBar {
  double ucon[4], ucov[4], bcon[4], bcov[4];
};

HARM_HD void statebar_zero(StateBar& b) {
#pragma unroll
  for (int m = 0; m < 4; m++) b.ucon[m] = b.ucov[m] = b.bcon[m] = b.bcov[m] = 0.;
}

// w = G v with G symmetric (lower(), phys.c:633-652): vb += G wb
HARM_HD void lower_rev(const GeoU& g, const double* wb, double* vb) {
  vb[0] += g.g00 * wb[0] + g.g01 * wb[1] + g.g02 * wb[2] + g.g03 * wb[3];
  vb[1] += g.g01 * wb[0] + g.g11 * wb[1] + g.g12 * wb[2] + g.g13 * wb[3];
  vb[2] += g.g02 * wb[0] + g.g12 * wb[1] + g.g22 * wb[2] + g.g23 * wb[3];
  vb[3] += g.g03 * wb[0] + g.g13 * wb[1] + g.g23 * wb[2] + g.g33 * wb[3];
}

// get_state (ucon from the relative 4-velocity, ucov, bcon, bcov); qb is
// consumed (modified), pb accumulates.
HARM_HD void get_state_rev(const double* pr, const GeoU& g, const State& q, StateBar& qb, double* pb) {
  const double alpha = 1. / sqrt(-g.n00);
  const double beta[4] = {0., g.n01 * alpha * alpha, g.n02 * alpha * alpha, g.n03 * alpha * alpha};
  const double gamma = q.ucon[0] * alpha;  // ucon0 = gamma/alpha
  // bcov = G bcon
  lower_rev(g, qb.bcov, qb.bcon);
  // bcon_i = (B_i + bcon0 ucon_i) / ucon0
  double bcon0b = qb.bcon[0];
#pragma unroll
  for (int i = 1; i < 4; i++) {
    const double t = qb.bcon[i] / q.ucon[0];
    pb[B1 + i - 1] += t;
    bcon0b += t * q.ucon[i];
    qb.ucon[i] += t * q.bcon[0];
    qb.ucon[0] -= t * q.bcon[i];
  }
  // bcon0 = B^i ucov_i
#pragma unroll
  for (int i = 1; i < 4; i++) {
    pb[B1 + i - 1] += bcon0b * q.ucov[i];
    qb.ucov[i] += bcon0b * pr[B1 + i - 1];
  }
  // ucov = G ucon
  lower_rev(g, qb.ucov, qb.ucon);
  // ucon0 = gamma/alpha, ucon_i = u_i - gamma beta_i / alpha
  double gammab = qb.ucon[0] / alpha;
#pragma unroll
  for (int i = 1; i < 4; i++) {
    pb[U1 + i - 1] += qb.ucon[i];
    gammab -= qb.ucon[i] * beta[i] / alpha;
  }
  // gamma = sqrt(1 + g_ij u^i u^j)
  const double qsqb = gammab * 0.5 / gamma;
  const double u1 = pr[U1], u2 = pr[U2], u3 = pr[U3];
  pb[U1] += qsqb * 2. * (g.g11 * u1 + g.g12 * u2 + g.g13 * u3);
  pb[U2] += qsqb * 2. * (g.g12 * u1 + g.g22 * u2 + g.g23 * u3);
  pb[U3] += qsqb * 2. * (g.g13 * u1 + g.g23 * u2 + g.g33 * u3);
}

// mhd_calc: T^dir_j (j=0..3) from Tb
HARM_HD void mhd_rev(const KParams& P, const double* pr, int dir, const State& q, const double* Tb, StateBar& qb,
                     double* pb) {
  double w = 0.;
  if (P.eomtype == 0) {
    const double Pg = (P.gam - 1.) * pr[UU];
    w = Pg + pr[RHO] + pr[UU];
  }
  const double bsq = dot4(q.bcon, q.bcov);
  const double eta = w + bsq;
  double etab = 0.;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    etab += Tb[j] * q.ucon[dir] * q.ucov[j];
    qb.ucon[dir] += Tb[j] * eta * q.ucov[j];
    qb.ucov[j] += Tb[j] * eta * q.ucon[dir];
    qb.bcon[dir] -= Tb[j] * q.bcov[j];
    qb.bcov[j] -= Tb[j] * q.bcon[dir];
  }
  const double ptotb = Tb[dir];
  const double bsqb = etab + 0.5 * ptotb;
  if (P.eomtype == 0) {
    const double wb = etab, Pgb = ptotb + wb;
    pb[RHO] += wb;
    pb[UU] += wb + (P.gam - 1.) * Pgb;
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    qb.bcon[j] += bsqb * q.bcov[j];
    qb.bcov[j] += bsqb * q.bcon[j];
  }
}

// primtoflux (dir 0 gives U) including the gdet factor and the rest-mass shift
HARM_HD void primtoflux_rev(const KParams& P, const double* pr, const State& q, int dir, double gdet,
                            const double* Fb, StateBar& qb, double* pb) {
  double fb[NPR];
#pragma unroll
  for (int k = 0; k < NPR; k++) fb[k] = Fb[k] * gdet;
  if (P.eomtype == 0) fb[RHO] += fb[UU];
  mhd_rev(P, pr, dir, q, &fb[UU], qb, pb);
  pb[RHO] += fb[RHO] * q.ucon[dir];
  qb.ucon[dir] += fb[RHO] * pr[RHO];
  if (dir > 0) {
    const double Bd = pr[B1 + dir - 1];
#pragma unroll
    for (int k = 1; k < 4; k++) {
      const double f = (pr[B1 + k - 1] * q.ucon[dir] - Bd * q.ucon[k]) / q.ucon[0];
      const double t = fb[B1 + k - 1] / q.ucon[0];
      pb[B1 + k - 1] += t * q.ucon[dir];
      qb.ucon[dir] += t * pr[B1 + k - 1];
      pb[B1 + dir - 1] -= t * q.ucon[k];
      qb.ucon[k] -= t * Bd;
      qb.ucon[0] -= t * f;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 3; k++) pb[B1 + k] += fb[B1 + k];
  }
}

// vchar + simplefast: adjoints of (vmax, vmin).  Returns 1 where the primal
// sits on a non-differentiable point (a clamped cms2 or a zero discriminant);
// the derivative there is taken as that of the clamped branch.
HARM_HD int vchar_rev(const KParams& P, const double* pr, const State& q, int dir, const GeoU& g, double vmaxb,
                      double vminb, StateBar& qb, double* pb) {
  const double gam = P.gam;
  const double bsq = dot4(q.bcon, q.bcov);
  const double rho = pr[RHO], u = pr[UU];
  double WW = rho + gam * u, cs2 = 0., EF, va2 = 0.;
  const bool wwpos = WW > 0.0;
  if (wwpos) cs2 = gam * (gam - 1.) * u / WW;
  else WW = 0.0;
  EF = bsq + WW;
  const bool efpos = EF > 0.0;
  if (efpos) va2 = bsq / EF;
  double cms2 = cs2 + va2 - cs2 * va2;
  bool clamped = false;
  if (cms2 < 0.0) { cms2 = 0.0; clamped = true; }
  if (cms2 > 1.) { cms2 = 1.0; clamped = true; }
  const double Asq = g.ndd, Bsq = g.n00, AB = (dir == 1) ? g.n01 : g.n02;
  const double Au = q.ucon[dir], Bu = q.ucon[0];
  const double Au2 = Au * Au, Bu2 = Bu * Bu, AuBu = Au * Bu;
  const double A = Bu2 - (Bsq + Bu2) * cms2;
  const double B = 2. * (AuBu - (AB + AuBu) * cms2);
  const double C = Au2 - (Asq + Au2) * cms2;
  double discr = B * B - 4. * A * C;
  bool dclamp = false;
  if (discr <= 0.0) { discr = 0.0; dclamp = true; }
  const double sd = sqrt(discr);
  const double vm = -(-B + sd) / (2. * A);
  const double vp = -(-B - sd) / (2. * A);
  double vmb, vpb;
  if (vp < vm) { vmb = vmaxb; vpb = vminb; }
  else { vpb = vmaxb; vmb = vminb; }
  double Bb = (vmb + vpb) / (2. * A);
  const double sdb = (vpb - vmb) / (2. * A);
  double Ab = -(vmb * vm + vpb * vp) / A;
  const double discrb = dclamp ? 0. : sdb * 0.5 / sd;
  Bb += 2. * B * discrb;
  Ab -= 4. * C * discrb;
  const double Cb = -4. * A * discrb;
  double Bub = Ab * (2. * Bu - 2. * Bu * cms2);
  double cms2b = -Ab * (Bsq + Bu2);
  const double AuBub = 2. * Bb * (1. - cms2);
  cms2b -= 2. * Bb * (AB + AuBu);
  double Aub = Cb * 2. * Au * (1. - cms2);
  cms2b -= Cb * (Asq + Au2);
  Aub += AuBub * Bu;
  Bub += AuBub * Au;
  qb.ucon[dir] += Aub;
  qb.ucon[0] += Bub;
  double cs2b = 0., va2b = 0.;
  if (!clamped) {
    cs2b = cms2b * (1. - va2);
    va2b = cms2b * (1. - cs2);
  }
  double bsqb = 0., WWb = 0.;
  if (efpos) {
    bsqb += va2b / EF;
    const double EFb = -va2b * bsq / (EF * EF);
    bsqb += EFb;
    WWb += EFb;
  }
  if (wwpos) {
    pb[UU] += cs2b * gam * (gam - 1.) / WW;
    WWb -= cs2b * cs2 / WW;
    pb[RHO] += WWb;
    pb[UU] += gam * WWb;
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    qb.bcon[j] += bsqb * q.bcov[j];
    qb.bcov[j] += bsqb * q.bcon[j];
  }
  return (clamped || dclamp) ? 1 : 0;
}

// The face flux of fluxcalc() (BDIRCONT averaging of the normal field,
// limit_gamma, states, fluxes, wave speeds, HLL/LAXF) as a function of the
// reconstructed left/right states p_l, p_r.  Fo_b is the adjoint of the face
// flux; plb, prb (16 values: left then right) accumulate.  fix_flux is not
// part of this function (the adjoint is defined without FIXUPFLUX).  Returns
// a count of non-smooth points met (limit_gamma active, clamps).
template <int DIR>
HARM_HD int face_flux_rev(const KParams& P, const double* pl_in, const double* pr_in, const GeoU& g,
                          const double* Fob, double* plb, double* prb) {
  int ns = 0;
  double p_l[NPR], p_r[NPR];
#pragma unroll
  for (int k = 0; k < NPR; k++) { p_l[k] = pl_in[k]; p_r[k] = pr_in[k]; }
  constexpr int kb = B1 + DIR - 1;
  const double bavg = 0.5 * (p_l[kb] + p_r[kb]);
  p_l[kb] = bavg;
  p_r[kb] = bavg;
  if (limit_gamma(P.gammamax, p_l, g) != 0) ns++;
  if (limit_gamma(P.gammamax, p_r, g) != 0) ns++;
  State sl, sr;
  get_state(p_l, g, sl);
  get_state(p_r, g, sr);
  double F_l[NPR], F_r[NPR], U_l[NPR], U_r[NPR];
  primtoflux(P, p_l, sl, DIR, g.gdet, F_l);
  primtoflux(P, p_r, sr, DIR, g.gdet, F_r);
  primtoflux(P, p_l, sl, 0, g.gdet, U_l);
  primtoflux(P, p_r, sr, 0, g.gdet, U_r);
  double cmax_l, cmin_l, cmax_r, cmin_r;
  vchar(P, p_l, sl, DIR, g, cmax_l, cmin_l);
  vchar(P, p_r, sr, DIR, g, cmax_r, cmin_r);
  const double lmin = dmin(cmin_l, cmin_r), lmax = dmax(cmax_l, cmax_r);
  const double cmin = fabs(dmax(0., -lmin)), cmax = fabs(dmax(0., lmax));
  const double ctop = dmax(cmax, cmin);
  double F_lb[NPR], F_rb[NPR], U_lb[NPR], U_rb[NPR];
  double cmaxb = 0., cminb = 0., ctopb = 0.;
  const bool hll = (P.fluxmethod == 1) && (cmax + cmin != 0.0);
  const double D = cmax + cmin;
#pragma unroll
  for (int k = 0; k < NPR; k++) {
    const double dUk = U_r[k] - U_l[k];
    if (hll) {
      const double N = cmax * F_l[k] + cmin * F_r[k] - cmax * cmin * dUk;
      const double Nb = Fob[k] / D;
      const double Db = -Fob[k] * (N / D) / D;
      F_lb[k] = cmax * Nb;
      F_rb[k] = cmin * Nb;
      U_rb[k] = -cmax * cmin * Nb;
      U_lb[k] = cmax * cmin * Nb;
      cmaxb += Nb * (F_l[k] - cmin * dUk) + Db;
      cminb += Nb * (F_r[k] - cmax * dUk) + Db;
    } else {
      F_lb[k] = 0.5 * Fob[k];
      F_rb[k] = 0.5 * Fob[k];
      U_rb[k] = -0.5 * ctop * Fob[k];
      U_lb[k] = 0.5 * ctop * Fob[k];
      ctopb -= 0.5 * Fob[k] * dUk;
    }
  }
  if (cmax > cmin) cmaxb += ctopb;
  else cminb += ctopb;
  const double lmaxb = (0. > lmax) ? 0. : cmaxb;
  const double lminb = (0. > -lmin) ? 0. : -cminb;
  double vmaxb_l = 0., vmaxb_r = 0., vminb_l = 0., vminb_r = 0.;
  if (cmax_l > cmax_r) vmaxb_l = lmaxb;
  else vmaxb_r = lmaxb;
  if (cmin_l > cmin_r) vminb_r = lminb;
  else vminb_l = lminb;
  StateBar bl, br;
  statebar_zero(bl);
  statebar_zero(br);
  double pbl[NPR] = {0., 0., 0., 0., 0., 0., 0., 0.}, pbr[NPR] = {0., 0., 0., 0., 0., 0., 0., 0.};
  ns += vchar_rev(P, p_l, sl, DIR, g, vmaxb_l, vminb_l, bl, pbl);
  ns += vchar_rev(P, p_r, sr, DIR, g, vmaxb_r, vminb_r, br, pbr);
  primtoflux_rev(P, p_l, sl, DIR, g.gdet, F_lb, bl, pbl);
  primtoflux_rev(P, p_l, sl, 0, g.gdet, U_lb, bl, pbl);
  primtoflux_rev(P, p_r, sr, DIR, g.gdet, F_rb, br, pbr);
  primtoflux_rev(P, p_r, sr, 0, g.gdet, U_rb, br, pbr);
  get_state_rev(p_l, g, sl, bl, pbl);
  get_state_rev(p_r, g, sr, br, pbr);
  // normal-field averaging
  const double bb = 0.5 * (pbl[kb] + pbr[kb]);
  pbl[kb] = bb;
  pbr[kb] = bb;
#pragma unroll
  for (int k = 0; k < NPR; k++) {
    plb[k] += pbl[k];
    prb[k] += pbr[k];
  }
  return ns;
}

// The slope dq of slope_lim_3points on the branch the primal takes, as the
// coefficients of dq = cl*yl + cc*yc + cr*yr (exact for the piecewise-linear
// limiters; the tangent of VANL's smooth branch).
HARM_HD void slope3_coef(int lim, double yl, double yc, double yr, double& cl, double& cc, double& cr) {
  cl = cc = cr = 0.;
  if (lim == 3) {  // MC: minmod(Dqc, minmod(Dqm, Dqp))
    const double Dqm = 2.0 * (yc - yl), Dqp = 2.0 * (yr - yc), Dqc = 0.5 * (yr - yl);
    int inner;  // 0: zero, 1: Dqm, 2: Dqp
    double iv;
    if (Dqm * Dqp <= 0) { inner = 0; iv = 0.0; }
    else if (fabs(Dqm) < fabs(Dqp)) { inner = 1; iv = Dqm; }
    else { inner = 2; iv = Dqp; }
    if (Dqc * iv <= 0) return;
    if (fabs(Dqc) < fabs(iv)) { cl = -0.5; cr = 0.5; return; }
    if (inner == 1) { cl = -2.0; cc = 2.0; }
    else if (inner == 2) { cc = -2.0; cr = 2.0; }
  } else if (lim == 1) {  // VANL: 2 Dqm Dqp / (Dqm + Dqp)
    const double Dqm = (yc - yl), Dqp = (yr - yc), s = Dqm * Dqp;
    if (s <= 0.) return;
    const double den = Dqm + Dqp;
    const double dm = 2.0 * Dqp * Dqp / (den * den), dp = 2.0 * Dqm * Dqm / (den * den);
    cl = -dm;
    cc = dm - dp;
    cr = dp;
  } else if (lim == 2) {  // MINM
    const double a = yc - yl, b = yr - yc;
    if (a * b <= 0) return;
    if (fabs(a) < fabs(b)) { cl = -1.; cc = 1.; }
    else { cc = -1.; cr = 1.; }
  } else if (lim == 101) { cl = -0.5; cr = 0.5; }
  else if (lim == 102) { cl = -1.; cc = 1.; }
  else if (lim == 103) { cc = -1.; cr = 1.; }
}

// Solve A^T x = b for a 5x5 A (row-major) by Gaussian elimination with
// partial pivoting; returns 1 if A is singular.  Every loop is unrolled and
// the pivot rows are exchanged by compare-and-swap, so all indices are
// compile-time constants and the system stays in registers.
HARM_HD int solve5_transpose(const double A[5][5], const double* b, double* x) {
  double M[5][6];
#pragma unroll
  for (int r = 0; r < 5; r++) {
#pragma unroll
    for (int c = 0; c < 5; c++) M[r][c] = A[c][r];
    M[r][5] = b[r];
  }
  int singular = 0;
#pragma unroll
  for (int c = 0; c < 5; c++) {
#pragma unroll
    for (int r = c + 1; r < 5; r++) {
      const bool sw = fabs(M[r][c]) > fabs(M[c][c]);
#pragma unroll
      for (int k = 0; k < 6; k++) {
        const double t = M[c][k];
        M[c][k] = sw ? M[r][k] : t;
        M[r][k] = sw ? t : M[r][k];
      }
    }
    if (M[c][c] == 0.0) singular = 1;
#pragma unroll
    for (int r = c + 1; r < 5; r++) {
      const double f = M[r][c] / M[c][c];
#pragma unroll
      for (int k = c; k < 6; k++) M[r][k] -= f * M[c][k];
    }
  }
#pragma unroll
  for (int r = 4; r >= 0; r--) {
    double s = M[r][5];
#pragma unroll
    for (int k = r + 1; k < 5; k++) s -= M[r][k] * x[k];
    x[r] = s / M[r][r];
  }
  return singular;
}

// Implicit-function adjoint of the inversion U -> p at its converged
// solution p: with J = dU/dp (U = primtoU(p) incl. gdet), returns
// Ub = J^{-T} pb.  J = [[A, C], [0, gdet I]] (U_B = gdet B), so
// A^T Ub[0..4] = pb[0..4] and Ub[B] = (pb[B] - C^T Ub[0..4]) / gdet.
HARM_HD int inversion_rev(const KParams& P, const double* p, const GeoU& g, const double* pbar, double* Ub) {
  State q;
  get_state(p, g, q);
  double A[5][5], C[5][3];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    double e[NPR] = {0., 0., 0., 0., 0., 0., 0., 0.};
    e[k] = 1.;
    StateBar qb;
    statebar_zero(qb);
    double row[NPR] = {0., 0., 0., 0., 0., 0., 0., 0.};
    primtoflux_rev(P, p, q, 0, g.gdet, e, qb, row);
    get_state_rev(p, g, q, qb, row);
#pragma unroll
    for (int m = 0; m < 5; m++) A[k][m] = row[m];
#pragma unroll
    for (int m = 0; m < 3; m++) C[k][m] = row[5 + m];
  }
  double x[5];
  if (solve5_transpose(A, pbar, x)) return 1;
#pragma unroll
  for (int m = 0; m < 5; m++) Ub[m] = x[m];
#pragma unroll
  for (int m = 0; m < 3; m++) {
    double s = pbar[5 + m];
#pragma unroll
    for (int k = 0; k < 5; k++) s -= C[k][m] * x[k];
    Ub[5 + m] = s / g.gdet;
  }
  return 0;
}

// source_conn: dU[UU+l] = gdet sum_jk T^j_k Gamma^k_{lj} (conn[k][l][j]);
// element e of the zone's 64 is cn(e)
template <class CN>
HARM_HD void source_rev_at(const KParams& P, const double* pr, const GeoU& g, const CN& cn, const double* dUb,
                           double* pb) {
  State q;
  get_state(pr, g, q);
  StateBar qb;
  statebar_zero(qb);
  for (int jj = 0; jj < 4; jj++) {
    double Tb[4];
    for (int kk = 0; kk < 4; kk++) {
      double s = 0.;
      for (int l = 0; l < 4; l++) s += dUb[UU + l] * cn((kk * 4 + l) * 4 + jj);
      Tb[kk] = g.gdet * s;
    }
    mhd_rev(P, pr, jj, q, Tb, qb, pb);
  }
  get_state_rev(pr, g, q, qb, pb);
}

// the same with element e at conn[e] (the zone's own row of 64)
HARM_HD void source_rev(const KParams& P, const double* pr, const GeoU& g, const double* conn, const double* dUb,
                        double* pb) {
  source_rev_at(P, pr, g, [&](int e) { return conn[e]; }, dUb, pb);
}

// d(0.5 b^2 gdet dV)/dp at a zone (the magnetic-energy functional)
HARM_HD void bsq_rev(const double* pr, const GeoU& g, double wgt, double* pb) {
  State q;
  get_state(pr, g, q);
  StateBar qb;
  statebar_zero(qb);
  const double bsqb = 0.5 * wgt;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    qb.bcon[j] += bsqb * q.bcov[j];
    qb.bcov[j] += bsqb * q.bcon[j];
  }
  get_state_rev(pr, g, q, qb, pb);
}

}  // namespace harm
