// harm_device.h -- per-zone / per-face FP64 physics of HARM's conservative
// step as __host__ __device__ functions (the GPU path runs them on gfx950).
//
// Every function keeps the floating-point operation order of the reference
// function it cites, and the library is compiled with -ffp-contract=off, so
// that a step on the GPU is bit-identical to the reference CPU step on the
// same inputs (IEEE add/mul/div/sqrt are correctly rounded on CDNA4).
//
// Geometry reaches these functions as register structs filled from the packed
// device tables (see harm_kernels.hip):
//   GeoU : the entries that get_state()/primtoflux()/vchar() read -- the upper
//          triangle of gcov (phys.c:633-652 lower() reads only that), gcon^{0mu},
//          the one diagonal gcon^{dd} of the flux direction, and gdet.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

// host+device: the same functions also build for the CPU (a host-only
// compile checks the adjoint's local derivatives against finite differences)
#define HARM_HD __host__ __device__ __forceinline__

namespace harm {

constexpr int NPR = 8;
constexpr int RHO = 0, UU = 1, U1 = 2, U2 = 3, U3 = 4, B1 = 5, B2 = 6, B3 = 7;
constexpr double NUMEPSILON = 2.2204460492503131e-16;

// kernel-side parameter block (by value in kernel arguments)
struct KParams {
  int n1, n2, N2M;
  int eomtype, lim, fluxmethod, fluxb;
  int bc[4];
  int fixupflux, poledeath, checksolution, eno;
  int boundsfile, extrap;  // bounds.*.c (HARM_BOUNDS_*), the outflow fill's EXTRAP
  const double* rscale;    // EXTRAP 2: rescale() factors [nz][8] (harm_set_radius)
  double gam, cour, defcon, safe, gammamax, gammafail, rhominlimit, uuminlimit;
  double dx1, dx2;
  // floor/failure accounting (diag_fixup), set only while it is on: dVF,
  // the zone states before fixup1zone, the conserved state each force-free
  // inversion took (its dissipation account), and the event slots of a substage
  double dVF;
  double* acct_pre;
  double* acct_uin;
  int* acct_flag;
  double* acct_val;
};

// The physics below is written once for a scalar type T: double for the
// step itself, and a value-plus-tangent type for the tangent-linear step
// (harm_tangent.h), whose values are the double step's to the bit.
template <class T>
struct GeoT {
  T g00, g01, g02, g03, g11, g12, g13, g22, g23, g33;  // gcov upper
  T n00, n01, n02, n03;                                // gcon^{0 mu}
  T ndd;                                               // gcon^{dir dir}
  T gdet;
};
using GeoU = GeoT<double>;

template <class T>
struct StateT {
  T ucon[4], ucov[4], bcon[4], bcov[4];
};
using State = StateT<double>;

template <class T>
HARM_HD T dmax(T a, T b) { return (a > b) ? a : b; }  // phys.c:1580
template <class T>
HARM_HD T dmin(T a, T b) { return (a > b) ? b : a; }  // phys.c:1587
template <class T>
HARM_HD T dsign(T a) {                                // phys.c:1562
  return (a > 0.0) ? T(1.0) : ((a < 0.0) ? T(-1.0) : T(0.0));
}

// phys.c:848-877 gamma_calc
template <class T>
HARM_HD T gamma_calc(const T* pr, const GeoT<T>& g) {
  T qsq = g.g11 * pr[U1] * pr[U1] + g.g22 * pr[U2] * pr[U2] + g.g33 * pr[U3] * pr[U3] +
               2. * (g.g12 * pr[U1] * pr[U2] + g.g13 * pr[U1] * pr[U3] + g.g23 * pr[U2] * pr[U3]);
  return sqrt(1. + qsq);
}

// phys.c:633-652 lower
template <class T>
HARM_HD void lower(const T* v, const GeoT<T>& g, T* w) {
  w[0] = g.g00 * v[0] + g.g01 * v[1] + g.g02 * v[2] + g.g03 * v[3];
  w[1] = g.g01 * v[0] + g.g11 * v[1] + g.g12 * v[2] + g.g13 * v[3];
  w[2] = g.g02 * v[0] + g.g12 * v[1] + g.g22 * v[2] + g.g23 * v[3];
  w[3] = g.g03 * v[0] + g.g13 * v[1] + g.g23 * v[2] + g.g33 * v[3];
}

// phys.c:716-727 get_state with ucon_calc_rel4vel :830-845 and bcon_calc :285-294
template <class T>
HARM_HD void get_state(const T* pr, const GeoT<T>& g, StateT<T>& q) {
  T alpha = 1. / sqrt(-g.n00);
  T beta1 = g.n01 * alpha * alpha;
  T beta2 = g.n02 * alpha * alpha;
  T beta3 = g.n03 * alpha * alpha;
  T gamma = gamma_calc(pr, g);
  q.ucon[0] = gamma / alpha;
  q.ucon[1] = pr[U1] - gamma * beta1 / alpha;
  q.ucon[2] = pr[U2] - gamma * beta2 / alpha;
  q.ucon[3] = pr[U3] - gamma * beta3 / alpha;
  lower(q.ucon, g, q.ucov);
  q.bcon[0] = pr[B1] * q.ucov[1] + pr[B2] * q.ucov[2] + pr[B3] * q.ucov[3];
  q.bcon[1] = (pr[B1] + q.bcon[0] * q.ucon[1]) / q.ucon[0];
  q.bcon[2] = (pr[B2] + q.bcon[0] * q.ucon[2]) / q.ucon[0];
  q.bcon[3] = (pr[B3] + q.bcon[0] * q.ucon[3]) / q.ucon[0];
  lower(q.bcon, g, q.bcov);
}

template <class T>
HARM_HD T dot4(const T* a, const T* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}

// phys.c:381-406 mhd_calc_0: T^dir_j, j=0..3 into mhd[0..3]
template <class T>
HARM_HD void mhd_calc(const KParams& P, const T* pr, int dir, const StateT<T>& q, T* mhd) {
  T r, u, Pg, w;
  if (P.eomtype == 0) {
    r = pr[RHO];
    u = pr[UU];
    Pg = (P.gam - 1.) * u;
    w = Pg + r + u;
  } else {
    r = u = Pg = w = 0.0;
  }
  (void)r;
  T bsq = dot4(q.bcon, q.bcov);
  T eta = w + bsq;
  T ptot = Pg + bsq * 0.5;
  T ed = eta * q.ucon[dir];
  T bd = q.bcon[dir];
#pragma unroll
  for (int j = 0; j < 4; j++) mhd[j] = ed * q.ucov[j] + ptot * ((dir == j) ? 1. : 0.) - bd * q.bcov[j];
}

// phys.c:16-56 primtoflux (massflux :59, PRIMMAXWELL dual Faraday :155-185,
// UtoU UNOTHING->UEVOLVE :219-265 with e[k]=gdet)
template <class T>
HARM_HD void primtoflux(const KParams& P, const T* pr, const StateT<T>& q, int dir, T gdet, T* F) {
  T fin[NPR];
  fin[RHO] = pr[RHO] * q.ucon[dir];
  mhd_calc(P, pr, dir, q, &fin[UU]);
  if (dir > 0) {
    T Bd = pr[B1 + dir - 1];
    fin[B1] = (pr[B1] * q.ucon[dir] - Bd * q.ucon[1]) / q.ucon[0];
    fin[B2] = (pr[B2] * q.ucon[dir] - Bd * q.ucon[2]) / q.ucon[0];
    fin[B3] = (pr[B3] * q.ucon[dir] - Bd * q.ucon[3]) / q.ucon[0];
  } else {
    fin[B1] = pr[B1];
    fin[B2] = pr[B2];
    fin[B3] = pr[B3];
  }
  if (P.eomtype == 0) fin[UU] += fin[RHO];
#pragma unroll
  for (int k = 0; k < NPR; k++) F[k] = fin[k] * gdet;
}

// vchar.c:37-135 vchar + :166-234 simplefast; returns 1 on the reference's
// fail() conditions (COEFF_NEG, COEFF_SUP, VCHAR_DISCR).  Written as
// selects between values computed on every path (the same operations as the
// reference's branches, so the same bits), which keeps the FP64 code of the
// calling kernels in straight runs the scheduler can interleave.
template <class T>
HARM_HD int vchar(const KParams& P, const T* pr, const StateT<T>& q, int dir, const GeoT<T>& g, T& vmax, T& vmin) {
  double gam = P.gam;
  T bsq = dot4(q.bcon, q.bcov);
  T rho = pr[RHO], u = pr[UU];
  T WW0 = rho + gam * u;
  const bool wpos = WW0 > 0.0;
  const T cs2w = gam * (gam - 1.) * u / WW0;
  const T WW = wpos ? WW0 : T(0.0), cs2 = wpos ? cs2w : T(0.0);
  T EF0 = bsq + WW;
  const bool epos = EF0 > 0.0;
  const T va2e = bsq / EF0;
  const T va2 = epos ? va2e : T(0.0);
  T cms2 = cs2 + va2 - cs2 * va2;
  const T rneg = cms2 / (cs2 + va2);
  const bool neg = cms2 < 0.0, negok = rneg > -NUMEPSILON * 100.0;
  const bool sup = cms2 > 1., supok = cms2 < 1.0 + NUMEPSILON * 100.0;
  int fail = (neg && !negok) || (!neg && sup && !supok);
  cms2 = (neg && negok) ? T(0.0) : cms2;
  cms2 = (!neg && sup && supok) ? T(1.0) : cms2;
  // simplefast: Acov=e_dir, Bcov=e_t raised with gcon (raise() keeps only the
  // selected gcon entries: the remaining terms are x*0)
  T Asq = g.ndd, Bsq = g.n00, AB = (dir == 1) ? g.n01 : g.n02;
  T Au = q.ucon[dir], Bu = q.ucon[0];
  T Au2 = Au * Au, Bu2 = Bu * Bu, AuBu = Au * Bu;
  T A = Bu2 - (Bsq + Bu2) * cms2;
  T B = 2. * (AuBu - (AB + AuBu) * cms2);
  T C = Au2 - (Asq + Au2) * cms2;
  T discr = B * B - 4. * A * C;
  // (-1e-10, 0): zeroed; below -1e-10: fail and zeroed; -1e-10 itself kept
  const bool dlow = discr < -1E-10, dsmall = (discr < 0.0) && (discr > -1E-10);
  fail |= dlow;
  discr = (dsmall || dlow) ? T(0.0) : discr;
  discr = sqrt(discr);
  T vm = -(-B + discr) / (2. * A);
  T vp = -(-B - discr) / (2. * A);
  const bool sw = vp < vm;
  vmax = sw ? vm : vp;
  vmin = sw ? vp : vm;
  return fail;
}

// fixup.c:1298-1354 limit_gamma: -1 changed, 1 failure, 0 untouched (as
// selects: the factor is formed on every path, applied where the reference
// applies it)
template <class T>
HARM_HD int limit_gamma(double gammamax, T* pr, const GeoT<T>& g) {
  if (gammamax <= 1.0) return 0;
  T gamma = gamma_calc(pr, g);
  const bool over = (gamma > gammamax && (gamma != 1.0));
  T pref = (gammamax * gammamax - 1.) / (gamma * gamma - 1.);
  const bool bad = pref < 0.0;
  T f = sqrt(pref);
  const bool apply = over && !bad;
  const T u1 = pr[U1] * f, u2 = pr[U2] * f, u3 = pr[U3] * f;
  pr[U1] = apply ? u1 : pr[U1];
  pr[U2] = apply ? u2 : pr[U2];
  pr[U3] = apply ? u3 : pr[U3];
  return over ? (bad ? 1 : -1) : 0;
}

// The face flux of fluxcalc() (step_ch.c:930-1065) from the reconstructed
// left/right states: BDIRCONT averaging of the normal field, limit_gamma on
// both states (H6), get_state, F and U of both states, vchar, cminmax_calc
// (:1095-1148) and HLL/LAXF (:1047-1054); dtij = cour dx / max(cmax, cmin)
// (:1062-1065).  p_l, p_r are modified (averaging, limit_gamma).
template <int DIR, class T>
HARM_HD void face_flux(const KParams& P, T* p_l, T* p_r, const GeoT<T>& g, T* Fo, T& dtij, int& fail) {
  {
    constexpr int kb = B1 + DIR - 1;
    T bavg = 0.5 * (p_l[kb] + p_r[kb]);
    p_l[kb] = bavg;
    p_r[kb] = bavg;
  }
  if (limit_gamma(P.gammamax, p_l, g) >= 1) fail = 1;
  if (limit_gamma(P.gammamax, p_r, g) >= 1) fail = 1;
  StateT<T> sl, sr;
  get_state(p_l, g, sl);
  get_state(p_r, g, sr);
  T F_l[NPR], F_r[NPR], U_l[NPR], U_r[NPR];
  primtoflux(P, p_l, sl, DIR, g.gdet, F_l);
  primtoflux(P, p_r, sr, DIR, g.gdet, F_r);
  primtoflux(P, p_l, sl, 0, g.gdet, U_l);
  primtoflux(P, p_r, sr, 0, g.gdet, U_r);
  T cmax_l, cmin_l, cmax_r, cmin_r;
  if (vchar(P, p_l, sl, DIR, g, cmax_l, cmin_l)) fail = 1;
  if (vchar(P, p_r, sr, DIR, g, cmax_r, cmin_r)) fail = 1;
  T lmin = dmin(cmin_l, cmin_r), lmax = dmax(cmax_l, cmax_r);
  T cmin = fabs(dmax(T(0.), -lmin)), cmax = fabs(dmax(T(0.), lmax));
  T ctop = dmax(cmax, cmin);
  double hllfrac = (P.fluxmethod == 1) ? 1.0 : 0.0;
  double laxffrac = (P.fluxmethod == 0) ? 1.0 : 0.0;
  const bool spread = (cmax + cmin != 0.0);
#pragma unroll
  for (int k = 0; k < NPR; k++) {
    const T lax = (0.5 * (F_l[k] + F_r[k] - ctop * (U_r[k] - U_l[k])));
    const T mix = hllfrac * ((cmax * F_l[k] + cmin * F_r[k] - cmax * cmin * (U_r[k] - U_l[k])) / (cmax + cmin)) +
                       laxffrac * lax;
    Fo[k] = spread ? mix : lax;
  }
  T cm = dmax(cmax, cmin);
  dtij = P.cour * ((DIR == 1) ? P.dx1 : P.dx2) / cm;
  if (!(dtij < 1.e9)) dtij = 1.e9;
}

// source_conn (phys.c:475-619, WITHGDET): dU[UU+l] = gdet sum_jk T^j_k
// Gamma^k_{lj} with the connection stored as conn[k][l][j]; element e of the
// zone's 64 is cn(e) (an accessor: the zone's own row, or on the device the
// component-major table)
template <class T, class CN>
HARM_HD void source_conn_at(const KParams& P, const T* pr, const GeoT<T>& g, const CN& cn, T* dU) {
  StateT<T> qb;
  get_state(pr, g, qb);
  T mhd[4][4];
#pragma unroll
  for (int jj = 0; jj < 4; jj++) mhd_calc(P, pr, jj, qb, mhd[jj]);
  T dUc[4];
#pragma unroll
  for (int l = 0; l < 4; l++) {
    T acc = 0;
#pragma unroll
    for (int jj = 0; jj < 4; jj++)
#pragma unroll
      for (int kk = 0; kk < 4; kk++) acc += mhd[jj][kk] * cn((kk * 4 + l) * 4 + jj);
    dUc[l] = acc;
  }
  dU[RHO] = g.gdet * 0.;
  dU[UU] = g.gdet * dUc[0];
  dU[U1] = g.gdet * dUc[1];
  dU[U2] = g.gdet * dUc[2];
  dU[U3] = g.gdet * dUc[3];
  dU[B1] = g.gdet * 0.;
  dU[B2] = g.gdet * 0.;
  dU[B3] = g.gdet * 0.;
}

// the same with element e at cn[e * cs] (cs = 1: the zone's own row of 64)
template <class T>
HARM_HD void source_conn(const KParams& P, const T* pr, const GeoT<T>& g, const T* cn, T* dU, size_t cs = 1) {
  source_conn_at(P, pr, g, [&](int e) { return cn[(size_t)e * cs]; }, dU);
}

// conn[k][l][j] is symmetric in (l, j) (set_grid() computes it so, to the
// bit): the device table keeps each k's ten entries with l <= j, component-
// major ([NCONN][zones]).  Element e = (k*4 + l)*4 + j of the full table is
// entry conn_sym(e) there.
constexpr int NCONN = 40;
HARM_HD constexpr int conn_sym(int e) {
  const int k = e >> 4, l = (e >> 2) & 3, j = e & 3;
  const int a = (l < j) ? l : j, b = (l < j) ? j : l;
  return k * 10 + a * 4 - (a * (a - 1)) / 2 + (b - a);
}

// source_conn from the symmetric component-major table (cs = the number of zones)
template <class T>
HARM_HD void source_conn_sym(const KParams& P, const T* pr, const GeoT<T>& g, const T* cn, T* dU, size_t cs) {
  source_conn_at(P, pr, g, [&](int e) { return cn[(size_t)conn_sym(e) * cs]; }, dU);
}

// ---------------- reconstruction (interp.c) ----------------

// interp.h:13-14
template <class T>
HARM_HD T minmod(T a, T b) {
  const T m = (fabs(a) < fabs(b)) ? (a) : (b);
  return ((a) * (b) <= 0) ? T(0.0) : m;
}

// interp.c:106-149 slope_lim_3points
template <class T>
HARM_HD T slope3(int lim, T yl, T yc, T yr) {
  if (lim == 3) {  // MC
    T Dqm = 2.0 * (yc - yl), Dqp = 2.0 * (yr - yc), Dqc = 0.5 * (yr - yl);
    return minmod(Dqc, minmod(Dqm, Dqp));
  } else if (lim == 1) {  // VANL
    T Dqm = (yc - yl), Dqp = (yr - yc), s = Dqm * Dqp;
    const T v = (2.0 * s / (Dqm + Dqp));
    return (s <= 0.) ? T(0.) : v;
  } else if (lim == 2) {  // MINM
    return minmod(yc - yl, yr - yc);
  } else if (lim == 101) return 0.5 * (yr - yl);
  else if (lim == 102) return (yc - yl);
  else if (lim == 103) return (yr - yc);
  return T(0.0);  // DONOR
}

// interp.c:157-205 csslope
HARM_HD double csslope(int kk, double x1, double x2, double x3, double x4, double x5) {
  double Dqp = (x4 - x3), Dqpp = (x5 - x4), Dqm = (x3 - x2), Dqmm = (x2 - x1);
  double Dqc = 0.5 * (x4 - x2), Dqcm = 0.5 * (x3 - x1), Dqcp = 0.5 * (x5 - x3);
  double s = 0.5 * (dsign(Dqp) + dsign(Dqm));
  double sp = 0.5 * (dsign(Dqpp) + dsign(Dqp));
  double sm = 0.5 * (dsign(Dqm) + dsign(Dqmm));
  double alpha;
  if (kk == RHO) alpha = 2.0;
  else if ((kk >= U1) && (kk <= U3)) alpha = 1.25;
  else if (kk == UU) alpha = 1.0;
  else alpha = 1.25;
  double Dql = alpha * dmin(fabs(Dqp), fabs(Dqm));
  double Dqlp = alpha * dmin(fabs(Dqpp), fabs(Dqp));
  double Dqlm = alpha * dmin(fabs(Dqm), fabs(Dqmm));
  double Dqbp = sp * dmin(Dqlp, fabs(Dqcp));
  double Dqbm = sm * dmin(Dqlm, fabs(Dqcm));
  return s * dmin(fabs((1.3333333333333333333333333333333333) * Dqc -
                       (0.1666666666666666666666666666666666) * (Dqbp + Dqbm)),
                  Dql);
}

// interp.c:449-538 para4 (PARA2LIM=MC): left/right edge values of the middle
// zone.  Every branch of the reference is a select between values formed on
// every path (same operations, same bits), so a kernel's 16 reconstructions
// form one straight run of independent FP64 work.
template <class T>
HARM_HD void para4(T y0, T y1, T y2, T y3, T y4, T& lout, T& rout) {
  T y[5] = {y0, y1, y2, y3, y4}, dq[4];
#pragma unroll
  for (int i = 1; i < 4; i++) {
    T Dqm = 2.0 * (y[i] - y[i - 1]);
    T Dqp = 2.0 * (y[i + 1] - y[i]);
    T Dqc = 0.5 * (y[i + 1] - y[i - 1]);
    T s = Dqm * Dqp;
    const T m = dmin(fabs(Dqc), dmin(fabs(Dqm), fabs(Dqp))) * dsign(Dqc);
    dq[i] = (s <= 0.) ? T(0.) : m;
  }
  T l = 0.5 * (y[2] + y[1]) - (dq[2] - dq[1]) / 8.0;
  T r = 0.5 * (y[3] + y[2]) - (dq[3] - dq[2]) / 8.0;
  l = dmax(dmin(y[2], y[1]), l);
  l = dmin(dmax(y[2], y[1]), l);
  r = dmax(dmin(y[2], y[3]), r);
  r = dmin(dmax(y[2], y[3]), r);
  T qa = (r - y[2]) * (y[2] - l);
  T qd = (r - l);
  T qe = 8.0 * (y[2] - 0.5 * (l + r));
  const T l3 = 3.0 * y[2] - 2.0 * r;
  const T l1 = (qd * (qd - qe) < 0.0) ? l3 : l;
  const T r3 = 3.0 * y[2] - 2.0 * l1;
  const T r1 = (qd * (qd + qe) < 0.0) ? r3 : r;
  const bool flat = qa <= 0.;
  lout = flat ? y[2] : l1;
  rout = flat ? y[2] : r1;
}

// ---------------- GRMHD inversion (utoprim_2d_final.c) ----------------

struct U2P {
  double Bsq, QdotBsq, Qtsq, Qdotn, D, wglobal, gam, gammafail;
};

// utoprim_2d_final.c:356-370 vsq_calc
HARM_HD double vsq_calc(const U2P& s, double W) {
  double Wsq = W * W;
  double Xsq = (s.Bsq + W) * (s.Bsq + W);
  return ((Wsq * s.Qtsq + s.QdotBsq * (s.Bsq + 2. * W)) / (Wsq * Xsq));
}

// utoprim_2d_final.c:1082-1155 func_vsq2 (+ pressure_W_vsq :447, dpdW_calc_vsq :465, dpdvsq_calc :479)
HARM_HD void func_vsq2(const U2P& s, double W, double vsq, double& dx0, double& dx1,
                                          double& f, double& df) {
  double gam = s.gam, Wsq = W * W, gtmp = 1. - vsq;
  double p_tmp = (gam - 1.) * (W * gtmp - s.D * sqrt(gtmp)) / gam;
  double dPdW = (gam - 1.) * (1. - vsq) / gam;
  double dPdvsq = (gam - 1.) * (0.5 * s.D / sqrt(1. - vsq) - W) / gam;
  double Bsq = s.Bsq, QdotBsq = s.QdotBsq;
  double t2 = -0.5 * Bsq + dPdvsq;
  double t3 = Bsq + W;
  double t4 = t3 * t3;
  double t9 = 1 / Wsq;
  double t11 = s.Qtsq - vsq * t4 + QdotBsq * (Bsq + 2.0 * W) * t9;
  double t16 = QdotBsq * t9;
  double t18 = -s.Qdotn - 0.5 * Bsq * (1.0 + vsq) + 0.5 * t16 - W + p_tmp;
  double t21 = 1 / t3;
  double t23 = 1 / W;
  double t24 = t16 * t23;
  double t25 = -1.0 + dPdW - t24;
  double t35 = t25 * t3 + (Bsq - 2.0 * dPdvsq) * (QdotBsq + vsq * Wsq * W) * t9 * t23;
  double t36 = 1 / t35;
  dx0 = -(t2 * t11 + t4 * t18) * t21 * t36;
  double t40 = (vsq + t24) * t3;
  dx1 = -(-t25 * t11 - 2.0 * t40 * t18) * t21 * t36;
  df = -t11 * t11 - t18 * t18;
  f = -0.5 * (df);
}

// utoprim_2d_final.c:644-1012 general_newton_raphson without line search
// (USE_LINE_SEARCH=-2), NEWCONVERGE=1, EXTRA_NEWT_ITER=2, MAX_NEWT_ITER=200
HARM_HD int newton2d(const U2P& s, double& x0, double& x1, int& niter) {
  const double NEWT_TOL = 1.0e-15, MIN_NEWT_TOL = 1.0e-10;
  double VSQ_TOO_BIG = (1.0 - 1.0 / ((s.gammafail - 1.0) + 1.0));
  double GAMMASQ_TOO_BIG = (s.gammafail * s.gammafail);
  double errx = 1., f = 1., df = 1., dx0, dx1;
  int n_iter = 0, i_extra = 0, doing_extra = 0;
  bool keep = true;
  while (keep) {
    func_vsq2(s, x0, x1, dx0, dx1, f, df);
    double xo0 = x0;
    x0 += dx0;
    x1 += dx1;
    errx = (x0 == 0.) ? fabs(dx0) : fabs(dx0 / x0);
    x0 = fabs(x0);
    x0 = (x0 / s.wglobal > GAMMASQ_TOO_BIG) ? xo0 : x0;
    x1 = (x1 < 0.) ? 0.0 : x1;
    x1 = (x1 > 1.) ? VSQ_TOO_BIG : x1;
    if ((fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) doing_extra = 1;
    if (doing_extra == 1) i_extra++;
    if (((fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) || (i_extra > 2) || (n_iter >= (200 - 1))) keep = false;
    n_iter++;
  }
  niter = n_iter;
  if (!isfinite(f) || !isfinite(df) || !isfinite(x0) || !isfinite(x1)) return 1;
  if (fabs(errx) > MIN_NEWT_TOL) return 1;
  return 0;
}

}  // namespace harm
