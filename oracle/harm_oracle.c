/*
 * harm_oracle.c -- CPU restatement of HARM's conservative timestep (FP64).
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
 * product.  Every function restates the reference function it cites with the
 * same floating-point operation order, so that with -ffp-contract=off it is
 * bit-identical to the compiled reference on the same inputs.
 *
 * Compile-time flags of the reference that this restatement fixes:
 *   WHICHVEL=VELREL4, WHICHEOM=WITHGDET, REMOVERESTMASSFROMUU=1, NPR=8,
 *   MAXWELL=PRIMMAXWELL, TRUEFAST=0, HYPERHLL=0, BDIRCONT=1, EVOLVECHECKS=1,
 *   RESCALEINTERP=0, LIMADJUST=FLUXADJUST=0, CORNGDETVERSION=1,
 *   FIXUPZONES=FIXUP1ZONE, UTOPRIMVERSION=UTOPRIM2DFINAL,
 *   UTOPRIMTRYAGAIN=0, rescaletype=1, EXTRAP=0; CHECKSOLUTION and DOENO are
 *   run-time parameters (oracle_params.checksolution, .eno).
 */
#include "harm_oracle.h"
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NPR 8
#define NDIM 4
#define NG 3
#define RHO 0
#define UU 1
#define U1 2
#define U2 3
#define U3 4
#define B1 5
#define B2 6
#define B3 7
#define FACE1 0
#define FACE2 1
#define CORN 2
#define CENT 3
#define NUMEPSILON (2.2204460492503131e-16)

/* u2p_defs.h:40-78 with PRECISEINVERSION=1 */
#define MAX_NEWT_ITER 200
#define NEWT_TOL 1.0e-15
#define MIN_NEWT_TOL 1.0e-10
#define EXTRA_NEWT_ITER 2
#define MAXNEGUTSQ (1E-10)
#define MAXNEGVSQ (1.0 - 1.0 / (MAXNEGUTSQ + 1.0))
#define FAIL_VAL 1.e30

/* failure codes (global.h:525-536) */
#define UTOPRIMNOFAIL 0
#define UTOPRIMFAILCONVUTSQ 2
#define UTOPRIMFAILCONVW 3
#define UTOPRIMFAILCONVUTSQ2 4
#define UTOPRIMFAILRHONEG 5
#define UTOPRIMFAILUNEG 6
#define UTOPRIMFAILRHOUNEG 7
#define NUMPFLAGS 5

struct oracle_ctx {
  oracle_params p;
  int N1M, N2M, nz;
  double *gcov, *gcon, *gdet, *conn, *floors;
  double *arr[12];
  int *pflag;
  double *gammacheck; /* fixup_checksolution's static gammacheck[N1+2][N2+2] */
  double *eno[3];     /* uenotmp0..2 (defs.h:80-82), DOENO only */
  double t, dt, tf;
  long nstroke;
  int failed;
  int acct;          /* conservation accounting on */
  int ffde_pr0_zero; /* force-free fixup1zone measures from pr0 = 0 (see or_acct_ffde_pr0_zero) */
  int finalstep;     /* the substage being advanced is the step's last */
  or_acct A;
  double dx3;
  double *frdot;     /* [n1][8] */
  double *lumvsr;    /* [n1] */
  int cur;           /* current (jcon, fcon) computed every step */
  double *cf;        /* cfaraday [nz][5][3] */
  double *fcon;      /* [nz][6] */
  double *jcon;      /* [nz][4] */
  long long calls;   /* current_calc_1's statics */
  double lastdt;
  double *panalytic; /* [nz][8], FIXED boundaries and bound_disk */
  double *omegaf;    /* [nz][8], omegafanalytic (component 0 used) */
  double *radius;    /* [nz], r of the zone centres (EXTRAP 2) */
  int disk_first;    /* bound_disk's static firsttime */
};

typedef struct {
  const double (*gcov)[NDIM];
  const double (*gcon)[NDIM];
  double g;
  double e[NPR];
  int i, j, p;
} geom_t;

typedef struct {
  double ucon[NDIM], ucov[NDIM], bcon[NDIM], bcov[NDIM];
} state_t;

static inline int zi(const oracle_ctx *c, int i, int j) {
  return (i + NG) * c->N2M + (j + NG);
}
#define PA(c, a, i, j) ((a) + (size_t)NPR * zi((c), (i), (j)))

/* phys.c:1562-1592 */
static inline double sign_(double a) { if (a > 0.0) return 1.0; else if (a < 0.0) return -1.0; else return 0.0; }
static inline double max_(double a, double b) { if (a > b) return a; else return b; }
static inline double min_(double a, double b) { if (a > b) return b; else return a; }
/* interp.h:13-14 */
#define MINMODB(a, b) ((fabs(a) < fabs(b)) ? (a) : (b))
#define MINMOD(a, b) (((a) * (b) <= 0) ? 0.0 : MINMODB(a, b))

/* phys.c:730-765 get_geometry + assign_eomfunc (WITHGDET: e[k]=gdet) */
static void get_geometry(const oracle_ctx *c, int i, int j, int loc, geom_t *g) {
  size_t z = (size_t)zi(c, i, j);
  g->gcov = (const double (*)[NDIM])(c->gcov + (z * 4 + loc) * 16);
  g->gcon = (const double (*)[NDIM])(c->gcon + (z * 4 + loc) * 16);
  g->g = c->gdet[z * 4 + loc];
  for (int k = 0; k < NPR; k++) g->e[k] = g->g;
  g->i = i; g->j = j; g->p = loc;
}

/* phys.c:848-877 gamma_calc */
static double gamma_calc(const double *pr, const geom_t *g) {
  const double (*gc)[NDIM] = g->gcov;
  double qsq = gc[1][1] * pr[U1] * pr[U1] + gc[2][2] * pr[U2] * pr[U2] +
               gc[3][3] * pr[U3] * pr[U3] +
               2. * (gc[1][2] * pr[U1] * pr[U2] + gc[1][3] * pr[U1] * pr[U3] +
                     gc[2][3] * pr[U2] * pr[U3]);
  return sqrt(1. + qsq);
}

/* phys.c:830-845 ucon_calc_rel4vel */
static void ucon_calc(const double *pr, const geom_t *g, double *ucon) {
  double alpha = 1. / sqrt(-g->gcon[0][0]);
  double beta[NDIM];
  for (int j = 1; j < NDIM; j++) beta[j] = g->gcon[0][j] * alpha * alpha;
  double gamma = gamma_calc(pr, g);
  ucon[0] = gamma / alpha;
  for (int j = 1; j < NDIM; j++) ucon[j] = pr[U1 + j - 1] - gamma * beta[j] / alpha;
}

/* phys.c:633-652 lower (upper triangle of gcov) */
static void lower(const double *ucon, const geom_t *g, double *ucov) {
  const double (*gc)[NDIM] = g->gcov;
  ucov[0] = gc[0][0] * ucon[0] + gc[0][1] * ucon[1] + gc[0][2] * ucon[2] + gc[0][3] * ucon[3];
  ucov[1] = gc[0][1] * ucon[0] + gc[1][1] * ucon[1] + gc[1][2] * ucon[2] + gc[1][3] * ucon[3];
  ucov[2] = gc[0][2] * ucon[0] + gc[1][2] * ucon[1] + gc[2][2] * ucon[2] + gc[2][3] * ucon[3];
  ucov[3] = gc[0][3] * ucon[0] + gc[1][3] * ucon[1] + gc[2][3] * ucon[2] + gc[3][3] * ucon[3];
}

/* phys.c:285-294 bcon_calc */
static void bcon_calc(const double *pr, const double *ucon, const double *ucov, double *bcon) {
  bcon[0] = pr[B1] * ucov[1] + pr[B2] * ucov[2] + pr[B3] * ucov[3];
  for (int j = 1; j <= 3; j++) bcon[j] = (pr[B1 - 1 + j] + bcon[0] * ucon[j]) / ucon[0];
}

/* phys.c:716-727 get_state */
static void get_state(const double *pr, const geom_t *g, state_t *q) {
  ucon_calc(pr, g, q->ucon);
  lower(q->ucon, g, q->ucov);
  bcon_calc(pr, q->ucon, q->ucov, q->bcon);
  lower(q->bcon, g, q->bcov);
}

#define DOT(a, b) ((a)[0] * (b)[0] + (a)[1] * (b)[1] + (a)[2] * (b)[2] + (a)[3] * (b)[3])

/* phys.c:381-406 mhd_calc_0 (T^dir_j) */
static void mhd_calc_0(const oracle_ctx *c, const double *pr, int dir, const state_t *q, double *mhd) {
  double r, u, P, w, bsq, eta, ptot;
  if (c->p.eomtype == OR_GRMHD) {
    r = pr[RHO]; u = pr[UU]; P = (c->p.gam - 1.) * u; w = P + r + u;
  } else {
    r = u = P = w = 0.0;
  }
  (void)r;
  bsq = DOT(q->bcon, q->bcov);
  eta = w + bsq;
  ptot = P + bsq * 0.5;
  for (int j = 0; j < NDIM; j++)
    mhd[j] = eta * q->ucon[dir] * q->ucov[j] + ptot * ((dir == j) ? 1. : 0.) - q->bcon[dir] * q->bcov[j];
}

/* phys.c:16-56 primtoflux, with massflux_calc :59, dualfullfaraday_calc
   PRIMMAXWELL :155-185 and UtoU(UNOTHING->UEVOLVE) :219-265 */
static void primtoflux(const oracle_ctx *c, const double *pr, const state_t *q, int dir,
                       const geom_t *g, double *flux) {
  double fin[NPR];
  fin[RHO] = pr[RHO] * q->ucon[dir];
  mhd_calc_0(c, pr, dir, q, &fin[UU]);
  if (dir > 0) {
    fin[B1] = (pr[B1] * q->ucon[dir] - pr[B1 + dir - 1] * q->ucon[1]) / q->ucon[0];
    fin[B2] = (pr[B2] * q->ucon[dir] - pr[B1 + dir - 1] * q->ucon[2]) / q->ucon[0];
    fin[B3] = (pr[B3] * q->ucon[dir] - pr[B1 + dir - 1] * q->ucon[3]) / q->ucon[0];
  } else {
    fin[B1] = pr[B1]; fin[B2] = pr[B2]; fin[B3] = pr[B3];
  }
  if (c->p.eomtype != OR_FFDE) fin[UU] += fin[RHO];
  for (int k = 0; k < NPR; k++) flux[k] = fin[k] * g->e[k];
}

/* vchar.c:166-234 simplefast, with raise() of unit covectors reduced to
   the gcon entries it selects (exact: the other terms are x*0 added to it) */
static int simplefast(int dir, const geom_t *g, const state_t *q, double cms2, double *vmin, double *vmax) {
  double Asq = g->gcon[dir][dir];
  double Bsq = g->gcon[0][0];
  double Au = q->ucon[dir];
  double Bu = q->ucon[0];
  double AB = g->gcon[0][dir];
  double Au2 = Au * Au, Bu2 = Bu * Bu, AuBu = Au * Bu;
  double A = Bu2 - (Bsq + Bu2) * cms2;
  double B = 2. * (AuBu - (AB + AuBu) * cms2);
  double C = Au2 - (Asq + Au2) * cms2;
  double discr = B * B - 4. * A * C;
  int fail = 0;
  if ((discr < 0.0) && (discr > -1E-10)) discr = 0.0;
  else if (discr < -1E-10) { fail = 1; discr = 0.; }
  discr = sqrt(discr);
  double vm = -(-B + discr) / (2. * A);
  double vp = -(-B - discr) / (2. * A);
  if (vp < vm) { *vmax = vm; *vmin = vp; } else { *vmax = vp; *vmin = vm; }
  return fail;
}

/* vchar.c:37-135 vchar (TRUEFAST=0, no smoother, no group velocity) */
static int vchar(const oracle_ctx *c, const double *pr, const state_t *q, int dir, const geom_t *g,
                 double *vmax, double *vmin) {
  double gam = c->p.gam;
  double bsq = DOT(q->bcon, q->bcov);
  double rho = pr[RHO], u = pr[UU];
  double WW = rho + gam * u, cs2, EF, va2, cms2;
  int fail = 0;
  if (WW > 0.0) cs2 = gam * (gam - 1.) * u / WW;
  else WW = cs2 = 0.0;
  EF = bsq + WW;
  if (EF > 0.0) va2 = bsq / EF;
  else EF = va2 = 0.0;
  cms2 = cs2 + va2 - cs2 * va2;
  if (cms2 < 0.0) {
    if (cms2 / (cs2 + va2) > -NUMEPSILON * 100.0) cms2 = 0.0;
    else fail = 1;
  }
  if (cms2 > 1.) {
    if (cms2 < 1.0 + NUMEPSILON * 100.0) cms2 = 1.0;
    else fail = 1;
  }
  if (fail) return 1;
  return simplefast(dir, g, q, cms2, vmin, vmax);
}

/* step_ch.c:1095-1148 cminmax_calc (HYPERHLL=0) */
static void cminmax_calc(double cmin_l, double cmin_r, double cmax_l, double cmax_r,
                         double *cmin, double *cmax, double *ctop) {
  double lmin = min_(cmin_l, cmin_r);
  double lmax = max_(cmax_l, cmax_r);
  lmin = fabs(max_(0., -lmin));
  lmax = fabs(max_(0., lmax));
  *cmin = lmin; *cmax = lmax; *ctop = max_(lmax, lmin);
}

/* fixup.c:1298-1354 limit_gamma; returns -1 if changed, 1 on failure */
static int limit_gamma(double gammamax, double *pr, const geom_t *g) {
  double realgammamax = gammamax;
  if (realgammamax <= 1.0) return 0;
  double gamma = gamma_calc(pr, g);
  if ((gamma > realgammamax && (gamma != 1.0))) {
    double pref = (realgammamax * realgammamax - 1.) / (gamma * gamma - 1.);
    if (pref < 0.0) return 1;
    double f = sqrt(pref);
    pr[U1] *= f; pr[U2] *= f; pr[U3] *= f;
    return -1;
  }
  return 0;
}

/* ---------------- conservation accounting ---------------- */

/* failure/floor categories (global.h:536-544) */
#define COUNTUTOPRIMFAILCONV 0
#define COUNTFLOORACT 1
#define COUNTLIMITGAMMAACT 2
#define COUNTUTOPRIMFAILRHONEG 4
#define COUNTUTOPRIMFAILUNEG 5
#define COUNTUTOPRIMFAILRHOUNEG 6
#define COUNTGAMMAPERC 7
#define COUNTUPERC 8

static int in_region(const int *ep, int i, int j) {
  return (i >= ep[OR_X1DN]) && (i <= ep[OR_X1UP]) && (j >= ep[OR_X2DN]) && (j <= ep[OR_X2UP]);
}

/* primtoU with returntype UDIAG: the geometry-free conserved quantities of
   phys.c:16-56 (rest mass kept in UU) times gdet (UtoU UNOTHING->UDIAG,
   phys.c:219-265) */
static void primtoU_diag(const oracle_ctx *c, const double *pr, const state_t *q, const geom_t *g, double *U) {
  double fin[NPR];
  fin[RHO] = pr[RHO] * q->ucon[0];
  mhd_calc_0(c, pr, 0, q, &fin[UU]);
  fin[B1] = pr[B1]; fin[B2] = pr[B2]; fin[B3] = pr[B3];
  for (int k = 0; k < NPR; k++) U[k] = fin[k] * g->g;
}

/* UtoU(UEVOLVE -> UDIAG) (phys.c:219-265): divide out e[k], add the rest
   mass back into UU (REMOVERESTMASSFROMUU, not force-free), times gdet */
static void evolve_to_diag(const oracle_ctx *c, const geom_t *g, const double *Uin, double *Uout) {
  double u[NPR];
  for (int k = 0; k < NPR; k++) u[k] = Uin[k] / g->e[k];
  if (c->p.eomtype != OR_FFDE) u[UU] += -u[RHO];
  for (int k = 0; k < NPR; k++) Uout[k] = u[k] * g->g;
}

/* fixup.c:174-229 diag_fixup: dVF (U(pr) - U(pr0)) in every region holding
   zone (i,j) of the geometry, on an accounting (final) substage */
static void diag_fixup(oracle_ctx *c, const double *pr0, const double *pr, const geom_t *g, int finalstep,
                       int whocalled) {
  if (!c->acct || !finalstep) return;
  for (int r = 0; r < OR_NREG; r++) {
    if (!in_region(c->A.enerpos[r], g->i, g->j)) continue;
    state_t q;
    double Ui[NPR], Uf[NPR];
    get_state(pr0, g, &q);
    primtoU_diag(c, pr0, &q, g, Ui);
    get_state(pr, g, &q);
    primtoU_diag(c, pr, &q, g, Uf);
    for (int k = 0; k < NPR; k++) {
      double ftemp = c->A.dVF * (Uf[k] - Ui[k]);
      c->A.fladdterms[r][whocalled][k] += ftemp;
      c->A.fladd[r][k] += ftemp;
    }
  }
}

/* limit_gamma with its accounting (fixup.c:1340-1350): every caller on the
   path passes finalstep=-1.0, which counts */
static int limit_gamma_acct(oracle_ctx *c, double gammamax, double *pr, const geom_t *g) {
  double pr0[NPR];
  if (c->acct) memcpy(pr0, pr, sizeof(pr0));
  int r = limit_gamma(gammamax, pr, g);
  if (r == -1) diag_fixup(c, pr0, pr, g, 1, COUNTLIMITGAMMAACT);
  return r;
}

/* diag.c:1231-1265 diag_source on the accounting substage (Dt == dt), with
   the cooling sources (sc >= 1) zero */
static void diag_source(oracle_ctx *c, const geom_t *g, const double *dUgeomsrc, double Dt) {
  for (int r = 0; r < OR_NREG; r++) {
    if (!in_region(c->A.enerpos[r], g->i, g->j)) continue;
    for (int sc = 0; sc < OR_NUMSOURCES; sc++) {
      double ftemp[NPR], fdiag[NPR];
      for (int k = 0; k < NPR; k++) ftemp[k] = Dt * c->A.dVF * (sc == 0 ? dUgeomsrc[k] : 0.0);
      evolve_to_diag(c, g, ftemp, fdiag);
      for (int k = 0; k < NPR; k++) {
        c->A.sourceaddterms[r][sc][k] += fdiag[k];
        c->A.sourceadd[r][k] += fdiag[k];
        if (k == UU && r == 0) c->lumvsr[g->i] += fdiag[k];
      }
    }
  }
}

/* ---------------- reconstruction (interp.c) ---------------- */

/* interp.c:106-149 slope_lim_3points */
static double slope_lim_3points(int lim, double yl, double yc, double yr) {
  double Dqm, Dqp, Dqc, s;
  if (lim == OR_MC) {
    Dqm = 2.0 * (yc - yl); Dqp = 2.0 * (yr - yc); Dqc = 0.5 * (yr - yl);
    return MINMOD(Dqc, MINMOD(Dqm, Dqp));
  } else if (lim == OR_VANL) {
    Dqm = (yc - yl); Dqp = (yr - yc); s = Dqm * Dqp;
    if (s <= 0.) return 0.;
    return (2.0 * s / (Dqm + Dqp));
  } else if (lim == OR_MINM) {
    Dqm = (yc - yl); Dqp = (yr - yc);
    return MINMOD(Dqm, Dqp);
  } else if (lim == OR_NLIMCENT) return 0.5 * (yr - yl);
  else if (lim == OR_NLIMUP) return (yc - yl);
  else if (lim == OR_NLIMDOWN) return (yr - yc);
  return 0.0; /* DONOR */
}

/* interp.c:157-205 csslope */
static double csslope(int kk, double x1, double x2, double x3, double x4, double x5) {
  double s, sm, sp, Dql, Dqlm, Dqlp, Dqp, Dqpp, Dqm, Dqmm, Dqc, Dqcm, Dqcp, Dqbp, Dqbm, alpha;
  Dqp = (x4 - x3); Dqpp = (x5 - x4); Dqm = (x3 - x2); Dqmm = (x2 - x1);
  Dqc = 0.5 * (x4 - x2); Dqcm = 0.5 * (x3 - x1); Dqcp = 0.5 * (x5 - x3);
  s = 0.5 * (sign_(Dqp) + sign_(Dqm));
  sp = 0.5 * (sign_(Dqpp) + sign_(Dqp));
  sm = 0.5 * (sign_(Dqm) + sign_(Dqmm));
  if (kk == RHO) alpha = 2.0;
  else if ((kk >= U1) && (kk <= U3)) alpha = 1.25;
  else if (kk == UU) alpha = 1.0;
  else if ((kk >= B1) && (kk <= B3)) alpha = 1.25;
  else alpha = 2.0;
  Dql = alpha * min_(fabs(Dqp), fabs(Dqm));
  Dqlp = alpha * min_(fabs(Dqpp), fabs(Dqp));
  Dqlm = alpha * min_(fabs(Dqm), fabs(Dqmm));
  Dqbp = sp * min_(Dqlp, fabs(Dqcp));
  Dqbm = sm * min_(Dqlm, fabs(Dqcm));
  /* FOURTHIRD, SIXTH are long double constants (global.h:54-55): the
     difference is formed in long double and rounded to double by fabs() */
  return s * min_(fabs((double)((1.3333333333333333333333333333333333L) * Dqc -
                                (0.1666666666666666666666666666666666L) * (Dqbp + Dqbm))), Dql);
}

/* interp.c:449-538 para4 with PARA2LIM=MC */
static void para4(double x1, double x2, double x3, double x4, double x5, double *lout, double *rout) {
  double y[5], dq[5];
  double Dqm, Dqc, Dqp, aDqm, aDqp, aDqc, s, l, r, qa, qd, qe;
  y[0] = x1; y[1] = x2; y[2] = x3; y[3] = x4; y[4] = x5;
  for (int i = 1; i < 4; i++) {
    Dqm = 2.0 * (y[i] - y[i - 1]);
    Dqp = 2.0 * (y[i + 1] - y[i]);
    Dqc = 0.5 * (y[i + 1] - y[i - 1]);
    aDqm = fabs(Dqm); aDqp = fabs(Dqp); aDqc = fabs(Dqc);
    s = Dqm * Dqp;
    if (s <= 0.) dq[i] = 0.;
    else dq[i] = min_(aDqc, min_(aDqm, aDqp)) * sign_(Dqc);
  }
  l = 0.5 * (y[2] + y[1]) - (dq[2] - dq[1]) / 8.0;
  r = 0.5 * (y[3] + y[2]) - (dq[3] - dq[2]) / 8.0;
  l = max_(min_(y[2], y[1]), l);
  l = min_(max_(y[2], y[1]), l);
  r = max_(min_(y[2], y[3]), r);
  r = min_(max_(y[2], y[3]), r);
  qa = (r - y[2]) * (y[2] - l);
  qd = (r - l);
  qe = 8.0 * (y[2] - 0.5 * (l + r));
  if (qa <= 0.) { l = y[2]; r = y[2]; }
  else {
    if (qd * (qd - qe) < 0.0) l = 3.0 * y[2] - 2.0 * r;
    if (qd * (qd + qe) < 0.0) r = 3.0 * y[2] - 2.0 * l;
  }
  *lout = l; *rout = r;
}

/* interp.c:8-44 slope_lim (LIMADJUST=0, WHICHPARA=PARA4) */
void or_slope_lim(int lim, int k, double yll, double yl, double yc, double yr, double yrr,
                  double *dq, double *left, double *right) {
  if (lim == OR_PARA) para4(yll, yl, yc, yr, yrr, left, right);
  else if (lim == OR_CSSLOPE) *dq = csslope(k, yll, yl, yc, yr, yrr);
  else *dq = slope_lim_3points(lim, yl, yc, yr);
}

/* ---------------- fluxes (step_ch.c) ---------------- */

/* step_ch.c:741-1082 fluxcalc */
int or_fluxcalc(oracle_ctx *c, int pr_sel, int dir, double *ndt) {
  const double *pr = c->arr[pr_sel];
  double *F = c->arr[dir == 1 ? OR_F1 : OR_F2];
  double *dq = c->arr[dir == 1 ? OR_DQ1 : OR_DQ2];
  double *pleft = c->arr[OR_PLEFT], *pright = c->arr[OR_PRIGHT];
  int idel = (dir == 1), jdel = (dir == 2), face = (dir == 1) ? FACE1 : FACE2;
  int n1 = c->p.n1, n2 = c->p.n2, lim = c->p.lim;
  double dxd = (dir == 1) ? c->p.dx1 : c->p.dx2;
  double hllfrac = (c->p.fluxmethod == OR_HLL) ? 1.0 : 0.0;
  double laxffrac = (c->p.fluxmethod == OR_LAXF) ? 1.0 : 0.0;
  int fail = 0;

  /* DQZLOOP PLOOP slope_lim (step_ch.c:812-823) */
#pragma omp parallel for schedule(static)
  for (int i = -1; i <= n1; i++)
    for (int j = -1; j <= n2; j++)
      for (int k = 0; k < NPR; k++)
        or_slope_lim(lim, k, PA(c, pr, i - 2 * idel, j - 2 * jdel)[k], PA(c, pr, i - idel, j - jdel)[k],
                     PA(c, pr, i, j)[k], PA(c, pr, i + idel, j + jdel)[k], PA(c, pr, i + 2 * idel, j + 2 * jdel)[k],
                     &PA(c, dq, i, j)[k], &PA(c, pleft, i, j)[k], &PA(c, pright, i, j)[k]);

  double mindt = 1.e9;
  /* FZLOOP(-jdel,-idel) (step_ch.c:836-1069) */
#pragma omp parallel for schedule(static) reduction(min : mindt) reduction(| : fail) if (!c->acct)
  for (int i = -jdel; i <= n1; i++)
    for (int j = -idel; j <= n2; j++) {
      double p_l[NPR], p_r[NPR], F_l[NPR], F_r[NPR], U_l[NPR], U_r[NPR];
      double cmax_l, cmax_r, cmin_l, cmin_r, cmax, cmin, ctop, dtij;
      state_t sl, sr;
      geom_t g;
      for (int k = 0; k < NPR; k++) {
        if (lim != OR_PARA) {
          p_l[k] = PA(c, pr, i - idel, j - jdel)[k] + 0.5 * PA(c, dq, i - idel, j - jdel)[k];
          p_r[k] = PA(c, pr, i, j)[k] - 0.5 * PA(c, dq, i, j)[k];
        } else {
          p_l[k] = PA(c, pright, i - idel, j - jdel)[k];
          p_r[k] = PA(c, pleft, i, j)[k];
        }
      }
      get_geometry(c, i, j, face, &g);
      if (dir == 1) p_l[B1] = p_r[B1] = 0.5 * (p_l[B1] + p_r[B1]);
      else p_l[B2] = p_r[B2] = 0.5 * (p_l[B2] + p_r[B2]);
      if (limit_gamma_acct(c, c->p.gammamax, p_l, &g) >= 1) fail |= 1;
      if (limit_gamma_acct(c, c->p.gammamax, p_r, &g) >= 1) fail |= 1;
      get_state(p_l, &g, &sl);
      get_state(p_r, &g, &sr);
      primtoflux(c, p_l, &sl, dir, &g, F_l);
      primtoflux(c, p_r, &sr, dir, &g, F_r);
      primtoflux(c, p_l, &sl, 0, &g, U_l);
      primtoflux(c, p_r, &sr, 0, &g, U_r);
      if (vchar(c, p_l, &sl, dir, &g, &cmax_l, &cmin_l)) fail |= 1;
      if (vchar(c, p_r, &sr, dir, &g, &cmax_r, &cmin_r)) fail |= 1;
      cminmax_calc(cmin_l, cmin_r, cmax_l, cmax_r, &cmin, &cmax, &ctop);
      double *Fz = PA(c, F, i, j);
      if (cmax + cmin != 0.0) {
        for (int k = 0; k < NPR; k++)
          Fz[k] = hllfrac * ((cmax * F_l[k] + cmin * F_r[k] - cmax * cmin * (U_r[k] - U_l[k])) / (cmax + cmin)) +
                  laxffrac * (0.5 * (F_l[k] + F_r[k] - ctop * (U_r[k] - U_l[k])));
      } else {
        for (int k = 0; k < NPR; k++) Fz[k] = (0.5 * (F_l[k] + F_r[k] - ctop * (U_r[k] - U_l[k])));
      }
      cmax = max_(cmax, cmin);
      dtij = c->p.cour * dxd / cmax;
      if (dtij < mindt) mindt = dtij;
    }
  *ndt = mindt;
  if (fail) c->failed = 1;
  return fail;
}

/* fixup.c:1357-1410 fix_flux (single rank: mycpupos=0, ncpux=1) */
void or_fix_flux(oracle_ctx *c) {
  double *F1 = c->arr[OR_F1], *F2 = c->arr[OR_F2];
  int n1 = c->p.n1, n2 = c->p.n2;
  if (c->p.bc[OR_X2DN] == OR_POLARAXIS) {
    for (int i = -1; i < n1; i++) {
      PA(c, F1, i, -1)[B2] = -PA(c, F1, i, 0)[B2];
      for (int k = 0; k < NPR; k++) if (k != U2) PA(c, F2, i, 0)[k] = 0.;
    }
  }
  if (c->p.bc[OR_X2UP] == OR_POLARAXIS) {
    for (int i = -1; i < n1; i++) {
      PA(c, F1, i, n2)[B2] = -PA(c, F1, i, n2 - 1)[B2];
      for (int k = 0; k < NPR; k++) if (k != U2) PA(c, F2, i, n2)[k] = 0.;
    }
  }
  for (int j = -NG; j < n2 + NG; j++)
    for (int i = -NG; i < n1 + NG; i++)
      if ((i >= 0) && (i < n1) && (abs(j - (n2 - 1) / 2) <= 0))
        for (int k = 0; k < NPR; k++) if (k != U2) PA(c, F2, i, j)[k] = 0.;
  if (c->p.bc[OR_X1DN] == OR_OUTFLOW)
    for (int j = -1; j < n2; j++) if (PA(c, F1, 0, j)[RHO] > 0) PA(c, F1, 0, j)[RHO] = 0;
  if (c->p.bc[OR_X1UP] == OR_OUTFLOW)
    for (int j = -1; j < n2; j++) if (PA(c, F1, n1, j)[RHO] < 0) PA(c, F1, n1, j)[RHO] = 0;
}

/* step_ch.c:1159-1427 flux_ct (FLUXCTTOTH/FLUXCD/FLUXCTHLL, CORNGDETVERSION=1) */
int or_flux_ct(oracle_ctx *c) {
  double *F1 = c->arr[OR_F1], *F2 = c->arr[OR_F2], *emf = c->arr[OR_EMF];
  int n1 = c->p.n1, n2 = c->p.n2;
  if (c->p.fluxb == OR_FLUXCTHLL) return 0;
#define EMF(i, j) emf[zi(c, (i), (j))]
  if (c->p.fluxb == OR_FLUXCTTOTH) {
    for (int i = 0; i <= n1; i++)
      for (int j = 0; j <= n2; j++) {
        double ec = c->gdet[(size_t)zi(c, i, j) * 4 + CORN];
        double e1 = c->gdet[(size_t)zi(c, i, j) * 4 + FACE1];
        double e2 = c->gdet[(size_t)zi(c, i, j - 1) * 4 + FACE1];
        double e3 = c->gdet[(size_t)zi(c, i, j) * 4 + FACE2];
        double e4 = c->gdet[(size_t)zi(c, i - 1, j) * 4 + FACE2];
        EMF(i, j) = 0.25 * (ec) * (PA(c, F1, i, j)[B2] / (e1) + PA(c, F1, i, j - 1)[B2] / (e2) -
                                   PA(c, F2, i, j)[B1] / (e3) - PA(c, F2, i - 1, j)[B1] / (e4));
      }
    for (int i = 0; i <= n1; i++)
      for (int j = 0; j <= n2 - 1; j++) {
        PA(c, F1, i, j)[B1] = 0.;
        PA(c, F1, i, j)[B2] = 0.5 * (EMF(i, j) + EMF(i, j + 1));
      }
    for (int i = 0; i <= n1 - 1; i++)
      for (int j = 0; j <= n2; j++) {
        PA(c, F2, i, j)[B1] = -0.5 * (EMF(i, j) + EMF(i + 1, j));
        PA(c, F2, i, j)[B2] = 0.;
      }
  } else if (c->p.fluxb == OR_FLUXCD) {
    for (int i = 0; i <= n1; i++)
      for (int j = 0; j <= n2; j++)
        EMF(i, j) = 0.125 * (-PA(c, F1, i, j)[B2] - PA(c, F1, i + 1, j)[B2] + PA(c, F2, i, j)[B1] + PA(c, F2, i, j + 1)[B1]);
    for (int i = 0; i <= n1; i++)
      for (int j = 0; j <= n2 - 1; j++) { PA(c, F1, i, j)[B1] = 0.0; PA(c, F1, i, j)[B2] = EMF(i, j); }
    for (int i = 0; i <= n1 - 1; i++)
      for (int j = 0; j <= n2; j++) { PA(c, F2, i, j)[B1] = -EMF(i, j); PA(c, F2, i, j)[B2] = 0.; }
  }
#undef EMF
  return 0;
}

/* ---------------- sources (phys.c) ---------------- */

/* phys.c:475-510 source + :520-619 source_conn (WITHGDET, cooling=0);
   dUcomp0 (may be NULL) receives dUcomp[GEOMSOURCE] */
static void source(const oracle_ctx *c, const double *pr, const geom_t *g, double *dU, double *dUcomp0) {
  state_t q;
  double mhd[NDIM][NDIM], dUconn[NPR];
  const double *cn = c->conn + (size_t)zi(c, g->i, g->j) * 64;
  get_state(pr, g, &q);
  for (int j = 0; j < NDIM; j++) mhd_calc_0(c, pr, j, &q, mhd[j]);
  for (int k = 0; k < NPR; k++) dUconn[k] = 0;
  for (int l = 0; l < NDIM; l++)
    for (int j = 0; j < NDIM; j++)
      for (int k = 0; k < NDIM; k++)
        dUconn[UU + l] += mhd[j][k] * cn[(k * 4 + l) * 4 + j];
  /* dUcomp[GEOMSOURCE]=e*dUconn, all other components 0; dU = sum over them */
  for (int k = 0; k < NPR; k++) dU[k] = g->e[k] * dUconn[k];
  if (dUcomp0)
    for (int k = 0; k < NPR; k++) dUcomp0[k] = 0.0 * g->e[k] + g->e[k] * dUconn[k];
}

/* ---------------- inversion (utoprim_2d_final.c, u2p_util.c) ---------------- */

typedef struct {
  double Bsq, QdotBsq, Qtsq, Qdotn, D, wglobal, gam, gammafail;
  long nstroke;
} u2p_t;

/* utoprim_2d_final.c:356-370 vsq_calc */
static double vsq_calc(const u2p_t *s, double W) {
  double Wsq = W * W;
  double Xsq = (s->Bsq + W) * (s->Bsq + W);
  return ((Wsq * s->Qtsq + s->QdotBsq * (s->Bsq + 2. * W)) / (Wsq * Xsq));
}

/* utoprim_2d_final.c:1082-1155 func_vsq2 with pressure_W_vsq :447, dpdW_calc_vsq :465,
   dpdvsq_calc :479 (GAMMA=gam) */
static void func_vsq2(const u2p_t *s, const double x[2], double dx[2], double *f, double *df) {
  double gam = s->gam;
  double W = x[0], vsq = x[1], Wsq = W * W;
  double gtmp = 1. - vsq;
  double p_tmp = (gam - 1.) * (W * gtmp - s->D * sqrt(gtmp)) / gam;
  double dPdW = (gam - 1.) * (1. - vsq) / gam;
  double dPdvsq = (gam - 1.) * (0.5 * s->D / sqrt(1. - vsq) - W) / gam;
  double Bsq = s->Bsq, QdotBsq = s->QdotBsq;
  double t2 = -0.5 * Bsq + dPdvsq;
  double t3 = Bsq + W;
  double t4 = t3 * t3;
  double t9 = 1 / Wsq;
  double t11 = s->Qtsq - vsq * t4 + QdotBsq * (Bsq + 2.0 * W) * t9;
  double t16 = QdotBsq * t9;
  double t18 = -s->Qdotn - 0.5 * Bsq * (1.0 + vsq) + 0.5 * t16 - W + p_tmp;
  double t21 = 1 / t3;
  double t23 = 1 / W;
  double t24 = t16 * t23;
  double t25 = -1.0 + dPdW - t24;
  double t35 = t25 * t3 + (Bsq - 2.0 * dPdvsq) * (QdotBsq + vsq * Wsq * W) * t9 * t23;
  double t36 = 1 / t35;
  dx[0] = -(t2 * t11 + t4 * t18) * t21 * t36;
  double t40 = (vsq + t24) * t3;
  dx[1] = -(-t25 * t11 - 2.0 * t40 * t18) * t21 * t36;
  *df = -t11 * t11 - t18 * t18;
  *f = -0.5 * (*df);
}

/* utoprim_2d_final.c:644-1012 general_newton_raphson, do_line_search=USE_LINE_SEARCH=-2
   (no line search), NEWCONVERGE=1, CHECK_FOR_STALL=0, OPTIMIZED=1 */
static int general_newton_raphson(u2p_t *s, double x[2]) {
  double f, df, dx[2], x_old[2];
  double errx;
  int n_iter = 0, i_extra = 0, doing_extra = 0, keep_iterating = 1;
  double VSQ_TOO_BIG = (1.0 - 1.0 / ((s->gammafail - 1.0) + 1.0));
  double GAMMASQ_TOO_BIG = (s->gammafail * s->gammafail);
  errx = 1.;
  df = f = 1.;
  while (keep_iterating) {
    s->nstroke++;
    func_vsq2(s, x, dx, &f, &df);
    errx = 0.;
    x_old[0] = x[0]; x_old[1] = x[1];
    x[0] += dx[0]; x[1] += dx[1];
    errx = (x[0] == 0.) ? fabs(dx[0]) : fabs(dx[0] / x[0]);
    /* validate_x (utoprim_2d_final.c:617-633) */
    x[0] = fabs(x[0]);
    x[0] = (x[0] / s->wglobal > GAMMASQ_TOO_BIG) ? x_old[0] : x[0];
    x[1] = (x[1] < 0.) ? 0.0 : x[1];
    x[1] = (x[1] > 1.) ? VSQ_TOO_BIG : x[1];
    if ((fabs(errx) <= NEWT_TOL) && (doing_extra == 0) && (EXTRA_NEWT_ITER > 0)) doing_extra = 1;
    if (doing_extra == 1) i_extra++;
    if (((fabs(errx) <= NEWT_TOL) && (doing_extra == 0)) || (i_extra > EXTRA_NEWT_ITER) ||
        (n_iter >= (MAX_NEWT_ITER - 1)))
      keep_iterating = 0;
    n_iter++;
  }
  if (!isfinite(f) || !isfinite(df) || !isfinite(x[0]) || !isfinite(x[1])) return 1;
  if (fabs(errx) > MIN_NEWT_TOL) return 1;
  return 0;
}

/* utoprim_2d_final.c:136-396 Utoprim_new_body (+ find_root_2D_gen :514-576, x1_of_x0 :594) */
static int Utoprim_new_body(const oracle_ctx *c, const double U[NPR], const geom_t *g, double prim[NPR],
                            int *pflag, long *nstroke) {
  double QdotB, Bcon[4], Bcov[4], Qcov[4], Qcon[4], ncov[4], ncon[4], Qsq, Qtcon[4];
  double rho0, u, p, w, gammasq, gamma, gtmp, W_last, W, utsq, vsq;
  int i, j, retval, i_increase;
  u2p_t s;
  double gam = c->p.gam, gammafail = c->p.gammafail;
  double UTSQ_TOO_BIG = ((gammafail - 1.0) * (gammafail - 1.0));
  double VSQ_TOO_BIG = (1.0 - 1.0 / ((gammafail - 1.0) + 1.0));
  double GAMMASQ_TOO_BIG = (gammafail * gammafail);
  s.gam = gam; s.gammafail = gammafail; s.nstroke = 0;

  *pflag = UTOPRIMNOFAIL;
  for (i = B1; i <= B3; i++) prim[i] = U[i];
  Bcon[0] = 0.;
  for (i = 1; i < 4; i++) Bcon[i] = U[B1 + i - 1];
  for (i = 0; i < 4; i++) { Bcov[i] = 0.; for (j = 0; j < 4; j++) Bcov[i] += g->gcov[i][j] * Bcon[j]; }
  for (i = 0; i < 4; i++) Qcov[i] = U[UU + i];
  for (i = 0; i < 4; i++) { Qcon[i] = 0.; for (j = 0; j < 4; j++) Qcon[i] += g->gcon[i][j] * Qcov[j]; }
  s.Bsq = 0.;
  for (i = 1; i < 4; i++) s.Bsq += Bcon[i] * Bcov[i];
  QdotB = 0.;
  for (i = 0; i < 4; i++) QdotB += Qcov[i] * Bcon[i];
  s.QdotBsq = QdotB * QdotB;
  {
    double lapse = sqrt(-1. / g->gcon[0][0]);
    ncov[0] = -lapse; ncov[1] = 0.; ncov[2] = 0.; ncov[3] = 0.;
  }
  for (i = 0; i < 4; i++) { ncon[i] = 0.; for (j = 0; j < 4; j++) ncon[i] += g->gcon[i][j] * ncov[j]; }
  s.Qdotn = Qcon[0] * ncov[0];
  Qsq = 0.;
  for (i = 0; i < 4; i++) Qsq += Qcov[i] * Qcon[i];
  s.Qtsq = Qsq + s.Qdotn * s.Qdotn;
  s.D = U[RHO];

  utsq = 0.;
  for (i = 1; i < 4; i++)
    for (j = 1; j < 4; j++) utsq += g->gcov[i][j] * prim[U1 + i - 1] * prim[U1 + j - 1];
  if ((utsq < 0.) && (fabs(utsq) < MAXNEGUTSQ)) utsq = 0.0;
  if (utsq < 0. || utsq > UTSQ_TOO_BIG) { *pflag = UTOPRIMFAILCONVUTSQ; return 1; }

  gammasq = 1. + utsq;
  gamma = sqrt(gammasq);
  rho0 = s.D / gamma;
  u = prim[UU];
  p = (gam - 1.) * u;
  w = rho0 + u + p;
  W_last = w * gammasq;
  i_increase = 0;
  while (((W_last * W_last * W_last * (W_last + 2. * s.Bsq) - s.QdotBsq * (2. * W_last + s.Bsq)) <=
          W_last * W_last * (s.Qtsq - s.Bsq * s.Bsq)) &&
         (i_increase < 10)) {
    W_last *= 10.;
    i_increase++;
  }
  s.wglobal = w;
  {
    double x[2], vsq0;
    x[0] = fabs(W_last);
    vsq0 = fabs(vsq_calc(&s, W_last));
    x[1] = (vsq0 > 1.) ? VSQ_TOO_BIG : vsq0;
    retval = general_newton_raphson(&s, x);
    W = x[0];
  }
  *nstroke += s.nstroke;
  if ((retval != 0) || (W == FAIL_VAL)) { *pflag = retval * 100 + 1; return retval; }
  else {
    double Wtest = W / s.wglobal;
    if (Wtest <= 0. || Wtest > GAMMASQ_TOO_BIG) { *pflag = UTOPRIMFAILCONVW; return retval; }
  }
  vsq = vsq_calc(&s, W);
  if ((vsq < 0.) && (fabs(vsq) < MAXNEGVSQ)) vsq = 0.0;
  if ((vsq > VSQ_TOO_BIG) || (vsq < 0.0)) { *pflag = UTOPRIMFAILCONVUTSQ2; return retval; }
  gtmp = sqrt(1. - vsq);
  gamma = 1. / gtmp;
  rho0 = s.D * gtmp;
  w = W * (1. - vsq);
  p = (gam - 1.) * (w - rho0) / gam;
  u = w - (rho0 + p);
  if ((rho0 <= 0.) || (u < 0.)) {
    if ((rho0 <= 0.) && (u >= 0.)) *pflag = UTOPRIMFAILRHONEG;
    if ((rho0 > 0.) && (u < 0.)) *pflag = UTOPRIMFAILUNEG;
    if ((rho0 <= 0.) && (u < 0.)) *pflag = UTOPRIMFAILRHOUNEG;
  }
  prim[RHO] = rho0;
  prim[UU] = u;
  for (i = 1; i < 4; i++) Qtcon[i] = Qcon[i] + ncon[i] * s.Qdotn;
  for (i = 1; i < 4; i++) prim[U1 + i - 1] = gamma / (W + s.Bsq) * (Qtcon[i] + QdotB * Bcon[i] / W);
  for (i = B1; i <= B3; i++) prim[i] = U[i];
  return retval;
}

/* utoprim_2d_final.c:26-101 Utoprim_2d_final */
static int Utoprim_2d_final(const oracle_ctx *c, const double U[NPR], const geom_t *g, double prim[NPR],
                            int *pflag, long *nstroke) {
  double U_tmp[NPR], prim_tmp[NPR], alpha;
  int i, ret;
  for (i = B1; i <= B3; i++) prim[i] = U[i];
  alpha = 1.0 / sqrt(-g->gcon[0][0]);
  for (i = RHO; i <= B3; i++) U_tmp[i] = alpha * U[i];
  for (i = 0; i < B1; i++) prim_tmp[i] = prim[i];
  for (i = B1; i <= B3; i++) prim_tmp[i] = alpha * prim[i];
  ret = Utoprim_new_body(c, U_tmp, g, prim_tmp, pflag, nstroke);
  if (ret == 0)
    for (i = 0; i < B1; i++) prim[i] = prim_tmp[i];
  return 0;
}

/* phys.ffde.c:141-372 Utoprim_ffde (CHECKBSQMINUSESQ=1, AVOIDUTINF=2,
   AVOIDCS=0, FIXB3=0, REMOVET=0).  U is geometry-free: T^t_mu in U[UU..U3],
   B^i in U[B1..B3].  pr[RHO], pr[UU] are not touched.  ACCOUNTDISS only feeds
   diagnostics (diss_account, phys.ffde.c:572-607) and is left out.
   Returns 1 on the space-like failure of vcon2pr (phys.ffde.c:309-312). */
/* vcon2pr(VELREL4) -> ucon_calc_3vel (transforms.c:348-396, phys.c:885-955):
   u^t from the normalisation of the coordinate velocity vcon; the reference
   returns u^i itself here (transforms.c:381-386).  1 if vcon is space-like. */
static int vcon2pr_rel4(const geom_t *g, const double vcon[NDIM], double *pr) {
  double discr, ucon[NDIM];
  const double (*gc)[NDIM] = g->gcov;
  discr = +gc[0][0] + gc[1][1] * vcon[1] * vcon[1] + gc[2][2] * vcon[2] * vcon[2] + gc[3][3] * vcon[3] * vcon[3] +
          2. * (gc[0][1] * vcon[1] + gc[0][2] * vcon[2] + gc[0][3] * vcon[3] + gc[1][2] * vcon[1] * vcon[2] +
                gc[1][3] * vcon[1] * vcon[3] + gc[2][3] * vcon[2] * vcon[3]);
  if (discr > 0.) return 1;
  ucon[0] = 1. / sqrt(-discr);
  for (int j = 1; j < NDIM; j++) ucon[j] = vcon[j] * ucon[0];
  pr[U1] = ucon[1];
  pr[U2] = ucon[2];
  pr[U3] = ucon[3];
  return 0;
}

static int Utoprim_ffde(const oracle_ctx *c, const double U[NPR], const geom_t *g, double *pr) {
  double alphasq, BconOalpha[NDIM], BcovOalpha[NDIM], BsqOalphasq, T0[NDIM];
  double EconOalpha[NDIM], EcovOalpha[NDIM], EsqOalphasq, BsqmEsqOalphasq, realBsqOalphasq;
  double gammasq, Gsq, oBsqOalphasqgeom, oBsqgeom, vcon[NDIM];
  int j;
  alphasq = 1. / (-g->gcon[0][0]);
  /* U2BBsq (phys.ffde.c:1261-1283) */
  BconOalpha[0] = 0.0;
  for (j = 1; j < NDIM; j++) BconOalpha[j] = U[B1 + j - 1];
  lower(BconOalpha, g, BcovOalpha);
  BsqOalphasq = 0.0;
  for (j = 1; j < NDIM; j++) BsqOalphasq += BcovOalpha[j] * BconOalpha[j];
  for (j = 0; j < NDIM; j++) T0[j] = U[UU + j];
  /* TtoEconOalpha (phys.ffde.c:1287-1317) */
  oBsqOalphasqgeom = 1.0 / (BsqOalphasq * g->g);
  EconOalpha[0] = 0;
  EconOalpha[1] = (BcovOalpha[2] * T0[3] - BcovOalpha[3] * T0[2]) * (oBsqOalphasqgeom);
  EconOalpha[2] = (BcovOalpha[3] * T0[1] - BcovOalpha[1] * T0[3]) * (oBsqOalphasqgeom);
  EconOalpha[3] = (BcovOalpha[1] * T0[2] - BcovOalpha[2] * T0[1]) * (oBsqOalphasqgeom);
  /* B^2-E^2 check with the Lorentz-factor cap (phys.ffde.c:198-231) */
  realBsqOalphasq = BsqOalphasq;
  lower(EconOalpha, g, EcovOalpha);
  EsqOalphasq = 0.0;
  for (j = 1; j < NDIM; j++) EsqOalphasq += EconOalpha[j] * EcovOalpha[j];
  BsqmEsqOalphasq = BsqOalphasq - EsqOalphasq;
  gammasq = BsqOalphasq / BsqmEsqOalphasq;
  if ((gammasq > c->p.gammamax * c->p.gammamax) || (gammasq < 1.0)) {
    Gsq = 1.0 - 1.0 / (c->p.gammamax * c->p.gammamax);
    realBsqOalphasq = (sqrt(EsqOalphasq * BsqOalphasq / Gsq));
  }
  /* EBcovtovcon (phys.ffde.c:551-560) */
  oBsqgeom = 1.0 / (realBsqOalphasq * g->g);
  vcon[1] = alphasq * (-g->gcon[0][1] + (EcovOalpha[2] * BcovOalpha[3] - EcovOalpha[3] * BcovOalpha[2]) * oBsqgeom);
  vcon[2] = alphasq * (-g->gcon[0][2] + (EcovOalpha[3] * BcovOalpha[1] - EcovOalpha[1] * BcovOalpha[3]) * oBsqgeom);
  vcon[3] = alphasq * (-g->gcon[0][3] + (EcovOalpha[1] * BcovOalpha[2] - EcovOalpha[2] * BcovOalpha[1]) * oBsqgeom);
  if (vcon2pr_rel4(g, vcon, pr)) return 1;
  pr[B1] = U[B1];
  pr[B2] = U[B2];
  pr[B3] = U[B3];
  return 0;
}

/* step_ch.c:468-552 Utoprimgen: UtoU(UEVOLVE->UNOTHING) then the inversion
   of the configured equations of motion.  Returns 1 on a fatal failure
   (FFDE space-like drift), which the reference propagates out of advance(). */
static int Utoprimgen(const oracle_ctx *c, const double *U, const geom_t *g, double *pr, int *pflag,
                      long *nstroke) {
  double Ug[NPR];
  for (int k = 0; k < NPR; k++) Ug[k] = U[k] / g->e[k];
  if (c->p.eomtype != OR_FFDE) Ug[UU] += -Ug[RHO];
  if (c->p.eomtype == OR_FFDE) {
    *pflag = 0;
    return Utoprim_ffde(c, Ug, g, pr);
  }
  return Utoprim_2d_final(c, Ug, g, pr, pflag, nstroke);
}

/* fixup.c:274-377 fixup1zone (floors from the precomputed table, then
   limit_gamma); a change is accounted as COUNTFLOORACT on the final substage
   (measured from the state before the floors).  Force-free leaves pr0
   unset in the reference (fixup.c:293-313); here it is the incoming state,
   or zero with ffde_pr0_zero (the reference with zero-initialised locals). */
static int fixup1zone(oracle_ctx *c, double *pr, const geom_t *g, int finalstep) {
  double pr0[NPR];
  int changed = 0;
  memcpy(pr0, pr, sizeof(pr0));
  if (c->p.eomtype == OR_FFDE && c->ffde_pr0_zero) memset(pr0, 0, sizeof(pr0));
  if (c->p.eomtype == OR_GRMHD) {
    const double *fl = c->floors + (size_t)zi(c, g->i, g->j) * 2;
    double prfloor[2] = {fl[0], fl[1]};
    if (prfloor[0] < c->p.rhominlimit) prfloor[0] = c->p.rhominlimit;
    if (prfloor[1] < c->p.uuminlimit) prfloor[1] = c->p.uuminlimit;
    if (prfloor[0] > pr[RHO]) { pr[RHO] = prfloor[0]; changed = 1; }
    if (prfloor[1] > pr[UU]) { pr[UU] = prfloor[1]; changed = 1; }
  }
  int r = limit_gamma_acct(c, c->p.gammamax, pr, g);
  if (r >= 1) return 1;
  if (r == -1) changed = 1;
  if (changed) diag_fixup(c, pr0, pr, g, finalstep, COUNTFLOORACT);
  return 0;
}

/* ---------------- ENO point <-> zone-average conversion ----------------
 * reconstructeno.c / reconstructeno.h (DOENO=1 reconstructs U, DOENO=2 dU).
 * Weights: row s holds the weights of the stencil's cells for the zone at
 * position s of the stencil (the stencil starts s zones to its left).  Their
 * values follow from exact polynomial reconstruction of degree order-1 (cell
 * averages to the centre value, and back); the 3- and 5-cell tables are
 * reconstructeno.h:18-58. */
enum { ENO_A2C = 0, ENO_C2A = 1 };
static const double eno_a2c3[3][3] = {
    {23. / 24., 1. / 12., -1. / 24.}, {-1. / 24., 13. / 12., -1. / 24.}, {-1. / 24., 1. / 12., 23. / 24.}};
static const double eno_c2a3[3][3] = {
    {25. / 24., -1. / 12., 1. / 24.}, {1. / 24., 11. / 12., 1. / 24.}, {1. / 24., -1. / 12., 25. / 24.}};
static const double eno_a2c5[5][5] = {
    {563. / 640., 57. / 160., -373. / 960., 91. / 480., -71. / 1920.},
    {-71. / 1920., 511. / 480., -13. / 960., -3. / 160., 3. / 640.},
    {3. / 640., -29. / 480., 1067. / 960., -29. / 480., 3. / 640.},
    {3. / 640., -3. / 160., -13. / 960., 511. / 480., -71. / 1920.},
    {-71. / 1920., 91. / 480., -373. / 960., 57. / 160., 563. / 640.}};
static const double eno_c2a5[5][5] = {
    {6463. / 5760., -523. / 1440., 383. / 960., -283. / 1440., 223. / 5760.},
    {223. / 5760., 1337. / 1440., 23. / 960., 17. / 1440., -17. / 5760.},
    {-17. / 5760., 77. / 1440., 863. / 960., 77. / 1440., -17. / 5760.},
    {-17. / 5760., 17. / 1440., 23. / 960., 1337. / 1440., 223. / 5760.},
    {223. / 5760., -283. / 1440., 383. / 960., -523. / 1440., 6463. / 5760.}};

/* reconstructeno.h:47-58 cvt_matrices: only orders 3 and 5 have a table */
static const double *eno_matrix(int type, int order) {
  if (order == 3) return type == ENO_A2C ? &eno_a2c3[0][0] : &eno_c2a3[0][0];
  if (order == 5) return type == ENO_A2C ? &eno_a2c5[0][0] : &eno_c2a5[0][0];
  return NULL;
}

double or_eno_weight(int type, int order, int s, int d) {
  const double *m = eno_matrix(type, order);
  if (!m || s < 0 || s >= order || d < 0 || d >= order) return 0.0;
  return m[s * order + d];
}

/* reconstructeno.c:17-56 newtdiff: the order-th undivided Newton difference
   of the values u[s*st], u[(s+1)*st], ..., by the same recursion */
static double eno_newtdiff(const double *u, ptrdiff_t st, int order, int s) {
  if (order <= 0) return u[s * st];
  return eno_newtdiff(u, st, order - 1, s + 1) - eno_newtdiff(u, st, order - 1, s);
}

/* reconstructeno.c:58-131 choose_stencil on the line u[0..maxi] (min_index
   0): grows the stencil from the zone pos towards the smoother side, with the
   bias factor 2 towards the centred stencil; returns the left shift and may
   lower *order when the line is shorter than the stencil */
static int eno_choose_stencil(const double *u, ptrdiff_t st, int *order, int maxi, int pos) {
  const double bias_factor = 2.0;
  int is = pos, shift = 0;
  if (maxi + 1 < *order) *order = maxi + 1;
  const int index_pref = pos - (*order) / 2;
  for (int m = 1; m < *order; m++) {
    if (is == 0) continue;         /* cannot expand to the left */
    if (is + m - 1 == maxi) {      /* cannot expand to the right */
      is--;
      shift++;
      continue;
    }
    double leftdiff = eno_newtdiff(u, st, m, is - 1);
    double rightdiff = eno_newtdiff(u, st, m, is);
    double fac = (is <= index_pref) ? 1 / bias_factor : bias_factor;
    if (fabs(leftdiff) < fac * fabs(rightdiff)) {
      is--;
      shift++;
    }
  }
  return shift;
}

/* reconstructeno.c:133-235 cvtij for one component: u is the line through
   the zone along the conversion direction (index 0 = first interior zone,
   stride st), maxi = N-1 */
static double eno_cvt(int type, const double *u, ptrdiff_t st, int maxi, int pos, int *order_out, int *shift_out) {
  int order = 5; /* default_order */
  int shift = eno_choose_stencil(u, st, &order, maxi, pos);
  if (order < 3) { /* MIN_CVT_ORDER: leave the value as it is */
    *order_out = order;
    *shift_out = shift;
    return u[pos * st];
  }
  const double *mat;
  for (;;) { /* lower the order to one with a table, keeping the stencil near centred */
    mat = eno_matrix(type, order);
    if (mat) break;
    --order;
    if (shift > order / 2) shift--;
  }
  double out = 0.0;
  for (int d = 0; d < order; d++) out += mat[shift * order + d] * u[(pos - shift + d) * st];
  *order_out = order;
  *shift_out = shift;
  return out;
}

double or_eno_line(int type, const double *line, int n, int pos, int *order, int *shift) {
  return eno_cvt(type, line, 1, n - 1, pos, order, shift);
}

/* cvtij (reconstructeno.c:133-235) of one zone along dir */
static void eno_cvtij(const oracle_ctx *c, int dir, int type, const double *uin, int i, int j, double *uout) {
  const int nk = (c->p.eomtype == OR_FFDE) ? U3 + 1 : NPR;
  const ptrdiff_t st = (dir == 1) ? (ptrdiff_t)NPR * c->N2M : NPR;
  for (int k = 0; k < nk; k++) {
    const double *line = (dir == 1) ? PA(c, uin, 0, j) + k : PA(c, uin, i, 0) + k;
    int order, shift;
    uout[k] = eno_cvt(type, line, st, (dir == 1 ? c->p.n1 : c->p.n2) - 1, dir == 1 ? i : j, &order, &shift);
  }
  for (int k = nk; k < NPR; k++) uout[k] = PA(c, uin, i, j)[k]; /* EOMFFDE keeps B */
}

/* a2c()/c2a() over every interior zone along dir (CZLOOP cvtij) */
static void eno_convert(const oracle_ctx *c, int dir, int type, const double *uin, double *uout) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < c->p.n1; i++)
    for (int j = 0; j < c->p.n2; j++) eno_cvtij(c, dir, type, uin, i, j, PA(c, uout, i, j));
}

/* flux2dUavg (advance.c:204-265, FLUXB != FLUXCD) */
static void flux2dUavg(const oracle_ctx *c, int i, int j, double *dU1, double *dU2) {
  const double *F1 = c->arr[OR_F1], *F2 = c->arr[OR_F2];
  double idx1 = 1.0 / c->p.dx1, idx2 = 1.0 / c->p.dx2;
  for (int k = 0; k < NPR; k++) {
    dU1[k] = (-(PA(c, F1, i + 1, j)[k] - PA(c, F1, i, j)[k]) * idx1);
    dU2[k] = (-(PA(c, F2, i, j + 1)[k] - PA(c, F2, i, j)[k]) * idx2);
  }
}

/* dUtoU (advance.c:273-286) */
static void dUtoU(const double *dUgeom, const double *dUriemann, const double *CUf, const double *Cunew,
                  const double *Ui, double *ul, double *Uf, double *un, double dt) {
  for (int k = 0; k < NPR; k++)
    ul[k] = Uf[k] = CUf[0] * Ui[k] + CUf[1] * ul[k] + CUf[2] * dt * (dUriemann[k] + dUgeom[k]);
  for (int k = 0; k < NPR; k++)
    un[k] += Cunew[0] * Ui[k] + Cunew[1] * dt * (dUriemann[k] + dUgeom[k]) + Cunew[2] * Uf[k];
}

/* fixup.c:234-270 diag_fixup_U: dVF (Uf - Ui) in every region holding the
   zone, on every substage (its finalstep test is forced true) */
static void diag_fixup_U(oracle_ctx *c, const double *Ui, const double *Uf, const geom_t *g, int whocalled) {
  for (int r = 0; r < OR_NREG; r++) {
    if (!in_region(c->A.enerpos[r], g->i, g->j)) continue;
    for (int k = 0; k < NPR; k++) {
      double ftemp = c->A.dVF * (Uf[k] - Ui[k]);
      c->A.fladdterms[r][whocalled][k] += ftemp;
      c->A.fladd[r][k] += ftemp;
    }
  }
}

/* The force-free inversion's dissipation account (phys.ffde.c:354-364 with
   ACCOUNTDISS 1, diss_account :572-607 with ACCOUNTNONCONS 1), for the
   geometry-free Ug the inversion took and the state pr it returned.  The
   inversion's E/alpha and its Lorentz-factor test are formed again from Ug
   as Utoprim_ffde forms them.  Initial T^t_mu from E and B alone (FitUtoT,
   :917-935: Max_con :1472-1552, ffdestresstensor :1430-1462, lower_A
   :1654-1682), final U from pr (primtoU UDIAG), the energy lost by the
   non-conservative inversion added back; COUNTUTOPRIMFAILCONV when the cap
   acted, COUNTFLOORACT otherwise. */
static void diss_account(oracle_ctx *c, const double *Ug, const geom_t *g, const double *pr) {
  const double (*gcv)[NDIM] = g->gcov;
  const double (*gcn)[NDIM] = g->gcon;
  double BconOalpha[NDIM], BcovOalpha[NDIM], BsqOalphasq, EconOalpha[NDIM], EcovOalpha[NDIM], EsqOalphasq;
  int j, k, i, modified = 0;
  BconOalpha[0] = 0.0;
  for (j = 1; j < NDIM; j++) BconOalpha[j] = Ug[B1 + j - 1];
  lower(BconOalpha, g, BcovOalpha);
  BsqOalphasq = 0.0;
  for (j = 1; j < NDIM; j++) BsqOalphasq += BcovOalpha[j] * BconOalpha[j];
  {
    double o = 1.0 / (BsqOalphasq * g->g);
    EconOalpha[0] = 0;
    EconOalpha[1] = (BcovOalpha[2] * Ug[UU + 3] - BcovOalpha[3] * Ug[UU + 2]) * (o);
    EconOalpha[2] = (BcovOalpha[3] * Ug[UU + 1] - BcovOalpha[1] * Ug[UU + 3]) * (o);
    EconOalpha[3] = (BcovOalpha[1] * Ug[UU + 2] - BcovOalpha[2] * Ug[UU + 1]) * (o);
  }
  lower(EconOalpha, g, EcovOalpha);
  EsqOalphasq = 0.0;
  for (j = 1; j < NDIM; j++) EsqOalphasq += EconOalpha[j] * EcovOalpha[j];
  {
    double gammasq = BsqOalphasq / (BsqOalphasq - EsqOalphasq);
    if ((gammasq > c->p.gammamax * c->p.gammamax) || (gammasq < 1.0)) modified = 1;
  }
  /* FitUtoT: prffde U1..U3 = F^{it} = -E^i/alpha, B1..B3 = -B^i */
  double Fit[NDIM], Bm[NDIM], etacov[NDIM], etacon[NDIM], Econ[NDIM], Ecov[NDIM];
  double Mcon[NDIM][NDIM], Mcov[NDIM][NDIM], T[NDIM][NDIM], Msq;
  double alpha = 1. / sqrt(-gcn[0][0]);
  etacov[0] = -alpha;
  for (j = 1; j < NDIM; j++) etacov[j] = 0.0;
  etacon[0] = gcn[0][0] * etacov[0] + gcn[0][1] * etacov[1] + gcn[0][2] * etacov[2] + gcn[0][3] * etacov[3];
  etacon[1] = gcn[0][1] * etacov[0] + gcn[1][1] * etacov[1] + gcn[1][2] * etacov[2] + gcn[1][3] * etacov[3];
  etacon[2] = gcn[0][2] * etacov[0] + gcn[1][2] * etacov[1] + gcn[2][2] * etacov[2] + gcn[2][3] * etacov[3];
  etacon[3] = gcn[0][3] * etacov[0] + gcn[1][3] * etacov[1] + gcn[2][3] * etacov[2] + gcn[3][3] * etacov[3];
  Fit[0] = 0.0;
  for (j = 1; j < NDIM; j++) Fit[j] = -EconOalpha[j];
  Bm[0] = 0.0;
  for (j = 1; j < NDIM; j++) Bm[j] = etacov[0] * (-Ug[B1 + j - 1]);
  for (j = 0; j < NDIM; j++) Econ[j] = etacov[0] * Fit[j];
  lower(Econ, g, Ecov);
  Mcon[0][0] = Mcon[1][1] = Mcon[2][2] = Mcon[3][3] = 0.;
  Mcon[1][2] = -etacon[1] * Bm[2] + etacon[2] * Bm[1] - (etacov[0] / g->g) * Ecov[3];
  Mcon[2][3] = -etacon[2] * Bm[3] + etacon[3] * Bm[2] - (etacov[0] / g->g) * Ecov[1];
  Mcon[3][1] = -etacon[3] * Bm[1] + etacon[1] * Bm[3] - (etacov[0] / g->g) * Ecov[2];
  Mcon[2][1] = -Mcon[1][2];
  Mcon[3][2] = -Mcon[2][3];
  Mcon[1][3] = -Mcon[3][1];
  for (j = 1; j < NDIM; j++) {
    Mcon[0][j] = Bm[j] / etacov[0];
    Mcon[j][0] = -Mcon[0][j];
  }
  /* lower_A: the six upper entries summed over (j, k) in loop order */
  for (j = 0; j < NDIM; j++)
    for (k = 0; k < NDIM; k++) Mcov[j][k] = 0.;
  for (j = 0; j < NDIM; j++)
    for (k = 0; k < NDIM; k++) {
      Mcov[0][1] += (gcv[0][j]) * (gcv[1][k]) * Mcon[j][k];
      Mcov[0][2] += (gcv[0][j]) * (gcv[2][k]) * Mcon[j][k];
      Mcov[0][3] += (gcv[0][j]) * (gcv[3][k]) * Mcon[j][k];
      Mcov[1][2] += (gcv[1][j]) * (gcv[2][k]) * Mcon[j][k];
      Mcov[2][3] += (gcv[2][j]) * (gcv[3][k]) * Mcon[j][k];
      Mcov[3][1] += (gcv[3][j]) * (gcv[1][k]) * Mcon[j][k];
    }
  Mcov[1][0] = -Mcov[0][1];
  Mcov[2][0] = -Mcov[0][2];
  Mcov[3][0] = -Mcov[0][3];
  Mcov[2][1] = -Mcov[1][2];
  Mcov[3][2] = -Mcov[2][3];
  Mcov[1][3] = -Mcov[3][1];
  /* ffdestresstensor: T^a_b = M^ac M_bc - 1/4 delta (M_cd M^cd) */
  Msq = 0.;
  for (j = 0; j < NDIM; j++)
    for (k = 0; k < NDIM; k++) Msq += Mcov[j][k] * Mcon[j][k];
  for (j = 0; j < NDIM; j++)
    for (k = 0; k < NDIM; k++) {
      T[j][k] = 0.;
      for (i = 0; i < NDIM; i++) T[j][k] += Mcon[j][i] * Mcov[k][i];
    }
  for (j = 0; j < NDIM; j++) T[j][j] -= 0.25 * Msq;
  double Ui[NPR], Uf[NPR];
  state_t q;
  Ui[RHO] = 0.0;
  for (j = 0; j < NDIM; j++) Ui[UU + j] = g->g * T[0][j];
  for (j = 1; j < NDIM; j++) Ui[B1 + j - 1] = g->g * Ug[B1 + j - 1];
  get_state(pr, g, &q);
  primtoU_diag(c, pr, &q, g, Uf);
  Uf[RHO] = 0.0;
  Uf[UU] -= (g->g * Ug[UU] - Ui[UU]);
  diag_fixup_U(c, Ui, Uf, g, modified ? COUNTUTOPRIMFAILCONV : COUNTFLOORACT);
}

/* Utoprimgen of U into pf (guess: pf as it stands) and fixup1zone
   (advance.c:164-180 / 460-476); returns the failure flag */
static int invert_zone(oracle_ctx *c, int i, int j, const geom_t *g, const double *U, double *pfz, long *nstroke) {
  int pfl, fail = Utoprimgen(c, U, g, pfz, &pfl, nstroke);
  if (c->acct && c->p.eomtype == OR_FFDE && !fail) {
    double Ug[NPR];
    for (int k = 0; k < NPR; k++) Ug[k] = U[k] / g->e[k];
    diss_account(c, Ug, g, pfz);
  }
  c->pflag[(size_t)zi(c, i, j) * NUMPFLAGS + 0] = pfl;
  if (fixup1zone(c, pfz, g, c->finalstep)) fail |= 1;
  return fail;
}

/* advance.c:298-487 advance_eno_u (DOENO=1): Ui(pi) at zone centres is
   converted to zone averages (c2a along x1 then x2), updated with the flux
   divergence and source, converted back (a2c along x1 then x2) and inverted */
static int advance_eno_u(oracle_ctx *c, double *pi, double *pb, const double *CUf, double *pf, const double *Cunew,
                         int timeorder, int numtimeorders) {
  double *ulast = c->arr[OR_ULAST], *unew = c->arr[OR_UNEW];
  double *e0 = c->eno[0], *e1 = c->eno[1], *e2 = c->eno[2];
  const int n1 = c->p.n1, n2 = c->p.n2;
  const double dt = c->dt;
  long nstroke = 0;
  int fail = 0;
  /* pf = pi and Uicenter = U(pi) (uenotmp0) */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n1; i++)
    for (int j = 0; j < n2; j++) {
      geom_t g;
      state_t q;
      double pin[NPR];
      for (int k = 0; k < NPR; k++) pin[k] = PA(c, pi, i, j)[k];
      for (int k = 0; k < NPR; k++) PA(c, pf, i, j)[k] = pin[k];
      get_geometry(c, i, j, CENT, &g);
      get_state(pin, &g, &q);
      primtoflux(c, pin, &q, 0, &g, PA(c, e0, i, j));
    }
  eno_convert(c, 1, ENO_C2A, e0, e1); /* Uiavgtmp */
  eno_convert(c, 2, ENO_C2A, e1, e2); /* Uiavg */
  /* average Uf (or unew on the last substage) into Ufavg (uenotmp0) */
#pragma omp parallel for schedule(static) if (!c->acct)
  for (int i = 0; i < n1; i++)
    for (int j = 0; j < n2; j++) {
      geom_t g;
      double Ui[NPR], Uf[NPR], dUgeom[NPR], d1[NPR], d2[NPR], dUriemann[NPR], pbz[NPR], dUc0[NPR];
      double *ul = PA(c, ulast, i, j), *un = PA(c, unew, i, j);
      get_geometry(c, i, j, CENT, &g);
      for (int k = 0; k < NPR; k++) pbz[k] = PA(c, pb, i, j)[k];
      source(c, pbz, &g, dUgeom, dUc0);
      if (c->acct && c->finalstep) diag_source(c, &g, dUc0, dt);
      for (int k = 0; k < NPR; k++) Ui[k] = PA(c, e2, i, j)[k];
      if (timeorder == 0) for (int k = 0; k < NPR; k++) ul[k] = un[k] = 0.0;
      flux2dUavg(c, i, j, d1, d2);
      for (int k = 0; k < NPR; k++) dUriemann[k] = d1[k] + d2[k];
      dUtoU(dUgeom, dUriemann, CUf, Cunew, Ui, ul, Uf, un, dt);
      const double *src = (timeorder == numtimeorders - 1) ? un : Uf;
      for (int k = 0; k < NPR; k++) PA(c, e0, i, j)[k] = src[k];
    }
  eno_convert(c, 1, ENO_A2C, e0, e1); /* Ufcentertmp */
  eno_convert(c, 2, ENO_A2C, e1, e2); /* Ufcenter */
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : nstroke) reduction(| : fail) if (!c->acct)
  for (int i = 0; i < n1; i++)
    for (int j = 0; j < n2; j++) {
      geom_t g;
      get_geometry(c, i, j, CENT, &g);
      fail |= invert_zone(c, i, j, &g, PA(c, e2, i, j), PA(c, pf, i, j), &nstroke);
    }
  c->nstroke += nstroke;
  return fail;
}

/* advance.c:46-197 advance_standard (FIXUPZONES=FIXUP1ZONE, STEPOVERNEGU=0);
   DOENO=2 converts each direction's averaged flux divergence to its zone
   centre first (advance.c:101-108, 142-158); DOENO=1 is advance_eno_u */
int or_advance(oracle_ctx *c, int pi_sel, int pb_sel, const double *CUf, int pf_sel, const double *Cunew,
               int timeorder, int numtimeorders, double *ndt) {
  double ndt1, ndt2;
  double *pi = c->arr[pi_sel], *pb = c->arr[pb_sel], *pf = c->arr[pf_sel];
  double *ulast = c->arr[OR_ULAST], *unew = c->arr[OR_UNEW];
  int n1 = c->p.n1, n2 = c->p.n2;
  double dt = c->dt;
  long nstroke = 0;
  int fail = 0;
  c->finalstep = (timeorder == numtimeorders - 1);
  fail |= or_fluxcalc(c, pb_sel, 1, &ndt1);
  fail |= or_fluxcalc(c, pb_sel, 2, &ndt2);
  /* advance_standard calls fix_flux under FIXUPFLUX (advance.c:89-91),
     advance_eno_u always (advance.c:355) */
  if (c->p.fixupflux || c->p.eno == 1) or_fix_flux(c);
  or_flux_ct(c);
  if (c->p.eno == 1) {
    fail |= advance_eno_u(c, pi, pb, CUf, pf, Cunew, timeorder, numtimeorders);
  } else {
    if (c->p.eno == 2) {
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n1; i++)
        for (int j = 0; j < n2; j++) flux2dUavg(c, i, j, PA(c, c->eno[0], i, j), PA(c, c->eno[1], i, j));
    }
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : nstroke) reduction(| : fail) if (!c->acct)
    for (int i = 0; i < n1; i++)
      for (int j = 0; j < n2; j++) {
        double Uf[NPR], Ui[NPR], dUgeom[NPR], dUriemann[NPR], d1[NPR], d2[NPR], dUc0[NPR];
        geom_t g;
        state_t q;
        double *pfz = PA(c, pf, i, j), *piz = PA(c, pi, i, j);
        double *ul = PA(c, ulast, i, j), *un = PA(c, unew, i, j);
        double pin[NPR];
        for (int k = 0; k < NPR; k++) pin[k] = piz[k];
        for (int k = 0; k < NPR; k++) pfz[k] = pin[k];
        if (timeorder == 0) for (int k = 0; k < NPR; k++) ul[k] = un[k] = 0.0;
        get_geometry(c, i, j, CENT, &g);
        get_state(pin, &g, &q);
        primtoflux(c, pin, &q, 0, &g, Ui);
        {
          double pbz[NPR];
          for (int k = 0; k < NPR; k++) pbz[k] = PA(c, pb, i, j)[k];
          source(c, pbz, &g, dUgeom, dUc0);
          /* advance.c:138-140: diag_source with dt4diag = dt on the last substage */
          if (c->acct && c->finalstep) diag_source(c, &g, dUc0, dt);
        }
        if (c->p.eno == 2) {
          eno_cvtij(c, 1, ENO_A2C, c->eno[0], i, j, d1);
          eno_cvtij(c, 2, ENO_A2C, c->eno[1], i, j, d2);
        } else {
          flux2dUavg(c, i, j, d1, d2);
        }
        for (int k = 0; k < NPR; k++) dUriemann[k] = d1[k] + d2[k];
        dUtoU(dUgeom, dUriemann, CUf, Cunew, Ui, ul, Uf, un, dt);
        fail |= invert_zone(c, i, j, &g, (timeorder == numtimeorders - 1) ? un : Uf, pfz, &nstroke);
      }
    c->nstroke += nstroke;
  }
  *ndt = c->p.defcon * 1. / (1. / ndt1 + 1. / ndt2);
  if (fail) c->failed = 1;
  return fail;
}

/* fixup.c:1733-1841 inflow_check_rel4vel (dir=1, CENT: iin=-1, iout=N1) */
static int inflow_check_rel4vel_x1(const oracle_ctx *c, double *pr, const geom_t *g) {
  double ucon[NDIM];
  int dofix = 0;
  ucon_calc(pr, g, ucon);
  if (((g->i <= -1) && (c->p.bc[OR_X1DN] == OR_OUTFLOW) && (ucon[1] > 0.)) ||
      ((g->i >= c->p.n1) && (c->p.bc[OR_X1UP] == OR_OUTFLOW) && (ucon[1] < 0.)))
    dofix = 1;
  if (!dofix) return 0;
  double gamma = gamma_calc(pr, g);
  pr[U1] /= gamma; pr[U2] /= gamma; pr[U3] /= gamma;
  double alpha = 1. / sqrt(-g->gcon[0][0]);
  double betacon = g->gcon[0][1] * alpha * alpha;
  pr[U1] = betacon / alpha;
  double vsq = 0.;
  for (int j = 1; j < NDIM; j++)
    for (int k = 1; k < NDIM; k++) vsq += g->gcov[j][k] * pr[U1 + j - 1] * pr[U1 + k - 1];
  if (vsq < 0.0) {
    if (vsq > -NUMEPSILON * 100.0) vsq = 0.0;
    else return 1;
  }
  if (vsq >= 1.0) vsq = 1.0 - 1.0 / (c->p.gammamax * c->p.gammamax);
  gamma = 1. / sqrt(1. - vsq);
  pr[U1] *= gamma; pr[U2] *= gamma; pr[U3] *= gamma;
  return -1;
}

/* the outflow ghost (i, j) from edge zone (ri, j): EXTRAP 0 copies,
   EXTRAP 1 scales the densities and B1 by gdet and the other components
   linearly (bounds.fishmon.c:37-56 and 194-217, bounds.c:40-56 and 129-146;
   the up side's U1 with twice the slope) */
/* interp.c:559-620 rescale(), dir 1, VARTOINTERP PRIMTOINTERP: the factors
   at zone (i, j) of CENT from its r */
static void rescale_factors(const oracle_ctx *c, int i, int j, double *scale) {
  const double r = c->radius[zi(c, i, j)];
  if (c->p.rescaletype == 0) scale[RHO] = pow(r, -1.5);
  else scale[RHO] = pow(r, -2.7);
  scale[UU] = scale[RHO] / r;
  scale[U1] = scale[RHO];
  scale[U2] = 1.0;
  scale[U3] = 1.0 / (r * r);
  if (c->p.rescaletype == 0) scale[B1] = scale[U3];
  else scale[B1] = pow(r, -2.4);
  scale[B2] = scale[B1];
  scale[B3] = scale[B1];
}

static void outflow_value(const oracle_ctx *c, double *pr, int i, int j, int ri) {
  const double *src = PA(c, pr, ri, j);
  double *dst = PA(c, pr, i, j);
  if (!c->p.extrap) {
    memcpy(dst, src, sizeof(double) * NPR);
    return;
  }
  if (c->p.extrap == 2) { /* bounds.fishmon.c:56-65, 219-228: rescale(1) at the edge, rescale(-1) here */
    double sr[NPR], si[NPR], pre[NPR];
    rescale_factors(c, ri, j, sr);
    for (int k = 0; k < NPR; k++) pre[k] = src[k] / sr[k];
    rescale_factors(c, i, j, si);
    for (int k = 0; k < NPR; k++) dst[k] = pre[k] * si[k];
    return;
  }
  const double gr = c->gdet[(size_t)zi(c, ri, j) * 4 + CENT], gi = c->gdet[(size_t)zi(c, i, j) * 4 + CENT];
  const double dx1 = c->p.dx1;
  const int up = ri > 0;
  for (int k = RHO; k <= UU; k++) dst[k] = src[k] * gr / gi;
  dst[U1] = up ? src[U1] * (1. - 2 * (i - ri) * dx1) : src[U1] * (1. - (i - ri) * dx1);
  for (int k = U2; k <= U3; k++) dst[k] = up ? src[k] * (1. - (i - ri) * dx1) : src[k] * (1. + (i - ri) * dx1);
  dst[B1] = src[B1] * gr / gi;
  for (int k = B2; k <= B3; k++) dst[k] = up ? src[k] * (1. - (i - ri) * dx1) : src[k] * (1. + (i - ri) * dx1);
}

#define DISK_DELTAJ 2      /* bounds.c:378 */
#define DISK_TSTATIC 15.0  /* bounds.c:387 */
/* bounds.c:350-760 bound_disk(): the equatorial conductor of the disk
   problem, on the DELTAJ zones each side of the equator, every i (ghosts
   included): after the first call, the state made exactly symmetric across
   the equator (B1, B3, U2 antisymmetric); B1, B3 from panalytic until
   t > TSTATIC, then from the zone DELTAJ+1 away from the equator; B2 from
   panalytic; the velocity from E_phi = 0 with the field line's angular
   velocity omegafanalytic (get_omegastar, init.c:739, PROBLEMTYPE 7) through
   vcon2pr, or U2 = 0 if that fails (which also clears the failure flag, as
   the reference's failed=0 does).  RDISKINNER = 0 and RDISKOUTER = 1e30
   make the reference's radius tests true for every zone with 0 < r < 1e30,
   which holds on every cylindrical-Minkowski grid (R > (Rin e^npow -
   Rout)/(e^npow - 1) > 0), and its RNEARBLACKHOLE branch (r <= 0) is never
   taken: neither is restated.  Needs an even N2. */
static void bound_disk(oracle_ctx *c, double *pr) {
  const int n1 = c->p.n1, n2 = c->p.n2, half = n2 / 2;
  if (!c->disk_first) {
    for (int j = -NG; j < n2 + NG; j++)
      for (int i = -NG; i < n1 + NG; i++) {
        if (!((j - half < 0) && (j - half >= -DISK_DELTAJ))) continue;
        const int rj = half - j + (n2 - 1) / 2;
        double *a = PA(c, pr, i, j), *b = PA(c, pr, i, rj);
        for (int k = 0; k < NPR; k++) {
          if ((k == B1) || (k == B3) || (k == U2)) {
            a[k] = 0.5 * (a[k] - b[k]);
            b[k] = 0.5 * (b[k] - a[k]);
          } else {
            a[k] = 0.5 * (a[k] + b[k]);
            b[k] = 0.5 * (a[k] + b[k]);
          }
        }
      }
  }
  double fun = (c->t <= DISK_TSTATIC) ? 0.0 : 1.0;
  for (int j = -NG; j < n2 + NG; j++)
    for (int i = -NG; i < n1 + NG; i++) {
      int rj;
      if ((j - half < 0) && (j - half >= -DISK_DELTAJ)) rj = -(DISK_DELTAJ + 1) + half;
      else if ((j - half <= DISK_DELTAJ - 1) && (j - half >= 0)) rj = DISK_DELTAJ + half;
      else continue;
      double *a = PA(c, pr, i, j);
      const double *an = PA(c, c->panalytic, i, j);
      a[B1] = (1.0 - fun) * an[B1] + fun * PA(c, pr, i, rj)[B1];
      a[B3] = (1.0 - fun) * an[B3] + fun * PA(c, pr, i, rj)[B3];
      a[B2] = an[B2];
    }
  for (int j = -NG; j < n2 + NG; j++)
    for (int i = -NG; i < n1 + NG; i++) {
      if (!((j - half <= DISK_DELTAJ - 1) && (j - half >= -DISK_DELTAJ))) continue;
      geom_t g;
      get_geometry(c, i, j, CENT, &g);
      double *a = PA(c, pr, i, j);
      double Bcon[NDIM], Bcov[NDIM], Bsq, vcon[NDIM], primtest[NPR];
      Bcon[0] = 0;
      Bcon[1] = a[B1];
      Bcon[2] = a[B2];
      Bcon[3] = a[B3];
      lower(Bcon, &g, Bcov);
      Bsq = 0.0;
      for (int k = 0; k <= 3; k++) Bsq += Bcon[k] * Bcov[k];
      const double Om = c->omegaf[(size_t)zi(c, i, j) * NPR + 0];
      vcon[1] = -Bcon[1] * (Bcov[0] + Om * Bcov[3]) / Bsq;
      vcon[2] = -Bcon[2] * (Bcov[0] + Om * Bcov[3]) / Bsq;
      vcon[3] = Om - Bcon[3] * (Bcov[0] + Om * Bcov[3]) / Bsq;
      int sanity = 1;
      if (vcon2pr_rel4(&g, vcon, primtest)) {
        c->failed = 0;
        sanity = 0;
      } else {
        for (int k = U1; k <= U3; k++)
          if (isnan(primtest[k])) sanity = 0;
        if (sanity) {
          a[U1] = primtest[U1];
          a[U2] = primtest[U2];
          a[U3] = primtest[U3];
        }
      }
      if (!sanity) a[U2] = 0.0;
      if (sanity) c->pflag[(size_t)zi(c, i, j) * NUMPFLAGS + 0] = 0;
    }
  c->disk_first = 0;
}

/* ghost fill: bounds.fishmon.c:13-355 (outflow, polar axis + POLEDEATH; a
   FIXED side left alone) or bounds.c:11-343 (outflow, FIXED from panalytic,
   then bound_disk), and the periodic wrap (the single-rank equivalent of
   the reference's MPI halo, boundmpi.c:5-130: x1 faces first, then x2
   faces over the full x1 range) */
int or_bound_prim(oracle_ctx *c, int sel) {
  double *pr = c->arr[sel];
  int n1 = c->p.n1, n2 = c->p.n2, fail = 0;
  const int fixed_from_analytic = (c->p.boundsfile == 1);
  if (c->p.extrap == 2 && !c->radius) { /* EXTRAP 2 without or_set_radius */
    c->failed = 1;
    return 1;
  }
  /* x1 */
  if (c->p.bc[OR_X1DN] == OR_PERIODIC) {
    for (int j = 0; j < n2; j++)
      for (int i = -NG; i < 0; i++) memcpy(PA(c, pr, i, j), PA(c, pr, i + n1, j), sizeof(double) * NPR);
  } else if (c->p.bc[OR_X1DN] == OR_OUTFLOW) {
    for (int j = 0; j < n2; j++) {
      for (int i = -1; i >= -NG; i--) outflow_value(c, pr, i, j, 0);
      for (int i = -1; i >= -NG; i--) {
        geom_t g;
        get_geometry(c, i, j, CENT, &g);
        if (inflow_check_rel4vel_x1(c, PA(c, pr, i, j), &g) >= 1) fail = 1;
        if (limit_gamma(c->p.gammamax, PA(c, pr, i, j), &g) >= 1) fail = 1;
      }
    }
  } else if (c->p.bc[OR_X1DN] == OR_FIXED && fixed_from_analytic) {
    for (int j = -NG; j < n2 + NG; j++)
      for (int i = -1; i >= -NG; i--) memcpy(PA(c, pr, i, j), PA(c, c->panalytic, i, j), sizeof(double) * NPR);
  }
  if (c->p.bc[OR_X1UP] == OR_PERIODIC) {
    for (int j = 0; j < n2; j++)
      for (int i = n1; i < n1 + NG; i++) memcpy(PA(c, pr, i, j), PA(c, pr, i - n1, j), sizeof(double) * NPR);
  } else if (c->p.bc[OR_X1UP] == OR_OUTFLOW) {
    for (int j = 0; j < n2; j++) {
      for (int i = n1; i <= n1 - 1 + NG; i++) outflow_value(c, pr, i, j, n1 - 1);
      for (int i = n1; i <= n1 - 1 + NG; i++) {
        geom_t g;
        get_geometry(c, i, j, CENT, &g);
        if (inflow_check_rel4vel_x1(c, PA(c, pr, i, j), &g) >= 1) fail = 1;
        if (limit_gamma(c->p.gammamax, PA(c, pr, i, j), &g) >= 1) fail = 1;
      }
    }
  } else if (c->p.bc[OR_X1UP] == OR_FIXED && fixed_from_analytic) {
    for (int j = -NG; j < n2 + NG; j++)
      for (int i = n1; i <= n1 - 1 + NG; i++)
        memcpy(PA(c, pr, i, j), PA(c, c->panalytic, i, j), sizeof(double) * NPR);
  }
  /* x2 */
  if (c->p.bc[OR_X2DN] == OR_PERIODIC) {
    for (int i = -NG; i < n1 + NG; i++)
      for (int j = -NG; j < 0; j++) memcpy(PA(c, pr, i, j), PA(c, pr, i, j + n2), sizeof(double) * NPR);
  } else if (c->p.bc[OR_X2DN] == OR_POLARAXIS) {
    for (int i = -NG; i <= n1 - 1 + NG; i++)
      for (int j = -NG; j <= -1; j++) memcpy(PA(c, pr, i, j), PA(c, pr, i, 0 + (0 - j - 1)), sizeof(double) * NPR);
  } else if (c->p.bc[OR_X2DN] == OR_FIXED && fixed_from_analytic) {
    for (int i = -NG; i <= n1 - 1 + NG; i++)
      for (int j = -NG; j <= -1; j++) memcpy(PA(c, pr, i, j), PA(c, c->panalytic, i, j), sizeof(double) * NPR);
  }
  if (c->p.bc[OR_X2UP] == OR_PERIODIC) {
    for (int i = -NG; i < n1 + NG; i++)
      for (int j = n2; j < n2 + NG; j++) memcpy(PA(c, pr, i, j), PA(c, pr, i, j - n2), sizeof(double) * NPR);
  } else if (c->p.bc[OR_X2UP] == OR_POLARAXIS) {
    for (int i = -NG; i <= n1 - 1 + NG; i++)
      for (int j = n2; j <= n2 - 1 + NG; j++)
        memcpy(PA(c, pr, i, j), PA(c, pr, i, (n2 - 1) + ((n2 - 1) - j + 1)), sizeof(double) * NPR);
  } else if (c->p.bc[OR_X2UP] == OR_FIXED && fixed_from_analytic) {
    for (int i = -NG; i <= n1 - 1 + NG; i++)
      for (int j = n2; j <= n2 - 1 + NG; j++) memcpy(PA(c, pr, i, j), PA(c, c->panalytic, i, j), sizeof(double) * NPR);
  }
  /* bounds.c:214-233, 268-278: its polar axis makes u^2 and b^2
     antisymmetric across the pole (bounds.fishmon.c leaves them, POSDEFMETRIC 0) */
  if (c->p.boundsfile == 1) {
    if (c->p.bc[OR_X2DN] == OR_POLARAXIS)
      for (int i = -NG; i <= n1 - 1 + NG; i++)
        for (int j = -NG; j < 0; j++) {
          double *z = PA(c, pr, i, j);
          z[U2] *= -1.; z[U3] *= 1.; z[B2] *= -1.; z[B3] *= 1.;
        }
    if (c->p.bc[OR_X2UP] == OR_POLARAXIS)
      for (int i = -NG; i <= n1 - 1 + NG; i++)
        for (int j = n2; j <= n2 - 1 + NG; j++) {
          double *z = PA(c, pr, i, j);
          z[U1] *= 1.; z[U2] *= -1.; z[U3] *= 1.; z[B1] *= 1.; z[B2] *= -1.; z[B3] *= 1.;
        }
  }
  if (c->p.poledeath) {
    if (c->p.bc[OR_X2DN] == OR_POLARAXIS)
      for (int i = -NG; i <= n1 - 1 + NG; i++)
        for (int j = -NG; j < 0 + 1; j++) { PA(c, pr, i, j)[U2] *= 0; PA(c, pr, i, j)[B2] *= 0.; }
    if (c->p.bc[OR_X2UP] == OR_POLARAXIS)
      for (int i = -NG; i <= n1 - 1 + NG; i++)
        for (int j = n2 - 1; j <= n2 - 1 + NG; j++) { PA(c, pr, i, j)[U2] *= 0.; PA(c, pr, i, j)[B2] *= 0.; }
  }
  /* bounds.c:330-339: the disk's equatorial conditions after the fills */
  if (c->p.boundsfile == 1) bound_disk(c, pr);
  if (fail) c->failed = 1;
  return fail;
}

void or_set_radius(oracle_ctx *c, const double *r) {
  if (!c->radius) c->radius = (double *)calloc((size_t)c->nz, sizeof(double));
  memcpy(c->radius, r, sizeof(double) * (size_t)c->nz);
}

void or_set_analytic(oracle_ctx *c, const double *panalytic, const double *omegaf) {
  if (!c->panalytic) c->panalytic = (double *)calloc((size_t)c->nz * NPR, sizeof(double));
  if (!c->omegaf) c->omegaf = (double *)calloc((size_t)c->nz * NPR, sizeof(double));
  memcpy(c->panalytic, panalytic, sizeof(double) * (size_t)c->nz * NPR);
  if (omegaf) memcpy(c->omegaf, omegaf, sizeof(double) * (size_t)c->nz * NPR);
}

/* boundsint.c:4-60 bound_pflag (FLAGUTOPRIMFAIL only): x1 outflow copies over
   j in [-3,N2+2], then polar reflection over the full i range; periodic
   wraps x1 (interior j) then x2 (all i), as the primitives do. */
/* boundsint.c:4-60 bound_pflag: every pflag slot (FLOOP) of the ghost zones */
static void bound_pflag(oracle_ctx *c) {
  int n1 = c->p.n1, n2 = c->p.n2;
#define PFC(i, j, i2, j2) memcpy(&c->pflag[(size_t)zi(c, (i), (j)) * NUMPFLAGS], \
                                 &c->pflag[(size_t)zi(c, (i2), (j2)) * NUMPFLAGS], sizeof(int) * NUMPFLAGS)
  if (c->p.bc[OR_X1DN] == OR_OUTFLOW)
    for (int j = -NG; j <= n2 - 1 + NG; j++)
      for (int i = -1; i >= -NG; i--) PFC(i, j, 0, j);
  if (c->p.bc[OR_X1UP] == OR_OUTFLOW)
    for (int j = -NG; j <= n2 - 1 + NG; j++)
      for (int i = n1; i <= n1 - 1 + NG; i++) PFC(i, j, n1 - 1, j);
  if (c->p.bc[OR_X1DN] == OR_PERIODIC)
    for (int j = 0; j < n2; j++) {
      for (int i = -NG; i < 0; i++) PFC(i, j, i + n1, j);
      for (int i = n1; i < n1 + NG; i++) PFC(i, j, i - n1, j);
    }
  /* the polar reflection is unconditional in the reference (boundsint.c:37-52) */
  if (c->p.bc[OR_X2DN] != OR_PERIODIC)
    for (int i = -NG; i <= n1 - 1 + NG; i++)
      for (int j = -NG; j <= -1; j++) PFC(i, j, i, 0 + (0 - j - 1));
  if (c->p.bc[OR_X2UP] != OR_PERIODIC)
    for (int i = -NG; i <= n1 - 1 + NG; i++)
      for (int j = n2; j <= n2 - 1 + NG; j++) PFC(i, j, i, (n2 - 1) + ((n2 - 1) - j + 1));
  if (c->p.bc[OR_X2DN] == OR_PERIODIC)
    for (int i = -NG; i < n1 + NG; i++) {
      for (int j = -NG; j < 0; j++) PFC(i, j, i, j + n2);
      for (int j = n2; j < n2 + NG; j++) PFC(i, j, i, j - n2);
    }
#undef PFC
}

/* initbase.c:213-214 */
#define BSQORHOLIMIT 20.0
#define BSQOULIMIT 100.0
/* global.h:264-265, 558-559; fixup.c:380 */
#define GAMMAPERCDIFFMAX 2.0
#define UPERCDIFFMAX 10.0
#define UTOPRIMFAILGAMMAPERC 8
#define UTOPRIMFAILUPERC 9
#define MAXVOTES 8
#define FLAGBSQORHO 2
#define FLAGBSQOU 3

/* fixup.c:1197-1295 get_bsqflags (LIMADJUST=FLUXADJUST=0): b^2/rho and
   b^2/u flags over DQZLOOP (i in [-1,N1], j in [-1,N2]) */
static void get_bsqflags(oracle_ctx *c, const double *pv) {
  if (!c->p.checksolution) return;
  for (int i = -1; i <= c->p.n1; i++)
    for (int j = -1; j <= c->p.n2; j++) {
      int *pf = &c->pflag[(size_t)zi(c, i, j) * NUMPFLAGS];
      int f0 = 0, f3 = 0;
      if (c->p.eomtype == OR_GRMHD) {
        const double *pr = PA(c, pv, i, j);
        geom_t g;
        state_t q;
        get_geometry(c, i, j, CENT, &g);
        get_state(pr, &g, &q);
        double bsq = DOT(q.bcon, q.bcov);
        pf[FLAGBSQORHO] = 0;
        if (bsq / pr[RHO] > BSQORHOLIMIT) f0 = 2;
        else if (bsq / pr[RHO] > 0.9 * BSQORHOLIMIT) f0 = 1;
        else f0 = 0;
        if (bsq / pr[UU] > BSQOULIMIT) f3 = 2;
        else if (bsq / pr[UU] > 0.9 * BSQOULIMIT) f3 = 1;
        else f3 = 0;
      }
      pf[FLAGBSQORHO] = f0;
      pf[FLAGBSQOU] = f3;
    }
}

/* fixup.c:388-488 fixup_checksolution: zones with a b^2 flag whose inversion
   succeeded are failed (UTOPRIMFAILGAMMAPERC / UTOPRIMFAILUPERC) when most of
   their neighbours' gamma or u differ by large factors.  gammacheck persists
   across calls (a function-static array in the reference) and is refreshed
   only at flagged zones; the interior loop runs in (i, j) order and sees the
   flags it has already changed. */
static void fixup_checksolution(oracle_ctx *c, const double *pv) {
  int n1 = c->p.n1, n2 = c->p.n2;
#define PFS(i, j, s) c->pflag[(size_t)zi(c, (i), (j)) * NUMPFLAGS + (s)]
#define GC(i, j) c->gammacheck[(size_t)((i) + 1) * (n2 + 2) + (j) + 1]
#define FLAGGED(i, j) ((PFS(i, j, FLAGBSQORHO) || PFS(i, j, FLAGBSQOU)) && (PFS(i, j, 0) == UTOPRIMNOFAIL))
  for (int i = -1; i <= n1; i++)
    for (int j = -1; j <= n2; j++)
      if (FLAGGED(i, j)) {
        geom_t g;
        get_geometry(c, i, j, CENT, &g);
        GC(i, j) = gamma_calc(PA(c, pv, i, j), &g);
      }
  for (int i = 0; i < n1; i++)
    for (int j = 0; j < n2; j++) {
      if (!FLAGGED(i, j)) continue;
      const int di[MAXVOTES] = {0, 0, 1, -1, 1, 1, 1, -1}, dj[MAXVOTES] = {1, -1, 0, 0, 1, -1, 1, -1};
      double percdiff[MAXVOTES], percdiff2[MAXVOTES];
      for (int l = 0; l < MAXVOTES; l++) {
        int ok = (PFS(i + di[l], j + dj[l], 0) == UTOPRIMNOFAIL);
        percdiff[l] = ok ? (GC(i + di[l], j + dj[l]) / GC(i, j)) : -1;
        if (c->p.eomtype == OR_GRMHD)
          percdiff2[l] = ok ? (PA(c, pv, i + di[l], j + dj[l])[UU] / PA(c, pv, i, j)[UU]) : -1;
      }
      int vote0 = 0, vote1 = 0, numvotes = 0;
      for (int l = 0; l < MAXVOTES; l++) {
        if (percdiff[0] >= 0.0) {
          if ((fabs(percdiff[l]) > GAMMAPERCDIFFMAX) || (fabs(percdiff[l]) < 1.0 / GAMMAPERCDIFFMAX)) vote0++;
          if ((c->p.eomtype == OR_GRMHD) &&
              ((fabs(percdiff2[l]) > UPERCDIFFMAX) || (fabs(percdiff2[l]) < 1.0 / UPERCDIFFMAX)))
            vote1++;
          numvotes++;
        }
      }
      if (vote0 > numvotes * 0.5) PFS(i, j, 0) = UTOPRIMFAILGAMMAPERC;
      if ((c->p.eomtype == OR_GRMHD) && (vote1 > numvotes * 0.5)) PFS(i, j, 0) = UTOPRIMFAILUPERC;
    }
#undef FLAGGED
#undef GC
#undef PFS
}

/* fixup.c:527-920 fixup_utoprim (UTOPRIMADJUST=UTOPRIMAVG, ORDERINDEPENDENT=1,
   HANDLEUNEG=0, TODONOGOOD=NOGOODAVERAGE, STEPOVERNEGU=0): every zone whose
   inversion failed takes the average of its good neighbours from a copy made
   before the loop, its gamma capped at its own old gamma (or GAMMAMAX); then
   its flag becomes UTOPRIMFAILFIXED. */
static int fixup_utoprim(oracle_ctx *c, double *pv) {
  int n1 = c->p.n1, n2 = c->p.n2, nfixed = 0;
  /* cp: the reference's prc copy (fixup.c:544-552) */
  double *cp = (double *)malloc(sizeof(double) * (size_t)c->nz * NPR);
  memcpy(cp, pv, sizeof(double) * (size_t)c->nz * NPR);
#define PF(i, j) c->pflag[(size_t)zi(c, (i), (j)) * NUMPFLAGS]
#define PR(i, j, k) PA(c, cp, (i), (j))[k]
#define GOOD(i, j) (PF((i), (j)) == UTOPRIMNOFAIL)
#define AVG4_1(i, j, k) (0.25 * (PR(i, j + 1, k) + PR(i, j - 1, k) + PR(i - 1, j, k) + PR(i + 1, j, k)))
#define AVG4_2(i, j, k) (0.25 * (PR(i + 1, j + 1, k) + PR(i + 1, j - 1, k) + PR(i - 1, j + 1, k) + PR(i - 1, j - 1, k)))
  for (int i = 0; i < n1; i++)
    for (int j = 0; j < n2; j++) {
      int mypflag = PF(i, j);
      if (mypflag == 0) continue;
      double *pz = PA(c, pv, i, j);
      double pr0[NPR];
      memcpy(pr0, pz, sizeof(pr0));
      geom_t g;
      get_geometry(c, i, j, CENT, &g);
      double gamma = gamma_calc(pz, &g);
      if (gamma > c->p.gammamax) gamma = c->p.gammamax;
      int k;
      if (GOOD(i, j + 1) && GOOD(i, j - 1) && GOOD(i + 1, j) && GOOD(i - 1, j)) {
        for (k = 0; k <= U3; k++) pz[k] = AVG4_1(i, j, k);
      } else if (GOOD(i + 1, j + 1) && GOOD(i + 1, j - 1) && GOOD(i - 1, j + 1) && GOOD(i - 1, j - 1)) {
        for (k = 0; k <= U3; k++) pz[k] = AVG4_2(i, j, k);
      } else if (GOOD(i, j + 1) && GOOD(i, j - 1)) {
        for (k = 0; k <= U3; k++) pz[k] = (0.5 * (PR(i, j + 1, k) + PR(i, j - 1, k)));
      } else if (GOOD(i + 1, j) && GOOD(i - 1, j)) {
        for (k = 0; k <= U3; k++) pz[k] = (0.5 * (PR(i + 1, j, k) + PR(i - 1, j, k)));
      } else if (GOOD(i + 1, j + 1) && GOOD(i - 1, j - 1)) {
        for (k = 0; k <= U3; k++) pz[k] = (0.5 * (PR(i + 1, j + 1, k) + PR(i - 1, j - 1, k)));
      } else if (GOOD(i + 1, j - 1) && GOOD(i - 1, j + 1)) {
        for (k = 0; k <= U3; k++) pz[k] = (0.5 * (PR(i + 1, j - 1, k) + PR(i - 1, j + 1, k)));
      } else if (GOOD(i + 1, j + 1)) {
        for (k = 0; k <= U3; k++) pz[k] = PR(i + 1, j + 1, k);
      } else if (GOOD(i + 1, j)) {
        for (k = 0; k <= U3; k++) pz[k] = PR(i + 1, j, k);
      } else if (GOOD(i + 1, j - 1)) {
        for (k = 0; k <= U3; k++) pz[k] = PR(i + 1, j - 1, k);
      } else if (GOOD(i, j - 1)) {
        for (k = 0; k <= U3; k++) pz[k] = PR(i, j - 1, k);
      } else if (GOOD(i - 1, j - 1)) {
        for (k = 0; k <= U3; k++) pz[k] = PR(i - 1, j - 1, k);
      } else if (GOOD(i - 1, j)) {
        for (k = 0; k <= U3; k++) pz[k] = PR(i - 1, j, k);
      } else if (GOOD(i - 1, j + 1)) {
        for (k = 0; k <= U3; k++) pz[k] = PR(i - 1, j + 1, k);
      } else if (GOOD(i, j + 1)) {
        for (k = 0; k <= U3; k++) pz[k] = PR(i, j + 1, k);
      } else {
        for (k = 0; k <= U3; k++) pz[k] = 0.5 * (AVG4_1(i, j, k) + AVG4_2(i, j, k));
      }
      if (limit_gamma_acct(c, gamma, pz, &g) >= 1) c->failed = 1;
      /* accounting (fixup.c:870-915): the failure type's category
         (STEPOVERNEGU=0), then the flag becomes UTOPRIMFAILFIXED */
      {
        int type;
        if (mypflag == 1 || mypflag == UTOPRIMFAILCONVUTSQ || mypflag >= 101 || mypflag == UTOPRIMFAILCONVW ||
            mypflag == UTOPRIMFAILCONVUTSQ2)
          type = COUNTUTOPRIMFAILCONV;
        else if (mypflag == UTOPRIMFAILRHONEG) type = COUNTUTOPRIMFAILRHONEG;
        else if (mypflag == UTOPRIMFAILUNEG) type = COUNTUTOPRIMFAILUNEG;
        else if (mypflag == UTOPRIMFAILRHOUNEG) type = COUNTUTOPRIMFAILRHOUNEG;
        else if (mypflag == 8) type = COUNTGAMMAPERC;
        else type = COUNTUPERC;
        diag_fixup(c, pr0, pz, &g, c->finalstep, type);
      }
      PF(i, j) = -1;
      nfixed++;
    }
#undef AVG4_1
#undef AVG4_2
#undef GOOD
#undef PR
#undef PF
  free(cp);
  return nfixed;
}

/* fixup.c:37-94 post_fixup: bound_pflag then fixup_utoprim (GRMHD);
   returns the number of zones repaired */
/* fixup.c:37-94 post_fixup: bound_pflag, fixup_checksolution (CHECKSOLUTION),
   fixup_utoprim (GRMHD) */
int or_post_fixup(oracle_ctx *c, int sel) {
  bound_pflag(c);
  if (c->p.checksolution) fixup_checksolution(c, c->arr[sel]);
  if (c->p.eomtype != OR_GRMHD) return 0;
  return fixup_utoprim(c, c->arr[sel]);
}

/* step_ch.c:35-188 step_ch_simplempi + setup_rktimestep :365-460 */
int or_step(oracle_ctx *c) {
  double CUf[4][3], Cunew[4][3], ndt = 0;
  int pii[4], pbb[4], pff[4], nto = c->p.timeorder, fail = 0;
  memset(CUf, 0, sizeof(CUf)); memset(Cunew, 0, sizeof(Cunew));
  if (nto == 4) {
    double cu[4][3] = {{1.0, 0.0, 0.5}, {1.0, 0.0, 0.5}, {1.0, 0.0, 1.0}, {1.0, 0.0, 1.0}};
    double cn[4][3] = {{1.0, 1.0 / 6.0, 0.0}, {0.0, 1.0 / 3.0, 0.0}, {0.0, 1.0 / 3.0, 0.0}, {0.0, 1.0 / 6.0, 0.0}};
    memcpy(CUf, cu, sizeof(cu)); memcpy(Cunew, cn, sizeof(cn));
    int a[4] = {OR_P, OR_P, OR_P, OR_P}, b[4] = {OR_P, OR_PK0, OR_PK1, OR_PK0}, f[4] = {OR_PK0, OR_PK1, OR_PK0, OR_P};
    memcpy(pii, a, sizeof(a)); memcpy(pbb, b, sizeof(b)); memcpy(pff, f, sizeof(f));
  } else if (nto == 3) {
    double cu[3][3] = {{1.0, 0.0, 1.0}, {3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0}, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
    double cn[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 1.0}};
    memcpy(CUf, cu, sizeof(cu)); memcpy(Cunew, cn, sizeof(cn));
    int a[3] = {OR_P, OR_P, OR_P}, b[3] = {OR_P, OR_PK0, OR_PK1}, f[3] = {OR_PK0, OR_PK1, OR_P};
    memcpy(pii, a, sizeof(a)); memcpy(pbb, b, sizeof(b)); memcpy(pff, f, sizeof(f));
  } else if (nto == 2) {
    double cu[2][3] = {{1.0, 0.0, 0.5}, {1.0, 0.0, 1.0}};
    double cn[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 1.0}};
    memcpy(CUf, cu, sizeof(cu)); memcpy(Cunew, cn, sizeof(cn));
    int a[2] = {OR_P, OR_P}, b[2] = {OR_P, OR_PK0}, f[2] = {OR_PK0, OR_P};
    memcpy(pii, a, sizeof(a)); memcpy(pbb, b, sizeof(b)); memcpy(pff, f, sizeof(f));
  } else {
    double cu[1][3] = {{1.0, 0.0, 1.0}};
    double cn[1][3] = {{0.0, 0.0, 1.0}};
    memcpy(CUf, cu, sizeof(cu)); memcpy(Cunew, cn, sizeof(cn));
    pii[0] = pbb[0] = pff[0] = OR_P;
    nto = 1;
  }
  for (int to = 0; to < nto; to++) {
    get_bsqflags(c, c->arr[pii[to]]); /* pre_fixup (step_ch.c:92, fixup.c:25-33) */
    fail |= or_advance(c, pii[to], pbb[to], CUf[to], pff[to], Cunew[to], to, nto, &ndt);
    fail |= or_bound_prim(c, pff[to]);
    or_post_fixup(c, pff[to]);
    fail |= or_bound_prim(c, pff[to]);
  }
  /* step_ch.c:143-153 the current (WHICHCURRENTCALC 1), :160 diag_flux,
     both with the step's dt */
  if (c->cur) or_current(c);
  if (c->acct) or_diag_flux(c, c->dt);
  c->t += c->dt;
  if (ndt > c->p.safe * c->dt) ndt = c->p.safe * c->dt;
  c->dt = ndt;
  if (c->t + c->dt > c->tf) c->dt = c->tf - c->t;
  return fail;
}

/* ---------------- single-zone entry points ---------------- */

int or_utoprim_2d_final(oracle_ctx *c, int i, int j, const double *Ugeomfree, double *pr, int *pflag) {
  geom_t g;
  long ns = 0;
  get_geometry(c, i, j, CENT, &g);
  Utoprim_2d_final(c, Ugeomfree, &g, pr, pflag, &ns);
  return (int)ns;
}

int or_utoprimgen(oracle_ctx *c, int i, int j, const double *U, double *pr, int *pflag) {
  geom_t g;
  long ns = 0;
  get_geometry(c, i, j, CENT, &g);
  if (Utoprimgen(c, U, &g, pr, pflag, &ns)) *pflag = 50; /* FFDE space-like drift */
  return (int)ns;
}

int or_primtoU(oracle_ctx *c, int i, int j, int loc, const double *pr, double *U) {
  return or_primtoflux(c, i, j, loc, pr, 0, U);
}

int or_primtoflux(oracle_ctx *c, int i, int j, int loc, const double *pr, int dir, double *F) {
  geom_t g;
  state_t q;
  get_geometry(c, i, j, loc, &g);
  get_state(pr, &g, &q);
  primtoflux(c, pr, &q, dir, &g, F);
  return 0;
}

int or_vchar(oracle_ctx *c, int i, int j, int loc, const double *pr, int dir, double *vmax, double *vmin) {
  geom_t g;
  state_t q;
  get_geometry(c, i, j, loc, &g);
  get_state(pr, &g, &q);
  return vchar(c, pr, &q, dir, &g, vmax, vmin);
}

int or_source(oracle_ctx *c, int i, int j, const double *pr, double *dU) {
  geom_t g;
  get_geometry(c, i, j, CENT, &g);
  source(c, pr, &g, dU, NULL);
  return 0;
}

/* ---------------- context ---------------- */

oracle_ctx *or_create(const oracle_params *prm) {
  oracle_ctx *c = (oracle_ctx *)calloc(1, sizeof(oracle_ctx));
  c->p = *prm;
  c->N1M = prm->n1 + 2 * NG;
  c->N2M = prm->n2 + 2 * NG;
  c->nz = c->N1M * c->N2M;
  c->gcov = (double *)calloc((size_t)c->nz * 64, sizeof(double));
  c->gcon = (double *)calloc((size_t)c->nz * 64, sizeof(double));
  c->gdet = (double *)calloc((size_t)c->nz * 4, sizeof(double));
  c->conn = (double *)calloc((size_t)c->nz * 64, sizeof(double));
  c->floors = (double *)calloc((size_t)c->nz * 2, sizeof(double));
  for (int a = 0; a < 12; a++) c->arr[a] = (double *)calloc((size_t)c->nz * NPR, sizeof(double));
  c->pflag = (int *)calloc((size_t)c->nz * NUMPFLAGS, sizeof(int));
  /* set_arrays.c:86: every zone's flags start at -100.  The interior's
     failure flags are written by each step's inversion before anything reads
     them, and bound_pflag overwrites the ghosts of every side but a FIXED x1
     one, whose ghosts keep this value. */
  if (c->pflag)
    for (int i = -NG; i < c->p.n1 + NG; i++)
      for (int j = -NG; j < c->p.n2 + NG; j++)
        if (i < 0 || i >= c->p.n1) c->pflag[(size_t)zi(c, i, j) * NUMPFLAGS + 0] = -100;
  c->gammacheck = (double *)calloc((size_t)(prm->n1 + 2) * (prm->n2 + 2), sizeof(double));
  for (int a = 0; a < 3; a++) c->eno[a] = (double *)calloc((size_t)c->nz * NPR, sizeof(double));
  c->frdot = (double *)calloc((size_t)prm->n1 * NPR, sizeof(double));
  c->lumvsr = (double *)calloc((size_t)prm->n1, sizeof(double));
  c->tf = 1e30;
  c->disk_first = prm->disk_first;
#ifdef _OPENMP
  if (prm->nthreads > 0) omp_set_num_threads(prm->nthreads);
#endif
  return c;
}

void or_destroy(oracle_ctx *c) {
  if (!c) return;
  free(c->gcov); free(c->gcon); free(c->gdet); free(c->conn); free(c->floors); free(c->gammacheck);
  for (int a = 0; a < 3; a++) free(c->eno[a]);
  for (int a = 0; a < 12; a++) free(c->arr[a]);
  free(c->pflag);
  free(c->frdot); free(c->lumvsr);
  free(c->cf); free(c->fcon); free(c->jcon);
  free(c->panalytic); free(c->omegaf); free(c->radius);
  free(c);
}

void or_set_geometry(oracle_ctx *c, const double *gcov, const double *gcon, const double *gdet,
                     const double *conn, const double *floors) {
  memcpy(c->gcov, gcov, sizeof(double) * (size_t)c->nz * 64);
  memcpy(c->gcon, gcon, sizeof(double) * (size_t)c->nz * 64);
  memcpy(c->gdet, gdet, sizeof(double) * (size_t)c->nz * 4);
  memcpy(c->conn, conn, sizeof(double) * (size_t)c->nz * 64);
  memcpy(c->floors, floors, sizeof(double) * (size_t)c->nz * 2);
}

double *or_array(oracle_ctx *c, int sel) { return c->arr[sel]; }
int *or_pflag(oracle_ctx *c) { return c->pflag; }
void or_set_time(oracle_ctx *c, double t, double dt, double tf) { c->t = t; c->dt = dt; c->tf = tf; }
void or_get_time(const oracle_ctx *c, double *t, double *dt) { *t = c->t; *dt = c->dt; }
long or_nstroke(const oracle_ctx *c) { return c->nstroke; }
int or_failed(const oracle_ctx *c) { return c->failed; }

/* ---------------- conservation accounting: entry points ---------------- */

void or_acct_setup(oracle_ctx *c, const int *enerpos, const int *doflux, double dx3) {
  memset(&c->A, 0, sizeof(c->A));
  memcpy(c->A.enerpos, enerpos, sizeof(c->A.enerpos));
  memcpy(c->A.doflux, doflux, sizeof(c->A.doflux));
  c->dx3 = dx3;
  c->A.dVF = (c->p.dx1 * c->p.dx2) * dx3; /* set_grid.c:17-19 dV, dVF */
  memset(c->frdot, 0, sizeof(double) * (size_t)c->p.n1 * NPR);
  memset(c->lumvsr, 0, sizeof(double) * (size_t)c->p.n1);
  c->acct = 1;
}

void or_acct_off(oracle_ctx *c) { c->acct = 0; }

void or_acct_ffde_pr0_zero(oracle_ctx *c, int on) { c->ffde_pr0_zero = on; }

void or_acct_get(const oracle_ctx *c, or_acct *out, double *frdot, double *lumvsr) {
  if (out) *out = c->A;
  if (frdot) memcpy(frdot, c->frdot, sizeof(double) * (size_t)c->p.n1 * NPR);
  if (lumvsr) memcpy(lumvsr, c->lumvsr, sizeof(double) * (size_t)c->p.n1);
}

/* diag.c:576-834 diag_flux (JETBSQORHO 3.162; the step had no failed zone,
   so get_state runs on every zone) */
void or_diag_flux(oracle_ctx *c, double Dt) {
  const double *prim = c->arr[OR_P];
  or_acct *A = &c->A;
  const double gam = c->p.gam;
  for (int r = 0; r < OR_NREG; r++) {
    const int *doflux = A->doflux[r], *ep = A->enerpos[r];
    for (int dir = 0; dir < 4; dir++)
      for (int k = 0; k < NPR; k++) {
        A->pdot[r][dir][k] = 0;
        for (int fl = 0; fl < OR_NUMFLUXTERMS; fl++) A->pdotterms[r][dir][fl][k] = 0;
        if (r == 0)
          for (int fl = 0; fl < OR_NUMFLUXTERMS; fl++) A->pdottermsjet2[dir][fl][k] = 0;
      }
    for (int dir = 0; dir < 4; dir++) {
      if (doflux[dir] < 0) continue;
      int start1, stop1, start2, stop2, loc, fluxdir;
      double surface;
      const double *flux;
      if (dir == OR_X1UP || dir == OR_X1DN) {
        surface = c->p.dx2 * c->dx3;
        start1 = stop1 = doflux[dir];
        start2 = ep[OR_X2DN]; stop2 = ep[OR_X2UP];
        flux = c->arr[OR_F1]; loc = FACE1; fluxdir = 1;
      } else {
        surface = c->p.dx1 * c->dx3;
        start2 = stop2 = doflux[dir];
        start1 = ep[OR_X1DN]; stop1 = ep[OR_X1UP];
        flux = c->arr[OR_F2]; loc = FACE2; fluxdir = 2;
      }
      double *pdot = A->pdot[r][dir];
      double (*terms)[8] = A->pdotterms[r][dir];
      double (*jet2)[8] = A->pdottermsjet2[dir];
      for (int i = start1; i <= stop1; i++)
        for (int j = start2; j <= stop2; j++) {
          geom_t g;
          state_t q;
          double Ftemp[NPR], Fd[NPR], pr[NPR];
          get_geometry(c, i, j, loc, &g);
          for (int k = 0; k < NPR; k++) Ftemp[k] = PA(c, flux, i, j)[k] * surface;
          evolve_to_diag(c, &g, Ftemp, Fd);
          for (int k = 0; k < NPR; k++) pdot[k] += Fd[k];
          get_geometry(c, i, j, CENT, &g);
          for (int k = 0; k < NPR; k++) pr[k] = PA(c, prim, i, j)[k];
          get_state(pr, &g, &q);
          double pgas = (gam - 1.0) * pr[UU];
          double bsq = 0;
          for (int l = 0; l < NDIM; l++) bsq += (q.bcon[l]) * (q.bcov[l]);
          int condjet2 = 0;
          if (r == 0) {
            double bsqorho = bsq / pr[RHO];
            if (dir == OR_X1DN) condjet2 = (bsqorho > 3.162);
            else if (dir == OR_X1UP) condjet2 = ((q.ucon[1] > 0.0) && (-q.ucov[0] - 1.0 > 0.0));
          }
          double surgdet = (g.g) * surface;
          double f[OR_NUMFLUXTERMS];
          for (int k = 0; k < NPR; k++) {
            int nf = 0;
            if (k == RHO) {
              f[0] = pr[k] * (q.ucon[fluxdir]) * surgdet;
              nf = 1;
            } else if (k >= UU && k <= U3) {
              int l = k - UU;
              double d = (fluxdir == l) ? 1.0 : 0.0; /* delta(), phys.c */
              f[0] = pgas * (q.ucon[fluxdir]) * (q.ucov[l]) * surgdet;
              f[1] = PA(c, prim, i, j)[RHO] * (q.ucon[fluxdir]) * (q.ucov[l]) * surgdet;
              f[2] = PA(c, prim, i, j)[UU] * (q.ucon[fluxdir]) * (q.ucov[l]) * surgdet;
              f[3] = bsq * (q.ucon[fluxdir]) * (q.ucov[l]) * surgdet;
              f[4] = pgas * d * surgdet;
              f[5] = 0.5 * bsq * d * surgdet;
              f[6] = -(q.bcon[fluxdir]) * (q.bcov[l]) * surgdet;
              nf = 7;
            } else {
              int m = k - B1 + 1;
              f[0] = (q.bcon[m]) * (q.ucon[fluxdir]) * surgdet;
              f[1] = -(q.bcon[fluxdir]) * (q.ucon[m]) * surgdet;
              nf = 2;
            }
            for (int fl = 0; fl < nf; fl++) {
              terms[fl][k] += f[fl];
              if (condjet2) jet2[fl][k] += f[fl];
            }
          }
        }
      for (int k = 0; k < NPR; k++) A->pcum[r][dir][k] += pdot[k] * Dt;
    }
  }
  /* radial flux vs radius (diag.c:816-827) */
  const double surface = c->p.dx2 * c->dx3;
  for (int i = 0; i < c->p.n1; i++) {
    double *fr = c->frdot + (size_t)i * NPR;
    for (int k = 0; k < NPR; k++) fr[k] = 0;
    for (int j = 0; j < c->p.n2; j++) {
      geom_t g;
      double Ftemp[NPR], Fd[NPR];
      get_geometry(c, i, j, FACE1, &g);
      for (int k = 0; k < NPR; k++) Ftemp[k] = PA(c, c->arr[OR_F1], i, j)[k] * surface;
      evolve_to_diag(c, &g, Ftemp, Fd);
      for (int k = 0; k < NPR; k++) fr[k] += Fd[k];
    }
  }
}

/* SETFDIVBFLUXCT (global.h:1152-1175): |div B| at corner (i,j), zero on the
   rows j=0 and j=N2 */
static double corner_divb(const oracle_ctx *c, const double *p, int i, int j) {
  const int n2 = c->p.n2;
  const double dx1 = c->p.dx1, dx2 = c->p.dx2;
  if (!((j != 0) && (j != n2))) return 0.;
#define GC(i, j) c->gdet[(size_t)zi(c, (i), (j)) * 4 + CENT]
#define BB(i, j, k) PA(c, p, (i), (j))[k]
  double num = (((dx1) > (dx2)) ? (dx1) : (dx2)) *
               (0.5 * (BB(i, j, B1) * GC(i, j) + BB(i, j - 1, B1) * GC(i, j - 1) - BB(i - 1, j, B1) * GC(i - 1, j) -
                       BB(i - 1, j - 1, B1) * GC(i - 1, j - 1)) /
                    dx1 +
                0.5 * (BB(i, j, B2) * GC(i, j) + BB(i - 1, j, B2) * GC(i - 1, j) - BB(i, j - 1, B2) * GC(i, j - 1) -
                       BB(i - 1, j - 1, B2) * GC(i - 1, j - 1)) /
                    dx2);
  double sum = BB(i, j, B1) * GC(i, j) + BB(i, j - 1, B1) * GC(i, j - 1) + BB(i - 1, j, B1) * GC(i - 1, j) +
               BB(i - 1, j - 1, B1) * GC(i - 1, j - 1) + BB(i, j, B2) * GC(i, j) + BB(i - 1, j, B2) * GC(i - 1, j) +
               BB(i, j - 1, B2) * GC(i, j - 1) + BB(i - 1, j - 1, B2) * GC(i - 1, j - 1);
#undef GC
#undef BB
  return fabs(num / (c->gdet[(size_t)zi(c, i, j) * 4 + CORN] * 0.125 * fabs(fabs(sum) + 1.e-20)));
}

/* dump_ener.c:449-486 divbmaxavg with SETFDIVBFLUXCT (global.h:1152-1175):
   the corner divergence over LOOPDIVB (i in [-2,N1+2), j in [-2,N2+2)),
   zero on the rows j=0 and j=N2; the first largest and the sum over
   N1*N2 */
void or_divbmaxavg(const oracle_ctx *c, int sel, double *divbmax_out, double *divbavg_out, int *imax_out,
                   int *jmax_out) {
  const double *p = c->arr[sel];
  const int n1 = c->p.n1, n2 = c->p.n2;
  const double dx1 = c->p.dx1, dx2 = c->p.dx2;
  int imax = 0, jmax = 0;
  double divbmax = 0, divbavg = 0;
  (void)dx1; (void)dx2;
  for (int i = -2; i < n1 + 2; i++)
    for (int j = -2; j < n2 + 2; j++) {
      double divb = corner_divb(c, p, i, j);
      if (divb > divbmax) { imax = i; jmax = j; divbmax = divb; }
      divbavg += divb;
    }
  divbavg /= (double)(n1 * n2);
  *divbmax_out = divbmax; *divbavg_out = divbavg;
  if (imax_out) *imax_out = imax;
  if (jmax_out) *jmax_out = jmax;
}

/* dump_ener.c:562-590 constotal: sum over the region's zones of
   primtoU(UDIAG) * dVF, in ZLOOP order */
void or_constotal(const oracle_ctx *c, int sel, int region, double *vars) {
  const double *p = c->arr[sel];
  const int *ep = c->A.enerpos[region];
  for (int k = 0; k < NPR; k++) vars[k] = 0.0;
  for (int i = 0; i < c->p.n1; i++)
    for (int j = 0; j < c->p.n2; j++) {
      if (!in_region(ep, i, j)) continue;
      geom_t g;
      state_t q;
      double U[NPR], pr[NPR];
      get_geometry(c, i, j, CENT, &g);
      for (int k = 0; k < NPR; k++) pr[k] = PA(c, p, i, j)[k];
      get_state(pr, &g, &q);
      primtoU_diag(c, pr, &q, &g, U);
      for (int k = 0; k < NPR; k++) {
        double ftemp = U[k] * c->A.dVF;
        vars[k] += ftemp;
      }
    }
}

/* ---------------- current and the 56-column dump ---------------- */

void or_current_enable(oracle_ctx *c, int on) {
  if (!c->cf) {
    c->cf = (double *)calloc((size_t)c->nz * 15, sizeof(double));
    c->fcon = (double *)calloc((size_t)c->nz * 6, sizeof(double));
    c->jcon = (double *)calloc((size_t)c->nz * 4, sizeof(double));
  } else {
    memset(c->cf, 0, sizeof(double) * (size_t)c->nz * 15);
    memset(c->fcon, 0, sizeof(double) * (size_t)c->nz * 6);
    memset(c->jcon, 0, sizeof(double) * (size_t)c->nz * 4);
  }
  c->calls = 0;
  c->lastdt = 0.0;
  c->cur = on;
}

/* phys.c:1124-1218 current_doprecalc(0..3) + current_precalc with
   WHICHCURRENTCALC 1 (face CENT, every zone), then :1335-1460
   current_calc_1 over LOOPP11 x LOOPP12, with its statics calls, lastdt */
void or_current(oracle_ctx *c) {
  const double *p = c->arr[OR_P];
  const int n1 = c->p.n1, n2 = c->p.n2;
  for (int i = -NG; i < n1 + NG; i++)
    for (int j = -NG; j < n2 + NG; j++) {
      geom_t g;
      state_t q;
      get_geometry(c, i, j, CENT, &g);
      get_state(PA(c, p, i, j), &g, &q);
      double (*f)[3] = (double (*)[3])(c->cf + (size_t)zi(c, i, j) * 15);
      const double *bc = q.bcov, *uc = q.ucov;
      f[3][0] = f[4][0]; f[3][1] = f[4][1]; f[3][2] = f[4][2];
      f[4][0] = f[0][0]; f[4][1] = f[0][1]; f[4][2] = f[0][2];
      f[0][0] = -1.0 / g.g * (-bc[3] * uc[2] + bc[2] * uc[3]);
      f[0][1] = -1.0 / g.g * (bc[3] * uc[1] - bc[1] * uc[3]);
      f[0][2] = -1.0 / g.g * (-bc[2] * uc[1] + bc[1] * uc[2]);
      f[1][0] = -1.0 / g.g * (bc[3] * uc[2] - bc[2] * uc[3]);
      f[1][1] = -1.0 / g.g * (-bc[3] * uc[0] + bc[0] * uc[3]);
      f[1][2] = -1.0 / g.g * (bc[2] * uc[0] - bc[0] * uc[2]);
      f[2][0] = -1.0 / g.g * (-bc[3] * uc[1] + bc[1] * uc[3]);
      f[2][1] = -1.0 / g.g * (bc[3] * uc[0] - bc[0] * uc[3]);
      f[2][2] = -1.0 / g.g * (-bc[1] * uc[0] + bc[0] * uc[1]);
      double *fc = c->fcon + (size_t)zi(c, i, j) * 6;
      fc[0] = -1.0 / g.g * (bc[3] * uc[2] - bc[2] * uc[3]);
      fc[1] = -1.0 / g.g * (-bc[3] * uc[1] + bc[1] * uc[3]);
      fc[2] = -1.0 / g.g * (bc[2] * uc[1] - bc[1] * uc[2]);
      fc[3] = -1.0 / g.g * (bc[3] * uc[0] - bc[0] * uc[3]);
      fc[4] = -1.0 / g.g * (-bc[2] * uc[0] + bc[0] * uc[2]);
      fc[5] = -1.0 / g.g * (bc[1] * uc[0] - bc[0] * uc[1]);
    }
  const double dtl = c->lastdt, dtr = c->dt;
  const double idx1 = 1.0 / (2.0 * c->p.dx1), idx2 = 1.0 / (2.0 * c->p.dx2);
#define CF(i, j, s, m) c->cf[(size_t)zi(c, (i), (j)) * 15 + (s) * 3 + (m)]
#define GD(i, j) c->gdet[(size_t)zi(c, (i), (j)) * 4 + CENT]
  for (int i = -2; i < n1 + 2; i++)
    for (int j = -2; j < n2 + 2; j++) {
      const double gt = GD(i, j), gr = GD(i, j), grp1 = GD(i + 1, j), ghp1 = GD(i, j + 1), grm1 = GD(i - 1, j),
                   ghm1 = GD(i, j - 1);
      double *jc = c->jcon + (size_t)zi(c, i, j) * 4;
      jc[0] = 1. / gt * (grp1 * CF(i + 1, j, 1, 0) - grm1 * CF(i - 1, j, 1, 0)) * idx1 +
              1. / gt * (ghp1 * CF(i, j + 1, 2, 0) - ghm1 * CF(i, j - 1, 2, 0)) * idx2;
      double derf[3];
      for (int m = 0; m < 3; m++) {
        if (c->calls >= 2) {
          const double fl = CF(i, j, 3, m), f0 = CF(i, j, 4, m), fr = CF(i, j, 0, m);
          const double AA = (dtr * (fl - f0) + dtl * (fr - f0)) / (dtl * dtr * (dtl + dtr));
          const double BB = (-dtr * dtr * (fl - f0) + dtl * dtl * (fr - f0)) / (dtl * dtr * (dtl + dtr));
          derf[m] = 2.0 * AA * dtr + BB;
        } else {
          derf[m] = 0;
        }
      }
      jc[1] = derf[0] + 1. / gt * (ghp1 * CF(i, j + 1, 2, 1) - ghm1 * CF(i, j - 1, 2, 1)) * idx2;
      /* the reference differences F^{hr} with geomr (zone i) at i-1 */
      jc[2] = derf[1] + 1. / gt * (grp1 * CF(i + 1, j, 1, 1) - gr * CF(i - 1, j, 1, 1)) * idx1;
      jc[3] = derf[2] + 1. / gt * (grp1 * CF(i + 1, j, 1, 2) - grm1 * CF(i - 1, j, 1, 2)) * idx1 +
              1. / gt * (ghp1 * CF(i, j + 1, 2, 2) - ghm1 * CF(i, j - 1, 2, 2)) * idx2;
    }
#undef CF
#undef GD
  c->calls++;
  c->lastdt = c->dt;
}

void or_current_get(const oracle_ctx *c, double *jcon, double *fcon) {
  if (jcon) memcpy(jcon, c->jcon, sizeof(double) * (size_t)c->nz * 4);
  if (fcon) memcpy(fcon, c->fcon, sizeof(double) * (size_t)c->nz * 6);
}

/* phys.c:655-688 lowerf */
static void lowerf(const double *fcon, const geom_t *g, double *fcov) {
  double m[NDIM][NDIM], o[NDIM][NDIM];
  m[0][0] = m[1][1] = m[2][2] = m[3][3] = 0;
  m[0][1] = fcon[0]; m[0][2] = fcon[1]; m[0][3] = fcon[2];
  m[1][2] = fcon[3]; m[1][3] = fcon[4]; m[2][3] = fcon[5];
  m[1][0] = -fcon[0]; m[2][0] = -fcon[1]; m[3][0] = -fcon[2];
  m[2][1] = -fcon[3]; m[3][1] = -fcon[4]; m[3][2] = -fcon[5];
  for (int j = 0; j < NDIM; j++)
    for (int k = 0; k < NDIM; k++) {
      o[j][k] = 0;
      for (int jp = 0; jp < NDIM; jp++)
        for (int kp = 0; kp < NDIM; kp++) o[j][k] += m[jp][kp] * g->gcov[j][jp] * g->gcov[k][kp];
    }
  fcov[0] = o[0][1]; fcov[1] = o[0][2]; fcov[2] = o[0][3];
  fcov[3] = o[1][2]; fcov[4] = o[1][3]; fcov[5] = o[2][3];
}

/* dump.c:82-191 dump_content without its six coordinate columns (i, j,
   X1, X2, r, th): p, divb, ucon, ucov, bcon, bcov, vmin/vmax in x1 and x2,
   gdet, jcon, jcov, fcon, fcov -- 50 values per interior zone, out[n1][n2][50] */
void or_dump_content(const oracle_ctx *c, double *out) {
  const double *p = c->arr[OR_P];
  for (int i = 0; i < c->p.n1; i++)
    for (int j = 0; j < c->p.n2; j++) {
      geom_t g;
      state_t q;
      double vmax1, vmin1, vmax2, vmin2, jcov[NDIM], fcov[6];
      const double *pz = PA(c, p, i, j);
      get_geometry(c, i, j, CENT, &g);
      get_state(pz, &g, &q);
      vchar(c, pz, &q, 1, &g, &vmax1, &vmin1);
      vchar(c, pz, &q, 2, &g, &vmax2, &vmin2);
      double *o = out + ((size_t)i * c->p.n2 + j) * 50;
      int n = 0;
      for (int k = 0; k < NPR; k++) o[n++] = pz[k];
      o[n++] = corner_divb(c, p, i, j);
      for (int k = 0; k < NDIM; k++) o[n++] = q.ucon[k];
      for (int k = 0; k < NDIM; k++) o[n++] = q.ucov[k];
      for (int k = 0; k < NDIM; k++) o[n++] = q.bcon[k];
      for (int k = 0; k < NDIM; k++) o[n++] = q.bcov[k];
      o[n++] = vmin1; o[n++] = vmax1; o[n++] = vmin2; o[n++] = vmax2;
      o[n++] = g.g;
      const double *jc = c->cur ? c->jcon + (size_t)zi(c, i, j) * 4 : NULL;
      const double *fc = c->cur ? c->fcon + (size_t)zi(c, i, j) * 6 : NULL;
      double zero[6] = {0, 0, 0, 0, 0, 0};
      if (!jc) jc = zero;
      if (!fc) fc = zero;
      lower(jc, &g, jcov);
      for (int k = 0; k < NDIM; k++) o[n++] = jc[k];
      for (int k = 0; k < NDIM; k++) o[n++] = jcov[k];
      lowerf(fc, &g, fcov);
      for (int k = 0; k < 6; k++) o[n++] = fc[k];
      for (int k = 0; k < 6; k++) o[n++] = fcov[k];
    }
}
