/*
 * harm_problem.c -- host-side geometry tables and initial states.
 *
 * The reference computes its metric tables once on the host (set_grid.c:6-88)
 * and the kernels only read them; this file produces the same tables for the
 * two coordinate systems the benchmark problems use, with the reference's
 * operation order so that the values agree bit for bit where the libm calls
 * agree:
 *   - Kerr-Schild metric in modified Kerr-Schild coordinates (defcoord 0),
 *   - Minkowski metric on the unit square (Cartesian),
 *   - cylindrical Minkowski in (R, z) on the bi-exponential grid defcoord 666
 *     (the reference's default problem's grid).
 * It also writes the two initial states (Fishbone-Moncrief torus,
 * Orszag-Tang vortex).  No GPU use.
 */
#include "../../include/harm_amd.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NDIM 4
#define NPR 8
#define NG 3
#define FACE1 0
#define FACE2 1
#define CORN 2
#define CENT 3
#define RHO 0
#define UU 1
#define U1 2
#define U2 3
#define U3 4
#define B1 5
#define B2 6
#define B3 7
#define SINGSMALL (1E-20)
#define CONNDELTA 1E-5

static size_t zidx(const harm_grid *g, int i, int j) {
  return (size_t)(i + NG) * (size_t)(g->n2 + 2 * NG) + (size_t)(j + NG);
}

/* coord.c:920-939 coord() (single rank: startpos=0) */
static void coord(const harm_grid *g, int i, int j, int loc, double *X) {
  if (loc == FACE1) {
    X[1] = g->startx1 + (i + 0) * g->dx1;
    X[2] = g->startx2 + ((j + 0) + 0.5) * g->dx2;
  } else if (loc == FACE2) {
    X[1] = g->startx1 + ((i + 0) + 0.5) * g->dx1;
    X[2] = g->startx2 + (j + 0) * g->dx2;
  } else if (loc == CENT) {
    X[1] = g->startx1 + ((i + 0) + 0.5) * g->dx1;
    X[2] = g->startx2 + ((j + 0) + 0.5) * g->dx2;
  } else {
    X[1] = g->startx1 + (i + 0) * g->dx1;
    X[2] = g->startx2 + (j + 0) * g->dx2;
  }
  X[0] = X[3] = 0.0;
}

/* coord.c:334-455 bl_coord(), defcoord 0 and 666 (POSDEFMETRIC=0,
   COORDSINGFIX=1); for 666 r is the cylindrical R and th the height z */
static void bl_coord(const harm_grid *g, const double *X, double *r, double *th) {
  if (g->defcoord == 666) {
    *r = ((g->Rout - g->Rin) * exp(g->npow * X[1]) + g->Rin * exp(g->npow) - g->Rout) / (exp(g->npow) - 1.0);
    if (X[2] > 0.0)
      *th = ((g->Zout - 0) * exp(g->npow * fabs(X[2])) + 0 * exp(g->npow) - g->Zout) / (exp(g->npow) - 1.0);
    else
      *th = ((g->Zin - 0) * exp(g->npow * fabs(X[2])) + 0 * exp(g->npow) - g->Zin) / (exp(g->npow) - 1.0);
  } else {
    *r = g->R0 + exp(X[1]);
    *th = M_PI * X[2] + ((1. - g->hslope) / 2.) * sin(2. * M_PI * X[2]);
  }
  if (fabs(*th) < SINGSMALL) {
    if (*th >= 0) *th = SINGSMALL;
    if (*th < 0) *th = -SINGSMALL;
  }
}

/* coord.c:647-678 dxdxp_numerical (GAMMIEDERIVATIVE, DXDELTA 1e-5), the
   Jacobian for defcoord > 10 */
#define DXDELTA 1E-5
static void dxdxp_numerical(const harm_grid *g, const double *X, double dxdxp[NDIM][NDIM]) {
  for (int k = 0; k < NDIM; k++)
    for (int j = 0; j < NDIM; j++) {
      if ((j == 0) || (k == 0) || (j == 3) || (k == 3)) {
        dxdxp[j][k] = (j != k) ? 0.0 : 1.0;
      } else {
        double Xh[NDIM], Xl[NDIM], ch[NDIM], cl[NDIM];
        for (int l = 0; l < NDIM; l++) Xl[l] = Xh[l] = X[l];
        Xh[k] += DXDELTA;
        Xl[k] -= DXDELTA;
        bl_coord(g, Xh, &ch[1], &ch[2]);
        bl_coord(g, Xl, &cl[1], &cl[2]);
        dxdxp[j][k] = (ch[j] - cl[j]) / (Xh[k] - Xl[k]);
      }
    }
}

/* coord.c:462-476 dxdxprim: coord.c:476-482 dxdxp_analytic() for defcoord 0 */
static void dxdxprim(const harm_grid *g, const double *X, double r, double th, double dxdxp[NDIM][NDIM]) {
  (void)th;
  if (g->defcoord > 10) {
    dxdxp_numerical(g, X, dxdxp);
    return;
  }
  for (int j = 0; j < NDIM; j++)
    for (int k = 0; k < NDIM; k++) dxdxp[j][k] = 0.0;
  for (int j = 0; j < NDIM; j++) dxdxp[j][j] = 1.0;
  dxdxp[1][1] = r - g->R0;
  dxdxp[2][2] = M_PI + (1. - g->hslope) * M_PI * cos(2. * M_PI * X[2]);
}

/* metric.c:540-599 set_gcov_ksmetric */
static void gcov_ks(double a, double r, double th, double gcov[NDIM][NDIM]) {
  double cth = cos(th), sth = sin(th);
  double s2 = sth * sth;
  double rho2 = r * r + a * a * cth * cth;
  gcov[0][0] = (-1. + 2. * r / rho2);
  gcov[0][1] = (2. * r / rho2);
  gcov[0][2] = 0;
  gcov[0][3] = (-2. * a * r * s2 / rho2);
  gcov[1][0] = gcov[0][1];
  gcov[1][1] = (1. + 2. * r / rho2);
  gcov[1][2] = 0;
  gcov[1][3] = (-a * s2 * (1. + 2. * r / rho2));
  gcov[2][0] = 0;
  gcov[2][1] = 0;
  gcov[2][2] = (rho2);
  gcov[2][3] = 0;
  gcov[3][0] = gcov[0][3];
  gcov[3][1] = gcov[1][3];
  gcov[3][2] = 0;
  gcov[3][3] = (s2 * (rho2 + a * a * s2 * (1. + 2. * r / rho2)));
}

/* metric.c:666-723 set_gcon_ksmetric */
static void gcon_ks(double a, double r, double th, double gcon[NDIM][NDIM]) {
  double cth = cos(th), sth = sin(th);
  double s2 = sth * sth;
  double rho2 = r * r + a * a * cth * cth;
  gcon[0][0] = (-(1. + 2. * r / rho2));
  gcon[0][1] = (2. * r / rho2);
  gcon[0][2] = 0;
  gcon[0][3] = 0;
  gcon[1][0] = gcon[0][1];
  gcon[1][1] = ((r * (r - 2.) + a * a) / rho2);
  gcon[1][2] = 0;
  gcon[1][3] = (a / rho2);
  gcon[2][0] = gcon[0][2];
  gcon[2][1] = gcon[1][2];
  gcon[2][2] = (1. / rho2);
  gcon[2][3] = 0;
  gcon[3][0] = gcon[0][3];
  gcon[3][1] = gcon[1][3];
  gcon[3][2] = gcon[2][3];
  gcon[3][3] = (1. / (rho2 * s2));
}

/* metric.c:468-505 set_gcov_cylminkmetric (Mass = 0, RSTAR = 0.1) in (R, z) */
static void gcov_cylmink(double R, double z, double gcov[NDIM][NDIM]) {
  double Mass = 0.0, RSTAR = 0.1, phi;
  double r = sqrt(R * R + z * z);
  if (r < RSTAR) phi = (Mass / (2.0 * RSTAR)) * ((r / RSTAR) * (r / RSTAR) - 3.0);
  else phi = -Mass / r;
  gcov[0][0] = -1.0 - 2.0 * phi;
  gcov[0][1] = 0.0;
  gcov[0][2] = 0.0;
  gcov[0][3] = 0.0;
  gcov[1][0] = 0.0;
  gcov[1][1] = 1.0 - 2.0 * phi * R * R / (r * r);
  gcov[1][2] = -2.0 * phi * R * z / (r * r);
  gcov[1][3] = 0.0;
  gcov[2][0] = gcov[0][2];
  gcov[2][1] = gcov[1][2];
  gcov[2][2] = 1.0 - 2.0 * phi * z * z / (r * r);
  gcov[2][3] = 0.0;
  gcov[3][0] = gcov[0][3];
  gcov[3][1] = gcov[1][3];
  gcov[3][2] = gcov[2][3];
  gcov[3][3] = R * R;
}

/* metric.c:441-466 set_gcov_minkcartmetric */
static void gcov_mink(double gcov[NDIM][NDIM]) {
  for (int j = 0; j < NDIM; j++)
    for (int k = 0; k < NDIM; k++) gcov[j][k] = 0.0;
  gcov[0][0] = -1.0;
  gcov[1][1] = 1.0;
  gcov[2][2] = 1.0;
  gcov[3][3] = 1.0;
}

/* metric.c:1057-1083 gcov2gcovprim */
static void gcov2gcovprim(const double dxdxp[NDIM][NDIM], double gcov[NDIM][NDIM]) {
  double tmp[NDIM][NDIM];
  for (int j = 0; j < NDIM; j++)
    for (int k = 0; k < NDIM; k++) tmp[j][k] = 0.;
  for (int j = 0; j < NDIM; j++)
    for (int k = 0; k < NDIM; k++)
      for (int l = 0; l < NDIM; l++)
        for (int m = 0; m < NDIM; m++) tmp[j][k] += gcov[l][m] * dxdxp[l][j] * dxdxp[m][k];
  memcpy(gcov, tmp, sizeof(tmp));
}

/* Gauss-Jordan inverse with full pivoting (Numerical Recipes in C 2nd ed.
   section 2.1, the algorithm metric.c:1146 matrix_inverse -> gaussj.c uses). */
static int gauss_jordan_inverse(const double in[NDIM][NDIM], double out[NDIM][NDIM]) {
  double a[NDIM][NDIM];
  int indxc[NDIM], indxr[NDIM], ipiv[NDIM];
  int icol = 0, irow = 0;
  memcpy(a, in, sizeof(a));
  for (int j = 0; j < NDIM; j++) ipiv[j] = 0;
  for (int i = 0; i < NDIM; i++) {
    double big = 0.0;
    for (int j = 0; j < NDIM; j++)
      if (ipiv[j] != 1)
        for (int k = 0; k < NDIM; k++)
          if (ipiv[k] == 0) {
            if (fabs(a[j][k]) >= big) { big = fabs(a[j][k]); irow = j; icol = k; }
          } else if (ipiv[k] > 1) return 1;
    ++(ipiv[icol]);
    if (irow != icol)
      for (int l = 0; l < NDIM; l++) { double t = a[irow][l]; a[irow][l] = a[icol][l]; a[icol][l] = t; }
    indxr[i] = irow;
    indxc[i] = icol;
    if (a[icol][icol] == 0.0) return 1;
    double pivinv = 1.0 / a[icol][icol];
    a[icol][icol] = 1.0;
    for (int l = 0; l < NDIM; l++) a[icol][l] *= pivinv;
    for (int ll = 0; ll < NDIM; ll++)
      if (ll != icol) {
        double dum = a[ll][icol];
        a[ll][icol] = 0.0;
        for (int l = 0; l < NDIM; l++) a[ll][l] -= a[icol][l] * dum;
      }
  }
  for (int l = NDIM - 1; l >= 0; l--)
    if (indxr[l] != indxc[l])
      for (int k = 0; k < NDIM; k++) { double t = a[k][indxr[l]]; a[k][indxr[l]] = a[k][indxc[l]]; a[k][indxc[l]] = t; }
  memcpy(out, a, sizeof(a));
  return 0;
}

/* metric.c:1085-1105 gcon2gconprim */
static void gcon2gconprim(const double dxdxp[NDIM][NDIM], double gcon[NDIM][NDIM]) {
  double idxdxp[NDIM][NDIM], tmp[NDIM][NDIM];
  gauss_jordan_inverse(dxdxp, idxdxp);
  for (int j = 0; j < NDIM; j++)
    for (int k = 0; k < NDIM; k++) tmp[j][k] = 0.;
  for (int j = 0; j < NDIM; j++)
    for (int k = 0; k < NDIM; k++)
      for (int l = 0; l < NDIM; l++)
        for (int m = 0; m < NDIM; m++) tmp[j][k] += idxdxp[j][l] * idxdxp[k][m] * gcon[l][m];
  memcpy(gcon, tmp, sizeof(tmp));
}

/* LU decomposition with implicit partial pivoting (Numerical Recipes in C 2nd
   ed. section 2.3, the algorithm ludcmp.c restates); determinant as
   metric.c:1109-1140 gdet_func forms it: sqrt(-d). */
static double gdet_lu(const double gcov[NDIM][NDIM]) {
  double a[NDIM + 1][NDIM + 1], vv[NDIM + 1], d = 1.0;
  int imax = 0;
  for (int j = 0; j < NDIM; j++)
    for (int k = 0; k < NDIM; k++) a[j + 1][k + 1] = gcov[j][k];
  for (int i = 1; i <= NDIM; i++) {
    double big = 0.0, temp;
    for (int j = 1; j <= NDIM; j++)
      if ((temp = fabs(a[i][j])) > big) big = temp;
    if (big == 0.0) return -1.0;
    vv[i] = 1.0 / big;
  }
  for (int j = 1; j <= NDIM; j++) {
    double big, dum, sum;
    for (int i = 1; i < j; i++) {
      sum = a[i][j];
      for (int k = 1; k < i; k++) sum -= a[i][k] * a[k][j];
      a[i][j] = sum;
    }
    big = 0.0;
    for (int i = j; i <= NDIM; i++) {
      sum = a[i][j];
      for (int k = 1; k < j; k++) sum -= a[i][k] * a[k][j];
      a[i][j] = sum;
      if ((dum = vv[i] * fabs(sum)) >= big) { big = dum; imax = i; }
    }
    if (j != imax) {
      for (int k = 1; k <= NDIM; k++) { dum = a[imax][k]; a[imax][k] = a[j][k]; a[j][k] = dum; }
      d = -(d);
      vv[imax] = vv[j];
    }
    if (a[j][j] == 0.0) a[j][j] = 1.0e-30;
    if (j != NDIM) {
      dum = 1.0 / (a[j][j]);
      for (int i = j + 1; i <= NDIM; i++) a[i][j] *= dum;
    }
  }
  for (int j = 1; j <= NDIM; j++) d *= a[j][j];
  if (d >= 0) return -1.0;
  return sqrt(-d);
}

/* metric.c:11-73 gcov_func(getprim=1, MCOORD, X) */
static void gcov_func(const harm_grid *g, const double *X, double gcov[NDIM][NDIM]) {
  double r, th, dxdxp[NDIM][NDIM];
  if (g->metric == HARM_METRIC_MINK) { gcov_mink(gcov); return; }
  bl_coord(g, X, &r, &th);
  if (g->metric == HARM_METRIC_CYLMINK) gcov_cylmink(r, th, gcov);
  else gcov_ks(g->a, r, th, gcov);
  dxdxprim(g, X, r, th, dxdxp);
  gcov2gcovprim(dxdxp, gcov);
}

/* metric.c:75-115 gcon_func(getprim=1, MCOORD, X, gcov) */
static void gcon_func(const harm_grid *g, const double *X, const double gcov[NDIM][NDIM], double gcon[NDIM][NDIM]) {
  double r, th, dxdxp[NDIM][NDIM];
  /* no analytic gcon: the numerical inverse of the grid-coordinate gcov */
  if (g->metric == HARM_METRIC_MINK || g->metric == HARM_METRIC_CYLMINK) {
    gauss_jordan_inverse(gcov, gcon);
    return;
  }
  bl_coord(g, X, &r, &th);
  gcon_ks(g->a, r, th, gcon);
  dxdxprim(g, X, r, th, dxdxp);
  gcon2gconprim(dxdxp, gcon);
}

/* metric.c:1188-1262 conn_func_numerical1 (GAMMIEDERIVATIVE, WITHGDET) */
static void conn_numerical(const harm_grid *g, const double *X, const double gcon[NDIM][NDIM],
                           double conn[NDIM][NDIM][NDIM]) {
  double tmp[NDIM][NDIM][NDIM], Xh[NDIM], Xl[NDIM], gh[NDIM][NDIM], gl[NDIM][NDIM];
  if (g->metric == HARM_METRIC_MINK) { /* metric.c:824 set_conn_minkcartmetric: all zero */
    memset(conn, 0, sizeof(double) * 64);
    return;
  }
  for (int k = 0; k < NDIM; k++) {
    for (int l = 0; l < NDIM; l++) { Xh[l] = X[l]; Xl[l] = X[l]; }
    Xh[k] += CONNDELTA;
    Xl[k] -= CONNDELTA;
    gcov_func(g, Xh, gh);
    gcov_func(g, Xl, gl);
    for (int i = 0; i < NDIM; i++)
      for (int j = 0; j < NDIM; j++) conn[i][j][k] = (gh[i][j] - gl[i][j]) / (Xh[k] - Xl[k]);
  }
  for (int i = 0; i < NDIM; i++)
    for (int j = 0; j < NDIM; j++)
      for (int k = 0; k < NDIM; k++) tmp[i][j][k] = 0.5 * (conn[j][i][k] + conn[k][i][j] - conn[k][j][i]);
  for (int i = 0; i < NDIM; i++)
    for (int j = 0; j < NDIM; j++)
      for (int k = 0; k < NDIM; k++) {
        conn[i][j][k] = 0.;
        for (int l = 0; l < NDIM; l++) conn[i][j][k] += gcon[i][l] * tmp[l][j][k];
      }
}

/* coord.c:817-842 setihor()/setRin() for defcoord 0 */
int harm_grid_mks(harm_grid *g, int n1, int n2, double a, double Rout, double hslope, double R0) {
  if (!g || n1 < 2 || n2 < 2) return 1;
  memset(g, 0, sizeof(*g));
  g->n1 = n1; g->n2 = n2;
  g->metric = HARM_METRIC_KS;
  g->defcoord = 0;
  g->a = a; g->R0 = R0; g->Rout = Rout; g->hslope = hslope;
  double Rhor = (1. + sqrt(1. - a * a));            /* phys.c rhor_calc(0) */
  int ihor = (n1 * (0.1) > 10) ? 10 : (int)((double)n1 * (double)(0.1));
  double ihoradjust = ((double)ihor) + (0.25);
  double ftemp = ihoradjust / (double)n1;
  g->Rin = (R0 + pow((Rhor - R0) / pow(Rout - R0, ftemp), 1.0 / (1.0 - ftemp)));
  /* coord.c:720-726 set_points() */
  g->startx1 = log(g->Rin - R0);
  g->startx2 = 0.;
  g->dx1 = log((Rout - R0) / (g->Rin - R0)) / n1;
  g->dx2 = 1. / n2;
  g->rhomin = 1.e-4;   /* initbase.c:207-208 */
  g->uumin = 1.e-6;
  return 0;
}

/* init.c:54-60 init_grid() (CYLMINKMETRIC), coord.c:134-136
   set_coord_parms() (npow 10) and coord.c:798-804 set_points() */
int harm_grid_cylmink(harm_grid *g, int n1, int n2, double Rin, double Rout, double Zin, double Zout) {
  if (!g || n1 < 2 || n2 < 2) return 1;
  memset(g, 0, sizeof(*g));
  g->n1 = n1; g->n2 = n2;
  g->metric = HARM_METRIC_CYLMINK;
  g->defcoord = 666;
  g->Rin = Rin; g->Rout = Rout; g->Zin = Zin; g->Zout = Zout;
  g->npow = 10.0;
  g->hslope = 0.3;   /* init_defgrid (initbase.c:112); unused by defcoord 666 */
  g->startx1 = 0.;
  g->startx2 = -1.0;
  g->dx1 = 1.0 / n1;
  g->dx2 = 2.0 / n2;
  g->rhomin = 1.e-4;
  g->uumin = 1.e-6;
  return 0;
}

int harm_grid_cartesian(harm_grid *g, int n1, int n2) {
  if (!g || n1 < 2 || n2 < 2) return 1;
  memset(g, 0, sizeof(*g));
  g->n1 = n1; g->n2 = n2;
  g->metric = HARM_METRIC_MINK;
  g->defcoord = 0;
  g->R0 = 0.0; g->Rin = 1.0; g->Rout = exp(1.0); g->hslope = 1.0;
  g->startx1 = 0.0; g->startx2 = 0.0;
  g->dx1 = 1.0 / n1; g->dx2 = 1.0 / n2;
  g->rhomin = 1.e-4;
  g->uumin = 1.e-6;
  return 0;
}

/* set_grid.c:6-88 over FULLLOOP and the 4 positions, plus the floor values
   fixup.c:1025-1028 (rescaletype 1) at CENT. */
int harm_geometry(const harm_grid *g, double *gcov, double *gcon, double *gdet, double *conn, double *floors) {
  if (!g) return 1;
  for (int i = -NG; i < g->n1 + NG; i++)
    for (int j = -NG; j < g->n2 + NG; j++) {
      size_t z = zidx(g, i, j);
      for (int loc = 3; loc >= 0; loc--) {
        double X[NDIM], gc[NDIM][NDIM], gn[NDIM][NDIM];
        coord(g, i, j, loc, X);
        gcov_func(g, X, gc);
        double gd = gdet_lu(gc);
        if (gd < 0) return 2;
        gcon_func(g, X, gc, gn);
        memcpy(gcov + (z * 4 + loc) * 16, gc, sizeof(gc));
        memcpy(gcon + (z * 4 + loc) * 16, gn, sizeof(gn));
        gdet[z * 4 + loc] = gd;
        if (loc == CENT) {
          double cn[NDIM][NDIM][NDIM];
          conn_numerical(g, X, gn, cn);
          memcpy(conn + z * 64, cn, sizeof(cn));
          double r, th;
          bl_coord(g, X, &r, &th);
          floors[z * 2 + 0] = g->rhomin * pow(r, -2.7);
          floors[z * 2 + 1] = g->uumin * pow(r, -2.7);
        }
      }
    }
  return 0;
}

/* ---------------- initial states ---------------- */

/* splitmix64 -> uniform [0,1); the perturbation generator (the reference uses
   its own ranc(), ranc.c:31) */
static double urand(unsigned long long *s) {
  unsigned long long z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

/* init.fishmon.c:389-420 lfish_calc */
static double lfish_calc(double a, double r) {
  return (((pow(a, 2) - 2. * a * sqrt(r) + pow(r, 2)) *
           ((-2. * a * r * (pow(a, 2) - 2. * a * sqrt(r) + pow(r, 2))) / sqrt(2. * a * sqrt(r) + (-3. + r) * r) +
            ((a + (-2. + r) * sqrt(r)) * (pow(r, 3) + pow(a, 2) * (2. + r))) /
                sqrt(1 + (2. * a) / pow(r, 1.5) - 3. / r))) /
          (pow(r, 3) * sqrt(2. * a * sqrt(r) + (-3. + r) * r) * (pow(a, 2) + (-2. + r) * r)));
}

/* metric.c:601-664 set_gcov_blmetric */
static void gcov_bl(double a, double r, double th, double gcov[NDIM][NDIM]) {
  double cth = cos(th), sth = sin(th), s2 = sth * sth, a2 = a * a, r2 = r * r;
  double DD = 1. - 2. / r + a2 / r2;
  double mu = 1. + a2 * cth * cth / r2;
  for (int j = 0; j < NDIM; j++)
    for (int k = 0; k < NDIM; k++) gcov[j][k] = 0.;
  gcov[0][0] = (-(1. - 2. / (r * mu)));
  gcov[0][3] = gcov[3][0] = (-2. * a * s2 / (r * mu));
  gcov[1][1] = (mu / DD);
  gcov[2][2] = (r2 * mu);
  gcov[3][3] = (r2 * sth * sth * (1. + a2 / r2 + 2. * a2 * s2 / (r2 * r * mu)));
}

/* phys.c:957-994 ucon_calc_4vel (u^t from spatial u^i) */
static int ucon_4vel(const double gcov[NDIM][NDIM], double *ucon) {
  double AA = gcov[0][0];
  double BB = 2. * (gcov[0][1] * ucon[1] + gcov[0][2] * ucon[2] + gcov[0][3] * ucon[3]);
  double CC = 1. + gcov[1][1] * ucon[1] * ucon[1] + gcov[2][2] * ucon[2] * ucon[2] +
              gcov[3][3] * ucon[3] * ucon[3] +
              2. * (gcov[1][2] * ucon[1] * ucon[2] + gcov[1][3] * ucon[1] * ucon[3] + gcov[2][3] * ucon[2] * ucon[3]);
  double discr = BB * BB - 4. * AA * CC;
  if (discr < 0.) return 1;
  ucon[0] = (-BB - sqrt(discr)) / (2. * AA);
  return 0;
}

/* phys.c:830-845 (u^mu from relative 4-velocity) for the field normalisation */
static void state_bsq(const double *pr, const double gcov[NDIM][NDIM], const double gcon[NDIM][NDIM], double *bsq) {
  double alpha = 1. / sqrt(-gcon[0][0]);
  double qsq = gcov[1][1] * pr[U1] * pr[U1] + gcov[2][2] * pr[U2] * pr[U2] + gcov[3][3] * pr[U3] * pr[U3] +
               2. * (gcov[1][2] * pr[U1] * pr[U2] + gcov[1][3] * pr[U1] * pr[U3] + gcov[2][3] * pr[U2] * pr[U3]);
  double gamma = sqrt(1. + qsq), ucon[4], ucov[4], bcon[4], bcov[4];
  ucon[0] = gamma / alpha;
  for (int j = 1; j < 4; j++) ucon[j] = pr[U1 + j - 1] - gamma * (gcon[0][j] * alpha * alpha) / alpha;
  for (int j = 0; j < 4; j++) { ucov[j] = 0; for (int k = 0; k < 4; k++) ucov[j] += gcov[j][k] * ucon[k]; }
  bcon[0] = pr[B1] * ucov[1] + pr[B2] * ucov[2] + pr[B3] * ucov[3];
  for (int j = 1; j < 4; j++) bcon[j] = (pr[B1 - 1 + j] + bcon[0] * ucon[j]) / ucon[0];
  for (int j = 0; j < 4; j++) { bcov[j] = 0; for (int k = 0; k < 4; k++) bcov[j] += gcov[j][k] * bcon[k]; }
  *bsq = bcon[0] * bcov[0] + bcon[1] * bcov[1] + bcon[2] * bcov[2] + bcon[3] * bcov[3];
}

int harm_init_fishmon(const harm_grid *g, const double *gcov, const double *gcon, const double *gdet, double gam,
                      double rin, double rmax, double kappa, double beta, double randfact, unsigned long seed,
                      double *p) {
  if (!g || g->metric != HARM_METRIC_KS) return 1;
  double a = g->a;
  double l = lfish_calc(a, rmax);
  unsigned long long rs = (unsigned long long)seed * 0x2545F4914F6CDD1DULL + 7ULL;
  int n1 = g->n1, n2 = g->n2;
  double rhomax = 0, umax = 0, bsq_max = 0;
  double *A = (double *)calloc((size_t)(n1 + 2 * NG) * (n2 + 2 * NG), sizeof(double));
  if (!A) return 3;
  /* init.fishmon.c:115-215 init_dsandvels + transforms.c:9-23 bl2met2metp2v */
  for (int i = -NG; i < n1 + NG; i++)
    for (int j = -NG; j < n2 + NG; j++) {
      size_t z = zidx(g, i, j);
      double *pr = p + z * NPR;
      const double(*gc)[4] = (const double(*)[4])(gcov + (z * 4 + CENT) * 16);
      const double(*gn)[4] = (const double(*)[4])(gcon + (z * 4 + CENT) * 16);
      double X[NDIM], r, th;
      coord(g, i, j, CENT, X);
      bl_coord(g, X, &r, &th);
      double sth = sin(th), cth = cos(th);
      double DD = r * r - 2. * r + a * a;
      double AA = (r * r + a * a) * (r * r + a * a) - DD * a * a * sth * sth;
      double SS = r * r + a * a * cth * cth;
      double thin = M_PI / 2., sthin = sin(thin), cthin = cos(thin);
      double DDin = rin * rin - 2. * rin + a * a;
      double AAin = (rin * rin + a * a) * (rin * rin + a * a) - DDin * a * a * sthin * sthin;
      double SSin = rin * rin + a * a * cthin * cthin;
      double lnh;
      if (r >= rin) {
        lnh = 0.5 * log((1. + sqrt(1. + 4. * (l * l * SS * SS) * DD / (AA * sth * AA * sth))) / (SS * DD / AA)) -
              0.5 * sqrt(1. + 4. * (l * l * SS * SS) * DD / (AA * AA * sth * sth)) - 2. * a * r * l / AA -
              (0.5 * log((1. + sqrt(1. + 4. * (l * l * SSin * SSin) * DDin / (AAin * AAin * sthin * sthin))) /
                         (SSin * DDin / AAin)) -
               0.5 * sqrt(1. + 4. * (l * l * SSin * SSin) * DDin / (AAin * AAin * sthin * sthin)) -
               2. * a * rin * l / AAin);
      } else
        lnh = 1.;
      for (int k = 0; k < NPR; k++) pr[k] = 0.;
      if (lnh < 0. || r < rin) {
        /* fixup.c:1119-1196 set_atmosphere(-1, VELREL4) */
        pr[RHO] = g->rhomin * pow(r, -1.5);
        pr[UU] = g->uumin * pow(r, -2.5);
      } else {
        double hm1 = exp(lnh) - 1.;
        double rho = pow(hm1 * (gam - 1.) / (kappa * gam), 1. / (gam - 1.));
        double u = kappa * pow(rho, gam) / (gam - 1.);
        double expm2chi = SS * SS * DD / (AA * AA * sth * sth);
        double up1 = sqrt((-1. + sqrt(1. + 4. * l * l * expm2chi)) / 2.);
        double up = 2. * a * r * sqrt(1. + up1 * up1) / sqrt(AA * SS * DD) + sqrt(SS / AA) * up1 / sth;
        pr[RHO] = rho;
        pr[UU] = u * (1. + randfact * (urand(&rs) - 0.5));
        /* BL 4-velocity -> KS (transforms.c:109-160) -> KS' (mettometp) -> VELREL4 (ucon2pr) */
        double gbl[NDIM][NDIM], ucon[NDIM] = {0, 0., 0., up}, tmp[NDIM];
        gcov_bl(a, r, th, gbl);
        if (ucon_4vel(gbl, ucon)) { free(A); return 4; }
        double trans01 = (2. * r / (r * r - 2. * r + a * a)), trans31 = (a / (r * r - 2. * r + a * a));
        tmp[0] = ucon[0] + trans01 * ucon[1];
        tmp[1] = ucon[1];
        tmp[2] = ucon[2];
        tmp[3] = trans31 * ucon[1] + ucon[3];
        double dxdxp[NDIM][NDIM], idxdxp[NDIM][NDIM];
        dxdxprim(g, X, r, th, dxdxp);
        gauss_jordan_inverse(dxdxp, idxdxp);
        for (int jj = 0; jj < NDIM; jj++) {
          ucon[jj] = 0.;
          for (int kk = 0; kk < NDIM; kk++) ucon[jj] += idxdxp[jj][kk] * tmp[kk];
        }
        double alphasq = 1. / (-gn[0][0]);
        double gammaoalpha = ucon[0];
        for (int jj = 1; jj < NDIM; jj++) pr[U1 + jj - 1] = ucon[jj] + (alphasq * gn[0][jj]) * gammaoalpha;
      }
      (void)gc;
    }
  /* init.fishmon.c:283-311 normalize_densities */
  for (int i = 0; i < n1; i++)
    for (int j = 0; j < n2; j++) {
      double X[NDIM], r, th;
      coord(g, i, j, CENT, X);
      bl_coord(g, X, &r, &th);
      double *pr = p + zidx(g, i, j) * NPR;
      if (pr[RHO] > rhomax) rhomax = pr[RHO];
      if (pr[UU] > umax && r > rin) umax = pr[UU];
    }
  for (int i = 0; i < n1; i++)
    for (int j = 0; j < n2; j++) {
      double *pr = p + zidx(g, i, j) * NPR;
      pr[RHO] /= rhomax;
      pr[UU] /= rhomax;
    }
  umax /= rhomax;
  rhomax = 1.;
  /* init_atmosphere (init.fishmon.c:96-111) and 1-zone floors (fixup.c:274-377) */
  for (int i = 0; i < n1; i++)
    for (int j = 0; j < n2; j++) {
      double X[NDIM], r, th;
      coord(g, i, j, CENT, X);
      bl_coord(g, X, &r, &th);
      double *pr = p + zidx(g, i, j) * NPR;
      double ratm = g->rhomin * pow(r, -1.5);
      if (pr[RHO] < ratm) {
        pr[RHO] = ratm;
        pr[UU] = g->uumin * pow(r, -2.5);
        pr[U1] = pr[U2] = pr[U3] = 0.0;
      }
      double fr = g->rhomin * pow(r, -2.7), fu = g->uumin * pow(r, -2.7);
      if (fr < 1e-20) fr = 1e-20;
      if (fu < 1e-20) fu = 1e-20;
      if (fr > pr[RHO]) pr[RHO] = fr;
      if (fu > pr[UU]) pr[UU] = fu;
    }
  /* init.fishmon.c:222-246 init_vpot at CORN: A = max(rho_av/rhomax - 0.2, 0) */
  for (int i = -NG + 1; i < n1 + NG; i++)
    for (int j = -NG + 1; j < n2 + NG; j++) {
      double rho_av = 0.25 * (p[zidx(g, i, j) * NPR + RHO] + p[zidx(g, i - 1, j) * NPR + RHO] +
                              p[zidx(g, i, j - 1) * NPR + RHO] + p[zidx(g, i - 1, j - 1) * NPR + RHO]);
      double q = rho_av / rhomax - 0.2;
      A[zidx(g, i, j)] = (q > 0.) ? q : 0.;
    }
  /* initbase.c:565-618 vpot2field (interior + first ghost layer) */
  for (int i = -NG + 1; i < n1 + NG - 1; i++)
    for (int j = -NG + 1; j < n2 + NG - 1; j++) {
      double gd = gdet[zidx(g, i, j) * 4 + CENT];
      double *pr = p + zidx(g, i, j) * NPR;
      pr[B1] = (-A[zidx(g, i, j)] + A[zidx(g, i, j + 1)] - A[zidx(g, i + 1, j)] + A[zidx(g, i + 1, j + 1)]) /
               (2. * g->dx2 * gd);
      pr[B2] = -(-A[zidx(g, i, j)] - A[zidx(g, i, j + 1)] + A[zidx(g, i + 1, j)] + A[zidx(g, i + 1, j + 1)]) /
               (2. * g->dx1 * gd);
      pr[B3] = 0.;
    }
  free(A);
  /* init.fishmon.c:315-375 normalize_field */
  for (int i = 0; i < n1; i++)
    for (int j = 0; j < n2; j++) {
      size_t z = zidx(g, i, j);
      double bsq;
      state_bsq(p + z * NPR, (const double(*)[4])(gcov + (z * 4 + CENT) * 16),
                (const double(*)[4])(gcon + (z * 4 + CENT) * 16), &bsq);
      if (bsq > bsq_max) bsq_max = bsq;
    }
  if (bsq_max > 0) {
    double beta_act = (gam - 1.) * umax / (0.5 * bsq_max);
    double norm = sqrt(beta_act / beta);
    for (int i = -NG; i < n1 + NG; i++)
      for (int j = -NG; j < n2 + NG; j++) {
        double *pr = p + zidx(g, i, j) * NPR;
        pr[B1] *= norm; pr[B2] *= norm; pr[B3] *= norm;
      }
  }
  return 0;
}

/* Relativistic Orszag-Tang vortex: rho=25/(36 pi), p=5/(12 pi), v=(-0.5 sin 2pi y,
   0.5 sin 2pi x), B=(-B0 sin 2pi y, B0 sin 4pi x), B0=1/sqrt(4 pi).  In Minkowski
   the relative 4-velocity is gamma*v. */
int harm_init_orszag_tang(const harm_grid *g, double gam, double *p) {
  if (!g || g->metric != HARM_METRIC_MINK) return 1;
  double rho = 25. / (36. * M_PI), pres = 5. / (12. * M_PI), B0 = 1. / sqrt(4. * M_PI);
  for (int i = -NG; i < g->n1 + NG; i++)
    for (int j = -NG; j < g->n2 + NG; j++) {
      double X[NDIM];
      coord(g, i, j, CENT, X);
      double x = X[1], y = X[2];
      double v1 = -0.5 * sin(2. * M_PI * y), v2 = 0.5 * sin(2. * M_PI * x);
      double gamma = 1. / sqrt(1. - v1 * v1 - v2 * v2);
      double *pr = p + zidx(g, i, j) * NPR;
      pr[RHO] = rho;
      pr[UU] = pres / (gam - 1.);
      pr[U1] = gamma * v1;
      pr[U2] = gamma * v2;
      pr[U3] = 0.;
      pr[B1] = -B0 * sin(2. * M_PI * y);
      pr[B2] = B0 * sin(4. * M_PI * x);
      pr[B3] = 0.;
    }
  return 0;
}

/* Split monopole for the force-free equations (init.bz.c:146-155 and
   init.c:187-196, PROBLEMTYPE 0): rho = u = 0, zero velocity,
   B^r = +-B0 gdet(ihor, j) / gdet(i, j) with the sign of the hemisphere
   (theta < pi/2 positive).  The reference leaves `horizoni` undefined; it is
   the horizon zone index setihor() gives (coord.c:817-822), where setRin()
   puts the horizon on the inner face. */
int harm_init_monopole(const harm_grid *g, const double *gdet, double B0, double *p) {
  if (!g || g->metric != HARM_METRIC_KS || !gdet || !p) return 1;
  int ihor = (g->n1 * (0.1) > 10) ? 10 : (int)((double)g->n1 * (double)(0.1));
  for (int i = -NG; i < g->n1 + NG; i++)
    for (int j = -NG; j < g->n2 + NG; j++) {
      double X[NDIM], r, th;
      coord(g, i, j, CENT, X);
      bl_coord(g, X, &r, &th);
      double *pr = p + zidx(g, i, j) * NPR;
      double gh = gdet[zidx(g, ihor, j) * 4 + CENT], gz = gdet[zidx(g, i, j) * 4 + CENT];
      pr[RHO] = pr[UU] = 0.0;
      pr[U1] = pr[U2] = pr[U3] = 0.0;
      pr[B2] = pr[B3] = 0;
      if (th < M_PI * 0.5) pr[B1] = B0 * gh / (gz);
      else pr[B1] = -B0 * gh / (gz);
    }
  return 0;
}

/* initbase.c:736-780 setflux() and :788-980 setjetflux() on one rank
   (mycpupos = 0, ncpux = 1): region 0 the grid, region 1 the funnel
   theta < pi/2 - h_over_r_jet, region 2 the funnel theta > pi/2 + h_over_r_jet,
   with the reference's edge indices (region 2 starts at the zone below the
   first x2 face past its edge). */
int harm_ener_regions(const harm_grid *g, double h_over_r_jet, int enerpos[HARM_NUMENERREGIONS][4],
                      int doflux[HARM_NUMENERREGIONS][4]) {
  if (!g || !enerpos || !doflux || g->n1 < 1 || g->n2 < 1) return 1;
  const int N1 = g->n1, N2 = g->n2;
  double X[NDIM], r, th;
  int *ep, *df, jetedge;
  /* setflux */
  ep = enerpos[0]; df = doflux[0];
  ep[HARM_X1DN] = 0; ep[HARM_X1UP] = N1 - 1; ep[HARM_X2DN] = 0; ep[HARM_X2UP] = N2 - 1;
  df[HARM_X1DN] = 0; df[HARM_X1UP] = N1; df[HARM_X2DN] = 0; df[HARM_X2UP] = N2;
  /* setjetflux */
  const double thetajet = M_PI * 0.5 - h_over_r_jet;
  coord(g, 0, 0, FACE2, X);
  bl_coord(g, X, &r, &th);
  const double startth = th;
  coord(g, 0, N2, FACE2, X);
  bl_coord(g, X, &r, &th);
  const double endth = th;
  if ((fabs(startth - thetajet) / thetajet < 1E-8) || (fabs(endth - thetajet) / thetajet < 1E-8) ||
      (fabs(startth - (M_PI - thetajet)) / thetajet < 1E-8) || (fabs(endth - (M_PI - thetajet)) / thetajet < 1E-8))
    return 2;
  /* inner jet */
  ep = enerpos[1]; df = doflux[1];
  if ((startth <= thetajet) && (endth <= thetajet)) {
    ep[HARM_X1DN] = 0; ep[HARM_X1UP] = N1 - 1; ep[HARM_X2DN] = 0; ep[HARM_X2UP] = N2 - 1;
    jetedge = -100;
  } else if ((startth < thetajet) && (endth > thetajet)) {
    ep[HARM_X1DN] = 0; ep[HARM_X1UP] = N1 - 1; ep[HARM_X2DN] = 0; ep[HARM_X2UP] = N2 - 1;
    jetedge = -100;
    for (int j = 0; j <= N2; j++) {
      coord(g, 0, j, FACE2, X);
      bl_coord(g, X, &r, &th);
      if (th > thetajet) { ep[HARM_X2UP] = j - 1; jetedge = j; break; }
    }
  } else if ((startth >= thetajet) && (endth >= thetajet)) {
    ep[HARM_X1DN] = ep[HARM_X1UP] = ep[HARM_X2DN] = ep[HARM_X2UP] = -100;
    jetedge = -100;
  } else {
    return 2;
  }
  df[HARM_X1DN] = (ep[HARM_X1DN] != -100) ? 0 : -100;
  df[HARM_X1UP] = (ep[HARM_X1DN] != -100) ? N1 : -100;
  df[HARM_X2DN] = 0;
  df[HARM_X2UP] = (jetedge != -100) ? jetedge : -100;
  /* outer jet */
  ep = enerpos[2]; df = doflux[2];
  if ((startth <= M_PI - thetajet) && (endth <= M_PI - thetajet)) {
    ep[HARM_X1DN] = ep[HARM_X1UP] = ep[HARM_X2DN] = ep[HARM_X2UP] = -100;
    jetedge = -100;
  } else if ((startth < M_PI - thetajet) && (endth > M_PI - thetajet)) {
    ep[HARM_X1DN] = 0; ep[HARM_X1UP] = N1 - 1; ep[HARM_X2DN] = 0;
    jetedge = -100;
    for (int j = 0; j <= N2; j++) {
      coord(g, 0, j, FACE2, X);
      bl_coord(g, X, &r, &th);
      if (th > M_PI - thetajet) { ep[HARM_X2DN] = j - 1; jetedge = j - 1; break; }
    }
    ep[HARM_X2UP] = N2 - 1;
  } else if ((startth >= M_PI - thetajet) && (endth >= M_PI - thetajet)) {
    ep[HARM_X1DN] = 0; ep[HARM_X1UP] = N1 - 1; ep[HARM_X2DN] = 0; ep[HARM_X2UP] = N2 - 1;
    jetedge = -100;
  } else {
    return 2;
  }
  df[HARM_X1DN] = (ep[HARM_X1DN] != -100) ? 0 : -100;
  df[HARM_X1UP] = (ep[HARM_X1DN] != -100) ? N1 : -100;
  df[HARM_X2DN] = (jetedge != -100) ? jetedge : -100;
  df[HARM_X2UP] = N2;
  return 0;
}

/* coord() and bl_coord() at a zone centre (coord.c:920-939, 334-455): the
   coordinate columns of the dump (dump.c:102-103) */
int harm_zone_coords(const harm_grid *g, int i, int j, double *X1, double *X2, double *r, double *th) {
  if (!g) return 1;
  double X[NDIM];
  coord(g, i, j, CENT, X);
  bl_coord(g, X, r, th);
  *X1 = X[1];
  *X2 = X[2];
  return 0;
}
