// adjoint_shim.cpp -- TEST INFRASTRUCTURE.  A host-only build of the
// product's per-zone physics (harm_device.h), its reverse-mode counterparts
// (harm_adjoint.h) and the ENO conversion (harm_eno.h), so that the CPU tests
// can check each vector-Jacobian product against central finite differences
// of the forward function, and the conversion against the oracle's.  Built by
// the tests with
//   hipcc -x hip --offload-host-only -O2 -ffp-contract=off -shared -fPIC
#include "../harm_ramt_amd/csrc/harm_adjoint.h"
#include "../harm_ramt_amd/csrc/harm_eno.h"

using namespace harm;

static KParams mkP(const double* prm) {
  // prm: gam, eomtype, fluxmethod, gammamax, lim, cour, dx1, dx2
  KParams P = {};
  P.gam = prm[0];
  P.eomtype = (int)prm[1];
  P.fluxmethod = (int)prm[2];
  P.gammamax = prm[3];
  P.lim = (int)prm[4];
  P.cour = prm[5];
  P.dx1 = prm[6];
  P.dx2 = prm[7];
  return P;
}

// gcov[16], gcon[16] full, gdet; ndd for the given direction
static GeoU mkG(const double* gcov, const double* gcon, double gdet, int dir) {
  GeoU g;
  g.g00 = gcov[0]; g.g01 = gcov[1]; g.g02 = gcov[2]; g.g03 = gcov[3];
  g.g11 = gcov[5]; g.g12 = gcov[6]; g.g13 = gcov[7];
  g.g22 = gcov[10]; g.g23 = gcov[11]; g.g33 = gcov[15];
  g.n00 = gcon[0]; g.n01 = gcon[1]; g.n02 = gcon[2]; g.n03 = gcon[3];
  g.ndd = gcon[dir * 5];
  g.gdet = gdet;
  return g;
}

extern "C" {

void shim_primtoflux(const double* prm, const double* gcov, const double* gcon, double gdet, const double* pr,
                     int dir, double* F) {
  KParams P = mkP(prm);
  GeoU g = mkG(gcov, gcon, gdet, dir);
  State q;
  get_state(pr, g, q);
  primtoflux(P, pr, q, dir, gdet, F);
}

void shim_primtoflux_vjp(const double* prm, const double* gcov, const double* gcon, double gdet, const double* pr,
                         int dir, const double* Fb, double* pb) {
  KParams P = mkP(prm);
  GeoU g = mkG(gcov, gcon, gdet, dir);
  State q;
  get_state(pr, g, q);
  StateBar qb;
  statebar_zero(qb);
  for (int k = 0; k < NPR; k++) pb[k] = 0.;
  primtoflux_rev(P, pr, q, dir, gdet, Fb, qb, pb);
  get_state_rev(pr, g, q, qb, pb);
}

void shim_vchar(const double* prm, const double* gcov, const double* gcon, double gdet, const double* pr, int dir,
                double* v) {
  KParams P = mkP(prm);
  GeoU g = mkG(gcov, gcon, gdet, dir);
  State q;
  get_state(pr, g, q);
  vchar(P, pr, q, dir, g, v[0], v[1]);
}

int shim_vchar_vjp(const double* prm, const double* gcov, const double* gcon, double gdet, const double* pr,
                   int dir, const double* vb, double* pb) {
  KParams P = mkP(prm);
  GeoU g = mkG(gcov, gcon, gdet, dir);
  State q;
  get_state(pr, g, q);
  StateBar qb;
  statebar_zero(qb);
  for (int k = 0; k < NPR; k++) pb[k] = 0.;
  int ns = vchar_rev(P, pr, q, dir, g, vb[0], vb[1], qb, pb);
  get_state_rev(pr, g, q, qb, pb);
  return ns;
}

void shim_face_flux(const double* prm, const double* gcov, const double* gcon, double gdet, const double* pl,
                    const double* pr, int dir, double* F) {
  KParams P = mkP(prm);
  GeoU g = mkG(gcov, gcon, gdet, dir);
  double a[NPR], b[NPR], dtij;
  int fail = 0;
  for (int k = 0; k < NPR; k++) { a[k] = pl[k]; b[k] = pr[k]; }
  if (dir == 1) face_flux<1>(P, a, b, g, F, dtij, fail);
  else face_flux<2>(P, a, b, g, F, dtij, fail);
}

int shim_face_flux_vjp(const double* prm, const double* gcov, const double* gcon, double gdet, const double* pl,
                       const double* pr, int dir, const double* Fb, double* plb, double* prb) {
  KParams P = mkP(prm);
  GeoU g = mkG(gcov, gcon, gdet, dir);
  for (int k = 0; k < NPR; k++) plb[k] = prb[k] = 0.;
  if (dir == 1) return face_flux_rev<1>(P, pl, pr, g, Fb, plb, prb);
  return face_flux_rev<2>(P, pl, pr, g, Fb, plb, prb);
}

void shim_source(const double* prm, const double* gcov, const double* gcon, double gdet, const double* conn,
                 const double* pr, double* dU) {
  KParams P = mkP(prm);
  GeoU g = mkG(gcov, gcon, gdet, 1);
  source_conn(P, pr, g, conn, dU);
}

void shim_source_vjp(const double* prm, const double* gcov, const double* gcon, double gdet, const double* conn,
                     const double* pr, const double* dUb, double* pb) {
  KParams P = mkP(prm);
  GeoU g = mkG(gcov, gcon, gdet, 1);
  for (int k = 0; k < NPR; k++) pb[k] = 0.;
  source_rev(P, pr, g, conn, dUb, pb);
}

double shim_bsq(const double* gcov, const double* gcon, double gdet, const double* pr) {
  GeoU g = mkG(gcov, gcon, gdet, 1);
  State q;
  get_state(pr, g, q);
  return dot4(q.bcon, q.bcov);
}

void shim_bsq_vjp(const double* gcov, const double* gcon, double gdet, const double* pr, double* pb) {
  GeoU g = mkG(gcov, gcon, gdet, 1);
  for (int k = 0; k < NPR; k++) pb[k] = 0.;
  bsq_rev(pr, g, 2.0, pb);  // d(b^2)/dp
}

int shim_inversion_vjp(const double* prm, const double* gcov, const double* gcon, double gdet, const double* p,
                       const double* pbar, double* Ub) {
  KParams P = mkP(prm);
  GeoU g = mkG(gcov, gcon, gdet, 1);
  return inversion_rev(P, p, g, pbar, Ub);
}

double shim_slope3(int lim, double yl, double yc, double yr) { return slope3(lim, yl, yc, yr); }

void shim_slope3_coef(int lim, double yl, double yc, double yr, double* c) {
  slope3_coef(lim, yl, yc, yr, c[0], c[1], c[2]);
}

// the device conversion of the value at pos of a line of n values
double shim_eno_line(int type, const double* line, int n, int pos) {
  double v[9];
  for (int o = 0; o < 9; o++) {
    const int q = pos - 4 + o;
    v[o] = (q >= 0 && q < n) ? line[q] : 0.;
  }
  return type == ENO_A2C ? eno_cvt_line<ENO_A2C>(v, pos, n) : eno_cvt_line<ENO_C2A>(v, pos, n);
}

double shim_eno_weight(int type, int order, int s, int d) { return eno_weight(type, order, s, d); }

}  // extern "C"
