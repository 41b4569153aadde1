// harm_kernels.hip -- HIP kernels for gfx950 and the C-ABI (include/harm_amd.h)
// of HARM's conservative timestep.
//
// One RK substage is five kernels on one stream:
//   k_flux<1>, k_flux<2>  fluxcalc() per face: reconstruction, limit_gamma,
//                         states, fluxes, vchar, HLL/LAXF, fix_flux edge rules,
//                         wavefront-reduced Courant minimum      (step_ch.c:741)
//   k_emf                 flux_ct() corner EMFs (Toth)            (step_ch.c:1159)
//   k_update              advance() per zone: U(pi), source(pb), flux
//                         divergence with the CT fluxes formed from the EMFs,
//                         RK combination, Newton con-to-prim, fixup1zone
//                                                                  (advance.c:46)
//   k_bound_x1, k_bound_x2  bound_prim() ghost fill                 (bounds.fishmon.c)
// and one single-thread kernel after the last substage updates t and dt on
// the device (step_ch.c:174-183), so a step never synchronises with the host
// and a whole step is captured once into a HIP graph and replayed.
//
// Device layout (HBM):
//   primitive-shaped arrays: the reference's AoS p[N1M][N2M][8] (64 B/zone);
//   gF[2][nz][16]  FACE1/FACE2: gcov upper 10, gcon^{00..03}, gcon^{dd}, gdet
//   gC[nz][40]     CENT: gcov 16, gcon 16, gdet, floor rho, floor u, gdet(CORN)
//   gConn[nz][64]  CENT connection Gamma^k_{lj} as conn[k][l][j]
#include <hip/hip_runtime.h>
#include <string.h>
#include <stdio.h>
#include <stdarg.h>
#include <stdlib.h>
#include <vector>
#include <mutex>
#include "harm_device.h"
#include "../../include/harm_amd.h"

using namespace harm;

namespace {

constexpr int NG = 3;
constexpr int GF = 16;   // doubles per face geometry record
constexpr int GCN = 40;  // doubles per CENT geometry record

struct DevScalars {
  unsigned long long ndt_acc[2];  // running min of dtij bits, per direction
  double ndt_last[2];             // snapshot taken after both flux kernels
  double t, dt, tf;
  int failed;
  int nfail;
  unsigned long long nstroke;
};

thread_local char g_err[512];

int seterr(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return 1;
}

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) return seterr("%s: %s", #x, hipGetErrorString(e_));      \
  } while (0)

__device__ __forceinline__ int zidx(const KParams& P, int i, int j) { return (i + NG) * P.N2M + (j + NG); }

__device__ __forceinline__ void load8(const double* __restrict__ a, int z, double* v) {
  const double2* a2 = reinterpret_cast<const double2*>(a + (size_t)z * 8);
#pragma unroll
  for (int m = 0; m < 4; m++) {
    double2 t = a2[m];
    v[2 * m] = t.x;
    v[2 * m + 1] = t.y;
  }
}
__device__ __forceinline__ void store8(double* __restrict__ a, int z, const double* v) {
  double2* a2 = reinterpret_cast<double2*>(a + (size_t)z * 8);
#pragma unroll
  for (int m = 0; m < 4; m++) a2[m] = make_double2(v[2 * m], v[2 * m + 1]);
}

__device__ __forceinline__ void load_geo_face(const double* __restrict__ gF, int z, GeoU& g) {
  const double2* a = reinterpret_cast<const double2*>(gF + (size_t)z * GF);
  double2 t0 = a[0], t1 = a[1], t2 = a[2], t3 = a[3], t4 = a[4], t5 = a[5], t6 = a[6], t7 = a[7];
  g.g00 = t0.x; g.g01 = t0.y; g.g02 = t1.x; g.g03 = t1.y; g.g11 = t2.x; g.g12 = t2.y;
  g.g13 = t3.x; g.g22 = t3.y; g.g23 = t4.x; g.g33 = t4.y;
  g.n00 = t5.x; g.n01 = t5.y; g.n02 = t6.x; g.n03 = t6.y; g.ndd = t7.x; g.gdet = t7.y;
}

// upper-triangle view of a full CENT record
__device__ __forceinline__ void geo_from_cent(const double* c, GeoU& g) {
  g.g00 = c[0]; g.g01 = c[1]; g.g02 = c[2]; g.g03 = c[3];
  g.g11 = c[5]; g.g12 = c[6]; g.g13 = c[7];
  g.g22 = c[10]; g.g23 = c[11]; g.g33 = c[15];
  g.n00 = c[16]; g.n01 = c[17]; g.n02 = c[18]; g.n03 = c[19];
  g.ndd = c[21];
  g.gdet = c[32];
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = dmin(v, __shfl_xor(v, off));
  return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// ---------------------------------------------------------------------------
// fluxcalc(): one thread per face.  dir 1 faces: i in [0,N1], j in [-1,N2];
// dir 2 faces: i in [-1,N1], j in [0,N2] (FZLOOP(-jdel,-idel), mympi.h:287).
// ---------------------------------------------------------------------------
template <int DIR>
__global__ __launch_bounds__(256) void k_flux(KParams P, const double* __restrict__ pr, double* __restrict__ F,
                                              const double* __restrict__ gF, DevScalars* __restrict__ ds) {
  constexpr int idel = (DIR == 1), jdel = (DIR == 2);
  const int j = blockIdx.x * blockDim.x + threadIdx.x - idel;
  const int i = blockIdx.y * blockDim.y + threadIdx.y - jdel;
  const bool valid = (i <= P.n1) && (j <= P.n2);
  double dtij = 1.e9;
  int fail = 0;
  if (valid) {
    // stencil along DIR: zones m = -3..+2 relative to the face's right zone
    const int zstride = (DIR == 1) ? P.N2M : 1;
    const int z0 = zidx(P, i, j);
    double p_l[NPR], p_r[NPR];
    if (P.lim == 4) {  // PARA: pright[i-1], pleft[i] (step_ch.c:896-903)
      double s[6][NPR];
#pragma unroll
      for (int m = 0; m < 6; m++) load8(pr, z0 + (m - 3) * zstride, s[m]);
#pragma unroll
      for (int k = 0; k < NPR; k++) {
        double l, r, l2, r2;
        para4(s[0][k], s[1][k], s[2][k], s[3][k], s[4][k], l, r);
        para4(s[1][k], s[2][k], s[3][k], s[4][k], s[5][k], l2, r2);
        p_l[k] = r;
        p_r[k] = l2;
      }
    } else {  // slope limiters: p[i-1]+dq/2, p[i]-dq/2
      double s[5][NPR];
#pragma unroll
      for (int m = 0; m < 5; m++) load8(pr, z0 + (m - 3) * zstride, s[m]);
      double s5[NPR];
      load8(pr, z0 + 2 * zstride, s5);
#pragma unroll
      for (int k = 0; k < NPR; k++) {
        double dql, dqr;
        if (P.lim == 90) {
          dql = csslope(k, s[0][k], s[1][k], s[2][k], s[3][k], s[4][k]);
          dqr = csslope(k, s[1][k], s[2][k], s[3][k], s[4][k], s5[k]);
        } else {
          dql = slope3(P.lim, s[1][k], s[2][k], s[3][k]);
          dqr = slope3(P.lim, s[2][k], s[3][k], s[4][k]);
        }
        p_l[k] = s[2][k] + 0.5 * dql;
        p_r[k] = s[3][k] - 0.5 * dqr;
      }
    }
    GeoU g;
    load_geo_face(gF, z0, g);
    // BDIRCONT (step_ch.c:944-947)
    {
      constexpr int kb = B1 + DIR - 1;
      double bavg = 0.5 * (p_l[kb] + p_r[kb]);
      p_l[kb] = bavg;
      p_r[kb] = bavg;
    }
    if (limit_gamma(P.gammamax, p_l, g) >= 1) fail = 1;
    if (limit_gamma(P.gammamax, p_r, g) >= 1) fail = 1;
    State sl, sr;
    get_state(p_l, g, sl);
    get_state(p_r, g, sr);
    double F_l[NPR], F_r[NPR], U_l[NPR], U_r[NPR];
    primtoflux(P, p_l, sl, DIR, g.gdet, F_l);
    primtoflux(P, p_r, sr, DIR, g.gdet, F_r);
    primtoflux(P, p_l, sl, 0, g.gdet, U_l);
    primtoflux(P, p_r, sr, 0, g.gdet, U_r);
    double cmax_l, cmin_l, cmax_r, cmin_r;
    if (vchar(P, p_l, sl, DIR, g, cmax_l, cmin_l)) fail = 1;
    if (vchar(P, p_r, sr, DIR, g, cmax_r, cmin_r)) fail = 1;
    // cminmax_calc (step_ch.c:1095-1148)
    double lmin = dmin(cmin_l, cmin_r), lmax = dmax(cmax_l, cmax_r);
    double cmin = fabs(dmax(0., -lmin)), cmax = fabs(dmax(0., lmax));
    double ctop = dmax(cmax, cmin);
    double hllfrac = (P.fluxmethod == 1) ? 1.0 : 0.0;
    double laxffrac = (P.fluxmethod == 0) ? 1.0 : 0.0;
    double Fo[NPR];
    if (cmax + cmin != 0.0) {
#pragma unroll
      for (int k = 0; k < NPR; k++)
        Fo[k] = hllfrac * ((cmax * F_l[k] + cmin * F_r[k] - cmax * cmin * (U_r[k] - U_l[k])) / (cmax + cmin)) +
                laxffrac * (0.5 * (F_l[k] + F_r[k] - ctop * (U_r[k] - U_l[k])));
    } else {
#pragma unroll
      for (int k = 0; k < NPR; k++) Fo[k] = (0.5 * (F_l[k] + F_r[k] - ctop * (U_r[k] - U_l[k])));
    }
    double cm = dmax(cmax, cmin);
    dtij = P.cour * ((DIR == 1) ? P.dx1 : P.dx2) / cm;
    if (!(dtij < 1.e9)) dtij = 1.e9;  // the reference's min ignores NaN and values >= 1e9
    // fix_flux() edge rules (fixup.c:1357-1410), applied by the face that owns them
    if (P.fixupflux) {
      if (DIR == 1) {
        if (i == 0 && P.bc[HARM_X1DN] == HARM_OUTFLOW && j >= -1 && j < P.n2 && Fo[RHO] > 0) Fo[RHO] = 0;
        if (i == P.n1 && P.bc[HARM_X1UP] == HARM_OUTFLOW && j >= -1 && j < P.n2 && Fo[RHO] < 0) Fo[RHO] = 0;
      } else {
        bool polar_dn = (j == 0 && P.bc[HARM_X2DN] == HARM_POLARAXIS && i >= -1 && i < P.n1);
        bool polar_up = (j == P.n2 && P.bc[HARM_X2UP] == HARM_POLARAXIS && i >= -1 && i < P.n1);
        bool equator = (i >= 0 && i < P.n1 && j == (P.n2 - 1) / 2);
        if (polar_dn || polar_up || equator) {
#pragma unroll
          for (int k = 0; k < NPR; k++)
            if (k != U2) Fo[k] = 0.;
        }
      }
    }
    bool skip_b2 = false;
    if (DIR == 1 && P.fixupflux && i >= -1 && i < P.n1) {
      // F1[i][-1][B2] = -F1[i][0][B2] and F1[i][N2][B2] = -F1[i][N2-1][B2]
      if (j == 0 && P.bc[HARM_X2DN] == HARM_POLARAXIS) F[(size_t)zidx(P, i, -1) * 8 + B2] = -Fo[B2];
      if (j == P.n2 - 1 && P.bc[HARM_X2UP] == HARM_POLARAXIS) F[(size_t)zidx(P, i, P.n2) * 8 + B2] = -Fo[B2];
      if ((j == -1 && P.bc[HARM_X2DN] == HARM_POLARAXIS) || (j == P.n2 && P.bc[HARM_X2UP] == HARM_POLARAXIS))
        skip_b2 = true;
    }
    if (skip_b2) {
      double* Fz = F + (size_t)z0 * 8;
#pragma unroll
      for (int k = 0; k < NPR; k++)
        if (k != B2) Fz[k] = Fo[k];
    } else {
      store8(F, z0, Fo);
    }
  }
  // Courant limit: wavefront min, then one 64-bit atomic per wavefront
  double wm = wave_min(dtij);
  if ((threadIdx.x & 63) == 0 && wm < 1.e9)
    atomicMin(&ds->ndt_acc[DIR - 1], (unsigned long long)__double_as_longlong(wm));
  if (fail) atomicOr(&ds->failed, 1);
}

// ---------------------------------------------------------------------------
// flux_ct() EMF at corners i in [0,N1], j in [0,N2] (step_ch.c:1223-1234).
// Thread (0,0) also snapshots and resets the Courant accumulators.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_emf(KParams P, const double* __restrict__ F1, const double* __restrict__ F2,
                                             const double* __restrict__ gF, const double* __restrict__ gC,
                                             double* __restrict__ emf, DevScalars* __restrict__ ds) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i == 0 && j == 0) {
    ds->ndt_last[0] = __longlong_as_double((long long)ds->ndt_acc[0]);
    ds->ndt_last[1] = __longlong_as_double((long long)ds->ndt_acc[1]);
    ds->ndt_acc[0] = (unsigned long long)__double_as_longlong(1.e9);
    ds->ndt_acc[1] = (unsigned long long)__double_as_longlong(1.e9);
  }
  if (i > P.n1 || j > P.n2) return;
  const size_t nz = (size_t)(P.n1 + 2 * NG) * P.N2M;
  const double* gF1 = gF;
  const double* gF2 = gF + nz * GF;
  int z = zidx(P, i, j), zjm = zidx(P, i, j - 1), zim = zidx(P, i - 1, j);
  double ec = gC[(size_t)z * GCN + 35];
  double e1 = gF1[(size_t)z * GF + 15], e2 = gF1[(size_t)zjm * GF + 15];
  double e3 = gF2[(size_t)z * GF + 15], e4 = gF2[(size_t)zim * GF + 15];
  emf[z] = 0.25 * (ec) * (F1[(size_t)z * 8 + B2] / (e1) + F1[(size_t)zjm * 8 + B2] / (e2) -
                          F2[(size_t)z * 8 + B1] / (e3) - F2[(size_t)zim * 8 + B1] / (e4));
}

// flux_ct() rewrite of the B fluxes (step_ch.c:1405-1412); used by the
// standalone harm_flux_ct() entry point (the step forms them in k_update)
__global__ void k_ctapply(KParams P, double* __restrict__ F1, double* __restrict__ F2,
                          const double* __restrict__ emf) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i <= P.n1 && j <= P.n2 - 1) {
    int z = zidx(P, i, j);
    F1[(size_t)z * 8 + B1] = 0.;
    F1[(size_t)z * 8 + B2] = 0.5 * (emf[z] + emf[zidx(P, i, j + 1)]);
  }
  if (i <= P.n1 - 1 && j <= P.n2) {
    int z = zidx(P, i, j);
    F2[(size_t)z * 8 + B1] = -0.5 * (emf[z] + emf[zidx(P, i + 1, j)]);
    F2[(size_t)z * 8 + B2] = 0.;
  }
}

struct RKCoef {
  double cuf0, cuf1, cuf2, cn0, cn1, cn2;
  int timeorder, last;
};

// GRMHD inversion of one zone: Utoprimgen (step_ch.c:468-552) ->
// Utoprim_2d_final (utoprim_2d_final.c:26-101) -> Utoprim_new_body (:136-396).
// U is the evolved conserved state, pr the guess in / solution out.
__device__ int utoprimgen(const KParams& P, const double* U, const double* gc, double* pr, int& niter) {
  const double* gcov = gc;
  const double* gcon = gc + 16;
  double gdet = gc[32];
  double Ug[NPR];
#pragma unroll
  for (int k = 0; k < NPR; k++) Ug[k] = U[k] / gdet;
  Ug[UU] += -Ug[RHO];
  // Utoprim_2d_final
  double prim[NPR];
#pragma unroll
  for (int k = 0; k < NPR; k++) prim[k] = pr[k];
  prim[B1] = Ug[B1];
  prim[B2] = Ug[B2];
  prim[B3] = Ug[B3];
  double alpha = 1.0 / sqrt(-gcon[0]);
  double Ut[NPR], pt[NPR];
#pragma unroll
  for (int k = 0; k < NPR; k++) Ut[k] = alpha * Ug[k];
#pragma unroll
  for (int k = 0; k < B1; k++) pt[k] = prim[k];
  // Utoprim_new_body on (Ut, pt)
  int pflag = 0, retval = 0;
  niter = 0;
  double Bcon[4] = {0., Ut[B1], Ut[B2], Ut[B3]}, Bcov[4], Qcov[4], Qcon[4], ncov[4], ncon[4];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    Bcov[a] = 0.;
#pragma unroll
    for (int b = 0; b < 4; b++) Bcov[a] += gcov[a * 4 + b] * Bcon[b];
  }
#pragma unroll
  for (int a = 0; a < 4; a++) Qcov[a] = Ut[UU + a];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    Qcon[a] = 0.;
#pragma unroll
    for (int b = 0; b < 4; b++) Qcon[a] += gcon[a * 4 + b] * Qcov[b];
  }
  U2P s;
  s.gam = P.gam;
  s.gammafail = P.gammafail;
  s.Bsq = 0.;
#pragma unroll
  for (int a = 1; a < 4; a++) s.Bsq += Bcon[a] * Bcov[a];
  double QdotB = 0.;
#pragma unroll
  for (int a = 0; a < 4; a++) QdotB += Qcov[a] * Bcon[a];
  s.QdotBsq = QdotB * QdotB;
  double lapse = sqrt(-1. / gcon[0]);
  ncov[0] = -lapse;
  ncov[1] = 0.;
  ncov[2] = 0.;
  ncov[3] = 0.;
#pragma unroll
  for (int a = 0; a < 4; a++) {
    ncon[a] = 0.;
#pragma unroll
    for (int b = 0; b < 4; b++) ncon[a] += gcon[a * 4 + b] * ncov[b];
  }
  s.Qdotn = Qcon[0] * ncov[0];
  double Qsq = 0.;
#pragma unroll
  for (int a = 0; a < 4; a++) Qsq += Qcov[a] * Qcon[a];
  s.Qtsq = Qsq + s.Qdotn * s.Qdotn;
  s.D = Ut[RHO];
  const double gammafail = P.gammafail;
  const double UTSQ_TOO_BIG = ((gammafail - 1.0) * (gammafail - 1.0));
  const double VSQ_TOO_BIG = (1.0 - 1.0 / ((gammafail - 1.0) + 1.0));
  const double GAMMASQ_TOO_BIG = (gammafail * gammafail);
  const double MAXNEGUTSQ = 1E-10;
  const double MAXNEGVSQ = (1.0 - 1.0 / (MAXNEGUTSQ + 1.0));
  double utsq = 0.;
#pragma unroll
  for (int a = 1; a < 4; a++)
#pragma unroll
    for (int b = 1; b < 4; b++) utsq += gcov[a * 4 + b] * pt[U1 + a - 1] * pt[U1 + b - 1];
  if ((utsq < 0.) && (fabs(utsq) < MAXNEGUTSQ)) utsq = 0.0;
  bool solved = false;
  double W = 0.;
  if (utsq < 0. || utsq > UTSQ_TOO_BIG) {
    pflag = 2;
    retval = 1;
  } else {
    double gammasq = 1. + utsq;
    double gamma = sqrt(gammasq);
    double rho0 = s.D / gamma;
    double u = pt[UU];
    double p = (P.gam - 1.) * u;
    double w = rho0 + u + p;
    double W_last = w * gammasq;
    int i_increase = 0;
    while (((W_last * W_last * W_last * (W_last + 2. * s.Bsq) - s.QdotBsq * (2. * W_last + s.Bsq)) <=
            W_last * W_last * (s.Qtsq - s.Bsq * s.Bsq)) &&
           (i_increase < 10)) {
      W_last *= 10.;
      i_increase++;
    }
    s.wglobal = w;
    double x0 = fabs(W_last);
    double vsq0 = fabs(vsq_calc(s, W_last));
    double x1 = (vsq0 > 1.) ? VSQ_TOO_BIG : vsq0;
    retval = newton2d(s, x0, x1, niter);
    W = x0;
    if ((retval != 0) || (W == 1.e30)) {
      pflag = retval * 100 + 1;
    } else {
      double Wtest = W / s.wglobal;
      if (Wtest <= 0. || Wtest > GAMMASQ_TOO_BIG) {
        pflag = 3;
      } else {
        double vsq = vsq_calc(s, W);
        if ((vsq < 0.) && (fabs(vsq) < MAXNEGVSQ)) vsq = 0.0;
        if ((vsq > VSQ_TOO_BIG) || (vsq < 0.0)) {
          pflag = 4;
        } else {
          double gtmp = sqrt(1. - vsq);
          double gam2 = 1. / gtmp;
          double rho1 = s.D * gtmp;
          double w1 = W * (1. - vsq);
          double p1 = (P.gam - 1.) * (w1 - rho1) / P.gam;
          double u1 = w1 - (rho1 + p1);
          if ((rho1 <= 0.) || (u1 < 0.)) {
            if ((rho1 <= 0.) && (u1 >= 0.)) pflag = 5;
            if ((rho1 > 0.) && (u1 < 0.)) pflag = 6;
            if ((rho1 <= 0.) && (u1 < 0.)) pflag = 7;
          }
          pt[RHO] = rho1;
          pt[UU] = u1;
#pragma unroll
          for (int a = 1; a < 4; a++) {
            double Qtcon = Qcon[a] + ncon[a] * s.Qdotn;
            pt[U1 + a - 1] = gam2 / (W + s.Bsq) * (Qtcon + QdotB * Bcon[a] / W);
          }
          solved = true;
        }
      }
    }
  }
  (void)solved;
  // Utoprim_2d_final copies rho,u,u^i back when Utoprim_new_body returned 0
  if (retval == 0) {
#pragma unroll
    for (int k = 0; k < B1; k++) prim[k] = pt[k];
  }
#pragma unroll
  for (int k = 0; k < NPR; k++) pr[k] = prim[k];
  return pflag;
}

// ---------------------------------------------------------------------------
// advance_standard() per-zone loop (advance.c:114-181), zones i<N1, j<N2.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_update(KParams P, RKCoef C, const double* __restrict__ pi,
                                                const double* __restrict__ pb, double* __restrict__ pf,
                                                double* __restrict__ ulast, double* __restrict__ unew,
                                                const double* __restrict__ F1, const double* __restrict__ F2,
                                                const double* __restrict__ emf, const double* __restrict__ gC,
                                                const double* __restrict__ gConn, int* __restrict__ pflag,
                                                DevScalars* __restrict__ ds) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  int niter = 0, failz = 0, fail = 0;
  if (i < P.n1 && j < P.n2) {
    const int z = zidx(P, i, j);
    const double dt = ds->dt;
    double gc[36];
    {
      const double2* a = reinterpret_cast<const double2*>(gC + (size_t)z * GCN);
#pragma unroll
      for (int m = 0; m < 18; m++) {
        double2 t = a[m];
        gc[2 * m] = t.x;
        gc[2 * m + 1] = t.y;
      }
    }
    GeoU g;
    geo_from_cent(gc, g);
    double pin[NPR], pbz[NPR];
    load8(pi, z, pin);
    load8(pb, z, pbz);
    // Ui = U(pi)
    State q;
    double Ui[NPR];
    get_state(pin, g, q);
    primtoflux(P, pin, q, 0, g.gdet, Ui);
    // source(pb): T^j_k Gamma^k_{lj} contraction (phys.c:475-619)
    double dUg[NPR];
    {
      State qb;
      get_state(pbz, g, qb);
      double mhd[4][4];
#pragma unroll
      for (int jj = 0; jj < 4; jj++) mhd_calc(P, pbz, jj, qb, mhd[jj]);
      const double* cn = gConn + (size_t)z * 64;
      double dUc[4];
#pragma unroll
      for (int l = 0; l < 4; l++) {
        double acc = 0;
#pragma unroll
        for (int jj = 0; jj < 4; jj++)
#pragma unroll
          for (int kk = 0; kk < 4; kk++) acc += mhd[jj][kk] * cn[(kk * 4 + l) * 4 + jj];
        dUc[l] = acc;
      }
      dUg[RHO] = g.gdet * 0.;
      dUg[UU] = g.gdet * dUc[0];
      dUg[U1] = g.gdet * dUc[1];
      dUg[U2] = g.gdet * dUc[2];
      dUg[U3] = g.gdet * dUc[3];
      dUg[B1] = g.gdet * 0.;
      dUg[B2] = g.gdet * 0.;
      dUg[B3] = g.gdet * 0.;
    }
    // flux divergence (flux2dUavg, advance.c:204-265) with the CT B fluxes
    double f1a[NPR], f1b[NPR], f2a[NPR], f2b[NPR];
    load8(F1, z, f1a);
    load8(F1, zidx(P, i + 1, j), f1b);
    load8(F2, z, f2a);
    load8(F2, zidx(P, i, j + 1), f2b);
    if (P.fluxb == HARM_FLUXCTTOTH) {
      double e00 = emf[z], e01 = emf[zidx(P, i, j + 1)], e10 = emf[zidx(P, i + 1, j)],
             e11 = emf[zidx(P, i + 1, j + 1)];
      f1a[B1] = 0.;
      f1a[B2] = 0.5 * (e00 + e01);
      f1b[B1] = 0.;
      f1b[B2] = 0.5 * (e10 + e11);
      f2a[B1] = -0.5 * (e00 + e10);
      f2a[B2] = 0.;
      f2b[B1] = -0.5 * (e01 + e11);
      f2b[B2] = 0.;
    }
    const double idx1 = 1.0 / P.dx1, idx2 = 1.0 / P.dx2;
    double ul[NPR], un[NPR], Uf[NPR];
    if (C.timeorder == 0) {
#pragma unroll
      for (int k = 0; k < NPR; k++) ul[k] = un[k] = 0.0;
    } else {
      load8(ulast, z, ul);
      load8(unew, z, un);
    }
#pragma unroll
    for (int k = 0; k < NPR; k++) {
      double d1 = (-(f1b[k] - f1a[k]) * idx1);
      double d2 = (-(f2b[k] - f2a[k]) * idx2);
      double dUr = d1 + d2;
      // dUtoU (advance.c:273-286)
      double dtdU = (dUr + dUg[k]);
      Uf[k] = C.cuf0 * Ui[k] + C.cuf1 * ul[k] + C.cuf2 * dt * dtdU;
      un[k] += C.cn0 * Ui[k] + C.cn1 * dt * dtdU + C.cn2 * Uf[k];
      ul[k] = Uf[k];
    }
    store8(ulast, z, ul);
    store8(unew, z, un);
    // inversion, guess = pi (pf=pi assignment, advance.c:121)
    double pout[NPR];
#pragma unroll
    for (int k = 0; k < NPR; k++) pout[k] = pin[k];
    int pfl = utoprimgen(P, C.last ? un : Uf, gc, pout, niter);
    pflag[z] = pfl;
    failz = (pfl != 0);
    // fixup1zone (fixup.c:274-377): floors then limit_gamma
    if (P.eomtype == 0) {
      double fr = gc[33], fu = gc[34];
      if (fr < P.rhominlimit) fr = P.rhominlimit;
      if (fu < P.uuminlimit) fu = P.uuminlimit;
      if (fr > pout[RHO]) pout[RHO] = fr;
      if (fu > pout[UU]) pout[UU] = fu;
    }
    if (limit_gamma(P.gammamax, pout, g) >= 1) fail = 1;
    store8(pf, z, pout);
  }
  unsigned long long ns = wave_sum_u64((unsigned long long)niter);
  unsigned long long nf = wave_sum_u64((unsigned long long)failz);
  if ((threadIdx.x & 63) == 0) {
    if (ns) atomicAdd(&ds->nstroke, ns);
    if (nf) atomicAdd(&ds->nfail, (int)nf);
  }
  if (fail) atomicOr(&ds->failed, 1);
}

// inflow_check_rel4vel for x1 ghosts at CENT (fixup.c:1733-1841)
__device__ int inflow_check_x1(const KParams& P, int i, double* pr, const double* gc, const GeoU& g) {
  double alpha = 1. / sqrt(-g.n00);
  double gamma = gamma_calc(pr, g);
  double ucon1 = pr[U1] - gamma * (g.n01 * alpha * alpha) / alpha;
  bool dofix = ((i <= -1) && (P.bc[HARM_X1DN] == HARM_OUTFLOW) && (ucon1 > 0.)) ||
               ((i >= P.n1) && (P.bc[HARM_X1UP] == HARM_OUTFLOW) && (ucon1 < 0.));
  if (!dofix) return 0;
  pr[U1] /= gamma;
  pr[U2] /= gamma;
  pr[U3] /= gamma;
  double alpha2 = 1. / sqrt(-gc[16]);
  double betacon = gc[17] * alpha2 * alpha2;
  pr[U1] = betacon / alpha2;
  double vsq = 0.;
#pragma unroll
  for (int a = 1; a < 4; a++)
#pragma unroll
    for (int b = 1; b < 4; b++) vsq += gc[a * 4 + b] * pr[U1 + a - 1] * pr[U1 + b - 1];
  if (vsq < 0.0) {
    if (vsq > -NUMEPSILON * 100.0) vsq = 0.0;
    else return 1;
  }
  if (vsq >= 1.0) vsq = 1.0 - 1.0 / (P.gammamax * P.gammamax);
  gamma = 1. / sqrt(1. - vsq);
  pr[U1] *= gamma;
  pr[U2] *= gamma;
  pr[U3] *= gamma;
  return -1;
}

// x1 ghosts, j in [0,N2): thread (gi, j), gi in 0..5 -> i = -3..-1, N1..N1+2
__global__ void k_bound_x1(KParams P, double* __restrict__ pr, const double* __restrict__ gC,
                           DevScalars* __restrict__ ds) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int gi = blockIdx.y;
  if (j >= P.n2) return;
  const bool dn = gi < NG;
  const int i = dn ? (gi - NG) : (P.n1 + gi - NG);
  const int bc = dn ? P.bc[HARM_X1DN] : P.bc[HARM_X1UP];
  double v[NPR];
  int fail = 0;
  if (bc == HARM_PERIODIC) {
    load8(pr, zidx(P, dn ? i + P.n1 : i - P.n1, j), v);
    store8(pr, zidx(P, i, j), v);
  } else if (bc == HARM_OUTFLOW) {
    load8(pr, zidx(P, dn ? 0 : P.n1 - 1, j), v);
    const int z = zidx(P, i, j);
    double gc[36];
#pragma unroll
    for (int m = 0; m < 36; m++) gc[m] = gC[(size_t)z * GCN + m];
    GeoU g;
    geo_from_cent(gc, g);
    if (inflow_check_x1(P, i, v, gc, g) >= 1) fail = 1;
    if (limit_gamma(P.gammamax, v, g) >= 1) fail = 1;
    store8(pr, z, v);
  }
  if (fail) atomicOr(&ds->failed, 1);
}

// x2 ghosts over i in [-3, N1+2] (corners included), then POLEDEATH zeroing of
// U2,B2 on j in [-3,0] and [N2-1,N2+2] (bounds.fishmon.c:263-355)
__global__ void k_bound_x2(KParams P, double* __restrict__ pr) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) - NG;
  const int gj = blockIdx.y;  // 0..7: 0..2 dn ghosts, 3 row 0, 4 row N2-1, 5..7 up ghosts
  if (i > P.n1 + NG - 1) return;
  const bool dn = gj <= 3;
  const int bc = dn ? P.bc[HARM_X2DN] : P.bc[HARM_X2UP];
  if (gj == 3 || gj == 4) {  // interior pole rows: only POLEDEATH
    if (bc == HARM_POLARAXIS && P.poledeath) {
      int j = (gj == 3) ? 0 : P.n2 - 1;
      double* a = pr + (size_t)zidx(P, i, j) * 8;
      a[U2] *= 0.;
      a[B2] *= 0.;
    }
    return;
  }
  const int j = dn ? (gj - NG) : (P.n2 + gj - 5);
  double v[NPR];
  if (bc == HARM_PERIODIC) {
    load8(pr, zidx(P, i, dn ? j + P.n2 : j - P.n2), v);
    store8(pr, zidx(P, i, j), v);
  } else if (bc == HARM_POLARAXIS) {
    int js = dn ? (0 + (0 - j - 1)) : ((P.n2 - 1) + ((P.n2 - 1) - j + 1));
    load8(pr, zidx(P, i, js), v);
    if (P.poledeath) {
      v[U2] *= 0.;
      v[B2] *= 0.;
    }
    store8(pr, zidx(P, i, j), v);
  }
}

// step_ch.c:174-183 (t, dt update) with advance.c:190 (ndt from the last substage)
__global__ void k_dt(KParams P, DevScalars* ds) {
  double ndt = P.defcon * 1. / (1. / ds->ndt_last[0] + 1. / ds->ndt_last[1]);
  double dt = ds->dt;
  ds->t += dt;
  if (ndt > P.safe * dt) ndt = P.safe * dt;
  dt = ndt;
  if (ds->t + dt > ds->tf) dt = ds->tf - ds->t;
  ds->dt = dt;
}

__global__ void k_utoprim_batch(KParams P, int n, const int* __restrict__ zone, const double* __restrict__ U,
                                double* __restrict__ pr, int* __restrict__ pfl, const double* __restrict__ gC) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int z = zidx(P, zone[2 * t], zone[2 * t + 1]);
  double gc[36];
#pragma unroll
  for (int m = 0; m < 36; m++) gc[m] = gC[(size_t)z * GCN + m];
  double u[NPR], p[NPR];
#pragma unroll
  for (int k = 0; k < NPR; k++) {
    u[k] = U[(size_t)t * 8 + k];
    p[k] = pr[(size_t)t * 8 + k];
  }
  int niter;
  pfl[t] = utoprimgen(P, u, gc, p, niter);
#pragma unroll
  for (int k = 0; k < NPR; k++) pr[(size_t)t * 8 + k] = p[k];
}

__global__ void k_pflag_expand(int nz, const int* __restrict__ pf, int* __restrict__ out) {
  int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= nz) return;
  out[(size_t)z * 5] = pf[z];
#pragma unroll
  for (int k = 1; k < 5; k++) out[(size_t)z * 5 + k] = 0;
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================

struct harm_ctx {
  harm_params prm;
  KParams kp;
  int N1M, N2M, nz;
  double* arr[HARM_NARRAYS];
  double* emf;
  int* pflag;
  double* gF;
  double* gC;
  double* gConn;
  DevScalars* ds;
  hipStream_t stream;
  hipStream_t own_stream;
  hipGraphExec_t graph;
  bool graph_ok;
  bool geometry_set;
  bool prof;
  hipEvent_t ev0, ev1;
  double ktime[HARM_K_COUNT];
  long long kcount[HARM_K_COUNT];
};

extern "C" const char* harm_last_error(void) { return g_err; }

static int check_prm(const harm_params* p) {
  if (!p) return seterr("null params");
  if (p->n1 < 4 || p->n2 < 4) return seterr("n1,n2 must be >= 4 (got %d,%d)", p->n1, p->n2);
  if ((long long)(p->n1 + 6) * (p->n2 + 6) > (1LL << 28)) return seterr("grid too large");
  if (p->eomtype != HARM_EOMGRMHD)
    return seterr("eomtype %d: only EOMGRMHD is implemented on the GPU path", p->eomtype);
  int l = p->lim;
  if (!(l == HARM_DONOR || l == HARM_VANL || l == HARM_MINM || l == HARM_MC || l == HARM_PARA ||
        l == HARM_CSSLOPE || l == HARM_NLIMCENT || l == HARM_NLIMUP || l == HARM_NLIMDOWN))
    return seterr("unknown slope limiter %d", l);
  if (p->fluxmethod != HARM_LAXFFLUX && p->fluxmethod != HARM_HLLFLUX)
    return seterr("unknown flux method %d", p->fluxmethod);
  if (p->fluxb != HARM_FLUXCTTOTH && p->fluxb != HARM_FLUXCTHLL)
    return seterr("FLUXB %d not supported (FLUXCTTOTH=1, FLUXCTHLL=0)", p->fluxb);
  if (p->timeorder < 1 || p->timeorder > 4) return seterr("timeorder must be 1..4");
  for (int k = 0; k < 4; k++)
    if (p->bc[k] != HARM_OUTFLOW && p->bc[k] != HARM_POLARAXIS && p->bc[k] != HARM_PERIODIC)
      return seterr("boundary type %d not supported", p->bc[k]);
  if ((p->bc[HARM_X1DN] == HARM_POLARAXIS) || (p->bc[HARM_X1UP] == HARM_POLARAXIS))
    return seterr("POLARAXIS is an x2 boundary");
  if ((p->bc[HARM_X2DN] == HARM_OUTFLOW) || (p->bc[HARM_X2UP] == HARM_OUTFLOW))
    return seterr("x2 OUTFLOW is not implemented (x2 boundaries: POLARAXIS or PERIODIC)");
  if ((p->bc[HARM_X1DN] == HARM_PERIODIC) != (p->bc[HARM_X1UP] == HARM_PERIODIC) ||
      (p->bc[HARM_X2DN] == HARM_PERIODIC) != (p->bc[HARM_X2UP] == HARM_PERIODIC))
    return seterr("periodic boundaries must be set on both sides");
  if (!(p->dx1 > 0) || !(p->dx2 > 0)) return seterr("dx1, dx2 must be positive");
  return 0;
}

extern "C" int harm_create(const harm_params* prm, harm_ctx** out) {
  if (!out) return seterr("null out");
  *out = nullptr;
  if (check_prm(prm)) return 1;
  harm_ctx* c = (harm_ctx*)calloc(1, sizeof(harm_ctx));
  if (!c) return seterr("out of host memory");
  c->prm = *prm;
  c->N1M = prm->n1 + 2 * NG;
  c->N2M = prm->n2 + 2 * NG;
  c->nz = c->N1M * c->N2M;
  KParams& k = c->kp;
  k.n1 = prm->n1; k.n2 = prm->n2; k.N2M = c->N2M;
  k.eomtype = prm->eomtype; k.lim = prm->lim; k.fluxmethod = prm->fluxmethod; k.fluxb = prm->fluxb;
  for (int b = 0; b < 4; b++) k.bc[b] = prm->bc[b];
  k.fixupflux = prm->fixupflux; k.poledeath = prm->poledeath;
  k.gam = prm->gam; k.cour = prm->cour; k.defcon = prm->defcon; k.safe = prm->safe;
  k.gammamax = prm->gammamax; k.gammafail = prm->gammafail;
  k.rhominlimit = prm->rhominlimit; k.uuminlimit = prm->uuminlimit;
  k.dx1 = prm->dx1; k.dx2 = prm->dx2;
  size_t nz = (size_t)c->nz;
  hipError_t e = hipSuccess;
  for (int a = 0; a < HARM_NARRAYS && e == hipSuccess; a++) {
    e = hipMalloc(&c->arr[a], nz * 8 * sizeof(double));
    if (e == hipSuccess) e = hipMemset(c->arr[a], 0, nz * 8 * sizeof(double));
  }
  if (e == hipSuccess) e = hipMalloc(&c->emf, nz * sizeof(double));
  if (e == hipSuccess) e = hipMemset(c->emf, 0, nz * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&c->pflag, nz * sizeof(int));
  if (e == hipSuccess) e = hipMemset(c->pflag, 0, nz * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&c->gF, 2 * nz * GF * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&c->gC, nz * GCN * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&c->gConn, nz * 64 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&c->ds, sizeof(DevScalars));
  if (e == hipSuccess) e = hipStreamCreate(&c->own_stream);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e != hipSuccess) {
    seterr("harm_create: %s", hipGetErrorString(e));
    harm_destroy(c);
    return 1;
  }
  c->stream = c->own_stream;
  DevScalars h;
  memset(&h, 0, sizeof(h));
  double big = 1.e9;
  memcpy(&h.ndt_acc[0], &big, 8);
  memcpy(&h.ndt_acc[1], &big, 8);
  h.ndt_last[0] = h.ndt_last[1] = big;
  h.tf = 1e30;
  e = hipMemcpy(c->ds, &h, sizeof(h), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    seterr("harm_create: %s", hipGetErrorString(e));
    harm_destroy(c);
    return 1;
  }
  *out = c;
  return 0;
}

extern "C" int harm_destroy(harm_ctx* c) {
  if (!c) return 0;
  if (c->graph_ok) (void)hipGraphExecDestroy(c->graph);
  for (int a = 0; a < HARM_NARRAYS; a++) if (c->arr[a]) (void)hipFree(c->arr[a]);
  if (c->emf) (void)hipFree(c->emf);
  if (c->pflag) (void)hipFree(c->pflag);
  if (c->gF) (void)hipFree(c->gF);
  if (c->gC) (void)hipFree(c->gC);
  if (c->gConn) (void)hipFree(c->gConn);
  if (c->ds) (void)hipFree(c->ds);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  free(c);
  return 0;
}

extern "C" int harm_set_stream(harm_ctx* c, void* s) {
  if (!c) return seterr("null ctx");
  c->stream = s ? (hipStream_t)s : c->own_stream;
  if (c->graph_ok) {
    (void)hipGraphExecDestroy(c->graph);
    c->graph_ok = false;
  }
  return 0;
}

extern "C" int harm_set_geometry(harm_ctx* c, const double* gcov, const double* gcon, const double* gdet,
                                 const double* conn, const double* floors) {
  if (!c || !gcov || !gcon || !gdet || !conn || !floors) return seterr("null argument");
  size_t nz = (size_t)c->nz;
  std::vector<double> hF(2 * nz * GF), hC(nz * GCN, 0.0);
  for (size_t z = 0; z < nz; z++) {
    for (int pos = 0; pos < 2; pos++) {
      const double* gc = gcov + (z * 4 + pos) * 16;
      const double* gn = gcon + (z * 4 + pos) * 16;
      double* o = &hF[((size_t)pos * nz + z) * GF];
      o[0] = gc[0]; o[1] = gc[1]; o[2] = gc[2]; o[3] = gc[3];
      o[4] = gc[5]; o[5] = gc[6]; o[6] = gc[7];
      o[7] = gc[10]; o[8] = gc[11]; o[9] = gc[15];
      o[10] = gn[0]; o[11] = gn[1]; o[12] = gn[2]; o[13] = gn[3];
      o[14] = (pos == 0) ? gn[5] : gn[10];
      o[15] = gdet[z * 4 + pos];
    }
    double* o = &hC[z * GCN];
    memcpy(o, gcov + (z * 4 + 3) * 16, 16 * sizeof(double));
    memcpy(o + 16, gcon + (z * 4 + 3) * 16, 16 * sizeof(double));
    o[32] = gdet[z * 4 + 3];
    o[33] = floors[z * 2 + 0];
    o[34] = floors[z * 2 + 1];
    o[35] = gdet[z * 4 + 2];
  }
  HIPCHK(hipMemcpy(c->gF, hF.data(), hF.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->gC, hC.data(), hC.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->gConn, conn, nz * 64 * sizeof(double), hipMemcpyHostToDevice));
  c->geometry_set = true;
  return 0;
}

extern "C" int harm_set_array(harm_ctx* c, int sel, const double* host) {
  if (!c || !host) return seterr("null argument");
  if (sel < 0 || sel >= HARM_NARRAYS) return seterr("bad array selector %d", sel);
  HIPCHK(hipMemcpyAsync(c->arr[sel], host, (size_t)c->nz * 8 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int harm_get_array(harm_ctx* c, int sel, double* host) {
  if (!c || !host) return seterr("null argument");
  if (sel < 0 || sel >= HARM_NARRAYS) return seterr("bad array selector %d", sel);
  HIPCHK(hipMemcpyAsync(host, c->arr[sel], (size_t)c->nz * 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int harm_array_ptr(harm_ctx* c, int sel, double** dev) {
  if (!c || !dev) return seterr("null argument");
  if (sel < 0 || sel >= HARM_NARRAYS) return seterr("bad array selector %d", sel);
  *dev = c->arr[sel];
  return 0;
}

extern "C" int harm_get_pflag(harm_ctx* c, int* host) {
  if (!c || !host) return seterr("null argument");
  int* tmp;
  HIPCHK(hipMalloc(&tmp, (size_t)c->nz * 5 * sizeof(int)));
  hipLaunchKernelGGL(k_pflag_expand, dim3((c->nz + 255) / 256), dim3(256), 0, c->stream, c->nz, c->pflag, tmp);
  hipError_t e = hipMemcpyAsync(host, tmp, (size_t)c->nz * 5 * sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return seterr("harm_get_pflag: %s", hipGetErrorString(e));
  return 0;
}

extern "C" int harm_set_time(harm_ctx* c, double t, double dt, double tf) {
  if (!c) return seterr("null ctx");
  double v[3] = {t, dt, tf};
  HIPCHK(hipMemcpyAsync(&c->ds->t, v, 3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int harm_get_time(harm_ctx* c, double* t, double* dt) {
  if (!c) return seterr("null ctx");
  double v[3];
  HIPCHK(hipMemcpyAsync(v, &c->ds->t, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (t) *t = v[0];
  if (dt) *dt = v[1];
  return 0;
}

// ---- launch helpers (optionally timed with events on the launch stream) ----

static inline dim3 grid2(int nj, int ni, dim3 b) { return dim3((nj + b.x - 1) / b.x, (ni + b.y - 1) / b.y); }

static int prof_begin(harm_ctx* c) {
  if (c->prof) HIPCHK(hipEventRecord(c->ev0, c->stream));
  return 0;
}
static int prof_end(harm_ctx* c, int kid) {
  if (c->prof) {
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->ktime[kid] += ms;
    c->kcount[kid] += 1;
  }
  return 0;
}

static int launch_flux(harm_ctx* c, int pr_sel, int dir) {
  const dim3 b(64, 4);
  if (prof_begin(c)) return 1;
  if (dir == 1)
    hipLaunchKernelGGL(k_flux<1>, grid2(c->kp.n2 + 2, c->kp.n1 + 1, b), b, 0, c->stream, c->kp, c->arr[pr_sel],
                       c->arr[HARM_F1], c->gF, c->ds);
  else
    hipLaunchKernelGGL(k_flux<2>, grid2(c->kp.n2 + 1, c->kp.n1 + 2, b), b, 0, c->stream, c->kp, c->arr[pr_sel],
                       c->arr[HARM_F2], c->gF + (size_t)c->nz * GF, c->ds);
  HIPCHK(hipGetLastError());
  return prof_end(c, dir == 1 ? HARM_K_FLUX1 : HARM_K_FLUX2);
}

static int launch_emf(harm_ctx* c) {
  const dim3 b(64, 4);
  if (prof_begin(c)) return 1;
  hipLaunchKernelGGL(k_emf, grid2(c->kp.n2 + 1, c->kp.n1 + 1, b), b, 0, c->stream, c->kp, c->arr[HARM_F1],
                     c->arr[HARM_F2], c->gF, c->gC, c->emf, c->ds);
  HIPCHK(hipGetLastError());
  return prof_end(c, HARM_K_EMF);
}

static int launch_update(harm_ctx* c, int pi, int pb, int pf, const RKCoef& C) {
  const dim3 b(64, 4);
  if (prof_begin(c)) return 1;
  hipLaunchKernelGGL(k_update, grid2(c->kp.n2, c->kp.n1, b), b, 0, c->stream, c->kp, C, c->arr[pi], c->arr[pb],
                     c->arr[pf], c->arr[HARM_ULAST], c->arr[HARM_UNEW], c->arr[HARM_F1], c->arr[HARM_F2], c->emf,
                     c->gC, c->gConn, c->pflag, c->ds);
  HIPCHK(hipGetLastError());
  return prof_end(c, HARM_K_UPDATE);
}

static int launch_bound(harm_ctx* c, int sel) {
  if (prof_begin(c)) return 1;
  hipLaunchKernelGGL(k_bound_x1, dim3((c->kp.n2 + 63) / 64, 2 * NG), dim3(64), 0, c->stream, c->kp, c->arr[sel],
                     c->gC, c->ds);
  hipLaunchKernelGGL(k_bound_x2, dim3((c->kp.n1 + 2 * NG + 63) / 64, 8), dim3(64), 0, c->stream, c->kp,
                     c->arr[sel]);
  HIPCHK(hipGetLastError());
  return prof_end(c, HARM_K_BOUND);
}

static void rk_setup(int timeorder, int* nto, double CUf[4][3], double Cunew[4][3], int pii[4], int pbb[4],
                     int pff[4]) {
  // step_ch.c:365-460 setup_rktimestep
  memset(CUf, 0, sizeof(double) * 12);
  memset(Cunew, 0, sizeof(double) * 12);
  if (timeorder == 4) {
    *nto = 4;
    double cu[4][3] = {{1.0, 0.0, 0.5}, {1.0, 0.0, 0.5}, {1.0, 0.0, 1.0}, {1.0, 0.0, 1.0}};
    double cn[4][3] = {{1.0, 1.0 / 6.0, 0.0}, {0.0, 1.0 / 3.0, 0.0}, {0.0, 1.0 / 3.0, 0.0}, {0.0, 1.0 / 6.0, 0.0}};
    memcpy(CUf, cu, sizeof(cu));
    memcpy(Cunew, cn, sizeof(cn));
    int a[4] = {HARM_P, HARM_P, HARM_P, HARM_P}, b[4] = {HARM_P, HARM_PK0, HARM_PK1, HARM_PK0},
        f[4] = {HARM_PK0, HARM_PK1, HARM_PK0, HARM_P};
    memcpy(pii, a, sizeof(a)); memcpy(pbb, b, sizeof(b)); memcpy(pff, f, sizeof(f));
  } else if (timeorder == 3) {
    *nto = 3;
    double cu[3][3] = {{1.0, 0.0, 1.0}, {3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0}, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
    double cn[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 1.0}};
    memcpy(CUf, cu, sizeof(cu));
    memcpy(Cunew, cn, sizeof(cn));
    int a[3] = {HARM_P, HARM_P, HARM_P}, b[3] = {HARM_P, HARM_PK0, HARM_PK1}, f[3] = {HARM_PK0, HARM_PK1, HARM_P};
    memcpy(pii, a, sizeof(a)); memcpy(pbb, b, sizeof(b)); memcpy(pff, f, sizeof(f));
  } else if (timeorder == 2) {
    *nto = 2;
    double cu[2][3] = {{1.0, 0.0, 0.5}, {1.0, 0.0, 1.0}};
    double cn[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 1.0}};
    memcpy(CUf, cu, sizeof(cu));
    memcpy(Cunew, cn, sizeof(cn));
    int a[2] = {HARM_P, HARM_P}, b[2] = {HARM_P, HARM_PK0}, f[2] = {HARM_PK0, HARM_P};
    memcpy(pii, a, sizeof(a)); memcpy(pbb, b, sizeof(b)); memcpy(pff, f, sizeof(f));
  } else {
    *nto = 1;
    CUf[0][0] = 1.0; CUf[0][1] = 0.0; CUf[0][2] = 1.0;
    Cunew[0][0] = 0.0; Cunew[0][1] = 0.0; Cunew[0][2] = 1.0;
    pii[0] = pbb[0] = pff[0] = HARM_P;
  }
}

// one substage without host synchronisation (fluxcalc x2, emf, update)
static int substage(harm_ctx* c, int pi, int pb, int pf, const double* CUf, const double* Cunew, int to, int nto) {
  RKCoef C;
  C.cuf0 = CUf[0]; C.cuf1 = CUf[1]; C.cuf2 = CUf[2];
  C.cn0 = Cunew[0]; C.cn1 = Cunew[1]; C.cn2 = Cunew[2];
  C.timeorder = to;
  C.last = (to == nto - 1);
  if (launch_flux(c, pb, 1)) return 1;
  if (launch_flux(c, pb, 2)) return 1;
  if (launch_emf(c)) return 1;
  if (launch_update(c, pi, pb, pf, C)) return 1;
  return 0;
}

static int enqueue_step(harm_ctx* c) {
  double CUf[4][3], Cunew[4][3];
  int pii[4], pbb[4], pff[4], nto;
  rk_setup(c->prm.timeorder, &nto, CUf, Cunew, pii, pbb, pff);
  for (int to = 0; to < nto; to++) {
    if (substage(c, pii[to], pbb[to], pff[to], CUf[to], Cunew[to], to, nto)) return 1;
    if (to == nto - 1) {
      hipLaunchKernelGGL(k_dt, dim3(1), dim3(1), 0, c->stream, c->kp, c->ds);
      HIPCHK(hipGetLastError());
    }
    // bound_prim; post_fixup (fixup_utoprim) is not applied on the GPU path yet,
    // so the second bound_prim of step_ch.c:124 would be a repeat of this one
    if (launch_bound(c, pff[to])) return 1;
  }
  return 0;
}

static int need_geom(harm_ctx* c) {
  if (!c) return seterr("null ctx");
  if (!c->geometry_set) return seterr("geometry not set (harm_set_geometry)");
  return 0;
}

extern "C" int harm_fluxcalc(harm_ctx* c, int pr_sel, int dir, double* ndt) {
  if (need_geom(c)) return 1;
  if (pr_sel < 0 || pr_sel > HARM_PK1) return seterr("pr_sel must be HARM_P/PK0/PK1");
  if (dir != 1 && dir != 2) return seterr("dir must be 1 or 2");
  if (launch_flux(c, pr_sel, dir)) return 1;
  unsigned long long bits;
  HIPCHK(hipMemcpyAsync(&bits, &c->ds->ndt_acc[dir - 1], 8, hipMemcpyDeviceToHost, c->stream));
  unsigned long long big;
  double b = 1.e9;
  memcpy(&big, &b, 8);
  HIPCHK(hipMemcpyAsync(&c->ds->ndt_acc[dir - 1], &big, 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (ndt) memcpy(ndt, &bits, 8);
  return 0;
}

extern "C" int harm_fix_flux(harm_ctx* c) {
  (void)c;
  // the fix_flux() edge rules are applied inside the flux kernels when
  // params.fixupflux is set; nothing is left to do here
  return need_geom(c);
}

extern "C" int harm_flux_ct(harm_ctx* c) {
  if (need_geom(c)) return 1;
  if (c->prm.fluxb == HARM_FLUXCTHLL) return 0;
  // keep the Courant accumulators: k_emf snapshots/resets them, so restore after
  DevScalars save;
  HIPCHK(hipMemcpyAsync(&save, c->ds, sizeof(save), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (launch_emf(c)) return 1;
  HIPCHK(hipMemcpyAsync(c->ds, &save, sizeof(save), hipMemcpyHostToDevice, c->stream));
  const dim3 b(64, 4);
  hipLaunchKernelGGL(k_ctapply, grid2(c->kp.n2 + 1, c->kp.n1 + 1, b), b, 0, c->stream, c->kp, c->arr[HARM_F1],
                     c->arr[HARM_F2], c->emf);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int harm_advance(harm_ctx* c, int pi, int pb, const double CUf[3], int pf, const double Cunew[3],
                            int timeorder, int numtimeorders, double* ndt) {
  if (need_geom(c)) return 1;
  if (pi < 0 || pi > HARM_PK1 || pb < 0 || pb > HARM_PK1 || pf < 0 || pf > HARM_PK1)
    return seterr("pi/pb/pf must be HARM_P/PK0/PK1");
  if (!CUf || !Cunew) return seterr("null coefficients");
  if (substage(c, pi, pb, pf, CUf, Cunew, timeorder, numtimeorders)) return 1;
  double last[2];
  HIPCHK(hipMemcpyAsync(last, c->ds->ndt_last, sizeof(last), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (ndt) *ndt = c->prm.defcon * 1. / (1. / last[0] + 1. / last[1]);
  return 0;
}

extern "C" int harm_bound_prim(harm_ctx* c, int sel) {
  if (need_geom(c)) return 1;
  if (sel < 0 || sel > HARM_PK1) return seterr("sel must be HARM_P/PK0/PK1");
  if (launch_bound(c, sel)) return 1;
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int harm_step(harm_ctx* c) {
  if (need_geom(c)) return 1;
  if (enqueue_step(c)) return 1;
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

static int build_graph(harm_ctx* c) {
  if (c->graph_ok) return 0;
  hipGraph_t g;
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_step(c);
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) return 1;
  if (e != hipSuccess) return seterr("graph capture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return seterr("graph instantiate: %s", hipGetErrorString(e));
  c->graph_ok = true;
  return 0;
}

extern "C" int harm_step_n(harm_ctx* c, int nsteps) {
  if (need_geom(c)) return 1;
  if (nsteps < 0) return seterr("nsteps < 0");
  if (c->prof) {  // timed per kernel: eager launches
    for (int s = 0; s < nsteps; s++)
      if (enqueue_step(c)) return 1;
  } else {
    if (build_graph(c)) return 1;
    for (int s = 0; s < nsteps; s++) HIPCHK(hipGraphLaunch(c->graph, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int harm_utoprimgen(harm_ctx* c, int n, const int* zone, const double* U, double* pr, int* pflag) {
  if (need_geom(c)) return 1;
  if (n <= 0) return 0;
  if (!zone || !U || !pr || !pflag) return seterr("null argument");
  for (int t = 0; t < n; t++) {
    int i = zone[2 * t], j = zone[2 * t + 1];
    if (i < -NG || i >= c->prm.n1 + NG || j < -NG || j >= c->prm.n2 + NG)
      return seterr("zone %d (%d,%d) out of range", t, i, j);
  }
  int* dz;
  double *dU, *dp;
  int* dfl;
  HIPCHK(hipMalloc(&dz, (size_t)n * 2 * sizeof(int)));
  HIPCHK(hipMalloc(&dU, (size_t)n * 8 * sizeof(double)));
  HIPCHK(hipMalloc(&dp, (size_t)n * 8 * sizeof(double)));
  HIPCHK(hipMalloc(&dfl, (size_t)n * sizeof(int)));
  hipError_t e = hipMemcpy(dz, zone, (size_t)n * 2 * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dU, U, (size_t)n * 8 * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dp, pr, (size_t)n * 8 * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_utoprim_batch, dim3((n + 127) / 128), dim3(128), 0, c->stream, c->kp, n, dz, dU, dp, dfl,
                       c->gC);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(pr, dp, (size_t)n * 8 * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(pflag, dfl, (size_t)n * sizeof(int), hipMemcpyDeviceToHost);
  (void)hipFree(dz);
  (void)hipFree(dU);
  (void)hipFree(dp);
  (void)hipFree(dfl);
  if (e != hipSuccess) return seterr("harm_utoprimgen: %s", hipGetErrorString(e));
  return 0;
}

extern "C" int harm_get_failures(harm_ctx* c, int* nfailzones, int* failed) {
  if (!c) return seterr("null ctx");
  DevScalars h;
  HIPCHK(hipMemcpyAsync(&h, c->ds, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (nfailzones) *nfailzones = h.nfail;
  if (failed) *failed = h.failed;
  return 0;
}

extern "C" int harm_get_nstroke(harm_ctx* c, long long* nstroke) {
  if (!c || !nstroke) return seterr("null argument");
  unsigned long long v;
  HIPCHK(hipMemcpyAsync(&v, &c->ds->nstroke, 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *nstroke = (long long)v;
  return 0;
}

extern "C" int harm_prof_enable(harm_ctx* c, int on) {
  if (!c) return seterr("null ctx");
  c->prof = (on != 0);
  return 0;
}

extern "C" int harm_prof_get(harm_ctx* c, int kernel, double* total_ms, long long* count) {
  if (!c) return seterr("null ctx");
  if (kernel < 0 || kernel >= HARM_K_COUNT) return seterr("bad kernel id %d", kernel);
  if (total_ms) *total_ms = c->ktime[kernel];
  if (count) *count = c->kcount[kernel];
  return 0;
}

extern "C" int harm_prof_reset(harm_ctx* c) {
  if (!c) return seterr("null ctx");
  for (int k = 0; k < HARM_K_COUNT; k++) {
    c->ktime[k] = 0;
    c->kcount[k] = 0;
  }
  return 0;
}
