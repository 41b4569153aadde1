// harm_kernels.hip -- HIP kernels for gfx950 and the C-ABI (include/harm_amd.h)
// of HARM's conservative timestep.
//
// One RK substage is three stages on one stream:
//   flux       fluxcalc() for both directions in one launch: reconstruction,
//              limit_gamma, states, fluxes, vchar, HLL/LAXF, fix_flux edge
//              rules, wavefront-reduced Courant minimum      (step_ch.c:741)
//              k_flux_single (thread per face) or k_flux (lane pair per face)
//   k_update   advance() per zone: U(pi), source(pb), flux divergence with
//              the flux-CT fluxes formed from the corner EMFs (step_ch.c:1159)
//              inline, RK combination, Newton con-to-prim, fixup1zone, and
//              the list of zones whose inversion failed      (advance.c:46)
//   boundary   bound_prim(), post_fixup() on the failed list, bound_prim(),
//              and after the last substage the t,dt update (step_ch.c:85-183):
//              k_boundary (one workgroup) or k_bx1/k_bx2/k_fixlist/k_fixapply
// All variants of a stage give bit-identical results; the first step times
// them and keeps the fastest.  A step never synchronises with the host, and
// a whole step is captured once into a HIP graph and replayed.
//
// Device layout (HBM):
//   primitive-shaped arrays: the reference's AoS p[N1M][N2M][8] (64 B/zone);
//   gF[2][nz][16]  FACE1/FACE2: gcov upper 10, gcon^{00..03}, gcon^{dd}, gdet
//   gC[nz][30]     CENT: gcov upper 10 (gcov is symmetric to the bit), gcon 16
//                  (whose row 0 can hold -0 where its column holds +0), gdet,
//                  floor rho, floor u, gdet(CORN)
//   gConn[40][nz]  CENT connection Gamma^k_{lj}, l <= j (conn_sym), component-major
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <string.h>
#include <stdio.h>
#include <stdarg.h>
#include <stdlib.h>
#include <vector>
#include <mutex>
#include <algorithm>
#include <stddef.h>
#include "harm_device.h"
#include "harm_tangent.h"
#include "harm_adjoint.h"
#include "harm_eno.h"
#include "../../include/harm_amd.h"

using namespace harm;

namespace {

constexpr int NG = 3;
constexpr int GF = 16;   // doubles per face geometry record
constexpr int GCN = 30;  // doubles per CENT geometry record
constexpr int GC_GDET = 26, GC_GDETC = 29;  // gdet at CENT and at CORN in the record

// Accumulators that every wavefront of a launch updates are striped over
// NSTRIPE slots, one 128-byte line apart: thousands of same-address atomics
// from one launch serialise (measured on MI355X: 2,586 waves adding to one
// address cost 20-27 us more than to 64 addresses), a slot per wavefront
// group does not.  Readers combine the slots.
constexpr int NSTRIPE = 64;
constexpr int SSTR = 16;  // slot stride in 8-byte words

struct DevScalars {
  unsigned long long ndt_acc[2][NSTRIPE * SSTR];  // running min of dtij bits, per direction, striped
  double ndt_last[2];             // snapshot taken after both flux kernels
  double t, dt, tf;
  int failed;
  int nfail;
  int nlist;  // entries of the failed-zone list of the current substage
  int cs_changed;  // fixup_checksolution: the second pass differs from the first
  double dt_step;  // dt of the step just completed (diag_flux's Dt)
  unsigned long long nstroke[NSTRIPE * SSTR];  // Newton iterations, striped; word 8 of a
                                               // slot: the sum of each wavefront's largest count
  // HARM_BOUND_INUPDATE with N1 >= 5: the x1 ghost columns of the pole rows'
  // outflow edge zones as the second bound_prim() computes them (from the row
  // after POLEDEATH), written by k_update, copied by k_fixall where the edge
  // zone was not repaired; [e][g]: e&1 the x1 side, e&2 the x2 side, column g+1
  double pole_x1[4][NG][8];
  int pole_fail[4][NG];
  int disk_first;  // bound_disk()'s firsttime: its next call skips the symmetrisation
};

// the stripe of the calling wavefront (any spread of a launch's wavefronts)
__device__ __forceinline__ int stripe_of(const dim3 bi) {
  const unsigned w = (threadIdx.z * blockDim.y + threadIdx.y) * blockDim.x + threadIdx.x;
  return (int)((bi.x * 7u + bi.y * 13u + bi.z * 29u + (w >> 6) * 5u) & (NSTRIPE - 1));
}
__device__ __forceinline__ void ndt_push(DevScalars* __restrict__ ds, int dir, double wm, const dim3 bi) {
  atomicMin(&ds->ndt_acc[dir - 1][stripe_of(bi) * SSTR], (unsigned long long)__double_as_longlong(wm));
}
// the Courant limits of the substage's flux launch into ndt_last, every
// slot reset for the next one (between the flux launches): one wavefront,
// lane k taking slot k of both directions (one round trip, the minimum by
// butterflies; a minimum does not depend on its order)
static_assert(NSTRIPE == 64, "one slot per lane of a wavefront");
__device__ __forceinline__ void ndt_take_wave(DevScalars* __restrict__ ds, int lane) {
  const unsigned long long big = (unsigned long long)__double_as_longlong(1.e9);
  unsigned long long m[2] = {ds->ndt_acc[0][lane * SSTR], ds->ndt_acc[1][lane * SSTR]};
#pragma unroll
  for (int d = 0; d < 2; d++) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(m[d], off);
      m[d] = (o < m[d]) ? o : m[d];
    }
    ds->ndt_acc[d][lane * SSTR] = big;
  }
  if (lane == 0) {
    ds->ndt_last[0] = __longlong_as_double((long long)m[0]);
    ds->ndt_last[1] = __longlong_as_double((long long)m[1]);
  }
}
// the first wavefront of workgroup (0, 0) of a zone launch takes them
__device__ __forceinline__ void ndt_take_first_wave(DevScalars* __restrict__ ds, const dim3 bi) {
  const int t = threadIdx.x + threadIdx.y * blockDim.x;
  if (bi.x == 0 && bi.y == 0 && t < 64) ndt_take_wave(ds, t);
}

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
template <class T>
__device__ __forceinline__ void geo_from_cent(const T* c, GeoT<T>& g) {
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
__device__ __forceinline__ int wave_max_i32(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
  return v;
}

// ---------------------------------------------------------------------------
// fluxcalc() for both directions in one launch (blockIdx.z = dir-1).  Each
// face is handled by a lane pair: the even lane forms the left state (right
// edge of the zone below), the odd lane the right state (left edge of the
// zone above); the two half-computations (limit_gamma, get_state, primtoflux
// for F and U, vchar) run in parallel and are exchanged with __shfl_xor, then
// each lane writes half of the flux components.  Face ranges follow
// FZLOOP(-jdel,-idel) (mympi.h:287): dir 1 i in [0,N1], j in [-1,N2];
// dir 2 i in [-1,N1], j in [0,N2].
// Besides F the kernel writes q = F1[B2]/gdet(FACE1) (dir 1) and
// q = F2[B1]/gdet(FACE2) (dir 2): the four quotients of the corner EMF
// (step_ch.c:1228-1230), so the EMF is a 5-flop sum wherever it is needed.
// ---------------------------------------------------------------------------
template <int DIR>
__device__ __forceinline__ void flux_face(const KParams& P, const double* __restrict__ pr, double* __restrict__ F,
                                          double* __restrict__ q, const double* __restrict__ gF, int i, int j,
                                          int side, bool valid, double& dtij, int& fail) {
  constexpr int idel = (DIR == 1), jdel = (DIR == 2);
  constexpr int kb = B1 + DIR - 1;
  const int zstride = (DIR == 1) ? P.N2M : 1;
  const int z0 = valid ? zidx(P, i, j) : 0;
  double p[NPR];
  GeoU g;
  if (valid) {
    // this lane's 5-zone stencil along DIR starts at zone (face - 3 + side)
    double s[5][NPR];
#pragma unroll
    for (int m = 0; m < 5; m++) load8(pr, z0 + (m - 3 + side) * zstride, s[m]);
    if (P.lim == 4) {  // PARA: left state = pright[i-1], right state = pleft[i] (step_ch.c:896-903)
#pragma unroll
      for (int k = 0; k < NPR; k++) {
        double l, r;
        para4(s[0][k], s[1][k], s[2][k], s[3][k], s[4][k], l, r);
        p[k] = side ? l : r;
      }
    } else {  // p[i-1] + dq[i-1]/2, p[i] - dq[i]/2 (step_ch.c:894-895)
#pragma unroll
      for (int k = 0; k < NPR; k++) {
        double dq = (P.lim == 90) ? csslope(k, s[0][k], s[1][k], s[2][k], s[3][k], s[4][k])
                                  : slope3(P.lim, s[1][k], s[2][k], s[3][k]);
        p[k] = side ? (s[2][k] - 0.5 * dq) : (s[2][k] + 0.5 * dq);
      }
    }
    load_geo_face(gF, z0, g);
  } else {
#pragma unroll
    for (int k = 0; k < NPR; k++) p[k] = 0.0;
    g = GeoU{-1, 0, 0, 0, 1, 0, 0, 1, 0, 1, -1, 0, 0, 0, 1, 1};
  }
  // BDIRCONT (step_ch.c:944-947): both lanes form 0.5*(p_l+p_r) in that order
  {
    double other = __shfl_xor(p[kb], 1);
    double pl = side ? other : p[kb], prr = side ? p[kb] : other;
    p[kb] = 0.5 * (pl + prr);
  }
  if (limit_gamma(P.gammamax, p, g) >= 1) fail = 1;
  State st;
  get_state(p, g, st);
  double Fs[NPR], Us[NPR], cmax_s, cmin_s;
  primtoflux(P, p, st, DIR, g.gdet, Fs);
  primtoflux(P, p, st, 0, g.gdet, Us);
  if (vchar(P, p, st, DIR, g, cmax_s, cmin_s)) fail = 1;
  if (!valid) fail = 0;
  // exchange the two half-states; each lane keeps the half of the components
  // it will write (k = 4*side + m), selected with compile-time indices
  double Fmine[4], Umine[4], Fthat[4], Uthat[4];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    // the partner needs the other half: send components 4*(1-side)+m
    double fs = side ? Fs[m] : Fs[4 + m], us = side ? Us[m] : Us[4 + m];
    Fthat[m] = __shfl_xor(fs, 1);   // partner's value of component 4*side+m
    Uthat[m] = __shfl_xor(us, 1);
    Fmine[m] = side ? Fs[4 + m] : Fs[m];
    Umine[m] = side ? Us[4 + m] : Us[m];
  }
  double cmax_o = __shfl_xor(cmax_s, 1), cmin_o = __shfl_xor(cmin_s, 1);
  const double cmax_l = side ? cmax_o : cmax_s, cmin_l = side ? cmin_o : cmin_s;
  const double cmax_r = side ? cmax_s : cmax_o, cmin_r = side ? cmin_s : cmin_o;
  // cminmax_calc (step_ch.c:1095-1148)
  double lmin = dmin(cmin_l, cmin_r), lmax = dmax(cmax_l, cmax_r);
  double cmin = fabs(dmax(0., -lmin)), cmax = fabs(dmax(0., lmax));
  double ctop = dmax(cmax, cmin);
  double hllfrac = (P.fluxmethod == 1) ? 1.0 : 0.0;
  double laxffrac = (P.fluxmethod == 0) ? 1.0 : 0.0;
  double Fh[4];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    double Fl = side ? Fthat[m] : Fmine[m], Fr = side ? Fmine[m] : Fthat[m];
    double Ul = side ? Uthat[m] : Umine[m], Ur = side ? Umine[m] : Uthat[m];
    if (cmax + cmin != 0.0)
      Fh[m] = hllfrac * ((cmax * Fl + cmin * Fr - cmax * cmin * (Ur - Ul)) / (cmax + cmin)) +
              laxffrac * (0.5 * (Fl + Fr - ctop * (Ur - Ul)));
    else
      Fh[m] = (0.5 * (Fl + Fr - ctop * (Ur - Ul)));
  }
  double cm = dmax(cmax, cmin);
  dtij = P.cour * ((DIR == 1) ? P.dx1 : P.dx2) / cm;
  if (!(dtij < 1.e9) || !valid) dtij = 1.e9;  // the reference's running min ignores NaN and >= 1e9
  if (!valid) return;
  // fix_flux() edge rules (fixup.c:1357-1410), applied by the face that owns them
  bool skip_b2 = false;
  if (P.fixupflux) {
    if (DIR == 1) {
      if (side == 0) {
        if (i == 0 && P.bc[HARM_X1DN] == HARM_OUTFLOW && j >= -1 && j < P.n2 && Fh[RHO] > 0) Fh[RHO] = 0;
        if (i == P.n1 && P.bc[HARM_X1UP] == HARM_OUTFLOW && j >= -1 && j < P.n2 && Fh[RHO] < 0) Fh[RHO] = 0;
      } else if (i >= -1 && i < P.n1) {
        // F1[i][-1][B2] = -F1[i][0][B2], F1[i][N2][B2] = -F1[i][N2-1][B2], with their EMF quotients
        if (j == 0 && P.bc[HARM_X2DN] == HARM_POLARAXIS) {
          int zg = zidx(P, i, -1);
          F[(size_t)zg * 8 + B2] = -Fh[B2 - 4];
          q[zg] = (-Fh[B2 - 4]) / gF[(size_t)zg * GF + 15];
        }
        if (j == P.n2 - 1 && P.bc[HARM_X2UP] == HARM_POLARAXIS) {
          int zg = zidx(P, i, P.n2);
          F[(size_t)zg * 8 + B2] = -Fh[B2 - 4];
          q[zg] = (-Fh[B2 - 4]) / gF[(size_t)zg * GF + 15];
        }
        skip_b2 = (j == -1 && P.bc[HARM_X2DN] == HARM_POLARAXIS) || (j == P.n2 && P.bc[HARM_X2UP] == HARM_POLARAXIS);
      }
    } else {
      bool polar_dn = (j == 0 && P.bc[HARM_X2DN] == HARM_POLARAXIS && i >= -1 && i < P.n1);
      bool polar_up = (j == P.n2 && P.bc[HARM_X2UP] == HARM_POLARAXIS && i >= -1 && i < P.n1);
      bool equator = (i >= 0 && i < P.n1 && j == (P.n2 - 1) / 2);
      if (polar_dn || polar_up || equator) {
#pragma unroll
        for (int m = 0; m < 4; m++)
          if (4 * side + m != U2) Fh[m] = 0.;
      }
    }
  }
  double* Fz = F + (size_t)z0 * 8 + 4 * side;
  if (skip_b2) {
    Fz[0] = Fh[0];
    Fz[1] = Fh[1];
    Fz[3] = Fh[3];
  } else {
    reinterpret_cast<double2*>(Fz)[0] = make_double2(Fh[0], Fh[1]);
    reinterpret_cast<double2*>(Fz)[1] = make_double2(Fh[2], Fh[3]);
    if (side == 1) q[z0] = (DIR == 1 ? Fh[B2 - 4] : Fh[B1 - 4]) / g.gdet;
  }
}

__device__ __forceinline__ void flux_body(KParams P, const double* __restrict__ pr, double* __restrict__ F1,
                                              double* __restrict__ F2, double* __restrict__ q1,
                                              double* __restrict__ q2, const double* __restrict__ gF,
                                              DevScalars* __restrict__ ds, int dir0, const dim3 bi, const dim3 gd) {
  const int dir = dir0 + bi.z;
  const int side = threadIdx.x & 1;
  const int jf = bi.x * (blockDim.x >> 1) + (threadIdx.x >> 1);
  const int ifc = bi.y * blockDim.y + threadIdx.y;
  double dtij = 1.e9;
  int fail = 0;
  const size_t nz = (size_t)(P.n1 + 2 * NG) * P.N2M;
  if (dir == 1) {
    const int i = ifc, j = jf - 1;
    flux_face<1>(P, pr, F1, q1, gF, i, j, side, (i <= P.n1) && (j <= P.n2), dtij, fail);
  } else {
    const int i = ifc - 1, j = jf;
    flux_face<2>(P, pr, F2, q2, gF + nz * GF, i, j, side, (i <= P.n1) && (j <= P.n2), dtij, fail);
  }
  double wm = wave_min(dtij);
  if ((threadIdx.x & 63) == 0 && wm < 1.e9) ndt_push(ds, dir, wm, bi);
  if (fail) atomicOr(&ds->failed, 1);
}

__global__ __launch_bounds__(256) void k_flux(KParams P, const double* __restrict__ pr, double* __restrict__ F1,
                                              double* __restrict__ F2, double* __restrict__ q1,
                                              double* __restrict__ q2, const double* __restrict__ gF,
                                              DevScalars* __restrict__ ds, int dir0) {
  flux_body(P, pr, F1, F2, q1, q2, gF, ds, dir0, dim3(blockIdx.x, blockIdx.y, blockIdx.z),
      dim3(gridDim.x, gridDim.y, gridDim.z));
}

template <int DIR>
__device__ __forceinline__ void face_finish(const KParams& P, double* p_l, double* p_r, double* __restrict__ F,
                                            double* __restrict__ q, const double* __restrict__ gF, int i, int j,
                                            double& dtij, int& fail);

// the left and right states of face (i,j) of direction DIR (step_ch.c:812-903)
template <int DIR>
__device__ __forceinline__ void face_recon(const KParams& P, const double* __restrict__ pr, int i, int j,
                                           double* p_l, double* p_r) {
  const int zstride = (DIR == 1) ? P.N2M : 1;
  const int z0 = zidx(P, i, j);
  if (P.lim == 4) {
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
  } else {
    double s[6][NPR];
#pragma unroll
    for (int m = 0; m < 6; m++) load8(pr, z0 + (m - 3) * zstride, s[m]);
#pragma unroll
    for (int k = 0; k < NPR; k++) {
      double dql, dqr;
      if (P.lim == 90) {
        dql = csslope(k, s[0][k], s[1][k], s[2][k], s[3][k], s[4][k]);
        dqr = csslope(k, s[1][k], s[2][k], s[3][k], s[4][k], s[5][k]);
      } else {
        dql = slope3(P.lim, s[1][k], s[2][k], s[3][k]);
        dqr = slope3(P.lim, s[2][k], s[3][k], s[4][k]);
      }
      p_l[k] = s[2][k] + 0.5 * dql;
      p_r[k] = s[3][k] - 0.5 * dqr;
    }
  }
}

// fluxcalc() with one thread per face (the whole face on one lane); same
// arithmetic as flux_face<DIR>, different work split.
template <int DIR>
__device__ __forceinline__ void flux_face_single(const KParams& P, const double* __restrict__ pr,
                                                 double* __restrict__ F, double* __restrict__ q,
                                                 const double* __restrict__ gF, int i, int j, double& dtij,
                                                 int& fail) {
  double p_l[NPR], p_r[NPR];
  face_recon<DIR>(P, pr, i, j, p_l, p_r);
  face_finish<DIR>(P, p_l, p_r, F, q, gF, i, j, dtij, fail);
}

// everything of a face after the reconstruction: face geometry, fluxes and
// wave speeds, HLL/LAXF, the fix_flux edge rules, F and the EMF quotient
template <int DIR>
__device__ __forceinline__ void face_finish(const KParams& P, double* p_l, double* p_r, double* __restrict__ F,
                                            double* __restrict__ q, const double* __restrict__ gF, int i, int j,
                                            double& dtij, int& fail) {
  const int z0 = zidx(P, i, j);
  GeoU g;
  load_geo_face(gF, z0, g);
  double Fo[NPR];
  face_flux<DIR>(P, p_l, p_r, g, Fo, dtij, fail);
  bool skip_b2 = false;
  if (P.fixupflux) {
    if (DIR == 1) {
      if (i == 0 && P.bc[HARM_X1DN] == HARM_OUTFLOW && j >= -1 && j < P.n2 && Fo[RHO] > 0) Fo[RHO] = 0;
      if (i == P.n1 && P.bc[HARM_X1UP] == HARM_OUTFLOW && j >= -1 && j < P.n2 && Fo[RHO] < 0) Fo[RHO] = 0;
      if (i >= -1 && i < P.n1) {
        if (j == 0 && P.bc[HARM_X2DN] == HARM_POLARAXIS) {
          int zg = zidx(P, i, -1);
          F[(size_t)zg * 8 + B2] = -Fo[B2];
          q[zg] = (-Fo[B2]) / gF[(size_t)zg * GF + 15];
        }
        if (j == P.n2 - 1 && P.bc[HARM_X2UP] == HARM_POLARAXIS) {
          int zg = zidx(P, i, P.n2);
          F[(size_t)zg * 8 + B2] = -Fo[B2];
          q[zg] = (-Fo[B2]) / gF[(size_t)zg * GF + 15];
        }
        skip_b2 = (j == -1 && P.bc[HARM_X2DN] == HARM_POLARAXIS) || (j == P.n2 && P.bc[HARM_X2UP] == HARM_POLARAXIS);
      }
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
  if (skip_b2) {
    double* Fz = F + (size_t)z0 * 8;
#pragma unroll
    for (int k = 0; k < NPR; k++)
      if (k != B2) Fz[k] = Fo[k];
  } else {
    store8(F, z0, Fo);
    q[z0] = (DIR == 1 ? Fo[B2] : Fo[B1]) / g.gdet;
  }
}

__device__ __forceinline__ void flux_single_body(KParams P, const double* __restrict__ pr,
                                                     double* __restrict__ F1, double* __restrict__ F2,
                                                     double* __restrict__ q1, double* __restrict__ q2,
                                                     const double* __restrict__ gF, DevScalars* __restrict__ ds,
                                                     int dir0, const dim3 bi, const dim3 gd) {
  const int dir = dir0 + bi.z;
  const int jf = bi.x * blockDim.x + threadIdx.x;
  const int ifc = bi.y * blockDim.y + threadIdx.y;
  double dtij = 1.e9;
  int fail = 0;
  const size_t nz = (size_t)(P.n1 + 2 * NG) * P.N2M;
  if (dir == 1) {
    const int i = ifc, j = jf - 1;
    if (i <= P.n1 && j <= P.n2) flux_face_single<1>(P, pr, F1, q1, gF, i, j, dtij, fail);
  } else {
    const int i = ifc - 1, j = jf;
    if (i <= P.n1 && j <= P.n2) flux_face_single<2>(P, pr, F2, q2, gF + nz * GF, i, j, dtij, fail);
  }
  double wm = wave_min(dtij);
  if ((threadIdx.x & 63) == 0 && wm < 1.e9) ndt_push(ds, dir, wm, bi);
  if (fail) atomicOr(&ds->failed, 1);
}

// The scheme parameters of the common configurations as compile-time
// constants (a negative template value keeps the runtime parameter): the
// inlined physics then has no branches on the limiter, the equations of
// motion or the flux method, and the scheduler sees longer straight runs.
template <int LIMC, int EOMC, int FMC>
__device__ __forceinline__ KParams specialize(KParams P) {
  if (LIMC >= 0) P.lim = LIMC;
  if (EOMC >= 0) P.eomtype = EOMC;
  if (FMC >= 0) P.fluxmethod = FMC;
  return P;
}

template <int LIMC, int EOMC, int FMC>
__global__ __launch_bounds__(256) void k_flux_single(KParams P, const double* __restrict__ pr,
                                                     double* __restrict__ F1, double* __restrict__ F2,
                                                     double* __restrict__ q1, double* __restrict__ q2,
                                                     const double* __restrict__ gF, DevScalars* __restrict__ ds,
                                                     int dir0) {
  flux_single_body(specialize<LIMC, EOMC, FMC>(P), pr, F1, F2, q1, q2, gF, ds, dir0,
                   dim3(blockIdx.x, blockIdx.y, blockIdx.z), dim3(gridDim.x, gridDim.y, gridDim.z));
}

using FluxKernel = void (*)(KParams, const double*, double*, double*, double*, double*, const double*, DevScalars*,
                            int);
// the instantiation for a context's scheme: the torus (PARA, GRMHD, HLL), the
// force-free problems (MC, FFDE, LAXF) and Orszag-Tang (MC, GRMHD, HLL), or
// the general kernel
FluxKernel flux_single_kernel(const KParams& P) {
  if (P.lim == HARM_PARA && P.eomtype == HARM_EOMGRMHD && P.fluxmethod == HARM_HLLFLUX)
    return k_flux_single<HARM_PARA, HARM_EOMGRMHD, HARM_HLLFLUX>;
  if (P.lim == HARM_MC && P.eomtype == HARM_EOMFFDE && P.fluxmethod == HARM_LAXFFLUX)
    return k_flux_single<HARM_MC, HARM_EOMFFDE, HARM_LAXFFLUX>;
  if (P.lim == HARM_MC && P.eomtype == HARM_EOMGRMHD && P.fluxmethod == HARM_HLLFLUX)
    return k_flux_single<HARM_MC, HARM_EOMGRMHD, HARM_HLLFLUX>;
  return k_flux_single<-1, -1, -1>;
}

// Edge states of one zone along DIR from its 5-zone stencil s[0..4]
// (s[2] the zone): l = value at its lower face, r = at its upper face, with
// the expressions of flux_face_single (PARA: para4 edges; slopes: y -+ dq/2)
__device__ __forceinline__ void zone_edges(const KParams& P, const double (&s)[5][NPR], double* l, double* r) {
#pragma unroll
  for (int k = 0; k < NPR; k++) {
    if (P.lim == 4) {
      para4(s[0][k], s[1][k], s[2][k], s[3][k], s[4][k], l[k], r[k]);
    } else {
      const double dq = (P.lim == 90) ? csslope(k, s[0][k], s[1][k], s[2][k], s[3][k], s[4][k])
                                      : slope3(P.lim, s[1][k], s[2][k], s[3][k]);
      r[k] = s[2][k] + 0.5 * dq;
      l[k] = s[2][k] - 0.5 * dq;
    }
  }
}

// fluxcalc() with one thread per zone and direction: each zone's edge states
// are reconstructed once and shared with the face below it -- along x2
// between the lanes of a wavefront (63 faces per 64 lanes), along x1 through
// LDS between the rows of a workgroup (3 face rows per 4).  Workgroup 64x4.
constexpr int FSH_ROWS = 4;
__host__ __device__ __forceinline__ void flux_shared_grid(int n1, int n2, int& gx, int& gy) {
  const int gx2 = (n2 + 1 + 62) / 63, gy2 = (n1 + 2 + FSH_ROWS - 1) / FSH_ROWS;           // x2 faces
  const int gx1 = (n2 + 2 + 63) / 64, gy1 = (n1 + 1 + FSH_ROWS - 2) / (FSH_ROWS - 1);     // x1 faces
  gx = gx2 > gx1 ? gx2 : gx1;
  gy = gy2 > gy1 ? gy2 : gy1;
}

__global__ __launch_bounds__(64 * FSH_ROWS) void k_flux_shared(KParams P, const double* __restrict__ pr,
                                                              double* __restrict__ F1, double* __restrict__ F2,
                                                              double* __restrict__ q1, double* __restrict__ q2,
                                                              const double* __restrict__ gF,
                                                              DevScalars* __restrict__ ds, int dir0) {
  __shared__ double rsh[FSH_ROWS][64][NPR];
  const int dir = dir0 + blockIdx.z;
  const int tx = threadIdx.x, ty = threadIdx.y;
  double dtij = 1.e9;
  int fail = 0;
  double s[5][NPR], l[NPR], r[NPR];
  if (dir == 2) {
    // lanes along x2: zone jz, row i; face (i, jz) takes r of lane tx-1
    const int jz = 63 * (int)blockIdx.x - 1 + tx, i = FSH_ROWS * (int)blockIdx.y + ty - 1;
    const bool zone_ok = (jz >= -1 && jz <= P.n2 && i >= -1 && i <= P.n1);
    if (zone_ok) {
#pragma unroll
      for (int m = 0; m < 5; m++) load8(pr, zidx(P, i, jz - 2 + m), s[m]);
      zone_edges(P, s, l, r);
    } else {
#pragma unroll
      for (int k = 0; k < NPR; k++) l[k] = r[k] = 0.;
    }
    double p_l[NPR];
#pragma unroll
    for (int k = 0; k < NPR; k++) p_l[k] = __shfl_up(r[k], 1);
    if (tx >= 1 && zone_ok && jz >= 0) {
      const size_t nz = (size_t)(P.n1 + 2 * NG) * P.N2M;
      face_finish<2>(P, p_l, l, F2, q2, gF + nz * GF, i, jz, dtij, fail);
    }
  } else {
    // lanes along x2, rows along x1: zone iz of row ty; face (iz, j) takes r
    // of row ty-1 through LDS
    const int j = 64 * (int)blockIdx.x + tx - 1, iz = (FSH_ROWS - 1) * (int)blockIdx.y + ty - 1;
    const bool zone_ok = (j >= -1 && j <= P.n2 && iz >= -1 && iz <= P.n1);
    if (zone_ok) {
#pragma unroll
      for (int m = 0; m < 5; m++) load8(pr, zidx(P, iz - 2 + m, j), s[m]);
      zone_edges(P, s, l, r);
    } else {
#pragma unroll
      for (int k = 0; k < NPR; k++) l[k] = r[k] = 0.;
    }
#pragma unroll
    for (int k = 0; k < NPR; k++) rsh[ty][tx][k] = r[k];
    __syncthreads();
    if (ty >= 1 && zone_ok && iz >= 0) {
      double p_l[NPR];
#pragma unroll
      for (int k = 0; k < NPR; k++) p_l[k] = rsh[ty - 1][tx][k];
      face_finish<1>(P, p_l, l, F1, q1, gF, iz, j, dtij, fail);
    }
  }
  double wm = wave_min(dtij);
  if ((tx & 63) == 0 && wm < 1.e9) ndt_push(ds, dir, wm, blockIdx);
  if (fail) atomicOr(&ds->failed, 1);
}

// corner EMF of flux_ct() (step_ch.c:1223-1234, CORNGDETVERSION=1) from the
// per-face quotients written by k_flux
__device__ __forceinline__ double emf_at(const KParams& P, const double* __restrict__ q1,
                                         const double* __restrict__ q2, const double* __restrict__ gC, int i, int j) {
  const int z = zidx(P, i, j);
  return 0.25 * (gC[(size_t)z * GCN + GC_GDETC]) *
         (q1[z] + q1[zidx(P, i, j - 1)] - q2[z] - q2[zidx(P, i - 1, j)]);
}

// flux_ct() materialised into F1/F2 (step_ch.c:1223-1234, 1405-1412), for the
// standalone harm_flux_ct() entry point: EMFs first, then the B fluxes
__global__ void k_emf(KParams P, const double* __restrict__ q1, const double* __restrict__ q2,
                      const double* __restrict__ gC, double* __restrict__ emf) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > P.n1 || j > P.n2) return;
  emf[zidx(P, i, j)] = emf_at(P, q1, q2, gC, i, j);
}

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

// fix_flux() (fixup.c:1357-1410) on the fluxes a context holds, for the
// standalone harm_fix_flux() entry point (the step applies the same rules
// inside the flux kernels).  One thread per face position (i in [-1,N1],
// j in [-1,N2]) handles its F1 and its F2 face; no rule reads a value another
// rule writes, so the reference's sequential order needs no ordering here.
// The EMF quotients of every face whose B flux a rule changes follow F.
__global__ void k_fixflux(KParams P, double* __restrict__ F1, double* __restrict__ F2, double* __restrict__ q1,
                          double* __restrict__ q2, const double* __restrict__ gF) {
  const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x) - 1;
  const int i = (int)(blockIdx.y * blockDim.y + threadIdx.y) - 1;
  if (i > P.n1 || j > P.n2) return;
  const size_t nz = (size_t)(P.n1 + 2 * NG) * P.N2M;
  const int z = zidx(P, i, j);
  // F1[i][-1][B2] = -F1[i][0][B2], F1[i][N2][B2] = -F1[i][N2-1][B2], i in [-1,N1)
  if (i < P.n1 && ((j == -1 && P.bc[HARM_X2DN] == HARM_POLARAXIS) || (j == P.n2 && P.bc[HARM_X2UP] == HARM_POLARAXIS))) {
    const int zs = zidx(P, i, j == -1 ? 0 : P.n2 - 1);
    const double v = -F1[(size_t)zs * 8 + B2];
    F1[(size_t)z * 8 + B2] = v;
    q1[z] = v / gF[(size_t)z * GF + 15];
  }
  // F2 at the polar faces (i in [-1,N1)) and the equator (i in [0,N1)):
  // every component but U2 zeroed
  const bool polar = i < P.n1 && ((j == 0 && P.bc[HARM_X2DN] == HARM_POLARAXIS) ||
                                  (j == P.n2 && P.bc[HARM_X2UP] == HARM_POLARAXIS));
  const bool equator = i >= 0 && i < P.n1 && j == (P.n2 - 1) / 2;
  if (polar || equator) {
    double* f = F2 + (size_t)z * 8;
#pragma unroll
    for (int k = 0; k < NPR; k++)
      if (k != U2) f[k] = 0.;
    q2[z] = f[B1] / gF[(nz + z) * GF + 15];
  }
  // no mass flux into the grid through an outflow x1 face, j in [-1,N2)
  if (j < P.n2) {
    double* f = F1 + (size_t)z * 8;
    if (i == 0 && P.bc[HARM_X1DN] == HARM_OUTFLOW && f[RHO] > 0) f[RHO] = 0;
    if (i == P.n1 && P.bc[HARM_X1UP] == HARM_OUTFLOW && f[RHO] < 0) f[RHO] = 0;
  }
}

// EMF quotients F1[B2]/gdet(FACE1), F2[B1]/gdet(FACE2) of the fluxes a
// context holds (step_ch.c:1228-1230), for the standalone harm_flux_ct():
// x1 faces i in [0,N1], j in [-1,N2]; x2 faces i in [-1,N1], j in [0,N2]
__global__ void k_qfromF(KParams P, const double* __restrict__ F1, const double* __restrict__ F2,
                         double* __restrict__ q1, double* __restrict__ q2, const double* __restrict__ gF) {
  const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x) - 1;
  const int i = (int)(blockIdx.y * blockDim.y + threadIdx.y) - 1;
  if (i > P.n1 || j > P.n2) return;
  const size_t nz = (size_t)(P.n1 + 2 * NG) * P.N2M;
  const int z = zidx(P, i, j);
  if (i >= 0) q1[z] = F1[(size_t)z * 8 + B2] / gF[(size_t)z * GF + 15];
  if (j >= 0) q2[z] = F2[(size_t)z * 8 + B1] / gF[(nz + z) * GF + 15];
}

struct RKCoef {
  double cuf0, cuf1, cuf2, cn0, cn1, cn2;
  int timeorder, last;
};

// lower() (phys.c:633-652, upper triangle of gcov) on a full CENT record
template <class T>
__device__ __forceinline__ void lower_gc(const T* gc, const T* a, T* b) {
  b[0] = gc[0] * a[0] + gc[1] * a[1] + gc[2] * a[2] + gc[3] * a[3];
  b[1] = gc[1] * a[0] + gc[5] * a[1] + gc[6] * a[2] + gc[7] * a[3];
  b[2] = gc[2] * a[0] + gc[6] * a[1] + gc[10] * a[2] + gc[11] * a[3];
  b[3] = gc[3] * a[0] + gc[7] * a[1] + gc[11] * a[2] + gc[15] * a[3];
}

// Force-free inversion of one zone, phys.ffde.c:141-372 Utoprim_ffde
// (CHECKBSQMINUSESQ=1, AVOIDUTINF=2, AVOIDCS=0, REMOVET=0; ACCOUNTDISS feeds
// diagnostics only).  Ug is geometry-free: T^t_mu in Ug[UU..U3], B^i in
// Ug[B1..B3].  The drift velocity's u^i goes into pr[U1..U3] (vcon2pr,
// transforms.c:381-386); pr[RHO], pr[UU] are not touched.  Returns 1 on the
// space-like failure (phys.ffde.c:309-312), which is fatal in the reference.
template <class T>
__device__ int utoprim_ffde(const KParams& P, const T* Ug, const T* gc, T* pr) {
  const T* gcon = gc + 16;
  const T gdet = gc[32];
  const T alphasq = 1. / (-gcon[0]);
  T Bcon[4] = {T(0.0), Ug[B1], Ug[B2], Ug[B3]}, Bcov[4];
  lower_gc(gc, Bcon, Bcov);
  T Bsq = 0.0;
#pragma unroll
  for (int j = 1; j < 4; j++) Bsq += Bcov[j] * Bcon[j];
  const T T1 = Ug[UU + 1], T2 = Ug[UU + 2], T3 = Ug[UU + 3];
  const T o = 1.0 / (Bsq * gdet);
  T Econ[4], Ecov[4];
  Econ[0] = 0;
  Econ[1] = (Bcov[2] * T3 - Bcov[3] * T2) * (o);
  Econ[2] = (Bcov[3] * T1 - Bcov[1] * T3) * (o);
  Econ[3] = (Bcov[1] * T2 - Bcov[2] * T1) * (o);
  T realBsq = Bsq;
  lower_gc(gc, Econ, Ecov);
  T Esq = 0.0;
#pragma unroll
  for (int j = 1; j < 4; j++) Esq += Econ[j] * Ecov[j];
  const T BsqmEsq = Bsq - Esq;
  const T gammasq = Bsq / BsqmEsq;
  if ((gammasq > P.gammamax * P.gammamax) || (gammasq < 1.0)) {
    const double Gsq = 1.0 - 1.0 / (P.gammamax * P.gammamax);
    realBsq = (sqrt(Esq * Bsq / Gsq));
  }
  const T ob = 1.0 / (realBsq * gdet);
  T v1 = alphasq * (-gcon[1] + (Ecov[2] * Bcov[3] - Ecov[3] * Bcov[2]) * ob);
  T v2 = alphasq * (-gcon[2] + (Ecov[3] * Bcov[1] - Ecov[1] * Bcov[3]) * ob);
  T v3 = alphasq * (-gcon[3] + (Ecov[1] * Bcov[2] - Ecov[2] * Bcov[1]) * ob);
  // ucon_calc_3vel (phys.c:885-955)
  T discr = +gc[0] + gc[5] * v1 * v1 + gc[10] * v2 * v2 + gc[15] * v3 * v3 +
                 2. * (gc[1] * v1 + gc[2] * v2 + gc[3] * v3 + gc[6] * v1 * v2 + gc[7] * v1 * v3 + gc[11] * v2 * v3);
  if (discr > 0.) return 1;
  const T ut = 1. / sqrt(-discr);
  pr[U1] = v1 * ut;
  pr[U2] = v2 * ut;
  pr[U3] = v3 * ut;
  pr[B1] = Ug[B1];
  pr[B2] = Ug[B2];
  pr[B3] = Ug[B3];
  return 0;
}

// Inversion of one zone: Utoprimgen (step_ch.c:468-552) -> for GRMHD
// Utoprim_2d_final (utoprim_2d_final.c:26-101) -> Utoprim_new_body (:136-396),
// for FFDE Utoprim_ffde.  U is the evolved conserved state, pr the guess in /
// solution out.  Returns the pflag value, or HARM_FFDE_SPACELIKE.
__device__ int utoprimgen(const KParams& P, const double* U, const double* gc, double* pr, int& niter) {
  const double* gcov = gc;
  const double* gcon = gc + 16;
  double gdet = gc[32];
  double Ug[NPR];
#pragma unroll
  for (int k = 0; k < NPR; k++) Ug[k] = U[k] / gdet;
  niter = 0;
  if (P.eomtype == HARM_EOMFFDE) return utoprim_ffde(P, Ug, gc, pr) ? HARM_FFDE_SPACELIKE : 0;
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
// advance() per zone: advance_standard (advance.c:114-181) and the two
// DOENO forms (advance.c:101-158 and 298-487), zones i<N1, j<N2.
// ---------------------------------------------------------------------------

// the full CENT record (gcov 16, gcon 16, gdet, floor rho, floor u,
// gdet(CORN)) from the stored one, whose element m is r(m)
template <class T, class R>
__device__ __forceinline__ void cent_expand(const R& r, T* gc) {
  constexpr int up[16] = {0, 1, 2, 3, 1, 4, 5, 6, 2, 5, 7, 8, 3, 6, 8, 9};  // [a][b] -> upper entry
#pragma unroll
  for (int m = 0; m < 16; m++) {
    gc[m] = r(up[m]);
    gc[16 + m] = r(10 + m);
  }
#pragma unroll
  for (int m = 0; m < 4; m++) gc[32 + m] = r(GC_GDET + m);
}

// full CENT geometry record of zone z
__device__ __forceinline__ void load_cent(const double* __restrict__ gC, int z, double* gc) {
  const double2* a = reinterpret_cast<const double2*>(gC + (size_t)z * GCN);
  double v[GCN];
#pragma unroll
  for (int m = 0; m < GCN / 2; m++) {
    const double2 t = a[m];
    v[2 * m] = t.x;
    v[2 * m + 1] = t.y;
  }
  cent_expand(
      [&](int m) { return v[m]; }, gc);
}

// flux2dUavg (advance.c:204-265): each direction's flux difference, with the
// flux-CT B fluxes formed from the corner EMFs (step_ch.c:1223-1234, 1405-1412)
__device__ __forceinline__ void flux_div(const KParams& P, const double* __restrict__ F1,
                                         const double* __restrict__ F2, const double* __restrict__ q1,
                                         const double* __restrict__ q2, const double* __restrict__ gC, int i, int j,
                                         double* d1, double* d2) {
  const int z = zidx(P, i, j);
  double f1a[NPR], f1b[NPR], f2a[NPR], f2b[NPR];
  load8(F1, z, f1a);
  load8(F1, zidx(P, i + 1, j), f1b);
  load8(F2, z, f2a);
  load8(F2, zidx(P, i, j + 1), f2b);
  if (P.fluxb == HARM_FLUXCTTOTH) {
    double e00 = emf_at(P, q1, q2, gC, i, j), e01 = emf_at(P, q1, q2, gC, i, j + 1),
           e10 = emf_at(P, q1, q2, gC, i + 1, j), e11 = emf_at(P, q1, q2, gC, i + 1, j + 1);
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
#pragma unroll
  for (int k = 0; k < NPR; k++) {
    d1[k] = (-(f1b[k] - f1a[k]) * idx1);
    d2[k] = (-(f2b[k] - f2a[k]) * idx2);
  }
}

// cvtij (reconstructeno.c:133-235) of zone (i,j) along DIR for every
// component of the AoS array u; EOMFFDE converts RHO..U3 and keeps B
template <int TYPE, int DIR>
__device__ __forceinline__ void eno_cvt_zone(const KParams& P, const double* __restrict__ u, int i, int j,
                                             double* out) {
  const int pos = (DIR == 1) ? i : j, n = (DIR == 1) ? P.n1 : P.n2;
  double v[NPR][9];
#pragma unroll
  for (int o = 0; o < 9; o++) {
    const int q = pos - 4 + o;
    double t[NPR];
    if (q >= 0 && q < n) {
      load8(u, (DIR == 1) ? zidx(P, q, j) : zidx(P, i, q), t);
    } else {
#pragma unroll
      for (int k = 0; k < NPR; k++) t[k] = 0.;
    }
#pragma unroll
    for (int k = 0; k < NPR; k++) v[k][o] = t[k];
  }
  const int nk = (P.eomtype == HARM_EOMFFDE) ? U3 + 1 : NPR;
#pragma unroll
  for (int k = 0; k < NPR; k++) out[k] = (k < nk) ? eno_cvt_line<TYPE>(v[k], pos, n) : v[k][4];
}

// ---- floor/failure accounting (diag_fixup, fixup.c:174-229) ----
// Every change a fixup makes to a zone or face state is an event: dVF (U(after)
// - U(before)) in UDIAG form, booked in the energy regions holding that zone
// or face.  The events of a substage have fixed slots in the order the
// reference makes them: x1 faces (FZLOOP, left state then right), x2 faces,
// zones (the force-free inversion's dissipation account, limit_gamma, then
// floors), failed zones (limit_gamma, then the repair).
// A flag (category + 1) marks a slot in use; the slots are compacted in order
// and summed after the substage.

// primtoU with returntype UDIAG (phys.c:16-56, UtoU UNOTHING -> UDIAG):
// rest mass kept in UU, times gdet
__device__ __forceinline__ void primtoU_diag(const KParams& P, const double* pr, const State& q, double gdet,
                                             double* U) {
  double fin[NPR];
  fin[RHO] = pr[RHO] * q.ucon[0];
  mhd_calc(P, pr, 0, q, &fin[UU]);
  fin[B1] = pr[B1];
  fin[B2] = pr[B2];
  fin[B3] = pr[B3];
#pragma unroll
  for (int k = 0; k < NPR; k++) U[k] = fin[k] * gdet;
}

// one event into slot `slot`: fixup.c:199-217
__device__ void acct_event(const KParams& P, const GeoU& g, const double* pr0, const double* pr, int slot, int cat) {
  State q;
  double Ui[NPR], Uf[NPR];
  get_state(pr0, g, q);
  primtoU_diag(P, pr0, q, g.gdet, Ui);
  get_state(pr, g, q);
  primtoU_diag(P, pr, q, g.gdet, Uf);
  double* o = P.acct_val + (size_t)slot * 8;
#pragma unroll
  for (int k = 0; k < NPR; k++) o[k] = P.dVF * (Uf[k] - Ui[k]);
  P.acct_flag[slot] = cat + 1;
}

// slot bases of the four event classes of a substage: x1 and x2 faces (two
// states each), zones (three: the force-free inversion's dissipation account,
// limit_gamma, the floors), failed zones (two: limit_gamma, the repair)
__host__ __device__ __forceinline__ int acct_base(int n1, int n2, int cls) {
  const int nA = (n1 + 1) * (n2 + 2) * 2, nB = (n1 + 2) * (n2 + 1) * 2, nC = n1 * n2 * 3, nD = n1 * n2 * 2;
  return cls == 0 ? 0 : cls == 1 ? nA : cls == 2 ? nA + nB : cls == 3 ? nA + nB + nC : nA + nB + nC + nD;
}

// Utoprimgen of U into pf[z] from the guess pin (the pf=pi assignment,
// advance.c:121), fixup1zone (fixup.c:274-377), the zone's pflag and the
// failed-zone list read by fixup_utoprim; with accounting on, the state
// before fixup1zone is kept for its events
__device__ __forceinline__ void invert_zone(const KParams& P, const double* U, const double* pin, const double* gc,
                                            const GeoU& g, int z, double* __restrict__ pf,
                                            int* __restrict__ pflag, int* __restrict__ faillist,
                                            DevScalars* __restrict__ ds, int& niter, int& failz, int& fail) {
  double pout[NPR];
#pragma unroll
  for (int k = 0; k < NPR; k++) pout[k] = pin[k];
  int pfl = utoprimgen(P, U, gc, pout, niter);
  if (pfl == HARM_FFDE_SPACELIKE) {  // fatal in the reference (advance() returns 1)
    fail = 1;
    pfl = 0;
  }
  pflag[z] = pfl;
  failz = (pfl != 0);
  if (failz) {  // zones for fixup_utoprim (k_boundary)
    int slot = atomicAdd(&ds->nlist, 1);
    if (slot < P.n1 * P.n2) faillist[slot] = z;
  }
  if (P.acct_pre) {
    store8(P.acct_pre, z, pout);
    if (P.eomtype == HARM_EOMFFDE) store8(P.acct_uin, z, U);
  }
  // fixup1zone: floors then limit_gamma
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

// Newton iterations, failed zones and the fatal flag of a wavefront
__device__ __forceinline__ void count_zone(DevScalars* __restrict__ ds, int niter, int failz, int fail, const dim3 bi) {
  unsigned long long ns = wave_sum_u64((unsigned long long)niter);
  unsigned long long nf = wave_sum_u64((unsigned long long)failz);
  const int wm = wave_max_i32(niter);  // the wavefront runs its slowest zone's iterations
  if ((threadIdx.x & 63) == 0) {
    if (ns) {
      atomicAdd(&ds->nstroke[stripe_of(bi) * SSTR], ns);
      atomicAdd(&ds->nstroke[stripe_of(bi) * SSTR + 8], (unsigned long long)wm);
    }
    if (nf) atomicAdd(&ds->nfail, (int)nf);
  }
  if (fail) atomicOr(&ds->failed, 1);
}

// The per-zone update.  ENO (= P.eno) selects the form:
//   0  advance_standard: Ui = U(pi), flux divergence, inversion
//   1  advance_eno_u, middle loop (advance.c:401-438): Ui = average of U(pi)
//      (c2a along x2 of eB, which holds c2a along x1); the averaged Uf (or
//      unew on the last substage) goes to eA, inverted later by k_eno_invert
//   2  DOENO=2 (advance.c:142-158): the flux divergence of each direction is
//      converted average -> centre (a2c of eA along x1, of eB along x2)
constexpr int NIMG = 1 + 2 * NG;  // image parts of an interior zone (bound_image_part)
__device__ void bound_image_part(const KParams& P, double* __restrict__ pv, const double* __restrict__ gC, int i,
                                 int j, const double* v, int part, int& fail);
__device__ __forceinline__ bool pole_row(const KParams& P, int j);
__device__ __forceinline__ bool near_edge(const KParams& P, int i, int j);
__device__ __forceinline__ int x1_image_column(const KParams& P, int i, int part);
__device__ __forceinline__ void x1_value(const KParams& P, const double* __restrict__ gC, int i, int j, double* v,
                                         int& fail);

// The edge block of a k_update launch (64x4 workgroups) whose other three
// rows write the x1 ghost columns of its outflow edge row: 0 the down side
// (row 0), 1 the up side (row N1-1), -1 none (both edges in one block: the
// edge zones write their columns themselves).
__device__ __forceinline__ int edge_block_side(const KParams& P, const dim3 bi) {
  const bool dn = P.bc[HARM_X1DN] == HARM_OUTFLOW && bi.y == 0;
  const bool up = P.bc[HARM_X1UP] == HARM_OUTFLOW && (int)bi.y == (P.n1 - 1) / 4;
  return (dn && up) ? -1 : dn ? 0 : up ? 1 : -1;
}
static_assert(NG == 3, "an edge block has NG rows besides the edge row");

// BND: the first bound_prim() of the substage done here (ENO 0 only): each
// zone near an edge writes its ghost images after its inversion; an outflow
// edge row hands its values to the other rows of its workgroup (LDS, one
// barrier), which write the NG x1 ghost columns, one column per row
template <int ENO, int BND = 0>
__device__ __forceinline__ void update_body(KParams P, RKCoef C, const double* __restrict__ pi,
                                                const double* __restrict__ pb, double* __restrict__ pf,
                                                double* __restrict__ ulast, double* __restrict__ unew,
                                                const double* __restrict__ F1, const double* __restrict__ F2,
                                                const double* __restrict__ q1, const double* __restrict__ q2,
                                                const double* __restrict__ gC, const double* __restrict__ gConn,
                                                int* __restrict__ pflag, int* __restrict__ faillist,
                                                DevScalars* __restrict__ ds, double* eA, const double* eB,
                                                const dim3 bi, const dim3 gd) {
  const int j = bi.x * blockDim.x + threadIdx.x;
  const int i = bi.y * blockDim.y + threadIdx.y;
  int niter = 0, failz = 0, fail = 0;
  __shared__ double edge_v[BND ? 64 * NPR : 1];
  const int eside = BND ? edge_block_side(P, bi) : -1;
  ndt_take_first_wave(ds, bi);  // this substage's Courant limits (advance.c:190 reads them)
  if (i < P.n1 && j < P.n2) {
    const int z = zidx(P, i, j);
    const double dt = ds->dt;
    double gc[36];
    load_cent(gC, z, gc);
    GeoU g;
    geo_from_cent(gc, g);
    double pin[NPR], pbz[NPR];
    load8(pi, z, pin);
    load8(pb, z, pbz);
    double Ui[NPR];
    if constexpr (ENO == 1) {
      eno_cvt_zone<ENO_C2A, 2>(P, eB, i, j, Ui);
    } else {  // Ui = U(pi)
      State q;
      get_state(pin, g, q);
      primtoflux(P, pin, q, 0, g.gdet, Ui);
    }
    // source(pb): T^j_k Gamma^k_{lj} contraction (phys.c:475-619)
    double dUg[NPR];
    source_conn_sym(P, pbz, g, gConn + z, dUg, (size_t)(P.n1 + 2 * NG) * P.N2M);
    double d1[NPR], d2[NPR];
    if constexpr (ENO == 2) {
      eno_cvt_zone<ENO_A2C, 1>(P, eA, i, j, d1);
      eno_cvt_zone<ENO_A2C, 2>(P, eB, i, j, d2);
    } else {
      flux_div(P, F1, F2, q1, q2, gC, i, j, d1, d2);
    }
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
      double dUr = d1[k] + d2[k];
      // dUtoU (advance.c:273-286)
      double dtdU = (dUr + dUg[k]);
      Uf[k] = C.cuf0 * Ui[k] + C.cuf1 * ul[k] + C.cuf2 * dt * dtdU;
      un[k] += C.cn0 * Ui[k] + C.cn1 * dt * dtdU + C.cn2 * Uf[k];
      ul[k] = Uf[k];
    }
    store8(ulast, z, ul);
    store8(unew, z, un);
    if constexpr (ENO == 1) {
      store8(eA, z, C.last ? un : Uf);
    } else {
      invert_zone(P, C.last ? un : Uf, pin, gc, g, z, pf, pflag, faillist, ds, niter, failz, fail);
      if constexpr (BND == 1) {  // (the edge block's other rows continue below)
        if (near_edge(P, i, j)) {
          double v[NPR];
          load8(pf, z, v);
          const bool handed = eside >= 0 && i == (eside ? P.n1 - 1 : 0);
          for (int part = 0; part < NIMG; part++)
            if (!(handed && part > 0 && (part > NG) == (eside == 1)))
              bound_image_part(P, pf, gC, i, j, v, part, fail);
          if (handed) {
#pragma unroll
            for (int k = 0; k < NPR; k++) edge_v[threadIdx.x * NPR + k] = v[k];
          }
          if (pole_row(P, j)) {  // POLEDEATH of the pole row, after its images
            pf[(size_t)z * 8 + U2] *= 0.;
            pf[(size_t)z * 8 + B2] *= 0.;
          }
        }
      }
    }
  }
  if constexpr (BND == 1) {
    if (eside >= 0) {  // block-uniform
      __syncthreads();  // the edge row's values in edge_v
      const int orow = eside ? (P.n1 - 1) - (int)(bi.y * blockDim.y) : 0;
      const int ty = threadIdx.y, src = eside ? P.n1 - 1 : 0;
      if (ty != orow && j < P.n2) {  // the first bound_prim()'s column, one per row
        const int part = (eside ? NG + 1 : 1) + (ty < orow ? ty : ty - 1);
        double v[NPR];
#pragma unroll
        for (int k = 0; k < NPR; k++) v[k] = edge_v[threadIdx.x * NPR + k];
        bound_image_part(P, pf, gC, src, j, v, part, fail);
      } else if (ty == orow && threadIdx.x >= 1 && threadIdx.x <= 2 * NG && P.n1 >= 5) {
        // the edge row's own lanes 1..2NG: the second bound_prim()'s value of
        // column g of pole row p (0: j = 0, 1: j = N2-1), from the zeroed row
        const int p = (threadIdx.x - 1) / NG, g = (threadIdx.x - 1) % NG;
        const int jp = p ? P.n2 - 1 : 0, lane = jp - (int)(bi.x * blockDim.x);
        if (lane >= 0 && lane < (int)blockDim.x && pole_row(P, jp)) {
          double v[NPR];
#pragma unroll
          for (int k = 0; k < NPR; k++) v[k] = edge_v[lane * NPR + k];
          const int part = (eside ? NG + 1 : 1) + g, e = eside | (p ? 2 : 0);
          int f2 = 0;
          v[U2] *= 0.;
          v[B2] *= 0.;
          x1_value(P, gC, x1_image_column(P, src, part), jp, v, f2);
          v[U2] *= 0.;
          v[B2] *= 0.;
#pragma unroll
          for (int k = 0; k < NPR; k++) ds->pole_x1[e][g][k] = v[k];
          ds->pole_fail[e][g] = f2;
        }
      }
    }
  }
  count_zone(ds, niter, failz, fail, bi);
}

// EOMC >= 0: the equations of motion as a compile-time constant (the
// inversion and the fluxes without the other branch)
template <int ENO, int EOMC = -1, int BND = 0>
__global__ __launch_bounds__(256) void k_update(KParams P, RKCoef C, const double* __restrict__ pi,
                                                const double* __restrict__ pb, double* __restrict__ pf,
                                                double* __restrict__ ulast, double* __restrict__ unew,
                                                const double* __restrict__ F1, const double* __restrict__ F2,
                                                const double* __restrict__ q1, const double* __restrict__ q2,
                                                const double* __restrict__ gC, const double* __restrict__ gConn,
                                                int* __restrict__ pflag, int* __restrict__ faillist,
                                                DevScalars* __restrict__ ds, double* eA, const double* eB) {
  update_body<ENO, BND>(specialize<-1, EOMC, -1>(P), C, pi, pb, pf, ulast, unew, F1, F2, q1, q2, gC, gConn, pflag,
                   faillist, ds, eA, eB, dim3(blockIdx.x, blockIdx.y, blockIdx.z),
                   dim3(gridDim.x, gridDim.y, gridDim.z));
}

// ---- DOENO kernels (interior zones, thread per zone) ----

// advance_eno_u first loop (advance.c:363-384): U(pi) at zone centres
__global__ __launch_bounds__(256) void k_eno_ucenter(KParams P, const double* __restrict__ pi,
                                                     const double* __restrict__ gC, double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= P.n1 || j >= P.n2) return;
  const int z = zidx(P, i, j);
  double gc[36];
  load_cent(gC, z, gc);
  GeoU g;
  geo_from_cent(gc, g);
  double pin[NPR], U[NPR];
  load8(pi, z, pin);
  State q;
  get_state(pin, g, q);
  primtoflux(P, pin, q, 0, g.gdet, U);
  store8(out, z, U);
}

// a2c()/c2a() along x1 of every interior zone (CZLOOP cvtij)
template <int TYPE>
__global__ __launch_bounds__(256) void k_eno_cvt1(KParams P, const double* __restrict__ in,
                                                  double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= P.n1 || j >= P.n2) return;
  double o[NPR];
  eno_cvt_zone<TYPE, 1>(P, in, i, j, o);
  store8(out, zidx(P, i, j), o);
}

// DOENO=2 (advance.c:101-108): each direction's averaged flux divergence
__global__ __launch_bounds__(256) void k_eno_fluxdiv(KParams P, const double* __restrict__ F1,
                                                     const double* __restrict__ F2, const double* __restrict__ q1,
                                                     const double* __restrict__ q2, const double* __restrict__ gC,
                                                     double* __restrict__ d1out, double* __restrict__ d2out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= P.n1 || j >= P.n2) return;
  double d1[NPR], d2[NPR];
  flux_div(P, F1, F2, q1, q2, gC, i, j, d1, d2);
  const int z = zidx(P, i, j);
  store8(d1out, z, d1);
  store8(d2out, z, d2);
}

// advance_eno_u last loop (advance.c:459-477): the averaged update converted
// to zone centres (a2c along x2 of in, which holds a2c along x1), inverted
// and fixed up
__global__ __launch_bounds__(256) void k_eno_invert(KParams P, const double* __restrict__ pi,
                                                    double* __restrict__ pf, const double* __restrict__ in,
                                                    const double* __restrict__ gC, int* __restrict__ pflag,
                                                    int* __restrict__ faillist, DevScalars* __restrict__ ds) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  int niter = 0, failz = 0, fail = 0;
  if (i < P.n1 && j < P.n2) {
    const int z = zidx(P, i, j);
    double gc[36];
    load_cent(gC, z, gc);
    GeoU g;
    geo_from_cent(gc, g);
    double pin[NPR], U[NPR];
    load8(pi, z, pin);
    eno_cvt_zone<ENO_A2C, 2>(P, in, i, j, U);
    invert_zone(P, U, pin, gc, g, z, pf, pflag, faillist, ds, niter, failz, fail);
  }
  count_zone(ds, niter, failz, fail, blockIdx);
}

// inflow_check_rel4vel for x1 ghosts at CENT (fixup.c:1733-1841)
template <class T>
__device__ int inflow_check_x1(const KParams& P, int i, T* pr, const T* gc, const GeoT<T>& g) {
  T alpha = 1. / sqrt(-g.n00);
  T gamma = gamma_calc(pr, g);
  T ucon1 = pr[U1] - gamma * (g.n01 * alpha * alpha) / alpha;
  bool dofix = ((i <= -1) && (P.bc[HARM_X1DN] == HARM_OUTFLOW) && (ucon1 > 0.)) ||
               ((i >= P.n1) && (P.bc[HARM_X1UP] == HARM_OUTFLOW) && (ucon1 < 0.));
  if (!dofix) return 0;
  pr[U1] /= gamma;
  pr[U2] /= gamma;
  pr[U3] /= gamma;
  T alpha2 = 1. / sqrt(-gc[16]);
  T betacon = gc[17] * alpha2 * alpha2;
  pr[U1] = betacon / alpha2;
  T vsq = 0.;
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

// x1 ghost cell (gi in 0..5 -> i = -3..-1, N1..N1+2; j in [0,N2)):
// bounds.fishmon.c:26-110,190-262 (EXTRAP=0 copy, inflow check, limit_gamma)
// or the periodic wrap
// the interior zone an x1 ghost (i, j) is filled from
__device__ __forceinline__ int x1_source(const KParams& P, int i) {
  const bool dn = i < 0;
  const int bc = dn ? P.bc[HARM_X1DN] : P.bc[HARM_X1UP];
  if (bc == HARM_PERIODIC) return dn ? i + P.n1 : i - P.n1;
  return dn ? 0 : P.n1 - 1;
}

// the value of x1 ghost (i, j) from the given source state v (in/out)
__device__ __forceinline__ void x1_value(const KParams& P, const double* __restrict__ gC, int i, int j, double* v,
                                         int& fail) {
  const int bc = (i < 0) ? P.bc[HARM_X1DN] : P.bc[HARM_X1UP];
  if (bc != HARM_OUTFLOW) return;  // periodic: the copy itself
  double gc[36];
  load_cent(gC, zidx(P, i, j), gc);
  GeoU g;
  geo_from_cent(gc, g);
  if (inflow_check_x1(P, i, v, gc, g) >= 1) fail = 1;
  if (limit_gamma(P.gammamax, v, g) >= 1) fail = 1;
}

__device__ void bound_x1_cell(const KParams& P, double* __restrict__ pr, const double* __restrict__ gC, int gi,
                              int j, int& fail) {
  const bool dn = gi < NG;
  const int i = dn ? (gi - NG) : (P.n1 + gi - NG);
  const int bc = dn ? P.bc[HARM_X1DN] : P.bc[HARM_X1UP];
  if (bc != HARM_PERIODIC && bc != HARM_OUTFLOW) return;
  double v[NPR];
  load8(pr, zidx(P, x1_source(P, i), j), v);
  x1_value(P, gC, i, j, v, fail);
  store8(pr, zidx(P, i, j), v);
}

// x2 ghost cell (i in [-3,N1+2], gj in 0..5 -> j = -3..-1, N2..N2+2):
// polar reflection with POLEDEATH on the ghost's U2,B2 (bounds.fishmon.c:263-355)
// or the periodic wrap
__device__ void bound_x2_cell(const KParams& P, double* __restrict__ pr, int i, int gj) {
  const bool dn = gj < NG;
  const int j = dn ? (gj - NG) : (P.n2 + gj - NG);
  const int bc = dn ? P.bc[HARM_X2DN] : P.bc[HARM_X2UP];
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

// bound_prim() from the writer's side: every ghost zone whose value the
// composed rules above (x1 first, a corner from its x1 value) take from
// interior zone (i, j), given that zone's value v before POLEDEATH zeroes it.
// The images come in parts, each a set of ghosts of its own: part 0 the x2
// images of the zone (polar reflections with POLEDEATH's zeroing, periodic
// wraps), part 1..NG the x1 ghost column -part (down side), part
// NG+1..2NG the column N1-1+part-NG (up side), each with its own x2 images
// (the corners).  An x1 part exists when zone i is that column's source
// (outflow: the edge zone, through x1_value; periodic: the wrapped zone).
// Each ghost has one source, so writers of different (zone, part) pairs
// never collide; only x1 OUTFLOW/PERIODIC and x2 POLARAXIS/PERIODIC are
// covered (bound_images_ok).  The zeroing of the pole row itself is the
// caller's, after every part has read v.
__host__ __device__ __forceinline__ bool bound_images_ok(const KParams& P) {
  if (P.extrap != 0 || P.boundsfile != HARM_BOUNDS_FISHMON) return false;
  for (int s = 0; s < 2; s++) {
    const int b1 = P.bc[s ? HARM_X1UP : HARM_X1DN], b2 = P.bc[s ? HARM_X2UP : HARM_X2DN];
    if (b1 != HARM_OUTFLOW && b1 != HARM_PERIODIC) return false;
    if (b2 != HARM_POLARAXIS && b2 != HARM_PERIODIC) return false;
  }
  return true;
}
__device__ __forceinline__ bool pole_row(const KParams& P, int j) {
  return P.poledeath && ((j == 0 && P.bc[HARM_X2DN] == HARM_POLARAXIS) ||
                         (j == P.n2 - 1 && P.bc[HARM_X2UP] == HARM_POLARAXIS));
}
__device__ __forceinline__ bool near_edge(const KParams& P, int i, int j) {
  return i < NG || i >= P.n1 - NG || j < NG || j >= P.n2 - NG;
}
// the x2 images of (ii, j) holding value u
__device__ __forceinline__ void bound_x2_images(const KParams& P, double* __restrict__ pv, int ii, int j,
                                                const double* u) {
  for (int s = 0; s < 2; s++) {
    const int bc = s ? P.bc[HARM_X2UP] : P.bc[HARM_X2DN];
    int jg;
    if (bc == HARM_POLARAXIS) {
      if (s ? (j < P.n2 - NG) : (j >= NG)) continue;
      jg = s ? 2 * P.n2 - 1 - j : -1 - j;
    } else {  // periodic
      if (s ? (j >= NG) : (j < P.n2 - NG)) continue;
      jg = s ? j + P.n2 : j - P.n2;
    }
    double t[NPR];
#pragma unroll
    for (int k = 0; k < NPR; k++) t[k] = u[k];
    if (bc == HARM_POLARAXIS && P.poledeath) {
      t[U2] *= 0.;
      t[B2] *= 0.;
    }
    store8(pv, zidx(P, ii, jg), t);
  }
}
// the x1 ghost column of image part 1..2NG if zone column i is its source, else -1000
__device__ __forceinline__ int x1_image_column(const KParams& P, int i, int part) {
  const bool up = part > NG;
  const int g = up ? part - NG : part;  // 1..NG
  const int ig = up ? P.n1 - 1 + g : -g;
  const int bc = up ? P.bc[HARM_X1UP] : P.bc[HARM_X1DN];
  if (bc == HARM_OUTFLOW) return (i == (up ? P.n1 - 1 : 0)) ? ig : -1000;
  return (i == (up ? ig - P.n1 : ig + P.n1)) ? ig : -1000;  // periodic
}
// one part of the images of interior zone (i, j)
__device__ void bound_image_part(const KParams& P, double* __restrict__ pv, const double* __restrict__ gC, int i,
                                 int j, const double* v, int part, int& fail) {
  if (part == 0) {
    bound_x2_images(P, pv, i, j, v);
    return;
  }
  const int ig = x1_image_column(P, i, part);
  if (ig == -1000) return;
  double w[NPR];
#pragma unroll
  for (int k = 0; k < NPR; k++) w[k] = v[k];
  x1_value(P, gC, ig, j, w, fail);
  if (pole_row(P, j)) {
    w[U2] *= 0.;
    w[B2] *= 0.;
  }
  store8(pv, zidx(P, ig, j), w);
  bound_x2_images(P, pv, ig, j, w);
}

// bound_prim() by one workgroup: x1 faces, then x2 faces over all i (corners
// from the x1 values), then POLEDEATH of the two interior pole rows, in the
// reference's order (barriers between the phases)
__device__ void bound_block(const KParams& P, double* __restrict__ pr, const double* __restrict__ gC, int& fail) {
  const int tid = threadIdx.x, nt = blockDim.x, ni = P.n1 + 2 * NG;
  for (int t = tid; t < 2 * NG * P.n2; t += nt) bound_x1_cell(P, pr, gC, t / P.n2, t % P.n2, fail);
  __syncthreads();
  for (int t = tid; t < 2 * NG * ni; t += nt) bound_x2_cell(P, pr, t % ni - NG, t / ni);
  __syncthreads();
  if (P.poledeath) {
    for (int t = tid; t < 2 * ni; t += nt) {
      const int i = t % ni - NG;
      const bool dn = t < ni;
      if ((dn ? P.bc[HARM_X2DN] : P.bc[HARM_X2UP]) != HARM_POLARAXIS) continue;
      double* a = pr + (size_t)zidx(P, i, dn ? 0 : P.n2 - 1) * 8;
      a[U2] *= 0.;
      a[B2] *= 0.;
    }
  }
  __syncthreads();
}

// pflag of a zone at distance <= 1 from the interior, as bound_pflag()
// (boundsint.c:4-60) leaves it: x2 mapping first (polar reflection / periodic
// wrap over all i), then x1 (outflow copy of the edge zone / periodic wrap).
// (A FIXED x1 side is left alone there: fixed_x1_ghost.)
__device__ __forceinline__ int pflag_zone(const KParams& P, int i, int j) {
  if (j < 0) j = (P.bc[HARM_X2DN] == HARM_PERIODIC) ? j + P.n2 : (0 + (0 - j - 1));
  else if (j >= P.n2) j = (P.bc[HARM_X2UP] == HARM_PERIODIC) ? j - P.n2 : ((P.n2 - 1) + ((P.n2 - 1) - j + 1));
  if (i < 0) i = (P.bc[HARM_X1DN] == HARM_PERIODIC) ? i + P.n1 : 0;
  else if (i >= P.n1) i = (P.bc[HARM_X1UP] == HARM_PERIODIC) ? i - P.n1 : P.n1 - 1;
  return zidx(P, i, j);
}
// A ghost column of a FIXED x1 side: bound_pflag does nothing there ("if
// fixed BC: do nothing", boundsint.c:34) and nothing else writes its flags,
// so they keep set_arrays()'s initial -100 (set_arrays.c:86), which the
// repair counts as a failed neighbour.
constexpr int PFLAG_INIT = -100;
__device__ __forceinline__ bool fixed_x1_ghost(const KParams& P, int i) {
  return (i < 0 && P.bc[HARM_X1DN] == HARM_FIXED) || (i >= P.n1 && P.bc[HARM_X1UP] == HARM_FIXED);
}
__device__ __forceinline__ int pflag_at(const KParams& P, const int* __restrict__ pflag, int i, int j) {
  return pflag[pflag_zone(P, i, j)];
}
__device__ __forceinline__ bool interior(const KParams& P, int i, int j) {
  return i >= 0 && i < P.n1 && j >= 0 && j < P.n2;
}

// fixup_utoprim() for one failed zone (fixup.c:527-920; UTOPRIMAVG,
// ORDERINDEPENDENT, HANDLEUNEG=0, TODONOGOOD=NOGOODAVERAGE): the average of its
// good neighbours from the bounded array (read-only during this phase), gamma
// capped at its own previous gamma or GAMMAMAX; result to pfix.
// cs (nullable): zones failed by fixup_checksolution in this pass.  Ghost
// neighbours carry the flags bound_pflag copied before that check, so a
// ghost whose source zone was failed by it still counts as good.
// PF: the neighbours' densities and velocities (components RHO..U3, the ones
// the repair averages) loaded with the flags (one memory round trip instead
// of two; the values are those the branch then reads)
template <bool PF = false>
__device__ void fixup_zone(const KParams& P, const double* __restrict__ pv, const int* __restrict__ pflag,
                           const int* __restrict__ cs, double* __restrict__ pfix, const double* __restrict__ gC,
                           int z, int& fail, double* keep = nullptr) {
  const int i = z / P.N2M - NG, j = z % P.N2M - NG;
  constexpr int NA = U3 + 1;
  double nb[PF ? 9 : 1][PF ? NA : 1];
  if constexpr (PF) {
#pragma unroll
    for (int di = -1; di <= 1; di++)
#pragma unroll
      for (int dj = -1; dj <= 1; dj++) {
        const double* a = pv + (size_t)zidx(P, i + di, j + dj) * 8;
        double* o = nb[(di + 1) * 3 + (dj + 1)];
#pragma unroll
        for (int k = 0; k < NA; k++) o[k] = a[k];
      }
  }
  double gc[36];
  load_cent(gC, z, gc);
  GeoU g;
  geo_from_cent(gc, g);
  double pz[NPR];
  load8(pv, z, pz);
  double gamma = gamma_calc(pz, g);
  if (gamma > P.gammamax) gamma = P.gammamax;
  auto good = [&](int di, int dj) {
    const int a = i + di, b = j + dj, zm = pflag_zone(P, a, b);
    int f = fixed_x1_ghost(P, a) ? PFLAG_INIT : pflag[zm];
    if (cs && !interior(P, a, b) && cs[zm]) f = 0;
    return f == 0;
  };
  auto val = [&](int di, int dj, int k) {
    if constexpr (PF) return nb[(di + 1) * 3 + (dj + 1)][k];
    else return pv[(size_t)zidx(P, i + di, j + dj) * 8 + k];
  };
  auto avg41 = [&](int k) { return 0.25 * (val(0, 1, k) + val(0, -1, k) + val(-1, 0, k) + val(1, 0, k)); };
  auto avg42 = [&](int k) { return 0.25 * (val(1, 1, k) + val(1, -1, k) + val(-1, 1, k) + val(-1, -1, k)); };
  int mode;
  if (good(0, 1) && good(0, -1) && good(1, 0) && good(-1, 0)) mode = 0;
  else if (good(1, 1) && good(1, -1) && good(-1, 1) && good(-1, -1)) mode = 1;
  else if (good(0, 1) && good(0, -1)) mode = 2;
  else if (good(1, 0) && good(-1, 0)) mode = 3;
  else if (good(1, 1) && good(-1, -1)) mode = 4;
  else if (good(1, -1) && good(-1, 1)) mode = 5;
  else if (good(1, 1)) mode = 6;
  else if (good(1, 0)) mode = 7;
  else if (good(1, -1)) mode = 8;
  else if (good(0, -1)) mode = 9;
  else if (good(-1, -1)) mode = 10;
  else if (good(-1, 0)) mode = 11;
  else if (good(-1, 1)) mode = 12;
  else if (good(0, 1)) mode = 13;
  else mode = 14;
  for (int k = 0; k <= U3; k++) {
    double v;
    switch (mode) {
      case 0: v = avg41(k); break;
      case 1: v = avg42(k); break;
      case 2: v = (0.5 * (val(0, 1, k) + val(0, -1, k))); break;
      case 3: v = (0.5 * (val(1, 0, k) + val(-1, 0, k))); break;
      case 4: v = (0.5 * (val(1, 1, k) + val(-1, -1, k))); break;
      case 5: v = (0.5 * (val(1, -1, k) + val(-1, 1, k))); break;
      case 6: v = val(1, 1, k); break;
      case 7: v = val(1, 0, k); break;
      case 8: v = val(1, -1, k); break;
      case 9: v = val(0, -1, k); break;
      case 10: v = val(-1, -1, k); break;
      case 11: v = val(-1, 0, k); break;
      case 12: v = val(-1, 1, k); break;
      case 13: v = val(0, 1, k); break;
      default: v = 0.5 * (avg41(k) + avg42(k)); break;
    }
    pz[k] = v;
  }
  const int slot = acct_base(P.n1, P.n2, 3) + (i * P.n2 + j) * 2;
  if (P.acct_flag) {  // the states of fixup.c:850's limit_gamma event: before it ...
    double* o = P.acct_val + (size_t)slot * 8;
#pragma unroll
    for (int k = 0; k < NPR; k++) o[k] = (k <= U3) ? pz[k] : pv[(size_t)z * 8 + k];
  }
  const int lg = limit_gamma(gamma, pz, g);
  if (lg >= 1) fail = 1;
  if (P.acct_flag) {  // ... after it, and the repair's (fixup.c:870-915): the state before and its failure type
    double* o = P.acct_val + (size_t)(slot + 1) * 8;
#pragma unroll
    for (int k = 0; k < NPR; k++) o[k] = (k <= U3) ? pz[k] : pv[(size_t)z * 8 + k];
#pragma unroll
    for (int k = 0; k < NPR; k++) P.acct_pre[(size_t)z * 8 + k] = pv[(size_t)z * 8 + k];
    const int f = pflag[z];
    const int cat = (f == 5) ? 4 : (f == 6) ? 5 : (f == 7) ? 6 : (f == 8) ? 7 : (f == 9) ? 8 : 0;
    P.acct_flag[slot] = (lg == -1) ? -1 : -2;  // states stored; k_acct_fixg makes the events
    P.acct_flag[slot + 1] = cat + 1;
  }
  double* o = pfix + (size_t)z * 8;  // the whole repaired state (the fields unchanged)
#pragma unroll
  for (int k = 0; k < NPR; k++) o[k] = pz[k];
  if (keep) {
#pragma unroll
    for (int k = 0; k < NPR; k++) keep[k] = pz[k];
  }
}

constexpr int KB_BOUND = 1;     // bound_prim()
constexpr int KB_FIXUP = 2;     // post_fixup(): bound_pflag + fixup_utoprim on the failed-zone list
constexpr int KB_BOUND2 = 8;    // bound_prim() after the fixup (step_ch.c:124)
constexpr int KB_DT = 4;        // t,dt update (step_ch.c:174-183)
constexpr int KB_INUPDATE = 16; // the first bound_prim() done by k_update (HARM_BOUND_INUPDATE)

// step_ch.c:174-183 with advance.c:190: the new dt from the last substage's
// Courant limits, clipped to SAFE*dt and to tf
__device__ inline void dt_update(const KParams& P, DevScalars* ds) {
  double ndt = P.defcon * 1. / (1. / ds->ndt_last[0] + 1. / ds->ndt_last[1]);
  double dt = ds->dt;
  ds->dt_step = dt;
  ds->t += dt;
  if (ndt > P.safe * dt) ndt = P.safe * dt;
  dt = ndt;
  if (ds->t + dt > ds->tf) dt = ds->tf - ds->t;
  ds->dt = dt;
}

// Everything between two advance() calls, as one workgroup with barriers
// between the phases: bound_prim, post_fixup, bound_prim, and after the last
// substage the time/dt update.  The failed-zone list is filled by k_update.
__device__ __forceinline__ void boundary_body(KParams P, double* __restrict__ pv, const double* __restrict__ gC,
                                                   int* __restrict__ pflag, const int* __restrict__ faillist,
                                                   double* __restrict__ pfix, DevScalars* __restrict__ ds, int mode, const dim3 bi, const dim3 gd) {
  int fail = 0;
  if (mode & KB_BOUND) bound_block(P, pv, gC, fail);
  if (mode & KB_FIXUP) {
    int n = ds->nlist;
    if (n > P.n1 * P.n2) n = P.n1 * P.n2;
    for (int t = threadIdx.x; t < n; t += blockDim.x) fixup_zone(P, pv, pflag, nullptr, pfix, gC, faillist[t], fail);
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
      const int z = faillist[t];
#pragma unroll
      for (int k = 0; k <= U3; k++) pv[(size_t)z * 8 + k] = pfix[(size_t)z * 8 + k];
      pflag[z] = -1;  // UTOPRIMFAILFIXED (global.h:524)
    }
    __syncthreads();
    if (threadIdx.x == 0) ds->nlist = 0;
  }
  if (mode & KB_BOUND2) bound_block(P, pv, gC, fail);
  if ((mode & KB_DT) && threadIdx.x == 0) dt_update(P, ds);
  if (fail) atomicOr(&ds->failed, 1);
}

__global__ __launch_bounds__(1024) void k_boundary(KParams P, double* __restrict__ pv, const double* __restrict__ gC,
                                                   int* __restrict__ pflag, const int* __restrict__ faillist,
                                                   double* __restrict__ pfix, DevScalars* __restrict__ ds, int mode) {
  boundary_body(P, pv, gC, pflag, faillist, pfix, ds, mode, dim3(blockIdx.x, blockIdx.y, blockIdx.z),
      dim3(gridDim.x, gridDim.y, gridDim.z));
}

// ---- the same boundary work as separate wide launches (bit-identical) ----

// bound_prim() x1 faces: thread (j, gi)
__device__ __forceinline__ void bx1_body(KParams P, double* __restrict__ pv, const double* __restrict__ gC,
                      DevScalars* __restrict__ ds, int reset_list, const dim3 bi, const dim3 gd) {
  const int j = bi.x * blockDim.x + threadIdx.x;
  if (reset_list && bi.x == 0 && bi.y == 0 && threadIdx.x == 0) ds->nlist = 0;
  if (j >= P.n2) return;
  int fail = 0;
  bound_x1_cell(P, pv, gC, bi.y, j, fail);
  if (fail) atomicOr(&ds->failed, 1);
}

__global__ void k_bx1(KParams P, double* __restrict__ pv, const double* __restrict__ gC,
                      DevScalars* __restrict__ ds, int reset_list) {
  bx1_body(P, pv, gC, ds, reset_list, dim3(blockIdx.x, blockIdx.y, blockIdx.z),
      dim3(gridDim.x, gridDim.y, gridDim.z));
}

// bound_prim() x2 faces over all i (rows gj 0..5) and POLEDEATH of the two
// interior pole rows (gj 6,7).  A ghost copied from a pole row reads its U2,B2
// either before or after that row's x*0 and multiplies them by 0 itself, which
// gives the same signed zero either way.  Optionally the t,dt update.
__device__ __forceinline__ void bx2_body(KParams P, double* __restrict__ pv, DevScalars* __restrict__ ds, int do_dt, const dim3 bi, const dim3 gd) {
  const int i = (int)(bi.x * blockDim.x + threadIdx.x) - NG;
  const int gj = bi.y;
  if (do_dt && bi.x == 0 && gj == 0 && threadIdx.x == 0) dt_update(P, ds);
  if (i > P.n1 + NG - 1) return;
  if (gj < 2 * NG) {
    bound_x2_cell(P, pv, i, gj);
  } else if (P.poledeath) {
    const bool dn = (gj == 2 * NG);
    if ((dn ? P.bc[HARM_X2DN] : P.bc[HARM_X2UP]) != HARM_POLARAXIS) return;
    double* a = pv + (size_t)zidx(P, i, dn ? 0 : P.n2 - 1) * 8;
    a[U2] *= 0.;
    a[B2] *= 0.;
  }
}

__global__ void k_bx2(KParams P, double* __restrict__ pv, DevScalars* __restrict__ ds, int do_dt) {
  bx2_body(P, pv, ds, do_dt, dim3(blockIdx.x, blockIdx.y, blockIdx.z),
      dim3(gridDim.x, gridDim.y, gridDim.z));
}

// bound_prim() in one launch, equal to k_bx1 then k_bx2.  Every ghost zone
// is computed from the interior by composing the rules the reference applies
// in turn (a corner: the x2 rule applied to the x1 rule's value).  POLEDEATH
// zeroes U2,B2 of the interior pole rows, which the x1 rule must read first:
// the x1 ghosts of a pole row, and the corners reflecting them, are filled
// by the first 4*NG threads of workgroup 0 (one per pole row and ghost
// column), which read all their sources, meet at a barrier, and only then
// zero the sources.  Every other reader of a pole row is an x2 ghost, which
// multiplies the same U2,B2 by 0 itself, so it gets the same signed zero
// whether it reads before or after the zeroing.
// Threads: [pole-row x1 ghosts 4*NG][x1 ghosts 2*NG*N2][x2 ghosts 2*NG*(N1+2NG)][pole rows 2*N1]
constexpr int BOUND_NP = 4 * NG;
__host__ __device__ __forceinline__ int bound_threads(const KParams& P) {
  return BOUND_NP + 2 * NG * P.n2 + 2 * NG * (P.n1 + 2 * NG) + 2 * P.n1;
}

// interior zone i of a pole row is read by the x1 rule (a pole-row thread's source)
__device__ __forceinline__ bool x1_reads(const KParams& P, int i) {
  const int bdn = P.bc[HARM_X1DN], bup = P.bc[HARM_X1UP];
  if (bdn == HARM_OUTFLOW && i == 0) return true;
  if (bup == HARM_OUTFLOW && i == P.n1 - 1) return true;
  if (bdn == HARM_PERIODIC && i >= P.n1 - NG) return true;
  if (bup == HARM_PERIODIC && i < NG) return true;
  return false;
}

__device__ __forceinline__ void bound_fused_body(KParams P, double* __restrict__ pv, const double* __restrict__ gC,
                                                 DevScalars* __restrict__ ds, int reset_list, int do_dt, int t) {
  if (t == 0) {
    if (reset_list) ds->nlist = 0;
    if (do_dt) dt_update(P, ds);
  }
  const int NI = P.n1 + 2 * NG;
  const bool pdn = P.poledeath && P.bc[HARM_X2DN] == HARM_POLARAXIS;
  const bool pup = P.poledeath && P.bc[HARM_X2UP] == HARM_POLARAXIS;
  int fail = 0;
  // pole-row x1 ghosts and the corners reflecting them (workgroup 0 only;
  // the barrier is reached by every thread of workgroup 0)
  if (blockIdx.x == 0) {
    const int row = threadIdx.x / (2 * NG), g = threadIdx.x % (2 * NG);  // g < NG: i = -1-g, else N1+g-NG
    const int i = (g < NG) ? -1 - g : P.n1 + g - NG;
    const int xb = (g < NG) ? P.bc[HARM_X1DN] : P.bc[HARM_X1UP];
    const bool mine = threadIdx.x < BOUND_NP && (row == 0 ? pdn : pup) &&
                      (xb == HARM_OUTFLOW || xb == HARM_PERIODIC);
    const int jp = (row == 0) ? 0 : P.n2 - 1, jc = (row == 0) ? -1 : P.n2;
    double v[NPR];
    if (mine) load8(pv, zidx(P, x1_source(P, i), jp), v);
    __syncthreads();  // every source read before any is zeroed
    if (mine) {
      x1_value(P, gC, i, jp, v, fail);
      v[U2] *= 0.;
      v[B2] *= 0.;
      store8(pv, zidx(P, i, jp), v);
      store8(pv, zidx(P, i, jc), v);
      double* a = pv + (size_t)zidx(P, x1_source(P, i), jp) * 8;
      a[U2] *= 0.;
      a[B2] *= 0.;
    }
  }
  t -= BOUND_NP;
  const int nA = 2 * NG * P.n2, nB = 2 * NG * NI;
  if (t < 0) {
    // a pole-row thread
  } else if (t < nA) {  // x1 ghosts off the pole rows
    const int gi = t / P.n2, j = t % P.n2;
    if (!((j == 0 && pdn) || (j == P.n2 - 1 && pup))) bound_x1_cell(P, pv, gC, gi, j, fail);
  } else if ((t -= nA) < nB) {  // x2 ghosts
    const int i = t % NI - NG, gj = t / NI;
    const bool dn = gj < NG;
    const int j = dn ? (gj - NG) : (P.n2 + gj - NG);
    const int bc = dn ? P.bc[HARM_X2DN] : P.bc[HARM_X2UP];
    int js;
    if (bc == HARM_PERIODIC) {
      js = dn ? j + P.n2 : j - P.n2;
    } else if (bc == HARM_POLARAXIS) {
      js = dn ? (0 + (0 - j - 1)) : ((P.n2 - 1) + ((P.n2 - 1) - j + 1));
    } else {
      js = -1000;
    }
    const bool zero = (bc == HARM_POLARAXIS) && P.poledeath;
    if (js != -1000) {
      double v[NPR];
      bool done = true;
      if (i >= 0 && i < P.n1) {
        load8(pv, zidx(P, i, js), v);
      } else if (zero && js == (dn ? 0 : P.n2 - 1)) {
        done = false;  // a pole-row thread fills it
      } else {
        const int xb = (i < 0) ? P.bc[HARM_X1DN] : P.bc[HARM_X1UP];
        if (xb != HARM_OUTFLOW && xb != HARM_PERIODIC) {
          done = false;
        } else {
          load8(pv, zidx(P, x1_source(P, i), js), v);
          x1_value(P, gC, i, js, v, fail);
        }
      }
      if (done) {
        if (zero) {
          v[U2] *= 0.;
          v[B2] *= 0.;
        }
        store8(pv, zidx(P, i, j), v);
      }
    }
  } else if ((t -= nB) < 2 * P.n1) {  // POLEDEATH of the other interior pole-row zones
    const bool dn = t < P.n1;
    const int i = t % P.n1;
    if (!(dn ? pdn : pup) || x1_reads(P, i)) return;
    double* a = pv + (size_t)zidx(P, i, dn ? 0 : P.n2 - 1) * 8;
    a[U2] *= 0.;
    a[B2] *= 0.;
  }
  if (fail) atomicOr(&ds->failed, 1);
}

__global__ __launch_bounds__(256) void k_bound(KParams P, double* __restrict__ pv, const double* __restrict__ gC,
                                               DevScalars* __restrict__ ds, int reset_list, int do_dt) {
  bound_fused_body(P, pv, gC, ds, reset_list, do_dt, blockIdx.x * blockDim.x + threadIdx.x);
}

// fixup_utoprim() over the failed-zone list: compute, then apply
__device__ __forceinline__ void fixlist_body(KParams P, const double* __restrict__ pv, const int* __restrict__ pflag,
                          const int* __restrict__ cs, const int* __restrict__ faillist, double* __restrict__ pfix,
                          const double* __restrict__ gC, DevScalars* __restrict__ ds, const dim3 bi, const dim3 gd) {
  int n = ds->nlist, fail = 0;
  if (n > P.n1 * P.n2) n = P.n1 * P.n2;
  for (int t = bi.x * blockDim.x + threadIdx.x; t < n; t += gd.x * blockDim.x)
    fixup_zone(P, pv, pflag, cs, pfix, gC, faillist[t], fail);
  if (fail) atomicOr(&ds->failed, 1);
}

__global__ void k_fixlist(KParams P, const double* __restrict__ pv, const int* __restrict__ pflag,
                          const int* __restrict__ cs, const int* __restrict__ faillist, double* __restrict__ pfix,
                          const double* __restrict__ gC, DevScalars* __restrict__ ds) {
  fixlist_body(P, pv, pflag, cs, faillist, pfix, gC, ds, dim3(blockIdx.x, blockIdx.y, blockIdx.z),
      dim3(gridDim.x, gridDim.y, gridDim.z));
}

// the scalar tail of a fixup-only boundary pass: clear the list, update dt
__global__ void k_tail(KParams P, DevScalars* __restrict__ ds, int do_dt) {
  ds->nlist = 0;
  if (do_dt) dt_update(P, ds);
}

__device__ __forceinline__ void fixapply_body(KParams P, double* __restrict__ pv, int* __restrict__ pflag,
                           const int* __restrict__ faillist, const double* __restrict__ pfix,
                           const DevScalars* __restrict__ ds, const dim3 bi, const dim3 gd) {
  int n = ds->nlist;
  if (n > P.n1 * P.n2) n = P.n1 * P.n2;
  for (int t = bi.x * blockDim.x + threadIdx.x; t < n; t += gd.x * blockDim.x) {
    const int z = faillist[t];
#pragma unroll
    for (int k = 0; k <= U3; k++) pv[(size_t)z * 8 + k] = pfix[(size_t)z * 8 + k];
    pflag[z] = -1;  // UTOPRIMFAILFIXED (global.h:524)
  }
}

__global__ void k_fixapply(KParams P, double* __restrict__ pv, int* __restrict__ pflag,
                           const int* __restrict__ faillist, const double* __restrict__ pfix,
                           const DevScalars* __restrict__ ds) {
  fixapply_body(P, pv, pflag, faillist, pfix, ds, dim3(blockIdx.x, blockIdx.y, blockIdx.z),
      dim3(gridDim.x, gridDim.y, gridDim.z));
}

// fixup_utoprim over the whole failed-zone list in one workgroup: every
// repaired state from the unrepaired array, a barrier, then all written (the
// reference reads its copy prc, fixup.c:527-920), then the list is cleared
// and, after the last substage without a second bound_prim, the t,dt update
__global__ __launch_bounds__(1024) void k_fixall(KParams P, double* __restrict__ pv, int* __restrict__ pflag,
                                                 const int* __restrict__ cs, const int* __restrict__ faillist,
                                                 double* __restrict__ pfix, const double* __restrict__ gC,
                                                 DevScalars* __restrict__ ds, int do_dt) {
  int n = ds->nlist, fail = 0;
  if (n > P.n1 * P.n2) n = P.n1 * P.n2;
  for (int t = threadIdx.x; t < n; t += blockDim.x) fixup_zone(P, pv, pflag, cs, pfix, gC, faillist[t], fail);
  __syncthreads();
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    const int z = faillist[t];
#pragma unroll
    for (int k = 0; k <= U3; k++) pv[(size_t)z * 8 + k] = pfix[(size_t)z * 8 + k];
    pflag[z] = -1;  // UTOPRIMFAILFIXED (global.h:524)
  }
  if (fail) atomicOr(&ds->failed, 1);
  if (threadIdx.x == 0) {  // every thread read nlist before the barrier
    ds->nlist = 0;
    if (do_dt) dt_update(P, ds);
  }
}

// The rest of a substage after a k_update that did the first bound_prim()
// (HARM_BOUND_INUPDATE), in one workgroup.  FIX: post_fixup's repair of the
// failed-zone list, every repaired state computed from the unrepaired array
// (kept in LDS, the first FIXKEEP of them, else read back from pfix), then,
// after one barrier, all at once: each repaired zone written (with
// POLEDEATH's zeroing on a pole row), and the second bound_prim() as the
// ghost images of what the repair and POLEDEATH can change: the repaired
// zones' images from their repaired states, and the x1 columns of the pole
// rows' outflow edge zones, which the reference recomputes from the zeroed
// rows (copied from what k_update computed for them when N1 >= 5, else
// computed here).  Every other ghost keeps the value it holds.  Without FIX
// (no post_fixup: force-free) only the pole rows' columns, the list left as
// it is.  Then the t,dt update after the last substage.
constexpr int FIXKEEP = 512;
template <bool FIX>
__global__ __launch_bounds__(256) void k_fixbound(KParams P, double* __restrict__ pv, int* __restrict__ pflag,
                                                  const int* __restrict__ faillist, double* __restrict__ pfix,
                                                  const double* __restrict__ gC, DevScalars* __restrict__ ds,
                                                  int do_dt) {
  __shared__ double keep[FIX ? FIXKEEP * NPR : 1];
  __shared__ int keepz[FIX ? FIXKEEP : 1];
  // the t, dt update after the last substage reads only the step's scalars:
  // one thread of the last wavefront, beside the repair
  if (do_dt && threadIdx.x == blockDim.x - 1) dt_update(P, ds);
  const int cap = P.n1 * P.n2;
  int n = FIX ? ds->nlist : 0, fail = 0;
  // the list's first entries read beside its length (used only below n)
  const int z0 = (FIX && (int)threadIdx.x < cap) ? faillist[threadIdx.x] : 0;
  if (n > cap) n = cap;
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    const int z = (t == (int)threadIdx.x) ? z0 : faillist[t];
    fixup_zone<true>(P, pv, pflag, nullptr, pfix, gC, z, fail, t < FIXKEEP ? keep + (size_t)t * NPR : nullptr);
    if (t < FIXKEEP) keepz[t] = z;
  }
  __syncthreads();  // every repaired state computed from the unrepaired array
  auto entry = [&](int q, double* v) {
    const int z = (q < FIXKEEP) ? keepz[q] : faillist[q];
    if (q < FIXKEEP) {
#pragma unroll
      for (int k = 0; k < NPR; k++) v[k] = keep[q * NPR + k];
    } else {
      load8(pfix, z, v);
    }
    return z;
  };
  const int nimg = n * NIMG;
  for (int t = threadIdx.x; t < n + nimg + 4 * NG; t += blockDim.x) {
    double v[NPR];
    if (t < n) {  // the repaired zone
      const int z = entry(t, v);
      const bool pole = pole_row(P, z % P.N2M - NG);
#pragma unroll
      for (int k = 0; k <= U3; k++) pv[(size_t)z * 8 + k] = (pole && k == U2) ? v[k] * 0. : v[k];
      if (pole) pv[(size_t)z * 8 + B2] = v[B2] * 0.;
      pflag[z] = -1;  // UTOPRIMFAILFIXED (global.h:524)
    } else if (t < n + nimg) {  // one part of its images
      const int q = (t - n) / NIMG, part = (t - n) % NIMG;
      const int z = entry(q, v);
      bound_image_part(P, pv, gC, z / P.N2M - NG, z % P.N2M - NG, v, part, fail);
    } else {  // column g of pole-row edge zone e (e&1 the x1 side, e&2 the x2 side)
      const int e = (t - n - nimg) / NG, g = (t - n - nimg) % NG;
      const int i = (e & 1) ? P.n1 - 1 : 0, j = (e & 2) ? P.n2 - 1 : 0;
      if (!pole_row(P, j) || P.bc[(e & 1) ? HARM_X1UP : HARM_X1DN] != HARM_OUTFLOW) continue;
      const int ze = zidx(P, i, j);
      bool repaired = false;  // then its images are the list's
      for (int q = 0; q < n && !repaired; q++) repaired = ((q < FIXKEEP) ? keepz[q] : faillist[q]) == ze;
      if (repaired) continue;
      const int ig = (e & 1) ? P.n1 + g : -1 - g;
      if (P.n1 >= 5) {  // computed by k_update from the zeroed row
#pragma unroll
        for (int k = 0; k < NPR; k++) v[k] = ds->pole_x1[e][g][k];
        if (ds->pole_fail[e][g]) fail = 1;
        store8(pv, zidx(P, ig, j), v);
        bound_x2_images(P, pv, ig, j, v);
      } else {
        load8(pv, ze, v);
        bound_image_part(P, pv, gC, i, j, v, (e & 1) ? NG + 1 + g : 1 + g, fail);
      }
    }
  }
  if (fail) atomicOr(&ds->failed, 1);
  if (FIX && threadIdx.x == 0) ds->nlist = 0;  // every thread read nlist before the barrier
}

// ---- bound_prim() with the general rules: the outflow extrapolation
// (EXTRAP 1), FIXED sides, and bounds.c's bound_disk().  One launch per
// phase of the reference's loops, in its order: x1 faces, x2 faces,
// POLEDEATH, the disk.  Taken for every schedule variant when the context
// uses any of them (general_bounds()). ----

// the outflow ghost (i, j) from edge zone (ri, j) before the inflow check:
// EXTRAP 0 a copy; EXTRAP 1 the densities and B1 scaled by gdet, the other
// components extrapolated linearly, the up side's U1 with twice the slope
// (bounds.fishmon.c:37-56, 194-217; bounds.c:40-56, 129-146)
__device__ __forceinline__ void outflow_value(const KParams& P, const double* __restrict__ pv,
                                              const double* __restrict__ gC, int i, int j, int ri, double* v) {
  const double* src = pv + (size_t)zidx(P, ri, j) * 8;
#pragma unroll
  for (int k = 0; k < NPR; k++) v[k] = src[k];
  if (!P.extrap) return;
  if (P.extrap == 2) {  // rescale(1, ...) of the edge zone, rescale(-1, ...) at the ghost
    const double* sr = P.rscale + (size_t)zidx(P, ri, j) * 8;
    const double* si = P.rscale + (size_t)zidx(P, i, j) * 8;
#pragma unroll
    for (int k = 0; k < NPR; k++) v[k] = (src[k] / sr[k]) * si[k];
    return;
  }
  const double gr = gC[(size_t)zidx(P, ri, j) * GCN + GC_GDET], gi = gC[(size_t)zidx(P, i, j) * GCN + GC_GDET];
  const bool up = ri > 0;
  const double dx1 = P.dx1;
  for (int k = RHO; k <= UU; k++) v[k] = src[k] * gr / gi;
  v[U1] = up ? src[U1] * (1. - 2 * (i - ri) * dx1) : src[U1] * (1. - (i - ri) * dx1);
  for (int k = U2; k <= U3; k++) v[k] = up ? src[k] * (1. - (i - ri) * dx1) : src[k] * (1. + (i - ri) * dx1);
  v[B1] = src[B1] * gr / gi;
  for (int k = B2; k <= B3; k++) v[k] = up ? src[k] * (1. - (i - ri) * dx1) : src[k] * (1. + (i - ri) * dx1);
}

// x1 ghost columns: thread (column g, j), j over [-NG, N2+NG): outflow and
// periodic fill interior j; FIXED copies panalytic over every j with the
// bounds.c rules and leaves the ghosts alone with bounds.fishmon.c's
__global__ void k_bg_x1(KParams P, double* __restrict__ pv, const double* __restrict__ gC,
                        const double* __restrict__ pan, DevScalars* __restrict__ ds) {
  const int nj = P.n2 + 2 * NG;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * NG * nj) return;
  const int g = t / nj, j = t % nj - NG;
  const bool up = g >= NG;
  const int i = up ? P.n1 + (g - NG) : -1 - g;
  const int bc = up ? P.bc[HARM_X1UP] : P.bc[HARM_X1DN];
  double v[NPR];
  int fail = 0;
  if (bc == HARM_FIXED) {
    if (P.boundsfile != HARM_BOUNDS_DISK) return;
    load8(pan, zidx(P, i, j), v);
  } else if (j < 0 || j >= P.n2) {
    return;
  } else if (bc == HARM_PERIODIC) {
    load8(pv, zidx(P, up ? i - P.n1 : i + P.n1, j), v);
  } else if (bc == HARM_OUTFLOW) {
    outflow_value(P, pv, gC, i, j, up ? P.n1 - 1 : 0, v);
    x1_value(P, gC, i, j, v, fail);
  } else {
    return;
  }
  store8(pv, zidx(P, i, j), v);
  if (fail) atomicOr(&ds->failed, 1);
}

// x2 ghost rows over i in [-NG, N1+NG): periodic wrap, polar reflection
// (bounds.fishmon.c:263-279), FIXED from panalytic (bounds.c rules)
__global__ void k_bg_x2(KParams P, double* __restrict__ pv, const double* __restrict__ pan) {
  const int ni = P.n1 + 2 * NG;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * NG * ni) return;
  const int g = t / ni, i = t % ni - NG;
  const bool up = g >= NG;
  const int j = up ? P.n2 + (g - NG) : -1 - g;
  const int bc = up ? P.bc[HARM_X2UP] : P.bc[HARM_X2DN];
  double v[NPR];
  if (bc == HARM_PERIODIC) {
    load8(pv, zidx(P, i, up ? j - P.n2 : j + P.n2), v);
  } else if (bc == HARM_POLARAXIS) {
    load8(pv, zidx(P, i, up ? (P.n2 - 1) + ((P.n2 - 1) - j + 1) : 0 + (0 - j - 1)), v);
    if (P.boundsfile == HARM_BOUNDS_DISK) {  // bounds.c:214-233, 268-278: u and b antisymmetric across the pole
      v[U2] *= -1.;
      v[U3] *= 1.;
      v[B2] *= -1.;
      v[B3] *= 1.;
    }
  } else if (bc == HARM_FIXED && P.boundsfile == HARM_BOUNDS_DISK) {
    load8(pan, zidx(P, i, j), v);
  } else {
    return;
  }
  store8(pv, zidx(P, i, j), v);
}

// POLEDEATH (bounds.fishmon.c:322-355): U2, B2 zeroed on the ghost rows and
// the pole row of a polar side, every i
__global__ void k_bg_pole(KParams P, double* __restrict__ pv) {
  const int ni = P.n1 + 2 * NG;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * (NG + 1) * ni) return;
  const int r = t / ni, i = t % ni - NG;
  const bool up = r > NG;
  if ((up ? P.bc[HARM_X2UP] : P.bc[HARM_X2DN]) != HARM_POLARAXIS) return;
  const int j = up ? P.n2 - 1 + (r - NG - 1) : -NG + r;
  double* a = pv + (size_t)zidx(P, i, j) * 8;
  a[U2] *= 0.;
  a[B2] *= 0.;
}

constexpr int DISK_DELTAJ = 2;        // bounds.c:378
constexpr double DISK_TSTATIC = 15.0;  // bounds.c:387

// bounds.c:350-760 bound_disk(), one thread per (i, pair of zones mirrored
// across the equator, DELTAJ pairs): after the first call the pair made
// exactly symmetric (B1, B3, U2 antisymmetric); then per zone B1, B3 from
// panalytic until t > TSTATIC and from the zone DELTAJ+1 off the equator
// after, B2 from panalytic; the velocity from E_phi = 0 with the field
// line's angular velocity omegafanalytic[.][0] (get_omegastar, init.c:739),
// through vcon2pr, or U2 = 0 when that fails (clearing the failure flag as
// the reference's failed=0 does).  The rows DELTAJ+1 off the equator are
// only read, so the pairs are independent.  RDISKINNER 0, RDISKOUTER 1e30:
// the reference's radius tests hold for every zone of a cylindrical-
// Minkowski grid and are not restated.  N2 even (checked at harm_create).
__global__ void k_bg_disk(KParams P, double* __restrict__ pv, const double* __restrict__ pan,
                          const double* __restrict__ omf, const double* __restrict__ gC, int* __restrict__ pflag,
                          DevScalars* __restrict__ ds) {
  const int ni = P.n1 + 2 * NG;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= DISK_DELTAJ * ni) return;
  const int i = t % ni - NG, q = t / ni;
  const int half = P.n2 / 2;
  const int ja = half - DISK_DELTAJ + q, jb = half - ja + (P.n2 - 1) / 2;
  double* a = pv + (size_t)zidx(P, i, ja) * 8;
  double* b = pv + (size_t)zidx(P, i, jb) * 8;
  if (!ds->disk_first) {
#pragma unroll
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
  const double fun = (ds->t <= DISK_TSTATIC) ? 0.0 : 1.0;
  for (int side = 0; side < 2; side++) {
    const int j = side ? jb : ja;
    const int rj = side ? DISK_DELTAJ + half : -(DISK_DELTAJ + 1) + half;
    double* p = side ? b : a;
    const double* an = pan + (size_t)zidx(P, i, j) * 8;
    const double* rp = pv + (size_t)zidx(P, i, rj) * 8;
    p[B1] = (1.0 - fun) * an[B1] + fun * rp[B1];
    p[B3] = (1.0 - fun) * an[B3] + fun * rp[B3];
    p[B2] = an[B2];
    double gc[36];
    load_cent(gC, zidx(P, i, j), gc);
    GeoU g;
    geo_from_cent(gc, g);
    double Bcon[4], Bcov[4], Bsq;
    Bcon[0] = 0;
    Bcon[1] = p[B1];
    Bcon[2] = p[B2];
    Bcon[3] = p[B3];
    lower(Bcon, g, Bcov);
    Bsq = 0.0;
    for (int k = 0; k <= 3; k++) Bsq += Bcon[k] * Bcov[k];
    const double Om = omf[(size_t)zidx(P, i, j) * 8];
    const double v1 = -Bcon[1] * (Bcov[0] + Om * Bcov[3]) / Bsq;
    const double v2 = -Bcon[2] * (Bcov[0] + Om * Bcov[3]) / Bsq;
    const double v3 = Om - Bcon[3] * (Bcov[0] + Om * Bcov[3]) / Bsq;
    // vcon2pr(VELREL4): ucon_calc_3vel (phys.c:885-955), u^i returned
    const double discr = +gc[0] + gc[5] * v1 * v1 + gc[10] * v2 * v2 + gc[15] * v3 * v3 +
                         2. * (gc[1] * v1 + gc[2] * v2 + gc[3] * v3 + gc[6] * v1 * v2 + gc[7] * v1 * v3 +
                               gc[11] * v2 * v3);
    bool sane = true;
    if (discr > 0.) {
      ds->failed = 0;
      sane = false;
    } else {
      const double ut = 1. / sqrt(-discr);
      const double w1 = v1 * ut, w2 = v2 * ut, w3 = v3 * ut;
      if (isnan(w1) || isnan(w2) || isnan(w3)) {
        sane = false;
      } else {
        p[U1] = w1;
        p[U2] = w2;
        p[U3] = w3;
      }
    }
    if (!sane) p[U2] = 0.0;
    if (sane) pflag[zidx(P, i, j)] = 0;
  }
}

// after a bound_disk() call: the next one symmetrises
__global__ void k_bg_diskdone(DevScalars* __restrict__ ds) { ds->disk_first = 0; }

__global__ __launch_bounds__(128) void k_utoprim_batch(KParams P, int n, const int* __restrict__ zone, const double* __restrict__ U,
                                double* __restrict__ pr, int* __restrict__ pfl, const double* __restrict__ gC) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int z = zidx(P, zone[2 * t], zone[2 * t + 1]);
  double gc[36];
  load_cent(gC, z, gc);
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

// Staging memory of harm_utoprimgen_noalloc: the zone indices, U, the
// guesses/solutions and the flags of up to HARM_UTOPRIM_NOALLOC_MAX zones,
// laid out as [zones int2][U 8 doubles][pr 8 doubles][pflag int] per block.
// Host code never passes the array's name to a kernel: it resolves the
// array's device address per device once (hipGetSymbolAddress) and uses that.
constexpr size_t NOALLOC_ZONE_OFF = 0;
constexpr size_t NOALLOC_U_OFF = NOALLOC_ZONE_OFF + (size_t)HARM_UTOPRIM_NOALLOC_MAX * 2 * sizeof(int);
constexpr size_t NOALLOC_PR_OFF = NOALLOC_U_OFF + (size_t)HARM_UTOPRIM_NOALLOC_MAX * 8 * sizeof(double);
constexpr size_t NOALLOC_FL_OFF = NOALLOC_PR_OFF + (size_t)HARM_UTOPRIM_NOALLOC_MAX * 8 * sizeof(double);
constexpr size_t NOALLOC_BYTES = NOALLOC_FL_OFF + (size_t)HARM_UTOPRIM_NOALLOC_MAX * sizeof(int);
static_assert(NOALLOC_U_OFF % 256 == 0 && NOALLOC_PR_OFF % 256 == 0 && NOALLOC_FL_OFF % 256 == 0,
              "staging blocks 256-byte aligned");

// ---------------------------------------------------------------------------
// CHECKSOLUTION: get_bsqflags (pre_fixup) and fixup_checksolution (post_fixup)
// ---------------------------------------------------------------------------
constexpr double BSQORHOLIMIT = 20.0;     // initbase.c:213
constexpr double BSQOULIMIT = 100.0;      // initbase.c:214
constexpr double GAMMAPERCDIFFMAX = 2.0;  // global.h:264
constexpr double UPERCDIFFMAX = 10.0;     // global.h:265
constexpr int UTOPRIMFAILGAMMAPERC = 8;   // global.h:558
constexpr int UTOPRIMFAILUPERC = 9;       // global.h:559

// get_bsqflags (fixup.c:1197-1295, LIMADJUST=FLUXADJUST=0): b^2/rho and
// b^2/u flags (0, 1, 2) packed as f_rho | f_u << 8.  The reference also flags
// the ghost ring, but bound_pflag overwrites those with interior copies
// before anything reads them, so only the interior is computed.
__global__ void k_bsqflags(KParams P, const double* __restrict__ pv, const double* __restrict__ gC,
                           int* __restrict__ bsqf) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= P.n1 || j >= P.n2) return;
  const int z = zidx(P, i, j);
  int f0 = 0, f3 = 0;
  if (P.eomtype == HARM_EOMGRMHD) {
    double gc[36];
    load_cent(gC, z, gc);
    GeoU g;
    geo_from_cent(gc, g);
    double pr[NPR];
    load8(pv, z, pr);
    State q;
    get_state(pr, g, q);
    const double bsq = dot4(q.bcon, q.bcov);
    if (bsq / pr[RHO] > BSQORHOLIMIT) f0 = 2;
    else if (bsq / pr[RHO] > 0.9 * BSQORHOLIMIT) f0 = 1;
    if (bsq / pr[UU] > BSQOULIMIT) f3 = 2;
    else if (bsq / pr[UU] > 0.9 * BSQOULIMIT) f3 = 1;
  }
  bsqf[z] = f0 | (f3 << 8);
}

// fixup.c:400-413: gamma at the flagged zones whose inversion succeeded, over
// DQZLOOP (i in [-1,N1], j in [-1,N2]; ghost flags are the bound_pflag copies).
// gammacheck persists between calls like the reference's static array.
__global__ void k_gammacheck(KParams P, const double* __restrict__ pv, const double* __restrict__ gC,
                             const int* __restrict__ pflag, const int* __restrict__ bsqf,
                             double* __restrict__ gammacheck) {
  const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x) - 1;
  const int i = (int)(blockIdx.y * blockDim.y + threadIdx.y) - 1;
  if (i > P.n1 || j > P.n2) return;
  const int zm = pflag_zone(P, i, j);
  if (bsqf[zm] == 0 || pflag[zm] != 0) return;
  const int z = zidx(P, i, j);
  double gc[36];
  load_cent(gC, z, gc);
  GeoU g;
  geo_from_cent(gc, g);
  double pr[NPR];
  load8(pv, z, pr);
  gammacheck[z] = gamma_calc(pr, g);
}

// fixup.c:416-485 for one flagged interior zone: 0, UTOPRIMFAILGAMMAPERC or
// UTOPRIMFAILUPERC.  Neighbour failure flags: the reference loop runs in
// (i, j) order, so the three neighbours it has already passed ((i,j-1),
// (i-1,j), (i-1,j-1)) show the flags it set there; those come from `prev`
// (nullable), the others are the flags before the check.
__device__ int cs_vote(const KParams& P, const double* __restrict__ pv, const int* __restrict__ pflag,
                       const double* __restrict__ gammacheck, const int* __restrict__ prev, int i, int j) {
  const int di[8] = {0, 0, 1, -1, 1, 1, 1, -1}, dj[8] = {1, -1, 0, 0, 1, -1, 1, -1};
  const int z = zidx(P, i, j);
  const double gz = gammacheck[z], uz = pv[(size_t)z * 8 + UU];
  double pd[8], pd2[8];
#pragma unroll
  for (int l = 0; l < 8; l++) {
    const int a = i + di[l], b = j + dj[l];
    const int zn = zidx(P, a, b);
    int f = pflag[pflag_zone(P, a, b)];
    const bool earlier = interior(P, a, b) && (a < i || (a == i && b < j));
    if (prev && earlier && prev[zn]) f = prev[zn];
    const bool ok = (f == 0);
    pd[l] = ok ? (gammacheck[zn] / gz) : -1;
    pd2[l] = ok ? (pv[(size_t)zn * 8 + UU] / uz) : -1;
  }
  int vote0 = 0, vote1 = 0, numvotes = 0;
  const bool grmhd = (P.eomtype == HARM_EOMGRMHD);
#pragma unroll
  for (int l = 0; l < 8; l++) {
    if (pd[0] >= 0.0) {
      if ((fabs(pd[l]) > GAMMAPERCDIFFMAX) || (fabs(pd[l]) < 1.0 / GAMMAPERCDIFFMAX)) vote0++;
      if (grmhd && ((fabs(pd2[l]) > UPERCDIFFMAX) || (fabs(pd2[l]) < 1.0 / UPERCDIFFMAX))) vote1++;
      numvotes++;
    }
  }
  int r = 0;
  if (vote0 > numvotes * 0.5) r = UTOPRIMFAILGAMMAPERC;
  if (grmhd && (vote1 > numvotes * 0.5)) r = UTOPRIMFAILUPERC;
  return r;
}

// One parallel pass of the check: dst = outcome of every flagged zone with
// the earlier neighbours' outcomes taken from src (first pass: none).  The
// sequential result is the unique fixed point of this map, so a second pass
// equal to the first is exact; otherwise cs_changed is set.
__global__ void k_cs_eval(KParams P, const double* __restrict__ pv, const int* __restrict__ pflag,
                          const int* __restrict__ bsqf, const double* __restrict__ gammacheck,
                          const int* __restrict__ src, int* __restrict__ dst, DevScalars* __restrict__ ds,
                          int first) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (first && i == 0 && j == 0) ds->cs_changed = 0;
  if (i >= P.n1 || j >= P.n2) return;
  const int z = zidx(P, i, j);
  int r = 0;
  if (bsqf[z] != 0 && pflag[z] == 0) r = cs_vote(P, pv, pflag, gammacheck, first ? nullptr : src, i, j);
  dst[z] = r;
  if (!first && r != src[z]) atomicOr(&ds->cs_changed, 1);
}

// When the two passes differ: the reference's loop itself, in order, on one
// thread (exact, and rare: it needs a chain of failed neighbours).
__global__ void k_cs_sequential(KParams P, const double* __restrict__ pv, const int* __restrict__ pflag,
                                const int* __restrict__ bsqf, const double* __restrict__ gammacheck,
                                int* __restrict__ dst, const DevScalars* __restrict__ ds, int always) {
  if (!always && !ds->cs_changed) return;
  for (int i = 0; i < P.n1; i++)
    for (int j = 0; j < P.n2; j++) dst[zidx(P, i, j)] = 0;
  for (int i = 0; i < P.n1; i++)
    for (int j = 0; j < P.n2; j++) {
      const int z = zidx(P, i, j);
      if (bsqf[z] != 0 && pflag[z] == 0) dst[z] = cs_vote(P, pv, pflag, gammacheck, dst, i, j);
    }
}

// set the failure flags found by the check and put those zones on the list
// fixup_utoprim repairs
__global__ void k_cs_apply(KParams P, const int* __restrict__ cs, int* __restrict__ pflag,
                           int* __restrict__ faillist, DevScalars* __restrict__ ds) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= P.n1 || j >= P.n2) return;
  const int z = zidx(P, i, j);
  if (!cs[z]) return;
  pflag[z] = cs[z];
  const int slot = atomicAdd(&ds->nlist, 1);
  if (slot < P.n1 * P.n2) faillist[slot] = z;
  atomicAdd(&ds->nfail, 1);
}

__global__ void k_pflag_expand(int nz, const int* __restrict__ pf, const int* __restrict__ bsqf,
                               int* __restrict__ out) {
  int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= nz) return;
  out[(size_t)z * 5] = pf[z];
  out[(size_t)z * 5 + 1] = 0;                  // FLAGREALLIM (LIMADJUST=0)
  out[(size_t)z * 5 + 2] = bsqf[z] & 0xff;     // FLAGBSQORHO
  out[(size_t)z * 5 + 3] = bsqf[z] >> 8;       // FLAGBSQOU
  out[(size_t)z * 5 + 4] = 0;                  // FLAGREALFLUX (FLUXADJUST=0)
}


// ---------------------------------------------------------------------------
// Batched runs: many independent grids advanced by one launch per stage.
// Each run is described by a RunDesc in device memory; a per-kernel block
// table maps every block of the launch to (run, local block x, y, z), so the
// per-run kernel bodies run unchanged on each run's own arrays and scalars.
// ---------------------------------------------------------------------------
struct RunDesc {
  KParams P;
  double* arr[HARM_NARRAYS];
  double *q1, *q2, *pfix;
  const double *gF, *gC, *gConn;
  int *pflag, *faillist;
  DevScalars* ds;
};

constexpr int FIXBLK = 2;  // blocks per run for the failed-list kernels

__global__ __launch_bounds__(256) void kb_flux(const RunDesc* __restrict__ runs, const int4* __restrict__ tab,
                                               int pr_sel, int dir0) {
  const int4 e = tab[blockIdx.x];
  const RunDesc& r = runs[e.x];
  flux_single_body(r.P, r.arr[pr_sel], r.arr[HARM_F1], r.arr[HARM_F2], r.q1, r.q2, r.gF, r.ds, dir0,
                   dim3(e.y, e.z, e.w), dim3(0, 0, 0));
}

__global__ __launch_bounds__(256) void kb_update(const RunDesc* __restrict__ runs, const int4* __restrict__ tab,
                                                 RKCoef C, int pi, int pb, int pf) {
  const int4 e = tab[blockIdx.x];
  const RunDesc& r = runs[e.x];
  update_body<0>(r.P, C, r.arr[pi], r.arr[pb], r.arr[pf], r.arr[HARM_ULAST], r.arr[HARM_UNEW], r.arr[HARM_F1],
              r.arr[HARM_F2], r.q1, r.q2, r.gC, r.gConn, r.pflag, r.faillist, r.ds, nullptr, nullptr,
              dim3(e.y, e.z, e.w), dim3(0, 0, 0));
}

__global__ __launch_bounds__(1024) void kb_boundary(const RunDesc* __restrict__ runs, int sel, int mode) {
  const RunDesc& r = runs[blockIdx.x];
  const int m = (r.P.eomtype == HARM_EOMGRMHD) ? mode : (mode & ~KB_FIXUP);
  boundary_body(r.P, r.arr[sel], r.gC, r.pflag, r.faillist, r.pfix, r.ds, m, dim3(0, 0, 0), dim3(1, 1, 1));
}

__global__ void kb_bx1(const RunDesc* __restrict__ runs, const int4* __restrict__ tab, int sel, int reset_list) {
  const int4 e = tab[blockIdx.x];
  const RunDesc& r = runs[e.x];
  bx1_body(r.P, r.arr[sel], r.gC, r.ds, reset_list, dim3(e.y, e.z, e.w), dim3(0, 0, 0));
}

__global__ void kb_bx2(const RunDesc* __restrict__ runs, const int4* __restrict__ tab, int sel, int do_dt) {
  const int4 e = tab[blockIdx.x];
  const RunDesc& r = runs[e.x];
  bx2_body(r.P, r.arr[sel], r.ds, do_dt, dim3(e.y, e.z, e.w), dim3(0, 0, 0));
}

// failed-list kernels: FIXBLK blocks per run, block b of run r is blockIdx.x = r*FIXBLK+b
__global__ void kb_fixlist(const RunDesc* __restrict__ runs, int sel) {
  const RunDesc& r = runs[blockIdx.x / FIXBLK];
  fixlist_body(r.P, r.arr[sel], r.pflag, nullptr, r.faillist, r.pfix, r.gC, r.ds, dim3(blockIdx.x % FIXBLK, 0, 0),
               dim3(FIXBLK, 1, 1));
}

__global__ void kb_fixapply(const RunDesc* __restrict__ runs, int sel) {
  const RunDesc& r = runs[blockIdx.x / FIXBLK];
  fixapply_body(r.P, r.arr[sel], r.pflag, r.faillist, r.pfix, r.ds, dim3(blockIdx.x % FIXBLK, 0, 0),
                dim3(FIXBLK, 1, 1));
}

}  // namespace

// The staging array of harm_utoprimgen_noalloc (layout above); every GPU the
// library runs on holds its own copy until the process ends.
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

// its device address on each device, resolved once per device with that
// device current (the array's name is a host-side stand-in, not an address),
// and a lock per device that keeps calls sharing one device's copy apart
constexpr int NOALLOC_MAXDEV = 64;
static unsigned char* g_scratch_ptr[NOALLOC_MAXDEV];
static std::mutex g_scratch_mu[NOALLOC_MAXDEV];

static int scratch_resolve(int dev, unsigned char** out) {
  if (dev < 0 || dev >= NOALLOC_MAXDEV) return seterr("device %d beyond the staging table (%d)", dev, NOALLOC_MAXDEV);
  std::lock_guard<std::mutex> lk(g_scratch_mu[dev]);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return seterr("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = (unsigned char*)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

// ===========================================================================
// C-ABI
// ===========================================================================

namespace {
struct AcctState;  // conservation accounting (harm_acct.inl)
struct CurState;   // the current J, F (harm_diag.inl)
}

struct harm_ctx {
  harm_params prm;
  KParams kp;
  int N1M, N2M, nz;
  double* arr[HARM_NARRAYS];
  double* q1;     // F1[B2]/gdet(FACE1) per face
  double* q2;     // F2[B1]/gdet(FACE2) per face
  double* emf;    // corner EMFs (standalone flux_ct only)
  double* pfix;   // repaired states of failed zones
  int* pflag;     // FLAGUTOPRIMFAIL per zone
  int* faillist;  // zones whose inversion failed in the current substage
  int* bsqf;      // get_bsqflags: b^2/rho | b^2/u << 8
  double* gammacheck;  // fixup_checksolution's persistent gamma table
  int* csA;       // fixup_checksolution outcomes, first and second pass
  int* csB;
  double* enoA;   // DOENO scratch (uenotmp0..2 of defs.h:80-82, two suffice)
  double* enoB;
  double* gF;
  double* gC;
  double* gConn;
  DevScalars* ds;
  hipStream_t stream;
  hipStream_t own_stream;
  hipGraphExec_t graph;             // one step_ch() as a graph
  bool graph_ok;
  hipGraphExec_t graphK;            // STEPS_PER_GRAPH steps in one graph (harm_step_n)
  bool graphK_ok;
  bool geometry_set;
  bool prof;
  int flux_variant;   // HARM_FLUX_* chosen by the autotuner or the caller
  int bound_variant;  // HARM_BOUND_*
  bool tuned;
  std::vector<hipEvent_t>* evpool;  // pairs (start, end) recorded without waiting
  std::vector<int>* evkid;          // kernel id of each recorded pair
  int evused;
  int pkid;                         // stage whose kernels are being timed, -1: none
  double* pan;                      // panalytic [nz][8] (harm_set_analytic), FIXED and bound_disk
  double* omf;                      // omegafanalytic [nz][8], bound_disk
  double* rscale;                   // EXTRAP 2 rescale() factors [nz][8] (harm_set_radius)
  int perr;                         // a timed launch could not get its events
  double ktime[HARM_K_COUNT];
  long long kcount[HARM_K_COUNT];
  AcctState* acct;  // null until accounting or a dump-time reduction is first used
  CurState* cur;    // null until the current is first enabled
  int device;               // the GPU the context was created on
  unsigned char* scratch;   // that device's staging array (harm_utoprimgen_noalloc)
};

static void cur_free(harm_ctx* c);
static bool cur_on(const harm_ctx* c);
static void cur_set_on(harm_ctx* c, bool on);
static int cur_step(harm_ctx* c);

static void acct_free(harm_ctx* c);
static bool acct_on(const harm_ctx* c);
static void acct_set_on(harm_ctx* c, bool on);
static int acct_source(harm_ctx* c, int pi, int pb, int pf);
static int acct_flux(harm_ctx* c, double Dt);
static int acct_faces(harm_ctx* c, int pb);
static int acct_zones(harm_ctx* c, int finalstep);
static int acct_merge(harm_ctx* c, int finalstep);

extern "C" const char* harm_last_error(void) { return g_err; }

static int check_prm(const harm_params* p) {
  if (!p) return seterr("null params");
  if (p->n1 < 4 || p->n2 < 4) return seterr("n1,n2 must be >= 4 (got %d,%d)", p->n1, p->n2);
  if ((long long)(p->n1 + 6) * (p->n2 + 6) > (1LL << 28)) return seterr("grid too large");
  if (p->eomtype != HARM_EOMGRMHD && p->eomtype != HARM_EOMFFDE) return seterr("bad eomtype %d", p->eomtype);
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
    if (p->bc[k] != HARM_OUTFLOW && p->bc[k] != HARM_POLARAXIS && p->bc[k] != HARM_PERIODIC &&
        p->bc[k] != HARM_FIXED)
      return seterr("boundary type %d not supported", p->bc[k]);
  if (p->boundsfile != HARM_BOUNDS_FISHMON && p->boundsfile != HARM_BOUNDS_DISK)
    return seterr("boundsfile must be HARM_BOUNDS_FISHMON or HARM_BOUNDS_DISK");
  // bounds.fishmon.c has no FIXED branch: the reference leaves such ghosts
  // alone, so its Runge-Kutta stage buffers keep set_arrays()'s zeros there
  for (int k = 0; k < 4; k++)
    if (p->bc[k] == HARM_FIXED && p->boundsfile != HARM_BOUNDS_DISK)
      return seterr("FIXED sides need HARM_BOUNDS_DISK (bounds.c fills them from panalytic)");
  if (p->extrap < 0 || p->extrap > 2) return seterr("extrap must be 0, 1 or 2");
  if (p->extrap == 2 && p->rescaletype != 0 && p->rescaletype != 1) return seterr("rescaletype must be 0 or 1");
  if (p->boundsfile == HARM_BOUNDS_DISK && p->n2 % 2)
    return seterr("bound_disk needs an even n2 (an equator between two rows)");
  if ((p->bc[HARM_X1DN] == HARM_POLARAXIS) || (p->bc[HARM_X1UP] == HARM_POLARAXIS))
    return seterr("POLARAXIS is an x2 boundary");
  if ((p->bc[HARM_X2DN] == HARM_OUTFLOW) || (p->bc[HARM_X2UP] == HARM_OUTFLOW))
    return seterr("x2 OUTFLOW is not implemented (x2 boundaries: POLARAXIS or PERIODIC)");
  if ((p->bc[HARM_X1DN] == HARM_PERIODIC) != (p->bc[HARM_X1UP] == HARM_PERIODIC) ||
      (p->bc[HARM_X2DN] == HARM_PERIODIC) != (p->bc[HARM_X2UP] == HARM_PERIODIC))
    return seterr("periodic boundaries must be set on both sides");
  if (!(p->dx1 > 0) || !(p->dx2 > 0)) return seterr("dx1, dx2 must be positive");
  if (p->checksolution != 0 && p->checksolution != 1) return seterr("checksolution must be 0 or 1");
  if (p->eno < 0 || p->eno > 2) return seterr("eno must be 0, 1 or 2");
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
  // advance_eno_u calls fix_flux whatever FIXUPFLUX says (advance.c:355)
  k.fixupflux = prm->fixupflux || prm->eno == 1;
  k.poledeath = prm->poledeath; k.checksolution = prm->checksolution; k.eno = prm->eno;
  k.gam = prm->gam; k.cour = prm->cour; k.defcon = prm->defcon; k.safe = prm->safe;
  k.gammamax = prm->gammamax; k.gammafail = prm->gammafail;
  k.rhominlimit = prm->rhominlimit; k.uuminlimit = prm->uuminlimit;
  k.dx1 = prm->dx1; k.dx2 = prm->dx2;
  k.boundsfile = prm->boundsfile; k.extrap = prm->extrap;
  c->evpool = new std::vector<hipEvent_t>();
  c->evkid = new std::vector<int>();
  c->pkid = -1;
  c->perr = 0;
  size_t nz = (size_t)c->nz;
  hipError_t e = hipSuccess;
  for (int a = 0; a < HARM_NARRAYS && e == hipSuccess; a++) {
    e = hipMalloc(&c->arr[a], nz * 8 * sizeof(double));
    if (e == hipSuccess) e = hipMemset(c->arr[a], 0, nz * 8 * sizeof(double));
  }
  double** dbufs[4] = {&c->q1, &c->q2, &c->emf, &c->pfix};
  size_t dsz[4] = {nz, nz, nz, nz * 8};
  for (int b = 0; b < 4 && e == hipSuccess; b++) {
    e = hipMalloc(dbufs[b], dsz[b] * sizeof(double));
    if (e == hipSuccess) e = hipMemset(*dbufs[b], 0, dsz[b] * sizeof(double));
  }
  if (e == hipSuccess) e = hipMalloc(&c->pflag, nz * sizeof(int));
  if (e == hipSuccess) e = hipMemset(c->pflag, 0, nz * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&c->bsqf, nz * sizeof(int));
  if (e == hipSuccess) e = hipMemset(c->bsqf, 0, nz * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&c->gammacheck, nz * sizeof(double));
  if (e == hipSuccess) e = hipMemset(c->gammacheck, 0, nz * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&c->csA, nz * sizeof(int));
  if (e == hipSuccess) e = hipMemset(c->csA, 0, nz * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&c->csB, nz * sizeof(int));
  if (e == hipSuccess) e = hipMemset(c->csB, 0, nz * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&c->faillist, nz * sizeof(int));
  if (prm->eno) {
    if (e == hipSuccess) e = hipMalloc(&c->enoA, nz * 8 * sizeof(double));
    if (e == hipSuccess) e = hipMemset(c->enoA, 0, nz * 8 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&c->enoB, nz * 8 * sizeof(double));
    if (e == hipSuccess) e = hipMemset(c->enoB, 0, nz * 8 * sizeof(double));
  }
  if (e == hipSuccess) e = hipMalloc(&c->gF, 2 * nz * GF * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&c->gC, nz * GCN * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&c->gConn, nz * NCONN * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&c->ds, sizeof(DevScalars));
  if (e == hipSuccess) e = hipStreamCreate(&c->own_stream);
  if (e == hipSuccess) e = hipGetDevice(&c->device);
  if (e != hipSuccess) {
    seterr("harm_create: %s", hipGetErrorString(e));
    harm_destroy(c);
    return 1;
  }
  if (scratch_resolve(c->device, &c->scratch)) {
    harm_destroy(c);
    return 1;
  }
  c->stream = c->own_stream;
  DevScalars h;
  memset(&h, 0, sizeof(h));
  double big = 1.e9;
  for (int d = 0; d < 2; d++)
    for (int k = 0; k < NSTRIPE * SSTR; k++) memcpy(&h.ndt_acc[d][k], &big, 8);
  h.ndt_last[0] = h.ndt_last[1] = big;
  h.tf = 1e30;
  h.disk_first = prm->disk_first ? 1 : 0;
  e = hipMemcpy(c->ds, &h, sizeof(h), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    seterr("harm_create: %s", hipGetErrorString(e));
    harm_destroy(c);
    return 1;
  }
  *out = c;
  return 0;
}

extern "C" int harm_set_analytic(harm_ctx* c, const double* panalytic, const double* omegaf) {
  if (!c) return seterr("null ctx");
  if (!panalytic) return seterr("null panalytic");
  const size_t bytes = (size_t)c->nz * 8 * sizeof(double);
  if (!c->pan) HIPCHK(hipMalloc(&c->pan, bytes));
  HIPCHK(hipMemcpy(c->pan, panalytic, bytes, hipMemcpyHostToDevice));
  if (omegaf) {
    if (!c->omf) HIPCHK(hipMalloc(&c->omf, bytes));
    HIPCHK(hipMemcpy(c->omf, omegaf, bytes, hipMemcpyHostToDevice));
  }
  return 0;
}

// captured launches hold the parameters, stream and variants of their capture
static void drop_graphs(harm_ctx* c) {
  if (c->graph_ok) (void)hipGraphExecDestroy(c->graph);
  if (c->graphK_ok) (void)hipGraphExecDestroy(c->graphK);
  c->graph_ok = c->graphK_ok = false;
}

// interp.c:559-620 rescale(), dir 1, VARTOINTERP PRIMTOINTERP: the factors of
// every zone from its r, in the reference's operations (glibc pow on the host)
extern "C" int harm_set_radius(harm_ctx* c, const double* r) {
  if (!c) return seterr("null ctx");
  if (!r) return seterr("null r");
  const int rt = c->prm.rescaletype;
  if (rt != 0 && rt != 1) return seterr("rescaletype must be 0 or 1");
  std::vector<double> h((size_t)c->nz * 8);
  for (size_t z = 0; z < (size_t)c->nz; z++) {
    const double rr = r[z];
    double* s = &h[z * 8];
    s[RHO] = (rt == 0) ? pow(rr, -1.5) : pow(rr, -2.7);
    s[UU] = s[RHO] / rr;
    s[U1] = s[RHO];
    s[U2] = 1.0;
    s[U3] = 1.0 / (rr * rr);
    s[B1] = (rt == 0) ? s[U3] : pow(rr, -2.4);
    s[B2] = s[B1];
    s[B3] = s[B1];
  }
  if (!c->rscale) HIPCHK(hipMalloc(&c->rscale, h.size() * sizeof(double)));
  HIPCHK(hipMemcpy(c->rscale, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  c->kp.rscale = c->rscale;
  drop_graphs(c);
  return 0;
}

extern "C" int harm_destroy(harm_ctx* c) {
  if (!c) return 0;
  drop_graphs(c);
  for (int a = 0; a < HARM_NARRAYS; a++)
    if (c->arr[a]) (void)hipFree(c->arr[a]);
  double* dbufs[4] = {c->q1, c->q2, c->emf, c->pfix};
  for (double* b : dbufs)
    if (b) (void)hipFree(b);
  if (c->pflag) (void)hipFree(c->pflag);
  if (c->bsqf) (void)hipFree(c->bsqf);
  if (c->enoA) (void)hipFree(c->enoA);
  if (c->enoB) (void)hipFree(c->enoB);
  if (c->gammacheck) (void)hipFree(c->gammacheck);
  if (c->csA) (void)hipFree(c->csA);
  if (c->csB) (void)hipFree(c->csB);
  if (c->faillist) (void)hipFree(c->faillist);
  if (c->gF) (void)hipFree(c->gF);
  if (c->gC) (void)hipFree(c->gC);
  if (c->gConn) (void)hipFree(c->gConn);
  if (c->ds) (void)hipFree(c->ds);
  if (c->pan) (void)hipFree(c->pan);
  if (c->omf) (void)hipFree(c->omf);
  if (c->rscale) (void)hipFree(c->rscale);
  if (c->evpool) {
    for (hipEvent_t e : *c->evpool) (void)hipEventDestroy(e);
    delete c->evpool;
  }
  delete c->evkid;
  acct_free(c);
  cur_free(c);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  free(c);
  return 0;
}

extern "C" int harm_set_stream(harm_ctx* c, void* s) {
  if (!c) return seterr("null ctx");
  c->stream = s ? (hipStream_t)s : c->own_stream;
  drop_graphs(c);
  return 0;
}

// the reference's set_grid() tables packed into the device records: face
// records hF [2][nz][GF], CENT records hC [nz][GCN], the connection
// component-major hT [NCONN][nz] (floors may be null: zero).  The records keep
// one triangle of gcov and of the connection's lower index pair, so those
// must be symmetric to the bit (set_grid()'s are); returns false if not.
static bool pack_geometry(size_t nz, const double* gcov, const double* gcon, const double* gdet, const double* conn,
                          const double* floors, std::vector<double>& hF, std::vector<double>& hC,
                          std::vector<double>& hT) {
  static const int ua[10] = {0, 0, 0, 0, 1, 1, 1, 2, 2, 3}, ub[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};
  auto same = [](double x, double y) { return memcmp(&x, &y, sizeof(double)) == 0; };
  hF.assign(2 * nz * GF, 0.0);
  hC.assign(nz * GCN, 0.0);
  hT.assign(nz * NCONN, 0.0);
  for (size_t z = 0; z < nz; z++) {
    for (int pos = 0; pos < 4; pos++)
      for (int a = 0; a < 4; a++)
        for (int b = a + 1; b < 4; b++) {
          const size_t m = (z * 4 + pos) * 16;
          if (!same(gcov[m + a * 4 + b], gcov[m + b * 4 + a])) return false;
        }
    for (int k = 0; k < 4; k++)
      for (int l = 0; l < 4; l++)
        for (int j = l + 1; j < 4; j++)
          if (!same(conn[z * 64 + (k * 4 + l) * 4 + j], conn[z * 64 + (k * 4 + j) * 4 + l])) return false;
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
    const double* gc = gcov + (z * 4 + 3) * 16;
    const double* gn = gcon + (z * 4 + 3) * 16;
    for (int u = 0; u < 10; u++) o[u] = gc[ua[u] * 4 + ub[u]];
    for (int m = 0; m < 16; m++) o[10 + m] = gn[m];
    o[GC_GDET] = gdet[z * 4 + 3];
    o[GC_GDET + 1] = floors ? floors[z * 2 + 0] : 0.0;
    o[GC_GDET + 2] = floors ? floors[z * 2 + 1] : 0.0;
    o[GC_GDETC] = gdet[z * 4 + 2];
  }
  // the connection component-major ([NCONN][zones]): the update's lanes
  // (adjacent zones) then read adjacent words of each component
  for (size_t z = 0; z < nz; z++)
    for (int k = 0; k < 4; k++)
      for (int u = 0; u < 10; u++) hT[(size_t)(k * 10 + u) * nz + z] = conn[z * 64 + (k * 4 + ua[u]) * 4 + ub[u]];
  return true;
}

extern "C" int harm_set_geometry(harm_ctx* c, const double* gcov, const double* gcon, const double* gdet,
                                 const double* conn, const double* floors) {
  if (!c || !gcov || !gcon || !gdet || !conn || !floors) return seterr("null argument");
  std::vector<double> hF, hC, hT;
  if (!pack_geometry((size_t)c->nz, gcov, gcon, gdet, conn, floors, hF, hC, hT))
    return seterr("geometry tables: gcov and the connection's lower index pair must be symmetric to the bit");
  HIPCHK(hipMemcpy(c->gF, hF.data(), hF.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->gC, hC.data(), hC.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->gConn, hT.data(), hT.size() * sizeof(double), hipMemcpyHostToDevice));
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

extern "C" int harm_rdump_write(harm_ctx* c, const char* path, int binary, const harm_restart_header* h) {
  if (!c || !path || !h) return seterr("null argument");
  std::vector<double> p((size_t)c->nz * 8);
  if (harm_get_array(c, HARM_P, p.data())) return 1;
  const int rc = harm_rdump_write_host(path, binary, h, c->kp.n1, c->kp.n2, p.data());
  if (rc == 2) return seterr("restart header grid %dx%d differs from the context's %dx%d", h->totalsize[0],
                             h->totalsize[1], c->kp.n1, c->kp.n2);
  if (rc) return seterr("harm_rdump_write %s: error %d", path, rc);
  return 0;
}

extern "C" int harm_rdump_read(harm_ctx* c, const char* path, int binary, harm_restart_header* h) {
  if (!c || !path || !h) return seterr("null argument");
  std::vector<double> p((size_t)c->nz * 8);
  if (harm_get_array(c, HARM_P, p.data())) return 1;
  const int rc = harm_rdump_read_host(path, binary, h, c->kp.n1, c->kp.n2, p.data());
  if (rc == 2) return seterr("restart file grid %dx%d differs from the context's %dx%d", h->totalsize[0],
                             h->totalsize[1], c->kp.n1, c->kp.n2);
  if (rc) return seterr("harm_rdump_read %s: error %d", path, rc);
  return harm_set_array(c, HARM_P, p.data());
}

extern "C" int harm_get_pflag(harm_ctx* c, int* host) {
  if (!c || !host) return seterr("null argument");
  int* tmp;
  HIPCHK(hipMalloc(&tmp, (size_t)c->nz * 5 * sizeof(int)));
  hipLaunchKernelGGL(k_pflag_expand, dim3((c->nz + 255) / 256), dim3(256), 0, c->stream, c->nz, c->pflag, c->bsqf,
                     tmp);
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

// ---- launch helpers ----

static inline dim3 grid2(int nj, int ni, dim3 b) { return dim3((nj + b.x - 1) / b.x, (ni + b.y - 1) / b.y); }

// Per-kernel timing: while a timed stage is open (prof_begin .. prof_end),
// each of its kernels is launched with its own start/stop event pair through
// hipExtLaunchKernelGGL, whose events take the kernel's own begin and end
// timestamps (the dispatch gap before a kernel is not counted).  A stage's
// time is the sum over its kernels, its count one per stage call.  Pairs are
// drained (synchronise + elapsed time) when the pool fills or when the
// totals are read.
static int prof_drain(harm_ctx* c) {
  if (!c->evused) return 0;
  HIPCHK(hipEventSynchronize((*c->evpool)[2 * (c->evused - 1) + 1]));
  for (int k = 0; k < c->evused; k++) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, (*c->evpool)[2 * k], (*c->evpool)[2 * k + 1]));
    c->ktime[(*c->evkid)[k]] += ms;
  }
  c->evused = 0;
  return 0;
}
static int prof_begin(harm_ctx* c, int kid) {
  if (!c->prof) return 0;
  c->pkid = kid;
  c->kcount[kid] += 1;
  return 0;
}
static int prof_end(harm_ctx* c) {
  c->pkid = -1;
  if (c->perr) {
    c->perr = 0;
    return 1;
  }
  return 0;
}
// the next event pair of the pool for a kernel of the open stage
static bool prof_pair(harm_ctx* c, hipEvent_t* e0, hipEvent_t* e1) {
  if (c->evused == 2048 && prof_drain(c)) return false;
  if ((int)c->evpool->size() < 2 * (c->evused + 1)) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return false;
    if (hipEventCreate(&b) != hipSuccess) {
      (void)hipEventDestroy(a);
      return false;
    }
    c->evpool->push_back(a);
    c->evpool->push_back(b);
    c->evkid->push_back(0);
  }
  (*c->evkid)[c->evused] = c->pkid;
  *e0 = (*c->evpool)[2 * c->evused];
  *e1 = (*c->evpool)[2 * c->evused + 1];
  c->evused++;
  return true;
}
// a launch on the context's stream, timed when a stage is open
template <typename F, typename... A>
static void tlaunch(harm_ctx* c, F kernel, dim3 g, dim3 b, unsigned shm, A... args) {
  hipEvent_t e0, e1;
  if (c->prof && c->pkid >= 0) {
    if (prof_pair(c, &e0, &e1)) {
      hipExtLaunchKernelGGL(kernel, g, b, shm, c->stream, e0, e1, 0, args...);
      return;
    }
    if (!c->perr) seterr("per-kernel timing: event pool");
    c->perr = 1;
  }
  hipLaunchKernelGGL(kernel, g, b, shm, c->stream, args...);
}

// fluxcalc for dir0..dir0+ndir-1 (both directions in one launch when ndir=2);
// fix: apply fix_flux()'s edge rules in the same launch when the context
// has them on (the step), or never (the standalone fluxcalc)
static int launch_flux(harm_ctx* c, int pr_sel, int dir0, int ndir, bool fix = true) {
  const int nfj = c->kp.n2 + 2, nfi = c->kp.n1 + 2;  // covers both directions' face ranges
  KParams kp = c->kp;
  if (!fix) kp.fixupflux = 0;
  if (prof_begin(c, HARM_K_FLUX)) return 1;
  switch (c->flux_variant) {
    case HARM_FLUX_PAIR: {
      const dim3 b(64, 4);
      tlaunch(c,k_flux, dim3((nfj + 31) / 32, (nfi + 3) / 4, ndir), b, 0, kp, c->arr[pr_sel],
                         c->arr[HARM_F1], c->arr[HARM_F2], c->q1, c->q2, c->gF, c->ds, dir0);
      break;
    }
    case HARM_FLUX_SHARED: {
      int gx, gy;
      flux_shared_grid(c->kp.n1, c->kp.n2, gx, gy);
      tlaunch(c,k_flux_shared, dim3(gx, gy, ndir), dim3(64, FSH_ROWS), 0, kp,
                         c->arr[pr_sel], c->arr[HARM_F1], c->arr[HARM_F2], c->q1, c->q2, c->gF, c->ds, dir0);
      break;
    }
    default: {
      const dim3 b = (c->flux_variant == HARM_FLUX_FACE_64x2) ? dim3(64, 2) : dim3(64, 4);
      tlaunch(c,flux_single_kernel(kp), dim3((nfj + 63) / 64, (nfi + b.y - 1) / b.y, ndir), b, 0, kp, c->arr[pr_sel], c->arr[HARM_F1], c->arr[HARM_F2], c->q1, c->q2, c->gF,
                         c->ds, dir0);
      break;
    }
  }
  HIPCHK(hipGetLastError());
  return prof_end(c);
}

// advance() after the fluxes.  DOENO=1 (advance_eno_u): U(pi) -> A, c2a
// along x1 A -> B, update (c2a along x2 of B; averaged Uf -> A), a2c along
// x1 A -> B, inversion (a2c along x2 of B).  DOENO=2: the two flux
// divergences -> A, B, then the update converts them.
static int launch_update(harm_ctx* c, int pi, int pb, int pf, const RKCoef& C, bool bnd = false) {
  const dim3 b(64, 4);
  const dim3 gz = grid2(c->kp.n2, c->kp.n1, b);
  // the reference sets pf = pi before it reads pb zone by zone (advance.c:121,
  // 367): with pf the same array as pb, the source term sees pi
  const int pbs = (pf == pb) ? pi : pb;
  if (prof_begin(c, HARM_K_UPDATE)) return 1;
  if (c->kp.eno == 1) {
    tlaunch(c,k_eno_ucenter, gz, b, 0, c->kp, c->arr[pi], c->gC, c->enoA);
    tlaunch(c,k_eno_cvt1<ENO_C2A>, gz, b, 0, c->kp, c->enoA, c->enoB);
  } else if (c->kp.eno == 2) {
    tlaunch(c,k_eno_fluxdiv, gz, b, 0, c->kp, c->arr[HARM_F1], c->arr[HARM_F2], c->q1, c->q2,
                       c->gC, c->enoA, c->enoB);
  }
  auto* kup = (c->kp.eno == 1)                        ? k_update<1>
              : (c->kp.eno == 2)                      ? k_update<2>
              : (c->kp.eomtype == HARM_EOMGRMHD)      ? (bnd ? k_update<0, HARM_EOMGRMHD, 1>
                                                             : k_update<0, HARM_EOMGRMHD>)
              : (c->kp.eomtype == HARM_EOMFFDE)       ? (bnd ? k_update<0, HARM_EOMFFDE, 1>
                                                             : k_update<0, HARM_EOMFFDE>)
                                                      : (bnd ? k_update<0, -1, 1> : k_update<0>);
  tlaunch(c,kup, gz, b, 0, c->kp, C, c->arr[pi], c->arr[pbs], c->arr[pf], c->arr[HARM_ULAST],
                     c->arr[HARM_UNEW], c->arr[HARM_F1], c->arr[HARM_F2], c->q1, c->q2, c->gC, c->gConn, c->pflag,
                     c->faillist, c->ds, c->enoA, c->enoB);
  if (c->kp.eno == 1) {
    tlaunch(c,k_eno_cvt1<ENO_A2C>, gz, b, 0, c->kp, c->enoA, c->enoB);
    tlaunch(c,k_eno_invert, gz, b, 0, c->kp, c->arr[pi], c->arr[pf], c->enoB, c->gC, c->pflag,
                       c->faillist, c->ds);
  }
  HIPCHK(hipGetLastError());
  return prof_end(c);
}

static void launch_bsqflags(harm_ctx* c, int sel) {
  const dim3 b(64, 4);
  hipLaunchKernelGGL(k_bsqflags, grid2(c->kp.n2, c->kp.n1, b), b, 0, c->stream, c->kp, c->arr[sel], c->gC, c->bsqf);
}

// the ghost fill needs the general rules (outflow extrapolation, a FIXED
// side or bounds.c's disk): then every schedule variant runs them
static bool general_bounds(const harm_ctx* c) {
  bool fixed = false;
  for (int b = 0; b < 4; b++) fixed = fixed || (c->prm.bc[b] == HARM_FIXED);
  return fixed || c->prm.extrap != 0 || c->prm.boundsfile != HARM_BOUNDS_FISHMON;
}

static void launch_general_fill(harm_ctx* c, int sel) {
  const int ni = c->kp.n1 + 2 * NG, nj = c->kp.n2 + 2 * NG;
  tlaunch(c, k_bg_x1, dim3((2 * NG * nj + 255) / 256), dim3(256), 0, c->kp, c->arr[sel], c->gC, c->pan, c->ds);
  tlaunch(c, k_bg_x2, dim3((2 * NG * ni + 255) / 256), dim3(256), 0, c->kp, c->arr[sel], c->pan);
  if (c->kp.poledeath) tlaunch(c, k_bg_pole, dim3((2 * (NG + 1) * ni + 255) / 256), dim3(256), 0, c->kp, c->arr[sel]);
  if (c->prm.boundsfile == HARM_BOUNDS_DISK) {
    tlaunch(c, k_bg_disk, dim3((DISK_DELTAJ * ni + 255) / 256), dim3(256), 0, c->kp, c->arr[sel], c->pan, c->omf,
            c->gC, c->pflag, c->ds);
    tlaunch(c, k_bg_diskdone, dim3(1), dim3(1), 0, c->ds);
  }
}

static int launch_boundary(harm_ctx* c, int sel, int mode) {
  if (c->prm.eomtype != HARM_EOMGRMHD) mode &= ~KB_FIXUP;
  if (general_bounds(c)) {
    bool needs_analytic = c->prm.boundsfile == HARM_BOUNDS_DISK;
    if (needs_analytic && (!c->pan || !c->omf))
      return seterr("bound_prim: the analytic solution is not set (harm_set_analytic)");
    if (c->prm.extrap == 2 && !c->rscale) return seterr("bound_prim: EXTRAP 2 needs the zone radii (harm_set_radius)");
    if (c->kp.checksolution && (mode & KB_FIXUP))
      return seterr("checksolution with extrapolated, FIXED or disk boundaries is not built");
    if (prof_begin(c, HARM_K_BOUNDARY)) return 1;
    if (mode & KB_BOUND) launch_general_fill(c, sel);
    if (mode & KB_FIXUP)
      tlaunch(c, k_fixall, dim3(1), dim3(1024), 0, c->kp, c->arr[sel], c->pflag, nullptr, c->faillist, c->pfix,
              c->gC, c->ds, 0);
    if (mode & KB_BOUND2) launch_general_fill(c, sel);
    if (mode & KB_DT) tlaunch(c, k_tail, dim3(1), dim3(1), 0, c->kp, c->ds, 1);
    HIPCHK(hipGetLastError());
    return prof_end(c);
  }
  if (prof_begin(c, HARM_K_BOUNDARY)) return 1;
  if (mode & KB_INUPDATE) {  // the first bound_prim() was k_update's
    const int do_dt = (mode & KB_DT) ? 1 : 0;
    if (mode & KB_FIXUP) {
      tlaunch(c, k_fixbound<true>, dim3(1), dim3(256), 0, c->kp, c->arr[sel], c->pflag, c->faillist, c->pfix,
              c->gC, c->ds, do_dt);
    } else if ((mode & KB_BOUND2) || do_dt) {
      tlaunch(c, k_fixbound<false>, dim3(1), dim3(256), 0, c->kp, c->arr[sel], c->pflag, c->faillist, c->pfix,
              c->gC, c->ds, do_dt);
    }
  } else if (c->bound_variant == HARM_BOUND_BLOCK && !c->kp.checksolution) {
    tlaunch(c,k_boundary, dim3(1), dim3(1024), 0, c->kp, c->arr[sel], c->gC, c->pflag,
                       c->faillist, c->pfix, c->ds, mode);
  } else {
    const dim3 gx1((c->kp.n2 + 63) / 64, 2 * NG), gx2((c->kp.n1 + 2 * NG + 63) / 64, 2 * NG + 2);
    const dim3 gb((bound_threads(c->kp) + 255) / 256);
    const bool bound2 = (mode & KB_BOUND2) != 0, fix = (mode & KB_FIXUP) != 0;
    // the in-update fill's fallback where it does not apply is the fused one
    const bool fused = (c->bound_variant == HARM_BOUND_FUSED) || (c->bound_variant == HARM_BOUND_INUPDATE);
    if ((mode & KB_BOUND) && fused) {
      tlaunch(c,k_bound, gb, dim3(256), 0, c->kp, c->arr[sel], c->gC, c->ds, 0,
                         (!fix && !bound2 && (mode & KB_DT)) ? 1 : 0);
    } else if (mode & KB_BOUND) {
      tlaunch(c,k_bx1, gx1, dim3(64), 0, c->kp, c->arr[sel], c->gC, c->ds, 0);
      tlaunch(c,k_bx2, gx2, dim3(64), 0, c->kp, c->arr[sel], c->ds,
                         (!fix && !bound2 && (mode & KB_DT)) ? 1 : 0);
    }
    if (fix && c->kp.checksolution) {  // fixup_checksolution (fixup.c:388-488)
      const dim3 b(64, 4);
      const dim3 gi = grid2(c->kp.n2, c->kp.n1, b);
      tlaunch(c,k_gammacheck, grid2(c->kp.n2 + 2, c->kp.n1 + 2, b), b, 0, c->kp, c->arr[sel],
                         c->gC, c->pflag, c->bsqf, c->gammacheck);
      tlaunch(c,k_cs_eval, gi, b, 0, c->kp, c->arr[sel], c->pflag, c->bsqf, c->gammacheck,
                         c->csB, c->csA, c->ds, 1);
      tlaunch(c,k_cs_eval, gi, b, 0, c->kp, c->arr[sel], c->pflag, c->bsqf, c->gammacheck,
                         c->csA, c->csB, c->ds, 0);
      tlaunch(c,k_cs_sequential, dim3(1), dim3(1), 0, c->kp, c->arr[sel], c->pflag, c->bsqf,
                         c->gammacheck, c->csB, c->ds, c->bound_variant == HARM_BOUND_KERNELS_SEQCHECK ? 1 : 0);
      tlaunch(c,k_cs_apply, gi, b, 0, c->kp, c->csB, c->pflag, c->faillist, c->ds);
    }
    const bool fixone = (c->bound_variant == HARM_BOUND_FIXONE) || fused;
    if (fix && fixone) {  // list repaired and cleared by one workgroup
      tlaunch(c,k_fixall, dim3(1), dim3(1024), 0, c->kp, c->arr[sel], c->pflag,
                         c->kp.checksolution ? c->csB : nullptr, c->faillist, c->pfix, c->gC, c->ds,
                         (!bound2 && (mode & KB_DT)) ? 1 : 0);
    } else if (fix) {
      tlaunch(c,k_fixlist, dim3(32), dim3(64), 0, c->kp, c->arr[sel], c->pflag,
                         c->kp.checksolution ? c->csB : nullptr, c->faillist, c->pfix, c->gC, c->ds);
      tlaunch(c,k_fixapply, dim3(32), dim3(64), 0, c->kp, c->arr[sel], c->pflag, c->faillist,
                         c->pfix, c->ds);
    }
    if (bound2 && fused) {
      tlaunch(c,k_bound, gb, dim3(256), 0, c->kp, c->arr[sel], c->gC, c->ds, 0,
                         (mode & KB_DT) ? 1 : 0);
    } else if (bound2) {
      tlaunch(c,k_bx1, gx1, dim3(64), 0, c->kp, c->arr[sel], c->gC, c->ds,
                         (fix && !fixone) ? 1 : 0);
      tlaunch(c,k_bx2, gx2, dim3(64), 0, c->kp, c->arr[sel], c->ds, (mode & KB_DT) ? 1 : 0);
    } else if (fix && !fixone) {
      tlaunch(c,k_tail, dim3(1), dim3(1), 0, c->kp, c->ds, (mode & KB_DT) ? 1 : 0);
    }
  }
  HIPCHK(hipGetLastError());
  return prof_end(c);
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

// advance(): fluxcalc both directions (+ fix_flux, flux_ct formed in k_update), update
static int substage(harm_ctx* c, int pi, int pb, int pf, const double* CUf, const double* Cunew, int to, int nto,
                    bool bnd = false) {
  RKCoef C;
  C.cuf0 = CUf[0]; C.cuf1 = CUf[1]; C.cuf2 = CUf[2];
  C.cn0 = Cunew[0]; C.cn1 = Cunew[1]; C.cn2 = Cunew[2];
  C.timeorder = to;
  C.last = (to == nto - 1);
  if (launch_flux(c, pb, 1, 2)) return 1;
  if (launch_update(c, pi, pb, pf, C, bnd)) return 1;
  return 0;
}

// HARM_BOUND_INUPDATE where it applies: the plain update (no ENO), no
// fixup_checksolution between the fills, no accounting (which reads the
// updated zones before bound_prim), boundary kinds the images cover
static bool bound_in_update(const harm_ctx* c) {
  return c->bound_variant == HARM_BOUND_INUPDATE && c->kp.eno == 0 && !c->kp.checksolution && !acct_on(c) &&
         bound_images_ok(c->kp);
}

static int enqueue_step(harm_ctx* c) {
  double CUf[4][3], Cunew[4][3];
  int pii[4], pbb[4], pff[4], nto;
  rk_setup(c->prm.timeorder, &nto, CUf, Cunew, pii, pbb, pff);
  for (int to = 0; to < nto; to++) {
    // pre_fixup (step_ch.c:92) flags every substage from pii = p, which no
    // substage changes before its own update: once per step is the same
    if (to == 0 && c->kp.checksolution) launch_bsqflags(c, pii[to]);
    const bool acct = acct_on(c);
    // advance.c:138-140: diag_source on the last substage
    if (to == nto - 1 && acct && acct_source(c, pii[to], pbb[to], pff[to])) return 1;
    if (acct && acct_faces(c, pbb[to])) return 1;
    const bool bnd = bound_in_update(c);
    if (substage(c, pii[to], pbb[to], pff[to], CUf[to], Cunew[to], to, nto, bnd)) return 1;
    if (acct && acct_zones(c, to == nto - 1)) return 1;
    // step_ch.c:104-124 bound_prim, post_fixup, bound_prim; then :174-183 after the last substage
    int mode = KB_BOUND | KB_FIXUP | KB_BOUND2 | ((to == nto - 1) ? KB_DT : 0) | (bnd ? KB_INUPDATE : 0);
    if (launch_boundary(c, pff[to], mode)) return 1;
    if (acct && acct_merge(c, to == nto - 1)) return 1;
  }
  // step_ch.c:143-153 the current, :160 diag_flux, with the step's dt
  if (cur_on(c) && cur_step(c)) return 1;
  if (acct_on(c) && acct_flux(c, -1.0)) return 1;
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
  // the fluxes as fluxcalc() leaves them: fix_flux() is a call of its own
  if (launch_flux(c, pr_sel, dir, 1, false)) return 1;
  // the striped running minimum: combined here, then every slot reset
  std::vector<unsigned long long> slots(NSTRIPE * SSTR);
  const size_t sbytes = slots.size() * sizeof(unsigned long long);
  HIPCHK(hipMemcpyAsync(slots.data(), &c->ds->ndt_acc[dir - 1][0], sbytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  unsigned long long big;
  double b = 1.e9;
  memcpy(&big, &b, 8);
  unsigned long long bits = big;
  for (int k = 0; k < NSTRIPE; k++) bits = std::min(bits, slots[(size_t)k * SSTR]);
  std::fill(slots.begin(), slots.end(), big);
  HIPCHK(hipMemcpyAsync(&c->ds->ndt_acc[dir - 1][0], slots.data(), sbytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (ndt) memcpy(ndt, &bits, 8);
  return 0;
}

extern "C" int harm_fix_flux(harm_ctx* c) {
  if (need_geom(c)) return 1;
  const dim3 b(64, 4);
  hipLaunchKernelGGL(k_fixflux, grid2(c->kp.n2 + 2, c->kp.n1 + 2, b), b, 0, c->stream, c->kp, c->arr[HARM_F1],
                     c->arr[HARM_F2], c->q1, c->q2, c->gF);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int harm_flux_ct(harm_ctx* c) {
  if (need_geom(c)) return 1;
  // FLUXCTHLL: the reference's flux_ct() leaves the fluxes as they are
  if (c->prm.fluxb == HARM_FLUXCTHLL) return 0;
  const dim3 b(64, 4);
  if (prof_begin(c, HARM_K_EMF)) return 1;
  // the EMFs from the fluxes the context holds (edits through harm_set_array
  // included), as step_ch.c:1223-1234 reads F1, F2
  tlaunch(c,k_qfromF, grid2(c->kp.n2 + 2, c->kp.n1 + 2, b), b, 0, c->kp, c->arr[HARM_F1],
                     c->arr[HARM_F2], c->q1, c->q2, c->gF);
  tlaunch(c,k_emf, grid2(c->kp.n2 + 1, c->kp.n1 + 1, b), b, 0, c->kp, c->q1, c->q2, c->gC,
                     c->emf);
  HIPCHK(hipGetLastError());
  if (prof_end(c)) return 1;
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
  if (launch_boundary(c, sel, KB_BOUND)) return 1;
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int harm_pre_fixup(harm_ctx* c, int sel) {
  if (need_geom(c)) return 1;
  if (sel < 0 || sel > HARM_PK1) return seterr("sel must be HARM_P/PK0/PK1");
  launch_bsqflags(c, sel);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int harm_post_fixup(harm_ctx* c, int sel) {
  if (need_geom(c)) return 1;
  if (sel < 0 || sel > HARM_PK1) return seterr("sel must be HARM_P/PK0/PK1");
  // a repair outside a step is not accounted (no merge follows it)
  const bool acct = acct_on(c);
  if (acct) acct_set_on(c, false);
  const int rc = launch_boundary(c, sel, KB_FIXUP);
  if (acct) acct_set_on(c, true);
  if (rc) return 1;
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

static int autotune(harm_ctx* c);

extern "C" int harm_step(harm_ctx* c) {
  if (need_geom(c)) return 1;
  if (autotune(c)) return 1;
  if (enqueue_step(c)) return 1;
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

static int capture_step(harm_ctx* c, hipGraphExec_t* ex, int nsteps = 1) {
  hipGraph_t g;
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int s = 0; s < nsteps && !rc; s++) rc = enqueue_step(c);
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) return 1;
  if (e != hipSuccess) return seterr("graph capture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return seterr("graph instantiate: %s", hipGetErrorString(e));
  return 0;
}

static int build_graph(harm_ctx* c) {
  if (c->graph_ok) return 0;
  if (capture_step(c, &c->graph)) return 1;
  c->graph_ok = true;
  return 0;
}

// harm_step_n replays STEPS_PER_GRAPH steps per graph launch: a launch of a
// whole graph costs ~8 us before its first kernel, a kernel inside one none
constexpr int STEPS_PER_GRAPH = 8;
static int build_graphK(harm_ctx* c) {
  if (c->graphK_ok) return 0;
  if (capture_step(c, &c->graphK, STEPS_PER_GRAPH)) return 1;
  c->graphK_ok = true;
  return 0;
}

// Autotuner.  The kernel variants give bit-identical results; each
// combination is captured as a one-step graph and timed over a few replays
// from the live state, which is restored before every replay pass and
// afterwards: P, the device scalars and, with checksolution, the persistent
// gamma table (written only at flagged zones, read at their neighbours).
// Every other array is overwritten within a step before it is read;
// accounting and the current are paused while the trials run.
static int time_variant(harm_ctx* c, const double* keep, const double* keep_gc, const DevScalars* keep_ds, float* ms) {
  const size_t pbytes = (size_t)c->nz * 8 * sizeof(double);
  hipGraphExec_t ex;
  if (capture_step(c, &ex)) return 1;
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  hipError_t e = hipSuccess;
  for (int pass = 0; pass < 2 && e == hipSuccess; pass++) {
    e = hipMemcpyAsync(c->arr[HARM_P], keep, pbytes, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->ds, keep_ds, sizeof(DevScalars), hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess && keep_gc)
      e = hipMemcpyAsync(c->gammacheck, keep_gc, (size_t)c->nz * sizeof(double), hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess && pass == 1) e = hipEventRecord(a, c->stream);
    for (int r = 0; r < (pass ? 4 : 1) && e == hipSuccess; r++) e = hipGraphLaunch(ex, c->stream);
    if (e == hipSuccess && pass == 1) e = hipEventRecord(b, c->stream);
  }
  if (e == hipSuccess) e = hipEventSynchronize(b);
  if (e == hipSuccess) e = hipEventElapsedTime(ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipGraphExecDestroy(ex);
  if (e != hipSuccess) return seterr("autotune: %s", hipGetErrorString(e));
  return 0;
}

static int autotune(harm_ctx* c) {
  if (c->tuned) return 0;
  const size_t pbytes = (size_t)c->nz * 8 * sizeof(double);
  double* keep = nullptr;
  double* keep_gc = nullptr;
  DevScalars* keep_ds = nullptr;  // the device scalars, kept on the device
  HIPCHK(hipMalloc(&keep, pbytes));
  int rc = 0;
  if (hipMalloc(&keep_ds, sizeof(DevScalars)) != hipSuccess ||
      hipMemcpyAsync(keep_ds, c->ds, sizeof(DevScalars), hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(keep, c->arr[HARM_P], pbytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
    rc = seterr("autotune: copy");
  const size_t gbytes = (size_t)c->nz * sizeof(double);
  if (!rc && c->kp.checksolution &&
      (hipMalloc(&keep_gc, gbytes) != hipSuccess ||
       hipMemcpyAsync(keep_gc, c->gammacheck, gbytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess))
    rc = seterr("autotune: gamma table copy");
  const bool prof = c->prof;
  c->prof = false;
  // the timed replays must not accumulate nor advance the current's
  // history: accounting and the current are off meanwhile
  const bool acct = acct_on(c);
  if (acct) acct_set_on(c, false);
  const bool cur = cur_on(c);
  if (cur) cur_set_on(c, false);
  float best = 1e30f;
  int bf = HARM_FLUX_FACE_64x4, bb = HARM_BOUND_KERNELS;
  const int fv[4] = {HARM_FLUX_FACE_64x4, HARM_FLUX_FACE_64x2, HARM_FLUX_PAIR, HARM_FLUX_SHARED};
  // the one-workgroup pass has no fixup_checksolution and the fill inside
  // the update does not apply with it: both are tried last
  const int bv[5] = {HARM_BOUND_KERNELS, HARM_BOUND_FIXONE, HARM_BOUND_FUSED, HARM_BOUND_INUPDATE, HARM_BOUND_BLOCK};
  const int nbv = c->kp.checksolution ? 3 : 5;
  for (int f = 0; f < 4 && !rc; f++)
    for (int v = 0; v < nbv && !rc; v++) {
      float ms = 0.f;
      c->flux_variant = fv[f];
      c->bound_variant = bv[v];
      rc = time_variant(c, keep, keep_gc, keep_ds, &ms);
      if (!rc && ms < best) { best = ms; bf = fv[f]; bb = bv[v]; }
    }
  c->prof = prof;
  if (acct) acct_set_on(c, true);
  if (cur) cur_set_on(c, true);
  c->flux_variant = bf;
  c->bound_variant = bb;
  hipError_t e = hipMemcpyAsync(c->arr[HARM_P], keep, pbytes, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess && keep_ds)
    e = hipMemcpyAsync(c->ds, keep_ds, sizeof(DevScalars), hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess && keep_gc) e = hipMemcpyAsync(c->gammacheck, keep_gc, gbytes, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(keep);
  if (keep_ds) (void)hipFree(keep_ds);
  if (keep_gc) (void)hipFree(keep_gc);
  if (rc) return 1;
  if (e != hipSuccess) return seterr("autotune: %s", hipGetErrorString(e));
  c->tuned = true;
  return 0;
}

extern "C" int harm_step_n(harm_ctx* c, int nsteps) {
  if (need_geom(c)) return 1;
  if (nsteps < 0) return seterr("nsteps < 0");
  if (autotune(c)) return 1;
  if (c->prof) {  // timed per kernel: eager launches
    for (int s = 0; s < nsteps; s++)
      if (enqueue_step(c)) return 1;
  } else {
    if (build_graph(c)) return 1;
    int s = 0;
    if (nsteps >= STEPS_PER_GRAPH) {
      if (build_graphK(c)) return 1;
      for (; s + STEPS_PER_GRAPH <= nsteps; s += STEPS_PER_GRAPH) HIPCHK(hipGraphLaunch(c->graphK, c->stream));
    }
    for (; s < nsteps; s++) HIPCHK(hipGraphLaunch(c->graph, c->stream));
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

// harm_utoprimgen with its staging in the device's static array: no
// allocation, at most HARM_UTOPRIM_NOALLOC_MAX zones.  Every staged value a
// call reads it has written first, so nothing needs clearing; calls sharing a
// device are serialised by that device's lock.
extern "C" int harm_utoprimgen_noalloc(harm_ctx* c, int n, const int* zone, const double* U, double* pr,
                                       int* pflag) {
  if (need_geom(c)) return 1;
  if (n <= 0) return 0;
  if (!zone || !U || !pr || !pflag) return seterr("null argument");
  if (n > HARM_UTOPRIM_NOALLOC_MAX)
    return seterr("harm_utoprimgen_noalloc: %d zones exceed the staging capacity of %d (use harm_utoprimgen)", n,
                  HARM_UTOPRIM_NOALLOC_MAX);
  if (!c->scratch) return seterr("harm_utoprimgen_noalloc: no staging memory on device %d", c->device);
  for (int t = 0; t < n; t++) {
    int i = zone[2 * t], j = zone[2 * t + 1];
    if (i < -NG || i >= c->prm.n1 + NG || j < -NG || j >= c->prm.n2 + NG)
      return seterr("zone %d (%d,%d) out of range", t, i, j);
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu[c->device]);
  int* dz = (int*)(c->scratch + NOALLOC_ZONE_OFF);
  double* dU = (double*)(c->scratch + NOALLOC_U_OFF);
  double* dp = (double*)(c->scratch + NOALLOC_PR_OFF);
  int* dfl = (int*)(c->scratch + NOALLOC_FL_OFF);
  hipError_t e = hipMemcpyAsync(dz, zone, (size_t)n * 2 * sizeof(int), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dU, U, (size_t)n * 8 * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dp, pr, (size_t)n * 8 * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_utoprim_batch, dim3((n + 127) / 128), dim3(128), 0, c->stream, c->kp, n, dz, dU, dp, dfl,
                       c->gC);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(pr, dp, (size_t)n * 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(pflag, dfl, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return seterr("harm_utoprimgen_noalloc: %s", hipGetErrorString(e));
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
  std::vector<unsigned long long> v(NSTRIPE * SSTR);
  HIPCHK(hipMemcpyAsync(v.data(), &c->ds->nstroke[0], v.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  unsigned long long sum = 0;
  for (int k = 0; k < NSTRIPE; k++) sum += v[(size_t)k * SSTR];
  *nstroke = (long long)sum;
  return 0;
}

extern "C" int harm_get_newton_spread(harm_ctx* c, long long* nstroke, long long* wave_max_sum) {
  if (!c || !nstroke || !wave_max_sum) return seterr("null argument");
  std::vector<unsigned long long> v(NSTRIPE * SSTR);
  HIPCHK(hipMemcpyAsync(v.data(), &c->ds->nstroke[0], v.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  unsigned long long a = 0, b = 0;
  for (int k = 0; k < NSTRIPE; k++) {
    a += v[(size_t)k * SSTR];
    b += v[(size_t)k * SSTR + 8];
  }
  *nstroke = (long long)a;
  *wave_max_sum = (long long)b;
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
  if (prof_drain(c)) return 1;
  if (total_ms) *total_ms = c->ktime[kernel];
  if (count) *count = c->kcount[kernel];
  return 0;
}

extern "C" int harm_prof_reset(harm_ctx* c) {
  if (!c) return seterr("null ctx");
  if (prof_drain(c)) return 1;
  for (int k = 0; k < HARM_K_COUNT; k++) {
    c->ktime[k] = 0;
    c->kcount[k] = 0;
  }
  return 0;
}

extern "C" int harm_set_tuning(harm_ctx* c, int flux_variant, int bound_variant) {
  if (!c) return seterr("null ctx");
  if (flux_variant < 0 || flux_variant > HARM_FLUX_SHARED) return seterr("bad flux variant %d", flux_variant);
  if (bound_variant < 0 || bound_variant > HARM_BOUND_INUPDATE)
    return seterr("bad boundary variant %d", bound_variant);
  c->flux_variant = flux_variant;
  c->bound_variant = bound_variant;
  c->tuned = true;
  drop_graphs(c);
  return 0;
}

extern "C" int harm_get_tuning(harm_ctx* c, int* flux_variant, int* bound_variant) {
  if (!c) return seterr("null ctx");
  if (flux_variant) *flux_variant = c->flux_variant;
  if (bound_variant) *bound_variant = c->bound_variant;
  return 0;
}

// ===========================================================================
// Batched runs (C-ABI): contexts advanced together, one launch per stage
// ===========================================================================
enum { BT_FLUX4 = 0, BT_FLUX2, BT_UPDATE, BT_BX1, BT_BX2, BT_N };

struct harm_batch {
  int n;
  std::vector<harm_ctx*>* ctx;
  RunDesc* d_runs;
  int4* d_tab[BT_N];
  int ntab[BT_N];
  hipStream_t stream;
  hipGraphExec_t graph;
  bool graph_ok;
  bool tuned;
  int flux_variant;   // HARM_FLUX_FACE_64x4 or HARM_FLUX_FACE_64x2
  int bound_variant;  // HARM_BOUND_*
  int timeorder;
};

static void add_blocks(std::vector<int4>& v, int run, dim3 g) {
  for (unsigned z = 0; z < g.z; z++)
    for (unsigned y = 0; y < g.y; y++)
      for (unsigned x = 0; x < g.x; x++) v.push_back(make_int4(run, (int)x, (int)y, (int)z));
}

static int batch_flux(harm_batch* b, int pr_sel) {
  const int t = (b->flux_variant == HARM_FLUX_FACE_64x2) ? BT_FLUX2 : BT_FLUX4;
  const dim3 blk = (t == BT_FLUX2) ? dim3(64, 2) : dim3(64, 4);
  hipLaunchKernelGGL(kb_flux, dim3(b->ntab[t]), blk, 0, b->stream, b->d_runs, b->d_tab[t], pr_sel, 1);
  HIPCHK(hipGetLastError());
  return 0;
}

static int batch_boundary(harm_batch* b, int sel, int mode) {
  if (b->bound_variant == HARM_BOUND_BLOCK) {
    hipLaunchKernelGGL(kb_boundary, dim3(b->n), dim3(1024), 0, b->stream, b->d_runs, sel, mode);
  } else {
    const bool fix = (mode & KB_FIXUP) != 0, bound2 = (mode & KB_BOUND2) != 0;
    if (mode & KB_BOUND) {
      hipLaunchKernelGGL(kb_bx1, dim3(b->ntab[BT_BX1]), dim3(64), 0, b->stream, b->d_runs, b->d_tab[BT_BX1], sel, 0);
      hipLaunchKernelGGL(kb_bx2, dim3(b->ntab[BT_BX2]), dim3(64), 0, b->stream, b->d_runs, b->d_tab[BT_BX2], sel,
                         (!fix && !bound2 && (mode & KB_DT)) ? 1 : 0);
    }
    if (fix) {
      hipLaunchKernelGGL(kb_fixlist, dim3(b->n * FIXBLK), dim3(64), 0, b->stream, b->d_runs, sel);
      hipLaunchKernelGGL(kb_fixapply, dim3(b->n * FIXBLK), dim3(64), 0, b->stream, b->d_runs, sel);
    }
    if (!bound2) return seterr("batch boundary pass without the second bound_prim");
    hipLaunchKernelGGL(kb_bx1, dim3(b->ntab[BT_BX1]), dim3(64), 0, b->stream, b->d_runs, b->d_tab[BT_BX1], sel,
                       fix ? 1 : 0);
    hipLaunchKernelGGL(kb_bx2, dim3(b->ntab[BT_BX2]), dim3(64), 0, b->stream, b->d_runs, b->d_tab[BT_BX2], sel,
                       (mode & KB_DT) ? 1 : 0);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// one step_ch() of every run (each with its own t, dt and Courant limit)
static int batch_enqueue_step(harm_batch* b) {
  double CUf[4][3], Cunew[4][3];
  int pii[4], pbb[4], pff[4], nto;
  rk_setup(b->timeorder, &nto, CUf, Cunew, pii, pbb, pff);
  for (int to = 0; to < nto; to++) {
    RKCoef C;
    C.cuf0 = CUf[to][0]; C.cuf1 = CUf[to][1]; C.cuf2 = CUf[to][2];
    C.cn0 = Cunew[to][0]; C.cn1 = Cunew[to][1]; C.cn2 = Cunew[to][2];
    C.timeorder = to;
    C.last = (to == nto - 1);
    if (batch_flux(b, pbb[to])) return 1;
    hipLaunchKernelGGL(kb_update, dim3(b->ntab[BT_UPDATE]), dim3(64, 4), 0, b->stream, b->d_runs,
                       b->d_tab[BT_UPDATE], C, pii[to], pbb[to], pff[to]);
    HIPCHK(hipGetLastError());
    if (batch_boundary(b, pff[to], KB_BOUND | KB_FIXUP | KB_BOUND2 | ((to == nto - 1) ? KB_DT : 0))) return 1;
  }
  return 0;
}

static int batch_capture(harm_batch* b, hipGraphExec_t* ex) {
  hipGraph_t g;
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = batch_enqueue_step(b);
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) return 1;
  if (e != hipSuccess) return seterr("batch graph capture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return seterr("batch graph instantiate: %s", hipGetErrorString(e));
  return 0;
}

extern "C" int harm_batch_destroy(harm_batch* b) {
  if (!b) return 0;
  if (b->graph_ok) (void)hipGraphExecDestroy(b->graph);
  for (int t = 0; t < BT_N; t++) (void)hipFree(b->d_tab[t]);
  (void)hipFree(b->d_runs);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b->ctx;
  delete b;
  return 0;
}

extern "C" int harm_batch_create(harm_ctx* const* ctxs, int n, harm_batch** out) {
  if (!out) return seterr("null output pointer");
  *out = nullptr;
  if (!ctxs || n < 1) return seterr("batch needs at least one context");
  for (int r = 0; r < n; r++) {
    if (!ctxs[r]) return seterr("context %d is null", r);
    if (!ctxs[r]->geometry_set) return seterr("context %d: geometry not set", r);
    if (ctxs[r]->prm.timeorder != ctxs[0]->prm.timeorder)
      return seterr("context %d: every run of a batch must use the same TIMEORDER", r);
    if (ctxs[r]->prm.checksolution) return seterr("context %d: CHECKSOLUTION is not batched", r);
    if (ctxs[r]->prm.eno) return seterr("context %d: DOENO is not batched", r);
    if (general_bounds(ctxs[r]))
      return seterr("context %d: extrapolated, FIXED or disk boundaries are not batched", r);
  }
  HIPCHK(hipDeviceSynchronize());
  harm_batch* b = new harm_batch();
  memset(b, 0, sizeof(*b));
  b->n = n;
  b->ctx = new std::vector<harm_ctx*>(ctxs, ctxs + n);
  b->timeorder = ctxs[0]->prm.timeorder;
  b->flux_variant = HARM_FLUX_FACE_64x4;
  b->bound_variant = HARM_BOUND_KERNELS;
  std::vector<RunDesc> hr(n);
  std::vector<int4> tab[BT_N];
  for (int r = 0; r < n; r++) {
    harm_ctx* c = ctxs[r];
    RunDesc& d = hr[r];
    d.P = c->kp;
    d.P.acct_pre = nullptr;  // a batch does no accounting
    d.P.acct_uin = nullptr;
    d.P.acct_flag = nullptr;
    d.P.acct_val = nullptr;
    for (int a = 0; a < HARM_NARRAYS; a++) d.arr[a] = c->arr[a];
    d.q1 = c->q1; d.q2 = c->q2; d.pfix = c->pfix;
    d.gF = c->gF; d.gC = c->gC; d.gConn = c->gConn;
    d.pflag = c->pflag; d.faillist = c->faillist; d.ds = c->ds;
    const int n1 = c->kp.n1, n2 = c->kp.n2, nfj = n2 + 2, nfi = n1 + 2;
    add_blocks(tab[BT_FLUX4], r, dim3((nfj + 63) / 64, (nfi + 3) / 4, 2));
    add_blocks(tab[BT_FLUX2], r, dim3((nfj + 63) / 64, (nfi + 1) / 2, 2));
    add_blocks(tab[BT_UPDATE], r, grid2(n2, n1, dim3(64, 4)));
    add_blocks(tab[BT_BX1], r, dim3((n2 + 63) / 64, 2 * NG));
    add_blocks(tab[BT_BX2], r, dim3((n1 + 2 * NG + 63) / 64, 2 * NG + 2));
  }
  hipError_t e = hipMalloc(&b->d_runs, sizeof(RunDesc) * n);
  if (e == hipSuccess) e = hipMemcpy(b->d_runs, hr.data(), sizeof(RunDesc) * n, hipMemcpyHostToDevice);
  for (int t = 0; t < BT_N && e == hipSuccess; t++) {
    b->ntab[t] = (int)tab[t].size();
    e = hipMalloc(&b->d_tab[t], sizeof(int4) * tab[t].size());
    if (e == hipSuccess)
      e = hipMemcpy(b->d_tab[t], tab[t].data(), sizeof(int4) * tab[t].size(), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    harm_batch_destroy(b);
    return seterr("harm_batch_create: %s", hipGetErrorString(e));
  }
  *out = b;
  return 0;
}

// Batch autotuner: the same four variant combinations as one-step batch
// graphs, timed from the live state of every run, which is then restored.
static int batch_autotune(harm_batch* b) {
  if (b->tuned) return 0;
  const int n = b->n;
  std::vector<DevScalars*> save(n, nullptr);  // device copies of each run's scalars
  std::vector<double*> keep(n, nullptr);
  int rc = 0;
  for (int r = 0; r < n && !rc; r++) {
    harm_ctx* c = (*b->ctx)[r];
    if (hipMalloc(&save[r], sizeof(DevScalars)) != hipSuccess ||
        hipMemcpy(save[r], c->ds, sizeof(DevScalars), hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMalloc(&keep[r], (size_t)c->nz * 8 * sizeof(double)) != hipSuccess ||
        hipMemcpy(keep[r], c->arr[HARM_P], (size_t)c->nz * 8 * sizeof(double), hipMemcpyDeviceToDevice) != hipSuccess)
      rc = seterr("batch autotune: save");
  }
  auto restore = [&]() -> hipError_t {
    hipError_t e = hipSuccess;
    for (int r = 0; r < n && e == hipSuccess; r++) {
      harm_ctx* c = (*b->ctx)[r];
      e = hipMemcpyAsync(c->arr[HARM_P], keep[r], (size_t)c->nz * 8 * sizeof(double), hipMemcpyDeviceToDevice,
                         b->stream);
      if (e == hipSuccess) e = hipMemcpyAsync(c->ds, save[r], sizeof(DevScalars), hipMemcpyDeviceToDevice, b->stream);
    }
    return e;
  };
  float best = 1e30f;
  int bf = HARM_FLUX_FACE_64x4, bb = HARM_BOUND_KERNELS;
  const int fv[2] = {HARM_FLUX_FACE_64x4, HARM_FLUX_FACE_64x2};
  const int bv[2] = {HARM_BOUND_KERNELS, HARM_BOUND_BLOCK};
  hipEvent_t ea = nullptr, eb = nullptr;
  if (!rc && (hipEventCreate(&ea) != hipSuccess || hipEventCreate(&eb) != hipSuccess)) rc = seterr("batch autotune: events");
  for (int f = 0; f < 2 && !rc; f++)
    for (int v = 0; v < 2 && !rc; v++) {
      b->flux_variant = fv[f];
      b->bound_variant = bv[v];
      hipGraphExec_t ex;
      if (batch_capture(b, &ex)) { rc = 1; break; }
      hipError_t e = restore();
      if (e == hipSuccess) e = hipGraphLaunch(ex, b->stream);  // warm-up
      if (e == hipSuccess) e = restore();
      if (e == hipSuccess) e = hipEventRecord(ea, b->stream);
      for (int k = 0; k < 3 && e == hipSuccess; k++) e = hipGraphLaunch(ex, b->stream);
      if (e == hipSuccess) e = hipEventRecord(eb, b->stream);
      if (e == hipSuccess) e = hipEventSynchronize(eb);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, ea, eb);
      (void)hipGraphExecDestroy(ex);
      if (e != hipSuccess) { rc = seterr("batch autotune: %s", hipGetErrorString(e)); break; }
      if (ms < best) { best = ms; bf = fv[f]; bb = bv[v]; }
    }
  b->flux_variant = bf;
  b->bound_variant = bb;
  if (!rc) {
    hipError_t e = restore();
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) rc = seterr("batch autotune: %s", hipGetErrorString(e));
  }
  if (ea) (void)hipEventDestroy(ea);
  if (eb) (void)hipEventDestroy(eb);
  for (int r = 0; r < n; r++) {
    (void)hipFree(keep[r]);
    if (save[r]) (void)hipFree(save[r]);
  }
  if (!rc) b->tuned = true;
  return rc;
}

extern "C" int harm_batch_step_n(harm_batch* b, int nsteps) {
  if (!b) return seterr("null batch");
  if (nsteps < 0) return seterr("nsteps < 0");
  if (batch_autotune(b)) return 1;
  if (!b->graph_ok) {
    if (batch_capture(b, &b->graph)) return 1;
    b->graph_ok = true;
  }
  for (int s = 0; s < nsteps; s++) HIPCHK(hipGraphLaunch(b->graph, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

extern "C" int harm_batch_set_tuning(harm_batch* b, int flux_variant, int bound_variant) {
  if (!b) return seterr("null batch");
  if (flux_variant != HARM_FLUX_FACE_64x4 && flux_variant != HARM_FLUX_FACE_64x2)
    return seterr("batch flux variant %d: only the thread-per-face variants are batched", flux_variant);
  if (bound_variant != HARM_BOUND_KERNELS && bound_variant != HARM_BOUND_BLOCK)
    return seterr("bad boundary variant %d", bound_variant);
  b->flux_variant = flux_variant;
  b->bound_variant = bound_variant;
  b->tuned = true;
  if (b->graph_ok) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph_ok = false;
  }
  return 0;
}

extern "C" int harm_batch_get_tuning(harm_batch* b, int* flux_variant, int* bound_variant) {
  if (!b) return seterr("null batch");
  if (flux_variant) *flux_variant = b->flux_variant;
  if (bound_variant) *bound_variant = b->bound_variant;
  return 0;
}

// ===========================================================================
// Discrete adjoint (C-ABI): kernels in the anonymous namespace, then the host
// side, in harm_adjoint.inl
// ===========================================================================
namespace {
#include "harm_adjoint.inl"
#include "harm_acct.inl"
#include "harm_diag.inl"
#include "harm_tangent.inl"
