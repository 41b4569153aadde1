// harm_tangent.inl -- the tangent-linear (forward-mode) step on the GPU
// (included at the end of harm_kernels.hip).
//
// A tangent-linear run advances a context's state p together with its
// derivative dp/ds with respect to one parameter s, given the derivative of
// the geometry tables (dg/ds) and of the initial state.  The kernels run the
// same physics functions as the step, instantiated on the value-plus-tangent
// scalar of harm_tangent.h: the values they compute are the step's to the
// bit (the run replaces the plain step), the tangents follow by the chain
// rule on the branches the values take.  The time steps are the primal's
// (adaptive dt from the values), so the derivative is that of the run with
// its dt sequence held fixed.
//
// Scheme support: EOMFFDE (the force-free inversion is an explicit map; the
// GRMHD Newton inversion has no tangent here), every limiter but CSSLOPE,
// HLL/LAXF, FLUXCTTOTH or FLUXCTHLL, OUTFLOW/POLARAXIS/PERIODIC boundaries with
// or without POLEDEATH, FIXUPFLUX on or off; no DOENO, no accounting, no
// current (harm_tl_create reports anything else).

namespace {

__device__ __forceinline__ void load8t(const double* __restrict__ a, const double* __restrict__ da, int z, Dual* v) {
  double x[NPR], dx[NPR];
  load8(a, z, x);
  load8(da, z, dx);
#pragma unroll
  for (int k = 0; k < NPR; k++) v[k] = Dual(x[k], dx[k]);
}
__device__ __forceinline__ void store8t(double* __restrict__ a, double* __restrict__ da, int z, const Dual* v) {
  double x[NPR], dx[NPR];
#pragma unroll
  for (int k = 0; k < NPR; k++) {
    x[k] = v[k].v;
    dx[k] = v[k].d;
  }
  store8(a, z, x);
  store8(da, z, dx);
}

__device__ __forceinline__ void load_geo_face_t(const double* __restrict__ gF, const double* __restrict__ dgF, int z,
                                                GeoT<Dual>& g) {
  GeoU a, b;
  load_geo_face(gF, z, a);
  load_geo_face(dgF, z, b);
  g.g00 = Dual(a.g00, b.g00); g.g01 = Dual(a.g01, b.g01); g.g02 = Dual(a.g02, b.g02); g.g03 = Dual(a.g03, b.g03);
  g.g11 = Dual(a.g11, b.g11); g.g12 = Dual(a.g12, b.g12); g.g13 = Dual(a.g13, b.g13);
  g.g22 = Dual(a.g22, b.g22); g.g23 = Dual(a.g23, b.g23); g.g33 = Dual(a.g33, b.g33);
  g.n00 = Dual(a.n00, b.n00); g.n01 = Dual(a.n01, b.n01); g.n02 = Dual(a.n02, b.n02); g.n03 = Dual(a.n03, b.n03);
  g.ndd = Dual(a.ndd, b.ndd);
  g.gdet = Dual(a.gdet, b.gdet);
}

__device__ __forceinline__ void load_cent_t(const double* __restrict__ gC, const double* __restrict__ dgC, int z,
                                            Dual* gc) {
  cent_expand([&](int m) { return Dual(gC[(size_t)z * GCN + m], dgC[(size_t)z * GCN + m]); }, gc);
}

// ---- fluxcalc: one thread per face (flux_face_single / face_finish) ----
template <int DIR>
__device__ __forceinline__ void tl_face(const KParams& P, const double* __restrict__ pr, const double* __restrict__ dpr,
                                        double* __restrict__ F, double* __restrict__ dF, double* __restrict__ q,
                                        double* __restrict__ dq, const double* __restrict__ gF,
                                        const double* __restrict__ dgF, int i, int j, double& dtij, int& fail) {
  const int zstride = (DIR == 1) ? P.N2M : 1;
  const int z0 = zidx(P, i, j);
  Dual s[6][NPR], p_l[NPR], p_r[NPR];
#pragma unroll
  for (int m = 0; m < 6; m++) load8t(pr, dpr, z0 + (m - 3) * zstride, s[m]);
#pragma unroll
  for (int k = 0; k < NPR; k++) {
    if (P.lim == HARM_PARA) {
      Dual l, r, l2, r2;
      para4(s[0][k], s[1][k], s[2][k], s[3][k], s[4][k], l, r);
      para4(s[1][k], s[2][k], s[3][k], s[4][k], s[5][k], l2, r2);
      p_l[k] = r;
      p_r[k] = l2;
    } else {
      const Dual dql = slope3(P.lim, s[1][k], s[2][k], s[3][k]);
      const Dual dqr = slope3(P.lim, s[2][k], s[3][k], s[4][k]);
      p_l[k] = s[2][k] + 0.5 * dql;
      p_r[k] = s[3][k] - 0.5 * dqr;
    }
  }
  GeoT<Dual> g;
  load_geo_face_t(gF, dgF, z0, g);
  Dual Fo[NPR], dt;
  face_flux<DIR>(P, p_l, p_r, g, Fo, dt, fail);
  dtij = dt.v;
  // fix_flux() edge rules (fixup.c:1357-1410), as the step's flux kernel applies them
  bool skip_b2 = false;
  if (P.fixupflux) {
    if (DIR == 1) {
      if (i == 0 && P.bc[HARM_X1DN] == HARM_OUTFLOW && j >= -1 && j < P.n2 && Fo[RHO] > 0) Fo[RHO] = 0;
      if (i == P.n1 && P.bc[HARM_X1UP] == HARM_OUTFLOW && j >= -1 && j < P.n2 && Fo[RHO] < 0) Fo[RHO] = 0;
      if (i >= -1 && i < P.n1) {
        if ((j == 0 && P.bc[HARM_X2DN] == HARM_POLARAXIS) || (j == P.n2 - 1 && P.bc[HARM_X2UP] == HARM_POLARAXIS)) {
          const int zg = zidx(P, i, j == 0 ? -1 : P.n2);
          const Dual v = -Fo[B2];
          const Dual qg = v / Dual(gF[(size_t)zg * GF + 15], dgF[(size_t)zg * GF + 15]);
          F[(size_t)zg * 8 + B2] = v.v;
          dF[(size_t)zg * 8 + B2] = v.d;
          q[zg] = qg.v;
          dq[zg] = qg.d;
        }
        skip_b2 = (j == -1 && P.bc[HARM_X2DN] == HARM_POLARAXIS) || (j == P.n2 && P.bc[HARM_X2UP] == HARM_POLARAXIS);
      }
    } else {
      const bool polar_dn = (j == 0 && P.bc[HARM_X2DN] == HARM_POLARAXIS && i >= -1 && i < P.n1);
      const bool polar_up = (j == P.n2 && P.bc[HARM_X2UP] == HARM_POLARAXIS && i >= -1 && i < P.n1);
      const bool equator = (i >= 0 && i < P.n1 && j == (P.n2 - 1) / 2);
      if (polar_dn || polar_up || equator) {
#pragma unroll
        for (int k = 0; k < NPR; k++)
          if (k != U2) Fo[k] = 0.;
      }
    }
  }
  double* Fz = F + (size_t)z0 * 8;
  double* dFz = dF + (size_t)z0 * 8;
#pragma unroll
  for (int k = 0; k < NPR; k++)
    if (!(skip_b2 && k == B2)) {
      Fz[k] = Fo[k].v;
      dFz[k] = Fo[k].d;
    }
  if (!skip_b2) {
    const Dual qq = (DIR == 1 ? Fo[B2] : Fo[B1]) / g.gdet;
    q[z0] = qq.v;
    dq[z0] = qq.d;
  }
}

__global__ __launch_bounds__(128) void k_tl_flux(KParams P, const double* __restrict__ pr,
                                                 const double* __restrict__ dpr, double* __restrict__ F1,
                                                 double* __restrict__ dF1, double* __restrict__ F2,
                                                 double* __restrict__ dF2, double* __restrict__ q1,
                                                 double* __restrict__ dq1, double* __restrict__ q2,
                                                 double* __restrict__ dq2, const double* __restrict__ gF,
                                                 const double* __restrict__ dgF, DevScalars* __restrict__ ds) {
  const int dir = 1 + blockIdx.z;
  const int jf = blockIdx.x * blockDim.x + threadIdx.x;
  const int ifc = blockIdx.y * blockDim.y + threadIdx.y;
  double dtij = 1.e9;
  int fail = 0;
  const size_t nz = (size_t)(P.n1 + 2 * NG) * P.N2M;
  if (dir == 1) {
    const int i = ifc, j = jf - 1;
    if (i <= P.n1 && j <= P.n2) tl_face<1>(P, pr, dpr, F1, dF1, q1, dq1, gF, dgF, i, j, dtij, fail);
  } else {
    const int i = ifc - 1, j = jf;
    if (i <= P.n1 && j <= P.n2) tl_face<2>(P, pr, dpr, F2, dF2, q2, dq2, gF + nz * GF, dgF + nz * GF, i, j, dtij, fail);
  }
  double wm = wave_min(dtij);
  if ((threadIdx.x & 63) == 0 && wm < 1.e9) ndt_push(ds, dir, wm, blockIdx);
  if (fail) atomicOr(&ds->failed, 1);
}

// ---- advance(): one thread per interior zone (update_body<0> + invert_zone) ----
__device__ __forceinline__ Dual tl_emf(const KParams& P, const double* __restrict__ q1, const double* __restrict__ dq1,
                                       const double* __restrict__ q2, const double* __restrict__ dq2,
                                       const double* __restrict__ gC, const double* __restrict__ dgC, int i, int j) {
  const int z = zidx(P, i, j), zj = zidx(P, i, j - 1), zi = zidx(P, i - 1, j);
  const Dual gc(gC[(size_t)z * GCN + GC_GDETC], dgC[(size_t)z * GCN + GC_GDETC]);
  return 0.25 * gc * (Dual(q1[z], dq1[z]) + Dual(q1[zj], dq1[zj]) - Dual(q2[z], dq2[z]) - Dual(q2[zi], dq2[zi]));
}

struct TLArrays {
  const double *pi, *dpi, *pb, *dpb;
  double *pf, *dpf, *ul, *dul, *un, *dun;
  const double *F1, *dF1, *F2, *dF2, *q1, *dq1, *q2, *dq2;
  const double *gC, *dgC, *gConn, *dgConn;
  double ddx1, ddx2;  // tangents of the grid spacings dx1, dx2
};

__global__ __launch_bounds__(256) void k_tl_update(KParams P, RKCoef C, TLArrays A, int* __restrict__ pflag,
                                                   DevScalars* __restrict__ ds) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  int fail = 0;
  ndt_take_first_wave(ds, blockIdx);  // this substage's Courant limits (advance.c:190 reads them)
  if (i < P.n1 && j < P.n2) {
    const int z = zidx(P, i, j);
    const double dt = ds->dt;
    Dual gc[36];
    load_cent_t(A.gC, A.dgC, z, gc);
    GeoT<Dual> g;
    geo_from_cent(gc, g);
    Dual pin[NPR], pbz[NPR], Ui[NPR];
    load8t(A.pi, A.dpi, z, pin);
    load8t(A.pb, A.dpb, z, pbz);
    {
      StateT<Dual> q;
      get_state(pin, g, q);
      primtoflux(P, pin, q, 0, g.gdet, Ui);
    }
    // phases kept in order (each phase's loads issued where it starts): the
    // value-plus-tangent state of the whole update does not fit in registers
    __builtin_amdgcn_sched_barrier(0);
    // source(pb) with the component-major connection and its tangent
    const size_t nzz = (size_t)(P.n1 + 2 * NG) * P.N2M;
    Dual dUg[NPR];
    source_conn_at(P, pbz, g,
                   [&](int e) {
                     const size_t o = (size_t)conn_sym(e) * nzz + z;
                     return Dual(A.gConn[o], A.dgConn[o]);
                   },
                   dUg);
    __builtin_amdgcn_sched_barrier(0);
    // flux2dUavg with the flux-CT B fluxes from the corner EMFs
    Dual f1a[NPR], f1b[NPR], f2a[NPR], f2b[NPR];
    load8t(A.F1, A.dF1, z, f1a);
    load8t(A.F1, A.dF1, zidx(P, i + 1, j), f1b);
    load8t(A.F2, A.dF2, z, f2a);
    load8t(A.F2, A.dF2, zidx(P, i, j + 1), f2b);
    if (P.fluxb == HARM_FLUXCTTOTH) {
      const Dual e00 = tl_emf(P, A.q1, A.dq1, A.q2, A.dq2, A.gC, A.dgC, i, j),
                 e01 = tl_emf(P, A.q1, A.dq1, A.q2, A.dq2, A.gC, A.dgC, i, j + 1),
                 e10 = tl_emf(P, A.q1, A.dq1, A.q2, A.dq2, A.gC, A.dgC, i + 1, j),
                 e11 = tl_emf(P, A.q1, A.dq1, A.q2, A.dq2, A.gC, A.dgC, i + 1, j + 1);
      f1a[B1] = 0.;
      f1a[B2] = 0.5 * (e00 + e01);
      f1b[B1] = 0.;
      f1b[B2] = 0.5 * (e10 + e11);
      f2a[B1] = -0.5 * (e00 + e10);
      f2a[B2] = 0.;
      f2b[B1] = -0.5 * (e01 + e11);
      f2b[B2] = 0.;
    }
    __builtin_amdgcn_sched_barrier(0);
    const Dual idx1 = 1.0 / Dual(P.dx1, A.ddx1), idx2 = 1.0 / Dual(P.dx2, A.ddx2);
    Dual ul[NPR], un[NPR], Uf[NPR];
    if (C.timeorder == 0) {
#pragma unroll
      for (int k = 0; k < NPR; k++) ul[k] = un[k] = 0.0;
    } else {
      load8t(A.ul, A.dul, z, ul);
      load8t(A.un, A.dun, z, un);
    }
#pragma unroll
    for (int k = 0; k < NPR; k++) {
      const Dual d1 = (-(f1b[k] - f1a[k]) * idx1), d2 = (-(f2b[k] - f2a[k]) * idx2);
      const Dual dUr = d1 + d2;
      const Dual dtdU = (dUr + dUg[k]);
      Uf[k] = C.cuf0 * Ui[k] + C.cuf1 * ul[k] + C.cuf2 * dt * dtdU;
      un[k] += C.cn0 * Ui[k] + C.cn1 * dt * dtdU + C.cn2 * Uf[k];
      ul[k] = Uf[k];
    }
    store8t(A.ul, A.dul, z, ul);
    store8t(A.un, A.dun, z, un);
    __builtin_amdgcn_sched_barrier(0);
    // Utoprimgen (force-free) into pf from the guess pi, then fixup1zone
    Dual Ug[NPR], pout[NPR];
#pragma unroll
    for (int k = 0; k < NPR; k++) {
      Ug[k] = (C.last ? un[k] : Uf[k]) / gc[32];
      pout[k] = pin[k];
    }
    if (utoprim_ffde(P, Ug, gc, pout)) fail = 1;  // space-like drift: fatal in the reference
    pflag[z] = 0;
    if (limit_gamma(P.gammamax, pout, g) >= 1) fail = 1;
    store8t(A.pf, A.dpf, z, pout);
  }
  if (fail) atomicOr(&ds->failed, 1);
}

// ---- bound_prim() by one workgroup (bound_block), value and tangent ----
__device__ void tl_bound_x1(const KParams& P, double* __restrict__ pr, double* __restrict__ dpr,
                            const double* __restrict__ gC, const double* __restrict__ dgC, int gi, int j, int& fail) {
  const bool dn = gi < NG;
  const int i = dn ? (gi - NG) : (P.n1 + gi - NG);
  const int bc = dn ? P.bc[HARM_X1DN] : P.bc[HARM_X1UP];
  if (bc != HARM_PERIODIC && bc != HARM_OUTFLOW) return;
  Dual v[NPR];
  load8t(pr, dpr, zidx(P, x1_source(P, i), j), v);
  if (bc == HARM_OUTFLOW) {
    Dual gc[36];
    load_cent_t(gC, dgC, zidx(P, i, j), gc);
    GeoT<Dual> g;
    geo_from_cent(gc, g);
    if (inflow_check_x1(P, i, v, gc, g) >= 1) fail = 1;
    if (limit_gamma(P.gammamax, v, g) >= 1) fail = 1;
  }
  store8t(pr, dpr, zidx(P, i, j), v);
}

__device__ void tl_bound_x2(const KParams& P, double* __restrict__ pr, double* __restrict__ dpr, int i, int gj) {
  const bool dn = gj < NG;
  const int j = dn ? (gj - NG) : (P.n2 + gj - NG);
  const int bc = dn ? P.bc[HARM_X2DN] : P.bc[HARM_X2UP];
  Dual v[NPR];
  if (bc == HARM_PERIODIC) {
    load8t(pr, dpr, zidx(P, i, dn ? j + P.n2 : j - P.n2), v);
    store8t(pr, dpr, zidx(P, i, j), v);
  } else if (bc == HARM_POLARAXIS) {
    const int js = dn ? (0 + (0 - j - 1)) : ((P.n2 - 1) + ((P.n2 - 1) - j + 1));
    load8t(pr, dpr, zidx(P, i, js), v);
    if (P.poledeath) {
      v[U2] *= 0.;
      v[B2] *= 0.;
    }
    store8t(pr, dpr, zidx(P, i, j), v);
  }
}

__device__ void tl_bound_block(const KParams& P, double* __restrict__ pr, double* __restrict__ dpr,
                               const double* __restrict__ gC, const double* __restrict__ dgC, int& fail) {
  const int tid = threadIdx.x, nt = blockDim.x, ni = P.n1 + 2 * NG;
  for (int t = tid; t < 2 * NG * P.n2; t += nt) tl_bound_x1(P, pr, dpr, gC, dgC, t / P.n2, t % P.n2, fail);
  __syncthreads();
  for (int t = tid; t < 2 * NG * ni; t += nt) tl_bound_x2(P, pr, dpr, t % ni - NG, t / ni);
  __syncthreads();
  if (P.poledeath) {
    for (int t = tid; t < 2 * ni; t += nt) {
      const int i = t % ni - NG;
      const bool dn = t < ni;
      if ((dn ? P.bc[HARM_X2DN] : P.bc[HARM_X2UP]) != HARM_POLARAXIS) continue;
      const size_t o = (size_t)zidx(P, i, dn ? 0 : P.n2 - 1) * 8;
      pr[o + U2] *= 0.;
      pr[o + B2] *= 0.;
      dpr[o + U2] *= 0.;
      dpr[o + B2] *= 0.;
    }
  }
  __syncthreads();
}

// bound_prim, (force-free: no post_fixup), bound_prim, and after the last
// substage the t,dt update -- boundary_body's order
__global__ __launch_bounds__(1024) void k_tl_boundary(KParams P, double* __restrict__ pv, double* __restrict__ dpv,
                                                      const double* __restrict__ gC, const double* __restrict__ dgC,
                                                      DevScalars* __restrict__ ds, int bound2, int do_dt) {
  int fail = 0;
  tl_bound_block(P, pv, dpv, gC, dgC, fail);
  if (bound2) tl_bound_block(P, pv, dpv, gC, dgC, fail);
  if (threadIdx.x == 0) {
    ds->nlist = 0;
    if (do_dt) dt_update(P, ds);
  }
  if (fail) atomicOr(&ds->failed, 1);
}

// the outgoing energy flux through zone row i, L = -2 pi dx2 sum_j gdet T^1_0
// (phys.c primtoflux, dir 1, component UU for the force-free equations), and
// its tangent: one lane per j, then one lane adds them in j order
__global__ void k_tl_eflux(KParams P, const double* __restrict__ p, const double* __restrict__ dp,
                           const double* __restrict__ gC, const double* __restrict__ dgC, int i,
                           double* __restrict__ terms, double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < P.n2) {
    const int z = zidx(P, i, j);
    Dual gc[36], pr[NPR], F[NPR];
    load_cent_t(gC, dgC, z, gc);
    GeoT<Dual> g;
    geo_from_cent(gc, g);
    load8t(p, dp, z, pr);
    StateT<Dual> q;
    get_state(pr, g, q);
    primtoflux(P, pr, q, 1, g.gdet, F);
    terms[2 * j] = F[UU].v;
    terms[2 * j + 1] = F[UU].d;
  }
}
__global__ void k_tl_eflux_sum(KParams P, const double* __restrict__ terms, double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Dual acc = 0.;
  for (int j = 0; j < P.n2; j++) acc += Dual(terms[2 * j], terms[2 * j + 1]);
  const Dual L = -(2. * M_PI) * P.dx2 * acc;
  out[0] = L.v;
  out[1] = L.d;
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct harm_tl {
  harm_ctx* c;
  double* dA[3];  // tangents of P, PK0, PK1
  double *dF1, *dF2, *dq1, *dq2, *dul, *dun;
  double *dgF, *dgC, *dgConn;
  double* red;    // energy-flux terms and result
  double ddx[2];  // tangents of dx1, dx2
};

static int tl_check(const harm_ctx* c) {
  const harm_params& p = c->prm;
  if (p.eomtype != HARM_EOMFFDE) return seterr("tangent-linear step: only EOMFFDE (the force-free inversion)");
  if (p.lim == HARM_CSSLOPE) return seterr("tangent-linear step: CSSLOPE has no tangent here");
  if (p.eno) return seterr("tangent-linear step: DOENO has no tangent");
  if (p.checksolution) return seterr("tangent-linear step: CHECKSOLUTION has no tangent");
  if (acct_on(c) || cur_on(c)) return seterr("tangent-linear step: accounting and the current must be off");
  if (general_bounds(c)) return seterr("tangent-linear step: extrapolated, FIXED or disk boundaries have no tangent");
  return 0;
}

extern "C" int harm_tl_destroy(harm_tl* t) {
  if (!t) return 0;
  double* bufs[] = {t->dA[0], t->dA[1], t->dA[2], t->dF1, t->dF2, t->dq1, t->dq2, t->dul, t->dun,
                    t->dgF, t->dgC, t->dgConn, t->red};
  for (double* b : bufs)
    if (b) (void)hipFree(b);
  delete t;
  return 0;
}

extern "C" int harm_tl_create(harm_ctx* c, harm_tl** out) {
  if (!out) return seterr("null output pointer");
  *out = nullptr;
  if (need_geom(c) || tl_check(c)) return 1;
  harm_tl* t = new harm_tl();
  memset(t, 0, sizeof(*t));
  t->c = c;
  const size_t nz = (size_t)c->nz;
  struct { double** p; size_t n; } bufs[] = {
      {&t->dA[0], nz * 8}, {&t->dA[1], nz * 8}, {&t->dA[2], nz * 8}, {&t->dF1, nz * 8}, {&t->dF2, nz * 8},
      {&t->dq1, nz}, {&t->dq2, nz}, {&t->dul, nz * 8}, {&t->dun, nz * 8}, {&t->dgF, 2 * nz * GF},
      {&t->dgC, nz * GCN}, {&t->dgConn, nz * NCONN}, {&t->red, 2 * (size_t)c->kp.n2 + 2}};
  hipError_t e = hipSuccess;
  for (auto& b : bufs) {
    if (e == hipSuccess) e = hipMalloc(b.p, b.n * sizeof(double));
    if (e == hipSuccess) e = hipMemset(*b.p, 0, b.n * sizeof(double));
  }
  if (e != hipSuccess) {
    harm_tl_destroy(t);
    return seterr("harm_tl_create: %s", hipGetErrorString(e));
  }
  *out = t;
  return 0;
}

extern "C" int harm_tl_set_geometry_tangent(harm_tl* t, const double* dgcov, const double* dgcon, const double* dgdet,
                                            const double* dconn) {
  if (!t || !dgcov || !dgcon || !dgdet || !dconn) return seterr("null argument");
  std::vector<double> hF, hC, hT;
  if (!pack_geometry((size_t)t->c->nz, dgcov, dgcon, dgdet, dconn, nullptr, hF, hC, hT))
    return seterr("geometry tangents: gcov and the connection's lower index pair must be symmetric to the bit");
  HIPCHK(hipMemcpy(t->dgF, hF.data(), hF.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(t->dgC, hC.data(), hC.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(t->dgConn, hT.data(), hT.size() * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int harm_tl_set_spacing_tangent(harm_tl* t, double ddx1, double ddx2) {
  if (!t) return seterr("null tangent-linear run");
  t->ddx[0] = ddx1;
  t->ddx[1] = ddx2;
  return 0;
}

extern "C" int harm_tl_set_state_tangent(harm_tl* t, const double* dp) {
  if (!t || !dp) return seterr("null argument");
  HIPCHK(hipMemcpy(t->dA[0], dp, (size_t)t->c->nz * 8 * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int harm_tl_get_state_tangent(harm_tl* t, double* dp) {
  if (!t || !dp) return seterr("null argument");
  HIPCHK(hipStreamSynchronize(t->c->stream));
  HIPCHK(hipMemcpy(dp, t->dA[0], (size_t)t->c->nz * 8 * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

// one step_ch() of value and tangent
static int tl_enqueue_step(harm_tl* t) {
  harm_ctx* c = t->c;
  double CUf[4][3], Cunew[4][3];
  int pii[4], pbb[4], pff[4], nto;
  rk_setup(c->prm.timeorder, &nto, CUf, Cunew, pii, pbb, pff);
  const int nfj = c->kp.n2 + 2, nfi = c->kp.n1 + 2;
  const dim3 bf(64, 2), bu(64, 4);
  for (int to = 0; to < nto; to++) {
    RKCoef C;
    C.cuf0 = CUf[to][0]; C.cuf1 = CUf[to][1]; C.cuf2 = CUf[to][2];
    C.cn0 = Cunew[to][0]; C.cn1 = Cunew[to][1]; C.cn2 = Cunew[to][2];
    C.timeorder = to;
    C.last = (to == nto - 1);
    const int pb = pbb[to], pi = pii[to], pf = pff[to];
    hipLaunchKernelGGL(k_tl_flux, dim3((nfj + 63) / 64, (nfi + 1) / 2, 2), bf, 0, c->stream, c->kp, c->arr[pb],
                       t->dA[pb], c->arr[HARM_F1], t->dF1, c->arr[HARM_F2], t->dF2, c->q1, t->dq1, c->q2, t->dq2,
                       c->gF, t->dgF, c->ds);
    const int pbs = (pf == pb) ? pi : pb;  // advance.c:121: pf = pi before pb is read
    TLArrays A;
    A.pi = c->arr[pi]; A.dpi = t->dA[pi]; A.pb = c->arr[pbs]; A.dpb = t->dA[pbs];
    A.pf = c->arr[pf]; A.dpf = t->dA[pf];
    A.ul = c->arr[HARM_ULAST]; A.dul = t->dul; A.un = c->arr[HARM_UNEW]; A.dun = t->dun;
    A.F1 = c->arr[HARM_F1]; A.dF1 = t->dF1; A.F2 = c->arr[HARM_F2]; A.dF2 = t->dF2;
    A.q1 = c->q1; A.dq1 = t->dq1; A.q2 = c->q2; A.dq2 = t->dq2;
    A.gC = c->gC; A.dgC = t->dgC; A.gConn = c->gConn; A.dgConn = t->dgConn;
    A.ddx1 = t->ddx[0]; A.ddx2 = t->ddx[1];
    hipLaunchKernelGGL(k_tl_update, grid2(c->kp.n2, c->kp.n1, bu), bu, 0, c->stream, c->kp, C, A, c->pflag, c->ds);
    hipLaunchKernelGGL(k_tl_boundary, dim3(1), dim3(1024), 0, c->stream, c->kp, c->arr[pf], t->dA[pf], c->gC,
                       t->dgC, c->ds, 1, C.last ? 1 : 0);
    HIPCHK(hipGetLastError());
  }
  return 0;
}

extern "C" int harm_tl_step_n(harm_tl* t, int nsteps) {
  if (!t) return seterr("null tangent-linear run");
  if (nsteps < 0) return seterr("nsteps < 0");
  if (tl_check(t->c)) return 1;
  for (int s = 0; s < nsteps; s++)
    if (tl_enqueue_step(t)) return 1;
  HIPCHK(hipStreamSynchronize(t->c->stream));
  return 0;
}

extern "C" int harm_tl_energy_flux(harm_tl* t, int i, double* L, double* dL) {
  if (!t) return seterr("null tangent-linear run");
  harm_ctx* c = t->c;
  if (i < 0 || i >= c->kp.n1) return seterr("zone row %d outside [0,%d)", i, c->kp.n1);
  const int n2 = c->kp.n2;
  hipLaunchKernelGGL(k_tl_eflux, dim3((n2 + 63) / 64), dim3(64), 0, c->stream, c->kp, c->arr[HARM_P], t->dA[0],
                     c->gC, t->dgC, i, t->red, t->red + 2 * n2);
  hipLaunchKernelGGL(k_tl_eflux_sum, dim3(1), dim3(64), 0, c->stream, c->kp, t->red, t->red + 2 * n2);
  HIPCHK(hipGetLastError());
  double v[2];
  HIPCHK(hipMemcpyAsync(v, t->red + 2 * n2, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (L) *L = v[0];
  if (dL) *dL = v[1];
  return 0;
}
