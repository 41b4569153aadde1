// harm_adjoint.inl -- discrete adjoint of step_ch() on the GPU (included at
// the end of harm_kernels.hip: it drives the forward step's own launchers).
//
// The differentiated map is the conservative step with the time-step
// sequence held fixed (dt_n recorded by the forward run).  Within a step:
//   per RK substage s: F = flux(pb_s) (reconstruction on the branch its
//   limiter took, HLL/LAXF on the branch max/min took), flux-CT, divergence,
//   source; Uf_s = a0 Ui + a1 ulast + a2 dt dU_s; unew += b0 Ui + b1 dt dU_s +
//   b2 Uf_s; pf_s = P(Uf_s) (last substage: P(unew)), the inversion P being
//   differentiated through the implicit-function theorem at its converged
//   solution; floors on their active branch; periodic bound_prim transposed.
// Reverse sweep: checkpoints every K steps; each K-step segment is recomputed
// forward once, storing every substage state, then reversed step by step --
// except the last segment, whose states the forward run stores as it goes
// (K = nsteps: no recompute at all, one state per substage kept in HBM).
// Supported: GRMHD, PERIODIC boundaries, FIXUPFLUX off, limiters DONOR, VANL,
// MINM, MC, NLIM*; HLL or LAXF; FLUXCTTOTH or no CT.  Points where the
// primal is not differentiable (limit_gamma active, failed inversions, clamped
// wave speeds) are counted and reported.

struct AdjRK {
  double a0, a1, a2, b0, b1, b2, dt;
  int last;
};

__device__ __forceinline__ double adj_ld(const KParams& P, const double* ldU, int i, int j, int k) {
  return (i >= 0 && i < P.n1 && j >= 0 && j < P.n2) ? ldU[(size_t)zidx(P, i, j) * 8 + k] : 0.;
}

// The adjoint's steps are replayed as HIP graphs; the step index lives on
// the device (ctr[0] = step n, ctr[1] = first step of the segment), so one
// captured graph serves every step.

// dt of step ctr[0]: record (before the step) or impose
__global__ void k_adj_dt(DevScalars* ds, double* dts, const int* ctr, int impose) {
  const int n = ctr[0];
  if (impose) ds->dt = dts[n];
  else dts[n] = ds->dt;
}

__global__ void k_adj_setctr(int* ctr, int n, int n0) {
  ctr[0] = n;
  ctr[1] = n0;
}

__global__ void k_adj_next(int* ctr, int delta) { ctr[0] += delta; }

// substage state s of step ctr[0] within its segment's store
__device__ __forceinline__ size_t adj_state(const int* ctr, int nto, int s, size_t nz8) {
  return ((size_t)(ctr[0] - ctr[1]) * nto + s) * nz8;
}

// store the state after substage `to` (state to+1 of the step)
__global__ void k_adj_store(const double* __restrict__ src, double* __restrict__ seg, const int* __restrict__ ctr,
                            int nto, int to, size_t nz8) {
  double* dst = seg + adj_state(ctr, nto, to + 1, nz8);
  const size_t n2 = nz8 / 2;
  const double2* s2 = reinterpret_cast<const double2*>(src);
  double2* d2 = reinterpret_cast<double2*>(dst);
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n2; t += (size_t)gridDim.x * blockDim.x)
    d2[t] = s2[t];
}

// slot[0] = pb of substage s (the state before it), slot[1] = its pf
__global__ void k_adj_slot(const double* seg, const int* ctr, int nto, int s, size_t nz8, const double** slot) {
  slot[0] = seg + adj_state(ctr, nto, s, nz8);
  slot[1] = slot[0] + nz8;
}

// periodic bound_prim transposed: x2 ghosts (all i) fold into their sources,
// then x1 ghosts (interior j) fold into theirs; ghost adjoints are cleared
__global__ void k_adj_bt_x2(KParams P, double* lp) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) - NG;
  const int gj = blockIdx.y;
  if (i > P.n1 + NG - 1) return;
  const int j = (gj < NG) ? (gj - NG) : (P.n2 + gj - NG);
  const int js = (gj < NG) ? (j + P.n2) : (j - P.n2);
  double* g = lp + (size_t)zidx(P, i, j) * 8;
  double* s = lp + (size_t)zidx(P, i, js) * 8;
#pragma unroll
  for (int k = 0; k < NPR; k++) {
    s[k] += g[k];
    g[k] = 0.;
  }
}
__global__ void k_adj_bt_x1(KParams P, double* lp) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int gi = blockIdx.y;
  if (j >= P.n2) return;
  const int i = (gi < NG) ? (gi - NG) : (P.n1 + gi - NG);
  const int is = (gi < NG) ? (i + P.n1) : (i - P.n1);
  double* g = lp + (size_t)zidx(P, i, j) * 8;
  double* s = lp + (size_t)zidx(P, is, j) * 8;
#pragma unroll
  for (int k = 0; k < NPR; k++) {
    s[k] += g[k];
    g[k] = 0.;
  }
}

// seed: lambda_N = d(sum 0.5 b^2 gdet dx1 dx2)/dp_N on the interior, and the
// functional's row sums (summed on the host in a fixed order)
__global__ void k_adj_seed_bsq(KParams P, const double* p, const double* gC, double* lp, double* rowsum) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= P.n1 + NG || j >= P.n2 + NG || i < -NG) return;
  const int z = zidx(P, i, j);
  double l[NPR] = {0., 0., 0., 0., 0., 0., 0., 0.};
  if (i >= 0 && i < P.n1 && j >= 0 && j < P.n2) {
    double gc[36];
    load_cent(gC, z, gc);
    GeoU g;
    geo_from_cent(gc, g);
    double pr[NPR];
    load8(p, z, pr);
    bsq_rev(pr, g, g.gdet * P.dx1 * P.dx2, l);
  }
  store8(lp, z, l);
  if (j == 0 && i >= 0 && i < P.n1) {
    double acc = 0.;
    for (int jj = 0; jj < P.n2; jj++) {
      const int zz = zidx(P, i, jj);
      double gc[36];
      load_cent(gC, zz, gc);
      GeoU g;
      geo_from_cent(gc, g);
      double pr[NPR];
      load8(p, zz, pr);
      State q;
      get_state(pr, g, q);
      acc += 0.5 * dot4(q.bcon, q.bcov) * g.gdet * P.dx1 * P.dx2;
    }
    rowsum[i] = acc;
  }
}

// inversion (implicit-function theorem) + RK combination, interior zones
template <int LIMC, int FMC>
__global__ __launch_bounds__(256) void k_adj_zone(KParams P0, AdjRK R, const double* const* __restrict__ slot, const double* __restrict__ dts,
                           const int* __restrict__ ctr, const double* __restrict__ gC,
                           const double* __restrict__ lp, double* __restrict__ lunew, double* __restrict__ lulast,
                           double* __restrict__ lUi, double* __restrict__ ldU, int* __restrict__ nonsmooth) {
  const KParams P = specialize<LIMC, HARM_EOMGRMHD, FMC>(P0);
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= P.n1 || j >= P.n2) return;
  const double* __restrict__ pf = slot[1];
  R.dt = dts[ctr[0]];
  const int z = zidx(P, i, j);
  double gc[36];
  load_cent(gC, z, gc);
  GeoU g;
  geo_from_cent(gc, g);
  double p[NPR], lpf[NPR];
  load8(pf, z, p);
  load8(lp, z, lpf);
  int ns = 0;
  if (P.eomtype == 0) {  // fixup1zone floors on their active branch
    double fr = gc[33], fu = gc[34];
    if (fr < P.rhominlimit) fr = P.rhominlimit;
    if (fu < P.uuminlimit) fu = P.uuminlimit;
    if (p[RHO] == fr) lpf[RHO] = 0.;
    if (p[UU] == fu) lpf[UU] = 0.;
  }
  if (gamma_calc(p, g) >= P.gammamax) ns++;
  double lX[NPR];
  if (inversion_rev(P, p, g, lpf, lX)) {
    ns++;
#pragma unroll
    for (int k = 0; k < NPR; k++) lX[k] = 0.;
  }
  double lun[NPR], lul[NPR], lui[NPR], ldu[NPR];
  load8(lunew, z, lun);
  load8(lulast, z, lul);
  load8(lUi, z, lui);
#pragma unroll
  for (int k = 0; k < NPR; k++) {
    if (R.last) lun[k] += lX[k];
    const double lUf = (R.last ? 0. : lX[k]) + R.b2 * lun[k] + lul[k];
    lui[k] += R.b0 * lun[k] + R.a0 * lUf;
    ldu[k] = R.b1 * R.dt * lun[k] + R.a2 * R.dt * lUf;
    lul[k] = R.a1 * lUf;
  }
  store8(lunew, z, lun);
  store8(lulast, z, lul);
  store8(lUi, z, lui);
  store8(ldU, z, ldu);
  if (ns) atomicAdd(nonsmooth, ns);
}

// corner EMF adjoints (flux-CT transposed), corners i in [0,N1], j in [0,N2]
__global__ void k_adj_emf(KParams P, const double* __restrict__ ldU, double* __restrict__ lemf) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > P.n1 || j > P.n2) return;
  const double idx1 = 1.0 / P.dx1, idx2 = 1.0 / P.dx2;
  auto f1b = [&](int a, int b) { return (adj_ld(P, ldU, a, b, B2) - adj_ld(P, ldU, a - 1, b, B2)) * idx1; };
  auto f2b = [&](int a, int b) { return (adj_ld(P, ldU, a, b, B1) - adj_ld(P, ldU, a, b - 1, B1)) * idx2; };
  lemf[zidx(P, i, j)] = 0.5 * (f1b(i, j) + f1b(i, j - 1)) - 0.5 * (f2b(i, j) + f2b(i - 1, j));
}

// face adjoints: flux adjoint from the divergence (and CT), then the face
// flux reversed to the two reconstructed states; lf[z][0..7] left state,
// [8..15] right state.  blockIdx.z = dir-1.
template <int DIR>
__device__ __forceinline__ void adj_face(const KParams& P, const double* __restrict__ pb,
                                         const double* __restrict__ gF, const double* __restrict__ gC,
                                         const double* __restrict__ ldU, const double* __restrict__ lemf,
                                         double* __restrict__ lf, int i, int j, int& ns) {
  const int zstride = (DIR == 1) ? P.N2M : 1;
  const int z0 = zidx(P, i, j);
  const double idx = 1.0 / ((DIR == 1) ? P.dx1 : P.dx2);
  double Fob[NPR];
#pragma unroll
  for (int k = 0; k < NPR; k++)
    Fob[k] = (DIR == 1) ? (adj_ld(P, ldU, i, j, k) - adj_ld(P, ldU, i - 1, j, k)) * idx
                        : (adj_ld(P, ldU, i, j, k) - adj_ld(P, ldU, i, j - 1, k)) * idx;
  GeoU g;
  load_geo_face(gF, z0, g);
  if (P.fluxb == HARM_FLUXCTTOTH) {
    auto le = [&](int a, int b) {
      return (a >= 0 && a <= P.n1 && b >= 0 && b <= P.n2) ? gC[(size_t)zidx(P, a, b) * GCN + GC_GDETC] * lemf[zidx(P, a, b)]
                                                           : 0.;
    };
    if (DIR == 1) {
      Fob[B1] = 0.;
      Fob[B2] = 0.25 * (le(i, j) + le(i, j + 1)) / g.gdet;
    } else {
      Fob[B2] = 0.;
      Fob[B1] = -0.25 * (le(i, j) + le(i + 1, j)) / g.gdet;
    }
  }
  double s[5][NPR];
#pragma unroll
  for (int m = 0; m < 5; m++) load8(pb, z0 + (m - 3) * zstride, s[m]);
  double p_l[NPR], p_r[NPR];
#pragma unroll
  for (int k = 0; k < NPR; k++) {
    const double dql = slope3(P.lim, s[1][k], s[2][k], s[3][k]);
    const double dqr = slope3(P.lim, s[2][k], s[3][k], s[4][k]);
    p_l[k] = s[2][k] + 0.5 * dql;
    p_r[k] = s[3][k] - 0.5 * dqr;
  }
  double plb[NPR] = {0., 0., 0., 0., 0., 0., 0., 0.}, prb[NPR] = {0., 0., 0., 0., 0., 0., 0., 0.};
  ns += face_flux_rev<DIR>(P, p_l, p_r, g, Fob, plb, prb);
  store8(lf, 2 * z0, plb);
  store8(lf, 2 * z0 + 1, prb);
}

template <int LIMC, int FMC>
__global__ __launch_bounds__(256) void k_adj_face(KParams P0, const double* const* __restrict__ slot,
                                                  const double* __restrict__ gF, const double* __restrict__ gC,
                                                  const double* __restrict__ ldU, const double* __restrict__ lemf,
                                                  double* __restrict__ lf1, double* __restrict__ lf2,
                                                  int* __restrict__ nonsmooth) {
  const KParams P = specialize<LIMC, HARM_EOMGRMHD, FMC>(P0);
  const double* __restrict__ pb = slot[0];
  const int dir = 1 + blockIdx.z;
  const int jf = blockIdx.x * blockDim.x + threadIdx.x;
  const int ifc = blockIdx.y * blockDim.y + threadIdx.y;
  const size_t nz = (size_t)(P.n1 + 2 * NG) * P.N2M;
  int ns = 0;
  if (dir == 1) {
    const int i = ifc, j = jf - 1;
    if (i <= P.n1 && j <= P.n2) adj_face<1>(P, pb, gF, gC, ldU, lemf, lf1, i, j, ns);
  } else {
    const int i = ifc - 1, j = jf;
    if (i <= P.n1 && j <= P.n2) adj_face<2>(P, pb, gF + nz * GF, gC, ldU, lemf, lf2, i, j, ns);
  }
  if (ns) atomicAdd(nonsmooth, ns);
}

// gather: adjoint of every zone of pb from the face-state adjoints of the
// faces whose reconstruction stencil holds it (pl(f) = p(f-1) + dq(f-1)/2,
// pr(f) = p(f) - dq(f)/2, dq on the limiter's branch), plus the source term
template <int LIMC, int FMC>
__global__ __launch_bounds__(256) void k_adj_gather(KParams P0, const double* const* __restrict__ slot,
                                                    const double* __restrict__ gC, const double* __restrict__ gConn,
                                                    const double* __restrict__ ldU, const double* __restrict__ lf1,
                                                    const double* __restrict__ lf2, double* __restrict__ lp) {
  const KParams P = specialize<LIMC, HARM_EOMGRMHD, FMC>(P0);
  const double* __restrict__ pb = slot[0];
  const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x) - NG;
  const int i = (int)(blockIdx.y * blockDim.y + threadIdx.y) - NG;
  if (i > P.n1 + NG - 1 || j > P.n2 + NG - 1) return;
  const int z = zidx(P, i, j);
  double l[NPR];
#pragma unroll
  for (int k = 0; k < NPR; k++) l[k] = 0.;
#pragma unroll
  for (int dir = 1; dir <= 2; dir++) {
    const double* lf = (dir == 1) ? lf1 : lf2;
    const int c = (dir == 1) ? i : j;                 // position along the line
    const int nmax = ((dir == 1) ? P.n1 : P.n2) + NG;  // one past the last zone
    const int stride = (dir == 1) ? P.N2M : 1;
    // lpl(f) = lf[2 z(f)], lpr(f) = lf[2 z(f) + 1]; faces exist for f in
    // [-NG, nmax).  The line's zones c-2..c+2 and faces c-1..c+2 are loaded
    // once, as whole records; out-of-range entries are never used.
    double pz[5][NPR], fl[4][NPR], fr[4][NPR];
#pragma unroll
    for (int m = 0; m < 5; m++) {
      const int w = c - 2 + m;
      if (w >= -NG && w < nmax) {
        load8(pb, z + (w - c) * stride, pz[m]);
      } else {
#pragma unroll
        for (int k = 0; k < NPR; k++) pz[m][k] = 0.;
      }
    }
#pragma unroll
    for (int m = 0; m < 4; m++) {
      const int f = c - 1 + m;
      if (f >= -NG && f < nmax) {
        load8(lf, 2 * (z + (f - c) * stride), fl[m]);
        load8(lf, 2 * (z + (f - c) * stride) + 1, fr[m]);
      } else {
#pragma unroll
        for (int k = 0; k < NPR; k++) fl[m][k] = fr[m][k] = 0.;
      }
    }
#pragma unroll
    for (int k = 0; k < NPR; k++) {
      double acc = fl[2][k] + fr[1][k];  // lpl(c+1) + lpr(c)
#pragma unroll
      for (int dw = -1; dw <= 1; dw++) {
        const int w = c + dw;
        if (w - 1 < -NG || w + 1 >= nmax) continue;
        const double ldq = 0.5 * fl[dw + 2][k] - 0.5 * fr[dw + 1][k];  // lpl(w+1), lpr(w)
        if (ldq == 0.) continue;
        double cl, cc, cr;
        slope3_coef(P.lim, pz[dw + 1][k], pz[dw + 2][k], pz[dw + 3][k], cl, cc, cr);
        acc += ((dw == -1) ? cr : ((dw == 0) ? cc : cl)) * ldq;
      }
      l[k] += acc;
    }
  }
  if (i >= 0 && i < P.n1 && j >= 0 && j < P.n2) {
    double gc[36];
    load_cent(gC, z, gc);
    GeoU g;
    geo_from_cent(gc, g);
    double pr[NPR], ldu[NPR];
    load8(pb, z, pr);
    load8(ldU, z, ldu);
    {
      const size_t cs = (size_t)(P.n1 + 2 * NG) * P.N2M;
      source_rev_at(P, pr, g, [&](int e) { return gConn[(size_t)conn_sym(e) * cs + z]; }, ldu, l);
    }
  }
  store8(lp, z, l);
}

// Ui = U(p_n) transposed into the adjoint of p_n; clears lUi, lunew, lulast
// (the adjoints of the substage accumulators, which start each step at zero)
__global__ __launch_bounds__(256) void k_adj_ui(KParams P, const double* const* __restrict__ slot, const double* __restrict__ gC,
                         double* __restrict__ lUi, double* __restrict__ lunew, double* __restrict__ lulast,
                         double* __restrict__ lp) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= P.n1 || j >= P.n2) return;
  const double* __restrict__ pn = slot[0];
  const int z = zidx(P, i, j);
  double gc[36];
  load_cent(gC, z, gc);
  GeoU g;
  geo_from_cent(gc, g);
  double pr[NPR], lui[NPR], l[NPR];
  load8(pn, z, pr);
  load8(lUi, z, lui);
  load8(lp, z, l);
  State q;
  get_state(pr, g, q);
  StateBar qb;
  statebar_zero(qb);
  primtoflux_rev(P, pr, q, 0, g.gdet, lui, qb, l);
  get_state_rev(pr, g, q, qb, l);
  store8(lp, z, l);
  const double zero[NPR] = {0., 0., 0., 0., 0., 0., 0., 0.};
  store8(lUi, z, zero);
  store8(lunew, z, zero);
  store8(lulast, z, zero);
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct harm_adj {
  harm_ctx* c;
  int nsteps, K, nseg;
  double* d_dts;
  std::vector<double>* dts;        // host copy of the dt sequence
  std::vector<double*>* ckpt;      // P at steps 0, K, 2K, ...
  DevScalars* d_ckds;              // device scalars at the checkpoints
  double* pN;                      // P after the forward run
  DevScalars* d_dsN;
  double* seg;                     // K*nto+1 states of one segment
  int nto;
  double *lp, *lunew, *lulast, *lUi, *ldU, *lemf, *lf1, *lf2, *rowsum;
  int* d_ns;
  int* d_ctr;                      // step index, first step of the segment
  const double** d_slot;           // pb and pf of the substage being reversed
  hipGraphExec_t gx[5];            // steps: record dt, impose dt, impose + store, reverse, record + store
  bool gx_ok[5];
  hipGraphExec_t gx8[5];           // the same, ADJ_STEPS_PER_GRAPH steps per graph
  bool gx8_ok[5];
  bool forward_done, seeded;
  bool kept_last;                  // seg holds the last segment's states from the forward run
  int recompute_failures;
};

static int adj_check(const harm_ctx* c) {
  const harm_params& p = c->prm;
  if (p.eomtype != HARM_EOMGRMHD) return seterr("adjoint: only EOMGRMHD");
  for (int k = 0; k < 4; k++)
    if (p.bc[k] != HARM_PERIODIC) return seterr("adjoint: only PERIODIC boundaries");
  if (p.fixupflux) return seterr("adjoint: FIXUPFLUX must be off");
  if (p.eno) return seterr("adjoint: DOENO has no adjoint");
  if (!(p.lim == HARM_DONOR || p.lim == HARM_VANL || p.lim == HARM_MINM || p.lim == HARM_MC ||
        p.lim == HARM_NLIMCENT || p.lim == HARM_NLIMUP || p.lim == HARM_NLIMDOWN))
    return seterr("adjoint: limiter %d has no adjoint (DONOR, VANL, MINM, MC, NLIM*)", p.lim);
  return 0;
}

extern "C" int harm_adj_destroy(harm_adj* a) {
  if (!a) return 0;
  for (int k = 0; k < 5; k++) {
    if (a->gx_ok[k]) (void)hipGraphExecDestroy(a->gx[k]);
    if (a->gx8_ok[k]) (void)hipGraphExecDestroy(a->gx8[k]);
  }
  if (a->d_ctr) (void)hipFree(a->d_ctr);
  if (a->d_slot) (void)hipFree((void*)a->d_slot);
  if (a->ckpt)
    for (double* pp : *a->ckpt) (void)hipFree(pp);
  double* bufs[] = {a->d_dts, a->pN, a->seg, a->lp, a->lunew, a->lulast, a->lUi, a->ldU, a->lemf, a->lf1, a->lf2,
                    a->rowsum};
  for (double* b : bufs) (void)hipFree(b);
  (void)hipFree(a->d_ckds);
  (void)hipFree(a->d_dsN);
  (void)hipFree(a->d_ns);
  delete a->dts;
  delete a->ckpt;
  delete a;
  return 0;
}

extern "C" int harm_adj_create(harm_ctx* c, int nsteps, int checkpoint_every, harm_adj** out) {
  if (!out) return seterr("null output pointer");
  *out = nullptr;
  if (need_geom(c)) return 1;
  if (nsteps < 1 || checkpoint_every < 1) return seterr("adjoint: nsteps and checkpoint_every must be >= 1");
  if (acct_on(c)) return seterr("adjoint: switch conservation accounting off first (harm_acct_off)");
  if (adj_check(c)) return 1;
  harm_adj* a = new harm_adj();
  memset(a, 0, sizeof(*a));
  a->c = c;
  a->nsteps = nsteps;
  a->K = checkpoint_every < nsteps ? checkpoint_every : nsteps;
  a->nseg = (nsteps + a->K - 1) / a->K;
  double CUf[4][3], Cunew[4][3];
  int pii[4], pbb[4], pff[4];
  rk_setup(c->prm.timeorder, &a->nto, CUf, Cunew, pii, pbb, pff);
  a->dts = new std::vector<double>(nsteps, 0.0);
  a->ckpt = new std::vector<double*>(a->nseg, nullptr);
  const size_t sb = (size_t)c->nz * 8 * sizeof(double);
  hipError_t e = hipMalloc(&a->d_dts, sizeof(double) * nsteps);
  for (int m = 0; m < a->nseg && e == hipSuccess; m++) e = hipMalloc(&(*a->ckpt)[m], sb);
  if (e == hipSuccess) e = hipMalloc(&a->d_ckds, sizeof(DevScalars) * a->nseg);
  if (e == hipSuccess) e = hipMalloc(&a->pN, sb);
  if (e == hipSuccess) e = hipMalloc(&a->d_dsN, sizeof(DevScalars));
  if (e == hipSuccess) e = hipMalloc(&a->seg, sb * ((size_t)a->K * a->nto + 1));
  double** lb[] = {&a->lp, &a->lunew, &a->lulast, &a->lUi, &a->ldU};
  for (double** b : lb)
    if (e == hipSuccess && (e = hipMalloc(b, sb)) == hipSuccess) e = hipMemset(*b, 0, sb);
  if (e == hipSuccess && (e = hipMalloc(&a->lemf, (size_t)c->nz * sizeof(double))) == hipSuccess)
    e = hipMemset(a->lemf, 0, (size_t)c->nz * sizeof(double));
  if (e == hipSuccess && (e = hipMalloc(&a->lf1, 2 * sb)) == hipSuccess) e = hipMemset(a->lf1, 0, 2 * sb);
  if (e == hipSuccess && (e = hipMalloc(&a->lf2, 2 * sb)) == hipSuccess) e = hipMemset(a->lf2, 0, 2 * sb);
  if (e == hipSuccess) e = hipMalloc(&a->rowsum, sizeof(double) * c->kp.n1);
  if (e == hipSuccess && (e = hipMalloc(&a->d_ns, sizeof(int))) == hipSuccess) e = hipMemset(a->d_ns, 0, sizeof(int));
  if (e == hipSuccess && (e = hipMalloc(&a->d_ctr, 2 * sizeof(int))) == hipSuccess)
    e = hipMemset(a->d_ctr, 0, 2 * sizeof(int));
  if (e == hipSuccess && (e = hipMalloc((void**)&a->d_slot, 2 * sizeof(double*))) == hipSuccess)
    e = hipMemset((void*)a->d_slot, 0, 2 * sizeof(double*));
  if (e != hipSuccess) {
    harm_adj_destroy(a);
    return seterr("harm_adj_create: %s", hipGetErrorString(e));
  }
  *out = a;
  return 0;
}

// one step with the dt of step n imposed (or recorded), eager launches
// one forward step of step ctr[0] (dt recorded or imposed); with store, the
// state after every substage goes to the segment store; then ctr[0]++
static int adj_step(harm_adj* a, int impose, bool store) {
  harm_ctx* c = a->c;
  hipLaunchKernelGGL(k_adj_dt, dim3(1), dim3(1), 0, c->stream, c->ds, a->d_dts, a->d_ctr, impose);
  double CUf[4][3], Cunew[4][3];
  int pii[4], pbb[4], pff[4], nto;
  rk_setup(c->prm.timeorder, &nto, CUf, Cunew, pii, pbb, pff);
  const size_t nz8 = (size_t)c->nz * 8;
  for (int to = 0; to < nto; to++) {
    const bool bnd = bound_in_update(c);  // the step's own boundary schedule
    if (substage(c, pii[to], pbb[to], pff[to], CUf[to], Cunew[to], to, nto, bnd)) return 1;
    if (launch_boundary(c, pff[to],
                        KB_BOUND | KB_FIXUP | KB_BOUND2 | ((to == nto - 1) ? KB_DT : 0) | (bnd ? KB_INUPDATE : 0)))
      return 1;
    if (store)
      hipLaunchKernelGGL(k_adj_store, dim3(512), dim3(256), 0, c->stream, c->arr[pff[to]], a->seg, a->d_ctr, nto, to,
                         nz8);
  }
  hipLaunchKernelGGL(k_adj_next, dim3(1), dim3(1), 0, c->stream, a->d_ctr, 1);
  HIPCHK(hipGetLastError());
  return 0;
}

static int adj_boundT(harm_adj* a);

// the reverse of step ctr[0] (its substage states in the segment store),
// then ctr[0]--
// the reverse kernels specialised for the Orszag-Tang scheme (MC with HLL or
// LAXF), or the general ones
template <int LIMC, int FMC>
static void adj_launch_substage(harm_adj* a, const AdjRK& R, bool first) {
  harm_ctx* c = a->c;
  const dim3 b(64, 4);
  const dim3 gz = grid2(c->kp.n2, c->kp.n1, b);
  const dim3 gc2 = grid2(c->kp.n2 + 1, c->kp.n1 + 1, b);
  const dim3 gf((c->kp.n2 + 2 + 63) / 64, (c->kp.n1 + 2 + 3) / 4, 2);
  const dim3 gall = grid2(c->kp.n2 + 2 * NG, c->kp.n1 + 2 * NG, b);
  hipLaunchKernelGGL((k_adj_zone<LIMC, FMC>), gz, b, 0, c->stream, c->kp, R, a->d_slot, a->d_dts, a->d_ctr, c->gC,
                     a->lp, a->lunew, a->lulast, a->lUi, a->ldU, a->d_ns);
  if (c->prm.fluxb == HARM_FLUXCTTOTH)
    hipLaunchKernelGGL(k_adj_emf, gc2, b, 0, c->stream, c->kp, a->ldU, a->lemf);
  hipLaunchKernelGGL((k_adj_face<LIMC, FMC>), gf, b, 0, c->stream, c->kp, a->d_slot, c->gF, c->gC, a->ldU, a->lemf,
                     a->lf1, a->lf2, a->d_ns);
  hipLaunchKernelGGL((k_adj_gather<LIMC, FMC>), gall, b, 0, c->stream, c->kp, a->d_slot, c->gC, c->gConn, a->ldU,
                     a->lf1, a->lf2, a->lp);
  if (first)
    hipLaunchKernelGGL(k_adj_ui, gz, b, 0, c->stream, c->kp, a->d_slot, c->gC, a->lUi, a->lunew, a->lulast, a->lp);
}

static int adj_reverse_step(harm_adj* a) {
  harm_ctx* c = a->c;
  const size_t nz8 = (size_t)c->nz * 8;
  double CUf[4][3], Cunew[4][3];
  int pii[4], pbb[4], pff[4], nto;
  rk_setup(c->prm.timeorder, &nto, CUf, Cunew, pii, pbb, pff);
  for (int s = nto - 1; s >= 0; s--) {
    AdjRK R;
    R.a0 = CUf[s][0]; R.a1 = CUf[s][1]; R.a2 = CUf[s][2];
    R.b0 = Cunew[s][0]; R.b1 = Cunew[s][1]; R.b2 = Cunew[s][2];
    R.dt = 0.;  // read from the dt sequence on the device
    R.last = (s == nto - 1);
    if (adj_boundT(a)) return 1;
    hipLaunchKernelGGL(k_adj_slot, dim3(1), dim3(1), 0, c->stream, a->seg, a->d_ctr, nto, s, nz8, a->d_slot);
    if (c->kp.lim == HARM_MC && c->kp.fluxmethod == HARM_HLLFLUX)
      adj_launch_substage<HARM_MC, HARM_HLLFLUX>(a, R, s == 0);
    else if (c->kp.lim == HARM_MC && c->kp.fluxmethod == HARM_LAXFFLUX)
      adj_launch_substage<HARM_MC, HARM_LAXFFLUX>(a, R, s == 0);
    else
      adj_launch_substage<-1, -1>(a, R, s == 0);
  }
  hipLaunchKernelGGL(k_adj_next, dim3(1), dim3(1), 0, c->stream, a->d_ctr, -1);
  HIPCHK(hipGetLastError());
  return 0;
}

// graph k: 0 forward step recording dt, 1 forward step imposing dt,
// 2 recompute step (imposed dt, states stored), 3 reverse step, 4 forward
// step recording dt with its states stored
// (each step advances the device step counter itself, so a graph of several
// steps is that many steps)
constexpr int ADJ_STEPS_PER_GRAPH = 8;
static int adj_graph(harm_adj* a, int k, bool multi = false) {
  hipGraphExec_t* ex = multi ? &a->gx8[k] : &a->gx[k];
  bool* ok = multi ? &a->gx8_ok[k] : &a->gx_ok[k];
  if (*ok) return 0;
  harm_ctx* c = a->c;
  hipGraph_t g;
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int s = 0; s < (multi ? ADJ_STEPS_PER_GRAPH : 1) && !rc; s++)
    rc = (k == 3) ? adj_reverse_step(a) : adj_step(a, (k == 0 || k == 4) ? 0 : 1, k == 2 || k == 4);
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) {
    if (e == hipSuccess) (void)hipGraphDestroy(g);
    return 1;
  }
  if (e != hipSuccess) return seterr("adjoint graph capture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return seterr("adjoint graph instantiate: %s", hipGetErrorString(e));
  *ok = true;
  return 0;
}

// count consecutive steps of graph k: multi-step replays, then single ones
static int adj_launch_steps(harm_adj* a, int k, int count) {
  harm_ctx* c = a->c;
  int n = 0;
  if (count >= ADJ_STEPS_PER_GRAPH) {
    if (adj_graph(a, k, true)) return 1;
    for (; n + ADJ_STEPS_PER_GRAPH <= count; n += ADJ_STEPS_PER_GRAPH)
      if (hipGraphLaunch(a->gx8[k], c->stream) != hipSuccess) return seterr("adjoint: graph launch");
  }
  if (n < count && adj_graph(a, k)) return 1;
  for (; n < count; n++)
    if (hipGraphLaunch(a->gx[k], c->stream) != hipSuccess) return seterr("adjoint: graph launch");
  return 0;
}

extern "C" int harm_adj_forward(harm_adj* a, const double* dts) {
  if (!a) return seterr("null adjoint");
  harm_ctx* c = a->c;
  if (acct_on(c)) return seterr("adjoint: switch conservation accounting off first (harm_acct_off)");
  if (autotune(c)) return 1;
  const bool was_prof = c->prof;
  c->prof = false;
  const size_t sb = (size_t)c->nz * 8 * sizeof(double);
  if (dts) {
    memcpy(a->dts->data(), dts, sizeof(double) * a->nsteps);
    HIPCHK(hipMemcpy(a->d_dts, dts, sizeof(double) * a->nsteps, hipMemcpyHostToDevice));
  }
  const int gk = dts ? 1 : 0, gks = dts ? 2 : 4;  // plain steps, and the last segment's (states stored)
  const int n0last = (a->nseg - 1) * a->K;
  int rc = adj_graph(a, gk);
  if (!rc) rc = adj_graph(a, gks);
  if (!rc) hipLaunchKernelGGL(k_adj_setctr, dim3(1), dim3(1), 0, c->stream, a->d_ctr, 0, 0);
  for (int n = 0; n < a->nsteps && !rc;) {  // one checkpoint segment at a time
    const int m = n / a->K;
    if (hipMemcpyAsync((*a->ckpt)[m], c->arr[HARM_P], sb, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(a->d_ckds + m, c->ds, sizeof(DevScalars), hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
      rc = seterr("adjoint forward: checkpoint copy");
    if (!rc && n == n0last) {  // the last segment: its initial state, then every substage's
      if (hipMemcpyAsync(a->seg, c->arr[HARM_P], sb, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
        rc = seterr("adjoint forward: segment store");
      hipLaunchKernelGGL(k_adj_setctr, dim3(1), dim3(1), 0, c->stream, a->d_ctr, n0last, n0last);
    }
    const int next = (n + a->K < a->nsteps) ? n + a->K : a->nsteps;
    if (!rc) rc = adj_launch_steps(a, n >= n0last ? gks : gk, next - n);
    n = next;
  }
  c->prof = was_prof;
  if (rc) return 1;
  HIPCHK(hipMemcpyAsync(a->pN, c->arr[HARM_P], sb, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(a->d_dsN, c->ds, sizeof(DevScalars), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (!dts) HIPCHK(hipMemcpy(a->dts->data(), a->d_dts, sizeof(double) * a->nsteps, hipMemcpyDeviceToHost));
  a->forward_done = true;
  a->seeded = false;
  a->kept_last = true;
  return 0;
}

extern "C" int harm_adj_get_dts(harm_adj* a, double* dts) {
  if (!a || !dts) return seterr("null argument");
  if (!a->forward_done) return seterr("adjoint: run harm_adj_forward first");
  memcpy(dts, a->dts->data(), sizeof(double) * a->nsteps);
  return 0;
}

extern "C" int harm_adj_seed_magnetic_energy(harm_adj* a, double* J) {
  if (!a) return seterr("null adjoint");
  if (!a->forward_done) return seterr("adjoint: run harm_adj_forward first");
  harm_ctx* c = a->c;
  const dim3 b(64, 4);
  hipLaunchKernelGGL(k_adj_seed_bsq, grid2(c->kp.n2 + NG, c->kp.n1 + NG, b), b, 0, c->stream, c->kp, a->pN, c->gC,
                     a->lp, a->rowsum);
  HIPCHK(hipGetLastError());
  // ghost rows/columns below 0 are not covered by that grid: clear them
  std::vector<double> rs(c->kp.n1);
  HIPCHK(hipStreamSynchronize(c->stream));
  {
    std::vector<double> h((size_t)c->nz * 8);
    HIPCHK(hipMemcpy(h.data(), a->lp, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < c->N1M; i++)
      for (int j = 0; j < c->N2M; j++)
        if (i < NG || j < NG)
          for (int k = 0; k < 8; k++) h[((size_t)i * c->N2M + j) * 8 + k] = 0.;
    HIPCHK(hipMemcpy(a->lp, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(rs.data(), a->rowsum, sizeof(double) * c->kp.n1, hipMemcpyDeviceToHost));
  double s = 0.;
  for (int i = 0; i < c->kp.n1; i++) s += rs[i];
  if (J) *J = s;
  a->seeded = true;
  return 0;
}

extern "C" int harm_adj_seed(harm_adj* a, const double* lambda) {
  if (!a || !lambda) return seterr("null argument");
  if (!a->forward_done) return seterr("adjoint: run harm_adj_forward first");
  HIPCHK(hipMemcpy(a->lp, lambda, (size_t)a->c->nz * 8 * sizeof(double), hipMemcpyHostToDevice));
  a->seeded = true;
  return 0;
}

static int adj_boundT(harm_adj* a) {
  harm_ctx* c = a->c;
  hipLaunchKernelGGL(k_adj_bt_x2, dim3((c->kp.n1 + 2 * NG + 63) / 64, 2 * NG), dim3(64), 0, c->stream, c->kp, a->lp);
  hipLaunchKernelGGL(k_adj_bt_x1, dim3((c->kp.n2 + 63) / 64, 2 * NG), dim3(64), 0, c->stream, c->kp, a->lp);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int harm_adj_backward(harm_adj* a) {
  if (!a) return seterr("null adjoint");
  if (!a->seeded) return seterr("adjoint: seed the final-state adjoint first");
  harm_ctx* c = a->c;
  if (acct_on(c)) return seterr("adjoint: switch conservation accounting off first (harm_acct_off)");
  const bool was_prof = c->prof;
  c->prof = false;
  const size_t sb = (size_t)c->nz * 8 * sizeof(double);
  HIPCHK(hipMemsetAsync(a->d_ns, 0, sizeof(int), c->stream));
  int rc = adj_graph(a, 2);
  if (!rc) rc = adj_graph(a, 3);
  a->recompute_failures = 0;
  for (int m = a->nseg - 1; m >= 0 && !rc; m--) {
    const int n0 = m * a->K, n1 = (n0 + a->K < a->nsteps) ? n0 + a->K : a->nsteps;
    if (m == a->nseg - 1 && a->kept_last) {  // stored by the forward run
      hipLaunchKernelGGL(k_adj_setctr, dim3(1), dim3(1), 0, c->stream, a->d_ctr, n1 - 1, n0);
      rc = adj_launch_steps(a, 3, n1 - n0);
      continue;
    }
    // recompute the segment, storing every substage state
    HIPCHK(hipMemcpyAsync(c->arr[HARM_P], (*a->ckpt)[m], sb, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->ds, a->d_ckds + m, sizeof(DevScalars), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(a->seg, c->arr[HARM_P], sb, hipMemcpyDeviceToDevice, c->stream));
    int nf0 = 0, nf1 = 0, dummy = 0;
    if (harm_get_failures(c, &nf0, &dummy)) return 1;
    hipLaunchKernelGGL(k_adj_setctr, dim3(1), dim3(1), 0, c->stream, a->d_ctr, n0, n0);
    rc = adj_launch_steps(a, 2, n1 - n0);
    if (rc) break;
    if (harm_get_failures(c, &nf1, &dummy)) return 1;
    a->recompute_failures += nf1 - nf0;
    // reverse the steps of the segment, last first
    hipLaunchKernelGGL(k_adj_setctr, dim3(1), dim3(1), 0, c->stream, a->d_ctr, n1 - 1, n0);
    rc = adj_launch_steps(a, 3, n1 - n0);
  }
  c->prof = was_prof;
  if (a->nseg > 1) a->kept_last = false;  // the earlier segments' recomputes overwrote it
  if (rc) return 1;
  if (adj_boundT(a)) return 1;
  // leave the context at the end of the forward run
  HIPCHK(hipMemcpyAsync(c->arr[HARM_P], a->pN, sb, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->ds, a->d_dsN, sizeof(DevScalars), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int harm_adj_get_lambda(harm_adj* a, double* host) {
  if (!a || !host) return seterr("null argument");
  HIPCHK(hipMemcpy(host, a->lp, (size_t)a->c->nz * 8 * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

// the branch a slope limiter takes on (yl, yc, yr) (interp.c:106-149): for
// MC, which argument each of its two minmods returned (0: the zero branch);
// MINM the same for its one minmod; VANL whether the slopes agree in sign;
// 0 for the limiters without branches
__device__ __forceinline__ int slope_branch(int lim, double yl, double yc, double yr) {
  auto mm = [](double a, double b, double& m) {
    if (a * b <= 0) {
      m = 0.0;
      return 0;
    }
    m = (fabs(a) < fabs(b)) ? a : b;
    return (fabs(a) < fabs(b)) ? 1 : 2;
  };
  double m, m2;
  if (lim == HARM_MC) {
    const int in = mm(2.0 * (yc - yl), 2.0 * (yr - yc), m);
    return in | (mm(0.5 * (yr - yl), m, m2) << 2);
  }
  if (lim == HARM_MINM) return mm(yc - yl, yr - yc, m);
  if (lim == HARM_VANL) return ((yc - yl) * (yr - yc) <= 0.) ? 0 : 1;
  return 0;
}

// limiter branches that differ between two runs' reconstruction inputs
// (state z of each store): every zone whose slope a face uses (i in -1..N1
// along x1, j in -1..N2 along x2), every component
__global__ __launch_bounds__(256) void k_adj_branch_diff(KParams P, const double* __restrict__ sa,
                                                         const double* __restrict__ sb, size_t nz8,
                                                         unsigned long long* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x - 1;
  const int i = blockIdx.y * blockDim.y + threadIdx.y - 1;
  const double* a = sa + blockIdx.z * nz8;
  const double* b = sb + blockIdx.z * nz8;
  unsigned long long n = 0, c = 0;
  if (i <= P.n1 && j <= P.n2) {
#pragma unroll
    for (int dir = 1; dir <= 2; dir++) {
      if ((dir == 1 && (j < 0 || j >= P.n2)) || (dir == 2 && (i < 0 || i >= P.n1))) continue;
      const int zl = (dir == 1) ? zidx(P, i - 1, j) : zidx(P, i, j - 1), z = zidx(P, i, j);
      const int zr = (dir == 1) ? zidx(P, i + 1, j) : zidx(P, i, j + 1);
      for (int k = 0; k < NPR; k++) {
        const int ba = slope_branch(P.lim, a[(size_t)zl * 8 + k], a[(size_t)z * 8 + k], a[(size_t)zr * 8 + k]);
        const int bb = slope_branch(P.lim, b[(size_t)zl * 8 + k], b[(size_t)z * 8 + k], b[(size_t)zr * 8 + k]);
        n += (ba != bb);
        c++;
      }
    }
  }
  n = wave_sum_u64(n);
  c = wave_sum_u64(c);
  if ((threadIdx.x & 63) == 0 && c) {
    atomicAdd(&out[0], n);
    atomicAdd(&out[1], c);
  }
}

extern "C" int harm_adj_branch_switches(harm_adj* a, harm_adj* b, long long* switches, long long* compared) {
  if (!a || !b) return seterr("null adjoint");
  const harm_ctx *ca = a->c, *cb = b->c;
  if (!a->forward_done || !b->forward_done || !a->kept_last || !b->kept_last || a->nseg != 1 || b->nseg != 1)
    return seterr("branch switches: both runs need a forward run that kept every state (checkpoint_every = nsteps)");
  if (ca->prm.n1 != cb->prm.n1 || ca->prm.n2 != cb->prm.n2 || ca->prm.lim != cb->prm.lim || a->nsteps != b->nsteps ||
      a->nto != b->nto)
    return seterr("branch switches: the runs differ in grid, limiter or steps");
  unsigned long long* d = nullptr;
  HIPCHK(hipMalloc(&d, 2 * sizeof(unsigned long long)));
  HIPCHK(hipMemsetAsync(d, 0, 2 * sizeof(unsigned long long), ca->stream));
  const size_t nz8 = (size_t)ca->nz * 8;
  const int ns = a->nsteps * a->nto;  // the substages' reconstruction inputs: states 0 .. ns-1
  const dim3 blk(64, 4);
  for (int s0 = 0; s0 < ns; s0 += 1024) {
    const dim3 g((ca->prm.n2 + 2 + 63) / 64, (ca->prm.n1 + 2 + 3) / 4, std::min(1024, ns - s0));
    hipLaunchKernelGGL(k_adj_branch_diff, g, blk, 0, ca->stream, ca->kp, a->seg + s0 * nz8, b->seg + s0 * nz8, nz8, d);
  }
  unsigned long long h[2] = {0, 0};
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, ca->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ca->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return seterr("branch switches: %s", hipGetErrorString(e));
  if (switches) *switches = (long long)h[0];
  if (compared) *compared = (long long)h[1];
  return 0;
}

extern "C" int harm_adj_get_nonsmooth(harm_adj* a, int* count, int* recompute_failures) {
  if (!a) return seterr("null adjoint");
  int ns = 0;
  HIPCHK(hipMemcpy(&ns, a->d_ns, sizeof(int), hipMemcpyDeviceToHost));
  if (count) *count = ns;
  if (recompute_failures) *recompute_failures = a->recompute_failures;
  return 0;
}
