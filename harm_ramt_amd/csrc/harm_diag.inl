// harm_diag.inl -- the current and the 56-column dump on the device
// (included by harm_kernels.hip after harm_acct.inl).
//
// The reference's step computes, after its last substage, the Faraday
// tensor of every zone and the current J^mu = F^{mu nu}_{;nu}
// (WHICHCURRENTCALC 1: current_doprecalc + current_calc_1, phys.c:1124-1460),
// with the time derivative of F^{i t} from its last three values once three
// steps have run.  Every value is per zone or a 3-point difference of
// neighbours, so both phases are one thread per zone; the call count and the
// previous dt live in device memory so that a step stays one graph.
// dump_content (dump.c:82-191) is one thread per zone too.

namespace {

struct CurState {
  bool on;
  double* cf;    // cfaraday [nz][5][3]
  double* fcon;  // [nz][6]
  double* jcon;  // [nz][4]
  double* st;    // current_calc_1's statics: call count, last dt
};

// current_doprecalc(0..3) with current_precalc (phys.c:1124-1218), every zone
__global__ void k_cur_pre(KParams P, const double* __restrict__ p, const double* __restrict__ gC,
                          double* __restrict__ cf, double* __restrict__ fcon) {
  const int jj = blockIdx.x * blockDim.x + threadIdx.x;
  const int ii = blockIdx.y * blockDim.y + threadIdx.y;
  if (ii >= P.n1 + 2 * NG || jj >= P.N2M) return;
  const int z = ii * P.N2M + jj;
  double gc[36];
  load_cent(gC, z, gc);
  GeoU g;
  geo_from_cent(gc, g);
  double pr[NPR];
  load8(p, z, pr);
  State q;
  get_state(pr, g, q);
  const double* bc = q.bcov;
  const double* uc = q.ucov;
  double* f = cf + (size_t)z * 15;  // [slot][3]
  f[9] = f[12];
  f[10] = f[13];
  f[11] = f[14];
  f[12] = f[0];
  f[13] = f[1];
  f[14] = f[2];
  f[0] = -1.0 / g.gdet * (-bc[3] * uc[2] + bc[2] * uc[3]);
  f[1] = -1.0 / g.gdet * (bc[3] * uc[1] - bc[1] * uc[3]);
  f[2] = -1.0 / g.gdet * (-bc[2] * uc[1] + bc[1] * uc[2]);
  f[3] = -1.0 / g.gdet * (bc[3] * uc[2] - bc[2] * uc[3]);
  f[4] = -1.0 / g.gdet * (-bc[3] * uc[0] + bc[0] * uc[3]);
  f[5] = -1.0 / g.gdet * (bc[2] * uc[0] - bc[0] * uc[2]);
  f[6] = -1.0 / g.gdet * (-bc[3] * uc[1] + bc[1] * uc[3]);
  f[7] = -1.0 / g.gdet * (bc[3] * uc[0] - bc[0] * uc[3]);
  f[8] = -1.0 / g.gdet * (-bc[1] * uc[0] + bc[0] * uc[1]);
  double* fc = fcon + (size_t)z * 6;
  fc[0] = -1.0 / g.gdet * (bc[3] * uc[2] - bc[2] * uc[3]);
  fc[1] = -1.0 / g.gdet * (-bc[3] * uc[1] + bc[1] * uc[3]);
  fc[2] = -1.0 / g.gdet * (bc[2] * uc[1] - bc[1] * uc[2]);
  fc[3] = -1.0 / g.gdet * (bc[3] * uc[0] - bc[0] * uc[3]);
  fc[4] = -1.0 / g.gdet * (-bc[2] * uc[0] + bc[0] * uc[2]);
  fc[5] = -1.0 / g.gdet * (bc[1] * uc[0] - bc[0] * uc[1]);
}

// current_calc_1 (phys.c:1335-1460) over LOOPP11 x LOOPP12; dtr is the
// step's dt, dtl the previous step's
__global__ void k_cur_calc(KParams P, const double* __restrict__ cf, const double* __restrict__ gC,
                           double* __restrict__ jcon, const DevScalars* __restrict__ ds,
                           const double* __restrict__ st) {
  const int jj = blockIdx.x * blockDim.x + threadIdx.x;
  const int ii = blockIdx.y * blockDim.y + threadIdx.y;
  if (ii >= P.n1 + 4 || jj >= P.n2 + 4) return;
  const int i = ii - 2, j = jj - 2;
  const bool full = st[0] >= 2.0;
  const double dtl = st[1], dtr = ds->dt_step;
  const double idx1 = 1.0 / (2.0 * P.dx1), idx2 = 1.0 / (2.0 * P.dx2);
  auto GD = [&](int a, int b) { return gC[(size_t)zidx(P, a, b) * GCN + GC_GDET]; };
  auto CF = [&](int a, int b, int s, int m) { return cf[(size_t)zidx(P, a, b) * 15 + s * 3 + m]; };
  const double gt = GD(i, j), gr = GD(i, j), grp1 = GD(i + 1, j), ghp1 = GD(i, j + 1), grm1 = GD(i - 1, j),
               ghm1 = GD(i, j - 1);
  double derf[3];
#pragma unroll
  for (int m = 0; m < 3; m++) {
    if (full) {
      const double fl = CF(i, j, 3, m), f0 = CF(i, j, 4, m), fr = CF(i, j, 0, m);
      const double AA = (dtr * (fl - f0) + dtl * (fr - f0)) / (dtl * dtr * (dtl + dtr));
      const double BB = (-dtr * dtr * (fl - f0) + dtl * dtl * (fr - f0)) / (dtl * dtr * (dtl + dtr));
      derf[m] = 2.0 * AA * dtr + BB;
    } else {
      derf[m] = 0;
    }
  }
  double* jc = jcon + (size_t)zidx(P, i, j) * 4;
  jc[0] = 1. / gt * (grp1 * CF(i + 1, j, 1, 0) - grm1 * CF(i - 1, j, 1, 0)) * idx1 +
          1. / gt * (ghp1 * CF(i, j + 1, 2, 0) - ghm1 * CF(i, j - 1, 2, 0)) * idx2;
  jc[1] = derf[0] + 1. / gt * (ghp1 * CF(i, j + 1, 2, 1) - ghm1 * CF(i, j - 1, 2, 1)) * idx2;
  // the reference differences F^{hr} with the zone's own gdet at i-1
  jc[2] = derf[1] + 1. / gt * (grp1 * CF(i + 1, j, 1, 1) - gr * CF(i - 1, j, 1, 1)) * idx1;
  jc[3] = derf[2] + 1. / gt * (grp1 * CF(i + 1, j, 1, 2) - grm1 * CF(i - 1, j, 1, 2)) * idx1 +
          1. / gt * (ghp1 * CF(i, j + 1, 2, 2) - ghm1 * CF(i, j - 1, 2, 2)) * idx2;
}

__global__ void k_cur_tail(const DevScalars* __restrict__ ds, double* __restrict__ st) {
  st[0] += 1.0;
  st[1] = ds->dt_step;
}

// lowerf (phys.c:655-688) with the full CENT gcov
__device__ __forceinline__ void lowerf_gc(const double* gc, const double* fcon, double* fcov) {
  double m[4][4];
  m[0][0] = m[1][1] = m[2][2] = m[3][3] = 0;
  m[0][1] = fcon[0]; m[0][2] = fcon[1]; m[0][3] = fcon[2];
  m[1][2] = fcon[3]; m[1][3] = fcon[4]; m[2][3] = fcon[5];
  m[1][0] = -fcon[0]; m[2][0] = -fcon[1]; m[3][0] = -fcon[2];
  m[2][1] = -fcon[3]; m[3][1] = -fcon[4]; m[3][2] = -fcon[5];
  const int jo[6] = {0, 0, 0, 1, 1, 2}, ko[6] = {1, 2, 3, 2, 3, 3};
#pragma unroll
  for (int n = 0; n < 6; n++) {
    const int j = jo[n], k = ko[n];
    double o = 0;
#pragma unroll
    for (int jp = 0; jp < 4; jp++)
#pragma unroll
      for (int kp = 0; kp < 4; kp++) o += m[jp][kp] * gc[j * 4 + jp] * gc[k * 4 + kp];
    fcov[n] = o;
  }
}

// dump_content (dump.c:82-191) after the six coordinate columns, one thread
// per interior zone: out[n1][n2][50]
__global__ void k_dump(KParams P, const double* __restrict__ p, const double* __restrict__ gC,
                       const double* __restrict__ jcon, const double* __restrict__ fcon, double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= P.n1 || j >= P.n2) return;
  const int z = zidx(P, i, j);
  double gc[36];
  load_cent(gC, z, gc);
  GeoU g;
  geo_from_cent(gc, g);
  double pr[NPR];
  load8(p, z, pr);
  State q;
  get_state(pr, g, q);
  double vmax1, vmin1, vmax2, vmin2;
  vchar(P, pr, q, 1, g, vmax1, vmin1);
  GeoU g2 = g;
  g2.ndd = gc[16 + 10];  // gcon^{22}
  vchar(P, pr, q, 2, g2, vmax2, vmin2);
  double* o = out + ((size_t)i * P.n2 + j) * 50;
  int n = 0;
#pragma unroll
  for (int k = 0; k < NPR; k++) o[n++] = pr[k];
  o[n++] = corner_divb(P, p, gC, i, j);
#pragma unroll
  for (int k = 0; k < 4; k++) o[n++] = q.ucon[k];
#pragma unroll
  for (int k = 0; k < 4; k++) o[n++] = q.ucov[k];
#pragma unroll
  for (int k = 0; k < 4; k++) o[n++] = q.bcon[k];
#pragma unroll
  for (int k = 0; k < 4; k++) o[n++] = q.bcov[k];
  o[n++] = vmin1;
  o[n++] = vmax1;
  o[n++] = vmin2;
  o[n++] = vmax2;
  o[n++] = g.gdet;
  double jc[4] = {0, 0, 0, 0}, fc[6] = {0, 0, 0, 0, 0, 0}, jcov[4], fcov[6];
  if (jcon) {
#pragma unroll
    for (int k = 0; k < 4; k++) jc[k] = jcon[(size_t)z * 4 + k];
#pragma unroll
    for (int k = 0; k < 6; k++) fc[k] = fcon[(size_t)z * 6 + k];
  }
  lower(jc, g, jcov);
#pragma unroll
  for (int k = 0; k < 4; k++) o[n++] = jc[k];
#pragma unroll
  for (int k = 0; k < 4; k++) o[n++] = jcov[k];
  lowerf_gc(gc, fc, fcov);
#pragma unroll
  for (int k = 0; k < 6; k++) o[n++] = fc[k];
#pragma unroll
  for (int k = 0; k < 6; k++) o[n++] = fcov[k];
}

}  // namespace

static bool cur_on(const harm_ctx* c) { return c->cur && c->cur->on; }
static void cur_set_on(harm_ctx* c, bool on) {
  if (c->cur) c->cur->on = on;
}

static void cur_free(harm_ctx* c) {
  CurState* s = c->cur;
  if (!s) return;
  void* bufs[4] = {s->cf, s->fcon, s->jcon, s->st};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  delete s;
  c->cur = nullptr;
}

// step_ch.c:143-153, after the last substage's bound_prim
static int cur_step(harm_ctx* c) {
  CurState* s = c->cur;
  const dim3 b(64, 4);
  hipLaunchKernelGGL(k_cur_pre, grid2(c->N2M, c->N1M, b), b, 0, c->stream, c->kp, c->arr[HARM_P], c->gC, s->cf,
                     s->fcon);
  hipLaunchKernelGGL(k_cur_calc, grid2(c->kp.n2 + 4, c->kp.n1 + 4, b), b, 0, c->stream, c->kp, s->cf, c->gC, s->jcon,
                     c->ds, s->st);
  hipLaunchKernelGGL(k_cur_tail, dim3(1), dim3(1), 0, c->stream, c->ds, s->st);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int harm_current_enable(harm_ctx* c, int on) {
  if (!c) return seterr("null ctx");
  if (!c->cur) {
    CurState* s = new CurState();
    memset(s, 0, sizeof(*s));
    c->cur = s;
    hipError_t e = hipMalloc(&s->cf, (size_t)c->nz * 15 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&s->fcon, (size_t)c->nz * 6 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&s->jcon, (size_t)c->nz * 4 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&s->st, 2 * sizeof(double));
    if (e != hipSuccess) return seterr("current buffers: %s", hipGetErrorString(e));
  }
  CurState* s = c->cur;
  if (on && !s->on) {  // a fresh history, as the reference's statics and zeroed globals start
    HIPCHK(hipMemsetAsync(s->cf, 0, (size_t)c->nz * 15 * sizeof(double), c->stream));
    HIPCHK(hipMemsetAsync(s->fcon, 0, (size_t)c->nz * 6 * sizeof(double), c->stream));
    HIPCHK(hipMemsetAsync(s->jcon, 0, (size_t)c->nz * 4 * sizeof(double), c->stream));
    HIPCHK(hipMemsetAsync(s->st, 0, 2 * sizeof(double), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  s->on = on != 0;
  invalidate_graph(c);
  return 0;
}

extern "C" int harm_get_current(harm_ctx* c, double* jcon, double* fcon) {
  if (!c) return seterr("null ctx");
  if (!c->cur) return seterr("the current was never enabled (harm_current_enable)");
  if (jcon)
    HIPCHK(hipMemcpyAsync(jcon, c->cur->jcon, (size_t)c->nz * 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (fcon)
    HIPCHK(hipMemcpyAsync(fcon, c->cur->fcon, (size_t)c->nz * 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int harm_dump_write(harm_ctx* c, const harm_grid* grid, const char* path, int binary, long realnstep) {
  if (need_geom(c)) return 1;
  if (!grid || !path) return seterr("null argument");
  const int n1 = c->kp.n1, n2 = c->kp.n2;
  if (grid->n1 != n1 || grid->n2 != n2) return seterr("grid %dx%d differs from the context's %dx%d", grid->n1,
                                                      grid->n2, n1, n2);
  double* dev = nullptr;
  HIPCHK(hipMalloc(&dev, (size_t)n1 * n2 * 50 * sizeof(double)));
  const dim3 b(64, 4);
  const bool cur = c->cur != nullptr && c->cur->on;
  hipLaunchKernelGGL(k_dump, grid2(n2, n1, b), b, 0, c->stream, c->kp, c->arr[HARM_P], c->gC,
                     cur ? c->cur->jcon : nullptr, cur ? c->cur->fcon : nullptr, dev);
  std::vector<double> phys((size_t)n1 * n2 * 50);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(phys.data(), dev, phys.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dev);
  if (e != hipSuccess) return seterr("harm_dump_write: %s", hipGetErrorString(e));
  std::vector<double> cols((size_t)n1 * n2 * HARM_DUMP_COLUMNS);
  for (int i = 0; i < n1; i++)
    for (int j = 0; j < n2; j++) {
      double* o = &cols[((size_t)i * n2 + j) * HARM_DUMP_COLUMNS];
      o[0] = (double)i;
      o[1] = (double)j;
      harm_zone_coords(grid, i, j, &o[2], &o[3], &o[4], &o[5]);
      memcpy(o + 6, &phys[((size_t)i * n2 + j) * 50], 50 * sizeof(double));
    }
  harm_dump_header h;
  memset(&h, 0, sizeof(h));
  double t, dt;
  if (harm_get_time(c, &t, &dt)) return 1;
  h.t = t;
  h.totalsize[0] = n1;
  h.totalsize[1] = n2;
  h.startx[0] = grid->startx1;
  h.startx[1] = grid->startx2;
  h.dx[0] = grid->dx1;
  h.dx[1] = grid->dx2;
  h.realnstep = realnstep;
  h.gam = c->prm.gam;
  h.a = grid->a;
  h.R0 = grid->R0;
  h.Rin = grid->Rin;
  h.Rout = grid->Rout;
  h.hslope = grid->hslope;
  h.dt = dt;
  h.defcoord = grid->defcoord;
  const int rc = harm_dump_write_host(path, binary, &h, n1, n2, cols.data());
  if (rc) return seterr("harm_dump_write %s: error %d", path, rc);
  return 0;
}
