// harm_acct.inl -- conservation accounting of the step on the device
// (included by harm_kernels.hip after the context and step code).
//
// The reference accumulates, every step, the fluxes through the boundaries
// of each energy region (diag_flux, diag.c:576-834, after the step) and the
// volume source (diag_source, diag.c:1231-1265, on the last substage), and
// at dump time reduces the state to div B extremes (divbmaxavg,
// dump_ener.c:449-486) and conserved totals (constotal, dump_ener.c:562-590).
// Its sums are serial loops, so each total here is formed in two phases:
// every face or zone term in parallel (the same FP64 operations as the
// reference, one thread each), then one lane per accumulator adding the terms
// in the reference's loop order.  The serial phase is a single wavefront;
// its length is one boundary (diag_flux), one grid row (frdot, lumvsr) or the
// region's zones (sourceadd, constotal, divbavg).

namespace {

constexpr int ACC_NREG = HARM_NUMENERREGIONS;
constexpr int ACC_NFL = HARM_NUMFLUXTERMS;
constexpr int REG_D = (int)(sizeof(harm_ener_region) / sizeof(double));
constexpr int OFF_PDOT = (int)(offsetof(harm_ener_region, pdot) / sizeof(double));
constexpr int OFF_TERMS = (int)(offsetof(harm_ener_region, pdotterms) / sizeof(double));
constexpr int OFF_PCUM = (int)(offsetof(harm_ener_region, pcum) / sizeof(double));
constexpr int OFF_SRC = (int)(offsetof(harm_ener_region, sourceadd) / sizeof(double));
constexpr int OFF_FLADD = (int)(offsetof(harm_ener_region, fladd) / sizeof(double));
constexpr int OFF_FLT = (int)(offsetof(harm_ener_region, fladdterms) / sizeof(double));
constexpr int ACC_NFLOOR = HARM_NUMFAILFLOORFLAGS;
constexpr int OFF_SRCT = (int)(offsetof(harm_ener_region, sourceaddterms) / sizeof(double));
constexpr int OFF_JET2 = ACC_NREG * REG_D;             // pdottermsjet2[4][7][8] after the regions
constexpr int ACC_D = OFF_JET2 + 4 * ACC_NFL * 8;
constexpr int FACE_SLOTS = 8 + ACC_NFL * 8 + 1;        // per face: pdot term, flux terms, jet2 condition

// regions by value in kernel arguments
struct AcctRegions {
  int ep[ACC_NREG][4], df[ACC_NREG][4];
  int off[ACC_NREG][4], len[ACC_NREG][4];  // each boundary's run in the face list
  double dVF, dx3;
};

struct AcctState {
  AcctRegions R;
  bool on;
  double* acc;      // ACC_D doubles: harm_ener_region[3], pdottermsjet2
  double* frdot;    // [n1][8]
  double* lumvsr;   // [n1]
  double* scratch;  // per-term buffer of the first phase
  double* out;      // 4 doubles: dump-time results
  int* faces;       // (i, j) of each boundary face, boundaries in (region, dir) order
  size_t scratch_n;
  // fixup events of a substage (slots of acct_base) and their compaction
  double* pre;      // zone states before fixup1zone [nz][8]
  double* uin;      // the conserved state each force-free inversion took [nz][8]
  int* flag;        // [nslots], category + 1 of the event in the slot
  double* val;      // [nslots][8]
  int* list;        // [nslots][2]: slot, category of the events in order
  double* cval;     // [nslots][8]: their values in that order
  int* cnt;         // [nblocks] events per block of 256 slots, then their offsets
  int* off;
  int* total;
  int nslots, nblocks;
  bool tree;        // HARM_ACCT_TREE: source and fixup totals by reduction trees
  double* part;     // their per-block partial sums
};

__device__ __forceinline__ bool acct_in(const int* ep, int i, int j) {
  return (i >= ep[HARM_X1DN]) && (i <= ep[HARM_X1UP]) && (j >= ep[HARM_X2DN]) && (j <= ep[HARM_X2UP]);
}

// UtoU(UEVOLVE -> UDIAG), phys.c:219-265 (e[k] = gdet)
__device__ __forceinline__ void evolve_to_diag(const KParams& P, double gdet, const double* Uin, double* Uout) {
  double u[NPR];
#pragma unroll
  for (int k = 0; k < NPR; k++) u[k] = Uin[k] / gdet;
  if (P.eomtype == 0) u[UU] += -u[RHO];
#pragma unroll
  for (int k = 0; k < NPR; k++) Uout[k] = u[k] * gdet;
}

// a step's fluxes as the reference's F1/F2 hold them after flux_ct()
// (step_ch.c:1403-1412): the B fluxes of the flux-CT faces from the EMFs
__device__ __forceinline__ void final_flux(const KParams& P, const double* __restrict__ F, const double* __restrict__ q1,
                                          const double* __restrict__ q2, const double* __restrict__ gC, int dir, int i,
                                          int j, double* f) {
  load8(F, zidx(P, i, j), f);
  if (P.fluxb != HARM_FLUXCTTOTH) return;
  if (dir == 1 && i >= 0 && i <= P.n1 && j >= 0 && j <= P.n2 - 1) {
    f[B1] = 0.;
    f[B2] = 0.5 * (emf_at(P, q1, q2, gC, i, j) + emf_at(P, q1, q2, gC, i, j + 1));
  } else if (dir == 2 && i >= 0 && i <= P.n1 - 1 && j >= 0 && j <= P.n2) {
    f[B1] = -0.5 * (emf_at(P, q1, q2, gC, i, j) + emf_at(P, q1, q2, gC, i + 1, j));
    f[B2] = 0.;
  }
}

// The serial phase of a sum: wave 0 adds the items of a chunk in order from
// LDS while waves 1-3 stage the next chunk (double buffer), so that the
// chain of dependent adds never waits on memory.  load(item, w) gives
// component w of an item; consume(first, count, lds) runs on wave 0.
constexpr int STG_CH = 1024;
template <int W, class Load, class Consume>
__device__ __forceinline__ void staged_items(int nitems, double* lds, Load load, Consume consume) {
  const int tid = threadIdx.x;
  const int nch = (nitems + STG_CH - 1) / STG_CH;
  auto stage = [&](int c, double* b) {  // 16 loads in flight per thread
    const int step = blockDim.x - 64;
    for (int q0 = tid - 64; q0 < STG_CH * W; q0 += 16 * step) {
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int q = q0 + u * step, it = c * STG_CH + q / W;
        v[u] = (q < STG_CH * W && it < nitems) ? load(it, q % W) : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int q = q0 + u * step;
        if (q < STG_CH * W) b[q] = v[u];
      }
    }
  };
  if (tid >= 64 && nch > 0) stage(0, lds);
  __syncthreads();
  for (int c = 0; c < nch; c++) {
    if (tid >= 64) {
      if (c + 1 < nch) stage(c + 1, lds + ((c + 1) & 1) * STG_CH * W);
    } else {
      consume(c * STG_CH, min(STG_CH, nitems - c * STG_CH), lds + (c & 1) * STG_CH * W);
    }
    __syncthreads();
  }
}

// Adds column k of the chunk's zones that lie in the box [i0,i1] x [j0,j1]
// to the chains in ZLOOP order (zones i-major from `first`, n2 per row):
// whole row segments at a time, eight LDS values ahead of the adds.
template <int NCH>
__device__ __forceinline__ void box_rows(int first, int cnt, const double* b, int k, int n2, int i0, int i1, int j0,
                                         int j1, double (&acc)[NCH]) {
  int i = first / n2, j = first % n2, m = 0;
  while (m < cnt) {
    const int run = min(cnt - m, n2 - j);
    if (i >= i0 && i <= i1) {
      const int ja = max(j, j0), jb = min(j + run - 1, j1);
      const double* row = b + (ptrdiff_t)(m - j) * 8 + k;  // row[jj*8] is zone (i, jj)
      int jj = ja;
      for (; jj + 16 <= jb + 1; jj += 16) {  // full groups: no per-value test
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) v[u] = row[(jj + u) * 8];
#pragma unroll
        for (int u = 0; u < 16; u++)
#pragma unroll
          for (int c = 0; c < NCH; c++) acc[c] += v[u];
      }
      for (; jj <= jb; jj++)
#pragma unroll
        for (int c = 0; c < NCH; c++) acc[c] += row[jj * 8];
    }
    m += run;
    j += run;
    if (j == n2) {
      j = 0;
      i++;
    }
  }
}

// diag_source's term of each interior zone (diag.c:1247-1259 with the
// geometric source only; cooling is off): UDIAG form of Dt dVF dUcomp[0]
__global__ void k_acct_srcterm(KParams P, AcctRegions R, const double* __restrict__ pb, const double* __restrict__ gC,
                               const double* __restrict__ gConn, const DevScalars* __restrict__ ds,
                               double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= P.n1 || j >= P.n2) return;
  const int z = zidx(P, i, j);
  double gc[36];
  load_cent(gC, z, gc);
  GeoU g;
  geo_from_cent(gc, g);
  double pr[NPR], dU[NPR], ft[NPR], fd[NPR];
  load8(pb, z, pr);
  source_conn_sym(P, pr, g, gConn + z, dU, (size_t)(P.n1 + 2 * NG) * P.N2M);
  const double Dt = ds->dt;
#pragma unroll
  for (int k = 0; k < NPR; k++) ft[k] = Dt * R.dVF * (0.0 + dU[k]);  // dUcomp[GEOMSOURCE] = 0 e + e dUconn
  evolve_to_diag(P, g.gdet, ft, fd);
  store8(out, i * P.n2 + j, fd);
}

// sourceadd and sourceaddterms[0] (lane = region*8 + k) over the region's
// zones in ZLOOP order.  The reference also adds the six cooling sources,
// all +0.0 here; x + 0.0 differs from x only for x = -0.0, and a sum that
// starts at +0.0 never becomes -0.0 (a sum is -0.0 only if both terms are),
// so those adds change nothing and are left out.
__global__ __launch_bounds__(256) void k_acct_srcsum(KParams P, AcctRegions R, const double* __restrict__ term,
                                                     double* __restrict__ acc) {
  __shared__ double lds[2 * STG_CH * 8];
  const int l = threadIdx.x;
  const int r = l / 8, k = l % 8;
  const int rr = r < ACC_NREG ? r : 0;
  const int i0 = max(R.ep[rr][HARM_X1DN], 0), i1 = min(R.ep[rr][HARM_X1UP], P.n1 - 1);
  const int j0 = max(R.ep[rr][HARM_X2DN], 0), j1 = min(R.ep[rr][HARM_X2UP], P.n2 - 1);
  // sourceadd and sourceaddterms[GEOMSOURCE] receive the same terms from the
  // same start (both zeroed together): one chain serves both
  double ch[1] = {0};
  if (l < ACC_NREG * 8) ch[0] = acc[r * REG_D + OFF_SRC + k];
  const int n2 = P.n2;
  staged_items<8>(P.n1 * n2, lds, [&](int z, int w) { return term[(size_t)z * 8 + w]; },
                  [&](int first, int cnt, const double* b) {
                    if (l < ACC_NREG * 8) box_rows<1>(first, cnt, b, k, n2, i0, i1, j0, j1, ch);
                  });
  if (l < ACC_NREG * 8) {
    acc[r * REG_D + OFF_SRC + k] = ch[0];
    acc[r * REG_D + OFF_SRCT + k] = ch[0];
  }
}

// lumvsr (region 0, k = UU): one lane per x1 row, x2 in order
__global__ void k_acct_lumsum(KParams P, AcctRegions R, const double* __restrict__ term, double* __restrict__ lumvsr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int* ep = R.ep[0];
  if (i >= P.n1 || i < ep[HARM_X1DN] || i > ep[HARM_X1UP]) return;
  const int j0 = max(ep[HARM_X2DN], 0), j1 = min(ep[HARM_X2UP], P.n2 - 1);
  const double* row = term + (size_t)i * P.n2 * 8 + UU;
  double l = lumvsr[i];
  for (int j = j0; j <= j1; j += 16) {
    double v[16];
#pragma unroll
    for (int m = 0; m < 16; m++) v[m] = (j + m <= j1) ? row[(size_t)(j + m) * 8] : 0.0;
#pragma unroll
    for (int m = 0; m < 16; m++)
      if (j + m <= j1) l += v[m];
  }
  lumvsr[i] = l;
}

// diag_flux's terms of each boundary face (diag.c:649-803)
__global__ void k_acct_faceterm(KParams P, AcctRegions R, const int* __restrict__ faces, int nfaces,
                                const double* __restrict__ p, const double* __restrict__ F1,
                                const double* __restrict__ F2, const double* __restrict__ q1,
                                const double* __restrict__ q2, const double* __restrict__ gF,
                                const double* __restrict__ gC, double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nfaces) return;
  int r = 0, dir = 0;
  for (int rr = 0; rr < ACC_NREG; rr++)
    for (int d = 0; d < 4; d++)
      if (R.len[rr][d] > 0 && e >= R.off[rr][d] && e < R.off[rr][d] + R.len[rr][d]) { r = rr; dir = d; }
  const int i = faces[2 * e], j = faces[2 * e + 1];
  const bool x1 = (dir == HARM_X1UP || dir == HARM_X1DN);
  const int fluxdir = x1 ? 1 : 2;
  const double surface = x1 ? P.dx2 * R.dx3 : P.dx1 * R.dx3;
  const int z = zidx(P, i, j);
  const size_t nz = (size_t)(P.n1 + 2 * NG) * P.N2M;
  double* o = out + (size_t)e * FACE_SLOTS;
  {
    double f[NPR], ft[NPR], fd[NPR];
    final_flux(P, x1 ? F1 : F2, q1, q2, gC, fluxdir, i, j, f);
    const double fg = gF[((x1 ? 0 : nz) + z) * GF + 15];
#pragma unroll
    for (int k = 0; k < NPR; k++) ft[k] = f[k] * surface;
    evolve_to_diag(P, fg, ft, fd);
#pragma unroll
    for (int k = 0; k < NPR; k++) o[k] = fd[k];
  }
  double gc[36];
  load_cent(gC, z, gc);
  GeoU g;
  geo_from_cent(gc, g);
  double pr[NPR];
  load8(p, z, pr);
  State q;
  get_state(pr, g, q);
  const double pgas = (P.gam - 1.0) * pr[UU];
  double bsq = 0;
#pragma unroll
  for (int l = 0; l < 4; l++) bsq += (q.bcon[l]) * (q.bcov[l]);
  int cond = 0;
  if (r == 0) {
    const double bsqorho = bsq / pr[RHO];
    if (dir == HARM_X1DN) cond = (bsqorho > 3.162);  // JETBSQORHO (diag.c:572)
    else if (dir == HARM_X1UP) cond = ((q.ucon[1] > 0.0) && (-q.ucov[0] - 1.0 > 0.0));
  }
  const double surgdet = (g.gdet) * surface;
  double* tm = o + 8;  // [fl][k]
#pragma unroll
  for (int m = 0; m < ACC_NFL * 8; m++) tm[m] = 0.0;
  tm[0 * 8 + RHO] = pr[RHO] * (q.ucon[fluxdir]) * surgdet;
#pragma unroll
  for (int l = 0; l < 4; l++) {
    const int k = UU + l;
    const double d = (fluxdir == l) ? 1. : 0.;
    tm[0 * 8 + k] = pgas * (q.ucon[fluxdir]) * (q.ucov[l]) * surgdet;
    tm[1 * 8 + k] = pr[RHO] * (q.ucon[fluxdir]) * (q.ucov[l]) * surgdet;
    tm[2 * 8 + k] = pr[UU] * (q.ucon[fluxdir]) * (q.ucov[l]) * surgdet;
    tm[3 * 8 + k] = bsq * (q.ucon[fluxdir]) * (q.ucov[l]) * surgdet;
    tm[4 * 8 + k] = pgas * d * surgdet;
    tm[5 * 8 + k] = 0.5 * bsq * d * surgdet;
    tm[6 * 8 + k] = -(q.bcon[fluxdir]) * (q.bcov[l]) * surgdet;
  }
#pragma unroll
  for (int m = 1; m <= 3; m++) {
    tm[0 * 8 + B1 + m - 1] = (q.bcon[m]) * (q.ucon[fluxdir]) * surgdet;
    tm[1 * 8 + B1 + m - 1] = -(q.bcon[fluxdir]) * (q.ucon[m]) * surgdet;
  }
  o[FACE_SLOTS - 1] = (double)cond;
}

// the boundary sums of diag_flux, one lane per (region, dir, quantity):
// slots 0-7 pdot (then pcum += pdot Dt), 8-63 pdotterms, 64-119 jet2
// (region 0); boundaries with no flux are zeroed as the reference zeroes them
__global__ void k_acct_facesum(AcctRegions R, const double* __restrict__ term, double* __restrict__ acc,
                               const DevScalars* __restrict__ ds, double Dt, int dt_from_step) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  constexpr int NS = 8 + 2 * ACC_NFL * 8;
  if (t >= ACC_NREG * 4 * NS) return;
  const int r = t / (4 * NS), dir = (t / NS) % 4, slot = t % NS;
  const bool jet = slot >= 8 + ACC_NFL * 8;
  if (jet && r != 0) return;
  const int flux = R.df[r][dir] >= 0;
  const int off = R.off[r][dir], len = flux ? R.len[r][dir] : 0;
  const int col = jet ? slot - ACC_NFL * 8 : slot;  // column of the face record
  double s = 0.0;
  for (int e = off; e < off + len; e += 16) {
    double v[16], cnd[16];
#pragma unroll
    for (int m = 0; m < 16; m++) {
      const bool in = e + m < off + len;
      v[m] = in ? term[(size_t)(e + m) * FACE_SLOTS + col] : 0.0;
      cnd[m] = (in && jet) ? term[(size_t)(e + m) * FACE_SLOTS + FACE_SLOTS - 1] : 1.0;
    }
#pragma unroll
    for (int m = 0; m < 16; m++)
      if (e + m < off + len && cnd[m] != 0.0) s += v[m];
  }
  double* a = acc + r * REG_D;
  if (jet) {
    acc[OFF_JET2 + dir * ACC_NFL * 8 + (slot - 8 - ACC_NFL * 8)] = s;
  } else if (slot < 8) {
    a[OFF_PDOT + dir * 8 + slot] = s;
    if (flux) {
      const double dtv = dt_from_step ? ds->dt_step : Dt;
      a[OFF_PCUM + dir * 8 + slot] += s * dtv;
    }
  } else {
    a[OFF_TERMS + dir * ACC_NFL * 8 + (slot - 8)] = s;
  }
}

// frdot (diag.c:816-827): the x1 face terms of the interior, then one lane
// per (row, quantity) summing over x2
__global__ void k_acct_frterm(KParams P, AcctRegions R, const double* __restrict__ F1, const double* __restrict__ q1,
                              const double* __restrict__ q2, const double* __restrict__ gF,
                              const double* __restrict__ gC, double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= P.n1 || j >= P.n2) return;
  double f[NPR], ft[NPR], fd[NPR];
  final_flux(P, F1, q1, q2, gC, 1, i, j, f);
  const double surface = P.dx2 * R.dx3;
#pragma unroll
  for (int k = 0; k < NPR; k++) ft[k] = f[k] * surface;
  evolve_to_diag(P, gF[(size_t)zidx(P, i, j) * GF + 15], ft, fd);
  store8(out, i * P.n2 + j, fd);
}

__global__ void k_acct_frsum(KParams P, const double* __restrict__ term, double* __restrict__ frdot) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P.n1 * 8) return;
  const int i = t / 8, k = t % 8;
  const double* row = term + (size_t)i * P.n2 * 8 + k;
  double s = 0;
  for (int j = 0; j < P.n2; j += 16) {
    double v[16];
#pragma unroll
    for (int m = 0; m < 16; m++) v[m] = (j + m < P.n2) ? row[(size_t)(j + m) * 8] : 0.0;
#pragma unroll
    for (int m = 0; m < 16; m++)
      if (j + m < P.n2) s += v[m];
  }
  frdot[t] = s;
}

// |div B| at corner (i,j) (SETFDIVBFLUXCT, global.h:1152-1175), zero on
// the rows j=0 and j=N2
__device__ double corner_divb(const KParams& P, const double* __restrict__ p, const double* __restrict__ gC, int i,
                              int j) {
  double divb = 0.;
  if ((j != 0) && (j != P.n2)) {
    const double dx1 = P.dx1, dx2 = P.dx2;
    const int z00 = zidx(P, i, j), z0m = zidx(P, i, j - 1), zm0 = zidx(P, i - 1, j), zmm = zidx(P, i - 1, j - 1);
    const double g00 = gC[(size_t)z00 * GCN + GC_GDET], g0m = gC[(size_t)z0m * GCN + GC_GDET];
    const double gm0 = gC[(size_t)zm0 * GCN + GC_GDET], gmm = gC[(size_t)zmm * GCN + GC_GDET];
    const double b100 = p[(size_t)z00 * 8 + B1], b10m = p[(size_t)z0m * 8 + B1];
    const double b1m0 = p[(size_t)zm0 * 8 + B1], b1mm = p[(size_t)zmm * 8 + B1];
    const double b200 = p[(size_t)z00 * 8 + B2], b20m = p[(size_t)z0m * 8 + B2];
    const double b2m0 = p[(size_t)zm0 * 8 + B2], b2mm = p[(size_t)zmm * 8 + B2];
    const double num = ((dx1 > dx2) ? dx1 : dx2) *
                       (0.5 * (b100 * g00 + b10m * g0m - b1m0 * gm0 - b1mm * gmm) / dx1 +
                        0.5 * (b200 * g00 + b2m0 * gm0 - b20m * g0m - b2mm * gmm) / dx2);
    const double sum = b100 * g00 + b10m * g0m + b1m0 * gm0 + b1mm * gmm + b200 * g00 + b2m0 * gm0 + b20m * g0m +
                       b2mm * gmm;
    divb = fabs(num / (gC[(size_t)z00 * GCN + GC_GDETC] * 0.125 * fabs(fabs(sum) + 1.e-20)));
  }
  return divb;
}

// divbmaxavg: |div B| of each corner of LOOPDIVB, then the first maximum and
// the sum in loop order
__global__ void k_divb(KParams P, const double* __restrict__ p, const double* __restrict__ gC,
                       double* __restrict__ out) {
  const int jj = blockIdx.x * blockDim.x + threadIdx.x;
  const int ii = blockIdx.y * blockDim.y + threadIdx.y;
  if (ii >= P.n1 + 4 || jj >= P.n2 + 4) return;
  out[ii * (P.n2 + 4) + jj] = corner_divb(P, p, gC, ii - 2, jj - 2);
}

__global__ __launch_bounds__(256) void k_divb_reduce(KParams P, const double* __restrict__ divb,
                                                     double* __restrict__ res) {
  __shared__ double lds[2 * STG_CH];  // 16 KiB
  const int w = P.n2 + 4;
  double mx = 0, sum = 0;
  int imax = 0, jmax = 0;
  staged_items<1>((P.n1 + 4) * w, lds, [&](int it, int) { return divb[it]; },
                  [&](int first, int cnt, const double* b) {
                    if (threadIdx.x != 0) return;
                    for (int m = 0; m < cnt; m += 8) {
                      double d[8];
#pragma unroll
                      for (int u = 0; u < 8; u++) d[u] = (m + u < cnt) ? b[m + u] : 0.0;
#pragma unroll
                      for (int u = 0; u < 8; u++) {
                        if (m + u >= cnt) break;
                        if (d[u] > mx) {
                          imax = (first + m + u) / w - 2;
                          jmax = (first + m + u) % w - 2;
                          mx = d[u];
                        }
                        sum += d[u];
                      }
                    }
                  });
  if (threadIdx.x == 0) {
    res[0] = mx;
    res[1] = sum / (double)(P.n1 * P.n2);
    res[2] = (double)imax;
    res[3] = (double)jmax;
  }
}

// constotal: U (UDIAG) dVF of each interior zone, then lane k sums the
// region's zones in ZLOOP order
__global__ void k_cons_term(KParams P, AcctRegions R, const double* __restrict__ p, const double* __restrict__ gC,
                            double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= P.n1 || j >= P.n2) return;
  const int z = zidx(P, i, j);
  double gc[36];
  load_cent(gC, z, gc);
  GeoU g;
  geo_from_cent(gc, g);
  double pr[NPR], U[NPR], ft[NPR];
  load8(p, z, pr);
  State q;
  get_state(pr, g, q);
  primtoU_diag(P, pr, q, g.gdet, U);
#pragma unroll
  for (int k = 0; k < NPR; k++) ft[k] = U[k] * R.dVF;
  store8(out, i * P.n2 + j, ft);
}

__global__ __launch_bounds__(256) void k_cons_sum(KParams P, AcctRegions R, int region,
                                                  const double* __restrict__ term, double* __restrict__ res) {
  __shared__ double lds[2 * STG_CH * 8];
  const int k = threadIdx.x;
  const int i0 = max(R.ep[region][HARM_X1DN], 0), i1 = min(R.ep[region][HARM_X1UP], P.n1 - 1);
  const int j0 = max(R.ep[region][HARM_X2DN], 0), j1 = min(R.ep[region][HARM_X2UP], P.n2 - 1);
  const int n2 = P.n2;
  double ch[1] = {0.0};
  staged_items<8>(P.n1 * n2, lds, [&](int z, int w) { return term[(size_t)z * 8 + w]; },
                  [&](int first, int cnt, const double* b) {
                    if (k < NPR) box_rows<1>(first, cnt, b, k, n2, i0, i1, j0, j1, ch);
                  });
  if (k < NPR) res[k] = ch[0];
}

// limit_gamma on the face states of fluxcalc (step_ch.c:978-979): the event
// of each state it changes (faces and their order as the flux kernels)
__global__ __launch_bounds__(256) void k_acct_faceg(KParams P, const double* __restrict__ pr, const double* __restrict__ gF) {
  const int dir = 1 + blockIdx.z;
  const int jf = blockIdx.x * blockDim.x + threadIdx.x;
  const int ifc = blockIdx.y * blockDim.y + threadIdx.y;
  const int i = (dir == 1) ? ifc : ifc - 1, j = (dir == 1) ? jf - 1 : jf;
  if (i > P.n1 || j > P.n2) return;
  double p[2][NPR];
  if (dir == 1) face_recon<1>(P, pr, i, j, p[0], p[1]);
  else face_recon<2>(P, pr, i, j, p[0], p[1]);
  const int kb = B1 + dir - 1;
  const double bavg = 0.5 * (p[0][kb] + p[1][kb]);
  p[0][kb] = bavg;
  p[1][kb] = bavg;
  const size_t nz = (size_t)(P.n1 + 2 * NG) * P.N2M;
  GeoU g;
  load_geo_face(gF + (dir == 2 ? nz * GF : 0), zidx(P, i, j), g);
  const int slot = (dir == 1) ? acct_base(P.n1, P.n2, 0) + (i * (P.n2 + 2) + (j + 1)) * 2
                              : acct_base(P.n1, P.n2, 1) + ((i + 1) * (P.n2 + 1) + j) * 2;
  for (int side = 0; side < 2; side++) {
    double p0[NPR];
#pragma unroll
    for (int k = 0; k < NPR; k++) p0[k] = p[side][k];
    if (limit_gamma(P.gammamax, p[side], g) == -1) acct_event(P, g, p0, p[side], slot + side, 2);
  }
}

// The force-free inversion's dissipation account (phys.ffde.c:354-364,
// ACCOUNTDISS 1; diss_account :572-607, ACCOUNTNONCONS 1) into slot `slot`:
// the geometry-free Ug the inversion took and the state pr it returned.  E/alpha
// and the Lorentz-factor test are formed from Ug as utoprim_ffde forms them; the
// initial T^t_mu from E and B alone (FitUtoT :917-935: Max_con :1472-1552,
// lower_A :1654-1682, ffdestresstensor :1430-1462), the final U from pr
// (primtoU UDIAG), the energy the non-conservative inversion lost added back.
// COUNTUTOPRIMFAILCONV when the cap acted, else COUNTFLOORACT; it counts on
// every substage (diag_fixup_U, fixup.c:234-270, forces its finalstep test).
__device__ void acct_diss_event(const KParams& P, const double* gc, const GeoU& g, const double* Ug,
                                const double* pr, int slot) {
  const double* gcn = gc + 16;
  const double gdet = gc[32];
  double Bcon[4] = {0.0, Ug[B1], Ug[B2], Ug[B3]}, Bcov[4];
  lower_gc(gc, Bcon, Bcov);
  double Bsq = 0.0;
#pragma unroll
  for (int j = 1; j < 4; j++) Bsq += Bcov[j] * Bcon[j];
  const double o = 1.0 / (Bsq * gdet);
  double E[4], Ecv[4];
  E[0] = 0;
  E[1] = (Bcov[2] * Ug[UU + 3] - Bcov[3] * Ug[UU + 2]) * (o);
  E[2] = (Bcov[3] * Ug[UU + 1] - Bcov[1] * Ug[UU + 3]) * (o);
  E[3] = (Bcov[1] * Ug[UU + 2] - Bcov[2] * Ug[UU + 1]) * (o);
  lower_gc(gc, E, Ecv);
  double Esq = 0.0;
#pragma unroll
  for (int j = 1; j < 4; j++) Esq += E[j] * Ecv[j];
  const double gammasq = Bsq / (Bsq - Esq);
  const int modified = (gammasq > P.gammamax * P.gammamax) || (gammasq < 1.0);
  // Maxwell tensor M^{ab} of F^{it} = -E^i/alpha and M^{ti} = -B^i, eta_a = (-alpha, 0, 0, 0)
  const double alpha = 1. / sqrt(-gcn[0]);
  const double et[4] = {-alpha, 0.0, 0.0, 0.0};
  double en[4];
  en[0] = gcn[0] * et[0] + gcn[1] * et[1] + gcn[2] * et[2] + gcn[3] * et[3];
  en[1] = gcn[1] * et[0] + gcn[5] * et[1] + gcn[6] * et[2] + gcn[7] * et[3];
  en[2] = gcn[2] * et[0] + gcn[6] * et[1] + gcn[10] * et[2] + gcn[11] * et[3];
  en[3] = gcn[3] * et[0] + gcn[7] * et[1] + gcn[11] * et[2] + gcn[15] * et[3];
  double Fit[4], Bm[4], Em[4], Emc[4];
  Fit[0] = 0.0;
  Bm[0] = 0.0;
#pragma unroll
  for (int j = 1; j < 4; j++) {
    Fit[j] = -E[j];
    Bm[j] = et[0] * (-Ug[B1 + j - 1]);
  }
#pragma unroll
  for (int j = 0; j < 4; j++) Em[j] = et[0] * Fit[j];
  lower_gc(gc, Em, Emc);
  double M[4][4], Mc[4][4];
  M[0][0] = M[1][1] = M[2][2] = M[3][3] = 0.;
  M[1][2] = -en[1] * Bm[2] + en[2] * Bm[1] - (et[0] / gdet) * Emc[3];
  M[2][3] = -en[2] * Bm[3] + en[3] * Bm[2] - (et[0] / gdet) * Emc[1];
  M[3][1] = -en[3] * Bm[1] + en[1] * Bm[3] - (et[0] / gdet) * Emc[2];
  M[2][1] = -M[1][2];
  M[3][2] = -M[2][3];
  M[1][3] = -M[3][1];
#pragma unroll
  for (int j = 1; j < 4; j++) {
    M[0][j] = Bm[j] / et[0];
    M[j][0] = -M[0][j];
  }
  // lower_A: the six upper entries of M_ab, summed over (j, k) in loop order
  double a01 = 0., a02 = 0., a03 = 0., a12 = 0., a23 = 0., a31 = 0.;
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int k = 0; k < 4; k++) {
      a01 += (gc[0 * 4 + j]) * (gc[1 * 4 + k]) * M[j][k];
      a02 += (gc[0 * 4 + j]) * (gc[2 * 4 + k]) * M[j][k];
      a03 += (gc[0 * 4 + j]) * (gc[3 * 4 + k]) * M[j][k];
      a12 += (gc[1 * 4 + j]) * (gc[2 * 4 + k]) * M[j][k];
      a23 += (gc[2 * 4 + j]) * (gc[3 * 4 + k]) * M[j][k];
      a31 += (gc[3 * 4 + j]) * (gc[1 * 4 + k]) * M[j][k];
    }
  Mc[0][0] = Mc[1][1] = Mc[2][2] = Mc[3][3] = 0.;
  Mc[0][1] = a01; Mc[0][2] = a02; Mc[0][3] = a03; Mc[1][2] = a12; Mc[2][3] = a23; Mc[3][1] = a31;
  Mc[1][0] = -a01; Mc[2][0] = -a02; Mc[3][0] = -a03; Mc[2][1] = -a12; Mc[3][2] = -a23; Mc[1][3] = -a31;
  // ffdestresstensor, row t: T^t_b = M^{tc} M_{bc} - 1/4 delta^t_b (M_cd M^cd)
  double Msq = 0.;
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int k = 0; k < 4; k++) Msq += Mc[j][k] * M[j][k];
  double Tt[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    double t = 0.;
#pragma unroll
    for (int i = 0; i < 4; i++) t += M[0][i] * Mc[k][i];
    Tt[k] = t;
  }
  Tt[0] -= 0.25 * Msq;
  double Ui[NPR], Uf[NPR];
  Ui[RHO] = 0.0;
#pragma unroll
  for (int j = 0; j < 4; j++) Ui[UU + j] = gdet * Tt[j];
#pragma unroll
  for (int j = 1; j < 4; j++) Ui[B1 + j - 1] = gdet * Ug[B1 + j - 1];
  State q;
  get_state(pr, g, q);
  primtoU_diag(P, pr, q, gdet, Uf);
  Uf[RHO] = 0.0;
  Uf[UU] -= (gdet * Ug[UU] - Ui[UU]);
  double* out = P.acct_val + (size_t)slot * 8;
#pragma unroll
  for (int k = 0; k < NPR; k++) out[k] = P.dVF * (Uf[k] - Ui[k]);
  P.acct_flag[slot] = (modified ? 0 : 1) + 1;
}

// the zone's events in the reference's order: the force-free inversion's
// dissipation account, then fixup1zone's (fixup.c:274-377) from the state
// before it: limit_gamma (COUNTLIMITGAMMAACT), then the whole change
// (COUNTFLOORACT, final substage)
__global__ __launch_bounds__(256) void k_acct_zoneg(KParams P, const double* __restrict__ gC, int finalstep) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= P.n1 || j >= P.n2) return;
  const int z = zidx(P, i, j);
  double gc[36];
  load_cent(gC, z, gc);
  GeoU g;
  geo_from_cent(gc, g);
  double p0[NPR], p1[NPR], p2[NPR];
  load8(P.acct_pre, z, p0);
  const int slot = acct_base(P.n1, P.n2, 2) + (i * P.n2 + j) * 3;
  if (P.eomtype == HARM_EOMFFDE) {
    double U[NPR], Ug[NPR], pt[NPR];
    load8(P.acct_uin, z, U);
#pragma unroll
    for (int k = 0; k < NPR; k++) {
      Ug[k] = U[k] / gc[32];
      pt[k] = p0[k];
    }
    // no account when the inversion failed (space-like drift, fatal)
    if (!utoprim_ffde(P, Ug, gc, pt)) acct_diss_event(P, gc, g, Ug, p0, slot);
  }
#pragma unroll
  for (int k = 0; k < NPR; k++) p1[k] = p0[k];
  int changed = 0;
  if (P.eomtype == 0) {
    double fr = gc[33], fu = gc[34];
    if (fr < P.rhominlimit) fr = P.rhominlimit;
    if (fu < P.uuminlimit) fu = P.uuminlimit;
    if (fr > p1[RHO]) { p1[RHO] = fr; changed = 1; }
    if (fu > p1[UU]) { p1[UU] = fu; changed = 1; }
  }
#pragma unroll
  for (int k = 0; k < NPR; k++) p2[k] = p1[k];
  if (limit_gamma(P.gammamax, p2, g) == -1) {
    acct_event(P, g, p1, p2, slot + 1, 2);
    changed = 1;
  }
  if (changed && finalstep) acct_event(P, g, p0, p2, slot + 2, 1);
}

// the failed zones' events from the states fixup_zone kept: limit_gamma's
// (slot flag -1: it changed the averaged state) and the repair's
__global__ __launch_bounds__(256) void k_acct_fixg(KParams P, const double* __restrict__ gC) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= P.n1 || j >= P.n2) return;
  const int slot = acct_base(P.n1, P.n2, 3) + (i * P.n2 + j) * 2;
  const int f = P.acct_flag[slot];
  if (f >= 0) return;
  const int z = zidx(P, i, j);
  double gc[36];
  load_cent(gC, z, gc);
  GeoU g;
  geo_from_cent(gc, g);
  double pavg[NPR], pfin[NPR], pr0[NPR];
  load8(P.acct_val, slot, pavg);
  load8(P.acct_val, slot + 1, pfin);
  load8(P.acct_pre, z, pr0);
  if (f == -1) acct_event(P, g, pavg, pfin, slot, 2);
  else P.acct_flag[slot] = 0;
  acct_event(P, g, pr0, pfin, slot + 1, P.acct_flag[slot + 1] - 1);
}

// ordered compaction of the used slots: counts per block of 256, offsets,
// then each block writes its events at its offset in slot order
__global__ __launch_bounds__(256) void k_acct_count(int nslots, const int* __restrict__ flag, int* __restrict__ cnt) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  const int c = __syncthreads_count((s < nslots) && flag[s] != 0);
  if (threadIdx.x == 0) cnt[blockIdx.x] = c;
}

__global__ __launch_bounds__(256) void k_acct_scan(int nblocks, const int* __restrict__ cnt, int* __restrict__ off,
                                                   int* __restrict__ total) {
  __shared__ int part[256];
  const int t = threadIdx.x;
  const int per = (nblocks + 255) / 256, b0 = t * per, b1 = min(b0 + per, nblocks);
  int sum = 0;
  for (int b = b0; b < b1; b++) sum += cnt[b];
  part[t] = sum;
  __syncthreads();
  if (t == 0) {
    int run = 0;
    for (int u = 0; u < 256; u++) {
      const int v = part[u];
      part[u] = run;
      run += v;
    }
    *total = run;
  }
  __syncthreads();
  int run = part[t];
  for (int b = b0; b < b1; b++) {
    off[b] = run;
    run += cnt[b];
  }
}

__global__ __launch_bounds__(256) void k_acct_compact(int nslots, int* __restrict__ flag, const int* __restrict__ off,
                                                      int* __restrict__ list, const double* __restrict__ val,
                                                      double* __restrict__ cval) {
  __shared__ int wcount[4];
  const int s = blockIdx.x * 256 + threadIdx.x;
  const int f = (s < nslots) ? flag[s] : 0;
  const unsigned long long m = __ballot(f != 0);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) wcount[w] = __popcll(m);
  __syncthreads();
  if (f != 0) {
    int pos = off[blockIdx.x] + __popcll(m & ((1ull << lane) - 1ull));
    for (int v = 0; v < w; v++) pos += wcount[v];
    list[2 * pos] = s;
    list[2 * pos + 1] = f - 1;
#pragma unroll
    for (int m = 0; m < 8; m++) cval[(size_t)pos * 8 + m] = val[(size_t)s * 8 + m];
    flag[s] = 0;
  }
}

// an event's key: bit r set when region r holds it, category + 1 above the
// region bits; 0 when it does not count on this substage (floor and repair
// events count on the last substage only)
__device__ __forceinline__ int acct_slot_key(int n1, int n2, const AcctRegions& R, int slot, int cc, int finalstep) {
  const int bB = acct_base(n1, n2, 1), bC = acct_base(n1, n2, 2), bD = acct_base(n1, n2, 3);
  int i, j;
  if (slot < bB) {
    const int f = slot >> 1;
    i = f / (n2 + 2);
    j = f % (n2 + 2) - 1;
  } else if (slot < bC) {
    const int f = (slot - bB) >> 1;
    i = f / (n2 + 1) - 1;
    j = f % (n2 + 1);
  } else {
    // a zone's third slot (the floors) and a failed zone's second (the
    // repair) count on the accounting substage only
    const bool zone = slot < bD;
    const int rel = slot - (zone ? bC : bD);
    const int f = zone ? rel / 3 : rel >> 1, sub = zone ? rel - 3 * f : rel & 1;
    if (sub == (zone ? 2 : 1) && !finalstep) return 0;
    i = f / n2;
    j = f % n2;
  }
  int key = 0;
#pragma unroll
  for (int q = 0; q < ACC_NREG; q++)
    if (i >= R.ep[q][HARM_X1DN] && i <= R.ep[q][HARM_X1UP] && j >= R.ep[q][HARM_X2DN] && j <= R.ep[q][HARM_X2UP])
      key |= 1 << q;
  return key ? key | ((cc + 1) << ACC_NREG) : 0;
}

// accumulator lane t of the fixup sums: t < 24 fladd[r][k], then
// fladdterms[r][cat][k] (cat = -1 for fladd); the event keys it takes are
// those with (key & kmask) == kwant
struct FixLane {
  int r, cat, k, kmask, kwant;
  __device__ explicit FixLane(int t) {
    constexpr int NT = ACC_NREG * 8;
    if (t < NT) {
      r = t / 8;
      cat = -1;
      k = t % 8;
    } else {
      const int u = t - NT;
      r = u / (ACC_NFLOOR * 8);
      cat = (u / 8) % ACC_NFLOOR;
      k = u % 8;
    }
    kmask = (1 << r) | ((cat < 0) ? 0 : (int)(~0u << ACC_NREG));
    kwant = (1 << r) | ((cat < 0) ? 0 : ((cat + 1) << ACC_NREG));
  }
  __device__ int dst() const { return r * REG_D + ((cat < 0) ? OFF_FLADD + k : OFF_FLT + cat * 8 + k); }
};
constexpr int FIX_LANES = ACC_NREG * 8 * (1 + ACC_NFLOOR);

// the substage's events added in order: lanes 0-23 fladd[r][k], then
// fladdterms[r][cat][k]; floor and repair events count on the last substage
__global__ __launch_bounds__(512) void k_acct_merge(KParams P, AcctRegions R, const int* __restrict__ list,
                                                    const int* __restrict__ total, const double* __restrict__ cval,
                                                    double* __restrict__ acc, int finalstep) {
  __shared__ double ev_val[2][256 * 8];
  // an event's key: bit r set when region r holds it, category + 1 above the
  // region bits; 0 when it does not count on this substage
  __shared__ int ev_key[2][256];
  const int t = threadIdx.x;
  const bool lane = t < FIX_LANES;  // waves 0-3 add, waves 4-7 stage the next 256 events
  const FixLane L(lane ? t : 0);
  const int n1 = P.n1, n2 = P.n2;
  double* dst = acc + L.dst();
  double s = lane ? *dst : 0.0;
  const int E = *total;
  const int nch = (E + 255) / 256;
  const int kmask = L.kmask, kwant = L.kwant;
  // where an event lies and whether it counts on this substage
  auto stage = [&](int c, int h) {
    const int u = t - 256, e = c * 256 + u;
    if (u < 0 || e >= E) return;
    ev_key[h][u] = acct_slot_key(n1, n2, R, list[2 * e], list[2 * e + 1], finalstep);
#pragma unroll
    for (int m = 0; m < 8; m++) ev_val[h][u * 8 + m] = cval[(size_t)e * 8 + m];
  };
  if (nch > 0) stage(0, 0);
  __syncthreads();
  for (int c = 0; c < nch; c++) {
    const int h = c & 1;
    if (t >= 256) {
      if (c + 1 < nch) stage(c + 1, h ^ 1);
    } else if (lane) {
      const int cnt = min(256, E - c * 256);
      for (int m0 = 0; m0 < cnt; m0 += 8) {  // eight events read ahead of the adds
        int kk[8];
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int m = min(m0 + u, 255);
          kk[u] = (m0 + u < cnt) ? ev_key[h][m] : 0;
          v[u] = ev_val[h][m * 8 + L.k];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const bool take = (kk[u] & kmask) == kwant;
          if (take) s += v[u];
        }
      }
    }
    __syncthreads();
  }
  if (lane) *dst = s;
}

// ---- tree sums (HARM_ACCT_TREE) ----
// The same accumulators, each substage's total formed by a fixed tree: every
// block sums its items (one thread per item, then wavefront butterflies and
// the four wavefronts in order) into a row of partial sums, then one
// workgroup per accumulator sums the rows (strided per thread, then the same
// tree) and adds the total.  Reproducible run to run; not the reference's
// serial order.
constexpr int TREE_SRC_NB = 256;  // blocks of the source sum
constexpr int TREE_SRC_N = ACC_NREG * 8;

__device__ __forceinline__ double block_sum(double v, double* lds4) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) lds4[threadIdx.x >> 6] = v;
  __syncthreads();
  return ((lds4[0] + lds4[1]) + (lds4[2] + lds4[3]));
}

// source terms of the interior zones, lane (region, k) of each block
__global__ __launch_bounds__(256) void k_acct_srcsum_part(KParams P, AcctRegions R, const double* __restrict__ term,
                                                          double* __restrict__ part) {
  __shared__ double lds4[4];
  double s[TREE_SRC_N];
#pragma unroll
  for (int a = 0; a < TREE_SRC_N; a++) s[a] = 0.0;
  const int nzz = P.n1 * P.n2;
  for (int z = blockIdx.x * 256 + threadIdx.x; z < nzz; z += TREE_SRC_NB * 256) {
    const int i = z / P.n2, j = z - i * P.n2;
    double v[NPR];
    load8(term, z, v);
#pragma unroll
    for (int r = 0; r < ACC_NREG; r++) {
      const bool in = i >= R.ep[r][HARM_X1DN] && i <= R.ep[r][HARM_X1UP] && j >= R.ep[r][HARM_X2DN] &&
                      j <= R.ep[r][HARM_X2UP];
#pragma unroll
      for (int k = 0; k < NPR; k++)
        if (in) s[r * 8 + k] += v[k];
    }
  }
#pragma unroll
  for (int a = 0; a < TREE_SRC_N; a++) {
    const double b = block_sum(s[a], lds4);
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * TREE_SRC_N + a] = b;
  }
}

// the substage's fixup events, 256 slots per block: lane FixLane(t) of each
// block over the block's events in slot order (used slots cleared)
__global__ __launch_bounds__(256) void k_acct_merge_part(KParams P, AcctRegions R, int nslots, int* __restrict__ flag,
                                                         const double* __restrict__ val, double* __restrict__ part,
                                                         int finalstep) {
  __shared__ int key[256];
  __shared__ double ev[256 * 8];
  const int t = threadIdx.x, slot = blockIdx.x * 256 + t;
  const int f = (slot < nslots) ? flag[slot] : 0;
  int kk = 0;
  if (f != 0) {
    kk = acct_slot_key(P.n1, P.n2, R, slot, f - 1, finalstep);
#pragma unroll
    for (int m = 0; m < 8; m++) ev[t * 8 + m] = val[(size_t)slot * 8 + m];
    flag[slot] = 0;
  }
  key[t] = kk;
  const int any = __syncthreads_count(kk != 0);
  if (t < FIX_LANES) {
    const FixLane L(t);
    double s = 0.0;
    if (any)
      for (int m = 0; m < 256; m++)
        if ((key[m] & L.kmask) == L.kwant) s += ev[m * 8 + L.k];
    part[(size_t)blockIdx.x * FIX_LANES + t] = s;
  }
}

// one workgroup per accumulator a: the sum of its nb partial rows added to
// the accumulator (mode 0: sourceadd[r][k] and sourceaddterms[r][0][k];
// mode 1: fladd / fladdterms)
__global__ __launch_bounds__(256) void k_acct_tree_final(const double* __restrict__ part, int nb, int nacc, int mode,
                                                         double* __restrict__ acc) {
  __shared__ double lds4[4];
  const int a = blockIdx.x;
  double v = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) v += part[(size_t)b * nacc + a];
  const double tot = block_sum(v, lds4);
  if (threadIdx.x == 0) {
    if (mode == 0) {
      const int r = a / 8, k = a % 8;
      const double n = acc[r * REG_D + OFF_SRC + k] + tot;
      acc[r * REG_D + OFF_SRC + k] = n;
      acc[r * REG_D + OFF_SRCT + k] = n;
    } else {
      const int d = FixLane(a).dst();
      acc[d] = acc[d] + tot;
    }
  }
}

}  // namespace

static bool acct_on(const harm_ctx* c) { return c->acct && c->acct->on; }
// accounting on/off for the steps; the kernels see its buffers only while on
static void acct_set_on(harm_ctx* c, bool on) {
  AcctState* a = c->acct;
  if (a) a->on = on;
  const bool k = a && on;
  c->kp.acct_pre = k ? a->pre : nullptr;
  c->kp.acct_uin = k ? a->uin : nullptr;
  c->kp.acct_flag = k ? a->flag : nullptr;
  c->kp.acct_val = k ? a->val : nullptr;
}

static int acct_alloc(harm_ctx* c) {
  if (c->acct) return 0;
  AcctState* a = new AcctState();
  memset(a, 0, sizeof(*a));
  const int n1 = c->prm.n1, n2 = c->prm.n2;
  a->scratch_n = std::max((size_t)n1 * n2 * 8, (size_t)(n1 + 4) * (n2 + 4));
  a->scratch_n = std::max(a->scratch_n, (size_t)ACC_NREG * 2 * (n1 + n2 + 2) * FACE_SLOTS);
  hipError_t e = hipMalloc(&a->acc, ACC_D * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&a->frdot, (size_t)n1 * 8 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&a->lumvsr, (size_t)n1 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&a->scratch, a->scratch_n * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&a->out, 8 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&a->faces, (size_t)ACC_NREG * 2 * (n1 + n2 + 2) * 2 * sizeof(int));
  a->nslots = acct_base(n1, n2, 4);
  a->nblocks = (a->nslots + 255) / 256;
  if (e == hipSuccess) e = hipMalloc(&a->pre, (size_t)c->nz * 8 * sizeof(double));
  if (e == hipSuccess && c->prm.eomtype == HARM_EOMFFDE) e = hipMalloc(&a->uin, (size_t)c->nz * 8 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&a->flag, (size_t)a->nslots * sizeof(int));
  if (e == hipSuccess) e = hipMemset(a->flag, 0, (size_t)a->nslots * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&a->val, (size_t)a->nslots * 8 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&a->list, (size_t)a->nslots * 2 * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&a->cval, (size_t)a->nslots * 8 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&a->cnt, (size_t)a->nblocks * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&a->off, (size_t)a->nblocks * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&a->total, sizeof(int));
  if (e == hipSuccess)
    e = hipMalloc(&a->part, (size_t)std::max(TREE_SRC_NB * TREE_SRC_N, a->nblocks * FIX_LANES) * sizeof(double));
  c->acct = a;
  if (e != hipSuccess) return seterr("accounting buffers: %s", hipGetErrorString(e));
  return 0;
}

static void acct_free(harm_ctx* c) {
  AcctState* a = c->acct;
  if (!a) return;
  void* bufs[16] = {a->acc,  a->frdot, a->lumvsr, a->scratch, a->out, a->faces, a->pre,  a->uin,
                     a->flag, a->val,   a->list,   a->cval,    a->cnt, a->off,   a->total, a->part};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  delete a;
  c->acct = nullptr;
}

// diag_source on the last substage, before its update overwrites pf (the
// source is that of the array the update reads, advance.c:121-140)
static int acct_source(harm_ctx* c, int pi, int pb, int pf) {
  AcctState* a = c->acct;
  const int pbs = (pf == pb) ? pi : pb;
  const dim3 b(64, 4);
  hipLaunchKernelGGL(k_acct_srcterm, grid2(c->kp.n2, c->kp.n1, b), b, 0, c->stream, c->kp, a->R, c->arr[pbs], c->gC,
                     c->gConn, c->ds, a->scratch);
  if (a->tree) {
    hipLaunchKernelGGL(k_acct_srcsum_part, dim3(TREE_SRC_NB), dim3(256), 0, c->stream, c->kp, a->R, a->scratch,
                       a->part);
    hipLaunchKernelGGL(k_acct_tree_final, dim3(TREE_SRC_N), dim3(256), 0, c->stream, a->part, TREE_SRC_NB,
                       TREE_SRC_N, 0, a->acc);
  } else {
    hipLaunchKernelGGL(k_acct_srcsum, dim3(1), dim3(256), 0, c->stream, c->kp, a->R, a->scratch, a->acc);
  }
  hipLaunchKernelGGL(k_acct_lumsum, dim3((c->kp.n1 + 63) / 64), dim3(64), 0, c->stream, c->kp, a->R, a->scratch,
                     a->lumvsr);
  HIPCHK(hipGetLastError());
  return 0;
}

// the fixup events of one substage: faces (from pb, before the substage) ...
static int acct_faces(harm_ctx* c, int pb) {
  const dim3 b(64, 4);
  hipLaunchKernelGGL(k_acct_faceg, dim3((c->kp.n2 + 2 + 63) / 64, (c->kp.n1 + 2 + 3) / 4, 2), b, 0, c->stream,
                     c->kp, c->arr[pb], c->gF);
  HIPCHK(hipGetLastError());
  return 0;
}

// ... zones (after the update) ...
static int acct_zones(harm_ctx* c, int finalstep) {
  const dim3 b(64, 4);
  hipLaunchKernelGGL(k_acct_zoneg, grid2(c->kp.n2, c->kp.n1, b), b, 0, c->stream, c->kp, c->gC, finalstep);
  HIPCHK(hipGetLastError());
  return 0;
}

// ... and, after the failed zones' repair, all of them added in order
static int acct_merge(harm_ctx* c, int finalstep) {
  AcctState* a = c->acct;
  if (c->prm.eomtype == HARM_EOMGRMHD) {  // fixup_utoprim runs for GRMHD only
    const dim3 b(64, 4);
    hipLaunchKernelGGL(k_acct_fixg, grid2(c->kp.n2, c->kp.n1, b), b, 0, c->stream, c->kp, c->gC);
  }
  if (a->tree) {
    hipLaunchKernelGGL(k_acct_merge_part, dim3(a->nblocks), dim3(256), 0, c->stream, c->kp, a->R, a->nslots, a->flag,
                       a->val, a->part, finalstep);
    hipLaunchKernelGGL(k_acct_tree_final, dim3(FIX_LANES), dim3(256), 0, c->stream, a->part, a->nblocks, FIX_LANES,
                       1, a->acc);
    HIPCHK(hipGetLastError());
    return 0;
  }
  hipLaunchKernelGGL(k_acct_count, dim3(a->nblocks), dim3(256), 0, c->stream, a->nslots, a->flag, a->cnt);
  hipLaunchKernelGGL(k_acct_scan, dim3(1), dim3(256), 0, c->stream, a->nblocks, a->cnt, a->off, a->total);
  hipLaunchKernelGGL(k_acct_compact, dim3(a->nblocks), dim3(256), 0, c->stream, a->nslots, a->flag, a->off, a->list,
                     a->val, a->cval);
  static_assert(FIX_LANES <= 256, "one workgroup holds every accumulator lane");
  hipLaunchKernelGGL(k_acct_merge, dim3(1), dim3(512), 0, c->stream, c->kp, a->R, a->list, a->total, a->cval, a->acc,
                     finalstep);
  HIPCHK(hipGetLastError());
  return 0;
}

// diag_flux of array HARM_P and the last fluxes; Dt < 0: the step's dt
static int acct_flux(harm_ctx* c, double Dt) {
  AcctState* a = c->acct;
  int nfaces = 0;
  for (int r = 0; r < ACC_NREG; r++)
    for (int d = 0; d < 4; d++) nfaces = std::max(nfaces, a->R.off[r][d] + a->R.len[r][d]);
  if (nfaces > 0)
    hipLaunchKernelGGL(k_acct_faceterm, dim3((nfaces + 63) / 64), dim3(64), 0, c->stream, c->kp, a->R, a->faces,
                       nfaces, c->arr[HARM_P], c->arr[HARM_F1], c->arr[HARM_F2], c->q1, c->q2, c->gF, c->gC,
                       a->scratch);
  const int nsum = ACC_NREG * 4 * (8 + 2 * ACC_NFL * 8);
  hipLaunchKernelGGL(k_acct_facesum, dim3((nsum + 63) / 64), dim3(64), 0, c->stream, a->R, a->scratch, a->acc, c->ds,
                     Dt, Dt < 0 ? 1 : 0);
  const dim3 b(64, 4);
  hipLaunchKernelGGL(k_acct_frterm, grid2(c->kp.n2, c->kp.n1, b), b, 0, c->stream, c->kp, a->R, c->arr[HARM_F1],
                     c->q1, c->q2, c->gF, c->gC, a->scratch);
  hipLaunchKernelGGL(k_acct_frsum, dim3((c->kp.n1 * 8 + 63) / 64), dim3(64), 0, c->stream, c->kp, a->scratch,
                     a->frdot);
  HIPCHK(hipGetLastError());
  return 0;
}

static void invalidate_graph(harm_ctx* c) { drop_graphs(c); }

extern "C" int harm_acct_setup(harm_ctx* c, const int* enerpos, const int* doflux, double dx3) {
  if (!c || !enerpos || !doflux) return seterr("null argument");
  if (!(dx3 > 0.0)) return seterr("dx3 must be positive");
  const int n1 = c->prm.n1, n2 = c->prm.n2;
  AcctRegions R;
  memset(&R, 0, sizeof(R));
  std::vector<int> faces;
  for (int r = 0; r < ACC_NREG; r++)
    for (int d = 0; d < 4; d++) {
      R.ep[r][d] = enerpos[r * 4 + d];
      R.df[r][d] = doflux[r * 4 + d];
    }
  for (int r = 0; r < ACC_NREG; r++)
    for (int d = 0; d < 4; d++) {
      R.off[r][d] = (int)(faces.size() / 2);
      if (R.df[r][d] < 0) continue;
      const bool x1 = (d == HARM_X1UP || d == HARM_X1DN);
      const int fixed = R.df[r][d];
      const int a0 = x1 ? R.ep[r][HARM_X2DN] : R.ep[r][HARM_X1DN];
      const int a1 = x1 ? R.ep[r][HARM_X2UP] : R.ep[r][HARM_X1UP];
      if (fixed > (x1 ? n1 : n2)) return seterr("region %d dir %d: flux face %d off the grid", r, d, fixed);
      if (a0 <= a1 && (a0 < 0 || a1 > (x1 ? n2 : n1) - 1))
        return seterr("region %d dir %d: faces %d..%d off the grid", r, d, a0, a1);
      for (int m = a0; m <= a1; m++) {
        faces.push_back(x1 ? fixed : m);
        faces.push_back(x1 ? m : fixed);
      }
      R.len[r][d] = (int)(faces.size() / 2) - R.off[r][d];
    }
  R.dx3 = dx3;
  R.dVF = (c->prm.dx1 * c->prm.dx2) * dx3;  // set_grid.c:17-19
  if (acct_alloc(c)) return 1;
  AcctState* a = c->acct;
  if (faces.size() / 2 > (size_t)ACC_NREG * 2 * (n1 + n2 + 2)) return seterr("too many boundary faces");
  a->R = R;
  if (!faces.empty())
    HIPCHK(hipMemcpyAsync(a->faces, faces.data(), faces.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(a->acc, 0, ACC_D * sizeof(double), c->stream));
  HIPCHK(hipMemsetAsync(a->frdot, 0, (size_t)n1 * 8 * sizeof(double), c->stream));
  HIPCHK(hipMemsetAsync(a->lumvsr, 0, (size_t)n1 * sizeof(double), c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->kp.dVF = R.dVF;
  a->tree = false;  // the reference's order until harm_acct_set_order says otherwise
  acct_set_on(c, true);
  invalidate_graph(c);
  return 0;
}

extern "C" int harm_acct_set_order(harm_ctx* c, int order) {
  if (!c) return seterr("null ctx");
  if (!c->acct) return seterr("accounting was never set up (harm_acct_setup)");
  if (order != HARM_ACCT_ORDERED && order != HARM_ACCT_TREE) return seterr("bad accounting order %d", order);
  c->acct->tree = (order == HARM_ACCT_TREE);
  invalidate_graph(c);
  return 0;
}

extern "C" int harm_acct_off(harm_ctx* c) {
  if (!c) return seterr("null ctx");
  acct_set_on(c, false);
  invalidate_graph(c);
  return 0;
}

extern "C" int harm_acct_read(harm_ctx* c, harm_ener_region* regions, double* jet2, double* frdot, double* lumvsr) {
  if (!c) return seterr("null ctx");
  if (!c->acct) return seterr("accounting was never set up (harm_acct_setup)");
  AcctState* a = c->acct;
  std::vector<double> h(ACC_D);
  HIPCHK(hipMemcpyAsync(h.data(), a->acc, ACC_D * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (frdot)
    HIPCHK(hipMemcpyAsync(frdot, a->frdot, (size_t)c->prm.n1 * 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (lumvsr)
    HIPCHK(hipMemcpyAsync(lumvsr, a->lumvsr, (size_t)c->prm.n1 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (regions) memcpy(regions, h.data(), sizeof(harm_ener_region) * ACC_NREG);
  if (jet2) memcpy(jet2, h.data() + OFF_JET2, 4 * ACC_NFL * 8 * sizeof(double));
  return 0;
}

extern "C" int harm_diag_flux(harm_ctx* c, double Dt) {
  if (need_geom(c)) return 1;
  if (!acct_on(c)) return seterr("accounting is off (harm_acct_setup)");
  if (!(Dt >= 0.0)) return seterr("Dt must be >= 0");
  if (acct_flux(c, Dt)) return 1;
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int harm_divbmaxavg(harm_ctx* c, int sel, double* divbmax, double* divbavg, int* imax, int* jmax) {
  if (need_geom(c)) return 1;
  if (sel < 0 || sel >= HARM_NARRAYS) return seterr("bad array selector %d", sel);
  if (acct_alloc(c)) return 1;
  AcctState* a = c->acct;
  const dim3 b(64, 4);
  hipLaunchKernelGGL(k_divb, grid2(c->kp.n2 + 4, c->kp.n1 + 4, b), b, 0, c->stream, c->kp, c->arr[sel], c->gC,
                     a->scratch);
  hipLaunchKernelGGL(k_divb_reduce, dim3(1), dim3(256), 0, c->stream, c->kp, a->scratch, a->out);
  HIPCHK(hipGetLastError());
  double r[4];
  HIPCHK(hipMemcpyAsync(r, a->out, sizeof(r), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (divbmax) *divbmax = r[0];
  if (divbavg) *divbavg = r[1];
  if (imax) *imax = (int)r[2];
  if (jmax) *jmax = (int)r[3];
  return 0;
}

extern "C" int harm_constotal(harm_ctx* c, int sel, int region, double* vars) {
  if (need_geom(c)) return 1;
  if (!vars) return seterr("null argument");
  if (sel < 0 || sel >= HARM_NARRAYS) return seterr("bad array selector %d", sel);
  if (region < 0 || region >= ACC_NREG) return seterr("region must be 0..%d", ACC_NREG - 1);
  if (!c->acct || !c->acct->R.dx3) return seterr("regions not set (harm_acct_setup)");
  AcctState* a = c->acct;
  const dim3 b(64, 4);
  hipLaunchKernelGGL(k_cons_term, grid2(c->kp.n2, c->kp.n1, b), b, 0, c->stream, c->kp, a->R, c->arr[sel], c->gC,
                     a->scratch);
  hipLaunchKernelGGL(k_cons_sum, dim3(1), dim3(256), 0, c->stream, c->kp, a->R, region, a->scratch, a->out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(vars, a->out, 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}
