/*
 * ref_harness.c -- driver that runs the COMPILED REFERENCE (HARM's own
 * step_ch() and everything below it) on a given state, to produce golden
 * vectors for the oracle.  Test infrastructure only; built by
 * oracle/ref/Makefile into oracle/_ref/ from the reference sources in place.
 *
 * It replaces the reference's main.c/init() (main.c:12-90, initbase.c:297):
 * instead of a problem initializer it reads the primitive state from a file,
 * then calls bound_prim() and step_ch() nsteps times, exactly as main.c's loop
 * does, writing the state after every step.
 *
 * For periodic problems this file also supplies bound_prim(): the reference
 * has no single-rank periodic boundary (its periodicity is the MPI halo,
 * boundmpi.c), so the harness provides the problem boundary file, as a HARM
 * user does by swapping bounds.*.c.
 *
 * usage: ref_harness CONFIG   (key value lines, see read_config())
 * With "rdump 1" it also writes the final state's restart files, with
 * "dump 1" its 56-column dump, with "diag 1" the conservation accumulators
 * and dump-time reductions (see the end of main()).
 *
 * Built with HARNESS_REFINIT it runs the reference's own init() instead
 * (initbase.c:297-560 with the compiled problem file's init_grid(),
 * init_global(), initial state and analytic solution), writes what it set up
 * (geometry tables, the state, panalytic, omegafanalytic) and then steps from
 * it: the reference's default problem end to end.
 */
#include "defs.h"
#include "step_ch.h"

#ifdef HARNESS_PERIODIC
/* periodic ghost fill: x1 faces first (interior j), then x2 faces over the
   full x1 range (corners), the order boundmpi.c:91-112 exchanges in */
int bound_prim(int boundstage, FTYPE prim[][N2M][NPR])
{
  int i, j, k;
  for (j = 0; j < N2; j++) {
    for (i = -N1BND; i < 0; i++) PLOOP prim[i][j][k] = prim[i + N1][j][k];
    for (i = N1; i < N1 + N1BND; i++) PLOOP prim[i][j][k] = prim[i - N1][j][k];
  }
  for (i = -N1BND; i < N1 + N1BND; i++) {
    for (j = -N2BND; j < 0; j++) PLOOP prim[i][j][k] = prim[i][j + N2][k];
    for (j = N2; j < N2 + N2BND; j++) PLOOP prim[i][j][k] = prim[i][j - N2][k];
  }
  return (0);
}
#endif

static char cfg_key[64][64];
static char cfg_val[64][256];
static int ncfg = 0;

static void read_config(const char *path)
{
  FILE *f = fopen(path, "r");
  if (!f) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
  while (ncfg < 64 && fscanf(f, "%63s %255s", cfg_key[ncfg], cfg_val[ncfg]) == 2) ncfg++;
  fclose(f);
}

static const char *cfg(const char *k, const char *def)
{
  int n;
  for (n = 0; n < ncfg; n++) if (!strcmp(cfg_key[n], k)) return cfg_val[n];
  return def;
}
static double cfgd(const char *k, double def) { const char *v = cfg(k, NULL); return v ? atof(v) : def; }
static int cfgi(const char *k, int def) { const char *v = cfg(k, NULL); return v ? atoi(v) : def; }

static void readall(const char *path, void *buf, size_t bytes)
{
  FILE *f = fopen(path, "rb");
  if (!f || fread(buf, 1, bytes, f) != bytes) { fprintf(stderr, "cannot read %s (%lu bytes)\n", path, (unsigned long)bytes); exit(3); }
  fclose(f);
}

#ifdef HARNESS_REFINIT
/* The reference's init() (initbase.c:297-560) for a fresh start, in its
   order, with the reference's own functions.  Two things differ: MPI is
   initialised first (pre_init() calls makedirs(), whose MPI_Barrier comes
   before its own init_mpi() and aborts under MPICH, initbase.c:13-54 with
   diag.c:928-940), and the dump files init() writes (dump(9997..9999)) and
   the diagnostics set-up of post_init() (initbase.c:621-660: horizon,
   flux regions, the current's history) are skipped -- they write files or
   diagnostics and do not touch the state. */
static int refinit(char **margv)
{
  extern void filterffde(int i, int j, FTYPE *pr);
  extern int init_vpot2field(SFTYPE A[][N2M], FTYPE pr[][N2M][NPR]);
  extern int init_postfield(FTYPE pr[][N2M][NPR]);
  extern int normalize_densities(FTYPE p[][N2M][NPR]);
  extern int normalize_field(FTYPE p[][N2M][NPR]);
  extern int init_dsandvels(int *whichvel, int *whichcoord, int ii, int jj, FTYPE *pr);
  extern int init_vpot(int ii, int jj, FTYPE *A);
  extern int pre_init_specific_init(void);
  extern int post_init_specific_init(void);
  extern int init_dumps(void);
  int whichvel, whichcoord, i, j, k;
  FTYPE (*A)[N2M];

  ranc(0);
  binaryoutput = TEXTOUTPUT;
  sortedoutput = SORTED;
  init_mpi(3, margv);
  set_arrays();
  init_dumps();
  for (k = 0; k < NDIM; k++) FULLLOOP jcon[i][j][k] = 0.0;
  pre_init_specific_init();
  init_defglobal();
  init_defconsts();
  init_defgrid();
  init_grid();
  set_coord_parms();
  init_global();
  Rhor = rhor_calc(0);
  Risco = rmso_calc(PROGRADERISCO);
  set_grid();
  ZSLOOP(-N1BND, N1 - 1 + N1BND, -N2BND, N2 - 1 + N2BND) {
    if (init_dsandvels(&whichvel, &whichcoord, i, j, p[i][j]) > 0) return 1;
    if (bl2met2metp2v(whichvel, whichcoord, p[i][j], i, j) >= 1) return 1;
  }
  normalize_densities(p);
  A = emf;
  ZSLOOP(-N1BND, N1 - 1 + N1BND, -N2BND, N2 - 1 + N2BND) A[i][j] = 0.;
  ZSLOOP(-N1BND, N1 - 1 + N1BND, -N2BND, N2 - 1 + N2BND) init_vpot(i, j, &A[i][j]);
  init_vpot2field(A, p);
  normalize_field(p);
  init_postfield(p);
  ZSLOOP(-N1BND, N1 - 1 + N1BND, -N2BND, N2 - 1 + N2BND) filterffde(i, j, p[i][j]);
  normalize_field(p);
  init_postfield(p);
  ZSLOOP(-N1BND, N1 - 1 + N1BND, -N2BND, N2 - 1 + N2BND) filterffde(i, j, p[i][j]);
  ZSLOOP(-N1BND, N1 - 1 + N1BND, -N2BND, N2 - 1 + N2BND) PLOOP panalytic[i][j][k] = p[i][j][k];
  if (bound_prim(STAGEM1, panalytic) >= 1) return 1;
#if(FIXUPAFTERINIT)
  if (fixup(STAGEM1, p, 0) >= 1) return 1;
#endif
  if (bound_prim(STAGEM1, p) >= 1) return 1;
  if (pre_fixup(STAGEM1, p) >= 1) return 1;
  post_init_specific_init();
  return failed ? 1 : 0;
}
#endif

int main(int argc, char *argv[])
{
  char *margv[4];
  int s, nsteps, i, j, k;
  FILE *out;
  const char *geomsrc;
  double hdr[4];

  if (argc < 2) { fprintf(stderr, "usage: %s CONFIG\n", argv[0]); return 2; }
  read_config(argv[1]);

  margv[0] = argv[0]; margv[1] = "1"; margv[2] = "1"; margv[3] = NULL;
#ifdef HARNESS_REFINIT
  {
    FILE *f;
    if (refinit(margv) >= 1) { fprintf(stderr, "init failed\n"); return 4; }
    /* x2 boundary types other than init()'s, from the first step on (init()'s
       own bound_prim() calls have run with its types) */
    if (cfg("bc2", NULL)) BCtype[X2UP] = cfgi("bc2", BCtype[X2UP]);
    if (cfg("bc3", NULL)) BCtype[X2DN] = cfgi("bc3", BCtype[X2DN]);
    if (cfg("tables_out", NULL)) {
      f = fopen(cfg("tables_out", ""), "wb");
      fwrite(&a_gcov[0][0][0][0][0], sizeof(FTYPE), (size_t)N1M * N2M * NPG * NDIM * NDIM, f);
      fwrite(&a_gcon[0][0][0][0][0], sizeof(FTYPE), (size_t)N1M * N2M * NPG * NDIM * NDIM, f);
      fwrite(&a_gdet[0][0][0], sizeof(FTYPE), (size_t)N1M * N2M * NPG, f);
      fwrite(&a_conn[0][0][0][0][0], sizeof(FTYPE), (size_t)N1M * N2M * NDIM * NDIM * NDIM, f);
      fwrite(&Rin, sizeof(FTYPE), 1, f);
      fwrite(&startx[1], sizeof(FTYPE), 2, f);
      fwrite(&dx[1], sizeof(FTYPE), 2, f);
      fclose(f);
    }
    f = fopen(cfg("init_out", "init.bin"), "wb");
    fwrite(&a_p[0][0][0], sizeof(FTYPE), (size_t)N1M * N2M * NPR, f);
    fwrite(&a_panalytic[0][0][0], sizeof(FTYPE), (size_t)N1M * N2M * NPR, f);
    fwrite(&a_omegafanalytic[0][0][0], sizeof(FTYPE), (size_t)N1M * N2M * NPR, f);
    fclose(f);
    cour = cfgd("cour", cour);
    /* the run-time parameters init() chose */
    f = fopen(cfg("params_out", "params.txt"), "w");
    fprintf(f, "lim %d\nfluxmethod %d\ntimeorder %d\nFLUXB %d\ncour %.17g\ngam %.17g\n", lim, fluxmethod,
            TIMEORDER, FLUXB, cour, gam);
    fprintf(f, "bc0 %d\nbc1 %d\nbc2 %d\nbc3 %d\n", BCtype[0], BCtype[1], BCtype[2], BCtype[3]);
    fprintf(f, "Rin %.17g\nRout %.17g\nZin %.17g\nZout %.17g\ndefcoord %d\n", Rin, Rout, Zin, Zout, defcoord);
    fprintf(f, "GAMMAMAX %.17g\nGAMMAFAIL %.17g\nSAFE %.17g\ndefcon %.17g\ntf %.17g\ndt %.17g\n", GAMMAMAX,
            GAMMAFAIL, SAFE, defcon, tf, dt);
    fclose(f);
  }
#else
  /* output modes as init() sets them before init_mpi() (initbase.c:42-50,
     USEROMIO 0); init_mpi() picks the file-combine method from them */
  binaryoutput = TEXTOUTPUT;
  sortedoutput = SORTED;
  init_mpi(3, margv);
  set_arrays();
  init_defglobal();
  init_defconsts();
  init_defgrid();

  /* runtime globals (initbase.c:119-243 defaults, overridden per problem) */
  a = cfgd("a", 0.0);
  R0 = cfgd("R0", 0.0);
  Rout = cfgd("Rout", 40.0);
  hslope = cfgd("hslope", 0.3);
  defcoord = 0;
  gam = cfgd("gam", 4. / 3.);
  cour = cfgd("cour", 0.9);
  lim = cfgi("lim", PARA);
  fluxmethod = cfgi("fluxmethod", HLLFLUX);
  TIMEORDER = cfgi("timeorder", 4);
  FLUXB = FLUXCTTOTH;
  UTOPRIMVERSION = UTOPRIM2DFINAL;
  BCtype[X1UP] = cfgi("bc0", OUTFLOW);
  BCtype[X1DN] = cfgi("bc1", OUTFLOW);
  BCtype[X2UP] = cfgi("bc2", POLARAXIS);
  BCtype[X2DN] = cfgi("bc3", POLARAXIS);
  cooling = 0;
  DODIAGS = 0;
  if (cfg("gammamax", NULL)) {  /* initbase.c:218-219 sets both from GAMMAMAX */
    GAMMAMAX = cfgd("gammamax", 25.0);
    GAMMAFAIL = 100.0 * GAMMAMAX;
  }

  geomsrc = cfg("geom", "ref");
  if (!strcmp(geomsrc, "ref")) {
    /* the reference's own tables: init_grid()+set_grid() as init() does */
    Rhor = rhor_calc(0);
    Rin = setRin(setihor());
    set_coord_parms();
    set_grid();
    if (cfg("tables_out", NULL)) {
      FILE *f = fopen(cfg("tables_out", ""), "wb");
      fwrite(&a_gcov[0][0][0][0][0], sizeof(FTYPE), (size_t)N1M * N2M * NPG * NDIM * NDIM, f);
      fwrite(&a_gcon[0][0][0][0][0], sizeof(FTYPE), (size_t)N1M * N2M * NPG * NDIM * NDIM, f);
      fwrite(&a_gdet[0][0][0], sizeof(FTYPE), (size_t)N1M * N2M * NPG, f);
      fwrite(&a_conn[0][0][0][0][0], sizeof(FTYPE), (size_t)N1M * N2M * NDIM * NDIM * NDIM, f);
      fwrite(&Rin, sizeof(FTYPE), 1, f);
      fwrite(&startx[1], sizeof(FTYPE), 2, f);
      fwrite(&dx[1], sizeof(FTYPE), 2, f);
      fclose(f);
    }
  } else {
    /* tables from a file (reference layout), grid spacing from the config */
    Rin = cfgd("Rin", 1.0);
    set_coord_parms();
    set_points();
    startx[1] = cfgd("startx1", 0.0);
    startx[2] = cfgd("startx2", 0.0);
    dx[1] = cfgd("dx1", 1.0 / N1);
    dx[2] = cfgd("dx2", 1.0 / N2);
    readall(geomsrc, &a_gcov[0][0][0][0][0], sizeof(FTYPE) * (size_t)N1M * N2M * NPG * NDIM * NDIM);
    readall(cfg("geom_gcon", ""), &a_gcon[0][0][0][0][0], sizeof(FTYPE) * (size_t)N1M * N2M * NPG * NDIM * NDIM);
    readall(cfg("geom_gdet", ""), &a_gdet[0][0][0], sizeof(FTYPE) * (size_t)N1M * N2M * NPG);
    readall(cfg("geom_conn", ""), &a_conn[0][0][0][0][0], sizeof(FTYPE) * (size_t)N1M * N2M * NDIM * NDIM * NDIM);
    for (i = 0; i < N1M; i++) for (j = 0; j < N2M; j++) for (k = 0; k < NPG; k++) a_eomfunc[i][j][k] = a_gdet[i][j][k];
    for (i = 0; i < N1M; i++) for (j = 0; j < N2M; j++) for (k = 0; k < NDIM; k++) a_conn2[i][j][k] = 0.0;
  }

  readall(cfg("prim", ""), &a_p[0][0][0], sizeof(FTYPE) * (size_t)N1M * N2M * NPR);
  bound_prim(STAGEM1, p);
#endif

  /* conservation accounting (every step runs diag_flux, step_ch.c:160, and
     diag_source, advance.c:140, whatever this option says): with "diag 1"
     the energy regions and volume factors are set up as init() does --
     dV and dVF as set_grid.c:17-19 forms them, h_over_r_jet as the problem
     files set it (init.fishmon.c:21-23), setflux()/setjetflux()
     (initbase.c:642-643), the luminosity accumulator zeroed
     (initbase.c:89) -- and the accumulators are written after the steps */
  if (cfgi("diag", 0)) {
    extern int setflux(void);
    extern int setjetflux(void);
    int ii;
    dV = dx[1] * dx[2];
    dVF = dV * dx[3];
    h_over_r_jet = cfgd("h_over_r_jet", 0.4);
    setflux();
    if (DOJETDIAG) setjetflux();
    for (ii = 0; ii < ncpux1 * N1; ii++) lumvsr[ii] = 0.0;
  }

  t = cfgd("t0", 0.0);
  dt = cfgd("dt", 1e-4);
  tf = cfgd("tf", 1e30);
  nsteps = cfgi("nsteps", 1);

  out = fopen(cfg("out", "ref_out.bin"), "wb");
  hdr[0] = nsteps; hdr[1] = N1; hdr[2] = N2; hdr[3] = 0;
  fwrite(hdr, sizeof(double), 4, out);
  for (s = 0; s < nsteps; s++) {
    double rec[4];
    int nf = 0;
    nstroke = 0;
    failed = 0;
    step_ch();
    ZLOOP if (pflag[i][j][FLAGUTOPRIMFAIL]) nf++;
    rec[0] = t; rec[1] = dt; rec[2] = (double)nstroke; rec[3] = (double)(failed + 2 * nf);
    fwrite(rec, sizeof(double), 4, out);
    fwrite(&a_p[0][0][0], sizeof(FTYPE), (size_t)N1M * N2M * NPR, out);
  }
  fclose(out);
  /* the accumulators after the steps, then the reference's own dump-time
     reductions of the final state: divbmaxavg() (dump_ener.c:449) and
     constotal() of each region (dump_ener.c:562) */
  if (cfgi("diag", 0)) {
    extern void divbmaxavg(FTYPE p[][N2M][NPR], SFTYPE *ptrdivbmax, SFTYPE *ptrdivbavg);
    extern int constotal(int enerregion, SFTYPE *vars);
    SFTYPE divbmax, divbavg, vars[NPR];
    double d;
    int r, dir;
    FILE *f = fopen(cfg("diag_out", "diag.bin"), "wb");
    for (r = 0; r < NUMENERREGIONS; r++) for (dir = 0; dir < COMPDIM * 2; dir++) {
      d = enerposreg[r][dir]; fwrite(&d, sizeof(double), 1, f);
    }
    for (r = 0; r < NUMENERREGIONS; r++) for (dir = 0; dir < COMPDIM * 2; dir++) {
      d = dofluxreg[r][dir]; fwrite(&d, sizeof(double), 1, f);
    }
    fwrite(pdotreg, sizeof(pdotreg), 1, f);
    fwrite(pdottermsreg, sizeof(pdottermsreg), 1, f);
    fwrite(pcumreg, sizeof(pcumreg), 1, f);
    fwrite(fladdreg, sizeof(fladdreg), 1, f);
    fwrite(fladdtermsreg, sizeof(fladdtermsreg), 1, f);
    fwrite(sourceaddreg, sizeof(sourceaddreg), 1, f);
    fwrite(sourceaddtermsreg, sizeof(sourceaddtermsreg), 1, f);
    fwrite(pdottermsjet2, sizeof(pdottermsjet2), 1, f);
    fwrite(frdot, sizeof(frdot), 1, f);
    fwrite(lumvsr, sizeof(SFTYPE), (size_t)ncpux1 * N1, f);
    divbmaxavg(p, &divbmax, &divbavg);
    fwrite(&divbmax, sizeof(double), 1, f);
    fwrite(&divbavg, sizeof(double), 1, f);
    for (r = 0; r < NUMENERREGIONS; r++) {
      constotal(r, vars);
      fwrite(vars, sizeof(double), NPR, f);
    }
    d = dVF; fwrite(&d, sizeof(double), 1, f);
    fclose(f);
  }
  /* optional: the restart file of the final state, written by the
     reference's own restart_write() (restart.c:182, through dump_gen),
     as text (the default binaryoutput, initbase.c:48) into dumps/rdump-0 and
     as binary into dumps/rdump-1.bin.  init_dumps() (initbase.c:989) sets the
     column counts that init() would; the luminosity accumulator the header
     carries (malloc'd in set_arrays.c:145, filled only by diagnostics this
     harness does not run) is set to zero. */
  if (cfgi("rdump", 0)) {
    extern int init_dumps(void);
    extern int restart_write(long dump_cnt);
    int ii;
    init_dumps();
    for (ii = 0; ii < ncpux1 * N1; ii++) lumvsr_tot[ii] = 0.0;
    binaryoutput = TEXTOUTPUT;
    restart_write(0);
    binaryoutput = BINARYOUTPUT;
    restart_write(1);
  }
  /* optional: the 56-column dump of the final state, written by the
     reference's own dump() (dump.c:23-191, through dump_gen) as text into
     dumps/dump0000 and as binary into dumps/dump0001 (with dump_gen's
     binary suffix); the current and Faraday columns are what step_ch()
     computed every step (step_ch.c:143-153) */
  if (cfgi("dump", 0)) {
    extern int init_dumps(void);
    extern int dump(long dump_cnt);
    init_dumps();
    binaryoutput = TEXTOUTPUT;
    dump(0);
    binaryoutput = BINARYOUTPUT;
    dump(1);
  }
#if(USEMPI)
  MPI_Finalize();
#endif
  return 0;
}
