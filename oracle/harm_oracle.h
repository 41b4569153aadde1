/*
 * harm_oracle.h -- CPU restatement of HARM's conservative timestep.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP path is checked
 * against: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product (harm_ramt_amd/) never links or calls it.
 *
 * Layout follows the reference exactly (defs.h:4-83, set_arrays.c:116-139):
 *   primitive-shaped arrays  a[N1M][N2M][NPR]   N?M = N? + 6 (3 ghosts/side)
 *   gcov, gcon               a[N1M][N2M][NPG][4][4]   (NPG: FACE1,FACE2,CORN,CENT)
 *   gdet                     a[N1M][N2M][NPG]
 *   conn                     a[N1M][N2M][4][4][4]     (CENT only)
 *   floors                   a[N1M][N2M][2]           (RHO,UU floor at CENT;
 *                                                      RHOMIN r^-2.7, UUMIN r^-2.7,
 *                                                      fixup.c:1025-1028 precomputed)
 * Index (i,j) with i in [-3, N1+2], j in [-3, N2+2], stored at (i+3, j+3).
 */
#ifndef HARM_ORACLE_H
#define HARM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* reference mnemonics (step_ch.h, global.h) */
enum { OR_DONOR = 0, OR_VANL = 1, OR_MINM = 2, OR_MC = 3, OR_PARA = 4,
       OR_CSSLOPE = 90, OR_NLIMCENT = 101, OR_NLIMUP = 102, OR_NLIMDOWN = 103 };
enum { OR_LAXF = 0, OR_HLL = 1 };
enum { OR_FLUXCTHLL = 0, OR_FLUXCTTOTH = 1, OR_FLUXCD = 2 };
enum { OR_OUTFLOW = 0, OR_FIXED = 3, OR_POLARAXIS = 4, OR_PERIODIC = 100 };
enum { OR_X1UP = 0, OR_X1DN = 1, OR_X2UP = 2, OR_X2DN = 3 };
enum { OR_GRMHD = 0, OR_FFDE = 1 };

typedef struct {
  int n1, n2;
  int eomtype, lim, fluxmethod, fluxb, timeorder;
  int bc[4];
  int fixupflux;        /* FIXUPFLUX (step_ch.h:82) */
  int poledeath;        /* POLEDEATH (bounds.fishmon.c:328) */
  int checksolution;    /* CHECKSOLUTION (global.h:259) */
  int eno;              /* DOENO (global.h:347): 0 off, 1 convert U, 2 convert dU */
  double gam, cour, safe, defcon, gammamax, gammafail;
  double rhominlimit, uuminlimit;
  double dx1, dx2;
  int boundsfile;       /* 0 bounds.fishmon.c, 1 bounds.c (FIXED from panalytic, bound_disk) */
  int extrap;           /* EXTRAP of the outflow fill (bounds.fishmon.c:8, bounds.c:5) */
  int disk_first;       /* bound_disk's static firsttime (bounds.c:373) */
  int rescaletype;      /* rescaletype (initbase.c:200) of EXTRAP 2's rescale() */
  int nthreads;         /* OpenMP threads, 0 = default */
} oracle_params;

typedef struct oracle_ctx oracle_ctx;

oracle_ctx *or_create(const oracle_params *prm);
void or_destroy(oracle_ctx *c);
/* geometry tables in the reference layout (copied) */
void or_set_geometry(oracle_ctx *c, const double *gcov, const double *gcon,
                     const double *gdet, const double *conn, const double *floors);

/* array selectors */
enum { OR_P = 0, OR_PK0 = 1, OR_PK1 = 2, OR_F1 = 3, OR_F2 = 4, OR_ULAST = 5,
       OR_UNEW = 6, OR_EMF = 7, OR_PLEFT = 8, OR_PRIGHT = 9, OR_DQ1 = 10, OR_DQ2 = 11 };
double *or_array(oracle_ctx *c, int sel);
int *or_pflag(oracle_ctx *c);
void or_set_time(oracle_ctx *c, double t, double dt, double tf);
void or_get_time(const oracle_ctx *c, double *t, double *dt);
long or_nstroke(const oracle_ctx *c);
int or_failed(const oracle_ctx *c);

/* the path */
int or_fluxcalc(oracle_ctx *c, int pr_sel, int dir, double *ndt);
void or_fix_flux(oracle_ctx *c);
int or_flux_ct(oracle_ctx *c);
int or_advance(oracle_ctx *c, int pi_sel, int pb_sel, const double *CUf,
               int pf_sel, const double *Cunew, int timeorder,
               int numtimeorders, double *ndt);
int or_bound_prim(oracle_ctx *c, int sel);
/* panalytic and omegafanalytic ([N1M][N2M][NPR], omegaf may be NULL): the
   analytic solution FIXED boundaries and bound_disk read (copied) */
/* EXTRAP 2: r of every zone centre [N1M][N2M] (bl_coord of coord(i,j,CENT)) */
void or_set_radius(oracle_ctx *c, const double *r);
void or_set_analytic(oracle_ctx *c, const double *panalytic, const double *omegaf);
int or_post_fixup(oracle_ctx *c, int sel);
int or_step(oracle_ctx *c);

/* single-zone entry points (for per-function tests) */
int or_utoprim_2d_final(oracle_ctx *c, int i, int j, const double *Ugeomfree,
                        double *pr, int *pflag);
int or_utoprimgen(oracle_ctx *c, int i, int j, const double *U, double *pr,
                  int *pflag);
int or_primtoU(oracle_ctx *c, int i, int j, int loc, const double *pr, double *U);
int or_primtoflux(oracle_ctx *c, int i, int j, int loc, const double *pr, int dir,
                  double *F);
int or_vchar(oracle_ctx *c, int i, int j, int loc, const double *pr, int dir,
             double *vmax, double *vmin);
int or_source(oracle_ctx *c, int i, int j, const double *pr, double *dU);
void or_slope_lim(int lim, int k, double yll, double yl, double yc, double yr,
                  double yrr, double *dq, double *left, double *right);
/* ENO point <-> average conversion of one value along a line of n values
   (stride 1): type 0 average-to-centre, 1 centre-to-average; returns the
   converted value at position pos and the stencil's order and shift */
double or_eno_line(int type, const double *line, int n, int pos, int *order, int *shift);
/* the conversion weight of a stencil of `order` cells for the zone at
   position s of the stencil, cell d (0 when there is no such table) */
double or_eno_weight(int type, int order, int s, int d);

/* ---- conservation accounting (diag.c, fixup.c, dump_ener.c) ----
   Energy regions are zone boxes enerpos[r][X1UP,X1DN,X2UP,X2DN] with the
   boundary faces doflux[r][dir] (negative: no flux through that side), as
   setflux()/setjetflux() set them (initbase.c:736-980); dx3 is dx[3]
   (2 pi in the reference's 2D coordinates, coord.c:728). */
#define OR_NREG 3
#define OR_NUMFLUXTERMS 7
#define OR_NUMSOURCES 7
#define OR_NUMFAILFLOORFLAGS 9
typedef struct {
  int enerpos[OR_NREG][4], doflux[OR_NREG][4];
  double dVF;
  /* the accumulators of defs.h:218-239, per region */
  double pdot[OR_NREG][4][8], pdotterms[OR_NREG][4][OR_NUMFLUXTERMS][8], pcum[OR_NREG][4][8];
  double fladd[OR_NREG][8], fladdterms[OR_NREG][OR_NUMFAILFLOORFLAGS][8];
  double sourceadd[OR_NREG][8], sourceaddterms[OR_NREG][OR_NUMSOURCES][8];
  double pdottermsjet2[4][OR_NUMFLUXTERMS][8];
} or_acct;
/* set the regions, zero every accumulator and switch accounting on (the
   steps then run their zone loops in the reference's order, one thread) */
void or_acct_setup(oracle_ctx *c, const int *enerpos, const int *doflux, double dx3);
void or_acct_off(oracle_ctx *c);
/* Force-free fixup1zone's COUNTFLOORACT event measured from pr0 = 0 instead
   of the incoming state: the reference declares pr0 and didchangeprim without
   setting them under EOMFFDE (fixup.c:283-313), and with its locals
   zero-initialised (oracle/ref ffdezi build) that is what it books. */
void or_acct_ffde_pr0_zero(oracle_ctx *c, int on);
/* the accumulators, frdot[n1][8] and lumvsr[n1] */
void or_acct_get(const oracle_ctx *c, or_acct *out, double *frdot, double *lumvsr);
/* diag_flux() of the current state and fluxes with time step Dt (the step
   calls it itself when accounting is on) */
void or_diag_flux(oracle_ctx *c, double Dt);
/* dump-time reductions of array sel: divbmaxavg() and constotal() */
void or_divbmaxavg(const oracle_ctx *c, int sel, double *divbmax, double *divbavg, int *imax, int *jmax);
void or_constotal(const oracle_ctx *c, int sel, int region, double *vars);

/* the current J^mu and Faraday F^{mu nu} the reference's step computes
   (WHICHCURRENTCALC 1, phys.c:1124-1460): on, every step updates them
   (zeroed, with current_calc_1's call count, when switched on) */
void or_current_enable(oracle_ctx *c, int on);
void or_current(oracle_ctx *c);
void or_current_get(const oracle_ctx *c, double *jcon, double *fcon);  /* [nz][4], [nz][6] */
/* dump_content (dump.c:82-191) after its six coordinate columns: 50 values
   per interior zone, out[n1][n2][50] */
void or_dump_content(const oracle_ctx *c, double *out);

#ifdef __cplusplus
}
#endif
#endif
