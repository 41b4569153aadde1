/*
 * harm_amd.h -- C-ABI of the MI355X (gfx950) HARM conservative timestep.
 *
 * Drop-in boundary for the reference's step_ch()/advance()/fluxcalc()/
 * flux_ct()/Utoprimgen()/bound_prim() call surface (global.h:1383-1391,
 * 1525-1545).  The reference keeps every array and parameter in globals
 * (defs.h:4-83, 168-214); here they live in a context handle so the API is
 * re-entrant, and the arrays live in GPU memory (HBM).
 *
 * Array layout is the reference's (set_arrays.c:116-139):
 *   primitive-shaped      double a[N1M][N2M][NPR],  N?M = N? + 6, NPR = 8
 *   gcov, gcon            double a[N1M][N2M][NPG][4][4]
 *   gdet                  double a[N1M][N2M][NPG]
 *   conn                  double a[N1M][N2M][4][4][4]     (CENT)
 *   floors                double a[N1M][N2M][2]           (CENT; RHOMIN r^-2.7,
 *                                                          UUMIN r^-2.7: the
 *                                                          set_density_floors()
 *                                                          values, fixup.c:996-1110)
 *   pflag                 int    a[N1M][N2M][NUMPFLAGS=5]
 * Zone (i,j), i in [-3,N1+2], j in [-3,N2+2], is stored at (i+3,j+3).
 *
 * Errors: every entry point returns 0 on success, >0 on error (the reference's
 * MYFUN convention, global.h:1204-1210).  harm_last_error() gives the text.
 * Per-zone inversion failures are not errors: they are reported in pflag
 * (FLAGUTOPRIMFAIL, global.h:546) and counted by harm_get_failures().
 */
#ifndef HARM_AMD_H
#define HARM_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- mnemonics, identical values to the reference ---- */
/* slope limiters (step_ch.h:48-58) */
#define HARM_DONOR 0
#define HARM_VANL 1
#define HARM_MINM 2
#define HARM_MC 3
#define HARM_PARA 4
#define HARM_CSSLOPE 90
#define HARM_NLIMCENT 101
#define HARM_NLIMUP 102
#define HARM_NLIMDOWN 103
/* flux methods (step_ch.h:17-18) */
#define HARM_LAXFFLUX 0
#define HARM_HLLFLUX 1
/* divB methods (step_ch.h:22-26): FLUXCTTOTH and FLUXCTHLL are supported */
#define HARM_FLUXCTHLL 0
#define HARM_FLUXCTTOTH 1
/* equations of motion (global.h:199-200) */
#define HARM_EOMGRMHD 0
#define HARM_EOMFFDE 1
/* boundary types (global.h:616-622); PERIODIC is the single-GPU stand-in for
   the reference's periodic MPI halo (init_mpi.c:305-313) */
#define HARM_OUTFLOW 0
#define HARM_FIXED 3
#define HARM_POLARAXIS 4
#define HARM_PERIODIC 100
/* boundary files (the reference swaps bounds.*.c; harm_params.boundsfile) */
#define HARM_BOUNDS_FISHMON 0  /* bounds.fishmon.c: outflow, polar axis with
                                  POLEDEATH; it has no FIXED branch, so
                                  harm_create refuses FIXED sides with it */
#define HARM_BOUNDS_DISK 1     /* bounds.c (the reference's default file):
                                  FIXED from panalytic, outflow, and
                                  bound_disk() after every fill */
/* boundary slots (global.h:85-88) */
#define HARM_X1UP 0
#define HARM_X1DN 1
#define HARM_X2UP 2
#define HARM_X2DN 3
/* array selectors for the context's device arrays (defs.h names) */
#define HARM_P 0
#define HARM_PK0 1
#define HARM_PK1 2
#define HARM_F1 3
#define HARM_F2 4
#define HARM_ULAST 5
#define HARM_UNEW 6
#define HARM_NARRAYS 7

/* pflag value returned by harm_utoprimgen for an FFDE zone whose drift
   velocity is space-like (phys.ffde.c:309-312); inside a step the reference
   treats it as fatal (advance() returns 1), here it sets the failed flag */
#define HARM_FFDE_SPACELIKE 50

#define HARM_NPR 8
#define HARM_NG 3
#define HARM_NUMPFLAGS 5

/* Parameters that the reference keeps in globals (initbase.c:119-243). */
typedef struct {
  int n1, n2;          /* interior zones, N1 and N2 (global.h:126-127) */
  int eomtype;         /* EOMTYPE */
  int lim;             /* lim */
  int fluxmethod;      /* fluxmethod */
  int fluxb;           /* FLUXB */
  int timeorder;       /* TIMEORDER (1..4) */
  int bc[4];           /* BCtype[X1UP..X2DN] */
  int fixupflux;       /* FIXUPFLUX (step_ch.h:82): fix_flux() on/off */
  int poledeath;       /* POLEDEATH (bounds.fishmon.c:328) */
  int checksolution;   /* CHECKSOLUTION (global.h:259): get_bsqflags before each
                          substage, fixup_checksolution in post_fixup */
  int eno;             /* DOENO (global.h:347): 0 off; 1 advance_eno_u (U
                          converted point -> average -> point around the
                          update, advance.c:298); 2 the flux divergence
                          converted average -> point (advance.c:101-158) */
  double gam;          /* adiabatic index */
  double cour;         /* Courant factor */
  double safe;         /* SAFE: max dt growth per step */
  double defcon;       /* defcon */
  double gammamax;     /* GAMMAMAX */
  double gammafail;    /* GAMMAFAIL */
  double rhominlimit;  /* RHOMINLIMIT */
  double uuminlimit;   /* UUMINLIMIT */
  double dx1, dx2;     /* dx[1], dx[2] */
  int boundsfile;      /* HARM_BOUNDS_*: which bounds.*.c bound_prim() is */
  int extrap;          /* EXTRAP of the outflow fill (bounds.fishmon.c:8,
                          bounds.c:5): 0 copy, 1 gdet / linear extrapolation,
                          2 power laws in r (rescale(), interp.c:559-620;
                          needs harm_set_radius) */
  int disk_first;      /* bound_disk()'s function-static firsttime
                          (bounds.c:373): 1 if its next call is the first of
                          the process (no equatorial symmetrisation) */
  int rescaletype;     /* rescaletype (initbase.c:200, default 1): the
                          powers of EXTRAP 2's rescale() */
} harm_params;

typedef struct harm_ctx harm_ctx;

/* Context lifetime: replaces set_arrays() (set_arrays.c:109) plus the
   globals set by init_defglobal().  Allocates all arrays on the current GPU. */
int harm_create(const harm_params *prm, harm_ctx **out);
int harm_destroy(harm_ctx *ctx);
const char *harm_last_error(void);

/* Run on a caller-owned HIP stream (hipStream_t passed as void*); NULL means
   the context's own stream. */
int harm_set_stream(harm_ctx *ctx, void *hip_stream);

/* Geometry tables in the reference layout (set_grid.c:39-88 output).  Host
   pointers; copied and packed into the device layout once. */
int harm_set_geometry(harm_ctx *ctx, const double *gcov, const double *gcon,
                      const double *gdet, const double *conn,
                      const double *floors);

/* The analytic solution the FIXED boundaries and bound_disk() read
   (initbase.c:519-523: panalytic, the initial state after its own
   bound_prim(); init.c:663: omegafanalytic, the field-line angular velocity
   of each zone in component 0), both [N1M][N2M][NPR] host arrays; omegaf
   may be NULL when bound_disk() is not used.  Copied to the device. */
/* EXTRAP 2: r of every zone centre ([N1M][N2M], bl_coord of coord(i, j,
   CENT), e.g. from harm_zone_coords); the rescale() factors of every zone
   are formed from it on the host (interp.c:559-620, dir 1, PRIMTOINTERP,
   the context's rescaletype) */
int harm_set_radius(harm_ctx *ctx, const double *r);
int harm_set_analytic(harm_ctx *ctx, const double *panalytic, const double *omegaf);

/* Host <-> device copies of a context array (selector HARM_*). */
int harm_set_array(harm_ctx *ctx, int sel, const double *host);
int harm_get_array(harm_ctx *ctx, int sel, double *host);
/* Device pointer of a context array, for zero-copy use (e.g. torch). */
int harm_array_ptr(harm_ctx *ctx, int sel, double **dev);
int harm_get_pflag(harm_ctx *ctx, int *host);

/* Time state: t, dt, tf (defs.h:177-179). */
int harm_set_time(harm_ctx *ctx, double t, double dt, double tf);
int harm_get_time(harm_ctx *ctx, double *t, double *dt);

/* step_ch.c:741 fluxcalc(stage, pr, F, dir, &ndt): fluxes of array pr_sel into
   F1 (dir=1) or F2 (dir=2) and the Courant limit ndt.  Includes the
   reference's per-face limit_gamma, vchar, HLL/LAXF.  The fluxes are left as
   fluxcalc() leaves them: fix_flux() is harm_fix_flux (harm_step and
   harm_advance apply its rules inside their own flux launch). */
int harm_fluxcalc(harm_ctx *ctx, int pr_sel, int dir, double *ndt);
/* fixup.c:1357 fix_flux(pb, F1, F2) on the fluxes the context holds,
   whatever FIXUPFLUX says (the reference's caller decides): polar-axis x1
   ghost-face B2 reflection and x2 polar/equator faces, outflow mass flux. */
int harm_fix_flux(harm_ctx *ctx);
/* fixup.c:25 pre_fixup(stage, pv): get_bsqflags (b^2/rho, b^2/u flags
   read by fixup_checksolution) from array sel; a no-op in effect unless
   checksolution is set. */
int harm_pre_fixup(harm_ctx *ctx, int sel);
/* fixup.c:37 post_fixup(stage, pv, finalstep): bound_pflag,
   fixup_checksolution (when checksolution is set) and fixup_utoprim (repair
   of zones whose inversion failed or that the check failed) on array sel.
   A standalone call is never accounted: the reference books fixup_utoprim's
   repairs with diag_fixup when finalstep is set (fixup.c:908); with
   accounting on, only the post_fixup inside harm_step/harm_step_n does. */
int harm_post_fixup(harm_ctx *ctx, int sel);
/* step_ch.c:1159 flux_ct(stage, pb, F1, F2): Toth's corner EMFs formed from
   the F1/F2 the context holds (CORNGDETVERSION 1) and rewritten into their
   B components.  FLUXCTHLL: the fluxes are left unchanged, as the
   reference's flux_ct() leaves them. */
int harm_flux_ct(harm_ctx *ctx);
/* advance.c:15 advance(stage, pi, ulast, pb, CUf, pf, Cunew, unew, timeorder,
   numtimeorders, &ndt): one RK substage.  pi/pb/pf may alias (advance.c:5-14). */
int harm_advance(harm_ctx *ctx, int pi_sel, int pb_sel, const double CUf[3],
                 int pf_sel, const double Cunew[3], int timeorder,
                 int numtimeorders, double *ndt);
/* bounds.fishmon.c:13 bound_prim(stage, prim) on array sel (device ghost fill:
   outflow/polar/periodic, replacing bound_mpi's halo exchange). */
int harm_bound_prim(harm_ctx *ctx, int sel);
/* step_ch.c:16 step_ch(): one full timestep including the dt update
   (step_ch.c:174-183). */
int harm_step(harm_ctx *ctx);
/* nsteps calls of step_ch() with no host synchronisation in between (the dt
   update stays on the GPU); replays a captured HIP graph of one step. */
int harm_step_n(harm_ctx *ctx, int nsteps);
/* step_ch.c:468 Utoprimgen(EVOLVEUTOPRIM, UEVOLVE, U, geom, pr) for n zones:
   zone[2*k],zone[2*k+1] = (i,j); U[k][8] evolved conserved state; pr[k][8]
   guess in, solution out; pflag[k] out.  Host arrays. */
int harm_utoprimgen(harm_ctx *ctx, int n, const int *zone, const double *U,
                    double *pr, int *pflag);

/* (no reference counterpart) harm_utoprimgen for callers that cannot
   allocate device memory (code under HIP graph capture, tight serving loops,
   harnesses that forbid hipMalloc): the same arguments and results, staged
   through a fixed device array of the library instead of per-call
   allocations.  At most HARM_UTOPRIM_NOALLOC_MAX zones per call (more: error
   1 with a message).  Each GPU holds one such array (about 2.2 MB) until the
   process ends; calls on one GPU share it and are serialised by a lock. */
#define HARM_UTOPRIM_NOALLOC_MAX 16384
int harm_utoprimgen_noalloc(harm_ctx *ctx, int n, const int *zone, const double *U,
                            double *pr, int *pflag);

/* Failure reporting: number of zones whose last inversion set pflag, and the
   reference's global `failed` flag (fail.c:13). */
int harm_get_failures(harm_ctx *ctx, int *nfailzones, int *failed);
/* Newton-iteration counter (the reference's nstroke, main.c:72-78). */
int harm_get_nstroke(harm_ctx *ctx, long long *nstroke);

/* (no reference counterpart) Per-stage timing with HIP events: each kernel
   of a timed stage is launched with its own event pair on the launch stream
   (hipExtLaunchKernelGGL; the events hold the kernel's begin and end, launch
   gaps excluded), a stage's time is the sum over its kernels, its count one
   per stage call (eager launches while on), stage ids:
   HARM_K_FLUX     fluxcalc, both directions in one launch
                   (k_flux_single or k_flux, see harm_set_tuning)
   HARM_K_EMF      corner EMFs + CT fluxes (standalone harm_flux_ct only)
   HARM_K_UPDATE   k_update: advance() per zone incl. CT fluxes and inversion
   HARM_K_BOUNDARY bound_prim + post_fixup + bound_prim (+ dt)
                   (k_boundary, or k_bx1/k_bx2/k_fixlist/k_fixapply) */
#define HARM_K_FLUX 0
#define HARM_K_EMF 2
#define HARM_K_UPDATE 3
#define HARM_K_BOUNDARY 4
#define HARM_K_COUNT 6
int harm_prof_enable(harm_ctx *ctx, int on);
/* (no reference counterpart) Kernel variants, all bit-identical: the first
   harm_step/harm_step_n times them on the live state and keeps the fastest;
   harm_set_tuning forces a choice (and disables the timing). */
#define HARM_FLUX_FACE_64x4 0  /* one thread per face, 64x4 workgroups */
#define HARM_FLUX_FACE_64x2 1  /* one thread per face, 64x2 workgroups */
#define HARM_FLUX_PAIR 2       /* a lane pair per face (left/right state) */
#define HARM_FLUX_SHARED 3     /* a thread per zone and direction: the zone's
                                  edge states reconstructed once, shared with
                                  the face below (lanes along x2, LDS along x1) */
#define HARM_BOUND_KERNELS 0   /* boundary work as six wide launches */
#define HARM_BOUND_BLOCK 1     /* boundary work in one workgroup with barriers
                                  (not with checksolution: wide launches) */
#define HARM_BOUND_KERNELS_SEQCHECK 2  /* wide launches, fixup_checksolution
                                          always by its sequential loop */
#define HARM_BOUND_FIXONE 3    /* wide ghost fills, the failed-zone repair
                                  in one workgroup (compute, barrier, write) */
#define HARM_BOUND_FUSED 4     /* each bound_prim() in one launch, the repair
                                  in one workgroup */
#define HARM_BOUND_INUPDATE 5  /* the first bound_prim() inside the update
                                  (each edge zone writes its ghost images),
                                  the repair and the second bound_prim() (the
                                  images of what the repair and POLEDEATH can
                                  change) in one workgroup; HARM_BOUND_FUSED
                                  where it does not apply: ENO, checksolution,
                                  accounting on, other boundary kinds */
int harm_set_tuning(harm_ctx *ctx, int flux_variant, int bound_variant);
int harm_get_tuning(harm_ctx *ctx, int *flux_variant, int *bound_variant);
int harm_prof_get(harm_ctx *ctx, int kernel, double *total_ms, long long *count);
/* (no reference counterpart) The Newton iteration count of harm_get_nstroke
   and the sum over wavefronts (64 zones) of each one's largest count: a
   wavefront runs its slowest zone's iterations, so 64 wave_max_sum /
   nstroke >= 1 is the iterations issued per iteration needed. */
int harm_get_newton_spread(harm_ctx *ctx, long long *nstroke, long long *wave_max_sum);
int harm_prof_reset(harm_ctx *ctx);

/* ---- batched runs (no reference counterpart: the reference advances one
   grid per process) ----
   A batch advances many independent contexts -- grids of different sizes,
   metrics, spins and parameters, each with its own t, dt and Courant limit --
   with one launch per stage covering the tiles of every run, captured once
   as a HIP graph per batch step.  Each run's results are bit-identical to
   stepping its context alone.  Contexts must have their geometry set and
   share TIMEORDER; they stay owned by the caller (destroy the batch first)
   and their arrays are updated in place (harm_get_array etc. work as usual). */
typedef struct harm_batch harm_batch;
int harm_batch_create(harm_ctx *const *ctxs, int n, harm_batch **out);
int harm_batch_destroy(harm_batch *batch);
/* nsteps calls of step_ch() for every run (first call: variant autotuning) */
int harm_batch_step_n(harm_batch *batch, int nsteps);
/* flux variant HARM_FLUX_FACE_64x4 or HARM_FLUX_FACE_64x2, any boundary variant */
int harm_batch_set_tuning(harm_batch *batch, int flux_variant, int bound_variant);
int harm_batch_get_tuning(harm_batch *batch, int *flux_variant, int *bound_variant);

/* ---- discrete adjoint (no reference counterpart) ----
   Gradient of a functional of the state after nsteps calls of step_ch()
   with respect to the initial state, by a hand-written reverse sweep of the
   GPU step: the time-step sequence is held at the values the forward run
   recorded (or was given); limiters, max/min and floors are differentiated
   on the branch the forward run took; the inversion through the
   implicit-function theorem at its converged solution.  Checkpoints every
   checkpoint_every steps; each segment is recomputed once, keeping every
   RK substage state, and then reversed; the last segment's states are kept
   by the forward run itself (checkpoint_every = nsteps: no recompute, at
   the memory of nsteps x substages + 1 states).  Supported: EOMGRMHD, PERIODIC
   boundaries, FIXUPFLUX off, limiters DONOR/VANL/MINM/MC/NLIM*, HLL or
   LAXF, with or without flux-CT (harm_adj_create reports anything else).
   Typical use: set the state (harm_set_array + harm_bound_prim) and time,
   harm_adj_forward, a seed, harm_adj_backward, harm_adj_get_lambda.  The
   context is left at the final state of the forward run. */
typedef struct harm_adj harm_adj;
int harm_adj_create(harm_ctx *ctx, int nsteps, int checkpoint_every, harm_adj **out);
int harm_adj_destroy(harm_adj *adj);
/* nsteps steps from the context's current state; dts[nsteps] imposes the
   time steps, NULL records the adaptive ones */
int harm_adj_forward(harm_adj *adj, const double *dts);
int harm_adj_get_dts(harm_adj *adj, double *dts);
/* seed lambda_N = dJ/dp_N for J = sum over interior zones of
   0.5 b^2 gdet dx1 dx2 (the magnetic energy); returns J */
int harm_adj_seed_magnetic_energy(harm_adj *adj, double *J);
/* seed lambda_N from a host array [N1M][N2M][NPR] */
int harm_adj_seed(harm_adj *adj, const double *lambda);
int harm_adj_backward(harm_adj *adj);
/* lambda_0 = dJ/dp_0 [N1M][N2M][NPR]: interior zones (ghost zones, which
   bound_prim overwrites, carry zero) */
int harm_adj_get_lambda(harm_adj *adj, double *lambda);
/* non-differentiable points met in the reverse sweep (active limit_gamma,
   clamped wave speeds, singular inversion Jacobian) and failed inversions
   repaired in the recomputation: both 0 means the gradient is exact */
int harm_adj_get_nonsmooth(harm_adj *adj, int *count, int *recompute_failures);
/* Limiter branches (MC: which argument each of its minmods returned; MINM,
   VANL likewise) that differ between two forward runs of the same grid,
   limiter and steps, over every substage's reconstruction input, every
   zone whose slope a face uses and every component: where J(p) is not
   differentiable between the two inputs (a central difference straddling
   such a switch measures a secant, not the derivative).  Both runs need
   checkpoint_every = nsteps (every state kept).  compared: the slopes
   compared. */
int harm_adj_branch_switches(harm_adj *a, harm_adj *b, long long *switches, long long *compared);

/* ---- tangent-linear (forward-mode) step (no reference counterpart) ----
   Advances a context's state together with its derivative with respect to
   one parameter s (for instance the black-hole spin), given the derivative
   of the geometry tables and of the initial state: the step's own physics
   evaluated on value-plus-tangent numbers, so the values are the plain
   step's to the bit and the tangents follow by the chain rule on the
   branches the values take.  The time steps are the values' own (adaptive
   dt), i.e. the derivative of the run with its dt sequence held fixed.
   Supported: EOMFFDE (the explicit force-free inversion), every limiter but
   CSSLOPE, HLL/LAXF, flux-CT, OUTFLOW/POLARAXIS/PERIODIC boundaries,
   FIXUPFLUX on or off; no DOENO, CHECKSOLUTION, accounting or current. */
typedef struct harm_tl harm_tl;
int harm_tl_create(harm_ctx *ctx, harm_tl **out);
int harm_tl_destroy(harm_tl *tl);
/* d/ds of the set_grid() tables, in harm_set_geometry's layout (zero until set) */
int harm_tl_set_geometry_tangent(harm_tl *tl, const double *dgcov, const double *dgcon,
                                 const double *dgdet, const double *dconn);
/* d/ds of the grid spacings dx1, dx2 (zero until set): the flux divergence
   divides by them (the Courant limit, which sets dt, is not differentiated) */
int harm_tl_set_spacing_tangent(harm_tl *tl, double ddx1, double ddx2);
/* d/ds of the context's state HARM_P, [N1M][N2M][NPR] (zero until set) */
int harm_tl_set_state_tangent(harm_tl *tl, const double *dp);
int harm_tl_get_state_tangent(harm_tl *tl, double *dp);
/* nsteps calls of step_ch() on the context's state and its tangent */
int harm_tl_step_n(harm_tl *tl, int nsteps);
/* the outgoing energy flux through zone row i of the context's state,
   L = -2 pi dx2 sum_j gdet T^1_0 (zone-centred, over j in [0,N2)), and dL/ds */
int harm_tl_energy_flux(harm_tl *tl, int i, double *L, double *dL);

/* ---- host-side problem and geometry generation (no GPU use) ---- */

/* Grid/coordinate description (coord.c: set_points, bl_coord, dxdxprim). */
#define HARM_METRIC_MINK 0  /* MINKMETRIC (metric.c:441) */
#define HARM_METRIC_KS 2    /* KSCOORDS (metric.c:540,666) */
#define HARM_METRIC_CYLMINK 4  /* CYLMINKMETRIC (metric.h:7, metric.c:468) */
typedef struct {
  int n1, n2;
  int metric;          /* HARM_METRIC_* */
  int defcoord;        /* 0: MKS, r=R0+exp(X1), th=pi X2+(1-h)/2 sin(2 pi X2) */
  double a;            /* black-hole spin */
  double R0, Rin, Rout, hslope;
  double startx1, startx2, dx1, dx2;
  double rhomin, uumin;  /* floor coefficients RHOMIN, UUMIN */
  double Zin, Zout, npow;  /* defcoord 666: cylindrical height range and the
                              exponential rate (coord.c:134-136, 421-431) */
} harm_grid;

/* MKS grid for a Kerr hole as the reference's init_grid() does for the torus:
   Rin = setRin(setihor()) (coord.c:817-909), set_points() (coord.c:720). */
int harm_grid_mks(harm_grid *g, int n1, int n2, double a, double Rout,
                  double hslope, double R0);
/* Unit-square Cartesian grid with Minkowski metric (X in [0,1)^2). */
int harm_grid_cartesian(harm_grid *g, int n1, int n2);
/* init.c:54-60 init_grid() for CYLMINKMETRIC: cylindrical Minkowski in
   (R, z) on the bi-exponential grid defcoord 666 (coord.c:421-431, 798-804),
   R in [Rin, Rout], z in [Zin, Zout]; the metric's conversion to the grid
   coordinates uses the numerical Jacobian (coord.c:647-678) */
int harm_grid_cylmink(harm_grid *g, int n1, int n2, double Rin, double Rout, double Zin, double Zout);
/* Geometry tables of set_grid() (set_grid.c:6-88) for the grid. */
int harm_geometry(const harm_grid *g, double *gcov, double *gcon, double *gdet,
                  double *conn, double *floors);
/* Fishbone-Moncrief torus initial primitives (init.fishmon.c:115-352):
   rin, rmax, kappa, beta (normalisation (gam-1)u_max/(0.5 b^2_max)),
   randfact (u perturbation amplitude), seed.  Writes p[N1M][N2M][8]. */
int harm_init_fishmon(const harm_grid *g, const double *gcov, const double *gcon,
                      const double *gdet, double gam, double rin, double rmax,
                      double kappa, double beta, double randfact,
                      unsigned long seed, double *p);
/* Relativistic Orszag-Tang vortex on the unit square (Minkowski). */
int harm_init_orszag_tang(const harm_grid *g, double gam, double *p);
/* Force-free split monopole (init.bz.c:146-155, PROBLEMTYPE 0) on a Kerr
   grid: rho=u=0, v=0, B^r = +-B0 gdet(ihor,j)/gdet(i,j) by hemisphere. */
int harm_init_monopole(const harm_grid *g, const double *gdet, double B0, double *p);

/* ---- restart files (restart.c:182-264 restart_write / restart_read) ----
   Header of write_restart_header() (restart.c:554-682; text: the
   WRITERESTARTHEADER line of global.h:929-935), then the 8 primitives of
   every interior zone, j outer and i inner: raw doubles in binary, in text
   "%21.15g" separated by spaces with one zone per line (the reference's
   default, binaryoutput = TEXTOUTPUT, keeps 15 significant digits; binary is
   exact).  Field names are the reference's globals.  The accumulator arrays
   have the reference build's sizes: NUMENERREGIONS 3 (DOJETDIAG 1),
   COMPDIM*2 directions, NUMTSCALES 4, NUMFAILFLOORFLAGS 9, NPRDUMP 8. */
#define HARM_NUMENERREGIONS 3
#define HARM_NUMDIRS 4
#define HARM_NUMTSCALES 4
#define HARM_NUMFAILFLOORFLAGS 9
typedef struct {
  int totalsize[2];                 /* N1, N2 */
  double t, tf;
  long nstep;
  double a, gam, cour;
  double DTd, DTavg, DTener, DTi;
  long DTr;
  double DTdebug;
  long dump_cnt, image_cnt, rdump_cnt, avg_cnt, debug_cnt;
  double dt;
  int lim, TIMEORDER, fluxmethod, FLUXB, UTOPRIMVERSION, failed;
  double R0, Rin, Rout, hslope;
  int defcoord;
  int BCtype[4];                    /* X1UP, X1DN, X2UP, X2DN */
  long realnstep;
  int debugfail, whichrestart, cooling;
  long restartsteps[2];
  int GAMMIEDUMP, GAMMIEIMAGE, GAMMIEENER, DODIAGS, DOENERDIAG, DOGDUMPDIAG, DORDUMPDIAG, DODUMPDIAG,
      DOAVGDIAG, DOIMAGEDIAG, DOAREAMAPDIAG, POSDEFMETRIC, periodicx1, periodicx2, binaryoutput, sortedoutput;
  double defcon, SAFE, RHOMIN, UUMIN, RHOMINLIMIT, UUMINLIMIT, BSQORHOLIMIT, BSQOULIMIT, GAMMAMAX, GAMMADAMP,
      GAMMAFAIL;
  double pcumreg_tot[HARM_NUMENERREGIONS][HARM_NUMDIRS][8];
  double fladdreg_tot[HARM_NUMENERREGIONS][8];
  double sourceaddreg_tot[HARM_NUMENERREGIONS][8];
  double Ureg_init_tot[HARM_NUMENERREGIONS][8];
  double failfloorcount_tot[HARM_NUMTSCALES][HARM_NUMFAILFLOORFLAGS];
  double prMAX[8];
  int rescaletype;
  double *lumvsr_tot;               /* N1 values (DOLUMVSR 1); NULL: written as 0, skipped on read */
} harm_restart_header;

/* Host arrays p[N1+6][N2+6][8]: write the interior of p, or read it into p
   (ghost zones untouched).  Returns 0, 1 bad argument, 2 header grid size
   differs from n1,n2, 3 cannot open, 4 short read/write.  In binary the
   8-byte defcoord field holds the int and 4 zero bytes (the reference writes
   the 4 bytes that follow the int in its memory there). */
int harm_rdump_write_host(const char *path, int binary, const harm_restart_header *h, int n1, int n2,
                          const double *p);
int harm_rdump_read_host(const char *path, int binary, harm_restart_header *h, int n1, int n2, double *p);
/* The same on a context's array HARM_P (restart_write / restart_read).  As
   in restart_init() (restart.c:18-100), call harm_bound_prim after a read. */
int harm_rdump_write(harm_ctx *ctx, const char *path, int binary, const harm_restart_header *h);
int harm_rdump_read(harm_ctx *ctx, const char *path, int binary, harm_restart_header *h);

/* ---- conservation accounting (diag.c:576-834 diag_flux, :1231-1265
   diag_source; dump_ener.c:449-590 divbmaxavg, constotal) ----
   An energy region is a zone box enerpos[X1UP,X1DN,X2UP,X2DN] with the
   faces doflux[dir] its boundary fluxes are summed on (negative: that side
   is not summed), as setflux()/setjetflux() set them (initbase.c:736-980).
   With accounting on, every step does what the reference's step does:
   diag_source() on the last substage (advance.c:138-140) and diag_flux()
   after the step (step_ch.c:160), and diag_fixup() (fixup.c:174-229) for
   every change limit_gamma, the floors and the failed-zone repair make;
   each sum formed in the reference's order so that the totals are the
   reference's to the bit.  The adjoint refuses a context with accounting on. */
#define HARM_NUMFLUXTERMS 7
#define HARM_NUMSOURCES 7
typedef struct {                    /* defs.h:218-239, one region */
  double pdot[HARM_NUMDIRS][8];     /* this step's surface integrals (UDIAG form) */
  double pdotterms[HARM_NUMDIRS][HARM_NUMFLUXTERMS][8];
  double pcum[HARM_NUMDIRS][8];     /* sum over steps of pdot * dt */
  double fladd[8];
  double fladdterms[HARM_NUMFAILFLOORFLAGS][8];
  double sourceadd[8];              /* sum over steps of the volume source */
  double sourceaddterms[HARM_NUMSOURCES][8];
} harm_ener_region;

/* setflux() and setjetflux() (initbase.c:736-980) on a grid: region 0 the
   whole grid, regions 1 and 2 the polar funnels within h_over_r_jet of the
   poles (theta from bl_coord at the x2 faces).  Returns 0, 1 bad argument,
   2 a jet edge on a grid face (the reference's "move h_over_r_jet" exit). */
int harm_ener_regions(const harm_grid *g, double h_over_r_jet, int enerpos[HARM_NUMENERREGIONS][4],
                      int doflux[HARM_NUMENERREGIONS][4]);
/* Accounting on: regions (arrays [HARM_NUMENERREGIONS][4]) and dx[3]
   (2 pi for the reference's 2D coordinates, coord.c:728); zeroes every
   accumulator.  Face loops must lie on the grid (indices 0..N) and zone
   boxes are clipped to the interior as ZLOOP clips them. */
int harm_acct_setup(harm_ctx *ctx, const int *enerpos, const int *doflux, double dx3);
int harm_acct_off(harm_ctx *ctx);
/* No counterpart in the reference: the order of the accounting's sums.
   HARM_ACCT_ORDERED (what harm_acct_setup sets) adds every term in the
   reference's serial loop order, one dependent chain per accumulator: the
   reference's totals to the bit.  HARM_ACCT_TREE, for throughput runs, forms
   each substage's volume-source and fixup totals by a fixed reduction tree
   (per-block partial sums, then a tree per accumulator) and adds them to the
   accumulators: the same from run to run, within rounding of the ordered
   totals (tested to 1e-12 of each accumulator array's largest value).  The
   surface integrals (diag_flux), frdot and lumvsr keep the serial order. */
#define HARM_ACCT_ORDERED 0
#define HARM_ACCT_TREE 1
int harm_acct_set_order(harm_ctx *ctx, int order);
/* the accumulators: regions[HARM_NUMENERREGIONS], pdottermsjet2[4][7][8],
   frdot[N1][8] (radial flux per radius), lumvsr[N1]; any may be NULL */
int harm_acct_read(harm_ctx *ctx, harm_ener_region *regions, double *pdottermsjet2, double *frdot,
                   double *lumvsr);
/* diag_flux() of array HARM_P and the last fluxes with time step Dt (the
   steps call it themselves while accounting is on) */
int harm_diag_flux(harm_ctx *ctx, double Dt);
/* divbmaxavg() of array sel: the largest corner |div B| (first in loop
   order) with its corner, and the sum over the corners / (N1 N2) */
int harm_divbmaxavg(harm_ctx *ctx, int sel, double *divbmax, double *divbavg, int *imax, int *jmax);
/* constotal() of array sel over a region: sum of U (UDIAG form) dVF */
int harm_constotal(harm_ctx *ctx, int sel, int region, double *vars);

/* ---- the current and the 56-column dump (phys.c:1124-1460,
   dump.c:23-191) ----
   The reference's step also computes the current J^mu and the Faraday
   tensor F^{mu nu} (WHICHCURRENTCALC 1, step_ch.c:143-153: the time
   derivative from the last three steps once three steps have run).  On, the
   steps here do the same (from a zero history, as a fresh run). */
int harm_current_enable(harm_ctx *ctx, int on);
/* host arrays jcon[N1+6][N2+6][4], fcon[N1+6][N2+6][6] (either may be NULL) */
int harm_get_current(harm_ctx *ctx, double *jcon, double *fcon);

/* dump_content() columns: i, j, X1, X2, r, th, the 8 primitives, div B,
   u^mu, u_mu, b^mu, b_mu, vmin/vmax in x1 and x2, gdet, J^mu, J_mu, the 6
   F^{mu nu}, the 6 F_{mu nu} */
#define HARM_DUMP_COLUMNS 56
typedef struct {                    /* dump_header() fields (dump.c:48-80) */
  double t;
  int totalsize[2];
  double startx[2], dx[2];
  long realnstep;
  double gam, a, R0, Rin, Rout, hslope, dt;
  int defcoord;
} harm_dump_header;
/* cols[N1][N2][56] in zone order as the file holds them (j outer, i inner):
   text "%21.15g" separated by spaces, one zone per line, or raw doubles.
   Returns 0, 1 bad argument, 2 header grid size differs, 3 cannot open,
   4 short write. */
int harm_dump_write_host(const char *path, int binary, const harm_dump_header *h, int n1, int n2,
                         const double *cols);
/* dump() of the context's state (dump.c:23-46): the grid gives the
   coordinate columns and header fields, the context t, dt and gam;
   realnstep is the caller's step count.  J and F columns are zero unless
   harm_current_enable is on. */
int harm_dump_write(harm_ctx *ctx, const harm_grid *grid, const char *path, int binary, long realnstep);
/* coord()/bl_coord() at zone centre (i,j) */
int harm_zone_coords(const harm_grid *g, int i, int j, double *X1, double *X2, double *r, double *th);

#ifdef __cplusplus
}
#endif
#endif
