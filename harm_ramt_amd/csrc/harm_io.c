/*
 * harm_io.c -- restart and dump files in the reference's format (host side).
 *
 * restart_write()/restart_read() (restart.c:182-264) go through dump_gen()
 * (dump.c:641): the header of write_restart_header() (restart.c:554-682),
 * then the 8 primitives of every interior zone, j outer and i inner.  On one
 * rank with the default file-combine method (MPICOMBINEMINMEM, chosen by
 * init_mpi.c:59-63 since init() sets sortedoutput=SORTED) the zone data is
 * written by the combine loop (init_mpi.c:962-990): binary as raw doubles,
 * text as "%21.15g" separated by single spaces with a newline after each
 * zone.  The header in text is one line (the WRITERESTARTHEADER format of
 * global.h:929-935 and the accumulator arrays as "%21.15g "), in binary the
 * same fields with their C sizes.
 *
 * The file layout, field order and formats have to be the reference's for
 * its tools to read these files and for these functions to read its files.
 */
#include "../../include/harm_amd.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NG 3
#define NPR 8

/* global.h:929-935 */
static const char HDR_FMT[] =
    "%d %d "
    "%21.15g %21.15g %ld %21.15g %21.15g %21.15g "
    "%21.15g %21.15g %21.15g %21.15g %ld %21.15g %ld %ld %ld %ld %ld"
    "%21.15g %d %d %d %d %d %d "
    "%21.15g %21.15g %21.15g %21.15g %d "
    "%d %d %d %d "
    "%ld %d %d %d %ld %ld %d %d %d %d %d %d %d %d %d %d %d %d %d %d %d %d %21.15g %21.15g %21.15g %21.15g %21.15g "
    "%21.15g %21.15g %21.15g %21.15g %21.15g %21.15g ";

/* the reading counterpart (restart.c:389-410, REALTYPE DOUBLETYPE formats) */
static const char HDR_SCAN[] =
    "%d %d "
    "%lf %lf %ld %lf %lf %lf "
    "%lf %lf %lf %lf %ld %lf %ld %ld %ld %ld %ld "
    "%lf %d %d %d %d %d %d "
    "%lf %lf %lf %lf %d "
    "%d %d %d %d "
    "%ld %d %d %d %ld %ld %d %d %d %d %d %d %d %d %d %d %d %d %d %d %d %d %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf "
    "%lf";

static int write_header_text(FILE *f, const harm_restart_header *h) {
  int e, d, k, s, fl;
  if (fprintf(f, HDR_FMT, h->totalsize[0], h->totalsize[1], h->t, h->tf, h->nstep, h->a, h->gam, h->cour, h->DTd,
              h->DTavg, h->DTener, h->DTi, h->DTr, h->DTdebug, h->dump_cnt, h->image_cnt, h->rdump_cnt, h->avg_cnt,
              h->debug_cnt, h->dt, h->lim, h->TIMEORDER, h->fluxmethod, h->FLUXB, h->UTOPRIMVERSION, h->failed, h->R0,
              h->Rin, h->Rout, h->hslope, h->defcoord, h->BCtype[0], h->BCtype[1], h->BCtype[2], h->BCtype[3],
              h->realnstep, h->debugfail, h->whichrestart, h->cooling, h->restartsteps[0], h->restartsteps[1],
              h->GAMMIEDUMP, h->GAMMIEIMAGE, h->GAMMIEENER, h->DODIAGS, h->DOENERDIAG, h->DOGDUMPDIAG, h->DORDUMPDIAG,
              h->DODUMPDIAG, h->DOAVGDIAG, h->DOIMAGEDIAG, h->DOAREAMAPDIAG, h->POSDEFMETRIC, h->periodicx1,
              h->periodicx2, h->binaryoutput, h->sortedoutput, h->defcon, h->SAFE, h->RHOMIN, h->UUMIN,
              h->RHOMINLIMIT, h->UUMINLIMIT, h->BSQORHOLIMIT, h->BSQOULIMIT, h->GAMMAMAX, h->GAMMADAMP,
              h->GAMMAFAIL) < 0)
    return 1;
  /* restart.c:658-676 */
  for (e = 0; e < HARM_NUMENERREGIONS; e++)
    for (d = 0; d < HARM_NUMDIRS; d++)
      for (k = 0; k < NPR; k++) fprintf(f, "%21.15g ", h->pcumreg_tot[e][d][k]);
  for (e = 0; e < HARM_NUMENERREGIONS; e++)
    for (k = 0; k < NPR; k++) fprintf(f, "%21.15g ", h->fladdreg_tot[e][k]);
  for (e = 0; e < HARM_NUMENERREGIONS; e++)
    for (k = 0; k < NPR; k++) fprintf(f, "%21.15g ", h->sourceaddreg_tot[e][k]);
  for (e = 0; e < HARM_NUMENERREGIONS; e++)
    for (k = 0; k < NPR; k++) fprintf(f, "%21.15g ", h->Ureg_init_tot[e][k]);
  for (s = 0; s < HARM_NUMTSCALES; s++)
    for (fl = 0; fl < HARM_NUMFAILFLOORFLAGS; fl++) fprintf(f, "%21.15g ", h->failfloorcount_tot[s][fl]);
  for (k = 0; k < NPR; k++) fprintf(f, "%21.15g ", h->prMAX[k]);
  fprintf(f, "%d ", h->rescaletype);
  for (k = 0; k < h->totalsize[0]; k++) fprintf(f, "%21.15g ", h->lumvsr_tot ? h->lumvsr_tot[k] : 0.0);
  return fprintf(f, "\n") < 0;
}

/* The header readers return 0, 1 on a short or malformed header, or 2 when
   the file's grid differs from n1 x n2: then nothing past the fixed fields is
   read, so lumvsr_tot (n1 doubles) is never written beyond its n1 entries. */
static int read_header_text(FILE *f, harm_restart_header *h, int n1, int n2) {
  int e, d, k, s, fl, n = 0;
  double dummy;
  n = fscanf(f, HDR_SCAN, &h->totalsize[0], &h->totalsize[1], &h->t, &h->tf, &h->nstep, &h->a, &h->gam, &h->cour,
             &h->DTd, &h->DTavg, &h->DTener, &h->DTi, &h->DTr, &h->DTdebug, &h->dump_cnt, &h->image_cnt,
             &h->rdump_cnt, &h->avg_cnt, &h->debug_cnt, &h->dt, &h->lim, &h->TIMEORDER, &h->fluxmethod, &h->FLUXB,
             &h->UTOPRIMVERSION, &h->failed, &h->R0, &h->Rin, &h->Rout, &h->hslope, &h->defcoord, &h->BCtype[0],
             &h->BCtype[1], &h->BCtype[2], &h->BCtype[3], &h->realnstep, &h->debugfail, &h->whichrestart,
             &h->cooling, &h->restartsteps[0], &h->restartsteps[1], &h->GAMMIEDUMP, &h->GAMMIEIMAGE,
             &h->GAMMIEENER, &h->DODIAGS, &h->DOENERDIAG, &h->DOGDUMPDIAG, &h->DORDUMPDIAG, &h->DODUMPDIAG,
             &h->DOAVGDIAG, &h->DOIMAGEDIAG, &h->DOAREAMAPDIAG, &h->POSDEFMETRIC, &h->periodicx1, &h->periodicx2,
             &h->binaryoutput, &h->sortedoutput, &h->defcon, &h->SAFE, &h->RHOMIN, &h->UUMIN, &h->RHOMINLIMIT,
             &h->UUMINLIMIT, &h->BSQORHOLIMIT, &h->BSQOULIMIT, &h->GAMMAMAX, &h->GAMMADAMP, &h->GAMMAFAIL);
  if (n != 68) return 1;
  if (h->totalsize[0] != n1 || h->totalsize[1] != n2) return 2;
  for (e = 0; e < HARM_NUMENERREGIONS; e++)
    for (d = 0; d < HARM_NUMDIRS; d++)
      for (k = 0; k < NPR; k++) n += fscanf(f, "%lf", &h->pcumreg_tot[e][d][k]);
  for (e = 0; e < HARM_NUMENERREGIONS; e++)
    for (k = 0; k < NPR; k++) n += fscanf(f, "%lf", &h->fladdreg_tot[e][k]);
  for (e = 0; e < HARM_NUMENERREGIONS; e++)
    for (k = 0; k < NPR; k++) n += fscanf(f, "%lf", &h->sourceaddreg_tot[e][k]);
  for (e = 0; e < HARM_NUMENERREGIONS; e++)
    for (k = 0; k < NPR; k++) n += fscanf(f, "%lf", &h->Ureg_init_tot[e][k]);
  for (s = 0; s < HARM_NUMTSCALES; s++)
    for (fl = 0; fl < HARM_NUMFAILFLOORFLAGS; fl++) n += fscanf(f, "%lf", &h->failfloorcount_tot[s][fl]);
  for (k = 0; k < NPR; k++) n += fscanf(f, "%lf", &h->prMAX[k]);
  n += fscanf(f, "%d", &h->rescaletype);
  for (k = 0; k < h->totalsize[0]; k++) n += fscanf(f, "%lf", h->lumvsr_tot ? &h->lumvsr_tot[k] : &dummy);
  return n != 68 + HARM_NUMENERREGIONS * (HARM_NUMDIRS + 3) * NPR + HARM_NUMTSCALES * HARM_NUMFAILFLOORFLAGS + NPR +
                   1 + h->totalsize[0];
}

/* binary: each field with its C size (restart.c:564-650).  defcoord, an
   int, is written there from its address with sizeof(FTYPE): 8 bytes, the
   int and the 4 bytes that follow it in memory; here the int and 4 zero
   bytes, and a reader takes the first 4. */
#define W(x) ok &= fwrite(&(x), sizeof(x), 1, f) == 1
#define R(x) ok &= fread(&(x), sizeof(x), 1, f) == 1
#define HDR_FIELDS(X)                                                                                          \
  X(h->totalsize[0]); X(h->totalsize[1]); X(h->t); X(h->tf); X(h->nstep); X(h->a); X(h->gam); X(h->cour);   \
  X(h->DTd); X(h->DTavg); X(h->DTener); X(h->DTi); X(h->DTr); X(h->DTdebug); X(h->dump_cnt);                \
  X(h->image_cnt); X(h->rdump_cnt); X(h->avg_cnt); X(h->debug_cnt); X(h->dt); X(h->lim); X(h->TIMEORDER);     \
  X(h->fluxmethod); X(h->FLUXB); X(h->UTOPRIMVERSION); X(h->failed); X(h->R0); X(h->Rin); X(h->Rout);         \
  X(h->hslope);
#define HDR_FIELDS2(X)                                                                                         \
  X(h->BCtype[0]); X(h->BCtype[1]); X(h->BCtype[2]); X(h->BCtype[3]); X(h->realnstep); X(h->debugfail);        \
  X(h->whichrestart); X(h->cooling); X(h->restartsteps[0]); X(h->restartsteps[1]); X(h->GAMMIEDUMP);          \
  X(h->GAMMIEIMAGE); X(h->GAMMIEENER); X(h->DODIAGS); X(h->DOENERDIAG); X(h->DOGDUMPDIAG); X(h->DORDUMPDIAG);  \
  X(h->DODUMPDIAG); X(h->DOAVGDIAG); X(h->DOIMAGEDIAG); X(h->DOAREAMAPDIAG); X(h->POSDEFMETRIC);               \
  X(h->periodicx1); X(h->periodicx2); X(h->binaryoutput); X(h->sortedoutput); X(h->defcon); X(h->SAFE);      \
  X(h->RHOMIN); X(h->UUMIN); X(h->RHOMINLIMIT); X(h->UUMINLIMIT); X(h->BSQORHOLIMIT); X(h->BSQOULIMIT);       \
  X(h->GAMMAMAX); X(h->GAMMADAMP); X(h->GAMMAFAIL); X(h->pcumreg_tot); X(h->fladdreg_tot);                    \
  X(h->sourceaddreg_tot); X(h->Ureg_init_tot); X(h->failfloorcount_tot); X(h->prMAX); X(h->rescaletype);

static int write_header_binary(FILE *f, const harm_restart_header *h) {
  int ok = 1, k;
  int dc[2];
  HDR_FIELDS(W)
  dc[0] = h->defcoord;
  dc[1] = 0;
  W(dc);
  HDR_FIELDS2(W)
  for (k = 0; k < h->totalsize[0]; k++) {
    double v = h->lumvsr_tot ? h->lumvsr_tot[k] : 0.0;
    W(v);
  }
  return !ok;
}

static int read_header_binary(FILE *f, harm_restart_header *h, int n1, int n2) {
  int ok = 1, k;
  int dc[2];
  HDR_FIELDS(R)
  if (!ok) return 1;
  if (h->totalsize[0] != n1 || h->totalsize[1] != n2) return 2;
  R(dc);
  h->defcoord = dc[0];
  HDR_FIELDS2(R)
  for (k = 0; ok && k < h->totalsize[0]; k++) {
    double v;
    R(v);
    if (h->lumvsr_tot) h->lumvsr_tot[k] = v;
  }
  return !ok;
}

int harm_rdump_write_host(const char *path, int binary, const harm_restart_header *h, int n1, int n2,
                          const double *p) {
  if (!path || !h || !p) return 1;
  if (h->totalsize[0] != n1 || h->totalsize[1] != n2) return 2;
  FILE *f = fopen(path, binary ? "w" : "wt");
  if (!f) return 3;
  int rc = binary ? write_header_binary(f, h) : write_header_text(f, h);
  const int n2m = n2 + 2 * NG;
  for (int j = 0; j < n2 && !rc; j++)
    for (int i = 0; i < n1 && !rc; i++) {
      const double *z = p + ((size_t)(i + NG) * n2m + (j + NG)) * NPR;
      if (binary) {
        rc |= fwrite(z, sizeof(double), NPR, f) != NPR;
      } else {
        for (int k = 0; k < NPR; k++) rc |= fprintf(f, (k < NPR - 1) ? "%21.15g " : "%21.15g\n", z[k]) < 0;
      }
    }
  rc |= fclose(f) != 0;
  return rc ? 4 : 0;
}

int harm_rdump_read_host(const char *path, int binary, harm_restart_header *h, int n1, int n2, double *p) {
  if (!path || !h || !p) return 1;
  FILE *f = fopen(path, binary ? "r" : "rt");
  if (!f) return 3;
  int rc = binary ? read_header_binary(f, h, n1, n2) : read_header_text(f, h, n1, n2);
  const int n2m = n2 + 2 * NG;
  for (int j = 0; j < n2 && !rc; j++)
    for (int i = 0; i < n1 && !rc; i++) {
      double *z = p + ((size_t)(i + NG) * n2m + (j + NG)) * NPR;
      if (binary) {
        rc |= fread(z, sizeof(double), NPR, f) != NPR;
      } else {
        for (int k = 0; k < NPR; k++) rc |= fscanf(f, "%lf", &z[k]) != 1;
      }
    }
  fclose(f);
  return rc == 2 ? 2 : (rc ? 4 : 0);
}

/* dump.c:48-80 dump_header: 16 fields, text on one line, binary with their
   C sizes */
static const char DUMP_HDR_FMT[] =
    "%21.15g %d %d %21.15g %21.15g %21.15g %21.15g %ld %21.15g %21.15g %21.15g %21.15g %21.15g %21.15g %21.15g %d\n";

int harm_dump_write_host(const char *path, int binary, const harm_dump_header *h, int n1, int n2, const double *cols) {
  if (!path || !h || !cols) return 1;
  if (h->totalsize[0] != n1 || h->totalsize[1] != n2) return 2;
  FILE *f = fopen(path, binary ? "w" : "wt");
  if (!f) return 3;
  int rc = 0;
  if (binary) {
    rc |= fwrite(&h->t, sizeof(double), 1, f) != 1;
    rc |= fwrite(&h->totalsize[0], sizeof(int), 1, f) != 1;
    rc |= fwrite(&h->totalsize[1], sizeof(int), 1, f) != 1;
    rc |= fwrite(&h->startx[0], sizeof(double), 1, f) != 1;
    rc |= fwrite(&h->startx[1], sizeof(double), 1, f) != 1;
    rc |= fwrite(&h->dx[0], sizeof(double), 1, f) != 1;
    rc |= fwrite(&h->dx[1], sizeof(double), 1, f) != 1;
    rc |= fwrite(&h->realnstep, sizeof(long), 1, f) != 1;
    rc |= fwrite(&h->gam, sizeof(double), 1, f) != 1;
    rc |= fwrite(&h->a, sizeof(double), 1, f) != 1;
    rc |= fwrite(&h->R0, sizeof(double), 1, f) != 1;
    rc |= fwrite(&h->Rin, sizeof(double), 1, f) != 1;
    rc |= fwrite(&h->Rout, sizeof(double), 1, f) != 1;
    rc |= fwrite(&h->hslope, sizeof(double), 1, f) != 1;
    rc |= fwrite(&h->dt, sizeof(double), 1, f) != 1;
    rc |= fwrite(&h->defcoord, sizeof(int), 1, f) != 1;
  } else {
    rc |= fprintf(f, DUMP_HDR_FMT, h->t, h->totalsize[0], h->totalsize[1], h->startx[0], h->startx[1], h->dx[0],
                  h->dx[1], h->realnstep, h->gam, h->a, h->R0, h->Rin, h->Rout, h->hslope, h->dt, h->defcoord) < 0;
  }
  for (int j = 0; j < n2 && !rc; j++)
    for (int i = 0; i < n1 && !rc; i++) {
      const double *z = cols + ((size_t)i * n2 + j) * HARM_DUMP_COLUMNS;
      if (binary) {
        rc |= fwrite(z, sizeof(double), HARM_DUMP_COLUMNS, f) != HARM_DUMP_COLUMNS;
      } else {
        for (int k = 0; k < HARM_DUMP_COLUMNS; k++)
          rc |= fprintf(f, (k < HARM_DUMP_COLUMNS - 1) ? "%21.15g " : "%21.15g\n", z[k]) < 0;
      }
    }
  rc |= fclose(f) != 0;
  return rc ? 4 : 0;
}
