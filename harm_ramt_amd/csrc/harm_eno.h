// ENO conversion between zone-centre values and zone averages
// (reconstructeno.c:58-235 choose_stencil/cvtij, reconstructeno.h:18-58
// weights), shared by the device kernels and host-compiled tests.
//
// For every value the stencil grows from the converted zone, one cell at a
// time, towards the side whose undivided Newton difference is smaller (with a
// bias factor 2 towards the centred stencil), restricted to the interior
// zones [0, n) of the conversion line; then the fixed weights of that stencil
// are applied.  The weights are those of exact polynomial reconstruction of
// degree order-1 (cell averages -> centre value, and centre values -> cell
// average); only orders 3 and 5 have tables.
#pragma once
#include "harm_device.h"

constexpr int ENO_A2C = 0;  // reconstructeno.h:3 CVT_A2C
constexpr int ENO_C2A = 1;  // reconstructeno.h:4 CVT_C2A

// Weight of stencil cell d for the zone at position s of an order-cell
// stencil (the stencil starts s zones to the left of the converted zone).
HARM_HD constexpr double eno_weight(int type, int order, int s, int d) {
  constexpr double a2c3[3][3] = {{23. / 24., 1. / 12., -1. / 24.},
                                 {-1. / 24., 13. / 12., -1. / 24.},
                                 {-1. / 24., 1. / 12., 23. / 24.}};
  constexpr double c2a3[3][3] = {{25. / 24., -1. / 12., 1. / 24.},
                                 {1. / 24., 11. / 12., 1. / 24.},
                                 {1. / 24., -1. / 12., 25. / 24.}};
  constexpr double a2c5[5][5] = {{563. / 640., 57. / 160., -373. / 960., 91. / 480., -71. / 1920.},
                                 {-71. / 1920., 511. / 480., -13. / 960., -3. / 160., 3. / 640.},
                                 {3. / 640., -29. / 480., 1067. / 960., -29. / 480., 3. / 640.},
                                 {3. / 640., -3. / 160., -13. / 960., 511. / 480., -71. / 1920.},
                                 {-71. / 1920., 91. / 480., -373. / 960., 57. / 160., 563. / 640.}};
  constexpr double c2a5[5][5] = {{6463. / 5760., -523. / 1440., 383. / 960., -283. / 1440., 223. / 5760.},
                                 {223. / 5760., 1337. / 1440., 23. / 960., 17. / 1440., -17. / 5760.},
                                 {-17. / 5760., 77. / 1440., 863. / 960., 77. / 1440., -17. / 5760.},
                                 {-17. / 5760., 17. / 1440., 23. / 960., 1337. / 1440., 223. / 5760.},
                                 {223. / 5760., -283. / 1440., 383. / 960., -523. / 1440., 6463. / 5760.}};
  return (order == 3) ? (type == ENO_A2C ? a2c3[s][d] : c2a3[s][d])
                      : (type == ENO_A2C ? a2c5[s][d] : c2a5[s][d]);
}

// cvtij (reconstructeno.c:173-229) for one value.  v[0..8] holds the value
// at positions pos-4 .. pos+4 of a line of n interior zones; entries outside
// [0, n) are never read.  Returns the converted value at pos.
template <int TYPE>
HARM_HD double eno_cvt_line(const double (&v)[9], int pos, int n) {
  int order = 5;              // default_order
  if (n < order) order = n;   // choose_stencil: fewer zones than the stencil
  // undivided differences by the recursion of newtdiff (reconstructeno.c:17):
  // D[m-1][p] is the m-th difference starting at position pos-4+p
  double D[4][8];
#pragma unroll
  for (int p = 0; p < 8; p++) D[0][p] = v[p + 1] - v[p];
#pragma unroll
  for (int m = 1; m < 4; m++)
#pragma unroll
    for (int p = 0; p < 7 - m + 1; p++) D[m][p] = D[m - 1][p + 1] - D[m - 1][p];
  // choose_stencil (reconstructeno.c:58-131): is = pos - shift
  const int index_pref = pos - order / 2;
  int shift = 0;
#pragma unroll
  for (int m = 1; m < 5; m++) {
    if (m >= order) break;
    const int is = pos - shift;
    if (is == 0) continue;                                   // cannot expand left
    if (is + m - 1 == n - 1) { shift++; continue; }          // cannot expand right
    double leftdiff = 0., rightdiff = 0.;
#pragma unroll
    for (int s = 0; s < m; s++)
      if (shift == s) {
        leftdiff = D[m - 1][3 - s];   // starts at is-1
        rightdiff = D[m - 1][4 - s];  // starts at is
      }
    const double fac = (is <= index_pref) ? 0.5 : 2.0;  // 1/bias_factor or bias_factor
    if (fabs(leftdiff) < fac * fabs(rightdiff)) shift++;
  }
  if (order < 3) return v[4];  // MIN_CVT_ORDER: leave the value as it is
  if (order == 4) {            // no 4-cell table: drop to 3, keep near centred
    order = 3;
    if (shift > 1) shift--;
  }
  double out = 0.0;
  if (order == 5) {
#pragma unroll
    for (int s = 0; s < 5; s++)
      if (shift == s) {
#pragma unroll
        for (int d = 0; d < 5; d++) out += eno_weight(TYPE, 5, s, d) * v[4 - s + d];
      }
  } else {
#pragma unroll
    for (int s = 0; s < 3; s++)
      if (shift == s) {
#pragma unroll
        for (int d = 0; d < 3; d++) out += eno_weight(TYPE, 3, s, d) * v[4 - s + d];
      }
  }
  return out;
}
