"""The ENO point <-> zone-average conversion (DOENO, reconstructeno.c), on
the CPU: the oracle's restatement and the product's device function (built
host-only in the test shim) against what the mathematics fixes, and against
each other.

- The conversion weights are those of exact polynomial reconstruction: they
  are derived here in rational arithmetic and must equal both tables.
- Any stencil of the chosen order is exact for polynomials of one degree
  less, so the conversion of a quartic's centre values gives its exact cell
  averages (and back), whatever stencil the selection picks; near the line's
  ends the stencil is one-sided, on lines shorter than 5 the order drops.
- The stencil selection is non-oscillatory: on a step, every zone whose side
  of the step holds at least 5 zones is converted from that side only, so a
  constant stays that constant.
- Device function and oracle agree bit for bit on random, integer (tie-rich)
  and discontinuous lines of every length from 1 to 12 and at every zone.
"""
from fractions import Fraction as Fr

import numpy as np
import pytest

from oracle import oracle as O

A2C, C2A = O.ENO_A2C, O.ENO_C2A


def _solve(A, b):
    n = len(A)
    M = [row[:] + [b[i]] for i, row in enumerate(A)]
    for c in range(n):
        p = next(r for r in range(c, n) if M[r][c] != 0)
        M[c], M[p] = M[p], M[c]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c] / M[c][c]
                M[r] = [x - f * y for x, y in zip(M[r], M[c])]
    return [M[i][n] / M[i][i] for i in range(n)]


def _cell_avg_pow(x, m):
    """average of t^m over the unit cell centred on x"""
    h = Fr(1, 2)
    return ((x + h) ** (m + 1) - (x - h) ** (m + 1)) / (m + 1)


def derived_weights(kind, r):
    """row s: weights of cells 0..r-1 for the zone at position s"""
    rows = []
    for s in range(r):
        if kind == A2C:   # cell averages -> value at the centre of cell s
            A = [[_cell_avg_pow(Fr(d), m) for m in range(r)] for d in range(r)]
            e = [Fr(s) ** m for m in range(r)]
        else:             # centre values -> average over cell s
            A = [[Fr(d) ** m for m in range(r)] for d in range(r)]
            e = [_cell_avg_pow(Fr(s), m) for m in range(r)]
        At = [[A[d][m] for d in range(r)] for m in range(r)]
        rows.append(_solve(At, e))
    return rows


@pytest.mark.parametrize("kind", [A2C, C2A])
@pytest.mark.parametrize("order", [3, 5])
def test_weights_are_exact_polynomial_reconstruction(kind, order, shim):
    W = derived_weights(kind, order)
    for s in range(order):
        assert sum(W[s]) == 1
        for d in range(order):
            assert O.eno_weight(kind, order, s, d) == float(W[s][d])
            assert shim.shim_eno_weight(kind, order, s, d) == float(W[s][d])
    assert O.eno_weight(kind, 4, 0, 0) == 0.0   # no 4-cell table


def _dev_line(shim, kind, line, pos):
    a = np.ascontiguousarray(line, dtype=np.float64)
    return shim.shim_eno_line(kind, a.ctypes.data_as(O.ctypes.POINTER(O.ctypes.c_double)), len(a), pos)


def _quartic(deg, rng):
    c = rng.standard_normal(deg + 1)
    P = np.polynomial.Polynomial(c)
    Pi = P.integ()
    x = np.arange(40, dtype=np.float64)
    return P(x), Pi(x + 0.5) - Pi(x - 0.5)


@pytest.mark.parametrize("n", [3, 4, 5, 6, 9, 16])
def test_conversion_exact_for_polynomials(n, shim):
    rng = np.random.default_rng(n)
    deg = 4 if n >= 5 else 2        # order 5, or 3 on short lines
    for _ in range(5):
        pts, avgs = _quartic(deg, rng)
        pts, avgs = pts[:n], avgs[:n]
        scale = np.abs(pts).max() + np.abs(avgs).max()
        for pos in range(n):
            v, order, shift = O.eno_line(C2A, pts, pos)
            assert order == (5 if n >= 5 else 3) and 0 <= shift < order and shift <= pos
            assert abs(v - avgs[pos]) < 1e-12 * scale
            assert abs(_dev_line(shim, C2A, pts, pos) - v) == 0
            w, order, shift = O.eno_line(A2C, avgs, pos)
            assert abs(w - pts[pos]) < 1e-12 * scale
            assert abs(_dev_line(shim, A2C, avgs, pos) - w) == 0


def test_short_lines_are_left_alone(shim):
    for n in (1, 2):
        line = np.array([3.25, -1.5])[:n]
        for pos in range(n):
            for kind in (A2C, C2A):
                assert O.eno_line(kind, line, pos)[0] == line[pos]
                assert _dev_line(shim, kind, line, pos) == line[pos]


@pytest.mark.parametrize("jump", [5, 9, 13])
def test_stencil_does_not_cross_a_step(jump, shim):
    n = 18
    lo, hi = 0.75, 2.5
    line = np.where(np.arange(n) < jump, lo, hi)
    for pos in range(n):
        side = lo if pos < jump else hi
        room = jump if pos < jump else n - jump
        if room < 5:
            continue
        for kind in (A2C, C2A):
            v, order, shift = O.eno_line(kind, line, pos)
            first = pos - shift
            assert (first + order <= jump) if pos < jump else (first >= jump), (kind, pos, shift)
            assert abs(v - side) <= 4 * np.spacing(side)
            assert _dev_line(shim, kind, line, pos) == v


def test_device_matches_oracle_bitwise(shim):
    rng = np.random.default_rng(11)
    lines = []
    for n in list(range(1, 13)) + [17]:
        lines.append(rng.standard_normal(n))
        lines.append(rng.integers(-2, 3, n).astype(np.float64))            # exact ties
        lines.append(np.where(np.arange(n) < n // 2, 1.0, -3.0) + 1e-3 * rng.standard_normal(n))
        lines.append(np.sin(0.7 * np.arange(n)) * 10.0 ** rng.integers(-3, 4))
    for line in lines:
        for pos in range(len(line)):
            for kind in (A2C, C2A):
                v = O.eno_line(kind, line, pos)[0]
                d = _dev_line(shim, kind, line, pos)
                assert v == d or (np.isnan(v) and np.isnan(d)), (kind, list(line), pos)
