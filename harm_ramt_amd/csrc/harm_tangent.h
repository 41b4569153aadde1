// harm_tangent.h -- a value-plus-tangent scalar for the tangent-linear
// (forward-mode) step: every operation computes its value exactly as the
// double operation does (so a tangent-linear run carries the primal run's
// values to the bit) and the tangent by the chain rule.  Comparisons act on
// the value, so every piecewise definition (limiters, max/min, clamps) is
// differentiated on the branch the primal takes.  Host+device.
#pragma once
#include "harm_device.h"

namespace harm {

// the double functions stay visible next to the overloads below
using ::fabs;
using ::sqrt;

struct Dual {
  double v, d;
  Dual() = default;
  HARM_HD Dual(double x) : v(x), d(0.) {}
  HARM_HD Dual(double x, double dx) : v(x), d(dx) {}
};

HARM_HD Dual operator-(const Dual& a) { return Dual(-a.v, -a.d); }
HARM_HD Dual operator+(const Dual& a) { return a; }
HARM_HD Dual operator+(const Dual& a, const Dual& b) { return Dual(a.v + b.v, a.d + b.d); }
HARM_HD Dual operator-(const Dual& a, const Dual& b) { return Dual(a.v - b.v, a.d - b.d); }
HARM_HD Dual operator*(const Dual& a, const Dual& b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
HARM_HD Dual operator/(const Dual& a, const Dual& b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
HARM_HD Dual operator+(const Dual& a, double b) { return Dual(a.v + b, a.d); }
HARM_HD Dual operator+(double a, const Dual& b) { return Dual(a + b.v, b.d); }
HARM_HD Dual operator-(const Dual& a, double b) { return Dual(a.v - b, a.d); }
HARM_HD Dual operator-(double a, const Dual& b) { return Dual(a - b.v, -b.d); }
HARM_HD Dual operator*(const Dual& a, double b) { return Dual(a.v * b, a.d * b); }
HARM_HD Dual operator*(double a, const Dual& b) { return Dual(a * b.v, a * b.d); }
HARM_HD Dual operator/(const Dual& a, double b) { return Dual(a.v / b, a.d / b); }
HARM_HD Dual operator/(double a, const Dual& b) {
  const double q = a / b.v;
  return Dual(q, -q * b.d / b.v);
}
HARM_HD Dual& operator+=(Dual& a, const Dual& b) { return a = a + b; }
HARM_HD Dual& operator-=(Dual& a, const Dual& b) { return a = a - b; }
HARM_HD Dual& operator*=(Dual& a, const Dual& b) { return a = a * b; }
HARM_HD Dual& operator*=(Dual& a, double b) { return a = a * b; }
HARM_HD Dual& operator/=(Dual& a, const Dual& b) { return a = a / b; }

HARM_HD bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
HARM_HD bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
HARM_HD bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
HARM_HD bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
HARM_HD bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
HARM_HD bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }
HARM_HD bool operator<(const Dual& a, double b) { return a.v < b; }
HARM_HD bool operator>(const Dual& a, double b) { return a.v > b; }
HARM_HD bool operator<=(const Dual& a, double b) { return a.v <= b; }
HARM_HD bool operator>=(const Dual& a, double b) { return a.v >= b; }
HARM_HD bool operator==(const Dual& a, double b) { return a.v == b; }
HARM_HD bool operator!=(const Dual& a, double b) { return a.v != b; }
HARM_HD bool operator<(double a, const Dual& b) { return a < b.v; }
HARM_HD bool operator>(double a, const Dual& b) { return a > b.v; }

// sqrt: the tangent of sqrt(0) is taken as 0 (the primal's clamped branches
// feed sqrt(0) only where the argument was set to a constant)
HARM_HD Dual sqrt(const Dual& a) {
  const double s = ::sqrt(a.v);
  return Dual(s, (s > 0.) ? a.d / (2. * s) : 0.);
}
HARM_HD Dual fabs(const Dual& a) { return Dual(::fabs(a.v), (a.v < 0.) ? -a.d : a.d); }

}  // namespace harm
