// Forward-mode (tangent-linear) arithmetic: a value and its derivative along
// one direction.  Instantiating the per-point RHS, the stage update and the
// projection with Dual yields the exact Jacobian-vector product of the RK4
// step (no finite differences).  Comparisons and branches (upwind sign,
// W floor) follow the value, as the derivative of the branch taken.
#pragma once

#include <cmath>

#include "bssn_common.h"

namespace bssn {

struct Dual {
  double v, d;
  Dual() = default;
  BSSN_HD constexpr Dual(double x) : v(x), d(0.0) {}
  BSSN_HD constexpr Dual(double x, double dx) : v(x), d(dx) {}
  BSSN_HD Dual &operator+=(const Dual &o) { v += o.v; d += o.d; return *this; }
  BSSN_HD Dual &operator-=(const Dual &o) { v -= o.v; d -= o.d; return *this; }
  BSSN_HD Dual &operator*=(const Dual &o) { d = d * o.v + v * o.d; v *= o.v; return *this; }
};

BSSN_HD __forceinline__ Dual operator+(const Dual &a, const Dual &b) { return Dual(a.v + b.v, a.d + b.d); }
BSSN_HD __forceinline__ Dual operator-(const Dual &a, const Dual &b) { return Dual(a.v - b.v, a.d - b.d); }
BSSN_HD __forceinline__ Dual operator-(const Dual &a) { return Dual(-a.v, -a.d); }
BSSN_HD __forceinline__ Dual operator*(const Dual &a, const Dual &b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
BSSN_HD __forceinline__ Dual operator/(const Dual &a, const Dual &b) {
  const double r = 1.0 / b.v, q = a.v * r;
  return Dual(q, (a.d - q * b.d) * r);
}
BSSN_HD __forceinline__ Dual operator+(const Dual &a, double b) { return Dual(a.v + b, a.d); }
BSSN_HD __forceinline__ Dual operator+(double a, const Dual &b) { return Dual(a + b.v, b.d); }
BSSN_HD __forceinline__ Dual operator-(const Dual &a, double b) { return Dual(a.v - b, a.d); }
BSSN_HD __forceinline__ Dual operator-(double a, const Dual &b) { return Dual(a - b.v, -b.d); }
BSSN_HD __forceinline__ Dual operator*(const Dual &a, double b) { return Dual(a.v * b, a.d * b); }
BSSN_HD __forceinline__ Dual operator*(double a, const Dual &b) { return Dual(a * b.v, a * b.d); }
BSSN_HD __forceinline__ Dual operator/(const Dual &a, double b) { return Dual(a.v / b, a.d / b); }
BSSN_HD __forceinline__ bool operator<(const Dual &a, const Dual &b) { return a.v < b.v; }
BSSN_HD __forceinline__ bool operator>(const Dual &a, const Dual &b) { return a.v > b.v; }
BSSN_HD __forceinline__ bool operator<=(const Dual &a, const Dual &b) { return a.v <= b.v; }
BSSN_HD __forceinline__ bool operator>=(const Dual &a, const Dual &b) { return a.v >= b.v; }
BSSN_HD __forceinline__ Dual sqrt(const Dual &a) {
  const double s = ::sqrt(a.v);
  return Dual(s, a.d * 0.5 / s);
}
BSSN_HD __forceinline__ Dual cbrt(const Dual &a) {
  const double c = ::cbrt(a.v);
  return Dual(c, a.d * c / (3.0 * a.v));
}

}  // namespace bssn
