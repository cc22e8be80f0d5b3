// Forward-mode (tangent-linear) arithmetic: a value and its derivatives
// along ND directions.  Instantiating the per-point RHS, the stage update
// and the projection with DualN yields the exact Jacobian-vector products
// of the RK4 step along ND directions at once (no finite differences).
// Comparisons and branches (upwind sign, W floor) follow the value, as the
// derivative of the branch taken.
#pragma once

#include <cmath>

#include "bssn_common.h"

namespace bssn {

template <int ND>
struct DualN {
  double v, d[ND];
  DualN() = default;
  BSSN_HD constexpr DualN(double x) : v(x), d{} {}
  BSSN_HD DualN &operator+=(const DualN &o) {
    v += o.v;
#pragma unroll
    for (int l = 0; l < ND; ++l) d[l] += o.d[l];
    return *this;
  }
  BSSN_HD DualN &operator-=(const DualN &o) {
    v -= o.v;
#pragma unroll
    for (int l = 0; l < ND; ++l) d[l] -= o.d[l];
    return *this;
  }
  BSSN_HD DualN &operator*=(const DualN &o) {
#pragma unroll
    for (int l = 0; l < ND; ++l) d[l] = d[l] * o.v + v * o.d[l];
    v *= o.v;
    return *this;
  }
};

template <int ND>
BSSN_HD __forceinline__ DualN<ND> operator+(const DualN<ND> &a, const DualN<ND> &b) {
  DualN<ND> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int l = 0; l < ND; ++l) r.d[l] = a.d[l] + b.d[l];
  return r;
}
template <int ND>
BSSN_HD __forceinline__ DualN<ND> operator-(const DualN<ND> &a, const DualN<ND> &b) {
  DualN<ND> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int l = 0; l < ND; ++l) r.d[l] = a.d[l] - b.d[l];
  return r;
}
template <int ND>
BSSN_HD __forceinline__ DualN<ND> operator-(const DualN<ND> &a) {
  DualN<ND> r;
  r.v = -a.v;
#pragma unroll
  for (int l = 0; l < ND; ++l) r.d[l] = -a.d[l];
  return r;
}
template <int ND>
BSSN_HD __forceinline__ DualN<ND> operator*(const DualN<ND> &a, const DualN<ND> &b) {
  DualN<ND> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int l = 0; l < ND; ++l) r.d[l] = a.d[l] * b.v + a.v * b.d[l];
  return r;
}
template <int ND>
BSSN_HD __forceinline__ DualN<ND> operator/(const DualN<ND> &a, const DualN<ND> &b) {
  const double s = 1.0 / b.v;
  DualN<ND> r;
  r.v = a.v * s;
#pragma unroll
  for (int l = 0; l < ND; ++l) r.d[l] = (a.d[l] - r.v * b.d[l]) * s;
  return r;
}
template <int ND>
BSSN_HD __forceinline__ DualN<ND> operator+(const DualN<ND> &a, double b) {
  DualN<ND> r = a;
  r.v += b;
  return r;
}
template <int ND>
BSSN_HD __forceinline__ DualN<ND> operator+(double a, const DualN<ND> &b) {
  return b + a;
}
template <int ND>
BSSN_HD __forceinline__ DualN<ND> operator-(const DualN<ND> &a, double b) {
  DualN<ND> r = a;
  r.v -= b;
  return r;
}
template <int ND>
BSSN_HD __forceinline__ DualN<ND> operator-(double a, const DualN<ND> &b) {
  DualN<ND> r = -b;
  r.v += a;
  return r;
}
template <int ND>
BSSN_HD __forceinline__ DualN<ND> operator*(const DualN<ND> &a, double b) {
  DualN<ND> r;
  r.v = a.v * b;
#pragma unroll
  for (int l = 0; l < ND; ++l) r.d[l] = a.d[l] * b;
  return r;
}
template <int ND>
BSSN_HD __forceinline__ DualN<ND> operator*(double a, const DualN<ND> &b) {
  return b * a;
}
template <int ND>
BSSN_HD __forceinline__ DualN<ND> operator/(const DualN<ND> &a, double b) {
  DualN<ND> r;
  r.v = a.v / b;
#pragma unroll
  for (int l = 0; l < ND; ++l) r.d[l] = a.d[l] / b;
  return r;
}
template <int ND>
BSSN_HD __forceinline__ bool operator<(const DualN<ND> &a, const DualN<ND> &b) { return a.v < b.v; }
template <int ND>
BSSN_HD __forceinline__ bool operator>(const DualN<ND> &a, const DualN<ND> &b) { return a.v > b.v; }
template <int ND>
BSSN_HD __forceinline__ bool operator<=(const DualN<ND> &a, const DualN<ND> &b) { return a.v <= b.v; }
template <int ND>
BSSN_HD __forceinline__ bool operator>=(const DualN<ND> &a, const DualN<ND> &b) { return a.v >= b.v; }
template <int ND>
BSSN_HD __forceinline__ DualN<ND> sqrt(const DualN<ND> &a) {
  DualN<ND> r;
  r.v = ::sqrt(a.v);
  const double h = 0.5 / r.v;
#pragma unroll
  for (int l = 0; l < ND; ++l) r.d[l] = a.d[l] * h;
  return r;
}
template <int ND>
BSSN_HD __forceinline__ DualN<ND> cbrt(const DualN<ND> &a) {
  DualN<ND> r;
  r.v = ::cbrt(a.v);
  const double h = r.v / (3.0 * a.v);
#pragma unroll
  for (int l = 0; l < ND; ++l) r.d[l] = a.d[l] * h;
  return r;
}

}  // namespace bssn
