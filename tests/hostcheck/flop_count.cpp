// Operation count of the per-point RHS algebra (test infrastructure; the
// library never calls it).
//
// bssn_rhs_point is instantiated with a counting scalar and the stencil
// rows of PointProvider (the reference operators of evolution/
// derivatives.py as dot products over their non-zero taps), at an interior
// point of a generic state: every add/sub, mul and div the algebra
// performs is counted once (a fused multiply-add counts as a mul and an
// add); operations with an exact literal zero (an accumulator's T(0)
// start) are not counted.  The RK4 stage update and the algebraic
// projection are counted the same way by their own templates.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "stage_common.h"

using namespace bssn;

namespace {

struct Counts {
  long long add = 0, mul = 0, div = 0, sqrt = 0, cbrt = 0;
};
Counts g_counts;

// A double that counts the operations performed on it.  z marks a literal
// zero (T(0) from an integer constant, as the accumulators start), whose
// adds and muls are free; values converted from doubles (grid samples,
// coefficients) are ordinary operands even when they happen to be zero.
struct FC {
  double v;
  bool z;
  FC() : v(0.0), z(true) {}
  FC(int x) : v((double)x), z(x == 0) {}
  FC(double x) : v(x), z(false) {}
  FC(double x, bool zero) : v(x), z(zero) {}
  explicit operator double() const { return v; }
  FC &operator+=(const FC &b);
  FC &operator-=(const FC &b);
  FC &operator*=(const FC &b);
};

inline FC operator+(const FC &a, const FC &b) {
  if (a.z) return b;
  if (b.z) return a;
  ++g_counts.add;
  return FC(a.v + b.v, false);
}
inline FC operator-(const FC &a) { return FC(-a.v, a.z); }
inline FC operator-(const FC &a, const FC &b) {
  if (b.z) return a;
  if (a.z) return -b;
  ++g_counts.add;
  return FC(a.v - b.v, false);
}
inline FC operator*(const FC &a, const FC &b) {
  if (a.z || b.z) return FC(0.0, true);
  ++g_counts.mul;
  return FC(a.v * b.v, false);
}
inline FC operator/(const FC &a, const FC &b) {
  ++g_counts.div;
  return FC(a.v / b.v, false);
}
FC &FC::operator+=(const FC &b) { return *this = *this + b; }
FC &FC::operator-=(const FC &b) { return *this = *this - b; }
FC &FC::operator*=(const FC &b) { return *this = *this * b; }
inline bool operator<(const FC &a, const FC &b) { return a.v < b.v; }
inline bool operator>(const FC &a, const FC &b) { return a.v > b.v; }
inline bool operator!=(const FC &a, const FC &b) { return a.v != b.v; }
inline bool operator==(const FC &a, const FC &b) { return a.v == b.v; }
inline FC sqrt(const FC &a) {
  ++g_counts.sqrt;
  return FC(std::sqrt(a.v), false);
}
inline FC cbrt(const FC &a) {
  ++g_counts.cbrt;
  return FC(std::cbrt(a.v), false);
}

// Receives the RHS values and nothing else (the stage update is counted
// on its own).
struct NullEmitter {
  RegStash st;
  void prefetch(int) {}
  void put(int, double) {}
  void put_face(int, double) {}
  bool has_matter() const { return false; }
  double matter(int) const { return 0.0; }
  double extra_A(int) const { return 0.0; }
  void done() {}
};

void store(long long *out, const Counts &c) {
  out[0] = c.add;
  out[1] = c.mul;
  out[2] = c.div;
  out[3] = c.sqrt;
  out[4] = c.cbrt;
}

}  // namespace

// Counts of one RHS evaluation at point (x, y, z) of `state` (out[0..4]:
// add/sub, mul, div, sqrt, cbrt) and of one algebraic projection
// (out[5..9]).  Returns a status as the library does.
extern "C" int host_flop_count(const bssn_params *p, const bssn_grid *g, const double *state, int x, int y, int z,
                               long long *out, char *err, int errlen) {
  Ctx C;
  const int rc = resolve(p, g, &C, err, (size_t)errlen);
  if (rc) return rc;
  PointProvider<FC> pv;
  pv.setup(C, state, x, y, z);
  double u[NCOMP];
  for (int c = 0; c < NCOMP; ++c) u[c] = state[c * C.ncell + pv.idx];
  NullEmitter e;
  g_counts = Counts();
  bssn_rhs_point<FC>(pv, C, u, e);
  store(out, g_counts);

  FC gm[6], A[6];
  for (int c = 0; c < 6; ++c) {
    gm[c] = FC(u[GXX + c], false);
    A[c] = FC(u[AXX + c], false);
  }
  g_counts = Counts();
  project(gm, A);
  store(out + 5, g_counts);
  return 0;
}
