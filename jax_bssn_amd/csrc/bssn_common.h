// Shared definitions of the BSSN kernels: component indices, resolved
// per-axis boundary data and the launch context passed by value to every
// kernel.  Host and device code both include this file.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Functions shared by the kernels and the host build of the per-point code
// (used by the tests to check the algebra on the CPU).
#define BSSN_HD __host__ __device__

namespace bssn {

// Packed-state component indices (the reference field order,
// bssn/variables.py:8-16, symmetric tensors stored once).
enum Comp : int {
  GXX = 0, GXY, GXZ, GYY, GYZ, GZZ,
  CW = 6,
  AXX = 7, AXY, AXZ, AYY, AYZ, AZZ,
  CK = 13,
  GAM = 14,  // 14, 15, 16
  ALPHA = 17,
  BETA = 18, // 18, 19, 20
  NCOMP = 21
};

// sym(i, j): index 0..5 of the symmetric pair (i, j) in xx xy xz yy yz zz order.
__host__ __device__ constexpr int sym(int i, int j) {
  return (i > j) ? sym(j, i) : (i == 0 ? j : (i == 1 ? 2 + j : 5));
}

// One spatial axis after the reference's _boundary_data resolution
// (evolution/spatial_derivatives.py:18-41).
struct Axis {
  int n;        // array extent along the axis
  int stride;   // element stride of the axis in a component array
  int lbc, rbc; // closure active at start / end-1 (Sommerfeld and owned face)
  int start, end;  // owned range [start, end) (guard_cells, owned_shape)
  int reduced;  // derivatives along this axis are zero
  int somm_lo, somm_hi;  // Sommerfeld RHS replacement planes active
  int shell_lo, shell_hi;  // widths of the closure shells at start / end
  double q;     // MAD weight (forced to 1 on a Sommerfeld axis)
};

struct Ctx {
  Axis ax[3];
  int64_t ncell;
  double dx, inv_dx, inv_dx2;
  double ko;        // nu / 64 * dx^5 / dx^6
  double eta, g, kappa, dt;
  double xmin[3];
  int gauge, zero_shift;
  int any_shell;    // some axis has a closure shell
};

// Stage-update arguments (see bssn_rk4_stage in include/bssn_amd.h).
struct StageOut {
  const double *u0;
  double *ustar;
  double *unext;
  double *rhs;      // MODE_RHS output
  int mode;         // 0 = write the RHS, 1..4 = RK4 stage
  int reproject;
};

}  // namespace bssn
