// Shared definitions of the BSSN kernels: component indices, resolved
// per-axis boundary data and the launch context passed by value to every
// kernel.  Host and device code both include this file.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Functions shared by the kernels and the host build of the per-point code
// (used by the tests to check the algebra on the CPU).
#define BSSN_HD __host__ __device__

// Address spaces made explicit in device code so that LDS and global
// accesses compile to ds_* / global_* instructions rather than flat ones.
#if defined(__HIP_DEVICE_COMPILE__)
#define BSSN_LDS __attribute__((address_space(3)))
#define BSSN_GLOBAL __attribute__((address_space(1)))
#define BSSN_CONSTANT __attribute__((address_space(4)))
#else
#define BSSN_LDS
#define BSSN_GLOBAL
#define BSSN_CONSTANT
#endif

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

// Matter-source components: rho, S_x, S_y, S_z, S_xx S_xy S_xz S_yy S_yz S_zz.
constexpr int NMATTER = 10;

// sym(i, j): index 0..5 of the symmetric pair (i, j) in xx xy xz yy yz zz order.
// Closed form (no recursion) so that it folds to a constant in unrolled loops.
__host__ __device__ constexpr int sym(int i, int j) {
  return (i < j ? i : j) * (5 - (i < j ? i : j)) / 2 + (i < j ? j : i);
}

// Stash slots of the per-point algebra: 0..5 metric of the stage update,
// 6..11 A, 12..14 shift, 15..20 Lie terms of the A RHS, 21 partial K RHS.
constexpr int STASH_SLOTS = 22;
enum StashSlot : int { ST_GN = 0, ST_A = 6, ST_B = 12, ST_ALIE = 15, ST_RK = 21 };

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
  double dt6, dt3, dt2;  // RK4 weights dt/6, dt/3, dt/2
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
  const double *kap;     // kappa term of the A RHS (6 components) or null
  const double *matter;  // rho, S_i (3), S_ij (6, packed) or null
  // filled by the launcher from mode (see stage_update()):
  //   u_next = nbase + wn k;  ustar = sbase + ws k  (sbase null: no update)
  const double *nbase, *sbase;
  double wn, ws;
};

}  // namespace bssn
