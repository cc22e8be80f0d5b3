// Companion kernels of the axisymmetric (z-axis) Cartoon evolution.
//
// Compact storage (reference cartoon/axisymmetry/reconstruction.py:1-145):
// the y = 0 half-plane, extents (nx, 1, nz), x index 0..3 parity ghosts and
// 4..nx-1 the positive radial half-cell centres rho = (i - 3.5) dx.  The
// reflection through the axis is a rotation by pi about z: scalars keep
// their sign, vectors take (-1, -1, +1), tensors the outer product.
//
//  * fill: the four ghosts from the positive half (:131-145);
//  * reconstruct: the nine y planes the Cartesian stencils need (:255-292):
//    each component's radial profile (parity-reflected through the axis,
//    continued past the outer face by a three-sample 1/r buffer, :148-170)
//    is interpolated with six-point Lagrange (cartoon/interpolation.py:9-87)
//    at rho = sqrt(x^2 + y^2) and rotated by phi about z (:212-252);
//  * project: the centre plane's positive half of an RHS on that support,
//    ghosts rebuilt by parity (:295-304).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "bssn_amd.h"
#include "bssn_common.h"
#include "host_errors.h"

namespace bssn {
namespace {

constexpr int AX_GHOSTS = 4, AX_SUPPORT = 9, AX_CENTER = 4, AX_OUTER = 3;

// reflection sign of packed component c under the rotation by pi about z
__host__ __device__ __forceinline__ double axi_parity(int c) {
  switch (c) {
    case GXX + 2: case GXX + 4:          // gamma xz, yz
    case AXX + 2: case AXX + 4:          // A xz, yz
    case GAM: case GAM + 1:              // Gamma^x, Gamma^y
    case BETA: case BETA + 1:            // beta^x, beta^y
      return -1.0;
    default:
      return 1.0;
  }
}

// asymptotic value of the 1/r continuation: delta_ij, W = 1, alpha = 1
__device__ __forceinline__ double axi_asymptotic(int c) {
  return (c == GXX || c == GXX + 3 || c == GXX + 5 || c == CW || c == ALPHA) ? 1.0 : 0.0;
}

struct AxiGeom {
  int nx, nz;      // compact extents (y = 1)
  double dx, z_min;
};

__global__ void __launch_bounds__(256) k_axi_fill(AxiGeom G, double *__restrict__ u) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)AX_GHOSTS * G.nz;
  if (id >= total) return;
  const int k = (int)(id % G.nz), m = (int)(id / G.nz);
  const int64_t n = (int64_t)G.nx * G.nz;
#pragma unroll 1
  for (int c = 0; c < NCOMP; ++c)
    u[c * n + (int64_t)m * G.nz + k] = axi_parity(c) * u[c * n + (int64_t)(2 * AX_GHOSTS - 1 - m) * G.nz + k];
}

// Sample m of component c's radial source at z index k: parity ghosts,
// the positive half, then the 1/r buffer.
__device__ __forceinline__ double axi_source(const AxiGeom &G, const double *__restrict__ u, int c, int m, int k,
                                             double ratio_b[AX_OUTER]) {
  const int64_t n = (int64_t)G.nx * G.nz;
  const double *f = u + c * n;
  if (m < AX_GHOSTS) return axi_parity(c) * f[(int64_t)(2 * AX_GHOSTS - 1 - m) * G.nz + k];
  if (m < G.nx) return f[(int64_t)m * G.nz + k];
  const double a = axi_asymptotic(c);
  return a + (f[(int64_t)(G.nx - 1) * G.nz + k] - a) * ratio_b[m - G.nx];
}

// Rotation of the radial profile by phi (cos cs, sin sn) about z, stored at
// support point pos: v^i = L^i_a v^a, T_ij = L_ia L_jb T_ab with
// L = [[c, -s, 0], [s, c, 0], [0, 0, 1]].  Every loop over components is
// unrolled so that prof / out stay in registers (a runtime index would
// place them in scratch memory).
__device__ __forceinline__ void axi_rotate_store(const double prof[NCOMP], double cs, double sn,
                                                 double *__restrict__ sup, int64_t nsup, int64_t pos) {
  const double L[3][3] = {{cs, -sn, 0.0}, {sn, cs, 0.0}, {0.0, 0.0, 1.0}};
  double out[NCOMP];
#pragma unroll
  for (int c = 0; c < NCOMP; ++c) out[c] = prof[c];
  const int vecs[2] = {GAM, BETA};
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int b = vecs[t];
#pragma unroll
    for (int r = 0; r < 3; ++r) out[b + r] = L[r][0] * prof[b] + L[r][1] * prof[b + 1] + L[r][2] * prof[b + 2];
  }
  const int tens[2] = {GXX, AXX};
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int b = tens[t];
    double T[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int s = 0; s < 3; ++s) T[r][s] = prof[b + sym(r, s)];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int s = r; s < 3; ++s) {
        double acc = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int bb = 0; bb < 3; ++bb) acc += L[r][a] * L[s][bb] * T[a][bb];
        out[b + sym(r, s)] = acc;
      }
  }
#pragma unroll
  for (int c = 0; c < NCOMP; ++c) sup[c * nsup + pos] = out[c];
}

// Planes j and 8 - j lie at y and -y: the same rho (so the same radial
// profile) and opposite sin(phi).  One thread interpolates the profile once
// for the pair p = 0..3 (planes p and 8 - p) or for the centre plane p = 4,
// and stores each plane's rotation.
constexpr int AX_PAIRS = AX_CENTER + 1;

__global__ void __launch_bounds__(256, 2) k_axi_reconstruct(AxiGeom G, const double *__restrict__ u,
                                                         double *__restrict__ sup) {
  const int64_t nsup = (int64_t)G.nx * AX_SUPPORT * G.nz;
  const int64_t nthr = (int64_t)G.nx * AX_PAIRS * G.nz;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nthr) return;
  const int k = (int)(id % G.nz);
  const int j = (int)((id / G.nz) % AX_PAIRS);
  const int i = (int)(id / ((int64_t)G.nz * AX_PAIRS));
  const double x = (i - 3.5) * G.dx, y = (j - 4.0) * G.dx;
  const double rho = sqrt(x * x + y * y);
  const bool on_axis = rho <= 0.0;
  const double cs = on_axis ? 1.0 : x / rho, sn = on_axis ? 0.0 : y / rho;
  // outer 1/r buffer ratios at this z
  const int nr = G.nx - AX_GHOSTS;
  const double z = G.z_min + G.dx * k;
  const double rho_edge = (nr - 0.5) * G.dx;
  const double r_edge = sqrt(rho_edge * rho_edge + z * z);
  double ratio[AX_OUTER];
#pragma unroll
  for (int b = 0; b < AX_OUTER; ++b) {
    const double rb = rho_edge + G.dx * (b + 1);
    ratio[b] = r_edge / sqrt(rb * rb + z * z);
  }
  // six-point Lagrange stencil at q on the source of nx + 3 samples
  const int size = G.nx + AX_OUTER;
  const double q = rho / G.dx + AX_GHOSTS - 0.5;
  int start = (int)floor(q) - 2;
  start = start < 0 ? 0 : (start > size - 6 ? size - 6 : start);
  double w[6];
#pragma unroll
  for (int a = 0; a < 6; ++a) {
    double p = 1.0;
#pragma unroll
    for (int m = 0; m < 6; ++m)
      if (m != a) p *= (q - (double)(start + m)) / (double)(a - m);
    w[a] = p;
  }
  const bool valid = q >= 0.0 && q <= (double)(size - 1);
  double prof[NCOMP];
#pragma unroll
  for (int c = 0; c < NCOMP; ++c) {
    double v = 0.0;
#pragma unroll
    for (int a = 0; a < 6; ++a) v += w[a] * axi_source(G, u, c, start + a, k, ratio);
    prof[c] = valid ? v : NAN;
  }
  axi_rotate_store(prof, cs, sn, sup, nsup, ((int64_t)i * AX_SUPPORT + j) * G.nz + k);
  // the mirror plane 8 - j at -y: the same profile, sin(phi) of opposite
  // sign (bitwise what the plane's own rho and y / rho give)
  if (j < AX_CENTER)
    axi_rotate_store(prof, cs, -sn, sup, nsup, ((int64_t)i * AX_SUPPORT + (2 * AX_CENTER - j)) * G.nz + k);
}

__global__ void __launch_bounds__(256) k_axi_project(AxiGeom G, const double *__restrict__ sup,
                                                     double *__restrict__ out) {
  const int64_t n = (int64_t)G.nx * G.nz;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  const int k = (int)(id % G.nz), m = (int)(id / G.nz);
  const int src_i = m < AX_GHOSTS ? 2 * AX_GHOSTS - 1 - m : m;
  const int64_t nsup = (int64_t)G.nx * AX_SUPPORT * G.nz;
  const int64_t s = ((int64_t)src_i * AX_SUPPORT + AX_CENTER) * G.nz + k;
#pragma unroll 1
  for (int c = 0; c < NCOMP; ++c) out[c * n + id] = (m < AX_GHOSTS ? axi_parity(c) : 1.0) * sup[c * nsup + s];
}

int axi_geom(const bssn_grid *g, double dx, double z_min, AxiGeom *G) {
  if (!g) return fail(BSSN_ERR_INVALID, "null grid");
  if (g->ny != 1) return fail(BSSN_ERR_INVALID, "axisymmetric Cartoon state requires a singleton y plane");
  if (g->nx - AX_GHOSTS < 6) return fail(BSSN_ERR_INVALID, "axisymmetric Cartoon mode requires at least six radial points");
  if (g->nz < 9) return fail(BSSN_ERR_INVALID, "axisymmetric Cartoon mode requires at least nine z points");
  if (!(dx > 0.0) || !std::isfinite(dx)) return fail(BSSN_ERR_INVALID, "axisymmetric Cartoon requires finite dx > 0");
  if ((int64_t)g->nx * AX_SUPPORT * g->nz >= ((int64_t)1 << 31))
    return fail(BSSN_ERR_INVALID, "support grid has too many points");
  G->nx = g->nx, G->nz = g->nz, G->dx = dx, G->z_min = z_min;
  return BSSN_OK;
}

}  // namespace
}  // namespace bssn

using namespace bssn;

extern "C" {

int bssn_cartoon_axi_fill_ghosts(const bssn_grid *compact, double *state, void *stream) {
  AxiGeom G;
  int rc = axi_geom(compact, 1.0, 0.0, &G);
  if (rc) return rc;
  if (!state) return fail(BSSN_ERR_INVALID, "null argument");
  const int64_t total = (int64_t)AX_GHOSTS * G.nz;
  hipLaunchKernelGGL(k_axi_fill, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream, G,
                     state);
  return check_hip(hipGetLastError(), "axisymmetric ghost kernel launch");
}

int bssn_cartoon_axi_reconstruct(const bssn_grid *compact, double dx, double z_min, const double *state,
                                 double *support, void *stream) {
  AxiGeom G;
  int rc = axi_geom(compact, dx, z_min, &G);
  if (rc) return rc;
  if (!state || !support) return fail(BSSN_ERR_INVALID, "null argument");
  if (state == support) return fail(BSSN_ERR_INVALID, "state and support must not alias");
  const int64_t nthr = (int64_t)G.nx * AX_PAIRS * G.nz;
  hipLaunchKernelGGL(k_axi_reconstruct, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     G, state, support);
  return check_hip(hipGetLastError(), "axisymmetric reconstruction kernel launch");
}

int bssn_cartoon_axi_project(const bssn_grid *compact, const double *support_rhs, double *rhs, void *stream) {
  AxiGeom G;
  int rc = axi_geom(compact, 1.0, 0.0, &G);
  if (rc) return rc;
  if (!support_rhs || !rhs) return fail(BSSN_ERR_INVALID, "null argument");
  if (support_rhs == rhs) return fail(BSSN_ERR_INVALID, "support_rhs and rhs must not alias");
  const int64_t n = (int64_t)G.nx * G.nz;
  hipLaunchKernelGGL(k_axi_project, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, G,
                     support_rhs, rhs);
  return check_hip(hipGetLastError(), "axisymmetric projection kernel launch");
}

}  // extern "C"
