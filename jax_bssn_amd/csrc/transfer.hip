// Companion kernels of the evolution: fixed-mesh-refinement transfers
// between a parent level and a vertex-centred 2:1 child patch.
//
//  * prolongation: the child's guard shell (ghost points of every axis) from
//    the parent by separable degree-5 Lagrange interpolation on six parent
//    nodes per axis (reference tests/support/fmr_reference.py:84-143 and
//    fmr/transfer.py:57-69, 237-244);
//  * restriction: injection of the coincident active child vertices into
//    the parent (fmr_reference.py:146-160, fmr/transfer.py:247-252).
//
// States are packed (21 components, each C-order with z fastest).  Child
// allocations are the active vertices plus `ghost` points on each side.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "bssn_amd.h"
#include "bssn_common.h"
#include "host_errors.h"

namespace bssn {
namespace {

struct LevelGeom {
  int n[3];        // allocation extents
  int64_t ncell;
};

struct ProlongGeom {
  LevelGeom parent, child;
  int poff;        // parent ghost width: parent active index 0 sits at poff
  int periodic[3]; // parent axes that wrap
  int lo[3];       // child origin in parent active indices
  int g;           // child ghost width
  int active[3];   // child active extents
};

// Six-node Lagrange weights at parent coordinate t of child index i along
// axis a; the nodes are first, first + 1, .., first + 5 (unwrapped).
__device__ __forceinline__ void lagrange6(const ProlongGeom &G, int a, int i, int &first, double w[6]) {
  const double t = (double)(G.poff + G.lo[a]) + (double)(i - G.g) * 0.5;
  first = (int)floor(t) - 2;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    double p = 1.0;
#pragma unroll
    for (int m = 0; m < 6; ++m)
      if (m != k) p *= (t - (double)(first + m)) / (double)(k - m);
    w[k] = p;
  }
}

// Six Lagrange nodes and weights of child index i along one axis.
__device__ __forceinline__ void axis_stencil(const ProlongGeom &G, int a, int i, int node[6], double w[6]) {
  int first;
  lagrange6(G, a, i, first, w);
#pragma unroll
  for (int k = 0; k < 6; ++k) node[k] = first + k;
  const int n = G.parent.n[a];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    int v = node[k];
    if (G.periodic[a]) {
      v %= n;
      if (v < 0) v += n;
    }
    node[k] = v;
  }
}

// Points of the guard shell: the two x slabs (full y-z planes), the two y
// slabs inside the active x range, the two z slabs inside the active x-y
// range; consecutive threads run along z within each, so every wave is full.
__host__ __device__ __forceinline__ int64_t shell_points(const ProlongGeom &G) {
  const int64_t g2 = 2 * G.g;
  return g2 * G.child.n[1] * G.child.n[2] + (int64_t)G.active[0] * g2 * G.child.n[2] +
         (int64_t)G.active[0] * G.active[1] * g2;
}

// Shell points from r0 on (r0 = 0: the whole shell; r0 = the x- and y-slab
// count: the z slabs only).
__global__ void __launch_bounds__(256) k_fmr_prolong(ProlongGeom G, const double *__restrict__ parent,
                                                     double *__restrict__ child, int64_t r0) {
  int64_t r = r0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= shell_points(G)) return;
  const int g = G.g, g2 = 2 * G.g, NY = G.child.n[1], NZ = G.child.n[2];
  const int64_t nxs = (int64_t)g2 * NY * NZ, nys = (int64_t)G.active[0] * g2 * NZ;
  int i, j, k;
  if (r < nxs) {
    k = (int)(r % NZ);
    j = (int)((r / NZ) % NY);
    const int t = (int)(r / ((int64_t)NZ * NY));
    i = t < g ? t : t + G.active[0];
  } else if ((r -= nxs) < nys) {
    k = (int)(r % NZ);
    const int t = (int)((r / NZ) % g2);
    j = t < g ? t : t + G.active[1];
    i = g + (int)(r / ((int64_t)NZ * g2));
  } else {
    r -= nys;
    const int t = (int)(r % g2);
    k = t < g ? t : t + G.active[2];
    j = g + (int)((r / g2) % G.active[1]);
    i = g + (int)(r / ((int64_t)g2 * G.active[1]));
  }
  const int64_t id = ((int64_t)i * NY + j) * NZ + k;
  int nx[6], ny[6], nz[6];
  double wx[6], wy[6], wz[6];
  axis_stencil(G, 0, i, nx, wx);
  axis_stencil(G, 1, j, ny, wy);
  axis_stencil(G, 2, k, nz, wz);
  const int64_t sx = (int64_t)G.parent.n[1] * G.parent.n[2], sy = G.parent.n[2];
  for (int c = 0; c < NCOMP; ++c) {
    const double *P = parent + c * G.parent.ncell;
    double v = 0.0;
    for (int cz = 0; cz < 6; ++cz) {
      double vy = 0.0;
      for (int by = 0; by < 6; ++by) {
        const double *row = P + ny[by] * sy + nz[cz];
        double vx = 0.0;
#pragma unroll
        for (int ax = 0; ax < 6; ++ax) vx += wx[ax] * row[nx[ax] * sx];
        vy += wy[by] * vx;
      }
      v += wz[cz] * vy;
    }
    child[c * G.child.ncell + id] = v;
  }
}

// The rows of the x and y slabs are whole z rows of ghosts, with x and y
// weights common to the row.  One workgroup per row interpolates the parent
// in x and y once per parent z node of the row (into LDS), then every child
// point of the row takes its six z nodes from there: about 36 / 2 + 6
// parent reads per point and component instead of 216.  The sums run in
// the order of k_fmr_prolong (x innermost, then y, then z).
constexpr int ROW_THREADS = 64;
// A workgroup of the row (column) kernel takes ROW_COMPS (COL_COMPS) of the
// components, blockIdx.y picking which: one row or column group is a single
// wave, and the split gives the launch enough waves to hide the latency of
// the parent reads (measured best: three per row, one per column group).
constexpr int ROW_COMPS = 3, COL_COMPS = 1;
static_assert(NCOMP % ROW_COMPS == 0 && NCOMP % COL_COMPS == 0, "components split evenly");
constexpr int ROW_MAXV = 1024;  // parent z nodes of one row (NZ / 2 + 7 at most)

__global__ void __launch_bounds__(ROW_THREADS) k_fmr_prolong_rows(ProlongGeom G, const double *__restrict__ parent,
                                                                  double *__restrict__ child) {
  __shared__ double V[ROW_MAXV];
  const int q = (int)blockIdx.x;
  const int g = G.g, NY = G.child.n[1], NZ = G.child.n[2];
  const int rows_x = 2 * g * NY;
  int i, j;
  if (q < rows_x) {
    const int t = q / NY;
    j = q % NY;
    i = t < g ? t : t + G.active[0];
  } else {
    const int qq = q - rows_x, t = qq % (2 * g);
    i = g + qq / (2 * g);
    j = t < g ? t : t + G.active[1];
  }
  int nx[6], ny[6];
  double wx[6], wy[6];
  axis_stencil(G, 0, i, nx, wx);
  axis_stencil(G, 1, j, ny, wy);
  int zb0, zb1;
  double wz[6];
  lagrange6(G, 2, 0, zb0, wz);
  lagrange6(G, 2, NZ - 1, zb1, wz);
  const int nv = zb1 - zb0 + 6;  // parent z nodes zb0 .. zb1 + 5
  const int64_t sx = (int64_t)G.parent.n[1] * G.parent.n[2], sy = G.parent.n[2];
  const int pn = G.parent.n[2];
  const int64_t crow = ((int64_t)i * NY + j) * NZ;
  for (int c = (int)blockIdx.y * ROW_COMPS; c < ((int)blockIdx.y + 1) * ROW_COMPS; ++c) {
    const double *P = parent + c * G.parent.ncell;
    __syncthreads();  // every lane is done with the previous component's nodes
    for (int v = threadIdx.x; v < nv; v += ROW_THREADS) {
      int zp = zb0 + v;
      if (G.periodic[2]) {
        zp %= pn;
        if (zp < 0) zp += pn;
      }
      double vy = 0.0;
      for (int by = 0; by < 6; ++by) {
        const double *row = P + ny[by] * sy + zp;
        double vx = 0.0;
#pragma unroll
        for (int ax = 0; ax < 6; ++ax) vx += wx[ax] * row[nx[ax] * sx];
        vy += wy[by] * vx;
      }
      V[v] = vy;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < NZ; k += ROW_THREADS) {
      int first;
      lagrange6(G, 2, k, first, wz);
      double v = 0.0;
#pragma unroll
      for (int cz = 0; cz < 6; ++cz) v += wz[cz] * V[first - zb0 + cz];
      child[c * G.child.ncell + crow + k] = v;
    }
  }
}

// The z slabs are short z columns: 2 g ghosts per (i, j) of the active x-y
// range, g at each end.  Sixteen lanes take one column: lanes [0, nlo)
// interpolate the parent in x and y at the parent z nodes of the low ghosts
// (first node zlo), lanes [nlo, nlo + nhi) at those of the high ghosts
// (first node zhi), into LDS; then lanes [0, 2 g) combine six of them in z
// per ghost.  Four columns per 64-lane workgroup.  Same sum order as
// k_fmr_prolong.
constexpr int COL_LANES = 16;
constexpr int COL_PER_BLOCK = 64 / COL_LANES;

__global__ void __launch_bounds__(64) k_fmr_prolong_zcols(ProlongGeom G, const double *__restrict__ parent,
                                                          double *__restrict__ child, int zlo, int nlo, int zhi,
                                                          int nhi) {
  __shared__ double V[COL_PER_BLOCK][COL_LANES];
  const int lane = (int)threadIdx.x % COL_LANES, slot = (int)threadIdx.x / COL_LANES;
  const int col = (int)blockIdx.x * COL_PER_BLOCK + slot;
  const bool valid = col < G.active[0] * G.active[1];
  const int g = G.g, NY = G.child.n[1], NZ = G.child.n[2];
  const int i = g + (valid ? col / G.active[1] : 0), j = g + (valid ? col % G.active[1] : 0);
  int nx[6], ny[6];
  double wx[6], wy[6], wz[6];
  axis_stencil(G, 0, i, nx, wx);
  axis_stencil(G, 1, j, ny, wy);
  // node of this lane (phase 1) and ghost of this lane (phase 2)
  const bool node_lane = valid && lane < nlo + nhi;
  int zp = lane < nlo ? zlo + lane : zhi + (lane - nlo);
  const int pn = G.parent.n[2];
  if (G.periodic[2]) {
    zp %= pn;
    if (zp < 0) zp += pn;
  }
  const bool ghost_lane = valid && lane < 2 * g;
  const int k = lane < g ? lane : lane + G.active[2];
  int first = 0;
  if (ghost_lane) lagrange6(G, 2, k, first, wz);
  const int vbase = lane < g ? first - zlo : nlo + first - zhi;
  const int64_t sx = (int64_t)G.parent.n[1] * G.parent.n[2], sy = G.parent.n[2];
  const int64_t id = ((int64_t)i * NY + j) * NZ + k;
  for (int c = (int)blockIdx.y * COL_COMPS; c < ((int)blockIdx.y + 1) * COL_COMPS; ++c) {
    __syncthreads();  // every lane is done with the previous component's nodes
    if (node_lane) {
      const double *P = parent + c * G.parent.ncell;
      double vy = 0.0;
      for (int by = 0; by < 6; ++by) {
        const double *row = P + ny[by] * sy + zp;
        double vx = 0.0;
#pragma unroll
        for (int ax = 0; ax < 6; ++ax) vx += wx[ax] * row[nx[ax] * sx];
        vy += wy[by] * vx;
      }
      V[slot][lane] = vy;
    }
    __syncthreads();
    if (ghost_lane) {
      double v = 0.0;
#pragma unroll
      for (int cz = 0; cz < 6; ++cz) v += wz[cz] * V[slot][vbase + cz];
      child[c * G.child.ncell + id] = v;
    }
  }
}

struct RestrictGeom {
  LevelGeom parent, child;
  int poff, lo[3], g, count[3];  // count: coarse vertices covered per axis
};

__global__ void __launch_bounds__(256) k_fmr_restrict(RestrictGeom G, const double *__restrict__ child,
                                                      double *__restrict__ parent) {
  const int64_t total = (int64_t)G.count[0] * G.count[1] * G.count[2];
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= total) return;
  const int k = (int)(id % G.count[2]);
  const int j = (int)((id / G.count[2]) % G.count[1]);
  const int i = (int)(id / ((int64_t)G.count[2] * G.count[1]));
  const int64_t src = ((int64_t)(G.g + 2 * i) * G.child.n[1] + (G.g + 2 * j)) * G.child.n[2] + (G.g + 2 * k);
  const int64_t dst = ((int64_t)(G.poff + G.lo[0] + i) * G.parent.n[1] + (G.poff + G.lo[1] + j)) * G.parent.n[2] +
                      (G.poff + G.lo[2] + k);
#pragma unroll
  for (int c = 0; c < NCOMP; ++c) parent[c * G.parent.ncell + dst] = child[c * G.child.ncell + src];
}

int level_geom(const bssn_grid *g, LevelGeom *L, const char *what) {
  if (!g) return fail(BSSN_ERR_INVALID, "%s grid is null", what);
  L->n[0] = g->nx, L->n[1] = g->ny, L->n[2] = g->nz;
  for (int a = 0; a < 3; ++a)
    if (L->n[a] < 1) return fail(BSSN_ERR_INVALID, "%s grid extents must be positive", what);
  L->ncell = (int64_t)L->n[0] * L->n[1] * L->n[2];
  if (L->ncell >= ((int64_t)1 << 31)) return fail(BSSN_ERR_INVALID, "%s grid has too many points", what);
  return BSSN_OK;
}

// Child active extents from its allocation: 2 (hi - lo) + 1 vertices per axis.
int child_active(const LevelGeom &C, int g, int active[3]) {
  if (g < 0) return fail(BSSN_ERR_INVALID, "ghost width must be non-negative");
  for (int a = 0; a < 3; ++a) {
    active[a] = C.n[a] - 2 * g;
    if (active[a] < 1 || active[a] % 2 == 0)
      return fail(BSSN_ERR_INVALID, "child axis %d: %d points minus 2 x %d ghosts is not an odd vertex count", a,
                  C.n[a], g);
  }
  return BSSN_OK;
}

}  // namespace
}  // namespace bssn

using namespace bssn;

extern "C" {

int bssn_fmr_prolong(const bssn_grid *parent_grid, int32_t parent_ghost, const int32_t *periodic,
                     const double *parent, const bssn_grid *child_grid, const int32_t *lo, int32_t ghost,
                     double *child, void *stream) {
  ProlongGeom G;
  int rc = level_geom(parent_grid, &G.parent, "parent");
  if (!rc) rc = level_geom(child_grid, &G.child, "child");
  if (rc) return rc;
  if (!periodic || !lo || !parent || !child) return fail(BSSN_ERR_INVALID, "null argument");
  if (parent == child) return fail(BSSN_ERR_INVALID, "parent and child must not alias");
  if (parent_ghost < 0) return fail(BSSN_ERR_INVALID, "parent ghost width must be non-negative");
  if (parent_ghost > 0 && (periodic[0] || periodic[1] || periodic[2]))
    return fail(BSSN_ERR_INVALID, "a periodic parent wraps its whole allocation and carries no ghosts");
  if ((rc = child_active(G.child, ghost, G.active))) return rc;
  G.poff = parent_ghost;
  G.g = ghost;
  for (int a = 0; a < 3; ++a) {
    G.lo[a] = lo[a];
    G.periodic[a] = periodic[a] ? 1 : 0;
    const int pactive = G.parent.n[a] - 2 * parent_ghost;
    const int hi = lo[a] + (G.active[a] - 1) / 2;
    if (lo[a] < 0 || hi >= pactive)
      return fail(BSSN_ERR_INVALID, "child bounds [%d, %d] leave the parent's %d active vertices on axis %d", lo[a],
                  hi, pactive, a);
    if (!G.periodic[a]) {
      // extreme stencils: first and last child allocation points
      const double t0 = parent_ghost + lo[a] - 0.5 * ghost;
      const double t1 = parent_ghost + lo[a] + 0.5 * (G.child.n[a] - 1 - ghost);
      if ((int)std::floor(t0) - 2 < 0 || (int)std::floor(t1) + 3 >= G.parent.n[a])
        return fail(BSSN_ERR_INVALID, "child prolongation stencil exceeds parent allocation on axis %d", a);
    }
  }
  const int64_t total = shell_points(G);
  if (total == 0) return BSSN_OK;
  // x and y slabs row by row when a row's parent z nodes fit the LDS
  // buffer, then the z slabs column by column when a column's parent z
  // nodes fit its sixteen lanes; what is left point by point
  const int rows = 2 * G.g * G.child.n[1] + G.active[0] * 2 * G.g;
  int64_t r0 = 0;
  if (rows > 0 && G.child.n[2] / 2 + 7 <= ROW_MAXV) {
    hipLaunchKernelGGL(k_fmr_prolong_rows, dim3((unsigned)rows, NCOMP / ROW_COMPS), dim3(ROW_THREADS), 0,
                       (hipStream_t)stream, G, parent, child);
    if ((rc = check_hip(hipGetLastError(), "prolongation row kernel launch"))) return rc;
    r0 = (int64_t)rows * G.child.n[2];
    // first parent z node of child z index k (as lagrange6 forms it)
    auto zfirst = [&](int k) { return (int)std::floor((double)(G.poff + G.lo[2]) + (double)(k - G.g) * 0.5) - 2; };
    const int zlo = zfirst(0), nlo = zfirst(G.g - 1) - zlo + 6;
    const int zhi = zfirst(G.g + G.active[2]), nhi = zfirst(G.child.n[2] - 1) - zhi + 6;
    const int cols = G.active[0] * G.active[1];
    if (r0 + (int64_t)cols * 2 * G.g == total && 2 * G.g <= COL_LANES && nlo + nhi <= COL_LANES) {
      const unsigned blocks = (unsigned)((cols + COL_PER_BLOCK - 1) / COL_PER_BLOCK);
      hipLaunchKernelGGL(k_fmr_prolong_zcols, dim3(blocks, NCOMP / COL_COMPS), dim3(64), 0, (hipStream_t)stream, G,
                         parent, child, zlo, nlo, zhi, nhi);
      return check_hip(hipGetLastError(), "prolongation column kernel launch");
    }
  }
  const int64_t blocks = (total - r0 + 255) / 256;
  if (blocks == 0) return BSSN_OK;
  hipLaunchKernelGGL(k_fmr_prolong, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, G, parent, child, r0);
  return check_hip(hipGetLastError(), "prolongation kernel launch");
}

int bssn_fmr_restrict(const bssn_grid *child_grid, const int32_t *lo, int32_t ghost, const double *child,
                      const bssn_grid *parent_grid, int32_t parent_ghost, double *parent, void *stream) {
  RestrictGeom G;
  int rc = level_geom(parent_grid, &G.parent, "parent");
  if (!rc) rc = level_geom(child_grid, &G.child, "child");
  if (rc) return rc;
  if (!lo || !parent || !child) return fail(BSSN_ERR_INVALID, "null argument");
  if (parent == child) return fail(BSSN_ERR_INVALID, "parent and child must not alias");
  if (parent_ghost < 0) return fail(BSSN_ERR_INVALID, "parent ghost width must be non-negative");
  int active[3];
  if ((rc = child_active(G.child, ghost, active))) return rc;
  G.poff = parent_ghost;
  G.g = ghost;
  for (int a = 0; a < 3; ++a) {
    G.lo[a] = lo[a];
    G.count[a] = (active[a] - 1) / 2 + 1;
    const int pactive = G.parent.n[a] - 2 * parent_ghost;
    if (lo[a] < 0 || lo[a] + G.count[a] > pactive)
      return fail(BSSN_ERR_INVALID, "child bounds leave the parent's %d active vertices on axis %d", pactive, a);
  }
  const int64_t total = (int64_t)G.count[0] * G.count[1] * G.count[2];
  const int64_t blocks = (total + 255) / 256;
  hipLaunchKernelGGL(k_fmr_restrict, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, G, child, parent);
  return check_hip(hipGetLastError(), "restriction kernel launch");
}

}  // extern "C"
