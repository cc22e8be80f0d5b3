// MI355X (gfx950) BSSN evolution: fused RHS + RK4 stage kernels, ghost-free
// periodic wrap / Sommerfeld closure shells, algebraic projection,
// constraints with wavefront reductions, and the C ABI of include/bssn_amd.h.
//
// One RK4 stage is one launch of the tile kernel (all points whose stencils
// are interior) plus, when an axis has Sommerfeld faces, one launch of the
// shell kernel over the three-point closure shells next to those faces.
// Both evaluate the same templated per-point RHS (rhs_point.h) and finish
// with the same in-register stage update and projection.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "bssn_amd.h"
#include "bssn_common.h"
#include "host_errors.h"
#include "rhs_point.h"
#include "stage_common.h"
#include "stage_kernels.h"
#include "stencil_providers.h"
#include "tiles.h"


namespace bssn {

// ----------------------------------------------------------- shell logic

// ------------------------------------------------------ projection kernel

__global__ void __launch_bounds__(256) k_project(int64_t n, const double *__restrict__ in, double *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double g[6], A[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    g[c] = in[(GXX + c) * n + i];
    A[c] = in[(AXX + c) * n + i];
  }
  project(g, A);
  if (in != out) {
    for (int c = 0; c < NCOMP; ++c)
      if (c < GXX || (c > GZZ && c < AXX) || c > AZZ) out[c * n + i] = in[c * n + i];
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    out[(GXX + c) * n + i] = g[c];
    out[(AXX + c) * n + i] = A[c];
  }
}

// RK4 stage combination of an RHS computed elsewhere, then projection
// (weights and bases from stage_update()).
__global__ void __launch_bounds__(256) k_combine(int64_t n, StageOut O, const double *__restrict__ k) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double g[6], A[6];
  for (int c = 0; c < NCOMP; ++c) {
    const int64_t at = c * n + i;
    const double kc = k[at];
    const double un = O.nbase[at] + O.wn * kc;
    if (O.sbase) O.ustar[at] = O.sbase[at] + O.ws * kc;
    if (c >= GXX && c <= GZZ) {
      g[c - GXX] = un;
    } else if (c >= AXX && c <= AZZ) {
      A[c - AXX] = un;
    } else {
      O.unext[at] = un;
    }
  }
  project(g, A);
  if (O.mode == 4 && O.reproject) project(g, A);
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    O.unext[(GXX + c) * n + i] = g[c];
    O.unext[(AXX + c) * n + i] = A[c];
  }
}

// ------------------------------------------------------ constraint kernel

// Reference bssn/constraints.py:22-97 (momentum), 271-303 (Hamiltonian),
// 318-350 (det, trace), 353-384 (Gamma constraint).
// The constraints at the point (x, y, z) (index i) from the stencils of the
// provider pv (global-memory rows of one point, or the LDS boxes of a tile,
// where every thread of the tile stages every component and only threads
// with inside set store).
template <class PS>
__device__ __forceinline__ void constraints_eval(const Ctx &C, PS &pv, const double *__restrict__ us,
                                                 const double *__restrict__ matter, int x, int y, int z, int64_t i,
                                                 double *out, bool inside) {
  double u[NCOMP];
  load_point(C, us, i, u);
  double g[6], A[6], ig[6];
  for (int c = 0; c < 6; ++c) {
    g[c] = u[GXX + c];
    A[c] = u[AXX + c];
  }
  inverse_metric(g, ig);
  const double W = u[CW], W2 = W * W, K = u[CK];
  double R[6] = {0, 0, 0, 0, 0, 0};
  double dg[6][3], dW[3], dK[3];
  // W: second derivatives
  pv.stage_comp(CW);
  for (int j = 0; j < 3; ++j) dW[j] = pv.d1(j);
  {
    double s[6];
    second_derivatives(pv, s);
    const double lap = contract(ig, s);
    for (int c = 0; c < 6; ++c) R[c] += W * (s[c] + g[c] * lap);
  }
  pv.stage_comp(CK);
  for (int j = 0; j < 3; ++j) dK[j] = pv.d1(j);
  for (int m = 0; m < 3; ++m) {
    pv.stage_comp(GAM + m);
    double dG[3];
    for (int j = 0; j < 3; ++j) dG[j] = pv.d1(j);
    for (int c = 0; c < 6; ++c) {
      const int a = SYM_I[c], b = SYM_J[c];
      R[c] += W2 * 0.5 * (g[sym(m, a)] * dG[b] + g[sym(m, b)] * dG[a]);
    }
  }
  for (int c = 0; c < 6; ++c) {
    pv.stage_comp(GXX + c);
    for (int d = 0; d < 3; ++d) dg[c][d] = pv.d1(d);
    double s[6];
    second_derivatives(pv, s);
    R[c] += -0.5 * W2 * contract(ig, s);
  }
  double C1[3][6], C2[3][6];
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 6; ++c) {
      const int j = SYM_I[c], k = SYM_J[c];
      C1[a][c] = 0.5 * (dg[sym(a, j)][k] + dg[sym(a, k)][j] - dg[c][a]);
    }
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 6; ++c)
      C2[a][c] = ig[sym(a, 0)] * C1[0][c] + ig[sym(a, 1)] * C1[1][c] + ig[sym(a, 2)] * C1[2][c];
  const double G[3] = {u[GAM], u[GAM + 1], u[GAM + 2]};
  {
    const double gradW2 = contract_vec(ig, dW, dW);
    double conW = 0.0;
    for (int k = 0; k < 3; ++k) conW += contract(ig, C2[k]) * dW[k];
    for (int c = 0; c < 6; ++c) {
      const int a = SYM_I[c], b = SYM_J[c];
      double t3 = 0, t4 = 0, t5 = 0;
      for (int m = 0; m < 3; ++m) t3 += G[m] * (C1[a][sym(b, m)] + C1[b][sym(a, m)]);
      for (int m = 0; m < 3; ++m)
        for (int n2 = 0; n2 < 3; ++n2) {
          const double gmn = ig[sym(m, n2)];
          double acc4 = 0, acc5 = 0;
          for (int k = 0; k < 3; ++k) {
            acc4 += C2[k][sym(m, a)] * C1[b][sym(k, n2)] + C2[k][sym(m, b)] * C1[a][sym(k, n2)];
            acc5 += C2[k][sym(a, m)] * C1[k][sym(b, n2)];
          }
          t4 += gmn * acc4;
          t5 += gmn * acc5;
        }
      const double chW = C2[0][c] * dW[0] + C2[1][c] * dW[1] + C2[2][c] * dW[2];
      R[c] += W2 * (0.5 * t3 + t4 + t5) - W * (chW + g[c] * conW) - 2.0 * g[c] * gradW2;
    }
  }
  double Aup[6];
  for (int c = 0; c < 6; ++c) {
    const int a = SYM_I[c], b = SYM_J[c];
    double acc = 0;
    for (int k = 0; k < 3; ++k)
      for (int l2 = 0; l2 < 3; ++l2) acc += ig[sym(a, k)] * ig[sym(b, l2)] * A[sym(k, l2)];
    Aup[c] = acc;
  }
  const int64_t n = C.ncell;
  const double H = contract(ig, R) + 2.0 / 3.0 * K * K - contract(A, Aup) - (matter ? 16.0 * M_PI * matter[i] : 0.0);

  // momentum constraint: first[r] = d_j B_r^j with B_r^j = gamma^{jk} A_rk
  // taken as a grid function (the derivative rows applied to B at the
  // taps); the inverse metric of each tap point is formed once for all r
  const double Wf = W > 1.0e-12 ? W : 1.0e-12;
  double first[3] = {0.0, 0.0, 0.0};
  for (int j = 0; j < 3; ++j) {
    if (C.ax[j].reduced) continue;
    const Row8<double> row = row_d1<double>(C.ax[j], j == 0 ? x : (j == 1 ? y : z), C.inv_dx, false);
    double sj[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 8; ++k) {
      if (row.c[k] == 0.0) continue;
      int q[3] = {x, y, z};
      q[j] = wrap_index(q[j] + row.o0 + k, C.ax[j].n);
      const int64_t t = (int64_t)q[0] * C.ax[0].stride + (int64_t)q[1] * C.ax[1].stride + q[2];
      double gg[6], AA[6], ii[6];
      for (int c = 0; c < 6; ++c) {
        gg[c] = us[(GXX + c) * n + t];
        AA[c] = us[(AXX + c) * n + t];
      }
      inverse_metric(gg, ii);
      for (int r = 0; r < 3; ++r)
        sj[r] += row.c[k] * (ii[sym(j, 0)] * AA[sym(r, 0)] + ii[sym(j, 1)] * AA[sym(r, 1)] +
                             ii[sym(j, 2)] * AA[sym(r, 2)]);
    }
    for (int r = 0; r < 3; ++r) first[r] += sj[r];
  }
  double dA[6][3];
  for (int c = 0; c < 6; ++c) {
    pv.stage_comp(AXX + c);
    for (int d = 0; d < 3; ++d) dA[c][d] = pv.d1(d);
  }
  if (!inside) return;
  out[0 * n + i] = H;
  for (int r = 0; r < 3; ++r) {
    double Bw = 0.0;  // B_r^j d_j W
    for (int j = 0; j < 3; ++j) {
      double B = ig[sym(j, 0)] * A[sym(r, 0)] + ig[sym(j, 1)] * A[sym(r, 1)] + ig[sym(j, 2)] * A[sym(r, 2)];
      Bw += B * dW[j];
    }
    double second = 0.0;
    for (int c = 0; c < 6; ++c) second += (SYM_I[c] == SYM_J[c] ? 1.0 : 2.0) * ig[c] * dA[c][r];
    out[(1 + r) * n + i] = first[r] - 0.5 * second - 3.0 * Bw / Wf - 2.0 / 3.0 * dK[r] -
                           (matter ? 8.0 * M_PI * matter[(1 + r) * n + i] : 0.0);
  }
  const double det = g[0] * (g[3] * g[5] - g[4] * g[4]) - g[1] * (g[1] * g[5] - g[4] * g[2]) +
                     g[2] * (g[1] * g[4] - g[3] * g[2]);
  out[4 * n + i] = det - 1.0;
  out[5 * n + i] = contract(ig, A);
  for (int r = 0; r < 3; ++r) out[(6 + r) * n + i] = G[r] - contract(ig, C2[r]);
}

__global__ void __launch_bounds__(256) k_constraints(Ctx C, const double *__restrict__ us,
                                                     const double *__restrict__ matter, double *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C.ncell) return;
  const int z = (int)(i % C.ax[2].n);
  const int y = (int)((i / C.ax[2].n) % C.ax[1].n);
  const int x = (int)(i / ((int64_t)C.ax[2].n * C.ax[1].n));
  PointProvider<double> pv;
  pv.setup(C, us, x, y, z);
  constraints_eval(C, pv, us, matter, x, y, z, pv.idx, out, true);
}

// ------------------------------------------------- constraint damping

__global__ void __launch_bounds__(256) k_derived(int64_t n, const double *__restrict__ us, double *D) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) derived_point(n, i, us, D);
}

struct Box {
  int lo[3], hi[3];
};

// kappa term on the points of the box (the whole grid or a region)
__global__ void __launch_bounds__(256) k_kappa(Ctx C, const double *__restrict__ us, const double *__restrict__ D,
                                               const double *__restrict__ matter, double *kap, Box B) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int ex = B.hi[0] - B.lo[0], ey = B.hi[1] - B.lo[1], ez = B.hi[2] - B.lo[2];
  if (gid >= (int64_t)ex * ey * ez) return;
  const int z = B.lo[2] + (int)(gid % ez);
  const int y = B.lo[1] + (int)((gid / ez) % ey);
  const int x = B.lo[0] + (int)(gid / ((int64_t)ez * ey));
  kappa_point(C, us, D, matter, (int64_t)x * C.ax[0].stride + (int64_t)y * C.ax[1].stride + z, kap);
}

// The kappa term on tiles: the components of the state, the derived fields
// and the matter sources are staged one at a time in one LDS box (the tile
// stencils, closure rows on face tiles), instead of global-memory taps.
template <class P>
struct KappaTaps {
  P *tp;
  const double *src;
  int64_t ncell;
  __device__ __forceinline__ void stage_comp(int c) { tp->stage_from(src + c * ncell); }
  __device__ __forceinline__ double d1(int a) const { return tp->d1(a); }
  __device__ __forceinline__ double d2(int a) const { return tp->d2(a); }
  __device__ __forceinline__ double dmix(int a, int b) const { return tp->dmix(a, b); }
};

template <bool CL, int CLA = 7>
__global__ void __launch_bounds__(TX *TY *TZ) k_kappa_tile(Ctx C, const double *__restrict__ us,
                                                          const double *__restrict__ D,
                                                          const double *__restrict__ matter, double *kap,
                                                          TileSet S) {
  // lean tiles: the closure-free provider with the padded LDS rows of the
  // interior stage kernel (fewer bank conflicts on the mixed derivatives)
  using P = TileProvider<double, TX, TY, TZ, false, CL, CLA, !CL>;
  __shared__ double box[P::NBOXP];
  int bx, by, bz, slab;
  tile_of(S, xcd_remap((int)blockIdx.x, S.count), bx, by, bz, &slab);
  const int tid = threadIdx.x;
  const int tz = tid % TZ, ty = (tid / TZ) % TY, tx = tid / (TY * TZ);
  const int x = bx * TX + tx, y = by * TY + ty, z = bz * TZ + tz;
  const bool inside = x < C.ax[0].n && y < C.ax[1].n && z < C.ax[2].n;
  const int xc = x < C.ax[0].n ? x : C.ax[0].n - 1;
  const int yc = y < C.ax[1].n ? y : C.ax[1].n - 1;
  const int zc = z < C.ax[2].n ? z : C.ax[2].n - 1;
  const int64_t idx = (int64_t)xc * C.ax[0].stride + (int64_t)yc * C.ax[1].stride + zc;
  P tp;
  tp.setup(C, us, box, box, bx * TX, by * TY, bz * TZ, tx, ty, tz, false);  // no stash, no prefetch
  KappaTaps<P> ps{&tp, us, C.ncell}, pd{&tp, D, C.ncell}, pm{&tp, matter, C.ncell};
  kappa_eval(C, ps, pd, matter ? &pm : (KappaTaps<P> *)nullptr, us, D, matter, idx, kap, inside);
}

// The constraints on tiles: the components staged one at a time in one LDS
// box (closure rows on face tiles), as the kappa term; the momentum
// constraint's d_j B_r^j reads its taps from global memory.
template <bool CL>
__global__ void __launch_bounds__(TX *TY *TZ) k_constraints_tile(Ctx C, const double *__restrict__ us,
                                                                const double *__restrict__ matter, double *out,
                                                                TileSet S) {
  using P = TileProvider<double, TX, TY, TZ, false, CL>;
  __shared__ double box[P::NBOXP];
  int bx, by, bz, slab;
  tile_of(S, xcd_remap((int)blockIdx.x, S.count), bx, by, bz, &slab);
  const int tid = threadIdx.x;
  const int tz = tid % TZ, ty = (tid / TZ) % TY, tx = tid / (TY * TZ);
  const int x = bx * TX + tx, y = by * TY + ty, z = bz * TZ + tz;
  const bool inside = x < C.ax[0].n && y < C.ax[1].n && z < C.ax[2].n;
  const int xc = x < C.ax[0].n ? x : C.ax[0].n - 1;
  const int yc = y < C.ax[1].n ? y : C.ax[1].n - 1;
  const int zc = z < C.ax[2].n ? z : C.ax[2].n - 1;
  const int64_t idx = (int64_t)xc * C.ax[0].stride + (int64_t)yc * C.ax[1].stride + zc;
  P tp;
  tp.setup(C, us, box, box, bx * TX, by * TY, bz * TZ, tx, ty, tz, false);  // no stash, no prefetch
  KappaTaps<P> ps{&tp, us, C.ncell};
  constraints_eval(C, ps, us, matter, xc, yc, zc, idx, out, inside);
}

// The kappa term on a box one plane thick along y (the Cartoon centre
// plane), on the region tiles of stage_kernels.h.
__global__ void __launch_bounds__(RTX *RTY *RTZ) k_kappa_region(Ctx C, const double *__restrict__ us,
                                                               const double *__restrict__ D,
                                                               const double *__restrict__ matter, double *kap,
                                                               RegionTiles R) {
  using P = TileProvider<double, RTX, RTY, RTZ, false, true, 5>;
  __shared__ double box[P::NBOXP];
  const int tile = (int)blockIdx.x;
  const int x0 = R.lo[0] + (tile / R.nz) * RTX, y0 = R.lo[1], z0 = R.lo[2] + (tile % R.nz) * RTZ;
  const int tid = threadIdx.x;
  const int tz = tid % RTZ, tx = tid / RTZ;
  const int x = x0 + tx, z = z0 + tz;
  const bool inside = x < R.hi[0] && z < R.hi[2];
  const int xc = x < R.hi[0] ? x : R.hi[0] - 1;
  const int zc = z < R.hi[2] ? z : R.hi[2] - 1;
  const int64_t idx = (int64_t)xc * C.ax[0].stride + (int64_t)y0 * C.ax[1].stride + zc;
  P tp;
  tp.setup(C, us, box, box, x0, y0, z0, tx, 0, tz, false);  // no stash, no prefetch
  KappaTaps<P> ps{&tp, us, C.ncell}, pd{&tp, D, C.ncell}, pm{&tp, matter, C.ncell};
  kappa_eval(C, ps, pd, matter ? &pm : (KappaTaps<P> *)nullptr, us, D, matter, idx, kap, inside);
}

// ------------------------------------------------------------- reductions

constexpr int RED_THREADS = 256;
constexpr int RED_BLOCKS_MAX = 1024;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// partial[b*10 + q]: q = 0..4 sums of squares (H, M, det, trA, Gamma), 5..9 max |.|
__global__ void __launch_bounds__(RED_THREADS) k_norm_partial(int64_t n, const double *__restrict__ cf,
                                                               double *partial) {
  double acc[10];
  for (int q = 0; q < 10; ++q) acc[q] = 0.0;
  const int group[9] = {0, 1, 1, 1, 2, 3, 4, 4, 4};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int f = 0; f < 9; ++f) {
      const double v = cf[f * n + i];
      acc[group[f]] += v * v;
      acc[5 + group[f]] = fmax(acc[5 + group[f]], fabs(v));
    }
  }
  __shared__ double red[RED_THREADS / 64][10];
  const int lane = threadIdx.x % 64, wid = threadIdx.x / 64;
#pragma unroll
  for (int q = 0; q < 10; ++q) {
    const double v = q < 5 ? wave_sum(acc[q]) : wave_max(acc[q]);
    if (lane == 0) red[wid][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < 10) {
    const int q = threadIdx.x;
    double v = red[0][q];
    for (int w = 1; w < RED_THREADS / 64; ++w) v = q < 5 ? v + red[w][q] : fmax(v, red[w][q]);
    partial[blockIdx.x * 10 + q] = v;
  }
}

__global__ void __launch_bounds__(64) k_norm_final(int nblocks, int64_t n, const double *__restrict__ partial,
                                                   double *norms) {
  const int q = threadIdx.x;
  if (q >= 10) return;
  double v = 0.0;
  for (int b = 0; b < nblocks; ++b) v = q < 5 ? v + partial[b * 10 + q] : fmax(v, partial[b * 10 + q]);
  if (q < 5) {
    const double count = (q == 1 || q == 4) ? 3.0 * (double)n : (double)n;
    v = sqrt(v / count);
  }
  norms[q] = v;
}

// Device scratch of bssn_constraint_norms_noalloc: the reduction's partial
// sums (at most RED_BLOCKS_MAX blocks x 10 values, 80 KB).  One copy per
// device, held until the process ends.  In host code this name is only a
// stand-in for the symbol; the device address comes from
// hipGetSymbolAddress (scratch_address below).
static __device__ __align__(256) double scratch[RED_BLOCKS_MAX * 10];

}  // namespace bssn

// =================================================================== host

using namespace bssn;

namespace {
thread_local std::string g_error;
}  // namespace

int bssn::fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_error = buf;
  return code;
}

void bssn::set_error(const char *msg) { g_error = msg; }

namespace {

int resolve(const bssn_params *p, const bssn_grid *grid, Ctx *C) {
  char err[512] = {0};
  const int rc = bssn::resolve(p, grid, C, err, sizeof err);
  if (rc) g_error = err;
  return rc;
}

// Per-stage scratch (doubles) beyond the state buffers: the derived grid
// functions and the kappa term when kappa != 0.
int64_t stage_scratch(const Ctx &C) { return C.kappa != 0.0 ? (int64_t)(NDERIVED + 6) * C.ncell : 0; }

// Derived fields on the whole grid (the kappa stencils reach past the box),
// then the kappa term on the box [lo, hi) (whole grid when lo is null).
int launch_kappa(const Ctx &C, const double *us, const double *matter, double *work, hipStream_t st,
                 const int *lo = nullptr, const int *hi = nullptr) {
  const int64_t blocks = (C.ncell + 255) / 256;
  double *D = work, *kap = work + (int64_t)NDERIVED * C.ncell;
  hipLaunchKernelGGL(k_derived, dim3((unsigned)blocks), dim3(256), 0, st, C.ncell, us, D);
  int rc = check_hip(hipGetLastError(), "derived-field kernel launch");
  if (rc) return rc;
  Box B;
  int64_t pts = 1;
  for (int a = 0; a < 3; ++a) {
    B.lo[a] = lo ? lo[a] : 0;
    B.hi[a] = hi ? hi[a] : C.ax[a].n;
    pts *= B.hi[a] - B.lo[a];
  }
  if (pts <= 0) return BSSN_OK;
  // the whole grid on tiles when the stage kernels tile it (no MAD blend)
  const bool whole = pts == C.ncell;
  const bool mad = C.ax[0].q != 1.0 || C.ax[1].q != 1.0 || C.ax[2].q != 1.0;
  if (whole && !mad && (!C.any_shell || tiles_aligned(C))) {
    const dim3 block(TX * TY * TZ);
    if (!C.any_shell) {
      const TileSet S = tile_set(C, 0, TX, TY, TZ);
      hipLaunchKernelGGL(k_kappa_tile<false>, dim3(S.count), block, 0, st, C, us, D, matter, kap, S);
      return check_hip(hipGetLastError(), "kappa tile kernel launch");
    }
    const TileSet lean = tile_set(C, 1, TX, TY, TZ), face = tile_set(C, 2, TX, TY, TZ);
    if (lean.count > 0) {
      hipLaunchKernelGGL(k_kappa_tile<false>, dim3(lean.count), block, 0, st, C, us, D, matter, kap, lean);
      if ((rc = check_hip(hipGetLastError(), "kappa tile kernel launch"))) return rc;
    }
    // face tiles: one launch per slab with closures along its axis only,
    // one for edges and corners (as the stage kernels)
    if (face.count > 0) {
      const TileSet f3 = tile_set(C, 3, TX, TY, TZ), f4 = tile_set(C, 4, TX, TY, TZ),
                    f5 = tile_set(C, 5, TX, TY, TZ), f6 = tile_set(C, 6, TX, TY, TZ);
      if (f3.count > 0) hipLaunchKernelGGL((k_kappa_tile<true, 1>), dim3(f3.count), block, 0, st, C, us, D, matter, kap, f3);
      if (f4.count > 0) hipLaunchKernelGGL((k_kappa_tile<true, 2>), dim3(f4.count), block, 0, st, C, us, D, matter, kap, f4);
      if (f5.count > 0) hipLaunchKernelGGL((k_kappa_tile<true, 4>), dim3(f5.count), block, 0, st, C, us, D, matter, kap, f5);
      if (f6.count > 0) hipLaunchKernelGGL((k_kappa_tile<true, 7>), dim3(f6.count), block, 0, st, C, us, D, matter, kap, f6);
      rc = check_hip(hipGetLastError(), "kappa tile kernel launch");
    }
    return rc;
  }
  // one plane along y, no closures along y: region tiles
  if (region_tiled(C, B.lo, B.hi)) {
    int count = 0;
    const RegionTiles R = region_tiles(B.lo, B.hi, &count);
    hipLaunchKernelGGL(k_kappa_region, dim3((unsigned)count), dim3(RTX * RTY * RTZ), 0, st, C, us, D, matter, kap,
                       R);
    return check_hip(hipGetLastError(), "kappa region kernel launch");
  }
  hipLaunchKernelGGL(k_kappa, dim3((unsigned)((pts + 255) / 256)), dim3(256), 0, st, C, us, D, matter, kap, B);
  return check_hip(hipGetLastError(), "kappa kernel launch");
}

int launch_stage(const Ctx &C, int precision, const double *us, const StageOut &O0, hipStream_t st,
                 int parts = PART_TILES | PART_SHELL,
                 double *work = nullptr) {
  StageOut O = O0;
  stage_update(O, C.dt);
  if (C.kappa != 0.0) {
    if (!work) return fail(BSSN_ERR_INVALID, "kappa != 0 needs a work buffer of bssn_work_size() doubles");
    const int rc = launch_kappa(C, us, O.matter, work, st);
    if (rc) return rc;
    O.kap = work + (int64_t)NDERIVED * C.ncell;
  }
  if (precision == BSSN_PREC_F64) return launch_stage_f64(C, us, O, st, parts);
  if (precision == BSSN_PREC_F32RHS) return launch_stage_f32(C, us, O, st, parts);
  return fail(BSSN_ERR_INVALID, "unknown precision %d", precision);
}

// RHS on the points of one box [lo, hi) through the closure-capable point
// kernel (every other point of rhs untouched).
int launch_region(const Ctx &C, int precision, const double *us, double *rhs, const int *lo, const int *hi,
                  double *work, hipStream_t st) {
  StageOut O{nullptr, nullptr, nullptr, rhs, 0, 0, nullptr, nullptr};
  stage_update(O, C.dt);
  if (C.kappa != 0.0) {
    if (!work) return fail(BSSN_ERR_INVALID, "kappa != 0 needs a work buffer of bssn_work_size() doubles");
    const int rc = launch_kappa(C, us, nullptr, work, st, lo, hi);
    if (rc) return rc;
    O.kap = work + (int64_t)NDERIVED * C.ncell;
  }
  if (region_tiled(C, lo, hi) && (int64_t)(hi[0] - lo[0]) * (hi[2] - lo[2]) > 0) {
    if (precision == BSSN_PREC_F64) return launch_region_f64(C, us, O, lo, hi, st);
    if (precision == BSSN_PREC_F32RHS) return launch_region_f32(C, us, O, lo, hi, st);
    return fail(BSSN_ERR_INVALID, "unknown precision %d", precision);
  }
  Slabs S;
  std::memset(&S, 0, sizeof S);
  S.count = 1;
  for (int a = 0; a < 3; ++a) {
    S.lo[0][a] = lo[a];
    S.hi[0][a] = hi[a];
  }
  S.first[0] = 0;
  S.first[1] = (int64_t)(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
  if (S.first[1] == 0) return BSSN_OK;
  if (precision == BSSN_PREC_F64) return launch_slabs_f64(C, us, O, S, st);
  if (precision == BSSN_PREC_F32RHS) return launch_slabs_f32(C, us, O, S, st);
  return fail(BSSN_ERR_INVALID, "unknown precision %d", precision);
}

// Points updated by the kernels of one part (PART_* bits, one at a time).
int64_t stage_points(const Ctx &C, int part) {
  const bool cl = C.any_shell && tiles_aligned(C);
  if (part == PART_LEAN || part == PART_CLOSURE) {
    if (!cl) return part == PART_LEAN ? stage_points(C, PART_TILES) : 0;
    return tile_set_points(C, tile_set(C, part == PART_LEAN ? 1 : 2, TX, TY, TZ), TX, TY, TZ);
  }
  int64_t shell = 0;
  if (C.any_shell && !tiles_aligned(C)) {
    // inclusion-exclusion over the per-axis shells: count the complement
    int64_t inner = 1;
    for (int a = 0; a < 3; ++a) inner *= (int64_t)(C.ax[a].n - C.ax[a].shell_lo - C.ax[a].shell_hi);
    shell = C.ncell - inner;
  }
  return part == PART_SHELL ? shell : C.ncell - shell;
}

bool valid_part(int part) {
  return part == PART_TILES || part == PART_SHELL || part == PART_LEAN || part == PART_CLOSURE;
}

// Device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy of the array).
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[MAX_DEVICES] = {};

int scratch_address(double **out) {
  int dev = 0;
  int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  if (dev < 0 || dev >= MAX_DEVICES) return fail(BSSN_ERR_UNSUPPORTED, "device %d beyond the %d the scratch table holds", dev, MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(bssn::scratch)), "hipGetSymbolAddress(scratch)");
    if (rc) return rc;
    g_scratch_ptr[dev] = (double *)p;
  }
  *out = g_scratch_ptr[dev];
  return BSSN_OK;
}

int launch_norms(int64_t n, const double *constraints, double *partial, double *norms, hipStream_t st) {
  int64_t blocks = (n + RED_THREADS - 1) / RED_THREADS;
  if (blocks > RED_BLOCKS_MAX) blocks = RED_BLOCKS_MAX;
  hipLaunchKernelGGL(k_norm_partial, dim3((unsigned)blocks), dim3(RED_THREADS), 0, st, n, constraints, partial);
  int rc = check_hip(hipGetLastError(), "norm kernel launch");
  if (rc) return rc;
  hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(64), 0, st, (int)blocks, n, partial, norms);
  return check_hip(hipGetLastError(), "norm kernel launch");
}

int launch_project(int64_t n, const double *in, double *out, hipStream_t st) {
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_project, dim3((unsigned)blocks), dim3(256), 0, st, n, in, out);
  return check_hip(hipGetLastError(), "projection kernel launch");
}

}  // namespace

extern "C" {

const char *bssn_last_error_string(void) { return g_error.c_str(); }

const char *bssn_version(void) { return "bssn_amd 0.1.0 gfx950"; }

int64_t bssn_work_size(const bssn_params *p, const bssn_grid *grid, int what) {
  Ctx C;
  int rc = resolve(p, grid, &C);
  if (rc) return rc;
  if (what == 0) return stage_scratch(C);
  if (what == 1) return 3 * NCOMP * C.ncell + stage_scratch(C);
  return fail(BSSN_ERR_INVALID, "what must be 0 (bssn_rhs / bssn_rk4_stage) or 1 (bssn_evolve)");
}

int bssn_rhs(const bssn_params *p, const bssn_grid *grid, int precision, const double *state, double *rhs,
             double *work, void *stream) {
  Ctx C;
  int rc = resolve(p, grid, &C);
  if (rc) return rc;
  if (!state || !rhs) return fail(BSSN_ERR_INVALID, "null state or rhs pointer");
  if (state == rhs) return fail(BSSN_ERR_INVALID, "state and rhs must not alias");
  StageOut O{nullptr, nullptr, nullptr, rhs, 0, 0, nullptr, nullptr};
  return launch_stage(C, precision, state, O, (hipStream_t)stream, 3, work);
}

int bssn_rhs_with_matter(const bssn_params *p, const bssn_grid *grid, int precision, const double *state,
                         const double *matter, double *rhs, double *work, void *stream) {
  Ctx C;
  int rc = resolve(p, grid, &C);
  if (rc) return rc;
  if (!state || !rhs || !matter) return fail(BSSN_ERR_INVALID, "null state, matter or rhs pointer");
  if (state == rhs) return fail(BSSN_ERR_INVALID, "state and rhs must not alias");
  StageOut O{nullptr, nullptr, nullptr, rhs, 0, 0, nullptr, matter};
  return launch_stage(C, precision, state, O, (hipStream_t)stream, 3, work);
}

int bssn_enforce_constraints(const bssn_grid *grid, const double *in, double *out, void *stream) {
  if (!grid || !in || !out) return fail(BSSN_ERR_INVALID, "null argument");
  const int64_t n = (int64_t)grid->nx * grid->ny * grid->nz;
  if (n <= 0) return fail(BSSN_ERR_INVALID, "empty grid");
  return launch_project(n, in, out, (hipStream_t)stream);
}

int bssn_rhs_region(const bssn_params *p, const bssn_grid *grid, int precision, const int32_t *lo,
                    const int32_t *hi, const double *state, double *rhs, double *work, void *stream) {
  Ctx C;
  int rc = resolve(p, grid, &C);
  if (rc) return rc;
  if (!lo || !hi || !state || !rhs) return fail(BSSN_ERR_INVALID, "null argument");
  if (state == rhs) return fail(BSSN_ERR_INVALID, "state and rhs must not alias");
  int l[3], h[3];
  for (int a = 0; a < 3; ++a) {
    l[a] = lo[a], h[a] = hi[a];
    if (l[a] < 0 || h[a] > C.ax[a].n || l[a] > h[a])
      return fail(BSSN_ERR_INVALID, "region [%d, %d) does not fit axis %d of extent %d", l[a], h[a], a, C.ax[a].n);
  }
  return launch_region(C, precision, state, rhs, l, h, work, (hipStream_t)stream);
}

int bssn_rk4_combine(const bssn_grid *grid, double dt, int stage, const double *k, const double *u0, double *ustar,
                     double *u_next, int reproject, void *stream) {
  if (!grid || !k || !ustar || !u_next || (stage != 4 && !u0)) return fail(BSSN_ERR_INVALID, "null argument");
  if (stage < 1 || stage > 4) return fail(BSSN_ERR_INVALID, "stage must be 1..4");
  if (!(dt > 0.0) || !std::isfinite(dt)) return fail(BSSN_ERR_INVALID, "dt must be finite and positive");
  if (u_next == k || (stage != 4 && u_next == u0) || u_next == ustar)
    return fail(BSSN_ERR_INVALID, "u_next must not alias k, ustar or (before stage 4) u0");
  const int64_t n = (int64_t)grid->nx * grid->ny * grid->nz;
  if (grid->nx < 1 || grid->ny < 1 || grid->nz < 1 || n >= ((int64_t)1 << 31))
    return fail(BSSN_ERR_INVALID, "grid extents must be positive and below 2^31 points");
  StageOut O{u0, ustar, u_next, nullptr, stage, reproject, nullptr, nullptr};
  stage_update(O, dt);
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_combine, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, n, O, k);
  return check_hip(hipGetLastError(), "combination kernel launch");
}

int bssn_rk4_stage(const bssn_params *p, const bssn_grid *grid, int precision, int stage, const double *u_s,
                   const double *u0, double *ustar, double *u_next, int reproject, double *work, void *stream) {
  Ctx C;
  int rc = resolve(p, grid, &C);
  if (rc) return rc;
  if (stage < 1 || stage > 4) return fail(BSSN_ERR_INVALID, "stage must be 1..4");
  if (!u_s || !ustar || !u_next || (stage != 4 && !u0)) return fail(BSSN_ERR_INVALID, "null buffer");
  if (u_next == u_s) return fail(BSSN_ERR_INVALID, "u_next must not alias u_s");
  if (stage != 4 && u_next == u0) return fail(BSSN_ERR_INVALID, "u_next must not alias u0 before stage 4");
  StageOut O{stage == 1 ? u_s : u0, ustar, u_next, nullptr, stage, reproject, nullptr, nullptr};
  return launch_stage(C, precision, u_s, O, (hipStream_t)stream, 3, work);
}

int bssn_evolve(const bssn_params *p, const bssn_grid *grid, int precision, int nsteps, double *state,
                double *work, void *stream) {
  Ctx C;
  int rc = resolve(p, grid, &C);
  if (rc) return rc;
  if (nsteps < 0) return fail(BSSN_ERR_INVALID, "nsteps must be non-negative");
  if (!state || !work) return fail(BSSN_ERR_INVALID, "null buffer");
  if (nsteps == 0) return BSSN_OK;
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = C.ncell * NCOMP;
  double *ustar = work, *ua = work + n, *ub = work + 2 * n, *scratch = work + 3 * n;
  rc = launch_project(C.ncell, state, state, st);
  if (rc) return rc;
  for (int s = 0; s < nsteps; ++s) {
    const int last = s == nsteps - 1;
    StageOut O1{state, ustar, ua, nullptr, 1, 0, nullptr, nullptr};
    if ((rc = launch_stage(C, precision, state, O1, st, 3, scratch))) return rc;
    StageOut O2{state, ustar, ub, nullptr, 2, 0, nullptr, nullptr};
    if ((rc = launch_stage(C, precision, ua, O2, st, 3, scratch))) return rc;
    StageOut O3{state, ustar, ua, nullptr, 3, 0, nullptr, nullptr};
    if ((rc = launch_stage(C, precision, ub, O3, st, 3, scratch))) return rc;
    StageOut O4{nullptr, ustar, state, nullptr, 4, last ? 0 : 1, nullptr, nullptr};
    if ((rc = launch_stage(C, precision, ua, O4, st, 3, scratch))) return rc;
  }
  return BSSN_OK;
}

int bssn_constraints_with_matter(const bssn_params *p, const bssn_grid *grid, const double *state,
                                 const double *matter, double *out, void *stream) {
  Ctx C;
  bssn_params q = p ? *p : bssn_params{};
  q.kappa = 0.0;  // constraints do not depend on kappa
  int rc = resolve(&q, grid, &C);
  if (rc) return rc;
  if (!state || !out) return fail(BSSN_ERR_INVALID, "null buffer");
  const hipStream_t st = (hipStream_t)stream;
  // tiles where the stage kernels tile the grid: whole grid, no MAD blend,
  // no reduced axis, closures inside the tiles
  bool tiles = !C.any_shell || tiles_aligned(C);
  for (int a = 0; a < 3; ++a)
    tiles = tiles && C.ax[a].q == 1.0 && !C.ax[a].reduced && C.ax[a].start == 0 && C.ax[a].end == C.ax[a].n;
  if (tiles) {
    const dim3 block(TX * TY * TZ);
    if (!C.any_shell) {
      const TileSet S = tile_set(C, 0, TX, TY, TZ);
      hipLaunchKernelGGL(k_constraints_tile<false>, dim3(S.count), block, 0, st, C, state, matter, out, S);
      return check_hip(hipGetLastError(), "constraint tile kernel launch");
    }
    const TileSet lean = tile_set(C, 1, TX, TY, TZ), face = tile_set(C, 2, TX, TY, TZ);
    if (lean.count > 0) {
      hipLaunchKernelGGL(k_constraints_tile<false>, dim3(lean.count), block, 0, st, C, state, matter, out, lean);
      if ((rc = check_hip(hipGetLastError(), "constraint tile kernel launch"))) return rc;
    }
    if (face.count > 0) {
      hipLaunchKernelGGL(k_constraints_tile<true>, dim3(face.count), block, 0, st, C, state, matter, out, face);
      rc = check_hip(hipGetLastError(), "constraint tile kernel launch");
    }
    return rc;
  }
  const int64_t blocks = (C.ncell + 255) / 256;
  hipLaunchKernelGGL(k_constraints, dim3((unsigned)blocks), dim3(256), 0, st, C, state, matter, out);
  return check_hip(hipGetLastError(), "constraint kernel launch");
}

int bssn_constraints(const bssn_params *p, const bssn_grid *grid, const double *state, double *out, void *stream) {
  return bssn_constraints_with_matter(p, grid, state, nullptr, out, stream);
}

int64_t bssn_norm_partial_size(const bssn_grid *grid) {
  if (!grid) return fail(BSSN_ERR_INVALID, "null grid");
  return (int64_t)RED_BLOCKS_MAX * 10;
}

int bssn_constraint_norms(const bssn_grid *grid, const double *constraints, double *partial, double *norms,
                          void *stream) {
  if (!grid || !constraints || !partial || !norms) return fail(BSSN_ERR_INVALID, "null argument");
  const int64_t n = (int64_t)grid->nx * grid->ny * grid->nz;
  if (grid->nx < 1 || grid->ny < 1 || grid->nz < 1) return fail(BSSN_ERR_INVALID, "empty grid");
  return launch_norms(n, constraints, partial, norms, (hipStream_t)stream);
}

int bssn_constraint_norms_noalloc(const bssn_grid *grid, const double *constraints, double *norms, void *stream) {
  if (!grid || !constraints || !norms) return fail(BSSN_ERR_INVALID, "null argument");
  if (grid->nx < 1 || grid->ny < 1 || grid->nz < 1) return fail(BSSN_ERR_INVALID, "empty grid");
  const int64_t n = (int64_t)grid->nx * grid->ny * grid->nz;
  if (n >= ((int64_t)1 << 31))
    return fail(BSSN_ERR_INVALID, "the allocation-free norms take at most 2^31 - 1 points per component (got %lld)",
                (long long)n);
  double *partial = nullptr;
  const int rc = scratch_address(&partial);
  if (rc) return rc;
  // k_norm_partial writes every slot k_norm_final reads: nothing to clear
  return launch_norms(n, constraints, partial, norms, (hipStream_t)stream);
}

int bssn_time_stage_part(const bssn_params *p, const bssn_grid *grid, int precision, int stage, int part,
                         const double *u_s, const double *u0, double *ustar, double *u_next, int reps, float *ms,
                         double *work, void *stream) {
  Ctx C;
  int rc = resolve(p, grid, &C);
  if (rc) return rc;
  if (stage < 1 || stage > 4) return fail(BSSN_ERR_INVALID, "stage must be 1..4");
  if (!valid_part(part)) return fail(BSSN_ERR_INVALID, "part must be one of BSSN_PART_*");
  if (reps < 1 || !ms) return fail(BSSN_ERR_INVALID, "reps must be positive and ms non-null");
  if (!u_s || !ustar || !u_next || (stage != 4 && !u0)) return fail(BSSN_ERR_INVALID, "null buffer");
  if (u_next == u_s) return fail(BSSN_ERR_INVALID, "u_next must not alias u_s");
  hipStream_t st = (hipStream_t)stream;
  hipEvent_t e0, e1;
  if ((rc = check_hip(hipEventCreate(&e0), "hipEventCreate"))) return rc;
  if ((rc = check_hip(hipEventCreate(&e1), "hipEventCreate"))) {
    (void)hipEventDestroy(e0);
    return rc;
  }
  StageOut O{stage == 1 ? u_s : u0, ustar, u_next, nullptr, stage, 0, nullptr, nullptr};
  rc = check_hip(hipEventRecord(e0, st), "hipEventRecord");
  for (int r = 0; r < reps && !rc; ++r) rc = launch_stage(C, precision, u_s, O, st, part, work);
  if (!rc) rc = check_hip(hipEventRecord(e1, st), "hipEventRecord");
  if (!rc) rc = check_hip(hipEventSynchronize(e1), "hipEventSynchronize");
  float t = 0.0f;
  if (!rc) rc = check_hip(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
  if (!rc) *ms = t / (float)reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

int64_t bssn_stage_points(const bssn_params *p, const bssn_grid *grid, int part) {
  Ctx C;
  int rc = resolve(p, grid, &C);
  if (rc) return rc;
  if (!valid_part(part)) return fail(BSSN_ERR_INVALID, "part must be one of BSSN_PART_*");
  return stage_points(C, part);
}

}  // extern "C"
