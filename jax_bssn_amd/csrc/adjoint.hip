// Reverse-mode differentiation (the adjoint) of the RHS, of the Hamiltonian
// constraint and of the algebraic projection: the transposed linearisations
// the reverse sweep of an RK4 evolution is built from (jax_bssn_amd/
// adjoint.py composes them, with checkpoints, into the gradient of the
// evolution).
//
// bssn_rhs_vjp forms state_bar = J_F(state)^T rhs_bar + J_H(state)^T h_bar
// in three steps:
//  * the stage kernels in recording mode (record_f64.hip): the features of
//    every point (values, first, second and upwind derivatives of every
//    component, with the reference's stencil rows: closures, MAD, upwind
//    side by the sign of the shift at the point) as NFEAT grid functions;
//  * k_vjp_point: the adjoints of those features (adjoint_point.h);
//  * k_vjp_transpose: at every point q, each component's adjoint is the
//    sum, over every point p whose stencils read q, of the row coefficient
//    of p at the offset q - p times p's feature adjoint -- the transposed
//    stencils, rows evaluated at p (a one-sided closure row, or the upwind
//    side chosen by p's own shift).
// kappa = 0, vacuum, whole grids (no guard cells) only.
#include <hip/hip_runtime.h>

#include "adjoint_point.h"
#include "bssn_amd.h"
#include "host_errors.h"
#include "stage_common.h"
#include "stage_kernels.h"

namespace bssn {
namespace {

__device__ __forceinline__ void point_of(const Ctx &C, int64_t i, int (&q)[3]) {
  q[2] = (int)(i % C.ax[2].n);
  q[1] = (int)((i / C.ax[2].n) % C.ax[1].n);
  q[0] = (int)(i / ((int64_t)C.ax[2].n * C.ax[1].n));
}

// Features of one point in the recorded grid functions (k-th at p[k n]),
// and their adjoints: stored once each, or added to what is stored.
struct StridedFeatures {
  const double *p;
  int64_t n;
  __device__ __forceinline__ double operator()(int k) const { return p[k * n]; }
};
// The upwind and dissipation features enter the RHS linearly, with
// adjoints rbar_c b_a and rbar_c: they are not stored; the point kernel
// stores the effective rbar_c in the dissipation slots instead and the
// transpose applies the upwind rows times the shift at their point.
struct StridedAdjoints {
  double *p;
  int64_t n;
  __device__ __forceinline__ void put(int k, double v) {
    if (k < F_UP) p[k * n] = v;
  }
};
struct StridedAdd {
  double *p;
  int64_t n;
  __device__ __forceinline__ void put(int k, double v) { p[k * n] += v; }
};

// The feature adjoints of every point from its recorded features (feat,
// NFEAT grid functions written by the stage kernels in recording mode).
// The feature and adjoint pointers are not restrict-qualified on purpose:
// after each store of an adjoint the compiler reloads the features it
// needs again (cache hits) instead of keeping them live in registers.
__global__ void __launch_bounds__(64) k_vjp_point(Ctx C, const double *feat, const double *__restrict__ rbar,
                                                  const double *__restrict__ hbar, double *fbar) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C.ncell) return;
  const int64_t n = C.ncell;
  int q[3];
  point_of(C, i, q);
  const bool face = rbar && sommerfeld_point(C, q[0], q[1], q[2]);
  // the Sommerfeld RHS replaces the bulk one on an active face point: the
  // bulk adjoint there carries the H part only (zero RHS adjoints)
  double rb[NCOMP];
  if (rbar)
    for (int c = 0; c < NCOMP; ++c) rb[c] = face ? 0.0 : rbar[c * n + i];
  StridedAdjoints out{fbar + i, n};
  rhs_vjp_point(C, StridedFeatures{feat + i, n}, rbar ? rb : (const double *)nullptr, hbar ? hbar[i] : 0.0, out);
  for (int c = 0; c < NCOMP; ++c)
    fbar[(F_KO + c) * n + i] = (rbar && !(c >= BETA && C.zero_shift)) ? rb[c] : 0.0;
  if (face) {
    double xr[3], rf[NCOMP];
    for (int a = 0; a < 3; ++a) xr[a] = C.xmin[a] + C.dx * (double)q[a];
    for (int c = 0; c < NCOMP; ++c) rf[c] = rbar[c * n + i];
    StridedAdd add{fbar + i, n};
    face_vjp_point(rf, xr, add);
  }
}

// coefficient of the tap at offset d of a row (0 outside its reach)
__device__ __forceinline__ double tap(const Row8<double> &r, int d) {
  const int k = d - r.o0;
  return (k >= 0 && k < r.n) ? r.c[k] : 0.0;
}

// The centred rows of points away from closures and MAD blends, as taps
// at offsets -3 .. 3 (without the 1/dx, 1/dx^2 scales): D4 first and
// second derivatives, the sixth difference of the dissipation and the
// forward / backward upwind D4 rows (reference evolution/derivatives.py).
__device__ constexpr double D1_TAPS[7] = {0.0, 1.0 / 12.0, -2.0 / 3.0, 0.0, 2.0 / 3.0, -1.0 / 12.0, 0.0};
__device__ constexpr double D2_TAPS[7] = {0.0, -1.0 / 12.0, 16.0 / 12.0, -30.0 / 12.0, 16.0 / 12.0, -1.0 / 12.0, 0.0};
__device__ constexpr double D6_TAPS[7] = {1.0, -6.0, 15.0, -20.0, 15.0, -6.0, 1.0};
__device__ constexpr double UPF_TAPS[7] = {0.0, 0.0, -3.0 / 12.0, -10.0 / 12.0, 18.0 / 12.0, -6.0 / 12.0, 1.0 / 12.0};
__device__ constexpr double UPB_TAPS[7] = {-1.0 / 12.0, 6.0 / 12.0, -18.0 / 12.0, 10.0 / 12.0, 3.0 / 12.0, 0.0, 0.0};

// offsets of the points whose rows along axis a may read q: the one-sided
// rows reach seven samples inward from a face, the centred ones three (four
// for the MAD upwind rows)
__device__ __forceinline__ int reach(const Axis &A, int qa) {
  const bool near = (A.lbc && qa - A.start < 10) || (A.rbc && A.end - 1 - qa < 10);
  const int r = near ? 7 : (A.q != 1.0 ? 4 : 3);
  return r < A.n ? r : A.n - 1;
}

// Transposed row coefficients along each axis, for every position q and
// tap offset d in [-7, 7]: the coefficient at offset d of the rows of the
// point p = q - d -- first and second derivative, sixth difference times
// the dissipation factor, and the upwind row for a positive, negative and
// zero shift at p.  Rows depend on the position along one axis only, so the
// points near closures read them instead of building closure rows per tap.
// Layout: axis 0 positions, then axis 1, then axis 2, each (q, d + 7, kind).
constexpr int TR_D = 15, TR_K = 6;

__host__ __device__ inline int64_t table_doubles(int nx, int ny, int nz) {
  return (int64_t)(nx + ny + nz) * TR_D * TR_K;
}

__device__ __forceinline__ int table_row(const Ctx &C, int a, int qa) {
  return (a > 0 ? C.ax[0].n : 0) + (a > 1 ? C.ax[1].n : 0) + qa;
}

__global__ void __launch_bounds__(256) k_vjp_tables(Ctx C, double *__restrict__ tab) {
  const int id = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (id >= (C.ax[0].n + C.ax[1].n + C.ax[2].n) * TR_D) return;
  int qa = id / TR_D, a = 0;
  while (a < 2 && qa >= C.ax[a].n) qa -= C.ax[a++].n;
  const Axis &A = C.ax[a];
  const int d = id % TR_D - 7;
  double *T = tab + (int64_t)id * TR_K;
  if (A.reduced) {
    for (int k = 0; k < TR_K; ++k) T[k] = 0.0;
    return;
  }
  const int p = wrap_index(qa - d, A.n);
  T[0] = tap(row_d1<double>(A, p, C.inv_dx, false), d);
  T[1] = tap(row_d2<double>(A, p, C.inv_dx2), d);
  T[2] = C.ko * tap(row_d6<double>(A, p), d);
  T[3] = tap(row_upwind<double>(A, p, C.inv_dx, 1.0), d);
  T[4] = tap(row_upwind<double>(A, p, C.inv_dx, -1.0), d);
  T[5] = tap(row_upwind<double>(A, p, C.inv_dx, 0.0), d);
}

// Points of the transpose in 4 x 4 x 16 tiles (x, y, z; z fastest over
// the lanes), one tile per 256-thread workgroup, tiles in x-major order
// spread over the XCDs in contiguous ranges: the taps of neighbouring rows
// and planes are read by the same XCD, from its L2.
constexpr int TT_X = 4, TT_Y = 4, TT_Z = 16, TT_THREADS = TT_X * TT_Y * TT_Z;

__host__ __device__ inline int64_t transpose_tiles(const Ctx &C) {
  return (int64_t)((C.ax[0].n + TT_X - 1) / TT_X) * ((C.ax[1].n + TT_Y - 1) / TT_Y) * ((C.ax[2].n + TT_Z - 1) / TT_Z);
}

__global__ void __launch_bounds__(TT_THREADS) k_vjp_transpose(Ctx C, const double *__restrict__ us,
                                                              const double *__restrict__ fbar,
                                                              const double *__restrict__ tab,
                                                              double *__restrict__ ubar, int accumulate) {
  const int nty = (C.ax[1].n + TT_Y - 1) / TT_Y, ntz = (C.ax[2].n + TT_Z - 1) / TT_Z;
  const int t = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const int q[3] = {(t / (ntz * nty)) * TT_X + (int)(threadIdx.x / (TT_Y * TT_Z)),
                    ((t / ntz) % nty) * TT_Y + (int)((threadIdx.x / TT_Z) % TT_Y),
                    (t % ntz) * TT_Z + (int)(threadIdx.x % TT_Z)};
  if (q[0] >= C.ax[0].n || q[1] >= C.ax[1].n || q[2] >= C.ax[2].n) return;
  const int64_t n = C.ncell;
  const int64_t i = (int64_t)q[0] * C.ax[0].stride + (int64_t)q[1] * C.ax[1].stride + q[2];
  double acc[NCOMP];
  for (int c = 0; c < NCOMP; ++c) acc[c] = fbar[(F_VAL + c) * n + i];
  // first, pure second, upwind and Kreiss-Oliger rows along each axis
  for (int a = 0; a < 3; ++a) {
    const Axis &A = C.ax[a];
    if (A.reduced) continue;
    const int R = reach(A, q[a]);
    // no closure row and no MAD blend within reach: the centred rows
    const bool plain = R == 3 && A.q == 1.0;
    for (int d = -R; d <= R; ++d) {
      int p[3] = {q[0], q[1], q[2]};
      p[a] = wrap_index(q[a] - d, A.n);
      const int64_t j = (int64_t)p[0] * C.ax[0].stride + (int64_t)p[1] * C.ax[1].stride + p[2];
      const double bp = us[(BETA + a) * n + j];
      double c1, c2, c6, cu;
      if (plain) {
        const int k = d + 3;
        c1 = D1_TAPS[k] * C.inv_dx;
        c2 = D2_TAPS[k] * C.inv_dx2;
        c6 = C.ko * D6_TAPS[k];
        cu = (bp > 0.0 ? UPF_TAPS[k] : (bp < 0.0 ? UPB_TAPS[k] : D1_TAPS[k])) * C.inv_dx;
      } else {
        const double *T = tab + ((int64_t)table_row(C, a, q[a]) * TR_D + d + 7) * TR_K;
        c1 = T[0];
        c2 = T[1];
        c6 = T[2];
        cu = bp > 0.0 ? T[3] : (bp < 0.0 ? T[4] : T[5]);
      }
      if (c1 == 0.0 && c2 == 0.0 && c6 == 0.0 && cu == 0.0) continue;
      if (c1 != 0.0)
        for (int c = 0; c < NCOMP; ++c) acc[c] += c1 * fbar[(F_D1 + 3 * c + a) * n + j];
      // upwind (times the shift at p) and dissipation rows on the effective rbar at p
      const double cr = cu * bp + c6;
      if (cr != 0.0)
        for (int c = 0; c < NCOMP; ++c) acc[c] += cr * fbar[(F_KO + c) * n + j];
      if (c2 != 0.0)
        for (int c = 0; c < NCOMP; ++c) {
          const int s = d2_slot(c);
          if (s >= 0) acc[c] += c2 * fbar[(F_D2 + 6 * s + sym(a, a)) * n + j];
        }
    }
  }
  // mixed second derivatives: products of the two axes' first-derivative rows
  for (int a = 0; a < 3; ++a)
    for (int bx = a + 1; bx < 3; ++bx) {
      const Axis &A = C.ax[a], &B = C.ax[bx];
      if (A.reduced || B.reduced) continue;
      const int Ra = reach(A, q[a]), Rb = reach(B, q[bx]);
      const bool plain_a = Ra == 3 && A.q == 1.0, plain_b = Rb == 3 && B.q == 1.0;
      for (int da = -Ra; da <= Ra; ++da) {
        const int pa = wrap_index(q[a] - da, A.n);
        const double ca = plain_a ? D1_TAPS[da + 3] * C.inv_dx
                                  : tab[((int64_t)table_row(C, a, q[a]) * TR_D + da + 7) * TR_K];
        if (ca == 0.0) continue;
        for (int db = -Rb; db <= Rb; ++db) {
          const int pb = wrap_index(q[bx] - db, B.n);
          const double cb = plain_b ? D1_TAPS[db + 3] * C.inv_dx
                                    : tab[((int64_t)table_row(C, bx, q[bx]) * TR_D + db + 7) * TR_K];
          if (cb == 0.0) continue;
          int p[3] = {q[0], q[1], q[2]};
          p[a] = pa;
          p[bx] = pb;
          const int64_t j = (int64_t)p[0] * C.ax[0].stride + (int64_t)p[1] * C.ax[1].stride + p[2];
          for (int c = 0; c < NCOMP; ++c) {
            const int s = d2_slot(c);
            if (s >= 0) acc[c] += ca * cb * fbar[(F_D2 + 6 * s + sym(a, bx)) * n + j];
          }
        }
      }
    }
  for (int c = 0; c < NCOMP; ++c) ubar[c * n + i] = (accumulate ? ubar[c * n + i] : 0.0) + acc[c];
}

__global__ void __launch_bounds__(256) k_project_vjp(int64_t n, const double *__restrict__ us,
                                                     double *__restrict__ ubar) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double g[6], A[6], gb[6], Ab[6];
  for (int c = 0; c < 6; ++c) {
    g[c] = us[(GXX + c) * n + i];
    A[c] = us[(AXX + c) * n + i];
    gb[c] = ubar[(GXX + c) * n + i];
    Ab[c] = ubar[(AXX + c) * n + i];
  }
  project_vjp(g, A, gb, Ab);
  for (int c = 0; c < 6; ++c) {
    ubar[(GXX + c) * n + i] = gb[c];
    ubar[(AXX + c) * n + i] = Ab[c];
  }
}

int resolve_vjp(const bssn_params *p, const bssn_grid *grid, Ctx *C) {
  char err[512] = {0};
  const int rc = resolve(p, grid, C, err, sizeof err);
  if (rc) return fail(rc, "%s", err);
  if (C->kappa != 0.0) return fail(BSSN_ERR_UNSUPPORTED, "the adjoint takes kappa = 0");
  if (p->guard_cells != 0 || p->owned_shape[0] >= 0)
    return fail(BSSN_ERR_UNSUPPORTED, "the adjoint takes whole grids (no guard cells or owned range)");
  return BSSN_OK;
}

}  // namespace
}  // namespace bssn

using namespace bssn;

extern "C" {

int64_t bssn_vjp_work_size(const bssn_grid *grid) {
  if (!grid || grid->nx < 1 || grid->ny < 1 || grid->nz < 1) return fail(BSSN_ERR_INVALID, "invalid grid");
  // features, their adjoints and the transposed row tables
  return 2 * (int64_t)NFEAT * grid->nx * grid->ny * grid->nz + table_doubles(grid->nx, grid->ny, grid->nz);
}

int bssn_rhs_vjp(const bssn_params *p, const bssn_grid *grid, const double *state, const double *rhs_bar,
                 const double *h_bar, double *state_bar, int accumulate, double *work, void *stream) {
  Ctx C;
  int rc = resolve_vjp(p, grid, &C);
  if (rc) return rc;
  if (!state || !state_bar || !work) return fail(BSSN_ERR_INVALID, "null buffer");
  if (!rhs_bar && !h_bar) return fail(BSSN_ERR_INVALID, "rhs_bar and h_bar are both null");
  const hipStream_t st = (hipStream_t)stream;
  const int64_t b64 = (C.ncell + 63) / 64;
  double *feat = work, *fbar = work + (int64_t)NFEAT * C.ncell;
  if ((rc = launch_record_f64(C, state, feat, st))) return rc;
  hipLaunchKernelGGL(k_vjp_point, dim3((unsigned)b64), dim3(64), 0, st, C, feat, rhs_bar, h_bar, fbar);
  if ((rc = check_hip(hipGetLastError(), "adjoint point kernel launch"))) return rc;
  double *tab = fbar + (int64_t)NFEAT * C.ncell;
  const int nrow = (C.ax[0].n + C.ax[1].n + C.ax[2].n) * TR_D;
  hipLaunchKernelGGL(k_vjp_tables, dim3((unsigned)((nrow + 255) / 256)), dim3(256), 0, st, C, tab);
  if ((rc = check_hip(hipGetLastError(), "adjoint table kernel launch"))) return rc;
  hipLaunchKernelGGL(k_vjp_transpose, dim3((unsigned)transpose_tiles(C)), dim3(TT_THREADS), 0, st, C, state, fbar, tab,
                     state_bar, accumulate);
  return check_hip(hipGetLastError(), "adjoint transpose kernel launch");
}

int bssn_project_vjp(const bssn_grid *grid, const double *state, double *state_bar, void *stream) {
  if (!grid || grid->nx < 1 || grid->ny < 1 || grid->nz < 1) return fail(BSSN_ERR_INVALID, "invalid grid");
  if (!state || !state_bar) return fail(BSSN_ERR_INVALID, "null buffer");
  const int64_t n = (int64_t)grid->nx * grid->ny * grid->nz;
  hipLaunchKernelGGL(k_project_vjp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n, state,
                     state_bar);
  return check_hip(hipGetLastError(), "adjoint projection kernel launch");
}

}  // extern "C"
