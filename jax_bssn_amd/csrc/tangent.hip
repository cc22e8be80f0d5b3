// Forward-mode differentiation of the evolution: the exact tangent-linear
// RK4 step along ND = 1 or 2 directions at once, and the tangent of the
// Hamiltonian constraint.
//
// The per-point RHS (rhs_point.h), the stage update and the projection are
// the same templates the evolution kernels use, instantiated with DualN
// (value and ND directional derivatives).  Stencils are linear, so a
// derivative of a dual field is the dual of the value's and the tangents'
// derivatives: the tile kernel stages the value and the ND tangents of a
// component in LDS side by side and runs the ordinary tile stencils on each
// (the upwind side follows the value's shift).  Configurations the tile
// kernels do not cover (MAD blend, closure faces off the tile grid) take the
// point kernel, whose taps come straight from global memory.  With p
// initial-data parameters, ceil(p / 2) tangent runs give the exact gradient
// of any output of the evolution (the torch.autograd.Function in
// jax_bssn_amd/tangent.py).  Restricted to kappa = 0 and vacuum.
//
// This unit holds the point kernels, the projection, the Hamiltonian
// tangent and the entry points; the tile kernels are in tangent_common.h.
#include "tangent_common.h"

namespace bssn {
namespace {

// Stencil taps of a value state and its ND tangents (direction l at
// dsrc + l * dstride), through the same rows (closures, MAD blend, upwind
// replacement) as PointProvider.
template <int ND>
struct DualPointProvider {
  using V = DualN<ND>;
  StashT<V> st;
  const Ctx *C;
  const double *src, *dsrc, *f, *df;
  int64_t dstride;
  int p[3];
  int64_t idx;
  bool isface;

  __device__ __forceinline__ void setup(const Ctx &ctx, const double *u, const double *du, int x, int y, int z) {
    C = &ctx;
    src = f = u;
    dsrc = df = du;
    dstride = (int64_t)NCOMP * ctx.ncell;
    p[0] = x, p[1] = y, p[2] = z;
    idx = (int64_t)x * ctx.ax[0].stride + (int64_t)y * ctx.ax[1].stride + z;
    bool fc = false, owned = true;
    for (int a = 0; a < 3; ++a) {
      const Axis &A = ctx.ax[a];
      fc |= (A.somm_lo && p[a] == A.start) || (A.somm_hi && p[a] == A.end - 1);
      owned &= p[a] >= A.start && p[a] < A.end;
    }
    isface = fc && owned;
  }
  __device__ __forceinline__ bool face() const { return isface; }
  __device__ __forceinline__ void stage_comp(int comp) {
    f = src + comp * C->ncell;
    df = dsrc + comp * C->ncell;
  }
  __device__ __forceinline__ void stage(int v) { stage_comp(visit_comp(v)); }
  __device__ __forceinline__ V at_index(int64_t j) const {
    V r;
    r.v = f[j];
#pragma unroll
    for (int l = 0; l < ND; ++l) r.d[l] = df[l * dstride + j];
    return r;
  }
  __device__ __forceinline__ V ld(int a, int oa) const {
    int q[3] = {p[0], p[1], p[2]};
    q[a] = wrap_index(q[a] + oa, C->ax[a].n);
    return at_index((int64_t)q[0] * C->ax[0].stride + (int64_t)q[1] * C->ax[1].stride + q[2]);
  }
  __device__ __forceinline__ V ld2(int a, int oa, int b, int ob) const {
    int q[3] = {p[0], p[1], p[2]};
    q[a] = wrap_index(q[a] + oa, C->ax[a].n);
    q[b] = wrap_index(q[b] + ob, C->ax[b].n);
    return at_index((int64_t)q[0] * C->ax[0].stride + (int64_t)q[1] * C->ax[1].stride + q[2]);
  }
  __device__ __forceinline__ V apply(int a, const Row8<double> &r) const {
    V s(0.0);
    for (int k = 0; k < r.n; ++k)
      if (r.c[k] != 0.0) s += r.c[k] * ld(a, r.o0 + k);
    return s;
  }
  __device__ __forceinline__ V val() const { return at_index(idx); }
  __device__ __forceinline__ V d1(int a) const {
    if (C->ax[a].reduced) return V(0.0);
    return apply(a, row_d1<double>(C->ax[a], p[a], C->inv_dx, false));
  }
  __device__ __forceinline__ V d2(int a) const {
    if (C->ax[a].reduced) return V(0.0);
    return apply(a, row_d2<double>(C->ax[a], p[a], C->inv_dx2));
  }
  __device__ __forceinline__ V dmix(int a, int b) const {
    if (C->ax[a].reduced || C->ax[b].reduced) return V(0.0);
    const Row8<double> ra = row_d1<double>(C->ax[a], p[a], C->inv_dx, false);
    const Row8<double> rb = row_d1<double>(C->ax[b], p[b], C->inv_dx, false);
    V s(0.0);
    for (int j = 0; j < rb.n; ++j) {
      if (rb.c[j] == 0.0) continue;
      V in(0.0);
      for (int i = 0; i < ra.n; ++i)
        if (ra.c[i] != 0.0) in += ra.c[i] * ld2(a, ra.o0 + i, b, rb.o0 + j);
      s += rb.c[j] * in;
    }
    return s;
  }
  __device__ __forceinline__ V upw(int a, V coeff) const {
    if (C->ax[a].reduced) return V(0.0);
    return apply(a, row_upwind<double>(C->ax[a], p[a], C->inv_dx, coeff.v));
  }
  __device__ __forceinline__ V ko() const {
    V s(0.0);
    for (int a = 0; a < 3; ++a) {
      if (C->ax[a].reduced) continue;
      s += apply(a, row_d6<double>(C->ax[a], p[a]));
    }
    return C->ko * s;
  }
  __device__ __forceinline__ V somm(int comp, V uc) const {
    const V d0 = d1(0), d1v = d1(1), d2v = d1(2);
    const double X = C->xmin[0] + C->dx * (double)p[0];
    const double Y = C->xmin[1] + C->dx * (double)p[1];
    const double Z = C->xmin[2] + C->dx * (double)p[2];
    const double r = ::sqrt(X * X + Y * Y + Z * Z);
    const double rs = r > 0.0 ? r : 1.0;
    const bool unit = (comp == GXX || comp == GYY || comp == GZZ || comp == CW || comp == ALPHA);
    const double u0 = unit ? 1.0 : 0.0;
    return -((X * d0 + Y * d1v + Z * d2v) / rs + (uc - u0) / rs);
  }
};

template <int ND>
__global__ void __launch_bounds__(128) k_stage_tangent(Ctx C, const double *__restrict__ us,
                                                       const double *__restrict__ dus, StageOut O, TangentOut T) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C.ncell) return;
  const int z = (int)(i % C.ax[2].n);
  const int y = (int)((i / C.ax[2].n) % C.ax[1].n);
  const int x = (int)(i / ((int64_t)C.ax[2].n * C.ax[1].n));
  DualPointProvider<ND> pv;
  pv.setup(C, us, dus, x, y, z);
  DualN<ND> u[NCOMP];
  load_dual(C, us, dus, i, u);
  DualEmitter<ND> out(C, O, T, i, true, pv.face());
  bssn_rhs_point<DualN<ND>>(pv, C, u, out);
}

template <int ND>
__global__ void __launch_bounds__(256) k_project_tangent(int64_t n, double *u, double *du) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t ds = (int64_t)NCOMP * n;
  DualN<ND> g[6], A[6];
  for (int c = 0; c < 6; ++c) {
    g[c].v = u[(GXX + c) * n + i];
    A[c].v = u[(AXX + c) * n + i];
#pragma unroll
    for (int l = 0; l < ND; ++l) {
      g[c].d[l] = du[l * ds + (GXX + c) * n + i];
      A[c].d[l] = du[l * ds + (AXX + c) * n + i];
    }
  }
  project(g, A);
  for (int c = 0; c < 6; ++c) {
    u[(GXX + c) * n + i] = g[c].v;
    u[(AXX + c) * n + i] = A[c].v;
#pragma unroll
    for (int l = 0; l < ND; ++l) {
      du[l * ds + (GXX + c) * n + i] = g[c].d[l];
      du[l * ds + (AXX + c) * n + i] = A[c].d[l];
    }
  }
}

// Hamiltonian constraint H = gamma^ij W^2 R_ij + 2/3 K^2 - A_ij A^ij
// (reference bssn/constraints.py:271-303, W^2 R_ij of geometry.py:21-119),
// with its tangent.
template <int ND>
__global__ void __launch_bounds__(128) k_hamiltonian_tangent(Ctx C, const double *__restrict__ us,
                                                             const double *__restrict__ dus, double *H,
                                                             double *dH) {
  using Dual = DualN<ND>;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C.ncell) return;
  const int z = (int)(i % C.ax[2].n);
  const int y = (int)((i / C.ax[2].n) % C.ax[1].n);
  const int x = (int)(i / ((int64_t)C.ax[2].n * C.ax[1].n));
  const int64_t n = C.ncell;
  const int64_t ds = (int64_t)NCOMP * n;
  DualPointProvider<ND> pv;
  pv.setup(C, us, dus, x, y, z);
  auto ld = [&](int c) {
    Dual r;
    r.v = us[c * n + i];
#pragma unroll
    for (int l = 0; l < ND; ++l) r.d[l] = dus[l * ds + c * n + i];
    return r;
  };
  Dual g[6], A[6], ig[6];
  for (int c = 0; c < 6; ++c) {
    g[c] = ld(GXX + c);
    A[c] = ld(AXX + c);
  }
  inverse_metric(g, ig);
  const Dual W = ld(CW), W2 = W * W, K = ld(CK);
  Dual R[6], dW[3];
  for (int c = 0; c < 6; ++c) R[c] = Dual(0.0);
  pv.stage_comp(CW);
  for (int j = 0; j < 3; ++j) dW[j] = pv.d1(j);
  {
    Dual s[6];
    second_derivatives(pv, s);
    const Dual lap = contract(ig, s);
    for (int c = 0; c < 6; ++c) R[c] += W * (s[c] + g[c] * lap);
  }
  for (int m = 0; m < 3; ++m) {
    pv.stage_comp(GAM + m);
    Dual dG[3];
    for (int j = 0; j < 3; ++j) dG[j] = pv.d1(j);
    for (int c = 0; c < 6; ++c) {
      const int a = SYM_I[c], b = SYM_J[c];
      R[c] += W2 * 0.5 * (g[sym(m, a)] * dG[b] + g[sym(m, b)] * dG[a]);
    }
  }
  Dual dg[6][3];
  for (int c = 0; c < 6; ++c) {
    pv.stage_comp(GXX + c);
    for (int d = 0; d < 3; ++d) dg[c][d] = pv.d1(d);
    Dual s[6];
    second_derivatives(pv, s);
    R[c] += -0.5 * W2 * contract(ig, s);
  }
  Dual C1[3][6], C2[3][6];
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 6; ++c) {
      const int j = SYM_I[c], k = SYM_J[c];
      C1[a][c] = 0.5 * (dg[sym(a, j)][k] + dg[sym(a, k)][j] - dg[c][a]);
    }
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 6; ++c)
      C2[a][c] = ig[sym(a, 0)] * C1[0][c] + ig[sym(a, 1)] * C1[1][c] + ig[sym(a, 2)] * C1[2][c];
  Dual G[3] = {ld(GAM), ld(GAM + 1), ld(GAM + 2)};
  {
    const Dual gradW2 = contract_vec(ig, dW, dW);
    Dual conW(0.0);
    for (int k = 0; k < 3; ++k) conW += contract(ig, C2[k]) * dW[k];
    for (int c = 0; c < 6; ++c) {
      const int a = SYM_I[c], b = SYM_J[c];
      Dual t3(0.0), t4(0.0), t5(0.0);
      for (int m = 0; m < 3; ++m) t3 += G[m] * (C1[a][sym(b, m)] + C1[b][sym(a, m)]);
      for (int m = 0; m < 3; ++m)
        for (int n2 = 0; n2 < 3; ++n2) {
          const Dual gmn = ig[sym(m, n2)];
          Dual acc4(0.0), acc5(0.0);
          for (int k = 0; k < 3; ++k) {
            acc4 += C2[k][sym(m, a)] * C1[b][sym(k, n2)] + C2[k][sym(m, b)] * C1[a][sym(k, n2)];
            acc5 += C2[k][sym(a, m)] * C1[k][sym(b, n2)];
          }
          t4 += gmn * acc4;
          t5 += gmn * acc5;
        }
      const Dual chW = C2[0][c] * dW[0] + C2[1][c] * dW[1] + C2[2][c] * dW[2];
      R[c] += W2 * (0.5 * t3 + t4 + t5) - W * (chW + g[c] * conW) - 2.0 * g[c] * gradW2;
    }
  }
  Dual Aup[6];
  for (int c = 0; c < 6; ++c) {
    const int a = SYM_I[c], b = SYM_J[c];
    Dual acc(0.0);
    for (int k = 0; k < 3; ++k)
      for (int l2 = 0; l2 < 3; ++l2) acc += ig[sym(a, k)] * ig[sym(b, l2)] * A[sym(k, l2)];
    Aup[c] = acc;
  }
  const Dual h = contract(ig, R) + (2.0 / 3.0) * K * K - contract(A, Aup);
  H[i] = h.v;
#pragma unroll
  for (int l = 0; l < ND; ++l) dH[l * n + i] = h.d[l];
}

int resolve_tangent(const bssn_params *p, const bssn_grid *grid, Ctx *C) {
  char err[512] = {0};
  const int rc = resolve(p, grid, C, err, sizeof err);
  if (rc) return fail(rc, "%s", err);
  if (C->kappa != 0.0) return fail(BSSN_ERR_UNSUPPORTED, "the tangent evolution takes kappa = 0");
  return BSSN_OK;
}

// The tile kernels cover the FP64 configurations of the evolution's tile
// path without the MAD blend: every tile (no closure shells), or interior
// and face tiles when the closure faces sit on the tile grid.
bool tangent_tiles(const Ctx &C) {
  const bool mad = C.ax[0].q != 1.0 || C.ax[1].q != 1.0 || C.ax[2].q != 1.0;
  return !mad && (!C.any_shell || tiles_aligned(C));
}

// The point kernels carry one direction; a stack of directions takes one
// launch per direction, the first storing the value as well (the value
// outputs of a stage may alias its inputs, e.g. ustar in stages 2 and 3, so
// they are written exactly once).  A point-kernel instantiation carrying two
// directions returned a wrong K right-hand side for whole workgroups on the
// device (cause not identified), so none is built.
int launch_tangent_points(const Ctx &C, int nd, const double *us, const double *dus, const StageOut &O,
                          const TangentOut &T, hipStream_t st) {
  const int64_t blocks = (C.ncell + 127) / 128, ds = (int64_t)NCOMP * C.ncell;
  auto shift = [&](auto *q, int l) { return q ? q + l * ds : q; };
  for (int l = 0; l < nd; ++l) {
    const TangentOut Tl{shift(T.du0, l), shift(T.dustar, l), shift(T.dunext, l), shift(T.drhs, l), l > 0};
    hipLaunchKernelGGL(k_stage_tangent<1>, dim3((unsigned)blocks), dim3(128), 0, st, C, us, dus + l * ds, O, Tl);
    const int rc = check_hip(hipGetLastError(), "tangent stage kernel launch");
    if (rc) return rc;
  }
  return BSSN_OK;
}

int launch_tangent_stage(const Ctx &C, int nd, const double *us, const double *dus, StageOut O, const TangentOut &T,
                         hipStream_t st) {
  stage_update(O, C.dt);
  if (tangent_tiles(C))
    return nd == 1 ? launch_tangent_tiles_1(C, us, dus, O, T, st) : launch_tangent_tiles_2(C, us, dus, O, T, st);
  return launch_tangent_points(C, nd, us, dus, O, T, st);
}

int check_dirs(int ndir) {
  if (ndir < 1 || ndir > MAX_DIRS) return fail(BSSN_ERR_INVALID, "ndir must be 1 or %d (got %d)", MAX_DIRS, ndir);
  return BSSN_OK;
}

}  // namespace
}  // namespace bssn

using namespace bssn;

extern "C" {

int bssn_rhs_tangent(const bssn_params *p, const bssn_grid *grid, int ndir, const double *state,
                     const double *dstate, double *rhs, double *drhs, void *stream) {
  Ctx C;
  int rc = resolve_tangent(p, grid, &C);
  if (rc || (rc = check_dirs(ndir))) return rc;
  if (!state || !dstate || !rhs || !drhs) return fail(BSSN_ERR_INVALID, "null buffer");
  StageOut O{nullptr, nullptr, nullptr, rhs, 0, 0, nullptr, nullptr};
  TangentOut T{nullptr, nullptr, nullptr, drhs};
  return launch_tangent_stage(C, ndir, state, dstate, O, T, (hipStream_t)stream);
}

int64_t bssn_tangent_work_size(const bssn_grid *grid, int ndir) {
  if (!grid || grid->nx <= 0 || grid->ny <= 0 || grid->nz <= 0) return fail(BSSN_ERR_INVALID, "invalid grid");
  if (ndir < 1 || ndir > MAX_DIRS) return fail(BSSN_ERR_INVALID, "ndir must be 1 or %d (got %d)", MAX_DIRS, ndir);
  return 3 * (int64_t)(1 + ndir) * NCOMP * (int64_t)grid->nx * grid->ny * grid->nz;
}

int bssn_evolve_tangent(const bssn_params *p, const bssn_grid *grid, int nsteps, int ndir, double *state,
                        double *dstate, double *work, void *stream) {
  Ctx C;
  int rc = resolve_tangent(p, grid, &C);
  if (rc || (rc = check_dirs(ndir))) return rc;
  if (nsteps < 0) return fail(BSSN_ERR_INVALID, "nsteps must be non-negative");
  if (!state || !dstate || !work) return fail(BSSN_ERR_INVALID, "null buffer");
  if (nsteps == 0) return BSSN_OK;
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = C.ncell * NCOMP, dn = n * ndir;
  double *ustar = work, *ua = work + n, *ub = work + 2 * n;
  double *dustar = work + 3 * n, *dua = dustar + dn, *dub = dua + dn;
  const int64_t pblocks = (C.ncell + 255) / 256;
  if (ndir == 1)
    hipLaunchKernelGGL(k_project_tangent<1>, dim3((unsigned)pblocks), dim3(256), 0, st, C.ncell, state, dstate);
  else
    hipLaunchKernelGGL(k_project_tangent<2>, dim3((unsigned)pblocks), dim3(256), 0, st, C.ncell, state, dstate);
  if ((rc = check_hip(hipGetLastError(), "tangent projection launch"))) return rc;
  for (int s = 0; s < nsteps; ++s) {
    const int last = s == nsteps - 1;
    if ((rc = launch_tangent_stage(C, ndir, state, dstate,
                                   StageOut{state, ustar, ua, nullptr, 1, 0, nullptr, nullptr},
                                   TangentOut{dstate, dustar, dua, nullptr}, st)))
      return rc;
    if ((rc = launch_tangent_stage(C, ndir, ua, dua, StageOut{state, ustar, ub, nullptr, 2, 0, nullptr, nullptr},
                                   TangentOut{dstate, dustar, dub, nullptr}, st)))
      return rc;
    if ((rc = launch_tangent_stage(C, ndir, ub, dub, StageOut{state, ustar, ua, nullptr, 3, 0, nullptr, nullptr},
                                   TangentOut{dstate, dustar, dua, nullptr}, st)))
      return rc;
    if ((rc = launch_tangent_stage(C, ndir, ua, dua,
                                   StageOut{nullptr, ustar, state, nullptr, 4, last ? 0 : 1, nullptr, nullptr},
                                   TangentOut{nullptr, dustar, dstate, nullptr}, st)))
      return rc;
  }
  return BSSN_OK;
}

int bssn_hamiltonian_tangent(const bssn_params *p, const bssn_grid *grid, int ndir, const double *state,
                             const double *dstate, double *H, double *dH, void *stream) {
  Ctx C;
  bssn_params q = p ? *p : bssn_params{};
  q.kappa = 0.0;  // the constraint does not depend on kappa
  int rc = resolve_tangent(&q, grid, &C);
  if (rc || (rc = check_dirs(ndir))) return rc;
  if (!state || !dstate || !H || !dH) return fail(BSSN_ERR_INVALID, "null buffer");
  const int64_t blocks = (C.ncell + 127) / 128;
  for (int l = 0; l < ndir; ++l) {  // one direction per launch, as the point stage kernel
    hipLaunchKernelGGL(k_hamiltonian_tangent<1>, dim3((unsigned)blocks), dim3(128), 0, (hipStream_t)stream, C,
                       state, dstate + l * NCOMP * C.ncell, H, dH + l * C.ncell);
    if ((rc = check_hip(hipGetLastError(), "tangent Hamiltonian kernel launch"))) return rc;
  }
  return BSSN_OK;
}

}  // extern "C"
