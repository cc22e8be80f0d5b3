// Forward-mode differentiation of the evolution: the exact tangent-linear
// RK4 step and the tangent of the Hamiltonian constraint.
//
// The per-point RHS (rhs_point.h), the stage update and the projection are
// the same templates the evolution kernels use, instantiated with Dual
// (value, derivative along one direction).  Stencil taps come straight
// from global memory (the value state and its tangent), so every point,
// closure and Sommerfeld face included, runs one kernel.  With p initial-
// data parameters, p tangent runs give the exact gradient of any output of
// the evolution (the torch.autograd.Function in jax_bssn_amd/tangent.py).
// Restricted to kappa = 0 and vacuum.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "bssn_amd.h"
#include "bssn_common.h"
#include "dual.h"
#include "host_errors.h"
#include "rhs_point.h"
#include "stage_common.h"
#include "stencil_providers.h"

namespace bssn {
namespace {

template <typename V>
struct StashT {
  V v[STASH_SLOTS];
  BSSN_HD __forceinline__ void put(int slot, V x) { v[slot] = x; }
  BSSN_HD __forceinline__ V get(int slot) const { return v[slot]; }
};

// Stencil taps of a value state and its tangent, through the same rows
// (closures, MAD blend, upwind replacement) as PointProvider.
struct DualPointProvider {
  StashT<Dual> st;
  const Ctx *C;
  const double *src, *dsrc, *f, *df;
  int p[3];
  int64_t idx;
  bool isface;

  __device__ __forceinline__ void setup(const Ctx &ctx, const double *u, const double *du, int x, int y, int z) {
    C = &ctx;
    src = f = u;
    dsrc = df = du;
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
  __device__ __forceinline__ Dual at_index(int64_t j) const { return Dual(f[j], df[j]); }
  __device__ __forceinline__ Dual ld(int a, int oa) const {
    int q[3] = {p[0], p[1], p[2]};
    q[a] = wrap_index(q[a] + oa, C->ax[a].n);
    return at_index((int64_t)q[0] * C->ax[0].stride + (int64_t)q[1] * C->ax[1].stride + q[2]);
  }
  __device__ __forceinline__ Dual ld2(int a, int oa, int b, int ob) const {
    int q[3] = {p[0], p[1], p[2]};
    q[a] = wrap_index(q[a] + oa, C->ax[a].n);
    q[b] = wrap_index(q[b] + ob, C->ax[b].n);
    return at_index((int64_t)q[0] * C->ax[0].stride + (int64_t)q[1] * C->ax[1].stride + q[2]);
  }
  __device__ __forceinline__ Dual apply(int a, const Row8<double> &r) const {
    Dual s(0.0);
    for (int k = 0; k < r.n; ++k)
      if (r.c[k] != 0.0) s += r.c[k] * ld(a, r.o0 + k);
    return s;
  }
  __device__ __forceinline__ Dual val() const { return at_index(idx); }
  __device__ __forceinline__ Dual d1(int a) const {
    if (C->ax[a].reduced) return Dual(0.0);
    return apply(a, row_d1<double>(C->ax[a], p[a], C->inv_dx, false));
  }
  __device__ __forceinline__ Dual d2(int a) const {
    if (C->ax[a].reduced) return Dual(0.0);
    return apply(a, row_d2<double>(C->ax[a], p[a], C->inv_dx2));
  }
  __device__ __forceinline__ Dual dmix(int a, int b) const {
    if (C->ax[a].reduced || C->ax[b].reduced) return Dual(0.0);
    const Row8<double> ra = row_d1<double>(C->ax[a], p[a], C->inv_dx, false);
    const Row8<double> rb = row_d1<double>(C->ax[b], p[b], C->inv_dx, false);
    Dual s(0.0);
    for (int j = 0; j < rb.n; ++j) {
      if (rb.c[j] == 0.0) continue;
      Dual in(0.0);
      for (int i = 0; i < ra.n; ++i)
        if (ra.c[i] != 0.0) in += ra.c[i] * ld2(a, ra.o0 + i, b, rb.o0 + j);
      s += rb.c[j] * in;
    }
    return s;
  }
  __device__ __forceinline__ Dual upw(int a, Dual coeff) const {
    if (C->ax[a].reduced) return Dual(0.0);
    return apply(a, row_upwind<double>(C->ax[a], p[a], C->inv_dx, coeff.v));
  }
  __device__ __forceinline__ Dual ko() const {
    Dual s(0.0);
    for (int a = 0; a < 3; ++a) {
      if (C->ax[a].reduced) continue;
      s += apply(a, row_d6<double>(C->ax[a], p[a]));
    }
    return C->ko * s;
  }
  __device__ __forceinline__ Dual somm(int comp, Dual uc) const {
    const Dual d0 = d1(0), d1v = d1(1), d2v = d1(2);
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

// Tangent buffers of one stage (the value buffers and weights are a StageOut).
struct TangentOut {
  const double *du0;
  double *dustar, *dunext, *drhs;
};

struct DualEmitter {
  int64_t n, idx;
  double wn, ws;
  const double *nbase, *sbase, *dnbase, *dsbase;
  double *ustar, *unext, *rhs, *dustar, *dunext, *drhs;
  int mode, reproject;
  bool active, face;
  Dual gn[6], An[6];

  __device__ __forceinline__ DualEmitter(const Ctx &C, const StageOut &O, const TangentOut &T, int64_t i,
                                         bool on_face)
      : n(C.ncell), idx(i), wn(O.wn), ws(O.ws), nbase(O.nbase), sbase(O.sbase), ustar(O.ustar), unext(O.unext),
        rhs(O.rhs), dustar(T.dustar), dunext(T.dunext), drhs(T.drhs), mode(O.mode), reproject(O.reproject),
        active(true), face(on_face) {
    dnbase = O.mode == 4 ? T.dustar : T.du0;
    dsbase = O.mode == 1 ? T.du0 : ((O.mode == 2 || O.mode == 3) ? T.dustar : nullptr);
  }
  __device__ __forceinline__ void put(int c, Dual k) {
    if (!face) write(c, k);
  }
  __device__ __forceinline__ void put_face(int c, Dual k) { write(c, k); }
  __device__ __forceinline__ void write(int c, Dual k) {
    const int64_t at = c * n + idx;
    if (mode == 0) {
      rhs[at] = k.v;
      drhs[at] = k.d;
      return;
    }
    const Dual un = Dual(nbase[at], dnbase[at]) + wn * k;
    if (sbase) {
      const Dual us = Dual(sbase[at], dsbase[at]) + ws * k;
      ustar[at] = us.v;
      dustar[at] = us.d;
    }
    if (c <= GZZ) {
      gn[c] = un;
    } else if (c >= AXX && c <= AZZ) {
      An[c - AXX] = un;
    } else {
      unext[at] = un.v;
      dunext[at] = un.d;
    }
  }
  __device__ __forceinline__ void done() {
    if (mode == 0) return;
    project(gn, An);
    if (mode == 4 && reproject) project(gn, An);
    for (int c = 0; c < 6; ++c) {
      unext[(GXX + c) * n + idx] = gn[c].v;
      dunext[(GXX + c) * n + idx] = gn[c].d;
      unext[(AXX + c) * n + idx] = An[c].v;
      dunext[(AXX + c) * n + idx] = An[c].d;
    }
  }
  __device__ __forceinline__ bool has_matter() const { return false; }
  __device__ __forceinline__ double matter(int) const { return 0.0; }
  __device__ __forceinline__ double extra_A(int) const { return 0.0; }
};

__global__ void __launch_bounds__(128) k_stage_tangent(Ctx C, const double *__restrict__ us,
                                                       const double *__restrict__ dus, StageOut O, TangentOut T) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C.ncell) return;
  const int z = (int)(i % C.ax[2].n);
  const int y = (int)((i / C.ax[2].n) % C.ax[1].n);
  const int x = (int)(i / ((int64_t)C.ax[2].n * C.ax[1].n));
  DualPointProvider pv;
  pv.setup(C, us, dus, x, y, z);
  Dual u[NCOMP];
  for (int c = 0; c < NCOMP; ++c) u[c] = Dual(us[c * C.ncell + i], dus[c * C.ncell + i]);
  DualEmitter out(C, O, T, i, pv.face());
  bssn_rhs_point<Dual>(pv, C, u, out);
}

__global__ void __launch_bounds__(256) k_project_tangent(int64_t n, double *u, double *du) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Dual g[6], A[6];
  for (int c = 0; c < 6; ++c) {
    g[c] = Dual(u[(GXX + c) * n + i], du[(GXX + c) * n + i]);
    A[c] = Dual(u[(AXX + c) * n + i], du[(AXX + c) * n + i]);
  }
  project(g, A);
  for (int c = 0; c < 6; ++c) {
    u[(GXX + c) * n + i] = g[c].v;
    du[(GXX + c) * n + i] = g[c].d;
    u[(AXX + c) * n + i] = A[c].v;
    du[(AXX + c) * n + i] = A[c].d;
  }
}

// Hamiltonian constraint H = gamma^ij W^2 R_ij + 2/3 K^2 - A_ij A^ij
// (reference bssn/constraints.py:271-303, W^2 R_ij of geometry.py:21-119),
// with its tangent.
__global__ void __launch_bounds__(128) k_hamiltonian_tangent(Ctx C, const double *__restrict__ us,
                                                             const double *__restrict__ dus, double *H,
                                                             double *dH) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C.ncell) return;
  const int z = (int)(i % C.ax[2].n);
  const int y = (int)((i / C.ax[2].n) % C.ax[1].n);
  const int x = (int)(i / ((int64_t)C.ax[2].n * C.ax[1].n));
  const int64_t n = C.ncell;
  DualPointProvider pv;
  pv.setup(C, us, dus, x, y, z);
  auto ld = [&](int c) { return Dual(us[c * n + i], dus[c * n + i]); };
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
  dH[i] = h.d;
}

int resolve_tangent(const bssn_params *p, const bssn_grid *grid, Ctx *C) {
  char err[512] = {0};
  const int rc = resolve(p, grid, C, err, sizeof err);
  if (rc) return fail(rc, "%s", err);
  if (C->kappa != 0.0) return fail(BSSN_ERR_UNSUPPORTED, "the tangent evolution takes kappa = 0");
  return BSSN_OK;
}

int launch_tangent_stage(const Ctx &C, const double *us, const double *dus, StageOut O, const TangentOut &T,
                         hipStream_t st) {
  stage_update(O, C.dt);
  const int64_t blocks = (C.ncell + 127) / 128;
  hipLaunchKernelGGL(k_stage_tangent, dim3((unsigned)blocks), dim3(128), 0, st, C, us, dus, O, T);
  return check_hip(hipGetLastError(), "tangent stage kernel launch");
}

}  // namespace
}  // namespace bssn

using namespace bssn;

extern "C" {

int bssn_rhs_tangent(const bssn_params *p, const bssn_grid *grid, const double *state, const double *dstate,
                     double *rhs, double *drhs, void *stream) {
  Ctx C;
  int rc = resolve_tangent(p, grid, &C);
  if (rc) return rc;
  if (!state || !dstate || !rhs || !drhs) return fail(BSSN_ERR_INVALID, "null buffer");
  StageOut O{nullptr, nullptr, nullptr, rhs, 0, 0, nullptr, nullptr};
  TangentOut T{nullptr, nullptr, nullptr, drhs};
  return launch_tangent_stage(C, state, dstate, O, T, (hipStream_t)stream);
}

int bssn_evolve_tangent(const bssn_params *p, const bssn_grid *grid, int nsteps, double *state, double *dstate,
                        double *work, void *stream) {
  Ctx C;
  int rc = resolve_tangent(p, grid, &C);
  if (rc) return rc;
  if (nsteps < 0) return fail(BSSN_ERR_INVALID, "nsteps must be non-negative");
  if (!state || !dstate || !work) return fail(BSSN_ERR_INVALID, "null buffer");
  if (nsteps == 0) return BSSN_OK;
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = C.ncell * NCOMP;
  double *ustar = work, *ua = work + n, *ub = work + 2 * n;
  double *dustar = work + 3 * n, *dua = work + 4 * n, *dub = work + 5 * n;
  const int64_t pblocks = (C.ncell + 255) / 256;
  hipLaunchKernelGGL(k_project_tangent, dim3((unsigned)pblocks), dim3(256), 0, st, C.ncell, state, dstate);
  if ((rc = check_hip(hipGetLastError(), "tangent projection launch"))) return rc;
  for (int s = 0; s < nsteps; ++s) {
    const int last = s == nsteps - 1;
    if ((rc = launch_tangent_stage(C, state, dstate, StageOut{state, ustar, ua, nullptr, 1, 0, nullptr, nullptr},
                                   TangentOut{dstate, dustar, dua, nullptr}, st)))
      return rc;
    if ((rc = launch_tangent_stage(C, ua, dua, StageOut{state, ustar, ub, nullptr, 2, 0, nullptr, nullptr},
                                   TangentOut{dstate, dustar, dub, nullptr}, st)))
      return rc;
    if ((rc = launch_tangent_stage(C, ub, dub, StageOut{state, ustar, ua, nullptr, 3, 0, nullptr, nullptr},
                                   TangentOut{dstate, dustar, dua, nullptr}, st)))
      return rc;
    if ((rc = launch_tangent_stage(C, ua, dua,
                                   StageOut{nullptr, ustar, state, nullptr, 4, last ? 0 : 1, nullptr, nullptr},
                                   TangentOut{nullptr, dustar, dstate, nullptr}, st)))
      return rc;
  }
  return BSSN_OK;
}

int bssn_hamiltonian_tangent(const bssn_params *p, const bssn_grid *grid, const double *state,
                             const double *dstate, double *H, double *dH, void *stream) {
  Ctx C;
  bssn_params q = p ? *p : bssn_params{};
  q.kappa = 0.0;  // the constraint does not depend on kappa
  int rc = resolve_tangent(&q, grid, &C);
  if (rc) return rc;
  if (!state || !dstate || !H || !dH) return fail(BSSN_ERR_INVALID, "null buffer");
  const int64_t blocks = (C.ncell + 127) / 128;
  hipLaunchKernelGGL(k_hamiltonian_tangent, dim3((unsigned)blocks), dim3(128), 0, (hipStream_t)stream, C, state,
                     dstate, H, dH);
  return check_hip(hipGetLastError(), "tangent Hamiltonian kernel launch");
}

}  // extern "C"
