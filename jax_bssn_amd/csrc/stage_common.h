// Pieces of a stage shared by the kernels and the host build used by the
// tests: parameter resolution, the Sommerfeld shell geometry, the algebraic
// projection, the Sommerfeld face RHS and the stage-update emitter.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "bssn_amd.h"
#include "bssn_common.h"
#include "rhs_point.h"
#include "stencil_providers.h"

namespace bssn {

// ------------------------------------------------------------ projection

// Reference time_evolve.py:94-131: gamma <- gamma det^(-1/3); then
// A <- A - 1/3 gamma (gamma^ij A_ij) with the inverse of the rescaled metric.
template <typename V>
BSSN_HD __forceinline__ void project(V (&g)[6], V (&A)[6]) {
  const V det = g[0] * (g[3] * g[5] - g[4] * g[4]) - g[1] * (g[1] * g[5] - g[4] * g[2]) +
                g[2] * (g[1] * g[4] - g[3] * g[2]);
  const V s = V(1.0) / cbrt(det);
#pragma unroll
  for (int c = 0; c < 6; ++c) g[c] = g[c] * s;
  V ig[6];
  inverse_metric(g, ig);
  const V tr = contract(ig, A);
#pragma unroll
  for (int c = 0; c < 6; ++c) A[c] = A[c] - V(1.0 / 3.0) * g[c] * tr;
}

// ------------------------------------------------------------ stage update

// Receives each component's RHS as soon as it is final and applies the
// stage update of bssn_rk4_stage (include/bssn_amd.h) right away; the
// metric and A are held until done() so they can be projected together.
// The RK4 stage combination as two fused multiply-adds per component
// (reference time_evolve.py:165-215 with an accumulating ustar):
//   stage 1: ustar = u0 + dt/6 k,      u_next = u0 + dt/2 k
//   stage 2: ustar = ustar + dt/3 k,   u_next = u0 + dt/2 k
//   stage 3: ustar = ustar + dt/3 k,   u_next = u0 + dt k
//   stage 4:                           u_next = ustar + dt/6 k
inline void stage_update(StageOut &O, double dt) {
  const int m = O.mode;
  O.nbase = (m == 4) ? O.ustar : O.u0;
  O.sbase = (m == 1) ? O.u0 : ((m == 2 || m == 3) ? O.ustar : nullptr);
  O.wn = (m == 3) ? dt : ((m == 4) ? dt / 6.0 : 0.5 * dt);
  O.ws = (m == 1) ? dt / 6.0 : dt / 3.0;
}

template <class S>
struct Emitter {
  int64_t n, idx;
  double wn, ws;
  const BSSN_GLOBAL double *nbase, *sbase;
  BSSN_GLOBAL double *ustar, *unext, *rhs;
  const BSSN_GLOBAL double *kap;
  const BSSN_GLOBAL double *mat;
  int mode, reproject;
  bool active, face;
  S st;  // holds the metric of the stage update until done() (its own copy of the handle)
  double An[6];
  double pn[NCOMP], ps[NCOMP];  // stage-update operands loaded by prefetch()

  BSSN_HD __forceinline__ Emitter(const Ctx &C, const StageOut &O, int64_t i, bool act, bool on_face, const S &stash)
      : n(C.ncell), idx(i), wn(O.wn), ws(O.ws), nbase((const BSSN_GLOBAL double *)O.nbase),
        sbase((const BSSN_GLOBAL double *)O.sbase), ustar((BSSN_GLOBAL double *)O.ustar),
        unext((BSSN_GLOBAL double *)O.unext), rhs((BSSN_GLOBAL double *)O.rhs),
        kap((const BSSN_GLOBAL double *)O.kap), mat((const BSSN_GLOBAL double *)O.matter), mode(O.mode),
        reproject(O.reproject), active(act), face(on_face), st(stash) {}

  // Load the stage-update operands of component c at this point (the
  // u_next and ustar bases) ahead of its RHS, so that their latency
  // overlaps the component's stencils.  Every put / put_face of c follows
  // a prefetch(c).
  BSSN_HD __forceinline__ void prefetch(int c) {
    if (!active || mode == 0) return;
    const int64_t at = c * n + idx;
    pn[c] = nbase[at];
    ps[c] = sbase ? sbase[at] : 0.0;
  }

  // bulk RHS of component c (ignored on a Sommerfeld face point)
  BSSN_HD __forceinline__ void put(int c, double k) {
    if (!face) write(c, k);
  }
  // the Sommerfeld RHS of component c on a face point
  BSSN_HD __forceinline__ void put_face(int c, double k) { write(c, k); }

  BSSN_HD __forceinline__ void write(int c, double k) {
    if (!active) return;
    const int64_t at = c * n + idx;
    if (mode == 0) {
      rhs[at] = k;
      return;
    }
    const double un = pn[c] + wn * k;
    if (sbase) ustar[at] = ps[c] + ws * k;
    if (c <= GZZ) {
      st.put(ST_GN + c, un);
    } else if (c >= AXX && c <= AZZ) {
      An[c - AXX] = un;
    } else {
      unext[at] = un;
    }
  }

  // matter source k (rho, S_i, S_ij packed) at this point
  BSSN_HD __forceinline__ bool has_matter() const { return mat != nullptr; }
  BSSN_HD __forceinline__ double matter(int k) const { return mat[k * n + idx]; }

  // constraint-damping term of A_ij (component c = 0..5), zero without kappa
  BSSN_HD __forceinline__ double extra_A(int c) const { return kap ? kap[c * n + idx] : 0.0; }

  BSSN_HD __forceinline__ void done() {
    if (!active || mode == 0) return;
    double gn[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) gn[c] = st.get(ST_GN + c);
    project(gn, An);
    if (mode == 4 && reproject) project(gn, An);
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      unext[(GXX + c) * n + idx] = gn[c];
      unext[(AXX + c) * n + idx] = An[c];
    }
  }
};

BSSN_HD __forceinline__ bool in_shell(const Ctx &C, int x, int y, int z) {
  const int p[3] = {x, y, z};
  bool s = false;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const Axis &A = C.ax[a];
    const int lo = p[a] - A.start, hi = A.end - 1 - p[a];
    s |= (lo >= 0 && lo < A.shell_lo) || (hi >= 0 && hi < A.shell_hi);
  }
  return s;
}

BSSN_HD __forceinline__ bool sommerfeld_point(const Ctx &C, int x, int y, int z) {
  const int p[3] = {x, y, z};
  bool face = false, owned = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const Axis &A = C.ax[a];
    face |= (A.somm_lo && p[a] == A.start) || (A.somm_hi && p[a] == A.end - 1);
    owned &= (p[a] >= A.start && p[a] < A.end);
  }
  return face && owned;
}

struct Slabs {
  int count;
  int lo[6][3], hi[6][3];
  int64_t first[7];  // prefix sums of slab sizes
};


#define BSSN_FAIL(code, ...) (std::snprintf(err, errlen, __VA_ARGS__), (code))

// Reference evolution/spatial_derivatives.py:18-41 (_boundary_data),
// evolution/boundaries.py:44-72 (mask) and the validators of
// parallel/initialization.py:31-37 / parallel/layout.py:92-108.
inline int resolve(const bssn_params *p, const bssn_grid *grid, Ctx *C, char *err, size_t errlen) {
  if (!p || !grid || !C) return BSSN_FAIL(BSSN_ERR_INVALID, "null argument");
  const int n[3] = {grid->nx, grid->ny, grid->nz};
  for (int d = 0; d < 3; ++d)
    if (n[d] < 1) return BSSN_FAIL(BSSN_ERR_INVALID, "grid extent %d along axis %d must be positive", n[d], d);
  const int64_t ncell = (int64_t)n[0] * n[1] * n[2];
  if (ncell >= (int64_t(1) << 31)) return BSSN_FAIL(BSSN_ERR_INVALID, "grid of %lld points exceeds 2^31", (long long)ncell);
  if (!(p->dx > 0.0) || !std::isfinite(p->dx)) return BSSN_FAIL(BSSN_ERR_INVALID, "dx must be finite and positive");
  if (p->guard_cells < 0) return BSSN_FAIL(BSSN_ERR_INVALID, "guard_cells must be non-negative");
  std::memset(C, 0, sizeof(Ctx));
  const bool owned_given = p->owned_shape[0] >= 0;
  int any_shell = 0;
  for (int d = 0; d < 3; ++d) {
    Axis &A = C->ax[d];
    const int lc = p->bc[2 * d], rc = p->bc[2 * d + 1];
    if ((lc != 0 && lc != 1) || (rc != 0 && rc != 1))
      return BSSN_FAIL(BSSN_ERR_INVALID, "boundary codes must be 0 (periodic) or 1 (Sommerfeld)");
    const int count = owned_given ? p->owned_shape[d] : n[d];
    A.n = n[d];
    A.start = p->guard_cells;
    A.end = A.start + count;
    if (count < 1 || A.end > n[d])
      return BSSN_FAIL(BSSN_ERR_INVALID, "owned range [%d, %d) does not fit axis %d of extent %d", A.start, A.end, d, n[d]);
    const bool global_somm = lc == 1 || rc == 1;
    A.q = global_somm ? 1.0 : p->mad_q;
    A.lbc = (lc == 1 && p->owned_faces[2 * d]) ? 1 : 0;
    A.rbc = (rc == 1 && p->owned_faces[2 * d + 1]) ? 1 : 0;
    if (p->reduced_axes[0] >= 0)
      A.reduced = p->reduced_axes[d] ? 1 : 0;
    else
      A.reduced = (owned_given && count == 1) ? 1 : 0;
    A.somm_lo = (A.lbc && (!owned_given || count > 1)) ? 1 : 0;
    A.somm_hi = (A.rbc && (!owned_given || count > 1)) ? 1 : 0;
    if (A.reduced && global_somm)
      return BSSN_FAIL(BSSN_ERR_INVALID, "axis %d is reduced and cannot carry a Sommerfeld face", d);
    // the one-sided rows reach seven samples inward from an owned face; in
    // a tile they may read the neighbour-filled guards (reference
    // parallel/initialization.py:31-37 counts min(extent, tile + guards))
    {
      const int reach = A.lbc && A.rbc ? count : (A.lbc ? n[d] - A.start : (A.rbc ? A.end : 8));
      if (reach < 8)
        return BSSN_FAIL(BSSN_ERR_INVALID,
                         "Sommerfeld closures on axis %d need at least 8 accessible inward samples (have %d)", d,
                         reach);
    }
    A.shell_lo = A.lbc ? 3 : 0;
    A.shell_hi = A.rbc ? 3 : 0;
    any_shell |= A.lbc | A.rbc;
  }
  C->ax[0].stride = n[1] * n[2];
  C->ax[1].stride = n[2];
  C->ax[2].stride = 1;
  C->ncell = ncell;
  C->dx = p->dx;
  C->inv_dx = 1.0 / p->dx;
  C->inv_dx2 = 1.0 / (p->dx * p->dx);
  C->ko = p->nu / 64.0 * std::pow(p->dx, 5) / std::pow(p->dx, 6);
  C->eta = p->eta;
  C->g = p->g;
  C->kappa = p->kappa;
  C->dt = p->dt;
  C->dt6 = p->dt / 6.0;
  C->dt3 = p->dt / 3.0;
  C->dt2 = 0.5 * p->dt;
  C->xmin[0] = p->x_min;
  C->xmin[1] = p->y_min;
  C->xmin[2] = p->z_min;
  C->gauge = p->gauge;
  C->zero_shift = p->zero_shift;
  C->any_shell = any_shell;
  if (p->gauge != 0 && p->gauge != 1) return BSSN_FAIL(BSSN_ERR_INVALID, "gauge must be 0 (harmonic) or 1 (1+log)");
  return BSSN_OK;
}

inline Slabs make_slabs(const Ctx &C) {
  Slabs S;
  std::memset(&S, 0, sizeof S);
  int64_t total = 0;
  for (int a = 0; a < 3; ++a) {
    for (int side = 0; side < 2; ++side) {
      const Axis &A = C.ax[a];
      const int w = side == 0 ? A.shell_lo : A.shell_hi;
      if (!w) continue;
      const int s = S.count++;
      for (int b = 0; b < 3; ++b) {
        S.lo[s][b] = 0;
        S.hi[s][b] = C.ax[b].n;
      }
      if (side == 0) {
        S.lo[s][a] = A.start;
        S.hi[s][a] = A.start + w;
      } else {
        S.lo[s][a] = A.end - w;
        S.hi[s][a] = A.end;
      }
      S.first[s] = total;
      total += (int64_t)(S.hi[s][0] - S.lo[s][0]) * (S.hi[s][1] - S.lo[s][1]) * (S.hi[s][2] - S.lo[s][2]);
    }
  }
  S.first[S.count] = total;
  return S;
}

#undef BSSN_FAIL

// ------------------------------------------------- constraint damping

// Derived grid functions of the kappa term (reference
// bssn/constraints.py:100-122): D[0..5] = gamma^ij, D[6 + 3i + j] =
// B_i^j = gamma^jk A_ik, D[15 + 3i + j] = B_i^j / max(W, 1e-12).
constexpr int NDERIVED = 24;

BSSN_HD inline void derived_point(int64_t n, int64_t i, const double *__restrict__ us, double *D) {
  double g[6], A[6], ig[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    g[c] = us[(GXX + c) * n + i];
    A[c] = us[(AXX + c) * n + i];
  }
  inverse_metric(g, ig);
  const double W = us[CW * n + i];
  const double Wf = W > 1.0e-12 ? W : 1.0e-12;
#pragma unroll
  for (int c = 0; c < 6; ++c) D[c * n + i] = ig[c];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double B = ig[sym(j, 0)] * A[sym(r, 0)] + ig[sym(j, 1)] * A[sym(r, 1)] + ig[sym(j, 2)] * A[sym(r, 2)];
      D[(6 + 3 * r + j) * n + i] = B;
      D[(15 + 3 * r + j) * n + i] = B / Wf;
    }
}

// kappa/2 alpha (D_j M_i + D_i M_j) (reference bssn/extrinsic_curvature.py:
// 204-218) with M_i and the product-rule d_l M_i of bssn/constraints.py:
// 22-97 and 100-213.  Mixed second derivatives keep the reference
// orientation: d2X[l, inner] is the outer derivative l of the inner one.
// Written once for any provider with stage_comp / d1 / d2 / dmix: ps reads
// the state, pd the derived fields D, pm the matter sources (may be null).
// Every provider call is made by every lane (the tile providers stage
// through LDS with barriers); kap is written only when ``write``.
template <class P, class PM>
BSSN_HD __forceinline__ void kappa_eval(const Ctx &C, P &ps, P &pd, PM *pm, const double *__restrict__ us,
                                        const double *__restrict__ D, const double *__restrict__ matter, int64_t i,
                                        double *kap, bool write) {
  // The fields are staged one at a time and every term is added as soon as
  // its factors are known (A_jk paired with the derivative of gamma^jk that
  // multiplies it, the metric last for the Christoffel term), so that few
  // values stay live across the stagings.
  const int64_t n = C.ncell;
  double ig[6];
  for (int c = 0; c < 6; ++c) ig[c] = D[c * n + i];
  const double W = us[CW * n + i], al = us[ALPHA * n + i];
  const double Wf = W > 1.0e-12 ? W : 1.0e-12;
  auto second = [&](P &pv, int inner, int outer) -> double {
    return inner == outer ? pv.d2(outer) : pv.dmix(inner, outer);
  };
  double dW[3];
  double M[3] = {0, 0, 0}, dM[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  ps.stage_comp(CW);
  for (int j = 0; j < 3; ++j) dW[j] = ps.d1(j);
  // -3 (B/W)_ij d_l d_j W; -3 B_r^j d_j W / W
  for (int j = 0; j < 3; ++j)
    for (int l = 0; l < 3; ++l) {
      const double s2 = second(ps, j, l);
      for (int r = 0; r < 3; ++r) dM[r][l] -= 3.0 * D[(15 + 3 * r + j) * n + i] * s2;
    }
  for (int r = 0; r < 3; ++r) {
    double t3 = 0.0;
    for (int j = 0; j < 3; ++j) t3 += D[(6 + 3 * r + j) * n + i] * dW[j];
    M[r] -= 3.0 * t3 / Wf;
  }
  // -2/3 d_r K, -2/3 d_l d_r K
  ps.stage_comp(CK);
  for (int r = 0; r < 3; ++r) M[r] -= 2.0 / 3.0 * ps.d1(r);
  for (int r = 0; r < 3; ++r)
    for (int l = 0; l < 3; ++l) dM[r][l] -= 2.0 / 3.0 * second(ps, r, l);
  // -1/2 gamma^jk d_r A_jk, -1/2 gamma^jk d_l d_r A_jk - 1/2 d_l gamma^jk d_r A_jk
  for (int c = 0; c < 6; ++c) {
    const double mult = SYM_I[c] == SYM_J[c] ? 1.0 : 2.0, w = mult * ig[c];
    ps.stage_comp(AXX + c);
    double dA[3];
    for (int d = 0; d < 3; ++d) dA[d] = ps.d1(d);
    for (int r = 0; r < 3; ++r) {
      M[r] -= 0.5 * w * dA[r];
      for (int l = 0; l < 3; ++l) dM[r][l] -= 0.5 * w * second(ps, r, l);
    }
    pd.stage_comp(c);
    for (int l = 0; l < 3; ++l) {
      const double dinv = pd.d1(l);
      for (int r = 0; r < 3; ++r) dM[r][l] -= 0.5 * mult * dinv * dA[r];
    }
  }
  // d_j B_r^j, d_l d_j B_r^j - 3 d_l (B/W)_r^j d_j W
  for (int r = 0; r < 3; ++r) {
    for (int j = 0; j < 3; ++j) {
      pd.stage_comp(6 + 3 * r + j);
      M[r] += pd.d1(j);
      for (int l = 0; l < 3; ++l) dM[r][l] += second(pd, j, l);
      pd.stage_comp(15 + 3 * r + j);
      for (int l = 0; l < 3; ++l) dM[r][l] -= 3.0 * pd.d1(l) * dW[j];
    }
  }
  if (pm) {
    // M_i - 8 pi S_i and d_l M_i - 8 pi d_l S_i (bssn/constraints.py:216-259)
    for (int r = 0; r < 3; ++r) {
      pm->stage_comp(1 + r);
      M[r] -= 8.0 * M_PI * matter[(1 + r) * n + i];
      for (int l = 0; l < 3; ++l) dM[r][l] -= 8.0 * M_PI * pm->d1(l);
    }
  }
  // D_j M_i = d_j M_i - Gamma^k_ij M_k
  double dg[6][3];
  for (int c = 0; c < 6; ++c) {
    ps.stage_comp(GXX + c);
    for (int d = 0; d < 3; ++d) dg[c][d] = ps.d1(d);
  }
  double C1[3][6];
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 6; ++c) C1[a][c] = 0.5 * (dg[sym(a, SYM_I[c])][SYM_J[c]] + dg[sym(a, SYM_J[c])][SYM_I[c]] - dg[c][a]);
  double DM[3][3];
  for (int r = 0; r < 3; ++r)
    for (int j = 0; j < 3; ++j) {
      double gm = 0.0;
      for (int k = 0; k < 3; ++k) {
        const double c2 = ig[sym(k, 0)] * C1[0][sym(r, j)] + ig[sym(k, 1)] * C1[1][sym(r, j)] +
                          ig[sym(k, 2)] * C1[2][sym(r, j)];
        gm += c2 * M[k];
      }
      DM[r][j] = dM[r][j] - gm;
    }
  if (!write) return;
  for (int c = 0; c < 6; ++c) {
    const int a = SYM_I[c], b = SYM_J[c];
    kap[c * n + i] = C.kappa / 2.0 * al * (DM[a][b] + DM[b][a]);
  }
}

// The kappa term at point i with taps from global memory (any point).
BSSN_HD inline void kappa_point(const Ctx &C, const double *__restrict__ us, const double *__restrict__ D,
                                const double *__restrict__ matter, int64_t i, double *kap) {
  const int z = (int)(i % C.ax[2].n);
  const int y = (int)((i / C.ax[2].n) % C.ax[1].n);
  const int x = (int)(i / ((int64_t)C.ax[2].n * C.ax[1].n));
  PointProvider<double> ps, pd, pm;
  ps.setup(C, us, x, y, z);
  pd.setup(C, D, x, y, z);
  if (matter) pm.setup(C, matter, x, y, z);
  kappa_eval(C, ps, pd, matter ? &pm : (PointProvider<double> *)nullptr, us, D, matter, i, kap, true);
}

}  // namespace bssn
