// Host build of the point adjoint (jax_bssn_amd/csrc/adjoint_point.h) and
// of the forward maps it transposes, for the CPU tests: the RHS of a point
// from its features through the kernels' own bssn_rhs_point, and the
// Hamiltonian constraint from the features in the arrangement of the
// constraint kernel (t4 / t5 sums rather than the Q products of the
// adjoint's forward pass).  Test infrastructure only.
#include <hip/hip_runtime.h>

#include <cmath>

#include "adjoint_point.h"
#include "stage_common.h"

using namespace bssn;

namespace {

struct CaptureEmitter {
  double out[NCOMP];
  bool face;
  void prefetch(int) {}
  void put(int c, double k) {
    if (!face) out[c] = k;
  }
  void put_face(int c, double k) { out[c] = k; }
  bool has_matter() const { return false; }
  double matter(int) const { return 0.0; }
  double extra_A(int) const { return 0.0; }
  void done() {}
};

double hamiltonian_from_features(const double *f) {
  double g[6], A[6], ig[6];
  for (int c = 0; c < 6; ++c) {
    g[c] = f[F_VAL + GXX + c];
    A[c] = f[F_VAL + AXX + c];
  }
  inverse_metric(g, ig);
  const double W = f[F_VAL + CW], W2 = W * W, K = f[F_VAL + CK];
  double R[6] = {0, 0, 0, 0, 0, 0}, dW[3], dg[6][3], G[3];
  for (int j = 0; j < 3; ++j) {
    dW[j] = f[F_D1 + 3 * CW + j];
    G[j] = f[F_VAL + GAM + j];
  }
  {
    const double *s = f + F_D2 + 6 * d2_slot(CW);
    const double lap = contract(ig, *(const double(*)[6])s);
    for (int c = 0; c < 6; ++c) R[c] += W * (s[c] + g[c] * lap);
  }
  for (int m = 0; m < 3; ++m)
    for (int c = 0; c < 6; ++c) {
      const int a = SYM_I[c], b = SYM_J[c];
      R[c] += W2 * 0.5 * (g[sym(m, a)] * f[F_D1 + 3 * (GAM + m) + b] + g[sym(m, b)] * f[F_D1 + 3 * (GAM + m) + a]);
    }
  for (int c = 0; c < 6; ++c) {
    for (int d = 0; d < 3; ++d) dg[c][d] = f[F_D1 + 3 * (GXX + c) + d];
    R[c] += -0.5 * W2 * contract(ig, *(const double(*)[6])(f + F_D2 + 6 * d2_slot(GXX + c)));
  }
  double C1[3][6], C2[3][6];
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 6; ++c) {
      const int j = SYM_I[c], k = SYM_J[c];
      C1[a][c] = 0.5 * (dg[sym(a, j)][k] + dg[sym(a, k)][j] - dg[c][a]);
    }
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 6; ++c) C2[a][c] = ig[sym(a, 0)] * C1[0][c] + ig[sym(a, 1)] * C1[1][c] + ig[sym(a, 2)] * C1[2][c];
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
  double Aup[6];
  for (int c = 0; c < 6; ++c) {
    const int a = SYM_I[c], b = SYM_J[c];
    double acc = 0;
    for (int k = 0; k < 3; ++k)
      for (int l2 = 0; l2 < 3; ++l2) acc += ig[sym(a, k)] * ig[sym(b, l2)] * A[sym(k, l2)];
    Aup[c] = acc;
  }
  return contract(ig, R) + 2.0 / 3.0 * K * K - contract(A, Aup);
}

}  // namespace

// RHS (out, 21) and H (*h) of a point with features f; face selects the
// Sommerfeld RHS with the point at xr.
extern "C" int host_features_forward(const bssn_params *p, const bssn_grid *g, const double *f, int face,
                                     const double *xr, double *out, double *h, char *err, int errlen) {
  Ctx C;
  const int rc = resolve(p, g, &C, err, (size_t)errlen);
  if (rc) return rc;
  FeatureProvider<double> pv;
  pv.f = f;
  pv.c = 0;
  pv.isface = face != 0;
  for (int a = 0; a < 3; ++a) pv.xr[a] = xr[a];
  double u[NCOMP];
  for (int c = 0; c < NCOMP; ++c) u[c] = f[F_VAL + c];
  CaptureEmitter e;
  e.face = face != 0;
  bssn_rhs_point<double>(pv, C, u, e);
  for (int c = 0; c < NCOMP; ++c) out[c] = e.out[c];
  *h = hamiltonian_from_features(f);
  return 0;
}

// Sink that adds into an array and counts the puts of every slot.
struct CountingAdjoints {
  double *p;
  int count[NFEAT];
  void put(int k, double v) {
    p[k] += v;
    ++count[k];
  }
};

// The feature adjoints (fbar, NFEAT) of <rbar, RHS> + hbar H at a point.
// Returns 2 when the bulk adjoint does not put every feature adjoint
// exactly once (the kernels store rather than add them).
extern "C" int host_features_vjp(const bssn_params *p, const bssn_grid *g, const double *f, int face,
                                 const double *xr, const double *rbar, double hbar, double *fbar, char *err,
                                 int errlen) {
  Ctx C;
  const int rc = resolve(p, g, &C, err, (size_t)errlen);
  if (rc) return rc;
  for (int k = 0; k < NFEAT; ++k) fbar[k] = 0.0;
  CountingAdjoints out{fbar, {}};
  const ArrayFeatures<double> fr{f};
  rhs_vjp_point(C, fr, face ? (const double *)nullptr : rbar, hbar, out);
  for (int k = 0; k < NFEAT; ++k)
    if (out.count[k] != 1) {
      snprintf(err, (size_t)errlen, "feature adjoint %d put %d times", k, out.count[k]);
      return 2;
    }
  if (face) {
    double x[3] = {xr[0], xr[1], xr[2]};
    ArrayAdjoints<double> add{fbar};
    face_vjp_point(rbar, x, add);
  }
  return 0;
}

// Projection P of a point (g, A packed, in place) and the adjoint of P.
extern "C" void host_project(double *g, double *A) {
  double gg[6], AA[6];
  for (int c = 0; c < 6; ++c) gg[c] = g[c], AA[c] = A[c];
  project(gg, AA);
  for (int c = 0; c < 6; ++c) g[c] = gg[c], A[c] = AA[c];
}

extern "C" void host_project_vjp(const double *g, const double *A, double *gb, double *Ab) {
  double gg[6], AA[6], g_[6], A_[6];
  for (int c = 0; c < 6; ++c) gg[c] = g[c], AA[c] = A[c], g_[c] = gb[c], A_[c] = Ab[c];
  project_vjp(gg, AA, g_, A_);
  for (int c = 0; c < 6; ++c) gb[c] = g_[c], Ab[c] = A_[c];
}
