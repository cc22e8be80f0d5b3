// Reverse mode (the adjoint) of the per-point BSSN right-hand side and of
// the Hamiltonian constraint, written by hand on the point's features.
//
// The RHS at a point is a function of its features: the point values of
// the 21 components, their first derivatives, the second derivatives of
// the shift, metric, lapse and W, the upwind derivatives (side chosen by
// the sign of the shift at the point, held fixed as jax differentiates
// jnp.where) and the Kreiss-Oliger sums -- every stencil of
// bssn_rhs_point (rhs_point.h) applied to the state.  Given the adjoint
// rbar of the 21 RHS values (and hbar of H), rhs_vjp_point forms the
// adjoint of every feature; the transposed stencils (adjoint.hip) carry
// those back to the state.  On an active Sommerfeld face point the RHS is
// the radial Sommerfeld RHS of every component (reference
// evolution/boundaries.py:29-172), a function of the first derivatives and
// values only.
//
// The forward algebra is restated here on full 3 x 3 (x 3) arrays, the
// same equations as rhs_point.h (reference bssn/geometry.py:21-186,
// bssn/conformal_connection.py:19-162, bssn/extrinsic_curvature.py:31-224,
// bssn/spatial_metric.py:17-106, bssn/shift_and_lapse.py:65-132) and the
// Hamiltonian constraint of bssn/constraints.py:271-303; the reverse sweep
// follows it term by term.  Symmetric inputs enter as full matrices and
// their adjoints are folded back onto the stored upper triangle.  kappa = 0
// and vacuum only.
#pragma once

#include "bssn_common.h"
#include "rhs_point.h"
#include "stencil_providers.h"

namespace bssn {

// Feature slots of one point (NFEAT doubles).
constexpr int F_VAL = 0;   // 21 point values
constexpr int F_D1 = 21;   // d_a u_c, 21 x 3 (of A only on Sommerfeld face points)
constexpr int F_D2 = 84;   // 11 x 6: d_a d_b of shift (3), metric (6), lapse, W; slot sym(a, b)
constexpr int F_UP = 150;  // 21 x 3 upwind derivatives
constexpr int F_KO = 213;  // 21 Kreiss-Oliger sums (with the nu / 64 dx^-1 factor; not recorded by the
                           // stage kernels, whose adjoint buffer holds the effective rbar here)
constexpr int NFEAT = 234;

// slot of component c among the 11 with second derivatives, -1 otherwise
BSSN_HD __forceinline__ constexpr int d2_slot(int c) {
  return c >= BETA ? c - BETA : (c <= GZZ ? 3 + c : (c == ALPHA ? 9 : (c == CW ? 10 : -1)));
}

// fold the adjoint of a full symmetric matrix onto its stored upper triangle
template <typename T>
BSSN_HD __forceinline__ T fold_sym(const T (&M)[3][3], int c) {
  const int i = SYM_I[c], j = SYM_J[c];
  return i == j ? M[i][i] : M[i][j] + M[j][i];
}

template <typename T>
BSSN_HD __forceinline__ void unpack_sym(const T *p, int stride, T (&M)[3][3]) {
  for (int c = 0; c < 6; ++c) {
    const int i = SYM_I[c], j = SYM_J[c];
    M[i][j] = M[j][i] = p[c * stride];
  }
}

// Serves stored features to bssn_rhs_point: the RHS of a point from its
// features alone (the forward map whose adjoint rhs_vjp_point forms).
template <typename T>
struct FeatureProvider {
  RegStash st;
  const T *f;
  int c;
  bool isface;
  T xr[3];
  BSSN_HD __forceinline__ void stage(int v) { c = visit_comp(v); }
  BSSN_HD __forceinline__ bool face() const { return isface; }
  BSSN_HD __forceinline__ T val() const { return f[F_VAL + c]; }
  BSSN_HD __forceinline__ T d1(int a) const { return f[F_D1 + 3 * c + a]; }
  BSSN_HD __forceinline__ T d2(int a) const { return f[F_D2 + 6 * d2_slot(c) + sym(a, a)]; }
  BSSN_HD __forceinline__ T dmix(int a, int b) const { return f[F_D2 + 6 * d2_slot(c) + sym(a, b)]; }
  BSSN_HD __forceinline__ T upw(int a, T) const { return f[F_UP + 3 * c + a]; }
  BSSN_HD __forceinline__ T ko() const { return f[F_KO + c]; }
  BSSN_HD __forceinline__ T somm(int comp, double uc) const {
    const T r = sqrt(xr[0] * xr[0] + xr[1] * xr[1] + xr[2] * xr[2]);
    const T rs = r > T(0) ? r : T(1);
    const bool unit = (comp == GXX || comp == GYY || comp == GZZ || comp == CW || comp == ALPHA);
    const T u0 = unit ? T(1) : T(0);
    return -((xr[0] * d1(0) + xr[1] * d1(1) + xr[2] * d1(2)) / rs + ((T)uc - u0) / rs);
  }
};

// Feature access of the point adjoint: a reader f(k) returns feature k of
// the point, a sink out.put(k, v) receives the adjoint of feature k.  The
// kernels read the recorded feature grid functions and write the adjoint
// grid functions in place (strided by the number of points); the host
// checks use plain arrays.
template <typename T>
struct ArrayFeatures {
  const T *p;
  BSSN_HD __forceinline__ T operator()(int k) const { return p[k]; }
};

template <typename T>
struct ArrayAdjoints {  // adds into an array
  T *p;
  BSSN_HD __forceinline__ void put(int k, T v) { p[k] += v; }
};

template <typename T, typename FR>
BSSN_HD __forceinline__ void unpack_feature(const FR &f, int base, T (&M)[3][3]) {
  for (int c = 0; c < 6; ++c) {
    const int i = SYM_I[c], j = SYM_J[c];
    M[i][j] = M[j][i] = f(base + c);
  }
}

// Adjoint of the Sommerfeld face RHS -(x . d u_c + u_c - u0_c) / r of every
// component at a face point (xr: the point's coordinates), put into the
// value and first-derivative slots.
template <typename T, typename S>
BSSN_HD void face_vjp_point(const T *rbar, const T (&xr)[3], S &out) {
  const T r = sqrt(xr[0] * xr[0] + xr[1] * xr[1] + xr[2] * xr[2]);
  const T rs = r > T(0) ? r : T(1);
  for (int c = 0; c < NCOMP; ++c) {
    const T w = -rbar[c] / rs;
    out.put(F_VAL + c, w);
    for (int a = 0; a < 3; ++a) out.put(F_D1 + 3 * c + a, w * xr[a]);
  }
}

// Adjoint of the bulk RHS (rbar, 21 values, may be null) and of the
// Hamiltonian constraint (hbar) at a point with features f.  Puts every one
// of the NFEAT feature adjoints exactly once (zero for the first
// derivatives of A, which the bulk RHS does not read, and for the upwind
// and dissipation features when rbar is null), each as soon as it is
// final, so few adjoints are live at once.
template <typename T, typename FR, typename S>
BSSN_HD void rhs_vjp_point(const Ctx &C, const FR &f, const T *rbar, T hbar, S &out) {
  // ------------------------------------------------------------ inputs
  T g[3][3], A[3][3], G[3], b[3];
  unpack_feature(f, F_VAL + GXX, g);
  unpack_feature(f, F_VAL + AXX, A);
  for (int i = 0; i < 3; ++i) {
    G[i] = f(F_VAL + GAM + i);
    b[i] = f(F_VAL + BETA + i);
  }
  const T W = f(F_VAL + CW), K = f(F_VAL + CK), al = f(F_VAL + ALPHA);
  T db[3][3], dG[3][3], da[3], dW[3], dK[3], dg[3][3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      db[i][j] = f(F_D1 + 3 * (BETA + i) + j);
      dG[i][j] = f(F_D1 + 3 * (GAM + i) + j);
    }
  for (int k = 0; k < 3; ++k) {
    da[k] = f(F_D1 + 3 * ALPHA + k);
    dW[k] = f(F_D1 + 3 * CW + k);
    dK[k] = f(F_D1 + 3 * CK + k);
  }
  for (int c = 0; c < 6; ++c)
    for (int k = 0; k < 3; ++k) dg[SYM_I[c]][SYM_J[c]][k] = dg[SYM_J[c]][SYM_I[c]][k] = f(F_D1 + 3 * (GXX + c) + k);
  T ddW[3][3];
  unpack_feature(f, F_D2 + 6 * d2_slot(CW), ddW);

  // ----------------------------------------------------------- forward
  T gp[6], igp[6];
  for (int c = 0; c < 6; ++c) gp[c] = g[SYM_I[c]][SYM_J[c]];
  inverse_metric(gp, igp);
  T ig[3][3];
  unpack_sym(igp, 1, ig);
  const T divb = db[0][0] + db[1][1] + db[2][2];
  T C1[3][3][3], C2[3][3][3], Q[3][3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) C1[i][j][k] = T(0.5) * (dg[i][j][k] + dg[i][k][j] - dg[j][k][i]);
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        T acc = T(0);
        for (int l = 0; l < 3; ++l) acc += ig[k][l] * C1[l][i][j];
        C2[k][i][j] = acc;
      }
  for (int k = 0; k < 3; ++k)
    for (int n = 0; n < 3; ++n)
      for (int i = 0; i < 3; ++i) {
        T acc = T(0);
        for (int m = 0; m < 3; ++m) acc += ig[m][n] * C2[k][m][i];
        Q[k][n][i] = acc;
      }
  T lap = T(0), gradW2 = T(0), gWa = T(0), conW = T(0);
  for (int m = 0; m < 3; ++m)
    for (int n = 0; n < 3; ++n) {
      lap += ig[m][n] * ddW[m][n];
      gradW2 += ig[m][n] * dW[m] * dW[n];
      gWa += ig[m][n] * dW[m] * da[n];
    }
  T trC2[3];
  for (int k = 0; k < 3; ++k) {
    T acc = T(0);
    for (int m = 0; m < 3; ++m)
      for (int n = 0; n < 3; ++n) acc += ig[m][n] * C2[k][m][n];
    trC2[k] = acc;
    conW += acc * dW[k];
  }
  const T W2 = W * W;
  // R_ij and Hs_ij (W^2 D_i D_j alpha) with their partial derivatives in W
  // at fixed other inputs (RW, HW), so the reverse sweep needs neither the
  // second derivatives nor the sums again
  T R[3][3], Hs[3][3], RW[3][3], HW[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      T ddg[3][3];
      unpack_feature(f, F_D2 + 6 * d2_slot(GXX + sym(i, j)), ddg);
      const T dda = f(F_D2 + 6 * d2_slot(ALPHA) + sym(i, j));
      T r1 = T(0);
      for (int m = 0; m < 3; ++m)
        for (int n = 0; n < 3; ++n) r1 += ig[m][n] * ddg[m][n];
      T r3 = T(0), r4 = T(0), r6 = T(0), r7 = T(0), h3 = T(0);
      for (int m = 0; m < 3; ++m) {
        r3 += g[m][i] * dG[m][j] + g[m][j] * dG[m][i];
        r4 += G[m] * (C1[i][j][m] + C1[j][i][m]);
        r6 += C2[m][i][j] * dW[m];
        h3 += C2[m][i][j] * da[m];
      }
      for (int k = 0; k < 3; ++k)
        for (int n = 0; n < 3; ++n) r7 += Q[k][n][i] * (C1[j][k][n] + C1[k][j][n]) + Q[k][n][j] * C1[i][k][n];
      R[i][j] = T(-0.5) * W2 * r1 + W * (ddW[i][j] + g[i][j] * lap) + T(0.5) * W2 * r3 + T(0.5) * W2 * r4 -
                T(2) * g[i][j] * gradW2 - W * r6 + W2 * r7 - W * g[i][j] * conW;
      RW[i][j] = -W * r1 + ddW[i][j] + g[i][j] * lap + W * r3 + W * r4 - r6 + T(2) * W * r7 - g[i][j] * conW;
      const T dd = dW[i] * da[j] + dW[j] * da[i] - g[i][j] * gWa;
      Hs[i][j] = W2 * dda + W * dd - W2 * h3;
      HW[i][j] = T(2) * W * dda + dd - T(2) * W * h3;
    }
  T Aup[3][3], gA[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      T acc = T(0), acc2 = T(0);
      for (int k = 0; k < 3; ++k) {
        acc2 += ig[i][k] * A[k][j];
        for (int l = 0; l < 3; ++l) acc += ig[i][k] * ig[j][l] * A[k][l];
      }
      Aup[i][j] = acc;
      gA[i][j] = acc2;
    }
  T s1[3];
  for (int m = 0; m < 3; ++m) {
    T acc = T(0);
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) acc += C1[m][j][k] * Aup[j][k];
    s1[m] = acc;
  }
  T Xr[3][3], trX = T(0);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      Xr[i][j] = al * R[i][j] - Hs[i][j];
      trX += ig[i][j] * Xr[i][j];
    }
  const T Wf = W > T(1.0e-12) ? W : T(1.0e-12);

  // ----------------------------------------------------------- adjoints
  T g_[3][3] = {}, A_[3][3] = {}, ig_[3][3] = {}, R_[3][3] = {}, Hs_[3][3] = {}, Aup_[3][3] = {};
  T C1_[3][3][3] = {}, C2_[3][3][3] = {};
  T dG_[3][3] = {}, ddW_[3][3] = {};
  T G_[3] = {}, da_[3] = {}, dW_[3] = {};
  T W_ = T(0), K_ = T(0), al_ = T(0);
  T rGam[3] = {T(0), T(0), T(0)};
  {
    T db_[3][3] = {}, b_[3] = {}, dK_[3] = {};
    T divb_ = T(0);
    if (rbar) {
      // full output adjoints: the stored upper triangle of each symmetric RHS
      T rA[3][3] = {}, rg[3][3] = {};
      for (int c = 0; c < 6; ++c) {
        rA[SYM_I[c]][SYM_J[c]] = rbar[AXX + c];
        rg[SYM_I[c]][SYM_J[c]] = rbar[GXX + c];
      }
      const T rK = rbar[CK], rW = rbar[CW], ra = rbar[ALPHA];
      T rb[3];
      for (int i = 0; i < 3; ++i) {
        rGam[i] = rbar[GAM + i];
        rb[i] = C.zero_shift ? T(0) : rbar[BETA + i];
      }
      // advection and dissipation of every component: adv_c = sum_a b_a U_ca
      for (int c = 0; c < NCOMP; ++c) {
        const T rc = (c >= BETA && C.zero_shift) ? T(0) : rbar[c];
        T up[3];
        for (int a = 0; a < 3; ++a) up[a] = f(F_UP + 3 * c + a);
        for (int a = 0; a < 3; ++a) {
          out.put(F_UP + 3 * c + a, rc * b[a]);
          b_[a] += rc * up[a];
        }
        out.put(F_KO + c, rc);
      }
      // A: al K A - 2 al A gA + Xr - g trX / 3 + ALIE
      T Xr_[3][3], AgA_[3][3], trX_ = T(0), ALIE_[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          T AgA = T(0);
          for (int l = 0; l < 3; ++l) AgA += A[i][l] * gA[l][j];
          al_ += rA[i][j] * (K * A[i][j] - T(2) * AgA);
          K_ += al * rA[i][j] * A[i][j];
          A_[i][j] += al * K * rA[i][j];
          AgA_[i][j] = T(-2) * al * rA[i][j];
          Xr_[i][j] = rA[i][j];
          g_[i][j] += T(-1.0 / 3.0) * trX * rA[i][j];
          trX_ += T(-1.0 / 3.0) * g[i][j] * rA[i][j];
          ALIE_[i][j] = rA[i][j];
        }
      {
        T gA_[3][3] = {};
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            for (int l = 0; l < 3; ++l) {
              A_[i][l] += AgA_[i][j] * gA[l][j];
              gA_[l][j] += A[i][l] * AgA_[i][j];
            }
        for (int k = 0; k < 3; ++k)
          for (int j = 0; j < 3; ++j)
            for (int l = 0; l < 3; ++l) {
              ig_[k][l] += gA_[k][j] * A[l][j];
              A_[l][j] += ig[k][l] * gA_[k][j];
            }
      }
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          ig_[i][j] += trX_ * Xr[i][j];
          Xr_[i][j] += trX_ * ig[i][j];
          al_ += Xr_[i][j] * R[i][j];
          R_[i][j] += al * Xr_[i][j];
          Hs_[i][j] -= Xr_[i][j];
        }
      // ALIE = A_mi db_mj + A_mj db_mi - 2/3 A divb
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          for (int m = 0; m < 3; ++m) {
            A_[m][i] += ALIE_[i][j] * db[m][j];
            db_[m][j] += ALIE_[i][j] * A[m][i];
            A_[m][j] += ALIE_[i][j] * db[m][i];
            db_[m][i] += ALIE_[i][j] * A[m][j];
          }
          A_[i][j] += T(-2.0 / 3.0) * divb * ALIE_[i][j];
          divb_ += T(-2.0 / 3.0) * A[i][j] * ALIE_[i][j];
        }
      // K: - ig:Hs + al A:Aup + al K^2 / 3
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          ig_[i][j] -= rK * Hs[i][j];
          Hs_[i][j] -= rK * ig[i][j];
          al_ += rK * A[i][j] * Aup[i][j];
          A_[i][j] += rK * al * Aup[i][j];
          Aup_[i][j] += rK * al * A[i][j];
        }
      al_ += rK * K * K * T(1.0 / 3.0);
      K_ += rK * al * K * T(2.0 / 3.0);
      // connection: rG + 2 al ig s1 - 6 al / Wf Aup dW - 2 Aup da
      T s1_[3] = {};
      T Wf_ = T(0);
      for (int i = 0; i < 3; ++i) {
        T t2 = T(0), t3 = T(0);
        for (int j = 0; j < 3; ++j) {
          t2 += ig[i][j] * s1[j];
          t3 += Aup[i][j] * dW[j];
          ig_[i][j] += rGam[i] * T(2) * al * s1[j];
          s1_[j] += rGam[i] * T(2) * al * ig[i][j];
          Aup_[i][j] += rGam[i] * (T(-6) * al / Wf * dW[j] - T(2) * da[j]);
          dW_[j] += rGam[i] * T(-6) * al / Wf * Aup[i][j];
          da_[j] += rGam[i] * T(-2) * Aup[i][j];
        }
        al_ += rGam[i] * (T(2) * t2 - T(6) / Wf * t3);
        Wf_ += rGam[i] * T(6) * al / (Wf * Wf) * t3;
      }
      if (W > T(1.0e-12)) W_ += Wf_;
      for (int m = 0; m < 3; ++m)
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k) {
            C1_[m][j][k] += s1_[m] * Aup[j][k];
            Aup_[j][k] += s1_[m] * C1[m][j][k];
          }
      // rG: ig:s_i + 1/3 ig_im s_jmj + 2/3 G divb - G_j db_ij - 4/3 al ig dK
      // (the adjoints of the shift's second derivatives s are put below)
      for (int i = 0; i < 3; ++i) {
        const T r = rGam[i];
        T s[3][3];
        unpack_feature(f, F_D2 + 6 * d2_slot(BETA + i), s);
        for (int m = 0; m < 3; ++m)
          for (int n = 0; n < 3; ++n) ig_[m][n] += r * s[m][n];
        for (int j = 0; j < 3; ++j)
          for (int m = 0; m < 3; ++m) ig_[i][m] += T(1.0 / 3.0) * r * f(F_D2 + 6 * d2_slot(BETA + j) + sym(m, j));
        G_[i] += r * T(2.0 / 3.0) * divb;
        divb_ += r * T(2.0 / 3.0) * G[i];
        for (int j = 0; j < 3; ++j) {
          G_[j] -= r * db[i][j];
          db_[i][j] -= r * G[j];
          ig_[i][j] += T(-4.0 / 3.0) * al * r * dK[j];
          dK_[j] += T(-4.0 / 3.0) * al * r * ig[i][j];
          al_ += T(-4.0 / 3.0) * r * ig[i][j] * dK[j];
        }
      }
      // W: al W K / 3 - W divb / 3
      al_ += rW * W * K * T(1.0 / 3.0);
      W_ += rW * (al * K - divb) * T(1.0 / 3.0);
      K_ += rW * al * W * T(1.0 / 3.0);
      divb_ -= rW * W * T(1.0 / 3.0);
      // lapse: -al^2 K (harmonic) or -2 al K (1+log)
      if (C.gauge == 0) {
        al_ += ra * T(-2) * al * K;
        K_ += ra * (-(al * al));
      } else {
        al_ += ra * T(-2) * K;
        K_ += ra * T(-2) * al;
      }
      // metric: g_mi db_mj + g_mj db_mi - 2/3 g divb - 2 al A
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          const T r = rg[i][j];
          for (int m = 0; m < 3; ++m) {
            g_[m][i] += r * db[m][j];
            db_[m][j] += r * g[m][i];
            g_[m][j] += r * db[m][i];
            db_[m][i] += r * g[m][j];
          }
          g_[i][j] += T(-2.0 / 3.0) * divb * r;
          divb_ += T(-2.0 / 3.0) * g[i][j] * r;
          al_ += T(-2) * A[i][j] * r;
          A_[i][j] += T(-2) * al * r;
        }
      // shift: g G - eta b (when evolved)
      for (int i = 0; i < 3; ++i) {
        G_[i] += rb[i] * (T)C.g;
        b_[i] -= rb[i] * (T)C.eta;
      }
    } else {
      for (int c = 0; c < NCOMP; ++c) {
        for (int a = 0; a < 3; ++a) out.put(F_UP + 3 * c + a, T(0));
        out.put(F_KO + c, T(0));
      }
    }
    // the shift, its first derivatives and those of K are read by the RHS alone
    for (int i = 0; i < 3; ++i) db_[i][i] += divb_;
    for (int i = 0; i < 3; ++i) {
      out.put(F_VAL + BETA + i, b_[i]);
      out.put(F_D1 + 3 * CK + i, dK_[i]);
      for (int j = 0; j < 3; ++j) out.put(F_D1 + 3 * (BETA + i) + j, db_[i][j]);
    }
  }
  // shift second derivatives: s_j,mn = rG_j ig^mn + delta_nj ig^im rG_i / 3
  {
    T v[3];
    for (int m = 0; m < 3; ++m) v[m] = T(1.0 / 3.0) * (rGam[0] * ig[0][m] + rGam[1] * ig[1][m] + rGam[2] * ig[2][m]);
    for (int j = 0; j < 3; ++j) {
      T sb[3][3];
      for (int m = 0; m < 3; ++m)
        for (int n = 0; n < 3; ++n) sb[m][n] = rGam[j] * ig[m][n] + (n == j ? v[m] : T(0));
      for (int p = 0; p < 6; ++p) out.put(F_D2 + 6 * d2_slot(BETA + j) + p, fold_sym(sb, p));
    }
  }
  // first derivatives of A are features of Sommerfeld face points only
  for (int c = 0; c < 6; ++c)
    for (int k = 0; k < 3; ++k) out.put(F_D1 + 3 * (AXX + c) + k, T(0));

  // Hamiltonian constraint: ig:R + 2/3 K^2 - A:Aup
  if (hbar != T(0)) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        R_[i][j] += hbar * ig[i][j];
        ig_[i][j] += hbar * R[i][j];
        A_[i][j] -= hbar * Aup[i][j];
        Aup_[i][j] -= hbar * A[i][j];
      }
    K_ += hbar * T(4.0 / 3.0) * K;
  }
  out.put(F_VAL + CK, K_);
  out.put(F_VAL + ALPHA, al_);

  // Aup = ig A ig
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
          const T w = Aup_[i][j];
          ig_[i][k] += w * ig[j][l] * A[k][l];
          ig_[j][l] += w * ig[i][k] * A[k][l];
          A_[k][l] += w * ig[i][k] * ig[j][l];
        }
  for (int c = 0; c < 6; ++c) out.put(F_VAL + AXX + c, fold_sym(A_, c));

  // Hs = W^2 dda + W (dW da + da dW - g gWa) - W^2 C2^k da_k
  T gWa_ = T(0);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const T h = Hs_[i][j];
      W_ += h * HW[i][j];
      dW_[i] += W * h * da[j];
      da_[j] += W * h * dW[i];
      dW_[j] += W * h * da[i];
      da_[i] += W * h * dW[j];
      g_[i][j] -= W * gWa * h;
      gWa_ -= W * g[i][j] * h;
      for (int k = 0; k < 3; ++k) {
        C2_[k][i][j] -= W2 * h * da[k];
        da_[k] -= W2 * h * C2[k][i][j];
      }
    }
  for (int p = 0; p < 6; ++p) out.put(F_D2 + 6 * d2_slot(ALPHA) + p, W2 * fold_sym(Hs_, p));
  for (int m = 0; m < 3; ++m)
    for (int n = 0; n < 3; ++n) {
      ig_[m][n] += gWa_ * dW[m] * da[n];
      dW_[m] += gWa_ * ig[m][n] * da[n];
      da_[n] += gWa_ * ig[m][n] * dW[m];
    }
  for (int k = 0; k < 3; ++k) out.put(F_D1 + 3 * ALPHA + k, da_[k]);

  // R
  T lap_ = T(0), gradW2_ = T(0), conW_ = T(0);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const T r = R_[i][j];
      for (int m = 0; m < 3; ++m) {
        g_[m][i] += T(0.5) * W2 * r * dG[m][j];
        dG_[m][j] += T(0.5) * W2 * r * g[m][i];
        g_[m][j] += T(0.5) * W2 * r * dG[m][i];
        dG_[m][i] += T(0.5) * W2 * r * g[m][j];
        G_[m] += T(0.5) * W2 * r * (C1[i][j][m] + C1[j][i][m]);
        C1_[i][j][m] += T(0.5) * W2 * r * G[m];
        C1_[j][i][m] += T(0.5) * W2 * r * G[m];
        C2_[m][i][j] -= W * r * dW[m];
        dW_[m] -= W * r * C2[m][i][j];
      }
      for (int k = 0; k < 3; ++k)
        for (int n = 0; n < 3; ++n) {
          C1_[j][k][n] += W2 * r * Q[k][n][i];
          C1_[k][j][n] += W2 * r * Q[k][n][i];
          C1_[i][k][n] += W2 * r * Q[k][n][j];
        }
      W_ += r * RW[i][j];
      ddW_[i][j] += W * r;
      g_[i][j] += r * (W * lap - T(2) * gradW2 - W * conW);
      lap_ += W * r * g[i][j];
      gradW2_ += T(-2) * r * g[i][j];
      conW_ += -W * r * g[i][j];
    }
  for (int i = 0; i < 3; ++i) {
    out.put(F_VAL + GAM + i, G_[i]);
    for (int j = 0; j < 3; ++j) out.put(F_D1 + 3 * (GAM + i) + j, dG_[i][j]);
  }
  out.put(F_VAL + CW, W_);
  // the metric's second derivatives enter R_ij as -W^2/2 ig^mn d_m d_n g_ij
  for (int c = 0; c < 6; ++c) {
    const T w = T(-0.5) * W2 * fold_sym(R_, c);
    T M[3][3];
    unpack_feature(f, F_D2 + 6 * d2_slot(GXX + c), M);
    for (int m = 0; m < 3; ++m)
      for (int n = 0; n < 3; ++n) ig_[m][n] += w * M[m][n];
    for (int p = 0; p < 6; ++p) out.put(F_D2 + 6 * d2_slot(GXX + c) + p, w * fold_sym(ig, p));
  }
  // Q^k_n^i = ig^mn C2^k_mi, whose adjoint is formed one entry at a time
  for (int k = 0; k < 3; ++k)
    for (int n = 0; n < 3; ++n)
      for (int x = 0; x < 3; ++x) {
        T q = T(0);
        for (int j = 0; j < 3; ++j) q += R_[x][j] * (C1[j][k][n] + C1[k][j][n]) + R_[j][x] * C1[j][k][n];
        q *= W2;
        for (int m = 0; m < 3; ++m) {
          ig_[m][n] += q * C2[k][m][x];
          C2_[k][m][x] += q * ig[m][n];
        }
      }
  for (int m = 0; m < 3; ++m)
    for (int n = 0; n < 3; ++n) {
      ig_[m][n] += lap_ * ddW[m][n] + gradW2_ * dW[m] * dW[n];
      ddW_[m][n] += lap_ * ig[m][n];
      dW_[m] += gradW2_ * ig[m][n] * dW[n];
      dW_[n] += gradW2_ * ig[m][n] * dW[m];
    }
  for (int p = 0; p < 6; ++p) out.put(F_D2 + 6 * d2_slot(CW) + p, fold_sym(ddW_, p));
  // conW = sum_k (ig:C2^k) dW_k
  for (int k = 0; k < 3; ++k) {
    dW_[k] += conW_ * trC2[k];
    for (int m = 0; m < 3; ++m)
      for (int n = 0; n < 3; ++n) {
        ig_[m][n] += conW_ * C2[k][m][n] * dW[k];
        C2_[k][m][n] += conW_ * ig[m][n] * dW[k];
      }
  }
  for (int k = 0; k < 3; ++k) out.put(F_D1 + 3 * CW + k, dW_[k]);
  // C2^k_ij = ig^kl C1_lij
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int l = 0; l < 3; ++l) {
          ig_[k][l] += C2_[k][i][j] * C1[l][i][j];
          C1_[l][i][j] += C2_[k][i][j] * ig[k][l];
        }
  // C1_ijk = (d_k g_ij + d_j g_ik - d_i g_jk) / 2
  {
    T dg_[3][3][3] = {};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          const T c = T(0.5) * C1_[i][j][k];
          dg_[i][j][k] += c;
          dg_[i][k][j] += c;
          dg_[j][k][i] -= c;
        }
    for (int c = 0; c < 6; ++c) {
      const int i = SYM_I[c], j = SYM_J[c];
      for (int k = 0; k < 3; ++k) out.put(F_D1 + 3 * (GXX + c) + k, i == j ? dg_[i][i][k] : dg_[i][j][k] + dg_[j][i][k]);
    }
  }
  // ig = g^-1: g_ -= ig ig_ ig
  for (int a = 0; a < 3; ++a)
    for (int bb = 0; bb < 3; ++bb) {
      T acc = T(0);
      for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) acc += ig[a][c] * ig_[c][d] * ig[d][bb];
      g_[a][bb] -= acc;
    }
  for (int c = 0; c < 6; ++c) out.put(F_VAL + GXX + c, fold_sym(g_, c));
}

// Adjoint of the algebraic projection P (reference time_evolve.py:94-131:
// gamma <- gamma det^(-1/3), then A <- A - 1/3 gamma (gamma^ij A_ij) with
// the inverse of the rescaled metric) at one point: given the adjoints
// (gb, Ab) of P's outputs, the adjoints of its inputs (g, A), in place.
template <typename T>
BSSN_HD void project_vjp(const T (&gp)[6], const T (&Ap)[6], T (&gb)[6], T (&Ab)[6]) {
  T g[3][3], A[3][3], go_[3][3] = {}, Ao_[3][3] = {};
  unpack_sym(gp, 1, g);
  unpack_sym(Ap, 1, A);
  for (int c = 0; c < 6; ++c) {
    go_[SYM_I[c]][SYM_J[c]] = gb[c];
    Ao_[SYM_I[c]][SYM_J[c]] = Ab[c];
  }
  const T det = gp[0] * (gp[3] * gp[5] - gp[4] * gp[4]) - gp[1] * (gp[1] * gp[5] - gp[4] * gp[2]) +
                gp[2] * (gp[1] * gp[4] - gp[3] * gp[2]);
  const T sc = T(1) / cbrt(det);
  T g1p[6], ig1p[6];
  for (int c = 0; c < 6; ++c) g1p[c] = gp[c] * sc;
  inverse_metric(g1p, ig1p);
  T g1[3][3], ig1[3][3];
  unpack_sym(g1p, 1, g1);
  unpack_sym(ig1p, 1, ig1);
  T t = T(0);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t += ig1[i][j] * A[i][j];
  // A' = A - g1 t / 3
  T g1_[3][3], A_[3][3], t_ = T(0);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      A_[i][j] = Ao_[i][j];
      g1_[i][j] = go_[i][j] + T(-1.0 / 3.0) * t * Ao_[i][j];
      t_ += T(-1.0 / 3.0) * g1[i][j] * Ao_[i][j];
    }
  // t = ig1 : A
  T ig1_[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      ig1_[i][j] = t_ * A[i][j];
      A_[i][j] += t_ * ig1[i][j];
    }
  // ig1 = g1^-1
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      T acc = T(0);
      for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) acc += ig1[a][c] * ig1_[c][d] * ig1[d][b];
      g1_[a][b] -= acc;
    }
  // g1 = g det^(-1/3): d det / d g = det g^-1
  T ig[3][3], igp[6];
  inverse_metric(gp, igp);
  unpack_sym(igp, 1, ig);
  T s_ = T(0);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) s_ += g1_[i][j] * g[i][j];
  const T det_ = s_ * T(-1.0 / 3.0) * sc / det;
  T g_[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) g_[i][j] = sc * g1_[i][j] + det_ * det * ig[i][j];
  for (int c = 0; c < 6; ++c) {
    gb[c] = fold_sym(g_, c);
    Ab[c] = fold_sym(A_, c);
  }
}

}  // namespace bssn
