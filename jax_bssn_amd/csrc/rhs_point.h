// The BSSN right-hand side at one grid point, written once for every
// stencil provider and arithmetic type.
//
// Components are visited one at a time (each visit may restage an LDS
// tile); every second-derivative contribution is folded into an
// accumulator as soon as it is formed, so only the 36 first derivatives the
// nonlinear terms need stay live until the final algebra.
//
// Equations (reference file:line):
//   conformal metric   bssn/spatial_metric.py:17-63
//   conformal factor   bssn/spatial_metric.py:66-106
//   traceless A        bssn/extrinsic_curvature.py:119-224 (vacuum)
//   trace K            bssn/extrinsic_curvature.py:31-92   (vacuum)
//   connection         bssn/conformal_connection.py:19-162
//   lapse, shift       bssn/shift_and_lapse.py:65-132, evolution/time_evolve.py:30-38
//   W^2 R_ij           bssn/geometry.py:21-119
//   W^2 D_i D_j alpha  bssn/geometry.py:122-186
//   Cholesky inverse   bssn/tensor_algebra.py:20-71
#pragma once

#include "bssn_common.h"

namespace bssn {

// Reference tensor_algebra.py:20-71 (explicit Cholesky, upper triangle).
template <typename T>
BSSN_HD __forceinline__ void inverse_metric(const T (&g)[6], T (&ig)[6]) {
  T L00 = sqrt(g[0]);
  T L10 = g[1] / L00;
  T L20 = g[2] / L00;
  T L11 = sqrt(g[3] - L10 * L10);
  T L21 = (g[4] - L20 * L10) / L11;
  T L22 = sqrt(g[5] - L20 * L20 - L21 * L21);
  T M00 = T(1) / L00;
  T M10 = -L10 / (L00 * L11);
  T M11 = T(1) / L11;
  T M20 = (L10 * L21 - L20 * L11) / (L00 * L11 * L22);
  T M21 = -L21 / (L11 * L22);
  T M22 = T(1) / L22;
  ig[0] = M00 * M00 + M10 * M10 + M20 * M20;
  ig[1] = M10 * M11 + M20 * M21;
  ig[2] = M20 * M22;
  ig[3] = M11 * M11 + M21 * M21;
  ig[4] = M21 * M22;
  ig[5] = M22 * M22;
}

// Full contraction X^{ij} Y_ij of two symmetric tensors.
template <typename T>
BSSN_HD __forceinline__ T contract(const T (&X)[6], const T (&Y)[6]) {
  return X[0] * Y[0] + X[3] * Y[3] + X[5] * Y[5] + T(2) * (X[1] * Y[1] + X[2] * Y[2] + X[4] * Y[4]);
}

// v_m X^{mn} w_n for a symmetric X.
template <typename T>
BSSN_HD __forceinline__ T contract_vec(const T (&X)[6], const T (&v)[3], const T (&w)[3]) {
  T acc = T(0);
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int n = 0; n < 3; ++n) acc += X[sym(m, n)] * v[m] * w[n];
  return acc;
}

constexpr int SYM_I[6] = {0, 0, 0, 1, 1, 2};
constexpr int SYM_J[6] = {0, 1, 2, 1, 2, 2};

// Second derivatives of the staged component: s[sym(m, n)].
template <typename T, class P>
BSSN_HD __forceinline__ void second_derivatives(const P &pv, T (&s)[6]) {
  s[0] = pv.d2(0);
  s[3] = pv.d2(1);
  s[5] = pv.d2(2);
  s[1] = pv.dmix(0, 1);
  s[2] = pv.dmix(0, 2);
  s[4] = pv.dmix(1, 2);
}

template <typename T, class P>
BSSN_HD __forceinline__ T advection(const P &pv, const T (&b)[3]) {
  return b[0] * pv.upw(0, b[0]) + b[1] * pv.upw(1, b[1]) + b[2] * pv.upw(2, b[2]);
}

// RHS of every component at this point, visiting the components in the
// order of visit_comp(); out.put(c, value) is called as soon
// as component c is final (so its stage update leaves the registers early)
// and out.done() after the last one.  u[] are the point values.
template <typename T, class P, class E>
BSSN_HD __forceinline__ void bssn_rhs_point(P &pv, const Ctx &C, const double (&u)[NCOMP], E &out) {
  T g[6], A[6], G[3], b[3];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    g[c] = (T)u[GXX + c];
    A[c] = (T)u[AXX + c];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    G[i] = (T)u[GAM + i];
    b[i] = (T)u[BETA + i];
  }
  const T W = (T)u[CW], K = (T)u[CK], al = (T)u[ALPHA];
  const T W2 = W * W;
  T ig[6];
  inverse_metric(g, ig);

  T R[6], Hs[6], rG[3];
#pragma unroll
  for (int c = 0; c < 6; ++c) { R[c] = T(0); Hs[c] = T(0); }
#pragma unroll
  for (int i = 0; i < 3; ++i) rG[i] = T(0);

  // ---- shift: d_j beta^i, gamma^mn d_m d_n beta^i, 1/3 gamma^im d_m div(beta)
  T db[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    pv.stage(i);
#pragma unroll
    for (int j = 0; j < 3; ++j) db[i][j] = pv.d1(j);
    T s[6];
    second_derivatives(pv, s);
    rG[i] += contract(ig, s);
    // d_m d_n beta^n with n = i contributes to the ninth term of every row
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      T acc = ig[sym(r, 0)] * s[sym(0, i)] + ig[sym(r, 1)] * s[sym(1, i)] + ig[sym(r, 2)] * s[sym(2, i)];
      rG[r] += T(1.0 / 3.0) * acc;
    }
    const T tr = advection(pv, b) + pv.ko();
    const T rb = C.zero_shift ? T(0) * b[i] : (T)C.g * G[i] + (tr - (T)C.eta * b[i]);
    out.put(BETA + i, (double)rb);
  }
  const T divb = db[0][0] + db[1][1] + db[2][2];

  // ---- lapse: first derivatives and W^2 d_i d_j alpha
  T da[3];
  {
    pv.stage(3);
#pragma unroll
    for (int j = 0; j < 3; ++j) da[j] = pv.d1(j);
    T s[6];
    second_derivatives(pv, s);
#pragma unroll
    for (int c = 0; c < 6; ++c) Hs[c] += W2 * s[c];
    const T slicing = (C.gauge == 0) ? -(al * al) * K : T(-2) * al * K;
    out.put(ALPHA, (double)(slicing + advection(pv, b) + pv.ko()));
  }

  // ---- conformal factor: W (d_i d_j W + gamma_ij gamma^mn d_m d_n W)
  T dW[3];
  {
    pv.stage(4);
#pragma unroll
    for (int j = 0; j < 3; ++j) dW[j] = pv.d1(j);
    T s[6];
    second_derivatives(pv, s);
    const T lap = contract(ig, s);
#pragma unroll
    for (int c = 0; c < 6; ++c) R[c] += W * (s[c] + g[c] * lap);
    out.put(CW, (double)(advection(pv, b) + T(1.0 / 3.0) * al * W * K - T(1.0 / 3.0) * W * divb + pv.ko()));
  }

  // ---- trace K: first derivatives
  T dK[3], rK;
  {
    pv.stage(5);
#pragma unroll
    for (int j = 0; j < 3; ++j) dK[j] = pv.d1(j);
    rK = advection(pv, b) + pv.ko();
  }

  // ---- connection: W^2 (gamma_m(i d_j) Gamma^m)
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    pv.stage(6 + m);
    T dG[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) dG[j] = pv.d1(j);
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const int i = SYM_I[c], j = SYM_J[c];
      R[c] += W2 * T(0.5) * (g[sym(m, i)] * dG[j] + g[sym(m, j)] * dG[i]);
    }
    rG[m] += advection(pv, b) + pv.ko();
  }

  // ---- conformal metric: d_k gamma_ij and -1/2 W^2 gamma^mn d_m d_n gamma_ij
  T dg[6][3];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    pv.stage(9 + c);
#pragma unroll
    for (int d = 0; d < 3; ++d) dg[c][d] = pv.d1(d);
    T s[6];
    second_derivatives(pv, s);
    R[c] += T(-0.5) * W2 * contract(ig, s);
    const int i = SYM_I[c], j = SYM_J[c];
    T lie = T(0);
#pragma unroll
    for (int m = 0; m < 3; ++m) lie += g[sym(m, i)] * db[m][j] + g[sym(m, j)] * db[m][i];
    out.put(GXX + c, (double)(advection(pv, b) + lie - T(2.0 / 3.0) * g[c] * divb - T(2) * al * A[c] + pv.ko()));
  }

  // ---- traceless A: transport and dissipation
  T rA[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    pv.stage(15 + c);
    rA[c] = advection(pv, b) + pv.ko();
  }

  // ---- Christoffel symbols: C1[i][jk] = Gamma_ijk, C2[i][jk] = Gamma^i_jk
  T C1[3][6], C2[3][6];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const int j = SYM_I[c], k = SYM_J[c];
      C1[i][c] = T(0.5) * (dg[sym(i, j)][k] + dg[sym(i, k)][j] - dg[c][i]);
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int c = 0; c < 6; ++c)
      C2[i][c] = ig[sym(i, 0)] * C1[0][c] + ig[sym(i, 1)] * C1[1][c] + ig[sym(i, 2)] * C1[2][c];

  // ---- W^2 R_ij: term_3 + term_4 + term_5 and the Christoffel parts of D_i D_j W.
  // With Q^{kn}_i = gamma^mn Gamma^k_mi:  term_4_ij = X_ij + X_ji,
  // X_ij = Q^{kn}_i Gamma_jkn;  term_5_ij = Q^{kn}_i Gamma_kjn.
  {
    T X[3][3], Y[6];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      T Q[3][3];
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int n = 0; n < 3; ++n)
          Q[k][n] = ig[sym(0, n)] * C2[k][sym(0, i)] + ig[sym(1, n)] * C2[k][sym(1, i)] +
                    ig[sym(2, n)] * C2[k][sym(2, i)];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        T x = T(0), y = T(0);
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int n = 0; n < 3; ++n) {
            x += Q[k][n] * C1[j][sym(k, n)];
            if (j >= i) y += Q[k][n] * C1[k][sym(j, n)];
          }
        X[i][j] = x;
        if (j >= i) Y[sym(i, j)] = y;
      }
    }
    const T gradW2 = contract_vec(ig, dW, dW);
    T conW = T(0);  // Gamma^k d_k W with Gamma^k = gamma^mn Gamma^k_mn
#pragma unroll
    for (int k = 0; k < 3; ++k) conW += contract(ig, C2[k]) * dW[k];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const int i = SYM_I[c], j = SYM_J[c];
      T t3 = T(0);
#pragma unroll
      for (int m = 0; m < 3; ++m) t3 += G[m] * (C1[i][sym(j, m)] + C1[j][sym(i, m)]);
      const T chW = C2[0][c] * dW[0] + C2[1][c] * dW[1] + C2[2][c] * dW[2];
      R[c] += W2 * (T(0.5) * t3 + (X[i][j] + X[j][i]) + Y[c]) - W * (chW + g[c] * conW) - T(2) * g[c] * gradW2;
    }
  }

  // ---- W^2 D_i D_j alpha: Christoffel and gradient parts
  {
    const T gWa = contract_vec(ig, dW, da);
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const int i = SYM_I[c], j = SYM_J[c];
      const T cha = C2[0][c] * da[0] + C2[1][c] * da[1] + C2[2][c] * da[2];
      Hs[c] += -W2 * cha + W * (dW[i] * da[j] + dW[j] * da[i] - g[c] * gWa);
    }
  }

  // ---- A^ij
  T Aup[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const int i = SYM_I[c], j = SYM_J[c];
    T acc = T(0);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const T row = ig[sym(j, 0)] * A[sym(k, 0)] + ig[sym(j, 1)] * A[sym(k, 1)] + ig[sym(j, 2)] * A[sym(k, 2)];
      acc += ig[sym(i, k)] * row;
    }
    Aup[c] = acc;
  }

  // ---- trace K
  out.put(CK, (double)(rK - contract(ig, Hs) + al * contract(A, Aup) + al * K * K * T(1.0 / 3.0)));

  // ---- connection
  {
    const T Wf = W > T(1.0e-12) ? W : T(1.0e-12);
    const T six_al_W = T(6) * al / Wf;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      T t1 = T(0), t3 = T(0), t4 = T(0), t7 = T(0);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        t1 += ig[sym(i, j)] * dK[j];
        t3 += Aup[sym(i, j)] * dW[j];
        t4 += Aup[sym(i, j)] * da[j];
        t7 += G[j] * db[i][j];
      }
      const T t2 = contract(C2[i], Aup);
      out.put(GAM + i, (double)(rG[i] + T(-4.0 / 3.0) * al * t1 + T(2) * al * t2 - six_al_W * t3 - T(2) * t4 +
                                T(2.0 / 3.0) * G[i] * divb - t7));
    }
  }

  // ---- traceless A
  {
    T Xr[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) Xr[c] = al * R[c] - Hs[c];
    const T trX = contract(ig, Xr);
    T gA[3][3];  // gamma^kl A_lj
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        gA[k][j] = ig[sym(k, 0)] * A[sym(0, j)] + ig[sym(k, 1)] * A[sym(1, j)] + ig[sym(k, 2)] * A[sym(2, j)];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const int i = SYM_I[c], j = SYM_J[c];
      const T AgA = A[sym(i, 0)] * gA[0][j] + A[sym(i, 1)] * gA[1][j] + A[sym(i, 2)] * gA[2][j];
      T lie = T(0);
#pragma unroll
      for (int m = 0; m < 3; ++m) lie += A[sym(m, i)] * db[m][j] + A[sym(m, j)] * db[m][i];
      out.put(AXX + c, (double)(rA[c] + al * K * A[c] - T(2) * al * AgA + (Xr[c] - T(1.0 / 3.0) * g[c] * trX) +
                                lie - T(2.0 / 3.0) * A[c] * divb));
    }
  }
  out.done();
}

}  // namespace bssn
