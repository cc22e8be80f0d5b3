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

// Inverse of the symmetric conformal metric from its upper triangle (the
// reference factors it by Cholesky, tensor_algebra.py:20-71; the cofactor
// form used here needs one reciprocal instead of nine divisions and agrees
// with it to roundoff).
template <typename T>
BSSN_HD __forceinline__ void inverse_metric(const T (&g)[6], T (&ig)[6]) {
  const T c00 = g[3] * g[5] - g[4] * g[4];
  const T c01 = g[2] * g[4] - g[1] * g[5];
  const T c02 = g[1] * g[4] - g[2] * g[3];
  const T idet = T(1) / (g[0] * c00 + g[1] * c01 + g[2] * c02);
  ig[0] = c00 * idet;
  ig[1] = c01 * idet;
  ig[2] = c02 * idet;
  ig[3] = (g[0] * g[5] - g[2] * g[2]) * idet;
  ig[4] = (g[1] * g[2] - g[0] * g[4]) * idet;
  ig[5] = (g[0] * g[3] - g[1] * g[1]) * idet;
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
// order of visit_comp() (shift, metric, lapse, W, K, connection, A).  Each
// component's RHS is handed to out.put(c, value) as soon as it is final
// (so its stage update leaves the registers early); out.done() follows the
// last one.  On an active Sommerfeld face point the provider reports face()
// and every component's RHS is replaced by the radial Sommerfeld RHS
// (reference evolution/boundaries.py:75-172) via out.put_face().
// u[] are the point values.
template <typename T, class P, class E, typename V>
BSSN_HD __forceinline__ void bssn_rhs_point(P &pv, const Ctx &C, const V (&u)[NCOMP], E &out) {
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
  const bool face = pv.face();
  T ig[6];
  inverse_metric(g, ig);

  T R[6], rG[3];
#pragma unroll
  for (int c = 0; c < 6; ++c) R[c] = T(0);
#pragma unroll
  for (int i = 0; i < 3; ++i) rG[i] = T(0);

  // ---- shift (visits 0..2): d_j beta^i; gamma^mn d_m d_n beta^i and
  //      1/3 gamma^im d_m d_n beta^n (connection terms 8, 9)
  T db[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    out.prefetch(BETA + i);
    pv.stage(i);
    if (face) out.put_face(BETA + i, V(pv.somm(BETA + i, u[BETA + i])));
#pragma unroll
    for (int j = 0; j < 3; ++j) db[i][j] = pv.d1(j);
    T s[6];
    second_derivatives(pv, s);
    rG[i] += contract(ig, s);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      T acc = ig[sym(r, 0)] * s[sym(0, i)] + ig[sym(r, 1)] * s[sym(1, i)] + ig[sym(r, 2)] * s[sym(2, i)];
      rG[r] += T(1.0 / 3.0) * acc;
    }
    const T tr = advection(pv, b) + pv.ko();
    const T rb = C.zero_shift ? T(0) * b[i] : (T)C.g * G[i] + (tr - (T)C.eta * b[i]);
    out.put(BETA + i, V(rb));
  }
  const T divb = db[0][0] + db[1][1] + db[2][2];
  // Lie terms of A and the shift terms of the connection, while d beta is live
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const int i = SYM_I[c], j = SYM_J[c];
    T lie = T(0);
#pragma unroll
    for (int m = 0; m < 3; ++m) lie += A[sym(m, i)] * db[m][j] + A[sym(m, j)] * db[m][i];
    pv.st.put(ST_ALIE + c, V(lie - T(2.0 / 3.0) * A[c] * divb));
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
    rG[i] += T(2.0 / 3.0) * G[i] * divb - (G[0] * db[i][0] + G[1] * db[i][1] + G[2] * db[i][2]);

  // ---- conformal metric (visits 3..8): d_k gamma_ij, -1/2 W^2 gamma^mn d_m d_n gamma_ij
  T dg[6][3];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    out.prefetch(GXX + c);
    pv.stage(3 + c);
    if (face) out.put_face(GXX + c, V(pv.somm(GXX + c, u[GXX + c])));
#pragma unroll
    for (int d = 0; d < 3; ++d) dg[c][d] = pv.d1(d);
    T s[6];
    second_derivatives(pv, s);
    R[c] += T(-0.5) * W2 * contract(ig, s);
    const int i = SYM_I[c], j = SYM_J[c];
    T lie = T(0);
#pragma unroll
    for (int m = 0; m < 3; ++m) lie += g[sym(m, i)] * db[m][j] + g[sym(m, j)] * db[m][i];
    out.put(GXX + c, V(advection(pv, b) + lie - T(2.0 / 3.0) * g[c] * divb - T(2) * al * A[c] + pv.ko()));
  }

  // A and the shift are idle until after the Christoffel algebra
#pragma unroll
  for (int c = 0; c < 6; ++c) pv.st.put(ST_A + c, V(A[c]));

  // ---- lapse (visit 9): d_i alpha and W^2 d_i d_j alpha
  T da[3], Hs[6];
  {
    out.prefetch(ALPHA);
    pv.stage(9);
    if (face) out.put_face(ALPHA, V(pv.somm(ALPHA, u[ALPHA])));
#pragma unroll
    for (int j = 0; j < 3; ++j) da[j] = pv.d1(j);
    T s[6];
    second_derivatives(pv, s);
#pragma unroll
    for (int c = 0; c < 6; ++c) Hs[c] = W2 * s[c];
    const T slicing = (C.gauge == 0) ? -(al * al) * K : T(-2) * al * K;
    out.put(ALPHA, V(slicing + advection(pv, b) + pv.ko()));
  }

  // ---- conformal factor (visit 10): W (d_i d_j W + gamma_ij gamma^mn d_m d_n W)
  T dW[3];
  {
    out.prefetch(CW);
    pv.stage(10);
    if (face) out.put_face(CW, V(pv.somm(CW, u[CW])));
#pragma unroll
    for (int j = 0; j < 3; ++j) dW[j] = pv.d1(j);
    T s[6];
    second_derivatives(pv, s);
    const T lap = contract(ig, s);
#pragma unroll
    for (int c = 0; c < 6; ++c) R[c] += W * (s[c] + g[c] * lap);
    out.put(CW, V(advection(pv, b) + T(1.0 / 3.0) * al * W * K - T(1.0 / 3.0) * W * divb + pv.ko()));
  }

  // ---- trace K (visit 11): -4/3 alpha gamma^ij d_j K goes straight to the connection
  {
    out.prefetch(CK);
    pv.stage(11);
    if (face) out.put_face(CK, V(pv.somm(CK, u[CK])));
    T dK[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) dK[j] = pv.d1(j);
#pragma unroll
    for (int i = 0; i < 3; ++i)
      rG[i] += T(-4.0 / 3.0) * al * (ig[sym(i, 0)] * dK[0] + ig[sym(i, 1)] * dK[1] + ig[sym(i, 2)] * dK[2]);
    pv.st.put(ST_RK, V(advection(pv, b) + pv.ko()));
  }

  // ---- connection (visits 12..14): W^2 gamma_m(i d_j) Gamma^m
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    out.prefetch(GAM + m);
    pv.stage(12 + m);
    if (face) out.put_face(GAM + m, V(pv.somm(GAM + m, u[GAM + m])));
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
#pragma unroll
  for (int i = 0; i < 3; ++i) pv.st.put(ST_B + i, V(b[i]));

  // ---- Christoffel algebra.  C1[i][jk] = Gamma_ijk.  For each upper index k
  // the row Gamma^k_ij = gamma^km Gamma_mij is formed, used and dropped.
  T C1[3][6];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const int j = SYM_I[c], k = SYM_J[c];
      C1[i][c] = T(0.5) * (dg[sym(i, j)][k] + dg[sym(i, k)][j] - dg[c][i]);
    }
  }
  {
    const T gradW2 = contract_vec(ig, dW, dW);
    const T gWa = contract_vec(ig, dW, da);
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const int i = SYM_I[c], j = SYM_J[c];
      // W (d_i W d_j alpha + d_j W d_i alpha - gamma_ij gamma^mn d_m W d_n alpha)
      Hs[c] += W * (dW[i] * da[j] + dW[j] * da[i] - g[c] * gWa);
      // term_3 = Gamma^m Gamma_(ij)m and -2 gamma_ij |dW|^2
      T t3 = T(0);
#pragma unroll
      for (int m = 0; m < 3; ++m) t3 += G[m] * (C1[i][sym(j, m)] + C1[j][sym(i, m)]);
      R[c] += W2 * T(0.5) * t3 - T(2) * g[c] * gradW2;
    }
  }
  T conW = T(0);  // gamma^mn Gamma^k_mn d_k W
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    T C2k[6];
#pragma unroll
    for (int c = 0; c < 6; ++c)
      C2k[c] = ig[sym(k, 0)] * C1[0][c] + ig[sym(k, 1)] * C1[1][c] + ig[sym(k, 2)] * C1[2][c];
    conW += contract(ig, C2k) * dW[k];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      R[c] -= W * C2k[c] * dW[k];
      Hs[c] -= W2 * C2k[c] * da[k];
    }
    // term_4 = gamma^mn (Gamma^k_mi Gamma_jkn + Gamma^k_mj Gamma_ikn), term_5 = gamma^mn Gamma^k_im Gamma_kjn,
    // one n at a time with Q_i = gamma^mn Gamma^k_mi
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      T Q[3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
        Q[i] = ig[sym(0, n)] * C2k[sym(0, i)] + ig[sym(1, n)] * C2k[sym(1, i)] + ig[sym(2, n)] * C2k[sym(2, i)];
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        const int i = SYM_I[c], j = SYM_J[c];
        R[c] += W2 * (Q[i] * (C1[j][sym(k, n)] + C1[k][sym(j, n)]) + Q[j] * C1[i][sym(k, n)]);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) R[c] -= W * g[c] * conW;

  const T Wf = W > T(1.0e-12) ? W : T(1.0e-12);

  // operands of the four updates that follow (fetched again here rather
  // than held across the Christoffel algebra since their visits)
#pragma unroll
  for (int i = 0; i < 3; ++i) out.prefetch(GAM + i);
  out.prefetch(CK);

  // ---- A^ij and the connection term 2 alpha Gamma^i_jk A^jk = 2 alpha gamma^im Gamma_mjk A^jk
#pragma unroll
  for (int c = 0; c < 6; ++c) A[c] = (T)pv.st.get(ST_A + c);
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
  {
    T s1[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) s1[m] = contract(C1[m], Aup);
    const T six_al_W = T(6) * al / Wf;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      T t2 = T(0), t3 = T(0), t4 = T(0);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        t2 += ig[sym(i, j)] * s1[j];
        t3 += Aup[sym(i, j)] * dW[j];
        t4 += Aup[sym(i, j)] * da[j];
      }
      T src = T(0);  // -16 pi alpha gamma^ij S_j (conformal_connection.py:165-178)
      if (out.has_matter())
        src = T(-16.0 * M_PI) * al *
              (ig[sym(i, 0)] * (T)out.matter(1) + ig[sym(i, 1)] * (T)out.matter(2) + ig[sym(i, 2)] * (T)out.matter(3));
      out.put(GAM + i, V(rG[i] + T(2) * al * t2 - six_al_W * t3 - T(2) * t4 + src));
    }
  }

  // ---- trace K (+ 4 pi alpha (rho + W^2 gamma^ij S_ij) with matter,
  //      reference extrinsic_curvature.py:78-87)
  {
    T rKv = (T)pv.st.get(ST_RK) - contract(ig, Hs) + al * contract(A, Aup) + al * K * K * T(1.0 / 3.0);
    if (out.has_matter()) {
      T S[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) S[c] = (T)out.matter(4 + c);
      rKv += T(4.0 * M_PI) * al * ((T)out.matter(0) + Wf * Wf * contract(ig, S));
    }
    out.put(CK, V(rKv));
  }

  // ---- everything of the A RHS except transport and dissipation
  T Arest[6];
  {
    T Xr[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) Xr[c] = al * R[c] - Hs[c];
    if (out.has_matter()) {  // - 8 pi alpha W^2 S_ij inside the trace-free part
#pragma unroll
      for (int c = 0; c < 6; ++c) Xr[c] -= T(8.0 * M_PI) * al * Wf * Wf * (T)out.matter(4 + c);
    }
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
      Arest[c] = al * K * A[c] - T(2) * al * AgA + (Xr[c] - T(1.0 / 3.0) * g[c] * trX) + (T)pv.st.get(ST_ALIE + c);
    }
  }

  // ---- traceless A (visits 15..20): transport and dissipation
#pragma unroll
  for (int i = 0; i < 3; ++i) b[i] = (T)pv.st.get(ST_B + i);
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    out.prefetch(AXX + c);
    pv.stage(15 + c);
    if (face) out.put_face(AXX + c, V(pv.somm(AXX + c, u[AXX + c])));
    out.put(AXX + c, V(Arest[c] + advection(pv, b) + pv.ko()) + out.extra_A(c));
  }
  out.done();
}

}  // namespace bssn
