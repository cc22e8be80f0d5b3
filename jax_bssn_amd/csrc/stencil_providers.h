// Stencil providers: where a kernel's finite-difference taps come from.
//
//  * TileProvider  -- interior points of a tile staged one component at a
//    time in LDS (coalesced loads, periodic wrap applied by the loader).
//    Stencil taps are compile-time LDS offsets from one base register.
//    Only interior stencils (no one-sided closures) are evaluated here.
//  * PointProvider -- any point, taps read straight from global memory
//    through per-axis stencil rows that carry the reference's one-sided
//    Sommerfeld closures, the MAD blend and the upwind replacement.
//    Used for the closure shells next to Sommerfeld faces and by the
//    constraint kernel.
//
// Both expose the same interface to the per-point RHS algebra:
//   stage(c), val(), d1(a), d2(a), dmix(a, b) [inner a, outer b],
//   upw(a, coeff), ko()
// Stencil coefficients follow reference evolution/derivatives.py:42-599 and
// evolution/spatial_derivatives.py:14-92.
#pragma once

#include "bssn_common.h"

namespace bssn {

// Visit order of the components in the per-point RHS (each visit may
// restage an LDS tile; the tile provider prefetches visit v + 1 during v).
__host__ __device__ constexpr int visit_comp(int v) {
  return v < 3 ? BETA + v : (v == 3 ? ALPHA : (v == 4 ? CW : (v == 5 ? CK : (v < 9 ? GAM + (v - 6) : (v < 15 ? GXX + (v - 9) : AXX + (v - 15))))));
}
constexpr int VISIT_FIRST = BETA;

__host__ __device__ inline int wrap_index(int i, int n) {
  int r = i % n;
  return r < 0 ? r + n : r;
}

// ---------------------------------------------------------------- LDS tile

template <typename T, int TX, int TY, int TZ, bool MAD>
struct TileProvider {
  static constexpr int H = MAD ? 4 : 3;
  static constexpr int NT = TX * TY * TZ;
  static constexpr int LX = TX + 2 * H, LY = TY + 2 * H, LZ = TZ + 2 * H;
  static constexpr int SX = LY * LZ, SY = LZ;
  static constexpr int NBOX = LX * LY * LZ;
  static constexpr int NSLOT = (NBOX + NT - 1) / NT;
  static constexpr int LDS_ELEMS = 2 * NBOX;  // double-buffered tile

  T *lds;        // two tile buffers
  const T *cur;  // buffer of the staged component
  const double *src;
  int64_t ncell;
  int goff[NSLOT];
  T pref[NSLOT];  // register prefetch of the next component's tile
  int c;
  int tid;
  T inv_dx, inv_dx2, ko_coef;
  T q[3];
  int red[3];

  __device__ static constexpr int stride(int a) { return a == 0 ? SX : (a == 1 ? SY : 1); }

  __device__ void fetch(int comp) {
    const double *f = src + comp * ncell;
#pragma unroll
    for (int k = 0; k < NSLOT; ++k) pref[k] = (T)f[goff[k]];
  }

  __device__ void setup(const Ctx &C, const double *u, T *smem, int x0, int y0, int z0,
                        int tx, int ty, int tz) {
    lds = smem;
    cur = smem;
    src = u;
    ncell = C.ncell;
    tid = (tx * TY + ty) * TZ + tz;
    c = ((tx + H) * LY + (ty + H)) * LZ + (tz + H);
    inv_dx = (T)C.inv_dx;
    inv_dx2 = (T)C.inv_dx2;
    ko_coef = (T)C.ko;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      q[a] = (T)C.ax[a].q;
      red[a] = C.ax[a].reduced;
    }
#pragma unroll
    for (int k = 0; k < NSLOT; ++k) {
      int e = tid + k * NT;
      if (e >= NBOX) e = NBOX - 1;
      int lx = e / SX, ly = (e / LZ) % LY, lz = e % LZ;
      int gx = wrap_index(x0 + lx - H, C.ax[0].n);
      int gy = wrap_index(y0 + ly - H, C.ax[1].n);
      int gz = wrap_index(z0 + lz - H, C.ax[2].n);
      goff[k] = gx * C.ax[0].stride + gy * C.ax[1].stride + gz;
    }
    fetch(VISIT_FIRST);
  }

  // Stage visit v: store the prefetched tile, one barrier, prefetch v + 1.
  // Buffer v & 1 was last read in visit v - 2, which every thread finished
  // before arriving at the barrier of visit v - 1.
  __device__ void stage(int v) {
    T *buf = lds + (v & 1) * NBOX;
#pragma unroll
    for (int k = 0; k < NSLOT; ++k) {
      int e = tid + k * NT;
      if (k < NSLOT - 1 || e < NBOX) buf[e] = pref[k];
    }
    __syncthreads();
    if (v + 1 < NCOMP) fetch(visit_comp(v + 1));
    cur = buf;
  }

  __device__ T at(int a, int o) const { return cur[c + o * stride(a)]; }
  __device__ T at2(int a, int oa, int b, int ob) const {
    return cur[c + oa * stride(a) + ob * stride(b)];
  }
  __device__ T val() const { return cur[c]; }

  // centred first derivative: D4, or q*D4 + (1-q)*D6 with MAD
  __device__ T d1(int a) const {
    if (red[a]) return T(0);
    T f1 = at(a, 1), b1 = at(a, -1), f2 = at(a, 2), b2 = at(a, -2);
    T d4 = (T(2.0 / 3.0) * (f1 - b1) - T(1.0 / 12.0) * (f2 - b2)) * inv_dx;
    if constexpr (!MAD) {
      return d4;
    } else {
      T f3 = at(a, 3), b3 = at(a, -3);
      T d6 = (T(3.0 / 4.0) * (f1 - b1) - T(3.0 / 20.0) * (f2 - b2) + T(1.0 / 60.0) * (f3 - b3)) * inv_dx;
      return q[a] * d4 + (T(1) - q[a]) * d6;
    }
  }

  __device__ T d2(int a) const {
    if (red[a]) return T(0);
    T f0 = val(), f1 = at(a, 1), b1 = at(a, -1), f2 = at(a, 2), b2 = at(a, -2);
    T d4 = (T(16.0) * (f1 + b1) - (f2 + b2) - T(30.0) * f0) * (inv_dx2 * T(1.0 / 12.0));
    if constexpr (!MAD) {
      return d4;
    } else {
      T f3 = at(a, 3), b3 = at(a, -3);
      T d6 = (T(1.0 / 90.0) * (f3 + b3) - T(3.0 / 20.0) * (f2 + b2) + T(3.0 / 2.0) * (f1 + b1)
              - T(49.0 / 18.0) * f0) * inv_dx2;
      return q[a] * d4 + (T(1) - q[a]) * d6;
    }
  }

  // mixed second derivative: inner centred D1 along a, outer along b
  __device__ T dmix(int a, int b) const {
    if (red[a] || red[b]) return T(0);
    auto inner = [&](int ob) -> T {
      T f1 = at2(a, 1, b, ob), b1 = at2(a, -1, b, ob);
      T f2 = at2(a, 2, b, ob), b2 = at2(a, -2, b, ob);
      T d4 = T(2.0 / 3.0) * (f1 - b1) - T(1.0 / 12.0) * (f2 - b2);
      if constexpr (!MAD) {
        return d4;
      } else {
        T f3 = at2(a, 3, b, ob), b3 = at2(a, -3, b, ob);
        T d6 = T(3.0 / 4.0) * (f1 - b1) - T(3.0 / 20.0) * (f2 - b2) + T(1.0 / 60.0) * (f3 - b3);
        return q[a] * d4 + (T(1) - q[a]) * d6;
      }
    };
    T p1 = inner(1), m1 = inner(-1), p2 = inner(2), m2 = inner(-2);
    T d4 = T(2.0 / 3.0) * (p1 - m1) - T(1.0 / 12.0) * (p2 - m2);
    if constexpr (!MAD) {
      return d4 * inv_dx2;
    } else {
      T p3 = inner(3), m3 = inner(-3);
      T d6 = T(3.0 / 4.0) * (p1 - m1) - T(3.0 / 20.0) * (p2 - m2) + T(1.0 / 60.0) * (p3 - m3);
      return (q[b] * d4 + (T(1) - q[b]) * d6) * inv_dx2;
    }
  }

  // upwind derivative selected by the sign of the RHS coefficient c
  __device__ T upw(int a, T coeff) const {
    if (red[a]) return T(0);
    T f0 = val(), f1 = at(a, 1), b1 = at(a, -1), f2 = at(a, 2), b2 = at(a, -2);
    T f3 = at(a, 3), b3 = at(a, -3);
    T fw = (T(-3.0) * b1 - T(10.0) * f0 + T(18.0) * f1 - T(6.0) * f2 + f3) * (inv_dx * T(1.0 / 12.0));
    T bw = (-b3 + T(6.0) * b2 - T(18.0) * b1 + T(10.0) * f0 + T(3.0) * f1) * (inv_dx * T(1.0 / 12.0));
    T ce = (T(2.0 / 3.0) * (f1 - b1) - T(1.0 / 12.0) * (f2 - b2)) * inv_dx;
    if constexpr (MAD) {
      T f4 = at(a, 4), b4 = at(a, -4);
      T fw6 = (T(2.0) * b2 - T(24.0) * b1 - T(35.0) * f0 + T(80.0) * f1 - T(30.0) * f2
               + T(8.0) * f3 - f4) * (inv_dx * T(1.0 / 60.0));
      T bw6 = (b4 - T(8.0) * b3 + T(30.0) * b2 - T(80.0) * b1 + T(35.0) * f0 + T(24.0) * f1
               - T(2.0) * f2) * (inv_dx * T(1.0 / 60.0));
      T ce6 = (T(3.0 / 4.0) * (f1 - b1) - T(3.0 / 20.0) * (f2 - b2) + T(1.0 / 60.0) * (f3 - b3)) * inv_dx;
      T qa = q[a], pa = T(1) - q[a];
      fw = qa * fw + pa * fw6;
      bw = qa * bw + pa * bw6;
      ce = qa * ce + pa * ce6;
    }
    return coeff > T(0) ? fw : (coeff < T(0) ? bw : ce);
  }

  // Kreiss-Oliger: nu/64 dx^5 sum_a D6_a
  __device__ T ko() const {
    T f0 = val();
    T s = T(0);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (red[a]) continue;
      T t = (at(a, 3) + at(a, -3)) - T(6.0) * (at(a, 2) + at(a, -2)) + T(15.0) * (at(a, 1) + at(a, -1))
            - T(20.0) * f0;
      s += t;
    }
    return ko_coef * s;
  }
};

// ------------------------------------------------------- global-memory rows

// A sparse one-dimensional stencil row: value = sum_k c[k] * f[p + off[k]].
struct Row {
  int n;
  int off[8];
  double c[8];
  BSSN_HD void push(int o, double v) {
    for (int k = 0; k < n; ++k)
      if (off[k] == o) { c[k] += v; return; }
    off[n] = o;
    c[n] = v;
    ++n;
  }
};

BSSN_HD inline void row_add(Row &r, const int *off, const double *c, int m, double scale) {
  for (int k = 0; k < m; ++k) r.push(off[k], c[k] * scale);
}

// Centred D1 row at position p of axis A (closures, MAD).
BSSN_HD inline Row row_d1(const Axis &A, int p, double inv_dx, bool force_d4) {
  Row r;
  r.n = 0;
  const int s = A.start, e = A.end;
  if (A.rbc && p == e - 1) {
    const int o[5] = {-4, -3, -2, -1, 0};
    const double c[5] = {1.0 / 4.0, -4.0 / 3.0, 3.0, -4.0, 25.0 / 12.0};
    row_add(r, o, c, 5, inv_dx);
  } else if (A.rbc && p == e - 2) {
    const int o[5] = {-3, -2, -1, 0, 1};
    const double c[5] = {-1.0 / 12.0, 1.0 / 2.0, -3.0 / 2.0, 5.0 / 6.0, 1.0 / 4.0};
    row_add(r, o, c, 5, inv_dx);
  } else if (A.lbc && p == s + 1) {
    const int o[5] = {-1, 0, 1, 2, 3};
    const double c[5] = {-1.0 / 4.0, -5.0 / 6.0, 3.0 / 2.0, -1.0 / 2.0, 1.0 / 12.0};
    row_add(r, o, c, 5, inv_dx);
  } else if (A.lbc && p == s) {
    const int o[5] = {0, 1, 2, 3, 4};
    const double c[5] = {-25.0 / 12.0, 4.0, -3.0, 4.0 / 3.0, -1.0 / 4.0};
    row_add(r, o, c, 5, inv_dx);
  } else {
    const int o4[4] = {-2, -1, 1, 2};
    const double c4[4] = {1.0 / 12.0, -2.0 / 3.0, 2.0 / 3.0, -1.0 / 12.0};
    if (force_d4 || A.q == 1.0) {
      row_add(r, o4, c4, 4, inv_dx);
    } else {
      const int o6[6] = {-3, -2, -1, 1, 2, 3};
      const double c6[6] = {-1.0 / 60.0, 3.0 / 20.0, -3.0 / 4.0, 3.0 / 4.0, -3.0 / 20.0, 1.0 / 60.0};
      row_add(r, o4, c4, 4, A.q * inv_dx);
      row_add(r, o6, c6, 6, (1.0 - A.q) * inv_dx);
    }
  }
  return r;
}

BSSN_HD inline Row row_d2(const Axis &A, int p, double inv_dx2) {
  Row r;
  r.n = 0;
  const int s = A.start, e = A.end;
  if (A.rbc && p == e - 1) {
    const int o[6] = {-5, -4, -3, -2, -1, 0};
    const double c[6] = {-5.0 / 6.0, 61.0 / 12.0, -13.0, 107.0 / 6.0, -77.0 / 6.0, 15.0 / 4.0};
    row_add(r, o, c, 6, inv_dx2);
  } else if (A.rbc && p == e - 2) {
    const int o[6] = {-4, -3, -2, -1, 0, 1};
    const double c[6] = {1.0 / 12.0, -1.0 / 2.0, 7.0 / 6.0, -1.0 / 3.0, -5.0 / 4.0, 5.0 / 6.0};
    row_add(r, o, c, 6, inv_dx2);
  } else if (A.lbc && p == s + 1) {
    const int o[6] = {-1, 0, 1, 2, 3, 4};
    const double c[6] = {5.0 / 6.0, -5.0 / 4.0, -1.0 / 3.0, 7.0 / 6.0, -1.0 / 2.0, 1.0 / 12.0};
    row_add(r, o, c, 6, inv_dx2);
  } else if (A.lbc && p == s) {
    const int o[6] = {0, 1, 2, 3, 4, 5};
    const double c[6] = {15.0 / 4.0, -77.0 / 6.0, 107.0 / 6.0, -13.0, 61.0 / 12.0, -5.0 / 6.0};
    row_add(r, o, c, 6, inv_dx2);
  } else {
    const int o4[5] = {-2, -1, 0, 1, 2};
    const double c4[5] = {-1.0 / 12.0, 16.0 / 12.0, -30.0 / 12.0, 16.0 / 12.0, -1.0 / 12.0};
    if (A.q == 1.0) {
      row_add(r, o4, c4, 5, inv_dx2);
    } else {
      const int o6[7] = {-3, -2, -1, 0, 1, 2, 3};
      const double c6[7] = {1.0 / 90.0, -3.0 / 20.0, 3.0 / 2.0, -49.0 / 18.0, 3.0 / 2.0, -3.0 / 20.0, 1.0 / 90.0};
      row_add(r, o4, c4, 5, A.q * inv_dx2);
      row_add(r, o6, c6, 7, (1.0 - A.q) * inv_dx2);
    }
  }
  return r;
}

BSSN_HD inline Row row_upwind(const Axis &A, int p, double inv_dx, double coeff) {
  const int s = A.start, e = A.end;
  const int w = A.n < 3 ? A.n : 3;
  if ((A.lbc && p >= s && p < s + w) || (A.rbc && p >= e - w && p < e))
    return row_d1(A, p, inv_dx, true);
  Row r;
  r.n = 0;
  const bool physical = A.lbc || A.rbc;
  const double q = physical ? 1.0 : A.q;
  const double s12 = inv_dx / 12.0, s60 = inv_dx / 60.0;
  if (coeff > 0.0) {
    const int o[5] = {-1, 0, 1, 2, 3};
    const double c[5] = {-3.0, -10.0, 18.0, -6.0, 1.0};
    row_add(r, o, c, 5, q * s12);
    if (q != 1.0) {
      const int o6[7] = {-2, -1, 0, 1, 2, 3, 4};
      const double c6[7] = {2.0, -24.0, -35.0, 80.0, -30.0, 8.0, -1.0};
      row_add(r, o6, c6, 7, (1.0 - q) * s60);
    }
  } else if (coeff < 0.0) {
    const int o[5] = {-3, -2, -1, 0, 1};
    const double c[5] = {-1.0, 6.0, -18.0, 10.0, 3.0};
    row_add(r, o, c, 5, q * s12);
    if (q != 1.0) {
      const int o6[7] = {-4, -3, -2, -1, 0, 1, 2};
      const double c6[7] = {1.0, -8.0, 30.0, -80.0, 35.0, 24.0, -2.0};
      row_add(r, o6, c6, 7, (1.0 - q) * s60);
    }
  } else {
    Axis B = A;
    B.q = q;
    return row_d1(B, p, inv_dx, false);
  }
  return r;
}

// Sixth difference row scaled by dx^6 (i.e. the bare integer stencil).
BSSN_HD inline Row row_d6(const Axis &A, int p) {
  const double L[3][8] = {{4.0, -27.0, 78.0, -125.0, 120.0, -69.0, 22.0, -3.0},
                          {3.0, -20.0, 57.0, -90.0, 85.0, -48.0, 15.0, -2.0},
                          {2.0, -13.0, 36.0, -55.0, 50.0, -27.0, 8.0, -1.0}};
  Row r;
  r.n = 0;
  const int s = A.start, e = A.end;
  for (int k = 0; k < 3; ++k) {
    if (A.rbc && p == e - 1 - k) {
      for (int t = 0; t < 8; ++t) r.push(k - t, L[k][t]);  // taps e-1-t of the point e-1-k
      return r;
    }
  }
  for (int k = 0; k < 3; ++k) {
    if (A.lbc && p == s + k) {
      for (int t = 0; t < 8; ++t) r.push(t - k, L[k][t]);
      return r;
    }
  }
  const int o[7] = {-3, -2, -1, 0, 1, 2, 3};
  const double c[7] = {1.0, -6.0, 15.0, -20.0, 15.0, -6.0, 1.0};
  row_add(r, o, c, 7, 1.0);
  return r;
}

template <typename T>
struct PointProvider {
  const Ctx *C;
  const double *src;
  const double *f;
  int p[3];
  int64_t idx;

  BSSN_HD void setup(const Ctx &ctx, const double *u, int x, int y, int z) {
    C = &ctx;
    src = u;
    f = u;
    p[0] = x; p[1] = y; p[2] = z;
    idx = (int64_t)x * ctx.ax[0].stride + (int64_t)y * ctx.ax[1].stride + z;
  }
  BSSN_HD void stage(int v) { f = src + visit_comp(v) * C->ncell; }
  BSSN_HD void stage_comp(int comp) { f = src + comp * C->ncell; }

  BSSN_HD T ld(int a, int oa) const {
    int q[3] = {p[0], p[1], p[2]};
    q[a] = wrap_index(q[a] + oa, C->ax[a].n);
    return (T)f[(int64_t)q[0] * C->ax[0].stride + (int64_t)q[1] * C->ax[1].stride + q[2]];
  }
  BSSN_HD T ld2(int a, int oa, int b, int ob) const {
    int q[3] = {p[0], p[1], p[2]};
    q[a] = wrap_index(q[a] + oa, C->ax[a].n);
    q[b] = wrap_index(q[b] + ob, C->ax[b].n);
    return (T)f[(int64_t)q[0] * C->ax[0].stride + (int64_t)q[1] * C->ax[1].stride + q[2]];
  }
  BSSN_HD T apply(int a, const Row &r) const {
    T s = T(0);
    for (int k = 0; k < r.n; ++k) s += (T)r.c[k] * ld(a, r.off[k]);
    return s;
  }
  BSSN_HD T val() const { return (T)f[idx]; }
  BSSN_HD T d1(int a) const {
    if (C->ax[a].reduced) return T(0);
    return apply(a, row_d1(C->ax[a], p[a], C->inv_dx, false));
  }
  BSSN_HD T d2(int a) const {
    if (C->ax[a].reduced) return T(0);
    return apply(a, row_d2(C->ax[a], p[a], C->inv_dx2));
  }
  BSSN_HD T dmix(int a, int b) const {
    if (C->ax[a].reduced || C->ax[b].reduced) return T(0);
    Row ra = row_d1(C->ax[a], p[a], C->inv_dx, false);
    Row rb = row_d1(C->ax[b], p[b], C->inv_dx, false);
    T s = T(0);
    for (int j = 0; j < rb.n; ++j) {
      T inner = T(0);
      for (int i = 0; i < ra.n; ++i) inner += (T)ra.c[i] * ld2(a, ra.off[i], b, rb.off[j]);
      s += (T)rb.c[j] * inner;
    }
    return s;
  }
  BSSN_HD T upw(int a, T coeff) const {
    if (C->ax[a].reduced) return T(0);
    return apply(a, row_upwind(C->ax[a], p[a], C->inv_dx, (double)coeff));
  }
  BSSN_HD T ko() const {
    T s = T(0);
    for (int a = 0; a < 3; ++a) {
      if (C->ax[a].reduced) continue;
      s += apply(a, row_d6(C->ax[a], p[a]));
    }
    return (T)C->ko * s;
  }
};

}  // namespace bssn
