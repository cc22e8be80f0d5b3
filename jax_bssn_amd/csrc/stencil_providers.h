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

// Components whose LDS boxes are held in registers ahead of their visit:
// in the FP64 closure-free tile kernels (one wave per SIMD, registers to
// spare) and in the others (two waves per SIMD).
#ifndef BSSN_PREFETCH_DEPTH_SEL
#define BSSN_PREFETCH_DEPTH_SEL 2
#endif
#ifndef BSSN_PREFETCH_DEPTH
#define BSSN_PREFETCH_DEPTH 1
#endif
// LDS z pitch of the boxes of the FP64 closure-free tile kernels (0: the
// box's own z extent).  A wave reads an 8 x 8 y-z plane of its box: with
// rows 24 doubles apart, four consecutive rows of eight cover the 64 banks
// once, where the natural pitch of 14 makes the rows collide in pairs.
#ifndef BSSN_LDS_PITCH_SEL
#define BSSN_LDS_PITCH_SEL 24
#endif
// The same for the other tile kernels (0: natural pitch).
#ifndef BSSN_LDS_PITCH
#define BSSN_LDS_PITCH 0
#endif

namespace bssn {

// Visit order of the components in the per-point RHS (each visit may
// restage an LDS tile; the tile provider prefetches visits v + 1 .. v + PD
// during v).
__host__ __device__ constexpr int visit_comp(int v) {
  // shift 0..2, metric 3..8, lapse 9, W 10, K 11, connection 12..14, A 15..20
  return v < 3 ? BETA + v
               : (v < 9 ? GXX + (v - 3)
                        : (v == 9 ? ALPHA
                                  : (v == 10 ? CW : (v == 11 ? CK : (v < 15 ? GAM + (v - 12) : AXX + (v - 15))))));
}

__host__ __device__ inline int wrap_index(int i, int n) {
  int r = i % n;
  return r < 0 ? r + n : r;
}

// ------------------------------------------------------------ stencil rows

// A one-dimensional stencil row as eight coefficients over the consecutive
// offsets o0 .. o0+7 from the point (unused taps are zero).  Every row of
// the reference operators fits: one-sided closures (5, 6 or 8 taps), the
// centred D4 / D6 / MAD rows and the biased upwind rows.  Coefficients are
// only ever indexed by constants, so rows live in registers.
template <typename T>
struct Row8 {
  int o0;
  int n;  // taps o0 .. o0 + n - 1 are the row's reach; c[k] = 0 for k >= n
  T c[8];
  // offset of tap k, clamped into the row's reach: a reader that walks all
  // eight taps never leaves the samples the row is defined on
  BSSN_HD __forceinline__ int off(int k) const { return o0 + (k < n ? k : n - 1); }
};

// a ? x : y on the bit patterns: exact, and never compiled into a branch
// (the upwind side differs from lane to lane).
template <typename T>
__device__ __forceinline__ T pick(bool a, T x, T y) {
  if constexpr (sizeof(T) == 8) {
    const long long m = -(long long)a;
    return __longlong_as_double((__double_as_longlong(x) & m) | (__double_as_longlong(y) & ~m));
  } else {
    const int m = -(int)a;
    return __int_as_float((__float_as_int(x) & m) | (__float_as_int(y) & ~m));
  }
}

template <typename T>
BSSN_HD __forceinline__ void row_clear(Row8<T> &r, int o0, int n = 8) {
  r.o0 = o0;
  r.n = n;
#pragma unroll
  for (int k = 0; k < 8; ++k) r.c[k] = T(0);
}

template <typename T>
BSSN_HD __forceinline__ void row_set5(Row8<T> &r, int o0, double a0, double a1, double a2, double a3, double a4,
                                      double s) {
  row_clear(r, o0, 5);
  r.c[0] = (T)(a0 * s);
  r.c[1] = (T)(a1 * s);
  r.c[2] = (T)(a2 * s);
  r.c[3] = (T)(a3 * s);
  r.c[4] = (T)(a4 * s);
}

// Centred first derivative at position p of axis A (reference
// evolution/derivatives.py:42-162): one-sided closures at the two outer
// points of an active face, else D4 (or the MAD blend q D4 + (1-q) D6).
template <typename T>
BSSN_HD __forceinline__ Row8<T> row_d1(const Axis &A, int p, double inv_dx, bool force_d4) {
  Row8<T> r;
  const int s = A.start, e = A.end;
  if (A.rbc && p == e - 1) {
    row_set5(r, -4, 1.0 / 4.0, -4.0 / 3.0, 3.0, -4.0, 25.0 / 12.0, inv_dx);
  } else if (A.rbc && p == e - 2) {
    row_set5(r, -3, -1.0 / 12.0, 1.0 / 2.0, -3.0 / 2.0, 5.0 / 6.0, 1.0 / 4.0, inv_dx);
  } else if (A.lbc && p == s + 1) {
    row_set5(r, -1, -1.0 / 4.0, -5.0 / 6.0, 3.0 / 2.0, -1.0 / 2.0, 1.0 / 12.0, inv_dx);
  } else if (A.lbc && p == s) {
    row_set5(r, 0, -25.0 / 12.0, 4.0, -3.0, 4.0 / 3.0, -1.0 / 4.0, inv_dx);
  } else if (force_d4 || A.q == 1.0) {
    row_set5(r, -2, 1.0 / 12.0, -2.0 / 3.0, 0.0, 2.0 / 3.0, -1.0 / 12.0, inv_dx);
  } else {
    const double q = A.q, w = 1.0 - q;
    row_clear(r, -3, 7);
    r.c[0] = (T)(w * (-1.0 / 60.0) * inv_dx);
    r.c[1] = (T)((q * (1.0 / 12.0) + w * (3.0 / 20.0)) * inv_dx);
    r.c[2] = (T)((q * (-2.0 / 3.0) + w * (-3.0 / 4.0)) * inv_dx);
    r.c[4] = (T)((q * (2.0 / 3.0) + w * (3.0 / 4.0)) * inv_dx);
    r.c[5] = (T)((q * (-1.0 / 12.0) + w * (-3.0 / 20.0)) * inv_dx);
    r.c[6] = (T)(w * (1.0 / 60.0) * inv_dx);
  }
  return r;
}

// Pure second derivative (reference derivatives.py:341-461).
template <typename T>
BSSN_HD __forceinline__ Row8<T> row_d2(const Axis &A, int p, double inv_dx2) {
  Row8<T> r;
  const int s = A.start, e = A.end;
  const double h = inv_dx2;
  if (A.rbc && p == e - 1) {
    row_clear(r, -5, 6);
    const double v[6] = {-5.0 / 6.0, 61.0 / 12.0, -13.0, 107.0 / 6.0, -77.0 / 6.0, 15.0 / 4.0};
#pragma unroll
    for (int k = 0; k < 6; ++k) r.c[k] = (T)(v[k] * h);
  } else if (A.rbc && p == e - 2) {
    row_clear(r, -4, 6);
    const double v[6] = {1.0 / 12.0, -1.0 / 2.0, 7.0 / 6.0, -1.0 / 3.0, -5.0 / 4.0, 5.0 / 6.0};
#pragma unroll
    for (int k = 0; k < 6; ++k) r.c[k] = (T)(v[k] * h);
  } else if (A.lbc && p == s + 1) {
    row_clear(r, -1, 6);
    const double v[6] = {5.0 / 6.0, -5.0 / 4.0, -1.0 / 3.0, 7.0 / 6.0, -1.0 / 2.0, 1.0 / 12.0};
#pragma unroll
    for (int k = 0; k < 6; ++k) r.c[k] = (T)(v[k] * h);
  } else if (A.lbc && p == s) {
    row_clear(r, 0, 6);
    const double v[6] = {15.0 / 4.0, -77.0 / 6.0, 107.0 / 6.0, -13.0, 61.0 / 12.0, -5.0 / 6.0};
#pragma unroll
    for (int k = 0; k < 6; ++k) r.c[k] = (T)(v[k] * h);
  } else if (A.q == 1.0) {
    row_set5(r, -2, -1.0 / 12.0, 16.0 / 12.0, -30.0 / 12.0, 16.0 / 12.0, -1.0 / 12.0, h);
  } else {
    const double q = A.q, w = 1.0 - q;
    row_clear(r, -3, 7);
    const double d4[7] = {0.0, -1.0 / 12.0, 16.0 / 12.0, -30.0 / 12.0, 16.0 / 12.0, -1.0 / 12.0, 0.0};
    const double d6[7] = {1.0 / 90.0, -3.0 / 20.0, 3.0 / 2.0, -49.0 / 18.0, 3.0 / 2.0, -3.0 / 20.0, 1.0 / 90.0};
#pragma unroll
    for (int k = 0; k < 7; ++k) r.c[k] = (T)((q * d4[k] + w * d6[k]) * h);
  }
  return r;
}

// Upwind-biased first derivative (reference derivatives.py:166-338): the
// sign of the RHS coefficient selects forward (> 0), backward (< 0) or the
// centred stencil (== 0); within three points of an active face the
// boundary-aware centred D4 replaces it.
template <typename T>
BSSN_HD __forceinline__ Row8<T> row_upwind(const Axis &A, int p, double inv_dx, double coeff) {
  const int s = A.start, e = A.end;
  const int wz = A.n < 3 ? A.n : 3;
  if ((A.lbc && p >= s && p < s + wz) || (A.rbc && p >= e - wz && p < e)) return row_d1<T>(A, p, inv_dx, true);
  const bool physical = A.lbc || A.rbc;
  const double q = physical ? 1.0 : A.q, w = 1.0 - q;
  Row8<T> r;
  const double s12 = inv_dx * (1.0 / 12.0), s60 = inv_dx * (1.0 / 60.0);
  if (coeff > 0.0) {
    // D4 forward on -1..3, D6 forward on -2..4
    row_clear(r, -2, 7);
    const double d4[7] = {0.0, -3.0, -10.0, 18.0, -6.0, 1.0, 0.0};
    const double d6[7] = {2.0, -24.0, -35.0, 80.0, -30.0, 8.0, -1.0};
#pragma unroll
    for (int k = 0; k < 7; ++k) r.c[k] = (T)(q * d4[k] * s12 + w * d6[k] * s60);
  } else if (coeff < 0.0) {
    // D4 backward on -3..1, D6 backward on -4..2
    row_clear(r, -4, 7);
    const double d4[7] = {0.0, -1.0, 6.0, -18.0, 10.0, 3.0, 0.0};
    const double d6[7] = {1.0, -8.0, 30.0, -80.0, 35.0, 24.0, -2.0};
#pragma unroll
    for (int k = 0; k < 7; ++k) r.c[k] = (T)(q * d4[k] * s12 + w * d6[k] * s60);
  } else {
    Axis B = A;
    B.q = q;
    return row_d1<T>(B, p, inv_dx, false);
  }
  return r;
}

// Sixth difference scaled by dx^6 (reference derivatives.py:463-599):
// 8-point one-sided rows at the three outer points of an active face.
template <typename T>
BSSN_HD __forceinline__ Row8<T> row_d6(const Axis &A, int p) {
  const double L[3][8] = {{4.0, -27.0, 78.0, -125.0, 120.0, -69.0, 22.0, -3.0},
                          {3.0, -20.0, 57.0, -90.0, 85.0, -48.0, 15.0, -2.0},
                          {2.0, -13.0, 36.0, -55.0, 50.0, -27.0, 8.0, -1.0}};
  Row8<T> r;
  const int s = A.start, e = A.end;
  const int kr = A.rbc ? e - 1 - p : -1;  // right rows: taps e-1-t -> offsets k - t
  const int kl = A.lbc ? p - s : -1;      // left rows: taps s+t -> offsets t - k
  if (kr >= 0 && kr < 3) {
    row_clear(r, kr - 7);
#pragma unroll
    for (int j = 0; j < 8; ++j) r.c[j] = (T)(kr == 0 ? L[0][7 - j] : (kr == 1 ? L[1][7 - j] : L[2][7 - j]));
  } else if (kl >= 0 && kl < 3) {
    row_clear(r, -kl);
#pragma unroll
    for (int j = 0; j < 8; ++j) r.c[j] = (T)(kl == 0 ? L[0][j] : (kl == 1 ? L[1][j] : L[2][j]));
  } else {
    row_clear(r, -3, 7);
    const double v[7] = {1.0, -6.0, 15.0, -20.0, 15.0, -6.0, 1.0};
#pragma unroll
    for (int k = 0; k < 7; ++k) r.c[k] = (T)v[k];
  }
  return r;
}

// Workgroup barrier that orders LDS only: every wave's LDS writes are done
// and visible after it, while its global loads stay in flight (a full
// __syncthreads() also drains them, so a prefetch issued before it could
// not overlap the next component's work).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ------------------------------------------------------------------ stash

// Per-thread parking space (slots in bssn_common.h) for values that stay
// live across the whole point but are idle during the register-heavy
// Christoffel algebra.

// In LDS, slot-major so that consecutive lanes hit consecutive banks.
template <int NT>
struct LdsStash {
  BSSN_LDS double *base;  // this thread's slot 0 (stride NT between slots)
  __device__ __forceinline__ void put(int slot, double v) { base[slot * NT] = v; }
  __device__ double get(int slot) const { return base[slot * NT]; }
};

// In registers (host build and the shell kernel).
struct RegStash {
  double v[STASH_SLOTS];
  BSSN_HD __forceinline__ void put(int slot, double x) { v[slot] = x; }
  BSSN_HD __forceinline__ double get(int slot) const { return v[slot]; }
};

// ---------------------------------------------------------------- LDS tile

// Box halo along an axis of tile extent t: 3 for the centred/upwind/KO
// stencils, 4 with MAD (sixth-order upwind).  With one-sided closures the
// taps of a face tile reach 7 points inward (the D6 closure rows), so the
// box must satisfy t + halo >= 8.
__host__ __device__ constexpr int tile_halo(int t, bool mad, bool cl) {
  return (cl && t + (mad ? 4 : 3) < 8) ? 8 - t : (mad ? 4 : 3);
}
static_assert(tile_halo(2, false, true) == 6 && tile_halo(4, false, true) == 4 && tile_halo(8, false, true) == 3,
              "closure halo");

// CL: the tile may hold closure points; CLA: the axes (bit a) along which
// they may lie (a face tile of one slab has closures along one axis only,
// which keeps the other axes' stencils compact and their halos narrow).
// SEL (closure-free tiles only): first derivatives and upwind derivatives
// without branches -- a reduced axis selects a zero scale and the upwind
// side is a bit select -- so that the stencils of one component form one
// basic block in which the scheduler issues the LDS taps early and shares
// them between derivatives.  Used by the FP64 stage kernels of interior
// tiles (about 15 % faster there); every other kernel keeps the branched
// form.
template <typename T, int TX, int TY, int TZ, bool MAD, bool CL, int CLA = 7, bool SEL = false>
struct TileProvider {
  static_assert(!(SEL && CL), "branch-free selects are for closure-free tiles");
  static constexpr bool CLX = CL && (CLA & 1), CLY = CL && (CLA & 2), CLZ = CL && (CLA & 4);
  static constexpr int HX = tile_halo(TX, MAD, CLX), HY = tile_halo(TY, MAD, CLY), HZ = tile_halo(TZ, MAD, CLZ);
  static constexpr int NT = TX * TY * TZ;
  static constexpr int LX = TX + 2 * HX, LY = TY + 2 * HY, LZ = TZ + 2 * HZ;
  // z pitch of the LDS box (rows of LZ samples; SEL kernels may pad it so
  // that the rows a wave reads fall in distinct banks)
  static constexpr int LZP = SEL ? (BSSN_LDS_PITCH_SEL > LZ ? BSSN_LDS_PITCH_SEL : LZ)
                                 : (BSSN_LDS_PITCH > LZ ? BSSN_LDS_PITCH : LZ);
  static constexpr int SX = LY * LZP, SY = LZP;
  static constexpr int NBOX = LX * LY * LZ;     // samples staged per component
  static constexpr int NBOXP = LX * LY * LZP;   // LDS elements of one box buffer
  static constexpr int NSLOT = (NBOX + NT - 1) / NT;
  // Double-buffered when the workgroups a CU holds (box buffers + stash)
  // still fit its 160 KB of LDS: one workgroup of the SEL kernels (one wave
  // per SIMD), two of the others; wider boxes (closure or MAD halos) keep
  // one buffer.
  static constexpr int NBUF =
      (2 * NBOXP * (int)sizeof(T) + STASH_SLOTS * NT * (int)sizeof(double) <= (SEL ? 160 : 80) * 1024) ? 2 : 1;
  static constexpr int LDS_ELEMS = NBUF * NBOXP;
  // LDS element of staged sample e (row-major over LX x LY rows of LZ)
  __device__ __forceinline__ static int lds_index(int e) { return LZP == LZ ? e : e + (e / LZ) * (LZP - LZ); }

  BSSN_LDS T *lds;        // two tile buffers
  const BSSN_LDS T *cur;  // buffer of the staged component
  LdsStash<TX * TY * TZ> st;
  const BSSN_GLOBAL double *src;
  int64_t ncell;
  int goff[NSLOT];
  int boff;       // FP32 offset sample of each component (the clamped tile origin)
  // register prefetch of the next PD components' tiles (ring indexed by the
  // visit number, a compile-time constant in the unrolled per-point code)
  static constexpr int PD = SEL ? BSSN_PREFETCH_DEPTH_SEL : BSSN_PREFETCH_DEPTH;
  T pref[PD][NSLOT];
  int c;
  int tid;
  T inv_dx, inv_dx2, ko_coef;
  T q[3];
  int red[3];
  // closure support (CL): block-uniform axis flags and the point's position
  Axis ax[3];
  int pos[3];
  int clz[3];     // tile touches a closure zone along the axis
  bool isface;
  T xr[3];        // coordinates for the Sommerfeld face RHS
  double uval;    // point value of the staged component (face RHS)

  __device__ __forceinline__ static constexpr int stride(int a) { return a == 0 ? SX : (a == 1 ? SY : 1); }

  // In FP32 the box holds each value minus the component's value at the
  // tile origin (one block-uniform FP64 load): every stencil's weights sum
  // to zero, so derivatives are unchanged, and FP32 keeps the digits of the
  // variation instead of those of a large constant background.
  __device__ __forceinline__ void fetch(int comp, int slot) {
    const BSSN_GLOBAL double *f = src + comp * ncell;
    const double base = sizeof(T) == sizeof(double) ? 0.0 : f[boff];
#pragma unroll
    for (int k = 0; k < NSLOT; ++k) pref[slot][k] = (T)(f[goff[k]] - base);
  }

  __device__ __forceinline__ void setup(const Ctx &C, const double *u, T *smem, double *stash_mem, int x0, int y0, int z0,
                        int tx, int ty, int tz, bool prefetch = true) {
    lds = (BSSN_LDS T *)smem;
    cur = lds;
    src = (const BSSN_GLOBAL double *)u;
    ncell = C.ncell;
    tid = threadIdx.x;  // the cooperative loads and the stash follow the lanes, not the point
    st.base = (BSSN_LDS double *)stash_mem + tid;
    c = ((tx + HX) * LY + (ty + HY)) * LZP + (tz + HZ);
    inv_dx = (T)C.inv_dx;
    inv_dx2 = (T)C.inv_dx2;
    ko_coef = (T)C.ko;
    const int o[3] = {x0, y0, z0};
    const int t[3] = {TX, TY, TZ};
    const int l[3] = {tx, ty, tz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      q[a] = (T)C.ax[a].q;
      red[a] = C.ax[a].reduced;
      ax[a] = C.ax[a];
      pos[a] = o[a] + l[a];
      clz[a] = CL && ((CLA >> a) & 1) &&
               ((C.ax[a].lbc && o[a] < C.ax[a].start + 3 && o[a] + t[a] > C.ax[a].start) ||
                (C.ax[a].rbc && o[a] + t[a] > C.ax[a].end - 3 && o[a] < C.ax[a].end));
    }
    isface = false;
    if constexpr (CL) {
      bool f = false, owned = true;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const Axis &A = C.ax[a];
        f |= (A.somm_lo && pos[a] == A.start) || (A.somm_hi && pos[a] == A.end - 1);
        owned &= pos[a] >= A.start && pos[a] < A.end;
        xr[a] = (T)(C.xmin[a] + C.dx * (double)pos[a]);
      }
      isface = f && owned;
    }
#pragma unroll
    for (int k = 0; k < NSLOT; ++k) {
      int e = tid + k * NT;
      if (e >= NBOX) e = NBOX - 1;
      int lx = e / (LY * LZ), ly = (e / LZ) % LY, lz = e % LZ;
      int gx = wrap_index(x0 + lx - HX, C.ax[0].n);
      int gy = wrap_index(y0 + ly - HY, C.ax[1].n);
      int gz = wrap_index(z0 + lz - HZ, C.ax[2].n);
      goff[k] = gx * C.ax[0].stride + gy * C.ax[1].stride + gz;
    }
    {
      // the FP32 offset sample: the tile origin clamped into the owned box,
      // so that it is within the stencils' reach of every owned point the
      // tile updates (a tile may start in guard samples beyond that reach)
      int b[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int o = a == 0 ? x0 : (a == 1 ? y0 : z0);
        b[a] = o < C.ax[a].start ? C.ax[a].start : (o >= C.ax[a].end ? C.ax[a].end - 1 : o);
      }
      boff = b[0] * C.ax[0].stride + b[1] * C.ax[1].stride + b[2];
    }
    if (prefetch) {
#pragma unroll
      for (int d = 0; d < PD; ++d) fetch(visit_comp(d), d);
    }
  }

  // Stage one component from an arbitrary array (component base f) into
  // the first box buffer, synchronously: loads, a barrier (every lane is
  // done with the previous box), stores, a barrier.  For passes that visit
  // components of several arrays in their own order (the kappa term).
  __device__ __forceinline__ void stage_from(const double *f) {
    const BSSN_GLOBAL double *g = (const BSSN_GLOBAL double *)f;
#pragma unroll
    for (int k = 0; k < NSLOT; ++k) pref[0][k] = (T)g[goff[k]];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NSLOT; ++k) {
      int e = tid + k * NT;
      if (k < NSLOT - 1 || e < NBOX) lds[lds_index(e)] = pref[0][k];
    }
    __syncthreads();
    cur = lds;
  }

  // Stage visit v: store the prefetched tile, one barrier, prefetch v + PD
  // into the ring slot just emptied.  Double-buffered, buffer v & 1 was last
  // read in visit v - 2, which every thread finished before arriving at the
  // barrier of visit v - 1; a single buffer needs a second barrier before it
  // is overwritten.
  __device__ __forceinline__ void stage(int v) {
    BSSN_LDS T *buf = lds + (NBUF == 2 ? (v & 1) * NBOXP : 0);
    const int slot = v % PD;
    if (NBUF == 1 && v > 0) lds_barrier();  // every thread is done with visit v - 1
#pragma unroll
    for (int k = 0; k < NSLOT; ++k) {
      int e = tid + k * NT;
      if (k < NSLOT - 1 || e < NBOX) buf[lds_index(e)] = pref[slot][k];
    }
    lds_barrier();
    if (v + PD < NCOMP) fetch(visit_comp(v + PD), slot);
    cur = buf;
  }

  __device__ __forceinline__ T at(int a, int o) const { return cur[c + o * stride(a)]; }
  __device__ __forceinline__ T at2(int a, int oa, int b, int ob) const {
    return cur[c + oa * stride(a) + ob * stride(b)];
  }
  __device__ __forceinline__ T val() const { return cur[c]; }

  // this lane sits in the closure zone of width w along axis a
  __device__ __forceinline__ bool zone(int a, int w) const {
    if constexpr (!CL) {
      return false;
    } else {
      if (!((CLA >> a) & 1) || !clz[a]) return false;
      const int lo = pos[a] - ax[a].start, hi = ax[a].end - 1 - pos[a];
      return (ax[a].lbc && lo >= 0 && lo < w) || (ax[a].rbc && hi >= 0 && hi < w);
    }
  }
  __device__ __forceinline__ T apply(int a, const Row8<T> &r) const {
    T s = T(0);
#pragma unroll
    for (int k = 0; k < 8; ++k) s += r.c[k] * at(a, r.off(k));
    return s;
  }

  __device__ __forceinline__ bool face() const { return CL && isface; }

  // centred first derivative: D4, or q*D4 + (1-q)*D6 with MAD
  __device__ __forceinline__ T d1(int a) const {
    if constexpr (!SEL) {
      if (red[a]) return T(0);
      if (zone(a, 2)) return apply(a, row_d1<T>(ax[a], pos[a], (double)inv_dx, false));
    }
    const T sc = (SEL && red[a]) ? T(0) : inv_dx;
    T f1 = at(a, 1), b1 = at(a, -1), f2 = at(a, 2), b2 = at(a, -2);
    T d4 = (T(2.0 / 3.0) * (f1 - b1) - T(1.0 / 12.0) * (f2 - b2)) * sc;
    if constexpr (!MAD) {
      return d4;
    } else {
      T f3 = at(a, 3), b3 = at(a, -3);
      T d6 = (T(3.0 / 4.0) * (f1 - b1) - T(3.0 / 20.0) * (f2 - b2) + T(1.0 / 60.0) * (f3 - b3)) * sc;
      return q[a] * d4 + (T(1) - q[a]) * d6;
    }
  }

  __device__ __forceinline__ T d2(int a) const {
    if (red[a]) return T(0);
    if (zone(a, 2)) return apply(a, row_d2<T>(ax[a], pos[a], (double)inv_dx2));
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

  // mixed second derivative: inner centred D1 along a, outer along b.  A
  // closure row is used only along an axis whose zone this lane is in; the
  // other axis keeps its compact centred (or MAD) formula.
  __device__ __forceinline__ T dmix(int a, int b) const {
    if (red[a] || red[b]) return T(0);
    // centred inner D1 along a (without 1/dx) at outer offset ob along b
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
    // centred outer D1 along b (without 1/dx) of values g(ob)
    auto outer = [&](auto &&g) -> T {
      T p1 = g(1), m1 = g(-1), p2 = g(2), m2 = g(-2);
      T d4 = T(2.0 / 3.0) * (p1 - m1) - T(1.0 / 12.0) * (p2 - m2);
      if constexpr (!MAD) {
        return d4;
      } else {
        T p3 = g(3), m3 = g(-3);
        T d6 = T(3.0 / 4.0) * (p1 - m1) - T(3.0 / 20.0) * (p2 - m2) + T(1.0 / 60.0) * (p3 - m3);
        return q[b] * d4 + (T(1) - q[b]) * d6;
      }
    };
    const bool za = zone(a, 2), zb = zone(b, 2);
    if (za && zb) {
      const Row8<T> ra = row_d1<T>(ax[a], pos[a], (double)inv_dx, false);
      const Row8<T> rb = row_d1<T>(ax[b], pos[b], (double)inv_dx, false);
      T s = T(0);
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        T in = T(0);
#pragma unroll
        for (int i = 0; i < 8; ++i) in += ra.c[i] * at2(a, ra.off(i), b, rb.off(j));
        s += rb.c[j] * in;
      }
      return s;
    }
    if (za) {  // closure row along a (1/dx in its weights), compact outer along b
      const Row8<T> ra = row_d1<T>(ax[a], pos[a], (double)inv_dx, false);
      auto row_a = [&](int ob) -> T {
        T in = T(0);
#pragma unroll
        for (int i = 0; i < 8; ++i) in += ra.c[i] * at2(a, ra.off(i), b, ob);
        return in;
      };
      return outer(row_a) * inv_dx;
    }
    if (zb) {  // compact inner along a, closure row along b
      const Row8<T> rb = row_d1<T>(ax[b], pos[b], (double)inv_dx, false);
      T s = T(0);
#pragma unroll
      for (int j = 0; j < 8; ++j) s += rb.c[j] * inner(rb.off(j));
      return s * inv_dx;
    }
    return outer(inner) * inv_dx2;
  }

  // upwind derivative selected by the sign of the RHS coefficient c
  __device__ __forceinline__ T upw(int a, T coeff) const {
    if constexpr (!SEL) {
      if (red[a]) return T(0);
      if (zone(a, 3)) return apply(a, row_upwind<T>(ax[a], pos[a], (double)inv_dx, (double)coeff));
    }
    const T sc = (SEL && red[a]) ? T(0) : inv_dx;
    T f0 = val(), f1 = at(a, 1), b1 = at(a, -1), f2 = at(a, 2), b2 = at(a, -2);
    T f3 = at(a, 3), b3 = at(a, -3);
    T fw = (T(-3.0) * b1 - T(10.0) * f0 + T(18.0) * f1 - T(6.0) * f2 + f3) * (sc * T(1.0 / 12.0));
    T bw = (-b3 + T(6.0) * b2 - T(18.0) * b1 + T(10.0) * f0 + T(3.0) * f1) * (sc * T(1.0 / 12.0));
    T ce = (T(2.0 / 3.0) * (f1 - b1) - T(1.0 / 12.0) * (f2 - b2)) * sc;
    if constexpr (MAD) {
      T f4 = at(a, 4), b4 = at(a, -4);
      T fw6 = (T(2.0) * b2 - T(24.0) * b1 - T(35.0) * f0 + T(80.0) * f1 - T(30.0) * f2
               + T(8.0) * f3 - f4) * (sc * T(1.0 / 60.0));
      T bw6 = (b4 - T(8.0) * b3 + T(30.0) * b2 - T(80.0) * b1 + T(35.0) * f0 + T(24.0) * f1
               - T(2.0) * f2) * (sc * T(1.0 / 60.0));
      T ce6 = (T(3.0 / 4.0) * (f1 - b1) - T(3.0 / 20.0) * (f2 - b2) + T(1.0 / 60.0) * (f3 - b3)) * sc;
      T qa = q[a], pa = T(1) - q[a];
      fw = qa * fw + pa * fw6;
      bw = qa * bw + pa * bw6;
      ce = qa * ce + pa * ce6;
    }
    if constexpr (SEL) {
      return pick(coeff > T(0), fw, pick(coeff < T(0), bw, ce));
    } else {
      return coeff > T(0) ? fw : (coeff < T(0) ? bw : ce);
    }
  }

  // Kreiss-Oliger: nu/64 dx^5 sum_a D6_a
  __device__ __forceinline__ T ko() const {
    T f0 = val();
    T s = T(0);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (red[a]) continue;
      if (zone(a, 3)) {
        s += apply(a, row_d6<T>(ax[a], pos[a]));
        continue;
      }
      T t = (at(a, 3) + at(a, -3)) - T(6.0) * (at(a, 2) + at(a, -2)) + T(15.0) * (at(a, 1) + at(a, -1))
            - T(20.0) * f0;
      s += t;
    }
    return ko_coef * s;
  }

  // Sommerfeld RHS of the staged component c with point value uc
  // (reference evolution/boundaries.py:29-92).
  __device__ __forceinline__ T somm(int comp, double uc) const {
    const T d0 = d1(0), d1v = d1(1), d2v = d1(2);
    const T r = sqrt(xr[0] * xr[0] + xr[1] * xr[1] + xr[2] * xr[2]);
    const T rs = r > T(0) ? r : T(1);
    const bool unit = (comp == GXX || comp == GYY || comp == GZZ || comp == CW || comp == ALPHA);
    const T u0 = unit ? T(1) : T(0);
    return -((xr[0] * d0 + xr[1] * d1v + xr[2] * d2v) / rs + ((T)uc - u0) / rs);
  }
};

template <typename T>
struct PointProvider {
  RegStash st;
  const Ctx *C;
  const BSSN_GLOBAL double *src;
  const BSSN_GLOBAL double *f;
  int p[3];
  int64_t idx;

  bool isface;

  BSSN_HD __forceinline__ void setup(const Ctx &ctx, const double *u, int x, int y, int z) {
    C = &ctx;
    src = (const BSSN_GLOBAL double *)u;
    f = src;
    p[0] = x; p[1] = y; p[2] = z;
    idx = (int64_t)x * ctx.ax[0].stride + (int64_t)y * ctx.ax[1].stride + z;
    bool fc = false, owned = true;
    for (int a = 0; a < 3; ++a) {
      const Axis &A = ctx.ax[a];
      fc |= (A.somm_lo && p[a] == A.start) || (A.somm_hi && p[a] == A.end - 1);
      owned &= p[a] >= A.start && p[a] < A.end;
    }
    isface = fc && owned;
  }
  BSSN_HD __forceinline__ bool face() const { return isface; }
  BSSN_HD __forceinline__ void stage(int v) { f = src + visit_comp(v) * C->ncell; }
  BSSN_HD __forceinline__ void stage_comp(int comp) { f = src + comp * C->ncell; }

  // FP32 taps are taken relative to the point's own FP64 value (stencil
  // weights sum to zero), as in the tile provider.
  BSSN_HD __forceinline__ double centre() const { return sizeof(T) == sizeof(double) ? 0.0 : f[idx]; }
  BSSN_HD __forceinline__ T ld(int a, int oa) const {
    int q[3] = {p[0], p[1], p[2]};
    q[a] = wrap_index(q[a] + oa, C->ax[a].n);
    return (T)(f[(int64_t)q[0] * C->ax[0].stride + (int64_t)q[1] * C->ax[1].stride + q[2]] - centre());
  }
  BSSN_HD __forceinline__ T ld2(int a, int oa, int b, int ob) const {
    int q[3] = {p[0], p[1], p[2]};
    q[a] = wrap_index(q[a] + oa, C->ax[a].n);
    q[b] = wrap_index(q[b] + ob, C->ax[b].n);
    return (T)(f[(int64_t)q[0] * C->ax[0].stride + (int64_t)q[1] * C->ax[1].stride + q[2]] - centre());
  }
  BSSN_HD __forceinline__ T apply(int a, const Row8<T> &r) const {
    T s = T(0);
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (r.c[k] != T(0)) s += r.c[k] * ld(a, r.o0 + k);
    return s;
  }
  BSSN_HD __forceinline__ T val() const { return (T)(f[idx] - centre()); }
  BSSN_HD __forceinline__ T d1(int a) const {
    if (C->ax[a].reduced) return T(0);
    return apply(a, row_d1<T>(C->ax[a], p[a], C->inv_dx, false));
  }
  BSSN_HD __forceinline__ T d2(int a) const {
    if (C->ax[a].reduced) return T(0);
    return apply(a, row_d2<T>(C->ax[a], p[a], C->inv_dx2));
  }
  BSSN_HD __forceinline__ T dmix(int a, int b) const {
    if (C->ax[a].reduced || C->ax[b].reduced) return T(0);
    const Row8<T> ra = row_d1<T>(C->ax[a], p[a], C->inv_dx, false);
    const Row8<T> rb = row_d1<T>(C->ax[b], p[b], C->inv_dx, false);
    T s = T(0);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if (rb.c[j] == T(0)) continue;
      T inner = T(0);
#pragma unroll
      for (int i = 0; i < 8; ++i)
        if (ra.c[i] != T(0)) inner += ra.c[i] * ld2(a, ra.o0 + i, b, rb.o0 + j);
      s += rb.c[j] * inner;
    }
    return s;
  }
  BSSN_HD __forceinline__ T upw(int a, T coeff) const {
    if (C->ax[a].reduced) return T(0);
    return apply(a, row_upwind<T>(C->ax[a], p[a], C->inv_dx, (double)coeff));
  }
  BSSN_HD __forceinline__ T ko() const {
    T s = T(0);
    for (int a = 0; a < 3; ++a) {
      if (C->ax[a].reduced) continue;
      s += apply(a, row_d6<T>(C->ax[a], p[a]));
    }
    return (T)C->ko * s;
  }
  // Sommerfeld RHS of the staged component (reference evolution/boundaries.py:29-92)
  BSSN_HD __forceinline__ T somm(int comp, double uc) const {
    const T d0 = d1(0), d1v = d1(1), d2v = d1(2);
    const T X = (T)(C->xmin[0] + C->dx * (double)p[0]);
    const T Y = (T)(C->xmin[1] + C->dx * (double)p[1]);
    const T Z = (T)(C->xmin[2] + C->dx * (double)p[2]);
    const T r = sqrt(X * X + Y * Y + Z * Z);
    const T rs = r > T(0) ? r : T(1);
    const bool unit = (comp == GXX || comp == GYY || comp == GZZ || comp == CW || comp == ALPHA);
    const T u0 = unit ? T(1) : T(0);
    return -((X * d0 + Y * d1v + Z * d2v) / rs + ((T)uc - u0) / rs);
  }
};

}  // namespace bssn
