// Forward-mode (tangent) stage kernels shared by the tangent translation
// units: dual-number stash and emitter, the dual LDS tile provider and the
// tile kernel, templated on the number of directions ND.  The tile kernels
// of each ND are instantiated in a translation unit of their own
// (tangent_tile1.hip, tangent_tile2.hip) so that they compile in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "bssn_amd.h"
#include "bssn_common.h"
#include "dual.h"
#include "host_errors.h"
#include "rhs_point.h"
#include "stage_common.h"
#include "stage_kernels.h"
#include "stencil_providers.h"
#include "tiles.h"

namespace bssn {

constexpr int MAX_DIRS = 2;

template <typename V>
struct StashT {
  V v[STASH_SLOTS];
  BSSN_HD __forceinline__ void put(int slot, V x) { v[slot] = x; }
  BSSN_HD __forceinline__ V get(int slot) const { return v[slot]; }
};

// The value and the ND tangents of one component staged side by side in
// LDS (ND + 1 boxes of the ordinary FP64 tile); every stencil is the
// ordinary tile stencil evaluated on each box.
template <int ND, int TX_, int TY_, int TZ_, bool CL>
struct DualTileProvider {
  using B = TileProvider<double, TX_, TY_, TZ_, false, CL>;
  using V = DualN<ND>;
  static constexpr int NBOX = B::NBOX, NBOXP = B::NBOXP, NSLOT = B::NSLOT, NT = B::NT;
  static constexpr int LDS_ELEMS = (ND + 1) * NBOXP;
  mutable B b;  // geometry, closure rows and stencils; b.cur selects the box
  StashT<V> st;
  BSSN_LDS double *lds;
  const BSSN_GLOBAL double *src, *dsrc;
  int64_t dstride;

  __device__ __forceinline__ void setup(const Ctx &C, const double *u, const double *du, double *smem, int x0,
                                        int y0, int z0, int tx, int ty, int tz) {
    b.setup(C, u, smem, smem, x0, y0, z0, tx, ty, tz, false);  // b's LDS stash is not used
    lds = (BSSN_LDS double *)smem;
    src = (const BSSN_GLOBAL double *)u;
    dsrc = (const BSSN_GLOBAL double *)du;
    dstride = (int64_t)NCOMP * C.ncell;
  }
  // Stage visit v: every box loaded between two barriers (the first waits
  // until every lane is done with visit v - 1).  The loads of all boxes are
  // issued before the first LDS store, so their latencies overlap.
  __device__ __forceinline__ void stage(int v) {
    const int64_t off = (int64_t)visit_comp(v) * b.ncell;
    double t[ND + 1][NSLOT];
#pragma unroll
    for (int l = 0; l <= ND; ++l) {
      const BSSN_GLOBAL double *f = (l == 0 ? src : dsrc + (l - 1) * dstride) + off;
#pragma unroll
      for (int k = 0; k < NSLOT; ++k) t[l][k] = f[b.goff[k]];
    }
    if (v > 0) __syncthreads();
#pragma unroll
    for (int l = 0; l <= ND; ++l) {
      BSSN_LDS double *buf = lds + l * NBOXP;
#pragma unroll
      for (int k = 0; k < NSLOT; ++k) {
        const int e = b.tid + k * NT;
        if (k < NSLOT - 1 || e < NBOX) buf[B::lds_index(e)] = t[l][k];
      }
    }
    __syncthreads();
  }
  template <class F>
  __device__ __forceinline__ V lanes(F &&fn) const {
    V r;
    b.cur = lds;
    r.v = fn();
#pragma unroll
    for (int l = 0; l < ND; ++l) {
      b.cur = lds + (l + 1) * NBOXP;
      r.d[l] = fn();
    }
    return r;
  }
  __device__ __forceinline__ bool face() const { return b.face(); }
  __device__ __forceinline__ V val() const {
    return lanes([&] { return b.val(); });
  }
  __device__ __forceinline__ V d1(int a) const {
    return lanes([&] { return b.d1(a); });
  }
  __device__ __forceinline__ V d2(int a) const {
    return lanes([&] { return b.d2(a); });
  }
  __device__ __forceinline__ V dmix(int a, int c) const {
    return lanes([&] { return b.dmix(a, c); });
  }
  __device__ __forceinline__ V upw(int a, V coeff) const {
    const double cv = coeff.v;
    return lanes([&] { return b.upw(a, cv); });
  }
  __device__ __forceinline__ V ko() const {
    return lanes([&] { return b.ko(); });
  }
  __device__ __forceinline__ V somm(int comp, V uc) const {
    const V d0 = d1(0), d1v = d1(1), d2v = d1(2);
    const double X = b.xr[0], Y = b.xr[1], Z = b.xr[2];
    const double r = ::sqrt(X * X + Y * Y + Z * Z);
    const double rs = r > 0.0 ? r : 1.0;
    const bool unit = (comp == GXX || comp == GYY || comp == GZZ || comp == CW || comp == ALPHA);
    const double u0 = unit ? 1.0 : 0.0;
    return -((X * d0 + Y * d1v + Z * d2v) / rs + (uc - u0) / rs);
  }
};

// Tangent buffers of one stage (the value buffers and weights are a
// StageOut); direction l of each sits at + l * NCOMP * ncell.
struct TangentOut {
  const double *du0;
  double *dustar, *dunext, *drhs;
  int skip_value;  // compute the value but store only the tangents
};

template <int ND>
struct DualEmitter {
  using V = DualN<ND>;
  int64_t n, idx, ds;
  double wn, ws;
  const double *nbase, *sbase, *dnbase, *dsbase;
  double *ustar, *unext, *rhs, *dustar, *dunext, *drhs;
  int mode, reproject;
  bool active, face, value_out;
  V gn[6], An[6];

  __device__ __forceinline__ DualEmitter(const Ctx &C, const StageOut &O, const TangentOut &T, int64_t i, bool act,
                                         bool on_face)
      : n(C.ncell), idx(i), ds((int64_t)NCOMP * C.ncell), wn(O.wn), ws(O.ws), nbase(O.nbase), sbase(O.sbase),
        ustar(O.ustar), unext(O.unext), rhs(O.rhs), dustar(T.dustar), dunext(T.dunext), drhs(T.drhs),
        mode(O.mode), reproject(O.reproject), active(act), face(on_face), value_out(!T.skip_value) {
    dnbase = O.mode == 4 ? T.dustar : T.du0;
    dsbase = O.mode == 1 ? T.du0 : ((O.mode == 2 || O.mode == 3) ? T.dustar : nullptr);
  }
  __device__ __forceinline__ void put(int c, V k) {
    if (!face) write(c, k);
  }
  __device__ __forceinline__ void put_face(int c, V k) { write(c, k); }
  __device__ __forceinline__ void prefetch(int) {}
  __device__ __forceinline__ void store(double *v, double *dv, int64_t at, const V &x) {
    if (value_out) v[at] = x.v;
#pragma unroll
    for (int l = 0; l < ND; ++l) dv[l * ds + at] = x.d[l];
  }
  __device__ __forceinline__ void write(int c, V k) {
    if (!active) return;
    const int64_t at = c * n + idx;
    if (mode == 0) {
      store(rhs, drhs, at, k);
      return;
    }
    V un;
    un.v = nbase[at] + wn * k.v;
#pragma unroll
    for (int l = 0; l < ND; ++l) un.d[l] = dnbase[l * ds + at] + wn * k.d[l];
    if (sbase) {
      V us;
      us.v = sbase[at] + ws * k.v;
#pragma unroll
      for (int l = 0; l < ND; ++l) us.d[l] = dsbase[l * ds + at] + ws * k.d[l];
      store(ustar, dustar, at, us);
    }
    if (c <= GZZ) {
      gn[c] = un;
    } else if (c >= AXX && c <= AZZ) {
      An[c - AXX] = un;
    } else {
      store(unext, dunext, at, un);
    }
  }
  __device__ __forceinline__ void done() {
    if (!active || mode == 0) return;
    project(gn, An);
    if (mode == 4 && reproject) project(gn, An);
    for (int c = 0; c < 6; ++c) {
      store(unext, dunext, (GXX + c) * n + idx, gn[c]);
      store(unext, dunext, (AXX + c) * n + idx, An[c]);
    }
  }
  __device__ __forceinline__ bool has_matter() const { return false; }
  __device__ __forceinline__ double matter(int) const { return 0.0; }
  __device__ __forceinline__ double extra_A(int) const { return 0.0; }
};

template <int ND>
__device__ __forceinline__ void load_dual(const Ctx &C, const double *us, const double *dus, int64_t i,
                                          DualN<ND> (&u)[NCOMP]) {
  const int64_t ds = (int64_t)NCOMP * C.ncell;
  for (int c = 0; c < NCOMP; ++c) {
    u[c].v = us[c * C.ncell + i];
#pragma unroll
    for (int l = 0; l < ND; ++l) u[c].d[l] = dus[l * ds + c * C.ncell + i];
  }
}

// The tile kernel of the tangent stage: tile sets, XCD remap and the
// face-tile wave layout of k_stage_tile.
template <int ND, bool CL>
__global__ void __launch_bounds__(TX *TY *TZ, 1)
    k_stage_tangent_tile(Ctx C, const double *__restrict__ us, const double *__restrict__ dus, StageOut O,
                         TangentOut T, TileSet S) {
  using P = DualTileProvider<ND, TX, TY, TZ, CL>;
  __shared__ double smem[P::LDS_ELEMS];
  int bx, by, bz, slab;
  tile_of(S, xcd_remap((int)blockIdx.x, S.count), bx, by, bz, &slab);
  const int tid = threadIdx.x;
  int tx, ty, tz;
  if (!CL || slab == 0) {
    tz = tid % TZ, ty = (tid / TZ) % TY, tx = tid / (TY * TZ);
  } else if (slab == 1) {
    tz = tid % TZ, tx = (tid / TZ) % TX, ty = tid / (TX * TZ);
  } else {
    ty = tid % TY, tx = (tid / TY) % TX, tz = tid / (TX * TY);
  }
  const int x = bx * TX + tx, y = by * TY + ty, z = bz * TZ + tz;
  const bool inside = x < C.ax[0].n && y < C.ax[1].n && z < C.ax[2].n;
  const int xc = x < C.ax[0].n ? x : C.ax[0].n - 1;
  const int yc = y < C.ax[1].n ? y : C.ax[1].n - 1;
  const int zc = z < C.ax[2].n ? z : C.ax[2].n - 1;
  const int64_t idx = (int64_t)xc * C.ax[0].stride + (int64_t)yc * C.ax[1].stride + zc;

  P pv;
  pv.setup(C, us, dus, smem, bx * TX, by * TY, bz * TZ, tx, ty, tz);
  DualN<ND> u[NCOMP];
  load_dual(C, us, dus, idx, u);
  DualEmitter<ND> out(C, O, T, idx, inside, pv.face());
  bssn_rhs_point<DualN<ND>>(pv, C, u, out);
}

template <int ND>
int launch_tangent_tiles_t(const Ctx &C, const double *us, const double *dus, const StageOut &O, const TangentOut &T,
                           hipStream_t st) {
  const dim3 block(TX * TY * TZ);
  if (!C.any_shell) {
    const TileSet S = tile_set(C, 0, TX, TY, TZ);
    hipLaunchKernelGGL((k_stage_tangent_tile<ND, false>), dim3(S.count), block, 0, st, C, us, dus, O, T, S);
    return check_hip(hipGetLastError(), "tangent tile kernel launch");
  }
  const TileSet lean = tile_set(C, 1, TX, TY, TZ);
  if (lean.count > 0) {
    hipLaunchKernelGGL((k_stage_tangent_tile<ND, false>), dim3(lean.count), block, 0, st, C, us, dus, O, T, lean);
    const int rc = check_hip(hipGetLastError(), "tangent tile kernel launch");
    if (rc) return rc;
  }
  const TileSet face = tile_set(C, 2, TX, TY, TZ);
  if (face.count > 0) {
    hipLaunchKernelGGL((k_stage_tangent_tile<ND, true>), dim3(face.count), block, 0, st, C, us, dus, O, T, face);
    return check_hip(hipGetLastError(), "tangent tile kernel launch");
  }
  return BSSN_OK;
}

// The tile launchers of one ND (tangent_tile1.hip / tangent_tile2.hip).
int launch_tangent_tiles_1(const Ctx &C, const double *us, const double *dus, const StageOut &O, const TangentOut &T,
                           hipStream_t st);
int launch_tangent_tiles_2(const Ctx &C, const double *us, const double *dus, const StageOut &O, const TangentOut &T,
                           hipStream_t st);

}  // namespace bssn
