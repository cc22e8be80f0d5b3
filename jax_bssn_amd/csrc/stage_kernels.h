// The fused RK4 stage kernels (tile kernels and closure-shell kernel) and
// their launchers, templated on the RHS arithmetic type.  Each precision is
// instantiated in a translation unit of its own (stage_f64.hip,
// stage_f32.hip) so that the two compile in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include <type_traits>

#include "adjoint_point.h"
#include "bssn_amd.h"
#include "bssn_common.h"
#include "host_errors.h"
#include "rhs_point.h"
#include "stage_common.h"
#include "stencil_providers.h"
#include "tiles.h"

// Minimum waves per SIMD a tile kernel is compiled for (caps its registers
// at 512 / waves per lane, VGPRs and AGPRs together).  The FP64 kernels
// without closure rows hold their whole point algebra in registers at one
// wave per SIMD (about 500, no scratch); at two they spill 600-1000 B per
// lane and run 12-40 % slower (256^3 puncture, MI355X).  The closure
// kernels and the FP32 ones run best at two.
#ifndef BSSN_F32_INTERIOR_MIN_WAVES
#define BSSN_F32_INTERIOR_MIN_WAVES 2
#endif
template <typename T, bool CL>
constexpr int tile_min_waves() {
  return CL ? 2 : (sizeof(T) == sizeof(double) ? 1 : BSSN_F32_INTERIOR_MIN_WAVES);
}

namespace bssn {

__device__ __forceinline__ void load_point(const Ctx &C, const double *__restrict__ u, int64_t idx,
                                           double (&v)[NCOMP]) {
  const BSSN_GLOBAL double *g = (const BSSN_GLOBAL double *)u;
#pragma unroll
  for (int c = 0; c < NCOMP; ++c) v[c] = g[c * C.ncell + idx];
}

// ------------------------------------------------------ feature recording

// A stencil provider that also stores the stencil values the per-point
// RHS asks of it -- the features of adjoint_point.h: first, second and
// upwind derivatives of the visited component, and on a Sommerfeld face
// point the first derivatives of the face RHS -- into feat[slot * n + idx].
// The Kreiss-Oliger sums enter the RHS linearly and their adjoint needs no
// feature, so they are not stored.  Run through the stage kernels with an
// emitter that keeps nothing, the compiler drops the algebra after the
// stencils (and the dissipation stencils altogether).
template <class P>
struct Recorder : P {
  // (names of their own: the providers have members such as idx)
  BSSN_GLOBAL double *rec_feat;
  int64_t rec_n, rec_idx;
  bool rec_on;
  int vc;  // component of the current visit
  __device__ __forceinline__ void record_to(double *f, int64_t ncell, int64_t i, bool on) {
    rec_feat = (BSSN_GLOBAL double *)f;
    rec_n = ncell;
    rec_idx = i;
    rec_on = on;
  }
  __device__ __forceinline__ void keep(int slot, double v) const {
    if (rec_on) rec_feat[slot * rec_n + rec_idx] = v;
  }
  __device__ __forceinline__ void stage(int v) {
    P::stage(v);
    vc = visit_comp(v);
  }
  __device__ __forceinline__ auto d1(int a) const {
    const auto v = P::d1(a);
    keep(F_D1 + 3 * vc + a, (double)v);
    return v;
  }
  __device__ __forceinline__ auto d2(int a) const {
    const auto v = P::d2(a);
    keep(F_D2 + 6 * d2_slot(vc) + sym(a, a), (double)v);
    return v;
  }
  __device__ __forceinline__ auto dmix(int a, int b) const {
    const auto v = P::dmix(a, b);
    keep(F_D2 + 6 * d2_slot(vc) + sym(a, b), (double)v);
    return v;
  }
  template <typename C>
  __device__ __forceinline__ auto upw(int a, C coeff) const {
    const auto v = P::upw(a, coeff);
    keep(F_UP + 3 * vc + a, (double)v);
    return v;
  }
  __device__ __forceinline__ auto somm(int comp, double uc) const {
    d1(0);
    d1(1);
    d1(2);
    return P::somm(comp, uc);
  }
};

// The emitter of the recording launches: the RHS values are not kept.
struct RecordEmitter {
  __device__ __forceinline__ void prefetch(int) {}
  __device__ __forceinline__ void put(int, double) {}
  __device__ __forceinline__ void put_face(int, double) {}
  __device__ __forceinline__ bool has_matter() const { return false; }
  __device__ __forceinline__ double matter(int) const { return 0.0; }
  __device__ __forceinline__ double extra_A(int) const { return 0.0; }
  __device__ __forceinline__ void done() {}
};

// Point values and the recorded stencils of one point (REC launches: the
// feature array is O.rhs).
template <typename T, class P>
__device__ __forceinline__ void record_point(P &pv, const Ctx &C, const StageOut &O, int64_t idx, bool mine,
                                             const double (&u)[NCOMP]) {
  pv.record_to(O.rhs, C.ncell, idx, mine);
  if (mine)
    for (int c = 0; c < NCOMP; ++c) ((BSSN_GLOBAL double *)O.rhs)[(F_VAL + c) * C.ncell + idx] = u[c];
  RecordEmitter out;
  bssn_rhs_point<T>(pv, C, u, out);
}

// ------------------------------------------------------------ tile kernel

template <typename T, int TX, int TY, int TZ, bool MAD, bool CL, int CLA = 7, bool REC = false>
__global__ void __launch_bounds__(TX *TY *TZ, (tile_min_waves<T, CL>()))
    k_stage_tile(Ctx C, const double *__restrict__ us, StageOut O, TileSet S) {
  using PB = TileProvider<T, TX, TY, TZ, MAD, CL, CLA, sizeof(T) == sizeof(double) && !CL>;
  using P = std::conditional_t<REC, Recorder<PB>, PB>;
  __shared__ T smem[P::LDS_ELEMS];
  __shared__ double stash[STASH_SLOTS * TX * TY * TZ];
  int bx, by, bz, slab;
  tile_of(S, xcd_remap((int)blockIdx.x, S.count), bx, by, bz, &slab);
  const int tid = threadIdx.x;
  // Waves are planes across the slowest of the three thread coordinates.
  // A face tile puts that coordinate along its closure axis, so a wave's
  // lanes share (or nearly share) one closure row instead of diverging.
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
  pv.setup(C, us, smem, stash, bx * TX, by * TY, bz * TZ, tx, ty, tz);
  double u[NCOMP];
  load_point(C, us, idx, u);
  // without in-tile closures the shell points belong to the shell kernel
  const bool mine = inside && (CL || !(C.any_shell && in_shell(C, x, y, z)));
  if constexpr (REC) {
    record_point<T>(pv, C, O, idx, mine, u);
  } else {
    Emitter<LdsStash<TX * TY * TZ>> out(C, O, idx, mine, pv.face(), pv.st);
    bssn_rhs_point<T>(pv, C, u, out);
  }
}

// ------------------------------------------------------- region tiles

// A box one plane thick along y (the Cartoon centre plane) as 16 x 1 x 16
// tiles over [lo, hi), tile t at lo + ((t / nz) * 16, 0, (t % nz) * 16),
// with closure rows along x and z (the box may end on those faces).  The
// square x-z footprint loads the fewest halo points per tile point (13.2
// against 14.5 for 8 x 1 x 32 over the seven planes of the y halo).
constexpr int RTX = 16, RTY = 1, RTZ = 16;

struct RegionTiles {
  int lo[3], hi[3];
  int nz;  // tiles along z
};

inline RegionTiles region_tiles(const int *lo, const int *hi, int *count) {
  RegionTiles R;
  for (int a = 0; a < 3; ++a) R.lo[a] = lo[a], R.hi[a] = hi[a];
  const int nx = (hi[0] - lo[0] + RTX - 1) / RTX;
  R.nz = (hi[2] - lo[2] + RTZ - 1) / RTZ;
  *count = nx * R.nz;
  return R;
}

// Region tiles apply: no MAD blend, one plane along y, no closures along y,
// and along x and z the tile grid anchored at lo meets every closure face
// on a tile boundary (then a face tile spans the face's last 8 or 32
// points, and its closure rows, which reach 7 points inward, stay inside
// the box of the tile plus its 3-point halo).
inline bool region_tiled(const Ctx &C, const int *lo, const int *hi) {
  const bool mad = C.ax[0].q != 1.0 || C.ax[1].q != 1.0 || C.ax[2].q != 1.0;
  if (mad || hi[1] - lo[1] != 1 || C.ax[1].lbc || C.ax[1].rbc || C.ax[1].reduced) return false;
  const int t[3] = {RTX, RTY, RTZ};
  for (int a = 0; a < 3; a += 2) {
    const Axis &A = C.ax[a];
    if (A.lbc && (lo[a] - A.start) % t[a] != 0) return false;
    if (A.rbc && (A.end - lo[a]) % t[a] != 0) return false;
  }
  return true;
}

template <typename T>
__global__ void __launch_bounds__(RTX *RTY *RTZ, (tile_min_waves<T, true>()))
    k_stage_region(Ctx C, const double *__restrict__ us, StageOut O, RegionTiles R) {
  using P = TileProvider<T, RTX, RTY, RTZ, false, true, 5>;
  __shared__ T smem[P::LDS_ELEMS];
  __shared__ double stash[STASH_SLOTS * RTX * RTY * RTZ];
  const int tile = (int)blockIdx.x;
  const int x0 = R.lo[0] + (tile / R.nz) * RTX, y0 = R.lo[1], z0 = R.lo[2] + (tile % R.nz) * RTZ;
  const int tid = threadIdx.x;
  const int tz = tid % RTZ, tx = tid / RTZ;
  const int x = x0 + tx, z = z0 + tz;
  const bool inside = x < R.hi[0] && z < R.hi[2];
  const int xc = x < R.hi[0] ? x : R.hi[0] - 1;
  const int zc = z < R.hi[2] ? z : R.hi[2] - 1;
  const int64_t idx = (int64_t)xc * C.ax[0].stride + (int64_t)y0 * C.ax[1].stride + zc;
  P pv;
  pv.setup(C, us, smem, stash, x0, y0, z0, tx, 0, tz);
  double u[NCOMP];
  load_point(C, us, idx, u);
  Emitter<LdsStash<RTX * RTY * RTZ>> out(C, O, idx, inside, pv.face(), pv.st);
  bssn_rhs_point<T>(pv, C, u, out);
}

template <typename T>
inline int launch_region_t(const Ctx &C, const double *us, const StageOut &O, const int *lo, const int *hi,
                           hipStream_t st) {
  int count = 0;
  const RegionTiles R = region_tiles(lo, hi, &count);
  if (count <= 0) return BSSN_OK;
  hipLaunchKernelGGL((k_stage_region<T>), dim3(count), dim3(RTX * RTY * RTZ), 0, st, C, us, O, R);
  return check_hip(hipGetLastError(), "region tile kernel launch");
}

// ----------------------------------------------------------- shell kernel

__device__ __forceinline__ bool in_slab(const Slabs &S, int s, int x, int y, int z) {
  return x >= S.lo[s][0] && x < S.hi[s][0] && y >= S.lo[s][1] && y < S.hi[s][1] && z >= S.lo[s][2] &&
         z < S.hi[s][2];
}

template <typename T, bool REC = false>
__global__ void __launch_bounds__(256) k_stage_shell(Ctx C, const double *__restrict__ us, StageOut O, Slabs S) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= S.first[S.count]) return;
  int s = 0;
  while (s + 1 < S.count && gid >= S.first[s + 1]) ++s;
  const int64_t l = gid - S.first[s];
  const int ez = S.hi[s][2] - S.lo[s][2], ey = S.hi[s][1] - S.lo[s][1];
  const int z = S.lo[s][2] + (int)(l % ez);
  const int y = S.lo[s][1] + (int)((l / ez) % ey);
  const int x = S.lo[s][0] + (int)(l / ((int64_t)ez * ey));
  for (int t = 0; t < s; ++t)
    if (in_slab(S, t, x, y, z)) return;

  std::conditional_t<REC, Recorder<PointProvider<T>>, PointProvider<T>> pv;
  pv.setup(C, us, x, y, z);
  double u[NCOMP];
  load_point(C, us, pv.idx, u);
  if constexpr (REC) {
    record_point<T>(pv, C, O, pv.idx, true, u);
  } else {
    Emitter<RegStash> out(C, O, pv.idx, true, pv.face(), pv.st);
    bssn_rhs_point<T>(pv, C, u, out);
  }
}

// Tile shape and the minimum waves per SIMD the tile kernels are compiled
// for (overridable at build time to produce tuning variants).
#ifndef BSSN_TILE_X
#define BSSN_TILE_X 4
#define BSSN_TILE_Y 8
#define BSSN_TILE_Z 8
#endif
constexpr int TX = BSSN_TILE_X, TY = BSSN_TILE_Y, TZ = BSSN_TILE_Z;

// launch parts of a stage (bssn_time_stage_part / bssn_stage_points)
constexpr int PART_TILES = BSSN_PART_TILES, PART_SHELL = BSSN_PART_SHELL, PART_LEAN = BSSN_PART_LEAN,
              PART_CLOSURE = BSSN_PART_CLOSURE;

// Closures can be evaluated inside the tiles when every closure axis has its
// owned range aligned to the tile grid (each face tile starts / ends on the
// face); otherwise the closure shells go to the shell kernel.
inline bool tiles_aligned(const Ctx &C) {
  const int t[3] = {TX, TY, TZ};
  for (int a = 0; a < 3; ++a) {
    const Axis &A = C.ax[a];
    if (!(A.lbc || A.rbc)) continue;
    if (A.start % t[a] != 0 || A.end % t[a] != 0) return false;
  }
  return true;
}

template <typename T, bool MAD, int CLA, bool REC>
inline int launch_face_tiles(const Ctx &C, const double *us, const StageOut &O, hipStream_t st, int kind) {
  const TileSet S = tile_set(C, kind, TX, TY, TZ);
  if (S.count <= 0) return BSSN_OK;
  hipLaunchKernelGGL((k_stage_tile<T, TX, TY, TZ, MAD, true, CLA, REC>), dim3(S.count), dim3(TX * TY * TZ), 0, st, C,
                     us, O, S);
  return check_hip(hipGetLastError(), "tile kernel launch");
}

template <typename T, bool MAD, bool CL, bool REC>
inline int launch_stage_t(const Ctx &C, const double *us, const StageOut &O, hipStream_t st, int parts) {
  int rc = BSSN_OK;
  // lean kernel: every tile, or with in-tile closures only the interior tiles
  if (parts & (PART_TILES | PART_LEAN)) {
    const TileSet S = tile_set(C, CL ? 1 : 0, TX, TY, TZ);
    if (S.count > 0) {
      hipLaunchKernelGGL((k_stage_tile<T, TX, TY, TZ, MAD, false, 7, REC>), dim3(S.count), dim3(TX * TY * TZ), 0, st,
                         C, us, O, S);
      rc = check_hip(hipGetLastError(), "tile kernel launch");
      if (rc) return rc;
    }
  }
  // closure-capable kernels on the face tiles: one per slab whose tiles
  // have closures along that axis only, one for edges and corners (with
  // MAD every face tile takes the general kernel)
  if (CL && (parts & (PART_TILES | PART_CLOSURE))) {
    if constexpr (MAD) {
      rc = launch_face_tiles<T, MAD, 7, REC>(C, us, O, st, 2);
    } else {
      if (!rc) rc = launch_face_tiles<T, MAD, 1, REC>(C, us, O, st, 3);
      if (!rc) rc = launch_face_tiles<T, MAD, 2, REC>(C, us, O, st, 4);
      if (!rc) rc = launch_face_tiles<T, MAD, 4, REC>(C, us, O, st, 5);
      if (!rc) rc = launch_face_tiles<T, MAD, 7, REC>(C, us, O, st, 6);
    }
    if (rc) return rc;
  }
  if ((parts & PART_SHELL) && C.any_shell && !CL) {
    Slabs S = make_slabs(C);
    const int64_t total = S.first[S.count];
    if (total > 0) {
      const int64_t blocks = (total + 255) / 256;
      hipLaunchKernelGGL((k_stage_shell<T, REC>), dim3((unsigned)blocks), dim3(256), 0, st, C, us, O, S);
      rc = check_hip(hipGetLastError(), "shell kernel launch");
    }
  }
  return rc;
}

template <typename T, bool REC = false>
inline int launch_stage_p(const Ctx &C, const double *us, const StageOut &O, hipStream_t st, int parts) {
  const bool mad = C.ax[0].q != 1.0 || C.ax[1].q != 1.0 || C.ax[2].q != 1.0;
  const bool cl = C.any_shell && tiles_aligned(C);
  if (mad)
    return cl ? launch_stage_t<T, true, true, REC>(C, us, O, st, parts)
              : launch_stage_t<T, true, false, REC>(C, us, O, st, parts);
  return cl ? launch_stage_t<T, false, true, REC>(C, us, O, st, parts)
            : launch_stage_t<T, false, false, REC>(C, us, O, st, parts);
}

// Launchers of one precision (defined in stage_f64.hip / stage_f32.hip).
int launch_stage_f64(const Ctx &C, const double *us, const StageOut &O, hipStream_t st, int parts);
// The FP64 stage kernels recording the features of every point into feat
// (NFEAT * ncell doubles) instead of updating (defined in record_f64.hip).
int launch_record_f64(const Ctx &C, const double *us, double *feat, hipStream_t st);
int launch_stage_f32(const Ctx &C, const double *us, const StageOut &O, hipStream_t st, int parts);
// the closure-capable point kernel over a list of slabs
int launch_slabs_f64(const Ctx &C, const double *us, const StageOut &O, const Slabs &S, hipStream_t st);
int launch_slabs_f32(const Ctx &C, const double *us, const StageOut &O, const Slabs &S, hipStream_t st);
// the region tile kernel (region_tiled() boxes)
int launch_region_f64(const Ctx &C, const double *us, const StageOut &O, const int *lo, const int *hi, hipStream_t st);
int launch_region_f32(const Ctx &C, const double *us, const StageOut &O, const int *lo, const int *hi, hipStream_t st);

template <typename T>
inline int launch_slabs_t(const Ctx &C, const double *us, const StageOut &O, const Slabs &S, hipStream_t st) {
  const int64_t total = S.first[S.count];
  if (total <= 0) return BSSN_OK;
  const int64_t blocks = (total + 255) / 256;
  hipLaunchKernelGGL((k_stage_shell<T>), dim3((unsigned)blocks), dim3(256), 0, st, C, us, O, S);
  return check_hip(hipGetLastError(), "point kernel launch");
}

}  // namespace bssn
