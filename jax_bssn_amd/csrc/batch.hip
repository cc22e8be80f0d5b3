// Batched evolution: many independent evolutions of uneven grid sizes
// advanced by shared launches per RK4 stage.
//
// A prepared batch lives in a caller-owned work buffer: per instance its RK4
// buffers and four stage descriptors (resolved parameters, stage buffers and
// weights, tile sets, first tile in each list), plus prefix sums of tiles
// and points.  Each stage is a launch of one workgroup per tile of
// the global list of lean tiles (tile t of instance i at prefix[i] + local
// index) and run the same per-point RHS and in-register stage update as the
// single-grid tile kernel, and, when some instance has Sommerfeld faces, four
// such launches over the global lists of its face tiles (slabs with closures
// along x, y or z only, edges and corners) with the closure-capable
// provider of that closure mask (one-sided rows and the radial face RHS).
// A step is a batched projection plus the stage launches, with no host work
// in between, so it can be captured in a HIP graph.
//
// Restricted to periodic or Sommerfeld faces on the tile grid of the
// chosen tile shape, mad_q = 1 and kappa = 0 (every other configuration
// runs through the single-grid entry points).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "bssn_amd.h"
#include "bssn_common.h"
#include "host_errors.h"
#include "rhs_point.h"
#include "stage_common.h"
#include "stencil_providers.h"
#include "tiles.h"

namespace bssn {
namespace {

struct BatchDesc {
  Ctx C;
  StageOut O;
  const double *us;
  TileSet lean;     // the instance's tiles without closure points
  TileSet face[4];  // with closures along x, y, z only, and edges and corners
  int lean_begin, face_begin[4];  // first global tile of the instance in each list
};

struct ProjDesc {
  double *state;
  int64_t ncell;
  int64_t cell_begin;
};

constexpr int BATCH_VARIANTS = 3;
__host__ __device__ constexpr int variant_tile(int v, int a) {
  // 0: 4 x 8 x 8, 1: 8 x 8 x 4, 2: 8 x 4 x 8 (256 threads each)
  return v == 0 ? (a == 0 ? 4 : 8) : (v == 1 ? (a == 2 ? 4 : 8) : (a == 1 ? 4 : 8));
}

// last instance whose first tile (prefix) is <= t (block-uniform)
__device__ __forceinline__ int batch_instance(const int *__restrict__ prefix, int ninst, int t) {
  int lo = 0, hi = ninst - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// The tile lists of a batch: lean tiles, then the face tiles with closures
// along x, y, z only and the edge and corner tiles (CLA = 1, 2, 4, 7).
__host__ __device__ constexpr int face_list(int cla) { return cla == 1 ? 0 : (cla == 2 ? 1 : (cla == 4 ? 2 : 3)); }

// The lean tiles, compiled for the occupancy of the single-grid interior
// tile kernel (one wave per SIMD in FP64, two in FP32); the face tiles (CL)
// with the closure-capable provider and the wave orientation of the
// single-grid face tiles.  One workgroup per tile: tile first + (blockIdx.x
// remapped per XCD) of a global list whose length is the last entry of its
// prefix sums.  (A
// loop over tiles in one workgroup keeps every thread-derived LDS offset
// live across iterations: 440 - 600 bytes of spills per lane, none without.)
template <typename T, int TX, int TY, int TZ, bool CL, int CLA>
__global__ void __launch_bounds__(TX *TY *TZ, (!CL && sizeof(T) == sizeof(double) ? 1 : 2))
    k_batch_tile(const BatchDesc *__restrict__ D, const int *__restrict__ prefix, int ninst, int first) {
  using P = TileProvider<T, TX, TY, TZ, false, CL, CLA, !CL && sizeof(T) == sizeof(double)>;
  __shared__ T smem[P::LDS_ELEMS];
  __shared__ double stash[STASH_SLOTS * TX * TY * TZ];
  const int tid = threadIdx.x;
  // each XCD takes a contiguous range of the launch's tiles (neighbouring
  // tiles of an instance share halos in one L2)
  const int t = first + xcd_remap((int)blockIdx.x, (int)gridDim.x);
  if (t >= prefix[ninst]) return;  // block-uniform, before any barrier
  // The descriptor table does not change while the launch runs: read it
  // through the constant address space, as a kernel argument is read, so
  // that the compiler may load a field again after the tile's barriers
  // instead of holding the whole context in registers across them.
  const BSSN_CONSTANT BatchDesc *Dk = (const BSSN_CONSTANT BatchDesc *)D;
  {
    const BatchDesc &d = *(const BatchDesc *)&Dk[batch_instance(prefix, ninst, t)];
    const Ctx &C = d.C;
    const StageOut &O = d.O;
    const double *const us = d.us;
    int bx, by, bz, slab = 0;
    if (CL) tile_of(d.face[face_list(CLA)], t - d.face_begin[face_list(CLA)], bx, by, bz, &slab);
    else tile_of(d.lean, t - d.lean_begin, bx, by, bz);
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
    {
      const BSSN_GLOBAL double *g = (const BSSN_GLOBAL double *)us;
#pragma unroll
      for (int c = 0; c < NCOMP; ++c) u[c] = g[c * C.ncell + idx];
    }
    Emitter<LdsStash<TX * TY * TZ>> out(C, O, idx, inside, pv.face(), pv.st);
    bssn_rhs_point<T>(pv, C, u, out);
  }
}

__global__ void __launch_bounds__(256) k_batch_project(const ProjDesc *__restrict__ D, int ninst, int64_t total) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  int lo = 0, hi = ninst - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (D[mid].cell_begin <= gid) lo = mid;
    else hi = mid - 1;
  }
  const ProjDesc d = D[lo];
  const int64_t n = d.ncell, i = gid - d.cell_begin;
  double g[6], A[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    g[c] = d.state[(GXX + c) * n + i];
    A[c] = d.state[(AXX + c) * n + i];
  }
  project(g, A);
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    d.state[(GXX + c) * n + i] = g[c];
    d.state[(AXX + c) * n + i] = A[c];
  }
}

// Work-buffer layout, in doubles, of a batch of n instances.
struct BatchLayout {
  int64_t desc, prefix, proj, bufs, total;
};

inline int64_t dwords(int64_t bytes) { return (bytes + 7) / 8; }

BatchLayout batch_layout(const bssn_grid *grids, int n) {
  BatchLayout L;
  L.desc = 0;
  L.prefix = L.desc + dwords((int64_t)4 * n * sizeof(BatchDesc));
  // prefix sums (n + 1 each) of the lean list, then of the four face lists
  L.proj = L.prefix + dwords((int64_t)5 * (n + 1) * sizeof(int));
  L.bufs = L.proj + dwords((int64_t)n * sizeof(ProjDesc));
  int64_t cells = 0;
  for (int i = 0; i < n; ++i) cells += (int64_t)grids[i].nx * grids[i].ny * grids[i].nz;
  L.total = L.bufs + 3 * NCOMP * cells;
  return L;
}

// tiles [first, first + blocks) of one list
template <typename T, bool CL, int CLA>
int launch_batch_t(int variant, const BatchDesc *D, const int *prefix, int ninst, int first, int blocks,
                   hipStream_t st) {
  switch (variant) {
    case 0:
      hipLaunchKernelGGL((k_batch_tile<T, 4, 8, 8, CL, CLA>), dim3(blocks), dim3(256), 0, st, D, prefix, ninst, first);
      break;
    case 1:
      hipLaunchKernelGGL((k_batch_tile<T, 8, 8, 4, CL, CLA>), dim3(blocks), dim3(256), 0, st, D, prefix, ninst, first);
      break;
    default:
      hipLaunchKernelGGL((k_batch_tile<T, 8, 4, 8, CL, CLA>), dim3(blocks), dim3(256), 0, st, D, prefix, ninst, first);
      break;
  }
  return check_hip(hipGetLastError(), "batched stage kernel launch");
}

// every tile of one list, in launches of at most max_blocks workgroups
template <typename T, bool CL, int CLA>
int launch_batch_list(int variant, const BatchDesc *D, const int *prefix, int ninst, int tiles, int max_blocks,
                      hipStream_t st) {
  int rc = BSSN_OK;
  for (int first = 0; first < tiles && !rc; first += max_blocks)
    rc = launch_batch_t<T, CL, CLA>(variant, D, prefix, ninst, first,
                                    tiles - first < max_blocks ? tiles - first : max_blocks, st);
  return rc;
}

// one stage: the lean list, then the four face lists (disjoint points)
template <typename T>
int launch_batch_stage(const bssn_batch_plan *plan, const BatchDesc *D, const int *prefix, hipStream_t st) {
  const int v = plan->tile_variant, n = plan->ninst, mb = plan->grid_blocks;
  const int *fp = prefix + (n + 1);
  const int *ft = plan->face_list_tiles;
  int rc = launch_batch_list<T, false, 7>(v, D, prefix, n, plan->total_tiles, mb, st);
  if (!rc) rc = launch_batch_list<T, true, 1>(v, D, fp, n, ft[face_list(1)], mb, st);
  if (!rc) rc = launch_batch_list<T, true, 2>(v, D, fp + (n + 1), n, ft[face_list(2)], mb, st);
  if (!rc) rc = launch_batch_list<T, true, 4>(v, D, fp + 2 * (n + 1), n, ft[face_list(4)], mb, st);
  if (!rc) rc = launch_batch_list<T, true, 7>(v, D, fp + 3 * (n + 1), n, ft[face_list(7)], mb, st);
  return rc;
}

}  // namespace
}  // namespace bssn

using namespace bssn;

extern "C" {

int64_t bssn_batch_work_size(const bssn_grid *grids, int n) {
  if (!grids || n < 1) return fail(BSSN_ERR_INVALID, "a batch needs at least one grid");
  return batch_layout(grids, n).total;
}

int bssn_batch_prepare(const bssn_params *params, const bssn_grid *grids, double *const *states, int n,
                       int precision, int tile_variant, int grid_blocks, double *work, bssn_batch_plan *plan,
                       void *stream) {
  if (!params || !grids || !states || !work || !plan || n < 1) return fail(BSSN_ERR_INVALID, "null argument");
  if (precision != BSSN_PREC_F64 && precision != BSSN_PREC_F32RHS)
    return fail(BSSN_ERR_INVALID, "unknown precision %d", precision);
  if (tile_variant < 0 || tile_variant >= BATCH_VARIANTS)
    return fail(BSSN_ERR_INVALID, "tile_variant must be 0..%d", BATCH_VARIANTS - 1);
  if (grid_blocks < 1) return fail(BSSN_ERR_INVALID, "grid_blocks must be positive");
  const BatchLayout L = batch_layout(grids, n);
  std::vector<BatchDesc> descs((size_t)4 * n);
  std::vector<int> prefix((size_t)5 * (n + 1));  // lean, then the four face lists
  std::vector<ProjDesc> proj((size_t)n);
  int64_t tiles = 0, ftiles[4] = {0, 0, 0, 0}, cells = 0, buf_off = L.bufs;
  const int t[3] = {variant_tile(tile_variant, 0), variant_tile(tile_variant, 1), variant_tile(tile_variant, 2)};
  for (int i = 0; i < n; ++i) {
    Ctx C;
    char err[512] = {0};
    int rc = resolve(&params[i], &grids[i], &C, err, sizeof err);
    if (rc) return fail(rc, "instance %d: %s", i, err);
    if (!states[i]) return fail(BSSN_ERR_INVALID, "instance %d: null state", i);
    if (C.kappa != 0.0 || C.ax[0].q != 1.0 || C.ax[1].q != 1.0 || C.ax[2].q != 1.0)
      return fail(BSSN_ERR_UNSUPPORTED, "instance %d: the batched launch takes mad_q = 1 and kappa = 0", i);
    bool aligned = true, whole = true;
    for (int a = 0; a < 3; ++a) {
      const Axis &A = C.ax[a];
      whole = whole && A.start == 0 && A.end == A.n && !A.reduced;
      if ((A.lbc || A.rbc) && (A.start % t[a] != 0 || A.end % t[a] != 0)) aligned = false;
    }
    if (C.any_shell && !(aligned && whole))
      return fail(BSSN_ERR_UNSUPPORTED,
                  "instance %d: Sommerfeld faces in a batch must lie on the %d x %d x %d tile grid (whole grids)", i,
                  t[0], t[1], t[2]);
    const TileSet lean = tile_set(C, C.any_shell ? 1 : 0, t[0], t[1], t[2]);
    TileSet face[4];
    for (int k = 0; k < 4; ++k) {
      face[k] = tile_set(C, 3 + k, t[0], t[1], t[2]);
      if (!C.any_shell) face[k].count = 0;
    }
    if (tiles + lean.count >= ((int64_t)1 << 31)) return fail(BSSN_ERR_INVALID, "batch has too many tiles");
    prefix[i] = (int)tiles;
    for (int k = 0; k < 4; ++k) {
      if (ftiles[k] + face[k].count >= ((int64_t)1 << 31)) return fail(BSSN_ERR_INVALID, "batch has too many tiles");
      prefix[(size_t)(k + 1) * (n + 1) + i] = (int)ftiles[k];
    }
    double *u = states[i];
    double *ustar = work + buf_off, *ua = ustar + NCOMP * C.ncell, *ub = ua + NCOMP * C.ncell;
    buf_off += 3 * NCOMP * C.ncell;
    const double *ins[4] = {u, ua, ub, ua};
    double *outs[4] = {ua, ub, ua, u};
    for (int s = 0; s < 4; ++s) {
      BatchDesc &d = descs[(size_t)s * n + i];
      std::memset(&d, 0, sizeof d);
      d.C = C;
      d.O = StageOut{s == 3 ? nullptr : u, ustar, outs[s], nullptr, s + 1, 0, nullptr, nullptr};
      stage_update(d.O, C.dt);
      d.us = ins[s];
      d.lean = lean;
      d.lean_begin = (int)tiles;
      for (int k = 0; k < 4; ++k) {
        d.face[k] = face[k];
        d.face_begin[k] = (int)ftiles[k];
      }
    }
    proj[i] = ProjDesc{u, C.ncell, cells};
    tiles += lean.count;
    for (int k = 0; k < 4; ++k) ftiles[k] += face[k].count;
    cells += C.ncell;
  }
  prefix[n] = (int)tiles;
  for (int k = 0; k < 4; ++k) prefix[(size_t)(k + 1) * (n + 1) + n] = (int)ftiles[k];
  const int64_t fsum = ftiles[0] + ftiles[1] + ftiles[2] + ftiles[3];
  if (fsum >= ((int64_t)1 << 31)) return fail(BSSN_ERR_INVALID, "batch has too many tiles");
  hipStream_t st = (hipStream_t)stream;
  char *base = (char *)work;
  int rc = check_hip(hipMemcpyAsync(base + L.desc * 8, descs.data(), descs.size() * sizeof(BatchDesc),
                                    hipMemcpyHostToDevice, st), "batch descriptor upload");
  if (!rc) rc = check_hip(hipMemcpyAsync(base + L.prefix * 8, prefix.data(), prefix.size() * sizeof(int),
                                         hipMemcpyHostToDevice, st), "batch prefix upload");
  if (!rc) rc = check_hip(hipMemcpyAsync(base + L.proj * 8, proj.data(), proj.size() * sizeof(ProjDesc),
                                         hipMemcpyHostToDevice, st), "batch projection upload");
  // the host vectors go out of scope on return: wait for the copies
  if (!rc) rc = check_hip(hipStreamSynchronize(st), "batch upload");
  if (rc) return rc;
  plan->ninst = n;
  plan->total_tiles = (int32_t)tiles;
  plan->face_tiles = (int32_t)fsum;
  for (int k = 0; k < 4; ++k) plan->face_list_tiles[k] = (int32_t)ftiles[k];
  plan->precision = precision;
  plan->tile_variant = tile_variant;
  plan->grid_blocks = grid_blocks;
  plan->total_cells = cells;
  plan->desc_offset = L.desc;
  plan->prefix_offset = L.prefix;
  plan->proj_offset = L.proj;
  return BSSN_OK;
}

int bssn_batch_step(const bssn_batch_plan *plan, double *work, void *stream) {
  if (!plan || !work) return fail(BSSN_ERR_INVALID, "null argument");
  const int32_t *ft = plan->face_list_tiles;
  if (plan->ninst < 1 || plan->total_tiles < 0 || ft[0] < 0 || ft[1] < 0 || ft[2] < 0 || ft[3] < 0 ||
      (int64_t)ft[0] + ft[1] + ft[2] + ft[3] != plan->face_tiles || plan->total_tiles + plan->face_tiles < 1 ||
      plan->grid_blocks < 1 || plan->tile_variant < 0 || plan->tile_variant >= BATCH_VARIANTS)
    return fail(BSSN_ERR_INVALID, "batch plan is not prepared");
  hipStream_t st = (hipStream_t)stream;
  const BatchDesc *D = (const BatchDesc *)(work + plan->desc_offset);
  const int *prefix = (const int *)(work + plan->prefix_offset);
  const ProjDesc *PD = (const ProjDesc *)(work + plan->proj_offset);
  const int64_t pblocks = (plan->total_cells + 255) / 256;
  hipLaunchKernelGGL(k_batch_project, dim3((unsigned)pblocks), dim3(256), 0, st, PD, plan->ninst, plan->total_cells);
  int rc = check_hip(hipGetLastError(), "batched projection launch");
  for (int s = 0; s < 4 && !rc; ++s) {
    const BatchDesc *Ds = D + (size_t)s * plan->ninst;
    rc = plan->precision == BSSN_PREC_F64 ? launch_batch_stage<double>(plan, Ds, prefix, st)
                                          : launch_batch_stage<float>(plan, Ds, prefix, st);
  }
  return rc;
}

}  // extern "C"
