// Intra-GPU ghost fill of a tiled field: the one-GPU form of the reference's
// tile halo refresh (parallel/communication.py:36-106, _sweep_tile).
//
// A tiled field holds counts[0] x counts[1] x counts[2] tiles, tile-major:
// tile t = (ix * cy + iy) * cz + iz is a block of ncomp component arrays of
// the allocated extent (width + 2 guard) per axis, C-order with z fastest.
// An axis with an upper endpoint gives its last tile one more owned sample,
// stored in the first nominal upper-guard slot.
//
// The sweep runs x, then y, then z (one launch each), so every slab carries
// the guards of the axes filled before it (edges and corners).  Per axis, a
// guard slab is either
//  * an exchange: the neighbour's nearest owned slab (the periodic wrap when
//    the axis has one tile; the upper endpoint shifts the slabs exactly as in
//    the reference), or
//  * a physical face (code 1 on the tile at that end): the exterior rule,
//    linear extrapolation from the face sample and its inward neighbour,
//    a constant copy of the face sample, or the slab left as it was.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "bssn_amd.h"
#include "host_errors.h"

namespace bssn {
namespace {

struct HaloGeom {
  int counts[3], width[3], alloc[3], endpoint[3];
  int g, ncomp, axis, exterior;
  int bc[6];
  int64_t tile_size;   // points of one tile (per component)
  int64_t slab_points; // guard points of one tile along the axis (both sides)
};

__global__ void __launch_bounds__(256) k_halo_axis(HaloGeom H, double *__restrict__ f) {
  const int ntiles = H.counts[0] * H.counts[1] * H.counts[2];
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)ntiles * H.slab_points) return;
  const int tile = (int)(id / H.slab_points);
  int64_t r = id % H.slab_points;
  const int a = H.axis, b = (a + 1) % 3, c = (a + 2) % 3;
  // position in the guard slab: (guard slot, coordinate along b, along c)
  const int pc = (int)(r % H.alloc[c]);
  r /= H.alloc[c];
  const int pb = (int)(r % H.alloc[b]);
  const int slot = (int)(r / H.alloc[b]);  // 0 .. 2g-1
  const int g = H.g, w = H.width[a];
  const bool upper = slot >= g;
  const int s = upper ? g + w + (slot - g) : slot;  // allocated index along a
  int ti[3];
  ti[2] = tile % H.counts[2];
  ti[1] = (tile / H.counts[2]) % H.counts[1];
  ti[0] = tile / (H.counts[2] * H.counts[1]);
  const int count = H.counts[a];
  const bool last = ti[a] == count - 1;
  const int e = (H.endpoint[a] && last) ? 1 : 0;
  auto at = [&](int t, int pa) -> int64_t {
    int p[3];
    p[a] = pa, p[b] = pb, p[c] = pc;
    return (int64_t)t * H.ncomp * H.tile_size + ((int64_t)p[0] * H.alloc[1] + p[1]) * H.alloc[2] + p[2];
  };
  const bool face = H.bc[2 * a + (upper ? 1 : 0)] == 1 && (upper ? last : ti[a] == 0);
  if (face) {
    if (H.exterior == 2) return;  // preserve
    const int base = upper ? g + w - 1 + e : g;
    const int inward = upper ? base - 1 : base + 1;
    const int64_t dst = at(tile, s), pbase = at(tile, base), pin = at(tile, inward);
    for (int k = 0; k < H.ncomp; ++k) {
      const int64_t o = (int64_t)k * H.tile_size;
      const double v0 = f[pbase + o];
      if (H.exterior == 1) {
        f[dst + o] = v0;
      } else {
        const double slope = upper ? v0 - f[pin + o] : f[pin + o] - v0;
        f[dst + o] = v0 + (double)(s - base) * slope;
      }
    }
    return;
  }
  int src_tile, src;
  if (count == 1) {
    const int owned = w + (H.endpoint[a] ? 1 : 0);
    int m = (s - g) % owned;
    if (m < 0) m += owned;
    src_tile = tile, src = g + m;
  } else {
    int nb[3] = {ti[0], ti[1], ti[2]};
    if (!upper) {
      nb[a] = (ti[a] - 1 + count) % count;
      const bool nb_end = H.endpoint[a] && nb[a] == count - 1;
      src = w + s + (nb_end ? 1 : 0);
    } else {
      nb[a] = (ti[a] + 1) % count;
      const int j = s - (g + w);
      if (e) {
        if (j == 0) return;  // the owned endpoint stays
        src = g + (j - 1);
      } else {
        src = g + j;
      }
    }
    src_tile = (nb[0] * H.counts[1] + nb[1]) * H.counts[2] + nb[2];
  }
  const int64_t dst = at(tile, s), from = at(src_tile, src);
  for (int k = 0; k < H.ncomp; ++k) f[dst + (int64_t)k * H.tile_size] = f[from + (int64_t)k * H.tile_size];
}

}  // namespace
}  // namespace bssn

using namespace bssn;

extern "C" {

int bssn_tiled_ghost_fill(const bssn_tile_layout *layout, const int32_t *bc, int exterior, int ncomp,
                          double *field, void *stream) {
  if (!layout || !bc || !field) return fail(BSSN_ERR_INVALID, "null argument");
  if (exterior < 0 || exterior > 2) return fail(BSSN_ERR_INVALID, "exterior must be 0 (linear), 1 (constant) or 2 (preserve)");
  if (ncomp < 1) return fail(BSSN_ERR_INVALID, "ncomp must be positive");
  HaloGeom H;
  H.g = layout->guard_cells;
  H.ncomp = ncomp;
  H.exterior = exterior;
  if (H.g < 1) return fail(BSSN_ERR_INVALID, "tiled fields require at least one guard cell");
  int64_t tile_size = 1, ntiles = 1;
  for (int a = 0; a < 3; ++a) {
    H.counts[a] = layout->counts[a];
    H.width[a] = layout->tile_shape[a];
    H.endpoint[a] = layout->upper_endpoints[a] ? 1 : 0;
    H.alloc[a] = H.width[a] + 2 * H.g;
    if (H.counts[a] < 1 || H.width[a] < 1)
      return fail(BSSN_ERR_INVALID, "tile counts and widths must be positive");
    if (H.counts[a] > 1 && H.width[a] < H.g)
      return fail(BSSN_ERR_INVALID, "a tile with neighbors must be at least as wide as its halo");
    if (H.endpoint[a] && H.g < 2)
      return fail(BSSN_ERR_INVALID, "an owned upper endpoint consumes a guard slot, so its axis needs at least two");
    tile_size *= H.alloc[a];
    ntiles *= H.counts[a];
  }
  for (int k = 0; k < 6; ++k) {
    if (bc[k] != 0 && bc[k] != 1) return fail(BSSN_ERR_INVALID, "boundary codes must be 0 (periodic) or 1 (Sommerfeld)");
    H.bc[k] = bc[k];
  }
  // The linear exterior rule takes its slope from the face sample and the
  // owned sample next to it; a face tile owning a single sample would read a
  // guard slot that the same launch writes.
  if (exterior == 0) {
    for (int a = 0; a < 3; ++a) {
      const int owned_lo = H.width[a] + (H.counts[a] == 1 ? H.endpoint[a] : 0);
      const int owned_hi = H.width[a] + H.endpoint[a];
      if ((H.bc[2 * a] == 1 && owned_lo < 2) || (H.bc[2 * a + 1] == 1 && owned_hi < 2))
        return fail(BSSN_ERR_INVALID,
                    "linear extrapolation on a physical face of axis %d needs at least two owned samples per face tile",
                    a);
    }
  }
  if (tile_size * ntiles * ncomp >= ((int64_t)1 << 40)) return fail(BSSN_ERR_INVALID, "tiled field too large");
  H.tile_size = tile_size;
  for (int a = 0; a < 3; ++a) {
    H.axis = a;
    H.slab_points = (int64_t)2 * H.g * (tile_size / H.alloc[a]);
    const int64_t total = ntiles * H.slab_points;
    hipLaunchKernelGGL(k_halo_axis, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream, H,
                       field);
    const int rc = check_hip(hipGetLastError(), "ghost fill kernel launch");
    if (rc) return rc;
  }
  return BSSN_OK;
}

}  // extern "C"
