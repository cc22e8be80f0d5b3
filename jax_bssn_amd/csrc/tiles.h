// Tile sets of the tile kernels: which tiles one launch covers and the
// block -> tile mapping (shared by the kernels, the launchers and the host
// build of the tests).
#pragma once

#include <algorithm>

#include "bssn_common.h"

namespace bssn {

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of tiles so tiles sharing halos share an L2 (a
// permutation of [0, n) for 0 <= i < n).
BSSN_HD __forceinline__ int xcd_remap(int i, int n) {
  constexpr int X = 8;
  const int per = n / X, rem = n % X;
  const int x = i % X, l = i / X;
  return x < rem ? x * (per + 1) + l : rem * (per + 1) + (x - rem) * per + l;
}

// A tile whose box meets a closure point (the three planes next to a
// Sommerfeld face) needs the closure-capable kernel.  Along each axis those
// tiles form at most two index ranges, [l0, l1) at the low face and
// [r0, r1) at the high face (l1 <= r0).  The tiles without closure points
// are then the product of the three per-axis complements, and the closure
// tiles the rest: the x slab (bx in the ranges), the y slab (bx outside,
// by inside) and the z slab (bx, by outside, bz inside).
//
// The closure tiles split further by the axes their closures lie along:
// kinds 3, 4, 5 the tiles with closures along x, y, z only (one slab's
// faces), kind 6 the tiles with closures along two or three axes (edges and
// corners), enumerated as x-and-y ranges (any z), x-and-z ranges (y
// outside) and y-and-z ranges (x outside).
struct TileSet {
  int n[3];                       // tiles per axis
  int l0[3], l1[3], r0[3], r1[3];  // closure-touching tile ranges per axis
  int kind;                       // 0 every tile, 1 tiles without closures, 2 with, 3..6 as above
  int count;                      // tiles launched (one workgroup each)
};

BSSN_HD __forceinline__ int range_size(const TileSet &S, int a) { return (S.l1[a] - S.l0[a]) + (S.r1[a] - S.r0[a]); }
// j-th tile index inside the closure ranges of axis a
BSSN_HD __forceinline__ int range_pick(const TileSet &S, int a, int j) {
  const int nl = S.l1[a] - S.l0[a];
  return j < nl ? S.l0[a] + j : S.r0[a] + (j - nl);
}
// j-th tile index outside the closure ranges of axis a
BSSN_HD __forceinline__ int compl_pick(const TileSet &S, int a, int j) {
  int i = j < S.l0[a] ? j : j + (S.l1[a] - S.l0[a]);
  return i < S.r0[a] ? i : i + (S.r1[a] - S.r0[a]);
}

// Tile coordinates of the k-th tile of the set, z fastest.
// ``slab`` (optional) receives the closure slab of a face tile: 0 x, 1 y,
// 2 z (0 for the other kinds).
BSSN_HD __forceinline__ void tile_of(const TileSet &S, int k, int &bx, int &by, int &bz, int *slab = nullptr) {
  if (slab) *slab = 0;
  if (S.kind == 0) {
    bz = k % S.n[2];
    by = (k / S.n[2]) % S.n[1];
    bx = k / (S.n[2] * S.n[1]);
    return;
  }
  const int cx = S.n[0] - range_size(S, 0), cy = S.n[1] - range_size(S, 1), cz = S.n[2] - range_size(S, 2);
  if (S.kind == 1) {
    // z fastest, then x, then y: a tile's x neighbours follow one z row
    // later (x is the thin axis of the 4 x 8 x 8 tiles); measured 1.5 %
    // faster than z, y, x and 9 % faster than x fastest
    bz = compl_pick(S, 2, k % cz);
    bx = compl_pick(S, 0, (k / cz) % cx);
    by = compl_pick(S, 1, k / (cz * cx));
    return;
  }
  const int sx = range_size(S, 0), sy = range_size(S, 1), sz = range_size(S, 2);
  if (S.kind >= 3 && S.kind <= 5) {
    const int a = S.kind - 3;  // the closure axis
    if (slab) *slab = a;
    const int ez = a == 2 ? sz : cz, ey = a == 1 ? sy : cy;
    const int jz = k % ez, jy = (k / ez) % ey, jx = k / (ez * ey);
    bz = a == 2 ? range_pick(S, 2, jz) : compl_pick(S, 2, jz);
    by = a == 1 ? range_pick(S, 1, jy) : compl_pick(S, 1, jy);
    bx = a == 0 ? range_pick(S, 0, jx) : compl_pick(S, 0, jx);
    return;
  }
  if (S.kind == 6) {
    const int nxy = sx * sy * S.n[2], nxz = sx * cy * sz;
    if (k < nxy) {
      bz = k % S.n[2];
      by = range_pick(S, 1, (k / S.n[2]) % sy);
      bx = range_pick(S, 0, k / (S.n[2] * sy));
    } else if ((k -= nxy) < nxz) {
      bz = range_pick(S, 2, k % sz);
      by = compl_pick(S, 1, (k / sz) % cy);
      bx = range_pick(S, 0, k / (sz * cy));
    } else {
      k -= nxz;
      if (slab) *slab = 1;
      bz = range_pick(S, 2, k % sz);
      by = range_pick(S, 1, (k / sz) % sy);
      bx = compl_pick(S, 0, k / (sz * sy));
    }
    return;
  }
  const int nx_slab = sx * S.n[1] * S.n[2], ny_slab = cx * sy * S.n[2];
  if (k < nx_slab) {
    bz = k % S.n[2];
    by = (k / S.n[2]) % S.n[1];
    bx = range_pick(S, 0, k / (S.n[2] * S.n[1]));
  } else if ((k -= nx_slab) < ny_slab) {
    if (slab) *slab = 1;
    bz = k % S.n[2];
    by = range_pick(S, 1, (k / S.n[2]) % sy);
    bx = compl_pick(S, 0, k / (S.n[2] * sy));
  } else {
    if (slab) *slab = 2;
    k -= ny_slab;
    bz = range_pick(S, 2, k % sz);
    by = compl_pick(S, 1, (k / sz) % cy);
    bx = compl_pick(S, 0, k / (sz * cy));
  }
}

// Reference predicate of the closure tiles (box of the tile at o with
// extent t meets a closure point).
BSSN_HD inline bool tile_has_closure(const Ctx &C, int ox, int oy, int oz, int tx, int ty, int tz) {
  const int o[3] = {ox, oy, oz};
  const int t[3] = {tx, ty, tz};
  bool touch = false;
  for (int a = 0; a < 3; ++a) {
    const Axis &A = C.ax[a];
    touch |= (A.lbc && o[a] < A.start + 3 && o[a] + t[a] > A.start) ||
             (A.rbc && o[a] + t[a] > A.end - 3 && o[a] < A.end);
  }
  return touch;
}

// The tile set of one launch: kind 0 every tile; 1 / 2 the tiles without /
// with closure points (tiles aligned to the Sommerfeld faces).
inline TileSet tile_set(const Ctx &C, int kind, int tx, int ty, int tz) {
  TileSet S;
  const int t[3] = {tx, ty, tz};
  int total = 1, lean = 1, ns[3], nc[3];
  for (int a = 0; a < 3; ++a) {
    const Axis &A = C.ax[a];
    const int n = (A.n + t[a] - 1) / t[a];
    S.n[a] = n;
    // tiles b with b t < start + 3 and (b + 1) t > start, resp. (b + 1) t > end - 3 and b t < end
    int l0 = 0, l1 = 0, r0 = n, r1 = n;
    if (kind != 0 && A.lbc) {
      l0 = std::max(0, A.start / t[a]);
      l1 = std::min(n, (A.start + 3 + t[a] - 1) / t[a]);
    }
    if (kind != 0 && A.rbc) {
      r0 = std::max(0, (A.end - 3) / t[a]);
      r1 = std::min(n, (A.end + t[a] - 1) / t[a]);
    }
    if (l1 <= l0) l0 = l1 = 0;
    if (r1 <= r0) r0 = r1 = n;
    if (l1 > r0 && l1 > l0 && r1 > r0) {  // the two ranges meet: one range
      l0 = std::min(l0, r0);
      l1 = std::max(l1, r1);
      r0 = r1 = n;
    }
    S.l0[a] = l0, S.l1[a] = l1, S.r0[a] = r0, S.r1[a] = r1;
    total *= n;
    ns[a] = (l1 - l0) + (r1 - r0);
    nc[a] = n - ns[a];
    lean *= nc[a];
  }
  S.kind = kind;
  switch (kind) {
    case 0: S.count = total; break;
    case 1: S.count = lean; break;
    case 2: S.count = total - lean; break;
    case 3: S.count = ns[0] * nc[1] * nc[2]; break;
    case 4: S.count = nc[0] * ns[1] * nc[2]; break;
    case 5: S.count = nc[0] * nc[1] * ns[2]; break;
    default: S.count = ns[0] * ns[1] * S.n[2] + ns[0] * nc[1] * ns[2] + nc[0] * ns[1] * ns[2]; break;
  }
  return S;
}

// Grid points covered by a tile set (tiles clipped to the grid).
inline int64_t tile_set_points(const Ctx &C, const TileSet &S, int tx, int ty, int tz) {
  int64_t pts = 0;
  for (int k = 0; k < S.count; ++k) {
    int bx, by, bz;
    tile_of(S, k, bx, by, bz);
    pts += (int64_t)std::min(tx, C.ax[0].n - bx * tx) * std::min(ty, C.ax[1].n - by * ty) *
           std::min(tz, C.ax[2].n - bz * tz);
  }
  return pts;
}

}  // namespace bssn
