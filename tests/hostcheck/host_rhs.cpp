// Host build of the per-point stage code (PointProvider path) for the CPU
// tests: the same rhs_point.h / stencil rows / projection / emitter the
// kernels run, looped over every point on the CPU.  Test infrastructure only;
// the library never calls it.
#include <hip/hip_runtime.h>

#include <vector>

#include "stage_common.h"
#include "tiles.h"

using namespace bssn;

namespace {

// Shares the provider's slots the way the kernels' LDS stash does.
struct SharedStash {
  RegStash *p;
  void put(int slot, double x) { p->put(slot, x); }
  double get(int slot) const { return p->get(slot); }
};

template <typename T>
int run(const bssn_params *p, const bssn_grid *g, int mode, const double *us, const double *u0, double *ustar,
        double *unext, double *rhs, int reproject, const double *matter, char *err, int errlen) {
  Ctx C;
  const int rc = resolve(p, g, &C, err, (size_t)errlen);
  if (rc) return rc;
  StageOut O{mode == 1 ? us : u0, ustar, unext, rhs, mode, reproject, nullptr, matter};
  stage_update(O, C.dt);
  std::vector<double> D, kap;
  if (C.kappa != 0.0) {
    D.resize((size_t)NDERIVED * C.ncell);
    kap.resize((size_t)6 * C.ncell);
    for (int64_t i = 0; i < C.ncell; ++i) derived_point(C.ncell, i, us, D.data());
    for (int64_t i = 0; i < C.ncell; ++i) kappa_point(C, us, D.data(), matter, i, kap.data());
    O.kap = kap.data();
  }
  for (int x = 0; x < C.ax[0].n; ++x)
    for (int y = 0; y < C.ax[1].n; ++y)
      for (int z = 0; z < C.ax[2].n; ++z) {
        PointProvider<T> pv;
        pv.setup(C, us, x, y, z);
        double u[NCOMP];
        for (int c = 0; c < NCOMP; ++c) u[c] = us[c * C.ncell + pv.idx];
        Emitter<SharedStash> out(C, O, pv.idx, true, pv.face(), SharedStash{&pv.st});
        bssn_rhs_point<T>(pv, C, u, out);
      }
  return 0;
}

}  // namespace

extern "C" int host_rhs(const bssn_params *p, const bssn_grid *g, const double *state, const double *matter,
                        double *rhs, char *err, int errlen) {
  return run<double>(p, g, 0, state, nullptr, nullptr, nullptr, rhs, 0, matter, err, errlen);
}

extern "C" int host_rhs_f32(const bssn_params *p, const bssn_grid *g, const double *state, double *rhs, char *err,
                            int errlen) {
  return run<float>(p, g, 0, state, nullptr, nullptr, nullptr, rhs, 0, nullptr, err, errlen);
}

extern "C" int host_stage(const bssn_params *p, const bssn_grid *g, int stage, const double *us, const double *u0,
                          double *ustar, double *unext, int reproject, char *err, int errlen) {
  return run<double>(p, g, stage, us, u0, ustar, unext, nullptr, reproject, nullptr, err, errlen);
}

// Tile sets against the brute-force predicate: every tile of kind 1 (2)
// enumerated exactly once iff it has no (has) closure points, the point
// count of the two sets adds up to the grid, and kinds 3..6 enumerate the
// closure tiles once each, split by the axes of their closures.  Returns the number of
// violations (or a negative status).
// The XCD remapping of workgroup indexes: count the n in [1, nmax] for
// which it is not a permutation of [0, n).
extern "C" int host_xcd_remap_check(int nmax) {
  int bad = 0;
  for (int n = 1; n <= nmax; ++n) {
    std::vector<int> seen((size_t)n, 0);
    bool ok = true;
    for (int i = 0; i < n; ++i) {
      const int t = xcd_remap(i, n);
      if (t < 0 || t >= n || seen[(size_t)t]++) ok = false;
    }
    bad += ok ? 0 : 1;
  }
  return bad;
}

extern "C" int host_tile_check(const bssn_params *p, const bssn_grid *g, int tx, int ty, int tz, char *err,
                               int errlen) {
  Ctx C;
  const int rc = resolve(p, g, &C, err, (size_t)errlen);
  if (rc) return rc;
  int bad = 0;
  int64_t pts = 0;
  for (int kind = 1; kind <= 2; ++kind) {
    const TileSet S = tile_set(C, kind, tx, ty, tz);
    std::vector<int> seen((size_t)S.n[0] * S.n[1] * S.n[2], 0);
    for (int k = 0; k < S.count; ++k) {
      int bx, by, bz;
      tile_of(S, k, bx, by, bz);
      if (bx < 0 || bx >= S.n[0] || by < 0 || by >= S.n[1] || bz < 0 || bz >= S.n[2]) {
        ++bad;
        continue;
      }
      seen[((size_t)bx * S.n[1] + by) * S.n[2] + bz] += 1;
    }
    for (int bx = 0; bx < S.n[0]; ++bx)
      for (int by = 0; by < S.n[1]; ++by)
        for (int bz = 0; bz < S.n[2]; ++bz) {
          const bool want = tile_has_closure(C, bx * tx, by * ty, bz * tz, tx, ty, tz) == (kind == 2);
          if (seen[((size_t)bx * S.n[1] + by) * S.n[2] + bz] != (want ? 1 : 0)) ++bad;
        }
    pts += tile_set_points(C, S, tx, ty, tz);
  }
  // kinds 3..6 partition the closure tiles by the axes their closures lie
  // along: 3, 4, 5 exactly x, y, z; 6 two or three axes
  const int t[3] = {tx, ty, tz};
  auto axes_of = [&](int bx, int by, int bz) {
    const int o[3] = {bx * tx, by * ty, bz * tz};
    int mask = 0;
    for (int a = 0; a < 3; ++a) {
      const Axis &A = C.ax[a];
      if ((A.lbc && o[a] < A.start + 3 && o[a] + t[a] > A.start) || (A.rbc && o[a] + t[a] > A.end - 3 && o[a] < A.end))
        mask |= 1 << a;
    }
    return mask;
  };
  const TileSet S2 = tile_set(C, 2, tx, ty, tz);
  std::vector<int> seen((size_t)S2.n[0] * S2.n[1] * S2.n[2], 0);
  int count = 0;
  for (int kind = 3; kind <= 6; ++kind) {
    const TileSet S = tile_set(C, kind, tx, ty, tz);
    count += S.count;
    for (int k = 0; k < S.count; ++k) {
      int bx, by, bz, slab;
      tile_of(S, k, bx, by, bz, &slab);
      if (bx < 0 || bx >= S.n[0] || by < 0 || by >= S.n[1] || bz < 0 || bz >= S.n[2]) {
        ++bad;
        continue;
      }
      seen[((size_t)bx * S.n[1] + by) * S.n[2] + bz] += 1;
      const int m = axes_of(bx, by, bz);
      const bool ok = kind == 6 ? (m != 0 && (m & (m - 1)) != 0) : (m == 1 << (kind - 3) && slab == kind - 3);
      if (!ok) ++bad;
    }
  }
  if (count != S2.count) ++bad;
  for (int bx = 0; bx < S2.n[0]; ++bx)
    for (int by = 0; by < S2.n[1]; ++by)
      for (int bz = 0; bz < S2.n[2]; ++bz)
        if (seen[((size_t)bx * S2.n[1] + by) * S2.n[2] + bz] != (axes_of(bx, by, bz) ? 1 : 0)) ++bad;
  return bad + (pts != C.ncell ? 1 : 0);
}

