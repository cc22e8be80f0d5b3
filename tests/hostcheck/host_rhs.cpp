// Host build of the per-point stage code (PointProvider path) for the CPU
// tests: the same rhs_point.h / stencil rows / projection / emitter the
// kernels run, looped over every point on the CPU.  Test infrastructure only;
// the library never calls it.
#include <hip/hip_runtime.h>

#include "stage_common.h"

using namespace bssn;

namespace {

int run(const bssn_params *p, const bssn_grid *g, int mode, const double *us, const double *u0, double *ustar,
        double *unext, double *rhs, int reproject, char *err, int errlen) {
  Ctx C;
  const int rc = resolve(p, g, &C, err, (size_t)errlen);
  if (rc) return rc;
  StageOut O{mode == 1 ? us : u0, ustar, unext, rhs, mode, reproject};
  for (int x = 0; x < C.ax[0].n; ++x)
    for (int y = 0; y < C.ax[1].n; ++y)
      for (int z = 0; z < C.ax[2].n; ++z) {
        PointProvider<double> pv;
        pv.setup(C, us, x, y, z);
        double u[NCOMP];
        for (int c = 0; c < NCOMP; ++c) u[c] = us[c * C.ncell + pv.idx];
        Emitter<RegStash> out(C, O, pv.idx, true, pv.face(), pv.st);
        bssn_rhs_point<double>(pv, C, u, out);
      }
  return 0;
}

}  // namespace

extern "C" int host_rhs(const bssn_params *p, const bssn_grid *g, const double *state, double *rhs, char *err,
                        int errlen) {
  return run(p, g, 0, state, nullptr, nullptr, nullptr, rhs, 0, err, errlen);
}

extern "C" int host_stage(const bssn_params *p, const bssn_grid *g, int stage, const double *us, const double *u0,
                          double *ustar, double *unext, int reproject, char *err, int errlen) {
  return run(p, g, stage, us, u0, ustar, unext, nullptr, reproject, err, errlen);
}
