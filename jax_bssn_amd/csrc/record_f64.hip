// The FP64 stage kernels in recording mode (stage_kernels.h, Recorder):
// the features of every point for the adjoint (adjoint.hip), through the
// same tile, face-tile and closure-shell kernels as the evolution.
#include "stage_kernels.h"

namespace bssn {

int launch_record_f64(const Ctx &C, const double *us, double *feat, hipStream_t st) {
  StageOut O{};
  O.rhs = feat;
  O.mode = 0;
  return launch_stage_p<double, true>(C, us, O, st, PART_TILES | PART_SHELL);
}

}  // namespace bssn
