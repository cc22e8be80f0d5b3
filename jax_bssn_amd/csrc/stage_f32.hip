// The float instantiation of the fused stage kernels (stage_kernels.h).
#include "stage_kernels.h"

namespace bssn {

int launch_stage_f32(const Ctx &C, const double *us, const StageOut &O, hipStream_t st, int parts) {
  return launch_stage_p<float>(C, us, O, st, parts);
}

int launch_slabs_f32(const Ctx &C, const double *us, const StageOut &O, const Slabs &S, hipStream_t st) {
  return launch_slabs_t<float>(C, us, O, S, st);
}

int launch_region_f32(const Ctx &C, const double *us, const StageOut &O, const int *lo, const int *hi,
                      hipStream_t st) {
  return launch_region_t<float>(C, us, O, lo, hi, st);
}

}  // namespace bssn
