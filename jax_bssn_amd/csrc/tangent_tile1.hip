// Tangent tile kernels carrying 1 direction (instantiation unit).
#include "tangent_common.h"

namespace bssn {

int launch_tangent_tiles_1(const Ctx &C, const double *us, const double *dus, const StageOut &O, const TangentOut &T,
                           hipStream_t st) {
  return launch_tangent_tiles_t<1>(C, us, dus, O, T, st);
}

}  // namespace bssn
