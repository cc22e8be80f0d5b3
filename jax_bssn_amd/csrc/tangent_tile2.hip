// Tangent tile kernels carrying 2 directions (instantiation unit).
#include "tangent_common.h"

namespace bssn {

int launch_tangent_tiles_2(const Ctx &C, const double *us, const double *dus, const StageOut &O, const TangentOut &T,
                           hipStream_t st) {
  return launch_tangent_tiles_t<2>(C, us, dus, O, T, st);
}

}  // namespace bssn
