// sqh_k_grad32.hip -- grad tile kernels, float arithmetic.
#include "sqh_launch_common.h"

namespace sqh {
hipError_t launch_tile_grad_float(const TileLaunch& t, hipStream_t st) { SQH_DISPATCH_N(launch_grad_n, float, t.N, t, st) }
}  // namespace sqh
