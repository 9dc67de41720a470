// sqh_k_grad64.hip -- grad tile kernels, double arithmetic.
#include "sqh_launch_common.h"

namespace sqh {
hipError_t launch_tile_grad_double(const TileLaunch& t, hipStream_t st) { SQH_DISPATCH_N(launch_grad_n, double, t.N, t, st) }
}  // namespace sqh
