// sqh_k_forward32.hip -- forward tile kernels, float arithmetic.
#include "sqh_launch_common.h"

namespace sqh {
hipError_t launch_tile_forward_float(const TileLaunch& t, hipStream_t st) { SQH_DISPATCH_N(launch_forward_n, float, t.N, t, st) }
}  // namespace sqh
