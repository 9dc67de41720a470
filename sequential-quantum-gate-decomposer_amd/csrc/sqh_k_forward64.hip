// sqh_k_forward64.hip -- forward tile kernels, double arithmetic.
#include "sqh_launch_common.h"

namespace sqh {
hipError_t launch_tile_forward_double(const TileLaunch& t, hipStream_t st) { SQH_DISPATCH_N(launch_forward_n, double, t.N, t, st) }
}  // namespace sqh
