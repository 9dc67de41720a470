// sqh_k_deriv.hip -- derivative-matrix tile kernels.
#include "sqh_launch_common.h"

namespace sqh {
hipError_t launch_tile_deriv_double(const TileLaunch& t, hipStream_t st) { SQH_DISPATCH_N(launch_deriv_n, double, t.N, t, st) }
hipError_t launch_tile_deriv_float(const TileLaunch& t, hipStream_t st) { SQH_DISPATCH_N(launch_deriv_n, float, t.N, t, st) }
}  // namespace sqh
