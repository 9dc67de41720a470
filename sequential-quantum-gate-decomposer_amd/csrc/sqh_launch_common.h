// sqh_launch_common.h -- launch helpers shared by the sqh_k_*.hip translation units.
#pragma once
#include "sqh_device.h"
#include "sqh_kernels.h"

namespace sqh {

static inline unsigned blocks_for(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

#define SQH_DISPATCH_N(FN, R, N, ...)         \
    switch (N) {                              \
    case 1: return FN<1, R>(__VA_ARGS__);     \
    case 2: return FN<2, R>(__VA_ARGS__);     \
    case 3: return FN<3, R>(__VA_ARGS__);     \
    case 4: return FN<4, R>(__VA_ARGS__);     \
    case 5: return FN<5, R>(__VA_ARGS__);     \
    case 6: return FN<6, R>(__VA_ARGS__);     \
    case 7: return FN<7, R>(__VA_ARGS__);     \
    case 8: return FN<8, R>(__VA_ARGS__);     \
    case 9: return FN<9, R>(__VA_ARGS__);     \
    case 10: return FN<10, R>(__VA_ARGS__);   \
    default: return hipErrorInvalidValue;     \
    }

template <int N, typename R>
hipError_t launch_forward_n(const TileLaunch& t, hipStream_t st) {
    using G = Geo<N>;
    const long long items = (long long)t.B * t.m_pad;
    const unsigned grid = blocks_for(items, G::CW);
    if (t.mode == 0)
        hipLaunchKernelGGL((k_tile_forward<N, R, 0>), dim3(grid), dim3(64), 0, st, (const cpx<R>*)t.in,
                           t.in_bstride, t.in_stride, (cpx<R>*)t.out, t.out_bstride, t.out_stride, t.m, t.m_pad,
                           t.B, t.ops, t.nops, t.lays, (const cpx<R>*)t.kern, t.kern_bstride, t.trace_off,
                           t.diag);
    else
        hipLaunchKernelGGL((k_tile_forward<N, R, 1>), dim3(grid), dim3(64), 0, st, (const cpx<R>*)t.in,
                           t.in_bstride, t.in_stride, (cpx<R>*)t.out, t.out_bstride, t.out_stride, t.m, t.m_pad,
                           t.B, t.ops, t.nops, t.lays, (const cpx<R>*)t.kern, t.kern_bstride, t.trace_off,
                           t.diag);
    return hipGetLastError();
}

template <int N, typename R>
hipError_t launch_grad_n(const TileLaunch& t, hipStream_t st) {
    using G = Geo<N>;
    const long long items = (long long)t.B * t.m_pad;
    const unsigned grid = blocks_for(items, G::CW);
    hipLaunchKernelGGL((k_tile_grad<N, R>), dim3(grid), dim3(64), 0, st, (const cpx<R>*)t.in, t.in_stride, t.m,
                       t.m_pad, t.B, t.ops, t.nops, t.lays, (const cpx<R>*)t.kern, t.kern_bstride, t.trace_off,
                       t.ns, t.diag, t.mpart);
    return hipGetLastError();
}

template <int N, typename R>
hipError_t launch_deriv_n(const TileLaunch& t, hipStream_t st) {
    using G = Geo<N>;
    const int wpb = t.m_pad >= G::CW ? t.m_pad / G::CW : 1;
    const unsigned grid = (unsigned)(wpb * t.ns);
    hipLaunchKernelGGL((k_tile_deriv<N, R>), dim3(grid), dim3(64), 0, st, (const cpx<R>*)t.in, t.in_stride, t.m,
                       t.m_pad, t.ops, t.nops, t.lays, (const cpx<R>*)t.kern, t.slots, t.params,
                       (cpx<R>*)t.out, t.out_bstride);
    return hipGetLastError();
}

}  // namespace sqh
