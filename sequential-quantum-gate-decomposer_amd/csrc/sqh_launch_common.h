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

// A wavefront holds one batch element when the padded column count is a multiple
// of its column count (only possible above 4 qubits).
template <int N>
inline bool uniform_waves(const TileLaunch& t) {
    return N > 4 && t.m_pad % Geo<N>::CW == 0;
}

// The uniform path groups 4 (else 2) wavefronts of one batch element per workgroup
// when the element spans a multiple of that many wavefronts and its kernel table
// fits this LDS budget.
constexpr size_t kKernelTableLdsBytes = 32 * 1024;

template <int N, typename R, int MODE, bool UNI, int WPG>
hipError_t launch_forward_nmu(const TileLaunch& t, hipStream_t st, unsigned waves) {
    const unsigned grid = (waves + WPG - 1) / WPG;
    const size_t shm = sizeof(cpx<R>) * ((size_t)WPG * Geo<N>::LDS + (WPG > 1 ? 4 * (size_t)t.ns : 0));
    hipLaunchKernelGGL((k_tile_forward<N, R, MODE, UNI, WPG>), dim3(grid), dim3(64 * WPG), shm, st,
                       (const cpx<R>*)t.in, t.in_bstride, t.in_stride, (cpx<R>*)t.out, t.out_bstride, t.out_stride,
                       t.m, t.m_pad, t.B, t.ops, t.nops, t.lays, (const cpx<R>*)t.kern, t.kern_bstride,
                       t.trace_off, t.diag, t.ns, t.ts, t.zslot, t.in_index);
    return hipGetLastError();
}

// Largest dynamic LDS a workgroup may use on the current device (cached per device).
inline size_t max_lds_per_group() {
    static size_t cache[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 64 * 1024;
    if (!cache[dev]) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
            v = 64 * 1024;
        cache[dev] = (size_t)v;
    }
    return cache[dev];
}

template <int N, typename R, int WPG>
inline bool group_fits(const TileLaunch& t) {
    const size_t shm = sizeof(cpx<R>) * ((size_t)WPG * Geo<N>::LDS + 4 * (size_t)t.ns);
    const int wpb = t.m_pad / Geo<N>::CW;
    return t.ns > 0 && wpb % WPG == 0 && sizeof(cpx<R>) * 4 * (size_t)t.ns <= kKernelTableLdsBytes &&
           shm <= max_lds_per_group();
}

template <int N, typename R, int MODE>
hipError_t launch_forward_nm(const TileLaunch& t, hipStream_t st, unsigned waves) {
    if constexpr (N > 4) {
        if (uniform_waves<N>(t) && t.wpg >= 0) {
            if ((t.wpg == 0 || t.wpg == 4) && group_fits<N, R, 4>(t))
                return launch_forward_nmu<N, R, MODE, true, 4>(t, st, waves);
            if ((t.wpg == 0 || t.wpg >= 2) && group_fits<N, R, 2>(t))
                return launch_forward_nmu<N, R, MODE, true, 2>(t, st, waves);
            return launch_forward_nmu<N, R, MODE, true, 1>(t, st, waves);
        }
    }
    return launch_forward_nmu<N, R, MODE, false, 1>(t, st, waves);
}

template <int N, typename R>
hipError_t launch_forward_n(const TileLaunch& t, hipStream_t st) {
    using G = Geo<N>;
    const long long items = (long long)t.B * t.m_pad;
    const unsigned waves = blocks_for(items, G::CW);
    return t.mode == 0 ? launch_forward_nm<N, R, 0>(t, st, waves) : launch_forward_nm<N, R, 1>(t, st, waves);
}

// Whether the forward pass of t would group its wavefronts WPG to a workgroup (one
// element per wavefront, the group's kernel table and LDS fit).
template <int N, typename R>
inline int forward_group(const TileLaunch& t, int wpg) {
    if constexpr (N > 4) {
        if (!uniform_waves<N>(t) || wpg < 0) return -1;
        if ((wpg == 0 || wpg == 4) && group_fits<N, R, 4>(t)) return 4;
        if ((wpg == 0 || wpg >= 2) && group_fits<N, R, 2>(t)) return 2;
        return 1;
    }
    return -1;
}

// LDS bytes of a staged program (kernel table, layouts, padded ops).
template <typename R>
inline size_t stage_bytes(const TileLaunch& t) {
    return sizeof(cpx<R>) * 4 * (size_t)t.ns + sizeof(Layout) * (size_t)t.stage_lays +
           sizeof(Op) * ((size_t)t.stage_ops + 2 * kOpPad);
}
constexpr size_t kStageLdsBytes = 48 * 1024;

template <int N, typename R, bool UNI, bool STAGE>
hipError_t launch_grad_nus(const TileLaunch& t, hipStream_t st, unsigned grid) {
    const size_t shm = STAGE ? stage_bytes<R>(t) : 0;
    hipLaunchKernelGGL((k_tile_grad<N, R, UNI, STAGE>), dim3(grid), dim3(64), shm, st, (const cpx<R>*)t.in,
                       t.in_stride, t.m, t.m_pad, t.B, t.ops, t.nops, t.lays, (const cpx<R>*)t.kern, t.kern_bstride,
                       t.trace_off, t.ns, t.diag, t.mpart, t.bi, t.stage_ops, t.stage_lays, t.ts);
    return hipGetLastError();
}

template <int N, typename R, bool UNI>
hipError_t launch_grad_nu(const TileLaunch& t, hipStream_t st, unsigned grid) {
    if constexpr (UNI) {
        if (t.stage != 0 && t.stage_ops > 0 && stage_bytes<R>(t) <= kStageLdsBytes)
            return launch_grad_nus<N, R, true, true>(t, st, grid);
    }
    return launch_grad_nus<N, R, UNI, false>(t, st, grid);
}

template <int N, typename R>
hipError_t launch_grad_n(const TileLaunch& t, hipStream_t st) {
    using G = Geo<N>;
    const long long items = (long long)t.B * t.m_pad;
    const unsigned grid = blocks_for(items, G::CW);
    // Per-run circuits need one element per wavefront (the host pads columns to a
    // whole wavefront for them, also at 4 qubits and below).
    if (uniform_waves<N>(t) || (t.bi.run_prog && t.m_pad % G::CW == 0)) return launch_grad_nu<N, R, true>(t, st, grid);
    if (t.bi.run_prog) return hipErrorInvalidValue;
    return launch_grad_nu<N, R, false>(t, st, grid);
}


}  // namespace sqh
