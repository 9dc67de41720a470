// sqh_k_misc.hip -- precision dispatch of the tile kernels and launchers of the
// small kernels (kernel tables, finish, streaming gates, copies).
#define SQH_DEFINE_MISC_KERNELS 1
#include "sqh_launch_common.h"

namespace sqh {

hipError_t launch_tile_forward_double(const TileLaunch& t, hipStream_t st);
hipError_t launch_tile_forward_float(const TileLaunch& t, hipStream_t st);
hipError_t launch_tile_grad_double(const TileLaunch& t, hipStream_t st);
hipError_t launch_tile_grad_float(const TileLaunch& t, hipStream_t st);

// Wavefronts per workgroup the forward pass of t runs with under option wpg (-1: the
// packed layout, several elements per wavefront).
int tile_forward_group(const TileLaunch& t, int wpg) {
#define SQH_FG_N(n) \
    case n: return t.f32 ? forward_group<n, float>(t, wpg) : forward_group<n, double>(t, wpg);
    switch (t.N) {
        SQH_FG_N(1) SQH_FG_N(2) SQH_FG_N(3) SQH_FG_N(4) SQH_FG_N(5) SQH_FG_N(6) SQH_FG_N(7) SQH_FG_N(8) SQH_FG_N(9)
        SQH_FG_N(10)
    default: return -1;
    }
#undef SQH_FG_N
}

bool tile_grad_can_stage(const TileLaunch& t) {
    const int cw = tile_columns_per_wave(t.N);
    if (t.N <= 4 || t.m_pad % cw != 0 || t.stage_ops <= 0) return false;
    return (t.f32 ? stage_bytes<float>(t) : stage_bytes<double>(t)) <= kStageLdsBytes;
}

hipError_t launch_tile_forward(const TileLaunch& t, hipStream_t st) {
    return t.f32 ? launch_tile_forward_float(t, st) : launch_tile_forward_double(t, st);
}
hipError_t launch_tile_grad(const TileLaunch& t, hipStream_t st) {
    return t.f32 ? launch_tile_grad_float(t, st) : launch_tile_grad_double(t, st);
}
size_t device_max_lds() { return max_lds_per_group(); }

int tile_columns_per_wave(int N) {
    if (N < 1 || N > kMaxTileQubits) return 0;
    const int K = N < 4 ? N : 4;
    return 64 >> (N - K);
}

hipError_t launch_build_kernels(bool f32, const double* params, int P, int B, const Slot* slots, int ns, void* kern,
                                hipStream_t st, BatchIndex bi) {
    const long long n = (long long)B * ns;
    if (n == 0) return hipSuccess;
    if (f32)
        hipLaunchKernelGGL(k_build_kernels<float>, dim3(blocks_for(n, 256)), dim3(256), 0, st, params, P, B, slots,
                           ns, (cpx<float>*)kern, bi);
    else
        hipLaunchKernelGGL(k_build_kernels<double>, dim3(blocks_for(n, 256)), dim3(256), 0, st, params, P, B, slots,
                           ns, (cpx<double>*)kern, bi);
    return hipGetLastError();
}

hipError_t launch_deriv_tables(const void* base, const Slot* slots, int ns, const double* params, int P,
                               void* tables, int* zslot, hipStream_t st) {
    if (P <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_deriv_tables, dim3(P), dim3(64), 0, st, (const cpx<double>*)base, slots, ns, params, P,
                       (cpx<double>*)tables, zslot);
    return hipGetLastError();
}

hipError_t launch_finish(const double* diag, int m, int B, int variant, double* costs, const double* mpart, int wpb,
                         int ns, const Slot* slots, const double* params, int P, double* grad, hipStream_t st,
                         BatchIndex bi, const TraceSpec& ts, double* seed_out) {
    hipLaunchKernelGGL(k_finish, dim3(B), dim3(256), 0, st, diag, m, variant, costs, mpart, wpb, ns, slots, params, P,
                       grad, bi, ts, seed_out);
    return hipGetLastError();
}

hipError_t launch_stream_gate(bool f32, void* mats, long long bstride, int stride, int rows, int m, int B, int kind,
                              int target, int control, const void* kern, long long kern_bstride, int slot,
                              hipStream_t st) {
    const long long n = (long long)(rows >> 1) * m * B;
    if (n == 0) return hipSuccess;
    if (f32)
        hipLaunchKernelGGL(k_stream_gate<float>, dim3(blocks_for(n, 256)), dim3(256), 0, st, (cpx<float>*)mats,
                           bstride, stride, rows, m, B, kind, target, control, (const cpx<float>*)kern, kern_bstride,
                           slot);
    else
        hipLaunchKernelGGL(k_stream_gate<double>, dim3(blocks_for(n, 256)), dim3(256), 0, st, (cpx<double>*)mats,
                           bstride, stride, rows, m, B, kind, target, control, (const cpx<double>*)kern,
                           kern_bstride, slot);
    return hipGetLastError();
}

hipError_t launch_extract_diag(bool f32, const void* mats, long long bstride, int stride, int m, int B,
                               int trace_off, double* diag, int nq, const TraceSpec& ts, hipStream_t st) {
    const long long n = (long long)B * m;
    if (f32)
        hipLaunchKernelGGL(k_extract_diag<float>, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                           (const cpx<float>*)mats, bstride, stride, m, B, trace_off, diag, nq, ts);
    else
        hipLaunchKernelGGL(k_extract_diag<double>, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                           (const cpx<double>*)mats, bstride, stride, m, B, trace_off, diag, nq, ts);
    return hipGetLastError();
}

hipError_t launch_broadcast_copy(bool f32, const void* src, int rows, int m, int sstride, void* dst,
                                 long long dbstride, int B, hipStream_t st) {
    const long long n = (long long)rows * m * B;
    if (f32)
        hipLaunchKernelGGL(k_broadcast_copy<float>, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                           (const cpx<float>*)src, rows, m, sstride, (cpx<float>*)dst, dbstride, B);
    else
        hipLaunchKernelGGL(k_broadcast_copy<double>, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                           (const cpx<double>*)src, rows, m, sstride, (cpx<double>*)dst, dbstride, B);
    return hipGetLastError();
}

hipError_t launch_convert(bool to_f32, const void* src, void* dst, long long nreal, hipStream_t st) {
    if (nreal == 0) return hipSuccess;
    if (to_f32)
        hipLaunchKernelGGL(k_c128_to_c64, dim3(blocks_for(nreal, 256)), dim3(256), 0, st, (const double*)src,
                           (float*)dst, nreal);
    else
        hipLaunchKernelGGL(k_c64_to_c128, dim3(blocks_for(nreal, 256)), dim3(256), 0, st, (const float*)src,
                           (double*)dst, nreal);
    return hipGetLastError();
}

}  // namespace sqh
