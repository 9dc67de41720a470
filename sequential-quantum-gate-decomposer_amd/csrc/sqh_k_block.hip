// sqh_k_block.hip -- dense 2^k x 2^k blocks on k qubits, applied in one HBM pass
// (apply_large_kernel_to_input, apply_large_kernel_to_input.cpp:82-102, 124+: local
// index bit i <-> the i-th smallest involved qubit), and the fusion of gate runs into
// such blocks.
//
// k = 4 runs on the matrix cores: a wavefront computes a 16 x 16 output tile Y = F X
// for 16 columns with 16 MFMAs (four real products, K = 16 in steps of 4):
// v_mfma_f64_16x16x4f64 for complex128, v_mfma_f32_16x16x4f32 (FP32 accumulation) for
// complex64; F stays in registers for the whole wavefront.  Other k use one thread per
// (group, column) with F in LDS.
#include <hip/hip_runtime.h>

#include "sqh_device.h"
#include "sqh_kernels.h"

namespace sqh {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA of one precision.  Operands: lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16]; result item r of lane l is D[row(r, l)][l % 16].
template <typename R>
struct Mfma16;
template <>
struct Mfma16<double> {
    typedef f64x4 V;
    static __device__ __forceinline__ V op(double a, double b, V c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int r, int lane) { return 4 * r + (lane >> 4); }
};
template <>
struct Mfma16<float> {
    typedef f32x4 V;
    static __device__ __forceinline__ V op(float a, float b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int r, int lane) { return 4 * (lane >> 4) + r; }
};

struct BlockQubits {
    int q[5];   // sorted ascending
    int k;
};

// Row of local index i in group g: insert the k involved bits (ascending) into g.
__device__ __forceinline__ long long block_base(long long g, const BlockQubits& bq) {
    for (int j = 0; j < bq.k; ++j) {
        const int q = bq.q[j];
        g = ((g >> q) << (q + 1)) | (g & ((1LL << q) - 1));
    }
    return g;
}

__device__ __forceinline__ int block_off(int i, const BlockQubits& bq) {
    int o = 0;
    for (int j = 0; j < bq.k; ++j)
        if ((i >> j) & 1) o |= 1 << bq.q[j];
    return o;
}

// Generic: thread -> (matrix b, group g, column c).  F: [B_F][D][D] complex (D = 2^k),
// per-matrix when fbstride != 0.
template <typename R, int K>
__global__ __launch_bounds__(256) void k_block_fma(cpx<R>* __restrict__ mats, long long bstride, int stride, int rows,
                                                   int m, int B, const cpx<R>* __restrict__ F, long long fbstride,
                                                   BlockQubits bq) {
    constexpr int D = 1 << K;
    __shared__ cpx<R> Fs[D * D];
    const long long groups = rows >> K;
    const long long per_b = groups * m;
    const long long tid0 = (long long)blockIdx.x * blockDim.x;
    const int b0 = (int)(tid0 / per_b);
    // All threads of a block share one batch element when per_b is a multiple of 256
    // (host guarantees: else fbstride == 0).
    const cpx<R>* Fb = F + (fbstride ? (long long)(b0 < B ? b0 : B - 1) * fbstride : 0);
    for (int i = threadIdx.x; i < D * D; i += blockDim.x) Fs[i] = Fb[i];
    __syncthreads();
    const long long tid = tid0 + threadIdx.x;
    if (tid >= per_b * B) return;
    const int b = (int)(tid / per_b);
    const long long rem = tid % per_b;
    const long long g = rem / m;
    const int col = (int)(rem % m);
    cpx<R>* M = mats + (long long)b * bstride;
    const long long base = block_base(g, bq);
    cpx<R> x[D];
#pragma unroll
    for (int i = 0; i < D; ++i) x[i] = M[(base + block_off(i, bq)) * stride + col];
#pragma unroll
    for (int i = 0; i < D; ++i) {
        cpx<R> y = mk<R>(0, 0);
#pragma unroll
        for (int j = 0; j < D; ++j) cacc(y, Fs[i * D + j], x[j]);
        M[(base + block_off(i, bq)) * stride + col] = y;
    }
}

// k = 4 on the matrix cores.  Wavefront -> (matrix b, group g, T consecutive 16-column
// tiles); the wavefront loads all its tiles before the first MFMA (the stores of one
// tile cannot be reordered before the loads of the next) and reuses F for each.
template <typename R, int T>
__global__ __launch_bounds__(256) void k_block_mfma4(cpx<R>* __restrict__ mats, long long bstride, int stride,
                                                     int rows, int m, int B, const cpx<R>* __restrict__ F,
                                                     long long fbstride, BlockQubits bq) {
    typedef Mfma16<R> MF;
    const int lane = threadIdx.x & 63;
    const long long wave = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const long long groups = rows >> 4;
    const long long tgroups = ((m + 15) / 16 + T - 1) / T;
    const long long per_b = groups * tgroups;
    if (wave >= per_b * B) return;
    const int b = (int)(wave / per_b);
    const long long rem = wave % per_b;
    const long long g = rem / tgroups;
    const int c0 = (int)(rem % tgroups) * 16 * T;
    const cpx<R>* Fb = F + (long long)b * fbstride;
    // A operand of step s: A[i][k] = F[lane%16][4s + lane/16]
    R ar[4], ai[4], ain[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const cpx<R> f = Fb[(lane & 15) * 16 + 4 * s + (lane >> 4)];
        ar[s] = f.re;
        ai[s] = f.im;
        ain[s] = -f.im;
    }
    cpx<R>* M = mats + (long long)b * bstride;
    const long long base = block_base(g, bq);
    long long roff[4];  // rows of the B operand: X[4s + lane/16][.]
#pragma unroll
    for (int s = 0; s < 4; ++s) roff[s] = (base + block_off(4 * s + (lane >> 4), bq)) * stride;
    // B operand of step s for tile t: B[k][n] = X[4s + lane/16][c0 + 16t + lane%16]
    R xr[T][4], xi[T][4];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int col = c0 + 16 * t + (lane & 15);
        const bool cok = col < m;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const cpx<R> v = cok ? M[roff[s] + col] : mk<R>(0, 0);
            xr[t][s] = v.re;
            xi[t][s] = v.im;
        }
    }
#pragma unroll
    for (int t = 0; t < T; ++t) {
        typename MF::V yr = {0, 0, 0, 0}, yi = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            yr = MF::op(ar[s], xr[t][s], yr);
            yr = MF::op(ain[s], xi[t][s], yr);
            yi = MF::op(ar[s], xi[t][s], yi);
            yi = MF::op(ai[s], xr[t][s], yi);
        }
        const int col = c0 + 16 * t + (lane & 15);
        if (col < m) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = MF::row(r, lane);
                M[(base + block_off(i, bq)) * stride + col] = mk<R>(yr[r], yi[r]);
            }
        }
    }
}

// Fused block matrices: the block's gates (FusedGate: local qubits 0..3, the block's
// slots) applied to the unit columns e_j.

// Gates staged per pass and batch elements per workgroup (x 16 block columns = one
// wavefront) of k_fuse_blocks.
constexpr int kFuseChunk = 32;
constexpr int kFuseElems = 4;

// Workgroup -> (block, 16 batch elements); thread -> (element, column j of the block).
// The block's gate records and the elements' 2x2 kernels of those gates are staged in
// LDS a chunk at a time (all loads in flight together), then every thread applies the
// gates to its unit column e_j.
template <typename R>
__global__ __launch_bounds__(16 * kFuseElems) void k_fuse_blocks(const FusedGate* __restrict__ fg,
                                                     const int* __restrict__ block_first, int nblocks,
                                                     const cpx<R>* __restrict__ kern, long long kern_bstride, int Bk,
                                                     cpx<R>* __restrict__ out, const int* __restrict__ count,
                                                     const int* __restrict__ zslot, const double* __restrict__ params,
                                                     int P, const Slot* __restrict__ slots) {
    __shared__ FusedGate srec[kFuseChunk];
    __shared__ cpx<R> skern[kFuseElems][kFuseChunk][4];
    const int blk = blockIdx.x % nblocks;
    const int b0 = (blockIdx.x / nblocks) * kFuseElems;
    if (count) Bk = min(*count, Bk);  // live elements (the optimiser's rounds)
    if (b0 >= Bk) return;             // the whole workgroup
    const int le = threadIdx.x >> 4, j = threadIdx.x & 15;
    // apply_derivate_to: the element's differentiated controlled gate (or -1) runs with
    // its control-0 block zeroed (the derivative of a controlled kernel)
    const int zs = (zslot && b0 + le < Bk) ? zslot[b0 + le] : -1;
    cpx<R> v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = mk<R>(i == j ? (R)1 : (R)0, (R)0);
    const int first = block_first[blk], last = block_first[blk + 1];
    for (int p0 = first; p0 < last; p0 += kFuseChunk) {
        const int cnt = min(kFuseChunk, last - p0);
        __syncthreads();  // the previous chunk is consumed
        for (int i = threadIdx.x; i < cnt; i += blockDim.x) srec[i] = fg[p0 + i];
        __syncthreads();
        if (params) {
            // the 2x2 kernels from the elements' parameters (as k_build_kernels computes them)
            for (int i = threadIdx.x; i < kFuseElems * cnt; i += blockDim.x) {
                const int e = i / cnt, gi = i % cnt;
                const int slot = srec[gi].slot;
                if (slot >= 0 && b0 + e < Bk) {
                    const Slot sl = slots[slot];
                    cpx<double> k4[4];
                    dev_gate_kernel(sl.type, params + (long long)(b0 + e) * P + sl.param_start, k4);
#pragma unroll
                    for (int k = 0; k < 4; ++k) skern[e][gi][k] = cvt<R>(k4[k]);
                }
            }
        } else {
            for (int i = threadIdx.x; i < kFuseElems * cnt * 4; i += blockDim.x) {
                const int e = i / (cnt * 4), gi = (i / 4) % cnt, k = i % 4;
                const int slot = srec[gi].slot;
                if (slot >= 0 && b0 + e < Bk) skern[e][gi][k] = kern[(long long)(b0 + e) * kern_bstride + 4 * slot + k];
            }
        }
        __syncthreads();
        for (int gi = 0; gi < cnt; ++gi) {
            const FusedGate gt = srec[gi];
            Op op;
            op.kind = (uint8_t)gt.kind;
            op.t = (uint8_t)gt.t;
            op.cmode = gt.c >= 0 ? CM_REG : CM_NONE;
            op.c = (uint8_t)(gt.c >= 0 ? gt.c : 0);
            op.idx = gt.slot;
            cpx<R> k[4] = {mk<R>(0, 0), mk<R>(0, 0), mk<R>(0, 0), mk<R>(0, 0)};
            if (gt.slot >= 0) {
#pragma unroll
                for (int e = 0; e < 4; ++e) k[e] = skern[le][gi][e];
            }
            if (gt.slot == zs && op.kind == OP_CU)
                exec_gate<4, true, R>(v, op, k, 0);
            else
                exec_gate<4, false, R>(v, op, k, 0);
        }
    }
    const int b = b0 + le;
    if (b >= Bk) return;
    // out[b][blk][i][j]
    cpx<R>* o = out + ((long long)b * nblocks + blk) * 256;
#pragma unroll
    for (int i = 0; i < 16; ++i) o[i * 16 + j] = v[i];
}

// Launchers -----------------------------------------------------------------------

template <typename R, int K>
static hipError_t launch_fma_k(void* mats, long long bstride, int stride, int rows, int m, int B, const void* F,
                               long long fbstride, const BlockQubits& bq, hipStream_t s) {
    const long long n = (long long)(rows >> K) * m * B;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL((k_block_fma<R, K>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (cpx<R>*)mats,
                       bstride, stride, rows, m, B, (const cpx<R>*)F, fbstride, bq);
    return hipGetLastError();
}

template <typename R>
static hipError_t launch_fma(void* mats, long long bstride, int stride, int rows, int m, int B, const void* F,
                             long long fbstride, const BlockQubits& bq, hipStream_t s) {
    switch (bq.k) {
    case 1: return launch_fma_k<R, 1>(mats, bstride, stride, rows, m, B, F, fbstride, bq, s);
    case 2: return launch_fma_k<R, 2>(mats, bstride, stride, rows, m, B, F, fbstride, bq, s);
    case 3: return launch_fma_k<R, 3>(mats, bstride, stride, rows, m, B, F, fbstride, bq, s);
    case 4: return launch_fma_k<R, 4>(mats, bstride, stride, rows, m, B, F, fbstride, bq, s);
    case 5: return launch_fma_k<R, 5>(mats, bstride, stride, rows, m, B, F, fbstride, bq, s);
    default: return hipErrorInvalidValue;
    }
}

// Dense block on the k sorted qubits of every matrix.  F holds one block, or one per
// matrix (fbstride = 4^k).  use_mfma selects the matrix-core kernel for k = 4.
template <typename R, int T>
static hipError_t launch_mfma4_t(void* mats, long long bstride, int stride, int rows, int m, int B, const void* F,
                                 long long fbstride, const BlockQubits& bq, int threads, hipStream_t s) {
    const long long waves = (long long)(rows >> 4) * (((m + 15) / 16 + T - 1) / T) * B;
    const int wpb = threads / 64;
    hipLaunchKernelGGL((k_block_mfma4<R, T>), dim3((unsigned)((waves + wpb - 1) / wpb)), dim3(threads), 0, s,
                       (cpx<R>*)mats, bstride, stride, rows, m, B, (const cpx<R>*)F, fbstride, bq);
    return hipGetLastError();
}

template <typename R>
static hipError_t launch_mfma4(void* mats, long long bstride, int stride, int rows, int m, int B, const void* F,
                               long long fbstride, const BlockQubits& bq, BlockCfg cfg, hipStream_t s) {
    switch (cfg.tiles) {
    case 1: return launch_mfma4_t<R, 1>(mats, bstride, stride, rows, m, B, F, fbstride, bq, cfg.threads, s);
    case 2: return launch_mfma4_t<R, 2>(mats, bstride, stride, rows, m, B, F, fbstride, bq, cfg.threads, s);
    case 4: return launch_mfma4_t<R, 4>(mats, bstride, stride, rows, m, B, F, fbstride, bq, cfg.threads, s);
    default: return hipErrorInvalidValue;
    }
}

bool block_cfg_valid(BlockCfg cfg) {
    return (cfg.tiles == 1 || cfg.tiles == 2 || cfg.tiles == 4) && (cfg.threads == 64 || cfg.threads == 256);
}

hipError_t launch_block_apply(bool f32, bool use_mfma, void* mats, long long bstride, int stride, int rows, int m,
                              int B, const void* F, long long fbstride, const int* qubits, int k, BlockCfg cfg,
                              hipStream_t s) {
    if (k < 1 || k > 5 || (1 << k) > rows) return hipErrorInvalidValue;
    BlockQubits bq;
    bq.k = k;
    for (int j = 0; j < 5; ++j) bq.q[j] = j < k ? qubits[j] : 0;
    if (use_mfma && k == 4) {
        if (!block_cfg_valid(cfg)) return hipErrorInvalidValue;
        return f32 ? launch_mfma4<float>(mats, bstride, stride, rows, m, B, F, fbstride, bq, cfg, s)
                   : launch_mfma4<double>(mats, bstride, stride, rows, m, B, F, fbstride, bq, cfg, s);
    }
    // The FMA kernel stages one block per workgroup: per-matrix blocks need each
    // workgroup inside one matrix.
    if (fbstride != 0 && (((long long)(rows >> k) * m) % 256) != 0) return hipErrorInvalidValue;
    return f32 ? launch_fma<float>(mats, bstride, stride, rows, m, B, F, fbstride, bq, s)
               : launch_fma<double>(mats, bstride, stride, rows, m, B, F, fbstride, bq, s);
}

hipError_t launch_fuse_blocks(bool f32, const void* fg, const int* block_first, int nblocks, const void* kern,
                              long long kern_bstride, int Bk, void* out, hipStream_t s, const int* count,
                              const int* zslot, const double* params, int P, const Slot* slots) {
    if (params && !slots) return hipErrorInvalidValue;
    if ((long long)Bk * nblocks == 0) return hipSuccess;
    const long long grid = (long long)nblocks * ((Bk + kFuseElems - 1) / kFuseElems);
    const int threads = 16 * kFuseElems;
    if (f32)
        hipLaunchKernelGGL(k_fuse_blocks<float>, dim3((unsigned)grid), dim3(threads), 0, s, (const FusedGate*)fg,
                           block_first, nblocks, (const cpx<float>*)kern, kern_bstride, Bk, (cpx<float>*)out, count, zslot, params, P, slots);
    else
        hipLaunchKernelGGL(k_fuse_blocks<double>, dim3((unsigned)grid), dim3(threads), 0, s, (const FusedGate*)fg,
                           block_first, nblocks, (const cpx<double>*)kern, kern_bstride, Bk, (cpx<double>*)out, count, zslot, params, P, slots);
    return hipGetLastError();
}

size_t fused_gate_bytes() { return sizeof(FusedGate); }

}  // namespace sqh
