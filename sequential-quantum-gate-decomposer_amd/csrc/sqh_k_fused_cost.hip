// sqh_k_fused_cost.hip -- batched cost of the circuit run as fused dense 4-qubit blocks
// on the matrix cores, with every column of the target resident on the chip from its
// first load to its trace entry.
//
// The circuit is the product of nb dense 16 x 16 blocks on sorted 4-qubit sets (the
// fusion plan; per batch element the block matrices F_b are built from its kernel
// tables).  A wavefront owns 64 "vector columns": W = 64 / O columns of the target,
// each split into O = 2^N / 16 outer groups of the 16 rows its block touches.  Per
// block it computes Y = F_b X for its 64 vector columns as four 16 x 16 output tiles,
// 16 MFMAs each (four real products, K = 16 in four steps):
// v_mfma_f64_16x16x4f64 for complex128, v_mfma_f32_16x16x4f32 (FP32 accumulation) for
// complex64.  Between blocks the tiles go through the wavefront's own LDS region in
// natural row order and come back in the next block's grouping; after the last block
// the lane holding row c + trace_off of column c writes that trace entry.  The target
// is read once per wavefront (it is shared by the batch and stays in L2); nothing
// else touches HBM but the block matrices and the per-column trace entries.
//
// Row j of a column under block b: j = base_b[o] + off_b[i] for local index i (bit q of
// i <-> the q-th smallest qubit of the block, as apply_large_kernel_to_input,
// apply_large_kernel_to_input.cpp:82-102) and outer index o (the other qubits in
// ascending order); both tables come from the host per block.
#include <hip/hip_runtime.h>

#include "sqh_device.h"
#include "sqh_kernels.h"

namespace sqh {

typedef double fc_f64x4 __attribute__((ext_vector_type(4)));
typedef float fc_f32x4 __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]; item r of
// lane l's result is D[row(r, l)][l % 16] (the two precisions order rows differently).
template <typename R>
struct FcMfma;
template <>
struct FcMfma<double> {
    typedef fc_f64x4 V;
    static __device__ __forceinline__ V op(double a, double b, V c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int r, int g) { return 4 * r + g; }
};
template <>
struct FcMfma<float> {
    typedef fc_f32x4 V;
    static __device__ __forceinline__ V op(float a, float b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int r, int g) { return 4 * g + r; }
};

// LDS position of row j of a column: the row's low nibble XOR its higher bits folded
// down, so the 16 lanes of a lane group that differ in the outer index land in
// distinct 16-element slots for most qubit sets.
template <int N>
__device__ __forceinline__ int fc_slot(int j) {
    if constexpr (N <= 4) {
        return j;
    } else {
        int h = j >> 4;
        if constexpr (N > 8) h ^= h >> 4;
        return j ^ (h & 15);
    }
}

// Wave-scope ordering of the LDS exchange: the stores of a remap are issued before its
// loads (LDS serves one wavefront's requests in order; this keeps the compiler from
// moving the loads up).
__device__ __forceinline__ void fc_wave_fence() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); }

// Wavefronts per workgroup: two keep the complex128 LDS regions (16 KiB each) and the
// row tables inside the 64 KiB a workgroup may allocate.
constexpr int kFcWaves = 2;

template <typename R, int N>
__global__ __launch_bounds__(64 * kFcWaves) void k_fused_cost(FusedCostArgs a) {
    typedef FcMfma<R> MF;
    constexpr int D = 1 << N;
    constexpr int O = D >> 4;  // outer groups per column
    constexpr int W = 64 / O;  // columns per wavefront
    extern __shared__ __align__(16) unsigned char fc_smem[];
    int* smap = reinterpret_cast<int*>(fc_smem);
    const int nmap = a.nb * kFusedMapStride;
    cpx<R>* region = reinterpret_cast<cpx<R>*>(fc_smem + (((size_t)nmap * sizeof(int) + 15) & ~(size_t)15));
    for (int i = threadIdx.x; i < nmap; i += blockDim.x) smap[i] = a.maps[i];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const long long wave = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (wave >= (long long)a.B * a.groups) return;
    const int e = (int)(wave / a.groups);
    const int c0 = (int)(wave % a.groups) * W;
    cpx<R>* lds = region + (threadIdx.x >> 6) * 1024;
    const int nl = lane & 15, g = lane >> 4;
    // vector column 16 t + nl of tile t: column c0 + w[t], outer group o[t]
    int w[4], o[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int n = 16 * t + nl;
        w[t] = n / O;
        o[t] = n % O;
    }
    const cpx<R>* U = static_cast<const cpx<R>*>(a.U);
    const int m = a.m;
    // B operand of step s, tile t: X[4 s + g][16 t + nl] (block 0's grouping, from U)
    R xr[4][4], xi[4][4];
    {
        const int* mp = smap;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int col = c0 + w[t];
            const int base = mp[16 + o[t]];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int j = base + mp[4 * s + g];
                const cpx<R> v = col < m ? U[(long long)j * m + col] : mk<R>(0, 0);
                xr[t][s] = v.re;
                xi[t][s] = v.im;
            }
        }
    }
    const cpx<R>* Fe = static_cast<const cpx<R>*>(a.F) + (long long)e * a.nb * 256;
    for (int b = 0; b < a.nb; ++b) {
        // A operand of step s: F_b[nl][4 s + g] (and its negated imaginary part)
        R ar[4], ai[4], ain[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const cpx<R> f = Fe[(long long)b * 256 + nl * 16 + 4 * s + g];
            ar[s] = f.re;
            ai[s] = f.im;
            ain[s] = -f.im;
        }
        typename MF::V yr[4], yi[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            yr[t] = typename MF::V{0, 0, 0, 0};
            yi[t] = typename MF::V{0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                yr[t] = MF::op(ar[s], xr[t][s], yr[t]);
                yr[t] = MF::op(ain[s], xi[t][s], yr[t]);
                yi[t] = MF::op(ar[s], xi[t][s], yi[t]);
                yi[t] = MF::op(ai[s], xr[t][s], yi[t]);
            }
        }
        const int* mp = smap + b * kFusedMapStride;
        if (b + 1 < a.nb) {
            // through LDS in natural row order, back in block b + 1's grouping
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int base = mp[16 + o[t]];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int j = base + mp[MF::row(r, g)];
                    lds[w[t] * D + fc_slot<N>(j)] = mk<R>(yr[t][r], yi[t][r]);
                }
            }
            fc_wave_fence();
            const int* mn = mp + kFusedMapStride;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int base = mn[16 + o[t]];
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const cpx<R> v = lds[w[t] * D + fc_slot<N>(base + mn[4 * s + g])];
                    xr[t][s] = v.re;
                    xi[t][s] = v.im;
                }
            }
            fc_wave_fence();
        } else {
            // trace entries: row c + trace_off of column c
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int col = c0 + w[t];
                const int base = mp[16 + o[t]];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int j = base + mp[MF::row(r, g)];
                    if (col < m && j == col + a.trace_off) {
                        double* d = a.diag + 2 * ((long long)e * m + col);
                        d[0] = (double)yr[t][r];
                        d[1] = (double)yi[t][r];
                    }
                }
            }
        }
    }
}

size_t fused_cost_lds_bytes(bool f32, int nb) {
    return (((size_t)nb * kFusedMapStride * sizeof(int) + 15) & ~(size_t)15) + (size_t)kFcWaves * 1024 * (f32 ? 8 : 16);
}

template <typename R>
static hipError_t launch_fc(int N, const FusedCostArgs& a, int grid, size_t shm, hipStream_t s) {
    switch (N) {
#define SQH_FC_CASE(n) \
    case n: hipLaunchKernelGGL((k_fused_cost<R, n>), dim3(grid), dim3(64 * kFcWaves), shm, s, a); break;
        SQH_FC_CASE(4)
        SQH_FC_CASE(5)
        SQH_FC_CASE(6)
        SQH_FC_CASE(7)
        SQH_FC_CASE(8)
        SQH_FC_CASE(9)
        SQH_FC_CASE(10)
#undef SQH_FC_CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fused_cost(bool f32, int N, FusedCostArgs a, hipStream_t s) {
    if (N < 4 || N > kFusedCostMaxQubits || a.nb <= 0 || a.m <= 0 || a.B <= 0) return hipErrorInvalidValue;
    const int W = 64 / ((1 << N) >> 4);
    a.groups = (a.m + W - 1) / W;
    const long long waves = (long long)a.B * a.groups;
    const int grid = (int)((waves + kFcWaves - 1) / kFcWaves);
    const size_t shm = fused_cost_lds_bytes(f32, a.nb);
    if (shm > 64 * 1024) return hipErrorInvalidValue;
    return f32 ? launch_fc<float>(N, a, grid, shm, s) : launch_fc<double>(N, a, grid, shm, s);
}

}  // namespace sqh
