// sqh_k_mlp.hip -- kernels of the warm-start network: a multilayer perceptron with
// BF16 weights/activations and FP32 accumulation on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16), GELU hidden layers, mean-squared-error loss and Adam.
//
// GEMM convention: C[M][N] = sum_k A(m,k) B(k,n) with A(m,k) = A[m*lda+k] (or
// A[k*lda+m] when AT) and B(k,n) = B[k*ldb+n] (or B[n*ldb+k] when BT).  A workgroup
// of 4 wavefronts computes a 64x64 tile, each wavefront a 32x32 quarter as 2x2
// MFMA 16x16 blocks, K advancing 32 per step through LDS.  M, N are multiples of
// 64 and K of 32 (the host pads every dimension).
#include <hip/hip_runtime.h>

#include <cmath>

namespace sqh {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

enum MlpEpi : int {
    EPI_HIDDEN = 0,   // z = acc + bias -> Z (fp32), H = bf16(gelu(z))
    EPI_OUTPUT = 1,   // y = acc + bias -> Y (fp32)
    EPI_BWD_DATA = 2, // dz = acc * gelu'(Zin) -> dZf (fp32), dZb (bf16)
    EPI_STORE = 3,    // C (fp32) = acc
};

struct EpiArgs {
    const float* bias;
    float* zf;        // fp32 output (Z, Y, dZ or C)
    bf16* hb;         // bf16 output (H or dZ)
    const float* zin; // pre-activation for gelu' (EPI_BWD_DATA)
    int ld;           // leading dimension of the outputs / zin
};

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float z) {
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

// One operand tile (64 rows of M or N x KS values of K) between global memory, the
// thread's registers and LDS ([row][k], k contiguous, rows padded by 8 values).  KMAJ:
// the operand is stored K-major (row k holds the 64 rows' values contiguously), so its
// 8-value vectors run along the rows and are transposed on the way into LDS.
template <bool KMAJ, int KS>
struct GemmTile {
    static constexpr int V = 64 * KS / 8 / 256;  // 8-value vectors per thread
    bf16x8 r[V];
    __device__ __forceinline__ void load(const bf16* __restrict__ X, int ld, int row0, int k0, int tid) {
#pragma unroll
        for (int i = 0; i < V; ++i) {
            const int v = tid + 256 * i;
            if constexpr (!KMAJ) {
                const int row = v / (KS / 8), kk = (v % (KS / 8)) * 8;
                r[i] = *reinterpret_cast<const bf16x8*>(&X[(size_t)(row0 + row) * ld + k0 + kk]);
            } else {
                const int k = v / 8, rr = (v % 8) * 8;
                r[i] = *reinterpret_cast<const bf16x8*>(&X[(size_t)(k0 + k) * ld + row0 + rr]);
            }
        }
    }
    __device__ __forceinline__ void store(bf16 (*S)[KS + 8], int tid) const {
#pragma unroll
        for (int i = 0; i < V; ++i) {
            const int v = tid + 256 * i;
            if constexpr (!KMAJ) {
                const int row = v / (KS / 8), kk = (v % (KS / 8)) * 8;
                *reinterpret_cast<bf16x8*>(&S[row][kk]) = r[i];
            } else {
                const int k = v / 8, rr = (v % 8) * 8;
#pragma unroll
                for (int j = 0; j < 8; ++j) S[rr + j][k] = r[i][j];
            }
        }
    }
};

// C[64 x 64 tile] = A B on the matrix cores: 4 wavefronts of 32 x 32 (2 x 2
// v_mfma_f32_16x16x32_bf16 tiles), K in stages of KS through double-buffered LDS with
// one barrier per stage; the next stage's global loads are in flight while the current
// stage's MFMAs run.  AT: A stored K-major (A^T given); BT: B stored N-major (B^T
// given).  Fused epilogues (EPI) write FP32 and BF16 results.
template <bool AT, bool BT, int EPI, int KS>
__global__ __launch_bounds__(256) void k_gemm_bf16(const bf16* __restrict__ A, int lda, const bf16* __restrict__ B,
                                                   int ldb, int K, EpiArgs e) {
    // [2 buffers][A, B][64][KS + 8] in dynamic LDS (up to 135 KiB at KS = 256)
    extern __shared__ __align__(16) unsigned char gemm_smem[];
    typedef bf16 Row[KS + 8];
    Row* const rows = reinterpret_cast<Row*>(gemm_smem);  // A0, B0, A1, B1: 64 rows each
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    GemmTile<AT, KS> ta;
    GemmTile<!BT, KS> tb;
    ta.load(A, lda, m0, 0, tid);
    tb.load(B, ldb, n0, 0, tid);
    ta.store(rows, tid);
    tb.store(rows + 64, tid);
    __syncthreads();
    const int S = K / KS;
    for (int st = 0; st < S; ++st) {
        const int cur = st & 1;
        Row* const As = rows + 128 * cur;
        Row* const Bs = rows + 128 * cur + 64;
        if (st + 1 < S) {
            ta.load(A, lda, m0, (st + 1) * KS, tid);
            tb.load(B, ldb, n0, (st + 1) * KS, tid);
        }
#pragma unroll
        for (int ks = 0; ks < KS; ks += 32) {
            bf16x8 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                a[i] = *reinterpret_cast<const bf16x8*>(&As[wm + i * 16 + (lane & 15)][ks + 8 * (lane >> 4)]);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                b[j] = *reinterpret_cast<const bf16x8*>(&Bs[wn + j * 16 + (lane & 15)][ks + 8 * (lane >> 4)]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (st + 1 < S) {
            ta.store(rows + 128 * (cur ^ 1), tid);
            tb.store(rows + 128 * (cur ^ 1) + 64, tid);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm + i * 16 + 4 * (lane >> 4) + r;
                const int col = n0 + wn + j * 16 + (lane & 15);
                const size_t o = (size_t)row * e.ld + col;
                const float v = acc[i][j][r];
                if constexpr (EPI == EPI_HIDDEN) {
                    const float z = v + e.bias[col];
                    e.zf[o] = z;
                    e.hb[o] = (bf16)gelu(z);
                } else if constexpr (EPI == EPI_OUTPUT) {
                    e.zf[o] = v + e.bias[col];
                } else if constexpr (EPI == EPI_BWD_DATA) {
                    const float dz = v * gelu_grad(e.zin[o]);
                    e.zf[o] = dz;
                    e.hb[o] = (bf16)dz;
                } else {
                    e.zf[o] = v;
                }
            }
}

// x (B x in, fp32, row stride in) -> padded bf16 (Bp x in_pad); rows >= B are zero.
__global__ void k_to_bf16_padded(const float* __restrict__ x, int B, int in, bf16* __restrict__ out, int Bp,
                                 int in_pad) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)Bp * in_pad) return;
    const int r = (int)(i / in_pad), c = (int)(i % in_pad);
    out[i] = (bf16)((r < B && c < in) ? x[(size_t)r * in + c] : 0.0f);
}

// Mean squared error over the B x out real entries; dY = 2 (y - t) / (B * out) (bf16 and
// fp32), padded entries get 0.  Per-block partial losses go to part[blockIdx].
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ y, int ldy, const float* __restrict__ t, int B,
                                             int out, int Bp, float* __restrict__ dyf, bf16* __restrict__ dyb,
                                             double* __restrict__ part) {
    __shared__ double sh[4];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (i < (size_t)Bp * ldy) {
        const int r = (int)(i / ldy), c = (int)(i % ldy);
        float g = 0.0f;
        if (r < B && c < out) {
            const float d = y[i] - t[(size_t)r * out + c];
            l = (double)d * d;
            g = 2.0f * d / ((float)B * (float)out);
        }
        dyf[i] = g;
        dyb[i] = (bf16)g;
    }
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// Output gradient supplied by the caller (dy: B x out, FP32) into the padded FP32 and
// BF16 buffers of the last layer; padding rows and columns are zero.
__global__ __launch_bounds__(256) void k_out_grad(const float* __restrict__ dy, int B, int out, int ldy, int Bp,
                                                  float* __restrict__ dyf, bf16* __restrict__ dyb) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)Bp * ldy) return;
    const int r = (int)(i / ldy), c = (int)(i % ldy);
    const float g = (r < B && c < out) ? dy[(size_t)r * out + c] : 0.0f;
    dyf[i] = g;
    dyb[i] = (bf16)g;
}

// db[n] = sum over the first B rows of dz[:, n]: a workgroup per 16 columns, its 64
// row slices each summing every 64th row, then combined in LDS in a fixed order (the
// sum is the same on every run).  64 row slices keep 64 workgroups of 16 wavefronts
// busy for a 1024-wide layer, where 4 slices per 64 columns kept 16 workgroups.
__global__ __launch_bounds__(1024) void k_bias_grad(const float* __restrict__ dz, int ld, int B, int N,
                                                    float* __restrict__ db) {
    __shared__ float part[64][17];
    const int cx = threadIdx.x & 15, ry = threadIdx.x >> 4;
    const int n = blockIdx.x * 16 + cx;
    float s = 0.0f;
    if (n < N)
        for (int r = ry; r < B; r += 64) s += dz[(size_t)r * ld + n];
    part[ry][cx] = s;
    __syncthreads();
    if (ry == 0 && n < N) {
        float t = 0.0f;
        for (int k = 0; k < 64; ++k) t += part[k][cx];
        db[n] = t;
    }
}

// Adam on fp32 master weights; refreshes the bf16 copy when wb is given.
__global__ void k_adam(float* __restrict__ w, bf16* __restrict__ wb, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, size_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = (bf16)wi;
}

__global__ void k_f32_to_bf16(const float* __restrict__ a, bf16* __restrict__ b, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (bf16)a[i];
}

// Host launchers ------------------------------------------------------------------

template <bool AT, bool BT, int EPI>
static hipError_t gemm(const void* A, int lda, const void* B, int ldb, int M, int N, int K, const EpiArgs& e,
                       hipStream_t s) {
    if (M % 64 || N % 64 || K % 32 || K <= 0) return hipErrorInvalidValue;
    // the widest K stage dividing K: fewer stages, each with more MFMAs per load wait
#define SQH_GEMM_KS(ks)                                                                                     \
    if (K % ks == 0) {                                                                                      \
        const size_t shm = 4 * 64 * (ks + 8) * sizeof(bf16);                                                \
        hipLaunchKernelGGL((k_gemm_bf16<AT, BT, EPI, ks>), dim3(N / 64, M / 64), dim3(256), shm, s,         \
                           (const bf16*)A, lda, (const bf16*)B, ldb, K, e);                                 \
        return hipGetLastError();                                                                           \
    }
    SQH_GEMM_KS(256)
    SQH_GEMM_KS(128)
    SQH_GEMM_KS(64)
    SQH_GEMM_KS(32)
#undef SQH_GEMM_KS
    return hipErrorInvalidValue;
}

// Z[M x N] = H[M x K] W[N x K]^T + b  (hidden: also H' = gelu(Z); output: Y)
hipError_t mlp_gemm_forward(bool hidden, const void* H, int K, const void* W, const float* bias, int M, int N,
                            float* Z, void* Hout, hipStream_t s) {
    EpiArgs e{bias, Z, (bf16*)Hout, nullptr, N};
    return hidden ? gemm<false, true, EPI_HIDDEN>(H, K, W, K, M, N, K, e, s)
                  : gemm<false, true, EPI_OUTPUT>(H, K, W, K, M, N, K, e, s);
}

// dH[M x K] = dZ[M x N] W[N x K], then dZprev = dH * gelu'(Zprev)
hipError_t mlp_gemm_backward_data(const void* dZ, int N, const void* W, int K, int M, const float* Zprev,
                                  float* dZprev_f, void* dZprev_b, hipStream_t s) {
    EpiArgs e{nullptr, dZprev_f, (bf16*)dZprev_b, Zprev, K};
    return gemm<false, false, EPI_BWD_DATA>(dZ, N, W, K, M, K, N, e, s);
}

// dW[N x K] = dZ[B x N]^T H[B x K]
hipError_t mlp_gemm_backward_weights(const void* dZ, int N, const void* H, int K, int B, float* dW, hipStream_t s) {
    EpiArgs e{nullptr, dW, nullptr, nullptr, K};
    return gemm<true, false, EPI_STORE>(dZ, N, H, K, N, K, B, e, s);
}

hipError_t mlp_to_bf16_padded(const float* x, int B, int in, void* out, int Bp, int in_pad, hipStream_t s) {
    const size_t n = (size_t)Bp * in_pad;
    hipLaunchKernelGGL(k_to_bf16_padded, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, B, in, (bf16*)out, Bp,
                       in_pad);
    return hipGetLastError();
}

hipError_t mlp_mse(const float* y, int ldy, const float* t, int B, int out, int Bp, float* dyf, void* dyb,
                   double* part, int* nblocks, hipStream_t s) {
    const size_t n = (size_t)Bp * ldy;
    const unsigned g = (unsigned)((n + 255) / 256);
    *nblocks = (int)g;
    hipLaunchKernelGGL(k_mse, dim3(g), dim3(256), 0, s, y, ldy, t, B, out, Bp, dyf, (bf16*)dyb, part);
    return hipGetLastError();
}

hipError_t mlp_out_grad(const float* dy, int B, int out, int ldy, int Bp, float* dyf, void* dyb, hipStream_t s) {
    const size_t n = (size_t)Bp * ldy;
    hipLaunchKernelGGL(k_out_grad, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dy, B, out, ldy, Bp, dyf,
                       (bf16*)dyb);
    return hipGetLastError();
}

hipError_t mlp_bias_grad(const float* dz, int ld, int B, int N, float* db, hipStream_t s) {
    hipLaunchKernelGGL(k_bias_grad, dim3((N + 15) / 16), dim3(1024), 0, s, dz, ld, B, N, db);
    return hipGetLastError();
}

hipError_t mlp_adam(float* w, void* wb, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2,
                    float eps, int step, hipStream_t s) {
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, w, (bf16*)wb, g, m, v, n, lr, b1,
                       b2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t mlp_f32_to_bf16(const float* a, void* b, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_f32_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, (bf16*)b, n);
    return hipGetLastError();
}

}  // namespace sqh
