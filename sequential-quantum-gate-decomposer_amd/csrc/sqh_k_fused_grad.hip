// sqh_k_fused_grad.hip -- batched cost and gradient of the circuit as fused dense
// 4-qubit blocks on the matrix cores, every column resident on the chip.
//
// Forward: as the fused cost kernel (sqh_k_fused_cost.hip), X_b = F_b X_{b-1} from the
// target, and the trace entries.  Backward, with Y the trace selector propagated
// through the blocks (Y_{b-1} = F_b^T Y_b, so Tr = sum Y_b * X_b for every b), per block
// from the last:
//   X_{b-1} = F_b^dagger X_b                                   (MFMA, F unitary)
//   M_b[i][k] += sum over the wavefront's (outer group, column) of Y_b[(i,o)] X_{b-1}[(k,o)]
//                                                              (MFMA on transposed tiles)
//   Y_{b-1} = F_b^T Y_b                                        (MFMA)
// so that dTr/dtheta = sum_ik (dF_b/dtheta)_ik M_b[i][k] for a parameter of block b.
// The wavefronts of a workgroup (one batch element) sum their M_b in LDS in a fixed
// order and write one partial per workgroup and block; k_fused_block_corr then sums the
// partials and runs the block's gates on 16-dimensional vectors (x from e_k, y from
// column k of M_b) backwards, producing each gate's 2x2 correlations in the layout of
// the register-tile adjoint kernel, from which k_finish forms the gradient.
//
// Layouts follow sqh_k_fused_cost.hip: row j of a column under block b is
// base_b[o] + off_b[i]; 16x16x4 MFMA operands A[l % 16][l / 16], B[l / 16][l % 16]; the
// B operand of step s holds rows 4 s + lane / 16, the D layout rows row(r, lane / 16).
#include <hip/hip_runtime.h>

#include "sqh_device.h"
#include "sqh_fused_chain.h"
#include "sqh_kernels.h"

namespace sqh {

typedef double fg_f64x4 __attribute__((ext_vector_type(4)));
typedef float fg_f32x4 __attribute__((ext_vector_type(4)));

template <typename R>
struct FgMfma;
template <>
struct FgMfma<double> {
    typedef fg_f64x4 V;
    static __device__ __forceinline__ V op(double a, double b, V c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int r, int g) { return 4 * r + g; }
};
template <>
struct FgMfma<float> {
    typedef fg_f32x4 V;
    static __device__ __forceinline__ V op(float a, float b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int r, int g) { return 4 * g + r; }
};

template <int N>
__device__ __forceinline__ int fg_slot(int j) {
    if constexpr (N <= 4) {
        return j;
    } else {
        int h = j >> 4;
        if constexpr (N > 8) h ^= h >> 4;
        return j ^ (h & 15);
    }
}

__device__ __forceinline__ void fg_wave_fence() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); }

// Entry (r, c) of a 16 x 16 tile in LDS, columns rotated by the row (XOR) so that both a
// row-wise and a column-wise sweep of 16 lanes touch 16 different bank groups.
__device__ __forceinline__ int fg_tile(int r, int c) { return r * 16 + (c ^ r); }

// Complex 16x16 tile product Y = A X for T tiles: A given by (ar, ai, ain = -ai) per
// step, X the B operand (xr, xi) per tile and step; result in the D layout.
template <typename R, int T>
__device__ __forceinline__ void fg_product(const R (&ar)[4], const R (&ai)[4], const R (&ain)[4], const R (&xr)[T][4],
                                           const R (&xi)[T][4], typename FgMfma<R>::V (&yr)[T],
                                           typename FgMfma<R>::V (&yi)[T]) {
    typedef FgMfma<R> MF;
#pragma unroll
    for (int t = 0; t < T; ++t) {
        yr[t] = typename MF::V{0, 0, 0, 0};
        yi[t] = typename MF::V{0, 0, 0, 0};
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
#pragma unroll
        for (int t = 0; t < T; ++t) {
            yr[t] = MF::op(ar[s], xr[t][s], yr[t]);
            yi[t] = MF::op(ar[s], xi[t][s], yi[t]);
        }
#pragma unroll
        for (int t = 0; t < T; ++t) {
            yr[t] = MF::op(ain[s], xi[t][s], yr[t]);
            yi[t] = MF::op(ai[s], xr[t][s], yi[t]);
        }
    }
}

// D-layout tiles (block b's grouping, tables mp) through the wavefront's LDS region into
// B-layout operands of the grouping of tables mn.
template <typename R, int N, int T>
__device__ __forceinline__ void fg_remap(const typename FgMfma<R>::V (&dr)[T], const typename FgMfma<R>::V (&di)[T],
                                         R (&xr)[T][4], R (&xi)[T][4], const int* mp, const int* mn, const int (&w)[T],
                                         const int (&o)[T], int g, cpx<R>* lds) {
    typedef FgMfma<R> MF;
    constexpr int D = 1 << N;
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int base = mp[16 + o[t]];
#pragma unroll
        for (int r = 0; r < 4; ++r) lds[w[t] * D + fg_slot<N>(base + mp[MF::row(r, g)])] = mk<R>(dr[t][r], di[t][r]);
    }
    fg_wave_fence();
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int base = mn[16 + o[t]];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const cpx<R> v = lds[w[t] * D + fg_slot<N>(base + mn[4 * s + g])];
            xr[t][s] = v.re;
            xi[t][s] = v.im;
        }
    }
    fg_wave_fence();
}

template <typename R, int N, int T>
__global__ __launch_bounds__(512) void k_fused_grad(FusedGradArgs a) {
    typedef FgMfma<R> MF;
    constexpr int D = 1 << N;
    constexpr int O = D >> 4;
    constexpr int W = 16 * T / O;
    static_assert(W >= 1, "a wavefront holds whole columns");
    // Per wavefront one LDS region of RG complex entries, used in turn by the remaps, the
    // two transposed tiles of the correlation product (2 x 256 entries) and the
    // workgroup's reduction of M_b (512 doubles).
    constexpr int RG = 16 * T * 16;
    static_assert(RG >= 512 && RG * sizeof(cpx<R>) >= 512 * sizeof(double), "the region holds the tiles");
    extern __shared__ __align__(16) unsigned char fg_smem[];
    const int nwv = blockDim.x >> 6;
    int* smap = reinterpret_cast<int*>(fg_smem);
    const int nmap = a.nb * kFusedMapStride;
    cpx<R>* region = reinterpret_cast<cpx<R>*>(fg_smem + (((size_t)nmap * sizeof(int) + 15) & ~(size_t)15));
    for (int i = threadIdx.x; i < nmap; i += blockDim.x) smap[i] = a.maps[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // XCD-ordered workgroups (the grid is a multiple of 8); with a live count (the
    // optimiser's rounds) the plain order, so the live elements spread over the XCDs
    int B = a.B;
    long long lwg = blockIdx.x;
    if (a.count) {
        B = min(*a.count, a.B);
    } else {
        const long long per_xcd = gridDim.x >> 3;
        lwg = (long long)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    }
    const long long wave = lwg * nwv + wv;
    // whole workgroups past the batch leave together (the workgroups of one element
    // hold whole column groups: groups % waves == 0)
    if (lwg * nwv >= (long long)B * a.groups) return;
    const int e = (int)(wave / a.groups);
    const int cg = (int)(wave % a.groups);
    const int c0 = cg * W;
    const int wgi = cg / nwv;  // the workgroup's index within its element
    cpx<R>* lds = region + (size_t)wv * RG;
    cpx<R>* ys = lds;
    cpx<R>* xs = lds + 256;
    const int nl = lane & 15, g = lane >> 4;
    int w[T], o[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int n = 16 * t + nl;
        w[t] = n / O;
        o[t] = n % O;
    }
    const cpx<R>* U = static_cast<const cpx<R>*>(a.U);
    if (a.count) U += (long long)a.run_target[a.elem_run[e]] * a.tgt_stride;
    const int m = a.m;
    R xr[T][4], xi[T][4];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int col = c0 + w[t];
        const int base = smap[16 + o[t]];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const cpx<R> v = col < m ? U[(long long)(base + smap[4 * s + g]) * m + col] : mk<R>(0, 0);
            xr[t][s] = v.re;
            xi[t][s] = v.im;
        }
    }
    const cpx<R>* Fe = static_cast<const cpx<R>*>(a.F) + (long long)e * a.nb * 256;
    typename MF::V dr[T], di[T];
    // forward
    for (int b = 0; b < a.nb; ++b) {
        R ar[4], ai[4], ain[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const cpx<R> f = Fe[(long long)b * 256 + nl * 16 + 4 * s + g];
            ar[s] = f.re;
            ai[s] = f.im;
            ain[s] = -f.im;
        }
        fg_product<R, T>(ar, ai, ain, xr, xi, dr, di);
        const int* mp = smap + b * kFusedMapStride;
        // next block's grouping, or for the last block its own B layout
        fg_remap<R, N, T>(dr, di, xr, xi, mp, b + 1 < a.nb ? mp + kFusedMapStride : mp, w, o, g, lds);
    }
    // trace entries from X (B layout of the last block) and the selector Y
    R yr_[T][4], yi_[T][4];
    {
        const int* mp = smap + (a.nb - 1) * kFusedMapStride;
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int col = c0 + w[t];
            const int base = mp[16 + o[t]];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const bool hit = col < m && base + mp[4 * s + g] == col + a.trace_off;
                if (hit) {
                    double* d = a.diag + 2 * ((long long)e * m + col);
                    d[0] = (double)xr[t][s];
                    d[1] = (double)xi[t][s];
                }
                yr_[t][s] = hit ? (R)1 : (R)0;
                yi_[t][s] = (R)0;
            }
        }
    }
    // backward
    for (int b = a.nb - 1; b >= 0; --b) {
        // F_b^T as the A operand: A[i][k] = F[k][i]; F^dagger: its conjugate
        R tr_[4], ti_[4], tin_[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const cpx<R> f = Fe[(long long)b * 256 + (4 * s + g) * 16 + nl];
            tr_[s] = f.re;
            ti_[s] = f.im;
            tin_[s] = -f.im;
        }
        // X_{b-1} = F^dagger X_b (A = conj(F^T): imaginary part negated)
        fg_product<R, T>(tr_, tin_, ti_, xr, xi, dr, di);
        // M_b += Y_b X_{b-1}^T over the wavefront's vector columns, tile by tile through
        // the transposed buffers: ys[i][vc] from Y (B layout), xs[k][vc] from X_{b-1}
        typename MF::V mr = typename MF::V{0, 0, 0, 0}, mi = typename MF::V{0, 0, 0, 0};
#pragma unroll
        for (int t = 0; t < T; ++t) {
#pragma unroll
            for (int s = 0; s < 4; ++s) ys[fg_tile(4 * s + g, nl)] = mk<R>(yr_[t][s], yi_[t][s]);
#pragma unroll
            for (int r = 0; r < 4; ++r) xs[fg_tile(MF::row(r, g), nl)] = mk<R>(dr[t][r], di[t][r]);
            fg_wave_fence();
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const cpx<R> yv = ys[fg_tile(nl, 4 * s + g)];
                const cpx<R> xv = xs[fg_tile(nl, 4 * s + g)];
                mr = MF::op(yv.re, xv.re, mr);
                mi = MF::op(yv.re, xv.im, mi);
                mr = MF::op(-yv.im, xv.im, mr);
                mi = MF::op(yv.im, xv.re, mi);
            }
            fg_wave_fence();
        }
        // the workgroup's wavefronts sum their M_b in a fixed order (each wavefront's
        // region as 512 doubles: [256][re, im])
        double* red = reinterpret_cast<double*>(lds);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int idx = MF::row(r, g) * 16 + nl;
            red[idx * 2] = (double)mr[r];
            red[idx * 2 + 1] = (double)mi[r];
        }
        __syncthreads();
        for (int i = threadIdx.x; i < 512; i += blockDim.x) {
            double sum = 0.0;
            for (int v = 0; v < nwv; ++v) sum += reinterpret_cast<const double*>(region + (size_t)v * RG)[i];
            a.mpart[(((long long)e * a.wgs + wgi) * a.nb + b) * 512 + i] = sum;
        }
        __syncthreads();
        if (b == 0) break;
        // Y_{b-1} = F^T Y_b, then both to the grouping of block b - 1
        typename MF::V er[T], ei[T];
        fg_product<R, T>(tr_, ti_, tin_, yr_, yi_, er, ei);
        const int* mp = smap + b * kFusedMapStride;
        const int* mn = mp - kFusedMapStride;
        fg_remap<R, N, T>(dr, di, xr, xi, mp, mn, w, o, g, lds);
        fg_remap<R, N, T>(er, ei, yr_, yi_, mp, mn, w, o, g, lds);
    }
}

// Cost + gradient with the adjoint vectors from the transposed chain (sqh_fused_chain.h).
// In the notation above, the adjoint after block b is Z_{b+1} = F_{b+1}^T ... F_{nb-1}^T
// e_rho (Z_nb = e_rho): exactly the chain's vectors, which touch only the rows inside the
// support of the blocks seen so far.  Phase 1 runs the chain down to Z_1 (block 0's
// operand), then Z_0 = F_0^T Z_1 gives the trace entries (dotted with the target's
// columns).  Phase 2 sweeps up: X_{-1} = the target's columns,
//   M_b += Z_{b+1} X_{b-1}^T,   X_b = F_b X_{b-1},   Z_{b+2} = conj(F_{b+1}) Z_{b+1}
// (F unitary: (F^T)^-1 = conj(F)), Z_nb set to the selector itself.  Block products:
// the chain's, then 3 nb - 3 (the 8-qubit 40-layer ansatz: 4 + 15 = 19 against the
// backward sweep's 23).  The correlations M_b and their partials are those of
// k_fused_grad.  Needs nb >= 2 and block 0's support to cover every qubit (the chain
// then ends in the forward layout).
template <typename R, int N, int T>
__global__ __launch_bounds__(512) void k_fused_grad_chain(FusedGradArgs a) {
    typedef FgMfma<R> MF;
    constexpr int D = 1 << N;
    constexpr int O = D >> 4;
    constexpr int W = 16 * T / O;
    static_assert(W >= 1, "a wavefront holds whole columns");
    constexpr int RG = 16 * T * 16;
    static_assert(RG >= 512 && RG * sizeof(cpx<R>) >= 512 * sizeof(double), "the region holds the tiles");
    extern __shared__ __align__(16) unsigned char fg_smem[];
    const int nwv = blockDim.x >> 6;
    int* smap = reinterpret_cast<int*>(fg_smem);
    int* tmap = smap + a.nb * kFusedMapStride;
    const int nmap = a.nb * (kFusedMapStride + kTraceMapStride);
    cpx<R>* region = reinterpret_cast<cpx<R>*>(fg_smem + (((size_t)nmap * sizeof(int) + 15) & ~(size_t)15));
    for (int i = threadIdx.x; i < a.nb * kFusedMapStride; i += blockDim.x) smap[i] = a.maps[i];
    for (int i = threadIdx.x; i < a.nb * kTraceMapStride; i += blockDim.x) tmap[i] = a.tmaps[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int B = a.B;
    long long lwg = blockIdx.x;
    if (a.count) {
        B = min(*a.count, a.B);
    } else {
        const long long per_xcd = gridDim.x >> 3;
        lwg = (long long)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    }
    const long long wave = lwg * nwv + wv;
    if (lwg * nwv >= (long long)B * a.groups) return;  // whole workgroups
    const int e = (int)(wave / a.groups);
    const int cg = (int)(wave % a.groups);
    const int c0 = cg * W;
    const int wgi = cg / nwv;
    cpx<R>* lds = region + (size_t)wv * RG;
    cpx<R>* ys = lds;
    cpx<R>* xs = lds + 256;
    const int nl = lane & 15, g = lane >> 4;
    int w[T], o[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        w[t] = (16 * t + nl) / O;
        o[t] = (16 * t + nl) % O;
    }
    const cpx<R>* U = static_cast<const cpx<R>*>(a.U);
    if (a.count) U += (long long)a.run_target[a.elem_run[e]] * a.tgt_stride;
    const int m = a.m;
    const cpx<R>* Fe = static_cast<const cpx<R>*>(a.F) + (long long)e * a.nb * 256;
    // phase 1: Z_1 in block 0's grouping (B layout)
    R zr[T][4], zi[T][4], far[4], fai[4];
    fused_chain_to_block0<R, N, T, false>(tmap, a.nb, lds, Fe, c0, m, a.trace_off, lane, zr, zi, far, fai);
    typename MF::V er[T], ei[T];
    {
        // Z_0 = F_0^T Z_1, times the target's columns: the trace entries
        R ar[4], ai[4], ain[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const cpx<R> f = Fe[(4 * s + g) * 16 + nl];
            ar[s] = f.re;
            ai[s] = f.im;
            ain[s] = -f.im;
        }
        fg_product<R, T>(ar, ai, ain, zr, zi, er, ei);
        constexpr int seg = O < 16 ? O : 16;
        double sr = 0, si = 0;
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int col = c0 + w[t];
            if (col < m) {
                const int base = smap[16 + o[t]];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const cpx<R> u = U[(long long)(base + smap[MF::row(r, g)]) * m + col];
                    sr += (double)er[t][r] * (double)u.re - (double)ei[t][r] * (double)u.im;
                    si += (double)er[t][r] * (double)u.im + (double)ei[t][r] * (double)u.re;
                }
            }
            if (O < 16 || ((t + 1) * 16) % O == 0) {
                sr += __shfl_xor(sr, 16, 64);
                si += __shfl_xor(si, 16, 64);
                sr += __shfl_xor(sr, 32, 64);
                si += __shfl_xor(si, 32, 64);
#pragma unroll
                for (int d = 1; d < seg; d <<= 1) {
                    sr += __shfl_xor(sr, d, 64);
                    si += __shfl_xor(si, d, 64);
                }
                if (g == 0 && (nl & (seg - 1)) == 0 && col < m) {
                    double* dg = a.diag + 2 * ((long long)e * m + col);
                    dg[0] = sr;
                    dg[1] = si;
                }
                sr = 0;
                si = 0;
            }
        }
    }
    // phase 2: X_{-1} = the target's columns in block 0's grouping (B layout)
    R xr[T][4], xi[T][4];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int col = c0 + w[t];
        const int base = smap[16 + o[t]];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const cpx<R> v = col < m ? U[(long long)(base + smap[4 * s + g]) * m + col] : mk<R>(0, 0);
            xr[t][s] = v.re;
            xi[t][s] = v.im;
        }
    }
    // M_bb += Y X^T over the wavefront's vector columns, Y in the B layout (ZB) or the D
    // layout (er / ei), X in the B layout; then the workgroup's partial of block bb
    auto correlate = [&](int bb, bool zb) __attribute__((always_inline)) {
        typename MF::V mr = typename MF::V{0, 0, 0, 0}, mi = typename MF::V{0, 0, 0, 0};
#pragma unroll
        for (int t = 0; t < T; ++t) {
            if (zb) {
#pragma unroll
                for (int s = 0; s < 4; ++s) ys[fg_tile(4 * s + g, nl)] = mk<R>(zr[t][s], zi[t][s]);
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) ys[fg_tile(MF::row(r, g), nl)] = mk<R>(er[t][r], ei[t][r]);
            }
#pragma unroll
            for (int s = 0; s < 4; ++s) xs[fg_tile(4 * s + g, nl)] = mk<R>(xr[t][s], xi[t][s]);
            fg_wave_fence();
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const cpx<R> yv = ys[fg_tile(nl, 4 * s + g)];
                const cpx<R> xv = xs[fg_tile(nl, 4 * s + g)];
                mr = MF::op(yv.re, xv.re, mr);
                mi = MF::op(yv.re, xv.im, mi);
                mr = MF::op(-yv.im, xv.im, mr);
                mi = MF::op(yv.im, xv.re, mi);
            }
            fg_wave_fence();
        }
        double* red = reinterpret_cast<double*>(lds);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int idx = MF::row(r, g) * 16 + nl;
            red[idx * 2] = (double)mr[r];
            red[idx * 2 + 1] = (double)mi[r];
        }
        __syncthreads();
        for (int i = threadIdx.x; i < 512; i += blockDim.x) {
            double sum = 0.0;
            for (int v = 0; v < nwv; ++v) sum += reinterpret_cast<const double*>(region + (size_t)v * RG)[i];
            a.mpart[(((long long)e * a.wgs + wgi) * a.nb + bb) * 512 + i] = sum;
        }
        __syncthreads();
    };
    correlate(0, true);
    typename MF::V dr[T], di[T];
    // one step up: X_b into block b + 1's grouping, Z_{b+2} (D layout, block b + 1's
    // grouping) from Z_{b+1} -- in the B layout of block 0 for the first step (zb), in
    // the D layout of block b after it -- then M_{b+1}.  The first step runs on its own
    // so the B-layout Z is not live through the loop.
    auto step = [&](int b, bool zb) __attribute__((always_inline)) {
        const int* mp = smap + b * kFusedMapStride;
        const int* mn = mp + kFusedMapStride;
        {
            R ar[4], ai[4], ain[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const cpx<R> f = Fe[(long long)b * 256 + nl * 16 + 4 * s + g];
                ar[s] = f.re;
                ai[s] = f.im;
                ain[s] = -f.im;
            }
            fg_product<R, T>(ar, ai, ain, xr, xi, dr, di);
            fg_remap<R, N, T>(dr, di, xr, xi, mp, mn, w, o, g, lds);
        }
        if (b + 2 < a.nb) {
            if (zb) {
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    const int base = mp[16 + o[t]];
#pragma unroll
                    for (int s = 0; s < 4; ++s)
                        lds[w[t] * D + fg_slot<N>(base + mp[4 * s + g])] = mk<R>(zr[t][s], zi[t][s]);
                }
                fg_wave_fence();
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    const int base = mn[16 + o[t]];
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const cpx<R> v = lds[w[t] * D + fg_slot<N>(base + mn[4 * s + g])];
                        zr[t][s] = v.re;
                        zi[t][s] = v.im;
                    }
                }
                fg_wave_fence();
            } else {
                fg_remap<R, N, T>(er, ei, zr, zi, mp, mn, w, o, g, lds);
            }
            R ar[4], ai[4], ain[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const cpx<R> f = Fe[(long long)(b + 1) * 256 + nl * 16 + 4 * s + g];
                ar[s] = f.re;
                ai[s] = -f.im;
                ain[s] = f.im;
            }
            fg_product<R, T>(ar, ai, ain, zr, zi, er, ei);
        } else {
            // Z_nb: the selector, row rho = c + trace_off of column c (D layout)
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const int col = c0 + w[t];
                const int base = mn[16 + o[t]];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    er[t][r] = (col < m && base + mn[MF::row(r, g)] == col + a.trace_off) ? (R)1 : (R)0;
                    ei[t][r] = (R)0;
                }
            }
        }
        correlate(b + 1, false);
    };
    step(0, true);
    for (int b = 1; b + 1 < a.nb; ++b) step(b, false);
}

// The workgroup partials of every (element, block) M summed in a fixed order, in place
// into the first partial: thread per double, consecutive threads on consecutive
// entries.
__global__ __launch_bounds__(256) void k_fused_sum_partials(double* __restrict__ mpart, int wgs, int nb,
                                                            long long total, const int* __restrict__ count) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;  // total = B * nb * 512
    const long long per_e = (long long)nb * 512;
    const long long e = i / per_e, r = i % per_e;
    if (count && e >= *count) return;
    double* base = mpart + e * wgs * per_e + r;
    double sum = base[0];
    for (int v = 1; v < wgs; ++v) sum += base[(long long)v * per_e];
    base[0] = sum;
}

// Gates staged per pass and batch elements per workgroup (x 16 block columns = one
// wavefront) of k_fused_block_corr.
constexpr int kCorrChunk = 32;
constexpr int kCorrElems = 4;

// Workgroup -> (block, 4 batch elements); thread -> (element, column k of the block).
// Thread k starts from x = F_b e_k (column k of the block matrix) and y = column k of
// M_b (the workgroup partials, summed into the first) and walks the block's gates
// backwards: x <- G^dagger x, the pair correlations of y and x, y <- G^T y.  The 16
// threads of an element sum their correlations in a fixed butterfly order, giving each
// parametric gate's 2x2 correlations in corr[e][slot][8] (the register-tile adjoint's
// layout).  Gate records and the elements' 2x2 kernels are staged in LDS a chunk at a
// time (from the block's end), so the walk issues no dependent global loads.
template <typename R>
__global__ __launch_bounds__(16 * kCorrElems) void k_fused_block_corr(const FusedGate* __restrict__ fg,
                                                         const int* __restrict__ block_first, int nb,
                                                         const cpx<R>* __restrict__ kern, long long kern_bstride,
                                                         const cpx<R>* __restrict__ F, const double* __restrict__ mpart,
                                                         int wgs, int B, double* __restrict__ corr, int ns,
                                                         const int* __restrict__ count) {
    __shared__ FusedGate srec[kCorrChunk];
    __shared__ cpx<R> skern[kCorrElems][kCorrChunk][4];
    const int blk = blockIdx.x % nb;
    const int e0 = (blockIdx.x / nb) * kCorrElems;
    if (count) B = min(*count, B);
    if (e0 >= B) return;  // the whole workgroup
    const int le = threadIdx.x >> 4, k = threadIdx.x & 15;
    const int e = e0 + le;
    const bool live = e < B;
    cpx<R> x[16], y[16];
    {
        const cpx<R>* Fb = F + ((long long)(live ? e : 0) * nb + blk) * 256;
        const double* Mb = mpart + ((long long)(live ? e : 0) * wgs * nb + blk) * 512;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            x[i] = live ? Fb[i * 16 + k] : mk<R>(0, 0);
            y[i] = live ? mk<R>((R)Mb[(i * 16 + k) * 2], (R)Mb[(i * 16 + k) * 2 + 1]) : mk<R>(0, 0);
        }
    }
    const int first = block_first[blk], last = block_first[blk + 1];
    for (int p1 = last; p1 > first; p1 -= kCorrChunk) {
        const int p0 = max(first, p1 - kCorrChunk), cnt = p1 - p0;
        __syncthreads();  // the previous chunk is consumed
        for (int i = threadIdx.x; i < cnt; i += blockDim.x) srec[i] = fg[p0 + i];
        __syncthreads();
        for (int i = threadIdx.x; i < kCorrElems * cnt * 4; i += blockDim.x) {
            const int el = i / (cnt * 4), gi = (i / 4) % cnt, t = i % 4;
            const int slot = srec[gi].slot;
            if (slot >= 0 && e0 + el < B) skern[el][gi][t] = kern[(long long)(e0 + el) * kern_bstride + 4 * slot + t];
        }
        __syncthreads();
        for (int gi = cnt - 1; gi >= 0; --gi) {
            const FusedGate gt = srec[gi];
            Op op;
            op.kind = (uint8_t)gt.kind;
            op.t = (uint8_t)gt.t;
            op.cmode = gt.c >= 0 ? CM_REG : CM_NONE;
            op.c = (uint8_t)(gt.c >= 0 ? gt.c : 0);
            op.idx = gt.slot;
            cpx<R> kk[4] = {mk<R>(0, 0), mk<R>(0, 0), mk<R>(0, 0), mk<R>(0, 0)};
            if (gt.slot >= 0) {
#pragma unroll
                for (int t = 0; t < 4; ++t) kk[t] = skern[le][gi][t];
            }
            if (op.kind == OP_X || op.kind == OP_Z) {
                // self-inverse real symmetric: G^dagger = G^T = G
                exec_gate<4, false, R>(x, op, kk, 0);
                exec_gate<4, false, R>(y, op, kk, 0);
                continue;
            }
            cpx<R> kd[4] = {mk(kk[0].re, -kk[0].im), mk(kk[2].re, -kk[2].im), mk(kk[1].re, -kk[1].im),
                            mk(kk[3].re, -kk[3].im)};
            cpx<R> kt[4] = {kk[0], kk[2], kk[1], kk[3]};
            cpx<R> acc[4] = {mk<R>(0, 0), mk<R>(0, 0), mk<R>(0, 0), mk<R>(0, 0)};
            switch (op.kind * 4 + op.t) {
                SQH_ARMS(OP_U, 4, (pairs_apply<T, false, false>(x, kd, 0u), pairs_corr<T>(y, x, acc, 0u, false),
                                   pairs_apply<T, false, false>(y, kt, 0u)))
                SQH_ARMS(OP_CU, 4,
                         (pairs_apply<T, true, false>(x, kd, ctrl_mask<4>(op.cmode, op.c, 0)),
                          pairs_corr<T>(y, x, acc, ctrl_mask<4>(op.cmode, op.c, 0), true),
                          pairs_apply<T, true, false>(y, kt, ctrl_mask<4>(op.cmode, op.c, 0))))
            default:
                break;
            }
            double r[8];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                r[2 * t] = (double)acc[t].re;
                r[2 * t + 1] = (double)acc[t].im;
            }
#pragma unroll
            for (int t = 0; t < 8; ++t) {
#pragma unroll
                for (int off = 8; off >= 1; off >>= 1) r[t] += __shfl_xor(r[t], off, 16);
            }
            if (live && k == 0 && gt.slot >= 0) {
                double* d = corr + ((long long)e * ns + gt.slot) * 8;
#pragma unroll
                for (int t = 0; t < 8; ++t) d[t] = r[t];
            }
        }
    }
}

size_t fused_grad_lds_bytes(bool f32, int nb, int T, int waves) {
    // both row tables (the chain kernel's too), then one region per wavefront
    const size_t esz = f32 ? 8 : 16;
    return (((size_t)nb * (kFusedMapStride + kTraceMapStride) * sizeof(int) + 15) & ~(size_t)15) +
           esz * (size_t)(16 * T * 16) * waves;
}

template <typename R>
static hipError_t launch_fg(int N, const FusedGradArgs& a, int T, int grid, int threads, size_t shm, hipStream_t s) {
#define SQH_FG_CASE(n, t)                                                                                    \
    if (N == n && T == t) {                                                                                  \
        if (a.tmaps)                                                                                         \
            hipLaunchKernelGGL((k_fused_grad_chain<R, n, t>), dim3(grid), dim3(threads), shm, s, a);         \
        else                                                                                                 \
            hipLaunchKernelGGL((k_fused_grad<R, n, t>), dim3(grid), dim3(threads), shm, s, a);               \
        return hipGetLastError();                                                                            \
    }
    SQH_FG_CASE(4, 4)
    SQH_FG_CASE(5, 4)
    SQH_FG_CASE(6, 4)
    SQH_FG_CASE(7, 4)
    SQH_FG_CASE(8, 4)
    SQH_FG_CASE(9, 4)
    SQH_FG_CASE(10, 4)
    SQH_FG_CASE(4, 2)
    SQH_FG_CASE(5, 2)
    SQH_FG_CASE(6, 2)
    SQH_FG_CASE(7, 2)
    SQH_FG_CASE(8, 2)
    SQH_FG_CASE(9, 2)
#undef SQH_FG_CASE
    return hipErrorInvalidValue;
}

// Column groups (one per wavefront) of an element: 16 T vector columns per wavefront,
// 2^N / 16 of them per matrix column.
static int fg_groups(int N, int m, int T) {
    const int W = 16 * T / ((1 << N) >> 4);
    return W < 1 ? 0 : (m + W - 1) / W;
}

bool fused_grad_shape_valid(bool f32, int N, int m, int nb, int T, int waves) {
    if (N < 4 || N > kFusedCostMaxQubits || (T != 2 && T != 4) || m <= 0 || nb <= 0) return false;
    if (waves != 1 && waves != 2 && waves != 4 && waves != 8) return false;
    const int groups = fg_groups(N, m, T);
    return groups > 0 && groups % waves == 0 && fused_grad_lds_bytes(f32, nb, T, waves) <= device_max_lds();
}

int fused_grad_wgs(int N, int m, int T, int waves) { return fg_groups(N, m, T) / waves; }

hipError_t launch_fused_grad(bool f32, int N, FusedGradArgs a, int T, int waves, hipStream_t s) {
    if (a.B <= 0 || !fused_grad_shape_valid(f32, N, a.m, a.nb, T, waves)) return hipErrorInvalidValue;
    a.groups = fg_groups(N, a.m, T);
    if (a.wgs != a.groups / waves) return hipErrorInvalidValue;  // the partials' layout the caller sized
    const long long wgs_total = (long long)a.B * a.wgs;
    const int grid = (int)((wgs_total + 7) / 8 * 8);
    const size_t shm = fused_grad_lds_bytes(f32, a.nb, T, waves);
    return f32 ? launch_fg<float>(N, a, T, grid, 64 * waves, shm, s)
               : launch_fg<double>(N, a, T, grid, 64 * waves, shm, s);
}

hipError_t launch_fused_block_corr(bool f32, const void* fg, const int* block_first, int nb, const void* kern,
                                   long long kern_bstride, const void* F, double* mpart, int wgs, int B, double* corr,
                                   int ns, hipStream_t s, const int* count) {
    const long long pairs = (long long)B * nb;
    if (pairs == 0) return hipSuccess;
    if (wgs > 1) {
        const long long total = pairs * 512;
        hipLaunchKernelGGL(k_fused_sum_partials, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                           mpart, wgs, nb, total, count);
    }
    const long long groups = (B + kCorrElems - 1) / kCorrElems;
    const unsigned grid = (unsigned)(groups * nb);
    if (f32)
        hipLaunchKernelGGL(k_fused_block_corr<float>, dim3(grid), dim3(16 * kCorrElems), 0, s, (const FusedGate*)fg, block_first,
                           nb, (const cpx<float>*)kern, kern_bstride, (const cpx<float>*)F, mpart, wgs, B, corr, ns, count);
    else
        hipLaunchKernelGGL(k_fused_block_corr<double>, dim3(grid), dim3(16 * kCorrElems), 0, s, (const FusedGate*)fg, block_first,
                           nb, (const cpx<double>*)kern, kern_bstride, (const cpx<double>*)F, mpart, wgs, B, corr, ns, count);
    return hipGetLastError();
}

}  // namespace sqh
