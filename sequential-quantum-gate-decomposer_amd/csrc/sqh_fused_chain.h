// sqh_fused_chain.h -- the fused 4-qubit blocks on the matrix cores: MFMA helpers
// (shared by the fused kernels) and the transposed chain from unit vectors as a device
// function for the cost + gradient kernel (k_fused_grad_chain).  The cost kernel
// (k_fused_trace) runs the same chain inline: called through this function, its loop's
// LDS loads went through extra register copies (149 against 51 v_mov per kernel) and
// it ran 4 % slower.
//
// Trace entry c is (V U)[rho, c] with rho = c + trace_off and V = F_{nb-1} ... F_0: row
// rho of V dotted with column c of U.  Row rho of V is column rho of
// V^T = F_0^T F_1^T ... F_{nb-1}^T, built from the unit vector e_rho: F_{nb-1}^T e_rho is
// row rho of F_{nb-1} on the rows that agree with rho outside the block's qubits
// (loaded, no arithmetic), and each later block only touches the rows that agree with
// rho outside the union S_b of the qubits seen so far (the support): 2^(|S_b| - 4) of a
// column's 2^N / 16 outer groups.
//
// Layout: a wavefront owns W = 16 T / O columns (O = 2^N / 16); block b's vector column
// n holds column w = n >> lg_b, support group k = n & (2^lg_b - 1), rows
// kb_b[k] | (rho_w & nS_b) | off_b[i] (kTraceMapStride ints per block: off[16], kb[64],
// lg, nS).  Only the first ceil(W 2^lg_b / 16) tiles carry work.  The wavefront's LDS
// region starts at zero and rows outside a support are never stored, so they read as
// zero (the supports only grow).  Once the support covers every qubit the blocks use the
// forward pass's layout (vector column 16 t + nl = column (16 t + nl) / O, outer group
// (16 t + nl) % O), with no row arithmetic.
#pragma once
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

// local index of row j in a block (bit q <-> the block's q-th smallest qubit), from the
// row offsets of local indices 1, 2, 4 and 8
__device__ __forceinline__ int ft_local(const int* mp, int j) {
    return ((j & mp[1]) ? 1 : 0) | ((j & mp[2]) ? 2 : 0) | ((j & mp[4]) ? 4 : 0) | ((j & mp[8]) ? 8 : 0);
}

// Y = F^T X on the first TS tiles (the forward kernel's step and product order)
template <typename R, int T, int TS>
__device__ __forceinline__ void ft_mma(const R (&ar)[4], const R (&ai)[4], const R (&ain)[4], const R (&xr)[T][4],
                                       const R (&xi)[T][4], typename FcMfma<R>::V (&yr)[T],
                                       typename FcMfma<R>::V (&yi)[T]) {
    typedef FcMfma<R> MF;
#pragma unroll
    for (int t = 0; t < TS; ++t) {
        yr[t] = typename MF::V{0, 0, 0, 0};
        yi[t] = typename MF::V{0, 0, 0, 0};
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
#pragma unroll
        for (int t = 0; t < TS; ++t) {
            yr[t] = MF::op(ar[s], xr[t][s], yr[t]);
            yi[t] = MF::op(ar[s], xi[t][s], yi[t]);
        }
#pragma unroll
        for (int t = 0; t < TS; ++t) {
            yr[t] = MF::op(ain[s], xi[t][s], yr[t]);
            yi[t] = MF::op(ai[s], xr[t][s], yi[t]);
        }
    }
}

// The chain F_{nb-1}^T, ..., F_1^T on e_rho for the wavefront's W columns c0 + w
// (rho = c0 + w + trace_off; columns past m take row trace_off).  On return xr / xi hold
// block 0's B operand in the chain's grouping of block 0 and, with PF, far / fai block
// 0's A operand (F_0^T[nl][4 s + g]).  smap: the chain tables in LDS; lds: the
// wavefront's region of W 2^N entries.
template <typename R, int N, int T, bool PF>
__device__ __forceinline__ void fused_chain_to_block0(const int* smap, int nb, cpx<R>* lds, const cpx<R>* Fe, int c0,
                                                      int m, int trace_off, int lane, R (&xr)[T][4], R (&xi)[T][4],
                                                      R (&far)[4], R (&fai)[4]) {
    typedef FcMfma<R> MF;
    constexpr int D = 1 << N;
    constexpr int O = D >> 4;
    constexpr int W = 16 * T / O;
    static_assert(W >= 1, "a wavefront holds whole columns");
    const int nl = lane & 15, g = lane >> 4;
    auto rho_of = [&](int w) { return (c0 + w < m ? c0 + w : 0) + trace_off; };
    // X of block bn (B operand of step s, tile t: X[4 s + g][16 t + nl] in its grouping)
    // from the LDS region, for the tiles holding its support's groups; vector columns
    // past the wavefront's columns read column 0 (their products are never stored)
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
        for (int s = 0; s < 4; ++s) xr[t][s] = xi[t][s] = 0;
    auto load_x = [&](int bn) {
        const int* mn = smap + bn * kTraceMapStride;
        const int lgn = mn[80], nSn = mn[81];
        const int tsn = ((W << lgn) + 15) >> 4;
#pragma unroll
        for (int t = 0; t < T; ++t) {
            if (t < tsn) {
                const int n = 16 * t + nl;
                const int w = n >> lgn < W ? n >> lgn : 0;
                const int base = mn[16 + (n & ((1 << lgn) - 1))] | (rho_of(w) & nSn);
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const cpx<R> v = lds[w * D + fc_slot<N>(base + mn[4 * s + g])];
                    xr[t][s] = v.re;
                    xi[t][s] = v.im;
                }
            }
        }
        fc_wave_fence();
    };
    // The region starts at zero (rows outside a block's support are never stored and
    // read as zero; the supports only grow), then holds F_{nb-1}^T e_rho per column:
    // row loc(rho) of F_{nb-1} on the rows that agree with rho outside its qubits.
    for (int i = lane; i < W * D; i += 64) lds[i] = mk<R>(0, 0);
    fc_wave_fence();
    {
        const int* mq = smap + (nb - 1) * kTraceMapStride;
        const cpx<R>* Fq = Fe + (long long)(nb - 1) * 256;
        for (int p = lane; p < 16 * W; p += 64) {
            const int w = p >> 4, i = p & 15;
            const int rho = rho_of(w);
            lds[w * D + fc_slot<N>((rho & mq[81]) | mq[i])] = Fq[ft_local(mq, rho) * 16 + i];
        }
    }
    fc_wave_fence();
    load_x(nb - 2);
    // A operand of step s: F_b^T[nl][4 s + g] = F_b[4 s + g][nl] (with PF, block b - 1's
    // is loaded while block b runs)
    if constexpr (PF) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const cpx<R> f = Fe[(long long)(nb - 2) * 256 + (4 * s + g) * 16 + nl];
            far[s] = f.re;
            fai[s] = f.im;
        }
    }
    // A operand of block b into ar / ai / ain (with PF from far / fai, block b - 1's then
    // loaded into them)
    auto a_operand = [&](int b, R (&ar)[4], R (&ai)[4], R (&ain)[4]) {
        if constexpr (PF) {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                ar[s] = far[s];
                ai[s] = fai[s];
                ain[s] = -fai[s];
            }
            if (b > 0) {
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const cpx<R> f = Fe[(long long)(b - 1) * 256 + (4 * s + g) * 16 + nl];
                    far[s] = f.re;
                    fai[s] = f.im;
                }
            }
        } else {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const cpx<R> f = Fe[(long long)b * 256 + (4 * s + g) * 16 + nl];
                ar[s] = f.re;
                ai[s] = f.im;
                ain[s] = -f.im;
            }
        }
    };
    // block b's output Y (its support's tiles) through LDS in natural row order, back in
    // block b - 1's grouping
    auto exchange = [&](int b, int lg, int nS, int ts, const typename MF::V (&yr)[T], const typename MF::V (&yi)[T]) {
        const int* mp = smap + b * kTraceMapStride;
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int n = 16 * t + nl;
            const int w = n >> lg;
            if (t < ts && w < W) {
                const int base = mp[16 + (n & ((1 << lg) - 1))] | (rho_of(w) & nS);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int j = base + mp[MF::row(r, g)];
                    lds[w * D + fc_slot<N>(j)] = mk<R>(yr[t][r], yi[t][r]);
                }
            }
        }
        fc_wave_fence();
        load_x(b - 1);
    };
    // The supports only grow along the chain: first the blocks whose support fills
    // fewer than T tiles (1, or 2 of 4), then the dense ones.  (Separate loops keep the
    // tile-count variants' registers apart.)
    int b = nb - 2;
    for (; b > 0; --b) {
        const int* mp = smap + b * kTraceMapStride;
        const int lg = mp[80], nS = mp[81];
        const int ts = ((W << lg) + 15) >> 4;
        if (ts >= T) break;
        R ar[4], ai[4], ain[4];
        a_operand(b, ar, ai, ain);
        typename MF::V yr[T], yi[T];
        if (T >= 4 && ts >= 2) ft_mma<R, T, (T >= 4 ? 2 : 1)>(ar, ai, ain, xr, xi, yr, yi);
        else ft_mma<R, T, 1>(ar, ai, ain, xr, xi, yr, yi);
        exchange(b, lg, nS, ts, yr, yi);
    }
    // Blocks of full support (no row outside it): vector column 16 t + nl is column
    // wd[t], outer group od[t], as in the forward pass -- no predicates, no row of V.
    int wd[T], od[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        wd[t] = (16 * t + nl) / O;
        od[t] = (16 * t + nl) % O;
    }
    for (; b > 0; --b) {
        const int* mp = smap + b * kTraceMapStride;
        const int* mn = mp - kTraceMapStride;
        R ar[4], ai[4], ain[4];
        a_operand(b, ar, ai, ain);
        typename MF::V yr[T], yi[T];
        ft_mma<R, T, T>(ar, ai, ain, xr, xi, yr, yi);
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int base = mp[16 + od[t]];
#pragma unroll
            for (int r = 0; r < 4; ++r)
                lds[wd[t] * D + fc_slot<N>(base + mp[MF::row(r, g)])] = mk<R>(yr[t][r], yi[t][r]);
        }
        fc_wave_fence();
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int base = mn[16 + od[t]];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const cpx<R> v = lds[wd[t] * D + fc_slot<N>(base + mn[4 * s + g])];
                xr[t][s] = v.re;
                xi[t][s] = v.im;
            }
        }
        fc_wave_fence();
    }
}

}  // namespace sqh
