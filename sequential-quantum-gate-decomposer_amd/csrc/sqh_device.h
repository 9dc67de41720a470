// sqh_device.h -- gfx950 device code (included by the sqh_k_*.hip translation units) for the gate-application / cost / gradient path.
//
// Register-tile engine (N <= 10 qubits): one wavefront holds a slab of CW columns of
// the d x m matrix (d = 2^N); each lane holds NR = 2^K complex entries (K = min(N,4)
// register qubits) of one column, L = 2^(N-K) lanes share a column.  Gates whose
// target is a register qubit are pure in-register FMA; the host-compiled program
// inserts REMAP ops (an LDS round trip inside the wavefront) so every target is a
// register qubit when its gate runs.  The whole circuit is applied with one HBM read
// of the input slab and one write (apply_to) or none (cost/gradient: only diagonal
// entries and per-gate 2x2 correlations leave the wavefront).
//
// Streaming kernels (any N): one launch per gate, a coalesced pass over the paired
// rows of every matrix in HBM -- the unfused path the reference's pair-row kernel
// (apply_kernel_to_input.cpp:35-115) describes.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

#include "sqh_internal.h"

namespace sqh {

template <typename R>
struct alignas(2 * sizeof(R)) cpx {
    R re, im;
};

template <typename R>
__device__ __forceinline__ cpx<R> mk(R re, R im) {
    cpx<R> r;
    r.re = re;
    r.im = im;
    return r;
}

// k0*a + k1*b, complex, fused multiply-adds.
template <typename R>
__device__ __forceinline__ cpx<R> mac2(cpx<R> k0, cpx<R> a, cpx<R> k1, cpx<R> b) {
    cpx<R> r;
    r.re = fma(k0.re, a.re, fma(-k0.im, a.im, fma(k1.re, b.re, -k1.im * b.im)));
    r.im = fma(k0.re, a.im, fma(k0.im, a.re, fma(k1.re, b.im, k1.im * b.re)));
    return r;
}

// acc += x*y (complex).
template <typename R>
__device__ __forceinline__ void cacc(cpx<R>& acc, cpx<R> x, cpx<R> y) {
    acc.re = fma(x.re, y.re, fma(-x.im, y.im, acc.re));
    acc.im = fma(x.re, y.im, fma(x.im, y.re, acc.im));
}

template <int N>
struct Geo {
    static constexpr int K = N < 4 ? N : 4;   // register qubits
    static constexpr int NR = 1 << K;         // complex registers per lane
    static constexpr int L = 1 << (N - K);    // lanes per column
    static constexpr int CW = 64 / L;         // columns per wavefront
    static constexpr int D = 1 << N;
    static constexpr int LDS = (N > K) ? D * CW : 1;
};

// ---------------------------------------------------------------------------
// Gate kernels from raw parameters (device restatement of the 2x2 builders:
// gate_kernel_templates.h:193-315, 509-601; U3.cpp:89-129; RX/RY/RZ/CRY.cpp).
// ---------------------------------------------------------------------------
__device__ inline void dev_gate_kernel(int type, const double* p, cpx<double> k[4]) {
    double s, c;
    switch (type) {
    case GT_U3: {
        double sp, cp, sl, cl;
        sincos(p[0], &s, &c);
        sincos(p[1], &sp, &cp);
        sincos(p[2], &sl, &cl);
        const double spl = sp * cl + cp * sl, cpl = cp * cl - sp * sl;
        k[0] = mk(c, 0.0);
        k[1] = mk(-s * cl, -s * sl);
        k[2] = mk(s * cp, s * sp);
        k[3] = mk(c * cpl, c * spl);
        return;
    }
    case GT_RX:
        sincos(p[0], &s, &c);
        k[0] = mk(c, 0.0); k[1] = mk(0.0, -s); k[2] = mk(0.0, -s); k[3] = mk(c, 0.0);
        return;
    case GT_RY:
    case GT_CRY:
        sincos(p[0], &s, &c);
        k[0] = mk(c, 0.0); k[1] = mk(-s, 0.0); k[2] = mk(s, 0.0); k[3] = mk(c, 0.0);
        return;
    case GT_RZ:
        sincos(p[0], &s, &c);
        k[0] = mk(c, -s); k[1] = mk(0.0, 0.0); k[2] = mk(0.0, 0.0); k[3] = mk(c, s);
        return;
    default:
        k[0] = k[1] = k[2] = k[3] = mk(0.0, 0.0);
    }
}

// Derivative w.r.t. raw parameter pidx.
__device__ inline void dev_deriv_kernel(int type, const double* p, int pidx, cpx<double> k[4]) {
    double s, c;
    switch (type) {
    case GT_U3: {
        double sp, cp, sl, cl;
        sincos(p[0], &s, &c);
        sincos(p[1], &sp, &cp);
        sincos(p[2], &sl, &cl);
        const double spl = sp * cl + cp * sl, cpl = cp * cl - sp * sl;
        if (pidx == 0) {
            k[0] = mk(-s, 0.0); k[1] = mk(-c * cl, -c * sl);
            k[2] = mk(c * cp, c * sp); k[3] = mk(-s * cpl, -s * spl);
        } else if (pidx == 1) {
            k[0] = mk(0.0, 0.0); k[1] = mk(0.0, 0.0);
            k[2] = mk(-s * sp, s * cp); k[3] = mk(-c * spl, c * cpl);
        } else {
            k[0] = mk(0.0, 0.0); k[1] = mk(s * sl, -s * cl);
            k[2] = mk(0.0, 0.0); k[3] = mk(-c * spl, c * cpl);
        }
        return;
    }
    case GT_RX:
        sincos(p[0], &s, &c);
        k[0] = mk(-s, 0.0); k[1] = mk(0.0, -c); k[2] = mk(0.0, -c); k[3] = mk(-s, 0.0);
        return;
    case GT_RY:
    case GT_CRY:
        sincos(p[0], &s, &c);
        k[0] = mk(-s, 0.0); k[1] = mk(-c, 0.0); k[2] = mk(c, 0.0); k[3] = mk(-s, 0.0);
        return;
    case GT_RZ:
        sincos(p[0], &s, &c);
        k[0] = mk(-s, -c); k[1] = mk(0.0, 0.0); k[2] = mk(0.0, 0.0); k[3] = mk(-s, c);
        return;
    default:
        k[0] = k[1] = k[2] = k[3] = mk(0.0, 0.0);
    }
}

template <typename R>
__device__ __forceinline__ cpx<R> cvt(cpx<double> a) {
    return mk<R>((R)a.re, (R)a.im);
}

// ---------------------------------------------------------------------------
// Register-tile primitives
// ---------------------------------------------------------------------------
// One switch arm per (op kind, local target): case KIND*4 + T runs STMT with a
// compile-time T.  A single jump table replaces nested dispatch.
#define SQH_ARMS(KIND, K, STMT)                                              \
    case (KIND) * 4 + 0: { if constexpr ((K) > 0) { constexpr int T = 0; STMT; } } break; \
    case (KIND) * 4 + 1: { if constexpr ((K) > 1) { constexpr int T = 1; STMT; } } break; \
    case (KIND) * 4 + 2: { if constexpr ((K) > 2) { constexpr int T = 2; STMT; } } break; \
    case (KIND) * 4 + 3: { if constexpr ((K) > 3) { constexpr int T = 3; STMT; } } break;

// Mask of registers whose row has the control bit set (bit i <-> register i).
template <int N>
__device__ __forceinline__ uint32_t ctrl_mask(int cmode, int c, int s) {
    constexpr int NR = Geo<N>::NR;
    constexpr uint32_t all = (NR >= 32) ? 0xffffffffu : ((1u << NR) - 1u);
    if (cmode == CM_NONE) return all;
    if (cmode == CM_LANE) return ((s >> c) & 1) ? all : 0u;
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < NR; ++i) m |= ((uint32_t)((i >> c) & 1)) << i;
    return m;
}

// v <- K v on the register pairs (i, i | 2^T) whose bit i is set in mask; with
// DERIV, pairs outside the mask are zeroed (apply_kernel_to_input.cpp:93-97).
template <int T, bool MASKED, bool DERIV, typename R, int NR>
__device__ __forceinline__ void pairs_apply(cpx<R> (&v)[NR], const cpx<R> k[4], uint32_t mask) {
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        if (i & (1 << T)) continue;
        const int j = i | (1 << T);
        if (MASKED && !((mask >> i) & 1u)) {
            if (DERIV) {
                v[i] = mk<R>(0, 0);
                v[j] = mk<R>(0, 0);
            }
            continue;
        }
        const cpx<R> a = v[i], b = v[j];
        v[i] = mac2(k[0], a, k[1], b);
        v[j] = mac2(k[2], a, k[3], b);
    }
}

// Kernel forms whose zero entries are known from the gate type (the terms they would
// contribute are exact zeros): U3 has a real (0,0) entry; RY / CRY are real; RZ is
// diagonal; RX has a real diagonal and an imaginary off-diagonal.  K^dagger and K^T
// keep the form.
enum KForm : int { KF_GEN = 0, KF_U3 = 1, KF_REAL = 2, KF_DIAG = 3, KF_RX = 4 };

template <int FORM, typename R>
__device__ __forceinline__ void form_mac(const cpx<R> k[4], cpx<R> a, cpx<R> b, cpx<R>& r0, cpx<R>& r1) {
    if constexpr (FORM == KF_U3) {
        r0.re = fma(k[0].re, a.re, fma(k[1].re, b.re, -k[1].im * b.im));
        r0.im = fma(k[0].re, a.im, fma(k[1].re, b.im, k[1].im * b.re));
        r1 = mac2(k[2], a, k[3], b);
    } else if constexpr (FORM == KF_REAL) {
        r0.re = fma(k[0].re, a.re, k[1].re * b.re);
        r0.im = fma(k[0].re, a.im, k[1].re * b.im);
        r1.re = fma(k[2].re, a.re, k[3].re * b.re);
        r1.im = fma(k[2].re, a.im, k[3].re * b.im);
    } else if constexpr (FORM == KF_DIAG) {
        r0.re = fma(k[0].re, a.re, -k[0].im * a.im);
        r0.im = fma(k[0].re, a.im, k[0].im * a.re);
        r1.re = fma(k[3].re, b.re, -k[3].im * b.im);
        r1.im = fma(k[3].re, b.im, k[3].im * b.re);
    } else if constexpr (FORM == KF_RX) {
        r0.re = fma(k[0].re, a.re, -k[1].im * b.im);
        r0.im = fma(k[0].re, a.im, k[1].im * b.re);
        r1.re = fma(k[3].re, b.re, -k[2].im * a.im);
        r1.im = fma(k[3].re, b.im, k[2].im * a.re);
    } else {
        r0 = mac2(k[0], a, k[1], b);
        r1 = mac2(k[2], a, k[3], b);
    }
}

// pairs_apply with a kernel of known form.
template <int FORM, int T, bool MASKED, typename R, int NR>
__device__ __forceinline__ void pairs_apply_f(cpx<R> (&v)[NR], const cpx<R> k[4], uint32_t mask) {
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        if (i & (1 << T)) continue;
        const int j = i | (1 << T);
        if (MASKED && !((mask >> i) & 1u)) continue;
        form_mac<FORM, R>(k, v[i], v[j], v[i], v[j]);
    }
}

template <int T, typename R, int NR>
__device__ __forceinline__ void pairs_swap(cpx<R> (&v)[NR], uint32_t mask) {
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        if (i & (1 << T)) continue;
        const int j = i | (1 << T);
        if (!((mask >> i) & 1u)) continue;
        const cpx<R> a = v[i];
        v[i] = v[j];
        v[j] = a;
    }
}

template <int T, typename R, int NR>
__device__ __forceinline__ void pairs_negate(cpx<R> (&v)[NR], uint32_t mask) {
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        if (i & (1 << T)) continue;
        const int j = i | (1 << T);
        if (!((mask >> i) & 1u)) continue;
        v[j].re = -v[j].re;
        v[j].im = -v[j].im;
    }
}

// Derivative of a controlled gate: its control-0 pairs (mask bit clear) become zero
// when z (apply_kernel_to_input.cpp:93-97).
template <int T, typename R, int NR>
__device__ __forceinline__ void pairs_zero_off(cpx<R> (&v)[NR], uint32_t mask, bool z) {
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        if (i & (1 << T)) continue;
        const int j = i | (1 << T);
        if ((mask >> i) & 1u) continue;
        if (z) {
            v[i] = mk<R>(0, 0);
            v[j] = mk<R>(0, 0);
        }
    }
}

// Correlation of the adjoint tile Y with the state X on the active pairs of target T:
// acc[a*2+b] += sum Y[r_a] X[r_b].
template <int T, typename R, int NR>
__device__ __forceinline__ void pairs_corr(const cpx<R> (&y)[NR], const cpx<R> (&x)[NR], cpx<R> acc[4],
                                           uint32_t mask, bool masked) {
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        if (i & (1 << T)) continue;
        const int j = i | (1 << T);
        if (masked && !((mask >> i) & 1u)) continue;
        cacc(acc[0], y[i], x[i]);
        cacc(acc[1], y[i], x[j]);
        cacc(acc[2], y[j], x[i]);
        cacc(acc[3], y[j], x[j]);
    }
}

template <int N>
__device__ __forceinline__ int sub_off(const Layout* l, int s) {
    int r = 0;
#pragma unroll
    for (int j = 0; j < N - Geo<N>::K; ++j) r |= ((s >> j) & 1) << l->sub_pos[j];
    return r;
}

template <int N>
__device__ __forceinline__ int lds_idx(int row, int cc) {
    return (row ^ ((row >> 4) & 15)) * Geo<N>::CW + cc;
}

// Exchange the slab through LDS from layout lo to layout ln (one wavefront).
template <int N, typename R>
__device__ __forceinline__ void remap(cpx<R> (&v)[Geo<N>::NR], const Layout* lo, const Layout* ln, int s,
                                      int cc, cpx<R>* lds) {
    if constexpr (N > Geo<N>::K) {
        const int so = sub_off<N>(lo, s), sn = sub_off<N>(ln, s);
#pragma unroll
        for (int i = 0; i < Geo<N>::NR; ++i) lds[lds_idx<N>(so + lo->reg_off[i], cc)] = v[i];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < Geo<N>::NR; ++i) v[i] = lds[lds_idx<N>(sn + ln->reg_off[i], cc)];
        __syncthreads();
    }
}

// Forward execution of one gate op with its 2x2 kernel k.
template <int N, bool DERIV, typename R>
__device__ __forceinline__ void exec_gate(cpx<R> (&v)[Geo<N>::NR], const Op op, const cpx<R> k[4], int s) {
    constexpr int K = Geo<N>::K;
    switch (op.kind * 4 + op.t) {
        SQH_ARMS(OP_U, K, (pairs_apply<T, false, false>(v, k, 0u)))
        SQH_ARMS(OP_CU, K, (pairs_apply<T, true, DERIV>(v, k, ctrl_mask<N>(op.cmode, op.c, s))))
        SQH_ARMS(OP_X, K, (pairs_swap<T>(v, ctrl_mask<N>(op.cmode, op.c, s))))
        SQH_ARMS(OP_Z, K, (pairs_negate<T>(v, ctrl_mask<N>(op.cmode, op.c, s))))
    default:
        break;
    }
}

// Kernel-table slot of an op (0 for ops without a kernel, so a prefetch of it is
// always in bounds).  Callers clamp look-ahead positions to the padding.
__device__ __forceinline__ int kslot(const Op op) { return op.kind <= OP_CU ? op.idx : 0; }

// Software-pipelined walk over a program: the op record two ahead and the kernel
// of the next op are loaded while the current op computes.
// LDSOPS: the op records come from LDS (a wavefront's staged copy), so each one is
// made wave-uniform with readfirstlane before it drives the dispatch.
template <typename R, bool LDSOPS = false>
struct OpStream {
    const Op* ops;
    const cpx<R>* kb;
    int nops;
    Op cur, nxt;
    cpx<R> k[4];
    __device__ __forceinline__ OpStream(const Op* ops_, const cpx<R>* kb_, int nops_, int first, int step)
        : ops(ops_), kb(kb_), nops(nops_) {
        cur = at(first);
        nxt = at(first + step);
        load(kslot(cur), k);
    }
    // ops[-kOpPad .. nops + kOpPad) are readable (no-op padding on both sides).
    __device__ __forceinline__ Op at(int p) const {
        Op o = ops[p];
        if constexpr (LDSOPS) {
            uint32_t w[2];
            __builtin_memcpy(w, &o, sizeof(w));
            w[0] = (uint32_t)__builtin_amdgcn_readfirstlane((int)w[0]);
            w[1] = (uint32_t)__builtin_amdgcn_readfirstlane((int)w[1]);
            __builtin_memcpy(&o, w, sizeof(w));
        }
        return o;
    }
    __device__ __forceinline__ void load(int slot, cpx<R> out[4]) const {
        const cpx<R>* kp = kb + 4 * slot;
        out[0] = kp[0];
        out[1] = kp[1];
        out[2] = kp[2];
        out[3] = kp[3];
    }
    // Issue the loads for the following ops; returns the kernel of the next op.
    __device__ __forceinline__ void prefetch(int p_after_next, cpx<R> knext[4], Op& after) const {
        after = at(p_after_next);
        load(kslot(nxt), knext);
    }
};

template <typename R>
__device__ __forceinline__ void load_k(const cpx<R>* kp, cpx<R> k[4]) {
    k[0] = kp[0];
    k[1] = kp[1];
    k[2] = kp[2];
    k[3] = kp[3];
}

struct LaneInfo {
    int s, cc, b, col;
    bool valid;   // the lane holds a real column of a real batch element
    bool bvalid;  // the lane's batch index is in range (before clamping)
};

// UNI: every lane of the wavefront belongs to one batch element (m_pad is a multiple
// of CW), so the batch index is made wave-uniform and per-element tables are read
// through the scalar cache.
template <int N, bool UNI = false>
__device__ __forceinline__ LaneInfo lane_info(int m, int m_pad, int B, int lane, long long wave) {
    constexpr int CW = Geo<N>::CW;
    LaneInfo li;
    li.s = lane / CW;
    li.cc = lane % CW;
    const long long item = wave * CW + li.cc;
    li.b = (int)(item / m_pad);
    li.col = (int)(item % m_pad);
    li.bvalid = li.b < B;
    li.valid = li.bvalid && (li.col < m);
    if (!li.bvalid) li.b = B - 1;
    if constexpr (UNI) li.b = __builtin_amdgcn_readfirstlane(li.b);
    return li;
}

// Trace channels of column li.col from the NR rows this lane holds (row so +
// lay->reg_off[i] in register i), tr = li.col + trace offset.  Without extra
// channels only the lane holding row tr writes; with them each lane sums its rows
// by popcount(row ^ tr), the column's L lanes (lane bits above log2 CW) combine
// their sums, and the sub-0 lane writes channels 0..nch.
template <int N, typename R>
__device__ __forceinline__ void emit_trace(const cpx<R>* v, const Layout* lay, int so, const LaneInfo& li, int tr,
                                           int m, double* __restrict__ diag, const TraceSpec& ts) {
    using G = Geo<N>;
    const long long at = (long long)li.b * m + li.col;
    if (ts.nch == 0) {
        // at most one register holds row tr: pick it with selects and store once
        // (a conditional store per register invites the compiler to sink earlier work
        // into the branches)
        R dre = 0, dim = 0;
        bool hit = false;
#pragma unroll
        for (int i = 0; i < G::NR; ++i) {
            const bool m = so + lay->reg_off[i] == tr;
            dre = m ? v[i].re : dre;
            dim = m ? v[i].im : dim;
            hit = hit || m;
        }
        if (li.valid && hit) {
            diag[2 * at] = (double)dre;
            diag[2 * at + 1] = (double)dim;
        }
        return;
    }
    double a[3][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
    for (int i = 0; i < G::NR; ++i) {
        const int pc = __popc((unsigned)((so + lay->reg_off[i]) ^ tr));
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (li.valid && pc == k && k <= ts.nch) {
                a[k][0] += (double)v[i].re;
                a[k][1] += (double)v[i].im;
            }
    }
#pragma unroll
    for (int o = G::CW; o < 64; o <<= 1)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            a[k][0] += __shfl_xor(a[k][0], o, 64);
            a[k][1] += __shfl_xor(a[k][1], o, 64);
        }
    if (li.valid && li.s == 0)
        for (int k = 0; k <= ts.nch; ++k) {
            diag[2 * (k * ts.chstride + at)] = a[k][0];
            diag[2 * (k * ts.chstride + at) + 1] = a[k][1];
        }
}

// Adjoint seed of column li.col: weight w[popcount(row ^ tr)] on the rows within
// nch bit flips of tr (w = (1, w1, w2), or the element's weights from ts.seed).
template <int N, typename R>
__device__ __forceinline__ void seed_trace(cpx<R>* y, const Layout* lay, int so, const LaneInfo& li, int tr,
                                           const TraceSpec& ts) {
    using G = Geo<N>;
    if (ts.nch == 0) {
#pragma unroll
        for (int i = 0; i < G::NR; ++i) {
            const bool hit = li.valid && so + lay->reg_off[i] == tr;
            y[i] = mk<R>(hit ? (R)1 : (R)0, (R)0);
        }
        return;
    }
    double w[3][2] = {{1.0, 0.0}, {ts.w1, 0.0}, {ts.w2, 0.0}};
    if (ts.seed) {
        const double* sw = ts.seed + 6 * (long long)li.b;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            w[k][0] = sw[2 * k];
            w[k][1] = sw[2 * k + 1];
        }
    }
#pragma unroll
    for (int i = 0; i < G::NR; ++i) {
        const int pc = __popc((unsigned)((so + lay->reg_off[i]) ^ tr));
        y[i] = mk<R>((R)0, (R)0);
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (li.valid && pc == k && k <= ts.nch) y[i] = mk<R>((R)w[k][0], (R)w[k][1]);
    }
}

// ---------------------------------------------------------------------------
// Forward tile kernel: apply the whole program, then either store the slab
// (Gates_block::apply_to) or emit the diagonal entries the trace needs.
// ---------------------------------------------------------------------------
// WPG > 1: a workgroup of WPG wavefronts that share one batch element (requires
// UNI and waves-per-element divisible by WPG); the element's kernel table (ns
// entries) is staged in LDS after the WPG remap buffers (dynamic shared memory).
template <int N, typename R, int MODE, bool UNI, int WPG>
__global__ __launch_bounds__(64 * WPG) void k_tile_forward(const cpx<R>* in, long long in_bstride,
                                                           int in_stride, cpx<R>* out,
                                                           long long out_bstride, int out_stride, int m, int m_pad,
                                                           int B, const Op* __restrict__ ops, int nops,
                                                           const Layout* __restrict__ lays,
                                                           const cpx<R>* __restrict__ kern, long long kern_bstride,
                                                           int trace_off, double* __restrict__ diag, int ns,
                                                           TraceSpec ts, const int* __restrict__ zslot,
                                                           const int* __restrict__ in_index) {
    using G = Geo<N>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<R>* smem = reinterpret_cast<cpx<R>*>(smem_raw);
    const int wave = (WPG > 1) ? (int)(threadIdx.x >> 6) : 0;
    cpx<R>* lds = smem + wave * G::LDS;
    if (((long long)blockIdx.x * WPG + wave) * G::CW / m_pad >= B) return;  // wavefront past the batch
    const LaneInfo li = lane_info<N, UNI>(m, m_pad, B, threadIdx.x & 63, (long long)blockIdx.x * WPG + wave);
    cpx<R> v[G::NR];
    {
        // input of element b: in + b * in_bstride, or in + in_index[b] * in_bstride
        const cpx<R>* src = in + (long long)(in_index ? in_index[li.b] : li.b) * in_bstride;
        const int so = sub_off<N>(&lays[0], li.s);
#pragma unroll
        for (int i = 0; i < G::NR; ++i) {
            const int row = so + lays[0].reg_off[i];
            v[i] = li.valid ? src[(long long)row * in_stride + li.col] : mk<R>(0, 0);
        }
    }
    const cpx<R>* kb = kern + (long long)li.b * kern_bstride;
    if constexpr (WPG > 1) {
        cpx<R>* ktab = smem + WPG * G::LDS;
        for (int i = threadIdx.x; i < 4 * ns; i += 64 * WPG) ktab[i] = kb[i];
        __syncthreads();
        kb = ktab;
    }
    int cur = 0;
    // derivative mode of one controlled slot per element (apply_derivate_to tables)
    const int zs = zslot ? zslot[li.b] : -1;
    OpStream<R> os(ops, kb, nops, 0, 1);
    for (int p = 0; p < nops; ++p) {
        cpx<R> kn[4];
        Op after;
        os.prefetch(p + 2, kn, after);
        const Op op = os.cur;
        if (op.kind == OP_REMAP) {
            const int o = op.idx >> 16, n = op.idx & 0xffff;
            remap<N, R>(v, &lays[o], &lays[n], li.s, li.cc, lds);
            cur = n;
        } else {
            exec_gate<N, false, R>(v, op, os.k, li.s);
            if (op.kind == OP_CU && zslot) {
                const uint32_t msk = ctrl_mask<N>(op.cmode, op.c, li.s);
                switch (op.t) {
                case 0: if constexpr (G::K > 0) pairs_zero_off<0>(v, msk, zs == op.idx); break;
                case 1: if constexpr (G::K > 1) pairs_zero_off<1>(v, msk, zs == op.idx); break;
                case 2: if constexpr (G::K > 2) pairs_zero_off<2>(v, msk, zs == op.idx); break;
                case 3: if constexpr (G::K > 3) pairs_zero_off<3>(v, msk, zs == op.idx); break;
                default: break;
                }
            }
        }
        os.cur = os.nxt;
        os.nxt = after;
#pragma unroll
        for (int e = 0; e < 4; ++e) os.k[e] = kn[e];
    }
    const int so = sub_off<N>(&lays[cur], li.s);
    if constexpr (MODE == 0) {
        cpx<R>* dst = out + (long long)li.b * out_bstride;
#pragma unroll
        for (int i = 0; i < G::NR; ++i) {
            const int row = so + lays[cur].reg_off[i];
            if (li.valid) dst[(long long)row * out_stride + li.col] = v[i];
        }
    } else {
        emit_trace<N, R>(v, &lays[cur], so, li, li.col + trace_off, m, diag, ts);
    }
}

// One DPP move of a double (two 32-bit halves); rows / banks outside the masks read 0.
template <int CTRL, int ROWS = 0xf, int BANKS = 0xf>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWS, BANKS, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, BANKS, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// Sum over all 64 lanes of the wavefront with DPP row operations (no LDS round
// trips): pairs within quads, half rows, rows, then rows 0+1 into 1 and 2+3 into 3
// by row broadcasts, and 0..2 into 3; the total lands in lane 63 and is read back
// into every lane.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_f64<0xb1>(v);         // quad_perm [1,0,3,2]
    v += dpp_f64<0x4e>(v);         // quad_perm [2,3,0,1]
    v += dpp_f64<0x141>(v);        // row_half_mirror
    v += dpp_f64<0x140>(v);        // row_mirror
    v += dpp_f64<0x142, 0xa>(v);   // row_bcast:15 -> rows 1, 3
    v += dpp_f64<0x143, 0xc>(v);   // row_bcast:31 -> rows 2, 3
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)b, 63);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// Segmented butterfly sum over the lanes that share a batch element: lanes
// lane ^ mk for mk < lim (columns of the element in the wavefront) and mk >= CW (the
// lanes of one column).  A wavefront of whole columns (CW = 64) with segments of up
// to 16 lanes, or one segment of 64, is summed with DPP row operations instead of
// LDS permutes (segments are aligned, so quad / half-row / row sums are exactly them).
template <int CW, typename T>
__device__ __forceinline__ T seg_sum(T x, int lim) {
    if constexpr (CW == 64 && sizeof(T) == 8) {
        if (lim == 64) return wave_sum(x);
        if (lim <= 16) {
            if (lim >= 2) x += dpp_f64<0xb1>(x);   // quad_perm [1,0,3,2]
            if (lim >= 4) x += dpp_f64<0x4e>(x);   // quad_perm [2,3,0,1]
            if (lim >= 8) x += dpp_f64<0x141>(x);  // row_half_mirror
            if (lim >= 16) x += dpp_f64<0x140>(x); // row_mirror
            return x;
        }
    }
#pragma unroll
    for (int mk = 1; mk < 64; mk <<= 1) {
        if (mk < lim || mk >= CW) x += __shfl_xor(x, mk, 64);
    }
    return x;
}

// ---------------------------------------------------------------------------
// Adjoint cost+gradient tile kernel.  Forward pass to X_N = C.U (diagonal written
// out), then with Y = E (the trace selector) walk the gates backwards:
//   X_j = G_j^dagger X_{j+1};  M_j += sum_pairs Y[r_a] X_j[r_b];  Y <- G_j^T Y.
// dTr/dp = sum_ab D_ab M_ab for each parameter's derivative kernel D, which the
// finish kernel evaluates.  This replaces the reference's per-parameter derivative
// matrices (Gates_block.cpp:1012-1150) with O(G d m) work.
// ---------------------------------------------------------------------------
// With mpart null the kernel stops after the forward pass and the trace channels.
// STAGE (with UNI): the wavefront first copies its program (ops with their padding,
// layouts) and its element's kernel table into LDS (dynamic shared memory, sized by
// the host for the largest program of the launch: ops_cap ops, lays_cap layouts),
// so the op walk waits on LDS instead of on scalar loads from L2/HBM.
template <int N, typename R, bool UNI, bool STAGE = false>
__global__ __launch_bounds__(64) void k_tile_grad(const cpx<R>* __restrict__ in, int in_stride, int m, int m_pad,
                                                  int B, const Op* __restrict__ ops, int nops,
                                                  const Layout* __restrict__ lays, const cpx<R>* __restrict__ kern,
                                                  long long kern_bstride, int trace_off, int ns,
                                                  double* __restrict__ diag, double* __restrict__ mpart,
                                                  BatchIndex bi, int ops_cap, int lays_cap, TraceSpec ts) {
    using G = Geo<N>;
    constexpr int CW = G::CW;
    __shared__ cpx<R> lds[G::LDS];
    if (bi.count) {
        B = *bi.count;
        if (B <= 0) return;
    }
    if ((long long)blockIdx.x * CW / m_pad >= B) return;  // whole wavefront past the live elements
    const LaneInfo li = lane_info<N, UNI>(m, m_pad, B, threadIdx.x, blockIdx.x);
    if (bi.count) in += (long long)bi.run_target[bi.elem_run[li.b]] * bi.tgt_stride;
    int nlays = lays_cap;
    if constexpr (UNI) {
        if (bi.run_prog) {  // this wavefront's run has its own circuit
            const ProgRef pg = bi.run_prog[bi.elem_run[li.b]];
            ops += pg.op_off;
            nops = pg.nops;
            lays += pg.lay_off;
            nlays = pg.nlays;
        }
    }
    const cpx<R>* kb = kern + (long long)li.b * kern_bstride;
    if constexpr (STAGE && UNI) {
        extern __shared__ __align__(16) unsigned char stage_raw[];
        cpx<R>* s_kern = reinterpret_cast<cpx<R>*>(stage_raw);
        Layout* s_lays = reinterpret_cast<Layout*>(s_kern + 4 * (size_t)ns);
        Op* s_ops = reinterpret_cast<Op*>(s_lays + lays_cap);
        const int lane = threadIdx.x;
        for (int i = lane; i < 4 * ns; i += 64) s_kern[i] = kb[i];
        {
            const int4* src = reinterpret_cast<const int4*>(lays);
            int4* dst = reinterpret_cast<int4*>(s_lays);
            const int n4 = nlays * (int)(sizeof(Layout) / sizeof(int4));
            for (int i = lane; i < n4; i += 64) dst[i] = src[i];
        }
        for (int i = lane; i < nops + 2 * kOpPad; i += 64) s_ops[i] = ops[i - kOpPad];
        __syncthreads();
        ops = s_ops + kOpPad;
        lays = s_lays;
        kb = s_kern;
    }
    (void)ops_cap;
    cpx<R> x[G::NR], y[G::NR];
    {
        const int so = sub_off<N>(&lays[0], li.s);
#pragma unroll
        for (int i = 0; i < G::NR; ++i) {
            const int row = so + lays[0].reg_off[i];
            x[i] = li.valid ? in[(long long)row * in_stride + li.col] : mk<R>(0, 0);
        }
    }
    int cur = 0;
    {
        OpStream<R, STAGE && UNI> os(ops, kb, nops, 0, 1);
        for (int p = 0; p < nops; ++p) {
            cpx<R> kn[4];
            Op after;
            os.prefetch(p + 2, kn, after);
            const Op op = os.cur;
            if (op.kind == OP_REMAP) {
                const int o = op.idx >> 16, n = op.idx & 0xffff;
                remap<N, R>(x, &lays[o], &lays[n], li.s, li.cc, lds);
                cur = n;
            } else {
                exec_gate<N, false, R>(x, op, os.k, li.s);
            }
            os.cur = os.nxt;
            os.nxt = after;
#pragma unroll
            for (int e = 0; e < 4; ++e) os.k[e] = kn[e];
        }
    }
    // Trace channels of C.U and the selector (adjoint seed) under the final layout.
    {
        const int so = sub_off<N>(&lays[cur], li.s);
        const int tr = li.col + trace_off;
        emit_trace<N, R>(x, &lays[cur], so, li, tr, m, diag, ts);
        if (!mpart) return;  // forward pass only: the trace channels (seed weights come next)
        seed_trace<N, R>(y, &lays[cur], so, li, tr, ts);
    }
    // Which lanes write the per-wave partials.
    const int lim = m_pad < CW ? m_pad : CW;
    const bool leader = li.bvalid && (li.s == 0) && (li.cc % lim == 0);
    const int wpb = m_pad >= CW ? m_pad / CW : 1;
    const long long item0 = (long long)blockIdx.x * CW + li.cc;
    const int wib = m_pad >= CW ? (int)((item0 % m_pad) / CW) : 0;
    double* mp = mpart + ((long long)li.b * wpb + wib) * (long long)ns * 8;

    OpStream<R, STAGE && UNI> os(ops, kb, nops, nops - 1, -1);
    for (int p = nops - 1; p >= 0; --p) {
        cpx<R> kn[4];
        Op after;
        os.prefetch(p - 2, kn, after);
        const Op op = os.cur;
        if (op.kind == OP_REMAP) {
            const int o = op.idx >> 16, n = op.idx & 0xffff;
            remap<N, R>(x, &lays[n], &lays[o], li.s, li.cc, lds);
            remap<N, R>(y, &lays[n], &lays[o], li.s, li.cc, lds);
        } else if (op.kind == OP_X || op.kind == OP_Z) {
            // Self-inverse real symmetric permutation / sign: G^dagger = G^T = G.
            exec_gate<N, false, R>(x, op, os.k, li.s);
            exec_gate<N, false, R>(y, op, os.k, li.s);
        } else {
            const cpx<R>* k = os.k;
            cpx<R> kd[4], kt[4];
            kd[0] = mk(k[0].re, -k[0].im);  // K^dagger
            kd[1] = mk(k[2].re, -k[2].im);
            kd[2] = mk(k[1].re, -k[1].im);
            kd[3] = mk(k[3].re, -k[3].im);
            kt[0] = k[0];                   // K^T
            kt[1] = k[2];
            kt[2] = k[1];
            kt[3] = k[3];
            cpx<R> acc[4] = {mk<R>(0, 0), mk<R>(0, 0), mk<R>(0, 0), mk<R>(0, 0)};
            constexpr int K = G::K;
            switch (op.kind * 4 + op.t) {
                SQH_ARMS(OP_U, K, (pairs_apply<T, false, false>(x, kd, 0u), pairs_corr<T>(y, x, acc, 0u, false),
                                   pairs_apply<T, false, false>(y, kt, 0u)))
                SQH_ARMS(OP_CU, K,
                         (pairs_apply<T, true, false>(x, kd, ctrl_mask<N>(op.cmode, op.c, li.s)),
                          pairs_corr<T>(y, x, acc, ctrl_mask<N>(op.cmode, op.c, li.s), true),
                          pairs_apply<T, true, false>(y, kt, ctrl_mask<N>(op.cmode, op.c, li.s))))
            default:
                break;
            }
            double r[8];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if constexpr (UNI) {  // one element per wavefront: a full-wave sum
                    r[2 * e] = wave_sum((double)acc[e].re);
                    r[2 * e + 1] = wave_sum((double)acc[e].im);
                } else {
                    r[2 * e] = seg_sum<CW>((double)acc[e].re, lim);
                    r[2 * e + 1] = seg_sum<CW>((double)acc[e].im, lim);
                }
            }
            if (leader) {
                double* d = mp + (long long)op.idx * 8;
#pragma unroll
                for (int e = 0; e < 8; ++e) d[e] = r[e];
            }
        }
        os.cur = os.nxt;
        os.nxt = after;
#pragma unroll
        for (int e = 0; e < 4; ++e) os.k[e] = kn[e];
    }
}

// ---------------------------------------------------------------------------
// Small kernels: kernel table, finish (cost + gradient), streaming gates.
// ---------------------------------------------------------------------------
template <typename R>
__global__ void k_build_kernels(const double* __restrict__ params, int P, int B, const Slot* __restrict__ slots,
                                int ns, cpx<R>* __restrict__ kern, BatchIndex bi) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (bi.count) B = *bi.count;
    if (i >= (long long)B * ns) return;
    const int b = (int)(i / ns), sl = (int)(i % ns);
    const int r = bi.count ? bi.elem_run[b] : b;
    if (bi.run_prog) {  // ns is the table stride; the run has pg.ns slots of its own
        const ProgRef pg = bi.run_prog[r];
        if (sl >= pg.ns) return;
        slots += pg.slot_off;
    }
    const Slot s = slots[sl];
    cpx<double> k[4];
    dev_gate_kernel(s.type, params + (long long)r * P + s.param_start, k);
    cpx<R>* o = kern + i * 4;
#pragma unroll
    for (int e = 0; e < 4; ++e) o[e] = cvt<R>(k[e]);
}

__device__ inline double block_sum(double v, double* sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    const int nw = blockDim.x / 64;
    for (int i = 0; i < nw; ++i) t += sh[i];
    return t;
}

// Cost variants (N_Qubit_Decomposition_Cost_Function.cpp:137-158, 524-549) and the
// gradient formulas of optimization_problem_combined (Optimization_Interface.cpp:1401-1450).
__device__ inline double cost_of(int variant, double tre, double tim, double m) {
    switch (variant) {
    case CV_FROBENIUS: return 1.0 - tre / m;
    case CV_HS: { const double d = 1.0 / m; return 1.0 - d * d * (tre * tre + tim * tim); }
    case CV_INFIDELITY: return 1.0 - ((tre * tre + tim * tim) / m + 1.0) / (m + 1.0);
    case CV_ABS_TRACE: return 1.0 - sqrt(tre * tre + tim * tim) / m;
    default: return __builtin_nan("");
    }
}

__device__ inline double grad_of(int variant, double tre, double tim, double kre, double kim, double m) {
    switch (variant) {
    case CV_FROBENIUS: return -kre / m;
    case CV_HS: { const double d = 1.0 / m; return -2.0 * d * d * tre * kre - 2.0 * d * d * tim * kim; }
    case CV_INFIDELITY: return -2.0 / m / (m + 1.0) * tre * kre - 2.0 / m / (m + 1.0) * tim * kim;
    case CV_ABS_TRACE: {
        const double a = sqrt(tre * tre + tim * tim);
        return a > 0.0 ? -(tre * kre + tim * kim) / (a * m) : 0.0;
    }
    default: return __builtin_nan("");
    }
}

// The correction variants (Optimization_Interface.cpp:676-716, 1391-1420) from the
// channel traces t[k] = (re, im): FROBENIUS_C* = 1 - (Re t0 + w1 Re t1 + w2 Re t2)/m
// (get_cost_function_with_correction[2]: N_Qubit_Decomposition_Cost_Function.cpp:191-380),
// HS_C* = 1 - (|t0|^2 + w1 |t1|^2 + w2 |t2|^2)/m^2 (get_trace_with_correction[2]: :564-664),
// w_k = sqrt(previous cost) * correction_k scale, w2 = 0 for the C1 variants.
__device__ inline double cost_corr(int variant, const double t[3][2], double w1, double w2, double m) {
    if (variant == CV_FROBENIUS_C1 || variant == CV_FROBENIUS_C2)
        return 1.0 - (t[0][0] + w1 * t[1][0] + w2 * t[2][0]) / m;
    const double d = 1.0 / m;
    return 1.0 - d * d *
                     (t[0][0] * t[0][0] + t[0][1] * t[0][1] + w1 * (t[1][0] * t[1][0] + t[1][1] * t[1][1]) +
                      w2 * (t[2][0] * t[2][0] + t[2][1] * t[2][1]));
}

// Gradient of a correction variant from kre + i kim = sum_ab D_ab M_ab, where M was
// built with the seed weights: (1, w1, w2) for FROBENIUS_C* (linear in C.U), and
// (conj t0, w1 conj t1, w2 conj t2) for HS_C*, so Re(sum D M) is the weighted sum of
// Re(conj(t_k) dt_k) of the reference's formula.
__device__ inline double grad_corr(int variant, double kre, double m) {
    if (variant == CV_FROBENIUS_C1 || variant == CV_FROBENIUS_C2) return -kre / m;
    return -2.0 / (m * m) * kre;
}

#ifdef SQH_DEFINE_MISC_KERNELS
// Per-parameter kernel tables of apply_derivate_to (one workgroup per parameter k,
// threads over the slots): the slot holding parameter k gets its derivative kernel
// (Gate::apply_derivate_to, Gate.cpp:645-706), every other slot its forward kernel.
__global__ __launch_bounds__(64) void k_deriv_tables(const cpx<double>* __restrict__ base,
                                                     const Slot* __restrict__ slots, int ns,
                                                     const double* __restrict__ params, int P,
                                                     cpx<double>* __restrict__ tables, int* __restrict__ zslot) {
    const int k = blockIdx.x;
    if (k >= P) return;
    cpx<double>* tk = tables + (long long)k * ns * 4;
    for (int j = threadIdx.x; j < ns; j += blockDim.x) {
        const Slot sl = slots[j];
        const int q = k - sl.param_start;
        if (q >= 0 && q < sl.nparam) {
            cpx<double> kd[4];
            dev_deriv_kernel(sl.type, params + sl.param_start, q, kd);
#pragma unroll
            for (int e = 0; e < 4; ++e) tk[4 * j + e] = kd[e];
            zslot[k] = sl.type == GT_CRY ? j : -1;
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) tk[4 * j + e] = base[4 * j + e];
        }
    }
}

// One workgroup per batch element: trace channels = sums of the per-column entries,
// cost, and (if mpart) the gradient from the per-gate correlations; with seed_out,
// also the HS_C* adjoint seed weights of the element.
__global__ __launch_bounds__(256) void k_finish(const double* __restrict__ diag, int m, int variant,
                                                double* __restrict__ costs, const double* __restrict__ mpart,
                                                int wpb, int ns, const Slot* __restrict__ slots,
                                                const double* __restrict__ params, int P,
                                                double* __restrict__ grad, BatchIndex bi, TraceSpec ts,
                                                double* __restrict__ seed_out) {
    __shared__ double sh[8];
    const int b = blockIdx.x;
    if (bi.count && b >= *bi.count) return;
    // Results go to the run's slots; params/costs/grad are indexed by run.
    const int r = bi.count ? bi.elem_run[b] : b;
    double t[3][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    for (int k = 0; k <= ts.nch; ++k) {
        const double* dk = diag + 2 * (k * ts.chstride + (long long)b * m);
        double re = 0.0, im = 0.0;
        for (int c = threadIdx.x; c < m; c += blockDim.x) {
            re += dk[2 * c];
            im += dk[2 * c + 1];
        }
        t[k][0] = block_sum(re, sh);
        t[k][1] = block_sum(im, sh);
    }
    const double tre = t[0][0], tim = t[0][1];
    const bool corr = ts.nch > 0;
    if (threadIdx.x == 0 && costs)
        costs[r] = corr ? cost_corr(variant, t, ts.w1, ts.w2, (double)m) : cost_of(variant, tre, tim, (double)m);
    if (threadIdx.x == 0 && seed_out) {
        const double w[3] = {1.0, ts.w1, ts.w2};
        for (int k = 0; k < 3; ++k) {
            seed_out[6 * (long long)b + 2 * k] = w[k] * t[k][0];
            seed_out[6 * (long long)b + 2 * k + 1] = -w[k] * t[k][1];
        }
    }
    if (!mpart) return;
    int nsr = ns;  // ns: stride of the correlation table
    if (bi.run_prog) {
        const ProgRef pg = bi.run_prog[r];
        slots += pg.slot_off;
        nsr = pg.ns;
    }
    for (int s = threadIdx.x; s < nsr; s += blockDim.x) {
        double M[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int w = 0; w < wpb; ++w) {
            const double* src = mpart + (((long long)b * wpb + w) * ns + s) * 8;
#pragma unroll
            for (int e = 0; e < 8; ++e) M[e] += src[e];
        }
        const Slot sl = slots[s];
        const double* pb = params + (long long)r * P;
        for (int q = 0; q < sl.nparam; ++q) {
            cpx<double> d[4];
            dev_deriv_kernel(sl.type, pb + sl.param_start, q, d);
            double kre = 0.0, kim = 0.0;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                kre += d[e].re * M[2 * e] - d[e].im * M[2 * e + 1];
                kim += d[e].re * M[2 * e + 1] + d[e].im * M[2 * e];
            }
            grad[(long long)r * P + sl.param_start + q] =
                corr ? grad_corr(variant, kre, (double)m) : grad_of(variant, tre, tim, kre, kim, (double)m);
        }
    }
}

#endif  // SQH_DEFINE_MISC_KERNELS

// Streaming one-gate pass over B matrices (rows = 2^n, m columns, row stride).
// Thread -> (matrix b, row pair, column); columns fastest so a wavefront moves
// contiguous row segments.  Pairs whose control bit is clear exit at once, so a
// controlled gate touches only the rows it changes.
template <typename R>
__global__ __launch_bounds__(256) void k_stream_gate(cpx<R>* __restrict__ mats, long long bstride, int stride,
                                                     int rows, int m, int B, int kind, int target, int control,
                                                     const cpx<R>* __restrict__ kern, long long kern_bstride,
                                                     int slot) {
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per_b = (long long)(rows >> 1) * m;
    if (tid >= per_b * B) return;
    const int b = (int)(tid / per_b);
    const long long rem = tid % per_b;
    const int pr = (int)(rem / m), col = (int)(rem % m);
    const int lo = pr & ((1 << target) - 1);
    const int r0 = ((pr >> target) << (target + 1)) | lo;
    const int r1 = r0 | (1 << target);
    if (control >= 0 && !((r0 >> control) & 1)) return;
    cpx<R>* M = mats + (long long)b * bstride;
    cpx<R>* a = M + (long long)r0 * stride + col;
    cpx<R>* c = M + (long long)r1 * stride + col;
    if (kind == OP_Z) {
        cpx<R> y = *c;
        y.re = -y.re;
        y.im = -y.im;
        *c = y;
        return;
    }
    const cpx<R> x0 = *a, x1 = *c;
    if (kind == OP_X) {
        *a = x1;
        *c = x0;
        return;
    }
    const cpx<R>* kp = kern + (long long)b * kern_bstride + 4 * slot;
    const cpx<R> k0 = kp[0], k1 = kp[1], k2 = kp[2], k3 = kp[3];
    *a = mac2(k0, x0, k1, x1);
    *c = mac2(k2, x0, k3, x1);
}

// Trace channels of column c of each matrix (rows = 2^n): channel k sums
// M[(c + off) ^ mask, c] over the masks of k set bits below 2^n, in the order of
// get_cost_function_with_correction2 (N_Qubit_Decomposition_Cost_Function.cpp:291-380).
template <typename R>
__global__ void k_extract_diag(const cpx<R>* __restrict__ mats, long long bstride, int stride, int m, int B,
                               int trace_off, double* __restrict__ diag, int n, TraceSpec ts) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)B * m) return;
    const int b = (int)(i / m), c = (int)(i % m);
    const cpx<R>* M = mats + (long long)b * bstride;
    const int tr = c + trace_off;
    const cpx<R> v = M[(long long)tr * stride + c];
    diag[2 * i] = (double)v.re;
    diag[2 * i + 1] = (double)v.im;
    if (ts.nch >= 1) {
        double re = 0.0, im = 0.0;
        for (int q = 0; q < n; ++q) {
            const cpx<R> u = M[(long long)(tr ^ (1 << q)) * stride + c];
            re += (double)u.re;
            im += (double)u.im;
        }
        diag[2 * (ts.chstride + i)] = re;
        diag[2 * (ts.chstride + i) + 1] = im;
    }
    if (ts.nch >= 2) {
        double re = 0.0, im = 0.0;
        for (int q = 0; q < n - 1; ++q)
            for (int q2 = q + 1; q2 < n; ++q2) {
                const cpx<R> u = M[(long long)(tr ^ (1 << q) ^ (1 << q2)) * stride + c];
                re += (double)u.re;
                im += (double)u.im;
            }
        diag[2 * (2 * ts.chstride + i)] = re;
        diag[2 * (2 * ts.chstride + i) + 1] = im;
    }
}

template <typename R>
__global__ void k_broadcast_copy(const cpx<R>* __restrict__ src, int rows, int m, int sstride,
                                 cpx<R>* __restrict__ dst, long long dbstride, int B) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per = (long long)rows * m;
    if (i >= per * B) return;
    const int b = (int)(i / per);
    const long long r = i % per;
    const int row = (int)(r / m), col = (int)(r % m);
    dst[(long long)b * dbstride + (long long)row * m + col] = src[(long long)row * sstride + col];
}

#ifdef SQH_DEFINE_MISC_KERNELS
__global__ void k_c128_to_c64(const double* __restrict__ src, float* __restrict__ dst, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (float)src[i];
}

__global__ void k_c64_to_c128(const float* __restrict__ src, double* __restrict__ dst, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (double)src[i];
}

#endif  // SQH_DEFINE_MISC_KERNELS

}  // namespace sqh
