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
// natural row order and come back in the next block's grouping.  Of the last block the
// cost needs only row c + trace_off of each column c: a 16-term dot product with that
// row of F_nb per column replaces the block product.  The target
// is read once per wavefront (it is shared by the batch and stays in L2); nothing
// else touches HBM but the block matrices and the per-column trace entries.
//
// Row j of a column under block b: j = base_b[o] + off_b[i] for local index i (bit q of
// i <-> the q-th smallest qubit of the block, as apply_large_kernel_to_input,
// apply_large_kernel_to_input.cpp:82-102) and outer index o (the other qubits in
// ascending order); both tables come from the host per block.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "sqh_device.h"
#include "sqh_fused_chain.h"
#include "sqh_kernels.h"

namespace sqh {

// One workgroup's launch shape (FusedCostCfg): T 16-column tiles per wavefront (2 or
// 4: 32 or 64 vector columns), H exchange groups per block (the tiles go through LDS in
// H independent groups, so the exchange of one overlaps the MFMAs of the next; a group
// must hold whole columns), PF: block b + 1's matrix loaded while block b runs (more
// registers, fewer waves), waves per workgroup at run time.  Every shape performs the
// same operations on every column in the same order: results are bit-identical.
//
// Workgroups are ordered for the XCDs: hardware workgroup i runs on XCD i mod 8, and the
// logical order gives each XCD a contiguous range of (element, column group) work, so
// an element's block matrices and the target stay in one XCD's L2.
template <typename R, int N, int T, int H, bool PF>
__global__ __launch_bounds__(256) void k_fused_cost(FusedCostArgs a) {
    typedef FcMfma<R> MF;
    constexpr int D = 1 << N;
    constexpr int O = D >> 4;      // outer groups per column
    constexpr int W = 16 * T / O;  // columns per wavefront
    constexpr int TH = T / H;      // tiles per exchange group
    static_assert(W >= 1 && (16 * TH) % O == 0, "an exchange group holds whole columns");
    extern __shared__ __align__(16) unsigned char fc_smem[];
    int* smap = reinterpret_cast<int*>(fc_smem);
    const int nmap = a.nb * kFusedMapStride;
    cpx<R>* region = reinterpret_cast<cpx<R>*>(fc_smem + (((size_t)nmap * sizeof(int) + 15) & ~(size_t)15));
    for (int i = threadIdx.x; i < nmap; i += blockDim.x) smap[i] = a.maps[i];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const long long per_xcd = gridDim.x >> 3;  // the grid is a multiple of 8
    const long long lwg = (long long)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    const long long wave = lwg * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (wave >= (long long)a.B * a.groups) return;
    const int e = (int)(wave / a.groups);
    const int c0 = (int)(wave % a.groups) * W;
    cpx<R>* lds = region + (threadIdx.x >> 6) * (16 * T * 16);
    const int nl = lane & 15, g = lane >> 4;
    // vector column 16 t + nl of tile t: column c0 + w[t], outer group o[t]
    int w[T], o[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int n = 16 * t + nl;
        w[t] = n / O;
        o[t] = n % O;
    }
    const cpx<R>* U =
        static_cast<const cpx<R>*>(a.U) + (a.in_index ? (long long)a.in_index[e] : (long long)e) * a.in_bstride;
    const int m = a.m, ist = a.in_stride;
    // B operand of step s, tile t: X[4 s + g][16 t + nl] (block 0's grouping, from U)
    R xr[T][4], xi[T][4];
    {
        const int* mp = smap;
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int col = c0 + w[t];
            const int base = mp[16 + o[t]];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int j = base + mp[4 * s + g];
                const cpx<R> v = col < m ? U[(long long)j * ist + col] : mk<R>(0, 0);
                xr[t][s] = v.re;
                xi[t][s] = v.im;
            }
        }
    }
    const cpx<R>* Fe = static_cast<const cpx<R>*>(a.F) + (long long)e * a.F_bstride;
    // A operand of step s: F_b[nl][4 s + g] (with PF, block b + 1's is loaded while
    // block b runs)
    R far[4], fai[4];
    if constexpr (PF) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const cpx<R> f = Fe[nl * 16 + 4 * s + g];
            far[s] = f.re;
            fai[s] = f.im;
        }
    }
    // the cost pass stops before the last block: of it only row c + trace_off of each
    // column c is needed (below)
    const int nbl = a.out ? a.nb : a.nb - 1;
    for (int b = 0; b < nbl; ++b) {
        R ar[4], ai[4], ain[4];
        if constexpr (PF) {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                ar[s] = far[s];
                ai[s] = fai[s];
                ain[s] = -fai[s];
            }
            if (b + 1 < nbl) {
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const cpx<R> f = Fe[(long long)(b + 1) * 256 + nl * 16 + 4 * s + g];
                    far[s] = f.re;
                    fai[s] = f.im;
                }
            }
        } else {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const cpx<R> f = Fe[(long long)b * 256 + nl * 16 + 4 * s + g];
                ar[s] = f.re;
                ai[s] = f.im;
                ain[s] = -f.im;
            }
        }
        const int* mp = smap + b * kFusedMapStride;
        const int* mn = mp + kFusedMapStride;
#pragma unroll
        for (int h = 0; h < H; ++h) {
            typename MF::V yr[TH], yi[TH];
#pragma unroll
            for (int u = 0; u < TH; ++u) {
                yr[u] = typename MF::V{0, 0, 0, 0};
                yi[u] = typename MF::V{0, 0, 0, 0};
            }
#pragma unroll
            for (int s = 0; s < 4; ++s) {
#pragma unroll
                for (int u = 0; u < TH; ++u) {
                    const int t = h * TH + u;
                    yr[u] = MF::op(ar[s], xr[t][s], yr[u]);
                    yi[u] = MF::op(ar[s], xi[t][s], yi[u]);
                }
#pragma unroll
                for (int u = 0; u < TH; ++u) {
                    const int t = h * TH + u;
                    yr[u] = MF::op(ain[s], xi[t][s], yr[u]);
                    yi[u] = MF::op(ai[s], xr[t][s], yi[u]);
                }
            }
            if (b + 1 < a.nb) {
                // through LDS in natural row order, back in block b + 1's grouping
#pragma unroll
                for (int u = 0; u < TH; ++u) {
                    const int t = h * TH + u;
                    const int base = mp[16 + o[t]];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int j = base + mp[MF::row(r, g)];
                        lds[w[t] * D + fc_slot<N>(j)] = mk<R>(yr[u][r], yi[u][r]);
                    }
                }
                fc_wave_fence();
#pragma unroll
                for (int u = 0; u < TH; ++u) {
                    const int t = h * TH + u;
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
                // apply_to: the transformed columns back to HBM (each wavefront writes
                // only the columns it read, so the pass may run in place)
                cpx<R>* O = static_cast<cpx<R>*>(a.out) + (long long)e * a.out_bstride;
#pragma unroll
                for (int u = 0; u < TH; ++u) {
                    const int t = h * TH + u;
                    const int col = c0 + w[t];
                    const int base = mp[16 + o[t]];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int j = base + mp[MF::row(r, g)];
                        if (col < m) O[(long long)j * a.out_stride + col] = mk<R>(yr[u][r], yi[u][r]);
                    }
                }
            }
        }
    }
    if (a.out) return;
    // cost pass, last block: one 16-term dot product with row c + trace_off of F per
    // column c.  X is in the block's grouping; the vector column whose outer group holds
    // the row contributes, its 4 lane groups summing their 4 entries each.
    {
        const int* mp = smap + (a.nb - 1) * kFusedMapStride;
        const int mask = mp[15];  // the block's four row bits (off of local index 15)
        const cpx<R>* Fl = Fe + (long long)(a.nb - 1) * 256;
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int col = c0 + w[t];
            const int jt = col + a.trace_off;
            const bool hit = col < m && (jt & ~mask) == mp[16 + o[t]];
            // local index of row jt: its bits at the block's qubits
            const int is = ((jt & mp[1]) ? 1 : 0) | ((jt & mp[2]) ? 2 : 0) | ((jt & mp[4]) ? 4 : 0) |
                           ((jt & mp[8]) ? 8 : 0);
            R sr = 0, si = 0;
            if (hit) {
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const cpx<R> f = Fl[is * 16 + 4 * s + g];
                    sr += f.re * xr[t][s] - f.im * xi[t][s];
                    si += f.re * xi[t][s] + f.im * xr[t][s];
                }
            }
            // the four lane groups of the vector column (lanes nl + 16 g)
            sr += __shfl_xor(sr, 16, 64);
            si += __shfl_xor(si, 16, 64);
            sr += __shfl_xor(sr, 32, 64);
            si += __shfl_xor(si, 32, 64);
            if (hit && g == 0) {
                double* d = a.diag + 2 * ((long long)e * m + col);
                d[0] = (double)sr;
                d[1] = (double)si;
            }
        }
    }
}

// ---------------------------------------------------------------------------------
// The cost pass as a transposed chain from unit vectors.
//
// Trace entry c is (V U)[rho, c] with rho = c + trace_off and V = F_{nb-1} ... F_0: row
// rho of V dotted with column c of U.  Row rho of V is column rho of
// V^T = F_0^T F_1^T ... F_{nb-1}^T, which this kernel builds from the unit vector
// e_rho: F_{nb-1}^T e_rho is row rho of F_{nb-1} on the rows that agree with rho
// outside the block's qubits (loaded, no arithmetic), and each later block only touches
// the rows that agree with rho outside the union S_b of the qubits seen so far (the
// support): 2^(|S_b| - 4) of a column's 2^N / 16 outer groups.  The 8-qubit 40-layer
// ansatz's chain has supports of 4, 7, 7, 8, 8, 8 qubits: 4 block products of work,
// where the forward pass over U's columns needs 5; the 10-qubit 60-layer ansatz's 9
// blocks need 6.25 instead of 8.  The last block's output is multiplied elementwise
// with column c of U and summed.
//
// Layout as k_fused_cost: a wavefront owns W = 16 T / O columns; block b's vector
// column n holds column w = n >> lg_b, support group k = n & (2^lg_b - 1), rows
// kb_b[k] | (rho_w & nS_b) | off_b[i] (kTraceMapStride ints per block: off[16],
// kb[64], lg, nS).  Only the first ceil(W 2^lg_b / 16) tiles carry work.  Rows outside
// the support are never stored; loads test the row against the previous block's
// support instead.  Every launch shape performs the same operations per column in the
// same order: results are bit-identical across shapes.

template <typename R, int N, int T, bool PF>
__global__ __launch_bounds__(256) void k_fused_trace(FusedCostArgs a) {
    typedef FcMfma<R> MF;
    constexpr int D = 1 << N;
    constexpr int O = D >> 4;      // outer groups per column
    constexpr int W = 16 * T / O;  // columns per wavefront
    static_assert(W >= 1, "a wavefront holds whole columns");
    extern __shared__ __align__(16) unsigned char fc_smem[];
    int* smap = reinterpret_cast<int*>(fc_smem);
    const int nmap = a.nb * kTraceMapStride;
    cpx<R>* region = reinterpret_cast<cpx<R>*>(fc_smem + (((size_t)nmap * sizeof(int) + 15) & ~(size_t)15));
    for (int i = threadIdx.x; i < nmap; i += blockDim.x) smap[i] = a.maps[i];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const long long per_xcd = gridDim.x >> 3;  // the grid is a multiple of 8
    const long long lwg = (long long)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    const long long wave = lwg * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (wave >= (long long)a.B * a.groups) return;
    const int e = (int)(wave / a.groups);
    const int c0 = (int)(wave % a.groups) * W;
    cpx<R>* lds = region + (threadIdx.x >> 6) * (16 * T * 16);
    const int nl = lane & 15, g = lane >> 4;
    const int m = a.m;
    // the row of V column c0 + w needs (columns past m take row trace_off and write
    // nothing; trace_off + m <= 2^N)
    auto rho_of = [&](int w) { return (c0 + w < m ? c0 + w : 0) + a.trace_off; };
    const cpx<R>* Fe = static_cast<const cpx<R>*>(a.F) + (long long)e * a.F_bstride;
    // X of block bn (B operand of step s, tile t: X[4 s + g][16 t + nl] in its grouping)
    // from the LDS region, for the tiles holding its support's groups; vector columns
    // past the wavefront's columns read column 0 (their products are never stored)
    R xr[T][4], xi[T][4];
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
        const int* mq = smap + (a.nb - 1) * kTraceMapStride;
        const cpx<R>* Fq = Fe + (long long)(a.nb - 1) * 256;
        for (int p = lane; p < 16 * W; p += 64) {
            const int w = p >> 4, i = p & 15;
            const int rho = rho_of(w);
            lds[w * D + fc_slot<N>((rho & mq[81]) | mq[i])] = Fq[ft_local(mq, rho) * 16 + i];
        }
    }
    fc_wave_fence();
    load_x(a.nb - 2);
    // A operand of step s: F_b^T[nl][4 s + g] = F_b[4 s + g][nl] (with PF, block b - 1's
    // is loaded while block b runs)
    R far[4], fai[4];
    if constexpr (PF) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const cpx<R> f = Fe[(long long)(a.nb - 2) * 256 + (4 * s + g) * 16 + nl];
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
    int b = a.nb - 2;
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
    // block 0, then row rho of V times column c of U: per lane its rows of the column,
    // then the lanes of the column (the four lane groups and, below 16 support groups,
    // the column's lanes of the tile); a column spanning several tiles accumulates them
    // first
    const int* mp = smap;
    const int lg0 = mp[80], nS = mp[81];
    typename MF::V yr[T], yi[T];
    {
        // all T tiles (those past the support's hold zeros when it stays short)
        R ar[4], ai[4], ain[4];
        a_operand(0, ar, ai, ain);
        ft_mma<R, T, T>(ar, ai, ain, xr, xi, yr, yi);
    }
    const cpx<R>* U =
        static_cast<const cpx<R>*>(a.U) + (a.in_index ? (long long)a.in_index[e] : (long long)e) * a.in_bstride;
    const int ist = a.in_stride;
    auto dot = [&](auto full_support) {
        constexpr bool FULL = decltype(full_support)::value;
        const int lg = FULL ? N - 4 : lg0;
        const int ng = 1 << lg, seg = ng < 16 ? ng : 16;
        const int ts = ((W << lg) + 15) >> 4;
        double sr = 0, si = 0;
#pragma unroll
        for (int t = 0; t < T; ++t) {
            if (t < ts) {
                const int n = 16 * t + nl;
                const int w = FULL ? wd[t] : n >> lg;
                const int col = c0 + w;
                if (w < W && col < m) {
                    const int base = FULL ? mp[16 + od[t]] : mp[16 + (n & (ng - 1))] | (rho_of(w) & nS);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int j = base + mp[MF::row(r, g)];
                        const cpx<R> u = U[(long long)j * ist + col];
                        sr += (double)yr[t][r] * (double)u.re - (double)yi[t][r] * (double)u.im;
                        si += (double)yr[t][r] * (double)u.im + (double)yi[t][r] * (double)u.re;
                    }
                }
                if (lg < 4 || (((t + 1) << 4) & (ng - 1)) == 0) {
                    sr += __shfl_xor(sr, 16, 64);
                    si += __shfl_xor(si, 16, 64);
                    sr += __shfl_xor(sr, 32, 64);
                    si += __shfl_xor(si, 32, 64);
#pragma unroll
                    for (int d = 1; d < 16; d <<= 1) {
                        if (d < seg) {
                            sr += __shfl_xor(sr, d, 64);
                            si += __shfl_xor(si, d, 64);
                        }
                    }
                    if (g == 0 && (nl & (seg - 1)) == 0 && w < W && col < m) {
                        double* dg = a.diag + 2 * ((long long)e * m + col);
                        dg[0] = sr;
                        dg[1] = si;
                    }
                    sr = 0;
                    si = 0;
                }
            }
        }
    };
    if (lg0 == N - 4) dot(std::true_type{});
    else dot(std::false_type{});
}

bool fused_trace_cfg_valid(int N, bool f32, int nb, FusedCostCfg cfg) {
    if (N < 4 || N > kFusedCostMaxQubits || nb < 2) return false;
    if ((cfg.tiles != 2 && cfg.tiles != 4) || cfg.groups != 1) return false;
    if (cfg.waves != 1 && cfg.waves != 2 && cfg.waves != 4) return false;
    if (16 * cfg.tiles < ((1 << N) >> 4)) return false;
    return fused_trace_lds_bytes(f32, nb, cfg) <= device_max_lds();
}

size_t fused_trace_lds_bytes(bool f32, int nb, FusedCostCfg cfg) {
    return (((size_t)nb * kTraceMapStride * sizeof(int) + 15) & ~(size_t)15) +
           (size_t)cfg.waves * cfg.tiles * 256 * (f32 ? 8 : 16);
}

template <typename R, int N>
static void launch_ft_n(const FusedCostArgs& a, FusedCostCfg cfg, int grid, size_t shm, hipStream_t s) {
    const int threads = 64 * cfg.waves;
    if constexpr (16 * 4 >= ((1 << N) >> 4)) {
        if (cfg.tiles == 4 && cfg.prefetch)
            hipLaunchKernelGGL((k_fused_trace<R, N, 4, true>), dim3(grid), dim3(threads), shm, s, a);
        else if (cfg.tiles == 4)
            hipLaunchKernelGGL((k_fused_trace<R, N, 4, false>), dim3(grid), dim3(threads), shm, s, a);
    }
    if constexpr (16 * 2 >= ((1 << N) >> 4)) {
        if (cfg.tiles == 2 && cfg.prefetch)
            hipLaunchKernelGGL((k_fused_trace<R, N, 2, true>), dim3(grid), dim3(threads), shm, s, a);
        else if (cfg.tiles == 2)
            hipLaunchKernelGGL((k_fused_trace<R, N, 2, false>), dim3(grid), dim3(threads), shm, s, a);
    }
}

template <typename R>
static hipError_t launch_ft(int N, const FusedCostArgs& a, FusedCostCfg cfg, int grid, size_t shm, hipStream_t s) {
    switch (N) {
    case 4: launch_ft_n<R, 4>(a, cfg, grid, shm, s); break;
    case 5: launch_ft_n<R, 5>(a, cfg, grid, shm, s); break;
    case 6: launch_ft_n<R, 6>(a, cfg, grid, shm, s); break;
    case 7: launch_ft_n<R, 7>(a, cfg, grid, shm, s); break;
    case 8: launch_ft_n<R, 8>(a, cfg, grid, shm, s); break;
    case 9: launch_ft_n<R, 9>(a, cfg, grid, shm, s); break;
    case 10: launch_ft_n<R, 10>(a, cfg, grid, shm, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fused_trace(bool f32, int N, FusedCostArgs a, FusedCostCfg cfg, hipStream_t s) {
    if (a.nb < 2 || a.m <= 0 || a.B <= 0 || a.out || !fused_trace_cfg_valid(N, f32, a.nb, cfg))
        return hipErrorInvalidValue;
    if (a.in_stride < a.m || a.trace_off < 0 || a.trace_off + a.m > (1 << N)) return hipErrorInvalidValue;
    const int W = 16 * cfg.tiles / ((1 << N) >> 4);
    a.groups = (a.m + W - 1) / W;
    const long long waves = (long long)a.B * a.groups;
    const long long wgs = (waves + cfg.waves - 1) / cfg.waves;
    const int grid = (int)((wgs + 7) / 8 * 8);
    const size_t shm = fused_trace_lds_bytes(f32, a.nb, cfg);
    return f32 ? launch_ft<float>(N, a, cfg, grid, shm, s) : launch_ft<double>(N, a, cfg, grid, shm, s);
}

bool fused_cost_cfg_valid(int N, bool f32, int nb, FusedCostCfg cfg) {
    if (N < 4 || N > kFusedCostMaxQubits) return false;
    const int O = (1 << N) >> 4;
    if (!((cfg.tiles == 4 && cfg.groups == 1) || (cfg.tiles == 4 && cfg.groups == 2 && cfg.prefetch) ||
          (cfg.tiles == 2 && cfg.groups == 1)))
        return false;
    if (cfg.waves != 1 && cfg.waves != 2 && cfg.waves != 4) return false;
    if (16 * cfg.tiles < O || (16 * cfg.tiles / cfg.groups) % O != 0) return false;
    return fused_cost_lds_bytes(f32, nb, cfg) <= device_max_lds();
}

size_t fused_cost_lds_bytes(bool f32, int nb, FusedCostCfg cfg) {
    return (((size_t)nb * kFusedMapStride * sizeof(int) + 15) & ~(size_t)15) +
           (size_t)cfg.waves * cfg.tiles * 256 * (f32 ? 8 : 16);
}

template <typename R, int N, int T, int H, bool PF>
static void launch_fc_shape(const FusedCostArgs& a, int grid, int threads, size_t shm, hipStream_t s) {
    if constexpr (16 * T >= ((1 << N) >> 4) && (16 * T / H) % ((1 << N) >> 4) == 0)
        hipLaunchKernelGGL((k_fused_cost<R, N, T, H, PF>), dim3(grid), dim3(threads), shm, s, a);
}

template <typename R, int N>
static void launch_fc_n(const FusedCostArgs& a, FusedCostCfg cfg, int grid, size_t shm, hipStream_t s) {
    const int threads = 64 * cfg.waves;
    if (cfg.tiles == 4 && cfg.groups == 2) launch_fc_shape<R, N, 4, 2, true>(a, grid, threads, shm, s);
    else if (cfg.tiles == 4 && cfg.prefetch) launch_fc_shape<R, N, 4, 1, true>(a, grid, threads, shm, s);
    else if (cfg.tiles == 4) launch_fc_shape<R, N, 4, 1, false>(a, grid, threads, shm, s);
    else if (cfg.prefetch) launch_fc_shape<R, N, 2, 1, true>(a, grid, threads, shm, s);
    else launch_fc_shape<R, N, 2, 1, false>(a, grid, threads, shm, s);
}

template <typename R>
static hipError_t launch_fc(int N, const FusedCostArgs& a, FusedCostCfg cfg, int grid, size_t shm, hipStream_t s) {
    switch (N) {
    case 4: launch_fc_n<R, 4>(a, cfg, grid, shm, s); break;
    case 5: launch_fc_n<R, 5>(a, cfg, grid, shm, s); break;
    case 6: launch_fc_n<R, 6>(a, cfg, grid, shm, s); break;
    case 7: launch_fc_n<R, 7>(a, cfg, grid, shm, s); break;
    case 8: launch_fc_n<R, 8>(a, cfg, grid, shm, s); break;
    case 9: launch_fc_n<R, 9>(a, cfg, grid, shm, s); break;
    case 10: launch_fc_n<R, 10>(a, cfg, grid, shm, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fused_cost(bool f32, int N, FusedCostArgs a, FusedCostCfg cfg, hipStream_t s) {
    if (a.nb <= 0 || a.m <= 0 || a.B <= 0 || !fused_cost_cfg_valid(N, f32, a.nb, cfg)) return hipErrorInvalidValue;
    if (a.in_stride < a.m || (a.out && a.out_stride < a.m)) return hipErrorInvalidValue;
    const int W = 16 * cfg.tiles / ((1 << N) >> 4);
    a.groups = (a.m + W - 1) / W;
    const long long waves = (long long)a.B * a.groups;
    const long long wgs = (waves + cfg.waves - 1) / cfg.waves;
    const int grid = (int)((wgs + 7) / 8 * 8);
    const size_t shm = fused_cost_lds_bytes(f32, a.nb, cfg);
    return f32 ? launch_fc<float>(N, a, cfg, grid, shm, s) : launch_fc<double>(N, a, cfg, grid, shm, s);
}

}  // namespace sqh
