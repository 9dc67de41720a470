// sqh_k_solve.hip -- persistent batched decomposition: one wavefront per run (four at
// six qubits), the whole optimisation in one launch.
//
// Each wavefront takes runs from a work counter and, for each, loops: advance the run's
// BFGS state machine (sqh_bfgs_core.h: BFGS_Powell with the Wolfe line search,
// iteration loops and outer restarts) until it asks for an evaluation, evaluate the
// cost and the full gradient at that point itself (register tile of the run's circuit
// over its target, adjoint sweep, correlations reduced in the wavefront), and go on --
// no kernel boundary, no host round trip, no per-round launch.  Runs finish at their
// own pace and the wavefront moves on to the next one, so uneven batches keep every
// wavefront busy.  The per-run vectors and Powell's factor stay in device memory; the
// evaluation point, the gradient, the 2x2 kernels and the correlations live in LDS.
#include "sqh_bfgs_core.h"
#include "sqh_device.h"
#include "sqh_kernels.h"
#include "sqh_solve.h"
#include "sqh_solve_loop.h"

namespace sqh {

// Cost and gradient of one run at xr (LDS) by the wavefront: f returned, gradient to gr
// (LDS).  kt: LDS table of ns 2x2 kernels (complex128 slots; complex64 uses the first
// half of each), mc: LDS [ns][8] correlations, lds: the remap buffer.
template <int N, typename R>
__device__ double wave_eval(const cpx<R>* __restrict__ U, int m, int trace_off, int variant, const Op* __restrict__ ops,
                            int nops, const Layout* __restrict__ lays, const Slot* __restrict__ slots, int ns,
                            const double* xr, double* gr, cpx<double>* kt_raw, double* mc, cpx<R>* lds) {
    using G = Geo<N>;
    constexpr int CW = G::CW;
    const int lane = threadIdx.x;
    cpx<R>* kt = reinterpret_cast<cpx<R>*>(kt_raw);
    for (int sl = lane; sl < ns; sl += 64) {
        const Slot s = slots[sl];
        cpx<double> k[4];
        dev_gate_kernel(s.type, xr + s.param_start, k);
#pragma unroll
        for (int e = 0; e < 4; ++e) kt[4 * sl + e] = cvt<R>(k[e]);
    }
    for (int i = lane; i < 8 * ns; i += 64) mc[i] = 0.0;
    __syncthreads();
    const int s_idx = lane / CW, cc = lane % CW;
    double tre = 0.0, tim = 0.0;  // this lane's trace entries
    for (int c0 = 0; c0 < m; c0 += CW) {
        const int col = c0 + cc;
        const bool valid = col < m;
        cpx<R> x[G::NR], y[G::NR];
        {
            const int so = sub_off<N>(&lays[0], s_idx);
#pragma unroll
            for (int i = 0; i < G::NR; ++i) {
                const int row = so + lays[0].reg_off[i];
                x[i] = valid ? U[(long long)row * m + col] : mk<R>(0, 0);
            }
        }
        int cur = 0;
        {
            OpStream<R> os(ops, kt, nops, 0, 1);
            for (int p = 0; p < nops; ++p) {
                cpx<R> kn[4];
                Op after;
                os.prefetch(p + 2, kn, after);
                const Op op = os.cur;
                if (op.kind == OP_REMAP) {
                    const int o = op.idx >> 16, n = op.idx & 0xffff;
                    remap<N, R>(x, &lays[o], &lays[n], s_idx, cc, lds);
                    cur = n;
                } else {
                    exec_gate<N, false, R>(x, op, os.k, s_idx);
                }
                os.cur = os.nxt;
                os.nxt = after;
#pragma unroll
                for (int e = 0; e < 4; ++e) os.k[e] = kn[e];
            }
        }
        {
            // the trace entry of column col and the selector seeding the adjoint sweep
            const int so = sub_off<N>(&lays[cur], s_idx);
            const int tr = col + trace_off;
#pragma unroll
            for (int i = 0; i < G::NR; ++i) {
                const bool hit = valid && so + lays[cur].reg_off[i] == tr;
                if (hit) {
                    tre += (double)x[i].re;
                    tim += (double)x[i].im;
                }
                y[i] = mk<R>(hit ? (R)1 : (R)0, (R)0);
            }
        }
        OpStream<R> os(ops, kt, nops, nops - 1, -1);
        for (int p = nops - 1; p >= 0; --p) {
            cpx<R> kn[4];
            Op after;
            os.prefetch(p - 2, kn, after);
            const Op op = os.cur;
            if (op.kind == OP_REMAP) {
                const int o = op.idx >> 16, n = op.idx & 0xffff;
                remap<N, R>(x, &lays[n], &lays[o], s_idx, cc, lds);
                remap<N, R>(y, &lays[n], &lays[o], s_idx, cc, lds);
            } else if (op.kind == OP_X || op.kind == OP_Z) {
                exec_gate<N, false, R>(x, op, os.k, s_idx);
                exec_gate<N, false, R>(y, op, os.k, s_idx);
            } else {
                const cpx<R>* k = os.k;
                cpx<R> kd[4], ktr[4];
                kd[0] = mk(k[0].re, -k[0].im);  // K^dagger
                kd[1] = mk(k[2].re, -k[2].im);
                kd[2] = mk(k[1].re, -k[1].im);
                kd[3] = mk(k[3].re, -k[3].im);
                ktr[0] = k[0];                  // K^T
                ktr[1] = k[2];
                ktr[2] = k[1];
                ktr[3] = k[3];
                cpx<R> acc[4] = {mk<R>(0, 0), mk<R>(0, 0), mk<R>(0, 0), mk<R>(0, 0)};
                constexpr int K = G::K;
                switch (op.kind * 4 + op.t) {
                    SQH_ARMS(OP_U, K, (pairs_apply<T, false, false>(x, kd, 0u), pairs_corr<T>(y, x, acc, 0u, false),
                                       pairs_apply<T, false, false>(y, ktr, 0u)))
                    SQH_ARMS(OP_CU, K,
                             (pairs_apply<T, true, false>(x, kd, ctrl_mask<N>(op.cmode, op.c, s_idx)),
                              pairs_corr<T>(y, x, acc, ctrl_mask<N>(op.cmode, op.c, s_idx), true),
                              pairs_apply<T, true, false>(y, ktr, ctrl_mask<N>(op.cmode, op.c, s_idx))))
                default:
                    break;
                }
                double r[8];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    r[2 * e] = wave_sum((double)acc[e].re);
                    r[2 * e + 1] = wave_sum((double)acc[e].im);
                }
                if (lane == 0) {
                    double* d = mc + (long long)op.idx * 8;
#pragma unroll
                    for (int e = 0; e < 8; ++e) d[e] += r[e];
                }
            }
            os.cur = os.nxt;
            os.nxt = after;
#pragma unroll
            for (int e = 0; e < 4; ++e) os.k[e] = kn[e];
        }
    }
    tre = wave_sum(tre);
    tim = wave_sum(tim);
    __syncthreads();
    // gradient from the correlations and each parameter's derivative kernel
    for (int sl = lane; sl < ns; sl += 64) {
        const Slot s = slots[sl];
        const double* M = mc + (long long)sl * 8;
        for (int q = 0; q < s.nparam; ++q) {
            cpx<double> d[4];
            dev_deriv_kernel(s.type, xr + s.param_start, q, d);
            double kre = 0.0, kim = 0.0;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                kre += d[e].re * M[2 * e] - d[e].im * M[2 * e + 1];
                kim += d[e].re * M[2 * e + 1] + d[e].im * M[2 * e];
            }
            gr[s.param_start + q] = grad_of(variant, tre, tim, kre, kim, (double)m);
        }
    }
    __syncthreads();
    return cost_of(variant, tre, tim, (double)m);
}

// The wavefront's column slabs through its own LDS buffer in two column groups (half
// the slab each, so the buffer is D x CW / 2): a remap permutes rows within each column,
// and a column's rows are held by the lanes of that column only.  Wave-scope ordering
// (LDS serves a wavefront's requests in order), no workgroup barrier.
template <int N, typename R>
__device__ __forceinline__ void remap_halves(cpx<R> (&v)[Geo<N>::NR], const Layout* lo, const Layout* ln, int s,
                                             int cc, cpx<R>* lds) {
    if constexpr (N > Geo<N>::K) {
        constexpr int H = Geo<N>::CW / 2;
        static_assert(H >= 1, "two column groups need two columns");
        const int so = sub_off<N>(lo, s), sn = sub_off<N>(ln, s);
        const int grp = cc / H, ch = cc % H;
#pragma unroll
        for (int gg = 0; gg < 2; ++gg) {
            if (grp == gg) {
#pragma unroll
                for (int i = 0; i < Geo<N>::NR; ++i) {
                    const int row = so + lo->reg_off[i];
                    lds[(row ^ ((row >> 4) & 15)) * H + ch] = v[i];
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            if (grp == gg) {
#pragma unroll
                for (int i = 0; i < Geo<N>::NR; ++i) {
                    const int row = sn + ln->reg_off[i];
                    v[i] = lds[(row ^ ((row >> 4) & 15)) * H + ch];
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        }
    }
}

// Cost and gradient of one run by NW wavefronts: wavefront w takes column slabs w,
// w + NW, ... (its own remap buffer and correlation table), the partial traces and
// correlations are summed over the wavefronts in a fixed order.  Same arithmetic per
// slab as wave_eval; sh holds the 2 NW partial traces.
template <int N, typename R, int NW>
__device__ double wave_eval_mw(const cpx<R>* __restrict__ U, int m, int trace_off, int variant,
                               const Op* __restrict__ ops, int nops, const Layout* __restrict__ lays,
                               const Slot* __restrict__ slots, int ns, const double* xr, double* gr,
                               cpx<double>* kt_raw, double* mc, cpx<R>* lds_all, double* sh) {
    using G = Geo<N>;
    constexpr int CW = G::CW, NT = 64 * NW;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    cpx<R>* kt = reinterpret_cast<cpx<R>*>(kt_raw);
    for (int sl = tid; sl < ns; sl += NT) {
        const Slot s = slots[sl];
        cpx<double> k[4];
        dev_gate_kernel(s.type, xr + s.param_start, k);
#pragma unroll
        for (int e = 0; e < 4; ++e) kt[4 * sl + e] = cvt<R>(k[e]);
    }
    for (int i = tid; i < 8 * ns * NW; i += NT) mc[i] = 0.0;
    __syncthreads();
    double* mcw = mc + (size_t)w * 8 * ns;
    cpx<R>* lds = lds_all + (size_t)w * (G::LDS / 2);
    const int s_idx = lane / CW, cc = lane % CW;
    double tre = 0.0, tim = 0.0;
    for (int c0 = w * CW; c0 < m; c0 += NW * CW) {
        const int col = c0 + cc;
        const bool valid = col < m;
        cpx<R> x[G::NR], y[G::NR];
        {
            const int so = sub_off<N>(&lays[0], s_idx);
#pragma unroll
            for (int i = 0; i < G::NR; ++i) {
                const int row = so + lays[0].reg_off[i];
                x[i] = valid ? U[(long long)row * m + col] : mk<R>(0, 0);
            }
        }
        int cur = 0;
        {
            OpStream<R> os(ops, kt, nops, 0, 1);
            for (int p = 0; p < nops; ++p) {
                cpx<R> kn[4];
                Op after;
                os.prefetch(p + 2, kn, after);
                const Op op = os.cur;
                if (op.kind == OP_REMAP) {
                    const int o = op.idx >> 16, n = op.idx & 0xffff;
                    remap_halves<N, R>(x, &lays[o], &lays[n], s_idx, cc, lds);
                    cur = n;
                } else {
                    exec_gate<N, false, R>(x, op, os.k, s_idx);
                }
                os.cur = os.nxt;
                os.nxt = after;
#pragma unroll
                for (int e = 0; e < 4; ++e) os.k[e] = kn[e];
            }
        }
        {
            const int so = sub_off<N>(&lays[cur], s_idx);
            const int tr = col + trace_off;
#pragma unroll
            for (int i = 0; i < G::NR; ++i) {
                const bool hit = valid && so + lays[cur].reg_off[i] == tr;
                if (hit) {
                    tre += (double)x[i].re;
                    tim += (double)x[i].im;
                }
                y[i] = mk<R>(hit ? (R)1 : (R)0, (R)0);
            }
        }
        OpStream<R> os(ops, kt, nops, nops - 1, -1);
        for (int p = nops - 1; p >= 0; --p) {
            cpx<R> kn[4];
            Op after;
            os.prefetch(p - 2, kn, after);
            const Op op = os.cur;
            if (op.kind == OP_REMAP) {
                const int o = op.idx >> 16, n = op.idx & 0xffff;
                remap_halves<N, R>(x, &lays[n], &lays[o], s_idx, cc, lds);
                remap_halves<N, R>(y, &lays[n], &lays[o], s_idx, cc, lds);
            } else if (op.kind == OP_X || op.kind == OP_Z) {
                exec_gate<N, false, R>(x, op, os.k, s_idx);
                exec_gate<N, false, R>(y, op, os.k, s_idx);
            } else {
                const cpx<R>* k = os.k;
                cpx<R> kd[4], ktr[4];
                kd[0] = mk(k[0].re, -k[0].im);
                kd[1] = mk(k[2].re, -k[2].im);
                kd[2] = mk(k[1].re, -k[1].im);
                kd[3] = mk(k[3].re, -k[3].im);
                ktr[0] = k[0];
                ktr[1] = k[2];
                ktr[2] = k[1];
                ktr[3] = k[3];
                cpx<R> acc[4] = {mk<R>(0, 0), mk<R>(0, 0), mk<R>(0, 0), mk<R>(0, 0)};
                constexpr int K = G::K;
                switch (op.kind * 4 + op.t) {
                    SQH_ARMS(OP_U, K, (pairs_apply<T, false, false>(x, kd, 0u), pairs_corr<T>(y, x, acc, 0u, false),
                                       pairs_apply<T, false, false>(y, ktr, 0u)))
                    SQH_ARMS(OP_CU, K,
                             (pairs_apply<T, true, false>(x, kd, ctrl_mask<N>(op.cmode, op.c, s_idx)),
                              pairs_corr<T>(y, x, acc, ctrl_mask<N>(op.cmode, op.c, s_idx), true),
                              pairs_apply<T, true, false>(y, ktr, ctrl_mask<N>(op.cmode, op.c, s_idx))))
                default:
                    break;
                }
                double r[8];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    r[2 * e] = wave_sum((double)acc[e].re);
                    r[2 * e + 1] = wave_sum((double)acc[e].im);
                }
                if (lane == 0) {
                    double* d = mcw + (long long)op.idx * 8;
#pragma unroll
                    for (int e = 0; e < 8; ++e) d[e] += r[e];
                }
            }
            os.cur = os.nxt;
            os.nxt = after;
#pragma unroll
            for (int e = 0; e < 4; ++e) os.k[e] = kn[e];
        }
    }
    tre = wave_sum(tre);
    tim = wave_sum(tim);
    if (lane == 0) {
        sh[2 * w] = tre;
        sh[2 * w + 1] = tim;
    }
    __syncthreads();
    tre = 0.0;
    tim = 0.0;
    for (int v = 0; v < NW; ++v) {
        tre += sh[2 * v];
        tim += sh[2 * v + 1];
    }
    for (int sl = tid; sl < ns; sl += NT) {
        const Slot s = slots[sl];
        double M[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) M[e] = 0.0;
        for (int v = 0; v < NW; ++v) {
            const double* src = mc + ((size_t)v * ns + sl) * 8;
#pragma unroll
            for (int e = 0; e < 8; ++e) M[e] += src[e];
        }
        for (int q = 0; q < s.nparam; ++q) {
            cpx<double> d[4];
            dev_deriv_kernel(s.type, xr + s.param_start, q, d);
            double kre = 0.0, kim = 0.0;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                kre += d[e].re * M[2 * e] - d[e].im * M[2 * e + 1];
                kim += d[e].re * M[2 * e + 1] + d[e].im * M[2 * e];
            }
            gr[s.param_start + q] = grad_of(variant, tre, tim, kre, kim, (double)m);
        }
    }
    __syncthreads();
    return cost_of(variant, tre, tim, (double)m);
}

// Wavefronts per run of the persistent solver: four at six qubits (a 64 x 64 target is
// four column slabs), one below.
constexpr int solve_waves(int N) { return N == 6 ? 4 : 1; }

template <int N>
__global__ __launch_bounds__(64 * solve_waves(N)) void k_solve_mw(SolveArgs a) {
    using G = Geo<N>;
    constexpr int NW = solve_waves(N);
    const int m = a.m, toff = a.trace_off, var = a.variant;
    solve_runs_mw<G::LDS / 2, NW>(
        a,
        [=](const cpx<double>* U, const Op* ops, int nops, const Layout* lays, const Slot* slots, int ns,
            const double* xr, double* gr, cpx<double>* kt, double* mc, cpx<double>* lds, double* sh) {
            return wave_eval_mw<N, double, NW>(U, m, toff, var, ops, nops, lays, slots, ns, xr, gr, kt, mc, lds, sh);
        },
        [=](const cpx<float>* U, const Op* ops, int nops, const Layout* lays, const Slot* slots, int ns,
            const double* xr, double* gr, cpx<double>* kt, double* mc, cpx<float>* lds, double* sh) {
            return wave_eval_mw<N, float, NW>(U, m, toff, var, ops, nops, lays, slots, ns, xr, gr, kt, mc, lds, sh);
        });
}

template <int N>
__global__ __launch_bounds__(64, (N <= 2) ? 2 : 1) void k_solve(SolveArgs a) {
    using G = Geo<N>;
    const int m = a.m, toff = a.trace_off, var = a.variant;
    solve_runs<G::LDS>(
        a,
        [=](const cpx<double>* U, const Op* ops, int nops, const Layout* lays, const Slot* slots, int ns,
            const double* xr, double* gr, cpx<double>* kt, double* mc, cpx<double>* lds) {
            return wave_eval<N, double>(U, m, toff, var, ops, nops, lays, slots, ns, xr, gr, kt, mc, lds);
        },
        [=](const cpx<float>* U, const Op* ops, int nops, const Layout* lays, const Slot* slots, int ns,
            const double* xr, double* gr, cpx<double>* kt, double* mc, cpx<float>* lds) {
            return wave_eval<N, float>(U, m, toff, var, ops, nops, lays, slots, ns, xr, gr, kt, mc, lds);
        });
}

// The runs of several groups of at most five qubits in one launch (one wavefront per
// run): a combined queue, each run dispatched to its group's code.  Separate launches
// per group would each need a stream of their own, and a process gets four hardware
// queues by default (one already taken by the default stream); launches on a shared
// queue run one after the other.
template <int N>
__device__ void solve_small_run(const SolveArgs& a, int r) {
    using G = Geo<N>;
    const int m = a.m, toff = a.trace_off, var = a.variant;
    solve_one_run<G::LDS>(
        a, r,
        [=](const cpx<double>* U, const Op* ops, int nops, const Layout* lays, const Slot* slots, int ns,
            const double* xr, double* gr, cpx<double>* kt, double* mc, cpx<double>* lds) {
            return wave_eval<N, double>(U, m, toff, var, ops, nops, lays, slots, ns, xr, gr, kt, mc, lds);
        },
        [=](const cpx<float>* U, const Op* ops, int nops, const Layout* lays, const Slot* slots, int ns,
            const double* xr, double* gr, cpx<double>* kt, double* mc, cpx<float>* lds) {
            return wave_eval<N, float>(U, m, toff, var, ops, nops, lays, slots, ns, xr, gr, kt, mc, lds);
        });
}

__global__ __launch_bounds__(64, 1) void k_solve_small(SolveSmallArgs s) {
    const int lane = threadIdx.x;
    for (;;) {
        int q = 0;
        if (lane == 0) q = atomicAdd(s.next, 1);
        q = __builtin_amdgcn_readfirstlane(__shfl(q, 0, 64));
        if (q >= s.n_runs) return;  // every wavefront leaves once the queue is empty
        const int r = s.queue[q];
        switch (s.queue_n[q]) {
        case 1: solve_small_run<1>(s.g[0], r); break;
        case 2: solve_small_run<2>(s.g[1], r); break;
        case 3: solve_small_run<3>(s.g[2], r); break;
        case 4: solve_small_run<4>(s.g[3], r); break;
        case 5: solve_small_run<5>(s.g[4], r); break;
        default: break;
        }
    }
}

hipError_t launch_solve_small(const SolveSmallArgs& s, int grid, size_t shm, hipStream_t st) {
    hipLaunchKernelGGL(k_solve_small, dim3(grid), dim3(64), shm, st, s);
    return hipGetLastError();
}

size_t solve_lds_bytes(int N, int Ps, int ns_max) {
    const int cw = tile_columns_per_wave(N);
    const int nw = solve_waves(N);
    if (nw > 1) {
        // NW remap buffers of half a slab, NW correlation tables, NW partial-direction
        // rows, the parked state and the queue slot
        const size_t remap = (size_t)(1 << N) * cw / 2 * nw;
        return sizeof(cpx<double>) * remap + sizeof(cpx<double>) * 4 * (size_t)ns_max +
               sizeof(double) * 8 * (size_t)ns_max * nw + sizeof(double) * (5 + nw + RV_COUNT) * (size_t)Ps +
               sizeof(double) * 8 + sizeof(RunState) + sizeof(int);
    }
    const size_t remap = N > 4 ? (size_t)(1 << N) * cw : 1;
    return sizeof(cpx<double>) * remap + sizeof(cpx<double>) * 4 * (size_t)ns_max + sizeof(double) * 8 * ns_max +
           sizeof(double) * (6 + RV_COUNT) * (size_t)Ps + sizeof(double) * 8 + sizeof(RunState);
}

hipError_t launch_solve(int N, const SolveArgs& a, int grid, hipStream_t s) {
    const size_t shm = solve_lds_bytes(N, a.o.P, a.ns_max);
    switch (N) {
#define SQH_SOLVE_CASE(n) \
    case n: hipLaunchKernelGGL(k_solve<n>, dim3(grid), dim3(64), shm, s, a); break;
        SQH_SOLVE_CASE(1)
        SQH_SOLVE_CASE(2)
        SQH_SOLVE_CASE(3)
        SQH_SOLVE_CASE(4)
        SQH_SOLVE_CASE(5)
#undef SQH_SOLVE_CASE
    case 6:
        hipLaunchKernelGGL(k_solve_mw<6>, dim3(grid), dim3(64 * solve_waves(6)), shm, s, a);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace sqh
