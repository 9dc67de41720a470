// sqh_k_solve.hip -- persistent batched decomposition: one wavefront per run, the
// whole optimisation in one launch.
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

size_t solve_lds_bytes(int N, int Ps, int ns_max) {
    const int cw = tile_columns_per_wave(N);
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
        SQH_SOLVE_CASE(6)
#undef SQH_SOLVE_CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace sqh
