// sqh_k_bfgs.hip -- kernels of the batched BFGS optimiser (state machine per run).
#include <cfloat>

#include "sqh_bfgs_device.h"
#include "sqh_kernels.h"

namespace sqh {

constexpr int kBfgsThreads = 256;
constexpr int kBfgsMaxP = 1024;
constexpr double kNumPrecision = 1.42e-14;  // Grad_Descend::num_precision (grad_descend.cpp:40)

__device__ inline double wg_sum(double v, double* sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < kBfgsThreads / 64; ++i) t += sh[i];
    return t;
}

// Sum over the wavefront, lane 0's result broadcast (so a branch on it is uniform).
__device__ inline double wave_total(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return __shfl(v, 0, 64);
}

__device__ inline double wg_min(double v, double* sh) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    double t = sh[0];
    for (int i = 1; i < kBfgsThreads / 64; ++i) t = fmin(t, sh[i]);
    return t;
}

// Ps: stride of the per-run vectors (the largest parameter count); run_P: each
// run's parameter count (null: all Ps).  counts: [2] finished runs, [4 + 2g + p]
// evaluation requests of group g in precision p.
__global__ void k_bfgs_init(RunState* __restrict__ st, double* __restrict__ vecs, const double* __restrict__ x0,
                            int B, int Ps, const int* __restrict__ run_P, int mixed, int* __restrict__ counts,
                            int n_counts) {
    const int r = blockIdx.x;
    if (r >= B) return;
    const int P = run_P ? run_P[r] : Ps;
    double* V = vecs + (size_t)r * RV_COUNT * Ps;
    for (int k = threadIdx.x; k < P; k += blockDim.x) V[RV_BEST * Ps + k] = x0[(size_t)r * Ps + k];
    if (threadIdx.x == 0) {
        RunState S;
        S.phase = PH_OUTER_START;
        S.status = BS_INITIAL;
        S.prec = mixed ? 1 : 0;
        S.outer = 0;
        S.ls_iter = 0;
        S.finished = 0;
        S.P = P;
        S.loop = 0;
        S.jit_used = S.pad1 = S.pad2 = S.pad3 = 0;
        S.calls = S.total_calls = S.calls32 = 0;
        S.f = S.fprev = 0.0;
        S.norm_Z = -1.0;
        S.step = S.f_lowest = S.f_highest = S.step_highest = S.step_lowest = S.step_best = S.f_best = 0.0;
        S.dg_abs_best = S.dg_highest = S.dg0 = S.maxstep = 0.0;
        S.cur_min = DBL_MAX;
        for (int i = 0; i < kMinWindow; ++i) S.minvec[i] = 0.0;
        st[r] = S;
    }
    if (r == 0)
        for (int i = threadIdx.x; i < n_counts; i += blockDim.x) counts[i] = 0;
}

// Clears the evaluation-request counters (counts[4 ..]) before each round.
__global__ void k_bfgs_reset_counts(int* __restrict__ counts, int n_groups) {
    for (int i = threadIdx.x; i < 2 * n_groups; i += blockDim.x) counts[4 + i] = 0;
}

// BFGS_Powell::BFGS_iteration (BFGS_Powell.cpp:202-386) across the workgroup.
// Z is P x P stored column-major (Z[row][c] at c * P + row), so a column is
// contiguous: the Givens sweep and the column updates read whole columns with
// consecutive threads on consecutive rows.  The rotation angles of the sweep depend
// only on the vector being folded (Z^T g), so one thread computes all of them first
// and every thread then applies the whole sequence to its own rows without barriers.
__device__ bool bfgs_update(double* __restrict__ Z, double* __restrict__ V, int P, int Ps, double& norm_Z,
                            double* sh, double* zs, double* rc, double* rs, double* pd) {
    const int tid = threadIdx.x, nt = blockDim.x;
    double* x = V + RV_X * Ps;
    double* g = V + RV_G * Ps;
    double* x0 = V + RV_X0 * Ps;
    double* g0 = V + RV_G0 * Ps;
    double* delta = V + RV_DELTA * Ps;
    double* gamma = V + RV_GAMMA * Ps;
    double* ztg = V + RV_ZTG * Ps;
    double* d = V + RV_D * Ps;
    double nd = 0.0, t0 = 0.0, dtg = 0.0;
    for (int k = tid; k < P; k += nt) {
        const double dk = x[k] - x0[k], gk = g[k] - g0[k];
        delta[k] = dk;
        gamma[k] = gk;
        nd += dk * dk;
        t0 += g0[k] * dk;
        dtg += gk * dk;
    }
    nd = wg_sum(nd, sh);
    t0 = wg_sum(t0, sh);
    dtg = wg_sum(dtg, sh);
    if (dtg < fabs(t0) * 0.1) return false;
    dtg = fmax(kNumPrecision, dtg);
    // Givens rotations folding Z^T g into its first entry (Powell 1987, Eq. 1.7):
    // angles first (sequential in c), then the column pairs (c-1, c) of every row.
    if (tid == 0) {
        double carry = ztg[P - 1];
        for (int c = P - 1; c > 0; --c) {
            const double zc = carry, zm = ztg[c - 1];
            if (zc == 0.0) {
                rc[c] = 1.0;
                rs[c] = 0.0;
                carry = zm;
                continue;
            }
            const double a = sqrt(zm * zm + zc * zc);
            rc[c] = zm / a;
            rs[c] = zc / a;
            carry = a;
        }
    }
    __syncthreads();
    for (int row = tid; row < P; row += nt) {
        double z1 = Z[(size_t)(P - 1) * P + row];
#pragma unroll 8
        for (int c = P - 1; c > 0; --c) {
            const double cv = rc[c], sv = rs[c];
            const double z0 = Z[(size_t)(c - 1) * P + row];
            if (sv == 0.0 && cv == 1.0) {
                Z[(size_t)c * P + row] = z1;
                z1 = z0;
                continue;
            }
            Z[(size_t)c * P + row] = cv * z1 - sv * z0;
            z1 = cv * z0 + sv * z1;
        }
        Z[row] = z1;
    }
    __syncthreads();
    if (norm_Z < 0.0) {
        norm_Z = nd / dtg;
    } else {
        const double t = sqrt(norm_Z * nd / dtg);
        norm_Z = fmin(norm_Z, t);
        norm_Z = fmax(norm_Z, t * 0.1);
    }
    const double sq = sqrt(dtg);
    for (int row = tid; row < P; row += nt) Z[row] = fabs(delta[row]) > kNumPrecision ? delta[row] / sq : 0.0;
    __syncthreads();
    // Column c (c >= 1) per wavefront, its rows over the lanes (contiguous, coalesced).
    const int lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
    for (int c = 1 + wave; c < P; c += nw) {
        const double* zc = Z + (size_t)c * P;
        double s = 0.0;
        for (int row = lane; row < P; row += 64) s += gamma[row] * zc[row];
        s = wave_total(s);
        if (lane == 0) zs[c] = s / dtg;
    }
    for (int i = tid; i < nw * P; i += nt) pd[i] = 0.0;
    __syncthreads();
    // The column updates, and in the same pass the next direction's products with the
    // new Z (PH_DIRECTION: Z^T g per column, and each column's share of Z (Z^T g),
    // summed over the wavefronts' partials afterwards).
    for (int c = wave; c < P; c += nw) {
        double* zc = Z + (size_t)c * P;
        if (c > 0) {
            const double gz = zs[c];
            double nz = 0.0;
            for (int row = lane; row < P; row += 64) {
                const double v = zc[row] - gz * delta[row];
                zc[row] = v;
                nz += v * v;
            }
            nz = wave_total(nz);
            if (nz < norm_Z && nz > kNumPrecision) {
                const double sc = sqrt(norm_Z / nz);
                for (int row = lane; row < P; row += 64) zc[row] *= sc;
            }
        }
        // each lane reads back only the rows it wrote (cache hits)
        double tg = 0.0;
        for (int row = lane; row < P; row += 64) tg += zc[row] * g[row];
        tg = wave_total(tg);
        if (lane == 0) ztg[c] = tg;
        for (int row = lane; row < P; row += 64) pd[wave * P + row] += zc[row] * tg;
    }
    __syncthreads();
    for (int row = tid; row < P; row += nt) {
        double sum = 0.0;
        for (int w = 0; w < nw; ++w) sum += pd[w * P + row];
        d[row] = -sum;
    }
    __syncthreads();
    return true;
}

__global__ __launch_bounds__(kBfgsThreads) void k_bfgs_advance(RunState* __restrict__ states,
                                                               double* __restrict__ vecs, double* __restrict__ Zs,
                                                               const double* __restrict__ fe,
                                                               const double* __restrict__ ge, double* __restrict__ xt,
                                                               int* __restrict__ lists, int* __restrict__ counts,
                                                               const int* __restrict__ run_group, BfgsOpts o,
                                                               int B) {
    __shared__ double sh[8];
    // dynamic LDS (sized by Ps): zs, rot_c, rot_s [Ps] and the wavefronts' partial
    // directions [kBfgsThreads / 64][Ps]
    extern __shared__ double bfgs_lds[];
    const int Ps = o.P;  // stride of the per-run vectors
    double* zs = bfgs_lds;
    double* rot_c = zs + Ps;
    double* rot_s = rot_c + Ps;
    double* pd = rot_s + Ps;
    const int r = blockIdx.x;
    if (r >= B) return;
    const int tid = threadIdx.x, nt = blockDim.x;
    RunState S = states[r];
    if (S.phase == PH_FINISHED) return;
    const int P = S.P;
    double* V = vecs + (size_t)r * RV_COUNT * Ps;
    double* x = V + RV_X * Ps;
    double* g = V + RV_G * Ps;
    double* d = V + RV_D * Ps;
    double* x0 = V + RV_X0 * Ps;
    double* g0 = V + RV_G0 * Ps;
    double* gbest = V + RV_GBEST * Ps;
    double* ztg = V + RV_ZTG * Ps;
    double* best = V + RV_BEST * Ps;
    double* Z = Zs + (size_t)r * Ps * Ps;
    double* xr = xt + (size_t)r * Ps;
    const double* gr = ge + (size_t)r * Ps;
    bool request = false;
    bool dir_ready = false;  // the update already formed d = -Z Z^T g
    for (int guard = 0; guard < 16 && !request && S.phase != PH_FINISHED; ++guard) {
        switch (S.phase) {
        case PH_OUTER_START:
        case PH_RUN_START: {
            // an outer iteration starts from the best point, a further iteration loop
            // from where the last run ended (jittered if it improved the minimum)
            const bool outer = S.phase == PH_OUTER_START;
            for (int k = tid; k < P; k += nt) {
                const double v = outer ? best[k] : x[k];
                x[k] = v;
                xr[k] = v;
            }
            for (size_t i = tid; i < (size_t)P * P; i += nt) Z[i] = (i / P == i % P) ? 1.0 : 0.0;
            S.calls = 0;
            S.norm_Z = -1.0;
            S.status = BS_VARS_INIT;
            S.phase = PH_INIT_WAIT;
            request = true;
            break;
        }
        case PH_INIT_WAIT: {
            S.f = fe[r];
            for (int k = tid; k < P; k += nt) g[k] = gr[k];
            S.calls++;
            S.fprev = fabs(S.f + S.f + 1.0);
            S.phase = PH_DIRECTION;
            __syncthreads();
            break;
        }
        case PH_DIRECTION: {
            if (S.prec == 1 && S.f < o.switch_cost) {
                S.status = BS_MINIMUM;
                S.phase = PH_BFGS_END;
                break;
            }
            // d = -Z Z^T g (BFGS_Powell.cpp:396-439); Z column-major: Z^T g with a
            // column per wavefront (coalesced rows), then Z (Z^T g) with a row per
            // thread -- unless the update just formed both
            if (!dir_ready) {
                const int lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
                for (int c = wave; c < P; c += nw) {
                    const double* zc = Z + (size_t)c * P;
                    double s = 0.0;
                    for (int row = lane; row < P; row += 64) s += zc[row] * g[row];
                    s = wave_total(s);
                    if (lane == 0) ztg[c] = s;
                }
                __syncthreads();
                for (int row = tid; row < P; row += nt) {
                    double s = 0.0;
#pragma unroll 8
                    for (int c = 0; c < P; ++c) s += Z[(size_t)c * P + row] * ztg[c];
                    d[row] = -s;
                }
                __syncthreads();
            }
            dir_ready = false;
            double dg = 0.0, gn = 0.0, ms = INFINITY;
            for (int k = tid; k < P; k += nt) {
                dg += d[k] * g[k];
                gn += g[k] * g[k];
                if (fabs(d[k]) >= 1e-5) ms = fmin(ms, fabs(2.0 * M_PI / d[k]));
            }
            dg = wg_sum(dg, sh);
            gn = wg_sum(gn, sh);
            ms = wg_min(ms, sh);
            double maxstep = 0.0;
            if (dg < 0.0) maxstep = (ms == INFINITY) ? 0.0 : ms;
            if (gn <= 1e-19 * 1e-19) { S.status = BS_MINIMUM; S.phase = PH_BFGS_END; break; }
            if (dg >= 0.0) { S.status = BS_NO_DESCENT; S.phase = PH_BFGS_END; break; }
            if (S.f >= S.fprev) { S.status = BS_MINIMUM; S.phase = PH_BFGS_END; break; }
            S.fprev = S.f;
            if (S.calls >= o.max_calls) { S.status = BS_MAX_ITER; S.phase = PH_BFGS_END; break; }
            // Line search set-up (grad_descend.cpp:127-160).
            S.maxstep = maxstep;
            S.step = fmin(1.0, maxstep);
            S.f_lowest = S.f_highest = S.f;
            S.step_highest = maxstep;
            S.step_lowest = 0.0;
            S.step_best = 0.0;
            S.f_best = S.f;
            S.dg_abs_best = fabs(dg);
            S.dg_highest = 0.0;
            S.dg0 = dg;
            S.ls_iter = 0;
            for (int k = tid; k < P; k += nt) {
                x0[k] = x[k];
                g0[k] = g[k];
                gbest[k] = g[k];
                xr[k] = x[k] + S.step * d[k];
            }
            S.phase = PH_LS_WAIT;
            request = true;
            break;
        }
        case PH_LS_WAIT: {
            // One probe of Grad_Descend::line_search (grad_descend.cpp:163-250).
            S.f = fe[r];
            double dgc = 0.0;
            for (int k = tid; k < P; k += nt) {
                x[k] = xr[k];
                g[k] = gr[k];
                dgc += d[k] * gr[k];
            }
            S.calls++;
            dgc = wg_sum(dgc, sh);
            if (S.f < S.f_best || fabs(dgc) < S.dg_abs_best) {
                S.step_best = S.step;
                S.f_best = S.f;
                S.dg_abs_best = fabs(dgc);
                for (int k = tid; k < P; k += nt) gbest[k] = g[k];
            }
            bool brk = false;
            if (S.calls == o.max_calls) {
                brk = true;
            } else {
                if (S.f < S.f_lowest + S.step * 0.1 * S.dg0) {
                    if (dgc >= S.dg0 * 0.7) {
                        brk = true;
                    } else {
                        S.step_lowest = S.step;
                        S.f_lowest = S.f;
                    }
                } else {
                    S.step_highest = S.step;
                    S.f_highest = S.f;
                    S.dg_highest = dgc;
                }
                if (!brk) {
                    if (S.ls_iter == 0 || S.step_lowest > 0.0) {
                        const double fit = (S.f_highest - S.f_lowest) / (S.step_highest - S.step_lowest) * 2.0 -
                                           S.dg_highest;
                        double scale = -fit * 0.5 / (S.dg_highest - fit);
                        scale = fmax(0.1, scale);
                        S.step = S.step_lowest + scale * (S.step_highest - S.step_lowest);
                    } else {
                        S.step = S.step * 0.1;
                        if (S.step < kNumPrecision) brk = true;
                    }
                }
                if (!brk) {
                    S.ls_iter++;
                    if (S.ls_iter >= 50) brk = true;
                }
            }
            __syncthreads();
            if (!brk) {
                for (int k = tid; k < P; k += nt) xr[k] = x0[k] + S.step * d[k];
                request = true;
                break;
            }
            // End of the line search (grad_descend.cpp:256-270).
            if (S.step != S.step_best) {
                S.step = S.step_best;
                S.f = S.f_best;
                for (int k = tid; k < P; k += nt) {
                    g[k] = gbest[k];
                    x[k] = x0[k] + S.step * d[k];
                }
            }
            __syncthreads();
            if (S.step == 0.0) {
                S.status = BS_ZERO_STEP;
                S.phase = PH_BFGS_END;
                break;
            }
            dir_ready = bfgs_update(Z, V, P, Ps, S.norm_Z, sh, zs, rot_c, rot_s, pd);
            S.phase = PH_DIRECTION;
            break;
        }
        case PH_BFGS_END: {
            S.total_calls += S.calls;
            if (S.prec == 1) {
                // The complex64 stage hands its point to the FP64 stage.
                for (int k = tid; k < P; k += nt) best[k] = x[k];
                S.calls32 += S.calls;
                S.prec = 0;
                S.phase = PH_OUTER_START;
                __syncthreads();
                break;
            }
            // BFGS.cpp:121-133: keep an improvement and jitter the guess by the next
            // U[0,2pi)/100 draws (consumed even when no further loop follows)
            if (S.cur_min > S.f) {
                S.cur_min = S.f;
                const bool more = S.loop + 1 < o.loops;
                const double* jt = (o.jitter && S.jit_used < o.jitter_per_run)
                                       ? o.jitter + ((size_t)r * o.jitter_per_run + S.jit_used) * Ps
                                       : nullptr;
                for (int k = tid; k < P; k += nt) {
                    best[k] = x[k];
                    if (more && jt) x[k] = x[k] + jt[k];
                }
                S.jit_used++;
            }
            S.loop++;
            if (S.loop < o.loops) {
                S.phase = PH_RUN_START;
                __syncthreads();
                break;
            }
            S.loop = 0;
            double mean = 0.0;
            for (int i = kMinWindow - 1; i > 0; --i) {
                S.minvec[i] = S.minvec[i - 1];
                mean += S.minvec[i - 1];
            }
            S.minvec[0] = S.cur_min;
            mean = (mean + S.cur_min) / kMinWindow;
            double sd = 0.0;
            for (int i = 0; i < kMinWindow; ++i) {
                const double t = S.minvec[i] - mean;
                sd += t * t;
            }
            sd = sqrt(sd / (kMinWindow - 1));
            S.outer++;
            __syncthreads();
            if (fabs(sd / S.minvec[kMinWindow - 1]) < o.conv_thr || fabs(S.cur_min) < o.tol || S.outer >= o.max_outer) {
                S.phase = PH_FINISHED;
                S.finished = 1;
                if (tid == 0) atomicAdd(&counts[2], 1);
            } else {
                S.phase = PH_OUTER_START;
            }
            break;
        }
        default:
            break;
        }
    }
    if (request && tid == 0) {
        const int q = 2 * (run_group ? run_group[r] : 0) + S.prec;
        const int j = atomicAdd(&counts[4 + q], 1);
        lists[(size_t)q * B + j] = r;
    }
    if (tid == 0) states[r] = S;
}

hipError_t launch_bfgs_init(void* st, double* vecs, const double* x0, int B, int Ps, const int* run_P, int mixed,
                            int* counts, int n_groups, hipStream_t s) {
    hipLaunchKernelGGL(k_bfgs_init, dim3(B), dim3(256), 0, s, (RunState*)st, vecs, x0, B, Ps, run_P, mixed, counts,
                       4 + 2 * n_groups);
    return hipGetLastError();
}

hipError_t launch_bfgs_reset_counts(int* counts, int n_groups, hipStream_t s) {
    hipLaunchKernelGGL(k_bfgs_reset_counts, dim3(1), dim3(64), 0, s, counts, n_groups);
    return hipGetLastError();
}

hipError_t launch_bfgs_advance(void* st, double* vecs, double* Z, const double* fe, const double* ge, double* xt,
                               int* lists, int* counts, const int* run_group, const BfgsOpts& o, int B,
                               hipStream_t s) {
    const size_t shm = sizeof(double) * (size_t)o.P * (3 + kBfgsThreads / 64);
    hipLaunchKernelGGL(k_bfgs_advance, dim3(B), dim3(kBfgsThreads), shm, s, (RunState*)st, vecs, Z, fe, ge, xt,
                       lists, counts, run_group, o, B);
    return hipGetLastError();
}

size_t bfgs_state_bytes() { return sizeof(RunState); }
int bfgs_max_params() { return kBfgsMaxP; }

}  // namespace sqh
