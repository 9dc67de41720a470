// sqh_bfgs_core.h -- one optimisation run's BFGS state machine, advanced by one
// workgroup: shared by the round kernel (k_bfgs_advance, one workgroup of 256 threads
// per run and round) and the persistent solver (k_solve, one wavefront per run,
// looping over its own evaluations).
#pragma once
#include "sqh_bfgs_device.h"
#include "sqh_device.h"

namespace sqh {

constexpr double kNumPrecision = 1.42e-14;  // Grad_Descend::num_precision (grad_descend.cpp:40)
constexpr double kTwoPi = 6.283185307179586476925286766559;

// A double the whole wavefront holds the same value of, as a wave-uniform (scalar)
// value: lane 0's bits.
__device__ inline double uniform_f64(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readfirstlane((int)b);
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the wavefront (DPP row operations), wave-uniform (so a branch on it is
// uniform and the optimiser's state derived from it can live in scalar registers).
__device__ inline double wave_total(double v) { return wave_sum(v); }

// Sum / minimum over the NT threads of the workgroup (every thread gets the result).
template <int NT>
__device__ inline double wg_sum(double v, double* sh) {
    if constexpr (NT == 64) {
        return wave_total(v);
    } else {
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
        __syncthreads();
        if (l == 0) sh[w] = v;
        __syncthreads();
        double t = 0.0;
        for (int i = 0; i < NT / 64; ++i) t += sh[i];
        return t;
    }
}

template <int NT>
__device__ inline double wg_min(double v, double* sh) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    if constexpr (NT == 64) {
        return uniform_f64(v);
    } else {
        const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
        __syncthreads();
        if (l == 0) sh[w] = v;
        __syncthreads();
        double t = sh[0];
        for (int i = 1; i < NT / 64; ++i) t = fmin(t, sh[i]);
        return t;
    }
}

// d = -Z (Z^T g) from ztg = Z^T g (BFGS_Powell.cpp:396-439), threads over the rows
// (coalesced column reads), two rows per thread for twice the loads in flight.
template <int NT>
__device__ inline void bfgs_direction(const double* __restrict__ Z, const double* __restrict__ ztg,
                                      double* __restrict__ d, int P) {
    for (int row = threadIdx.x; row < P; row += 2 * NT) {
        const int r2 = row + NT;
        const bool two = r2 < P;
        double s = 0.0, t = 0.0;
#pragma unroll 16
        for (int c = 0; c < P; ++c) {
            const double zc = ztg[c];
            s += Z[(size_t)c * P + row] * zc;
            t += (two ? Z[(size_t)c * P + r2] : 0.0) * zc;
        }
        d[row] = -s;
        if (two) d[r2] = -t;
    }
    __syncthreads();
}

// BFGS_Powell::BFGS_iteration (BFGS_Powell.cpp:202-386) across the workgroup.
// Z is P x P stored column-major (Z[row][c] at c * P + row), so a column is
// contiguous: the Givens sweep reads whole columns with consecutive threads on
// consecutive rows, the column updates give each thread whole columns.  The rotation angles of the sweep depend
// only on the vector being folded (Z^T g), so one thread computes all of them first
// and every thread then applies the whole sequence to its own rows without barriers.
template <int NT>
__device__ bool bfgs_update(double* __restrict__ Z, double* __restrict__ V, int P, int Ps, double& norm_Z,
                            double* sh, double* zs, double* rc, double* rs, double* pd) {
    const int tid = threadIdx.x, nt = NT;
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
    nd = wg_sum<NT>(nd, sh);
    t0 = wg_sum<NT>(t0, sh);
    dtg = wg_sum<NT>(dtg, sh);
    if (dtg < fabs(t0) * 0.1) return false;
    dtg = fmax(kNumPrecision, dtg);
    // Givens rotations folding Z^T g into its first entry (Powell 1987, Eq. 1.7):
    // angles first (sequential in c), then the column pairs (c-1, c) of every row.
    // The angles' recurrence runs through the carried norm only (a square root per
    // step); the divisions hang off it, so the loop is branch-free and unrolled to let
    // them overlap.  A zero carried entry is an identity rotation.
    if (tid == 0) {
        double carry = ztg[P - 1];
#pragma unroll 4
        for (int c = P - 1; c > 0; --c) {
            const double zc = carry, zm = ztg[c - 1];
            const bool skip = zc == 0.0;
            const double a = sqrt(zm * zm + zc * zc);
            rc[c] = skip ? 1.0 : zm / a;
            rs[c] = skip ? 0.0 : zc / a;
            carry = skip ? zm : a;
        }
    }
    __syncthreads();
    // Each thread rotates two rows at once (two independent chains, twice the loads in
    // flight); an identity rotation (1, 0) leaves the pair as the swap-free update does.
    for (int row = tid; row < P; row += 2 * nt) {
        const int r2 = row + nt;
        const bool two = r2 < P;
        double a1 = Z[(size_t)(P - 1) * P + row];
        double b1 = two ? Z[(size_t)(P - 1) * P + r2] : 0.0;
#pragma unroll 16
        for (int c = P - 1; c > 0; --c) {
            const double cv = rc[c], sv = rs[c];
            const double a0 = Z[(size_t)(c - 1) * P + row];
            const double b0 = two ? Z[(size_t)(c - 1) * P + r2] : 0.0;
            Z[(size_t)c * P + row] = cv * a1 - sv * a0;
            a1 = cv * a0 + sv * a1;
            if (two) Z[(size_t)c * P + r2] = cv * b1 - sv * b0;
            b1 = cv * b0 + sv * b1;
        }
        Z[row] = a1;
        if (two) Z[r2] = b1;
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
    // The column updates with threads over the columns: each thread streams its own
    // columns (contiguous in memory), so its dot products need no cross-lane
    // reductions, and reads delta, gamma and g at the same row as every other thread.
    // The same pass forms Z^T g for the next direction (PH_DIRECTION).
    for (int c = tid; c < P; c += nt) {
        double* zc = Z + (size_t)c * P;
        double tg = 0.0;
        if (c > 0) {
            double s = 0.0;
#pragma unroll 16
            for (int row = 0; row < P; ++row) s += gamma[row] * zc[row];
            const double gz = s / dtg;
            double nz = 0.0;
#pragma unroll 16
            for (int row = 0; row < P; ++row) {
                const double v = zc[row] - gz * delta[row];
                zc[row] = v;
                nz += v * v;
                tg += v * g[row];
            }
            if (nz < norm_Z && nz > kNumPrecision) {
                const double sc = sqrt(norm_Z / nz);
#pragma unroll 8
                for (int row = 0; row < P; ++row) zc[row] *= sc;
                tg *= sc;
            }
        } else {
#pragma unroll 8
            for (int row = 0; row < P; ++row) tg += zc[row] * g[row];
        }
        ztg[c] = tg;
    }
    __syncthreads();
    // d = -Z (Z^T g), threads over the rows (coalesced column reads), two rows each
    bfgs_direction<NT>(Z, ztg, d, P);
    return true;
}

// One run's optimiser, advanced by a workgroup of NT threads: consumes the cost fe_r
// and gradient gr of the point it asked for (phases PH_INIT_WAIT / PH_LS_WAIT), then
// runs until it asks for the next evaluation (its point in xr; returns true) or
// finishes (S.phase == PH_FINISHED; returns false).  V: the run's vectors, Z its
// Powell factor; zs, rot_c, rot_s [Ps] and pd [NT / 64][Ps] are LDS scratch.
template <int NT>
__device__ bool bfgs_advance_run(RunState& S, int r, double* __restrict__ V, double* __restrict__ Z, double fe_r,
                                 const double* __restrict__ gr, double* __restrict__ xr, const BfgsOpts& o,
                                 double* sh, double* zs, double* rot_c, double* rot_s, double* pd) {
    const int tid = threadIdx.x, nt = NT;
    const int Ps = o.P;
    const int P = S.P;
    double* x = V + RV_X * Ps;
    double* g = V + RV_G * Ps;
    double* d = V + RV_D * Ps;
    double* x0 = V + RV_X0 * Ps;
    double* g0 = V + RV_G0 * Ps;
    double* gbest = V + RV_GBEST * Ps;
    double* ztg = V + RV_ZTG * Ps;
    double* best = V + RV_BEST * Ps;
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
            S.f = fe_r;
            for (int k = tid; k < P; k += nt) g[k] = gr[k];
            S.calls++;
            S.fprev = fabs(S.f + S.f + 1.0);
            S.phase = PH_DIRECTION;
            __syncthreads();
            break;
        }
        case PH_DIRECTION: {
            if (S.prec == 1 && S.f < o.switch_cost) {
                // The complex64 descent hands over to FP64 inside the same BFGS run: the
                // point is evaluated again in FP64 and the run goes on with its Powell
                // factor (the inverse-Hessian estimate is kept, not rebuilt from the
                // identity), so the FP64 stage only finishes the descent.
                S.calls32 = S.calls;
                S.prec = 0;
                for (int k = tid; k < P; k += nt) xr[k] = x[k];
                S.phase = PH_INIT_WAIT;
                request = true;
                break;
            }
            // d = -Z Z^T g (BFGS_Powell.cpp:396-439); Z column-major: Z^T g with a
            // column per wavefront (coalesced rows), then Z (Z^T g) with a row per
            // thread -- unless the update just formed both
            if (!dir_ready) {
                for (int c = tid; c < P; c += nt) {  // threads over the columns
                    const double* zc = Z + (size_t)c * P;
                    double s = 0.0;
#pragma unroll 8
                    for (int row = 0; row < P; ++row) s += zc[row] * g[row];
                    ztg[c] = s;
                }
                __syncthreads();
                bfgs_direction<NT>(Z, ztg, d, P);
            }
            dir_ready = false;
            double dg = 0.0, gn = 0.0, ms = __builtin_inf();
            for (int k = tid; k < P; k += nt) {
                dg += d[k] * g[k];
                gn += g[k] * g[k];
                if (fabs(d[k]) >= 1e-5) ms = fmin(ms, fabs(kTwoPi / d[k]));
            }
            dg = wg_sum<NT>(dg, sh);
            gn = wg_sum<NT>(gn, sh);
            ms = wg_min<NT>(ms, sh);
            double maxstep = 0.0;
            if (dg < 0.0) maxstep = (ms == __builtin_inf()) ? 0.0 : ms;
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
            S.f = fe_r;
            double dgc = 0.0;
            for (int k = tid; k < P; k += nt) {
                x[k] = xr[k];
                g[k] = gr[k];
                dgc += d[k] * gr[k];
            }
            S.calls++;
            dgc = wg_sum<NT>(dgc, sh);
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
            dir_ready = bfgs_update<NT>(Z, V, P, Ps, S.norm_Z, sh, zs, rot_c, rot_s, pd);
            S.phase = PH_DIRECTION;
            break;
        }
        case PH_BFGS_END: {
            if (S.prec == 1) {
                // The complex64 stage stopped before reaching switch_cost: it hands over
                // to FP64 the same way, the point evaluated again and the run continued
                // with its Powell factor.
                S.calls32 = S.calls;
                S.prec = 0;
                for (int k = tid; k < P; k += nt) xr[k] = x[k];
                S.phase = PH_INIT_WAIT;
                request = true;
                break;
            }
            S.total_calls += S.calls;
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
            } else {
                S.phase = PH_OUTER_START;
            }
            break;
        }
        default:
            break;
        }
    }
    return request;
}

}  // namespace sqh
