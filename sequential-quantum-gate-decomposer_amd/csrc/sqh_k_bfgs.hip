// sqh_k_bfgs.hip -- kernels of the batched BFGS optimiser (state machine per run).
#include <cfloat>

#include "sqh_bfgs_core.h"
#include "sqh_kernels.h"

namespace sqh {

constexpr int kBfgsThreads = 256;
constexpr int kBfgsMaxP = 1024;

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
    const int tid = threadIdx.x;
    RunState S = states[r];
    if (S.phase == PH_FINISHED) return;
    double* V = vecs + (size_t)r * RV_COUNT * Ps;
    double* Z = Zs + (size_t)r * Ps * Ps;
    double* xr = xt + (size_t)r * Ps;
    const double* gr = ge + (size_t)r * Ps;
    const bool request = bfgs_advance_run<kBfgsThreads>(S, r, V, Z, fe[r], gr, xr, o, sh, zs, rot_c, rot_s, pd);
    if (S.phase == PH_FINISHED && tid == 0) atomicAdd(&counts[2], 1);
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
