// sqh_bfgs_device.h -- batched BFGS on the GPU: one workgroup per optimisation run.
//
// Each run is the reference's decomposition driver for one target: BFGS with
// Powell's Z update and the Wolfe line search (BFGS_Powell.cpp:70-386,
// grad_descend.cpp:127-272, 396-448), restarted from the best point for up to
// max_outer outer iterations with the 20-entry convergence window
// (Decomposition_Base.cpp:492-675, 742-754), each outer iteration running the BFGS
// engine's iteration loops with the U[0,2pi)/100 restart jitter after every
// improvement (BFGS.cpp:42-152; the jitter draws come from the host).
// The optimiser is written as a resumable state machine: k_bfgs_advance consumes the
// cost and gradient of the point the run asked for, advances until the next
// evaluation request (or the end), and appends the run to the evaluation list of
// its precision.  All runs move in lockstep, one evaluation per round, so every
// round is one batched cost+gradient launch.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

#include "sqh_internal.h"

namespace sqh {

enum BfgsPhase : int {
    PH_OUTER_START = 0,  // start a BFGS run from best_x
    PH_INIT_WAIT = 1,    // waiting for f, g at the start point
    PH_DIRECTION = 2,    // compute the search direction / convergence tests
    PH_LS_WAIT = 3,      // waiting for f, g at a line-search probe
    PH_BFGS_END = 4,     // a BFGS run ended: iteration-loop and outer-loop bookkeeping
    PH_FINISHED = 5,
    PH_RUN_START = 6,    // start a BFGS run from x (the next iteration loop)
};

// Status codes of the last BFGS run (grad_descend.h order).
enum BfgsStatus : int { BS_INITIAL = 0, BS_VARS_INIT = 1, BS_ZERO_STEP = 2, BS_MAX_ITER = 3, BS_NO_DESCENT = 4,
                        BS_MINIMUM = 5 };

constexpr int kMinWindow = 20;

struct BfgsOpts {
    long long max_calls;     // max_inner_iterations (evaluations per BFGS run)
    int max_outer;           // max_outer_iterations
    int mixed;               // start each run in complex64 / FP32
    double tol;              // optimization_tolerance on the FP64 cost
    double conv_thr;         // convergence_threshold of the minimum window
    double switch_cost;      // FP32 -> FP64 switch once the FP32 cost is below this
    int P;                   // stride of the per-run vectors (largest parameter count)
    int loops;               // BFGS iteration loops per outer iteration (BFGS.cpp:121)
    int jitter_per_run;      // jitter vectors per run in `jitter`
    const double* jitter;    // [run][jitter_per_run][P]: U[0,2pi)/100 draws, or null (zeros)
};

struct alignas(16) RunState {
    int phase, status, prec, outer;
    int ls_iter, finished, P, loop;  // P: this run's parameter count; loop: iteration loop index
    int jit_used, pad1, pad2, pad3;  // jitter vectors consumed
    long long calls, total_calls, calls32;
    double f, fprev, norm_Z;
    double step, f_lowest, f_highest, step_highest, step_lowest, step_best, f_best, dg_abs_best, dg_highest, dg0,
        maxstep;
    double cur_min;
    double minvec[kMinWindow];
};

// Per-run vectors, each P doubles.
enum RunVec : int { RV_X = 0, RV_G, RV_D, RV_X0, RV_G0, RV_GBEST, RV_ZTG, RV_BEST, RV_DELTA, RV_GAMMA, RV_COUNT };

}  // namespace sqh
