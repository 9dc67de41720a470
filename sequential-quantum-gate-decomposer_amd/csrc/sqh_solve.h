// sqh_solve.h -- arguments and launcher of the persistent batched solver (sqh_k_solve.hip).
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

#include "sqh_bfgs_device.h"
#include "sqh_internal.h"

namespace sqh {

// One launch serves the runs of one qubit count: queue[0 .. n_runs) lists them (null:
// runs 0 .. n_runs), taken one at a time by the wavefronts through the counter *next
// (zero before the launch).
struct SolveArgs {
    int* next = nullptr;
    int n_runs = 0;
    const int* queue = nullptr;
    // circuit: one program for every run, or per run through run_prog
    const Op* ops = nullptr;
    int nops = 0;
    const Layout* lays = nullptr;
    const Slot* slots = nullptr;
    int ns = 0;
    int ns_max = 1;  // largest slot count of the launch's programs (LDS tables)
    const ProgRef* run_prog = nullptr;
    const int* run_P = nullptr;       // per-run parameter count (null: o.P)
    // targets: run r reads target run_target[r] (null: 0), rows x m, tgt_stride entries apart
    const int* run_target = nullptr;
    const void* tgt64 = nullptr;
    const void* tgt32 = nullptr;
    long long tgt_stride = 0;
    int m = 0;
    int trace_off = 0;
    int variant = 0;
    // optimiser state (the round kernels' layout)
    RunState* states = nullptr;
    double* vecs = nullptr;
    double* Zs = nullptr;
    const double* x0 = nullptr;
    long long max_evals = 0;  // per run: a run still going after this many is left unfinished
    BfgsOpts o;
};

// Groups of 1..5 qubits in one launch: g[N - 1] describes the group of N qubits (its
// next / n_runs / queue unused); the combined queue lists the runs, queue_n[q] the
// qubit count of entry q.
struct SolveSmallArgs {
    SolveArgs g[5];
    int* next = nullptr;
    int n_runs = 0;
    const int* queue = nullptr;
    const int* queue_n = nullptr;
};

constexpr int kSolveSmallMaxQubits = 5;

#ifndef __HIPCC_RTC__
size_t solve_lds_bytes(int N, int Ps, int ns_max);
hipError_t launch_solve(int N, const SolveArgs& a, int grid, hipStream_t s);
hipError_t launch_solve_small(const SolveSmallArgs& s, int grid, size_t shm, hipStream_t st);
#endif

}  // namespace sqh
