// sqh_solve_loop.h -- the persistent solver's run loop (one wavefront per run), shared
// by the compiled kernels (k_solve, generic evaluation) and the run-time compiled ones
// (an evaluation specialised to the circuit).
#pragma once
#include "sqh_bfgs_core.h"
#include "sqh_solve.h"

namespace sqh {

// Takes runs from the work counter until the queue is empty; each run loops: advance
// the BFGS state machine until it asks for an evaluation, evaluate the cost and
// gradient at its point (eval64 / eval32 by the run's current precision), repeat.
// eval(U, ops, nops, lays, slots, ns, xr, gr, kt, mc, lds) returns the cost and
// writes the gradient to gr.  LDS (dynamic): the remap buffer (LDSN complex128
// entries), the kernel table, the correlations, the evaluation point and gradient,
// the optimiser's scratch and the run's state.
// One run to completion (the body of solve_runs): advance the BFGS state machine until
// it asks for an evaluation, evaluate the cost and gradient at its point (eval64 /
// eval32 by the run's current precision), repeat.  eval(U, ops, nops, lays, slots, ns,
// xr, gr, kt, mc, lds) returns the cost and writes the gradient to gr.  LDS (dynamic):
// the remap buffer (LDSN complex128 entries), the kernel table, the correlations, the
// evaluation point and gradient, the optimiser's scratch and the run's state.
template <int LDSN, class E64, class E32>
__device__ void solve_one_run(const SolveArgs& a, int r, E64 eval64, E32 eval32) {
    extern __shared__ __align__(16) unsigned char solve_lds[];
    const BfgsOpts& o = a.o;
    const int Ps = o.P;
    // LDS: remap buffer, kernel table, correlations, evaluation point and gradient,
    // the optimiser's scratch (zs, rot_c, rot_s, one partial-direction row), sh
    char* p = reinterpret_cast<char*>(solve_lds);
    cpx<double>* lds_remap = reinterpret_cast<cpx<double>*>(p);
    p += sizeof(cpx<double>) * (size_t)LDSN;
    cpx<double>* kt = reinterpret_cast<cpx<double>*>(p);
    p += sizeof(cpx<double>) * 4 * (size_t)a.ns_max;
    double* mc = reinterpret_cast<double*>(p);
    p += sizeof(double) * 8 * (size_t)a.ns_max;
    double* xr = reinterpret_cast<double*>(p);
    double* gr = xr + Ps;
    double* zs = gr + Ps;
    double* rot_c = zs + Ps;
    double* rot_s = rot_c + Ps;
    double* pd = rot_s + Ps;
    double* sh = pd + Ps;
    double* V = sh + 8;  // the run's vectors (RV_COUNT x Ps), in LDS while it runs
    RunState* s_state = reinterpret_cast<RunState*>(V + (size_t)RV_COUNT * Ps);
    const int lane = threadIdx.x;
    // the run's problem
    const Op* ops = a.ops;
    int nops = a.nops;
    const Layout* lays = a.lays;
    const Slot* slots = a.slots;
    int ns = a.ns;
    if (a.run_prog) {
        const ProgRef pg = a.run_prog[r];
        ops += pg.op_off;
        nops = pg.nops;
        lays += pg.lay_off;
        slots += pg.slot_off;
        ns = pg.ns;
    }
    const int P = a.run_P ? a.run_P[r] : Ps;
    const long long t = a.run_target ? a.run_target[r] : 0;
    const cpx<double>* U64 = static_cast<const cpx<double>*>(a.tgt64) + t * a.tgt_stride;
    const cpx<float>* U32 = static_cast<const cpx<float>*>(a.tgt32) + t * a.tgt_stride;
    double* Vg = a.vecs + (size_t)r * RV_COUNT * Ps;
    double* Z = a.Zs + (size_t)r * Ps * Ps;
    for (int k = lane; k < P; k += 64) V[RV_BEST * Ps + k] = a.x0[(size_t)r * Ps + k];
    // the run's state lives in LDS (every lane reads it; all lanes write the same
    // values), which leaves the registers to the evaluation and the vector loops
    __syncthreads();
    RunState& S = *s_state;
    S.phase = PH_OUTER_START;
    S.status = BS_INITIAL;
    S.prec = o.mixed ? 1 : 0;
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
    S.cur_min = __DBL_MAX__;
    for (int i = 0; i < kMinWindow; ++i) S.minvec[i] = 0.0;
    __syncthreads();
    double fe = 0.0;
    long long evals = 0;
    int idle = 0;
    for (;;) {
        const bool request = bfgs_advance_run<64>(S, r, V, Z, fe, gr, xr, o, sh, zs, rot_c, rot_s, pd);
        if (!request) {
            // a call either asks for an evaluation or finishes the run; the bound
            // only keeps a wavefront from spinning should neither happen
            if (S.phase == PH_FINISHED || ++idle >= 4) break;
            continue;
        }
        idle = 0;
        if (evals >= a.max_evals) break;  // cap: reported as unfinished
        ++evals;
        __syncthreads();
        const int prec = S.prec;
        if (prec == 1)
            fe = eval32(U32, ops, nops, lays, slots, ns, xr, gr, kt, mc, reinterpret_cast<cpx<float>*>(lds_remap));
        else
            fe = eval64(U64, ops, nops, lays, slots, ns, xr, gr, kt, mc, lds_remap);
    }
    __syncthreads();
    for (int k = lane; k < P; k += 64) Vg[RV_BEST * Ps + k] = V[RV_BEST * Ps + k];
    if (lane == 0) a.states[r] = S;
    __syncthreads();
}

// Takes runs from the work counter until the queue is empty and runs each to
// completion (one wavefront per run).
template <int LDSN, class E64, class E32>
__device__ void solve_runs(const SolveArgs& a, E64 eval64, E32 eval32) {
    const int lane = threadIdx.x;
    for (;;) {
        int q = 0;
        if (lane == 0) q = atomicAdd(a.next, 1);
        q = __builtin_amdgcn_readfirstlane(__shfl(q, 0, 64));
        if (q >= a.n_runs) return;  // every wavefront leaves once the queue is empty
        solve_one_run<LDSN>(a, a.queue ? a.queue[q] : q, eval64, eval32);
    }
}

}  // namespace sqh

namespace sqh {

// The same run loop with NW > 1 wavefronts per run (a workgroup of 64 NW threads per
// run): the optimiser's state machine runs on all NW wavefronts (as the round kernel's
// advance does: each thread holds the run state in registers, every thread computing
// the same values) and the evaluation splits the target's column slabs over the
// wavefronts.  Between evaluations the state is parked in LDS, so its registers are free
// while the evaluation runs.  eval(..., sh) as for solve_runs, with sh (>= 2 NW doubles)
// for the wavefronts' partial traces.  LDS (dynamic): NW remap buffers of LDSW complex128
// entries each, the kernel table, NW correlation tables, the evaluation point and
// gradient, the optimiser's scratch (zs, rot_c, rot_s, NW partial-direction rows), sh,
// the run's vectors, the parked state and the queue slot.
template <int LDSW, int NW, class E64, class E32>
__device__ void solve_runs_mw(const SolveArgs& a, E64 eval64, E32 eval32) {
    constexpr int NT = 64 * NW;
    extern __shared__ __align__(16) unsigned char solve_lds[];
    const BfgsOpts& o = a.o;
    const int Ps = o.P;
    char* p = reinterpret_cast<char*>(solve_lds);
    cpx<double>* lds_remap = reinterpret_cast<cpx<double>*>(p);
    p += sizeof(cpx<double>) * (size_t)LDSW * NW;
    cpx<double>* kt = reinterpret_cast<cpx<double>*>(p);
    p += sizeof(cpx<double>) * 4 * (size_t)a.ns_max;
    double* mc = reinterpret_cast<double*>(p);
    p += sizeof(double) * 8 * (size_t)a.ns_max * NW;
    double* xr = reinterpret_cast<double*>(p);
    double* gr = xr + Ps;
    double* zs = gr + Ps;
    double* rot_c = zs + Ps;
    double* rot_s = rot_c + Ps;
    double* pd = rot_s + Ps;
    double* sh = pd + (size_t)NW * Ps;
    double* V = sh + 8;
    RunState* s_state = reinterpret_cast<RunState*>(V + (size_t)RV_COUNT * Ps);
    int* s_q = reinterpret_cast<int*>(s_state + 1);
    const int tid = threadIdx.x;
    for (;;) {
        __syncthreads();  // the last run's reads of the queue slot are done
        if (tid == 0) *s_q = atomicAdd(a.next, 1);
        __syncthreads();
        const int q = *s_q;
        if (q >= a.n_runs) return;  // every workgroup leaves once the queue is empty
        const int r = a.queue ? a.queue[q] : q;
        const Op* ops = a.ops;
        int nops = a.nops;
        const Layout* lays = a.lays;
        const Slot* slots = a.slots;
        int ns = a.ns;
        if (a.run_prog) {
            const ProgRef pg = a.run_prog[r];
            ops += pg.op_off;
            nops = pg.nops;
            lays += pg.lay_off;
            slots += pg.slot_off;
            ns = pg.ns;
        }
        const int P = a.run_P ? a.run_P[r] : Ps;
        const long long t = a.run_target ? a.run_target[r] : 0;
        const cpx<double>* U64 = static_cast<const cpx<double>*>(a.tgt64) + t * a.tgt_stride;
        const cpx<float>* U32 = static_cast<const cpx<float>*>(a.tgt32) + t * a.tgt_stride;
        double* Vg = a.vecs + (size_t)r * RV_COUNT * Ps;
        double* Z = a.Zs + (size_t)r * Ps * Ps;
        for (int k = tid; k < P; k += NT) V[RV_BEST * Ps + k] = a.x0[(size_t)r * Ps + k];
        RunState S;
        S.phase = PH_OUTER_START;
        S.status = BS_INITIAL;
        S.prec = o.mixed ? 1 : 0;
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
        S.cur_min = __DBL_MAX__;
        for (int i = 0; i < kMinWindow; ++i) S.minvec[i] = 0.0;
        __syncthreads();
        double fe = 0.0;
        long long evals = 0;
        int idle = 0;
        for (;;) {
            const bool request = bfgs_advance_run<NT>(S, r, V, Z, fe, gr, xr, o, sh, zs, rot_c, rot_s, pd);
            if (!request) {
                if (S.phase == PH_FINISHED || ++idle >= 4) break;
                continue;
            }
            idle = 0;
            if (evals >= a.max_evals) break;  // cap: reported as unfinished
            ++evals;
            if (tid == 0) *s_state = S;
            __syncthreads();
            const int prec = S.prec;
            if (prec == 1)
                fe = eval32(U32, ops, nops, lays, slots, ns, xr, gr, kt, mc, reinterpret_cast<cpx<float>*>(lds_remap),
                            sh);
            else
                fe = eval64(U64, ops, nops, lays, slots, ns, xr, gr, kt, mc, lds_remap, sh);
            S = *s_state;
            __syncthreads();
        }
        __syncthreads();
        for (int k = tid; k < P; k += NT) Vg[RV_BEST * Ps + k] = V[RV_BEST * Ps + k];
        if (tid == 0) a.states[r] = S;
    }
}

}  // namespace sqh
