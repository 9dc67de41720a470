// sqh_bfgs.cpp -- host driver of the batched decomposition optimiser.
//
// n_runs independent decompositions (each a target index and a start point) run as
// lockstep rounds on the context's stream: advance every run's BFGS state machine
// (consuming the evaluation it asked for), then evaluate all requested points in one
// batched cost+gradient launch per precision.  rounds_per_graph rounds are captured
// once as a hipGraph and replayed until every run has finished.
#include <cstring>
#include <vector>

#include "sqh_bfgs_device.h"
#include "sqh_context.h"

namespace sqh {

hipError_t launch_bfgs_init(void* st, double* vecs, const double* x0, int B, int P, int mixed, int* counts,
                            hipStream_t s);
hipError_t launch_bfgs_reset_counts(int* counts, hipStream_t s);
hipError_t launch_bfgs_advance(void* st, double* vecs, double* Z, const double* fe, const double* ge, double* xt,
                               int* lists, int* counts, const BfgsOpts& o, int B, hipStream_t s);
size_t bfgs_state_bytes();
int bfgs_max_params();

namespace {

struct BatchBufs {
    DevBuf tgt64, tgt32, run_target, x0, states, vecs, Z, fe, ge, xt, lists, counts, kern64, kern32, diag, mpart;
    ~BatchBufs() {
        for (DevBuf* b : {&tgt64, &tgt32, &run_target, &x0, &states, &vecs, &Z, &fe, &ge, &xt, &lists, &counts,
                          &kern64, &kern32, &diag, &mpart})
            b->release();
    }
};

// One lockstep round: advance all runs, then evaluate the requested points.
int enqueue_round(sqh_ctx* c, BatchBufs& bb, const BfgsOpts& o, int B, int rows, int cols, bool mixed) {
    const Program& pr = c->prog;
    const int N = pr.N, P = pr.P, ns = (int)pr.slots.size();
    HIPTRY(c, launch_bfgs_reset_counts(bb.counts.as<int>(), c->stream));
    HIPTRY(c, launch_bfgs_advance(bb.states.p, bb.vecs.as<double>(), bb.Z.as<double>(), bb.fe.as<double>(),
                                  bb.ge.as<double>(), bb.xt.as<double>(), bb.lists.as<int>(), bb.counts.as<int>(), o,
                                  B, c->stream));
    const int off = uses_offset(c->variant) ? c->trace_offset : 0;
    for (int p = 0; p < (mixed ? 2 : 1); ++p) {
        const bool f32 = p == 1;
        TileLaunch t = base_launch(c, f32);
        void* kern = f32 ? bb.kern32.p : bb.kern64.p;
        t.kern = kern;
        t.in = f32 ? bb.tgt32.p : bb.tgt64.p;
        t.in_stride = cols;
        t.m = cols;
        t.m_pad = m_pad_for(N, cols);
        t.B = B;
        t.kern_bstride = (long long)ns * 4;
        t.trace_off = off;
        t.diag = bb.diag.as<double>();
        t.mpart = bb.mpart.as<double>();
        t.bi.count = bb.counts.as<int>() + p;
        t.bi.elem_run = bb.lists.as<int>() + (size_t)p * B;
        t.bi.run_target = bb.run_target.as<int>();
        t.bi.tgt_stride = (long long)rows * cols;
        const int wpb = waves_per_batch(N, t.m_pad);
        HIPTRY(c, launch_build_kernels(f32, bb.xt.as<double>(), P, B, pr.slots.empty() ? nullptr : c->d_slots.as<Slot>(),
                                       ns, kern, c->stream, t.bi));
        HIPTRY(c, launch_tile_grad(t, c->stream));
        HIPTRY(c, launch_finish(t.diag, cols, B, c->variant, bb.fe.as<double>(), t.mpart, wpb, ns,
                                c->d_slots.as<Slot>(), bb.xt.as<double>(), P, bb.ge.as<double>(), c->stream, t.bi));
    }
    return SQH_OK;
}

}  // namespace
}  // namespace sqh

using namespace sqh;

extern "C" int sqh_decompose_batch(sqh_ctx* c, const double* targets, int n_targets, int rows, int cols,
                                   const int* run_target, int n_runs, const double* x0,
                                   const sqh_decompose_options* opts, double* x_out, double* f_out, int* status_out,
                                   long long* evals_out, long long* rounds_out) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = require_circuit(c);
    if (rc) return rc;
    const Program& pr = c->prog;
    const int N = pr.N, P = pr.P, ns = (int)pr.slots.size();
    if (!targets || !run_target || !x0 || !opts || !x_out || !f_out) return fail(c, SQH_ERR_ARG, "null buffer");
    if (n_targets <= 0 || n_runs <= 0) return fail(c, SQH_ERR_ARG, "need at least one target and one run");
    if (rows != (1 << N) || cols <= 0 || cols > rows)
        return fail(c, SQH_ERR_ARG, "targets must be %d x m with m <= %d", 1 << N, 1 << N);
    if (N > kMaxTileQubits) return fail(c, SQH_ERR_ARG, "batched optimisation supports up to %d qubits", kMaxTileQubits);
    if (P <= 0 || P > bfgs_max_params())
        return fail(c, SQH_ERR_ARG, "batched optimisation needs 1..%d parameters", bfgs_max_params());
    const int off = uses_offset(c->variant) ? c->trace_offset : 0;
    if (cols + off > rows) return fail(c, SQH_ERR_ARG, "trace offset exceeds the target");
    for (int r = 0; r < n_runs; ++r)
        if (run_target[r] < 0 || run_target[r] >= n_targets) return fail(c, SQH_ERR_ARG, "run %d: bad target index", r);
    if (opts->max_inner_iterations <= 0 || opts->max_outer_iterations <= 0)
        return fail(c, SQH_ERR_ARG, "iteration limits must be positive");
    HIPTRY(c, hipSetDevice(c->device));

    const int B = n_runs;
    const bool mixed = opts->mixed_precision != 0;
    const size_t msz = (size_t)rows * cols;
    const int m_pad = m_pad_for(N, cols);
    const int wpb = waves_per_batch(N, m_pad);
    BatchBufs bb;
    HIPTRY(c, bb.tgt64.ensure((size_t)n_targets * msz * 16));
    HIPTRY(c, bb.tgt32.ensure((size_t)n_targets * msz * 8));
    HIPTRY(c, bb.run_target.ensure((size_t)B * sizeof(int)));
    HIPTRY(c, bb.x0.ensure((size_t)B * P * sizeof(double)));
    HIPTRY(c, bb.states.ensure((size_t)B * bfgs_state_bytes()));
    HIPTRY(c, bb.vecs.ensure((size_t)B * RV_COUNT * P * sizeof(double)));
    HIPTRY(c, bb.Z.ensure((size_t)B * P * P * sizeof(double)));
    HIPTRY(c, bb.fe.ensure((size_t)B * sizeof(double)));
    HIPTRY(c, bb.ge.ensure((size_t)B * P * sizeof(double)));
    HIPTRY(c, bb.xt.ensure((size_t)B * P * sizeof(double)));
    HIPTRY(c, bb.lists.ensure((size_t)2 * B * sizeof(int)));
    HIPTRY(c, bb.counts.ensure(4 * sizeof(int)));
    HIPTRY(c, bb.kern64.ensure((size_t)B * std::max(ns, 1) * 4 * 16));
    HIPTRY(c, bb.kern32.ensure((size_t)B * std::max(ns, 1) * 4 * 8));
    HIPTRY(c, bb.diag.ensure((size_t)B * cols * 2 * sizeof(double)));
    HIPTRY(c, bb.mpart.ensure((size_t)B * wpb * std::max(ns, 1) * 8 * sizeof(double)));
    HIPTRY(c, hipMemcpyAsync(bb.tgt64.p, targets, (size_t)n_targets * msz * 16, hipMemcpyHostToDevice, c->stream));
    HIPTRY(c, launch_convert(true, bb.tgt64.p, bb.tgt32.p, (long long)n_targets * msz * 2, c->stream));
    HIPTRY(c, hipMemcpyAsync(bb.run_target.p, run_target, (size_t)B * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPTRY(c, hipMemcpyAsync(bb.x0.p, x0, (size_t)B * P * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPTRY(c, launch_bfgs_init(bb.states.p, bb.vecs.as<double>(), bb.x0.as<double>(), B, P, mixed ? 1 : 0,
                               bb.counts.as<int>(), c->stream));

    BfgsOpts o;
    o.max_calls = opts->max_inner_iterations;
    o.max_outer = opts->max_outer_iterations;
    o.mixed = mixed ? 1 : 0;
    o.tol = opts->optimization_tolerance;
    o.conv_thr = opts->convergence_threshold;
    o.switch_cost = opts->switch_cost;
    o.P = P;

    const int R = opts->rounds_per_graph > 0 ? opts->rounds_per_graph : 16;
    const long long max_rounds =
        opts->max_rounds > 0 ? opts->max_rounds
                             : (long long)opts->max_outer_iterations * (opts->max_inner_iterations + 2) * (mixed ? 2 : 1);
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HIPTRY(c, hipStreamSynchronize(c->stream));
    HIPTRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < R; ++i) {
        rc = enqueue_round(c, bb, o, B, rows, cols, mixed);
        if (rc) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(c->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
    }
    HIPTRY(c, hipStreamEndCapture(c->stream, &graph));
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return fail(c, SQH_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    int* h_counts = nullptr;
    e = hipHostMalloc((void**)&h_counts, 4 * sizeof(int), hipHostMallocDefault);
    long long rounds = 0;
    while (e == hipSuccess) {
        e = hipGraphLaunch(exec, c->stream);
        if (e != hipSuccess) break;
        rounds += R;
        e = hipMemcpyAsync(h_counts, bb.counts.p, 4 * sizeof(int), hipMemcpyDeviceToHost, c->stream);
        if (e != hipSuccess) break;
        e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) break;
        if (h_counts[2] >= B || rounds >= max_rounds) break;
    }
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    if (h_counts) (void)hipHostFree(h_counts);
    if (e != hipSuccess) return fail(c, SQH_ERR_DEVICE, "batched optimisation: %s", hipGetErrorString(e));

    std::vector<char> st((size_t)B * bfgs_state_bytes());
    HIPTRY(c, hipMemcpy(st.data(), bb.states.p, st.size(), hipMemcpyDeviceToHost));
    std::vector<double> vecs((size_t)B * RV_COUNT * P);
    HIPTRY(c, hipMemcpy(vecs.data(), bb.vecs.p, vecs.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int r = 0; r < B; ++r) {
        RunState S;
        std::memcpy(&S, st.data() + (size_t)r * bfgs_state_bytes(), sizeof(RunState));
        std::memcpy(x_out + (size_t)r * P, vecs.data() + ((size_t)r * RV_COUNT + RV_BEST) * P, P * sizeof(double));
        f_out[r] = S.cur_min;
        if (status_out) status_out[r] = S.finished ? S.status : -1;
        if (evals_out) evals_out[r] = S.total_calls + (S.finished ? 0 : S.calls);
    }
    if (rounds_out) *rounds_out = rounds;
    return SQH_OK;
}
