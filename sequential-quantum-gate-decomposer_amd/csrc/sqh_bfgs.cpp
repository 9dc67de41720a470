// sqh_bfgs.cpp -- host driver of the batched decomposition optimiser.
//
// n_runs independent decompositions (each a circuit, a target and a start point) run
// as lockstep rounds on the context's stream: advance every run's BFGS state machine
// (consuming the evaluation it asked for), then evaluate all requested points with
// one batched cost+gradient launch per qubit-count group and precision.
// rounds_per_graph rounds are captured once as a hipGraph and replayed until every
// run has finished.
//
// sqh_decompose_batch: every run uses the context's circuit.
// sqh_decompose_problems: every run brings its own circuit (any qubit count up to
// the tile limit); runs with the same qubit count share a launch, each wavefront
// reading its run's program through a per-run table.
#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "sqh_bfgs_device.h"
#include "sqh_context.h"
#include "sqh_solve.h"

namespace sqh {

hipError_t launch_bfgs_init(void* st, double* vecs, const double* x0, int B, int Ps, const int* run_P, int mixed,
                            int* counts, int n_groups, hipStream_t s);
hipError_t launch_bfgs_reset_counts(int* counts, int n_groups, hipStream_t s);
hipError_t launch_bfgs_advance(void* st, double* vecs, double* Z, const double* fe, const double* ge, double* xt,
                               int* lists, int* counts, const int* run_group, const BfgsOpts& o, int B,
                               hipStream_t s);
size_t bfgs_state_bytes();
int bfgs_max_params();

namespace {

// One launch group: runs on the same number of qubits.
struct Group {
    int N = 0, rows = 0, cols = 0;
    int B = 0;          // runs in the group (grid bound of its launches)
    int n_targets = 0;  // target matrices of the group
    int ns = 0;         // slot stride of the kernel and correlation tables
    int m_pad = 0, wpb = 0;
    const Op* ops = nullptr;  // single circuit: its ops; per-run circuits: concatenation
    int nops = 0;
    const Layout* lays = nullptr;
    const Slot* slots = nullptr;
    const ProgRef* run_prog = nullptr;
    int ops_cap = 0, lays_cap = 0;  // largest program of the group (LDS staging)
    DevBuf tgt64, tgt32, d_ops, d_lays, d_slots;
    DevBuf kern64, kern32, diag, mpart;  // the group's own evaluation scratch
    DevBuf seed;                         // Hilbert-Schmidt corrections: per-element seed weights
    // the fused-block gradient (single circuit, where it applies): plan per precision
    // and its scratch (block matrices, correlation partials, gate correlations)
    bool fused[2] = {false, false};
    FusedGradPlan fplan[2];
    DevBuf ffused, fmpart, fcorr;
    ~Group() {
        for (DevBuf* b : {&tgt64, &tgt32, &d_ops, &d_lays, &d_slots, &kern64, &kern32, &diag, &mpart, &seed, &ffused,
                          &fmpart, &fcorr})
            b->release();
    }
};

struct BatchBufs {
    DevBuf run_target, run_group, run_P, run_prog, x0, states, vecs, Z, fe, ge, xt, lists, counts, jitter;
    ~BatchBufs() {
        for (DevBuf* b : {&run_target, &run_group, &run_P, &run_prog, &x0, &states, &vecs, &Z, &fe, &ge, &xt, &lists,
                          &counts, &jitter})
            b->release();
    }
};

// The evaluation launch of group g in precision p (f32 when p == 1).
TileLaunch group_launch(sqh_ctx* c, BatchBufs& bb, Group& gr, int g, int p, int B, bool grouped) {
    const bool f32 = p == 1;
    TileLaunch t;
    t.N = gr.N;
    t.f32 = f32;
    t.ops = gr.ops;
    t.nops = gr.nops;
    t.lays = gr.lays;
    t.slots = gr.slots;
    t.ns = gr.ns;
    t.kern = f32 ? gr.kern32.p : gr.kern64.p;
    t.in = f32 ? gr.tgt32.p : gr.tgt64.p;
    t.in_stride = gr.cols;
    t.m = gr.cols;
    t.m_pad = gr.m_pad;
    t.B = gr.B;
    t.kern_bstride = (long long)gr.ns * 4;
    t.trace_off = uses_offset(c->variant) && !grouped ? c->trace_offset : 0;
    t.diag = gr.diag.as<double>();
    t.ts = trace_spec(c, gr.B, gr.cols);
    t.mpart = gr.mpart.as<double>();
    t.bi.count = bb.counts.as<int>() + 4 + 2 * g + p;
    t.bi.elem_run = bb.lists.as<int>() + (size_t)(2 * g + p) * B;
    t.bi.run_target = bb.run_target.as<int>();
    t.bi.tgt_stride = (long long)gr.rows * gr.cols;
    t.bi.run_prog = gr.run_prog;
    t.stage_ops = gr.ops_cap;
    t.stage_lays = gr.lays_cap;
    return t;
}

// One lockstep round: advance all runs, then evaluate the requested points.  With
// several qubit-count groups the groups' evaluation chains are independent (disjoint
// runs, their own scratch), so the round forks them onto side streams after the
// advance and joins them before the next round's advance.
int enqueue_round(sqh_ctx* c, BatchBufs& bb, std::vector<Group>& groups, const BfgsOpts& o, int B, bool mixed,
                  bool grouped) {
    const int G = (int)groups.size();
    HIPTRY(c, launch_bfgs_reset_counts(bb.counts.as<int>(), G, c->stream));
    HIPTRY(c, launch_bfgs_advance(bb.states.p, bb.vecs.as<double>(), bb.Z.as<double>(), bb.fe.as<double>(),
                                  bb.ge.as<double>(), bb.xt.as<double>(), bb.lists.as<int>(), bb.counts.as<int>(),
                                  grouped ? bb.run_group.as<int>() : nullptr, o, B, c->stream));
    const bool fork = G > 1;
    if (fork) {
        HIPTRY(c, hipEventRecord(c->side_ev[0], c->stream));
        for (int g = 0; g < G; ++g) HIPTRY(c, hipStreamWaitEvent(c->side[g], c->side_ev[0], 0));
    }
    for (int g = 0; g < G; ++g) {
        Group& gr = groups[g];
        hipStream_t st = fork ? c->side[g] : c->stream;
        for (int p = 0; p < (mixed ? 2 : 1); ++p) {
            const TileLaunch t = group_launch(c, bb, gr, g, p, B, grouped);
            HIPTRY(c, launch_build_kernels(t.f32, bb.xt.as<double>(), o.P, gr.B, gr.slots, gr.ns,
                                           const_cast<void*>(t.kern), st, t.bi));
            if (gr.fused[p]) {
                // the fused blocks' adjoint: block matrices on the matrix cores forward and
                // backward, each block's correlations resolved to its gates
                const int rc = fused_grad_run(c, t.f32, gr.fplan[p], gr.B, gr.cols, t.kern, t.in, t.diag, t.trace_off,
                                              gr.ffused.p, gr.fmpart.as<double>(), gr.fcorr.as<double>(), t.bi, st,
                                              false);
                if (rc) return rc;
                HIPTRY(c, launch_finish(t.diag, gr.cols, gr.B, c->variant, bb.fe.as<double>(), gr.fcorr.as<double>(), 1,
                                        gr.ns, gr.slots, bb.xt.as<double>(), o.P, bb.ge.as<double>(), st, t.bi, t.ts));
                continue;
            }
            if (!grouped && jit_wants(c, 2)) {
                const int rc = jit_launch(c, t, 2, st);
                if (rc) return rc;
            } else if (t.ts.nch > 0 && !uses_offset(c->variant)) {
                // Hilbert-Schmidt corrections: the points' trace channels first, whose
                // weighted conjugates seed the adjoint sweep
                TileLaunch tf = t;
                tf.mpart = nullptr;
                HIPTRY(c, launch_tile_grad(tf, st));
                HIPTRY(c, launch_finish(t.diag, gr.cols, gr.B, c->variant, nullptr, nullptr, 0, gr.ns, gr.slots,
                                        bb.xt.as<double>(), o.P, nullptr, st, t.bi, t.ts, gr.seed.as<double>()));
                TileLaunch ts = t;
                ts.ts.seed = gr.seed.as<double>();
                HIPTRY(c, launch_tile_grad(ts, st));
            } else {
                HIPTRY(c, launch_tile_grad(t, st));
            }
            HIPTRY(c, launch_finish(t.diag, gr.cols, gr.B, c->variant, bb.fe.as<double>(), t.mpart, gr.wpb, gr.ns,
                                    gr.slots, bb.xt.as<double>(), o.P, bb.ge.as<double>(), st, t.bi, t.ts));
        }
    }
    if (fork) {
        for (int g = 0; g < G; ++g) {
            HIPTRY(c, hipEventRecord(c->side_ev[1 + g], c->side[g]));
            HIPTRY(c, hipStreamWaitEvent(c->stream, c->side_ev[1 + g], 0));
        }
    }
    return SQH_OK;
}

int loops_of(const sqh_decompose_options* opts) { return opts->iteration_loops > 0 ? opts->iteration_loops : 1; }

// The persistent solver (sqh_k_solve.hip) serves up to this many qubits, the cost
// variants without correction channels, and runs capped at this many evaluations.
constexpr int kSolveMaxQubits = 6;
constexpr int kSolveAutoQubits = 6;
constexpr long long kSolveMaxEvals = 1000000;

// Whether the persistent solver can run these groups ("" if so, else the reason).
const char* persistent_blocker(sqh_ctx* c, const std::vector<Group>& groups, int Ps, long long max_evals) {
    if (correction_channels(c->variant) > 0) return "the correction cost variants need the lockstep rounds";
    if (max_evals > kSolveMaxEvals) return "the iteration limits allow more evaluations per run than it runs";
    for (const Group& g : groups) {
        if (g.N > kSolveMaxQubits) return "it serves up to 6 qubits";
        if (solve_lds_bytes(g.N, Ps, std::max(g.ns, 1)) > device_max_lds()) return "the problem needs too much LDS";
    }
    return "";
}

// Every run to completion with persistent launches: groups of more than five qubits
// one launch each, the groups of at most five qubits together in one combined launch
// (when there are several).  The launches run side by side, the largest qubit count on
// the context's stream and every other launch on a forked stream; merging the small
// groups keeps the streams within the hardware queues a process gets by default (four,
// one taken by the default stream), where launches on a shared queue would run one
// after the other.
int solve_persistent(sqh_ctx* c, std::vector<Group>& groups, BatchBufs& bb, const BfgsOpts& o, int B,
                     const int* run_group, bool grouped, long long max_evals) {
    const int G = (int)groups.size();
    std::vector<std::vector<int>> queue(G);
    for (int r = 0; r < B; ++r) queue[run_group ? run_group[r] : 0].push_back(r);
    // the group's arguments (queue fields filled per launch below)
    auto group_args = [&](int g) {
        const Group& gr = groups[g];
        SolveArgs a;
        a.ops = gr.ops;
        a.nops = gr.nops;
        a.lays = gr.lays;
        a.slots = gr.slots;
        a.ns = gr.ns;
        a.ns_max = std::max(gr.ns, 1);
        a.run_prog = gr.run_prog;
        a.run_P = grouped ? bb.run_P.as<int>() : nullptr;
        a.run_target = bb.run_target.as<int>();
        a.tgt64 = gr.tgt64.p;
        a.tgt32 = gr.tgt32.p;
        a.tgt_stride = (long long)gr.rows * gr.cols;
        a.m = gr.cols;
        a.trace_off = uses_offset(c->variant) && !grouped ? c->trace_offset : 0;
        a.variant = c->variant;
        a.states = static_cast<RunState*>(bb.states.p);
        a.vecs = bb.vecs.as<double>();
        a.Zs = bb.Z.as<double>();
        a.x0 = bb.x0.as<double>();
        a.max_evals = max_evals;
        a.o = o;
        return a;
    };
    // launch units: a group on its own, or every small group together (unit `small`)
    std::vector<std::vector<int>> units;
    std::vector<int> small;
    for (int g = 0; g < G; ++g) {
        if (queue[g].empty()) continue;
        if (groups[g].N <= kSolveSmallMaxQubits) small.push_back(g);
        else units.push_back({g});
    }
    if (small.size() == 1) units.push_back(small);
    const bool merged = small.size() > 1;
    if (merged) units.push_back(small);
    auto unit_n = [&](const std::vector<int>& u) {
        int n = 0;
        for (int g : u) n = std::max(n, groups[g].N);
        return n;
    };
    std::stable_sort(units.begin(), units.end(),
                     [&](const std::vector<int>& x, const std::vector<int>& y) { return unit_n(x) > unit_n(y); });
    const int U = (int)units.size();
    // device queues: per unit, its runs (and for the merged unit each entry's qubit count)
    std::vector<int> flat, flat_n;
    std::vector<size_t> first(U);
    for (int u = 0; u < U; ++u) {
        first[u] = flat.size();
        for (int g : units[u])
            for (int r : queue[g]) {
                flat.push_back(r);
                flat_n.push_back(groups[g].N);
            }
    }
    DevBuf d_queue, d_queue_n, d_next;
    HIPTRY(c, d_queue.ensure(std::max<size_t>(flat.size(), 1) * sizeof(int)));
    HIPTRY(c, d_queue_n.ensure(std::max<size_t>(flat.size(), 1) * sizeof(int)));
    HIPTRY(c, d_next.ensure((size_t)std::max(U, 1) * sizeof(int)));
    if (!flat.empty()) {
        HIPTRY(c, hipMemcpyAsync(d_queue.p, flat.data(), flat.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPTRY(c, hipMemcpyAsync(d_queue_n.p, flat_n.data(), flat_n.size() * sizeof(int), hipMemcpyHostToDevice,
                                 c->stream));
    }
    HIPTRY(c, hipMemsetAsync(d_next.p, 0, (size_t)std::max(U, 1) * sizeof(int), c->stream));
    const bool fork = U > 1;
    if (fork) {
        HIPTRY(c, hipEventRecord(c->side_ev[0], c->stream));
        for (int u = 1; u < U; ++u) HIPTRY(c, hipStreamWaitEvent(c->side[u], c->side_ev[0], 0));
    }
    for (int u = 0; u < U; ++u) {
        hipStream_t st = u == 0 ? c->stream : c->side[u];
        const std::vector<int>& un = units[u];
        int n_runs = 0;
        for (int g : un) n_runs += (int)queue[g].size();
        const int grid = std::min(n_runs, 4096);
        if (merged && un == small) {
            SolveSmallArgs sa;
            size_t shm = 0;
            for (int g : un) {
                sa.g[groups[g].N - 1] = group_args(g);
                shm = std::max(shm, solve_lds_bytes(groups[g].N, o.P, std::max(groups[g].ns, 1)));
            }
            sa.next = d_next.as<int>() + u;
            sa.n_runs = n_runs;
            sa.queue = d_queue.as<int>() + first[u];
            sa.queue_n = d_queue_n.as<int>() + first[u];
            HIPTRY(c, launch_solve_small(sa, grid, shm, st));
            continue;
        }
        const int g = un[0];
        SolveArgs a = group_args(g);
        a.next = d_next.as<int>() + u;
        a.n_runs = n_runs;
        a.queue = d_queue.as<int>() + first[u];
        if (!grouped && jit_wants(c, kJitSolve)) {
            const int rc = jit_launch_solve(c, a, grid, solve_lds_bytes(groups[g].N, o.P, a.ns_max), st);
            if (rc) return rc;
        } else {
            HIPTRY(c, launch_solve(groups[g].N, a, grid, st));
        }
    }
    if (fork) {
        for (int u = 1; u < U; ++u) {
            HIPTRY(c, hipEventRecord(c->side_ev[1 + u], c->side[u]));
            HIPTRY(c, hipStreamWaitEvent(c->stream, c->side_ev[1 + u], 0));
        }
    }
    HIPTRY(c, hipStreamSynchronize(c->stream));
    return SQH_OK;
}

constexpr int kMaxJitter = 4096;  // restart-jitter vectors per run

// Jitter vectors per run the optimiser reads (0 when it runs one loop).
int jitter_count(const sqh_decompose_options* opts) {
    return (loops_of(opts) > 1 && opts->restart_jitter) ? opts->restart_jitter_count : 0;
}

// Reads the runs' results back.  rounds < 0: report the largest evaluation count of a
// run (the persistent solver has no lockstep rounds).
int collect(sqh_ctx* c, BatchBufs& bb, int B, int Ps, double* x_out, int x_stride, double* f_out, int* status_out,
            long long* evals_out, long long* rounds_out, long long rounds) {
    std::vector<char> st((size_t)B * bfgs_state_bytes());
    HIPTRY(c, hipMemcpy(st.data(), bb.states.p, st.size(), hipMemcpyDeviceToHost));
    std::vector<double> vecs((size_t)B * RV_COUNT * Ps);
    HIPTRY(c, hipMemcpy(vecs.data(), bb.vecs.p, vecs.size() * sizeof(double), hipMemcpyDeviceToHost));
    long long most = 0;
    for (int r = 0; r < B; ++r) {
        RunState S;
        std::memcpy(&S, st.data() + (size_t)r * bfgs_state_bytes(), sizeof(RunState));
        std::memcpy(x_out + (size_t)r * x_stride, vecs.data() + ((size_t)r * RV_COUNT + RV_BEST) * Ps,
                    (size_t)S.P * sizeof(double));
        f_out[r] = S.cur_min;
        const long long ev = S.total_calls + (S.finished ? 0 : S.calls);
        if (status_out) status_out[r] = S.finished ? S.status : -1;
        if (evals_out) evals_out[r] = ev;
        most = std::max(most, ev);
    }
    if (rounds_out) *rounds_out = rounds >= 0 ? rounds : most;
    return SQH_OK;
}

// Runs the optimiser over B runs in the given groups.  x0: B x Ps (row stride Ps);
// jitter: B x J x Ps restart-jitter vectors (J = max_outer * loops) or empty;
// run_target: index into the run's group targets; run_P / run_group: per run (null
// for the single-circuit case).  Outputs follow sqh_decompose_batch, with x_out
// rows of stride Ps.
int run_batch(sqh_ctx* c, std::vector<Group>& groups, int B, int Ps, const int* run_target, const int* run_P,
              const int* run_group, const double* x0, const std::vector<double>& jitter,
              const sqh_decompose_options* opts, double* x_out, int x_stride, double* f_out, int* status_out,
              long long* evals_out, long long* rounds_out) {
    const bool grouped = run_group != nullptr;
    const bool mixed = opts->mixed_precision != 0;
    const int G = (int)groups.size();
    for (Group& g : groups) {
        const size_t kern_elems = (size_t)g.B * std::max(g.ns, 1) * 4;
        HIPTRY(c, g.kern64.ensure(kern_elems * 16));
        HIPTRY(c, g.kern32.ensure(kern_elems * 8));
        HIPTRY(c, g.diag.ensure((size_t)(correction_channels(c->variant) + 1) * g.B * g.cols * 2 * sizeof(double)));
        HIPTRY(c, g.seed.ensure((size_t)g.B * 6 * sizeof(double)));
        HIPTRY(c, g.mpart.ensure((size_t)g.B * g.wpb * std::max(g.ns, 1) * 8 * sizeof(double)));
        // single circuit, fused path switched on (SQH_FUSED_COST_ON): the fused blocks'
        // adjoint where it applies (per precision used).  The automatic mode keeps the
        // tile kernels here: a round evaluates at most one point per run, and the fused
        // path's four extra small launches per round measured no faster on decompositions
        // of 7-8 qubits.
        size_t fb = 0, mb = 0, cb = 0;
        for (int p = 0; p < (mixed ? 2 : 1); ++p) {
            g.fused[p] = !grouped && c->fused_cost == SQH_FUSED_COST_ON &&
                         fused_grad_applies(c, trace_spec(c, g.B, g.cols), p == 1, g.B, g.cols, &g.fplan[p]);
            if (g.fused[p]) {
                fb = std::max(fb, g.fplan[p].fused_bytes);
                mb = std::max(mb, g.fplan[p].mpart_bytes);
                cb = std::max(cb, g.fplan[p].corr_bytes);
            }
        }
        if (fb) {
            HIPTRY(c, g.ffused.ensure(fb));
            HIPTRY(c, g.fmpart.ensure(mb));
            HIPTRY(c, g.fcorr.ensure(cb));
        }
    }
    if (G > 1) {  // side streams and fork/join events, kept by the context
        while ((int)c->side.size() < G) {
            hipStream_t st = nullptr;
            HIPTRY(c, hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            c->side.push_back(st);
        }
        while ((int)c->side_ev.size() < G + 1) {
            hipEvent_t e = nullptr;
            HIPTRY(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
            c->side_ev.push_back(e);
        }
    }
    BatchBufs bb;
    HIPTRY(c, bb.run_target.ensure((size_t)B * sizeof(int)));
    HIPTRY(c, bb.x0.ensure((size_t)B * Ps * sizeof(double)));
    HIPTRY(c, bb.states.ensure((size_t)B * bfgs_state_bytes()));
    HIPTRY(c, bb.vecs.ensure((size_t)B * RV_COUNT * Ps * sizeof(double)));
    HIPTRY(c, bb.Z.ensure((size_t)B * Ps * Ps * sizeof(double)));
    HIPTRY(c, bb.fe.ensure((size_t)B * sizeof(double)));
    HIPTRY(c, bb.ge.ensure((size_t)B * Ps * sizeof(double)));
    HIPTRY(c, bb.xt.ensure((size_t)B * Ps * sizeof(double)));
    HIPTRY(c, bb.lists.ensure((size_t)2 * G * B * sizeof(int)));
    HIPTRY(c, bb.counts.ensure((size_t)(4 + 2 * G) * sizeof(int)));
    HIPTRY(c, hipMemcpyAsync(bb.run_target.p, run_target, (size_t)B * sizeof(int), hipMemcpyHostToDevice, c->stream));
    if (grouped) {
        HIPTRY(c, bb.run_group.ensure((size_t)B * sizeof(int)));
        HIPTRY(c, bb.run_P.ensure((size_t)B * sizeof(int)));
        HIPTRY(c, hipMemcpyAsync(bb.run_group.p, run_group, (size_t)B * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPTRY(c, hipMemcpyAsync(bb.run_P.p, run_P, (size_t)B * sizeof(int), hipMemcpyHostToDevice, c->stream));
    }
    HIPTRY(c, hipMemcpyAsync(bb.x0.p, x0, (size_t)B * Ps * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPTRY(c, launch_bfgs_init(bb.states.p, bb.vecs.as<double>(), bb.x0.as<double>(), B, Ps,
                               grouped ? bb.run_P.as<int>() : nullptr, mixed ? 1 : 0, bb.counts.as<int>(), G,
                               c->stream));

    BfgsOpts o;
    o.max_calls = opts->max_inner_iterations;
    o.max_outer = opts->max_outer_iterations;
    o.mixed = mixed ? 1 : 0;
    o.tol = opts->optimization_tolerance;
    o.conv_thr = opts->convergence_threshold;
    o.switch_cost = opts->switch_cost;
    o.P = Ps;
    o.loops = loops_of(opts);
    o.jitter_per_run = 0;
    o.jitter = nullptr;
    if (o.loops > 1 && !jitter.empty()) {
        o.jitter_per_run = opts->restart_jitter_count;
        HIPTRY(c, bb.jitter.ensure(jitter.size() * sizeof(double)));
        HIPTRY(c, hipMemcpyAsync(bb.jitter.p, jitter.data(), jitter.size() * sizeof(double), hipMemcpyHostToDevice,
                                 c->stream));
        o.jitter = bb.jitter.as<double>();
    }

    const int R = opts->rounds_per_graph > 0 ? opts->rounds_per_graph : 16;
    const long long max_rounds =
        opts->max_rounds > 0 ? opts->max_rounds
                             : (long long)opts->max_outer_iterations * loops_of(opts) *
                                   (opts->max_inner_iterations + 2) * (mixed ? 2 : 1);
    const char* blocker = persistent_blocker(c, groups, Ps, max_rounds);
    if (opts->engine == SQH_ENGINE_PERSISTENT && *blocker)
        return fail(c, SQH_ERR_ARG, "persistent solver unavailable: %s", blocker);
    // Automatic choice: the persistent solver where it measured faster than the lockstep
    // rounds (every run on at most kSolveAutoQubits qubits; at 6 qubits with four
    // wavefronts per run, one per column slab), and where it applies (no correction
    // channels, the LDS a workgroup may allocate).
    bool auto_persistent = true;
    for (const Group& g : groups) auto_persistent = auto_persistent && g.N <= kSolveAutoQubits;
    const bool persistent = opts->engine == SQH_ENGINE_PERSISTENT ||
                            (opts->engine == SQH_ENGINE_AUTO && !*blocker && auto_persistent);
    if (persistent) {
        int rc = solve_persistent(c, groups, bb, o, B, run_group, grouped, max_rounds);
        if (rc) return rc;
        return collect(c, bb, B, Ps, x_out, x_stride, f_out, status_out, evals_out, rounds_out, -1);
    }
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (!grouped && jit_wants(c, 2)) {
        // compile the run-time kernels outside the capture (a module load is not a
        // stream operation)
        for (int p = 0; p < (mixed ? 2 : 1); ++p) {
            const int rc = jit_prepare(c, group_launch(c, bb, groups[0], 0, p, B, grouped), 2);
            if (rc) return rc;
        }
    }
    HIPTRY(c, hipStreamSynchronize(c->stream));
    HIPTRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < R; ++i) {
        const int rc = enqueue_round(c, bb, groups, o, B, mixed, grouped);
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

    return collect(c, bb, B, Ps, x_out, x_stride, f_out, status_out, evals_out, rounds_out, rounds);
}

int check_opts(sqh_ctx* c, const sqh_decompose_options* opts) {
    if (opts->max_inner_iterations <= 0 || opts->max_outer_iterations <= 0)
        return fail(c, SQH_ERR_ARG, "iteration limits must be positive");
    if (opts->iteration_loops < 0) return fail(c, SQH_ERR_ARG, "iteration_loops must be >= 0");
    if (opts->restart_jitter_count < 0 || opts->restart_jitter_count > kMaxJitter)
        return fail(c, SQH_ERR_ARG, "restart_jitter_count must be 0..%d", kMaxJitter);
    return SQH_OK;
}

// The restart-jitter vectors of one run (J x P from src) at stride Ps in dst.
void copy_jitter(std::vector<double>& dst, int r, int J, int Ps, const double* src, int P) {
    for (int j = 0; j < J; ++j)
        std::memcpy(&dst[((size_t)r * J + j) * Ps], src + (size_t)j * P, (size_t)P * sizeof(double));
}

// Uploads n_targets complex128 targets of a group and their complex64 copies.
int upload_targets(sqh_ctx* c, Group& g, const double* const* targets) {
    const size_t msz = (size_t)g.rows * g.cols;
    HIPTRY(c, g.tgt64.ensure((size_t)g.n_targets * msz * 16));
    HIPTRY(c, g.tgt32.ensure((size_t)g.n_targets * msz * 8));
    // targets adjacent in host memory go up in one copy (a batch of single-circuit
    // targets is one copy instead of one per target)
    for (int i = 0; i < g.n_targets;) {
        int j = i + 1;
        while (j < g.n_targets && targets[j] == targets[j - 1] + msz * 2) ++j;
        HIPTRY(c, hipMemcpyAsync(g.tgt64.as<char>() + (size_t)i * msz * 16, targets[i], (size_t)(j - i) * msz * 16,
                                 hipMemcpyHostToDevice, c->stream));
        i = j;
    }
    HIPTRY(c, launch_convert(true, g.tgt64.p, g.tgt32.p, (long long)g.n_targets * msz * 2, c->stream));
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
    if ((rc = check_opts(c, opts))) return rc;
    HIPTRY(c, hipSetDevice(c->device));

    std::vector<Group> groups(1);
    Group& g = groups[0];
    g.N = N;
    g.rows = rows;
    g.cols = cols;
    g.B = n_runs;
    g.n_targets = n_targets;
    g.ns = ns;
    g.m_pad = m_pad_for(N, cols);
    g.wpb = waves_per_batch(N, g.m_pad);
    g.ops = c->d_ops.as<Op>() + kOpPad;
    g.nops = (int)pr.ops.size();
    g.ops_cap = g.nops;
    g.lays_cap = (int)pr.layouts.size();
    g.lays = c->d_lays.as<Layout>();
    g.slots = pr.slots.empty() ? nullptr : c->d_slots.as<Slot>();
    std::vector<const double*> tp(n_targets);
    for (int i = 0; i < n_targets; ++i) tp[i] = targets + (size_t)i * rows * cols * 2;
    if ((rc = upload_targets(c, g, tp.data()))) return rc;
    std::vector<double> jitter;
    if (jitter_count(opts) > 0)
        jitter.assign(opts->restart_jitter, opts->restart_jitter + (size_t)n_runs * jitter_count(opts) * P);
    return run_batch(c, groups, n_runs, P, run_target, nullptr, nullptr, x0, jitter, opts, x_out, P, f_out,
                     status_out, evals_out, rounds_out);
}

extern "C" int sqh_decompose_problems(sqh_ctx* c, const sqh_problem* problems, int n_problems,
                                      const sqh_decompose_options* opts, double* x_out, double* f_out,
                                      int* status_out, long long* evals_out, long long* rounds_out) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    if (!problems || n_problems <= 0 || !opts || !x_out || !f_out)
        return fail(c, SQH_ERR_ARG, "need problems, options and output buffers");
    int rc = check_opts(c, opts);
    if (rc) return rc;
    // Compile every problem's circuit and sort the runs into qubit-count groups.
    std::vector<Program> progs(n_problems);
    std::map<int, std::vector<int>> by_n;
    int Ps = 1;
    for (int i = 0; i < n_problems; ++i) {
        const sqh_problem& q = problems[i];
        if (q.qbit_num < 1 || q.qbit_num > kMaxTileQubits)
            return fail(c, SQH_ERR_ARG, "problem %d: qubit count must be 1..%d", i, kMaxTileQubits);
        if (!q.target || !q.x0 || (q.n_gates > 0 && !q.gates))
            return fail(c, SQH_ERR_ARG, "problem %d: null gates, target or start point", i);
        std::vector<GateRec> gl(q.n_gates);
        for (int k = 0; k < q.n_gates; ++k)
            gl[k] = GateRec{q.gates[k].type, q.gates[k].target, q.gates[k].control, q.gates[k].param_start};
        std::string err;
        if (!compile_program(q.qbit_num, gl, progs[i], err)) return fail(c, SQH_ERR_GATE, "problem %d: %s", i, err.c_str());
        if (progs[i].P <= 0 || progs[i].P > bfgs_max_params())
            return fail(c, SQH_ERR_ARG, "problem %d: batched optimisation needs 1..%d parameters", i,
                        bfgs_max_params());
        Ps = std::max(Ps, progs[i].P);
        by_n[q.qbit_num].push_back(i);
    }
    HIPTRY(c, hipSetDevice(c->device));
    std::vector<Group> groups(by_n.size());
    std::vector<int> run_target(n_problems), run_group(n_problems), run_P(n_problems);
    std::vector<ProgRef> refs(n_problems);
    std::vector<double> x0((size_t)n_problems * Ps, 0.0);
    int gi = 0;
    for (auto& kv : by_n) {
        Group& g = groups[gi];
        const int N = kv.first;
        const std::vector<int>& members = kv.second;
        g.N = N;
        g.rows = g.cols = 1 << N;
        g.B = (int)members.size();
        g.n_targets = g.B;
        g.m_pad = m_pad_for(N, g.cols);
        const int cw = tile_columns_per_wave(N);
        g.m_pad = (g.m_pad + cw - 1) / cw * cw;  // one element per wavefront
        g.wpb = waves_per_batch(N, g.m_pad);
        Op nop;
        nop.kind = OP_NOP;
        nop.t = nop.cmode = nop.c = 0;
        nop.idx = 0;
        std::vector<Op> ops(kOpPad, nop);
        std::vector<Layout> lays;
        std::vector<Slot> slots;
        std::vector<const double*> tp;
        int ns = 1;
        for (int j = 0; j < g.B; ++j) {
            const int i = members[j];
            const Program& p = progs[i];
            ProgRef& pg = refs[i];
            pg.op_off = (int)ops.size();
            pg.nops = (int)p.ops.size();
            pg.lay_off = (int)lays.size();
            pg.slot_off = (int)slots.size();
            pg.ns = (int)p.slots.size();
            pg.nlays = (int)p.layouts.size();
            pg.pad1 = pg.pad2 = 0;
            g.ops_cap = std::max(g.ops_cap, pg.nops);
            g.lays_cap = std::max(g.lays_cap, pg.nlays);
            ops.insert(ops.end(), p.ops.begin(), p.ops.end());
            ops.insert(ops.end(), kOpPad, nop);
            lays.insert(lays.end(), p.layouts.begin(), p.layouts.end());
            slots.insert(slots.end(), p.slots.begin(), p.slots.end());
            ns = std::max(ns, pg.ns);
            run_target[i] = j;
            run_group[i] = gi;
            run_P[i] = p.P;
            std::memcpy(&x0[(size_t)i * Ps], problems[i].x0, (size_t)p.P * sizeof(double));
            tp.push_back(problems[i].target);
        }
        g.ns = ns;
        HIPTRY(c, g.d_ops.ensure(ops.size() * sizeof(Op)));
        HIPTRY(c, g.d_lays.ensure(std::max<size_t>(lays.size(), 1) * sizeof(Layout)));
        HIPTRY(c, g.d_slots.ensure(std::max<size_t>(slots.size(), 1) * sizeof(Slot)));
        HIPTRY(c, hipMemcpyAsync(g.d_ops.p, ops.data(), ops.size() * sizeof(Op), hipMemcpyHostToDevice, c->stream));
        if (!lays.empty())
            HIPTRY(c, hipMemcpyAsync(g.d_lays.p, lays.data(), lays.size() * sizeof(Layout), hipMemcpyHostToDevice,
                                     c->stream));
        if (!slots.empty())
            HIPTRY(c, hipMemcpyAsync(g.d_slots.p, slots.data(), slots.size() * sizeof(Slot), hipMemcpyHostToDevice,
                                     c->stream));
        g.ops = g.d_ops.as<Op>();  // offsets in the run table point past each program's leading pad
        g.nops = 0;
        g.lays = g.d_lays.as<Layout>();
        g.slots = g.d_slots.as<Slot>();
        if ((rc = upload_targets(c, g, tp.data()))) return rc;
        ++gi;
    }
    std::vector<double> jitter;
    if (jitter_count(opts) > 0) {
        const int J = jitter_count(opts);
        jitter.assign((size_t)n_problems * J * Ps, 0.0);
        size_t off = 0;
        for (int i = 0; i < n_problems; ++i) {
            copy_jitter(jitter, i, J, Ps, opts->restart_jitter + off, progs[i].P);
            off += (size_t)J * progs[i].P;
        }
    }
    DevBuf d_refs;
    HIPTRY(c, d_refs.ensure((size_t)n_problems * sizeof(ProgRef)));
    HIPTRY(c, hipMemcpyAsync(d_refs.p, refs.data(), (size_t)n_problems * sizeof(ProgRef), hipMemcpyHostToDevice,
                             c->stream));
    for (Group& g : groups) g.run_prog = d_refs.as<ProgRef>();
    std::vector<double> xo((size_t)n_problems * Ps);
    rc = run_batch(c, groups, n_problems, Ps, run_target.data(), run_P.data(), run_group.data(), x0.data(), jitter,
                   opts, xo.data(), Ps, f_out, status_out, evals_out, rounds_out);
    if (rc == SQH_OK) {
        size_t o = 0;
        for (int i = 0; i < n_problems; ++i) {
            std::memcpy(x_out + o, &xo[(size_t)i * Ps], (size_t)run_P[i] * sizeof(double));
            o += run_P[i];
        }
    }
    HIPTRY(c, hipStreamSynchronize(c->stream));
    d_refs.release();
    return rc;
}
