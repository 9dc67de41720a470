// sqh_api.cpp -- the C ABI (include/sqhip.h): contexts, device buffers and the
// host-side orchestration of the kernels for apply_to / apply_derivate_to /
// optimization_problem{,_batched,_combined,_combined_unitary}.
#include "sqh_context.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

using namespace sqh;

namespace sqh {

int fail(sqh_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

int m_pad_for(int N, int m) {
    const int cw = tile_columns_per_wave(N);
    if (cw <= 0) return m;
    if (m >= cw) return (m + cw - 1) / cw * cw;
    int p = 1;
    while (p < m) p <<= 1;
    return p;
}

int waves_per_batch(int N, int m_pad) {
    const int cw = tile_columns_per_wave(N);
    return m_pad >= cw ? m_pad / cw : 1;
}

// The FROBENIUS family reads the offset trace (N_Qubit_Decomposition_Cost_Function.cpp:
// 137-158, 191-380); the Hilbert-Schmidt family and INFIDELITY the plain one (:482-664).
bool uses_offset(int variant) {
    return variant == CV_FROBENIUS || variant == CV_FROBENIUS_C1 || variant == CV_FROBENIUS_C2 ||
           variant == CV_ABS_TRACE;
}

static bool valid_variant(int v) {
    return v == CV_FROBENIUS || v == CV_FROBENIUS_C1 || v == CV_FROBENIUS_C2 || v == CV_HS || v == CV_HS_C1 ||
           v == CV_HS_C2 || v == CV_INFIDELITY || v == CV_ABS_TRACE;
}

TraceSpec trace_spec(sqh_ctx* c, int B, int m) {
    TraceSpec ts;
    ts.nch = correction_channels(c->variant);
    ts.chstride = (long long)B * m;
    const double r = std::sqrt(c->prev_cost);
    ts.w1 = ts.nch >= 1 ? r * c->corr_scale1 : 0.0;
    ts.w2 = ts.nch >= 2 ? r * c->corr_scale2 : 0.0;
    return ts;
}

int require_circuit(sqh_ctx* c) {
    if (!c->have_circuit) return fail(c, SQH_ERR_STATE, "no circuit set");
    return SQH_OK;
}

int require_target(sqh_ctx* c) {
    if (!c->have_circuit) return fail(c, SQH_ERR_STATE, "no circuit set");
    if (!c->have_target) return fail(c, SQH_ERR_STATE, "no target unitary set");
    if (c->rows != (1 << c->prog.N))
        return fail(c, SQH_ERR_ARG, "target has %d rows but the circuit acts on %d qubits", c->rows, c->prog.N);
    const int off = uses_offset(c->variant) ? c->trace_offset : 0;
    if (c->cols + off > c->rows)
        return fail(c, SQH_ERR_ARG, "trace offset %d + %d columns exceeds %d rows", off, c->cols, c->rows);
    return SQH_OK;
}

int build_kernels(sqh_ctx* c, const double* d_params, int B, bool f32) {
    const int ns = (int)c->prog.slots.size();
    const size_t esz = f32 ? 8 : 16;
    HIPTRY(c, c->d_kern.ensure((size_t)B * std::max(ns, 1) * 4 * esz));
    HIPTRY(c, launch_build_kernels(f32, d_params, c->prog.P, B, c->d_slots.as<Slot>(), ns, c->d_kern.p, c->stream));
    return SQH_OK;
}

TileLaunch base_launch(sqh_ctx* c, bool f32) {
    TileLaunch t;
    t.N = c->prog.N;
    t.f32 = f32;
    t.ops = c->d_ops.as<Op>() + kOpPad;
    t.nops = (int)c->prog.ops.size();
    t.lays = c->d_lays.as<Layout>();
    t.kern = c->d_kern.p;
    t.ns = (int)c->prog.slots.size();
    t.slots = c->d_slots.as<Slot>();
    return t;
}

// One streaming HBM pass per gate over B matrices, in gate order.  kern holds the
// [B][ns][4] kernel tables (kern_bstride = 0 when the parameters are shared).
int stream_apply(sqh_ctx* c, void* mats, int rows, int cols, int stride, long long bstride, int B, bool f32,
                 long long kern_bstride) {
    for (size_t gi = 0; gi < c->gates.size(); ++gi) {
        const GateRec& g = c->gates[gi];
        int kind, control = -1;
        switch (g.type) {
        case GT_CNOT: kind = OP_X; control = g.control; break;
        case GT_CZ: kind = OP_Z; control = g.control; break;
        case GT_CRY: kind = OP_CU; control = g.control; break;
        default: kind = OP_U; break;
        }
        const int s = (kind == OP_U || kind == OP_CU) ? c->prog.gate_slot[gi] : 0;
        TimedLaunch tl(c);
        HIPTRY(c, launch_stream_gate(f32, mats, bstride, stride, rows, cols, B, kind, g.target, control, c->d_kern.p,
                                     kern_bstride, s, c->stream));
    }
    return SQH_OK;
}

// Row tables of the tile-resident fused cost: per block, the row offsets of the 16
// local indices (bit k <-> the k-th smallest qubit of the block) and the row bases of
// the 2^N / 16 outer indices (bit k <-> the k-th smallest other qubit).
std::vector<int> fused_row_maps(int N, const FusionPlan& fp) {
    const int nb = (int)fp.qubits.size();
    std::vector<int> maps((size_t)nb * kFusedMapStride, 0);
    for (int b = 0; b < nb; ++b) {
        int* mp = maps.data() + (size_t)b * kFusedMapStride;
        const auto& q = fp.qubits[b];
        int outer[16], no = 0;
        for (int k = 0; k < N; ++k)
            if (k != q[0] && k != q[1] && k != q[2] && k != q[3]) outer[no++] = k;
        for (int i = 0; i < 16; ++i) {
            int r = 0;
            for (int k = 0; k < 4; ++k)
                if ((i >> k) & 1) r |= 1 << q[k];
            mp[i] = r;
        }
        for (int o = 0; o < (1 << no); ++o) {
            int r = 0;
            for (int k = 0; k < no; ++k)
                if ((o >> k) & 1) r |= 1 << outer[k];
            mp[16 + o] = r;
        }
    }
    return maps;
}

// Row tables of the transposed-chain cost kernel: per block b, with the support
// S_b = Q_b u Q_{b+1} u ... u Q_{nb-1} (the qubits of the blocks the chain has run,
// from the last), the row offsets of the 16 local indices, the row bases of the
// support's outer indices (bit k <-> the k-th smallest qubit of S_b outside the block),
// log2 of their count and the row bits outside S_b.
std::vector<int> fused_trace_maps(int N, const FusionPlan& fp) {
    const int nb = (int)fp.qubits.size();
    std::vector<int> maps((size_t)nb * kTraceMapStride, 0);
    unsigned S = 0;
    for (int b = nb - 1; b >= 0; --b) {
        int* mp = maps.data() + (size_t)b * kTraceMapStride;
        const auto& q = fp.qubits[b];
        unsigned Q = 0;
        for (int k = 0; k < 4; ++k) Q |= 1u << q[k];
        S |= Q;
        for (int i = 0; i < 16; ++i) {
            int r = 0;
            for (int k = 0; k < 4; ++k)
                if ((i >> k) & 1) r |= 1 << q[k];
            mp[i] = r;
        }
        int outer[16], no = 0;
        for (int k = 0; k < N; ++k)
            if (((S >> k) & 1u) && !((Q >> k) & 1u)) outer[no++] = k;
        for (int o = 0; o < (1 << no); ++o) {
            int r = 0;
            for (int k = 0; k < no; ++k)
                if ((o >> k) & 1) r |= 1 << outer[k];
            mp[16 + o] = r;
        }
        mp[80] = no;
        mp[81] = (int)(~S & ((1u << N) - 1u));
    }
    return maps;
}

// The circuit fuses into 4-qubit blocks the fused kernels take (4 to 10 qubits).
bool fused_plan_fits(const sqh_ctx* c) {
    const int N = c->prog.N;
    return N >= 4 && N <= kFusedCostMaxQubits && c->fplan.k == 4 && !c->fplan.qubits.empty();
}

// Gate arithmetic of the register-tile kernels per matrix element: 14 flop per
// parametric one-qubit gate, 7 per CRY, none for CNOT / CZ.
double tile_gate_flops(const sqh_ctx* c) {
    double w = 0.0;
    for (const GateRec& g : c->gates) w += g.type == GT_CRY ? 7.0 : (g.type == GT_CNOT || g.type == GT_CZ) ? 0.0 : 14.0;
    return w;
}

// Whether the fused blocks do less arithmetic than the gate-by-gate kernels (128 flop
// per element per dense block; the cost + gradient: 4 nb - 1 block products against
// (3 + 8/7) x the forward gate arithmetic).  Shallow circuits, with few gates per
// block, stay on the tile kernels in the automatic mode.
bool fused_pays(const sqh_ctx* c, bool grad) {
    const double nb = (double)c->fplan.qubits.size(), w = tile_gate_flops(c);
    return grad ? (4.0 * nb - 1.0) * 128.0 <= (3.0 + 8.0 / 7.0) * w : nb * 128.0 <= w;
}

// Whether a pass (costs, or the apply_to store pass) runs through the tile-resident
// fused-block kernel: 4 <= N <= 10, the circuit has blocks, and the switch is on -- or
// automatic (the default) where it measured faster on MI355X than the register-tile
// kernels: from 7 qubits, or from 5 when the circuit-specialised kernels are off, and
// the blocks do no more arithmetic than the gates.  Costs also need no correction
// channels (fused_cost_applies).
bool fused_pass_applies(const sqh_ctx* c) {
    const int N = c->prog.N;
    if (!fused_plan_fits(c)) return false;
    if (c->fused_cost == SQH_FUSED_COST_AUTO) return (N >= 7 || (N >= 5 && !c->jit)) && fused_pays(c, false);
    return c->fused_cost == SQH_FUSED_COST_ON;
}

bool fused_cost_applies(const sqh_ctx* c, bool grad, const TraceSpec& ts) {
    return !grad && ts.nch == 0 && fused_pass_applies(c);
}

// Costs of B parameter vectors by the tile-resident fused-block kernels: block
// matrices (from the kernel table kern, or each gate's 2x2 kernel computed from d_params
// while staging) -> one launch over every column -> trace entries into diag.  U: the
// target (rows x m, row stride m), or with in_index element b's target at
// U + in_index[b] * in_bstride.
int fused_cost_eval(sqh_ctx* c, bool f32, int B, const void* kern, const double* d_params, double* diag, int trace_off,
                    const void* U, int m, const int* in_index, long long in_bstride) {
    const FusionPlan& fp = c->fplan;
    const int nb = (int)fp.qubits.size(), ns = (int)c->prog.slots.size();
    const size_t esz = f32 ? 8 : 16;
    HIPTRY(c, c->d_fused.ensure((size_t)B * nb * 256 * esz));
    // the block matrices from the kernel table kern, or with d_params straight from the
    // parameters (no kernel table)
    HIPTRY(c, launch_fuse_blocks(f32, c->d_fgates.p, c->d_ffirst.as<int>(), nb, kern, (long long)ns * 4, B,
                                 c->d_fused.p, c->stream, nullptr, nullptr, d_params, c->prog.P,
                                 c->d_slots.as<Slot>()));
    FusedCostArgs a;
    a.U = U;
    a.in_index = in_index;
    a.in_bstride = in_index ? in_bstride : 0;
    a.in_stride = m;
    a.F_bstride = (long long)nb * 256;
    a.m = m;
    a.trace_off = trace_off;
    a.B = B;
    a.nb = nb;
    a.F = c->d_fused.p;
    a.maps = c->d_fmaps.as<int>();
    a.diag = diag;
    TimedLaunch tl(c);
    return fused_cost_launch(c, f32, a);
}

// Launch plan of the fused-block gradient for the context's circuit, B elements of m
// columns: the shape set by sqh_set_fused_grad_config when it fits, else four 16-column
// tiles per wavefront from 9 qubits or from 64 elements (fewer, larger partials), two
// below (more wavefronts for small batches), and the most wavefronts per workgroup, up
// to four, that divide the column groups -- the fastest shapes measured on MI355X --
// with the scratch it needs.  Returns false when no shape fits the device's LDS.
bool fused_grad_plan(const sqh_ctx* c, bool f32, int B, int m, FusedGradPlan& p) {
    const int N = c->prog.N, nb = (int)c->fplan.qubits.size(), ns = (int)c->prog.slots.size();
    p.T = 0;
    if (c->fg_force_tiles && fused_grad_shape_valid(f32, N, m, nb, c->fg_force_tiles, c->fg_force_waves)) {
        p.T = c->fg_force_tiles;
        p.waves = c->fg_force_waves;
    } else {
        for (int t : {N >= 9 || B >= 64 ? 4 : 2, 4}) {
            for (int w : {4, 2, 1})
                if (fused_grad_shape_valid(f32, N, m, nb, t, w)) {
                    p.T = t;
                    p.waves = w;
                    break;
                }
            if (p.T) break;
        }
    }
    if (!p.T) return false;
    p.wgs = fused_grad_wgs(N, m, p.T, p.waves);
    p.fused_bytes = (size_t)B * nb * 256 * (f32 ? 8 : 16);
    p.mpart_bytes = (size_t)B * p.wgs * nb * 512 * sizeof(double);
    p.corr_bytes = (size_t)B * std::max(ns, 1) * 8 * sizeof(double);
    return true;
}

// Whether a cost + gradient evaluation of B elements of m columns runs through the
// fused blocks: the switch on (SQH_FUSED_COST_ON, or _AUTO from 6 qubits when the blocks
// do less arithmetic than the gates -- where it measured faster than the register-tile
// adjoint at every batch size), 4 <= N <= 10, no correction channels, and a launch
// shape fits.
bool fused_grad_applies(const sqh_ctx* c, const TraceSpec& ts, bool f32, int B, int m, FusedGradPlan* plan) {
    if (ts.nch != 0 || !fused_plan_fits(c)) return false;
    if (c->fused_cost == SQH_FUSED_COST_OFF ||
        (c->fused_cost == SQH_FUSED_COST_AUTO && (c->prog.N < 6 || !fused_pays(c, true))))
        return false;
    FusedGradPlan p;
    if (!fused_grad_plan(c, f32, B, m, p)) return false;
    if (plan) *plan = p;
    return true;
}

// Costs and gradients of B parameter vectors through the fused blocks on stream st:
// block matrices (into fused), the forward + backward kernel (trace entries into diag,
// per-workgroup correlation partials into mpart), then each block's gate correlations
// into corr ([B][ns][8]) for k_finish.  kern: the [B][ns][4] kernel tables; U: the
// target (m columns, row stride m), or with bi.count the optimiser's per-run targets
// (element e's at U + run_target[elem_run[e]] * tgt_stride, *count live elements).
// The buffers hold the plan's sizes; nothing is allocated.
// Whether the fused gradient takes its adjoint vectors from the transposed chain
// (k_fused_grad_chain): the chain switch on, at least two blocks, and the blocks'
// qubits covering the register (the chain then ends in the forward layout).
bool fused_grad_uses_chain(const sqh_ctx* c) {
    const FusionPlan& fp = c->fplan;
    if (!c->fused_chain || fp.qubits.size() < 2) return false;
    unsigned S = 0;
    for (const auto& q : fp.qubits)
        for (int k = 0; k < 4; ++k) S |= 1u << q[k];
    return S == (1u << c->prog.N) - 1u;
}

int fused_grad_run(sqh_ctx* c, bool f32, const FusedGradPlan& p, int B, int m, const void* kern, const void* U,
                   double* diag, int trace_off, void* fused, double* mpart, double* corr, const BatchIndex& bi,
                   hipStream_t st, bool timed) {
    const int N = c->prog.N, nb = (int)c->fplan.qubits.size(), ns = (int)c->prog.slots.size();
    HIPTRY(c, launch_fuse_blocks(f32, c->d_fgates.p, c->d_ffirst.as<int>(), nb, kern, (long long)ns * 4, B, fused, st,
                                 bi.count));
    FusedGradArgs a;
    a.U = U;
    a.m = m;
    a.trace_off = trace_off;
    a.B = B;
    a.nb = nb;
    a.F = fused;
    a.maps = c->d_fmaps.as<int>();
    a.diag = diag;
    a.mpart = mpart;
    a.wgs = p.wgs;
    a.count = bi.count;
    a.elem_run = bi.elem_run;
    a.run_target = bi.run_target;
    a.tgt_stride = bi.tgt_stride;
    if (fused_grad_uses_chain(c)) a.tmaps = c->d_tmaps.as<int>();
    if (timed) {
        TimedLaunch tl(c);
        HIPTRY(c, launch_fused_grad(f32, N, a, p.T, p.waves, st));
    } else {
        HIPTRY(c, launch_fused_grad(f32, N, a, p.T, p.waves, st));
    }
    HIPTRY(c, launch_fused_block_corr(f32, c->d_fgates.p, c->d_ffirst.as<int>(), nb, kern, (long long)ns * 4, fused,
                                      mpart, p.wgs, B, corr, ns, st, bi.count));
    return SQH_OK;
}

// The circuit as fused dense blocks (<= 4 qubits each): one HBM pass per block,
// matrix cores for the 4-qubit FP64 blocks.  kern holds the kernel tables built for
// this call (per matrix when kern_bstride != 0).
int fused_apply(sqh_ctx* c, void* mats, int rows, int cols, int stride, long long bstride, int B, bool f32,
                long long kern_bstride) {
    const FusionPlan& fp = c->fplan;
    const int nb = (int)fp.qubits.size();
    if (nb == 0) return SQH_OK;
    const int Bk = kern_bstride ? B : 1;
    const size_t esz = f32 ? 8 : 16;
    HIPTRY(c, c->d_fused.ensure((size_t)Bk * nb * 256 * esz));
    HIPTRY(c, launch_fuse_blocks(f32, c->d_fgates.p, c->d_ffirst.as<int>(), nb, c->d_kern.p, kern_bstride, Bk,
                                 c->d_fused.p, c->stream));
    BlockCfg cfg;
    int rc = block_config(c, f32, rows, cols, B, &cfg);
    if (rc) return rc;
    for (int blk = 0; blk < nb; ++blk) {
        const char* F = static_cast<const char*>(c->d_fused.p) + (size_t)blk * 256 * esz;
        const int D = 1 << fp.k;
        // blocks are stored 16 x 16; a smaller k uses the leading D x D corner only when k == 4
        if (fp.k != 4) return fail(c, SQH_ERR_ARG, "fused streaming needs at least 4 qubits");
        (void)D;
        TimedLaunch tl(c);
        HIPTRY(c, launch_block_apply(f32, true, mats, bstride, stride, rows, cols, B, F,
                                     Bk > 1 ? (long long)nb * 256 : 0, fp.qubits[blk].data(), fp.k, cfg, c->stream));
    }
    return SQH_OK;
}

void eval_scratch_bytes(sqh_ctx* c, int B, bool f32, bool grad, size_t bytes[4]) {
    const int N = c->prog.N, m = c->cols, ns = std::max((int)c->prog.slots.size(), 1);
    const TraceSpec ts = trace_spec(c, B, m);
    bytes[0] = (size_t)B * ns * 4 * (f32 ? 8 : 16);
    bytes[1] = (size_t)(ts.nch + 1) * B * m * 2 * sizeof(double);
    bytes[2] = grad ? (size_t)B * waves_per_batch(N, m_pad_for(N, m)) * ns * 8 * sizeof(double) : 0;
    bytes[3] = grad ? (size_t)B * 6 * sizeof(double) : 0;
}

// costs (and, with grad, the gradient) for B parameter vectors on the device; with
// `fixed` into the caller's scratch memory (tile path only, nothing allocated).
int eval_batch(sqh_ctx* c, const double* d_params, int B, int precision, double* d_costs, double* d_grad,
               const EvalScratch* fixed) {
    int rc = require_target(c);
    if (rc) return rc;
    if (B <= 0) return fail(c, SQH_ERR_ARG, "batch size must be positive");
    if (precision != SQH_PREC_F64 && precision != SQH_PREC_F32) return fail(c, SQH_ERR_ARG, "bad precision");
    const bool f32 = precision == SQH_PREC_F32;
    const int N = c->prog.N, m = c->cols, ns = (int)c->prog.slots.size();
    if (fixed && N > kMaxTileQubits)
        return fail(c, SQH_ERR_ARG, "evaluation without allocation supports up to %d qubits", kMaxTileQubits);
    const int off = uses_offset(c->variant) ? c->trace_offset : 0;
    const TraceSpec ts = trace_spec(c, B, m);
    const bool hs_corr = ts.nch > 0 && !uses_offset(c->variant);
    void* kern = nullptr;
    double* diag = nullptr;
    // cost-only batches on the fused blocks build their 2x2 kernels inside the block build
    const bool fused_costs = !fixed && fused_cost_applies(c, d_grad != nullptr, ts);
    if (fixed) {
        kern = fixed->kern;
        diag = fixed->diag;
        HIPTRY(c, launch_build_kernels(f32, d_params, c->prog.P, B, c->d_slots.as<Slot>(), ns, kern, c->stream));
    } else {
        if (!fused_costs && (rc = build_kernels(c, d_params, B, f32))) return rc;
        HIPTRY(c, c->d_diag.ensure((size_t)(ts.nch + 1) * B * m * 2 * sizeof(double)));
        kern = c->d_kern.p;
        diag = c->d_diag.as<double>();
    }
    // the run-time kernel when the context uses one (with fixed scratch: only if compiled)
    auto run = [&](const TileLaunch& t, int kind) -> int {
        if (jit_wants(c, kind)) {
            const int r = jit_launch(c, t, kind, c->stream, fixed == nullptr);
            if (r != kJitNotReady) return r;
        }
        return tile_launch(c, t, kind, fixed == nullptr);
    };
    const void* U = f32 ? c->d_U32.p : c->d_U64.p;
    FusedGradPlan fgp;
    if (d_grad && !fixed && fused_grad_applies(c, ts, f32, B, m, &fgp)) {
        HIPTRY(c, c->d_fused.ensure(fgp.fused_bytes));
        HIPTRY(c, c->d_mpart.ensure(fgp.mpart_bytes));
        HIPTRY(c, c->d_fcorr.ensure(fgp.corr_bytes));
        if ((rc = fused_grad_run(c, f32, fgp, B, m, kern, U, diag, off, c->d_fused.p, c->d_mpart.as<double>(),
                                 c->d_fcorr.as<double>(), BatchIndex(), c->stream, true)))
            return rc;
        HIPTRY(c, launch_finish(diag, m, B, c->variant, d_costs, c->d_fcorr.as<double>(), 1, ns, c->d_slots.as<Slot>(),
                                d_params, c->prog.P, d_grad, c->stream, BatchIndex(), ts));
        return SQH_OK;
    }
    if (fused_costs) {
        if ((rc = fused_cost_eval(c, f32, B, nullptr, d_params, diag, off, U, m, nullptr, 0))) return rc;
        HIPTRY(c, launch_finish(diag, m, B, c->variant, d_costs, nullptr, 0, ns, c->d_slots.as<Slot>(), d_params,
                                c->prog.P, nullptr, c->stream, BatchIndex(), ts));
        return SQH_OK;
    }
    if (N <= kMaxTileQubits) {
        TileLaunch t = base_launch(c, f32);
        t.kern = kern;
        t.in = U;
        t.in_bstride = 0;
        t.in_stride = m;
        t.m = m;
        t.m_pad = m_pad_for(N, m);
        t.B = B;
        t.kern_bstride = (long long)ns * 4;
        t.trace_off = off;
        t.diag = diag;
        t.ts = ts;
        const int wpb = waves_per_batch(N, t.m_pad);
        if (d_grad) {
            double* seed = nullptr;
            if (fixed) {
                t.mpart = fixed->mpart;
                seed = fixed->seed;
            } else {
                HIPTRY(c, c->d_mpart.ensure((size_t)B * wpb * std::max(ns, 1) * 8 * sizeof(double)));
                t.mpart = c->d_mpart.as<double>();
            }
            if (hs_corr) {
                // |t_k|^2 terms: a forward pass gives the channel traces, whose conjugates
                // (weighted) seed the adjoint sweep.
                if (!fixed) {
                    HIPTRY(c, c->d_seed.ensure((size_t)B * 6 * sizeof(double)));
                    seed = c->d_seed.as<double>();
                }
                TileLaunch tf = t;
                tf.mode = 1;
                tf.mpart = nullptr;
                HIPTRY(c, launch_tile_forward(tf, c->stream));
                HIPTRY(c, launch_finish(t.diag, m, B, c->variant, nullptr, nullptr, 0, ns, t.slots, d_params,
                                        c->prog.P, nullptr, c->stream, BatchIndex(), ts, seed));
                t.ts.seed = seed;
            }
            if (fixed) {
                if ((rc = run(t, 2))) return rc;
            } else {
                TimedLaunch tl(c);
                if ((rc = run(t, 2))) return rc;
            }
            HIPTRY(c, launch_finish(t.diag, m, B, c->variant, d_costs, t.mpart, wpb, ns, t.slots, d_params, c->prog.P,
                                    d_grad, c->stream, BatchIndex(), ts));
        } else {
            t.mode = 1;
            if (fixed) {
                if ((rc = run(t, 1))) return rc;
            } else {
                TimedLaunch tl(c);
                if ((rc = run(t, 1))) return rc;
            }
            HIPTRY(c, launch_finish(t.diag, m, B, c->variant, d_costs, nullptr, 0, ns, t.slots, d_params, c->prog.P,
                                    nullptr, c->stream, BatchIndex(), ts));
        }
        return SQH_OK;
    }
    if (d_grad) return fail(c, SQH_ERR_ARG, "gradient evaluation supports up to %d qubits", kMaxTileQubits);
    // Streaming path: B copies of the target, one pass per gate, then the diagonal.
    const size_t esz = f32 ? 8 : 16;
    const long long msz = (long long)c->rows * m;
    HIPTRY(c, c->d_mat.ensure((size_t)B * msz * esz));
    HIPTRY(c, launch_broadcast_copy(f32, U, c->rows, m, m, c->d_mat.p, msz, B, c->stream));
    if ((rc = fused_apply(c, c->d_mat.p, c->rows, m, m, msz, B, f32, (long long)ns * 4))) return rc;
    HIPTRY(c, launch_extract_diag(f32, c->d_mat.p, msz, m, m, B, off, c->d_diag.as<double>(), N, ts, c->stream));
    HIPTRY(c, launch_finish(c->d_diag.as<double>(), m, B, c->variant, d_costs, nullptr, 0, ns, c->d_slots.as<Slot>(),
                            d_params, c->prog.P, nullptr, c->stream, BatchIndex(), ts));
    return SQH_OK;
}

int eval_targets(sqh_ctx* c, const double* d_params, int B, const void* d_targets, int rows, int cols,
                 const int* d_index, double* d_costs) {
    int rc = require_circuit(c);
    if (rc) return rc;
    const int N = c->prog.N, m = cols, ns = (int)c->prog.slots.size();
    if (N > kMaxTileQubits) return fail(c, SQH_ERR_ARG, "per-target costs support up to %d qubits", kMaxTileQubits);
    if (B <= 0) return fail(c, SQH_ERR_ARG, "batch size must be positive");
    if (c->variant == CV_HS_C1 || c->variant == CV_HS_C2)
        return fail(c, SQH_ERR_ARG, "per-target costs do not take the corrected HS cost");
    const int off = uses_offset(c->variant) ? c->trace_offset : 0;
    if (rows != (1 << N) || cols <= 0 || cols + off > rows)
        return fail(c, SQH_ERR_ARG, "targets must be %d x m with m + trace offset <= %d", 1 << N, 1 << N);
    const TraceSpec ts = trace_spec(c, B, m);
    const bool fused_costs = fused_cost_applies(c, false, ts);
    if ((rc = build_kernels(c, d_params, B, false))) return rc;
    HIPTRY(c, c->d_diag.ensure((size_t)(ts.nch + 1) * B * m * 2 * sizeof(double)));
    TileLaunch t = base_launch(c, false);
    t.in = d_targets;
    t.in_bstride = (long long)rows * m;
    t.in_stride = m;
    t.in_index = d_index;
    t.m = m;
    t.m_pad = m_pad_for(N, m);
    t.B = B;
    t.kern_bstride = (long long)ns * 4;
    t.trace_off = off;
    t.diag = c->d_diag.as<double>();
    t.ts = ts;
    t.mode = 1;
    if (fused_costs) {
        if ((rc = fused_cost_eval(c, false, B, c->d_kern.p, nullptr, t.diag, off, d_targets, m, d_index,
                                  t.in_bstride)))
            return rc;
    } else {
        TimedLaunch tl(c);
        if (jit_wants(c, 1)) {
            if ((rc = jit_launch(c, t, 1, c->stream))) return rc;
        } else {
            HIPTRY(c, launch_tile_forward(t, c->stream));
        }
    }
    HIPTRY(c, launch_finish(t.diag, m, B, c->variant, d_costs, nullptr, 0, ns, t.slots, d_params, c->prog.P, nullptr,
                            c->stream, BatchIndex(), ts));
    return SQH_OK;
}

int apply_dev(sqh_ctx* c, const double* d_params, int pbstride, void* d_inout, int rows, int cols, int stride,
              long long bstride, int B, bool f32, int path) {
    int rc = require_circuit(c);
    if (rc) return rc;
    const int N = c->prog.N;
    if (rows != (1 << N)) return fail(c, SQH_ERR_ARG, "matrix has %d rows, the circuit needs %d", rows, 1 << N);
    if (cols <= 0 || stride < cols || B <= 0) return fail(c, SQH_ERR_ARG, "bad matrix shape");
    if (pbstride != 0 && pbstride != c->prog.P) return fail(c, SQH_ERR_ARG, "parameter batch stride must be 0 or P");
    const int ns = (int)c->prog.slots.size();
    const int Bk = pbstride ? B : 1;
    if ((rc = build_kernels(c, d_params, Bk, f32))) return rc;
    const long long kb = pbstride ? (long long)ns * 4 : 0;
    if (path == 0 && N > kMaxTileQubits) path = N >= 4 ? 2 : 1;
    if (path == 0 && fused_pass_applies(c)) {
        // the circuit's fused 4-qubit blocks on the matrix cores, every column resident
        const FusionPlan& fp = c->fplan;
        const int nb = (int)fp.qubits.size();
        const size_t esz = f32 ? 8 : 16;
        HIPTRY(c, c->d_fused.ensure((size_t)Bk * nb * 256 * esz));
        HIPTRY(c, launch_fuse_blocks(f32, c->d_fgates.p, c->d_ffirst.as<int>(), nb, c->d_kern.p, kb, Bk, c->d_fused.p,
                                     c->stream));
        FusedCostArgs a;
        a.U = d_inout;
        a.in_bstride = bstride;
        a.in_stride = stride;
        a.m = cols;
        a.B = B;
        a.nb = nb;
        a.F = c->d_fused.p;
        a.F_bstride = Bk > 1 ? (long long)nb * 256 : 0;
        a.maps = c->d_fmaps.as<int>();
        a.out = d_inout;
        a.out_bstride = bstride;
        a.out_stride = stride;
        TimedLaunch tl(c);
        return fused_cost_launch(c, f32, a);
    }
    if (path == 0) {
        TileLaunch t = base_launch(c, f32);
        t.mode = 0;
        t.in = d_inout;
        t.in_bstride = bstride;
        t.in_stride = stride;
        t.out = d_inout;
        t.out_bstride = bstride;
        t.out_stride = stride;
        t.m = cols;
        t.m_pad = m_pad_for(N, cols);
        t.B = B;
        t.kern_bstride = kb;
        TimedLaunch tl(c);
        if (jit_wants(c, 0)) return jit_launch(c, t, 0, c->stream);
        HIPTRY(c, launch_tile_forward(t, c->stream));
        return SQH_OK;
    }
    if (path == 2) return fused_apply(c, d_inout, rows, cols, stride, bstride, B, f32, kb);
    return stream_apply(c, d_inout, rows, cols, stride, bstride, B, f32, kb);
}

// Gates_block::apply_derivate_to: a forward store pass over P elements that share the
// input, element k running the circuit with parameter k's gate replaced by its
// derivative kernel (per-element kernel tables; a controlled gate's control-0 block
// zeroed) -- through the fused blocks where that path applies (fused_pass_applies),
// else the register-tile store kernel.  Output k is at d_out + k * rows * cols, row
// stride cols.
int deriv_dev(sqh_ctx* c, const double* d_params, const void* d_in, int rows, int cols, int stride, void* d_out) {
    int rc = require_circuit(c);
    if (rc) return rc;
    const int N = c->prog.N;
    if (rows != (1 << N)) return fail(c, SQH_ERR_ARG, "matrix has %d rows, the circuit needs %d", rows, 1 << N);
    if (cols <= 0 || stride < cols) return fail(c, SQH_ERR_ARG, "bad matrix shape");
    if (N > kMaxTileQubits)
        return fail(c, SQH_ERR_ARG, "apply_derivate_to supports up to %d qubits", kMaxTileQubits);
    if (c->prog.slots.empty()) return SQH_OK;
    const int ns = (int)c->prog.slots.size(), P = c->prog.P;
    bool controlled = false;
    for (const Op& op : c->prog.ops) controlled |= op.kind == OP_CU;
    if ((rc = build_kernels(c, d_params, 1, false))) return rc;
    HIPTRY(c, c->d_dtab.ensure((size_t)P * ns * 4 * 16));
    HIPTRY(c, c->d_zslot.ensure((size_t)P * sizeof(int)));
    HIPTRY(c, launch_deriv_tables(c->d_kern.p, c->d_slots.as<Slot>(), ns, d_params, P, c->d_dtab.p,
                                  c->d_zslot.as<int>(), c->stream));
    if (fused_pass_applies(c)) {
        // the fused blocks' store pass over the P elements (block matrices from the
        // per-parameter kernel tables), every column resident on the chip
        const int nb = (int)c->fplan.qubits.size();
        HIPTRY(c, c->d_fused.ensure((size_t)P * nb * 256 * 16));
        HIPTRY(c, launch_fuse_blocks(false, c->d_fgates.p, c->d_ffirst.as<int>(), nb, c->d_dtab.p, (long long)ns * 4, P,
                                     c->d_fused.p, c->stream, nullptr, controlled ? c->d_zslot.as<int>() : nullptr));
        FusedCostArgs a;
        a.U = d_in;
        a.in_stride = stride;
        a.m = cols;
        a.B = P;
        a.nb = nb;
        a.F = c->d_fused.p;
        a.F_bstride = (long long)nb * 256;
        a.maps = c->d_fmaps.as<int>();
        a.out = d_out;
        a.out_bstride = (long long)rows * cols;
        a.out_stride = cols;
        TimedLaunch tl(c);
        return fused_cost_launch(c, false, a);
    }
    TileLaunch t = base_launch(c, false);
    t.kern = c->d_dtab.p;
    t.kern_bstride = (long long)ns * 4;
    t.in = d_in;
    t.in_bstride = 0;
    t.in_stride = stride;
    t.out = d_out;
    t.out_bstride = (long long)rows * cols;
    t.out_stride = cols;
    t.m = cols;
    t.m_pad = m_pad_for(N, cols);
    t.B = P;
    t.mode = 0;
    t.zslot = controlled ? c->d_zslot.as<int>() : nullptr;
    TimedLaunch tl(c);
    if (jit_wants(c, 0)) return jit_launch(c, t, 0, c->stream);
    HIPTRY(c, launch_tile_forward(t, c->stream));
    return SQH_OK;
}

}  // namespace sqh

extern "C" {

const char* sqh_version(void) { return "sqhip 0.1 (gfx950)"; }

int sqh_create(int device, sqh_ctx** out) {
    if (!out) return SQH_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return SQH_ERR_DEVICE;
    if (device < 0 || device >= n) return SQH_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return SQH_ERR_DEVICE;
    sqh_ctx* c = new (std::nothrow) sqh_ctx();
    if (!c) return SQH_ERR_NOMEM;
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return SQH_ERR_DEVICE;
    }
    // the allocation-free entry point's scratch address on this device (its failure
    // is reported by that entry point, not here)
    (void)noalloc_scratch_init(device);
    *out = c;
    return SQH_OK;
}

void sqh_destroy(sqh_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    delete c;
}

const char* sqh_last_error(const sqh_ctx* c) { return c ? c->err.c_str() : "null context"; }

void* sqh_get_stream(sqh_ctx* c) { return c ? (void*)c->stream : nullptr; }

int sqh_synchronize(sqh_ctx* c) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    HIPTRY(c, hipSetDevice(c->device));
    HIPTRY(c, hipStreamSynchronize(c->stream));
    return SQH_OK;
}

int sqh_set_circuit(sqh_ctx* c, int qbit_num, const sqh_gate* gates, int n_gates) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    if (n_gates < 0 || (n_gates > 0 && !gates)) return fail(c, SQH_ERR_ARG, "bad gate array");
    std::vector<GateRec> g(n_gates);
    for (int i = 0; i < n_gates; ++i) g[i] = {gates[i].type, gates[i].target, gates[i].control, gates[i].param_start};
    Program p;
    std::string err;
    if (!compile_program(qbit_num, g, p, err)) {
        c->have_circuit = false;
        return fail(c, SQH_ERR_GATE, "%s", err.c_str());
    }
    HIPTRY(c, hipSetDevice(c->device));
    HIPTRY(c, hipStreamSynchronize(c->stream));
    // Device copy of the ops with kOpPad no-op records on each side.
    std::vector<Op> padded(p.ops.size() + 2 * kOpPad);
    Op nop;
    nop.kind = OP_NOP;
    nop.t = nop.cmode = nop.c = 0;
    nop.idx = 0;
    for (auto& o : padded) o = nop;
    std::copy(p.ops.begin(), p.ops.end(), padded.begin() + kOpPad);
    HIPTRY(c, c->d_ops.ensure(padded.size() * sizeof(Op)));
    HIPTRY(c, c->d_lays.ensure(p.layouts.size() * sizeof(Layout)));
    HIPTRY(c, c->d_slots.ensure(std::max<size_t>(p.slots.size(), 1) * sizeof(Slot)));
    HIPTRY(c, hipMemcpy(c->d_ops.p, padded.data(), padded.size() * sizeof(Op), hipMemcpyHostToDevice));
    HIPTRY(c, hipMemcpy(c->d_lays.p, p.layouts.data(), p.layouts.size() * sizeof(Layout), hipMemcpyHostToDevice));
    if (!p.slots.empty())
        HIPTRY(c, hipMemcpy(c->d_slots.p, p.slots.data(), p.slots.size() * sizeof(Slot), hipMemcpyHostToDevice));
    FusionPlan fp = plan_fusion(qbit_num, g, p.gate_slot);
    static_assert(sizeof(FusedRec) == 8, "fused record layout");
    HIPTRY(c, c->d_fgates.ensure(std::max<size_t>(fp.gates.size(), 1) * sizeof(FusedRec)));
    HIPTRY(c, c->d_ffirst.ensure(fp.first.size() * sizeof(int)));
    if (!fp.gates.empty())
        HIPTRY(c, hipMemcpy(c->d_fgates.p, fp.gates.data(), fp.gates.size() * sizeof(FusedRec), hipMemcpyHostToDevice));
    HIPTRY(c, hipMemcpy(c->d_ffirst.p, fp.first.data(), fp.first.size() * sizeof(int), hipMemcpyHostToDevice));
    if (qbit_num >= 4 && qbit_num <= kFusedCostMaxQubits && fp.k == 4 && !fp.qubits.empty()) {
        const std::vector<int> maps = fused_row_maps(qbit_num, fp);
        HIPTRY(c, c->d_fmaps.ensure(maps.size() * sizeof(int)));
        HIPTRY(c, hipMemcpy(c->d_fmaps.p, maps.data(), maps.size() * sizeof(int), hipMemcpyHostToDevice));
        const std::vector<int> tmaps = fused_trace_maps(qbit_num, fp);
        HIPTRY(c, c->d_tmaps.ensure(tmaps.size() * sizeof(int)));
        HIPTRY(c, hipMemcpy(c->d_tmaps.p, tmaps.data(), tmaps.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    c->fplan = std::move(fp);
    c->fused_tuned.clear();
    c->tile_tuned.clear();
    c->prog = std::move(p);
    c->gates = std::move(g);
    c->have_circuit = true;
    c->prog_version++;
    c->err.clear();
    return SQH_OK;
}

int sqh_get_parameter_num(const sqh_ctx* c) {
    if (!c || !c->have_circuit) return SQH_ERR_STATE;
    return c->prog.P;
}

int sqh_set_target(sqh_ctx* c, const double* U, int rows, int cols, int stride) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    if (!U || rows <= 0 || cols <= 0 || cols > rows || stride < cols || (rows & (rows - 1)))
        return fail(c, SQH_ERR_ARG, "target must be 2^n x m (m <= 2^n) with stride >= m");
    HIPTRY(c, hipSetDevice(c->device));
    HIPTRY(c, hipStreamSynchronize(c->stream));
    const size_t n = (size_t)rows * cols;
    HIPTRY(c, c->d_U64.ensure(n * 16));
    HIPTRY(c, c->d_U32.ensure(n * 8));
    HIPTRY(c, hipMemcpy2D(c->d_U64.p, (size_t)cols * 16, U, (size_t)stride * 16, (size_t)cols * 16, rows,
                          hipMemcpyHostToDevice));
    HIPTRY(c, launch_convert(true, c->d_U64.p, c->d_U32.p, (long long)n * 2, c->stream));
    HIPTRY(c, hipStreamSynchronize(c->stream));
    c->rows = rows;
    c->cols = cols;
    c->have_target = true;
    return SQH_OK;
}

int sqh_set_trace_offset(sqh_ctx* c, int trace_offset) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    if (trace_offset < 0) return fail(c, SQH_ERR_ARG, "trace offset must be non-negative");
    c->trace_offset = trace_offset;
    return SQH_OK;
}

int sqh_set_correction_weights(sqh_ctx* c, double prev_cost, double scale1, double scale2) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    if (!(prev_cost >= 0.0) || !std::isfinite(prev_cost) || !std::isfinite(scale1) || !std::isfinite(scale2))
        return fail(c, SQH_ERR_ARG, "correction weights must be finite and prev_cost non-negative");
    c->prev_cost = prev_cost;
    c->corr_scale1 = scale1;
    c->corr_scale2 = scale2;
    return SQH_OK;
}

int sqh_set_cost_function(sqh_ctx* c, int variant) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    if (!valid_variant(variant)) return fail(c, SQH_ERR_ARG, "cost function variant %d not supported", variant);
    c->variant = variant;
    return SQH_OK;
}

int sqh_apply_to(sqh_ctx* c, const void* params, void* inout, int rows, int cols, int stride, int is_f32) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = require_circuit(c);
    if (rc) return rc;
    if (!inout || (!params && c->prog.P > 0)) return fail(c, SQH_ERR_ARG, "null buffer");
    if (rows <= 0 || cols <= 0 || stride < cols) return fail(c, SQH_ERR_ARG, "bad matrix shape");
    HIPTRY(c, hipSetDevice(c->device));
    const int P = c->prog.P;
    std::vector<double> p64(std::max(P, 1));
    for (int i = 0; i < P; ++i) p64[i] = is_f32 ? (double)((const float*)params)[i] : ((const double*)params)[i];
    const size_t esz = is_f32 ? 8 : 16;
    HIPTRY(c, c->d_params.ensure(p64.size() * sizeof(double)));
    HIPTRY(c, c->d_mat.ensure((size_t)rows * cols * esz));
    HIPTRY(c, hipMemcpyAsync(c->d_params.p, p64.data(), p64.size() * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    HIPTRY(c, hipMemcpy2DAsync(c->d_mat.p, (size_t)cols * esz, inout, (size_t)stride * esz, (size_t)cols * esz, rows,
                               hipMemcpyHostToDevice, c->stream));
    if ((rc = apply_dev(c, c->d_params.as<double>(), 0, c->d_mat.p, rows, cols, cols, 0, 1, is_f32 != 0, 0)))
        return rc;
    HIPTRY(c, hipMemcpy2DAsync(inout, (size_t)stride * esz, c->d_mat.p, (size_t)cols * esz, (size_t)cols * esz, rows,
                               hipMemcpyDeviceToHost, c->stream));
    HIPTRY(c, hipStreamSynchronize(c->stream));
    return SQH_OK;
}

int sqh_apply_to_dev(sqh_ctx* c, const double* d_params, int pbstride, void* d_inout, int rows, int cols, int stride,
                     long long bstride, int B, int is_f32) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    HIPTRY(c, hipSetDevice(c->device));
    return apply_dev(c, d_params, pbstride, d_inout, rows, cols, stride, bstride, B, is_f32 != 0, 0);
}

int sqh_apply_to_stream_dev(sqh_ctx* c, const double* d_params, int pbstride, void* d_inout, int rows, int cols,
                            int stride, long long bstride, int B, int is_f32) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    HIPTRY(c, hipSetDevice(c->device));
    return apply_dev(c, d_params, pbstride, d_inout, rows, cols, stride, bstride, B, is_f32 != 0, 1);
}

int sqh_apply_to_fused_dev(sqh_ctx* c, const double* d_params, int pbstride, void* d_inout, int rows, int cols,
                           int stride, long long bstride, int B, int is_f32) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    HIPTRY(c, hipSetDevice(c->device));
    if (c->have_circuit && c->prog.N < 4) return fail(c, SQH_ERR_ARG, "fused streaming needs at least 4 qubits");
    return apply_dev(c, d_params, pbstride, d_inout, rows, cols, stride, bstride, B, is_f32 != 0, 2);
}

int sqh_apply_large_kernel_dev(sqh_ctx* c, const void* d_block, const int* qubits, int k, void* d_inout, int rows,
                               int cols, int stride, long long bstride, int B, int is_f32) {
    if (!c || !d_block || !qubits || !d_inout) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    if (k < 1 || k > 5 || rows < (1 << k) || (rows & (rows - 1)) || cols <= 0 || stride < cols || B <= 0)
        return fail(c, SQH_ERR_ARG, "bad block or matrix shape");
    int n = 0;
    while ((1 << n) < rows) ++n;
    for (int j = 0; j < k; ++j)
        if (qubits[j] < 0 || qubits[j] >= n || (j > 0 && qubits[j] <= qubits[j - 1]))
            return fail(c, SQH_ERR_ARG, "involved qubits must be distinct, sorted and below %d", n);
    HIPTRY(c, hipSetDevice(c->device));
    BlockCfg cfg;
    if (k == 4) {
        int rc = block_config(c, is_f32 != 0, rows, cols, B, &cfg);
        if (rc) return rc;
    }
    HIPTRY(c, launch_block_apply(is_f32 != 0, true, d_inout, bstride, stride, rows, cols, B, d_block, 0, qubits, k,
                                 cfg, c->stream));
    return SQH_OK;
}

int sqh_apply_derivate_to(sqh_ctx* c, const double* params, const double* in, int rows, int cols, int stride,
                          double* out) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = require_circuit(c);
    if (rc) return rc;
    if (!in || !out || (!params && c->prog.P > 0)) return fail(c, SQH_ERR_ARG, "null buffer");
    if (rows <= 0 || cols <= 0 || stride < cols) return fail(c, SQH_ERR_ARG, "bad matrix shape");
    HIPTRY(c, hipSetDevice(c->device));
    const int P = c->prog.P;
    if (P == 0) return SQH_OK;
    const size_t msz = (size_t)rows * cols;
    HIPTRY(c, c->d_params.ensure((size_t)P * sizeof(double)));
    HIPTRY(c, c->d_mat.ensure(msz * 16));
    HIPTRY(c, c->d_out.ensure((size_t)P * msz * 16));
    HIPTRY(c, hipMemcpyAsync(c->d_params.p, params, (size_t)P * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPTRY(c, hipMemcpy2DAsync(c->d_mat.p, (size_t)cols * 16, in, (size_t)stride * 16, (size_t)cols * 16, rows,
                               hipMemcpyHostToDevice, c->stream));
    if ((rc = deriv_dev(c, c->d_params.as<double>(), c->d_mat.p, rows, cols, cols, c->d_out.p))) return rc;
    HIPTRY(c, hipMemcpyAsync(out, c->d_out.p, (size_t)P * msz * 16, hipMemcpyDeviceToHost, c->stream));
    HIPTRY(c, hipStreamSynchronize(c->stream));
    return SQH_OK;
}

int sqh_apply_derivate_to_dev(sqh_ctx* c, const double* d_params, const double* d_in, int rows, int cols, int stride,
                              double* d_out) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    HIPTRY(c, hipSetDevice(c->device));
    return deriv_dev(c, d_params, d_in, rows, cols, stride, d_out);
}

int sqh_optimization_problem_batch_dev(sqh_ctx* c, const double* d_params, int B, int precision, double* d_costs) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    HIPTRY(c, hipSetDevice(c->device));
    return eval_batch(c, d_params, B, precision, d_costs, nullptr);
}

int sqh_optimization_problem_combined_batch_dev(sqh_ctx* c, const double* d_params, int B, int precision,
                                                double* d_f0, double* d_grad) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    HIPTRY(c, hipSetDevice(c->device));
    if (!d_grad) return fail(c, SQH_ERR_ARG, "null gradient buffer");
    return eval_batch(c, d_params, B, precision, d_f0, d_grad);
}

int sqh_optimization_problem_combined_targets_dev(sqh_ctx* c, const double* d_targets, int n_targets,
                                                  const int* d_target_index, const double* d_params, int B,
                                                  int precision, double* d_f0, double* d_grad) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = require_circuit(c);
    if (rc) return rc;
    const int N = c->prog.N, ns = (int)c->prog.slots.size();
    if (!d_targets || n_targets <= 0 || !d_target_index || B <= 0 || !d_f0 || !d_grad)
        return fail(c, SQH_ERR_ARG, "need targets, indices, a batch and output buffers");
    if (N > kMaxTileQubits) return fail(c, SQH_ERR_ARG, "gradient evaluation supports up to %d qubits", kMaxTileQubits);
    if (precision != SQH_PREC_F64 && precision != SQH_PREC_F32) return fail(c, SQH_ERR_ARG, "bad precision");
    if (correction_channels(c->variant) > 0)
        return fail(c, SQH_ERR_ARG, "per-target evaluation does not support the correction cost variants");
    const bool f32 = precision == SQH_PREC_F32;
    const int d = 1 << N;
    const long long msz = (long long)d * d;
    HIPTRY(c, hipSetDevice(c->device));
    const void* tg = d_targets;
    if (f32) {
        HIPTRY(c, c->d_mat.ensure((size_t)n_targets * msz * 8));
        HIPTRY(c, launch_convert(true, d_targets, c->d_mat.p, (long long)n_targets * msz * 2, c->stream));
        tg = c->d_mat.p;
    }
    // element b -> run b -> target d_target_index[b]
    if (c->idx_B != B) {
        std::vector<int> h(B + 1);
        h[0] = B;
        for (int b = 0; b < B; ++b) h[b + 1] = b;
        HIPTRY(c, c->d_idx.ensure(h.size() * sizeof(int)));
        HIPTRY(c, hipMemcpyAsync(c->d_idx.p, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPTRY(c, hipStreamSynchronize(c->stream));
        c->idx_B = B;
    }
    if ((rc = build_kernels(c, d_params, B, f32))) return rc;
    HIPTRY(c, c->d_diag.ensure((size_t)B * d * 2 * sizeof(double)));
    TileLaunch t = base_launch(c, f32);
    t.in = tg;
    t.in_stride = d;
    t.m = d;
    t.m_pad = m_pad_for(N, d);
    t.B = B;
    t.kern_bstride = (long long)ns * 4;
    t.trace_off = 0;
    t.diag = c->d_diag.as<double>();
    const int wpb = waves_per_batch(N, t.m_pad);
    HIPTRY(c, c->d_mpart.ensure((size_t)B * wpb * std::max(ns, 1) * 8 * sizeof(double)));
    t.mpart = c->d_mpart.as<double>();
    t.bi.count = c->d_idx.as<int>();
    t.bi.elem_run = c->d_idx.as<int>() + 1;
    t.bi.run_target = d_target_index;
    t.bi.tgt_stride = msz;
    {
        TimedLaunch tl(c);
        if (jit_wants(c, 2)) {
            if ((rc = jit_launch(c, t, 2, c->stream))) return rc;
        } else {
            HIPTRY(c, launch_tile_grad(t, c->stream));
        }
    }
    HIPTRY(c, launch_finish(t.diag, d, B, c->variant, d_f0, t.mpart, wpb, ns, t.slots, d_params, c->prog.P, d_grad,
                            c->stream, t.bi));
    return SQH_OK;
}

static int host_batch(sqh_ctx* c, const double* params, int B, int precision, double* costs, double* grad) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = require_target(c);
    if (rc) return rc;
    if (B <= 0 || !costs || (!params && c->prog.P > 0)) return fail(c, SQH_ERR_ARG, "bad batch arguments");
    HIPTRY(c, hipSetDevice(c->device));
    const int P = c->prog.P;
    const size_t np = (size_t)B * std::max(P, 1);
    HIPTRY(c, c->d_params.ensure(np * sizeof(double)));
    HIPTRY(c, c->d_res.ensure(((size_t)B + (grad ? np : 0)) * sizeof(double)));
    if (P > 0)
        HIPTRY(c, hipMemcpyAsync(c->d_params.p, params, (size_t)B * P * sizeof(double), hipMemcpyHostToDevice,
                                 c->stream));
    double* d_costs = c->d_res.as<double>();
    double* d_grad = grad ? d_costs + B : nullptr;
    if ((rc = eval_batch(c, c->d_params.as<double>(), B, precision, d_costs, d_grad))) return rc;
    HIPTRY(c, hipMemcpyAsync(costs, d_costs, (size_t)B * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (grad && P > 0)
        HIPTRY(c, hipMemcpyAsync(grad, d_grad, (size_t)B * P * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(c, hipStreamSynchronize(c->stream));
    return SQH_OK;
}

int sqh_optimization_problem_batch(sqh_ctx* c, const double* params, int B, int precision, double* costs) {
    return host_batch(c, params, B, precision, costs, nullptr);
}

int sqh_optimization_problem_combined_batch(sqh_ctx* c, const double* params, int B, int precision, double* f0,
                                            double* grad) {
    if (!grad) return c ? fail(c, SQH_ERR_ARG, "null gradient buffer") : SQH_ERR_ARG;
    return host_batch(c, params, B, precision, f0, grad);
}

int sqh_optimization_problem_combined_unitary(sqh_ctx* c, const double* params, double* out_unitary,
                                              double* out_derivs) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = require_target(c);
    if (rc) return rc;
    if (!out_unitary || (!params && c->prog.P > 0)) return fail(c, SQH_ERR_ARG, "null buffer");
    HIPTRY(c, hipSetDevice(c->device));
    const int P = c->prog.P, rows = c->rows, cols = c->cols;
    const size_t msz = (size_t)rows * cols;
    HIPTRY(c, c->d_params.ensure((size_t)std::max(P, 1) * sizeof(double)));
    HIPTRY(c, c->d_mat.ensure(msz * 16));
    if (P > 0)
        HIPTRY(c, hipMemcpyAsync(c->d_params.p, params, (size_t)P * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (P > 0 && out_derivs) {
        HIPTRY(c, c->d_out.ensure((size_t)P * msz * 16));
        if ((rc = deriv_dev(c, c->d_params.as<double>(), c->d_U64.p, rows, cols, cols, c->d_out.p))) return rc;
    }
    HIPTRY(c, hipMemcpyAsync(c->d_mat.p, c->d_U64.p, msz * 16, hipMemcpyDeviceToDevice, c->stream));
    if ((rc = apply_dev(c, c->d_params.as<double>(), 0, c->d_mat.p, rows, cols, cols, 0, 1, false, 0))) return rc;
    HIPTRY(c, hipMemcpyAsync(out_unitary, c->d_mat.p, msz * 16, hipMemcpyDeviceToHost, c->stream));
    if (P > 0 && out_derivs)
        HIPTRY(c, hipMemcpyAsync(out_derivs, c->d_out.p, (size_t)P * msz * 16, hipMemcpyDeviceToHost, c->stream));
    HIPTRY(c, hipStreamSynchronize(c->stream));
    return SQH_OK;
}

int sqh_set_jit(sqh_ctx* c, int enable) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    c->jit = enable != 0;
    return SQH_OK;
}

int sqh_set_fused_cost(sqh_ctx* c, int mode) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    if (mode != SQH_FUSED_COST_OFF && mode != SQH_FUSED_COST_ON && mode != SQH_FUSED_COST_AUTO)
        return fail(c, SQH_ERR_ARG, "fused cost mode must be 0 (off), 1 (on) or 2 (automatic)");
    c->fused_cost = mode;
    return SQH_OK;
}

int sqh_get_jit_stats(sqh_ctx* c, double* compile_ms, int* compiles) {
    if (!c || !compile_ms || !compiles) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    *compile_ms = jit_compile_ms(c, compiles);
    return SQH_OK;
}

int sqh_set_kernel_timing(sqh_ctx* c, int enable) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    c->timing = enable != 0;
    return SQH_OK;
}

int sqh_get_kernel_timing(sqh_ctx* c, double* total_ms, int* launches) {
    if (!c || !total_ms || !launches) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    HIPTRY(c, hipSetDevice(c->device));
    HIPTRY(c, hipStreamSynchronize(c->stream));
    double tot = 0.0;
    for (auto& e : c->events) {
        float ms = 0.0f;
        HIPTRY(c, hipEventElapsedTime(&ms, e.first, e.second));
        tot += ms;
    }
    *total_ms = tot;
    *launches = (int)c->events.size();
    for (auto& e : c->events) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    c->events.clear();
    return SQH_OK;
}

int sqh_inspect_program(int qbit_num, const sqh_gate* gates, int n_gates, void* ops_out, int ops_cap,
                        void* layouts_out, int layouts_cap, int* n_ops, int* n_layouts) {
    if (n_gates < 0 || (n_gates > 0 && !gates)) return SQH_ERR_ARG;
    std::vector<GateRec> g(n_gates);
    for (int i = 0; i < n_gates; ++i) g[i] = {gates[i].type, gates[i].target, gates[i].control, gates[i].param_start};
    Program p;
    std::string err;
    if (!compile_program(qbit_num, g, p, err)) return SQH_ERR_GATE;
    if (n_ops) *n_ops = (int)p.ops.size();
    if (n_layouts) *n_layouts = (int)p.layouts.size();
    if (ops_out && ops_cap > 0)
        std::memcpy(ops_out, p.ops.data(), sizeof(Op) * std::min<size_t>(p.ops.size(), (size_t)ops_cap));
    if (layouts_out && layouts_cap > 0)
        std::memcpy(layouts_out, p.layouts.data(),
                    sizeof(Layout) * std::min<size_t>(p.layouts.size(), (size_t)layouts_cap));
    return SQH_OK;
}

int sqh_inspect_fusion(int qbit_num, const sqh_gate* gates, int n_gates, int* block_first, int* block_qubits,
                       int blocks_cap, int* src, int src_cap, int* n_blocks) {
    if (n_gates < 0 || (n_gates > 0 && !gates)) return SQH_ERR_ARG;
    std::vector<GateRec> g(n_gates);
    for (int i = 0; i < n_gates; ++i) g[i] = {gates[i].type, gates[i].target, gates[i].control, gates[i].param_start};
    Program p;
    std::string err;
    if (!compile_program(qbit_num, g, p, err)) return SQH_ERR_GATE;
    const FusionPlan fp = plan_fusion(qbit_num, g, p.gate_slot);
    const int nb = (int)fp.qubits.size();
    if (n_blocks) *n_blocks = nb;
    for (int b = 0; b <= nb && block_first && b <= blocks_cap; ++b) block_first[b] = fp.first[b];
    for (int b = 0; b < nb && block_qubits && b < blocks_cap; ++b)
        for (int j = 0; j < 4; ++j) block_qubits[4 * b + j] = j < fp.k ? fp.qubits[b][j] : -1;
    for (int i = 0; i < (int)fp.src.size() && src && i < src_cap; ++i) src[i] = fp.src[i];
    return SQH_OK;
}

int sqh_inspect_slots(int qbit_num, const sqh_gate* gates, int n_gates, void* slots_out, int slots_cap,
                      int* n_slots) {
    if (n_gates < 0 || (n_gates > 0 && !gates)) return SQH_ERR_ARG;
    std::vector<GateRec> g(n_gates);
    for (int i = 0; i < n_gates; ++i) g[i] = {gates[i].type, gates[i].target, gates[i].control, gates[i].param_start};
    Program p;
    std::string err;
    if (!compile_program(qbit_num, g, p, err)) return SQH_ERR_GATE;
    if (n_slots) *n_slots = (int)p.slots.size();
    if (slots_out && slots_cap > 0)
        std::memcpy(slots_out, p.slots.data(), sizeof(Slot) * std::min<size_t>(p.slots.size(), (size_t)slots_cap));
    return SQH_OK;
}

}  // extern "C"
