// sqh_context.h -- the context behind the opaque sqh_ctx handle, and helpers shared
// by the C ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sqhip.h"
#include "sqh_kernels.h"
#include "sqh_program.h"

namespace sqh {

// Grow-only device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= n) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            p = nullptr;
            n = 0;
            if (e != hipSuccess) return e;
        }
        const size_t want = std::max<size_t>(bytes, 256);
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess) n = want;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

struct JitCache;
void jit_release(sqh_ctx* c);

}  // namespace sqh

struct sqh_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // Side streams of the batched optimiser: one per qubit-count group, forked from
    // and joined back into `stream` inside each captured round (default priority).
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> side_ev;
    std::mutex mu;
    std::string err;

    bool have_circuit = false;
    long long prog_version = 0;  // bumped by every sqh_set_circuit (keys the run-time kernels)
    sqh::Program prog;
    std::vector<sqh::GateRec> gates;
    sqh::DevBuf d_ops, d_lays, d_slots;

    bool have_target = false;
    int rows = 0, cols = 0;
    sqh::DevBuf d_U64, d_U32;
    int trace_offset = 0;
    int variant = sqh::CV_FROBENIUS;
    // correction-variant weights (Optimization_Interface.cpp:74-76 defaults)
    double prev_cost = 1.0, corr_scale1 = 1.0 / 1.7, corr_scale2 = 1.0 / 2.0;

    sqh::DevBuf d_params, d_kern, d_diag, d_mpart, d_res, d_mat, d_out;
    sqh::DevBuf d_seed;  // [B][3] complex adjoint seed weights of the HS correction variants
    sqh::DevBuf d_dtab, d_zslot;  // apply_derivate_to: per-parameter kernel tables, derivative slots
    sqh::DevBuf d_idx;  // [count][0..B) identity element list of the per-target evaluation
    int idx_B = -1;

    // Dense-block fusion of the circuit for the streaming path.
    sqh::FusionPlan fplan;
    sqh::DevBuf d_fgates, d_ffirst, d_fused;
    // Row tables of the tile-resident fused-block cost kernel ([blocks][kFusedMapStride];
    // present for 4 <= N <= 10) and its switch (sqh_set_fused_cost).
    sqh::DevBuf d_fmaps;
    // Row tables of the transposed-chain cost kernel ([blocks][kTraceMapStride]) and its
    // switch (sqh_set_fused_cost_chain; on by default): the cost pass builds the rows of
    // the circuit matrix the trace needs instead of running the target's columns forward.
    sqh::DevBuf d_tmaps;
    bool fused_chain = true;
    sqh::DevBuf d_fcorr;  // [B][ns][8] gate correlations of the fused-block gradient
    int fused_cost = SQH_FUSED_COST_AUTO;
    int fg_force_tiles = 0, fg_force_waves = 0;  // fused gradient launch shape (0: automatic)
    // its launch shape: forced (sqh_set_fused_cost_config) or picked per (precision,
    // qubits, columns, batch) by the run-time tuner
    sqh::FusedCostCfg fused_force;
    std::map<std::array<long long, 4>, sqh::FusedCostCfg> fused_tuned;

    // Run-time specialised tile kernels (sqh_set_jit).
    bool jit = false;
    sqh::JitCache* jit_cache = nullptr;

    // Launch shape of the 4-qubit matrix-core block kernel: forced (sqh_set_block_config)
    // or picked per (precision, rows, columns, batch) by the run-time tuner.
    bool autotune = true;
    sqh::BlockCfg block_force;
    std::map<std::array<long long, 4>, sqh::BlockCfg> block_tuned;
    std::map<std::array<long long, 4>, sqh::TileCfg> tile_tuned;  // (pass, precision, N, m_pad, B)
    sqh::TileCfg tile_force;                                       // forced options (tile_forced)
    bool tile_forced = false;
    double tune_ms = 0.0;

    // Kernel timers: HIP events around the dominant kernel of each call.
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;

    ~sqh_ctx() {
        sqh::jit_release(this);
        for (auto& e : events) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        for (sqh::DevBuf* b : {&d_ops, &d_lays, &d_slots, &d_U64, &d_U32, &d_params, &d_kern, &d_diag, &d_mpart,
                               &d_res, &d_mat, &d_out, &d_fgates, &d_ffirst, &d_fused, &d_fmaps, &d_tmaps, &d_fcorr, &d_idx, &d_seed, &d_dtab, &d_zslot})
            b->release();
        for (hipStream_t s : side) (void)hipStreamDestroy(s);
        for (hipEvent_t e : side_ev) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace sqh {

int fail(sqh_ctx* c, int code, const char* fmt, ...);

#define HIPTRY(ctx, expr)                                                                                  \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            return ::sqh::fail((ctx), e_ == hipErrorOutOfMemory ? SQH_ERR_NOMEM : SQH_ERR_DEVICE, "%s: %s", \
                               #expr, hipGetErrorString(e_));                                              \
    } while (0)

int m_pad_for(int N, int m);
int waves_per_batch(int N, int m_pad);
bool uses_offset(int variant);
// Trace channels of the context's cost variant for a launch over B elements of m
// columns (channel stride B * m), with the constant seed weights.
TraceSpec trace_spec(sqh_ctx* c, int B, int m);
int require_circuit(sqh_ctx* c);
int require_target(sqh_ctx* c);
int build_kernels(sqh_ctx* c, const double* d_params, int B, bool f32);
TileLaunch base_launch(sqh_ctx* c, bool f32);
int stream_apply(sqh_ctx* c, void* mats, int rows, int cols, int stride, long long bstride, int B, bool f32,
                 long long kern_bstride);
// Device memory an evaluation writes besides its outputs.  Given to eval_batch by a
// caller that cannot allocate: the evaluation then allocates nothing, compiles no
// run-time kernel (it uses one only if already compiled) and records no timers.
struct EvalScratch {
    void* kern = nullptr;     // [B][ns][4] 2x2 kernels (complex128, or complex64 in f32)
    double* diag = nullptr;   // [nch + 1][B][m] complex trace channels
    double* mpart = nullptr;  // [B][waves][ns][8] per-wavefront gradient correlations
    double* seed = nullptr;   // [B][3] complex adjoint seed weights (HS corrections)
};
// Bytes of each EvalScratch part for B elements of the context's circuit and target
// (kern, diag, mpart, seed; mpart and seed are 0 without a gradient).
void eval_scratch_bytes(sqh_ctx* c, int B, bool f32, bool grad, size_t bytes[4]);
int eval_batch(sqh_ctx* c, const double* d_params, int B, int precision, double* d_costs, double* d_grad,
               const EvalScratch* fixed = nullptr);
// Looks up, once per device (the device current), the address of the file-scope
// scratch array of the allocation-free entry point (sqh_noalloc.hip).
hipError_t noalloc_scratch_init(int device);
// FP64 costs of B parameter vectors, element b against the d_index[b]-th of the
// complex128 rows x cols targets at d_targets (contiguous, rows = 2^N); the context's
// circuit, cost variant and trace offset.  N <= 10.
int eval_targets(sqh_ctx* c, const double* d_params, int B, const void* d_targets, int rows, int cols,
                 const int* d_index, double* d_costs);
// path: 0 register tile (streaming above the tile limit), 1 one pass per gate,
// 2 fused dense blocks.
int apply_dev(sqh_ctx* c, const double* d_params, int pbstride, void* d_inout, int rows, int cols, int stride,
              long long bstride, int B, bool f32, int path);
int deriv_dev(sqh_ctx* c, const double* d_params, const void* d_in, int rows, int cols, int stride, void* d_out);
int fused_apply(sqh_ctx* c, void* mats, int rows, int cols, int stride, long long bstride, int B, bool f32,
                long long kern_bstride);
// Run-time specialised kernel for the current program: kind 0 forward store, 1
// forward diagonal, 2 gradient (sqh_jit.cpp).
// With compile = false a kernel that is not compiled yet is not compiled: the call
// returns kJitNotReady and launches nothing.
constexpr int kJitNotReady = 1;
int jit_launch(sqh_ctx* c, const TileLaunch& t, int kind, hipStream_t s, bool compile = true);
// Whether the context uses a run-time kernel of this kind for its program.
bool jit_wants(sqh_ctx* c, int kind);
// Run-time kernel kind of the persistent solver specialised to the circuit.
constexpr int kJitSolve = 3;
struct SolveArgs;
// Launches the circuit-specialised persistent solver (compiled on first use).
int jit_launch_solve(sqh_ctx* c, const SolveArgs& a, int grid, size_t shm, hipStream_t s);
// Compiles the kernel jit_launch would use for these shapes (before a graph capture).
int jit_prepare(sqh_ctx* c, const TileLaunch& t, int kind);
double jit_compile_ms(sqh_ctx* c, int* compiles);

hipError_t launch_block_apply(bool f32, bool use_mfma, void* mats, long long bstride, int stride, int rows, int m,
                              int B, const void* F, long long fbstride, const int* qubits, int k, BlockCfg cfg,
                              hipStream_t s);
// Launch shape of the 4-qubit block kernel for B matrices of rows x m: the forced one,
// the tuned one for this shape, or (tuning on, not under graph capture) tuned now by
// timing every candidate on a scratch batch of the same shape (sqh_tune.cpp).
int block_config(sqh_ctx* c, bool f32, int rows, int m, int B, BlockCfg* out);
// With params (and slots), each element's 2x2 kernels are computed from its P
// parameters while staging (kern unused): the cost pass needs no kernel table.
hipError_t launch_fuse_blocks(bool f32, const void* fg, const int* block_first, int nblocks, const void* kern,
                              long long kern_bstride, int Bk, void* out, hipStream_t s, const int* count = nullptr,
                              const int* zslot = nullptr, const double* params = nullptr, int P = 0,
                              const Slot* slots = nullptr);
size_t fused_gate_bytes();
// Fused-block cost + gradient (sqh_api.cpp): launch shape and scratch sizes, the
// applicability rule, and the enqueue (no allocation) on a stream.
struct FusedGradPlan {
    int T = 0, waves = 0, wgs = 0;
    size_t fused_bytes = 0, mpart_bytes = 0, corr_bytes = 0;
};
bool fused_grad_plan(const sqh_ctx* c, bool f32, int B, int m, FusedGradPlan& p);
bool fused_grad_applies(const sqh_ctx* c, const TraceSpec& ts, bool f32, int B, int m, FusedGradPlan* plan);
int fused_grad_run(sqh_ctx* c, bool f32, const FusedGradPlan& p, int B, int m, const void* kern, const void* U,
                   double* diag, int trace_off, void* fused, double* mpart, double* corr, const BatchIndex& bi,
                   hipStream_t st, bool timed);
// The fused-block cost launch with the context's launch shape (forced, tuned or default).
int fused_cost_launch(sqh_ctx* c, bool f32, const FusedCostArgs& a);
// A register-tile pass (kind 1: forward trace entries, 2: gradient) with the context's
// launch options: forced, tuned for this (pass, precision, N, m_pad, B), or -- with
// tune, tuning on, not under graph capture -- tuned now on the call's own inputs.
int tile_launch(sqh_ctx* c, const TileLaunch& t, int kind, bool tune);
// The launch t with options cfg applied (the gradient's staged program sized from the
// context's program).
TileLaunch tile_with(const sqh_ctx* c, const TileLaunch& t, int kind, const TileCfg& cfg);

// Brackets one kernel launch with timer events when timing is enabled.
struct TimedLaunch {
    sqh_ctx* c;
    hipEvent_t a = nullptr, b = nullptr;
    explicit TimedLaunch(sqh_ctx* ctx) : c(ctx) {
        if (c->timing && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess)
            (void)hipEventRecord(a, c->stream);
    }
    ~TimedLaunch() {
        if (a && b) {
            (void)hipEventRecord(b, c->stream);
            c->events.emplace_back(a, b);
        }
    }
};

}  // namespace sqh
