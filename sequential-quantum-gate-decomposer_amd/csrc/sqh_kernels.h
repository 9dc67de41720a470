// sqh_kernels.h -- host-side launchers for the kernels in sqh_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "sqh_internal.h"

namespace sqh {

// Launch shape of the 4-qubit matrix-core block kernel: 16-column tiles per wavefront
// (1, 2 or 4) and threads per workgroup (64 or 256).
struct BlockCfg {
    int tiles = 0, threads = 0;
};
bool block_cfg_valid(BlockCfg cfg);

// Arguments of one register-tile launch.  Matrices are interleaved complex
// (complex128, or complex64 when f32), row-major with a row stride in elements.
struct TileLaunch {
    int N = 0;                 // qubits (rows = 2^N)
    bool f32 = false;          // complex64 tiles with FP32 arithmetic
    int mode = 0;              // forward: 0 = store matrix, 1 = diagonal only
    const void* in = nullptr;  // input slab base
    long long in_bstride = 0;  // elements between batch inputs (0: shared input)
    int in_stride = 0;
    void* out = nullptr;       // forward store / derivative output
    long long out_bstride = 0; // elements between batch outputs (derivative: between matrices)
    int out_stride = 0;
    int m = 0;                 // columns
    int m_pad = 0;             // padded columns per batch element
    int B = 1;                 // batch elements
    const Op* ops = nullptr;
    int nops = 0;
    const Layout* lays = nullptr;
    const void* kern = nullptr;  // [B][ns][4] kernels
    long long kern_bstride = 0;  // complex entries between batch kernel tables (0: shared)
    int trace_off = 0;
    int ns = 0;                  // parametric gate count
    double* diag = nullptr;      // [nch + 1][B][m] complex trace-channel entries
    TraceSpec ts;                // trace channels and gradient seed (default: diagonal only)
    double* mpart = nullptr;     // [B][wpb][ns][8] gradient correlations
    const Slot* slots = nullptr;
    BatchIndex bi;             // optional element -> run -> target indirection
    const int* in_index = nullptr;  // forward: element b reads input in_index[b] (null: input b)
    const int* zslot = nullptr;  // forward store: per element, the controlled slot applied in
                                 // derivative mode (control-0 block zeroed), or -1
    int stage_ops = 0;         // > 0: gradient kernel stages the program in LDS (largest op count)
    int stage_lays = 0;        //      ... and up to this many layouts
    // Launch options that leave every column's arithmetic unchanged:
    int wpg = 0;     // forward, one element per wavefront: 4, 2 or 1 wavefronts per workgroup
                     // (falling back to fewer when the group does not fit), -1 the packed
                     // layout, 0 the largest group that fits
    int stage = -1;  // gradient: 0 never stage the program, otherwise stage when stage_ops > 0
};

// LDS a workgroup may allocate on the current device (160 KiB on MI355X).
size_t device_max_lds();

hipError_t launch_tile_forward(const TileLaunch& t, hipStream_t st);
int tile_forward_group(const TileLaunch& t, int wpg);
// Whether the gradient pass of t can stage its program in LDS (one element per
// wavefront; kernel table, layouts and ops within the staging budget).
bool tile_grad_can_stage(const TileLaunch& t);
// Launch options of the register-tile kernels the tuner picks among (TileLaunch::wpg,
// TileLaunch::stage).
struct TileCfg {
    int wpg = 0;
    int stage = -1;
};
hipError_t launch_tile_grad(const TileLaunch& t, hipStream_t st);
int tile_columns_per_wave(int N);

// Per-parameter kernel tables of apply_derivate_to: table k (ns entries) is the
// circuit's table base with the entry of parameter k's gate replaced by the gate's
// derivative kernel w.r.t. that parameter; zslot[k] is that gate's slot when it is
// controlled (its control-0 block must be zeroed), else -1.
hipError_t launch_deriv_tables(const void* base, const Slot* slots, int ns, const double* params, int P,
                               void* tables, int* zslot, hipStream_t st);
hipError_t launch_build_kernels(bool f32, const double* params, int P, int B, const Slot* slots, int ns, void* kern,
                                hipStream_t st, BatchIndex bi = BatchIndex());
hipError_t launch_finish(const double* diag, int m, int B, int variant, double* costs, const double* mpart, int wpb,
                         int ns, const Slot* slots, const double* params, int P, double* grad, hipStream_t st,
                         BatchIndex bi = BatchIndex(), const TraceSpec& ts = TraceSpec(),
                         double* seed_out = nullptr);
hipError_t launch_stream_gate(bool f32, void* mats, long long bstride, int stride, int rows, int m, int B, int kind,
                              int target, int control, const void* kern, long long kern_bstride, int slot,
                              hipStream_t st);
hipError_t launch_extract_diag(bool f32, const void* mats, long long bstride, int stride, int m, int B,
                               int trace_off, double* diag, int nq, const TraceSpec& ts, hipStream_t st);
hipError_t launch_broadcast_copy(bool f32, const void* src, int rows, int m, int sstride, void* dst,
                                 long long dbstride, int B, hipStream_t st);
hipError_t launch_convert(bool to_f32, const void* src, void* dst, long long nreal, hipStream_t st);

// Tile-resident pass of the circuit as fused 4-qubit blocks on the matrix cores
// (sqh_k_fused_cost.hip), 4 <= N <= kFusedCostMaxQubits: the trace entries of every
// column (cost) or, with out, the transformed matrices (apply_to).  maps: per block
// kFusedMapStride ints, the row offsets of the 16 local indices, then the row bases of
// the 2^N / 16 outer indices.  F: block matrices, element b's [nb][16][16] at
// F + b * F_bstride (0: shared).  diag receives the B x m trace entries (complex, as
// double pairs).
constexpr int kFusedMapStride = 16 + 64;
constexpr int kFusedCostMaxQubits = 10;
struct FusedCostArgs {
    const void* U = nullptr;  // input: element b's at U + (in_index ? in_index[b] : b) * in_bstride
    const int* in_index = nullptr;
    long long in_bstride = 0;
    int in_stride = 0;  // row stride of the input (elements)
    int m = 0, trace_off = 0, B = 0, nb = 0;
    const void* F = nullptr;
    long long F_bstride = 0;
    const int* maps = nullptr;
    double* diag = nullptr;
    void* out = nullptr;  // non-null: store element b's result at out + b * out_bstride (row stride out_stride)
    long long out_bstride = 0;
    int out_stride = 0;
    int groups = 0;  // column groups per element (set by the launcher)
};
// A gate of a fused block: OP_U / OP_CU / OP_X / OP_Z, local target, local control (-1
// none), parametric slot (kernel table index) or -1 (the host's FusedRec layout).
struct FusedGate {
    int8_t kind, t, c, pad;
    int32_t slot;
};

// Cost and gradient through the fused blocks (sqh_k_fused_grad.hip): the forward pass
// and trace entries as the cost kernel, then the blocks backwards with the trace
// selector, writing per workgroup and block the 16x16 correlation partials
// mpart[B][wgs][nb][256] (complex, as double pairs).  T: 16-column tiles per wavefront.
struct FusedGradArgs {
    const void* U = nullptr;  // target, rows = 2^N, m columns, row stride m
    int m = 0, trace_off = 0, B = 0, nb = 0;
    const void* F = nullptr;  // [B][nb][16][16] block matrices
    const int* maps = nullptr;
    double* diag = nullptr;
    double* mpart = nullptr;
    int wgs = 0;     // workgroups per element (the partials' layout)
    int groups = 0;  // column groups per element (set by the launcher)
    // optional optimiser indirection: live elements *count (of B), element e's target at
    // U + run_target[elem_run[e]] * tgt_stride
    const int* count = nullptr;
    const int* elem_run = nullptr;
    const int* run_target = nullptr;
    long long tgt_stride = 0;
    // non-null: the adjoint vectors from the transposed chain (k_fused_grad_chain), with
    // these chain tables ([nb][kTraceMapStride]); needs nb >= 2 and the blocks' qubits
    // covering the register
    const int* tmaps = nullptr;
};
// Launch shape: T 16-column tiles per wavefront (2 or 4; 4 at 10 qubits), waves 1, 2, 4
// or 8 per workgroup dividing the element's column groups; wgs = workgroups per element
// (the partials' count).
bool fused_grad_shape_valid(bool f32, int N, int m, int nb, int T, int waves);
int fused_grad_wgs(int N, int m, int T, int waves);
size_t fused_grad_lds_bytes(bool f32, int nb, int T, int waves);
hipError_t launch_fused_grad(bool f32, int N, FusedGradArgs a, int T, int waves, hipStream_t s);
// Per (element, block): the partials summed, then the block's gates run backwards on
// 16-dimensional vectors (from the columns of F, the block matrices [B][nb][16][16]),
// giving each parametric slot's 2x2 correlations in corr[B][ns][8] (the register-tile
// adjoint's layout, one partial per element).
hipError_t launch_fused_block_corr(bool f32, const void* fg, const int* block_first, int nb, const void* kern,
                                   long long kern_bstride, const void* F, double* mpart, int wgs, int B, double* corr,
                                   int ns, hipStream_t s, const int* count = nullptr);

// Launch shape: tiles per wavefront (2 or 4), LDS exchange groups per block (1, or 2
// with 4 tiles and prefetch), waves per workgroup (1, 2 or 4), next block's matrix
// prefetched (0 or 1).  Every valid shape gives bit-identical results.
struct FusedCostCfg {
    int tiles = 0, groups = 0, waves = 0, prefetch = 0;
};
bool fused_cost_cfg_valid(int N, bool f32, int nb, FusedCostCfg cfg);
size_t fused_cost_lds_bytes(bool f32, int nb, FusedCostCfg cfg);
hipError_t launch_fused_cost(bool f32, int N, FusedCostArgs a, FusedCostCfg cfg, hipStream_t s);
// The cost pass as a transposed chain from unit vectors (k_fused_trace, nb >= 2): the
// rows of the circuit matrix the trace needs, built block by block on the rows inside
// the support of the blocks seen so far, then dotted with the target's columns.  maps:
// per block kTraceMapStride ints (row offsets of the 16 local indices, row bases of the
// support's outer groups, log2 of their count, the row bits outside the support).  Shapes
// with one exchange group; every valid shape gives bit-identical results.
constexpr int kTraceMapStride = 16 + 64 + 2;
bool fused_trace_cfg_valid(int N, bool f32, int nb, FusedCostCfg cfg);
size_t fused_trace_lds_bytes(bool f32, int nb, FusedCostCfg cfg);
hipError_t launch_fused_trace(bool f32, int N, FusedCostArgs a, FusedCostCfg cfg, hipStream_t s);

}  // namespace sqh
