/*
 * sqhip.h -- C ABI of the MI355X (gfx950) gate-application / cost / gradient library.
 *
 * The library replaces the in-process calls the SQUANDER reference makes on its hot
 * path (paths relative to the reference checkout):
 *
 *   Gates_block::apply_to                 squander/src-cpp/gates/include/Gates_block.h:130
 *   Gates_block::apply_derivate_to        squander/src-cpp/gates/include/Gates_block.h:170-178
 *   Gates_block::apply_to_combined        squander/src-cpp/gates/include/Gates_block.h:182-196
 *   Optimization_Interface::optimization_problem          .../decomposition/include/Optimization_Interface.h:333
 *   Optimization_Interface::optimization_problem_batched  .../decomposition/include/Optimization_Interface.h:385
 *   Optimization_Interface::optimization_problem_grad     .../decomposition/include/Optimization_Interface.h:422
 *   Optimization_Interface::optimization_problem_combined .../decomposition/include/Optimization_Interface.h:434
 *   Optimization_Interface::optimization_problem_combined_unitary (Optimization_Interface.cpp:1525-1561)
 *
 * and follows the reference's own accelerator precedent (the dlopen'ed DFE C ABI,
 * squander/src-cpp/common/include/common_DFE.h:43-124): plain pointers and sizes,
 * integer status codes, one context per device with its own stream.
 *
 * Conventions (the reference's):
 *   - complex matrices are row-major, interleaved {re, im}; complex128 with float64
 *     parameters, or complex64 with float32 parameters where a call says so;
 *   - qubit q is bit q of the row index; gates run in list order, gate 0 first;
 *   - raw parameters: half-angles for RX/RY/RZ/CRY and for U3's theta;
 *   - gate types and cost variants use the reference's enum values.
 *
 * Every call returns SQH_OK (0) or a negative status; sqh_last_error() gives the
 * message (the reference throws std::string with the same information).
 * Calls on one context are serialised by a mutex; distinct contexts are independent.
 * "_dev" entry points take device pointers and are asynchronous on the context's
 * stream (sqh_get_stream); all others take host pointers and are synchronous.
 */
#ifndef SQHIP_H
#define SQHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SQH_OK 0
#define SQH_ERR_ARG (-1)      /* bad argument: sizes, dtypes, parameter count */
#define SQH_ERR_GATE (-2)     /* unsupported gate or qubit index */
#define SQH_ERR_STATE (-3)    /* circuit or target not set */
#define SQH_ERR_DEVICE (-4)   /* HIP runtime failure */
#define SQH_ERR_NOMEM (-5)    /* device allocation failed */

/* Reference gate_type values (squander/src-cpp/gates/include/Gate.h:39-51). */
#define SQH_GATE_CZ 4
#define SQH_GATE_CNOT 5
#define SQH_GATE_U3 7
#define SQH_GATE_RY 8
#define SQH_GATE_RX 9
#define SQH_GATE_RZ 10
#define SQH_GATE_CRY 14

/* Reference cost_function_type values (Optimization_Interface.h:43-45);
 * SQH_COST_ABS_TRACE = 1 - |Tr|/m has no reference counterpart. */
#define SQH_COST_FROBENIUS 0
#define SQH_COST_FROBENIUS_CORRECTION1 1   /* FROBENIUS_NORM_CORRECTION1 */
#define SQH_COST_FROBENIUS_CORRECTION2 2   /* FROBENIUS_NORM_CORRECTION2 */
#define SQH_COST_HILBERT_SCHMIDT 3
#define SQH_COST_HILBERT_SCHMIDT_CORRECTION1 4
#define SQH_COST_HILBERT_SCHMIDT_CORRECTION2 5
#define SQH_COST_INFIDELITY 9
#define SQH_COST_ABS_TRACE 100

/* Precision of a cost / gradient evaluation. */
#define SQH_PREC_F64 0   /* complex128 arithmetic */
#define SQH_PREC_F32 1   /* complex64 arithmetic, FP32 accumulation, FP64 trace sum */

typedef struct sqh_ctx sqh_ctx;

/* One gate: the reference's gate object reduced to what the path reads
 * (Gate.h: type, target_qbit, control_qbit, parameter_start_idx). control = -1 when
 * the gate has none. */
typedef struct sqh_gate {
    int32_t type;
    int32_t target;
    int32_t control;
    int32_t param_start;
} sqh_gate;

/* ---- context -------------------------------------------------------------- */
int sqh_create(int device, sqh_ctx** out);
void sqh_destroy(sqh_ctx* ctx);
const char* sqh_last_error(const sqh_ctx* ctx);
void* sqh_get_stream(sqh_ctx* ctx);            /* the context's hipStream_t */
int sqh_synchronize(sqh_ctx* ctx);
const char* sqh_version(void);

/* ---- circuit, target and cost configuration ------------------------------- */
/* Gates_block gate list (Gates_block.cpp:2500-2525 assigns param_start in order). */
int sqh_set_circuit(sqh_ctx* ctx, int qbit_num, const sqh_gate* gates, int n_gates);
int sqh_get_parameter_num(const sqh_ctx* ctx);
/* Decomposition target Umtx (rows = 2^qbit_num, cols <= rows), complex128 host
 * buffer with a row stride in elements (Decomposition_Base: Umtx). */
int sqh_set_target(sqh_ctx* ctx, const double* U, int rows, int cols, int stride);
int sqh_set_trace_offset(sqh_ctx* ctx, int trace_offset);       /* set_Trace_Offset */
int sqh_set_cost_function(sqh_ctx* ctx, int variant);           /* set_Cost_Function_Variant */
/* Weights of the correction variants (Optimization_Interface.h:83-89 members
 * prev_cost_fnv_val, correction1_scale, correction2_scale; defaults 1, 1/1.7, 1/2):
 * the single / double bit-flip traces enter with sqrt(prev_cost) * scale1 / scale2
 * (Optimization_Interface.cpp:676-716).  prev_cost must be >= 0. */
int sqh_set_correction_weights(sqh_ctx* ctx, double prev_cost, double scale1, double scale2);

/* ---- Gates_block::apply_to -------------------------------------------------- */
/* In place on a host matrix. is_f32: complex64 matrix + float32 parameters. */
int sqh_apply_to(sqh_ctx* ctx, const void* params, void* inout, int rows, int cols, int stride, int is_f32);
/* Device variant: B matrices at inout + b*bstride elements, parameters shared
 * (pbstride = 0) or per matrix (pbstride = P).  Up to 10 qubits it runs the register-tile
 * kernels, or -- where sqh_set_fused_cost selects it (by default from 7 qubits) -- the
 * tile-resident pass of the circuit's fused 4-qubit blocks on the matrix cores, which
 * writes each column back to where it read it; above 10 qubits the fused streaming
 * path. */
int sqh_apply_to_dev(sqh_ctx* ctx, const double* d_params, int pbstride, void* d_inout, int rows, int cols,
                     int stride, long long bstride, int B, int is_f32);
/* The same result computed by one streaming HBM pass per gate (the unfused path). */
int sqh_apply_to_stream_dev(sqh_ctx* ctx, const double* d_params, int pbstride, void* d_inout, int rows, int cols,
                            int stride, long long bstride, int B, int is_f32);

/* The same result with runs of gates fused into dense 16x16 blocks on <= 4 qubits
 * (one HBM pass per block on the matrix cores: FP64 MFMA for complex128, FP32 MFMA
 * with FP32 accumulation for complex64); needs >= 4 qubits. */
int sqh_apply_to_fused_dev(sqh_ctx* ctx, const double* d_params, int pbstride, void* d_inout, int rows, int cols,
                           int stride, long long bstride, int B, int is_f32);
/* apply_large_kernel_to_input (squander/src-cpp/gates/kernels/apply_large_kernel_to_input.cpp:82-102):
 * a dense 2^k x 2^k block (row-major complex128, or complex64 with is_f32) on the k
 * involved qubits (sorted ascending; local index bit i <-> qubits[i]) of B matrices,
 * 1 <= k <= 5, device pointers; qubits is a host array. */
int sqh_apply_large_kernel_dev(sqh_ctx* ctx, const void* d_block, const int* qubits, int k, void* d_inout, int rows,
                               int cols, int stride, long long bstride, int B, int is_f32);

/* Launch shape of the 4-qubit matrix-core block kernel (sqh_apply_to_fused_dev,
 * sqh_apply_large_kernel_dev with k = 4, the streaming cost path): sixteen-column tiles
 * per wavefront (1, 2, 4) and threads per workgroup (64, 256).  With tuning on (the
 * default) the first call of each (precision, rows, cols, B) shape times every
 * candidate on a scratch batch of that shape and keeps the fastest for the context;
 * sqh_set_block_config forces one shape (0, 0 returns to automatic);
 * sqh_get_block_config reports the shape used for a given call shape (tuning it if
 * needed) and the tuning time spent so far.  The shape never changes results.  No
 * counterpart in the reference (its AVX kernels have fixed blocking). */
int sqh_set_autotune(sqh_ctx* ctx, int enable);
int sqh_set_block_config(sqh_ctx* ctx, int tiles_per_wave, int threads);
int sqh_get_block_config(sqh_ctx* ctx, int is_f32, int rows, int cols, int B, int* tiles_per_wave, int* threads,
                         double* tune_ms);

/* ---- Gates_block::apply_derivate_to / apply_to_combined ---------------------- */
/* out receives P dense rows x cols complex128 matrices: out[k] = (dC/dp_k) . in.  One
 * store pass over P elements sharing the input, element k with parameter k's gate
 * replaced by its derivative: through the fused 4-qubit blocks where sqh_set_fused_cost
 * selects them (as for the cost pass), else the register-tile store kernel. */
int sqh_apply_derivate_to(sqh_ctx* ctx, const double* params, const double* in, int rows, int cols, int stride,
                          double* out);
int sqh_apply_derivate_to_dev(sqh_ctx* ctx, const double* d_params, const double* d_in, int rows, int cols,
                              int stride, double* d_out);

/* ---- Optimization_Interface: cost / gradient on the target ------------------ */
/* costs[b] for B parameter vectors (params is B x P row-major). */
int sqh_optimization_problem_batch(sqh_ctx* ctx, const double* params, int B, int precision, double* costs);
int sqh_optimization_problem_batch_dev(sqh_ctx* ctx, const double* d_params, int B, int precision,
                                       double* d_costs);
/* f0[b] and grad[b*P + k] (adjoint evaluation of d cost / d p_k). */
int sqh_optimization_problem_combined_batch(sqh_ctx* ctx, const double* params, int B, int precision,
                                            double* f0, double* grad);
int sqh_optimization_problem_combined_batch_dev(sqh_ctx* ctx, const double* d_params, int B, int precision,
                                                double* d_f0, double* d_grad);
/* Cost and gradient of B parameter vectors, element b against its own square target
 * targets[target_index[b]] (device arrays: n_targets x 2^n x 2^n complex128, B ints,
 * B x P parameters; trace offset 0).  The batched form of optimization_problem_combined
 * over many problems of one circuit (the caller of a learned start's training loop). */
int sqh_optimization_problem_combined_targets_dev(sqh_ctx* ctx, const double* d_targets, int n_targets,
                                                  const int* d_target_index, const double* d_params, int B,
                                                  int precision, double* d_f0, double* d_grad);
/* The same costs and gradients as sqh_optimization_problem_combined_batch_dev
 * (d_grad may be NULL: costs only) without allocating anything: no hipMalloc, no
 * hipFree, no allocating library routine and no run-time kernel compilation (a
 * circuit-specialised kernel is used only if an earlier call compiled it).  No
 * counterpart in the reference: it is for callers that cannot allocate while they
 * run (code under stream capture, serving loops, harnesses that forbid hipMalloc).
 * Its scratch memory (kernel tables, trace channels, gradient correlations) is one
 * 64 MiB device array per GPU, held until the process ends; a batch that needs more
 * is rejected with SQH_ERR_ARG (at 8 qubits with a gradient that is B > ~170, costs
 * only B > ~6800).  Up to 10 qubits.  All calls on one GPU share that array: calls
 * that overlap on a GPU race and must be serialised by the caller. */
int sqh_optimization_problem_combined_batch_noalloc_dev(sqh_ctx* ctx, const double* d_params, int B, int precision,
                                                        double* d_f0, double* d_grad);
/* optimization_problem_combined_unitary: C.Umtx (rows x cols dense) and the P
 * derivative matrices (P x rows x cols). */
int sqh_optimization_problem_combined_unitary(sqh_ctx* ctx, const double* params, double* out_unitary,
                                              double* out_derivs);

/* ---- batched decomposition (the optimiser that calls the path) -------------- */
/* Options of sqh_decompose_batch; the defaults of the reference's custom
 * decomposition at <= 5 qubits are max_inner_iterations = 10000,
 * max_outer_iterations = 4, convergence_threshold = 1e-5
 * (Optimization_Interface.cpp:1667-1670, N_Qubit_Decomposition_custom.cpp:59-76,
 * Decomposition_Base.cpp:131-224). */
typedef struct sqh_decompose_options {
    long long max_inner_iterations;   /* cost+gradient evaluations per BFGS run */
    int max_outer_iterations;         /* BFGS restarts from the best point */
    double optimization_tolerance;    /* stop once the FP64 cost is below this */
    double convergence_threshold;     /* relative spread of the last 20 minima */
    int mixed_precision;              /* 1: a first BFGS stage in complex64 / FP32 */
    double switch_cost;               /* ... handing over to FP64 below this FP32 cost */
    int rounds_per_graph;             /* lockstep rounds per captured hipGraph (0: 16) */
    long long max_rounds;             /* safety cap on rounds (0: derived from the limits) */
    /* BFGS iteration loops per outer iteration (BFGS.cpp:65-83,121-133; the reference
     * runs 3 at 2 qubits, else 1, or config max_iteration_loops[_bfgs]); 0 means 1.
     * After a loop improves the run's minimum the next loop starts from its end point
     * plus the next restart-jitter vector. */
    int iteration_loops;
    /* Restart-jitter vectors per run (J), and the host array holding them: the
     * reference's U[0,2pi)/100 draws in the order its generator yields them, J x P
     * values per run, run after run (sqh_decompose_problems: problem i's J x P_i block
     * after problem i-1's).  A run uses one vector per improvement; one that improves
     * more than J times restarts its later loops without jitter.  NULL or J = 0: no
     * jitter.  Read only when iteration_loops > 1. */
    int restart_jitter_count;
    const double* restart_jitter;
    /* Execution engine: SQH_ENGINE_AUTO picks the persistent solver when every run has
     * at most 6 qubits, the cost variant has no correction channels and the run's state
     * fits a workgroup's LDS, else lockstep rounds;
     * SQH_ENGINE_ROUNDS forces the lockstep rounds; SQH_ENGINE_PERSISTENT requires the
     * persistent solver (an error where it does not apply).  Both run the same
     * optimiser; results agree to rounding. */
    int engine;
} sqh_decompose_options;

#define SQH_ENGINE_AUTO 0
#define SQH_ENGINE_ROUNDS 1        /* every run advances one evaluation per lockstep round */
#define SQH_ENGINE_PERSISTENT 2    /* one wavefront per run, the whole optimisation in one launch */

/* n_runs independent decompositions of the context's circuit, run r on target
 * run_target[r] (targets: n_targets x rows x cols complex128, each the Umtx of
 * its problem) from start point x0[r] (n_runs x P).  Each run is the reference's
 * driver for a custom circuit: BFGS_Powell with the Wolfe line search
 * (BFGS_Powell.cpp:70-386, grad_descend.cpp:127-272) restarted from the best point
 * (Decomposition_Base.cpp:375-709) with the engine's iteration loops
 * (BFGS.cpp:42-152).  All runs
 * advance in lockstep on the GPU, one batched cost+gradient launch per round.
 * Outputs: best parameters, best FP64 cost, status of the last BFGS run (-1 if the
 * round cap stopped it), evaluations per run, and the number of rounds. */
int sqh_decompose_batch(sqh_ctx* ctx, const double* targets, int n_targets, int rows, int cols,
                        const int* run_target, int n_runs, const double* x0, const sqh_decompose_options* opts,
                        double* x_out, double* f_out, int* status_out, long long* evals_out, long long* rounds_out);

/* One decomposition problem of sqh_decompose_problems: its own circuit (gate list as
 * in sqh_set_circuit), its square 2^qbit_num target (the problem's Umtx, complex128,
 * row-major, stride 2^qbit_num) and its start point (the circuit's parameter count). */
typedef struct sqh_problem {
    int qbit_num;
    int n_gates;
    const sqh_gate* gates;
    const double* target;
    const double* x0;
} sqh_problem;

/* n_problems independent decompositions with different circuits and qubit counts
 * (1..10 qubits each) as one lockstep batch: the driver of sqh_decompose_batch per
 * problem (Decomposition_Base.cpp:375-709 for each), with the cost function the
 * context is set to (trace offset 0), and every problem's cost+gradient evaluated in
 * one launch per qubit count and precision per round.  x_out receives the problems'
 * parameter vectors one after another; the other outputs follow sqh_decompose_batch. */
int sqh_decompose_problems(sqh_ctx* ctx, const sqh_problem* problems, int n_problems,
                           const sqh_decompose_options* opts, double* x_out, double* f_out, int* status_out,
                           long long* evals_out, long long* rounds_out);

/* Options of sqh_agents; the reference defaults in brackets (AGENTS.cpp:98-232,
 * Optimization_Interface.cpp:62-64). */
typedef struct sqh_agents_options {
    int agent_num;                   /* agent_num [64] */
    long long max_iterations;        /* max_inner_iterations(_agent) */
    double optimization_tolerance;   /* optimization_tolerance(_agent) */
    int agent_lifetime;              /* agent_lifetime [1000] */
    double exploration_rate;         /* agent_exploration_rate [0.2] */
    double agent_randomization_rate; /* agent_randomization_rate [0.2] */
    double radius;                   /* Randomized_Radius [1.0] */
    double randomization_rate;       /* randomization_rate [0.3] */
    int convergence_length;          /* convergence_length [20] */
} sqh_agents_options;

/* Optimization_Interface::solve_layer_optimization_problem_AGENTS (AGENTS.cpp:42-870)
 * on the context's circuit and target, with the line search of its cost: FROBENIUS
 * (three-point sinusoid, AGENTS.cpp:395-443) or HILBERT_SCHMIDT_TEST (five-point fit
 * minimised by BFGS_Powell, AGENTS.cpp:544-675); other costs are rejected (the
 * reference has no line search for them).  The optimiser's std::mt19937 is seeded
 * with seed.  x: in the solution
 * guess, out optimized_parameters (P doubles, host).  *f_out: current_minimum;
 * *iters_out: iterations run; *evals_out: cost evaluations (all batched on the GPU). */
int sqh_agents(sqh_ctx* ctx, double* x, const sqh_agents_options* opts, unsigned int seed, double* f_out,
               long long* iters_out, long long* evals_out);

/* n_runs independent AGENTS runs on the context's circuit (FROBENIUS or
 * HILBERT_SCHMIDT_TEST cost, its trace offset, up to 10 qubits), run k against the k-th complex128 rows x cols target of
 * targets (host, contiguous, rows = 2^qbit_num), from the guess x + k*P with its
 * generator seeded with seeds[k].  Each run's results (x + k*P, f_out[k], iters_out[k],
 * evals_out[k]) equal those of sqh_agents on that target and seed; the runs advance in
 * lockstep and every stage's evaluations of all live runs are one launch.  The
 * reference runs one optimisation per target (AGENTS.cpp:42-870 per call); this entry
 * point batches those calls. */
int sqh_agents_batch(sqh_ctx* ctx, const double* targets, int n_runs, int rows, int cols, double* x,
                     const sqh_agents_options* opts, const unsigned int* seeds, double* f_out, long long* iters_out,
                     long long* evals_out);

/* Run-time specialised tile kernels.  With enable = 1 the context compiles its
 * circuit, on first use, into straight-line forward and gradient kernels (hiprtc,
 * cached per circuit) and uses them for apply_to, apply_derivate_to, the batched
 * costs and gradients and sqh_decompose_batch; results equal the generic kernels' to
 * rounding.  No
 * counterpart in the reference (which has one generic loop).  sqh_get_jit_stats
 * reports the compile time spent so far. */
int sqh_set_jit(sqh_ctx* ctx, int enable);
int sqh_get_jit_stats(sqh_ctx* ctx, double* compile_ms, int* compiles);
/* Cost-only batched evaluations (sqh_optimization_problem_batch[_dev], the AGENTS
 * populations, single and batched) through the circuit's fused dense 4-qubit blocks on
 * the matrix cores with every column resident on the chip (one launch: blocks applied
 * as MFMA tiles, LDS exchanges between blocks, trace entries at the end).  Applies for
 * 4 to 10 qubits and costs without correction channels; elsewhere the tile kernels
 * run.  Results agree with the tile kernels to rounding (the blocks reassociate the
 * products).  mode: SQH_FUSED_COST_OFF, _ON, or _AUTO (the default: on from 7 qubits,
 * or from 5 when the run-time specialised kernels are off -- where it measured faster
 * than the tile kernels -- when the blocks do no more arithmetic than the gates).  The
 * same mode selects the fused blocks' adjoint for cost + gradient batches
 * (sqh_optimization_problem_combined_batch[_dev]; _AUTO: from 6 qubits, with the same
 * arithmetic condition): forward and backward block products on the matrix cores, each
 * block's correlations then resolved to its gates.  sqh_decompose_batch's lockstep
 * rounds use the fused adjoint with _ON only.  apply_derivate_to follows the cost pass's
 * rule (one fused store pass over the P derivative elements).  No counterpart in the reference: it
 * replaces the per-gate loop of Gates_block::apply_to (Gates_block.cpp:606-710) inside
 * optimization_problem_batched (Optimization_Interface.cpp:940-1033) and the derivative
 * products of optimization_problem_combined (Optimization_Interface.cpp:1401-1450). */
#define SQH_FUSED_COST_OFF 0
#define SQH_FUSED_COST_ON 1
#define SQH_FUSED_COST_AUTO 2
int sqh_set_fused_cost(sqh_ctx* ctx, int mode);
/* Launch shape of the fused-block cost kernel: 16-column tiles per wavefront (2 or 4),
 * LDS exchange groups per block (1, or 2 with 4 tiles and prefetch: one group's
 * exchange overlaps the other's MFMAs), wavefronts per workgroup (1, 2 or 4), and
 * prefetch (1: the next block's matrix is loaded while a block runs, at the cost of
 * registers).  All zero (the default):
 * the first evaluation of each (precision, qubits, columns, batch) shape times the
 * valid shapes on its own inputs and keeps the fastest.  Every shape gives
 * bit-identical results.  sqh_get_fused_cost_config reports the shape in use for a
 * precision and batch size of the context's circuit and target. */
int sqh_set_fused_cost_config(sqh_ctx* ctx, int tiles_per_wave, int exchange_groups, int waves, int prefetch);
int sqh_get_fused_cost_config(sqh_ctx* ctx, int is_f32, int B, int* tiles_per_wave, int* exchange_groups,
                              int* waves, int* prefetch);
/* How the fused cost pass forms the trace entries.  on = 1 (the default): the
 * transposed chain -- entry c is row c + trace_offset of the circuit matrix V times
 * column c of the target, and that row is built from a unit vector through the blocks
 * transposed, last block first, touching only the rows inside the union of the blocks'
 * qubits seen so far (the 8-qubit 40-layer ansatz: 4 block products of work instead of
 * 5).  on = 0: the target's columns run forward through the blocks (the store pass
 * always does).  The two agree to rounding; the chain takes one exchange group
 * (sqh_set_fused_cost_config's groups = 1), and circuits of one block always run
 * forward.  The same switch selects where the fused gradient's adjoint vectors come
 * from: the chain, followed by one sweep up from the target (8-qubit ansatz: 19 block
 * products instead of 23; needs blocks whose qubits cover the register), or the
 * backward sweep.  No counterpart in the reference (the same trace and gradient as
 * Optimization_Interface.cpp:940-1033 and 1401-1450 compute). */
int sqh_set_fused_cost_chain(sqh_ctx* ctx, int on);
/* Launch shape of the fused-block cost + gradient kernel: 16-column tiles per
 * wavefront (2 or 4; 10 qubits need 4) and wavefronts per workgroup (1, 2, 4 or 8,
 * dividing the element's column groups).  Both zero (the default): a fixed rule per
 * qubit count.  A shape that does not fit the circuit, target or the device's LDS falls
 * back to the rule.  The shape sets how the correlation sums are grouped, so different
 * shapes agree to rounding (each shape is deterministic). */
int sqh_set_fused_grad_config(sqh_ctx* ctx, int tiles_per_wave, int waves);
/* Launch options of the register-tile kernels (the batched costs and gradients when
 * the fused-block path does not run).  waves_per_group: for the forward pass with one
 * batch element per wavefront (above 4 qubits), 4, 2 or 1 wavefronts of an element per
 * workgroup sharing its kernel table in LDS (fewer when a group does not fit), or -1
 * for the packed layout (several elements per wavefront).  stage_program: 1 stages the
 * gradient pass's program (ops, layouts, kernel table) in LDS where it fits, 0 reads
 * it through the scalar cache.  waves_per_group = 0 and stage_program = -1 (the
 * default): the first call of each (pass, precision, qubits, columns, batch) shape times
 * the distinct options on its own inputs and keeps the fastest.  Every option gives
 * bit-identical results.  sqh_get_tile_config reports, for kind 1 (forward) or 2
 * (gradient), the options in effect for the context's circuit, target and a batch of
 * B: wavefronts per workgroup (-1 packed) and whether the program is staged. */
int sqh_set_tile_config(sqh_ctx* ctx, int waves_per_group, int stage_program);
int sqh_get_tile_config(sqh_ctx* ctx, int kind, int is_f32, int B, int* waves_per_group, int* stage_program);
/* Introspection (no GPU needed): generate and compile, for gfx950, the run-time
 * kernel of a gate list (kind 0 forward store, 1 forward diagonal, 2 gradient, 3 the
 * persistent solver of sqh_decompose_batch, up to 4 qubits; uniform = one batch element
 * per wavefront).  Returns SQH_OK when it compiles;
 * log receives the compiler log. */
int sqh_jit_compile_check(int qbit_num, const sqh_gate* gates, int n_gates, int kind, int is_f32, int uniform,
                          char* log, int log_cap, long long* code_bytes, long long* source_bytes);
/* Introspection (no GPU needed): the generated source of the run-time kernel that
 * sqh_jit_compile_check compiles, into buf (at most cap - 1 bytes, NUL-terminated);
 * source_bytes receives its full length. */
int sqh_jit_source(int qbit_num, const sqh_gate* gates, int n_gates, int kind, int is_f32, int uniform, char* buf,
                   long long cap, long long* source_bytes);

/* ---- warm-start network ----------------------------------------------------- */
/* A multilayer perceptron in -> hidden (GELU) x n_hidden -> out with BF16 weights
 * and activations, FP32 accumulation on the matrix cores and FP32 master weights,
 * trained with Adam on mean-squared error.  It maps a target unitary to start angles
 * for the batched optimiser.  No counterpart in the reference (its nn module only
 * generates data: squander/src-cpp/nn/include/NN.h:58-146).  Parameters travel as a
 * flat float32 array: per layer W (out x in, row-major) then b (out). */
typedef struct sqh_mlp sqh_mlp;
int sqh_mlp_create(sqh_ctx* ctx, int in_dim, int hidden, int n_hidden, int out_dim, unsigned long long seed,
                   sqh_mlp** out);
void sqh_mlp_destroy(sqh_mlp* mlp);
long long sqh_mlp_param_count(const sqh_mlp* mlp);
int sqh_mlp_get_params(sqh_mlp* mlp, float* params);
int sqh_mlp_set_params(sqh_mlp* mlp, const float* params);
/* y (B x out, fp32) = network(x (B x in, fp32)); device pointers, asynchronous. */
int sqh_mlp_forward_dev(sqh_mlp* mlp, const float* d_x, int B, float* d_y);
/* One Adam step on the batch (x, targets t); *loss (if given) receives the batch loss
 * before the step (synchronises). */
int sqh_mlp_train_step_dev(sqh_mlp* mlp, const float* d_x, const float* d_t, int B, float lr, double* loss);
/* Loss and gradients (flat, like the parameters) of a host batch, without updating. */
/* One Adam step from a caller-supplied gradient of the loss with respect to the
 * network's output (d_dy: B x out_dim FP32, device), e.g. the gradient of the
 * decomposition cost at the predicted angles chained through the output coding. */
int sqh_mlp_backward_step_dev(sqh_mlp* mlp, const float* d_x, const float* d_dy, int B, float lr);
int sqh_mlp_gradients(sqh_mlp* mlp, const float* x, const float* t, int B, float* grads, double* loss);

/* ---- kernel timers --------------------------------------------------------- */
/* When enabled, every call brackets its dominant kernel (the register-tile kernel,
 * or the derivative kernel) with HIP events on the context's stream.
 * sqh_get_kernel_timing synchronises, returns the summed kernel time and the number
 * of timed launches since the last query, and resets the timers.  Counterpart of
 * the reference's circuit_simulation_time accumulator (Optimization_Interface.h:114-117). */
int sqh_set_kernel_timing(sqh_ctx* ctx, int enable);
int sqh_get_kernel_timing(sqh_ctx* ctx, double* total_ms, int* launches);

/* ---- introspection (host only, no device needed) ---------------------------- */
/* Compiles a gate list into the register-tile program the kernels execute and
 * copies out its op records (8 bytes each) and layout records (96 bytes each);
 * n_ops / n_layouts receive the counts.  Buffers may be NULL to query the counts.
 * No counterpart in the reference: it exists for tooling and the CPU test suite. */
int sqh_inspect_program(int qbit_num, const sqh_gate* gates, int n_gates, void* ops_out, int ops_cap,
                        void* layouts_out, int layouts_cap, int* n_ops, int* n_layouts);
/* The same program's parametric-gate records (16 bytes each: gate type, first
 * parameter, parameter count, op index), in the order the ops' slot indices use
 * (the compiler may run gates on disjoint qubits out of list order).  No counterpart
 * in the reference (tooling and the CPU test suite). */
int sqh_inspect_slots(int qbit_num, const sqh_gate* gates, int n_gates, void* slots_out, int slots_cap,
                      int* n_slots);
/* The dense-block fusion plan of the streaming path (sqh_apply_to_fused_dev): block b
 * runs gate-list entries src[block_first[b] .. block_first[b+1]) in that order on the
 * sorted qubits block_qubits[4b .. 4b+3].  Buffers may be NULL to query *n_blocks; sizes:
 * block_first n_blocks + 1, block_qubits 4 n_blocks, src n_gates.  No counterpart in the
 * reference (tooling and the CPU test suite). */
int sqh_inspect_fusion(int qbit_num, const sqh_gate* gates, int n_gates, int* block_first, int* block_qubits,
                       int blocks_cap, int* src, int src_cap, int* n_blocks);

#ifdef __cplusplus
}
#endif
#endif /* SQHIP_H */
