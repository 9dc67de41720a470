/*
 * sq_oracle.h -- CPU oracle for the gate-application / cost / gradient path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the SQUANDER
 * reference algorithms, used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  The product library never
 * links, loads or calls it.
 *
 * Conventions (all follow the reference):
 *   - complex128 matrices are row-major, interleaved {re, im}, with an
 *     explicit row stride in elements (QGDTypes.h:38-43, matrix_base.hpp:130-137);
 *   - qubit q is bit q of the row index (LSB = qubit 0);
 *   - gates are applied in list order, gate 0 first (Gates_block.cpp:683-708);
 *   - parameters are the reference's raw parameters: half-angles for the
 *     RX/RY/RZ/CRY angle and for the U3 theta (U3.cpp:49-51, RX.cpp:24).
 */
#ifndef SQ_ORACLE_H
#define SQ_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } or_c128;

/* Gate type ids use the reference's gate_type enum values (Gate.h:39-51). */
enum {
    OR_CZ = 4, OR_CNOT = 5, OR_U3 = 7, OR_RY = 8, OR_RX = 9, OR_RZ = 10, OR_CRY = 14
};

/* Cost-function variants use the reference's cost_function_type values
 * (Optimization_Interface.h:43-45).  OR_COST_ABS_TRACE (1 - |Tr|/m) has no
 * counterpart in the reference and gets an id outside its range. */
enum {
    OR_COST_FROBENIUS = 0, OR_COST_FROBENIUS_C1 = 1, OR_COST_FROBENIUS_C2 = 2,
    OR_COST_HS = 3, OR_COST_HS_C1 = 4, OR_COST_HS_C2 = 5,
    OR_COST_INFIDELITY = 9, OR_COST_ABS_TRACE = 100
};

typedef struct { int type, target, control, param_start; } or_gate;

int  or_param_num(int type);

/* 2x2 kernels, row-major [u00, u01, u10, u11]. */
void or_gate_kernel(int type, const double* p, or_c128 k[4]);
void or_deriv_kernel(int type, const double* p, int pidx, or_c128 k[4]);

/* Pair-row kernel on a rows x cols matrix (apply_kernel_to_input.cpp:35-115). */
void or_apply_kernel(or_c128* M, int rows, int cols, int stride, const or_c128 k[4],
                     int target, int control, int deriv);
/* CNOT row swap and CZ sign flip (apply_dedicated_gate_kernel_to_input.cpp:46-93, 221-254). */
void or_apply_x(or_c128* M, int rows, int cols, int stride, int target, int control);
void or_apply_z(or_c128* M, int rows, int cols, int stride, int target, int control);

/* One gate / the whole circuit, in place. Returns 0, or -1 on a bad gate. */
int or_apply_gate(const or_gate* g, const double* params, or_c128* M, int rows, int cols, int stride);
int or_apply_circuit(const or_gate* gates, int ng, const double* params,
                     or_c128* M, int rows, int cols, int stride);

/* apply_derivate_to: out holds P dense rows x cols matrices (row stride = cols),
 * out[k] = (dC/dp_k) . in.  Returns 0 / -1. */
int or_apply_derivate(const or_gate* gates, int ng, const double* params,
                      const or_c128* in, int rows, int cols, int stride, or_c128* out);

/* Trace helpers and the cost variants (N_Qubit_Decomposition_Cost_Function.cpp). */
or_c128 or_trace_offset(const or_c128* M, int cols, int stride, int trace_offset);
or_c128 or_trace_plain(const or_c128* M, int cols, int stride);
double or_cost(const or_c128* M, int cols, int stride, int variant, int trace_offset);
/* Bit-flip traces of the correction terms: sum over columns c and over the masks of
 * `bits` (1 or 2) set bits below 2^qbit_num of M[(c + trace_offset) ^ mask, c]. */
or_c128 or_trace_flips(const or_c128* M, int qbit_num, int cols, int stride, int trace_offset, int bits);
/* Every variant, the correction ones with w1 = sqrt(prev_cost) * scale1 and
 * w2 = sqrt(prev_cost) * scale2. */
double or_cost_corr(const or_c128* M, int qbit_num, int cols, int stride, int variant, int trace_offset,
                    double w1, double w2);

/* A decomposition problem: target (rows x cols), circuit, cost variant. */
typedef struct {
    const or_gate* gates; int ng; int P; int qbit_num;
    const or_c128* target; int rows, cols, stride;
    int variant, trace_offset;
    /* Optimization_Interface's prev_cost_fnv_val, correction1_scale, correction2_scale */
    double prev_cost, scale1, scale2;
} or_problem;

double or_optimization_problem(const or_problem* pr, const double* params);
/* f0 and grad (P entries) from the materialised derivative matrices
 * (Optimization_Interface.cpp:1380-1478). */
int or_optimization_problem_combined(const or_problem* pr, const double* params,
                                     double* f0, double* grad);
/* B costs, params is B x P row-major; nthreads <= 1 runs sequentially. */
void or_optimization_problem_batched(const or_problem* pr, const double* params, int B,
                                     int nthreads, double* costs);
/* B cost+grad evaluations, threaded over B. */
void or_optimization_problem_combined_batched(const or_problem* pr, const double* params, int B,
                                              int nthreads, double* f0, double* grad);

/* BFGS with Powell's Z update and the Wolfe line search
 * (BFGS_Powell.cpp:70-386, grad_descend.cpp:127-272, 396-448).
 * Runs on x in place, returns the final cost; *calls receives the number of
 * cost+gradient evaluations.  status codes mirror the reference's. */
double or_bfgs_powell(const or_problem* pr, double* x, long max_calls, long* calls, int* status);

/* The same optimiser on any objective: fn(user, x, &f, g) returns f(x) and its
 * gradient (n entries). */
typedef void (*or_fg_fn)(void* user, const double* x, double* f, double* g);
double or_bfgs_powell_fn(or_fg_fn fn, void* user, long n, double* x, long max_calls, long* calls,
                         int* status);

/* Observer of the optimiser's internal steps, called on the calling thread (a test
 * hook: the arrays are valid only during the call).
 *   OR_EV_DIRECTION  d = -Z Z^T g was formed: g, d, Z (n x n row-major), Ztg = Z^T g, dg = d.g
 *   OR_EV_PROBE      a line-search probe x0 + step*d: f, dg = d.g(x), f0, dg0, f_lowest
 *                    before the Armijo test, max_step; exit_curvature = the probe ends the
 *                    search by the Wolfe tests, exit_calls = the evaluation budget ended it
 *   OR_EV_LS_END     the accepted step (the best probe), f, g at it, g0 at x0, d
 *   OR_EV_UPDATE     Powell's update applied: delta, gamma, g0, dTg = delta.gamma (clamped
 *                    below at 1.42e-14), g0_dot_delta, Z_rot = Z after the Givens sweep,
 *                    Ztg = the rotated Z^T g0, Z = the updated factor, norm_Z
 *   OR_EV_SKIP       the update was skipped (delta.gamma < 0.1 |g0.delta|)
 *   OR_EV_START      a BFGS run starts at x
 *   OR_EV_END        a BFGS run ended at x with cost f; iter holds its status        */
enum { OR_EV_DIRECTION = 0, OR_EV_PROBE = 1, OR_EV_LS_END = 2, OR_EV_UPDATE = 3, OR_EV_SKIP = 4,
       OR_EV_START = 5, OR_EV_END = 6 };
typedef struct {
    int kind;
    long n, iter;
    double step, f, dg, f0, dg0, f_lowest, max_step;
    int exit_curvature, exit_calls;
    double dTg, g0_dot_delta, norm_Z;
    const double *g, *d, *g0, *delta, *gamma, *Z, *Z_rot, *Ztg, *x;
} or_bfgs_event;
typedef void (*or_bfgs_hook)(void* user, const or_bfgs_event* ev);
void or_bfgs_set_hook(or_bfgs_hook hook, void* user);

/* mt19937 + uniform_real_distribution(0, 2pi) as consumed by the reference's
 * decomposition driver (Decomposition_Base.cpp:32-42,426-436), libstdc++ semantics. */
typedef struct { unsigned int mt[624]; int idx; } or_mt19937;
void   or_mt_seed(or_mt19937* g, unsigned int seed);
unsigned int or_mt_next(or_mt19937* g);
double or_mt_uniform_2pi(or_mt19937* g);

/* N_Qubit_Decomposition_custom + solve_optimization_problem + the BFGS engine
 * (N_Qubit_Decomposition_custom.cpp:59-190, Decomposition_Base.cpp:375-709,
 * BFGS.cpp:42-152), whole circuit optimised as one block.  x receives the
 * optimised parameters; returns the best cost; *calls counts evaluations. */
double or_decompose_custom(const or_problem* pr, unsigned int seed, int initial_guess,
                           long max_outer_iterations, long iteration_loops,
                           long max_inner_iterations, double optimization_tolerance,
                           double convergence_threshold, double* x, long* calls);

/* libstdc++ distributions on mt19937 (GCC 11 bits/uniform_int_dist.h and
 * bits/random.tcc): uniform_int_distribution<int>(0, n-1) by the nearly
 * divisionless method on the 64-bit product of one 32-bit draw and n;
 * generate_canonical<double, 53> from two draws; uniform_real_distribution(a, b)
 * = a + (b - a) * canonical. */
int    or_mt_uniform_int(or_mt19937* g, int n);
double or_mt_canonical(or_mt19937* g);
double or_mt_uniform(or_mt19937* g, double a, double b);

/* Options of the AGENTS engine with the reference defaults in brackets
 * (AGENTS.cpp:98-232; Optimization_Interface.cpp:62-64). */
typedef struct {
    int agent_num;                   /* agent_num [64] */
    long max_iterations;             /* max_inner_iterations(_agent) */
    double tolerance;                /* optimization_tolerance(_agent) */
    int lifetime;                    /* agent_lifetime [1000] */
    double exploration_rate;         /* agent_exploration_rate [0.2] */
    double agent_randomization_rate; /* agent_randomization_rate [0.2] */
    double radius;                   /* Randomized_Radius [1.0] */
    double randomization_rate;       /* randomization_rate [0.3] */
    int convergence_length;          /* convergence_length [20] */
} or_agents_opts;

/* The five-point line search of the HILBERT_SCHMIDT_TEST cost (AGENTS.cpp:544-675).
 * fv = {f(p), f(p + pi/4), f(p + pi/2), f(p + pi), f(p + 3pi/2)} along one
 * parameter whose value is p.  model receives {kappa, xi + 2p, gamma, varphi + p,
 * offset} of f(p + s) = kappa sin(2s + model[1]) + gamma sin(s + model[3]) + offset
 * (Bayes_Opt.cpp:34-40), *shift0 the start shift (the larger term's minimum, fmod
 * 2pi).  or_agents_hs_search then minimises the model over s from *shift0 with
 * BFGS_Powell capped at 10 calls; *shift receives the shift, the model's value is
 * returned and *calls the calls made. */
void   or_agents_hs_fit(const double fv[5], double p, double model[5], double* shift0);
double or_agents_hs_search(const double fv[5], double p, double* shift, long* calls);

/* Optimization_Interface::solve_layer_optimization_problem_AGENTS
 * (AGENTS.cpp:42-870) for the FROBENIUS cost (the three-point sinusoid line search
 * of AGENTS.cpp:402-443) or the HILBERT_SCHMIDT_TEST cost (the five-point search
 * above) with the optimiser's mt19937 freshly seeded with seed and the solution
 * guess in x.  x receives optimized_parameters; returns current_minimum; *iters the
 * iterations run, *evals the cost evaluations.  Other cost variants: returns
 * DBL_MAX without running (the reference has no line search for them). */
double or_agents(const or_problem* pr, unsigned int seed, const or_agents_opts* o, double* x, long* iters,
                 long* evals);

#ifdef __cplusplus
}
#endif
#endif
