/*
 * sq_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY, see sq_oracle.h).
 *
 * A restatement, in plain C, of the SQUANDER reference algorithms for the
 * gate-application / cost / gradient path and its BFGS caller.  Each function
 * cites the reference file:line it follows (paths relative to the reference
 * checkout).  It is the checker for the HIP library and the timed CPU
 * baseline ("port"); it is never linked into the product.
 */
#include "sq_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline or_c128 cmul(or_c128 a, or_c128 b) {
    or_c128 r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static inline or_c128 cadd(or_c128 a, or_c128 b) {
    or_c128 r = {a.re + b.re, a.im + b.im};
    return r;
}
static inline or_c128 mk(double re, double im) {
    or_c128 r = {re, im};
    return r;
}

/* Gate::get_parameter_num for the seven gate classes (U3.cpp:36, RX/RY/RZ/CRY: 1). */
int or_param_num(int type) {
    switch (type) {
    case OR_U3: return 3;
    case OR_RX: case OR_RY: case OR_RZ: case OR_CRY: return 1;
    case OR_CNOT: case OR_CZ: return 0;
    default: return -1;
    }
}

/* 2x2 gate kernels from raw parameters.
 * U3:  gate_kernel_templates.h:509-521 (calc_one_qubit_u3_from_trig_to)
 * RX:  gate_kernel_templates.h:193-199, RY (and CRY, CRY.cpp:51-71): 237-243,
 * RZ:  gate_kernel_templates.h:281-287.  sin/cos of the raw parameter
 * (qgd_math.h:129-143, Gates_block.cpp:269-279). */
void or_gate_kernel(int type, const double* p, or_c128 k[4]) {
    switch (type) {
    case OR_U3: {
        double st = sin(p[0]), ct = cos(p[0]);
        double sp = sin(p[1]), cp = cos(p[1]);
        double sl = sin(p[2]), cl = cos(p[2]);
        double spl = sp * cl + cp * sl, cpl = cp * cl - sp * sl;
        k[0] = mk(ct, 0.0);
        k[1] = mk(-st * cl, -st * sl);
        k[2] = mk(st * cp, st * sp);
        k[3] = mk(ct * cpl, ct * spl);
        return;
    }
    case OR_RX: {
        double s = sin(p[0]), c = cos(p[0]);
        k[0] = mk(c, 0.0); k[1] = mk(0.0, -s); k[2] = mk(0.0, -s); k[3] = mk(c, 0.0);
        return;
    }
    case OR_RY: case OR_CRY: {
        double s = sin(p[0]), c = cos(p[0]);
        k[0] = mk(c, 0.0); k[1] = mk(-s, 0.0); k[2] = mk(s, 0.0); k[3] = mk(c, 0.0);
        return;
    }
    case OR_RZ: {
        double s = sin(p[0]), c = cos(p[0]);
        k[0] = mk(c, -s); k[1] = mk(0.0, 0.0); k[2] = mk(0.0, 0.0); k[3] = mk(c, s);
        return;
    }
    case OR_CNOT:   /* X kernel (CNOT.cpp:53-74) */
        k[0] = mk(0, 0); k[1] = mk(1, 0); k[2] = mk(1, 0); k[3] = mk(0, 0);
        return;
    case OR_CZ:     /* Z kernel (CZ.cpp:54-73) */
        k[0] = mk(1, 0); k[1] = mk(0, 0); k[2] = mk(0, 0); k[3] = mk(-1, 0);
        return;
    default:
        k[0] = k[1] = k[2] = k[3] = mk(0, 0);
    }
}

/* Derivative kernels w.r.t. the raw parameter.
 * U3: gate_kernel_templates.h:555-601 (theta, phi, lambda),
 * RX: 221-227, RY/CRY: 265-271, RZ: 309-315. */
void or_deriv_kernel(int type, const double* p, int pidx, or_c128 k[4]) {
    switch (type) {
    case OR_U3: {
        double st = sin(p[0]), ct = cos(p[0]);
        double sp = sin(p[1]), cp = cos(p[1]);
        double sl = sin(p[2]), cl = cos(p[2]);
        double spl = sp * cl + cp * sl, cpl = cp * cl - sp * sl;
        if (pidx == 0) {
            k[0] = mk(-st, 0.0);
            k[1] = mk(-ct * cl, -ct * sl);
            k[2] = mk(ct * cp, ct * sp);
            k[3] = mk(-st * cpl, -st * spl);
        } else if (pidx == 1) {
            k[0] = mk(0.0, 0.0);
            k[1] = mk(0.0, 0.0);
            k[2] = mk(-st * sp, st * cp);
            k[3] = mk(-ct * spl, ct * cpl);
        } else {
            k[0] = mk(0.0, 0.0);
            k[1] = mk(st * sl, -st * cl);
            k[2] = mk(0.0, 0.0);
            k[3] = mk(-ct * spl, ct * cpl);
        }
        return;
    }
    case OR_RX: {
        double s = sin(p[0]), c = cos(p[0]);
        k[0] = mk(-s, 0.0); k[1] = mk(0.0, -c); k[2] = mk(0.0, -c); k[3] = mk(-s, 0.0);
        return;
    }
    case OR_RY: case OR_CRY: {
        double s = sin(p[0]), c = cos(p[0]);
        k[0] = mk(-s, 0.0); k[1] = mk(-c, 0.0); k[2] = mk(c, 0.0); k[3] = mk(-s, 0.0);
        return;
    }
    case OR_RZ: {
        double s = sin(p[0]), c = cos(p[0]);
        k[0] = mk(-s, -c); k[1] = mk(0.0, 0.0); k[2] = mk(0.0, 0.0); k[3] = mk(-s, c);
        return;
    }
    default:
        k[0] = k[1] = k[2] = k[3] = mk(0, 0);
    }
}

/* apply_kernel_to_input_impl (apply_kernel_to_input.cpp:35-115): for every row
 * r with target bit clear and its partner r + 2^t, update all columns when the
 * control bit of r is set (or there is no control); with deriv set, rows whose
 * control bit is clear are zeroed instead. */
void or_apply_kernel(or_c128* M, int rows, int cols, int stride, const or_c128 k[4],
                     int target, int control, int deriv) {
    const int step = 1 << target;
    for (int base = 0; base < rows; base += 2 * step) {
        for (int i = 0; i < step; i++) {
            const int r0 = base + i, r1 = r0 + step;
            or_c128* a = M + (long)r0 * stride;
            or_c128* b = M + (long)r1 * stride;
            if (control < 0 || ((r0 >> control) & 1)) {
                for (int c = 0; c < cols; c++) {
                    or_c128 x = a[c], y = b[c];
                    a[c] = cadd(cmul(k[0], x), cmul(k[1], y));
                    b[c] = cadd(cmul(k[2], x), cmul(k[3], y));
                }
            } else if (deriv) {
                memset(a, 0, sizeof(or_c128) * (size_t)cols);
                memset(b, 0, sizeof(or_c128) * (size_t)cols);
            }
        }
    }
}

/* apply_X_kernel_to_input_impl (apply_dedicated_gate_kernel_to_input.cpp:46-93). */
void or_apply_x(or_c128* M, int rows, int cols, int stride, int target, int control) {
    const int step = 1 << target;
    for (int base = 0; base < rows; base += 2 * step) {
        for (int i = 0; i < step; i++) {
            const int r0 = base + i, r1 = r0 + step;
            if (control >= 0 && !((r0 >> control) & 1)) continue;
            or_c128* a = M + (long)r0 * stride;
            or_c128* b = M + (long)r1 * stride;
            for (int c = 0; c < cols; c++) { or_c128 t = a[c]; a[c] = b[c]; b[c] = t; }
        }
    }
}

/* apply_Z_kernel_to_input_impl (apply_dedicated_gate_kernel_to_input.cpp:221-254). */
void or_apply_z(or_c128* M, int rows, int cols, int stride, int target, int control) {
    const int step = 1 << target;
    for (int base = 0; base < rows; base += 2 * step) {
        for (int i = 0; i < step; i++) {
            const int r0 = base + i, r1 = r0 + step;
            if (control >= 0 && !((r0 >> control) & 1)) continue;
            or_c128* b = M + (long)r1 * stride;
            for (int c = 0; c < cols; c++) { b[c].re = -b[c].re; b[c].im = -b[c].im; }
        }
    }
}

/* Gate::apply_to_inner + Gate::apply_kernel_to dispatch (Gate.cpp:549-570, 1478-1768). */
int or_apply_gate(const or_gate* g, const double* params, or_c128* M, int rows, int cols, int stride) {
    or_c128 k[4];
    switch (g->type) {
    case OR_CNOT: or_apply_x(M, rows, cols, stride, g->target, g->control); return 0;
    case OR_CZ:   or_apply_z(M, rows, cols, stride, g->target, g->control); return 0;
    case OR_U3: case OR_RX: case OR_RY: case OR_RZ:
        or_gate_kernel(g->type, params + g->param_start, k);
        or_apply_kernel(M, rows, cols, stride, k, g->target, -1, 0);
        return 0;
    case OR_CRY:
        or_gate_kernel(g->type, params + g->param_start, k);
        or_apply_kernel(M, rows, cols, stride, k, g->target, g->control, 0);
        return 0;
    default:
        return -1;
    }
}

/* Gates_block::apply_to_inner gate loop (Gates_block.cpp:683-708). */
int or_apply_circuit(const or_gate* gates, int ng, const double* params,
                     or_c128* M, int rows, int cols, int stride) {
    for (int j = 0; j < ng; j++)
        if (or_apply_gate(&gates[j], params, M, rows, cols, stride)) return -1;
    return 0;
}

/* Gates_block::apply_derivate_to, prefix-only variant (Gates_block.cpp:1012-1150,
 * build_forward_inputs :430-470): the forward input of gate j is copied, the
 * derivative kernel of each of its parameters is applied with deriv zeroing
 * (Gate.cpp:645-706), then the remaining gates j+1.. are applied. */
int or_apply_derivate(const or_gate* gates, int ng, const double* params,
                      const or_c128* in, int rows, int cols, int stride, or_c128* out) {
    const size_t msz = (size_t)rows * cols;
    or_c128* X = (or_c128*)malloc(sizeof(or_c128) * msz);
    if (!X) return -1;
    for (int r = 0; r < rows; r++) memcpy(X + (size_t)r * cols, in + (size_t)r * stride, sizeof(or_c128) * cols);
    int rc = 0;
    for (int j = 0; j < ng && !rc; j++) {
        const or_gate* g = &gates[j];
        const int np = or_param_num(g->type);
        for (int q = 0; q < np; q++) {
            or_c128* D = out + (size_t)(g->param_start + q) * msz;
            or_c128 k[4];
            memcpy(D, X, sizeof(or_c128) * msz);
            or_deriv_kernel(g->type, params + g->param_start, q, k);
            or_apply_kernel(D, rows, cols, cols, k, g->target, g->type == OR_CRY ? g->control : -1, 1);
            for (int jj = j + 1; jj < ng; jj++)
                if (or_apply_gate(&gates[jj], params, D, rows, cols, cols)) rc = -1;
        }
        if (or_apply_gate(g, params, X, rows, cols, cols)) rc = -1;
    }
    free(X);
    return rc;
}

/* get_cost_function's trace with offset (N_Qubit_Decomposition_Cost_Function.cpp:137-158),
 * accumulated as a complex value. */
or_c128 or_trace_offset(const or_c128* M, int cols, int stride, int trace_offset) {
    or_c128 t = {0.0, 0.0};
    for (int i = 0; i < cols; i++) {
        const or_c128 v = M[(long)(i + trace_offset) * stride + i];
        t.re += v.re; t.im += v.im;
    }
    return t;
}

/* get_trace (N_Qubit_Decomposition_Cost_Function.cpp:482-499): plain diagonal, no offset. */
or_c128 or_trace_plain(const or_c128* M, int cols, int stride) {
    return or_trace_offset(M, cols, stride, 0);
}

/* Cost variants: FROBENIUS :137-158, HS :524-532, INFIDELITY :541-549; ABS_TRACE is
 * 1 - |Tr_offset|/m (no reference counterpart). */
double or_cost(const or_c128* M, int cols, int stride, int variant, int trace_offset) {
    const double m = (double)cols;
    switch (variant) {
    case OR_COST_FROBENIUS: {
        or_c128 t = or_trace_offset(M, cols, stride, trace_offset);
        return 1.0 - t.re / m;
    }
    case OR_COST_HS: {
        or_c128 t = or_trace_plain(M, cols, stride);
        const double d = 1.0 / m;
        return 1.0 - d * d * (t.re * t.re + t.im * t.im);
    }
    case OR_COST_INFIDELITY: {
        or_c128 t = or_trace_plain(M, cols, stride);
        return 1.0 - ((t.re * t.re + t.im * t.im) / m + 1.0) / (m + 1.0);
    }
    case OR_COST_ABS_TRACE: {
        or_c128 t = or_trace_offset(M, cols, stride, trace_offset);
        return 1.0 - sqrt(t.re * t.re + t.im * t.im) / m;
    }
    default:
        return NAN;
    }
}

/* The bit-flip sums of get_cost_function_with_correction[2]
 * (N_Qubit_Decomposition_Cost_Function.cpp:191-252, 291-380; loops over the qubit
 * (pair) outside, columns inside) and get_trace_with_correction[2] (:564-664, the
 * same sums with trace_offset 0, real and imaginary parts). */
or_c128 or_trace_flips(const or_c128* M, int qbit_num, int cols, int stride, int trace_offset, int bits) {
    or_c128 t = {0.0, 0.0};
    for (int q = 0; q < qbit_num; q++) {
        if (bits == 1) {
            const int mask = 1 << q;
            for (int c = 0; c < cols; c++) {
                const or_c128 v = M[(long)((c + trace_offset) ^ mask) * stride + c];
                t.re += v.re; t.im += v.im;
            }
        } else {
            for (int q2 = q + 1; q2 < qbit_num; q2++) {
                const int mask = (1 << q) + (1 << q2);
                for (int c = 0; c < cols; c++) {
                    const or_c128 v = M[(long)((c + trace_offset) ^ mask) * stride + c];
                    t.re += v.re; t.im += v.im;
                }
            }
        }
    }
    return t;
}

/* calculate_cost_function (Optimization_Interface.cpp:676-716): FROBENIUS_CORRECTION1/2
 * = get_cost_function - sqrt(prev) * (scale1 * flips1 + scale2 * flips2) / m;
 * HILBERT_SCHMIDT_TEST_CORRECTION1/2 = 1 - (|T|^2 + sqrt(prev) * (scale1 |T1|^2 +
 * scale2 |T2|^2)) / m^2 on the plain trace T and the offset-free flip traces. */
double or_cost_corr(const or_c128* M, int qbit_num, int cols, int stride, int variant, int trace_offset,
                    double w1, double w2) {
    const double m = (double)cols;
    switch (variant) {
    case OR_COST_FROBENIUS_C1:
    case OR_COST_FROBENIUS_C2: {
        const or_c128 t0 = or_trace_offset(M, cols, stride, trace_offset);
        const or_c128 t1 = or_trace_flips(M, qbit_num, cols, stride, trace_offset, 1);
        double c = (1.0 - t0.re / m) - w1 * (t1.re / m);
        if (variant == OR_COST_FROBENIUS_C2) {
            const or_c128 t2 = or_trace_flips(M, qbit_num, cols, stride, trace_offset, 2);
            c -= w2 * (t2.re / m);
        }
        return c;
    }
    case OR_COST_HS_C1:
    case OR_COST_HS_C2: {
        const or_c128 t0 = or_trace_plain(M, cols, stride);
        const or_c128 t1 = or_trace_flips(M, qbit_num, cols, stride, 0, 1);
        const double d = 1.0 / m;
        double s = t0.re * t0.re + t0.im * t0.im + w1 * (t1.re * t1.re + t1.im * t1.im);
        if (variant == OR_COST_HS_C2) {
            const or_c128 t2 = or_trace_flips(M, qbit_num, cols, stride, 0, 2);
            s += w2 * (t2.re * t2.re + t2.im * t2.im);
        }
        return 1.0 - d * d * s;
    }
    default:
        return or_cost(M, cols, stride, variant, trace_offset);
    }
}

static double pr_w1(const or_problem* pr) { return sqrt(pr->prev_cost) * pr->scale1; }
static double pr_w2(const or_problem* pr) { return sqrt(pr->prev_cost) * pr->scale2; }

/* Optimization_Interface::optimization_problem (Optimization_Interface.cpp:634-668):
 * copy the target, apply the circuit, reduce the cost. */
double or_optimization_problem(const or_problem* pr, const double* params) {
    const size_t msz = (size_t)pr->rows * pr->cols;
    or_c128* M = (or_c128*)malloc(sizeof(or_c128) * msz);
    if (!M) return NAN;
    for (int r = 0; r < pr->rows; r++)
        memcpy(M + (size_t)r * pr->cols, pr->target + (size_t)r * pr->stride, sizeof(or_c128) * pr->cols);
    or_apply_circuit(pr->gates, pr->ng, params, M, pr->rows, pr->cols, pr->cols);
    double c = or_cost_corr(M, pr->qbit_num, pr->cols, pr->cols, pr->variant, pr->trace_offset, pr_w1(pr),
                            pr_w2(pr));
    free(M);
    return c;
}

/* optimization_problem_combined_non_static, CPU branch (Optimization_Interface.cpp:1380-1478):
 * f0 from the forward product, gradient components from the derivative matrices. */
int or_optimization_problem_combined(const or_problem* pr, const double* params, double* f0, double* grad) {
    const int rows = pr->rows, cols = pr->cols, P = pr->P;
    const size_t msz = (size_t)rows * cols;
    or_c128* M = (or_c128*)malloc(sizeof(or_c128) * msz);
    or_c128* D = (or_c128*)malloc(sizeof(or_c128) * msz * (size_t)(P > 0 ? P : 1));
    if (!M || !D) { free(M); free(D); return -1; }
    for (int r = 0; r < rows; r++)
        memcpy(M + (size_t)r * cols, pr->target + (size_t)r * pr->stride, sizeof(or_c128) * cols);
    or_apply_derivate(pr->gates, pr->ng, params, M, rows, cols, cols, D);
    or_apply_circuit(pr->gates, pr->ng, params, M, rows, cols, cols);
    const double w1 = pr_w1(pr), w2 = pr_w2(pr);
    *f0 = or_cost_corr(M, pr->qbit_num, cols, cols, pr->variant, pr->trace_offset, w1, w2);
    const double m = (double)cols;
    or_c128 t = (pr->variant == OR_COST_FROBENIUS || pr->variant == OR_COST_ABS_TRACE)
                    ? or_trace_offset(M, cols, cols, pr->trace_offset)
                    : or_trace_plain(M, cols, cols);
    /* offset-free flip traces of C.U for the Hilbert-Schmidt corrections */
    const or_c128 t1 = or_trace_flips(M, pr->qbit_num, cols, cols, 0, 1);
    const or_c128 t2 = or_trace_flips(M, pr->qbit_num, cols, cols, 0, 2);
    for (int k = 0; k < P; k++) {
        const or_c128* Dk = D + (size_t)k * msz;
        switch (pr->variant) {
        case OR_COST_FROBENIUS:
            grad[k] = (1.0 - or_trace_offset(Dk, cols, cols, pr->trace_offset).re / m) - 1.0;
            break;
        /* Optimization_Interface.cpp:1396-1405: the derivative's correction cost minus 1 */
        case OR_COST_FROBENIUS_C1:
        case OR_COST_FROBENIUS_C2: {
            const or_c128 d0 = or_trace_offset(Dk, cols, cols, pr->trace_offset);
            const or_c128 d1 = or_trace_flips(Dk, pr->qbit_num, cols, cols, pr->trace_offset, 1);
            double g = (1.0 - d0.re / m) - w1 * (d1.re / m);
            if (pr->variant == OR_COST_FROBENIUS_C2)
                g -= w2 * (or_trace_flips(Dk, pr->qbit_num, cols, cols, pr->trace_offset, 2).re / m);
            grad[k] = g - 1.0;
            break;
        }
        /* Optimization_Interface.cpp:1412-1423 */
        case OR_COST_HS_C1:
        case OR_COST_HS_C2: {
            const or_c128 d0 = or_trace_plain(Dk, cols, cols);
            const or_c128 d1 = or_trace_flips(Dk, pr->qbit_num, cols, cols, 0, 1);
            const double d = 1.0 / m;
            double s = t.re * d0.re + t.im * d0.im + w1 * (t1.re * d1.re + t1.im * d1.im);
            if (pr->variant == OR_COST_HS_C2) {
                const or_c128 d2 = or_trace_flips(Dk, pr->qbit_num, cols, cols, 0, 2);
                s += w2 * (t2.re * d2.re + t2.im * d2.im);
            }
            grad[k] = -2.0 * d * d * s;
            break;
        }
        case OR_COST_HS: {
            or_c128 tk = or_trace_plain(Dk, cols, cols);
            const double d = 1.0 / m;
            grad[k] = -2.0 * d * d * t.re * tk.re - 2.0 * d * d * t.im * tk.im;
            break;
        }
        case OR_COST_INFIDELITY: {
            or_c128 tk = or_trace_plain(Dk, cols, cols);
            grad[k] = -2.0 / m / (m + 1.0) * t.re * tk.re - 2.0 / m / (m + 1.0) * t.im * tk.im;
            break;
        }
        case OR_COST_ABS_TRACE: {
            or_c128 tk = or_trace_offset(Dk, cols, cols, pr->trace_offset);
            const double a = sqrt(t.re * t.re + t.im * t.im);
            grad[k] = a > 0.0 ? -(t.re * tk.re + t.im * tk.im) / (a * m) : 0.0;
            break;
        }
        default:
            grad[k] = NAN;
        }
    }
    free(M);
    free(D);
    return 0;
}

/* optimization_problem_batched (Optimization_Interface.cpp:940-1033): independent
 * evaluations, parallel over the batch. */
void or_optimization_problem_batched(const or_problem* pr, const double* params, int B,
                                     int nthreads, double* costs) {
#ifdef _OPENMP
    if (nthreads > 1) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
        for (int b = 0; b < B; b++) costs[b] = or_optimization_problem(pr, params + (size_t)b * pr->P);
        return;
    }
#endif
    (void)nthreads;
    for (int b = 0; b < B; b++) costs[b] = or_optimization_problem(pr, params + (size_t)b * pr->P);
}

void or_optimization_problem_combined_batched(const or_problem* pr, const double* params, int B,
                                              int nthreads, double* f0, double* grad) {
#ifdef _OPENMP
    if (nthreads > 1) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
        for (int b = 0; b < B; b++)
            or_optimization_problem_combined(pr, params + (size_t)b * pr->P, f0 + b, grad + (size_t)b * pr->P);
        return;
    }
#endif
    (void)nthreads;
    for (int b = 0; b < B; b++)
        or_optimization_problem_combined(pr, params + (size_t)b * pr->P, f0 + b, grad + (size_t)b * pr->P);
}

/* ------------------------------------------------------------------------- */
/* BFGS with Powell's update of Z (Z Z^T approximates the inverse Hessian).  */
/* ------------------------------------------------------------------------- */

enum { ST_INITIAL = 0, ST_VARS_INIT = 1, ST_ZERO_STEP = 2, ST_MAX_ITER = 3,
       ST_NO_DESCENT = 4, ST_MINIMUM = 5 };

typedef struct {
    or_fg_fn fn;
    void* user;
    long n, calls, max_calls;
    int status;
    double* Z;      /* n x n row-major */
    double* Ztg;    /* Z^T g */
} bfgs_state;

/* Optional observer of the optimiser's internal steps (test hook, per thread). */
static _Thread_local or_bfgs_hook g_hook = 0;
static _Thread_local void* g_hook_user = 0;

void or_bfgs_set_hook(or_bfgs_hook hook, void* user) { g_hook = hook; g_hook_user = user; }

static void emit(or_bfgs_event* ev) { if (g_hook) g_hook(g_hook_user, ev); }

static void eval_fg(bfgs_state* s, const double* x, double* f, double* g) {
    s->calls++;
    s->fn(s->user, x, f, g);
}

static void problem_fg(void* user, const double* x, double* f, double* g) {
    or_optimization_problem_combined((const or_problem*)user, x, f, g);
}

/* Grad_Descend::get_Maximal_Line_Search_Step (grad_descend.cpp:396-425). */
static double maximal_step_of(const double* d, long n) {
    double ms = 0.0;
    for (long k = 0; k < n; k++) {
        if (fabs(d[k]) < 1e-5) continue;
        double b = fabs(2.0 * M_PI / d[k]);
        if (ms == 0.0 || b < ms) ms = b;
    }
    return ms;
}

/* Grad_Descend::line_search (grad_descend.cpp:127-272). */
static void line_search(bfgs_state* s, double* x, double* g, const double* dir, double* x0,
                        double* g0, double maximal_step, double dg0, double* f) {
    const long n = s->n;
    memcpy(x0, x, sizeof(double) * n);
    memcpy(g0, g, sizeof(double) * n);
    double* g_best = (double*)malloc(sizeof(double) * n);
    memcpy(g_best, g, sizeof(double) * n);
    const long max_loops = 50;
    double step = fmin(1.0, maximal_step);
    const double f0 = *f;
    double f_lowest = *f, f_highest = *f;
    double step_highest = maximal_step, step_lowest = 0.0;
    double step_best = 0.0, f_best = *f, dg_abs_best = fabs(dg0), dg_highest = 0.0;
    const double num_precision = 1.42e-14;

    for (long it = 0; it < max_loops; it++) {
        for (long k = 0; k < n; k++) x[k] = x0[k] + step * dir[k];
        eval_fg(s, x, f, g);
        double dg = 0.0;
        for (long k = 0; k < n; k++) dg += dir[k] * g[k];
        if (*f < f_best || fabs(dg) < dg_abs_best) {
            step_best = step; f_best = *f; dg_abs_best = fabs(dg);
            memcpy(g_best, g, sizeof(double) * n);
        }
        if (g_hook) {
            or_bfgs_event ev = {0};
            ev.kind = OR_EV_PROBE; ev.n = n; ev.iter = it; ev.step = step; ev.f = *f; ev.dg = dg;
            ev.f0 = f0; ev.dg0 = dg0; ev.f_lowest = f_lowest; ev.max_step = maximal_step;
            ev.exit_calls = (s->calls == s->max_calls);
            ev.exit_curvature = !ev.exit_calls && (*f < f_lowest + step * 0.1 * dg0) && (dg >= dg0 * 0.7);
            ev.g = g; ev.d = dir;
            emit(&ev);
        }
        if (s->calls == s->max_calls) break;
        if (*f < f_lowest + step * 0.1 * dg0) {
            if (dg >= dg0 * 0.7) break;
            step_lowest = step; f_lowest = *f;
        } else {
            step_highest = step; f_highest = *f; dg_highest = dg;
        }
        if (it == 0 || step_lowest > 0.0) {
            double dg_low_fit = (f_highest - f_lowest) / (step_highest - step_lowest) * 2.0 - dg_highest;
            double scale = -dg_low_fit * 0.5 / (dg_highest - dg_low_fit);
            scale = fmax(0.1, scale);
            step = step_lowest + scale * (step_highest - step_lowest);
        } else {
            step = step * 0.1;
            if (step < num_precision) break;
        }
    }
    if (step != step_best) {
        step = step_best;
        *f = f_best;
        memcpy(g, g_best, sizeof(double) * n);
        for (long k = 0; k < n; k++) x[k] = x0[k] + step * dir[k];
    }
    if (step == 0.0) s->status = ST_ZERO_STEP;
    if (g_hook) {
        or_bfgs_event ev = {0};
        ev.kind = OR_EV_LS_END; ev.n = n; ev.step = step; ev.f = *f; ev.f0 = f0; ev.dg0 = dg0;
        ev.max_step = maximal_step; ev.g = g; ev.d = dir; ev.g0 = g0;
        emit(&ev);
    }
    free(g_best);
}

/* BFGS_Powell::get_search_direction (BFGS_Powell.cpp:396-439): d = -Z Z^T g. */
static double search_direction(bfgs_state* s, const double* g, double* d) {
    const long n = s->n;
    memset(s->Ztg, 0, sizeof(double) * n);
    for (long r = 0; r < n; r++)
        for (long c = 0; c < n; c++) s->Ztg[c] += s->Z[r * n + c] * g[r];
    for (long r = 0; r < n; r++) {
        d[r] = 0.0;
        for (long c = 0; c < n; c++) d[r] = d[r] - s->Z[r * n + c] * s->Ztg[c];
    }
    double dg = 0.0;
    for (long k = 0; k < n; k++) dg += d[k] * g[k];
    if (g_hook) {
        or_bfgs_event ev = {0};
        ev.kind = OR_EV_DIRECTION; ev.n = n; ev.dg = dg; ev.g = g; ev.d = d; ev.Z = s->Z; ev.Ztg = s->Ztg;
        emit(&ev);
    }
    return dg;
}

/* BFGS_Powell::BFGS_iteration (BFGS_Powell.cpp:202-386). */
static void bfgs_iteration(bfgs_state* s, const double* x, const double* g, const double* x0,
                           const double* g0, double* norm_Z) {
    const long n = s->n;
    const double num_precision = 1.42e-14;
    double* delta = (double*)malloc(sizeof(double) * n);
    double* gamma = (double*)malloc(sizeof(double) * n);
    double norm_delta = 0.0, dTg = 0.0, tmp = 0.0;
    for (long k = 0; k < n; k++) {
        delta[k] = x[k] - x0[k];
        norm_delta += delta[k] * delta[k];
        tmp += g0[k] * delta[k];
        gamma[k] = g[k] - g0[k];
        dTg += gamma[k] * delta[k];
    }
    if (dTg < fabs(tmp) * 0.1) {
        if (g_hook) {
            or_bfgs_event ev = {0};
            ev.kind = OR_EV_SKIP; ev.n = n; ev.dTg = dTg; ev.g0_dot_delta = tmp;
            ev.delta = delta; ev.gamma = gamma; ev.g0 = g0; ev.Z = s->Z;
            emit(&ev);
        }
        free(delta); free(gamma); return;
    }
    dTg = fmax(num_precision, dTg);

    /* Givens rotations folding Z^T g into its first entry (Eq. 1.7 of Powell 1987). */
    for (long c = n - 1; c > 0; c--) {
        if (s->Ztg[c] == 0.0) continue;
        const long cm = c - 1;
        double a = sqrt(s->Ztg[cm] * s->Ztg[cm] + s->Ztg[c] * s->Ztg[c]);
        double cv = s->Ztg[cm] / a, sv = s->Ztg[c] / a;
        s->Ztg[cm] = a;
        for (long r = 0; r < n; r++) {
            double* z = s->Z + r * n + cm;
            double t = cv * z[1] - sv * z[0];
            z[0] = cv * z[0] + sv * z[1];
            z[1] = t;
        }
    }
    double* Z_rot = 0;
    if (g_hook) {
        Z_rot = (double*)malloc(sizeof(double) * n * n);
        memcpy(Z_rot, s->Z, sizeof(double) * n * n);
    }
    if (*norm_Z < 0.0) {
        *norm_Z = norm_delta / dTg;
    } else {
        tmp = sqrt(*norm_Z * norm_delta / dTg);
        *norm_Z = fmin(*norm_Z, tmp);
        *norm_Z = fmax(*norm_Z, tmp * 0.1);
    }
    tmp = sqrt(dTg);
    for (long k = 0; k < n; k++) s->Z[k * n] = fabs(delta[k]) > num_precision ? delta[k] / tmp : 0.0;

    double* gTZ = (double*)calloc((size_t)n, sizeof(double));
    double* nz = (double*)calloc((size_t)n, sizeof(double));
    for (long r = 0; r < n; r++)
        for (long c = 1; c < n; c++) gTZ[c] += gamma[r] * s->Z[r * n + c];
    for (long c = 1; c < n; c++) gTZ[c] = gTZ[c] / dTg;
    for (long r = 0; r < n; r++)
        for (long c = 1; c < n; c++) {
            double* z = s->Z + r * n + c;
            *z = *z - gTZ[c] * delta[r];
            nz[c] += (*z) * (*z);
        }
    for (long c = 1; c < n; c++) {
        if (nz[c] < *norm_Z && nz[c] > num_precision) {
            double sc = sqrt(*norm_Z / nz[c]);
            for (long r = 0; r < n; r++) s->Z[r * n + c] *= sc;
        }
    }
    if (g_hook) {
        or_bfgs_event ev = {0};
        ev.kind = OR_EV_UPDATE; ev.n = n; ev.dTg = dTg; ev.g0_dot_delta = 0.0;
        for (long k = 0; k < n; k++) ev.g0_dot_delta += g0[k] * delta[k];
        ev.norm_Z = *norm_Z; ev.delta = delta; ev.gamma = gamma; ev.g0 = g0; ev.Z = s->Z;
        ev.Z_rot = Z_rot; ev.Ztg = s->Ztg;
        emit(&ev);
        free(Z_rot);
    }
    free(gTZ); free(nz); free(delta); free(gamma);
}

/* BFGS_Powell::Optimize (BFGS_Powell.cpp:70-192) via Grad_Descend::Start_Optimization
 * (grad_descend.cpp:83-111). */
double or_bfgs_powell_fn(or_fg_fn fn, void* user, long n, double* x, long max_calls, long* calls,
                         int* status) {
    bfgs_state s;
    s.fn = fn; s.user = user; s.n = n; s.calls = 0; s.max_calls = max_calls; s.status = ST_INITIAL;
    if (n <= 0) { if (calls) *calls = 0; if (status) *status = ST_INITIAL; return DBL_MAX; }
    s.Z = (double*)calloc((size_t)(n * n), sizeof(double));
    s.Ztg = (double*)calloc((size_t)n, sizeof(double));
    for (long k = 0; k < n; k++) s.Z[k * n + k] = 1.0;
    s.status = ST_VARS_INIT;
    double* g = (double*)calloc((size_t)n, sizeof(double));
    double* d = (double*)calloc((size_t)n, sizeof(double));
    double* x0 = (double*)calloc((size_t)n, sizeof(double));
    double* g0 = (double*)calloc((size_t)n, sizeof(double));
    double f, norm_Z = -1.0;
    if (g_hook) {
        or_bfgs_event ev = {0};
        ev.kind = OR_EV_START; ev.n = n; ev.x = x;
        emit(&ev);
    }
    eval_fg(&s, x, &f, g);
    double fprev = fabs(f + f + 1.0);
    for (;;) {
        double ms = 0.0;
        double dg = search_direction(&s, g, d);
        if (dg < 0.0) ms = maximal_step_of(d, n);
        double gn = 0.0;
        for (long k = 0; k < n; k++) gn += g[k] * g[k];
        const double acc = 1e-19;
        if (gn <= acc * acc) { s.status = ST_MINIMUM; break; }
        if (dg >= 0.0) { s.status = ST_NO_DESCENT; break; }
        if (f >= fprev) { s.status = ST_MINIMUM; break; }
        fprev = f;
        if (s.calls >= s.max_calls) { s.status = ST_MAX_ITER; break; }
        line_search(&s, x, g, d, x0, g0, ms, dg, &f);
        if (s.status == ST_ZERO_STEP) break;
        bfgs_iteration(&s, x, g, x0, g0, &norm_Z);
    }
    if (g_hook) {
        or_bfgs_event ev = {0};
        ev.kind = OR_EV_END; ev.n = n; ev.x = x; ev.f = f; ev.iter = s.status;
        emit(&ev);
    }
    if (calls) *calls = s.calls;
    if (status) *status = s.status;
    free(s.Z); free(s.Ztg); free(g); free(d); free(x0); free(g0);
    return f;
}

double or_bfgs_powell(const or_problem* pr, double* x, long max_calls, long* calls, int* status) {
    return or_bfgs_powell_fn(problem_fg, (void*)pr, pr->P, x, max_calls, calls, status);
}

/* ------------------------------------------------------------------------- */
/* std::mt19937 and uniform_real_distribution<double>(0, 2pi), libstdc++.     */
/* ------------------------------------------------------------------------- */

void or_mt_seed(or_mt19937* g, unsigned int seed) {
    g->mt[0] = seed;
    for (int i = 1; i < 624; i++)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (unsigned int)i;
    g->idx = 624;
}

unsigned int or_mt_next(or_mt19937* g) {
    if (g->idx >= 624) {
        for (int i = 0; i < 624; i++) {
            unsigned int y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
            unsigned int v = g->mt[(i + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            g->mt[i] = v;
        }
        g->idx = 0;
    }
    unsigned int y = g->mt[g->idx++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

/* generate_canonical<double, 53> draws two 32-bit words, low word first. */
double or_mt_uniform_2pi(or_mt19937* g) {
    const double r = 4294967296.0;
    double sum = (double)or_mt_next(g);
    sum += (double)or_mt_next(g) * r;
    double u = sum / (r * r);
    if (u >= 1.0) u = nextafter(1.0, 0.0);
    return u * (2.0 * M_PI - 0.0) + 0.0;
}

static unsigned int mt_uniform_range(or_mt19937* g, unsigned int range);

/* Decomposition driver for a custom circuit optimised as one block:
 * initial guess (Decomposition_Base.cpp:415-446), outer loop with the 20-entry
 * convergence window (:492-675, 742-754), BFGS restarts with U[0,2pi)/100 jitter
 * after each improvement (BFGS.cpp:42-152). */
double or_decompose_custom(const or_problem* pr, unsigned int seed, int initial_guess,
                           long max_outer_iterations, long iteration_loops,
                           long max_inner_iterations, double optimization_tolerance,
                           double convergence_threshold, double* x, long* calls) {
    const int P = pr->P;
    or_mt19937 gen;
    or_mt_seed(&gen, seed);
    /* The decomposition object's constructor draws its id from the same generator
     * (Optimization_Interface.cpp:177-180: uniform_int_distribution<>(0, INT_MAX)). */
    (void)mt_uniform_range(&gen, 2147483648u);
    double* opt = (double*)calloc((size_t)(P > 0 ? P : 1), sizeof(double));
    double* guess = (double*)calloc((size_t)(P > 0 ? P : 1), sizeof(double));
    for (int k = 0; k < P; k++) {
        if (initial_guess == 0) opt[k] = 0.0;                      /* ZEROS */
        else if (initial_guess == 1) opt[k] = or_mt_uniform_2pi(&gen); /* RANDOM */
        else opt[k] = or_mt_uniform_2pi(&gen) / 100.0;             /* CLOSE_TO_ZERO */
    }
    enum { NWIN = 20 };
    double minvec[NWIN] = {0};
    double current_minimum = DBL_MAX;
    long total_calls = 0;
    if (iteration_loops < 1) iteration_loops = 1;
    double* best = (double*)calloc((size_t)(P > 0 ? P : 1), sizeof(double));
    int have_best = 0;
    for (long it = 0; it < max_outer_iterations; it++) {
        memcpy(guess, opt, sizeof(double) * P);
        have_best = 0;
        for (long l = 0; l < iteration_loops; l++) {
            long c = 0;
            int st = 0;
            double f = or_bfgs_powell(pr, guess, max_inner_iterations, &c, &st);
            total_calls += c;
            if (current_minimum > f) {
                current_minimum = f;
                memcpy(best, guess, sizeof(double) * P);
                have_best = 1;
                for (int k = 0; k < P; k++) guess[k] = guess[k] + or_mt_uniform_2pi(&gen) / 100.0;
            }
        }
        if (have_best) memcpy(opt, best, sizeof(double) * P);
        double mean = 0.0;
        for (int i = NWIN - 1; i > 0; i--) { minvec[i] = minvec[i - 1]; mean += minvec[i - 1]; }
        minvec[0] = current_minimum;
        mean = (mean + current_minimum) / NWIN;
        double sd = 0.0;
        for (int i = 0; i < NWIN; i++) { double t = minvec[i] - mean; sd += t * t; }
        sd = sqrt(sd / (NWIN - 1));
        if (fabs(sd / minvec[NWIN - 1]) < convergence_threshold) break;
        if (fabs(current_minimum - 0.0) < optimization_tolerance) break;
    }
    memcpy(x, opt, sizeof(double) * P);
    if (calls) *calls = total_calls;
    free(opt); free(guess); free(best);
    return current_minimum;
}

/* ---- libstdc++ distributions ------------------------------------------------ */

/* uniform_int_distribution::operator() with a 32-bit engine range: _S_nd<uint64>
 * (Lemire, "Fast random integer generation in an interval", 2019), for a range
 * of 1 .. 2^31 values. */
static unsigned int mt_uniform_range(or_mt19937* g, unsigned int range) {
    unsigned long long product = (unsigned long long)or_mt_next(g) * range;
    unsigned int low = (unsigned int)product;
    if (low < range) {
        const unsigned int threshold = (unsigned int)(-range) % range;
        while (low < threshold) {
            product = (unsigned long long)or_mt_next(g) * range;
            low = (unsigned int)product;
        }
    }
    return (unsigned int)(product >> 32);
}

int or_mt_uniform_int(or_mt19937* g, int n) { return (int)mt_uniform_range(g, (unsigned int)n); }

double or_mt_canonical(or_mt19937* g) {
    const double r = 4294967296.0;
    double sum = (double)or_mt_next(g);
    sum += (double)or_mt_next(g) * r;
    double u = sum / (r * r);
    if (u >= 1.0) u = nextafter(1.0, 0.0);
    return u;
}

double or_mt_uniform(or_mt19937* g, double a, double b) { return or_mt_canonical(g) * (b - a) + a; }

/* ---- AGENTS -------------------------------------------------------------------- */

/* Optimization_Interface::randomize_parameters (Optimization_Interface.cpp:576-602),
 * non-VQE branch. */
static void or_randomize_parameters(or_mt19937* g, const double* in, double* out, int P, double f0, double radius,
                                    double rate) {
    for (int j = 0; j < P; j++) {
        if (or_mt_canonical(g) <= rate)
            out[j] = in[j] + or_mt_uniform(g, -2.0 * M_PI, 2.0 * M_PI) * sqrt(f0) * radius;
        else
            out[j] = in[j];
    }
}

/* AGENTS.cpp:611-664: f(p') = kappa sin(2p' + xi) + gamma sin(p' + phi) + offset
 * from its values at p, p + pi/4, p + pi/2, p + pi and p + 3pi/2. */
void or_agents_hs_fit(const double fv[5], double p, double model[5], double* shift0) {
    const double f0 = fv[0], f_pi4 = fv[1], f_pi2 = fv[2], f_pi = fv[3], f_3pi2 = fv[4];
    const double f1 = f0 - f_pi;      /* 2 gamma sin(p + phi) */
    const double f2 = f_pi2 - f_3pi2; /* 2 gamma cos(p + phi) */
    const double gamma = sqrt(f1 * f1 + f2 * f2) * 0.5;
    const double varphi = atan2(f1, f2) - p;
    const double offset = 0.25 * (f0 + f_pi + f_pi2 + f_3pi2);
    const double f3 = 0.5 * (f0 + f_pi - 2 * offset);                           /* kappa sin(2p + xi) */
    const double f4 = f_pi4 - offset - gamma * sin(p + M_PI / 4 + varphi);     /* kappa cos(2p + xi) */
    const double kappa = sqrt(f3 * f3 + f4 * f4);
    const double xi = atan2(f3, f4) - 2 * p;
    model[0] = kappa;
    model[1] = xi + 2 * p;
    model[2] = gamma;
    model[3] = varphi + p;
    model[4] = offset;
    /* the larger term's minimiser; the double-frequency start subtracts p / 2, not p,
     * as the reference does (AGENTS.cpp:655-660) */
    double s = fabs(gamma) > fabs(kappa) ? 3 * M_PI / 2 - varphi - p : 3 * M_PI / 4 - xi / 2 - p / 2;
    *shift0 = fmod(s, M_PI * 2);
}

/* HS_partial_optimization_problem_combined (Bayes_Opt.cpp:34-40,58-66,83-88). */
static void hs_model_fg(void* user, const double* s, double* f, double* g) {
    const double* m = (const double*)user;
    *f = m[0] * sin(2 * s[0] + m[1]) + m[2] * sin(s[0] + m[3]) + m[4];
    g[0] = 2 * m[0] * cos(2 * s[0] + m[1]) + m[2] * cos(s[0] + m[3]);
}

/* AGENTS.cpp:664-667: BFGS_Powell on the model, Start_Optimization(shift, 10). */
double or_agents_hs_search(const double fv[5], double p, double* shift, long* calls) {
    double model[5];
    or_agents_hs_fit(fv, p, model, shift);
    return or_bfgs_powell_fn(hs_model_fg, model, 1, shift, 10, calls, 0);
}

double or_agents(const or_problem* pr, unsigned int seed, const or_agents_opts* o, double* x, long* iters,
                 long* evals) {
    const int P = pr->P, A = o->agent_num;
    const int five = pr->variant == OR_COST_HS; /* AGENTS.cpp:335-337 */
    if (pr->variant != OR_COST_FROBENIUS && !five) return DBL_MAX;
    or_mt19937 gen;
    or_mt_seed(&gen, seed);
    long n_eval = 0;
    double* best = (double*)malloc(sizeof(double) * (size_t)(P > 0 ? P : 1));
    double* ag = (double*)malloc(sizeof(double) * (size_t)A * (P > 0 ? P : 1));
    double* fmin_ag = (double*)malloc(sizeof(double) * (size_t)A);
    double* f_pi2 = (double*)malloc(sizeof(double) * (size_t)A);
    double* f_pi = (double*)malloc(sizeof(double) * (size_t)A);
    double* f_pi4 = (double*)malloc(sizeof(double) * (size_t)A);
    double* f_3pi2 = (double*)malloc(sizeof(double) * (size_t)A);
    double* saved = (double*)malloc(sizeof(double) * (size_t)A);
    int* pidx = (int*)malloc(sizeof(int) * (size_t)A);
    double* rnd = (double*)malloc(sizeof(double) * 2 * (size_t)A);
    double* win = (double*)calloc((size_t)o->convergence_length, sizeof(double));
    memcpy(best, x, sizeof(double) * P);
    /* AGENTS.cpp:98 */
    double current_minimum = or_optimization_problem(pr, best);
    n_eval++;
    /* AGENTS.cpp:280-296: initial parameter index of every agent */
    for (int a = 0; a < A; a++) pidx[a] = or_mt_uniform_int(&gen, P);
    /* AGENTS.cpp:300-340: agent 0 starts at the guess, the others randomised around it */
    for (int a = 0; a < A; a++) {
        if (a == 0) memcpy(ag, x, sizeof(double) * P);
        else or_randomize_parameters(&gen, best, ag + (size_t)a * P, P, current_minimum, o->radius,
                                     o->randomization_rate);
    }
    or_optimization_problem_batched(pr, ag, A, 1, fmin_ag);
    n_eval += A;
    double mean = 0.0, var = DBL_MAX;
    int widx = 0, most = 0, terminate = 0;
    long it;
    for (it = 0; it < o->max_iterations; it++) {
        /* AGENTS.cpp:372-390 */
        for (int a = 0; a < A; a++) {
            pidx[a] = or_mt_uniform_int(&gen, P);
            saved[a] = ag[(size_t)a * P + pidx[a]];
        }
        if (five) {
            /* five-point line search (AGENTS.cpp:544-675): the parameter at p + pi/4,
             * + pi/4, + pi/2, + pi/2, accumulated in place */
            double* fs[4] = {f_pi4, f_pi2, f_pi, f_3pi2};
            const double steps[4] = {M_PI / 4, M_PI / 4, M_PI / 2, M_PI / 2};
            for (int k = 0; k < 4; k++) {
                for (int a = 0; a < A; a++) ag[(size_t)a * P + pidx[a]] += steps[k];
                or_optimization_problem_batched(pr, ag, A, 1, fs[k]);
            }
            n_eval += 4 * A;
            for (int a = 0; a < A; a++) {
                const double fv[5] = {fmin_ag[a], f_pi4[a], f_pi2[a], f_pi[a], f_3pi2[a]};
                double shift;
                const double f = or_agents_hs_search(fv, saved[a], &shift, 0);
                ag[(size_t)a * P + pidx[a]] = saved[a] + shift;
                fmin_ag[a] = f;
            }
        } else {
            /* three-point line search on the sinusoid f(p) = offset + A cos(p - phi)
             * (AGENTS.cpp:395-443) */
            for (int a = 0; a < A; a++) ag[(size_t)a * P + pidx[a]] += M_PI / 2;
            or_optimization_problem_batched(pr, ag, A, 1, f_pi2);
            for (int a = 0; a < A; a++) ag[(size_t)a * P + pidx[a]] += M_PI / 2;
            or_optimization_problem_batched(pr, ag, A, 1, f_pi);
            n_eval += 2 * A;
            for (int a = 0; a < A; a++) {
                const double f0 = fmin_ag[a];
                const double a_cos = (f0 - f_pi[a]) / 2;
                const double offset = (f0 + f_pi[a]) / 2;
                const double a_sin = offset - f_pi2[a];
                const double phi0 = atan2(a_sin, a_cos);
                const double shift = phi0 > 0 ? M_PI - phi0 : -phi0 - M_PI;
                const double amp = sqrt(a_sin * a_sin + a_cos * a_cos);
                ag[(size_t)a * P + pidx[a]] = saved[a] + shift;
                fmin_ag[a] = offset - amp;
            }
        }
        /* AGENTS.cpp:696-703 */
        for (int i = 0; i < 2 * A; i++) rnd[i] = or_mt_canonical(&gen);
        const int period = (it % o->lifetime) == 0;
        if (period) {  /* AGENTS.cpp:737-742 */
            or_optimization_problem_batched(pr, ag, A, 1, fmin_ag);
            n_eval += A;
        }
        for (int a = 0; a < A; a++) {  /* AGENTS.cpp:747-850 */
            double* xa = ag + (size_t)a * P;
            if (fmin_ag[a] < o->tolerance) {
                terminate = 1;
                current_minimum = fmin_ag[a];
                most = a;
                memcpy(best, xa, sizeof(double) * P);
            }
            if (!period) continue;
            if (fmin_ag[a] <= current_minimum) {
                most = a;
                memcpy(best, xa, sizeof(double) * P);
                current_minimum = fmin_ag[a];
            } else if (rnd[2 * a] < o->exploration_rate && a != most) {
                fmin_ag[a] = fmin_ag[most];
                memcpy(xa, ag + (size_t)most * P, sizeof(double) * P);
                if (rnd[2 * a + 1] < o->agent_randomization_rate) {
                    or_randomize_parameters(&gen, best, xa, P, current_minimum, o->radius, o->randomization_rate);
                    fmin_ag[a] = or_optimization_problem(pr, xa);
                    n_eval++;
                }
            }
            if (a == 0) {  /* convergence window (AGENTS.cpp:818-836) */
                const int L = o->convergence_length;
                mean = mean + (current_minimum - win[widx]) / L;
                win[widx] = current_minimum;
                widx = (widx + 1) % L;
                var = 0.0;
                for (int i = 0; i < L; i++) var += (win[i] - mean) * (win[i] - mean);
                var = sqrt(var) / L;
                if (fabs(mean - current_minimum) < 1e-7 && var < 1e-7) terminate = 1;
            }
        }
        if (terminate) {
            it++;
            break;
        }
    }
    memcpy(x, best, sizeof(double) * P);
    if (iters) *iters = it;
    if (evals) *evals = n_eval;
    free(best); free(ag); free(fmin_ag); free(f_pi2); free(f_pi); free(f_pi4); free(f_3pi2); free(saved); free(pidx); free(rnd); free(win);
    return current_minimum;
}
