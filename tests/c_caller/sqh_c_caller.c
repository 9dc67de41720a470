/* sqh_c_caller.c -- a plain C caller of libsqhip.so: no Python, no PyTorch.
 *
 * The library's C ABI as a C or C++ host program sees it, with the HIP runtime the
 * library was linked against (/opt/rocm's libamdhip64).  Test program:
 *
 *   sqh_c_caller --version          prints sqh_version() (no device needed)
 *   sqh_c_caller IN OUT             one batched cost + gradient evaluation
 *
 * IN (native-endian): int32 qbit_num, n_gates, B, then n_gates x {type, target,
 * control, param_start} int32, then the 2^n x 2^n complex128 target (row-major),
 * then B x P float64 parameters.  OUT: B costs, then B x P gradient entries
 * (float64).  stdout names the libamdhip64 mapped into the process. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sqhip.h"

static int die(const char* what, sqh_ctx* c) {
    fprintf(stderr, "sqh_c_caller: %s%s%s\n", what, c ? ": " : "", c ? sqh_last_error(c) : "");
    return 1;
}

static void print_hip_runtime(void) {
    FILE* m = fopen("/proc/self/maps", "r");
    char line[4096];
    if (!m) return;
    while (fgets(line, sizeof line, m)) {
        char* p = strstr(line, "/");
        if (p && strstr(p, "libamdhip64")) {
            printf("hip runtime: %s", p);
            break;
        }
    }
    fclose(m);
}

int main(int argc, char** argv) {
    if (argc == 2 && strcmp(argv[1], "--version") == 0) {
        printf("%s\n", sqh_version());
        return 0;
    }
    if (argc != 3) {
        fprintf(stderr, "usage: %s --version | %s IN OUT\n", argv[0], argv[0]);
        return 2;
    }
    FILE* in = fopen(argv[1], "rb");
    if (!in) return die("cannot open input", NULL);
    int32_t hdr[3];
    if (fread(hdr, sizeof(int32_t), 3, in) != 3) return die("short header", NULL);
    const int n = hdr[0], ng = hdr[1], B = hdr[2];
    if (n < 1 || n > 16 || ng < 0 || B < 1) return die("bad header", NULL);
    const size_t d = (size_t)1 << n;
    sqh_gate* gates = (sqh_gate*)malloc(sizeof(sqh_gate) * (size_t)(ng > 0 ? ng : 1));
    double* U = (double*)malloc(sizeof(double) * 2 * d * d);
    if (fread(gates, sizeof(sqh_gate), (size_t)ng, in) != (size_t)ng) return die("short gate list", NULL);
    if (fread(U, sizeof(double) * 2, d * d, in) != d * d) return die("short target", NULL);

    sqh_ctx* c = NULL;
    if (sqh_create(0, &c) != SQH_OK) return die("sqh_create failed", c);
    if (sqh_set_circuit(c, n, gates, ng) != SQH_OK) return die("sqh_set_circuit", c);
    if (sqh_set_target(c, U, (int)d, (int)d, (int)d) != SQH_OK) return die("sqh_set_target", c);
    const int P = sqh_get_parameter_num(c);
    double* x = (double*)malloc(sizeof(double) * (size_t)B * (P > 0 ? P : 1));
    double* f = (double*)malloc(sizeof(double) * (size_t)B);
    double* g = (double*)malloc(sizeof(double) * (size_t)B * (P > 0 ? P : 1));
    if (fread(x, sizeof(double), (size_t)B * P, in) != (size_t)B * P) return die("short parameters", c);
    fclose(in);
    if (sqh_optimization_problem_combined_batch(c, x, B, SQH_PREC_F64, f, g) != SQH_OK)
        return die("sqh_optimization_problem_combined_batch", c);
    FILE* out = fopen(argv[2], "wb");
    if (!out) return die("cannot open output", c);
    fwrite(f, sizeof(double), (size_t)B, out);
    fwrite(g, sizeof(double), (size_t)B * P, out);
    fclose(out);
    print_hip_runtime();
    printf("parameters %d, batch %d, cost[0] %.17g\n", P, B, f[0]);
    sqh_destroy(c);
    free(gates); free(U); free(x); free(f); free(g);
    return 0;
}
