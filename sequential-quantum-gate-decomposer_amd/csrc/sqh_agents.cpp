// sqh_agents.cpp -- the AGENTS gradient-free population optimiser
// (Optimization_Interface::solve_layer_optimization_problem_AGENTS,
// AGENTS.cpp:42-870) on the batched cost of the context's circuit and target.
//
// Every iteration each agent moves one randomly chosen parameter to the minimum of
// the cost along it: for the FROBENIUS cost f(p) = offset + A cos(p - phi) exactly
// (all parameters enter the 2x2 kernels through sin/cos of the raw value), so the
// cost at p, p + pi/2 and p + pi fixes the sinusoid (AGENTS.cpp:395-443).  The two
// shifted populations are evaluated as batched launches on the GPU; the agents'
// bookkeeping (adoption of the best agent's state, re-randomisation, convergence
// window) stays on the host with the reference's generator: std::mt19937 with the
// libstdc++ distributions, consumed in the reference's order.
#include <cfloat>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "sqh_context.h"

namespace sqh {
namespace {

struct AgentsRun {
    sqh_ctx* c;
    int P;
    DevBuf d_x, d_f;
    long long evals = 0;
    ~AgentsRun() {
        d_x.release();
        d_f.release();
    }
    // FP64 costs of B parameter vectors (host in, host out).
    int eval(const double* x, int B, double* f) {
        HIPTRY(c, d_x.ensure((size_t)B * P * sizeof(double)));
        HIPTRY(c, d_f.ensure((size_t)B * sizeof(double)));
        HIPTRY(c, hipMemcpyAsync(d_x.p, x, (size_t)B * P * sizeof(double), hipMemcpyHostToDevice, c->stream));
        int rc = eval_batch(c, d_x.as<double>(), B, SQH_PREC_F64, d_f.as<double>(), nullptr);
        if (rc) return rc;
        HIPTRY(c, hipMemcpyAsync(f, d_f.p, (size_t)B * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPTRY(c, hipStreamSynchronize(c->stream));
        evals += B;
        return SQH_OK;
    }
};

// Optimization_Interface::randomize_parameters (Optimization_Interface.cpp:576-602).
void randomize(std::mt19937& gen, const double* in, double* out, int P, double f0, double radius, double rate) {
    std::uniform_real_distribution<> prob(0.0, 1.0);
    std::uniform_real_distribution<> real(-2 * M_PI, 2 * M_PI);
    for (int j = 0; j < P; ++j) out[j] = prob(gen) <= rate ? in[j] + real(gen) * std::sqrt(f0) * radius : in[j];
}

}  // namespace
}  // namespace sqh

using namespace sqh;

extern "C" int sqh_agents(sqh_ctx* c, double* x, const sqh_agents_options* o, unsigned int seed, double* f_out,
                          long long* iters_out, long long* evals_out) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = require_target(c);
    if (rc) return rc;
    const int P = c->prog.P;
    if (!x || !o || !f_out) return fail(c, SQH_ERR_ARG, "null buffer");
    if (P <= 0) return fail(c, SQH_ERR_ARG, "the circuit has no parameters");
    if (o->agent_num <= 0 || o->max_iterations < 0 || o->agent_lifetime <= 0 || o->convergence_length <= 0)
        return fail(c, SQH_ERR_ARG, "agent_num, agent_lifetime and convergence_length must be positive");
    if (c->variant != CV_FROBENIUS)
        return fail(c, SQH_ERR_ARG, "AGENTS runs the three-point sinusoid search of the FROBENIUS cost");
    HIPTRY(c, hipSetDevice(c->device));
    const int A = o->agent_num;
    AgentsRun run;
    run.c = c;
    run.P = P;
    std::mt19937 gen(seed);
    std::uniform_int_distribution<> pick(0, P - 1);
    std::uniform_real_distribution<> choose(0.0, 1.0);

    std::vector<double> best(x, x + P), ag((size_t)A * P), fa(A), f_pi2(A), f_pi(A), saved(A), rnd(2 * (size_t)A);
    std::vector<int> pidx(A);
    std::vector<double> win(o->convergence_length, 0.0);
    double current_minimum = 0.0;
    if ((rc = run.eval(best.data(), 1, &current_minimum))) return rc;
    for (int a = 0; a < A; ++a) pidx[a] = pick(gen);
    for (int a = 0; a < A; ++a) {
        if (a == 0) std::memcpy(ag.data(), x, (size_t)P * sizeof(double));
        else randomize(gen, best.data(), &ag[(size_t)a * P], P, current_minimum, o->radius, o->randomization_rate);
    }
    if ((rc = run.eval(ag.data(), A, fa.data()))) return rc;
    double mean = 0.0;
    int widx = 0, most = 0;
    bool terminate = false;
    long long it = 0;
    const int L = o->convergence_length;
    for (; it < o->max_iterations; ++it) {
        for (int a = 0; a < A; ++a) {
            pidx[a] = pick(gen);
            saved[a] = ag[(size_t)a * P + pidx[a]];
        }
        for (int a = 0; a < A; ++a) ag[(size_t)a * P + pidx[a]] += M_PI / 2;
        if ((rc = run.eval(ag.data(), A, f_pi2.data()))) return rc;
        for (int a = 0; a < A; ++a) ag[(size_t)a * P + pidx[a]] += M_PI / 2;
        if ((rc = run.eval(ag.data(), A, f_pi.data()))) return rc;
        for (int a = 0; a < A; ++a) {
            const double a_cos = (fa[a] - f_pi[a]) / 2;
            const double offset = (fa[a] + f_pi[a]) / 2;
            const double a_sin = offset - f_pi2[a];
            const double phi0 = std::atan2(a_sin, a_cos);
            const double shift = phi0 > 0 ? M_PI - phi0 : -phi0 - M_PI;
            ag[(size_t)a * P + pidx[a]] = saved[a] + shift;
            fa[a] = offset - std::sqrt(a_sin * a_sin + a_cos * a_cos);
        }
        for (int i = 0; i < 2 * A; ++i) rnd[i] = choose(gen);
        const bool period = it % o->agent_lifetime == 0;
        if (period && (rc = run.eval(ag.data(), A, fa.data()))) return rc;
        for (int a = 0; a < A; ++a) {
            double* xa = &ag[(size_t)a * P];
            if (fa[a] < o->optimization_tolerance) {
                terminate = true;
                current_minimum = fa[a];
                most = a;
                std::memcpy(best.data(), xa, (size_t)P * sizeof(double));
            }
            if (!period) continue;
            if (fa[a] <= current_minimum) {
                most = a;
                std::memcpy(best.data(), xa, (size_t)P * sizeof(double));
                current_minimum = fa[a];
            } else if (rnd[2 * a] < o->exploration_rate && a != most) {
                fa[a] = fa[most];
                std::memcpy(xa, &ag[(size_t)most * P], (size_t)P * sizeof(double));
                if (rnd[2 * a + 1] < o->agent_randomization_rate) {
                    randomize(gen, best.data(), xa, P, current_minimum, o->radius, o->randomization_rate);
                    if ((rc = run.eval(xa, 1, &fa[a]))) return rc;
                }
            }
            if (a == 0) {  // convergence window (AGENTS.cpp:818-836)
                mean = mean + (current_minimum - win[widx]) / L;
                win[widx] = current_minimum;
                widx = (widx + 1) % L;
                double var = 0.0;
                for (int i = 0; i < L; ++i) var += (win[i] - mean) * (win[i] - mean);
                var = std::sqrt(var) / L;
                if (std::fabs(mean - current_minimum) < 1e-7 && var < 1e-7) terminate = true;
            }
        }
        if (terminate) {
            ++it;
            break;
        }
    }
    std::memcpy(x, best.data(), (size_t)P * sizeof(double));
    *f_out = current_minimum;
    if (iters_out) *iters_out = it;
    if (evals_out) *evals_out = run.evals;
    return SQH_OK;
}
