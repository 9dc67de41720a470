// sqh_agents.cpp -- the AGENTS gradient-free population optimiser
// (Optimization_Interface::solve_layer_optimization_problem_AGENTS,
// AGENTS.cpp:42-870) on the batched cost of the context's circuit.
//
// Every iteration each agent moves one randomly chosen parameter to the minimum of
// the cost along it: for the FROBENIUS cost f(p) = offset + A cos(p - phi) exactly
// (all parameters enter the 2x2 kernels through sin/cos of the raw value), so the
// cost at p, p + pi/2 and p + pi fixes the sinusoid (AGENTS.cpp:395-443).  The
// agents' bookkeeping (adoption of the best agent's state, re-randomisation,
// convergence window) stays on the host with the reference's generator:
// std::mt19937 with the libstdc++ distributions, consumed in the reference's order.
//
// The host logic of a run is a state machine cut at every cost evaluation, so any
// number of runs (sqh_agents_batch: one per target) advance in lockstep and each
// stage's evaluations of all live runs go to the GPU in one launch.  A run's
// results do not depend on the others: the generator draws and the order of its
// host steps are those of a run on its own, and each cost is computed per element.
#include <cfloat>
#include <cmath>
#include <cstring>
#include <random>
#include <utility>
#include <vector>

#include "sqh_context.h"

namespace sqh {
namespace {

// Optimization_Interface::randomize_parameters (Optimization_Interface.cpp:576-602).
void randomize(std::mt19937& gen, const double* in, double* out, int P, double f0, double radius, double rate) {
    std::uniform_real_distribution<> prob(0.0, 1.0);
    std::uniform_real_distribution<> real(-2 * M_PI, 2 * M_PI);
    for (int j = 0; j < P; ++j) out[j] = prob(gen) <= rate ? in[j] + real(gen) * std::sqrt(f0) * radius : in[j];
}

struct AgentsRun {
    const sqh_agents_options* o = nullptr;
    int P = 0, A = 0, target = 0;
    std::mt19937 gen;
    std::uniform_int_distribution<> pick;
    std::uniform_real_distribution<> choose{0.0, 1.0};
    std::vector<double> x0, best, ag, fa, f_pi2, f_pi, saved, rnd, win;
    std::vector<int> pidx;
    double current_minimum = 0.0, mean = 0.0;
    int widx = 0, most = 0, stage = 0;
    bool terminate = false, done = false, period = false;
    long long it = 0, evals = 0;
    // evaluation requests of the current stage: P parameters -> one cost
    std::vector<std::pair<const double*, double*>> req;

    AgentsRun(const sqh_agents_options* opts, int P_, const double* x, unsigned int seed, int tgt)
        : o(opts), P(P_), A(opts->agent_num), target(tgt), gen(seed), pick(0, P_ - 1) {
        x0.assign(x, x + P);
        best = x0;
        ag.assign((size_t)A * P, 0.0);
        fa.assign(A, 0.0);
        f_pi2.assign(A, 0.0);
        f_pi.assign(A, 0.0);
        saved.assign(A, 0.0);
        rnd.assign(2 * (size_t)A, 0.0);
        win.assign(o->convergence_length, 0.0);
        pidx.assign(A, 0);
    }

    // Runs host steps until the run requests evaluations or finishes.
    void advance() {
        req.clear();
        while (!done) {
            switch (stage) {
            case 0:  // the guess' cost
                req.emplace_back(best.data(), &current_minimum);
                stage = 1;
                return;
            case 1:  // the initial agents (agent 0 = the guess)
                for (int a = 0; a < A; ++a) pidx[a] = pick(gen);
                for (int a = 0; a < A; ++a) {
                    if (a == 0) std::memcpy(ag.data(), x0.data(), (size_t)P * sizeof(double));
                    else randomize(gen, best.data(), &ag[(size_t)a * P], P, current_minimum, o->radius,
                                   o->randomization_rate);
                }
                for (int a = 0; a < A; ++a) req.emplace_back(&ag[(size_t)a * P], &fa[a]);
                stage = 2;
                return;
            case 2:  // one parameter per agent, shifted by pi/2
                if (it >= o->max_iterations) {
                    done = true;
                    return;
                }
                for (int a = 0; a < A; ++a) {
                    pidx[a] = pick(gen);
                    saved[a] = ag[(size_t)a * P + pidx[a]];
                }
                for (int a = 0; a < A; ++a) {
                    ag[(size_t)a * P + pidx[a]] += M_PI / 2;
                    req.emplace_back(&ag[(size_t)a * P], &f_pi2[a]);
                }
                stage = 3;
                return;
            case 3:  // ... and by pi
                for (int a = 0; a < A; ++a) {
                    ag[(size_t)a * P + pidx[a]] += M_PI / 2;
                    req.emplace_back(&ag[(size_t)a * P], &f_pi[a]);
                }
                stage = 4;
                return;
            case 4: {  // the sinusoid's minimum (AGENTS.cpp:425-443)
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
                period = it % o->agent_lifetime == 0;
                stage = 5;
                if (period) {
                    for (int a = 0; a < A; ++a) req.emplace_back(&ag[(size_t)a * P], &fa[a]);
                    return;
                }
                break;
            }
            case 5:  // adoption and re-randomisation (AGENTS.cpp:760-836)
                for (int a = 0; a < A; ++a) {
                    double* xa = &ag[(size_t)a * P];
                    if (fa[a] < o->optimization_tolerance) {
                        terminate = true;
                        current_minimum = fa[a];
                        most = a;
                        std::memcpy(best.data(), xa, (size_t)P * sizeof(double));
                    }
                    if (period) {
                        if (fa[a] <= current_minimum) {
                            most = a;
                            std::memcpy(best.data(), xa, (size_t)P * sizeof(double));
                            current_minimum = fa[a];
                        } else if (rnd[2 * a] < o->exploration_rate && a != most) {
                            fa[a] = fa[most];
                            std::memcpy(xa, &ag[(size_t)most * P], (size_t)P * sizeof(double));
                            if (rnd[2 * a + 1] < o->agent_randomization_rate) {
                                randomize(gen, best.data(), xa, P, current_minimum, o->radius,
                                          o->randomization_rate);
                                // its cost is read only from the next iteration on, so the
                                // evaluation waits for the end of the loop
                                req.emplace_back(xa, &fa[a]);
                            }
                        }
                    }
                    if (period && a == 0) {  // convergence window (AGENTS.cpp:818-836)
                        const int L = o->convergence_length;
                        mean = mean + (current_minimum - win[widx]) / L;
                        win[widx] = current_minimum;
                        widx = (widx + 1) % L;
                        double var = 0.0;
                        for (int i = 0; i < L; ++i) var += (win[i] - mean) * (win[i] - mean);
                        var = std::sqrt(var) / L;
                        if (std::fabs(mean - current_minimum) < 1e-7 && var < 1e-7) terminate = true;
                    }
                }
                stage = 6;
                if (!req.empty()) return;
                break;
            case 6:
                ++it;
                if (terminate) {
                    done = true;
                    return;
                }
                stage = 2;
                break;
            default:
                done = true;
                return;
            }
        }
    }
};

// Device buffers of one optimiser call, released on every exit path.
struct AgentsBufs {
    DevBuf x, f, idx, targets;
    ~AgentsBufs() {
        x.release();
        f.release();
        idx.release();
        targets.release();
    }
};

// Advances the runs in lockstep; each round's requests of all live runs are one
// launch: against the context's target (d_targets null) or, per run, its target.
int run_agents(sqh_ctx* c, std::vector<AgentsRun>& runs, AgentsBufs& d, const void* d_targets, int rows,
               int cols) {
    const int P = c->prog.P;
    std::vector<double> hx, hf;
    std::vector<int> hidx;
    std::vector<double*> dst;
    for (;;) {
        hx.clear();
        hidx.clear();
        dst.clear();
        for (AgentsRun& r : runs) {
            if (r.done) continue;
            r.advance();
            for (auto& q : r.req) {
                hx.insert(hx.end(), q.first, q.first + P);
                hidx.push_back(r.target);
                dst.push_back(q.second);
            }
            r.evals += (long long)r.req.size();
        }
        if (dst.empty()) return SQH_OK;
        const int E = (int)dst.size();
        hf.resize(E);
        HIPTRY(c, d.x.ensure((size_t)E * P * sizeof(double)));
        HIPTRY(c, d.f.ensure((size_t)E * sizeof(double)));
        HIPTRY(c, hipMemcpyAsync(d.x.p, hx.data(), (size_t)E * P * sizeof(double), hipMemcpyHostToDevice, c->stream));
        int rc;
        if (d_targets) {
            HIPTRY(c, d.idx.ensure((size_t)E * sizeof(int)));
            HIPTRY(c, hipMemcpyAsync(d.idx.p, hidx.data(), (size_t)E * sizeof(int), hipMemcpyHostToDevice, c->stream));
            rc = eval_targets(c, d.x.as<double>(), E, d_targets, rows, cols, d.idx.as<int>(), d.f.as<double>());
        } else {
            rc = eval_batch(c, d.x.as<double>(), E, SQH_PREC_F64, d.f.as<double>(), nullptr);
        }
        if (rc) return rc;
        HIPTRY(c, hipMemcpyAsync(hf.data(), d.f.p, (size_t)E * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPTRY(c, hipStreamSynchronize(c->stream));
        for (int e = 0; e < E; ++e) *dst[e] = hf[e];
    }
}

int check_options(sqh_ctx* c, const sqh_agents_options* o) {
    if (o->agent_num <= 0 || o->max_iterations < 0 || o->agent_lifetime <= 0 || o->convergence_length <= 0)
        return fail(c, SQH_ERR_ARG, "agent_num, agent_lifetime and convergence_length must be positive");
    if (c->variant != CV_FROBENIUS)
        return fail(c, SQH_ERR_ARG, "AGENTS runs the three-point sinusoid search of the FROBENIUS cost");
    return SQH_OK;
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
    if ((rc = check_options(c, o))) return rc;
    HIPTRY(c, hipSetDevice(c->device));
    AgentsBufs d;
    std::vector<AgentsRun> runs;
    runs.emplace_back(o, P, x, seed, 0);
    if ((rc = run_agents(c, runs, d, nullptr, 0, 0))) return rc;
    std::memcpy(x, runs[0].best.data(), (size_t)P * sizeof(double));
    *f_out = runs[0].current_minimum;
    if (iters_out) *iters_out = runs[0].it;
    if (evals_out) *evals_out = runs[0].evals;
    return SQH_OK;
}

extern "C" int sqh_agents_batch(sqh_ctx* c, const double* targets, int n_runs, int rows, int cols, double* x,
                                const sqh_agents_options* o, const unsigned int* seeds, double* f_out,
                                long long* iters_out, long long* evals_out) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = require_circuit(c);
    if (rc) return rc;
    const int P = c->prog.P;
    if (!targets || !x || !o || !seeds || !f_out) return fail(c, SQH_ERR_ARG, "null buffer");
    if (n_runs <= 0) return fail(c, SQH_ERR_ARG, "need at least one run");
    if (P <= 0) return fail(c, SQH_ERR_ARG, "the circuit has no parameters");
    if (c->prog.N > kMaxTileQubits)
        return fail(c, SQH_ERR_ARG, "batched AGENTS supports up to %d qubits", kMaxTileQubits);
    if ((rc = check_options(c, o))) return rc;
    const int off = uses_offset(c->variant) ? c->trace_offset : 0;
    if (rows != (1 << c->prog.N) || cols <= 0 || cols + off > rows)
        return fail(c, SQH_ERR_ARG, "targets must be %d x m with m + trace offset <= %d", 1 << c->prog.N,
                    1 << c->prog.N);
    HIPTRY(c, hipSetDevice(c->device));
    AgentsBufs d;
    const size_t tbytes = (size_t)n_runs * rows * cols * 16;
    HIPTRY(c, d.targets.ensure(tbytes));
    HIPTRY(c, hipMemcpyAsync(d.targets.p, targets, tbytes, hipMemcpyHostToDevice, c->stream));
    std::vector<AgentsRun> runs;
    runs.reserve(n_runs);
    for (int k = 0; k < n_runs; ++k) runs.emplace_back(o, P, x + (size_t)k * P, seeds[k], k);
    if ((rc = run_agents(c, runs, d, d.targets.p, rows, cols))) return rc;
    for (int k = 0; k < n_runs; ++k) {
        std::memcpy(x + (size_t)k * P, runs[k].best.data(), (size_t)P * sizeof(double));
        f_out[k] = runs[k].current_minimum;
        if (iters_out) iters_out[k] = runs[k].it;
        if (evals_out) evals_out[k] = runs[k].evals;
    }
    return SQH_OK;
}
