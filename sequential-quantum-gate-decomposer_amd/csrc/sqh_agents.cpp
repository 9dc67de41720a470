// sqh_agents.cpp -- the AGENTS gradient-free population optimiser
// (Optimization_Interface::solve_layer_optimization_problem_AGENTS,
// AGENTS.cpp:42-870) on the batched cost of the context's circuit.
//
// Every iteration each agent moves one randomly chosen parameter to the minimum of
// the cost along it.  All parameters enter the 2x2 kernels through sin/cos of the
// raw value, so the trace is a first-order trigonometric polynomial in each one:
// the FROBENIUS cost is f(p) = offset + A cos(p - phi) exactly, fixed by the cost at
// p, p + pi/2 and p + pi (AGENTS.cpp:395-443); the HILBERT_SCHMIDT_TEST cost (|Tr|^2)
// is kappa sin(2p + xi) + gamma sin(p + phi) + offset, fixed by five values
// (p, + pi/4, + pi/2, + pi, + 3pi/2) and minimised from the larger term's minimum by
// BFGS_Powell over the one shift (AGENTS.cpp:544-675).  The
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

// BFGS_Powell (BFGS_Powell.cpp:70-192 with Grad_Descend::line_search,
// grad_descend.cpp:127-272, and the step bound of grad_descend.cpp:396-425) for one
// variable: Powell's factor Z is a scalar and the update has no Givens sweep and no
// further columns.  fg(s, f, g) evaluates the function; returns the minimum found,
// the argument in s, after at most max_calls evaluations.
template <class FG>
double bfgs_powell_1d(FG fg, double& s, long max_calls) {
    constexpr double kPrecision = 1.42e-14;
    long calls = 0;
    auto eval = [&](double at, double& f, double& g) {
        ++calls;
        fg(at, f, g);
    };
    double f, g;
    eval(s, f, g);
    double z = 1.0, fprev = std::fabs(f + f + 1.0);
    for (;;) {
        const double d = -(z * (z * g));  // -Z Z^T g
        const double dg0 = d * g;
        const double max_step = (dg0 < 0.0 && std::fabs(d) >= 1e-5) ? std::fabs(2.0 * M_PI / d) : 0.0;
        if (g * g <= 1e-38 || dg0 >= 0.0 || f >= fprev) break;
        fprev = f;
        if (calls >= max_calls) break;
        // line search along d from s0 (Wolfe constants 0.1 and 0.7, at most 50 probes)
        const double s0 = s, g0 = g, f0 = f;
        double step = std::fmin(1.0, max_step);
        double f_low = f0, f_high = f0, dg_high = 0.0, step_low = 0.0, step_high = max_step;
        double step_best = 0.0, f_best = f0, g_best = g0, dg_abs_best = std::fabs(dg0);
        for (int probe = 0; probe < 50; ++probe) {
            s = s0 + step * d;
            eval(s, f, g);
            const double dg = d * g;
            if (f < f_best || std::fabs(dg) < dg_abs_best) {
                step_best = step;
                f_best = f;
                g_best = g;
                dg_abs_best = std::fabs(dg);
            }
            if (calls == max_calls) break;
            if (f < f_low + step * 0.1 * dg0) {
                if (dg >= dg0 * 0.7) break;
                step_low = step;
                f_low = f;
            } else {
                step_high = step;
                f_high = f;
                dg_high = dg;
            }
            if (probe == 0 || step_low > 0.0) {
                const double dg_fit = (f_high - f_low) / (step_high - step_low) * 2.0 - dg_high;
                step = step_low + std::fmax(0.1, -dg_fit * 0.5 / (dg_high - dg_fit)) * (step_high - step_low);
            } else {
                step *= 0.1;
                if (step < kPrecision) break;
            }
        }
        if (step != step_best) {
            step = step_best;
            f = f_best;
            g = g_best;
            s = s0 + step * d;
        }
        if (step == 0.0) break;
        // Powell's update of the scalar factor (skipped without enough curvature)
        const double delta = s - s0, gamma = g - g0;
        double dTg = gamma * delta;
        if (dTg < std::fabs(g0 * delta) * 0.1) continue;
        dTg = std::fmax(kPrecision, dTg);
        z = std::fabs(delta) > kPrecision ? delta / std::sqrt(dTg) : 0.0;
    }
    return f;
}

// The five-point fit of the HILBERT_SCHMIDT_TEST cost along one parameter of value p
// from fv = f(p), f(p + pi/4), f(p + pi/2), f(p + pi), f(p + 3pi/2), and the
// minimisation of the fitted model over the shift (AGENTS.cpp:626-667,
// HS_partial_optimization_problem_combined of Bayes_Opt.cpp:34-88).  Returns the
// model's minimum; shift receives the shift.
double hs_line_search(const double fv[5], double p, double& shift) {
    const double f1 = fv[0] - fv[3], f2 = fv[2] - fv[4];  // 2 gamma sin(p + phi), 2 gamma cos(p + phi)
    const double gamma = std::sqrt(f1 * f1 + f2 * f2) * 0.5;
    const double varphi = std::atan2(f1, f2) - p;
    const double offset = 0.25 * (fv[0] + fv[3] + fv[2] + fv[4]);
    const double f3 = 0.5 * (fv[0] + fv[3] - 2 * offset);                      // kappa sin(2p + xi)
    const double f4 = fv[1] - offset - gamma * std::sin(p + M_PI / 4 + varphi);  // kappa cos(2p + xi)
    const double kappa = std::sqrt(f3 * f3 + f4 * f4);
    const double xi = std::atan2(f3, f4) - 2 * p;
    const double a = xi + 2 * p, b = varphi + p;
    // the larger term's minimiser (the double-frequency start subtracts p / 2, as the
    // reference does), reduced by fmod 2pi
    shift = std::fmod(std::fabs(gamma) > std::fabs(kappa) ? 3 * M_PI / 2 - varphi - p : 3 * M_PI / 4 - xi / 2 - p / 2,
                      M_PI * 2);
    return bfgs_powell_1d(
        [&](double t, double& f, double& g) {
            f = kappa * std::sin(2 * t + a) + gamma * std::sin(t + b) + offset;
            g = 2 * kappa * std::cos(2 * t + a) + gamma * std::cos(t + b);
        },
        shift, 10);
}

struct AgentsRun {
    const sqh_agents_options* o = nullptr;
    int P = 0, A = 0, target = 0;
    std::mt19937 gen;
    std::uniform_int_distribution<> pick;
    std::uniform_real_distribution<> choose{0.0, 1.0};
    std::vector<double> x0, best, ag, fa, f_pi2, f_pi, f_pi4, f_3pi2, saved, rnd, win;
    std::vector<int> pidx;
    double current_minimum = 0.0, mean = 0.0;
    int widx = 0, most = 0, stage = 0;
    bool terminate = false, done = false, period = false, five = false;
    long long it = 0, evals = 0;
    // evaluation requests of the current stage: P parameters -> one cost
    std::vector<std::pair<const double*, double*>> req;

    AgentsRun(const sqh_agents_options* opts, int P_, const double* x, unsigned int seed, int tgt, bool hs)
        : o(opts), P(P_), A(opts->agent_num), target(tgt), gen(seed), pick(0, P_ - 1), five(hs) {
        x0.assign(x, x + P);
        best = x0;
        ag.assign((size_t)A * P, 0.0);
        fa.assign(A, 0.0);
        f_pi2.assign(A, 0.0);
        f_pi.assign(A, 0.0);
        f_pi4.assign(A, 0.0);
        f_3pi2.assign(A, 0.0);
        saved.assign(A, 0.0);
        rnd.assign(2 * (size_t)A, 0.0);
        win.assign(o->convergence_length, 0.0);
        pidx.assign(A, 0);
    }

    // Adds step to every agent's drawn parameter and requests the agents' costs into f.
    void shift_and_request(double step, std::vector<double>& f) {
        for (int a = 0; a < A; ++a) {
            ag[(size_t)a * P + pidx[a]] += step;
            req.emplace_back(&ag[(size_t)a * P], &f[a]);
        }
    }

    // After the line search: the agents' behaviour draws (AGENTS.cpp:696-703) and, every
    // agent_lifetime iterations, their costs re-evaluated (AGENTS.cpp:737-742).  True
    // when that evaluation was requested.
    bool end_line_search() {
        for (int i = 0; i < 2 * A; ++i) rnd[i] = choose(gen);
        period = it % o->agent_lifetime == 0;
        stage = 5;
        if (period) {
            for (int a = 0; a < A; ++a) req.emplace_back(&ag[(size_t)a * P], &fa[a]);
            return true;
        }
        return false;
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
            case 2:  // one parameter per agent, shifted by pi/2 (or pi/4 for the five points)
                if (it >= o->max_iterations) {
                    done = true;
                    return;
                }
                for (int a = 0; a < A; ++a) {
                    pidx[a] = pick(gen);
                    saved[a] = ag[(size_t)a * P + pidx[a]];
                }
                if (five) {
                    shift_and_request(M_PI / 4, f_pi4);
                    stage = 7;
                    return;
                }
                shift_and_request(M_PI / 2, f_pi2);
                stage = 3;
                return;
            case 3:  // ... and by pi
                shift_and_request(M_PI / 2, f_pi);
                stage = 4;
                return;
            case 4:  // the sinusoid's minimum (AGENTS.cpp:425-443)
                for (int a = 0; a < A; ++a) {
                    const double a_cos = (fa[a] - f_pi[a]) / 2;
                    const double offset = (fa[a] + f_pi[a]) / 2;
                    const double a_sin = offset - f_pi2[a];
                    const double phi0 = std::atan2(a_sin, a_cos);
                    const double shift = phi0 > 0 ? M_PI - phi0 : -phi0 - M_PI;
                    ag[(size_t)a * P + pidx[a]] = saved[a] + shift;
                    fa[a] = offset - std::sqrt(a_sin * a_sin + a_cos * a_cos);
                }
                if (end_line_search()) return;
                break;
            case 7:  // five points: p + pi/2, p + pi, p + 3pi/2 (the shifts accumulate in place)
                shift_and_request(M_PI / 4, f_pi2);
                stage = 8;
                return;
            case 8:
                shift_and_request(M_PI / 2, f_pi);
                stage = 9;
                return;
            case 9:
                shift_and_request(M_PI / 2, f_3pi2);
                stage = 10;
                return;
            case 10:  // the fitted model's minimum (AGENTS.cpp:611-675)
                for (int a = 0; a < A; ++a) {
                    const double fv[5] = {fa[a], f_pi4[a], f_pi2[a], f_pi[a], f_3pi2[a]};
                    double shift;
                    fa[a] = hs_line_search(fv, saved[a], shift);
                    ag[(size_t)a * P + pidx[a]] = saved[a] + shift;
                }
                if (end_line_search()) return;
                break;
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
    if (c->variant != CV_FROBENIUS && c->variant != CV_HS)
        return fail(c, SQH_ERR_ARG,
                    "AGENTS has line searches for the FROBENIUS (three points) and HILBERT_SCHMIDT_TEST (five "
                    "points) costs only");
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
    runs.emplace_back(o, P, x, seed, 0, c->variant == CV_HS);
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
    for (int k = 0; k < n_runs; ++k) runs.emplace_back(o, P, x + (size_t)k * P, seeds[k], k, c->variant == CV_HS);
    if ((rc = run_agents(c, runs, d, d.targets.p, rows, cols))) return rc;
    for (int k = 0; k < n_runs; ++k) {
        std::memcpy(x + (size_t)k * P, runs[k].best.data(), (size_t)P * sizeof(double));
        f_out[k] = runs[k].current_minimum;
        if (iters_out) iters_out[k] = runs[k].it;
        if (evals_out) evals_out[k] = runs[k].evals;
    }
    return SQH_OK;
}
