// sqh_tune.cpp -- run-time tuners of launch shapes: the 4-qubit block kernel of the
// streaming path, the tile-resident fused-block cost kernel, and the launch options of
// the register-tile forward and gradient kernels.
//
// Candidates: 1, 2 or 4 sixteen-column tiles per wavefront x 64 or 256 threads per
// workgroup.  The first fused-block application of a (precision, rows, columns, batch)
// shape times every candidate on a scratch batch of that shape (zeros; at most 1 GiB,
// fewer matrices when the real batch is larger) with HIP events, on blocks over the
// four lowest and the four highest qubits (the two ends of the row-stride range), and
// keeps the fastest for the context.  Every candidate computes the same arithmetic per
// element, so the choice changes timing only, never results.  Under stream capture, or
// with tuning off, the fixed per-precision default is used.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "sqh_context.h"

namespace sqh {
namespace {

const BlockCfg kCands[] = {{1, 256}, {2, 256}, {4, 256}, {1, 64}, {2, 64}, {4, 64}};
// Untuned defaults (the fastest shapes at 8 qubits x 256 columns x 512 matrices):
// complex128 one tile per wavefront, complex64 four (a 16-column tile of complex64 is
// only 128 bytes per row).
const BlockCfg kDefault64{1, 256}, kDefault32{4, 256};
constexpr size_t kTuneBytes = size_t(1) << 30;

struct TuneBufs {
    DevBuf mats, F;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~TuneBufs() {
        mats.release();
        F.release();
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

}  // namespace

int block_config(sqh_ctx* c, bool f32, int rows, int m, int B, BlockCfg* out) {
    if (block_cfg_valid(c->block_force)) {
        *out = c->block_force;
        return SQH_OK;
    }
    const std::array<long long, 4> key{f32 ? 1 : 0, rows, m, B};
    auto it = c->block_tuned.find(key);
    if (it != c->block_tuned.end()) {
        *out = it->second;
        return SQH_OK;
    }
    *out = f32 ? kDefault32 : kDefault64;
    if (!c->autotune || rows < 16 || m <= 0 || B <= 0) return SQH_OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPTRY(c, hipStreamIsCapturing(c->stream, &cs));
    if (cs != hipStreamCaptureStatusNone) return SQH_OK;

    const auto t_start = std::chrono::steady_clock::now();
    const size_t esz = f32 ? 8 : 16;
    const size_t mbytes = (size_t)rows * m * esz;
    long long Bt = B;
    if ((size_t)Bt * mbytes > kTuneBytes) Bt = std::max<long long>(1, (long long)(kTuneBytes / mbytes));
    TuneBufs t;
    if (t.mats.ensure((size_t)Bt * mbytes) != hipSuccess || t.F.ensure(256 * esz) != hipSuccess) {
        (void)hipGetLastError();  // no room for the scratch batch: keep the default
        return SQH_OK;
    }
    HIPTRY(c, hipMemsetAsync(t.mats.p, 0, (size_t)Bt * mbytes, c->stream));
    HIPTRY(c, hipMemsetAsync(t.F.p, 0, 256 * esz, c->stream));
    HIPTRY(c, hipEventCreate(&t.e0));
    HIPTRY(c, hipEventCreate(&t.e1));
    int n = 0;
    while ((1 << n) < rows) ++n;
    const int low[4] = {0, 1, 2, 3};
    const int high[4] = {n - 4, n - 3, n - 2, n - 1};
    float best_ms = 0.0f;
    for (const BlockCfg& cand : kCands) {
        auto run = [&]() -> hipError_t {
            for (const int* q : {low, high}) {
                hipError_t e = launch_block_apply(f32, true, t.mats.p, (long long)rows * m, m, rows, m, (int)Bt,
                                                  t.F.p, 0, q, 4, cand, c->stream);
                if (e != hipSuccess) return e;
            }
            return hipSuccess;
        };
        HIPTRY(c, run());  // warm-up
        HIPTRY(c, hipEventRecord(t.e0, c->stream));
        for (int r = 0; r < 3; ++r) HIPTRY(c, run());
        HIPTRY(c, hipEventRecord(t.e1, c->stream));
        HIPTRY(c, hipEventSynchronize(t.e1));
        float ms = 0.0f;
        HIPTRY(c, hipEventElapsedTime(&ms, t.e0, t.e1));
        if (&cand == kCands || ms < best_ms) {
            best_ms = ms;
            *out = cand;
        }
    }
    c->block_tuned[key] = *out;
    c->tune_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return SQH_OK;
}

// Fused-block kernels: candidates of tiles per wavefront x exchange groups x waves per
// workgroup x prefetch.  The first call of a (pass, precision, qubits, columns, batch)
// shape times every valid candidate on that call's own inputs (one warm-up and three
// timed launches each, HIP events) and keeps the fastest; the chosen one then runs for
// the call.  The cost pass runs the transposed chain (k_fused_trace, one exchange group)
// unless it is switched off or the circuit has a single block; its candidates all write
// the same trace entries (the shapes are bit-identical).  The store pass (apply_to) runs
// in place, so its candidates read the call's input and write a scratch copy of at most
// kFcTuneBytes instead (the first elements of the batch).  Under stream capture, or with
// tuning off, the default shape runs.
namespace {
constexpr size_t kFcTuneBytes = size_t(256) << 20;
const FusedCostCfg kFcCands[] = {{4, 1, 2, 0}, {4, 1, 2, 1}, {4, 2, 2, 1}, {4, 1, 1, 0}, {4, 2, 1, 1},
                                 {2, 1, 2, 0}, {2, 1, 4, 0}, {2, 1, 4, 1}};
const FusedCostCfg kFtCands[] = {{4, 1, 2, 0}, {4, 1, 2, 1}, {4, 1, 1, 0}, {4, 1, 4, 0},
                                 {2, 1, 2, 0}, {2, 1, 4, 0}, {2, 1, 4, 1}, {2, 1, 2, 1}};

FusedCostCfg fused_default(int N, bool f32, int nb, bool chain) {
    // complex128: all four tiles in one exchange (eight independent MFMA chains);
    // complex64: two exchange groups where the columns allow (the chain: one)
    const FusedCostCfg pref = f32 && !chain ? FusedCostCfg{4, 2, 2, 1} : FusedCostCfg{4, 1, 2, 0};
    auto valid = [&](FusedCostCfg cfg) {
        return chain ? fused_trace_cfg_valid(N, f32, nb, cfg) : fused_cost_cfg_valid(N, f32, nb, cfg);
    };
    if (valid(pref)) return pref;
    for (const FusedCostCfg& cfg : chain ? kFtCands : kFcCands)
        if (valid(cfg)) return cfg;
    return pref;
}

// whether a fused pass runs the transposed chain: the cost pass (no output), with the
// chain on and at least two blocks
bool fused_uses_chain(const sqh_ctx* c, bool store) {
    return !store && c->fused_chain && c->fplan.qubits.size() >= 2;
}
}  // namespace

int fused_cost_launch(sqh_ctx* c, bool f32, const FusedCostArgs& a_in) {
    const int N = c->prog.N, nb = a_in.nb;
    const bool chain = fused_uses_chain(c, a_in.out != nullptr);
    FusedCostArgs a = a_in;
    if (chain) a.maps = c->d_tmaps.as<int>();
    auto valid = [&](FusedCostCfg cfg) {
        return chain ? fused_trace_cfg_valid(N, f32, nb, cfg) : fused_cost_cfg_valid(N, f32, nb, cfg);
    };
    auto launch = [&](const FusedCostArgs& x, FusedCostCfg cfg) {
        return chain ? launch_fused_trace(f32, N, x, cfg, c->stream) : launch_fused_cost(f32, N, x, cfg, c->stream);
    };
    if (c->fused_force.tiles != 0) {
        if (!valid(c->fused_force))
            return fail(c, SQH_ERR_ARG,
                        "fused-block launch shape %d/%d/%d/%d does not fit %d qubits (a column must lie in one "
                        "exchange group; the transposed chain takes one group) or the LDS a workgroup may allocate",
                        c->fused_force.tiles, c->fused_force.groups, c->fused_force.waves, c->fused_force.prefetch, N);
        HIPTRY(c, launch(a, c->fused_force));
        return SQH_OK;
    }
    // the store pass (apply_to) moves the matrices through HBM, the chain is another
    // kernel: their own keys
    const std::array<long long, 4> key{(f32 ? 1 : 0) + (a.out ? 2 : 0) + (chain ? 4 : 0), N, a.m, a.B};
    auto it = c->fused_tuned.find(key);
    if (it != c->fused_tuned.end() && valid(it->second)) {
        HIPTRY(c, launch(a, it->second));
        return SQH_OK;
    }
    FusedCostCfg best = fused_default(N, f32, nb, chain);
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPTRY(c, hipStreamIsCapturing(c->stream, &cs));
    // the arguments the candidates run with: the call's own, or for the store pass a
    // scratch output (rows x m per element, row stride m)
    FusedCostArgs ta = a;
    struct Scratch {
        DevBuf b;
        hipStream_t s = nullptr;
        ~Scratch() {
            if (b.p) {
                (void)hipStreamSynchronize(s);  // released only after the candidates have run
                b.release();
            }
        }
    } scratch_guard;
    scratch_guard.s = c->stream;
    DevBuf& scratch = scratch_guard.b;
    bool can_tune = c->autotune && cs == hipStreamCaptureStatusNone;
    if (can_tune && a.out) {
        const size_t per = ((size_t)1 << N) * a.m * (f32 ? 8 : 16);
        ta.B = (int)std::max<size_t>(1, std::min<size_t>((size_t)a.B, kFcTuneBytes / per));
        if (scratch.ensure((size_t)ta.B * per) != hipSuccess) {
            (void)hipGetLastError();  // no room for the scratch copy: keep the default
            can_tune = false;
        }
        ta.out = scratch.p;
        ta.out_stride = a.m;
        ta.out_bstride = ((long long)1 << N) * a.m;
    }
    if (can_tune) {
        const auto t_start = std::chrono::steady_clock::now();
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIPTRY(c, hipEventCreate(&e0));
        hipError_t err = hipEventCreate(&e1);
        float best_ms = -1.0f;
        for (const FusedCostCfg& cand : chain ? kFtCands : kFcCands) {
            if (err != hipSuccess) break;
            if (!valid(cand)) continue;
            if ((err = launch(ta, cand)) != hipSuccess) break;  // warm-up
            if ((err = hipEventRecord(e0, c->stream)) != hipSuccess) break;
            for (int r = 0; r < 3 && err == hipSuccess; ++r) err = launch(ta, cand);
            if (err != hipSuccess || (err = hipEventRecord(e1, c->stream)) != hipSuccess ||
                (err = hipEventSynchronize(e1)) != hipSuccess)
                break;
            float ms = 0.0f;
            if ((err = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) break;
            if (best_ms < 0.0f || ms < best_ms) {
                best_ms = ms;
                best = cand;
            }
        }
        (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        HIPTRY(c, err);
        c->fused_tuned[key] = best;
        c->tune_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    }
    HIPTRY(c, launch(a, best));
    return SQH_OK;
}

// Register-tile kernels: launch options that leave every column's arithmetic unchanged,
// so every option gives bit-identical results.  Forward pass (trace entries) with one
// batch element per wavefront: 4, 2 or 1 wavefronts of an element per workgroup
// (sharing the element's kernel table in LDS), or the packed layout (several elements
// per wavefront, the table read through the caches).  Gradient pass with one element
// per wavefront: the program (ops, layouts, kernel table) staged in LDS, or read
// through the scalar cache.  The first call of a (pass, precision, qubits, padded
// columns, batch) shape times each distinct option on its own inputs (both passes
// write only trace entries and correlation partials, so repeating them changes
// nothing) and keeps the fastest.  The store pass (apply_to) keeps the rule.
TileLaunch tile_with(const sqh_ctx* c, const TileLaunch& t, int kind, const TileCfg& cfg) {
    TileLaunch u = t;
    if (kind == 2) {
        u.stage = cfg.stage;
        if (cfg.stage == 1 && u.stage_ops == 0) {
            u.stage_ops = u.nops;
            u.stage_lays = (int)c->prog.layouts.size();
        }
    } else {
        u.wpg = cfg.wpg;
    }
    return u;
}

int tile_launch(sqh_ctx* c, const TileLaunch& t, int kind, bool tune) {
    auto launch = [&](const TileLaunch& u) -> hipError_t {
        return kind == 2 ? launch_tile_grad(u, c->stream) : launch_tile_forward(u, c->stream);
    };
    if (c->tile_forced) {
        HIPTRY(c, launch(tile_with(c, t, kind, c->tile_force)));
        return SQH_OK;
    }
    const std::array<long long, 4> key{kind * 2 + (t.f32 ? 1 : 0), t.N, t.m_pad, t.B};
    auto it = c->tile_tuned.find(key);
    if (it != c->tile_tuned.end()) {
        HIPTRY(c, launch(tile_with(c, t, kind, it->second)));
        return SQH_OK;
    }
    // the distinct options of this launch, the rule's first
    std::vector<TileCfg> cands{TileCfg{}};
    if (kind == 2) {
        if (!t.bi.run_prog && tile_grad_can_stage(tile_with(c, t, kind, TileCfg{0, 1}))) cands.push_back(TileCfg{0, 1});
        cands[0].stage = t.stage_ops > 0 ? 1 : 0;
    } else if (t.mode == 1) {
        std::vector<int> groups{tile_forward_group(t, 0)};
        for (int w : {4, 2, 1, -1}) {
            const int g = tile_forward_group(t, w);
            if (std::find(groups.begin(), groups.end(), g) == groups.end()) {
                groups.push_back(g);
                cands.push_back(TileCfg{w, -1});
            }
        }
    }
    TileCfg best = cands[0];
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPTRY(c, hipStreamIsCapturing(c->stream, &cs));
    if (tune && c->autotune && cs == hipStreamCaptureStatusNone && cands.size() > 1) {
        const auto t_start = std::chrono::steady_clock::now();
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIPTRY(c, hipEventCreate(&e0));
        hipError_t err = hipEventCreate(&e1);
        float best_ms = -1.0f;
        for (const TileCfg& cand : cands) {
            if (err != hipSuccess) break;
            const TileLaunch u = tile_with(c, t, kind, cand);
            if ((err = launch(u)) != hipSuccess) break;  // warm-up
            if ((err = hipEventRecord(e0, c->stream)) != hipSuccess) break;
            for (int r = 0; r < 3 && err == hipSuccess; ++r) err = launch(u);
            if (err != hipSuccess || (err = hipEventRecord(e1, c->stream)) != hipSuccess ||
                (err = hipEventSynchronize(e1)) != hipSuccess)
                break;
            float ms = 0.0f;
            if ((err = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) break;
            if (best_ms < 0.0f || ms < best_ms) {
                best_ms = ms;
                best = cand;
            }
        }
        (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        HIPTRY(c, err);
        c->tile_tuned[key] = best;
        c->tune_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    }
    HIPTRY(c, launch(tile_with(c, t, kind, best)));
    return SQH_OK;
}

}  // namespace sqh

using namespace sqh;

extern "C" int sqh_set_tile_config(sqh_ctx* c, int waves_per_group, int stage_program) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    if (waves_per_group != 0 && waves_per_group != 1 && waves_per_group != 2 && waves_per_group != 4 &&
        waves_per_group != -1)
        return fail(c, SQH_ERR_ARG, "waves per group must be 1, 2, 4, -1 (packed) or 0 (tuner)");
    if (stage_program != -1 && stage_program != 0 && stage_program != 1)
        return fail(c, SQH_ERR_ARG, "stage_program must be 0, 1 or -1 (tuner)");
    c->tile_forced = !(waves_per_group == 0 && stage_program == -1);
    c->tile_force = TileCfg{waves_per_group, stage_program};
    return SQH_OK;
}

extern "C" int sqh_set_fused_cost_config(sqh_ctx* c, int tiles_per_wave, int exchange_groups, int waves,
                                         int prefetch) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    const FusedCostCfg cfg{tiles_per_wave, exchange_groups, waves, prefetch};
    if (!(tiles_per_wave == 0 && exchange_groups == 0 && waves == 0 && prefetch == 0) &&
        !((tiles_per_wave == 2 || tiles_per_wave == 4) && (exchange_groups == 1 || exchange_groups == 2) &&
          (waves == 1 || waves == 2 || waves == 4) && (prefetch == 0 || prefetch == 1) &&
          (exchange_groups == 1 || (tiles_per_wave == 4 && prefetch == 1))))
        return fail(c, SQH_ERR_ARG,
                    "tiles per wavefront must be 2 or 4, waves 1, 2 or 4, prefetch 0 or 1, exchange groups 1 "
                    "(or 2 with 4 tiles and prefetch); all 0 for the tuner");
    c->fused_force = cfg;
    return SQH_OK;
}

extern "C" int sqh_set_fused_cost_chain(sqh_ctx* c, int on) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    if (on != 0 && on != 1) return fail(c, SQH_ERR_ARG, "the chain switch is 0 or 1");
    c->fused_chain = on != 0;
    return SQH_OK;
}

extern "C" int sqh_set_fused_grad_config(sqh_ctx* c, int tiles_per_wave, int waves) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    if (!(tiles_per_wave == 0 && waves == 0) &&
        !((tiles_per_wave == 2 || tiles_per_wave == 4) && (waves == 1 || waves == 2 || waves == 4 || waves == 8)))
        return fail(c, SQH_ERR_ARG, "tiles per wavefront must be 2 or 4 and waves 1, 2, 4 or 8; both 0 for automatic");
    c->fg_force_tiles = tiles_per_wave;
    c->fg_force_waves = waves;
    return SQH_OK;
}

extern "C" int sqh_get_fused_cost_config(sqh_ctx* c, int is_f32, int B, int* tiles_per_wave, int* exchange_groups,
                                         int* waves, int* prefetch) {
    if (!c || !tiles_per_wave || !exchange_groups || !waves || !prefetch) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->have_circuit || !c->have_target) return fail(c, SQH_ERR_STATE, "set a circuit and a target first");
    const int N = c->prog.N, nb = (int)c->fplan.qubits.size();
    const bool chain = fused_uses_chain(c, false);
    FusedCostCfg cfg = fused_default(N, is_f32 != 0, nb, chain);
    if (chain ? fused_trace_cfg_valid(N, is_f32 != 0, nb, c->fused_force)
              : fused_cost_cfg_valid(N, is_f32 != 0, nb, c->fused_force)) {
        cfg = c->fused_force;
    } else {
        auto it = c->fused_tuned.find({(is_f32 ? 1 : 0) + (chain ? 4 : 0), N, c->cols, B});
        if (it != c->fused_tuned.end()) cfg = it->second;
    }
    *tiles_per_wave = cfg.tiles;
    *exchange_groups = cfg.groups;
    *waves = cfg.waves;
    *prefetch = cfg.prefetch;
    return SQH_OK;
}

extern "C" int sqh_get_tile_config(sqh_ctx* c, int kind, int is_f32, int B, int* waves_per_group,
                                   int* stage_program) {
    if (!c || !waves_per_group || !stage_program || (kind != 1 && kind != 2) || B <= 0) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->have_circuit || !c->have_target) return fail(c, SQH_ERR_STATE, "set a circuit and a target first");
    TileLaunch t = base_launch(c, is_f32 != 0);
    t.mode = 1;
    t.m = c->cols;
    t.m_pad = m_pad_for(t.N, t.m);
    t.B = B;
    TileCfg cfg;
    if (c->tile_forced) {
        cfg = c->tile_force;
    } else {
        auto it = c->tile_tuned.find({kind * 2 + (is_f32 ? 1 : 0), t.N, t.m_pad, B});
        if (it != c->tile_tuned.end()) cfg = it->second;
    }
    const TileLaunch u = tile_with(c, t, kind, cfg);
    *waves_per_group = kind == 1 ? tile_forward_group(u, u.wpg) : 1;
    *stage_program = kind == 2 ? (u.stage != 0 && tile_grad_can_stage(u) ? 1 : 0) : 0;
    return SQH_OK;
}

extern "C" int sqh_set_autotune(sqh_ctx* c, int enable) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    c->autotune = enable != 0;
    return SQH_OK;
}

extern "C" int sqh_set_block_config(sqh_ctx* c, int tiles_per_wave, int threads) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    const BlockCfg cfg{tiles_per_wave, threads};
    if (!(tiles_per_wave == 0 && threads == 0) && !block_cfg_valid(cfg))
        return fail(c, SQH_ERR_ARG, "tiles per wavefront must be 1, 2 or 4 and threads 64 or 256 (or both 0)");
    c->block_force = cfg;
    return SQH_OK;
}

extern "C" int sqh_get_block_config(sqh_ctx* c, int is_f32, int rows, int cols, int B, int* tiles_per_wave,
                                    int* threads, double* tune_ms) {
    if (!c || !tiles_per_wave || !threads) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    if (rows < 16 || (rows & (rows - 1)) || cols <= 0 || B <= 0) return fail(c, SQH_ERR_ARG, "bad matrix shape");
    HIPTRY(c, hipSetDevice(c->device));
    BlockCfg cfg;
    int rc = block_config(c, is_f32 != 0, rows, cols, B, &cfg);
    if (rc) return rc;
    *tiles_per_wave = cfg.tiles;
    *threads = cfg.threads;
    if (tune_ms) *tune_ms = c->tune_ms;
    return SQH_OK;
}
