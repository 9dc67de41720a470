// sqh_tune.cpp -- run-time tuner of the 4-qubit matrix-core block kernel's launch shape.
//
// Candidates: 1, 2 or 4 sixteen-column tiles per wavefront x 64 or 256 threads per
// workgroup.  The first fused-block application of a (precision, rows, columns, batch)
// shape times every candidate on a scratch batch of that shape (zeros; at most 1 GiB,
// fewer matrices when the real batch is larger) with HIP events, on blocks over the
// four lowest and the four highest qubits (the two ends of the row-stride range), and
// keeps the fastest for the context.  Every candidate computes the same arithmetic per
// element, so the choice changes timing only, never results.  Under stream capture, or
// with tuning off, the fixed per-precision default is used.
#include <chrono>
#include <cstring>

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

}  // namespace sqh

using namespace sqh;

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
