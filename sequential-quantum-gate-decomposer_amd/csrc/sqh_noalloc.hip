// sqh_noalloc.hip -- batched cost / gradient evaluation that allocates nothing.
//
// For callers that cannot allocate device memory while they run (code under stream
// capture, tight serving loops, harnesses that forbid hipMalloc): the evaluation's
// scratch -- the 2x2 kernel tables, the trace channels, the per-wavefront gradient
// correlations and the correction seeds -- lives in one file-scope device array of
// kScratchBytes per GPU, held until the process ends.  Host code never passes the
// array's name to a kernel: its device address is looked up once per device with
// hipGetSymbolAddress when a context is created (sqh_create) and used from then on.
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU (two
// contexts, two streams) race on it and must be serialised by the caller.
#include <mutex>

#include "sqh_context.h"

namespace sqh {
namespace {

constexpr size_t kScratchBytes = 64ull << 20;  // 64 MiB per GPU
constexpr int kMaxDevices = 64;
constexpr size_t kAlign = 256;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
hipError_t g_err[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};

size_t round_up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

}  // namespace

// The device address of this GPU's scratch array; looked up once per device with the
// device current.
hipError_t noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_done[device]) {
        void* p = nullptr;
        g_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_ptr[device] = g_err[device] == hipSuccess ? p : nullptr;
        g_done[device] = true;
    }
    return g_err[device];
}

static void* scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_done[device] ? g_ptr[device] : nullptr;
}

size_t noalloc_scratch_bytes() { return kScratchBytes; }

}  // namespace sqh

using namespace sqh;

extern "C" int sqh_optimization_problem_combined_batch_noalloc_dev(sqh_ctx* c, const double* d_params, int B,
                                                                   int precision, double* d_f0, double* d_grad) {
    if (!c) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = require_target(c);
    if (rc) return rc;
    if (!d_params || !d_f0) return fail(c, SQH_ERR_ARG, "null buffer");
    if (B <= 0) return fail(c, SQH_ERR_ARG, "batch size must be positive");
    if (c->prog.N > kMaxTileQubits)
        return fail(c, SQH_ERR_ARG, "evaluation without allocation supports up to %d qubits", kMaxTileQubits);
    char* base = static_cast<char*>(scratch_ptr(c->device));
    if (!base) return fail(c, SQH_ERR_DEVICE, "the scratch array of device %d has no address", c->device);
    size_t parts[4];
    eval_scratch_bytes(c, B, precision == SQH_PREC_F32, d_grad != nullptr, parts);
    size_t total = 0;
    for (size_t b : parts) total += round_up(b);
    if (total > kScratchBytes)
        return fail(c, SQH_ERR_ARG,
                    "batch of %d needs %zu bytes of scratch, above the %zu-byte limit of the allocation-free "
                    "entry point",
                    B, total, kScratchBytes);
    HIPTRY(c, hipSetDevice(c->device));
    EvalScratch s;
    size_t off = 0;
    s.kern = base + off;
    off += round_up(parts[0]);
    s.diag = reinterpret_cast<double*>(base + off);
    off += round_up(parts[1]);
    s.mpart = parts[2] ? reinterpret_cast<double*>(base + off) : nullptr;
    off += round_up(parts[2]);
    s.seed = parts[3] ? reinterpret_cast<double*>(base + off) : nullptr;
    return eval_batch(c, d_params, B, precision, d_f0, d_grad, &s);
}
