// sqh_mlp.cpp -- the warm-start network (no counterpart in the reference, whose
// nn module only generates data: squander/src-cpp/nn/include/NN.h:58-146).
//
// A multilayer perceptron in -> hidden (GELU) x n_hidden -> out, BF16 weights and
// activations with FP32 accumulation and FP32 master weights, trained with Adam on
// mean-squared error.  Every dimension is padded to the GEMM tile (64); padded
// weights, biases and gradients stay exactly zero.
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "sqh_context.h"

namespace sqh {
hipError_t mlp_gemm_forward(bool hidden, const void* H, int K, const void* W, const float* bias, int M, int N,
                            float* Z, void* Hout, hipStream_t s);
hipError_t mlp_gemm_backward_data(const void* dZ, int N, const void* W, int K, int M, const float* Zprev,
                                  float* dZprev_f, void* dZprev_b, hipStream_t s);
hipError_t mlp_gemm_backward_weights(const void* dZ, int N, const void* H, int K, int B, float* dW, hipStream_t s);
hipError_t mlp_to_bf16_padded(const float* x, int B, int in, void* out, int Bp, int in_pad, hipStream_t s);
hipError_t mlp_mse(const float* y, int ldy, const float* t, int B, int out, int Bp, float* dyf, void* dyb,
                   double* part, int* nblocks, hipStream_t s);
hipError_t mlp_bias_grad(const float* dz, int ld, int B, int N, float* db, hipStream_t s);
hipError_t mlp_out_grad(const float* dy, int B, int out, int ldy, int Bp, float* dyf, void* dyb, hipStream_t s);
hipError_t mlp_adam(float* w, void* wb, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2,
                    float eps, int step, hipStream_t s);
hipError_t mlp_f32_to_bf16(const float* a, void* b, size_t n, hipStream_t s);
}  // namespace sqh

using namespace sqh;

namespace {
inline int pad64(int v) { return (v + 63) / 64 * 64; }
}  // namespace

struct sqh_mlp {
    sqh_ctx* ctx = nullptr;
    int L = 0;                       // weight layers
    std::vector<int> dim, dpad;      // real and padded widths, L + 1 entries
    std::vector<DevBuf> W, Wb, b, mW, vW, mb, vb, gW, gb;
    std::vector<DevBuf> H, Z, dZf, dZb;  // activations per layer (H[0] = input)
    DevBuf part;
    int cap = 0;                     // padded batch the activation buffers hold
    int step = 0;

    ~sqh_mlp() {
        for (auto* v : {&W, &Wb, &b, &mW, &vW, &mb, &vb, &gW, &gb, &H, &Z, &dZf, &dZb})
            for (auto& x : *v) x.release();
        part.release();
    }
    hipStream_t st() const { return ctx->stream; }
};

namespace {

int ensure_batch(sqh_mlp* m, int Bp) {
    sqh_ctx* c = m->ctx;
    if (Bp <= m->cap) return SQH_OK;
    for (int l = 0; l <= m->L; ++l) HIPTRY(c, m->H[l].ensure((size_t)Bp * m->dpad[l] * 2));
    for (int l = 0; l < m->L; ++l) {
        HIPTRY(c, m->Z[l].ensure((size_t)Bp * m->dpad[l + 1] * 4));
        HIPTRY(c, m->dZf[l].ensure((size_t)Bp * m->dpad[l + 1] * 4));
        HIPTRY(c, m->dZb[l].ensure((size_t)Bp * m->dpad[l + 1] * 2));
    }
    HIPTRY(c, m->part.ensure((size_t)Bp * m->dpad[m->L] / 256 * 8 + 64));
    m->cap = Bp;
    return SQH_OK;
}

int forward(sqh_mlp* m, const float* d_x, int B, int Bp) {
    sqh_ctx* c = m->ctx;
    HIPTRY(c, mlp_to_bf16_padded(d_x, B, m->dim[0], m->H[0].p, Bp, m->dpad[0], m->st()));
    for (int l = 0; l < m->L; ++l) {
        const bool hidden = l + 1 < m->L;
        HIPTRY(c, mlp_gemm_forward(hidden, m->H[l].p, m->dpad[l], m->Wb[l].p, m->b[l].as<float>(), Bp, m->dpad[l + 1],
                                   m->Z[l].as<float>(), hidden ? m->H[l + 1].p : nullptr, m->st()));
    }
    return SQH_OK;
}

// Gradients into gW/gb from the output gradient already in dZf/dZb[L-1].
int backward_layers(sqh_mlp* m, int B, int Bp) {
    sqh_ctx* c = m->ctx;
    const int L = m->L;
    for (int l = L - 1; l >= 0; --l) {
        HIPTRY(c, mlp_gemm_backward_weights(m->dZb[l].p, m->dpad[l + 1], m->H[l].p, m->dpad[l], Bp,
                                            m->gW[l].as<float>(), m->st()));
        HIPTRY(c, mlp_bias_grad(m->dZf[l].as<float>(), m->dpad[l + 1], B, m->dpad[l + 1], m->gb[l].as<float>(),
                                m->st()));
        if (l > 0)
            HIPTRY(c, mlp_gemm_backward_data(m->dZb[l].p, m->dpad[l + 1], m->Wb[l].p, m->dpad[l], Bp,
                                             m->Z[l - 1].as<float>(), m->dZf[l - 1].as<float>(), m->dZb[l - 1].p,
                                             m->st()));
    }
    return SQH_OK;
}

int adam_step(sqh_mlp* m, float lr) {
    sqh_ctx* c = m->ctx;
    m->step++;
    for (int l = 0; l < m->L; ++l) {
        const size_t nw = (size_t)m->dpad[l + 1] * m->dpad[l];
        HIPTRY(c, mlp_adam(m->W[l].as<float>(), m->Wb[l].p, m->gW[l].as<float>(), m->mW[l].as<float>(),
                           m->vW[l].as<float>(), nw, lr, 0.9f, 0.999f, 1e-8f, m->step, m->st()));
        HIPTRY(c, mlp_adam(m->b[l].as<float>(), nullptr, m->gb[l].as<float>(), m->mb[l].as<float>(),
                           m->vb[l].as<float>(), (size_t)m->dpad[l + 1], lr, 0.9f, 0.999f, 1e-8f, m->step, m->st()));
    }
    return SQH_OK;
}

// Loss, gradients into gW/gb; the loss sum goes to *loss (synchronises).
int backward(sqh_mlp* m, const float* d_t, int B, int Bp, double* loss) {
    sqh_ctx* c = m->ctx;
    const int L = m->L;
    int nb = 0;
    HIPTRY(c, mlp_mse(m->Z[L - 1].as<float>(), m->dpad[L], d_t, B, m->dim[L], Bp, m->dZf[L - 1].as<float>(),
                      m->dZb[L - 1].p, m->part.as<double>(), &nb, m->st()));
    int rc = backward_layers(m, B, Bp);
    if (rc) return rc;
    if (loss) {
        std::vector<double> part(nb);
        HIPTRY(c, hipMemcpyAsync(part.data(), m->part.p, nb * sizeof(double), hipMemcpyDeviceToHost, m->st()));
        HIPTRY(c, hipStreamSynchronize(m->st()));
        double s = 0.0;
        for (double v : part) s += v;
        *loss = s / ((double)B * m->dim[L]);
    }
    return SQH_OK;
}

// Host flat layout of parameters: per layer W (real out x real in, row-major), then b.
template <typename F>
int for_each_layer(sqh_mlp* m, F&& f) {
    long long off = 0;
    for (int l = 0; l < m->L; ++l) {
        int rc = f(l, off);
        if (rc) return rc;
        off += (long long)m->dim[l + 1] * m->dim[l] + m->dim[l + 1];
    }
    return SQH_OK;
}

int copy_params(sqh_mlp* m, float* host, bool to_host, bool grads) {
    sqh_ctx* c = m->ctx;
    return for_each_layer(m, [&](int l, long long off) -> int {
        const int no = m->dim[l + 1], ni = m->dim[l], kp = m->dpad[l];
        DevBuf& Wd = grads ? m->gW[l] : m->W[l];
        DevBuf& bd = grads ? m->gb[l] : m->b[l];
        if (to_host) {
            HIPTRY(c, hipMemcpy2D(host + off, (size_t)ni * 4, Wd.p, (size_t)kp * 4, (size_t)ni * 4, no,
                                  hipMemcpyDeviceToHost));
            HIPTRY(c, hipMemcpy(host + off + (long long)no * ni, bd.p, (size_t)no * 4, hipMemcpyDeviceToHost));
        } else {
            HIPTRY(c, hipMemcpy2D(Wd.p, (size_t)kp * 4, host + off, (size_t)ni * 4, (size_t)ni * 4, no,
                                  hipMemcpyHostToDevice));
            HIPTRY(c, hipMemcpy(bd.p, host + off + (long long)no * ni, (size_t)no * 4, hipMemcpyHostToDevice));
            HIPTRY(c, mlp_f32_to_bf16(m->W[l].as<float>(), m->Wb[l].p, (size_t)m->dpad[l + 1] * kp, m->st()));
            HIPTRY(c, hipStreamSynchronize(m->st()));
        }
        return SQH_OK;
    });
}

}  // namespace

extern "C" {

int sqh_mlp_create(sqh_ctx* c, int in_dim, int hidden, int n_hidden, int out_dim, unsigned long long seed,
                   sqh_mlp** out) {
    if (!c || !out) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    *out = nullptr;
    if (in_dim <= 0 || hidden <= 0 || n_hidden < 0 || out_dim <= 0) return fail(c, SQH_ERR_ARG, "bad network shape");
    HIPTRY(c, hipSetDevice(c->device));
    sqh_mlp* m = new (std::nothrow) sqh_mlp();
    if (!m) return SQH_ERR_NOMEM;
    m->ctx = c;
    m->dim.push_back(in_dim);
    for (int i = 0; i < n_hidden; ++i) m->dim.push_back(hidden);
    m->dim.push_back(out_dim);
    m->L = (int)m->dim.size() - 1;
    for (int d : m->dim) m->dpad.push_back(pad64(d));
    for (auto* v : {&m->W, &m->Wb, &m->b, &m->mW, &m->vW, &m->mb, &m->vb, &m->gW, &m->gb, &m->Z, &m->dZf, &m->dZb})
        v->resize(m->L);
    m->H.resize(m->L + 1);
    std::mt19937_64 gen(seed);
    std::normal_distribution<float> nd(0.0f, 1.0f);
    for (int l = 0; l < m->L; ++l) {
        const int np = m->dpad[l + 1], kp = m->dpad[l];
        const size_t nw = (size_t)np * kp;
        bool ok = m->W[l].ensure(nw * 4) == hipSuccess && m->Wb[l].ensure(nw * 2) == hipSuccess &&
                  m->mW[l].ensure(nw * 4) == hipSuccess && m->vW[l].ensure(nw * 4) == hipSuccess &&
                  m->gW[l].ensure(nw * 4) == hipSuccess && m->b[l].ensure(np * 4) == hipSuccess &&
                  m->mb[l].ensure(np * 4) == hipSuccess && m->vb[l].ensure(np * 4) == hipSuccess &&
                  m->gb[l].ensure(np * 4) == hipSuccess;
        if (!ok) {
            delete m;
            return fail(c, SQH_ERR_NOMEM, "network allocation failed");
        }
        // He-style initialisation on the real block, zeros in the padding.
        std::vector<float> w(nw, 0.0f);
        const float sd = std::sqrt(2.0f / (float)m->dim[l]);
        for (int o = 0; o < m->dim[l + 1]; ++o)
            for (int i = 0; i < m->dim[l]; ++i) w[(size_t)o * kp + i] = sd * nd(gen);
        hipError_t e = hipMemcpy(m->W[l].p, w.data(), nw * 4, hipMemcpyHostToDevice);
        for (DevBuf* z : {&m->mW[l], &m->vW[l], &m->gW[l]})
            if (e == hipSuccess) e = hipMemset(z->p, 0, nw * 4);
        for (DevBuf* z : {&m->b[l], &m->mb[l], &m->vb[l], &m->gb[l]})
            if (e == hipSuccess) e = hipMemset(z->p, 0, (size_t)np * 4);
        if (e == hipSuccess) e = mlp_f32_to_bf16(m->W[l].as<float>(), m->Wb[l].p, nw, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            delete m;
            return fail(c, SQH_ERR_DEVICE, "network initialisation: %s", hipGetErrorString(e));
        }
    }
    *out = m;
    return SQH_OK;
}

void sqh_mlp_destroy(sqh_mlp* m) {
    if (!m) return;
    (void)hipStreamSynchronize(m->ctx->stream);
    delete m;
}

long long sqh_mlp_param_count(const sqh_mlp* m) {
    if (!m) return SQH_ERR_ARG;
    long long n = 0;
    for (int l = 0; l < m->L; ++l) n += (long long)m->dim[l + 1] * m->dim[l] + m->dim[l + 1];
    return n;
}

int sqh_mlp_get_params(sqh_mlp* m, float* out) {
    if (!m || !out) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(m->ctx->mu);
    HIPTRY(m->ctx, hipStreamSynchronize(m->st()));
    return copy_params(m, out, true, false);
}

int sqh_mlp_set_params(sqh_mlp* m, const float* in) {
    if (!m || !in) return SQH_ERR_ARG;
    std::lock_guard<std::mutex> lk(m->ctx->mu);
    HIPTRY(m->ctx, hipStreamSynchronize(m->st()));
    return copy_params(m, const_cast<float*>(in), false, false);
}

int sqh_mlp_forward_dev(sqh_mlp* m, const float* d_x, int B, float* d_y) {
    if (!m || !d_x || !d_y || B <= 0) return SQH_ERR_ARG;
    sqh_ctx* c = m->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    HIPTRY(c, hipSetDevice(c->device));
    const int Bp = pad64(B);
    int rc = ensure_batch(m, Bp);
    if (rc) return rc;
    if ((rc = forward(m, d_x, B, Bp))) return rc;
    HIPTRY(c, hipMemcpy2DAsync(d_y, (size_t)m->dim[m->L] * 4, m->Z[m->L - 1].p, (size_t)m->dpad[m->L] * 4,
                               (size_t)m->dim[m->L] * 4, B, hipMemcpyDeviceToDevice, m->st()));
    return SQH_OK;
}

int sqh_mlp_train_step_dev(sqh_mlp* m, const float* d_x, const float* d_t, int B, float lr, double* loss) {
    if (!m || !d_x || !d_t || B <= 0) return SQH_ERR_ARG;
    sqh_ctx* c = m->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    HIPTRY(c, hipSetDevice(c->device));
    const int Bp = pad64(B);
    int rc = ensure_batch(m, Bp);
    if (rc) return rc;
    if ((rc = forward(m, d_x, B, Bp))) return rc;
    if ((rc = backward(m, d_t, B, Bp, loss))) return rc;
    return adam_step(m, lr);
}

int sqh_mlp_backward_step_dev(sqh_mlp* m, const float* d_x, const float* d_dy, int B, float lr) {
    if (!m || !d_x || !d_dy || B <= 0) return SQH_ERR_ARG;
    sqh_ctx* c = m->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    HIPTRY(c, hipSetDevice(c->device));
    const int Bp = pad64(B);
    int rc = ensure_batch(m, Bp);
    if (rc) return rc;
    if ((rc = forward(m, d_x, B, Bp))) return rc;
    const int L = m->L;
    HIPTRY(c, mlp_out_grad(d_dy, B, m->dim[L], m->dpad[L], Bp, m->dZf[L - 1].as<float>(), m->dZb[L - 1].p,
                           m->st()));
    if ((rc = backward_layers(m, B, Bp))) return rc;
    return adam_step(m, lr);
}

int sqh_mlp_gradients(sqh_mlp* m, const float* x, const float* t, int B, float* grads, double* loss) {
    if (!m || !x || !t || !grads || B <= 0) return SQH_ERR_ARG;
    sqh_ctx* c = m->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    HIPTRY(c, hipSetDevice(c->device));
    const int Bp = pad64(B);
    int rc = ensure_batch(m, Bp);
    if (rc) return rc;
    DevBuf dx, dt;
    HIPTRY(c, dx.ensure((size_t)B * m->dim[0] * 4));
    HIPTRY(c, dt.ensure((size_t)B * m->dim[m->L] * 4));
    HIPTRY(c, hipMemcpyAsync(dx.p, x, (size_t)B * m->dim[0] * 4, hipMemcpyHostToDevice, m->st()));
    HIPTRY(c, hipMemcpyAsync(dt.p, t, (size_t)B * m->dim[m->L] * 4, hipMemcpyHostToDevice, m->st()));
    rc = forward(m, dx.as<float>(), B, Bp);
    if (!rc) rc = backward(m, dt.as<float>(), B, Bp, loss);
    if (!rc) {
        (void)hipStreamSynchronize(m->st());
        rc = copy_params(m, grads, true, true);
    }
    (void)hipStreamSynchronize(m->st());
    dx.release();
    dt.release();
    return rc;
}

}  // extern "C"
