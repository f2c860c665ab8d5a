// C-ABI of the seed network (declared in include/picca_hip.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "context.h"

namespace {

picca::MlpView mlp_view(picca_ctx *ctx) {
    return picca::MlpView{ctx->mlp_in, ctx->mlp_hidden, ctx->m_w1t.ptr, ctx->m_w1t_bf16.ptr,
                          ctx->m_b1.ptr, ctx->m_w2.ptr, ctx->m_b2.ptr};
}

picca::MlpGrads mlp_grads(picca_ctx *ctx, float *loss) {
    return picca::MlpGrads{ctx->g_w1t.ptr, ctx->g_b1.ptr, ctx->g_w2.ptr, ctx->g_b2.ptr, loss};
}

int64_t mlp_params(picca_ctx *ctx) {
    return (int64_t)ctx->mlp_hidden * ctx->mlp_in + 3 * (int64_t)ctx->mlp_hidden + 2;
}

int zero_f32(float *p, int64_t n, hipStream_t st) {
    if (n <= 0) return PICCA_OK;
    PICCA_HIP_CHECK(hipMemsetAsync(p, 0, sizeof(float) * (size_t)n, st));
    return PICCA_OK;
}

int zero_grads(picca_ctx *ctx) {
    hipStream_t st = ctx->stream;
    int s = zero_f32(ctx->g_w1t.ptr, (int64_t)ctx->mlp_hidden * ctx->mlp_in, st);
    if (s == PICCA_OK) s = zero_f32(ctx->g_b1.ptr, ctx->mlp_hidden, st);
    if (s == PICCA_OK) s = zero_f32(ctx->g_w2.ptr, 2 * (int64_t)ctx->mlp_hidden, st);
    if (s == PICCA_OK) s = zero_f32(ctx->g_b2.ptr, 2, st);
    return s;
}

// rows of scratch (h, y, dh) for batches of n rows
int reserve_scratch(picca_ctx *ctx, int64_t n) {
    PICCA_HIP_CHECK(ctx->m_h.reserve((size_t)(n * ctx->mlp_hidden)));
    PICCA_HIP_CHECK(ctx->m_dh.reserve((size_t)(n * ctx->mlp_hidden)));
    PICCA_HIP_CHECK(ctx->m_y.reserve((size_t)(2 * n)));
    return PICCA_OK;
}

bool finite(const float *p, int64_t n) {
    for (int64_t k = 0; k < n; ++k)
        if (!std::isfinite(p[k])) return false;
    return true;
}

}  // namespace

extern "C" {

int picca_mlp_init(picca_ctx *ctx, int32_t n_in, int32_t n_hidden, const float *w1,
                   const float *b1, const float *w2, const float *b2) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    if (n_in <= 0 || n_in > picca::MLP_MAX_IN || n_in % 32 != 0)
        return fail(PICCA_E_VALUE, "n_in must be a positive multiple of 32, at most 256");
    if (n_hidden <= 0 || n_hidden > picca::MLP_MAX_HIDDEN || n_hidden % 16 != 0)
        return fail(PICCA_E_VALUE, "n_hidden must be a positive multiple of 16, at most 128");
    if (!w1 || !b1 || !w2 || !b2) return fail(PICCA_E_VALUE, "parameters must not be NULL");
    const int64_t nw1 = (int64_t)n_in * n_hidden;
    if (!finite(w1, nw1) || !finite(b1, n_hidden) || !finite(w2, 2 * n_hidden) || !finite(b2, 2))
        return fail(PICCA_E_VALUE, "parameters must be finite");
    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    PICCA_HIP_CHECK(hipStreamSynchronize(st));
    std::vector<float> w1t((size_t)nw1);
    for (int32_t k = 0; k < n_in; ++k)
        for (int32_t j = 0; j < n_hidden; ++j) w1t[(size_t)j * n_in + k] = w1[(size_t)k * n_hidden + j];
    ctx->mlp_in = n_in;
    ctx->mlp_hidden = n_hidden;
    ctx->mlp_step = 0;
    PICCA_HIP_CHECK(ctx->m_w1t.upload(w1t.data(), (size_t)nw1, st));
    PICCA_HIP_CHECK(ctx->m_w1t_bf16.reserve((size_t)nw1));
    PICCA_HIP_CHECK(ctx->m_b1.upload(b1, (size_t)n_hidden, st));
    PICCA_HIP_CHECK(ctx->m_w2.upload(w2, (size_t)(2 * n_hidden), st));
    PICCA_HIP_CHECK(ctx->m_b2.upload(b2, 2, st));
    PICCA_HIP_CHECK(ctx->g_w1t.reserve((size_t)nw1));
    PICCA_HIP_CHECK(ctx->g_b1.reserve((size_t)n_hidden));
    PICCA_HIP_CHECK(ctx->g_w2.reserve((size_t)(2 * n_hidden)));
    PICCA_HIP_CHECK(ctx->g_b2.reserve(2));
    const int64_t np = mlp_params(ctx);
    PICCA_HIP_CHECK(ctx->a_m.reserve((size_t)np));
    PICCA_HIP_CHECK(ctx->a_v.reserve((size_t)np));
    int s = zero_f32(ctx->a_m.ptr, np, st);
    if (s == PICCA_OK) s = zero_f32(ctx->a_v.ptr, np, st);
    if (s == PICCA_OK) s = zero_grads(ctx);
    if (s != PICCA_OK) return s;
    PICCA_HIP_CHECK(picca::launch_mlp_pack(mlp_view(ctx), st));
    PICCA_HIP_CHECK(hipStreamSynchronize(st));
    return PICCA_OK;
}

int picca_mlp_get(picca_ctx *ctx, float *w1, float *b1, float *w2, float *b2) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    if (ctx->mlp_in == 0) return fail(PICCA_E_VALUE, "the network is not initialised");
    if (!w1 || !b1 || !w2 || !b2) return fail(PICCA_E_VALUE, "arrays must not be NULL");
    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int32_t n_in = ctx->mlp_in, n_hidden = ctx->mlp_hidden;
    std::vector<float> w1t((size_t)n_in * n_hidden);
    PICCA_HIP_CHECK(hipMemcpyAsync(w1t.data(), ctx->m_w1t.ptr, sizeof(float) * w1t.size(),
                                   hipMemcpyDeviceToHost, st));
    PICCA_HIP_CHECK(hipMemcpyAsync(b1, ctx->m_b1.ptr, sizeof(float) * n_hidden, hipMemcpyDeviceToHost, st));
    PICCA_HIP_CHECK(hipMemcpyAsync(w2, ctx->m_w2.ptr, sizeof(float) * 2 * n_hidden, hipMemcpyDeviceToHost, st));
    PICCA_HIP_CHECK(hipMemcpyAsync(b2, ctx->m_b2.ptr, sizeof(float) * 2, hipMemcpyDeviceToHost, st));
    PICCA_HIP_CHECK(hipStreamSynchronize(st));
    for (int32_t k = 0; k < n_in; ++k)
        for (int32_t j = 0; j < n_hidden; ++j) w1[(size_t)k * n_hidden + j] = w1t[(size_t)j * n_in + k];
    return PICCA_OK;
}

int picca_mlp_forward(picca_ctx *ctx, int64_t n, const float *x, float *y, double *kernel_ms) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    if (ctx->mlp_in == 0) return fail(PICCA_E_VALUE, "the network is not initialised");
    if (n < 0) return fail(PICCA_E_VALUE, "negative size");
    if (kernel_ms) *kernel_ms = 0.0;
    if (n == 0) return PICCA_OK;
    if (!x || !y) return fail(PICCA_E_VALUE, "arrays must not be NULL");
    if (!finite(x, n * ctx->mlp_in)) return fail(PICCA_E_VALUE, "inputs must be finite");
    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    DeviceBuffer<float> dx, dy;
    PICCA_HIP_CHECK(dx.upload(x, (size_t)(n * ctx->mlp_in), st));
    PICCA_HIP_CHECK(dy.reserve((size_t)(2 * n)));
    PICCA_HIP_CHECK(hipEventRecord(ctx->ev_start, st));
    PICCA_HIP_CHECK(picca::launch_mlp_forward(mlp_view(ctx), dx.ptr, nullptr, n, nullptr, dy.ptr, st));
    PICCA_HIP_CHECK(hipEventRecord(ctx->ev_stop, st));
    PICCA_HIP_CHECK(hipMemcpyAsync(y, dy.ptr, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, st));
    PICCA_HIP_CHECK(hipStreamSynchronize(st));
    float ms = 0.f;
    PICCA_HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    if (kernel_ms) *kernel_ms = (double)ms;
    return PICCA_OK;
}

int picca_mlp_gradients(picca_ctx *ctx, int64_t n, const float *x, const float *t,
                        float *g_w1, float *g_b1, float *g_w2, float *g_b2, float *loss) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    if (ctx->mlp_in == 0) return fail(PICCA_E_VALUE, "the network is not initialised");
    if (n <= 0) return fail(PICCA_E_VALUE, "the batch must not be empty");
    if (!x || !t || !g_w1 || !g_b1 || !g_w2 || !g_b2 || !loss)
        return fail(PICCA_E_VALUE, "arrays must not be NULL");
    if (!finite(x, n * ctx->mlp_in) || !finite(t, 2 * n))
        return fail(PICCA_E_VALUE, "inputs must be finite");
    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    DeviceBuffer<float> dx, dt, dloss;
    PICCA_HIP_CHECK(dx.upload(x, (size_t)(n * ctx->mlp_in), st));
    PICCA_HIP_CHECK(dt.upload(t, (size_t)(2 * n), st));
    PICCA_HIP_CHECK(dloss.reserve(1));
    int s = zero_f32(dloss.ptr, 1, st);
    if (s == PICCA_OK) s = zero_grads(ctx);
    if (s == PICCA_OK) s = reserve_scratch(ctx, n);
    if (s != PICCA_OK) return s;
    const picca::MlpView M = mlp_view(ctx);
    const picca::MlpGrads G = mlp_grads(ctx, dloss.ptr);
    PICCA_HIP_CHECK(picca::launch_mlp_forward(M, dx.ptr, nullptr, n, ctx->m_h.ptr, ctx->m_y.ptr, st));
    PICCA_HIP_CHECK(picca::launch_mlp_backward(M, G, dx.ptr, dt.ptr, nullptr, n, ctx->m_h.ptr,
                                               ctx->m_y.ptr, ctx->m_dh.ptr, st));
    const int32_t n_in = ctx->mlp_in, n_hidden = ctx->mlp_hidden;
    std::vector<float> gw1t((size_t)n_in * n_hidden);
    PICCA_HIP_CHECK(hipMemcpyAsync(gw1t.data(), ctx->g_w1t.ptr, sizeof(float) * gw1t.size(), hipMemcpyDeviceToHost, st));
    PICCA_HIP_CHECK(hipMemcpyAsync(g_b1, ctx->g_b1.ptr, sizeof(float) * n_hidden, hipMemcpyDeviceToHost, st));
    PICCA_HIP_CHECK(hipMemcpyAsync(g_w2, ctx->g_w2.ptr, sizeof(float) * 2 * n_hidden, hipMemcpyDeviceToHost, st));
    PICCA_HIP_CHECK(hipMemcpyAsync(g_b2, ctx->g_b2.ptr, sizeof(float) * 2, hipMemcpyDeviceToHost, st));
    PICCA_HIP_CHECK(hipMemcpyAsync(loss, dloss.ptr, sizeof(float), hipMemcpyDeviceToHost, st));
    PICCA_HIP_CHECK(hipStreamSynchronize(st));
    for (int32_t k = 0; k < n_in; ++k)
        for (int32_t j = 0; j < n_hidden; ++j) g_w1[(size_t)k * n_hidden + j] = gw1t[(size_t)j * n_in + k];
    return zero_grads(ctx) == PICCA_OK ? PICCA_OK : PICCA_E_HIP;
}

int picca_mlp_set_data(picca_ctx *ctx, int64_t n, const float *x, const float *t) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    if (ctx->mlp_in == 0) return fail(PICCA_E_VALUE, "the network is not initialised");
    if (n <= 0 || n > INT32_MAX) return fail(PICCA_E_VALUE, "training set size out of range");
    if (!x || !t) return fail(PICCA_E_VALUE, "arrays must not be NULL");
    if (!finite(x, n * ctx->mlp_in) || !finite(t, 2 * n))
        return fail(PICCA_E_VALUE, "training data must be finite");
    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    PICCA_HIP_CHECK(ctx->m_x.upload(x, (size_t)(n * ctx->mlp_in), st));
    PICCA_HIP_CHECK(ctx->m_t.upload(t, (size_t)(2 * n), st));
    PICCA_HIP_CHECK(hipStreamSynchronize(st));
    ctx->mlp_rows = n;
    return PICCA_OK;
}

int picca_mlp_train(picca_ctx *ctx, int64_t steps, int32_t batch, float lr, uint64_t seed,
                    float *losses) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    if (ctx->mlp_in == 0 || ctx->mlp_rows == 0)
        return fail(PICCA_E_VALUE, "the network or its training data is not set");
    if (steps <= 0 || batch <= 0 || batch > ctx->mlp_rows || !(lr > 0))
        return fail(PICCA_E_VALUE, "steps, batch (<= rows) and lr must be positive");
    if (!losses) return fail(PICCA_E_VALUE, "losses must not be NULL");
    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    // minibatch rows: consecutive slices of seeded permutations (splitmix64 + Fisher-Yates)
    const int64_t n = ctx->mlp_rows;
    std::vector<int32_t> rows((size_t)(steps * batch));
    std::vector<int32_t> perm((size_t)n);
    uint64_t state = seed;
    auto next = [&state]() {
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    int64_t pos = n;   // next unused position of the current permutation
    for (int64_t k = 0; k < steps; ++k) {
        if (pos + batch > n) {   // a new epoch: fresh permutation
            for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
            for (int64_t i = n - 1; i > 0; --i)
                std::swap(perm[i], perm[next() % (uint64_t)(i + 1)]);
            pos = 0;
        }
        for (int32_t b = 0; b < batch; ++b) rows[k * batch + b] = perm[pos++];
    }
    PICCA_HIP_CHECK(ctx->m_rows.upload(rows.data(), rows.size(), st));
    DeviceBuffer<float> dloss;
    PICCA_HIP_CHECK(dloss.reserve((size_t)steps));
    int s = zero_f32(dloss.ptr, steps, st);
    if (s == PICCA_OK) s = reserve_scratch(ctx, batch);
    if (s != PICCA_OK) return s;
    const picca::MlpView M = mlp_view(ctx);
    const picca::MlpAdam A{ctx->a_m.ptr, ctx->a_v.ptr};
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    for (int64_t k = 0; k < steps; ++k) {
        const int32_t *r = ctx->m_rows.ptr + k * batch;
        const picca::MlpGrads G = mlp_grads(ctx, dloss.ptr + k);
        PICCA_HIP_CHECK(picca::launch_mlp_forward(M, ctx->m_x.ptr, r, batch, ctx->m_h.ptr,
                                                  ctx->m_y.ptr, st));
        PICCA_HIP_CHECK(picca::launch_mlp_backward(M, G, ctx->m_x.ptr, ctx->m_t.ptr, r, batch,
                                                   ctx->m_h.ptr, ctx->m_y.ptr, ctx->m_dh.ptr, st));
        ctx->mlp_step += 1;
        const float corr1 = 1.f - std::pow(beta1, (float)ctx->mlp_step);
        const float corr2 = 1.f - std::pow(beta2, (float)ctx->mlp_step);
        PICCA_HIP_CHECK(picca::launch_mlp_adam(M, G, A, lr, beta1, beta2, eps, corr1, corr2, st));
    }
    PICCA_HIP_CHECK(hipMemcpyAsync(losses, dloss.ptr, sizeof(float) * steps, hipMemcpyDeviceToHost, st));
    PICCA_HIP_CHECK(hipStreamSynchronize(st));
    return PICCA_OK;
}

}  // extern "C"
