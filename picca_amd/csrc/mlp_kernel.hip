// Seed network of the continuum fit on gfx950: a two-layer perceptron
//     h = relu(x W1 + b1),  y = h W2 + b2
// (n_in inputs, n_hidden hidden units, 2 outputs), trained and evaluated with hand-written
// kernels.  The n_in x n_hidden products run on the matrix cores in BF16 with FP32
// accumulation (v_mfma_f32_16x16x32_bf16); the 2-wide output layer, the biases, the loss
// and Adam run in FP32 on the vector ALUs.
//
// MFMA operand layout (16x16x32, wave64): A is 16 rows x 32 k, lane l holds row l % 16 and
// k = 8 (l / 16) .. 8 (l / 16) + 7; B is 32 k x 16 columns, lane l holds column l % 16 and
// the same k range; the 16x16 FP32 result has lane l holding column l % 16, rows
// 4 (l / 16) .. 4 (l / 16) + 3.
//
// Layouts in HBM: x [n][n_in] FP32; W1 is kept transposed, w1t [n_hidden][n_in] (FP32
// master + BF16 copy), so both MFMA operands of the forward product are read as 16-byte
// runs along k; h [n][n_hidden] FP32 is kept for the backward pass.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine.h"

namespace picca {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bfloat16 bits, round to nearest even (inputs are finite)
__device__ __forceinline__ uint16_t bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ __bf16 to_bf16(float f) {
    return __builtin_bit_cast(__bf16, bf16_bits(f));
}

__device__ __forceinline__ f32x4 mfma(const bf16x8 &a, const bf16x8 &b, const f32x4 &c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

}  // namespace

// Forward pass over rows [0, n): one wave per 16 rows, all n_hidden columns.
// rows: optional gather indices (training minibatch), else identity.
__global__ __launch_bounds__(256) void mlp_forward_kernel(MlpView M, const float *__restrict__ x,
                                                          const int32_t *__restrict__ rows,
                                                          int64_t n, float *__restrict__ h,
                                                          float *__restrict__ y) {
    const int lane = threadIdx.x & 63;
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    if (r0 >= n) return;
    const int q = lane >> 4;          // k group / row group of the lane
    const int c = lane & 15;          // A row / B and C column of the lane
    const int64_t ra = r0 + c;        // row this lane loads for A
    const bool valid_a = ra < n;
    const int64_t src = valid_a ? (rows ? (int64_t)rows[ra] : ra) : 0;
    const int nt = M.n_hidden / 16;
    f32x4 acc[MLP_MAX_HIDDEN / 16];
#pragma unroll
    for (int t = 0; t < MLP_MAX_HIDDEN / 16; ++t) acc[t] = f32x4{0, 0, 0, 0};
    for (int k0 = 0; k0 < M.n_in; k0 += 32) {
        bf16x8 a;
        const float *xp = x + src * M.n_in + k0 + 8 * q;
#pragma unroll
        for (int e = 0; e < 8; ++e) a[e] = to_bf16(valid_a ? xp[e] : 0.f);
#pragma unroll
        for (int t = 0; t < MLP_MAX_HIDDEN / 16; ++t) {
            if (t < nt) {
                const bf16x8 b = *(const bf16x8 *)(M.w1t_bf16 + (int64_t)(t * 16 + c) * M.n_in +
                                                    k0 + 8 * q);
                acc[t] = mfma(a, b, acc[t]);
            }
        }
    }
    // epilogue: bias + relu -> h; output layer partial sums over this lane's columns
    float y0[4] = {0, 0, 0, 0}, y1[4] = {0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < MLP_MAX_HIDDEN / 16; ++t) {
        if (t >= nt) continue;
        const int col = t * 16 + c;
        const float bias = M.b1[col];
        const float w20 = M.w2[col * 2], w21 = M.w2[col * 2 + 1];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t row = r0 + 4 * q + i;
            float v = acc[t][i] + bias;
            v = v > 0.f ? v : 0.f;
            if (row < n && h) h[row * M.n_hidden + col] = v;
            y0[i] += v * w20;
            y1[i] += v * w21;
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        for (int off = 8; off > 0; off >>= 1) {
            y0[i] += __shfl_xor(y0[i], off, 16);
            y1[i] += __shfl_xor(y1[i], off, 16);
        }
    }
    if (c == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t row = r0 + 4 * q + i;
            if (row < n) {
                y[row * 2] = y0[i] + M.b2[0];
                y[row * 2 + 1] = y1[i] + M.b2[1];
            }
        }
    }
}

// Backward through the output layer and the relu, 64 rows per block:
// dy = (y - t) / n (loss = sum (y - t)^2 / (2 n)), dh = (dy W2^T) * (h > 0) -> dh,
// gradients of W2, b2, b1 and the loss accumulated with atomics.
__global__ __launch_bounds__(256) void mlp_backward_out_kernel(MlpView M, MlpGrads G,
                                                               const float *__restrict__ target,
                                                               const int32_t *__restrict__ rows,
                                                               int64_t n,
                                                               const float *__restrict__ h,
                                                               const float *__restrict__ y,
                                                               float *__restrict__ dh) {
    __shared__ float s_red[3 * MLP_MAX_HIDDEN];
    const int tid = threadIdx.x;
    const int col = tid % MLP_MAX_HIDDEN;
    const int half = tid / MLP_MAX_HIDDEN;            // 0 or 1
    const int64_t r0 = (int64_t)blockIdx.x * 64;
    const float inv_n = 1.0f / (float)n;
    float g_b1 = 0, g_w20 = 0, g_w21 = 0, loss = 0, g_b20 = 0, g_b21 = 0;
    const bool active = col < M.n_hidden;
    const float w20 = active ? M.w2[col * 2] : 0.f, w21 = active ? M.w2[col * 2 + 1] : 0.f;
    for (int i = half; i < 64; i += 2) {
        const int64_t row = r0 + i;
        if (row >= n) break;
        const int64_t src = rows ? (int64_t)rows[row] : row;
        const float e0 = y[row * 2] - target[src * 2];
        const float e1 = y[row * 2 + 1] - target[src * 2 + 1];
        const float d0 = e0 * inv_n, d1 = e1 * inv_n;
        if (col == 0) {
            loss += e0 * e0 + e1 * e1;
            g_b20 += d0;
            g_b21 += d1;
        }
        if (active) {
            const float hv = h[row * M.n_hidden + col];
            const float g = hv > 0.f ? d0 * w20 + d1 * w21 : 0.f;
            dh[row * M.n_hidden + col] = g;
            g_b1 += g;
            g_w20 += hv * d0;
            g_w21 += hv * d1;
        }
    }
    if (half == 1 && active) {
        s_red[col] = g_b1;
        s_red[MLP_MAX_HIDDEN + col] = g_w20;
        s_red[2 * MLP_MAX_HIDDEN + col] = g_w21;
    }
    __syncthreads();
    if (half == 0 && active) {
        atomicAdd(&G.b1[col], g_b1 + s_red[col]);
        atomicAdd(&G.w2[col * 2], g_w20 + s_red[MLP_MAX_HIDDEN + col]);
        atomicAdd(&G.w2[col * 2 + 1], g_w21 + s_red[2 * MLP_MAX_HIDDEN + col]);
    }
    if (col == 0) {
        atomicAdd(G.loss, 0.5f * loss * inv_n);
        atomicAdd(&G.b2[0], g_b20);
        atomicAdd(&G.b2[1], g_b21);
    }
}

// dW1^T [n_hidden][n_in] += dh^T x over the rows of this block (MLP_BW_ROWS rows), BF16
// operands staged transposed in LDS so that both fragments are 16-byte runs along the
// batch dimension.
__global__ __launch_bounds__(256) void mlp_backward_w1_kernel(MlpView M, MlpGrads G,
                                                              const float *__restrict__ x,
                                                              const int32_t *__restrict__ rows,
                                                              int64_t n,
                                                              const float *__restrict__ dh) {
    __shared__ __align__(16) uint16_t s_x[MLP_MAX_IN * 32];       // [k][32 rows]
    __shared__ __align__(16) uint16_t s_d[MLP_MAX_HIDDEN * 32];   // [j][32 rows]
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int q = lane >> 4, c = lane & 15;
    const int mt = M.n_hidden / 16;               // output row tiles (hidden units)
    const int ntl = M.n_in / 16;                  // output column tiles (inputs)
    const int tiles = mt * ntl;
    // each wave owns tiles wave, wave + 4, ...
    f32x4 acc[MLP_MAX_TILES_PER_WAVE];
    for (int k = 0; k < MLP_MAX_TILES_PER_WAVE; ++k) acc[k] = f32x4{0, 0, 0, 0};
    const int64_t row_begin = (int64_t)blockIdx.x * MLP_BW_ROWS;
    const int64_t row_end = row_begin + MLP_BW_ROWS < n ? row_begin + MLP_BW_ROWS : n;
    for (int64_t b0 = row_begin; b0 < row_end; b0 += 32) {
        // stage 32 rows of x and dh, transposed, as BF16 (rows past the end are 0)
        for (int e = tid; e < 32 * M.n_in; e += 256) {
            const int r = e / M.n_in, k = e - r * M.n_in;
            const int64_t row = b0 + r;
            float v = 0.f;
            if (row < row_end) v = x[(rows ? (int64_t)rows[row] : row) * M.n_in + k];
            s_x[k * 32 + r] = bf16_bits(v);
        }
        for (int e = tid; e < 32 * M.n_hidden; e += 256) {
            const int r = e / M.n_hidden, j = e - r * M.n_hidden;
            const int64_t row = b0 + r;
            s_d[j * 32 + r] = bf16_bits(row < row_end ? dh[row * M.n_hidden + j] : 0.f);
        }
        __syncthreads();
#pragma unroll
        for (int slot = 0; slot < MLP_MAX_TILES_PER_WAVE; ++slot) {
            const int tile = wave + 4 * slot;
            if (tile < tiles) {
                const int tm = tile / ntl, tn = tile - tm * ntl;
                const bf16x8 a = *(const bf16x8 *)(s_d + (tm * 16 + c) * 32 + 8 * q);
                const bf16x8 b = *(const bf16x8 *)(s_x + (tn * 16 + c) * 32 + 8 * q);
                acc[slot] = mfma(a, b, acc[slot]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int slot = 0; slot < MLP_MAX_TILES_PER_WAVE; ++slot) {
        const int tile = wave + 4 * slot;
        if (tile >= tiles) continue;
        const int tm = tile / ntl, tn = tile - tm * ntl;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int j = tm * 16 + 4 * q + i;   // hidden unit (row of dW1^T)
            const int k = tn * 16 + c;           // input (column of dW1^T)
            atomicAdd(&G.w1t[j * M.n_in + k], acc[slot][i]);
        }
    }
}

// Adam (Kingma & Ba 2015) on every parameter, then the BF16 copy of W1 and zeroed grads.
__global__ __launch_bounds__(256) void mlp_adam_kernel(MlpView M, MlpGrads G, MlpAdam A,
                                                       float lr, float beta1, float beta2,
                                                       float eps, float corr1, float corr2) {
    const int64_t n_w1 = (int64_t)M.n_hidden * M.n_in;
    const int64_t total = n_w1 + M.n_hidden + 2 * M.n_hidden + 2;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (int64_t)gridDim.x * blockDim.x) {
        float *w, *g;
        int64_t idx;
        if (k < n_w1) {
            w = M.w1t; g = G.w1t; idx = k;
        } else if (k < n_w1 + M.n_hidden) {
            w = M.b1; g = G.b1; idx = k - n_w1;
        } else if (k < n_w1 + 3 * M.n_hidden) {
            w = M.w2; g = G.w2; idx = k - n_w1 - M.n_hidden;
        } else {
            w = M.b2; g = G.b2; idx = k - n_w1 - 3 * M.n_hidden;
        }
        const float grad = g[idx];
        const float m = beta1 * A.m[k] + (1.f - beta1) * grad;
        const float v = beta2 * A.v[k] + (1.f - beta2) * grad * grad;
        A.m[k] = m;
        A.v[k] = v;
        const float step = lr * (m / corr1) / (sqrtf(v / corr2) + eps);
        w[idx] -= step;
        g[idx] = 0.f;
        if (k < n_w1) M.w1t_bf16[k] = bf16_bits(w[idx]);
    }
}

__global__ void mlp_pack_kernel(MlpView M) {
    const int64_t n_w1 = (int64_t)M.n_hidden * M.n_in;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n_w1;
         k += (int64_t)gridDim.x * blockDim.x)
        M.w1t_bf16[k] = bf16_bits(M.w1t[k]);
}

hipError_t launch_mlp_forward(const MlpView &M, const float *x, const int32_t *rows, int64_t n,
                              float *h, float *y, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int64_t blocks = (n + 63) / 64;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, M, x,
                       rows, n, h, y);
    return hipGetLastError();
}

hipError_t launch_mlp_backward(const MlpView &M, const MlpGrads &G, const float *x,
                               const float *target, const int32_t *rows, int64_t n,
                               const float *h, const float *y, float *dh, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_backward_out_kernel, dim3((unsigned)((n + 63) / 64)), dim3(256), 0,
                       stream, M, G, target, rows, n, h, y, dh);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(mlp_backward_w1_kernel,
                       dim3((unsigned)((n + MLP_BW_ROWS - 1) / MLP_BW_ROWS)), dim3(256), 0,
                       stream, M, G, x, rows, n, dh);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(const MlpView &M, const MlpGrads &G, const MlpAdam &A, float lr,
                           float beta1, float beta2, float eps, float corr1, float corr2,
                           hipStream_t stream) {
    hipLaunchKernelGGL(mlp_adam_kernel, dim3(256), dim3(256), 0, stream, M, G, A, lr, beta1,
                       beta2, eps, corr1, corr2);
    return hipGetLastError();
}

hipError_t launch_mlp_pack(const MlpView &M, hipStream_t stream) {
    hipLaunchKernelGGL(mlp_pack_kernel, dim3(128), dim3(256), 0, stream, M);
    return hipGetLastError();
}

}  // namespace picca
