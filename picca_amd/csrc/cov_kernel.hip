// Subsampling covariance of the per-healpix correlation rows on gfx950 (MI355X):
// compute_cov of py/picca/utils.py:100-127.
//
// With xi and weights of shape [num_samples][num_bins] (row-major, one row per healpix):
//     S[b] = sum_k w[k][b],  mean[b] = sum_k w[k][b] xi[k][b] / S[b]   (S[b] > 0),
//     Y[k][b] = w[k][b] (xi[k][b] - mean[b]),
//     cov[a][b] = sum_k Y[k][a] Y[k][b] / (S[a] S[b])                   (S[a] S[b] > 0).
// cov_prep_kernel forms S and the means with one thread per bin (sequential sums over the
// samples, as numpy's reduction over the leading axis); cov_syrk_kernel forms Y while
// staging it through LDS and the upper triangle of Y^T Y on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64) in 64 x 64 tiles, divides by S[a] S[b] and writes both halves.
#include <hip/hip_runtime.h>

#include "engine.h"

namespace picca {

namespace {

constexpr int COV_TILE = 64;    // output tile edge
constexpr int COV_KC = 16;      // samples per LDS stage
constexpr int COV_BLOCK = 256;  // 4 waves, each a 32 x 32 quadrant of the tile

typedef double double4_t __attribute__((ext_vector_type(4)));

}  // namespace

// S[b] and mean[b]: sequential sums over the samples (numpy's order for a reduction over
// the leading axis); 32 samples' loads are issued before their additions.
__global__ __launch_bounds__(256) void cov_prep_kernel(const double *__restrict__ xi,
                                                       const double *__restrict__ w,
                                                       int64_t num_samples, int64_t num_bins,
                                                       double *__restrict__ S,
                                                       double *__restrict__ mean) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= num_bins) return;
    double sw = 0.0, swx = 0.0;
    int64_t k = 0;
    for (; k + 32 <= num_samples; k += 32) {
        double wk[32], xk[32];
#pragma unroll
        for (int u = 0; u < 32; ++u) {
            wk[u] = w[(k + u) * num_bins + b];
            xk[u] = xi[(k + u) * num_bins + b];
        }
#pragma unroll
        for (int u = 0; u < 32; ++u) {
            swx += xk[u] * wk[u];
            sw += wk[u];
        }
    }
    for (; k < num_samples; ++k) {
        const double wk = w[k * num_bins + b];
        swx += xi[k * num_bins + b] * wk;
        sw += wk;
    }
    S[b] = sw;
    mean[b] = sw > 0.0 ? swx / sw : swx;
}

// Upper triangle of Y^T Y / (S S^T) with Y = w (xi - mean) formed while staging: each
// workgroup owns a 64 x 64 tile (ti <= tj); K runs in stages of 16 samples through LDS, the
// next stage's global loads issued before the current stage's matrix-core work.
__global__ __launch_bounds__(COV_BLOCK) void cov_syrk_kernel(const double *__restrict__ xi,
                                                             const double *__restrict__ w,
                                                             const double *__restrict__ S,
                                                             const double *__restrict__ mean,
                                                             int64_t num_samples,
                                                             int num_bins, int tiles,
                                                             double *__restrict__ cov) {
    __shared__ double As[COV_KC][COV_TILE];
    __shared__ double Bs[COV_KC][COV_TILE];
    int t = blockIdx.x, ti = 0;
    while (t >= tiles - ti) {
        t -= tiles - ti;
        ++ti;
    }
    const int tj = ti + t;
    const int a0 = ti * COV_TILE, b0 = tj * COV_TILE;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;

    // this thread stages column c = tid % 64 of rows r = tid / 64 + 4 q, q = 0..PER-1
    constexpr int PER = COV_KC * COV_TILE / COV_BLOCK;   // 4
    const int c = tid & (COV_TILE - 1);
    const int r0 = tid >> 6;
    const bool ain = a0 + c < num_bins, bin = b0 + c < num_bins;
    const double ma = ain ? mean[a0 + c] : 0.0;
    const double mb = bin ? mean[b0 + c] : 0.0;
    double xa[PER], wa[PER], xb[PER], wb[PER];
    auto fetch = [&](int64_t k0) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int64_t k = k0 + r0 + 4 * q;
            const bool kin = k < num_samples;
            xa[q] = (kin && ain) ? xi[k * num_bins + a0 + c] : 0.0;
            wa[q] = (kin && ain) ? w[k * num_bins + a0 + c] : 0.0;
            xb[q] = (kin && bin) ? xi[k * num_bins + b0 + c] : 0.0;
            wb[q] = (kin && bin) ? w[k * num_bins + b0 + c] : 0.0;
        }
    };

    double4_t acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = (double4_t){0.0, 0.0, 0.0, 0.0};

    fetch(0);
    for (int64_t k0 = 0; k0 < num_samples; k0 += COV_KC) {
        // Y of the staged samples (w (xi - mean), 0 past the ends)
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            As[r0 + 4 * q][c] = wa[q] * (xa[q] - ma);
            Bs[r0 + 4 * q][c] = wb[q] * (xb[q] - mb);
        }
        __syncthreads();
        if (k0 + COV_KC < num_samples) fetch(k0 + COV_KC);
#pragma unroll
        for (int kk = 0; kk < COV_KC; kk += 4) {
            double av[2], bv[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) av[x] = As[kk + lk][wr + 16 * x + li];
#pragma unroll
            for (int y = 0; y < 2; ++y) bv[y] = Bs[kk + lk][wc + 16 * y + li];
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
                    acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[x], bv[y], acc[x][y],
                                                                     0, 0, 0);
        }
        __syncthreads();
    }

    // the f64 16x16x4 result: register r of lane l holds row 4 r + l / 16, column l % 16
    // of each 16 x 16 block (the A and B operands: lane l holds A[l % 16][l / 16] and
    // B[l / 16][l % 16])
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int a = a0 + wr + 16 * x + 4 * r + lk;
                const int b = b0 + wc + 16 * y + li;
                if (a >= num_bins || b >= num_bins) continue;
                if (ti == tj && a > b) continue;   // the diagonal tile writes a <= b only
                double v = acc[x][y][r];
                const double ss = S[b] * S[a];
                if (ss > 0.0) v /= ss;
                cov[(int64_t)a * num_bins + b] = v;
                cov[(int64_t)b * num_bins + a] = v;
            }
}

hipError_t launch_cov_prep(const double *xi, const double *w, int64_t num_samples,
                           int num_bins, double *S, double *mean, hipStream_t stream) {
    if (num_bins <= 0) return hipSuccess;
    hipLaunchKernelGGL(cov_prep_kernel, dim3((num_bins + 255) / 256), dim3(256), 0, stream, xi,
                       w, num_samples, (int64_t)num_bins, S, mean);
    return hipGetLastError();
}

hipError_t launch_cov_syrk(const double *xi, const double *w, const double *S,
                           const double *mean, int64_t num_samples, int num_bins, double *cov,
                           hipStream_t stream) {
    if (num_bins <= 0) return hipSuccess;
    const int tiles = (num_bins + COV_TILE - 1) / COV_TILE;
    const int grid = tiles * (tiles + 1) / 2;
    hipLaunchKernelGGL(cov_syrk_kernel, dim3(grid), dim3(COV_BLOCK), 0, stream, xi, w, S, mean,
                       num_samples, num_bins, tiles, cov);
    return hipGetLastError();
}

}  // namespace picca
