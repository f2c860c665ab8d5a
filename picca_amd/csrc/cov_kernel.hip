// Subsampling covariance of the per-healpix correlation rows on gfx950 (MI355X):
// compute_cov of py/picca/utils.py:100-127.
//
// With xi and weights of shape [num_samples][num_bins] (row-major, one row per healpix):
//     S[b] = sum_k w[k][b],  mean[b] = sum_k w[k][b] xi[k][b] / S[b]   (S[b] > 0),
//     Y[k][b] = w[k][b] (xi[k][b] - mean[b]),
//     cov[a][b] = sum_k Y[k][a] Y[k][b] / (S[a] S[b])                   (S[a] S[b] > 0).
// cov_prep_kernel forms S, mean and Y with one thread per bin (sequential sums over the
// samples, as numpy's reduction over the leading axis); cov_syrk_kernel forms the upper
// triangle of Y^T Y on the FP64 matrix cores (v_mfma_f64_16x16x4f64) in 64 x 64 tiles,
// K staged through LDS 16 samples at a time, divides by S[a] S[b] and writes both halves.
#include <hip/hip_runtime.h>

#include "engine.h"

namespace picca {

namespace {

constexpr int COV_TILE = 64;    // output tile edge
constexpr int COV_KC = 16;      // samples per LDS stage
constexpr int COV_BLOCK = 256;  // 4 waves, each a 32 x 32 quadrant of the tile

typedef double double4_t __attribute__((ext_vector_type(4)));

}  // namespace

__global__ __launch_bounds__(256) void cov_prep_kernel(const double *__restrict__ xi,
                                                       const double *__restrict__ w,
                                                       int64_t num_samples, int64_t num_bins,
                                                       double *__restrict__ S,
                                                       double *__restrict__ Y) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= num_bins) return;
    double sw = 0.0, swx = 0.0;
    for (int64_t k = 0; k < num_samples; ++k) {
        const double wk = w[k * num_bins + b];
        swx += xi[k * num_bins + b] * wk;
        sw += wk;
    }
    double mean = swx;
    if (sw > 0.0) mean = swx / sw;
    S[b] = sw;
    for (int64_t k = 0; k < num_samples; ++k)
        Y[k * num_bins + b] = w[k * num_bins + b] * (xi[k * num_bins + b] - mean);
}

__global__ __launch_bounds__(COV_BLOCK) void cov_syrk_kernel(const double *__restrict__ Y,
                                                             const double *__restrict__ S,
                                                             int64_t num_samples,
                                                             int num_bins, int tiles,
                                                             double *__restrict__ cov) {
    __shared__ double As[COV_KC][COV_TILE];
    __shared__ double Bs[COV_KC][COV_TILE];
    // upper-triangle tile (ti <= tj) of this workgroup
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

    double4_t acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = (double4_t){0.0, 0.0, 0.0, 0.0};

    for (int64_t k0 = 0; k0 < num_samples; k0 += COV_KC) {
        // stage 16 samples of the tile's 64 + 64 columns (zeros past the ends)
#pragma unroll
        for (int e = tid; e < COV_KC * COV_TILE; e += COV_BLOCK) {
            const int r = e / COV_TILE, c = e - r * COV_TILE;
            const int64_t k = k0 + r;
            const bool kin = k < num_samples;
            As[r][c] = (kin && a0 + c < num_bins) ? Y[k * num_bins + a0 + c] : 0.0;
            Bs[r][c] = (kin && b0 + c < num_bins) ? Y[k * num_bins + b0 + c] : 0.0;
        }
        __syncthreads();
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

hipError_t launch_cov(const double *xi, const double *w, int64_t num_samples, int num_bins,
                      double *S, double *Y, double *cov, hipStream_t stream) {
    if (num_bins <= 0) return hipSuccess;
    hipLaunchKernelGGL(cov_prep_kernel, dim3((num_bins + 255) / 256), dim3(256), 0, stream, xi,
                       w, num_samples, (int64_t)num_bins, S, Y);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    const int tiles = (num_bins + COV_TILE - 1) / COV_TILE;
    const int grid = tiles * (tiles + 1) / 2;
    hipLaunchKernelGGL(cov_syrk_kernel, dim3(grid), dim3(COV_BLOCK), 0, stream, Y, S,
                       num_samples, num_bins, tiles, cov);
    return hipGetLastError();
}

}  // namespace picca
