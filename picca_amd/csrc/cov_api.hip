// C-ABI of the subsampling covariance (declared in include/picca_hip.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "context.h"

extern "C" {

int picca_cov_subsample(picca_ctx *ctx, int64_t num_samples, int64_t num_bins,
                        const double *xi, const double *weights, double *cov,
                        double *kernel_ms) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    if (num_samples < 0 || num_bins < 0) return fail(PICCA_E_VALUE, "negative size");
    if (num_bins > 65536) return fail(PICCA_E_VALUE, "num_bins above 65536");
    if (num_bins > 0 && (!cov || (num_samples > 0 && (!xi || !weights))))
        return fail(PICCA_E_VALUE, "arrays must not be NULL");
    const int64_t n = num_samples * num_bins;
    for (int64_t k = 0; k < n; ++k)
        if (!std::isfinite(xi[k]) || !std::isfinite(weights[k]))
            return fail(PICCA_E_VALUE, "xi and weights must be finite");
    if (num_bins == 0) return PICCA_OK;
    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    PICCA_HIP_CHECK(ctx->cv_xi.upload(xi, (size_t)n, ctx->stream));
    PICCA_HIP_CHECK(ctx->cv_w.upload(weights, (size_t)n, ctx->stream));
    PICCA_HIP_CHECK(ctx->cv_y.reserve((size_t)num_bins));
    PICCA_HIP_CHECK(ctx->cv_s.reserve((size_t)num_bins));
    PICCA_HIP_CHECK(ctx->cv_cov.reserve((size_t)(num_bins * num_bins)));
    PICCA_HIP_CHECK(hipEventRecord(ctx->ev_start, ctx->stream));
    PICCA_HIP_CHECK(picca::launch_cov_prep(ctx->cv_xi.ptr, ctx->cv_w.ptr, num_samples,
                                           (int)num_bins, ctx->cv_s.ptr, ctx->cv_y.ptr,
                                           ctx->stream));
    PICCA_HIP_CHECK(hipEventRecord(ctx->ev_mid, ctx->stream));
    PICCA_HIP_CHECK(picca::launch_cov_syrk(ctx->cv_xi.ptr, ctx->cv_w.ptr, ctx->cv_s.ptr,
                                           ctx->cv_y.ptr, num_samples, (int)num_bins,
                                           ctx->cv_cov.ptr, ctx->stream));
    PICCA_HIP_CHECK(hipEventRecord(ctx->ev_stop, ctx->stream));
    PICCA_HIP_CHECK(hipMemcpyAsync(cov, ctx->cv_cov.ptr,
                                   sizeof(double) * (size_t)(num_bins * num_bins),
                                   hipMemcpyDeviceToHost, ctx->stream));
    PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    float ms = 0.f, ms_syrk = 0.f;
    PICCA_HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    PICCA_HIP_CHECK(hipEventElapsedTime(&ms_syrk, ctx->ev_mid, ctx->ev_stop));
    if (kernel_ms) {
        kernel_ms[0] = (double)ms;
        kernel_ms[1] = (double)ms_syrk;
    }
    return PICCA_OK;
}

}  // extern "C"
