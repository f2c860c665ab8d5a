// C-ABI of the batched continuum fit (declared in include/picca_hip.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "context.h"

extern "C" {

int picca_cont_set_forests(picca_ctx *ctx, int64_t num_forests, const int64_t *offsets,
                           const double *flux, const double *ivar, const double *mean_cont,
                           const double *x, const double *eta, const double *var_lss,
                           const double *fudge, const int32_t *order) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    if (num_forests < 0 || num_forests > INT32_MAX)
        return fail(PICCA_E_VALUE, "num_forests out of range");
    if (!offsets || !flux || !ivar || !mean_cont || !x || !eta || !var_lss || !fudge || !order)
        return fail(PICCA_E_VALUE, "continuum arrays must not be NULL");
    if (offsets[0] != 0) return fail(PICCA_E_VALUE, "offsets[0] must be 0");
    for (int64_t k = 0; k < num_forests; ++k)
        if (offsets[k + 1] < offsets[k] || offsets[k + 1] - offsets[k] > INT32_MAX / 2)
            return fail(PICCA_E_VALUE, "offsets must be non-decreasing");
    const int64_t npix = offsets[num_forests];
    for (int64_t k = 0; k < npix; ++k) {
        if (!std::isfinite(flux[k]) || !std::isfinite(ivar[k]) || !std::isfinite(mean_cont[k]) ||
            !std::isfinite(x[k]) || !std::isfinite(eta[k]) || !std::isfinite(var_lss[k]) ||
            !std::isfinite(fudge[k]))
            return fail(PICCA_E_VALUE, "continuum inputs must be finite");
    }
    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    hipStream_t st = ctx->stream;
    const size_t n = (size_t)npix;
    PICCA_HIP_CHECK(ctx->c_off.upload(offsets, (size_t)num_forests + 1, st));
    PICCA_HIP_CHECK(ctx->c_flux.upload(flux, n, st));
    PICCA_HIP_CHECK(ctx->c_ivar.upload(ivar, n, st));
    PICCA_HIP_CHECK(ctx->c_mean.upload(mean_cont, n, st));
    PICCA_HIP_CHECK(ctx->c_x.upload(x, n, st));
    PICCA_HIP_CHECK(ctx->c_eta.upload(eta, n, st));
    PICCA_HIP_CHECK(ctx->c_lss.upload(var_lss, n, st));
    PICCA_HIP_CHECK(ctx->c_fudge.upload(fudge, n, st));
    PICCA_HIP_CHECK(ctx->c_order.upload(order, (size_t)num_forests, st));
    const size_t nf = (size_t)std::max<int64_t>(num_forests, 1);
    PICCA_HIP_CHECK(ctx->c_a0.reserve(nf));
    PICCA_HIP_CHECK(ctx->c_b0.reserve(nf));
    PICCA_HIP_CHECK(ctx->c_a.reserve(nf));
    PICCA_HIP_CHECK(ctx->c_b.reserve(nf));
    PICCA_HIP_CHECK(ctx->c_chi2.reserve(nf));
    PICCA_HIP_CHECK(ctx->c_ndata.reserve(nf));
    PICCA_HIP_CHECK(ctx->c_status.reserve(nf));
    PICCA_HIP_CHECK(ctx->c_evals.reserve(nf));
    PICCA_HIP_CHECK(hipStreamSynchronize(st));
    ctx->cont_forests = num_forests;
    ctx->cont_pixels = npix;
    ctx->cont_nbins = 0;
    ctx->cont_fitted = false;
    return PICCA_OK;
}

int picca_cont_fit(picca_ctx *ctx, const picca_cont_params *params, const double *a0,
                   const double *b0, double *a, double *b, double *chi2, int32_t *ndata,
                   int32_t *status, int32_t *evaluations, double *kernel_ms) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    if (!params) return fail(PICCA_E_VALUE, "params is NULL");
    if (params->max_iter < 1 || !(params->tol > 0) || !(params->tol_abs >= 0))
        return fail(PICCA_E_VALUE, "max_iter must be >= 1, tol > 0 and tol_abs >= 0");
    const int64_t nf = ctx->cont_forests;
    if (nf > 0 && (!a0 || !b0 || !a || !b || !chi2 || !ndata || !status || !evaluations))
        return fail(PICCA_E_VALUE, "arrays must not be NULL");
    for (int64_t k = 0; k < nf; ++k)
        if (!std::isfinite(a0[k]) || !std::isfinite(b0[k]))
            return fail(PICCA_E_VALUE, "start values must be finite");
    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    PICCA_HIP_CHECK(ctx->c_a0.upload(a0, (size_t)nf, st));
    PICCA_HIP_CHECK(ctx->c_b0.upload(b0, (size_t)nf, st));
    picca::ContConst K{params->max_iter, params->use_constant_weight, params->tol,
                       params->tol_abs};
    picca::ContIn I{ctx->c_off.ptr,  ctx->c_flux.ptr, ctx->c_ivar.ptr,  ctx->c_mean.ptr,
                    ctx->c_x.ptr,    ctx->c_eta.ptr,  ctx->c_lss.ptr,   ctx->c_fudge.ptr,
                    ctx->c_order.ptr, ctx->c_a0.ptr,  ctx->c_b0.ptr};
    picca::ContOut O{ctx->c_a.ptr, ctx->c_b.ptr, ctx->c_chi2.ptr, ctx->c_ndata.ptr,
                     ctx->c_status.ptr, ctx->c_evals.ptr};
    PICCA_HIP_CHECK(hipEventRecord(ctx->ev_start, st));
    PICCA_HIP_CHECK(picca::launch_cont_fit(K, I, O, nf, st));
    PICCA_HIP_CHECK(hipEventRecord(ctx->ev_stop, st));
    if (nf > 0) {
        PICCA_HIP_CHECK(hipMemcpyAsync(a, ctx->c_a.ptr, sizeof(double) * nf, hipMemcpyDeviceToHost, st));
        PICCA_HIP_CHECK(hipMemcpyAsync(b, ctx->c_b.ptr, sizeof(double) * nf, hipMemcpyDeviceToHost, st));
        PICCA_HIP_CHECK(hipMemcpyAsync(chi2, ctx->c_chi2.ptr, sizeof(double) * nf, hipMemcpyDeviceToHost, st));
        PICCA_HIP_CHECK(hipMemcpyAsync(ndata, ctx->c_ndata.ptr, sizeof(int32_t) * nf, hipMemcpyDeviceToHost, st));
        PICCA_HIP_CHECK(hipMemcpyAsync(status, ctx->c_status.ptr, sizeof(int32_t) * nf, hipMemcpyDeviceToHost, st));
        PICCA_HIP_CHECK(hipMemcpyAsync(evaluations, ctx->c_evals.ptr, sizeof(int32_t) * nf, hipMemcpyDeviceToHost, st));
    }
    PICCA_HIP_CHECK(hipStreamSynchronize(st));
    float ms = 0.f;
    PICCA_HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    if (kernel_ms) *kernel_ms = (double)ms;
    ctx->cont_fitted = true;
    return PICCA_OK;
}

int picca_cont_set_rest_frame(picca_ctx *ctx, const double *log_lambda_rest,
                              const double *transmission, const int32_t *bins,
                              int32_t num_bins) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    const int64_t npix = ctx->cont_pixels;
    if (num_bins < 2 || num_bins > PICCA_CONT_MAX_BINS)
        return fail(PICCA_E_VALUE, "num_bins must be in [2, " +
                                       std::to_string(PICCA_CONT_MAX_BINS) + "]");
    if (npix > 0 && (!log_lambda_rest || !transmission || !bins))
        return fail(PICCA_E_VALUE, "rest-frame arrays must not be NULL");
    for (int64_t k = 0; k < npix; ++k) {
        if (bins[k] < 0 || bins[k] >= num_bins)
            return fail(PICCA_E_VALUE, "a pixel's rest-frame bin is outside [0, num_bins)");
        if (!std::isfinite(log_lambda_rest[k]) || !std::isfinite(transmission[k]))
            return fail(PICCA_E_VALUE, "rest-frame inputs must be finite");
    }
    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    PICCA_HIP_CHECK(ctx->c_llr.upload(log_lambda_rest, (size_t)npix, st));
    PICCA_HIP_CHECK(ctx->c_trans.upload(transmission, (size_t)npix, st));
    PICCA_HIP_CHECK(ctx->c_bin.upload(bins, (size_t)npix, st));
    PICCA_HIP_CHECK(ctx->c_stack.reserve(2 * (size_t)num_bins));
    PICCA_HIP_CHECK(ctx->c_xs.reserve((size_t)num_bins));
    PICCA_HIP_CHECK(ctx->c_ys.reserve((size_t)num_bins));
    PICCA_HIP_CHECK(hipStreamSynchronize(st));
    ctx->cont_nbins = num_bins;
    return PICCA_OK;
}

int picca_cont_stack(picca_ctx *ctx, double *sum_ratio, double *sum_weight,
                     double *kernel_ms) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    if (ctx->cont_nbins == 0)
        return fail(PICCA_E_RUNTIME, "no rest frame set (picca_cont_set_rest_frame)");
    if (!ctx->cont_fitted) return fail(PICCA_E_RUNTIME, "no fit of the resident forests");
    if (!sum_ratio || !sum_weight) return fail(PICCA_E_VALUE, "outputs must not be NULL");
    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int32_t nbins = ctx->cont_nbins;
    double *d_ratio = ctx->c_stack.ptr, *d_weight = ctx->c_stack.ptr + nbins;
    PICCA_HIP_CHECK(picca::launch_zero_u64(reinterpret_cast<unsigned long long *>(d_ratio),
                                           2 * (int64_t)nbins, st));
    picca::ContIn I{ctx->c_off.ptr,  ctx->c_flux.ptr, ctx->c_ivar.ptr,  ctx->c_mean.ptr,
                    ctx->c_x.ptr,    ctx->c_eta.ptr,  ctx->c_lss.ptr,   ctx->c_fudge.ptr,
                    ctx->c_order.ptr, ctx->c_a0.ptr,  ctx->c_b0.ptr};
    PICCA_HIP_CHECK(hipEventRecord(ctx->ev_start, st));
    PICCA_HIP_CHECK(picca::launch_cont_stack(I, ctx->c_a.ptr, ctx->c_b.ptr, ctx->c_status.ptr,
                                             ctx->c_bin.ptr, nbins, d_ratio, d_weight,
                                             ctx->cont_forests, ctx->num_cus, st));
    PICCA_HIP_CHECK(hipEventRecord(ctx->ev_stop, st));
    PICCA_HIP_CHECK(hipMemcpyAsync(sum_ratio, d_ratio, sizeof(double) * nbins,
                                   hipMemcpyDeviceToHost, st));
    PICCA_HIP_CHECK(hipMemcpyAsync(sum_weight, d_weight, sizeof(double) * nbins,
                                   hipMemcpyDeviceToHost, st));
    PICCA_HIP_CHECK(hipStreamSynchronize(st));
    float ms = 0.f;
    PICCA_HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    if (kernel_ms) *kernel_ms = (double)ms;
    return PICCA_OK;
}

int picca_cont_set_mean_cont(picca_ctx *ctx, int32_t n, const double *x, const double *y,
                             double *kernel_ms) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    if (ctx->cont_nbins == 0)
        return fail(PICCA_E_RUNTIME, "no rest frame set (picca_cont_set_rest_frame)");
    if (n < 2 || n > ctx->cont_nbins)
        return fail(PICCA_E_VALUE, "the mean continuum needs 2 to num_bins nodes");
    if (!x || !y) return fail(PICCA_E_VALUE, "nodes must not be NULL");
    for (int32_t k = 0; k < n; ++k) {
        if (!std::isfinite(x[k]) || !std::isfinite(y[k]))
            return fail(PICCA_E_VALUE, "nodes must be finite");
        if (k > 0 && !(x[k] > x[k - 1]))
            return fail(PICCA_E_VALUE, "node positions must increase strictly");
    }
    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    PICCA_HIP_CHECK(ctx->c_xs.upload(x, (size_t)n, st));
    PICCA_HIP_CHECK(ctx->c_ys.upload(y, (size_t)n, st));
    PICCA_HIP_CHECK(hipEventRecord(ctx->ev_start, st));
    PICCA_HIP_CHECK(picca::launch_cont_interp(ctx->c_llr.ptr, ctx->c_trans.ptr, ctx->c_xs.ptr,
                                              ctx->c_ys.ptr, n, ctx->c_mean.ptr,
                                              ctx->cont_pixels, ctx->num_cus, st));
    PICCA_HIP_CHECK(hipEventRecord(ctx->ev_stop, st));
    PICCA_HIP_CHECK(hipStreamSynchronize(st));
    float ms = 0.f;
    PICCA_HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    if (kernel_ms) *kernel_ms = (double)ms;
    ctx->cont_fitted = false;   // the resident fit belongs to the old mean continuum
    return PICCA_OK;
}

}  // extern "C"
