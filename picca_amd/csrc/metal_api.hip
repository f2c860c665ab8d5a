// C-ABI of the metal distortion matrices (declared in include/picca_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "context.h"

extern "C" {

int picca_metal_set_absorber(picca_ctx *ctx, int32_t slot, const double *z,
                             const double *r_comov, const double *dist_m,
                             const double *weights, const double *zfac) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    if (slot < 0 || slot >= 4) return fail(PICCA_E_VALUE, "absorber slot must be 0..3");
    const int64_t npix = ctx->num_pixels;
    if (ctx->num_forests <= 0) return fail(PICCA_E_RUNTIME, "no forest store");
    if (npix > 0 && (!z || !r_comov || !dist_m || !weights || !zfac))
        return fail(PICCA_E_VALUE, "absorber arrays must not be NULL");
    for (int64_t k = 0; k < npix; ++k)
        if (!std::isfinite(z[k]) || !std::isfinite(r_comov[k]) || !std::isfinite(dist_m[k]) ||
            !std::isfinite(weights[k]) || !std::isfinite(zfac[k]))
            return fail(PICCA_E_VALUE, "absorber values must be finite");
    picca_ctx::MetalSlot &m = ctx->metal[slot];
    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    PICCA_HIP_CHECK(m.z.upload(z, (size_t)npix, ctx->stream));
    PICCA_HIP_CHECK(m.r.upload(r_comov, (size_t)npix, ctx->stream));
    PICCA_HIP_CHECK(m.d.upload(dist_m, (size_t)npix, ctx->stream));
    PICCA_HIP_CHECK(m.w.upload(weights, (size_t)npix, ctx->stream));
    PICCA_HIP_CHECK(m.f.upload(zfac, (size_t)npix, ctx->stream));
    PICCA_HIP_CHECK(m.rec.reserve((size_t)npix + picca::METAL_REC_PAD));
    const picca::MetalTable T{m.z.ptr, m.r.ptr, m.d.ptr, m.w.ptr, m.f.ptr, nullptr};
    PICCA_HIP_CHECK(picca::launch_metal_pack(ctx->r.ptr, ctx->d.ptr, T, npix, m.rec.ptr,
                                             ctx->stream));
    PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    m.set = true;
    m.pixels = npix;
    return PICCA_OK;
}

int picca_metal_pass(picca_ctx *ctx, const picca_cf_params *params, int32_t mode,
                     int32_t slot_a, int32_t slot_b, double zw_den, int64_t num_pairs,
                     const int32_t *pair_a, const int32_t *pair_b, const double *pair_ang,
                     const uint8_t *pair_flags, double *weights_dmat, double *dmat,
                     double *r_par_eff, double *r_trans_eff, double *z_eff,
                     double *weight_eff, picca_metal_report *report) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    if (!params) return fail(PICCA_E_VALUE, "params is NULL");
    if (mode != PICCA_METAL_AUTO && mode != PICCA_METAL_CROSS && mode != PICCA_METAL_XCF)
        return fail(PICCA_E_VALUE, "unknown metal mode");
    if (slot_a < 0 || slot_a >= 4 || slot_b < 0 || slot_b >= 4)
        return fail(PICCA_E_VALUE, "absorber slot must be 0..3");
    for (int32_t slot : {slot_a, slot_b})
        if (!ctx->metal[slot].set || ctx->metal[slot].pixels != ctx->num_pixels)
            return fail(PICCA_E_RUNTIME, "absorber slot not set for the current store");
    const int np = params->num_bins_r_par, nt = params->num_bins_r_trans;
    const int npm = params->num_model_bins_r_par, ntm = params->num_model_bins_r_trans;
    if (np <= 0 || nt <= 0 || npm <= 0 || ntm <= 0)
        return fail(PICCA_E_VALUE, "bin counts must be positive");
    if ((int64_t)np * nt > (1 << 24) || (int64_t)npm * ntm > (1 << 24) ||
        (int64_t)np * nt * ((int64_t)npm * ntm) > (1ll << 32))
        return fail(PICCA_E_VALUE, "too many bins");
    if (!(params->r_par_max > params->r_par_min) || !(params->r_trans_max > 0))
        return fail(PICCA_E_VALUE, "empty separation range");
    if (!(zw_den > 0) || !std::isfinite(zw_den))
        return fail(PICCA_E_VALUE, "zw_den must be positive and finite");
    if (num_pairs < 0) return fail(PICCA_E_VALUE, "negative size");
    if (num_pairs > 0 && (!pair_a || !pair_b || !pair_ang))
        return fail(PICCA_E_VALUE, "pair arrays must not be NULL");
    if (!weights_dmat || !dmat || !r_par_eff || !r_trans_eff || !z_eff || !weight_eff)
        return fail(PICCA_E_VALUE, "output arrays must not be NULL");
    const int64_t nf = ctx->num_forests;
    std::vector<int64_t> task_begin((size_t)std::max<int64_t>(num_pairs, 1));
    int64_t tasks = 0;
    for (int64_t k = 0; k < num_pairs; ++k) {
        if (pair_a[k] < 0 || pair_a[k] >= nf || pair_b[k] < 0 || pair_b[k] >= nf)
            return fail(PICCA_E_VALUE, "pair index out of range of the forest store");
        if (!(pair_ang[k] >= 0.0) || !(pair_ang[k] <= M_PI))
            return fail(PICCA_E_VALUE, "pair angle outside [0, pi]");
        task_begin[k] = tasks;
        const int64_t la = ctx->h_offsets[pair_a[k] + 1] - ctx->h_offsets[pair_a[k]];
        tasks += (la + 63) / 64;
    }
    if (tasks > INT32_MAX) return fail(PICCA_E_VALUE, "too many pixel chunks in one pass");
    std::vector<int32_t> task_pair((size_t)std::max<int64_t>(tasks, 1));
    for (int64_t k = 0; k < num_pairs; ++k) {
        const int64_t end = k + 1 < num_pairs ? task_begin[k + 1] : tasks;
        for (int64_t t = task_begin[k]; t < end; ++t) task_pair[t] = (int32_t)k;
    }
    std::vector<double> cs((size_t)std::max<int64_t>(2 * num_pairs, 1));
    std::vector<uint8_t> flags((size_t)std::max<int64_t>(num_pairs, 1), 0);
    for (int64_t k = 0; k < num_pairs; ++k) {
        cs[2 * k] = std::cos(pair_ang[k] / 2);
        cs[2 * k + 1] = std::sin(pair_ang[k] / 2);
        if (pair_flags) flags[k] = pair_flags[k];
    }

    picca::MetalConst C;
    std::memset(&C, 0, sizeof(C));
    C.np = np;
    C.nt = nt;
    C.nb = np * nt;
    C.npm = npm;
    C.ntm = ntm;
    C.nm = npm * ntm;
    C.mode = mode;
    C.shp = params->remove_same_half_plate_close_pairs != 0;
    C.rpmin = params->r_par_min;
    C.rpmax = params->r_par_max;
    C.rtmax = params->r_trans_max;
    C.span = params->r_par_max - params->r_par_min;
    C.width_p = C.span / np;
    C.inv_den = 1.0 / zw_den;
    C.inv_p = np / C.span;
    C.inv_t = nt / C.rtmax;
    C.inv_pm = npm / C.span;
    C.inv_tm = ntm / C.rtmax;
    // x * (n / den) and x / den * n differ by at most 4 ulp of |t| < n + 1 for any
    // coordinate that can land in range; the margin n 2^-46 + 1e-12 covers that 16 times over
    C.half_p = 0.5 - (np * 0x1p-46 + 1e-12);
    C.half_t = 0.5 - (nt * 0x1p-46 + 1e-12);
    C.half_pm = 0.5 - (npm * 0x1p-46 + 1e-12);
    C.half_tm = 0.5 - (ntm * 0x1p-46 + 1e-12);
    const int64_t nb = C.nb, nm = C.nm;

    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    PICCA_HIP_CHECK(ctx->dpa.upload(pair_a, (size_t)num_pairs, ctx->stream));
    PICCA_HIP_CHECK(ctx->dpb.upload(pair_b, (size_t)num_pairs, ctx->stream));
    PICCA_HIP_CHECK(ctx->dcs.upload(cs.data(), (size_t)(2 * num_pairs), ctx->stream));
    PICCA_HIP_CHECK(ctx->dflags.upload(flags.data(), (size_t)num_pairs, ctx->stream));
    PICCA_HIP_CHECK(ctx->m_task_begin.upload(task_begin.data(), (size_t)num_pairs, ctx->stream));
    PICCA_HIP_CHECK(ctx->m_task_pair.upload(task_pair.data(), (size_t)tasks, ctx->stream));
    PICCA_HIP_CHECK(ctx->o_wdm.reserve((size_t)nb));
    PICCA_HIP_CHECK(ctx->o_dmat.reserve((size_t)(nb * nm)));
    PICCA_HIP_CHECK(ctx->o_eff.reserve((size_t)(4 * nm)));
    PICCA_HIP_CHECK(ctx->m_stats.reserve(picca::METAL_NSTATS));
    PICCA_HIP_CHECK(ctx->queue.reserve(1));
    PICCA_HIP_CHECK(picca::launch_zero_u64((unsigned long long *)ctx->o_wdm.ptr, nb, ctx->stream));
    PICCA_HIP_CHECK(picca::launch_zero_u64((unsigned long long *)ctx->o_dmat.ptr, nb * nm,
                                           ctx->stream));
    PICCA_HIP_CHECK(picca::launch_zero_u64((unsigned long long *)ctx->o_eff.ptr, 4 * nm,
                                           ctx->stream));
    PICCA_HIP_CHECK(picca::launch_zero_u64(ctx->m_stats.ptr, picca::METAL_NSTATS, ctx->stream));
    PICCA_HIP_CHECK(picca::launch_zero_u64(ctx->queue.ptr, 1, ctx->stream));

    const picca_ctx::MetalSlot &sa = ctx->metal[slot_a];
    const picca_ctx::MetalSlot &sb = ctx->metal[slot_b];
    picca::MetalTable TA{sa.z.ptr, sa.r.ptr, sa.d.ptr, sa.w.ptr, sa.f.ptr, sa.rec.ptr};
    picca::MetalTable TB{sb.z.ptr, sb.r.ptr, sb.d.ptr, sb.w.ptr, sb.f.ptr, sb.rec.ptr};
    picca::MetalPairs P{ctx->dpa.ptr, ctx->dpb.ptr, ctx->dcs.ptr, ctx->dflags.ptr,
                        ctx->m_task_begin.ptr, num_pairs, tasks};
    picca::MetalOut O{ctx->o_wdm.ptr, ctx->o_dmat.ptr, ctx->o_eff.ptr, ctx->o_eff.ptr + nm,
                      ctx->o_eff.ptr + 2 * nm, ctx->o_eff.ptr + 3 * nm, ctx->queue.ptr,
                      ctx->m_stats.ptr};
    const size_t static_lds = 1024;
    const bool lds_hist = (int64_t)(picca::metal_lds_bytes(C) + static_lds) <= ctx->lds_limit;
    const int64_t waves = std::max<int64_t>(1, tasks);
    const int grid = (int)std::max<int64_t>(
        1, std::min<int64_t>((int64_t)ctx->num_cus * (lds_hist ? 1 : 2),
                             (waves + picca::METAL_BLOCK / 64 - 1) / (picca::METAL_BLOCK / 64)));
    PICCA_HIP_CHECK(hipEventRecord(ctx->ev_start, ctx->stream));
    PICCA_HIP_CHECK(picca::launch_metal(C, ctx->off.ptr, ctx->mono.ptr, ctx->r.ptr, ctx->d.ptr,
                                        TA, TB, P, ctx->m_task_pair.ptr, O, lds_hist, grid,
                                        ctx->stream));
    PICCA_HIP_CHECK(hipEventRecord(ctx->ev_stop, ctx->stream));

    std::vector<double> h_wdm((size_t)nb), h_dmat((size_t)(nb * nm)), h_eff((size_t)(4 * nm));
    unsigned long long h_stats[picca::METAL_NSTATS];
    PICCA_HIP_CHECK(hipMemcpyAsync(h_wdm.data(), ctx->o_wdm.ptr, sizeof(double) * nb,
                                   hipMemcpyDeviceToHost, ctx->stream));
    PICCA_HIP_CHECK(hipMemcpyAsync(h_dmat.data(), ctx->o_dmat.ptr, sizeof(double) * nb * nm,
                                   hipMemcpyDeviceToHost, ctx->stream));
    PICCA_HIP_CHECK(hipMemcpyAsync(h_eff.data(), ctx->o_eff.ptr, sizeof(double) * 4 * nm,
                                   hipMemcpyDeviceToHost, ctx->stream));
    PICCA_HIP_CHECK(hipMemcpyAsync(h_stats, ctx->m_stats.ptr, sizeof(h_stats),
                                   hipMemcpyDeviceToHost, ctx->stream));
    PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    PICCA_HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    for (int64_t k = 0; k < nb; ++k) weights_dmat[k] += h_wdm[k];
    parallel_for(nb * nm, [&](int64_t lo, int64_t hi) {
        for (int64_t k = lo; k < hi; ++k) dmat[k] += h_dmat[k];
    });
    for (int64_t k = 0; k < nm; ++k) {
        r_par_eff[k] += h_eff[k];
        r_trans_eff[k] += h_eff[nm + k];
        z_eff[k] += h_eff[2 * nm + k];
        weight_eff[k] += h_eff[3 * nm + k];
    }
    if (report) {
        report->data_pairs = (int64_t)h_stats[picca::METAL_STAT_DATA];
        report->model_pairs = (int64_t)h_stats[picca::METAL_STAT_MODEL];
        report->overflow_pairs = (int64_t)h_stats[picca::METAL_STAT_OVERFLOW];
        report->visited_pairs = (int64_t)h_stats[picca::METAL_STAT_VISITED];
        report->kernel_ms = (double)ms;
    }
    return PICCA_OK;
}

}  // extern "C"
