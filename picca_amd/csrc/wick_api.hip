// C-ABI of the Wick covariance terms (declared in include/picca_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "context.h"

extern "C" {

int picca_wick_t123(picca_ctx *ctx, const picca_cf_params *params, int64_t num_pairs,
                    const int32_t *pair_a, const int32_t *pair_b, const double *pair_ang,
                    const double *zev1, const double *zev2, const int64_t *wx_offsets,
                    int64_t wx_size, const double *wx, double *weights_wick,
                    int64_t *num_pairs_wick, double *t1, double *t2, double *t3,
                    picca_wick_report *report) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    if (!params) return fail(PICCA_E_VALUE, "params is NULL");
    if (ctx->num_forests <= 0) return fail(PICCA_E_RUNTIME, "no forest store");
    const int np = params->num_bins_r_par, nt = params->num_bins_r_trans;
    if (np <= 0 || nt <= 0) return fail(PICCA_E_VALUE, "bin counts must be positive");
    if ((int64_t)np * nt > 8192) return fail(PICCA_E_VALUE, "more than 8192 bins");
    if (!(params->r_par_max > params->r_par_min) || !(params->r_trans_max > 0))
        return fail(PICCA_E_VALUE, "empty separation range");
    if (num_pairs < 0 || wx_size < 0) return fail(PICCA_E_VALUE, "negative size");
    const int64_t nf = ctx->num_forests, npix = ctx->num_pixels;
    if (!zev1 || !zev2 || !wx_offsets || (wx_size > 0 && !wx))
        return fail(PICCA_E_VALUE, "input arrays must not be NULL");
    if (num_pairs > 0 && (!pair_a || !pair_b || !pair_ang))
        return fail(PICCA_E_VALUE, "pair arrays must not be NULL");
    if (!weights_wick || !num_pairs_wick || !t1 || !t2 || !t3)
        return fail(PICCA_E_VALUE, "output arrays must not be NULL");
    const std::vector<int64_t> &off = ctx->h_offsets;
    for (int64_t k = 0; k < num_pairs; ++k) {
        if (pair_a[k] < 0 || pair_a[k] >= nf || pair_b[k] < 0 || pair_b[k] >= nf)
            return fail(PICCA_E_VALUE, "pair index out of range of the forest store");
        if (!(pair_ang[k] >= 0.0) || !(pair_ang[k] <= M_PI))
            return fail(PICCA_E_VALUE, "pair angle outside [0, pi]");
        for (int32_t f : {pair_a[k], pair_b[k]}) {
            const int64_t n = off[f + 1] - off[f];
            if (n >= (1 << 20)) return fail(PICCA_E_VALUE, "forest longer than 2^20 pixels");
            if (wx_offsets[f] < 0 || wx_offsets[f] + n * n > wx_size)
                return fail(PICCA_E_VALUE, "weighted 1-D correlation missing for a forest");
        }
    }
    for (int64_t k = 0; k < npix; ++k)
        if (!std::isfinite(zev1[k]) || !std::isfinite(zev2[k]))
            return fail(PICCA_E_VALUE, "evolution factors must be finite");

    picca::WickConst C;
    std::memset(&C, 0, sizeof(C));
    C.np = np;
    C.nt = nt;
    C.nb = np * nt;
    C.xcorr = params->x_correlation != 0;
    C.rpmin = params->r_par_min;
    C.rpmax = params->r_par_max;
    C.rtmax = params->r_trans_max;
    C.span = params->r_par_max - params->r_par_min;
    const int64_t nb = C.nb;
    if ((int64_t)(picca::wick_row_lds(C) + 1024) > ctx->lds_limit ||
        (int64_t)(picca::wick_list_lds(C) + 2048) > ctx->lds_limit)
        return fail(PICCA_E_VALUE, "too many bins for the workgroup's LDS");

    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    PICCA_HIP_CHECK(ctx->wk_zev1.upload(zev1, (size_t)npix, ctx->stream));
    PICCA_HIP_CHECK(ctx->wk_zev2.upload(zev2, (size_t)npix, ctx->stream));
    PICCA_HIP_CHECK(ctx->wk_wx.upload(wx, (size_t)wx_size, ctx->stream));
    PICCA_HIP_CHECK(ctx->wk_wxoff.upload(wx_offsets, (size_t)nf, ctx->stream));
    PICCA_HIP_CHECK(ctx->o_wdm.reserve((size_t)nb));
    PICCA_HIP_CHECK(ctx->wk_nb.reserve((size_t)nb));
    PICCA_HIP_CHECK(ctx->o_dmat.reserve((size_t)(3 * nb * nb)));
    PICCA_HIP_CHECK(ctx->wk_stats.reserve(picca::WICK_NSTATS));
    PICCA_HIP_CHECK(ctx->wk_ntasks.reserve(1));
    PICCA_HIP_CHECK(picca::launch_zero_u64((unsigned long long *)ctx->o_wdm.ptr, nb, ctx->stream));
    PICCA_HIP_CHECK(picca::launch_zero_u64(ctx->wk_nb.ptr, nb, ctx->stream));
    PICCA_HIP_CHECK(picca::launch_zero_u64((unsigned long long *)ctx->o_dmat.ptr, 3 * nb * nb,
                                           ctx->stream));
    PICCA_HIP_CHECK(picca::launch_zero_u64(ctx->wk_stats.ptr, picca::WICK_NSTATS, ctx->stream));
    // row prefix sums of the weighted 1-D correlation of every forest of the pairs
    std::vector<int64_t> c1off((size_t)nf, 0);
    std::vector<int32_t> forests;
    {
        std::vector<char> used((size_t)nf, 0);
        for (int64_t k = 0; k < num_pairs; ++k) used[(size_t)pair_a[k]] = used[(size_t)pair_b[k]] = 1;
        int64_t total = 0;
        for (int64_t f = 0; f < nf; ++f) {
            if (!used[(size_t)f]) continue;
            const int64_t n = off[f + 1] - off[f];
            c1off[(size_t)f] = total;
            total += n * (n + 1);
            forests.push_back((int32_t)f);
        }
        PICCA_HIP_CHECK(ctx->wk_c1.reserve((size_t)std::max<int64_t>(total, 1)));
    }
    PICCA_HIP_CHECK(ctx->wk_c1off.upload(c1off.data(), c1off.size(), ctx->stream));
    PICCA_HIP_CHECK(ctx->wk_forests.upload(forests.data(), forests.size(), ctx->stream));
    PICCA_HIP_CHECK(picca::launch_wick_prefix(ctx->wk_wx.ptr, ctx->wk_wxoff.ptr, ctx->off.ptr,
                                              ctx->wk_forests.ptr, (int32_t)forests.size(),
                                              ctx->wk_c1.ptr, ctx->wk_c1off.ptr, ctx->stream));
    picca::WickInputs I{ctx->off.ptr,      ctx->r.ptr,        ctx->w.ptr,
                        ctx->wk_zev1.ptr,  ctx->wk_zev2.ptr,  ctx->wk_wx.ptr,
                        ctx->wk_wxoff.ptr, ctx->wk_c1.ptr,    ctx->wk_c1off.ptr};
    picca::WickOut O{ctx->o_wdm.ptr, ctx->wk_nb.ptr, ctx->o_dmat.ptr, ctx->o_dmat.ptr + nb * nb,
                     ctx->o_dmat.ptr + 2 * nb * nb, ctx->wk_stats.ptr};

    // batches of pairs whose lists (n1 n2 entries each, an upper bound) fit 1 GiB (and their
    // segment lists, 16 bytes per entry at most, 2 GiB)
    const int64_t list_budget = (int64_t)1 << 27;
    double total_ms = 0.0;
    int64_t batches = 0;
    std::vector<double> cs;
    std::vector<int64_t> base;
    for (int64_t p0 = 0; p0 < num_pairs;) {
        int64_t p1 = p0, entries = 0, max_tasks = 0;
        base.clear();
        while (p1 < num_pairs && p1 - p0 < INT32_MAX) {
            const int64_t n1 = off[pair_a[p1] + 1] - off[pair_a[p1]];
            const int64_t n2 = off[pair_b[p1] + 1] - off[pair_b[p1]];
            if (p1 > p0 && entries + n1 * n2 > list_budget) break;
            base.push_back(entries);
            entries += n1 * n2;
            max_tasks += std::min<int64_t>(nb, n1 * n2);
            ++p1;
        }
        const int64_t count = p1 - p0;
        cs.assign((size_t)(2 * count), 0.0);
        for (int64_t k = 0; k < count; ++k) {
            cs[2 * k] = std::cos(pair_ang[p0 + k] / 2);
            cs[2 * k + 1] = std::sin(pair_ang[p0 + k] / 2);
        }
        PICCA_HIP_CHECK(ctx->dpa.upload(pair_a + p0, (size_t)count, ctx->stream));
        PICCA_HIP_CHECK(ctx->dpb.upload(pair_b + p0, (size_t)count, ctx->stream));
        PICCA_HIP_CHECK(ctx->dcs.upload(cs.data(), cs.size(), ctx->stream));
        PICCA_HIP_CHECK(ctx->wk_base.upload(base.data(), base.size(), ctx->stream));
        PICCA_HIP_CHECK(ctx->wk_list.reserve((size_t)std::max<int64_t>(entries, 1)));
        PICCA_HIP_CHECK(ctx->wk_seg.reserve((size_t)std::max<int64_t>(entries, 1)));
        PICCA_HIP_CHECK(ctx->wk_count.reserve((size_t)count));
        PICCA_HIP_CHECK(ctx->wk_segcount.reserve((size_t)count));
        PICCA_HIP_CHECK(ctx->wk_task.reserve((size_t)(4 * std::max<int64_t>(max_tasks, 1))));
        PICCA_HIP_CHECK(picca::launch_zero_u64(ctx->wk_ntasks.ptr, 1, ctx->stream));
        int32_t *tk = ctx->wk_task.ptr;
        const int64_t mt = std::max<int64_t>(max_tasks, 1);
        picca::WickPairs P{ctx->dpa.ptr, ctx->dpb.ptr, ctx->dcs.ptr, (int32_t)count};
        picca::WickWork W{ctx->wk_list.ptr,  ctx->wk_base.ptr,     ctx->wk_count.ptr,
                          ctx->wk_seg.ptr,   ctx->wk_segcount.ptr, tk,
                          tk + mt,           tk + 2 * mt,          tk + 3 * mt,
                          ctx->wk_ntasks.ptr, max_tasks};
        PICCA_HIP_CHECK(hipEventRecord(ctx->ev_start, ctx->stream));
        PICCA_HIP_CHECK(picca::launch_wick_list(C, I, P, W, O, ctx->stream));
        unsigned long long tasks = 0;
        PICCA_HIP_CHECK(hipMemcpyAsync(&tasks, ctx->wk_ntasks.ptr, sizeof(tasks),
                                       hipMemcpyDeviceToHost, ctx->stream));
        PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if ((int64_t)tasks > max_tasks)
            return fail(PICCA_E_RUNTIME, "Wick task list overflow");
        PICCA_HIP_CHECK(picca::launch_wick_rows(C, I, P, W, O, (int64_t)tasks, ctx->stream));
        PICCA_HIP_CHECK(hipEventRecord(ctx->ev_stop, ctx->stream));
        PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        float ms = 0.f;
        PICCA_HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
        total_ms += ms;
        ++batches;
        p0 = p1;
    }

    std::vector<double> h_w((size_t)nb), h_t((size_t)(3 * nb * nb));
    std::vector<unsigned long long> h_n((size_t)nb);
    unsigned long long h_stats[picca::WICK_NSTATS];
    PICCA_HIP_CHECK(hipMemcpyAsync(h_w.data(), ctx->o_wdm.ptr, sizeof(double) * nb,
                                   hipMemcpyDeviceToHost, ctx->stream));
    PICCA_HIP_CHECK(hipMemcpyAsync(h_n.data(), ctx->wk_nb.ptr, sizeof(unsigned long long) * nb,
                                   hipMemcpyDeviceToHost, ctx->stream));
    PICCA_HIP_CHECK(hipMemcpyAsync(h_t.data(), ctx->o_dmat.ptr, sizeof(double) * 3 * nb * nb,
                                   hipMemcpyDeviceToHost, ctx->stream));
    PICCA_HIP_CHECK(hipMemcpyAsync(h_stats, ctx->wk_stats.ptr, sizeof(h_stats),
                                   hipMemcpyDeviceToHost, ctx->stream));
    PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (int64_t k = 0; k < nb; ++k) {
        weights_wick[k] += h_w[k];
        num_pairs_wick[k] += (int64_t)h_n[k];
    }
    for (int64_t k = 0; k < nb * nb; ++k) {
        t1[k] += h_t[k];
        t2[k] += h_t[nb * nb + k];
        t3[k] += h_t[2 * nb * nb + k];
    }
    if (report) {
        report->in_range_pairs = (int64_t)h_stats[picca::WICK_STAT_IN_RANGE];
        // the row kernel walks every listed pair a against every b != a, so it sees each
        // unordered pair {a, b} twice; the reference's loop (py/picca/cf.py:1573-1600)
        // visits index1 < index2 only, one product per unordered pair
        report->products = (int64_t)(h_stats[picca::WICK_STAT_PRODUCTS] / 2);
        report->overflow_pairs = (int64_t)h_stats[picca::WICK_STAT_OVERFLOW];
        report->batches = batches;
        report->kernel_ms = total_ms;
    }
    return PICCA_OK;
}

}  // extern "C"
