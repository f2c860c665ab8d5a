// C-ABI of the engine (declared in include/picca_hip.h): device context, resident forest
// store and pair lists, launches, result download and error reporting.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "context.h"
#include "xi_const.h"

namespace picca_host {

std::string &last_error() {
    static thread_local std::string message;
    return message;
}

}  // namespace picca_host

namespace {

bool forest_monotone(const double *x, int64_t n) {
    for (int64_t k = 1; k < n; ++k)
        if (!(x[k] >= x[k - 1])) return false;
    return true;
}

int build_xi_const(const picca_cf_params *p, picca::XiConst &C, int &mode) {
    const char *err = picca::fill_xi_const(p, C, mode);
    return err ? fail(PICCA_E_VALUE, err) : PICCA_OK;
}

}  // namespace

extern "C" {

int picca_abi_version(void) { return PICCA_HIP_ABI_VERSION; }

const char *picca_last_error(void) { return picca_host::last_error().c_str(); }

int picca_device_count(int32_t *count) {
    if (!count) return fail(PICCA_E_VALUE, "count is NULL");
    int n = 0;
    hipError_t err = hipGetDeviceCount(&n);
    if (err != hipSuccess) {
        *count = 0;
        return fail(PICCA_E_HIP, std::string("hipGetDeviceCount: ") + hipGetErrorString(err));
    }
    *count = n;
    return PICCA_OK;
}

int picca_device_info(int32_t device, char *name, int32_t name_len, int32_t *num_cus,
                      int64_t *lds_per_block, int64_t *global_mem) {
    hipDeviceProp_t prop;
    PICCA_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    if (name && name_len > 0) {
        std::snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (num_cus) *num_cus = prop.multiProcessorCount;
    if (lds_per_block) *lds_per_block = (int64_t)prop.maxSharedMemoryPerMultiProcessor;
    if (global_mem) *global_mem = (int64_t)prop.totalGlobalMem;
    return PICCA_OK;
}

int picca_create(int32_t device, picca_ctx **out) {
    if (!out) return fail(PICCA_E_VALUE, "out is NULL");
    *out = nullptr;
    int count = 0;
    PICCA_HIP_CHECK(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return fail(PICCA_E_VALUE, "device " + std::to_string(device) + " out of range");
    PICCA_HIP_CHECK(hipSetDevice(device));
    picca_ctx *ctx = new picca_ctx();
    ctx->device = device;
    hipDeviceProp_t prop;
    hipError_t err = hipGetDeviceProperties(&prop, device);
    if (err == hipSuccess) err = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipEventCreate(&ctx->ev_start);
    if (err == hipSuccess) err = hipEventCreate(&ctx->ev_stop);
    if (err == hipSuccess) err = hipEventCreate(&ctx->ev_mid);
    if (err != hipSuccess) {
        picca_destroy(ctx);
        return fail(PICCA_E_HIP, std::string("picca_create: ") + hipGetErrorString(err));
    }
    ctx->num_cus = prop.multiProcessorCount;
    ctx->lds_limit = (int64_t)prop.maxSharedMemoryPerMultiProcessor;
    // this GPU's copy of the no-allocation scratch array (device `device` is current)
    err = picca::xi_scratch_address(&ctx->scratch);
    if (err != hipSuccess) {
        picca_destroy(ctx);
        return fail(PICCA_E_HIP, std::string("picca_create: scratch address: ") +
                                     hipGetErrorString(err));
    }
    *out = ctx;
    return PICCA_OK;
}

int picca_destroy(picca_ctx *ctx) {
    if (!ctx) return PICCA_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->xi_graph_exec) (void)hipGraphExecDestroy(ctx->xi_graph_exec);
    if (ctx->xi_graph) (void)hipGraphDestroy(ctx->xi_graph);
    if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
    if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
    if (ctx->ev_mid) (void)hipEventDestroy(ctx->ev_mid);
    for (hipEvent_t e : ctx->dmat_events) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PICCA_OK;
}

int picca_set_forests(picca_ctx *ctx, int64_t num_forests, const int64_t *offsets,
                      const double *z, const double *r_comov, const double *dist_m,
                      const double *weights, const double *delta, const double *log_lambda,
                      const double *z_qso, const int32_t *order) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    if (num_forests < 0 || num_forests > INT32_MAX - 1)
        return fail(PICCA_E_VALUE, "num_forests out of range");
    if (!offsets || !z || !r_comov || !dist_m || !weights || !delta || !z_qso)
        return fail(PICCA_E_VALUE, "forest arrays must not be NULL");
    if (offsets[0] != 0) return fail(PICCA_E_VALUE, "offsets[0] must be 0");
    for (int64_t k = 0; k < num_forests; ++k) {
        if (offsets[k + 1] < offsets[k]) return fail(PICCA_E_VALUE, "offsets must be non-decreasing");
        if (offsets[k + 1] - offsets[k] > INT32_MAX / 2)
            return fail(PICCA_E_VALUE, "forest too long");
    }
    const int64_t npix = offsets[num_forests];
    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));

    std::vector<double> dw((size_t)npix);
    parallel_for(npix, [&](int64_t lo, int64_t hi) {
        for (int64_t k = lo; k < hi; ++k) dw[k] = delta[k] * weights[k];
    });
    std::vector<uint8_t> mono((size_t)std::max<int64_t>(num_forests, 1));
    parallel_for(num_forests, [&](int64_t lo, int64_t hi) {
        for (int64_t k = lo; k < hi; ++k) {
            const int64_t o = offsets[k], n = offsets[k + 1] - offsets[k];
            mono[k] = (uint8_t)((forest_monotone(r_comov + o, n) ? 1 : 0) |
                                (forest_monotone(dist_m + o, n) ? 2 : 0));
        }
    });
    double max_r = 0, max_d = 0;
    bool finite = true;
    for (int64_t k = 0; k < npix; ++k) {
        max_r = std::max(max_r, std::fabs(r_comov[k]));
        max_d = std::max(max_d, std::fabs(dist_m[k]));
        finite = finite && std::isfinite(z[k]) && std::isfinite(weights[k]) &&
                 std::isfinite(delta[k]) && std::isfinite(dw[k]);
    }
    if (!finite || !std::isfinite(max_r) || !std::isfinite(max_d))
        return fail(PICCA_E_VALUE, "forest arrays must be finite (z, r_comov, dist_m, "
                                   "weights, delta, delta*weights)");
    ctx->max_abs_r = max_r;
    ctx->max_abs_d = max_d;
    ctx->have_result = false;
    ctx->num_pairs = 0;
    ctx->h_offsets.assign(offsets, offsets + num_forests + 1);
    ctx->h_dmin.assign((size_t)num_forests, 0.0);
    ctx->h_dmax.assign((size_t)num_forests, 0.0);
    parallel_for(num_forests, [&](int64_t lo, int64_t hi) {
        for (int64_t k = lo; k < hi; ++k) {
            double a = INFINITY, b = -INFINITY;
            for (int64_t q = offsets[k]; q < offsets[k + 1]; ++q) {
                a = std::min(a, dist_m[q]);
                b = std::max(b, dist_m[q]);
            }
            ctx->h_dmin[k] = a;
            ctx->h_dmax[k] = b;
        }
    });
    PICCA_HIP_CHECK(ctx->off.upload(offsets, (size_t)num_forests + 1, ctx->stream));
    // the pair kernel reads pixel batches of four, up to three past a forest's end
    const size_t pad = picca::PIXEL_PAD;
    PICCA_HIP_CHECK(ctx->z.upload_padded(z, (size_t)npix, pad, ctx->stream));
    PICCA_HIP_CHECK(ctx->r.upload_padded(r_comov, (size_t)npix, pad, ctx->stream));
    PICCA_HIP_CHECK(ctx->d.upload_padded(dist_m, (size_t)npix, pad, ctx->stream));
    PICCA_HIP_CHECK(ctx->w.upload_padded(weights, (size_t)npix, pad, ctx->stream));
    PICCA_HIP_CHECK(ctx->dw.upload_padded(dw.data(), (size_t)npix, pad, ctx->stream));
    {
        std::vector<picca::PixelRecord> records((size_t)std::max<int64_t>(npix, 1));
        parallel_for(npix, [&](int64_t lo, int64_t hi) {
            for (int64_t k = lo; k < hi; ++k)
                records[k] = picca::PixelRecord{r_comov[k], dist_m[k], z[k], weights[k], dw[k]};
        });
        PICCA_HIP_CHECK(ctx->pix.upload_padded(records.data(), (size_t)npix, pad, ctx->stream));
        PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    }
    {
        // run walk of the xi kernel: (r, d) pairs, prefix sums (L + 1 entries per forest)
        // and the smallest pixel steps per forest
        std::vector<double2> rd((size_t)std::max<int64_t>(npix, 1));
        std::vector<picca::PrefixRecord> pre((size_t)(npix + num_forests));
        std::vector<double2> spc((size_t)std::max<int64_t>(num_forests, 1));
        parallel_for(npix, [&](int64_t lo, int64_t hi) {
            for (int64_t k = lo; k < hi; ++k) rd[k] = make_double2(r_comov[k], dist_m[k]);
        });
        parallel_for(num_forests, [&](int64_t lo, int64_t hi) {
            for (int64_t k = lo; k < hi; ++k) {
                const int64_t o = offsets[k], n = offsets[k + 1] - offsets[k];
                double sr, sd;
                picca::build_walk_forest(r_comov + o, dist_m + o, z + o, weights + o,
                                         dw.data() + o, n, pre.data() + o + k, sr, sd);
                spc[k] = make_double2(sr, sd);
            }
        });
        PICCA_HIP_CHECK(ctx->rd.upload_padded(rd.data(), (size_t)npix, pad, ctx->stream));
        PICCA_HIP_CHECK(ctx->pre.upload_padded(pre.data(), (size_t)(npix + num_forests), pad,
                                               ctx->stream));
        PICCA_HIP_CHECK(ctx->spc.upload(spc.data(), (size_t)num_forests, ctx->stream));
        PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        // typical pixel spacing of the store, for the automatic schedule of the xi kernel
        std::vector<double> steps;
        steps.reserve((size_t)num_forests);
        for (int64_t k = 0; k < num_forests; ++k) {
            const int64_t o = offsets[k], n = offsets[k + 1] - offsets[k];
            if (n >= 2) steps.push_back((r_comov[o + n - 1] - r_comov[o]) / (double)(n - 1));
        }
        ctx->median_step_r = 0.0;
        if (!steps.empty()) {
            std::nth_element(steps.begin(), steps.begin() + steps.size() / 2, steps.end());
            ctx->median_step_r = steps[steps.size() / 2];
        }
    }
    PICCA_HIP_CHECK(ctx->zq.upload(z_qso, (size_t)num_forests, ctx->stream));
    PICCA_HIP_CHECK(ctx->mono.upload(mono.data(), (size_t)num_forests, ctx->stream));
    ctx->has_log_lambda = log_lambda != nullptr;
    if (log_lambda) {
        PICCA_HIP_CHECK(ctx->ll.upload(log_lambda, (size_t)npix, ctx->stream));
        // per forest, sequential sums in pixel order as the reference's loops:
        // sum w, weighted mean log_lambda, dll = log_lambda - mean, sum w dll^2
        std::vector<double> dll((size_t)std::max<int64_t>(npix, 1));
        std::vector<double> sum_w((size_t)std::max<int64_t>(num_forests, 1));
        std::vector<double> sw_dll((size_t)std::max<int64_t>(num_forests, 1));
        parallel_for(num_forests, [&](int64_t lo, int64_t hi) {
            for (int64_t k = lo; k < hi; ++k) {
                const int64_t o = offsets[k], e = offsets[k + 1];
                double sw = 0, acc = 0;
                for (int64_t q = o; q < e; ++q) sw += weights[q];
                for (int64_t q = o; q < e; ++q) acc += log_lambda[q] * weights[q];
                const double mean = acc / sw;
                double s2 = 0;
                for (int64_t q = o; q < e; ++q) {
                    dll[q] = log_lambda[q] - mean;
                    s2 += weights[q] * (dll[q] * dll[q]);
                }
                sum_w[k] = sw;
                sw_dll[k] = s2;
            }
        });
        PICCA_HIP_CHECK(ctx->dll.upload_padded(dll.data(), (size_t)npix, pad, ctx->stream));
        PICCA_HIP_CHECK(ctx->sum_w.upload(sum_w.data(), (size_t)num_forests, ctx->stream));
        PICCA_HIP_CHECK(ctx->sw_dll.upload(sw_dll.data(), (size_t)num_forests, ctx->stream));
        PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        ctx->h_z.assign(z, z + npix);
    } else {
        ctx->h_z.clear();
    }
    ctx->has_order = order != nullptr;
    if (order) PICCA_HIP_CHECK(ctx->order.upload(order, (size_t)num_forests, ctx->stream));
    if (order)
        ctx->h_order.assign(order, order + num_forests);
    else
        ctx->h_order.clear();
    PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->num_forests = num_forests;
    ctx->num_pixels = npix;
    for (picca_ctx::MetalSlot &m : ctx->metal) m.set = false;   // tables of the old store
    return PICCA_OK;
}

int picca_xi_set_pairs(picca_ctx *ctx, int64_t num_pairs, const int32_t *pair_a,
                       const int32_t *pair_b, const double *pair_ang,
                       const uint8_t *pair_flags, const int32_t *pair_row, int32_t num_rows) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    if (num_pairs < 0 || num_rows < 0) return fail(PICCA_E_VALUE, "negative size");
    if (num_pairs > 0 && (!pair_a || !pair_b || !pair_ang || !pair_flags || !pair_row))
        return fail(PICCA_E_VALUE, "pair arrays must not be NULL");
    const int64_t nf = ctx->num_forests;
    // validation: indices, rows, contiguity of rows
    std::vector<uint8_t> row_seen((size_t)std::max(num_rows, 1), 0);
    for (int64_t k = 0; k < num_pairs; ++k) {
        if (pair_a[k] < 0 || pair_a[k] >= nf || pair_b[k] < 0 || pair_b[k] >= nf)
            return fail(PICCA_E_VALUE, "pair index out of range of the forest store");
        const int32_t row = pair_row[k];
        if (row < 0 || row >= num_rows) return fail(PICCA_E_VALUE, "pair row out of range");
        if (k == 0 || row != pair_row[k - 1]) {
            if (row_seen[row]) return fail(PICCA_E_VALUE, "pairs of one row must be contiguous");
            row_seen[row] = 1;
        }
        if (!(pair_ang[k] >= 0.0) || !(pair_ang[k] <= M_PI))
            return fail(PICCA_E_VALUE, "pair angle outside [0, pi]");
    }
    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));

    std::vector<double> cs((size_t)std::max<int64_t>(2 * num_pairs, 1));
    parallel_for(num_pairs, [&](int64_t lo, int64_t hi) {
        for (int64_t k = lo; k < hi; ++k) {
            cs[2 * k] = std::cos(pair_ang[k] / 2);
            cs[2 * k + 1] = std::sin(pair_ang[k] / 2);
        }
    });
    // groups: <= XI_GROUP_PAIRS consecutive pairs of one row
    ctx->h_group_begin.clear();
    ctx->h_group_row.clear();
    for (int64_t k = 0; k < num_pairs;) {
        int64_t end = k + 1;
        while (end < num_pairs && pair_row[end] == pair_row[k] && end - k < picca::XI_GROUP_PAIRS)
            ++end;
        ctx->h_group_begin.push_back(k);
        ctx->h_group_row.push_back(pair_row[k]);
        k = end;
    }
    ctx->h_group_begin.push_back(num_pairs);

    ctx->have_result = false;
    PICCA_HIP_CHECK(ctx->pa.upload(pair_a, (size_t)num_pairs, ctx->stream));
    PICCA_HIP_CHECK(ctx->pb.upload(pair_b, (size_t)num_pairs, ctx->stream));
    PICCA_HIP_CHECK(ctx->pang.upload(pair_ang, (size_t)num_pairs, ctx->stream));
    PICCA_HIP_CHECK(ctx->pcs.upload(cs.data(), (size_t)(2 * num_pairs), ctx->stream));
    PICCA_HIP_CHECK(ctx->pflags.upload(pair_flags, (size_t)num_pairs, ctx->stream));
    PICCA_HIP_CHECK(ctx->pchunks.reserve((size_t)std::max<int64_t>(2 * num_pairs, 2)));
    PICCA_HIP_CHECK(ctx->gbegin.upload(ctx->h_group_begin.data(), ctx->h_group_begin.size(),
                                       ctx->stream));
    PICCA_HIP_CHECK(ctx->grow.upload(ctx->h_group_row.data(), ctx->h_group_row.size(),
                                     ctx->stream));
    PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->num_pairs = num_pairs;
    ctx->num_rows = num_rows;
    return PICCA_OK;
}

namespace {

// picca_xi_execute's work with the output rows, the work-queue head and the counters in the
// given device buffers (out: num_rows * np*nt records; queue: 1 word; stats: XI_NSTATS
// words), all cleared here by kernels.  Allocates nothing.
int xi_execute_into(picca_ctx *ctx, const picca::XiConst &C0, int mode,
                    const picca_cf_params *params, picca::XiBin *out, unsigned long long *queue,
                    unsigned long long *stats_dev, bool allow_graph, picca_xi_report *report) {
    picca::XiConst C = C0;
    // the specialised kernel: standard binning, auto-correlation, r_par_min == 0, no cuts
    const bool general = !picca::xi_specialised(params, mode);
    // the automatic schedule keeps the pair loop's kernel (no walk) when the store's pixels
    // are too dense for runs of XI_WALK_MIN_DENSITY pixels per r_par bin
    const bool dense = ctx->median_step_r > 0 &&
                       C.width_p < picca::XI_WALK_MIN_DENSITY * ctx->median_step_r;
    picca::apply_xi_schedule(C, ctx->xi_schedule == PICCA_XI_SCHEDULE_AUTO && dense
                                    ? PICCA_XI_SCHEDULE_PAIRS
                                    : ctx->xi_schedule);
    // LDS: the histogram in its padded layout ((np + 1) x (nt + 1) bins of 48 B) when it
    // fits, plus one staging tile per wave for the loops that stage pixels (the fast loop
    // of the specialised kernel stages none)
    const size_t static_lds = 1024;
    const size_t waves = picca::XI_BLOCK / 64;
    const size_t hist_bytes = (size_t)(C.np + 1) * (size_t)C.ntp * picca::XI_HIST_BIN_BYTES;
    auto lds_bytes = [&](bool with_hist, int tile) {
        return (with_hist ? hist_bytes : 0) + waves * (size_t)(tile + 1) * picca::XI_TILE_PIXEL_BYTES;
    };
    bool lds_hist = false;
    size_t lds = 0;
    if (!general) {
        if ((int64_t)(lds_bytes(true, 0) + static_lds) <= ctx->lds_limit) {
            lds_hist = true;
            C.tile_len = 0;
            lds = lds_bytes(true, 0);
        }
    } else {
        for (int tile = 64; tile >= 16 && !lds_hist; tile /= 2) {
            if ((int64_t)(lds_bytes(true, tile) + static_lds) <= ctx->lds_limit) {
                lds_hist = true;
                C.tile_len = tile;
                lds = lds_bytes(true, tile);
            }
        }
    }
    if (!lds_hist) {  // histogram too large for LDS: runs go straight to the HBM rows
        C.tile_len = 64;
        lds = lds_bytes(false, 64);
    }
    const int64_t nrows = ctx->num_rows;
    const int64_t nout = nrows * C.nb;

    picca::ForestView F{ctx->pix.ptr, ctx->z.ptr, ctx->r.ptr, ctx->d.ptr, ctx->w.ptr, ctx->dw.ptr,
                        ctx->ll.ptr, ctx->off.ptr, ctx->zq.ptr, ctx->order.ptr, ctx->mono.ptr,
                        ctx->rd.ptr, ctx->pre.ptr, ctx->spc.ptr};
    picca::PairView P{ctx->pa.ptr, ctx->pb.ptr, ctx->pang.ptr, ctx->pcs.ptr, ctx->pflags.ptr,
                      ctx->pchunks.ptr};
    const int64_t ngroups = (int64_t)ctx->h_group_row.size();
    picca::GroupView G{ctx->gbegin.ptr, ctx->grow.ptr, ngroups};
    picca::XiOut O{out, queue, stats_dev};
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(ctx->num_cus, ngroups));
    const bool xcorr = params->x_correlation != 0;
    // the launch sequence: output, queue head and counters cleared by kernels (no memset
    // nodes in a graph), the chunk ranges of the pairs, then the pair kernel
    auto enqueue = [&](hipStream_t st) -> hipError_t {
        hipError_t e = picca::launch_zero_u64(reinterpret_cast<unsigned long long *>(out),
                                              nout * (int64_t)(sizeof(picca::XiBin) / 8), st);
        if (e == hipSuccess) e = picca::launch_zero_u64(queue, 1, st);
        if (e == hipSuccess) e = picca::launch_zero_u64(stats_dev, picca::XI_NSTATS, st);
        if (e == hipSuccess && ngroups > 0)
            e = picca::launch_xi_chunks(mode, C, F, P, ctx->num_pairs, st);
        if (e == hipSuccess && ngroups > 0)
            e = picca::launch_xi(mode, xcorr, general, lds_hist, C, F, P, G, O, grid, lds, st);
        return e;
    };
    PICCA_HIP_CHECK(hipEventRecord(ctx->ev_start, ctx->stream));
    if (allow_graph && ctx->xi_use_graph) {
        // one graph per set of launch arguments (everything the kernels take by value)
        std::vector<unsigned char> key;
        auto put = [&key](const void *v, size_t n) {
            const unsigned char *b = (const unsigned char *)v;
            key.insert(key.end(), b, b + n);
        };
        const int32_t flags[4] = {mode, xcorr ? 1 : 0, general ? 1 : 0, lds_hist ? 1 : 0};
        put(&C, sizeof(C));
        put(&F, sizeof(F));
        put(&P, sizeof(P));
        put(&G, sizeof(G));
        put(&O, sizeof(O));
        put(&grid, sizeof(grid));
        put(&lds, sizeof(lds));
        put(&nout, sizeof(nout));
        put(flags, sizeof(flags));
        if (!ctx->xi_graph_exec || key != ctx->xi_graph_key) {
            if (ctx->xi_graph_exec) PICCA_HIP_CHECK(hipGraphExecDestroy(ctx->xi_graph_exec));
            if (ctx->xi_graph) PICCA_HIP_CHECK(hipGraphDestroy(ctx->xi_graph));
            ctx->xi_graph_exec = nullptr;
            ctx->xi_graph = nullptr;
            ctx->xi_graph_key.clear();
            PICCA_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
            const hipError_t e = enqueue(ctx->stream);
            hipGraph_t graph = nullptr;
            const hipError_t e_end = hipStreamEndCapture(ctx->stream, &graph);
            if (e != hipSuccess || e_end != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                return fail(PICCA_E_HIP, std::string("xi graph capture: ") +
                                             hipGetErrorString(e != hipSuccess ? e : e_end));
            }
            ctx->xi_graph = graph;
            PICCA_HIP_CHECK(hipGraphInstantiate(&ctx->xi_graph_exec, graph, nullptr, nullptr, 0));
            ctx->xi_graph_key = key;
        }
        PICCA_HIP_CHECK(hipGraphLaunch(ctx->xi_graph_exec, ctx->stream));
    } else {
        PICCA_HIP_CHECK(enqueue(ctx->stream));
    }
    PICCA_HIP_CHECK(hipEventRecord(ctx->ev_stop, ctx->stream));
    unsigned long long stats[picca::XI_NSTATS];
    PICCA_HIP_CHECK(hipMemcpyAsync(stats, stats_dev, sizeof(stats), hipMemcpyDeviceToHost,
                                   ctx->stream));
    PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    PICCA_HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    if (report) {
        report->accepted_pairs = (int64_t)stats[picca::XI_STAT_ACCEPTED];
        report->edge_pairs = (int64_t)stats[picca::XI_STAT_EDGE];
        report->overflow_pairs = (int64_t)stats[picca::XI_STAT_OVERFLOW];
        report->visited_pairs = (int64_t)stats[picca::XI_STAT_VISITED];
        report->wave_steps = (int64_t)stats[picca::XI_STAT_STEPS];
        report->tasks = (int64_t)stats[picca::XI_STAT_TASKS];
        report->empty_tasks = (int64_t)stats[picca::XI_STAT_EMPTY_TASKS];
        report->runs = (int64_t)stats[picca::XI_STAT_RUNS];
        report->probes = (int64_t)stats[picca::XI_STAT_PROBES];
        report->walk_tasks = (int64_t)stats[picca::XI_STAT_WALK_TASKS];
        report->kernel_ms = (double)ms;
        report->launches = ngroups > 0 ? 1 : 0;
    }
    return PICCA_OK;
}

}  // namespace

int picca_xi_execute(picca_ctx *ctx, const picca_cf_params *params, picca_xi_report *report) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    picca::XiConst C;
    int mode = 0;
    int status = build_xi_const(params, C, mode);
    if (status != PICCA_OK) return status;
    {
        const char *err = picca::widen_xi_margins(C, ctx->max_abs_r, ctx->max_abs_d);
        if (err) return fail(PICCA_E_VALUE, err);
    }
    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    ctx->have_result = false;
    PICCA_HIP_CHECK(ctx->o_bins.reserve((size_t)ctx->num_rows * (size_t)C.nb));
    PICCA_HIP_CHECK(ctx->queue.reserve(1));
    PICCA_HIP_CHECK(ctx->stats.reserve(picca::XI_NSTATS));
    status = xi_execute_into(ctx, C, mode, params, ctx->o_bins.ptr, ctx->queue.ptr,
                             ctx->stats.ptr, true, report);
    if (status == PICCA_OK) {
        ctx->out_nb = C.nb;
        ctx->out_rows = ctx->num_rows;
        ctx->have_result = true;
    }
    return status;
}

int picca_xi_execute_noalloc(picca_ctx *ctx, const picca_cf_params *params, void *out_bins,
                             int64_t out_capacity, picca_xi_report *report) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    picca::XiConst C;
    int mode = 0;
    int status = build_xi_const(params, C, mode);
    if (status != PICCA_OK) return status;
    {
        const char *err = picca::widen_xi_margins(C, ctx->max_abs_r, ctx->max_abs_d);
        if (err) return fail(PICCA_E_VALUE, err);
    }
    const int64_t need = (int64_t)ctx->num_rows * C.nb;
    if (need > 0 && !out_bins) return fail(PICCA_E_VALUE, "out_bins is NULL");
    if (out_capacity < need)
        return fail(PICCA_E_VALUE, "out_bins holds " + std::to_string(out_capacity) +
                                       " records, num_rows * np * nt = " +
                                       std::to_string(need) + " are needed");
    if (!ctx->scratch)
        return fail(PICCA_E_RUNTIME, "the device scratch array of this GPU is not available");
    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    // scratch: word 0 the work-queue head, words 1.. the counters
    return xi_execute_into(ctx, C, mode, params, (picca::XiBin *)out_bins, ctx->scratch,
                           ctx->scratch + 1, false, report);
}

int picca_xi_use_graph(picca_ctx *ctx, int32_t on) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    ctx->xi_use_graph = on != 0;
    return PICCA_OK;
}

int picca_dmat_last_report(picca_ctx *ctx, picca_dmat_report *report) {
    if (!ctx || !report) return fail(PICCA_E_VALUE, "ctx or report is NULL");
    *report = ctx->dmat_report;
    return PICCA_OK;
}

int picca_xi_set_schedule(picca_ctx *ctx, int32_t schedule) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    if (schedule != PICCA_XI_SCHEDULE_AUTO && schedule != PICCA_XI_SCHEDULE_PAIRS &&
        schedule != PICCA_XI_SCHEDULE_RUNS)
        return fail(PICCA_E_VALUE, "unknown xi schedule");
    ctx->xi_schedule = schedule;
    return PICCA_OK;
}

int picca_xi_download(picca_ctx *ctx, double *weights, double *xi, double *r_par,
                      double *r_trans, double *z, int64_t *num_pairs) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    if (!ctx->have_result) return fail(PICCA_E_RUNTIME, "no xi result on the device");
    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t n = (size_t)ctx->out_rows * (size_t)ctx->out_nb;
    std::vector<picca::XiBin> host(std::max<size_t>(n, 1));
    if (n > 0)
        PICCA_HIP_CHECK(hipMemcpyAsync(host.data(), ctx->o_bins.ptr, n * sizeof(picca::XiBin),
                                       hipMemcpyDeviceToHost, ctx->stream));
    PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (size_t k = 0; k < n; ++k) {
        const picca::XiBin &b = host[k];
        if (weights) weights[k] = b.we;
        if (xi) xi[k] = b.xi;
        if (r_par) r_par[k] = b.rp;
        if (r_trans) r_trans[k] = b.rt;
        if (z) z[k] = b.z;
        if (num_pairs) num_pairs[k] = (int64_t)b.nb;
    }
    return PICCA_OK;
}

// Claim order of the cross-correlation's groups: runs of consecutive groups of one output
// row, each holding about half of the work still unclaimed over the grid (guided
// schedule, work = pixels of the forest x quasars of the group).  A workgroup moves its
// LDS histogram to HBM when its row changes, so whole-row claims early in the launch
// replace one flush per group by one per row, and the small claims at the end keep the
// grid balanced.  Returns the first group of each claim and, last, the group count.
static std::vector<int64_t> xcf_chunks(const std::vector<int64_t> &begin,
                                       const std::vector<int32_t> &forest,
                                       const std::vector<int32_t> &row,
                                       const std::vector<int64_t> &offsets, int grid) {
    const int64_t ngroups = (int64_t)forest.size();
    std::vector<double> work((size_t)ngroups);
    double total = 0;
    for (int64_t g = 0; g < ngroups; ++g) {
        const int32_t f = forest[(size_t)g];
        work[(size_t)g] = (double)(offsets[(size_t)f + 1] - offsets[(size_t)f]) *
                          (double)(begin[(size_t)g + 1] - begin[(size_t)g]);
        total += work[(size_t)g];
    }
    std::vector<int64_t> chunk;
    double left = total;
    const double share = 1.0 / (2.0 * std::max(grid, 1));
    for (int64_t g = 0; g < ngroups;) {
        chunk.push_back(g);
        const double target = left * share;
        double w = work[(size_t)g];
        int64_t e = g + 1;
        while (e < ngroups && row[(size_t)e] == row[(size_t)g] && w + work[(size_t)e] <= target)
            w += work[(size_t)e++];
        left -= w;
        g = e;
    }
    chunk.push_back(ngroups);
    return chunk;
}

int picca_xcf_set_pairs(picca_ctx *ctx, int64_t num_pairs, const int32_t *pair_forest,
                        const double *obj_r, const double *obj_d, const double *obj_z,
                        const double *obj_w, const double *pair_ang, const int32_t *pair_row,
                        int32_t num_rows, int32_t angular) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    if (num_pairs < 0 || num_rows < 0) return fail(PICCA_E_VALUE, "negative size");
    if (num_pairs > 0 && (!pair_forest || !obj_r || !obj_d || !obj_z || !obj_w || !pair_ang ||
                          !pair_row))
        return fail(PICCA_E_VALUE, "pair arrays must not be NULL");
    const int64_t nf = ctx->num_forests;
    // groups: maximal runs of pairs of one forest; rows contiguous
    std::vector<int64_t> begin;
    std::vector<int32_t> forest, row;
    std::vector<uint8_t> row_seen((size_t)std::max(num_rows, 1), 0);
    for (int64_t k = 0; k < num_pairs; ++k) {
        if (pair_forest[k] < 0 || pair_forest[k] >= nf)
            return fail(PICCA_E_VALUE, "pair forest out of range of the forest store");
        if (pair_row[k] < 0 || pair_row[k] >= num_rows)
            return fail(PICCA_E_VALUE, "pair row out of range");
        if (!std::isfinite(obj_r[k]) || !std::isfinite(obj_d[k]) || !std::isfinite(obj_z[k]) ||
            !std::isfinite(obj_w[k]) || !(pair_ang[k] >= 0.0) || !(pair_ang[k] <= M_PI))
            return fail(PICCA_E_VALUE, "quasar values must be finite, angles in [0, pi]");
        const bool new_group = k == 0 || pair_forest[k] != pair_forest[k - 1] ||
                               pair_row[k] != pair_row[k - 1];
        if (new_group) {
            if (k > 0 && pair_row[k] != pair_row[k - 1]) {
                if (row_seen[pair_row[k]])
                    return fail(PICCA_E_VALUE, "pairs of one row must be contiguous");
            }
            row_seen[pair_row[k]] = 1;
            begin.push_back(k);
            forest.push_back(pair_forest[k]);
            row.push_back(pair_row[k]);
        }
    }
    begin.push_back(num_pairs);
    // records; within a group sorted by the quasar's r (the kernel's pruning walks windows
    // of neighbouring r), which only reorders the additions of the sums
    std::vector<picca::XcfPairRecord> rec((size_t)std::max<int64_t>(num_pairs, 1));
    const int64_t ngroups = (int64_t)forest.size();
    ctx->hx_rt_lo.assign((size_t)num_pairs, 0.0);
    ctx->hx_rt_hi.assign((size_t)num_pairs, 0.0);
    parallel_for(ngroups, [&](int64_t lo, int64_t hi) {
        std::vector<int64_t> idx;
        for (int64_t g = lo; g < hi; ++g) {
            const int64_t b = begin[g], e = begin[g + 1];
            idx.resize((size_t)(e - b));
            for (int64_t k = b; k < e; ++k) idx[k - b] = k;
            if (!angular)
                std::stable_sort(idx.begin(), idx.end(),
                                 [&](int64_t x, int64_t y) { return obj_r[x] < obj_r[y]; });
            for (int64_t k = b; k < e; ++k) {
                const int64_t s = idx[k - b];
                picca::XcfPairRecord &r = rec[k];
                r.r = obj_r[s];
                r.d = obj_d[s];
                r.z = obj_z[s];
                r.w = obj_w[s];
                if (angular) {
                    r.c = pair_ang[s];
                    r.s = 0.0;
                } else {
                    r.c = std::cos(pair_ang[s] / 2);
                    r.s = std::sin(pair_ang[s] / 2);
                }
                const int32_t f = forest[g];
                ctx->hx_rt_lo[k] = (ctx->h_dmin[f] + r.d) * r.s;
                ctx->hx_rt_hi[k] = (ctx->h_dmax[f] + r.d) * r.s;
            }
        }
    });
    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->have_result = false;
    PICCA_HIP_CHECK(ctx->x_rec.upload(rec.data(), (size_t)num_pairs, ctx->stream));
    PICCA_HIP_CHECK(ctx->x_forest.upload(forest.data(), forest.size(), ctx->stream));
    PICCA_HIP_CHECK(ctx->x_row.upload(row.data(), row.size(), ctx->stream));
    PICCA_HIP_CHECK(ctx->x_begin.upload(begin.data(), begin.size(), ctx->stream));
    const std::vector<int64_t> chunk = xcf_chunks(begin, forest, row, ctx->h_offsets, ctx->num_cus);
    PICCA_HIP_CHECK(ctx->x_chunk.upload(chunk.data(), chunk.size(), ctx->stream));
    PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->xcf_chunks = (int64_t)chunk.size() - 1;
    ctx->xcf_pairs = num_pairs;
    ctx->xcf_rows = num_rows;
    ctx->xcf_groups = ngroups;
    ctx->hx_begin = std::move(begin);
    ctx->hx_forest = std::move(forest);
    ctx->hx_row = std::move(row);
    ctx->x_pairs_version += 1;
    return PICCA_OK;
}

// Row windows of the cross-correlation's fast loop: windows of `rows` r_trans rows, each
// with the pairs whose r_trans range (one row of margin on each side) reaches it.  A pair
// whose r_trans can round up to r_trans_max also goes to the first window: the reference
// bins it in bin 0 of the next r_par row.
static int xcf_plan(picca_ctx *ctx, const picca::XiConst &C, int rows) {
    if (ctx->x_plan_version == ctx->x_pairs_version && ctx->x_plan_nt == C.nt &&
        ctx->x_plan_rows == rows && ctx->x_plan_rtmax == C.rtmax)
        return PICCA_OK;
    const int nwin = (C.nt + rows - 1) / rows;
    ctx->x_windows.clear();
    for (int w = 0; w < nwin; ++w) ctx->x_windows.push_back(std::make_unique<picca_ctx::XcfWindow>());
    const int64_t ngroups = ctx->xcf_groups;
    const double inv_w = C.nt_d / C.rtmax;
    for (int w = 0; w < nwin; ++w) {
        picca_ctx::XcfWindow &win = *ctx->x_windows[(size_t)w];
        win.t0 = w * rows;
        win.rows = std::min(rows, C.nt - win.t0);
        const bool last = w == nwin - 1;
        std::vector<int64_t> idx, begin;
        std::vector<int32_t> forest, row;
        for (int64_t g = 0; g < ngroups; ++g) {
            const int64_t b = ctx->hx_begin[g], e = ctx->hx_begin[g + 1];
            const int64_t before = (int64_t)idx.size();
            for (int64_t k = b; k < e; ++k) {
                const double lo = ctx->hx_rt_lo[k] * inv_w * (1 - 1e-9) - 1.0;
                const double hi = ctx->hx_rt_hi[k] * inv_w * (1 + 1e-9) + 1.0;
                bool take = hi >= win.t0 && (last || lo < win.t0 + win.rows);
                if (w == 0 && hi >= C.nt_d) take = true;
                if (take) idx.push_back(k);
            }
            if ((int64_t)idx.size() > before) {
                begin.push_back(before);
                forest.push_back(ctx->hx_forest[g]);
                row.push_back(ctx->hx_row[g]);
            }
        }
        begin.push_back((int64_t)idx.size());
        win.groups = (int64_t)forest.size();
        const std::vector<int64_t> chunk = xcf_chunks(begin, forest, row, ctx->h_offsets,
                                                      ctx->num_cus);
        win.chunks = (int64_t)chunk.size() - 1;
        PICCA_HIP_CHECK(win.chunk.upload(chunk.data(), chunk.size(), ctx->stream));
        PICCA_HIP_CHECK(win.idx.upload(idx.data(), idx.size(), ctx->stream));
        PICCA_HIP_CHECK(win.begin.upload(begin.data(), begin.size(), ctx->stream));
        PICCA_HIP_CHECK(win.forest.upload(forest.data(), forest.size(), ctx->stream));
        PICCA_HIP_CHECK(win.row.upload(row.data(), row.size(), ctx->stream));
        PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    }
    ctx->x_plan_version = ctx->x_pairs_version;
    ctx->x_plan_nt = C.nt;
    ctx->x_plan_rows = rows;
    ctx->x_plan_rtmax = C.rtmax;
    return PICCA_OK;
}

int picca_xcf_execute(picca_ctx *ctx, const picca_cf_params *params, picca_xi_report *report) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    picca::XiConst C;
    int mode = 0;
    int status = build_xi_const(params, C, mode);
    if (status != PICCA_OK) return status;
    // the fast loop classifies pairs from (r1 - r2) (cos np / span) - r_par_min np / span and
    // (d1 + d2) (sin nt / r_trans_max): widen the edge margins by a bound of their rounding
    {
        const double u = std::ldexp(1.0, -53);
        const double err_p = 16.0 * u * ctx->max_abs_r * C.inv_p +
                             8.0 * u * (C.np_d + std::fabs(C.rpmin) * C.inv_p + 2.0);
        const double err_t = 16.0 * u * ctx->max_abs_d * C.inv_t + 8.0 * u * (C.nt_d + 2.0);
        C.fast_p += err_p;
        C.fast_t += err_t;
        C.half_p = 0.5 - C.fast_p;
        C.half_t = 0.5 - C.fast_t;
        if (!(C.half_p > 0.25 && C.half_t > 0.25))
            return fail(PICCA_E_VALUE, "edge_tol is too large compared with the bin width");
    }
    const size_t static_lds = 1024;
    const bool fast_mode = mode == picca::MODE_STANDARD && !params->has_z_min_pairs &&
                           !params->has_z_max_pairs;
    // rows of the padded window histogram that fit in LDS
    const int64_t per_row = (int64_t)(C.np + 2) * (int64_t)picca::XI_HIST_BIN_BYTES;
    const int64_t fit_rows = (ctx->lds_limit - (int64_t)static_lds) / per_row - 2;
    const bool fast = fast_mode && fit_rows >= 1;
    const size_t hist = (size_t)C.nb * picca::XI_HIST_BIN_BYTES;
    const bool lds_hist = fast || (int64_t)(hist + static_lds) <= ctx->lds_limit;
    int rows = C.nt;
    if (fast) {
        rows = (int)std::min<int64_t>(C.nt, fit_rows);
        // balance the windows: the same number of rows in each
        const int nwin = (C.nt + rows - 1) / rows;
        rows = (C.nt + nwin - 1) / nwin;
    }
    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    if (fast && rows < C.nt) {
        status = xcf_plan(ctx, C, rows);
        if (status != PICCA_OK) return status;
    }
    const int64_t nrows = ctx->xcf_rows;
    const int64_t nout = nrows * C.nb;
    PICCA_HIP_CHECK(ctx->o_bins.reserve((size_t)std::max<int64_t>(nout, 1)));
    PICCA_HIP_CHECK(ctx->queue.reserve(1));
    PICCA_HIP_CHECK(ctx->stats.reserve(picca::XI_NSTATS));
    PICCA_HIP_CHECK(picca::launch_zero_u64(reinterpret_cast<unsigned long long *>(ctx->o_bins.ptr),
                                           nout * (int64_t)(sizeof(picca::XiBin) / 8), ctx->stream));
    PICCA_HIP_CHECK(picca::launch_zero_u64(ctx->stats.ptr, picca::XI_NSTATS, ctx->stream));
    picca::ForestView F{ctx->pix.ptr, ctx->z.ptr, ctx->r.ptr, ctx->d.ptr, ctx->w.ptr, ctx->dw.ptr,
                        ctx->ll.ptr, ctx->off.ptr, ctx->zq.ptr, ctx->order.ptr, ctx->mono.ptr,
                        ctx->rd.ptr, ctx->pre.ptr, ctx->spc.ptr};
    picca::XiOut O{ctx->o_bins.ptr, ctx->queue.ptr, ctx->stats.ptr};
    PICCA_HIP_CHECK(hipEventRecord(ctx->ev_start, ctx->stream));
    int64_t launches = 0;
    auto run = [&](const picca::XcfView &X, const picca::XiConst &K) -> int {
        if (X.count <= 0) return PICCA_OK;
        ++launches;
        const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(ctx->num_cus, X.count));
        const size_t lds = fast ? (size_t)(K.np + 2) * (size_t)(K.win_rows + 2) *
                                      picca::XI_HIST_BIN_BYTES
                                : (lds_hist ? hist : 16);
        PICCA_HIP_CHECK(picca::launch_zero_u64(ctx->queue.ptr, 1, ctx->stream));
        PICCA_HIP_CHECK(picca::launch_xcf(mode, lds_hist, fast, K, F, X, O, grid, lds,
                                          ctx->stream));
        return PICCA_OK;
    };
    if (fast && rows < C.nt) {
        const int nwin = (int)ctx->x_windows.size();
        for (int w = 0; w < nwin; ++w) {
            const picca_ctx::XcfWindow &win = *ctx->x_windows[(size_t)w];
            picca::XiConst K = C;
            K.win_t0 = win.t0;
            K.win_rows = win.rows;
            K.win_last = w == nwin - 1;
            picca::XcfView X{ctx->x_rec.ptr, win.idx.ptr, win.forest.ptr, win.begin.ptr,
                             win.row.ptr, win.chunk.ptr, win.groups, win.chunks};
            status = run(X, K);
            if (status != PICCA_OK) return status;
        }
    } else {
        picca::XiConst K = C;
        K.win_t0 = 0;
        K.win_rows = C.nt;
        K.win_last = 1;
        picca::XcfView X{ctx->x_rec.ptr, nullptr, ctx->x_forest.ptr, ctx->x_begin.ptr,
                         ctx->x_row.ptr, ctx->x_chunk.ptr, ctx->xcf_groups, ctx->xcf_chunks};
        status = run(X, K);
        if (status != PICCA_OK) return status;
    }
    PICCA_HIP_CHECK(hipEventRecord(ctx->ev_stop, ctx->stream));
    unsigned long long stats[picca::XI_NSTATS];
    PICCA_HIP_CHECK(hipMemcpyAsync(stats, ctx->stats.ptr, sizeof(stats), hipMemcpyDeviceToHost,
                                   ctx->stream));
    PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    PICCA_HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    ctx->out_nb = C.nb;
    ctx->out_rows = (int32_t)nrows;
    ctx->have_result = true;
    if (report) {
        report->accepted_pairs = (int64_t)stats[picca::XI_STAT_ACCEPTED];
        report->edge_pairs = (int64_t)stats[picca::XI_STAT_EDGE];
        report->overflow_pairs = (int64_t)stats[picca::XI_STAT_OVERFLOW];
        report->visited_pairs = (int64_t)stats[picca::XI_STAT_VISITED];
        report->wave_steps = (int64_t)stats[picca::XI_STAT_STEPS];
        report->tasks = (int64_t)stats[picca::XI_STAT_TASKS];
        report->empty_tasks = (int64_t)stats[picca::XI_STAT_EMPTY_TASKS];
        report->kernel_ms = (double)ms;
        report->launches = launches;
    }
    return PICCA_OK;
}

int picca_diag_copy(picca_ctx *ctx, int64_t bytes, double *kernel_ms) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    if (bytes <= 0 || bytes % 16 != 0)
        return fail(PICCA_E_VALUE, "bytes must be a positive multiple of 16");
    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    DeviceBuffer<unsigned long long> src, dst;
    const int64_t words = bytes / 8;
    PICCA_HIP_CHECK(src.reserve((size_t)words));
    PICCA_HIP_CHECK(dst.reserve((size_t)words));
    PICCA_HIP_CHECK(picca::launch_zero_u64(src.ptr, words, ctx->stream));
    PICCA_HIP_CHECK(hipEventRecord(ctx->ev_start, ctx->stream));
    PICCA_HIP_CHECK(picca::launch_copy16(src.ptr, dst.ptr, bytes / 16, ctx->stream));
    PICCA_HIP_CHECK(hipEventRecord(ctx->ev_stop, ctx->stream));
    PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    PICCA_HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    if (kernel_ms) *kernel_ms = (double)ms;
    return PICCA_OK;
}

int picca_dmat_run(picca_ctx *ctx, const picca_cf_params *params, int64_t num_pairs,
                   const int32_t *pair_a, const int32_t *pair_b, const double *pair_ang,
                   const uint8_t *pair_flags, double *weights_dmat, double *dmat,
                   double *r_par_eff, double *r_trans_eff, double *z_eff,
                   double *weight_eff, int64_t *stats) {
    if (!ctx) return fail(PICCA_E_VALUE, "ctx is NULL");
    picca::XiConst X;
    int mode = 0;
    int status = build_xi_const(params, X, mode);
    if (status != PICCA_OK) return status;
    if (params->ang_correlation)
        return fail(PICCA_E_VALUE, "the distortion matrix has no angular mode");
    if (params->num_model_bins_r_par <= 0 || params->num_model_bins_r_trans <= 0)
        return fail(PICCA_E_VALUE, "num_model_bins_r_par and num_model_bins_r_trans must be "
                                   "positive");
    if (!ctx->has_log_lambda || !ctx->has_order)
        return fail(PICCA_E_VALUE, "the distortion matrix needs a forest store with "
                                   "log_lambda and order");
    if (num_pairs < 0) return fail(PICCA_E_VALUE, "negative size");
    if (num_pairs > 0 && (!pair_a || !pair_b || !pair_ang || !pair_flags))
        return fail(PICCA_E_VALUE, "pair arrays must not be NULL");
    if (!weights_dmat || !dmat || !r_par_eff || !r_trans_eff || !z_eff || !weight_eff)
        return fail(PICCA_E_VALUE, "output arrays must not be NULL");
    const int64_t nf = ctx->num_forests;
    for (int64_t k = 0; k < num_pairs; ++k) {
        if (pair_a[k] < 0 || pair_a[k] >= nf || pair_b[k] < 0 || pair_b[k] >= nf)
            return fail(PICCA_E_VALUE, "pair index out of range of the forest store");
        if (!(pair_ang[k] >= 0.0) || !(pair_ang[k] <= M_PI))
            return fail(PICCA_E_VALUE, "pair angle outside [0, pi]");
    }

    picca::DmatConst C;
    std::memset(&C, 0, sizeof(C));
    C.np = X.np;
    C.nt = X.nt;
    C.nb = X.nb;
    C.npm = params->num_model_bins_r_par;
    C.ntm = params->num_model_bins_r_trans;
    if ((int64_t)C.npm * C.ntm > (1 << 26) || (int64_t)C.nb * ((int64_t)C.npm * C.ntm) > (1ll << 32))
        return fail(PICCA_E_VALUE, "too many model bins");
    C.nm = C.npm * C.ntm;
    C.flags = X.flags & (picca::XI_HAS_ZMIN | picca::XI_HAS_ZMAX | picca::XI_HAS_ZERR |
                         picca::XI_REMOVE_SHP);
    if (params->redshift_evolution) C.flags |= picca::DMAT_REDSHIFT_EVOL;
    if (params->x_correlation) C.flags |= picca::DMAT_XCORR;
    if (params->rmu_binning) C.flags |= picca::DMAT_RMU;
    C.words_u = (C.nm + 31) / 32;
    C.rpmin = X.rpmin;
    C.rpmax = X.rpmax;
    C.rtmax = X.rtmax;
    C.span_p = X.span_p;
    C.width_p = X.width_p;
    C.zmin = X.zmin;
    C.zmax = X.zmax;
    C.zerr_ang = X.zerr_ang;
    C.zerr_kms = X.zerr_kms;
    C.np_d = C.np;
    C.nt_d = C.nt;
    C.npm_d = C.npm;
    C.ntm_d = C.ntm;
    C.inv_span_p = 1.0 / C.span_p;
    C.inv_rtmax = 1.0 / C.rtmax;
    // LDS of the pass kernel: U bitmap, row ranges, then bands of k r_trans rows
    // (4 sums per data bin, 8 per model bin)
    const size_t fixed = ((size_t)C.words_u * 4 + (size_t)(4 * C.nt + 4 * C.ntm) * 4 + 15) &
                         ~(size_t)15;
    const size_t lds_budget = std::min<size_t>(120 * 1024, (size_t)ctx->lds_limit);
    const size_t per_row = 32 * (size_t)C.np + 64 * (size_t)C.npm;
    if (fixed + per_row > lds_budget)
        return fail(PICCA_E_VALUE, "too many bins for the distortion-matrix pass kernel");
    // at most DMAT_PASS_ROWS rows per band: a pair's rows usually fit (one walk does census
    // and sums), and the smaller LDS lets several pair workgroups share a CU
    const int k_rows = (int)std::min<size_t>(
        {(lds_budget - fixed) / per_row, (size_t)std::max(C.nt, C.ntm),
         (size_t)picca::DMAT_PASS_ROWS});
    C.tb_max = std::min(C.nt, k_rows);
    C.ub_max = std::min(C.ntm, k_rows);
    const size_t pass_lds = fixed + 32 * (size_t)C.np * C.tb_max + 64 * (size_t)C.npm * C.ub_max;

    PICCA_HIP_CHECK(hipSetDevice(ctx->device));
    PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    const int64_t npix = ctx->num_pixels;
    const size_t pad = picca::PIXEL_PAD;
    {   // redshift-evolution factors of each pixel (py/picca/cf.py:680-685)
        std::vector<double> fa((size_t)std::max<int64_t>(npix, 1), 1.0);
        std::vector<double> fb((size_t)std::max<int64_t>(npix, 1), 1.0);
        if (params->redshift_evolution) {
            const double zr = 1 + params->z_ref;
            const double ea = params->alpha - 1, eb = params->alpha2 - 1;
            parallel_for(npix, [&](int64_t lo, int64_t hi) {
                for (int64_t k = lo; k < hi; ++k) {
                    fa[k] = std::pow((1 + ctx->h_z[k]) / zr, ea);
                    fb[k] = std::pow((1 + ctx->h_z[k]) / zr, eb);
                }
            });
        }
        PICCA_HIP_CHECK(ctx->zfa.upload_padded(fa.data(), (size_t)npix, pad, ctx->stream));
        PICCA_HIP_CHECK(ctx->zfb.upload_padded(fb.data(), (size_t)npix, pad, ctx->stream));
    }
    std::vector<double> cs((size_t)std::max<int64_t>(2 * num_pairs, 1));
    for (int64_t k = 0; k < num_pairs; ++k) {
        cs[2 * k] = std::cos(pair_ang[k] / 2);
        cs[2 * k + 1] = std::sin(pair_ang[k] / 2);
    }
    PICCA_HIP_CHECK(ctx->dpa.upload(pair_a, (size_t)num_pairs, ctx->stream));
    PICCA_HIP_CHECK(ctx->dpb.upload(pair_b, (size_t)num_pairs, ctx->stream));
    PICCA_HIP_CHECK(ctx->dang.upload(pair_ang, (size_t)num_pairs, ctx->stream));
    PICCA_HIP_CHECK(ctx->dcs.upload(cs.data(), (size_t)(2 * num_pairs), ctx->stream));
    PICCA_HIP_CHECK(ctx->dflags.upload(pair_flags, (size_t)num_pairs, ctx->stream));

    // outputs, accumulated over all batches on the device
    const int64_t nb = C.nb, nm = C.nm;
    PICCA_HIP_CHECK(ctx->o_wdm.reserve((size_t)nb));
    PICCA_HIP_CHECK(ctx->o_dmat.reserve((size_t)(nb * nm)));
    PICCA_HIP_CHECK(ctx->o_eff.reserve((size_t)(4 * nm)));
    PICCA_HIP_CHECK(ctx->dstats.reserve(picca::DMAT_NSTATS));
    PICCA_HIP_CHECK(picca::launch_zero_u64((unsigned long long *)ctx->o_wdm.ptr, nb, ctx->stream));
    PICCA_HIP_CHECK(picca::launch_zero_u64((unsigned long long *)ctx->o_dmat.ptr, nb * nm,
                                           ctx->stream));
    PICCA_HIP_CHECK(picca::launch_zero_u64((unsigned long long *)ctx->o_eff.ptr, 4 * nm,
                                           ctx->stream));
    PICCA_HIP_CHECK(picca::launch_zero_u64(ctx->dstats.ptr, picca::DMAT_NSTATS, ctx->stream));

    // workspace per slot
    const size_t slot_bytes = 8 * (4 * (size_t)C.nt * C.np + 4 * (size_t)C.ntm * C.npm) +
                              4 * (size_t)C.words_u + 16 + 16 * (size_t)(C.nt + C.ntm);
    const size_t ws_budget = (size_t)1536 << 20;
    const int64_t cap = std::max<int64_t>(
        1, std::min<int64_t>({num_pairs, (int64_t)(ws_budget / slot_bytes), (int64_t)1 << 20}));
    const int64_t touched_words = ((int64_t)C.nt * C.ntm + 63) / 64;
    PICCA_HIP_CHECK(ctx->ws_S.reserve((size_t)cap * 4 * C.nt * C.np));
    PICCA_HIP_CHECK(ctx->ws_eta.reserve((size_t)cap * 4 * C.ntm * C.npm));
    PICCA_HIP_CHECK(ctx->ws_U.reserve((size_t)cap * C.words_u));
    PICCA_HIP_CHECK(ctx->ws_band.reserve((size_t)cap * 4));
    PICCA_HIP_CHECK(ctx->ws_rng.reserve((size_t)cap * 4 * (C.nt + C.ntm)));
    PICCA_HIP_CHECK(ctx->ws_pix.reserve((size_t)cap));
    PICCA_HIP_CHECK(ctx->ws_touched.reserve((size_t)touched_words));

    picca::DmatForests F{ctx->r.ptr,   ctx->d.ptr,     ctx->z.ptr,     ctx->w.ptr,   ctx->dll.ptr,
                         ctx->zfa.ptr, ctx->zfb.ptr,   ctx->off.ptr,   ctx->zq.ptr,  ctx->sum_w.ptr,
                         ctx->sw_dll.ptr, ctx->order.ptr, ctx->mono.ptr};
    picca::DmatOut O{ctx->o_wdm.ptr, ctx->o_dmat.ptr, ctx->o_eff.ptr, ctx->o_eff.ptr + nm,
                     ctx->o_eff.ptr + 2 * nm, ctx->o_eff.ptr + 3 * nm};
    std::vector<int64_t> pix_off;
    int64_t batches = 0;
    for (int64_t p0 = 0; p0 < num_pairs; p0 += cap) {
        const int64_t count = std::min(cap, num_pairs - p0);
        pix_off.assign((size_t)count, 0);
        int64_t total = 0;
        for (int64_t k = 0; k < count; ++k) {
            pix_off[k] = total;
            const int32_t a = pair_a[p0 + k];
            total += ctx->h_offsets[a + 1] - ctx->h_offsets[a];
        }
        PICCA_HIP_CHECK(ctx->ws_pix.upload(pix_off.data(), (size_t)count, ctx->stream));
        PICCA_HIP_CHECK(ctx->ws_cnt.reserve((size_t)std::max<int64_t>(total, 1)));
        PICCA_HIP_CHECK(picca::launch_zero_u64(ctx->ws_touched.ptr, touched_words, ctx->stream));
        picca::DmatPairs P{ctx->dpa.ptr + p0, ctx->dpb.ptr + p0, ctx->dang.ptr + p0,
                           ctx->dcs.ptr + 2 * p0, ctx->dflags.ptr + p0, (int32_t)count};
        int32_t *rng = ctx->ws_rng.ptr;
        picca::DmatWork W{ctx->ws_S.ptr,
                          ctx->ws_eta.ptr,
                          ctx->ws_U.ptr,
                          ctx->ws_band.ptr,
                          rng,
                          rng + (size_t)count * 2 * C.nt,
                          rng + (size_t)count * (2 * C.nt + 2 * C.ntm),
                          rng + (size_t)count * (4 * C.nt + 2 * C.ntm),
                          ctx->ws_cnt.ptr,
                          ctx->ws_pix.ptr,
                          (uint32_t *)ctx->ws_touched.ptr,
                          ctx->dstats.ptr};
        const size_t ev = 3 * (size_t)batches++;
        while (ctx->dmat_events.size() < ev + 3) {
            hipEvent_t e = nullptr;
            PICCA_HIP_CHECK(hipEventCreate(&e));
            ctx->dmat_events.push_back(e);
        }
        PICCA_HIP_CHECK(hipEventRecord(ctx->dmat_events[ev], ctx->stream));
        PICCA_HIP_CHECK(picca::launch_dmat_pass(C, F, P, W, O, pass_lds, ctx->stream));
        PICCA_HIP_CHECK(hipEventRecord(ctx->dmat_events[ev + 1], ctx->stream));
        const int groups = (int)std::max<int64_t>(1, std::min<int64_t>(32, count / 16));
        bool quasar = true;   // every delta2 of the batch unprojected
        for (int64_t k = 0; k < count && quasar; ++k)
            quasar = ctx->h_order[(size_t)pair_b[p0 + k]] < 0;
        PICCA_HIP_CHECK(picca::launch_dmat_fill(C, F, P, W, O, groups, quasar, ctx->stream));
        PICCA_HIP_CHECK(hipEventRecord(ctx->dmat_events[ev + 2], ctx->stream));
    }

    std::vector<double> h_wdm((size_t)nb), h_dmat((size_t)(nb * nm)), h_eff((size_t)(4 * nm));
    unsigned long long h_stats[picca::DMAT_NSTATS];
    PICCA_HIP_CHECK(hipMemcpyAsync(h_wdm.data(), ctx->o_wdm.ptr, sizeof(double) * nb,
                                   hipMemcpyDeviceToHost, ctx->stream));
    PICCA_HIP_CHECK(hipMemcpyAsync(h_dmat.data(), ctx->o_dmat.ptr, sizeof(double) * nb * nm,
                                   hipMemcpyDeviceToHost, ctx->stream));
    PICCA_HIP_CHECK(hipMemcpyAsync(h_eff.data(), ctx->o_eff.ptr, sizeof(double) * 4 * nm,
                                   hipMemcpyDeviceToHost, ctx->stream));
    PICCA_HIP_CHECK(hipMemcpyAsync(h_stats, ctx->dstats.ptr, sizeof(h_stats),
                                   hipMemcpyDeviceToHost, ctx->stream));
    PICCA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->dmat_report = picca_dmat_report{0.0, 0.0, batches,
                                         (int64_t)h_stats[picca::DMAT_STAT_MFMA] * 2048};
    for (int64_t k = 0; k < batches; ++k) {
        float pass_ms = 0.f, fill_ms = 0.f;
        PICCA_HIP_CHECK(hipEventElapsedTime(&pass_ms, ctx->dmat_events[3 * k],
                                            ctx->dmat_events[3 * k + 1]));
        PICCA_HIP_CHECK(hipEventElapsedTime(&fill_ms, ctx->dmat_events[3 * k + 1],
                                            ctx->dmat_events[3 * k + 2]));
        ctx->dmat_report.pass_ms += pass_ms;
        ctx->dmat_report.fill_ms += fill_ms;
    }
    if (h_stats[picca::DMAT_STAT_NEGATIVE] > 0)
        return fail(PICCA_E_INDEX, "a selected pair has a negative data bin index");
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
    if (stats) {
        stats[0] = (int64_t)h_stats[picca::DMAT_STAT_IN_RANGE];
        stats[1] = (int64_t)h_stats[picca::DMAT_STAT_SELECTED];
        stats[2] = (int64_t)h_stats[picca::DMAT_STAT_BEYOND];
        stats[3] = (int64_t)h_stats[picca::DMAT_STAT_OVERFLOW];
    }
    return PICCA_OK;
}

}  // extern "C"
