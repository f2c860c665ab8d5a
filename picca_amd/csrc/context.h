// Device context of the C-ABI (shared by the entry points of the library).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <thread>
#include <memory>
#include <vector>

#include "../../include/picca_hip.h"
#include "engine.h"
#include "xi_walk.h"

namespace picca_host {

// message of the last failure on the calling thread (picca_last_error)
std::string &last_error();

inline int fail(int code, const std::string &msg) {
    last_error() = msg;
    return code;
}

#define PICCA_HIP_CHECK(expr)                                                              \
    do {                                                                                   \
        hipError_t err_ = (expr);                                                          \
        if (err_ != hipSuccess)                                                            \
            return picca_host::fail(PICCA_E_HIP,                                           \
                                    std::string(#expr) + ": " + hipGetErrorString(err_));  \
    } while (0)

template <typename T>
struct DeviceBuffer {
    T *ptr = nullptr;
    size_t count = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    ~DeviceBuffer() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
    hipError_t reserve(size_t n) {
        if (n <= count && ptr) return hipSuccess;
        release();
        size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
        hipError_t err = hipMalloc((void **)&ptr, bytes);
        if (err != hipSuccess) {
            ptr = nullptr;
            return err;
        }
        count = n;
        return hipSuccess;
    }
    hipError_t upload(const T *host, size_t n, hipStream_t stream) {
        hipError_t err = reserve(n);
        if (err != hipSuccess || n == 0) return err;
        return hipMemcpyAsync(ptr, host, n * sizeof(T), hipMemcpyHostToDevice, stream);
    }
    // n values followed by `pad` zeros (kernels may read a few elements past the end)
    hipError_t upload_padded(const T *host, size_t n, size_t pad, hipStream_t stream) {
        hipError_t err = reserve(n + pad);
        if (err != hipSuccess) return err;
        if (n > 0) err = hipMemcpyAsync(ptr, host, n * sizeof(T), hipMemcpyHostToDevice, stream);
        if (err != hipSuccess) return err;
        return hipMemsetAsync(ptr + n, 0, pad * sizeof(T), stream);
    }
};

// Runs body(begin, end) over [0, n) on up to hardware_concurrency threads.
template <typename F>
inline void parallel_for(int64_t n, F body) {
    unsigned hw = std::thread::hardware_concurrency();
    int64_t nthreads = std::max<int64_t>(1, std::min<int64_t>(hw ? hw : 1, 16));
    if (n < 65536) nthreads = 1;
    if (nthreads == 1) {
        body(0, n);
        return;
    }
    std::vector<std::thread> pool;
    int64_t step = (n + nthreads - 1) / nthreads;
    for (int64_t t = 0; t < nthreads; ++t) {
        int64_t lo = t * step, hi = std::min(n, lo + step);
        if (lo >= hi) break;
        pool.emplace_back([=] { body(lo, hi); });
    }
    for (auto &th : pool) th.join();
}

}  // namespace picca_host

using picca_host::DeviceBuffer;
using picca_host::fail;
using picca_host::parallel_for;

struct picca_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr, ev_mid = nullptr;
    // distortion matrix: events around each batch's pass and fill kernels (a pool that
    // grows with the batches of a run) and the last run's report
    std::vector<hipEvent_t> dmat_events;
    picca_dmat_report dmat_report = {0.0, 0.0, 0, 0};
    int num_cus = 0;
    int64_t lds_limit = 0;
    unsigned long long *scratch = nullptr;   // this GPU's no-allocation scratch array

    // forest store
    int64_t num_forests = 0;
    int64_t num_pixels = 0;
    double max_abs_r = 0;   // largest |r_comov| and |dist_m| of the store: bound the
    double max_abs_d = 0;   // rounding of the pre-scaled coordinates of the fast loop
    bool has_log_lambda = false;
    bool has_order = false;
    std::vector<int32_t> h_order;   // host copy of the fit orders (< 0: unprojected)
    std::vector<int64_t> h_offsets;
    std::vector<double> h_dmin, h_dmax;   // per forest: smallest / largest dist_m
    DeviceBuffer<double> z, r, d, w, dw, ll, zq;
    DeviceBuffer<picca::PixelRecord> pix;
    DeviceBuffer<int64_t> off;
    DeviceBuffer<int32_t> order;
    DeviceBuffer<uint8_t> mono;
    // run walk of the xi kernel (xi_walk.h)
    DeviceBuffer<double2> rd, spc;
    DeviceBuffer<picca::PrefixRecord> pre;
    int32_t xi_schedule = 0;   // PICCA_XI_SCHEDULE_*
    double median_step_r = 0;  // median over forests of the mean r_comov step between pixels
    // the xi launch sequence (zero kernels + pair kernel) as a HIP graph, replayed while
    // its launch arguments stay the same
    bool xi_use_graph = true;
    hipGraph_t xi_graph = nullptr;
    hipGraphExec_t xi_graph_exec = nullptr;
    std::vector<unsigned char> xi_graph_key;

    // xi pairs
    int64_t num_pairs = 0;
    int32_t num_rows = 0;
    DeviceBuffer<int32_t> pa, pb, pchunks;
    DeviceBuffer<double> pang, pcs;
    DeviceBuffer<uint8_t> pflags;
    std::vector<int64_t> h_group_begin;
    std::vector<int32_t> h_group_row;
    DeviceBuffer<int64_t> gbegin;
    DeviceBuffer<int32_t> grow;

    // distortion-matrix inputs (stores given with log_lambda): per-forest sums of
    // compute_dmat_forest_pairs_fast (py/picca/cf.py:577-594) and the pixel redshifts
    std::vector<double> h_z;
    DeviceBuffer<double> dll, sum_w, sw_dll, zfa, zfb;

    // dmat batches
    DeviceBuffer<int32_t> dpa, dpb;
    DeviceBuffer<double> dang, dcs;
    DeviceBuffer<uint8_t> dflags;
    DeviceBuffer<double> ws_S, ws_eta;
    DeviceBuffer<uint32_t> ws_U;
    DeviceBuffer<int32_t> ws_band, ws_rng, ws_cnt;
    DeviceBuffer<int64_t> ws_pix;
    DeviceBuffer<unsigned long long> ws_touched, dstats;
    DeviceBuffer<double> o_wdm, o_dmat, o_eff;

    // metal distortion matrix: absorber tables per slot (store order) and per-pass tasks
    struct MetalSlot {
        bool set = false;
        int64_t pixels = 0;
        DeviceBuffer<double> z, r, d, w, f;
        DeviceBuffer<picca::MetalRecord> rec;   // num_pixels + METAL_REC_PAD records
    };
    MetalSlot metal[4];
    DeviceBuffer<int64_t> m_task_begin;
    DeviceBuffer<int32_t> m_task_pair;
    DeviceBuffer<unsigned long long> m_stats;

    // subsampling covariance: rows, weights, per-bin means and weight sums, matrix
    DeviceBuffer<double> cv_xi, cv_w, cv_y, cv_s, cv_cov;

    // Wick terms: evolution factors, weighted 1-D correlations, per-batch lists and tasks
    DeviceBuffer<double> wk_zev1, wk_zev2, wk_wx, wk_c1;
    DeviceBuffer<int64_t> wk_wxoff, wk_base, wk_list, wk_c1off;
    DeviceBuffer<int32_t> wk_count, wk_task, wk_segcount, wk_forests;
    DeviceBuffer<int4> wk_seg;
    DeviceBuffer<unsigned long long> wk_nb, wk_stats, wk_ntasks;

    // continuum fit: resident spectra (forest pixels) and per-forest results
    int64_t cont_forests = 0;
    int64_t cont_pixels = 0;
    DeviceBuffer<int64_t> c_off;
    DeviceBuffer<double> c_flux, c_ivar, c_mean, c_x, c_eta, c_lss, c_fudge;
    DeviceBuffer<int32_t> c_order;
    DeviceBuffer<double> c_a0, c_b0, c_a, c_b, c_chi2;
    DeviceBuffer<int32_t> c_ndata, c_status, c_evals;
    // DR16 iteration: per pixel rest-frame log lambda, transmission correction, rest-frame
    // bin; the stacked sums; the mean continuum's interpolation nodes
    DeviceBuffer<double> c_llr, c_trans, c_stack, c_xs, c_ys;
    DeviceBuffer<int32_t> c_bin;
    int32_t cont_nbins = 0;     // 0: no rest frame set for the resident forests
    bool cont_fitted = false;   // a fit of the resident forests is on the device

    // seed network: parameters, Adam state, training data and scratch
    int32_t mlp_in = 0, mlp_hidden = 0;
    int64_t mlp_step = 0;   // Adam time step
    DeviceBuffer<float> m_w1t, m_b1, m_w2, m_b2;
    DeviceBuffer<uint16_t> m_w1t_bf16;
    DeviceBuffer<float> g_w1t, g_b1, g_w2, g_b2, g_loss;
    DeviceBuffer<float> a_m, a_v;
    DeviceBuffer<float> m_x, m_t, m_h, m_y, m_dh;
    DeviceBuffer<int32_t> m_rows;
    int64_t mlp_rows = 0;   // rows of the resident training data

    // forest x quasar pairs (cross-correlation)
    int64_t xcf_pairs = 0;
    int32_t xcf_rows = 0;
    DeviceBuffer<picca::XcfPairRecord> x_rec;
    DeviceBuffer<int32_t> x_forest, x_row;
    DeviceBuffer<int64_t> x_begin, x_chunk;
    int64_t xcf_groups = 0, xcf_chunks = 0;
    // host copies of the groups and, per pair in record order, the smallest and largest
    // r_trans any pixel of its forest can have with the quasar (for the row windows)
    std::vector<int64_t> hx_begin;
    std::vector<int32_t> hx_forest, hx_row;
    std::vector<double> hx_rt_lo, hx_rt_hi;
    // row windows of the last plan: per window, the pair indices and groups on the device
    struct XcfWindow {
        int32_t t0 = 0, rows = 0;
        int64_t groups = 0;
        int64_t chunks = 0;
        DeviceBuffer<int64_t> idx, begin, chunk;
        DeviceBuffer<int32_t> forest, row;
    };
    std::vector<std::unique_ptr<XcfWindow>> x_windows;
    int64_t x_plan_version = -1;
    int32_t x_plan_nt = 0, x_plan_rows = 0;
    double x_plan_rtmax = 0;
    int64_t x_pairs_version = 0;

    // xi results
    int64_t out_nb = 0;
    int32_t out_rows = 0;
    bool have_result = false;
    DeviceBuffer<picca::XiBin> o_bins;
    DeviceBuffer<unsigned long long> queue, stats;
};

