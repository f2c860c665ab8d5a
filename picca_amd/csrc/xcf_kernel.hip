// Pair kernel of the forest x quasar cross-correlation on gfx950 (MI355X).
//
// GPU restatement of compute_xi_forest_pairs_fast of py/picca/xcf.py:227-322 for every
// (forest, neighbouring quasar) pair of a run, grouped by output row (the healpix of the
// forest).  Same persistent structure as the auto-correlation kernel: one 1024-thread
// workgroup per CU keeps its current row's histogram in LDS; groups (one forest and all of
// its neighbours) come from an HBM queue; inside a group, waves take chunks of 64
// neighbours.  Lane l owns quasar 64 chunk + l (its coordinates and the pair's cos/sin in
// registers); the wave walks the forest's pixels in lockstep, each pixel a wave-uniform
// record read through the scalar cache.  Consecutive pixels move r_par monotonically, so a
// lane sums the pairs of one bin in registers and adds the run to LDS when the bin changes.
//
// Pruning (standard binning, forest with non-decreasing r_comov and dist_m): the host sorts
// each forest's neighbours by r_comov, and the wave only walks the pixels whose r_comov can
// give r_par in (r_par_min, r_par_max) and whose dist_m can give r_trans < r_trans_max for
// some lane.  Every pixel pair outside that window fails the reference's range test.
//
// Arithmetic: the reference's FP64 operations in its order (built with -ffp-contract=off);
// the bins are divide-then-multiply, as in the reference; delta w1 w2 is (delta w1) w2.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <limits.h>

#include "device_util.h"
#include "engine.h"
#include "pair_common.h"

namespace picca {

namespace {

template <int MODE>
__device__ __forceinline__ bool xcf_pair(const XiConst &C, const PixelRecord &q,
                                         const XcfPairRecord &o, int &bin, double &r_par,
                                         double &r_trans, double &z, uint32_t &edge,
                                         uint32_t &overflow) {
    z = (q.z + o.z) / 2;
    if ((C.flags & XI_HAS_ZMIN) && z < C.zmin) return false;
    if ((C.flags & XI_HAS_ZMAX) && z > C.zmax) return false;
    if (MODE == MODE_ANGULAR) {
        r_par = q.r / o.r;
        r_trans = o.c;
    } else {
        r_par = (q.r - o.r) * o.c;
        r_trans = (q.d + o.d) * o.s;
    }
    if (MODE == MODE_RMU) {
        r_trans = sqrt(r_trans * r_trans + r_par * r_par);
        r_par /= r_trans;
    }
    if (r_par >= C.rpmax || r_trans >= C.rtmax || r_par <= C.rpmin) {
        if ((r_par >= C.rpmax && r_par - C.rpmax < C.edge_tol) ||
            (r_trans >= C.rtmax && r_trans - C.rtmax < C.edge_tol) ||
            (r_par <= C.rpmin && C.rpmin - r_par < C.edge_tol))
            ++edge;
        return false;
    }
    const double t_par = (r_par - C.rpmin) / C.span_p * C.np_d;
    const double t_trans = r_trans / C.rtmax * C.nt_d;
    const double f_par = floor(t_par);
    const double f_trans = floor(t_trans);
    if ((r_par != C.rpmin && edge_distance(t_par, C.width_p) < C.edge_tol) ||
        (r_trans != 0 && edge_distance(t_trans, C.width_t) < C.edge_tol))
        ++edge;
    const double bin_f = f_trans + C.nt_d * f_par;
    if (!(bin_f >= 0.0 && bin_f < C.nb_d)) {
        ++overflow;
        return false;
    }
    bin = (int)bin_f;
    return true;
}

template <int MODE, typename Bin>
__device__ void xcf_task(const XiConst &C, const ForestView &F, const XcfView &X, int64_t p0,
                         int n, int chunk, int a, int lane, Bin *hist, LaneCounters &cnt,
                         uint64_t &n_steps) {
    const int k = chunk * 64 + lane;
    bool active = k < n;
    XcfPairRecord o{0, 0, 0, 0, 1, 0};
    if (active) {
        o = X.rec[p0 + k];
        active = o.w != 0;
    }
    if (__ballot(active) == 0) return;
    const int64_t oa = uniform_i64(F.off[a]);
    const int la = (int)(uniform_i64(F.off[a + 1]) - oa);
    int i_lo = 0, i_hi = la;
    if (MODE == MODE_STANDARD && (uniform_int(F.mono[a]) & 3) == 3) {
        const bool pos = active && o.c > 0;
        if (__ballot(active && !pos) == 0) {
            // r_par = (r1 - r2) cos in (r_par_min, r_par_max)  <=>
            // r1 in (r2 + r_par_min / cos, r2 + r_par_max / cos)
            double lo = wave_min(pos ? o.r + C.rpmin / o.c : INFINITY);
            double hi = wave_max(pos ? o.r + C.rpmax / o.c : -INFINITY);
            lo -= 1e-6 + 1e-12 * fabs(lo);
            hi += 1e-6 + 1e-12 * fabs(hi);
            i_lo = wave_search<true>(F.r + oa, la, lo, lane);
            i_hi = wave_search<false>(F.r + oa, la, hi, lane);
            // r_trans = (d1 + d2) sin < r_trans_max  =>  d1 < r_trans_max / sin - d2
            if (__ballot(active && !(o.s > 0)) == 0) {
                double d_lim = wave_max(active ? C.rtmax / o.s - o.d : -INFINITY);
                d_lim += 1e-6 + 1e-12 * fabs(d_lim);
                const int i_d = wave_search<false>(F.d + oa, la, d_lim, lane);
                if (i_d < i_hi) i_hi = i_d;
            }
        }
    }
    i_lo = uniform_int(i_lo);
    i_hi = uniform_int(i_hi);
    LaneRun run;
    run.bin = -1;
    run.count = 0;
    run.we = run.xi = run.rp = run.rt = run.z = 0.0;
    uint32_t edge = 0, overflow = 0;
    uint64_t visit = 0;
    const PixelRecord *rec = F.pix + oa;
    for (int i = i_lo; i < i_hi; ++i) {
        const PixelRecord q = load_record(rec + i);
        // the reference skips w1 == 0 (either sign) before the quasar loop
        if ((__double_as_longlong(q.w) << 1) == 0) continue;
        ++n_steps;
        if (!active) continue;
        ++visit;
        int bin = 0;
        double r_par, r_trans, z;
        if (xcf_pair<MODE>(C, q, o, bin, r_par, r_trans, z, edge, overflow))
            add_pair(run, hist, cnt, bin, q.w * o.w, q.dw * o.w, r_par, r_trans, z);
    }
    if (run.bin >= 0) {
        flush_run(run, hist);
        cnt.accept += run.count;
    }
    cnt.visit += visit;
    cnt.edge += edge;
    cnt.overflow += overflow;
}

}  // namespace

template <int MODE, bool LDS_HIST>
__global__ __launch_bounds__(XI_BLOCK) void xcf_kernel(XiConst C, ForestView F, XcfView X,
                                                       XiOut O) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int64_t s_group;
    __shared__ int s_task;
    __shared__ unsigned long long s_stats[XI_NSTATS];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int nb = C.nb;
    LocalBin *hist = (LocalBin *)reinterpret_cast<LdsBin *>(smem);
    if (LDS_HIST) {
        for (int k = tid; k < nb; k += XI_BLOCK) {
            hist[k].we = hist[k].xi = hist[k].rp = hist[k].rt = hist[k].z = 0.0;
            hist[k].nb = 0;
        }
    }
    if (tid < XI_NSTATS) s_stats[tid] = 0;
    LaneCounters cnt{0, 0, 0, 0};
    uint64_t n_steps = 0, n_tasks = 0, n_empty = 0;
    int current_row = -1;
    auto flush_hist = [&](int row) {
        XiBin *out = O.bins + (int64_t)row * nb;
        for (int k = tid; k < nb; k += XI_BLOCK) {
            const uint32_t c = hist[k].nb;
            if (c) {
                unsafeAtomicAdd(&out[k].we, hist[k].we);
                unsafeAtomicAdd(&out[k].xi, hist[k].xi);
                unsafeAtomicAdd(&out[k].rp, hist[k].rp);
                unsafeAtomicAdd(&out[k].rt, hist[k].rt);
                unsafeAtomicAdd(&out[k].z, hist[k].z);
                atomicAdd(&out[k].nb, (unsigned long long)c);
                hist[k].we = hist[k].xi = hist[k].rp = hist[k].rt = hist[k].z = 0.0;
                hist[k].nb = 0;
            }
        }
    };
    for (;;) {
        __syncthreads();
        if (tid == 0) {
            s_group = (int64_t)atomicAdd(O.queue, 1ull);
            s_task = 0;
        }
        __syncthreads();
        const int64_t g = s_group;
        if (g >= X.count) break;
        const int row = X.row[g];
        if (LDS_HIST && row != current_row && current_row >= 0) flush_hist(current_row);
        current_row = row;
        const int64_t p0 = X.begin[g];
        const int n = (int)(X.begin[g + 1] - p0);
        const int a = X.forest[g];
        const int chunks = (n + 63) >> 6;
        for (;;) {
            int t = 0;
            if (lane == 0) t = atomicAdd(&s_task, 1);
            t = uniform_int(__shfl(t, 0, 64));
            if (t >= chunks) break;
            const uint64_t before = n_steps;
            if (LDS_HIST)
                xcf_task<MODE>(C, F, X, p0, n, t, a, lane, hist, cnt, n_steps);
            else
                xcf_task<MODE>(C, F, X, p0, n, t, a, lane, O.bins + (int64_t)row * nb, cnt,
                               n_steps);
            ++n_tasks;
            if (n_steps == before) ++n_empty;
        }
    }
    if (LDS_HIST && current_row >= 0) flush_hist(current_row);
    const uint64_t v_visit = wave_sum_u64(cnt.visit);
    const uint64_t v_edge = wave_sum_u64(cnt.edge);
    const uint64_t v_over = wave_sum_u64(cnt.overflow);
    const uint64_t v_accept = wave_sum_u64(cnt.accept);
    if (lane == 0) {
        atomicAdd(&s_stats[XI_STAT_VISITED], (unsigned long long)v_visit);
        atomicAdd(&s_stats[XI_STAT_EDGE], (unsigned long long)v_edge);
        atomicAdd(&s_stats[XI_STAT_OVERFLOW], (unsigned long long)v_over);
        atomicAdd(&s_stats[XI_STAT_STEPS], (unsigned long long)n_steps);
        atomicAdd(&s_stats[XI_STAT_ACCEPTED], (unsigned long long)v_accept);
        atomicAdd(&s_stats[XI_STAT_TASKS], (unsigned long long)n_tasks);
        atomicAdd(&s_stats[XI_STAT_EMPTY_TASKS], (unsigned long long)n_empty);
    }
    __syncthreads();
    if (tid < XI_NSTATS) atomicAdd(O.stats + tid, s_stats[tid]);
}

template <int MODE, bool LDS_HIST>
static hipError_t launch_xcf_t(const XiConst &C, const ForestView &F, const XcfView &X,
                               const XiOut &O, int grid, size_t lds, hipStream_t stream) {
    auto kern = xcf_kernel<MODE, LDS_HIST>;
    hipError_t err = hipFuncSetAttribute((const void *)kern,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (err != hipSuccess) return err;
    kern<<<dim3(grid), dim3(XI_BLOCK), lds, stream>>>(C, F, X, O);
    return hipGetLastError();
}

hipError_t launch_xcf(int mode, bool lds_hist, const XiConst &C, const ForestView &F,
                      const XcfView &X, const XiOut &O, int grid, size_t lds,
                      hipStream_t stream) {
    if (mode == MODE_STANDARD)
        return lds_hist ? launch_xcf_t<MODE_STANDARD, true>(C, F, X, O, grid, lds, stream)
                        : launch_xcf_t<MODE_STANDARD, false>(C, F, X, O, grid, lds, stream);
    if (mode == MODE_RMU)
        return lds_hist ? launch_xcf_t<MODE_RMU, true>(C, F, X, O, grid, lds, stream)
                        : launch_xcf_t<MODE_RMU, false>(C, F, X, O, grid, lds, stream);
    return lds_hist ? launch_xcf_t<MODE_ANGULAR, true>(C, F, X, O, grid, lds, stream)
                    : launch_xcf_t<MODE_ANGULAR, false>(C, F, X, O, grid, lds, stream);
}

}  // namespace picca
