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
// Fast loop (standard binning, no pair-redshift cuts): the pixel records come two at a
// time, the next two already loading; the bin coordinates are taken as
// (r1 - r2) (cos np / span) - r_par_min np / span and (d1 + d2) (sin nt / r_trans_max), and
// any pair within the rounding margin of a bin edge or range limit is redone with the
// reference's divide-then-multiply (exact_xcf).  The LDS histogram is padded by one row and
// one column on each side: rejected pairs land there, so every pair is binned without a
// branch.  When the np x nt histogram does not fit in LDS the host runs one launch per
// window of r_trans rows, each with the pairs whose r_trans range reaches the window.
//
// General loop (r-mu and angular binning, pair-redshift cuts): the reference's FP64
// operations in its order for every pair (built with -ffp-contract=off); the bins are
// divide-then-multiply, as in the reference; delta w1 w2 is (delta w1) w2.
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
        // outside the range, but within edge_tol of it in both coordinates
        if (r_par - C.rpmax < C.edge_tol && C.rpmin - r_par < C.edge_tol &&
            r_trans - C.rtmax < C.edge_tol)
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
        o = X.rec[X.idx ? X.idx[p0 + k] : p0 + k];
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


typedef __attribute__((address_space(3))) unsigned long long LocalU64x;

// Run of the fast loop: consecutive pixels of one lane (quasar) in one padded bin; r_par and
// r_trans are summed before the lane's cos(ang/2) / sin(ang/2), applied per run.
struct XRun {
    int32_t bin;
    double count;
    double we, xi, rp, rt, z;
};

__device__ __forceinline__ void xflush(const XRun &run, LocalBin *hist, double c, double s) {
    LocalBin *b = (LocalBin *)((LocalByte *)hist +
                               __umul24((uint32_t)run.bin, (uint32_t)sizeof(LdsBin)));
    __hip_atomic_fetch_add(&b->we, run.we, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&b->xi, run.xi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&b->rp, run.rp * c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&b->rt, run.rt * s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&b->z, run.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&b->nb, (uint32_t)run.count, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ bool xzero_weight(double w) {
    const unsigned long long bits = (unsigned long long)__double_as_longlong(w);
    unsigned long long twice;
    asm volatile("s_lshl_b64 %0, %1, 1" : "=s"(twice) : "s"(bits) : "scc");
    return twice == 0;
}

__device__ __forceinline__ double xclamp(double v, double lo, double hi) {
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(v), "s"(lo));
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(r), "s"(hi));
    return r;
}

// The reference's arithmetic for one pixel-quasar pair (py/picca/xcf.py:289-322, standard
// binning, no redshift cut) as straight-line code with the constants read once: returns
// the padded LDS bin of the pair, or `trash`.  Edge and overflow pairs are counted when
// the pair's r_trans row belongs to this launch's window (the last window also owns the
// rows at and beyond its end), so that each pair is counted by exactly one window.
__device__ __forceinline__ int exact_xcf(const XiConst &K, double r1, double d1, double r2,
                                         double d2, double c, double s, int trash,
                                         LocalU64x *lds_stats) {
    const double rpmin = K.rpmin, rpmax = K.rpmax, rtmax = K.rtmax, tol = K.edge_tol;
    const double span = K.span_p, np_d = K.np_d, nt_d = K.nt_d, nb_d = K.nb_d;
    const double width_p = K.width_p, width_t = K.width_t;
    const int t0 = K.win_t0, rows = K.win_rows;
    const double r_par = (r1 - r2) * c;
    const double r_trans = (d1 + d2) * s;
    const bool out = (r_par >= rpmax) | (r_trans >= rtmax) | (r_par <= rpmin);
    const bool edge_out = (r_par - rpmax < tol) & (rpmin - r_par < tol) & (r_trans - rtmax < tol);
    const double t_par = (r_par - rpmin) / span * np_d;
    const double t_trans = r_trans / rtmax * nt_d;
    const double f_par = floor(t_par);
    const double f_trans = floor(t_trans);
    const bool edge_in = ((r_par != rpmin) & (edge_distance(t_par, width_p) < tol)) |
                         ((r_trans != 0) & (edge_distance(t_trans, width_t) < tol));
    const double bin_f = f_trans + nt_d * f_par;
    const bool over = !((bin_f >= 0.0) & (bin_f < nb_d));
    // r_trans row of the pair for the window ownership of its statistics
    const double row_f = r_trans >= rtmax ? nt_d : fmin(fmax(f_trans, 0.0), nt_d);
    const int row = (int)row_f;
    const bool mine = (row >= t0 && row < t0 + rows) || (K.win_last && row >= t0 + rows);
    const bool edge = out ? edge_out : edge_in;
    if (edge & mine)
        __hip_atomic_fetch_add(&lds_stats[XI_STAT_EDGE], 1ull, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
    if (!out & over & mine)
        __hip_atomic_fetch_add(&lds_stats[XI_STAT_OVERFLOW], 1ull, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
    // accepted: row f_par, local column f_trans - t0 (a bin whose column rounds to nt is
    // bin 0 of the next r_par row, as in the reference's flat index)
    double fp = f_par, ft = f_trans;
    if (ft >= nt_d) {
        ft -= nt_d;
        fp += 1.0;
    }
    const double lc = ft - (double)t0;
    const bool in_win = (lc >= 0.0) & (lc < (double)rows);
    const double pb = (fp + 1.0) * (double)(rows + 2) + (lc + 1.0);
    return (out | over | !in_win) ? trash : (int)fmin(fmax(pb, 0.0), (double)trash);
}

// One task of the fast loop: quasars [64 chunk, 64 chunk + 64) of group g against the
// pruned pixel window of forest a.
__device__ void xcf_task_fast(const XiConst &C, const XiConst &C_lds, const ForestView &F,
                              const XcfView &X, int64_t p0, int n, int chunk, int a, int lane,
                              LocalBin *hist, LaneCounters &cnt, LocalU64x *lds_stats,
                              uint64_t &n_steps) {
    const int k = chunk * 64 + lane;
    bool active = k < n;
    XcfPairRecord o{0, 0, 0, 0, 1, 0};
    if (active) {
        o = X.rec[X.idx ? X.idx[p0 + k] : p0 + k];
        active = o.w != 0;
    }
    if (__ballot(active) == 0) return;
    const int64_t oa = uniform_i64(F.off[a]);
    const int la = (int)(uniform_i64(F.off[a + 1]) - oa);
    int i_lo = 0, i_hi = la;
    if ((uniform_int(F.mono[a]) & 3) == 3) {
        const bool pos = active && o.c > 0;
        if (__ballot(active && !pos) == 0) {
            double lo = wave_min(pos ? o.r + C.rpmin / o.c : INFINITY);
            double hi = wave_max(pos ? o.r + C.rpmax / o.c : -INFINITY);
            lo -= 1e-6 + 1e-12 * fabs(lo);
            hi += 1e-6 + 1e-12 * fabs(hi);
            i_lo = wave_search<true>(F.r + oa, la, lo, lane);
            i_hi = wave_search<false>(F.r + oa, la, hi, lane);
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
    // per-lane factors: bin units per Mpc/h of r1 - r2 and d1 + d2; inactive lanes are
    // moved far beyond r_par_max so that their zero-weight pairs land in the padding
    const double cpl = o.c * C.inv_p;
    const double stl = o.s * C.inv_t;
    const double r2 = active ? o.r : -1e30;
    const double w2 = active ? o.w : 0.0;
    const double zh2 = 0.5 * o.z;
    const double gpo = 1.0 + C.fast_p - C.rpmin * C.inv_p;   // row 1 is r_par bin 0
    const double gto = 1.0 + C.fast_t - (double)C.win_t0;     // column 1 is row win_t0
    const double mp2 = 2.0 * C.fast_p, mt2 = 2.0 * C.fast_t;
    const double cap_p = C.np_d + 1.0, cap_t = (double)C.win_rows + 1.0;
    const int stride = C.win_rows + 2;
    const int trash = (C.np + 2) * stride - 1;
    XRun run;
    run.bin = trash;
    run.count = 0.0;
    run.we = run.xi = run.rp = run.rt = run.z = 0.0;
    uint32_t steps = 0;
    const PixelRecord *rec = F.pix + oa;
    PixelRecord c0 = load_record(rec + i_lo), c1 = load_record(rec + i_lo + 1);
    asm volatile("" ::"s"(c0.w), "s"(c0.z), "s"(c1.w), "s"(c1.z));
    auto step = [&](const PixelRecord &q) {
        // the reference skips w1 == 0 (either sign) before the quasar loop
        if (xzero_weight(q.w)) return;
        ++steps;
        const double diff = q.r - r2;
        const double sum = q.d + o.d;
        const double gp = fma(diff, cpl, gpo);
        const double gt = fma(sum, stl, gto);
        const bool near = (__builtin_amdgcn_fract(gp) < mp2) | (__builtin_amdgcn_fract(gt) < mt2);
        int bin = (int)__umul24((unsigned)stride, (unsigned)(int)xclamp(gp, 0.0, cap_p)) +
                  (int)xclamp(gt, 0.0, cap_t);
        if (active & near)
            bin = exact_xcf(C_lds, q.r, q.d, o.r, o.d, o.c, o.s, trash, lds_stats);
        const double w12 = q.w * w2;
        const double dw12 = q.dw * w2;
        const double z = fma(q.z, 0.5, zh2);
        const bool change = bin != run.bin;
        if (change) xflush(run, hist, o.c, o.s);
        const double keep = change ? 0.0 : 1.0;
        run.bin = bin;
        run.count = fma(run.count, keep, 1.0);
        run.we = fma(run.we, keep, w12);
        run.xi = fma(run.xi, keep, dw12);
        run.rp = fma(diff, w12, run.rp * keep);
        run.rt = fma(sum, w12, run.rt * keep);
        run.z = fma(z, w12, run.z * keep);
    };
    for (int i = i_lo; i < i_hi; i += 2) {
        const PixelRecord n0 = load_record(rec + i + 2), n1 = load_record(rec + i + 3);
        step(c0);
        if (i + 1 < i_hi) step(c1);
        c0 = n0;
        c1 = n1;
    }
    xflush(run, hist, o.c, o.s);
    n_steps += steps;
    if (active) cnt.visit += steps;
}

}  // namespace

// Adds the real bins of the padded window histogram to output row `row` in HBM and clears
// them; returns the pairs this thread moved.
__device__ __forceinline__ uint64_t xflush_window(const XiConst &C, LocalBin *hist, XiBin *out,
                                                  int tid) {
    uint64_t moved = 0;
    const int rows = C.win_rows, stride = rows + 2;
    for (int k = tid; k < C.np * rows; k += XI_BLOCK) {
        const int ip = k / rows, lc = k - ip * rows;
        const int kp = (ip + 1) * stride + lc + 1;
        const uint32_t c = hist[kp].nb;
        if (c) {
            XiBin *o = out + ip * C.nt + C.win_t0 + lc;
            unsafeAtomicAdd(&o->we, hist[kp].we);
            unsafeAtomicAdd(&o->xi, hist[kp].xi);
            unsafeAtomicAdd(&o->rp, hist[kp].rp);
            unsafeAtomicAdd(&o->rt, hist[kp].rt);
            unsafeAtomicAdd(&o->z, hist[kp].z);
            atomicAdd(&o->nb, (unsigned long long)c);
            hist[kp].we = hist[kp].xi = hist[kp].rp = hist[kp].rt = hist[kp].z = 0.0;
            hist[kp].nb = 0;
            moved += c;
        }
    }
    return moved;
}

// FAST: the fast loop (standard binning, no pair-redshift cut, padded window histogram in
// LDS).  Otherwise LDS_HIST selects the plain np x nt LDS histogram or the HBM rows.
template <int MODE, bool LDS_HIST, bool FAST>
__global__ __launch_bounds__(XI_BLOCK) void xcf_kernel(XiConst C, ForestView F, XcfView X,
                                                       XiOut O) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int64_t s_group;
    __shared__ int s_task;
    __shared__ unsigned long long s_stats[XI_NSTATS];
    __shared__ XiConst s_const;
    LocalU64x *lds_stats = (LocalU64x *)s_stats;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int nb = C.nb;
    const int nb_lds = FAST ? (C.np + 2) * (C.win_rows + 2) : nb;
    LocalBin *hist = (LocalBin *)reinterpret_cast<LdsBin *>(smem);
    if (LDS_HIST) {
        for (int k = tid; k < nb_lds; k += XI_BLOCK) {
            hist[k].we = hist[k].xi = hist[k].rp = hist[k].rt = hist[k].z = 0.0;
            hist[k].nb = 0;
        }
    }
    if (tid < XI_NSTATS) s_stats[tid] = 0;
    if (tid == 0) s_const = C;   // published by the first __syncthreads of the group loop
    LaneCounters cnt{0, 0, 0, 0};
    uint64_t accepted = 0;
    uint64_t n_steps = 0, n_tasks = 0, n_empty = 0;
    int current_row = -1;
    auto flush_hist = [&](int row) {
        XiBin *out = O.bins + (int64_t)row * nb;
        if (FAST) {
            accepted += xflush_window(C, hist, out, tid);
            return;
        }
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
    // thread 0's claimed run of groups [next, end): one queue ticket claims a run of groups
    // of one row (X.chunk), so the histogram moves to HBM at most once per run
    int64_t next = 0, end = 0;
    for (;;) {
        __syncthreads();
        if (tid == 0) {
            if (next >= end) {
                const int64_t c = (int64_t)atomicAdd(O.queue, 1ull);
                next = c < X.nchunks ? X.chunk[c] : X.count;
                end = c < X.nchunks ? X.chunk[c + 1] : X.count + 1;
            }
            s_group = next++;
            s_task = 0;
        }
        __syncthreads();
        const int64_t g = s_group;
        if (g >= X.count) break;
        const int row = X.row[g];
        if (LDS_HIST && row != current_row && current_row >= 0) {
            flush_hist(current_row);
            // no wave may add to a bin before every thread has moved and cleared its bins
            __syncthreads();
        }
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
            if (FAST)
                xcf_task_fast(C, s_const, F, X, p0, n, t, a, lane, hist, cnt, lds_stats,
                              n_steps);
            else if (LDS_HIST)
                xcf_task<MODE>(C, F, X, p0, n, t, a, lane, hist, cnt, n_steps);
            else
                xcf_task<MODE>(C, F, X, p0, n, t, a, lane, O.bins + (int64_t)row * nb, cnt,
                               n_steps);
            ++n_tasks;
            if (n_steps == before) ++n_empty;
        }
    }
    if (LDS_HIST && current_row >= 0) flush_hist(current_row);
    if (!FAST) accepted = cnt.accept;
    const uint64_t v_visit = wave_sum_u64(cnt.visit);
    const uint64_t v_edge = wave_sum_u64(cnt.edge);
    const uint64_t v_over = wave_sum_u64(cnt.overflow);
    const uint64_t v_accept = wave_sum_u64(accepted);
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

template <int MODE, bool LDS_HIST, bool FAST>
static hipError_t launch_xcf_t(const XiConst &C, const ForestView &F, const XcfView &X,
                               const XiOut &O, int grid, size_t lds, hipStream_t stream) {
    auto kern = xcf_kernel<MODE, LDS_HIST, FAST>;
    hipError_t err = hipFuncSetAttribute((const void *)kern,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (err != hipSuccess) return err;
    kern<<<dim3(grid), dim3(XI_BLOCK), lds, stream>>>(C, F, X, O);
    return hipGetLastError();
}

// fast: standard binning without pair-redshift cuts, window histogram in LDS (the host
// decides); otherwise lds_hist selects the plain LDS histogram or the HBM rows.
hipError_t launch_xcf(int mode, bool lds_hist, bool fast, const XiConst &C, const ForestView &F,
                      const XcfView &X, const XiOut &O, int grid, size_t lds,
                      hipStream_t stream) {
    if (mode == MODE_STANDARD) {
        if (fast) return launch_xcf_t<MODE_STANDARD, true, true>(C, F, X, O, grid, lds, stream);
        return lds_hist ? launch_xcf_t<MODE_STANDARD, true, false>(C, F, X, O, grid, lds, stream)
                        : launch_xcf_t<MODE_STANDARD, false, false>(C, F, X, O, grid, lds, stream);
    }
    if (mode == MODE_RMU)
        return lds_hist ? launch_xcf_t<MODE_RMU, true, false>(C, F, X, O, grid, lds, stream)
                        : launch_xcf_t<MODE_RMU, false, false>(C, F, X, O, grid, lds, stream);
    return lds_hist ? launch_xcf_t<MODE_ANGULAR, true, false>(C, F, X, O, grid, lds, stream)
                    : launch_xcf_t<MODE_ANGULAR, false, false>(C, F, X, O, grid, lds, stream);
}

}  // namespace picca
