// CPU harness of the pair kernel's run walk (test infrastructure, no product path).
//
// Runs the lane program of picca_amd/csrc/xi_walk.h -- the code the kernel runs -- on the
// host over a list of forest pairs, with the kernel's task split (64-pixel chunks of delta1,
// the pruned window of delta2, the same eligibility test) and every memory access checked
// against the forest's bounds.  Tasks the kernel would not walk are evaluated pair by pair
// with the reference's arithmetic.  The tests compare the result with the oracle, so the
// walk's keys, probes, prefix sums and edge pixels are checked on the CPU.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../picca_amd/csrc/xi_const.h"
#include "../../picca_amd/csrc/xi_walk.h"

using picca::PrefixRecord;
using picca::XiConst;

namespace {

enum { ST_EDGE, ST_OVER, ST_WALK_TASKS, ST_PAIR_TASKS, ST_RUNS, ST_PROBES, ST_OOB, ST_COVER,
       ST_N };

struct Hist {
    std::vector<double> we, xi, rp, rt, z;
    std::vector<int64_t> nb;
    explicit Hist(size_t n) : we(n), xi(n), rp(n), rt(n), z(n), nb(n) {}
};

// Memory of one walk task: delta2's pixels [0, lb) and prefix entries [0, lb]; an access
// outside them is counted and answered with a harmless value.
struct HostMem {
    const double *rp_, *dp_, *wp_, *zp_, *dwp_;
    const PrefixRecord *pre_;
    int64_t lb;
    Hist *hist;
    int64_t *stats;

    bool bad(int p, int64_t n) {
        if (p < 0 || p >= n) {
            ++stats[ST_OOB];
            return true;
        }
        return false;
    }
    void rd(int p, double &rv, double &dv) {
        if (bad(p, lb)) {
            rv = dv = 0.0;
            return;
        }
        rv = rp_[p];
        dv = dp_[p];
    }
    PrefixRecord prefix(int p) {
        if (bad(p, lb + 1)) return PrefixRecord{0, 0, 0, 0, 0, 0};
        return pre_[p];
    }
    double w(int p) { return bad(p, lb) ? 0.0 : wp_[p]; }
    double z(int p) { return bad(p, lb) ? 0.0 : zp_[p]; }
    double dw(int p) { return bad(p, lb) ? 0.0 : dwp_[p]; }
    void add(int bin, double we, double xi, double rp, double rt, double zz, uint32_t n) {
        if (bin < 0 || (size_t)bin >= hist->we.size()) {
            ++stats[ST_OOB];
            return;
        }
        hist->we[bin] += we;
        hist->xi[bin] += xi;
        hist->rp[bin] += rp;
        hist->rt[bin] += rt;
        hist->z[bin] += zz;
        hist->nb[bin] += n;
    }
    void flush(bool on, const picca::RunSums &r) {
        if (on) add(r.bin, r.we, r.xi, r.rp, r.rt, r.z, r.n);
    }
    bool any(bool v) { return v; }
    void pin(picca::RunSums &) {}
    void count_edge() { ++stats[ST_EDGE]; }
    void count_overflow() { ++stats[ST_OVER]; }
    void count_probe() { ++stats[ST_PROBES]; }
    void count_run(uint32_t n) {
        ++stats[ST_RUNS];
        stats[ST_COVER] += (int64_t)n;
    }
};

bool monotone(const double *x, int64_t n) {
    for (int64_t k = 1; k < n; ++k)
        if (!(x[k] >= x[k - 1])) return false;
    return true;
}

// first index in [0, n) with arr[idx] >= x (strict) or arr[idx] > x, arr non-decreasing
int lower(const double *arr, int n, double x, bool strict) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        const bool below = strict ? arr[mid] < x : arr[mid] <= x;
        if (below) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// One pixel pair with the reference's arithmetic (py/picca/cf.py:355-387), specialised
// case; same edge / overflow conventions as the kernel.
void exact_pair(const XiConst &C, double r1, double d1, double zh1, double w1, double dw1,
                double r2, double d2, double z2, double w2, double dw2, double cos_half,
                double sin_half, Hist &h, int64_t *stats) {
    const picca::WalkExact e = picca::walk_exact(C, r1, d1, r2, d2, cos_half, sin_half);
    if (e.edge) ++stats[ST_EDGE];
    if (e.over) ++stats[ST_OVER];
    const bool out = (e.r_par >= C.rpmax) | (e.r_trans >= C.rtmax);
    if (out || e.over) return;
    int pbin;
    if (e.wrap) {
        pbin = e.pbin;
    } else {
        const int row = e.m >= 0 ? e.m : -e.m - 1;
        pbin = row * C.ntp + e.c;
    }
    const double w12 = w1 * w2;
    h.we[pbin] += w12;
    h.xi[pbin] += dw1 * dw2;
    h.rp[pbin] += e.r_par * w12;
    h.rt[pbin] += e.r_trans * w12;
    h.z[pbin] += (zh1 + z2 * 0.5) * w12;
    h.nb[pbin] += 1;
}

}  // namespace

extern "C" {

// The walk's guess reciprocal (walk_rcp_guess), for its unit test.
double xi_walk_rcp_guess(double x) { return picca::walk_rcp_guess(x); }

// Correlation of the listed forest pairs (specialised case of the pair kernel only).
// Outputs are [num_rows * np * nt] each; stats = edge, overflow, walked tasks, pair-by-pair
// tasks, runs, probes, out-of-bounds accesses, pixel pairs covered by runs.  Returns 0, or
// -1 when the parameters are not the specialised case.
int xi_walk_host(const picca_cf_params *params, int64_t num_forests, const int64_t *off,
                 const double *z, const double *r, const double *d, const double *w,
                 const double *delta, int64_t num_pairs, const int32_t *pa, const int32_t *pb,
                 const double *pang, const int32_t *prow, int32_t num_rows, int32_t schedule,
                 double *o_we, double *o_xi, double *o_rp, double *o_rt, double *o_z,
                 int64_t *o_nb, int64_t *stats) {
    XiConst C;
    int mode = 0;
    if (picca::fill_xi_const(params, C, mode) || !picca::xi_specialised(params, mode))
        return -1;
    const int64_t npix = off[num_forests];
    double max_r = 0, max_d = 0;
    for (int64_t k = 0; k < npix; ++k) {
        max_r = std::fmax(max_r, std::fabs(r[k]));
        max_d = std::fmax(max_d, std::fabs(d[k]));
    }
    if (picca::widen_xi_margins(C, max_r, max_d)) return -1;
    picca::apply_xi_schedule(C, schedule);

    std::vector<double> dw((size_t)npix);
    for (int64_t k = 0; k < npix; ++k) dw[k] = delta[k] * w[k];
    std::vector<PrefixRecord> pre((size_t)(npix + num_forests));
    std::vector<double> spc_r((size_t)num_forests), spc_d((size_t)num_forests);
    std::vector<uint8_t> mono((size_t)num_forests);
    for (int64_t b = 0; b < num_forests; ++b) {
        const int64_t o = off[b], n = off[b + 1] - off[b];
        picca::build_walk_forest(r + o, d + o, z + o, w + o, dw.data() + o, n,
                                 pre.data() + o + b, spc_r[b], spc_d[b]);
        mono[b] = (uint8_t)((monotone(r + o, n) ? 1 : 0) | (monotone(d + o, n) ? 2 : 0));
    }
    std::memset(stats, 0, sizeof(int64_t) * ST_N);
    const size_t nbp = (size_t)(C.np + 1) * C.ntp;
    std::vector<Hist> rows;
    rows.reserve((size_t)num_rows);
    for (int32_t k = 0; k < num_rows; ++k) rows.emplace_back(nbp);

    for (int64_t p = 0; p < num_pairs; ++p) {
        const int a = pa[p], b = pb[p];
        const double ang = pang[p];
        const double cos_half = std::cos(ang / 2), sin_half = std::sin(ang / 2);
        Hist &h = rows[(size_t)prow[p]];
        const int64_t oa = off[a], ob = off[b];
        const int la = (int)(off[a + 1] - oa), lb = (int)(off[b + 1] - ob);
        // the kernel's task list: the chunks of delta1 xi_chunk_kernel keeps
        int first = 0, end = (la + 63) >> 6;
        if (cos_half > 0 && (mono[b] & 3) == 3 && (mono[a] & 1) && la > 0 && lb > 0) {
            const double reach = std::fmax(C.rpmax, std::fabs(C.rpmin)) / cos_half;
            const double slack = 8.0 * C.edge_tol / cos_half + 2e-6;
            double lo = r[ob] - reach, hi = r[ob + lb - 1] + reach;
            lo -= slack + 2e-12 * std::fabs(lo);
            hi += slack + 2e-12 * std::fabs(hi);
            const int i_lo = lower(r + oa, la, lo, true);
            int i_hi = lower(r + oa, la, hi, false);
            if (sin_half > 0 && (mono[a] & 2)) {
                double d_lim = (C.rtmax + 8.0 * C.edge_tol) / sin_half - d[ob];
                d_lim += 2e-6 + 2e-12 * std::fabs(d_lim);
                i_hi = std::min(i_hi, lower(d + oa, la, d_lim, false));
            }
            first = i_lo >= i_hi ? 0 : i_lo >> 6;
            end = i_lo >= i_hi ? 0 : (i_hi + 63) >> 6;
        }
        for (int chunk = first; chunk < end; ++chunk) {
            const int i0 = chunk * 64, i1 = std::min(la, i0 + 64);
            double r_min = INFINITY, r_max = -INFINITY, d_min = INFINITY;
            bool any = false;
            for (int i = i0; i < i1; ++i) {
                if (w[oa + i] == 0) continue;
                any = true;
                r_min = std::fmin(r_min, r[oa + i]);
                r_max = std::fmax(r_max, r[oa + i]);
                d_min = std::fmin(d_min, d[oa + i]);
            }
            if (!any) continue;
            int j_lo = 0, j_hi = lb;
            bool pruned = false;
            if (cos_half > 0 && (mono[b] & 3) == 3) {
                pruned = true;
                const double slack = 4.0 * C.edge_tol / cos_half + 1e-6;
                double lo_r = r_min - C.rpmax / cos_half;
                double hi_r = r_max + C.rpmax / cos_half;
                lo_r -= slack + 1e-12 * std::fabs(lo_r);
                hi_r += slack + 1e-12 * std::fabs(hi_r);
                j_lo = lower(r + ob, lb, lo_r, true);
                j_hi = lower(r + ob, lb, hi_r, false);
                if (sin_half > 0) {
                    double d_lim = (C.rtmax + 4.0 * C.edge_tol) / sin_half - d_min;
                    d_lim += 1e-6 + 1e-12 * std::fabs(d_lim);
                    const int j_d = lower(d + ob, lb, d_lim, false);
                    if (j_d < j_hi) j_hi = j_d;
                }
            }
            // an empty window: the kernel returns before choosing a loop
            if (j_lo >= j_hi) continue;
            const double cp = cos_half * C.inv_p, st = sin_half * C.inv_t;
            double q0 = 1.0;
            if (j_hi - j_lo >= 2) {
                const double span = (r[ob + j_hi - 1] - r[ob + j_lo]) * cp;
                if (span > 0) q0 = (double)(j_hi - 1 - j_lo) / span;
            }
            const bool walk = pruned && picca::walk_task(C, spc_r[b], spc_d[b], cp, st, q0);
            if (walk) {
                ++stats[ST_WALK_TASKS];
                picca::WalkTask T;
                T.cos_half = cos_half;
                T.sin_half = sin_half;
                T.cp = cp;
                T.st = st;
                T.rc = r[ob + (lb >> 1)];
                T.dc = d[ob + (lb >> 1)];
                T.j_lo = j_lo;
                T.j_hi = j_hi;
                T.q0 = q0;
                HostMem mem;
                mem.rp_ = r + ob;
                mem.dp_ = d + ob;
                mem.wp_ = w + ob;
                mem.zp_ = z + ob;
                mem.dwp_ = dw.data() + ob;
                mem.pre_ = pre.data() + ob + b;
                mem.lb = lb;
                mem.hist = &h;
                mem.stats = stats;
                for (int i = i0; i < i1; ++i) {
                    if (w[oa + i] == 0) continue;
                    const picca::WalkLane L{r[oa + i], d[oa + i], z[oa + i] * 0.5, w[oa + i],
                                            dw[oa + i]};
                    picca::walk_lane(C, T, L, mem, true);
                }
            } else {
                ++stats[ST_PAIR_TASKS];
                for (int i = i0; i < i1; ++i) {
                    if (w[oa + i] == 0) continue;
                    for (int j = j_lo; j < j_hi; ++j) {
                        if (w[ob + j] == 0) continue;
                        exact_pair(C, r[oa + i], d[oa + i], z[oa + i] * 0.5, w[oa + i],
                                   dw[oa + i], r[ob + j], d[ob + j], z[ob + j], w[ob + j],
                                   dw[ob + j], cos_half, sin_half, h, stats);
                    }
                }
            }
        }
    }
    // padded layout -> flat rows
    for (int32_t row = 0; row < num_rows; ++row) {
        const Hist &h = rows[(size_t)row];
        for (int k = 0; k < C.nb; ++k) {
            const int kp = k + k / C.nt;
            const size_t o = (size_t)row * C.nb + k;
            o_we[o] = h.we[kp];
            o_xi[o] = h.xi[kp];
            o_rp[o] = h.rp[kp];
            o_rt[o] = h.rt[kp];
            o_z[o] = h.z[kp];
            o_nb[o] = h.nb[kp];
        }
    }
    return 0;
}

}  // extern "C"
