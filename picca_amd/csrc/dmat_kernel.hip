// Distortion matrix on gfx950 (MI355X): compute_dmat_forest_pairs_fast
// (py/picca/cf.py:520-887) for a batch of kept forest pairs.
//
// The reference's fill loop (py/picca/cf.py:846-887) adds, for every selected pixel pair
// (i, j) in data bin A and every model bin k of the pair's unique model bins U,
//     w1 w2 (eta5 + eta6 dll2 + eta7 dll1 + eta8 dll1 dll2 - eta1[i] - eta2[j]
//            - eta3[i] dll2 - eta4[j] dll1)[k]
// to dmat[A, k], plus w1 w2 zfac to dmat[A, M(i, j)].  Grouping the sum by pixel gives
//     dmat[A, k] += S0[A] eta5[k] + Sj[A] eta6[k] + Si[A] eta7[k] + Sij[A] eta8[k]
//                 - sum_i (P[A, i] eta1[i, k] + P'[A, i] eta3[i, k])
//                 - sum_j (Q[A, j] eta2[j, k] + Q'[A, j] eta4[j, k])
// with P[A, i] = sum_j w1 w2, P' = sum_j w1 w2 dll2 over the selected pairs of pixel i in
// bin A (Q, Q' the same over i for pixel j) and S = the sums of P, P'.  eta1[i, .] and
// P[., i] are non-zero only in the one or two r_trans rows that pixel i's pairs reach, so
// the products are dense in r_par and banded in r_trans.
//
// Two kernels per batch:
//   dmat_pass_kernel  one workgroup per forest pair, lane per pixel i: selection, bins,
//                     eta5..eta8 and S in LDS bands, the effective-coordinate sums and
//                     weights_dmat, the unique model bins U (with the reference's
//                     count-pass truncation), and for each r_trans row the range of
//                     pixels that reach it;
//   dmat_fill_kernel  one workgroup per (data r_trans row t, model r_trans row u, slot
//                     group): an np x npm block of dmat held in the matrix cores'
//                     accumulators, adding the S x eta rank-4 term, and per four pixels i
//                     (then j) the vectors P, P', eta1, eta3 (Q, Q', eta2, eta4) built by
//                     walking those pixels' pairs, followed by P^T eta1 + P'^T eta3 as two
//                     K = 4 FP64 products (v_mfma_f64_16x16x4f64, K = the pixels); one
//                     atomic add per block entry.
// The pair pass walks every pixel pair once when the pair's r_trans rows fit its LDS bands
// (census and sums together), else twice (census, then the sums band by band).
// All arithmetic is FP64; bins follow the reference's divide-then-multiply.
//
// The forest x quasar distortion matrix (compute_dmat_forest_pairs_fast of
// py/picca/xcf.py:427-658) runs through the same kernels: the quasar is a one-pixel
// delta2 with order < 0, which removes every term of delta2's continuum fit (eta1, eta3,
// eta5..eta8) and leaves the identity, eta2 (w1 / sum w1) and eta4 terms of the
// reference; its pair flag PAIR_XCF_COUNT gives the count pass of xcf.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <limits.h>

#include "device_util.h"
#include "engine.h"

namespace picca {

namespace {

struct DmatPair {
    int a, b, la, lb;
    int64_t oa, ob;
    double ang, cos_half, sin_half;
    double zq1, zq2, sum_w1, sum_w2, sw1, sw2;
    // reciprocals of the continuum-fit normalisations: the eta terms multiply by these
    // instead of dividing per pixel pair (a few ulp from the reference's divisions)
    double inv_w1, inv_w2, inv_s1, inv_s2;
    bool shp_cut;   // remove_same_half_plate_close_pairs and the pair is on one half plate
    bool zerr_on;   // zerr cut active for this pair (ang < zerr_cut_deg)
    bool order1, order2;
    bool proj2;     // delta2 is projected (order >= 0); a quasar is not (order < 0)
    bool strict;    // cross-correlation count pass: r_par == r_par_min is not counted
    bool prune;     // standard binning, both forests monotone, cos(ang/2) > 0
};

__device__ DmatPair load_pair(const DmatConst &C, const DmatForests &F, const DmatPairs &P,
                              int s) {
    DmatPair q;
    q.a = uniform_int(P.a[s]);
    q.b = uniform_int(P.b[s]);
    q.oa = uniform_i64(F.off[q.a]);
    q.la = (int)(uniform_i64(F.off[q.a + 1]) - q.oa);
    q.ob = uniform_i64(F.off[q.b]);
    q.lb = (int)(uniform_i64(F.off[q.b + 1]) - q.ob);
    q.ang = uniform_f64(P.ang[s]);
    q.cos_half = uniform_f64(P.cs[2 * s]);
    q.sin_half = uniform_f64(P.cs[2 * s + 1]);
    q.zq1 = uniform_f64(F.zq[q.a]);
    q.zq2 = uniform_f64(F.zq[q.b]);
    q.sum_w1 = uniform_f64(F.sum_w[q.a]);
    q.sum_w2 = uniform_f64(F.sum_w[q.b]);
    q.sw1 = uniform_f64(F.sw_dll[q.a]);
    q.sw2 = uniform_f64(F.sw_dll[q.b]);
    q.inv_w1 = 1.0 / q.sum_w1;
    q.inv_w2 = 1.0 / q.sum_w2;
    q.inv_s1 = 1.0 / q.sw1;
    q.inv_s2 = 1.0 / q.sw2;
    q.shp_cut = (C.flags & XI_REMOVE_SHP) &&
                (uniform_int(P.flags[s]) & PAIR_SAME_HALF_PLATE) != 0;
    q.zerr_on = (C.flags & XI_HAS_ZERR) && (q.ang < C.zerr_ang);
    q.order1 = uniform_int(F.order[q.a]) == 1;
    q.order2 = uniform_int(F.order[q.b]) == 1;
    q.proj2 = uniform_int(F.order[q.b]) >= 0;
    q.strict = (uniform_int(P.flags[s]) & PAIR_XCF_COUNT) != 0;
    q.prune = !(C.flags & DMAT_RMU) && q.cos_half > 0 &&
              (uniform_int(F.mono[q.a]) & 3) == 3 && (uniform_int(F.mono[q.b]) & 3) == 3;
    return q;
}

// zerr test of py/picca/cf.py:628-635 / :647-654: true when the pixel is cut
__device__ __forceinline__ bool zerr_cut(const DmatConst &C, double z, double z_qso) {
    double z_qf = 0.5 * (z + z_qso);
    double dv_kms = fabs(z - z_qso) / (1 + z_qf);
    dv_kms *= SPEED_LIGHT_KMS;
    return dv_kms < C.zerr_kms;
}

// Data and model bins of one in-range pixel pair: the flat indices A = bt + nt bp and
// M = mt + ntm mp of the reference, and their row (r_par) and column (r_trans) parts as the
// reference's flat index reads them (a column bin that rounds to nt is column 0 of the next
// row).
struct PairBins {
    int A, M;
    int tA, rA;   // A % nt, A / nt
    int uM, mM;   // M % ntm, M / ntm
};

// Separations, range test and data/model bins of one pixel pair (py/picca/cf.py:656-700).
// The bin coordinates x np, y nt, x npm, y ntm (x = (r_par - r_par_min) / span, y =
// r_trans / r_trans_max) are taken with reciprocal multiplies, which differ from the
// reference's divisions by a few ulp; a coordinate t within 1e-12 (1 + t) of an integer
// (far above that difference) is redone with the divisions, so every pair gets the
// reference's bins.
// An exact zero is the same in both forms.
__device__ __forceinline__ bool pair_bins(const DmatConst &C, double r1, double d1, double r2,
                                          double d2, double cos_half, double sin_half,
                                          double &r_par, double &r_trans, PairBins &B) {
    r_par = (r1 - r2) * cos_half;
    r_trans = (d1 + d2) * sin_half;
    if (C.flags & DMAT_RMU) {
        r_trans = sqrt(r_trans * r_trans + r_par * r_par);
        r_par /= r_trans;
    }
    if (!(C.flags & DMAT_XCORR)) r_par = fabs(r_par);
    if (r_par >= C.rpmax || r_trans >= C.rtmax || r_par < C.rpmin) return false;
    double x = (r_par - C.rpmin) * C.inv_span_p;
    double y = r_trans * C.inv_rtmax;
    double tp = x * C.np_d, tt = y * C.nt_d, up = x * C.npm_d, ut = y * C.ntm_d;
    // the two forms differ by a few ulp of the coordinate (< 1e-15 t); t >= 0
    auto near_int = [](double t) { return fabs(t - rint(t)) < 1e-12 * (1.0 + t); };
    const bool near = (x != 0 && (near_int(tp) || near_int(up))) ||
                      (y != 0 && (near_int(tt) || near_int(ut)));
    if (near) {
        x = (r_par - C.rpmin) / C.span_p;
        y = r_trans / C.rtmax;
        tp = x * C.np_d;
        tt = y * C.nt_d;
        up = x * C.npm_d;
        ut = y * C.ntm_d;
    }
    const double bp = floor(tp), bt = floor(tt), mp = floor(up), mt = floor(ut);
    B.A = (int)(bt + C.nt_d * bp);
    B.M = (int)(mt + C.ntm_d * mp);
    // bt <= nt and mt <= ntm (y < 1 up to rounding); x >= 0 and y >= 0
    const bool wrap_t = bt >= C.nt_d, wrap_u = mt >= C.ntm_d;
    B.tA = wrap_t ? 0 : (int)bt;
    B.rA = (int)bp + (wrap_t ? 1 : 0);
    B.uM = wrap_u ? 0 : (int)mt;
    B.mM = (int)mp + (wrap_u ? 1 : 0);
    return true;
}

// Pair-redshift, zerr (pixel j) and same-half-plate parts of the selection
// (py/picca/cf.py:640-676); the zerr part of pixel i is passed in.
__device__ __forceinline__ bool pair_selected(const DmatConst &C, const DmatPair &q, bool i_sel,
                                              bool j_zerr_cut, double z, double r_par) {
    bool sel = i_sel && !j_zerr_cut;
    if ((C.flags & XI_HAS_ZMIN) && z < C.zmin) sel = false;
    if ((C.flags & XI_HAS_ZMAX) && z > C.zmax) sel = false;
    if (q.shp_cut && fabs(r_par) < C.width_p) sel = false;
    return sel;
}

// Window [lo, hi) of the other forest's pixels outside which every pair with this wave's
// coordinates fails the range test.  `r_lo`/`r_hi` bound the partner's r_comov, `d_lim`
// its dist_m; the slack is far above the rounding of the separations.
__device__ void window(const double *__restrict__ Or, const double *__restrict__ Od, int n,
                       double r_lo, double r_hi, double d_lim, bool use_d, int lane, int &lo,
                       int &hi) {
    r_lo -= 1e-6 + 1e-12 * fabs(r_lo);
    r_hi += 1e-6 + 1e-12 * fabs(r_hi);
    lo = wave_search<true>(Or, n, r_lo, lane);
    hi = wave_search<false>(Or, n, r_hi, lane);
    if (use_d) {
        d_lim += 1e-6 + 1e-12 * fabs(d_lim);
        const int h = wave_search<false>(Od, n, d_lim, lane);
        if (h < hi) hi = h;
    }
    lo = uniform_int(lo);
    hi = uniform_int(hi);
}

// pixels of delta2 that pixels of delta1 with r_comov in [r_min, r_max] and dist_m >= d_min
// can pair with
__device__ void window_b(const DmatConst &C, const DmatPair &q, const DmatForests &F,
                         double r_min, double r_max, double d_min, int lane, int &lo,
                         int &hi) {
    lo = 0;
    hi = q.lb;
    if (!q.prune) return;
    if (!(r_min <= r_max)) {
        hi = 0;
        return;
    }
    const double r_lo = r_min - C.rpmax / q.cos_half;
    const double r_hi = (C.flags & DMAT_XCORR) ? r_max - C.rpmin / q.cos_half
                                               : r_max + C.rpmax / q.cos_half;
    window(F.r + q.ob, F.d + q.ob, q.lb, r_lo, r_hi, C.rtmax / q.sin_half - d_min,
           q.sin_half > 0, lane, lo, hi);
}

// pixels of delta1 that pixel j of delta2 (r2, d2) can pair with
__device__ void window_a(const DmatConst &C, const DmatPair &q, const DmatForests &F, double r2,
                         double d2, int lane, int &lo, int &hi) {
    lo = 0;
    hi = q.la;
    if (!q.prune) return;
    const double r_lo = (C.flags & DMAT_XCORR) ? r2 + C.rpmin / q.cos_half
                                               : r2 - C.rpmax / q.cos_half;
    const double r_hi = r2 + C.rpmax / q.cos_half;
    window(F.r + q.oa, F.d + q.oa, q.la, r_lo, r_hi, C.rtmax / q.sin_half - d2,
           q.sin_half > 0, lane, lo, hi);
}

// Narrows [lo, hi) to the other forest's pixels whose pairs with a pixel at dist_m d_self
// can have r_trans in data row t or in model row u: r_trans = (d_self + d_other) sin(ang/2)
// grows with the other forest's (monotone) dist_m, so the two rows bound d_other.  The
// slack is far above the rounding of the row edges and of r_trans.
__device__ void row_window(const DmatConst &C, const DmatPair &q, const double *__restrict__ Od,
                           int n, double d_self, int t, int u, int lane, int &lo, int &hi) {
    if (!q.prune || !(q.sin_half > 0) || lo >= hi) return;
    const double r_lo = fmin(t * C.rtmax / C.nt_d, u * C.rtmax / C.ntm_d);
    const double r_hi = fmax((t + 1) * C.rtmax / C.nt_d, (u + 1) * C.rtmax / C.ntm_d);
    const double d_lo = r_lo / q.sin_half - d_self;
    const double d_hi = r_hi / q.sin_half - d_self;
    const double slack = 1e-6 + 1e-9 * (fabs(r_hi / q.sin_half) + fabs(d_self));
    const int l = uniform_int(wave_search<true>(Od, n, d_lo - slack, lane));
    const int h = uniform_int(wave_search<false>(Od, n, d_hi + slack, lane));
    lo = max(lo, l);
    hi = min(hi, h);
}

// The searches of window() and row_window() for one pixel, done together: thresholds 0, 1 on
// the partner's r_comov (first index >= x0, first index > x1), 2..4 on its dist_m (first
// > x2, first >= x3, first > x4; only the first `k` thresholds are searched).  One round of
// 64 probes serves every threshold, then one probe round per threshold inside its
// sub-interval, all issued together: two dependent memory rounds instead of two per
// threshold.  Forests longer than 4096 pixels take wave_search.
__device__ void search_rd(const double *__restrict__ R, const double *__restrict__ D, int n,
                          const double (&x)[5], int k, int lane, int (&res)[5]) {
    if (n > 4096) {
        res[0] = wave_search<true>(R, n, x[0], lane);
        res[1] = wave_search<false>(R, n, x[1], lane);
        if (k > 2) {
            res[2] = wave_search<false>(D, n, x[2], lane);
            res[3] = wave_search<true>(D, n, x[3], lane);
            res[4] = wave_search<false>(D, n, x[4], lane);
        }
        return;
    }
    const int stride = (n + 63) >> 6;   // <= 64
    const int probe = lane * stride;
    const bool ok = probe < n;
    double vr = 0.0, vd = 0.0;
    if (ok) {
        vr = R[probe];
        if (k > 2) vd = D[probe];
    }
    int sub_lo[5], sub_len[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        sub_lo[q] = 0;
        sub_len[q] = 0;
        if (q >= k) continue;
        const double v = q < 2 ? vr : vd;
        const bool below = ok && ((q == 0 || q == 3) ? v < x[q] : v <= x[q]);
        const int cnt = __popcll(__ballot(below));
        if (cnt > 0) {
            const int start = (cnt - 1) * stride + 1;
            const int stop = min(cnt * stride, n);
            sub_lo[q] = start;
            sub_len[q] = stop - start;   // < 64
        }
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        if (q >= k) continue;
        bool below = false;
        if (lane < sub_len[q]) {
            const double v = (q < 2 ? R : D)[sub_lo[q] + lane];
            below = (q == 0 || q == 3) ? v < x[q] : v <= x[q];
        }
        res[q] = uniform_int(sub_lo[q] + __popcll(__ballot(below)));
    }
}

// window_b / window_a followed by row_window for one pixel (r_comov r_self, dist_m d_self)
// of the other forest (Or, Od, n): the same [lo, hi), from search_rd.  `r_lo`, `r_hi` are
// the partner r_comov bounds of window_b / window_a before their slack.
__device__ void pixel_window(const DmatConst &C, const DmatPair &q, const double *__restrict__ Or,
                             const double *__restrict__ Od, int n, double r_lo, double r_hi,
                             double d_self, int t, int u, int lane, int &lo, int &hi) {
    lo = 0;
    hi = n;
    if (!q.prune) return;
    double x[5];
    x[0] = r_lo - (1e-6 + 1e-12 * fabs(r_lo));
    x[1] = r_hi + (1e-6 + 1e-12 * fabs(r_hi));
    const bool use_d = q.sin_half > 0;
    x[2] = x[3] = x[4] = 0.0;
    if (use_d) {
        const double d_lim = C.rtmax / q.sin_half - d_self;
        x[2] = d_lim + (1e-6 + 1e-12 * fabs(d_lim));
        const double r_row_lo = fmin(t * C.rtmax / C.nt_d, u * C.rtmax / C.ntm_d);
        const double r_row_hi = fmax((t + 1) * C.rtmax / C.nt_d, (u + 1) * C.rtmax / C.ntm_d);
        const double slack = 1e-6 + 1e-9 * (fabs(r_row_hi / q.sin_half) + fabs(d_self));
        x[3] = r_row_lo / q.sin_half - d_self - slack;
        x[4] = r_row_hi / q.sin_half - d_self + slack;
    }
    int res[5];
    search_rd(Or, Od, n, x, use_d ? 5 : 2, lane, res);
    lo = res[0];
    hi = res[1];
    if (use_d) {
        hi = min(hi, res[2]);
        lo = max(lo, res[3]);
        hi = min(hi, res[4]);
    }
}

__device__ __forceinline__ void lds_add(double *p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Records, for a run of consecutive partner pixels of lane pixel `self` that share row
// `row`, the pixel ranges of both forests.
__device__ __forceinline__ void note_run(int32_t *self_rng, int32_t *other_rng, int row,
                                         int self, int first, int last) {
    atomicMin(&self_rng[2 * row], self);
    atomicMax(&self_rng[2 * row + 1], self);
    atomicMin(&other_rng[2 * row], first);
    atomicMax(&other_rng[2 * row + 1], last);
}

struct RowRun {
    int row, first, last;
};

__device__ __forceinline__ void row_step(RowRun &run, int row, int j, int32_t *self_rng,
                                         int32_t *other_rng, int self) {
    if (row != run.row) {
        if (run.row >= 0) note_run(self_rng, other_rng, run.row, self, run.first, run.last);
        run.row = row;
        run.first = j;
    }
    run.last = j;
}

// Sums of one lane over consecutive pairs with the same bin, flushed to an LDS band.
struct RunA {
    int key;
    double s0, sj, si, sij;
};

struct RunM {
    int key;
    double e5, e6, e7, e8, rp, rt, z, w;
};

}  // namespace

__global__ __launch_bounds__(DMAT_PASS_BLOCK) void dmat_pass_kernel(DmatConst C, DmatForests F,
                                                                    DmatPairs P, DmatWork W,
                                                                    DmatOut O) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ unsigned long long s_cnt[6];   // in, sel, count_pass, overflow, negative
    __shared__ int s_band[4];

    const int s = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    constexpr int NWAVES = DMAT_PASS_BLOCK / 64;
    const DmatPair q = load_pair(C, F, P, s);

    uint32_t *lU = (uint32_t *)smem;
    int32_t *l_it = (int32_t *)(lU + C.words_u);
    int32_t *l_iu = l_it + 2 * C.nt;
    int32_t *l_jt = l_iu + 2 * C.ntm;
    int32_t *l_ju = l_jt + 2 * C.nt;
    const size_t fixed = ((size_t)C.words_u * 4 + (size_t)(4 * C.nt + 4 * C.ntm) * 4 + 15) & ~15;
    double *lS = (double *)(smem + fixed);            // [4][tb_max][np]
    double *lE = lS + 4 * C.tb_max * C.np;            // [8][ub_max][npm]

    for (int k = tid; k < C.words_u; k += DMAT_PASS_BLOCK) lU[k] = 0;
    for (int k = tid; k < 2 * C.nt + 2 * C.ntm; k += DMAT_PASS_BLOCK) {
        l_it[2 * k] = INT_MAX;
        l_it[2 * k + 1] = -1;
    }
    if (tid < 6) s_cnt[tid] = 0;
    __syncthreads();

    const double *__restrict__ Ar = F.r + q.oa;
    const double *__restrict__ Ad = F.d + q.oa;
    const double *__restrict__ Az = F.z + q.oa;
    const double *__restrict__ Aw = F.w + q.oa;
    const double *__restrict__ Al = F.dll + q.oa;
    const double *__restrict__ Af = F.zfa + q.oa;
    const double *__restrict__ Br = F.r + q.ob;
    const double *__restrict__ Bd = F.d + q.ob;
    const double *__restrict__ Bz = F.z + q.ob;
    const double *__restrict__ Bw = F.w + q.ob;
    const double *__restrict__ Bl = F.dll + q.ob;
    const double *__restrict__ Bf = F.zfb + q.ob;
    const bool evol = (C.flags & DMAT_REDSHIFT_EVOL) != 0;
    const int64_t pix = W.pix_off[s];

    // run flushes into the LDS bands that start at data row tc0 / model row uc0
    auto flush_a = [&](const RunA &ra, int tc0) {
        const int t = ra.key % C.nt;
        if (t < tc0 || t >= tc0 + C.tb_max) return;
        double *b = lS + (t - tc0) * C.np + ra.key / C.nt;
        const int stride = C.tb_max * C.np;
        lds_add(b, ra.s0);
        lds_add(b + stride, ra.sj);
        lds_add(b + 2 * stride, ra.si);
        lds_add(b + 3 * stride, ra.sij);
    };
    auto flush_m = [&](const RunM &rm, int uc0) {
        const int u = rm.key % C.ntm;
        if (u < uc0 || u >= uc0 + C.ub_max) return;
        double *b = lE + (u - uc0) * C.npm + rm.key / C.ntm;
        const int stride = C.ub_max * C.npm;
        lds_add(b, rm.e5);
        lds_add(b + stride, rm.e6);
        lds_add(b + 2 * stride, rm.e7);
        lds_add(b + 3 * stride, rm.e8);
        lds_add(b + 4 * stride, rm.rp);
        lds_add(b + 5 * stride, rm.rt);
        lds_add(b + 6 * stride, rm.z);
        lds_add(b + 7 * stride, rm.w);
    };

    // ---- one walk for census and sums when the pair's rows fit the LDS bands ------------
    // Without the same-half-plate truncation of U or the count pass of xcf, the census and
    // the accumulation can share one walk if every r_trans row the pair can reach is in the
    // bands from the start: r_trans = (d1 + d2) sin(ang/2) is non-decreasing in both
    // (monotone) dist_m, and pair_bins' row arithmetic is non-decreasing in r_trans, so the
    // rows of the first and last pixels' sums bound every pair's rows.
    int g_t0 = 0, g_u0 = 0;
    bool merged = false;
    if (!q.shp_cut && !q.strict && q.prune && q.la > 0 && q.lb > 0) {
        const double y_lo = (Ad[0] + Bd[0]) * q.sin_half / C.rtmax;
        const double y_hi = (Ad[q.la - 1] + Bd[q.lb - 1]) * q.sin_half / C.rtmax;
        const double t0 = fmax(floor(y_lo * C.nt_d), 0.0);
        const double t1 = fmin(floor(y_hi * C.nt_d), C.nt_d - 1);
        const double u0 = fmax(floor(y_lo * C.ntm_d), 0.0);
        const double u1 = fmin(floor(y_hi * C.ntm_d), C.ntm_d - 1);
        merged = t0 <= t1 && u0 <= u1 && t1 - t0 < C.tb_max && u1 - u0 < C.ub_max;
        if (merged) {
            g_t0 = (int)t0;
            g_u0 = (int)u0;
        }
    }
    if (merged) {
        for (int k = tid; k < 4 * C.tb_max * C.np; k += DMAT_PASS_BLOCK) lS[k] = 0;
        for (int k = tid; k < 8 * C.ub_max * C.npm; k += DMAT_PASS_BLOCK) lE[k] = 0;
        __syncthreads();
        uint64_t n_in = 0, n_sel = 0, n_over = 0, n_neg = 0;
        for (int c0 = wave * 64; c0 < q.la; c0 += NWAVES * 64) {
            const int i = c0 + lane;
            bool active = i < q.la;
            double r1 = 0, d1 = 0, z1 = 0, w1 = 0, l1 = 0, f1 = 1;
            if (active) {
                r1 = Ar[i];
                d1 = Ad[i];
                z1 = Az[i];
                w1 = Aw[i];
                l1 = Al[i];
                if (evol) f1 = Af[i];
                active = w1 != 0;
            }
            const bool i_sel = !(q.zerr_on && zerr_cut(C, z1, q.zq2));
            const double a7 = q.order1 ? w1 * l1 * q.inv_s1 : 0.0;
            int j_lo, j_hi;
            window_b(C, q, F, wave_min(active ? r1 : INFINITY),
                     wave_max(active ? r1 : -INFINITY), wave_min(active ? d1 : INFINITY), lane,
                     j_lo, j_hi);
            RowRun rt{-1, 0, 0}, ru{-1, 0, 0};
            int last_m = -1;
            RunA ra{-1, 0, 0, 0, 0};
            RunM rm{-1, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int j = j_lo; j < j_hi; ++j) {
                const double w2 = Bw[j];
                if (w2 == 0) continue;
                const double r2 = Br[j], d2 = Bd[j], z2 = Bz[j], l2 = Bl[j];
                const double f2 = evol ? Bf[j] : 1.0;
                const bool j_zerr = q.zerr_on && zerr_cut(C, z2, q.zq1);
                double r_par, r_trans;
                PairBins B;
                if (!active || !pair_bins(C, r1, d1, r2, d2, q.cos_half, q.sin_half, r_par,
                                          r_trans, B))
                    continue;
                const int A = B.A, M = B.M;
                ++n_in;   // the count pass takes every in-range pair here
                if (M < 0 || M >= C.nm) continue;
                const double z = (z1 + z2) / 2;
                if (M != last_m) {
                    atomicOr(&lU[M >> 5], 1u << (M & 31));
                    last_m = M;
                }
                row_step(ru, B.uM, j, l_iu, l_ju, i);
                const double w12 = w1 * w2;
                const double zfac = evol ? f1 * f2 : 1.0;
                const bool sel = pair_selected(C, q, i_sel, j_zerr, z, r_par);
                if (M != rm.key) {
                    if (rm.key >= 0) flush_m(rm, g_u0);
                    rm = RunM{M, 0, 0, 0, 0, 0, 0, 0, 0};
                }
                if (q.proj2) rm.e5 += zfac * w12 * q.inv_w1 * q.inv_w2;
                if (q.order2) rm.e6 += zfac * w1 * q.inv_w1 * (w2 * l2 * q.inv_s2);
                if (q.order1) {
                    if (q.proj2) rm.e7 += zfac * w2 * q.inv_w2 * a7;
                    if (q.order2) rm.e8 += zfac * w1 * l1 * w2 * l2 * q.inv_s1 * q.inv_s2;
                }
                if (!sel) continue;
                ++n_sel;
                rm.rp += w12 * r_par;
                rm.rt += w12 * r_trans;
                rm.z += w12 * z;
                rm.w += w12;
                if (A < 0) {
                    ++n_neg;
                    continue;
                }
                if (A >= C.nb) {
                    ++n_over;
                    continue;
                }
                row_step(rt, B.tA, j, l_it, l_jt, i);
                if (A != ra.key) {
                    if (ra.key >= 0) flush_a(ra, g_t0);
                    ra = RunA{A, 0, 0, 0, 0};
                }
                ra.s0 += w12;
                ra.sj += w12 * l2;
                ra.si += w12 * l1;
                ra.sij += w12 * l1 * l2;
            }
            if (rt.row >= 0) note_run(l_it, l_jt, rt.row, i, rt.first, rt.last);
            if (ru.row >= 0) note_run(l_iu, l_ju, ru.row, i, ru.first, ru.last);
            if (ra.key >= 0) flush_a(ra, g_t0);
            if (rm.key >= 0) flush_m(rm, g_u0);
        }
        n_in = wave_sum_u64(n_in);
        n_sel = wave_sum_u64(n_sel);
        n_over = wave_sum_u64(n_over);
        n_neg = wave_sum_u64(n_neg);
        if (lane == 0) {
            atomicAdd(&s_cnt[0], (unsigned long long)n_in);
            atomicAdd(&s_cnt[1], (unsigned long long)n_sel);
            atomicAdd(&s_cnt[2], (unsigned long long)n_in);
            atomicAdd(&s_cnt[3], (unsigned long long)n_over);
            atomicAdd(&s_cnt[4], (unsigned long long)n_neg);
        }
    }

    // ---- walk 0: census (counts, pixel ranges of every row, U unless truncated) ---------
    if (!merged) {
        uint64_t n_in = 0, n_sel = 0, n_cp = 0, n_over = 0, n_neg = 0;
        for (int c0 = wave * 64; c0 < q.la; c0 += NWAVES * 64) {
            const int i = c0 + lane;
            bool active = i < q.la;
            double r1 = 0, d1 = 0, z1 = 0, w1 = 0;
            if (active) {
                r1 = Ar[i];
                d1 = Ad[i];
                z1 = Az[i];
                w1 = Aw[i];
                active = w1 != 0;
            }
            const bool i_sel = !(q.zerr_on && zerr_cut(C, z1, q.zq2));
            int j_lo, j_hi;
            window_b(C, q, F, wave_min(active ? r1 : INFINITY),
                     wave_max(active ? r1 : -INFINITY), wave_min(active ? d1 : INFINITY), lane,
                     j_lo, j_hi);
            RowRun rt{-1, 0, 0}, ru{-1, 0, 0};
            int last_m = -1;
            uint32_t cnt_in_i = 0;
            for (int j = j_lo; j < j_hi; ++j) {
                const double w2 = Bw[j];
                if (w2 == 0) continue;
                const double r2 = Br[j], d2 = Bd[j], z2 = Bz[j];
                const bool j_zerr = q.zerr_on && zerr_cut(C, z2, q.zq1);
                double r_par, r_trans;
                PairBins B;
                if (!active || !pair_bins(C, r1, d1, r2, d2, q.cos_half, q.sin_half, r_par,
                                          r_trans, B))
                    continue;
                const int A = B.A, M = B.M;
                ++n_in;
                ++cnt_in_i;
                const bool close = q.shp_cut && fabs(r_par) < C.width_p;
                // the forest x quasar count pass rejects r_par <= r_par_min
                // (py/picca/xcf.py:464-466)
                if (!close && !(q.strict && r_par <= C.rpmin)) ++n_cp;
                if (M < 0 || M >= C.nm) continue;
                const double z = (z1 + z2) / 2;
                if (!q.shp_cut && M != last_m) {
                    atomicOr(&lU[M >> 5], 1u << (M & 31));
                    last_m = M;
                }
                row_step(ru, B.uM, j, l_iu, l_ju, i);
                if (!pair_selected(C, q, i_sel, j_zerr, z, r_par)) continue;
                ++n_sel;
                if (A < 0) {
                    ++n_neg;
                    continue;
                }
                if (A >= C.nb) {
                    ++n_over;
                    continue;
                }
                row_step(rt, B.tA, j, l_it, l_jt, i);
            }
            if (rt.row >= 0) note_run(l_it, l_jt, rt.row, i, rt.first, rt.last);
            if (ru.row >= 0) note_run(l_iu, l_ju, ru.row, i, ru.first, ru.last);
            if (q.shp_cut && i < q.la) W.cnt_i[pix + i] = (int32_t)cnt_in_i;
        }
        n_in = wave_sum_u64(n_in);
        n_sel = wave_sum_u64(n_sel);
        n_cp = wave_sum_u64(n_cp);
        n_over = wave_sum_u64(n_over);
        n_neg = wave_sum_u64(n_neg);
        if (lane == 0) {
            atomicAdd(&s_cnt[0], (unsigned long long)n_in);
            atomicAdd(&s_cnt[1], (unsigned long long)n_sel);
            atomicAdd(&s_cnt[2], (unsigned long long)n_cp);
            atomicAdd(&s_cnt[3], (unsigned long long)n_over);
            atomicAdd(&s_cnt[4], (unsigned long long)n_neg);
        }
    }
    __syncthreads();
    const unsigned long long count_pass = s_cnt[2];
    // The reference returns before its main pass when the count pass finds no pair
    // (py/picca/cf.py:572-573): such a forest pair adds nothing.
    const bool live = count_pass > 0;
    if (tid == 0) {
        if (live) {
            atomicAdd(&W.stats[DMAT_STAT_IN_RANGE], s_cnt[0]);
            atomicAdd(&W.stats[DMAT_STAT_SELECTED], s_cnt[1]);
            if (s_cnt[0] > count_pass)
                atomicAdd(&W.stats[DMAT_STAT_BEYOND], s_cnt[0] - count_pass);
            atomicAdd(&W.stats[DMAT_STAT_OVERFLOW], s_cnt[3]);
            atomicAdd(&W.stats[DMAT_STAT_NEGATIVE], s_cnt[4]);
        }
        int t_lo = INT_MAX, t_hi = -1, u_lo = INT_MAX, u_hi = -1;
        if (live) {
            for (int t = 0; t < C.nt; ++t)
                if (l_it[2 * t] <= l_it[2 * t + 1]) {
                    t_lo = min(t_lo, t);
                    t_hi = max(t_hi, t);
                }
            for (int u = 0; u < C.ntm; ++u)
                if (l_iu[2 * u] <= l_iu[2 * u + 1]) {
                    u_lo = min(u_lo, u);
                    u_hi = max(u_hi, u);
                }
        }
        if (t_hi < 0 || u_hi < 0) {   // nothing selected or no model bin: no fill
            t_lo = u_lo = 1;
            t_hi = u_hi = 0;
        }
        s_band[0] = t_lo;
        s_band[1] = t_hi;
        s_band[2] = u_lo;
        s_band[3] = u_hi;
        W.band[4 * s + 0] = t_lo;
        W.band[4 * s + 1] = t_hi;
        W.band[4 * s + 2] = u_lo;
        W.band[4 * s + 3] = u_hi;
    }
    // pixel ranges of every row
    for (int k = tid; k < 2 * C.nt; k += DMAT_PASS_BLOCK) {
        W.rng_it[(int64_t)s * 2 * C.nt + k] = l_it[k];
        W.rng_jt[(int64_t)s * 2 * C.nt + k] = l_jt[k];
    }
    for (int k = tid; k < 2 * C.ntm; k += DMAT_PASS_BLOCK) {
        W.rng_iu[(int64_t)s * 2 * C.ntm + k] = l_iu[k];
        W.rng_ju[(int64_t)s * 2 * C.ntm + k] = l_ju[k];
    }
    // (t, u) blocks this pair feeds
    if (live) {
        for (int k = tid; k < C.nt * C.ntm; k += DMAT_PASS_BLOCK) {
            const int t = k / C.ntm, u = k - t * C.ntm;
            if (l_it[2 * t] <= l_it[2 * t + 1] && l_iu[2 * u] <= l_iu[2 * u + 1])
                atomicOr(&W.touched[k >> 5], 1u << (k & 31));
        }
    }
    if (!live) return;   // uniform over the workgroup

    // ---- unique model bins of the first count_pass in-range pairs ----------------------
    // With same-half-plate removal the reference sizes its model-bin record by the count
    // pass, which skips close pairs, but fills it with every in-range pair in (i, j) order
    // and takes the unique values of the first count_pass entries (py/picca/cf.py:619,
    // 702-703, 874): U is the model bins of the first count_pass in-range pairs.
    if (q.shp_cut) {
        if (wave == 0) {   // exclusive prefix of the per-pixel in-range counts
            int64_t base = 0;
            for (int c0 = 0; c0 < q.la; c0 += 64) {
                const int i = c0 + lane;
                const int64_t v = i < q.la ? W.cnt_i[pix + i] : 0;
                int64_t incl = v;
                for (int off = 1; off < 64; off <<= 1) {
                    const int64_t o = __shfl_up(incl, off, 64);
                    if (lane >= off) incl += o;
                }
                if (i < q.la) W.cnt_i[pix + i] = (int32_t)(base + incl - v);
                base += __shfl(incl, 63, 64);
            }
        }
        __syncthreads();
        for (int c0 = wave * 64; c0 < q.la; c0 += NWAVES * 64) {
            const int i = c0 + lane;
            bool active = i < q.la;
            double r1 = 0, d1 = 0, w1 = 0;
            int64_t g = 0;
            if (active) {
                r1 = Ar[i];
                d1 = Ad[i];
                w1 = Aw[i];
                g = W.cnt_i[pix + i];
                active = w1 != 0 && g < (int64_t)count_pass;
            }
            int j_lo, j_hi;
            window_b(C, q, F, wave_min(active ? r1 : INFINITY),
                     wave_max(active ? r1 : -INFINITY), wave_min(active ? d1 : INFINITY), lane,
                     j_lo, j_hi);
            for (int j = j_lo; j < j_hi; ++j) {
                const double w2 = Bw[j];
                if (w2 == 0) continue;
                double r_par, r_trans;
                PairBins B;
                if (!active || !pair_bins(C, r1, d1, Br[j], Bd[j], q.cos_half, q.sin_half,
                                          r_par, r_trans, B))
                    continue;
                const int A = B.A, M = B.M;
                if (g < (int64_t)count_pass && M >= 0 && M < C.nm)
                    atomicOr(&lU[M >> 5], 1u << (M & 31));
                ++g;
            }
        }
    }
    __syncthreads();

    // ---- accumulation passes over LDS bands of r_trans rows ------------------------------
    const int t_lo = s_band[0], t_hi = s_band[1], u_lo = s_band[2], u_hi = s_band[3];
    const int n_t = t_hi >= t_lo ? t_hi - t_lo + 1 : 0;
    const int n_u = u_hi >= u_lo ? u_hi - u_lo + 1 : 0;
    // (the single walk above has filled the one band pass already)
    const int passes = merged ? (n_t > 0 || n_u > 0 ? 1 : 0)
                              : max((n_t + C.tb_max - 1) / C.tb_max,
                                    (n_u + C.ub_max - 1) / C.ub_max);
    for (int pass = 0; pass < passes; ++pass) {
        const int tc0 = merged ? g_t0 : t_lo + pass * C.tb_max;
        const int uc0 = merged ? g_u0 : u_lo + pass * C.ub_max;
        if (!merged) {
        for (int k = tid; k < 4 * C.tb_max * C.np; k += DMAT_PASS_BLOCK) lS[k] = 0;
        for (int k = tid; k < 8 * C.ub_max * C.npm; k += DMAT_PASS_BLOCK) lE[k] = 0;
        __syncthreads();
        for (int c0 = wave * 64; c0 < q.la; c0 += NWAVES * 64) {
            const int i = c0 + lane;
            bool active = i < q.la;
            double r1 = 0, d1 = 0, z1 = 0, w1 = 0, l1 = 0, f1 = 1;
            if (active) {
                r1 = Ar[i];
                d1 = Ad[i];
                z1 = Az[i];
                w1 = Aw[i];
                l1 = Al[i];
                if (evol) f1 = Af[i];
                active = w1 != 0;
            }
            const bool i_sel = !(q.zerr_on && zerr_cut(C, z1, q.zq2));
            // per-pixel factors of the eta sums (py/picca/cf.py:765-843)
            const double a7 = q.order1 ? w1 * l1 * q.inv_s1 : 0.0;    // w1 dll1 / sum w dll^2
            int j_lo, j_hi;
            window_b(C, q, F, wave_min(active ? r1 : INFINITY),
                     wave_max(active ? r1 : -INFINITY), wave_min(active ? d1 : INFINITY), lane,
                     j_lo, j_hi);
            RunA ra{-1, 0, 0, 0, 0};
            RunM rm{-1, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int j = j_lo; j < j_hi; ++j) {
                const double w2 = Bw[j];
                if (w2 == 0) continue;
                const double r2 = Br[j], d2 = Bd[j], z2 = Bz[j], l2 = Bl[j];
                const double f2 = evol ? Bf[j] : 1.0;
                const bool j_zerr = q.zerr_on && zerr_cut(C, z2, q.zq1);
                double r_par, r_trans;
                PairBins B;
                if (!active || !pair_bins(C, r1, d1, r2, d2, q.cos_half, q.sin_half, r_par,
                                          r_trans, B))
                    continue;
                const int A = B.A, M = B.M;
                if (M < 0 || M >= C.nm) continue;
                const double z = (z1 + z2) / 2;
                const double w12 = w1 * w2;
                const double zfac = evol ? f1 * f2 : 1.0;
                const bool sel = pair_selected(C, q, i_sel, j_zerr, z, r_par);
                if (M != rm.key) {
                    if (rm.key >= 0) flush_m(rm, uc0);
                    rm = RunM{M, 0, 0, 0, 0, 0, 0, 0, 0};
                }
                if (q.proj2) rm.e5 += zfac * w12 * q.inv_w1 * q.inv_w2;
                if (q.order2) rm.e6 += zfac * w1 * q.inv_w1 * (w2 * l2 * q.inv_s2);
                if (q.order1) {
                    if (q.proj2) rm.e7 += zfac * w2 * q.inv_w2 * a7;
                    if (q.order2) rm.e8 += zfac * w1 * l1 * w2 * l2 * q.inv_s1 * q.inv_s2;
                }
                if (!sel) continue;
                rm.rp += w12 * r_par;
                rm.rt += w12 * r_trans;
                rm.z += w12 * z;
                rm.w += w12;
                if (A < 0 || A >= C.nb) continue;
                if (A != ra.key) {
                    if (ra.key >= 0) flush_a(ra, tc0);
                    ra = RunA{A, 0, 0, 0, 0};
                }
                ra.s0 += w12;
                ra.sj += w12 * l2;
                ra.si += w12 * l1;
                ra.sij += w12 * l1 * l2;
            }
            if (ra.key >= 0) flush_a(ra, tc0);
            if (rm.key >= 0) flush_m(rm, uc0);
        }
        __syncthreads();
        }
        // band rows -> workspace (eta masked by U), weights_dmat and effective sums -> out
        const int t_end = min(t_hi + 1, tc0 + C.tb_max);
        for (int k = tid; k < (t_end - tc0) * C.np; k += DMAT_PASS_BLOCK) {
            const int tt = k / C.np, rp = k - tt * C.np, t = tc0 + tt;
            const double *b = lS + tt * C.np + rp;
            const int stride = C.tb_max * C.np;
            for (int c = 0; c < 4; ++c)
                W.S[(((int64_t)s * 4 + c) * C.nt + t) * C.np + rp] = b[c * stride];
            if (b[0] != 0) atomicAdd(&O.weights_dmat[rp * C.nt + t], b[0]);
        }
        const int u_end = min(u_hi + 1, uc0 + C.ub_max);
        for (int k = tid; k < (u_end - uc0) * C.npm; k += DMAT_PASS_BLOCK) {
            const int uu = k / C.npm, mrp = k - uu * C.npm, u = uc0 + uu;
            const double *b = lE + uu * C.npm + mrp;
            const int stride = C.ub_max * C.npm;
            const int kb = mrp * C.ntm + u;
            const double mask = ((lU[kb >> 5] >> (kb & 31)) & 1u) ? 1.0 : 0.0;
            for (int c = 0; c < 4; ++c)
                W.eta[(((int64_t)s * 4 + c) * C.ntm + u) * C.npm + mrp] = b[c * stride] * mask;
            if (b[4 * stride] != 0 || b[5 * stride] != 0 || b[6 * stride] != 0 ||
                b[7 * stride] != 0) {
                atomicAdd(&O.r_par_eff[kb], b[4 * stride]);
                atomicAdd(&O.r_trans_eff[kb], b[5 * stride]);
                atomicAdd(&O.z_eff[kb], b[6 * stride]);
                atomicAdd(&O.weight_eff[kb], b[7 * stride]);
            }
        }
        __syncthreads();
    }
    for (int k = tid; k < C.words_u; k += DMAT_PASS_BLOCK)
        W.U[(int64_t)s * C.words_u + k] = lU[k];
}

namespace {

// per-wave vectors of the fill kernel
constexpr int VEC_P = 0;                                  // P (or Q)       [CHUNK_RP]
constexpr int VEC_P2 = DMAT_CHUNK_RP;                     // P' (or Q')     [CHUNK_RP]
constexpr int VEC_E = 2 * DMAT_CHUNK_RP;                  // eta1 (or eta2) [CHUNK_MRP]
constexpr int VEC_E2 = 2 * DMAT_CHUNK_RP + DMAT_CHUNK_MRP;  // eta3 (or eta4)
constexpr int VEC_LEN = 2 * DMAT_CHUNK_RP + 2 * DMAT_CHUNK_MRP;
constexpr int FILL_WAVES = DMAT_FILL_BLOCK / 64;
constexpr int FILL_CB = DMAT_CHUNK_MRP / 16;   // 16-column blocks of model r_par bins
static_assert(DMAT_CHUNK_RP == 16 * FILL_WAVES, "one 16-row block of data bins per wave");
static_assert(DMAT_CHUNK_MRP % 16 == 0, "model bins in 16-column blocks");

typedef double double4_t __attribute__((ext_vector_type(4)));

constexpr size_t FILL_LDS_FIXED = 4 * DMAT_CHUNK_RP + 4 * DMAT_CHUNK_MRP  // S, eta rows
                                  + DMAT_CHUNK_MRP                          // U mask
                                  + (size_t)FILL_WAVES * VEC_LEN;          // vectors

// The pair-bin term w1 w2 zfac of a pair goes to its data bin A and its model bin M, which
// share x = (r_par - r_par_min) / span: the model r_par bin of data r_par bin a lies in
// [a npm / np, (a + 1) npm / np], so the fill keeps, per data bin, `diag_width` model bins
// from diag_lo(a) (one bin of slack each side) instead of the whole block; a term outside
// them (rounding beyond the slack: not expected) goes to dmat directly.
__host__ __device__ inline int diag_width(const DmatConst &C) {
    return min((C.npm + C.np - 1) / C.np + 3, DMAT_CHUNK_MRP);
}
__device__ __forceinline__ int diag_lo(const DmatConst &C, int rp) {
    return (int)(((int64_t)rp * C.npm) / C.np) - 1;
}

}  // namespace

// QUASAR: every delta2 of the batch is an unprojected quasar (order < 0), whose pairs have
// no eta1, eta3, eta5..eta8 terms; a separate instantiation keeps the forest x forest form's
// register budget.
template <bool QUASAR>
__global__ __launch_bounds__(DMAT_FILL_BLOCK, 3) void dmat_fill_kernel(DmatConst C, DmatForests F,
                                                                       DmatPairs P, DmatWork W,
                                                                       DmatOut O, int groups) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tile = blockIdx.x;
    if (!((W.touched[tile >> 5] >> (tile & 31)) & 1u)) return;
    const int t = tile / C.ntm;
    const int u = tile - t * C.ntm;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    // v_mfma_f64_16x16x4f64 operands: lane l holds A[l % 16][l / 16] and B[l / 16][l % 16];
    // register r of the result holds row 4 r + l / 16, column l % 16 of the 16 x 16 block.
    // Wave w owns the data r_par bins 16 w .. 16 w + 15 of the chunk and every model bin.
    const int li = lane & 15;
    const int lk = lane >> 4;
    const int row = 16 * wave + li;   // this lane's data bin in the A operands
    const bool evol = (C.flags & DMAT_REDSHIFT_EVOL) != 0;

    const int nd = diag_width(C);
    uint32_t n_mfma = 0;   // matrix-core instructions of this wave (the statistic)
    double *diag = (double *)smem;                              // [CHUNK_RP][nd]
    double *vS = diag + DMAT_CHUNK_RP * nd;                     // [4][CHUNK_RP]
    double *vE = vS + 4 * DMAT_CHUNK_RP;                        // [4][CHUNK_MRP]
    double *mask = vE + 4 * DMAT_CHUNK_MRP;                     // [CHUNK_MRP]
    double *vecs = mask + DMAT_CHUNK_MRP;                       // [FILL_WAVES][VEC_LEN]
    double *vec = vecs + wave * VEC_LEN;

    for (int rp0 = 0; rp0 < C.np; rp0 += DMAT_CHUNK_RP) {
        for (int mrp0 = 0; mrp0 < C.npm; mrp0 += DMAT_CHUNK_MRP) {
            const int nrp = min(DMAT_CHUNK_RP, C.np - rp0);
            const int nmrp = min(DMAT_CHUNK_MRP, C.npm - mrp0);
            double4_t acc[FILL_CB];
#pragma unroll
            for (int b = 0; b < FILL_CB; ++b) acc[b] = (double4_t){0.0, 0.0, 0.0, 0.0};
            for (int k = tid; k < DMAT_CHUNK_RP * nd; k += DMAT_FILL_BLOCK) diag[k] = 0;
            // pair-bin term of data bin r and model bin m of the chunk
            auto add_diag = [&](int r, int m, double v) {
                const int k = mrp0 + m - diag_lo(C, rp0 + r);
                if (k >= 0 && k < nd) {
                    lds_add(diag + r * nd + k, v);
                } else {
                    const int64_t A = (int64_t)(rp0 + r) * C.nt + t;
                    const int64_t M = (int64_t)(mrp0 + m) * C.ntm + u;
                    atomicAdd(&O.dmat[A * C.nm + M], v);
                }
            };
            for (int k = tid; k < FILL_WAVES * VEC_LEN; k += DMAT_FILL_BLOCK) vecs[k] = 0;
            __syncthreads();

            auto outer = [&](int nvec) {
                // acc -= sum over the first nvec waves' vectors of P x E + P2 x E2: two
                // K = 4 products on the matrix cores (K = the pixels), A = -P, B = E
                const bool kin = lk < nvec;
                const double *v = vecs + lk * VEC_LEN;
                const double a1 = kin ? -v[VEC_P + row] : 0.0;
                const double a2 = kin ? -v[VEC_P2 + row] : 0.0;
                // no pair of these pixels in this wave's data bins: nothing to add
                if (__ballot(a1 != 0.0 || a2 != 0.0) == 0) return;
                n_mfma += 2 * FILL_CB;
#pragma unroll
                for (int b = 0; b < FILL_CB; ++b) {
                    const double e1 = kin ? v[VEC_E + 16 * b + li] : 0.0;
                    const double e2 = kin ? v[VEC_E2 + 16 * b + li] : 0.0;
                    acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, e1, acc[b], 0, 0, 0);
                    acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, e2, acc[b], 0, 0, 0);
                }
            };

            for (int s = blockIdx.y; s < P.count; s += groups) {
                const int t_lo = W.band[4 * s], t_hi = W.band[4 * s + 1];
                const int u_lo = W.band[4 * s + 2], u_hi = W.band[4 * s + 3];
                if (t < t_lo || t > t_hi || u < u_lo || u > u_hi) continue;
                const DmatPair q = load_pair(C, F, P, s);

                // rank-4 term S[., t] x eta5..8[., u] (eta already masked by U); eta5..8
                // vanish when delta2 is not projected
                if (!QUASAR) {
                    for (int k = tid; k < 4 * DMAT_CHUNK_RP; k += DMAT_FILL_BLOCK) {
                        const int c = k / DMAT_CHUNK_RP, r = k - c * DMAT_CHUNK_RP;
                        vS[k] = r < nrp
                                    ? W.S[(((int64_t)s * 4 + c) * C.nt + t) * C.np + rp0 + r]
                                    : 0.0;
                    }
                    for (int k = tid; k < 4 * DMAT_CHUNK_MRP; k += DMAT_FILL_BLOCK) {
                        const int c = k / DMAT_CHUNK_MRP, m = k - c * DMAT_CHUNK_MRP;
                        vE[k] = m < nmrp ? W.eta[(((int64_t)s * 4 + c) * C.ntm + u) * C.npm +
                                                 mrp0 + m]
                                         : 0.0;
                    }
                }
                for (int m = tid; m < DMAT_CHUNK_MRP; m += DMAT_FILL_BLOCK) {
                    double v = 0.0;
                    if (m < nmrp) {
                        const int kb = (mrp0 + m) * C.ntm + u;
                        v = ((W.U[(int64_t)s * C.words_u + (kb >> 5)] >> (kb & 31)) & 1u) ? 1.0
                                                                                          : 0.0;
                    }
                    mask[m] = v;
                }
                __syncthreads();
                if (!QUASAR) {
                    // rank-4 term: one K = 4 product per block (K = the four sums)
                    const double sa = vS[lk * DMAT_CHUNK_RP + row];
                    if (__ballot(sa != 0.0) != 0) {
                        n_mfma += FILL_CB;
#pragma unroll
                        for (int b = 0; b < FILL_CB; ++b)
                            acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                sa, vE[lk * DMAT_CHUNK_MRP + 16 * b + li], acc[b], 0, 0, 0);
                    }
                }

                const double *__restrict__ Ar = F.r + q.oa;
                const double *__restrict__ Ad = F.d + q.oa;
                const double *__restrict__ Az = F.z + q.oa;
                const double *__restrict__ Aw = F.w + q.oa;
                const double *__restrict__ Al = F.dll + q.oa;
                const double *__restrict__ Af = F.zfa + q.oa;
                const double *__restrict__ Br = F.r + q.ob;
                const double *__restrict__ Bd = F.d + q.ob;
                const double *__restrict__ Bz = F.z + q.ob;
                const double *__restrict__ Bw = F.w + q.ob;
                const double *__restrict__ Bl = F.dll + q.ob;
                const double *__restrict__ Bf = F.zfb + q.ob;

                // ---- pixels i of delta1: P, P' (row t) and eta1, eta3 (row u) ----------
                // (eta1, eta3 vanish when delta2 is not projected: the QUASAR form adds the
                // identity term with the pixels j below)
                if (!QUASAR) {
                    const int32_t *it = W.rng_it + ((int64_t)s * C.nt + t) * 2;
                    const int32_t *iu = W.rng_iu + ((int64_t)s * C.ntm + u) * 2;
                    const int i0 = max(it[0], iu[0]);
                    const int i1 = min(it[1], iu[1]);
                    for (int base = i0; base <= i1; base += FILL_WAVES) {
                        const int i = base + wave;
                        if (i <= i1) {
                            const double w1 = Aw[i];
                            if (w1 != 0) {
                                const double r1 = Ar[i], d1 = Ad[i], z1 = Az[i], l1 = Al[i];
                                const double f1 = evol ? Af[i] : 1.0;
                                const bool i_sel = !(q.zerr_on && zerr_cut(C, z1, q.zq2));
                                int j_lo, j_hi;
                                pixel_window(C, q, Br, Bd, q.lb, r1 - C.rpmax / q.cos_half,
                                             (C.flags & DMAT_XCORR)
                                                 ? r1 - C.rpmin / q.cos_half
                                                 : r1 + C.rpmax / q.cos_half,
                                             d1, t, u, lane, j_lo, j_hi);
                                for (int j = j_lo + lane; j < j_hi; j += 64) {
                                    const double w2 = Bw[j];
                                    if (w2 == 0) continue;
                                    double r_par, r_trans;
                                    PairBins B;
                                    if (!pair_bins(C, r1, d1, Br[j], Bd[j], q.cos_half,
                                                   q.sin_half, r_par, r_trans, B))
                                        continue;
                                    const int A = B.A, M = B.M;
                                    if (M < 0 || M >= C.nm) continue;
                                    const double z2 = Bz[j], l2 = Bl[j];
                                    const double zfac = evol ? f1 * Bf[j] : 1.0;
                                    const int mu = B.uM;
                                    const int m = B.mM - mrp0;
                                    const bool in_m = mu == u && m >= 0 && m < nmrp;
                                    if (in_m && mask[m] != 0 && q.proj2) {
                                        lds_add(vec + VEC_E + m, zfac * w2 * q.inv_w2);
                                        if (q.order2)
                                            lds_add(vec + VEC_E2 + m, zfac * w2 * l2 * q.inv_s2);
                                    }
                                    const bool j_zerr = q.zerr_on && zerr_cut(C, z2, q.zq1);
                                    if (!pair_selected(C, q, i_sel, j_zerr, (z1 + z2) / 2, r_par))
                                        continue;
                                    if (A < 0 || A >= C.nb || B.tA != t) continue;
                                    const int r = B.rA - rp0;
                                    if (r < 0 || r >= nrp) continue;
                                    const double w12 = w1 * w2;
                                    lds_add(vec + VEC_P + r, w12);
                                    lds_add(vec + VEC_P2 + r, w12 * l2);
                                    if (in_m) add_diag(r, m, w12 * zfac);
                                }
                            }
                        }
                        __syncthreads();
                        outer(min(FILL_WAVES, i1 - base + 1));
                        __syncthreads();
                        for (int k = lane; k < VEC_LEN; k += 64) vec[k] = 0;
                    }
                }
                // ---- QUASAR: delta2 is an unprojected quasar; per pixel j, the whole
                // workgroup walks the pixels i that reach row t or row u into one shared
                // Q, Q', eta2, eta4, and adds the identity term of each selected pair -----
                if (QUASAR) {
                    const int32_t *it = W.rng_it + ((int64_t)s * C.nt + t) * 2;
                    const int32_t *iu = W.rng_iu + ((int64_t)s * C.ntm + u) * 2;
                    const int32_t *jt = W.rng_jt + ((int64_t)s * C.nt + t) * 2;
                    const int32_t *ju = W.rng_ju + ((int64_t)s * C.ntm + u) * 2;
                    const int ia = min(it[0], iu[0]), ib = max(it[1], iu[1]);
                    const int j0 = min(jt[0], ju[0]), j1 = max(jt[1], ju[1]);
                    double *sv = vecs;   // the vectors of wave 0, shared by the workgroup
                    for (int j = j0; ia <= ib && j <= j1; ++j) {
                        const double w2 = Bw[j];
                        if (w2 == 0) continue;
                        const double r2 = Br[j], d2 = Bd[j], z2 = Bz[j];
                        const double f2 = evol ? Bf[j] : 1.0;
                        const bool j_zerr = q.zerr_on && zerr_cut(C, z2, q.zq1);
                        int i_lo, i_hi;
                        window_a(C, q, F, r2, d2, lane, i_lo, i_hi);
                        i_lo = max(i_lo, ia);
                        i_hi = min(i_hi, ib + 1);
                        for (int i = i_lo + tid; i < i_hi; i += DMAT_FILL_BLOCK) {
                            const double w1 = Aw[i];
                            if (w1 == 0) continue;
                            double r_par, r_trans;
                            PairBins B;
                            if (!pair_bins(C, Ar[i], Ad[i], r2, d2, q.cos_half, q.sin_half,
                                           r_par, r_trans, B))
                                continue;
                            const int A = B.A, M = B.M;
                            if (M < 0 || M >= C.nm) continue;
                            const double z1 = Az[i], l1 = Al[i];
                            const double zfac = evol ? Af[i] * f2 : 1.0;
                            const int mu = B.uM;
                            const int m = B.mM - mrp0;
                            const bool in_m = mu == u && m >= 0 && m < nmrp;
                            if (in_m && mask[m] != 0) {
                                lds_add(sv + VEC_E + m, zfac * w1 * q.inv_w1);
                                if (q.order1) lds_add(sv + VEC_E2 + m, zfac * w1 * l1 * q.inv_s1);
                            }
                            const bool i_sel = !(q.zerr_on && zerr_cut(C, z1, q.zq2));
                            if (!pair_selected(C, q, i_sel, j_zerr, (z1 + z2) / 2, r_par))
                                continue;
                            if (A < 0 || A >= C.nb || B.tA != t) continue;
                            const int r = B.rA - rp0;
                            if (r < 0 || r >= nrp) continue;
                            const double w12 = w1 * w2;
                            lds_add(sv + VEC_P + r, w12);
                            lds_add(sv + VEC_P2 + r, w12 * l1);
                            if (in_m) add_diag(r, m, w12 * zfac);
                        }
                        __syncthreads();
                        outer(1);
                        __syncthreads();
                        for (int k = tid; k < VEC_LEN; k += DMAT_FILL_BLOCK) sv[k] = 0;
                        __syncthreads();
                    }
                }
                // ---- pixels j of delta2: Q, Q' (row t) and eta2, eta4 (row u) ----------
                if (!QUASAR) {
                    const int32_t *jt = W.rng_jt + ((int64_t)s * C.nt + t) * 2;
                    const int32_t *ju = W.rng_ju + ((int64_t)s * C.ntm + u) * 2;
                    const int j0 = max(jt[0], ju[0]);
                    const int j1 = min(jt[1], ju[1]);
                    for (int base = j0; base <= j1; base += FILL_WAVES) {
                        const int j = base + wave;
                        if (j <= j1) {
                            const double w2 = Bw[j];
                            if (w2 != 0) {
                                const double r2 = Br[j], d2 = Bd[j], z2 = Bz[j];
                                const double f2 = evol ? Bf[j] : 1.0;
                                const bool j_zerr = q.zerr_on && zerr_cut(C, z2, q.zq1);
                                int i_lo, i_hi;
                                pixel_window(C, q, Ar, Ad, q.la,
                                             (C.flags & DMAT_XCORR)
                                                 ? r2 + C.rpmin / q.cos_half
                                                 : r2 - C.rpmax / q.cos_half,
                                             r2 + C.rpmax / q.cos_half, d2, t, u, lane, i_lo,
                                             i_hi);
                                for (int i = i_lo + lane; i < i_hi; i += 64) {
                                    const double w1 = Aw[i];
                                    if (w1 == 0) continue;
                                    double r_par, r_trans;
                                    PairBins B;
                                    if (!pair_bins(C, Ar[i], Ad[i], r2, d2, q.cos_half,
                                                   q.sin_half, r_par, r_trans, B))
                                        continue;
                                    const int A = B.A, M = B.M;
                                    if (M < 0 || M >= C.nm) continue;
                                    const double z1 = Az[i], l1 = Al[i];
                                    const double zfac = evol ? Af[i] * f2 : 1.0;
                                    const int mu = B.uM;
                                    const int m = B.mM - mrp0;
                                    if (mu == u && m >= 0 && m < nmrp && mask[m] != 0) {
                                        lds_add(vec + VEC_E + m, zfac * w1 * q.inv_w1);
                                        if (q.order1)
                                            lds_add(vec + VEC_E2 + m, zfac * w1 * l1 * q.inv_s1);
                                    }
                                    const bool i_sel = !(q.zerr_on && zerr_cut(C, z1, q.zq2));
                                    if (!pair_selected(C, q, i_sel, j_zerr, (z1 + z2) / 2, r_par))
                                        continue;
                                    if (A < 0 || A >= C.nb || B.tA != t) continue;
                                    const int r = B.rA - rp0;
                                    if (r < 0 || r >= nrp) continue;
                                    const double w12 = w1 * w2;
                                    lds_add(vec + VEC_P + r, w12);
                                    lds_add(vec + VEC_P2 + r, w12 * l1);
                                }
                            }
                        }
                        __syncthreads();
                        outer(min(FILL_WAVES, j1 - base + 1));
                        __syncthreads();
                        for (int k = lane; k < VEC_LEN; k += 64) vec[k] = 0;
                    }
                }
                __syncthreads();   // vS / vE / mask are rewritten for the next pair
            }

            // block (+ the pair-bin term) -> dmat
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int r = 16 * wave + 4 * a + lk;
#pragma unroll
                for (int b = 0; b < FILL_CB; ++b) {
                    const int m = 16 * b + li;
                    if (r < nrp && m < nmrp) {
                        const int k = mrp0 + m - diag_lo(C, rp0 + r);
                        const double v =
                            acc[b][a] + ((k >= 0 && k < nd) ? diag[r * nd + k] : 0.0);
                        if (v != 0) {
                            const int64_t A = (int64_t)(rp0 + r) * C.nt + t;
                            const int64_t k = (int64_t)(mrp0 + m) * C.ntm + u;
                            atomicAdd(&O.dmat[A * C.nm + k], v);
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
    if (lane == 0 && n_mfma) atomicAdd(&W.stats[DMAT_STAT_MFMA], (unsigned long long)n_mfma);
}

size_t dmat_fill_lds(const DmatConst &C) {
    return (FILL_LDS_FIXED + (size_t)DMAT_CHUNK_RP * diag_width(C)) * sizeof(double);
}

hipError_t launch_dmat_pass(const DmatConst &C, const DmatForests &F, const DmatPairs &P,
                            const DmatWork &W, const DmatOut &O, size_t lds,
                            hipStream_t stream) {
    if (P.count <= 0) return hipSuccess;
    hipError_t err = hipFuncSetAttribute((const void *)dmat_pass_kernel,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(dmat_pass_kernel, dim3(P.count), dim3(DMAT_PASS_BLOCK), lds, stream, C,
                       F, P, W, O);
    return hipGetLastError();
}

hipError_t launch_dmat_fill(const DmatConst &C, const DmatForests &F, const DmatPairs &P,
                            const DmatWork &W, const DmatOut &O, int groups, bool quasar,
                            hipStream_t stream) {
    if (P.count <= 0) return hipSuccess;
    auto kern = quasar ? dmat_fill_kernel<true> : dmat_fill_kernel<false>;
    hipError_t err = hipFuncSetAttribute((const void *)kern,
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)dmat_fill_lds(C));
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(kern, dim3(C.nt * C.ntm, groups), dim3(DMAT_FILL_BLOCK),
                       dmat_fill_lds(C), stream, C, F, P, W, O, groups);
    return hipGetLastError();
}

}  // namespace picca
