// Pair kernel of the 3D correlation function on gfx950 (MI355X).
//
// GPU restatement of compute_xi_forest_pairs_fast (py/picca/cf.py:250-387) for a whole
// batch of forest pairs, grouped by output row (the healpix of delta1: picca_cf calls
// compute_xi once per healpix, py/picca/bin/picca_cf.py:452-458).
//
// Work decomposition
//   * one persistent workgroup of 1024 threads (16 wave64) per CU keeps the histogram of
//     its current output row in LDS, 48 B per bin (5 FP64 sums + u32 count), and adds it to
//     the row in HBM with FP64 atomics when the row changes.  The LDS histogram is padded
//     to (np + 1) x (nt + 1) bins (125 KB at 50x50): the extra row and column collect the
//     pairs the range test rejects, so the fast loop bins every pair without a branch.
//     When it does not fit in LDS the runs are added straight to the row in HBM;
//   * workgroups pull "groups" (<= 64 consecutive pairs of one row) from an HBM queue;
//   * inside a group, waves pull tasks = (pair, chunk of 64 pixels of delta1) from an
//     LDS counter.  Lane l owns pixel i = 64*chunk + l of delta1 in registers; the wave
//     walks the pixels j of delta2 in lockstep;
//   * specialised kernel (standard binning, auto-correlation, no cuts): pixel j is a
//     wave-uniform record read through the scalar cache;
//     general kernel: pixels j are staged per wave through LDS (lane l loads pixel j0 + l,
//     pixels the reference skips before any arithmetic are dropped, the rest compacted,
//     the next tile already loading) and read back with broadcast LDS reads;
//   * for the standard (r_par, r_trans) binning the j range is pruned to the window where
//     |r_par| < r_par_max and r_trans < r_trans_max can hold for some lane (r_comov and
//     dist_m are non-decreasing along a forest; forests where they are not are walked in
//     full).  The skipped pixel pairs are provably rejected by the reference's range test;
//   * a lane accumulates consecutive pairs that land in the same bin in registers
//     (r_par and r_trans move monotonically with j) and adds the run to the histogram
//     with six LDS atomics when its bin changes;
//   * run walk (xi_walk.h), the specialised kernel's schedule wherever delta2's pixels
//     increase strictly and a bin holds at least a few of them: each lane walks its own
//     runs of delta2 pixels with one bin, finds a run's end with about two probes and adds
//     the run from per-forest prefix sums -- one flush per run instead of one evaluation
//     per pair.
//
// Arithmetic: r_par, r_trans and the bin index follow the reference's FP64 operations
// (built with -ffp-contract=off).  The bin index is first taken with a reciprocal
// multiply; whenever the fractional part lies within the edge tolerance (plus a rounding
// margin) of an integer it is recomputed with the reference's divide-then-multiply, so
// every pair gets the reference's bin.  The mean pair redshift (z1 + z2) / 2 is formed as
// z1/2 + z2/2, which is the same double.  The fast loop sums |r1 - r2| w1 w2 and
// (d1 + d2) w1 w2 and multiplies by cos(ang/2), sin(ang/2) once per run, which changes
// the r_par / r_trans sums by rounding only (relative 1e-16 per pair).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_util.h"
#include "engine.h"
#include "pair_common.h"
#include "xi_walk.h"

namespace picca {

// Everything compute_xi_forest_pairs_fast does with one pixel pair after the weight and
// zerr skips, with the reference's arithmetic (py/picca/cf.py:336-387): pair-redshift cut,
// separations, range test, bins (reciprocal multiply, exact divide-then-multiply near an
// edge), same-half-plate cut, overflow guard.  Returns true and sets bin / r_par / r_trans
// for an accepted pair; pbin is the same bin in the padded LDS layout (row stride nt + 1).
template <int MODE, bool XCORR, bool GENERAL>
__device__ __forceinline__ bool exact_pair(const XiConst &C, double r1, double d1, double r2,
                                           double d2, double z, double ang, double cos_half,
                                           double sin_half, bool same_half_plate, int &bin,
                                           int &pbin, double &r_par, double &r_trans,
                                           uint32_t &edge, uint32_t &overflow) {
    if (GENERAL) {
        if ((C.flags & XI_HAS_ZMIN) && z < C.zmin) return false;
        if ((C.flags & XI_HAS_ZMAX) && z > C.zmax) return false;
    }
    if (MODE == MODE_ANGULAR) {
        r_par = r1 / r2;
        if (!XCORR && r_par < 1.0) r_par = 1.0 / r_par;
        r_trans = ang;
    } else {
        r_par = (r1 - r2) * cos_half;
        r_trans = (d1 + d2) * sin_half;
        if (MODE == MODE_RMU) {
            r_trans = sqrt(r_trans * r_trans + r_par * r_par);
            r_par /= r_trans;
        }
        if (!XCORR) r_par = fabs(r_par);
    }
    if (r_par >= C.rpmax || r_trans >= C.rtmax || r_par < C.rpmin) {
        // pairs just outside the range (within edge_tol of it in both coordinates) are edge
        // pairs too
        if (r_par - C.rpmax < C.edge_tol && C.rpmin - r_par < C.edge_tol &&
            r_trans - C.rtmax < C.edge_tol)
            ++edge;
        return false;
    }
    double t_par = (r_par - C.rpmin) * C.inv_p;
    double t_trans = r_trans * C.inv_t;
    double f_par = floor(t_par);
    double f_trans = floor(t_trans);
    if ((fabs(t_par - f_par - 0.5) > C.half_p) | (fabs(t_trans - f_trans - 0.5) > C.half_t)) {
        t_par = (r_par - C.rpmin) / C.span_p * C.np_d;
        t_trans = r_trans / C.rtmax * C.nt_d;
        f_par = floor(t_par);
        f_trans = floor(t_trans);
        // an exact zero (equal r_comov, or ang = 0) cannot move under rounding
        if ((r_par != 0 && edge_distance(t_par, C.width_p) < C.edge_tol) ||
            (r_trans != 0 && edge_distance(t_trans, C.width_t) < C.edge_tol))
            ++edge;
    }
    if (GENERAL && (C.flags & XI_REMOVE_SHP) && same_half_plate && fabs(r_par) < C.width_p)
        return false;
    const double bin_f = f_trans + C.nt_d * f_par;
    if (!(bin_f >= 0.0 && bin_f < C.nb_d)) {
        ++overflow;
        return false;
    }
    bin = (int)bin_f;
    // a pair whose r_trans bin rounds to nt lands in bin 0 of the next r_par row, as in the
    // reference's flat index
    pbin = (int)(f_trans + C.ntp_d * f_par) + (f_trans >= C.nt_d ? 1 : 0);
    return true;
}

// The exact re-evaluation of exact_pair for the fast loop (standard binning,
// auto-correlation, r_par_min == 0, no cuts) as straight-line code with every constant
// read once at the top: the fast loop runs it for a handful of pairs per billion, and a
// branchy version leaves loads outstanding on some of its paths, which the pixel loop
// would then have to wait for.  Returns the padded bin, or `trash` for a rejected pair;
// edge and overflow pairs are counted in LDS.
__device__ __forceinline__ int exact_fast(const XiConst &K, double r1, double d1, double r2,
                                          double d2, double cos_half, double sin_half,
                                          int trash, __attribute__((address_space(3)))
                                          unsigned long long *lds_stats) {
    const double rpmax = K.rpmax, rtmax = K.rtmax, tol = K.edge_tol, span = K.span_p;
    const double np_d = K.np_d, nt_d = K.nt_d, nb_d = K.nb_d, ntp_d = K.ntp_d;
    const double width_p = K.width_p, width_t = K.width_t;
    const double r_par = fabs((r1 - r2) * cos_half);
    const double r_trans = (d1 + d2) * sin_half;
    const bool out = (r_par >= rpmax) | (r_trans >= rtmax);
    // outside the range, but within edge_tol of it in both coordinates
    const bool edge_out = (r_par - rpmax < tol) & (r_trans - rtmax < tol);
    const double t_par = (r_par - 0.0) / span * np_d;
    const double t_trans = r_trans / rtmax * nt_d;
    const double f_par = floor(t_par);
    const double f_trans = floor(t_trans);
    const bool edge_in = ((r_par != 0) & (edge_distance(t_par, width_p) < tol)) |
                         ((r_trans != 0) & (edge_distance(t_trans, width_t) < tol));
    const double bin_f = f_trans + nt_d * f_par;
    const bool over = !((bin_f >= 0.0) & (bin_f < nb_d));
    const bool edge = out ? edge_out : edge_in;
    if (edge)
        __hip_atomic_fetch_add(&lds_stats[XI_STAT_EDGE], 1ull, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
    if (!out & over)
        __hip_atomic_fetch_add(&lds_stats[XI_STAT_OVERFLOW], 1ull, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
    // f_par, f_trans are small non-negative integers unless the pair is rejected
    const double pbin_f = f_trans + ntp_d * f_par + (f_trans >= nt_d ? 1.0 : 0.0);
    return (out | over) ? trash : (int)fmin(pbin_f, (double)trash);
}

// Run of the fast loop: consecutive pairs of one lane in one padded bin.  r_par and r_trans
// are summed before their cos(ang/2) / sin(ang/2) factors, which are constant over a task
// and applied when the run is added to the histogram.
struct FastRun {
    int32_t bin;
    double count;  // exact below 2^53
    double we, xi, rp, rt, z;
};

__device__ __forceinline__ void flush_fast(const FastRun &run, LocalBin *hist, double cos_half,
                                           double sin_half) {
    LocalBin *b = (LocalBin *)((LocalByte *)hist +
                               __umul24((uint32_t)run.bin, (uint32_t)sizeof(LdsBin)));
    __hip_atomic_fetch_add(&b->we, run.we, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&b->xi, run.xi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&b->rp, run.rp * cos_half, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&b->rt, run.rt * sin_half, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&b->z, run.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&b->nb, (uint32_t)run.count, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Wave-uniform test of a pixel weight for +0 or -0 on the scalar unit (a compiler-chosen
// FP64 compare would take a vector instruction per pixel).
__device__ __forceinline__ bool uniform_zero_weight(double w) {
    const unsigned long long bits = (unsigned long long)__double_as_longlong(w);
    unsigned long long twice;
    asm volatile("s_lshl_b64 %0, %1, 1" : "=s"(twice) : "s"(bits) : "scc");
    return twice == 0;
}

// min of two finite doubles as one v_min_f64 (fmin would first canonicalise an operand
// that the compiler cannot prove canonical, once per call)
__device__ __forceinline__ double min_f64(double a, double b) {
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "s"(b));
    return r;
}

typedef __attribute__((address_space(3))) unsigned long long LocalU64;

// Lanes whose closed runs the walk merges before its LDS atomics (a power of two, <= 16).
// On the 100k-forest survey: no merging 28.7 s, groups of 16 / 8 / 4 / 2 lanes 26.5 / 25.5 /
// 24.7 / 24.8 s; each step of the segmented sum costs ~20 VALU instructions per pass.
constexpr int WALK_FLUSH_GROUP = 4;

// Memory of the run walk on the device: delta2's pixels and prefix sums in HBM (read
// per lane), the LDS histogram of the output row, the LDS statistics.
struct DeviceWalkMem {
    const double2 *__restrict__ rdp;
    const PrefixRecord *__restrict__ pre;
    const double *__restrict__ wp;
    const double *__restrict__ zp;
    const double *__restrict__ dwp;
    LocalBin *hist;
    LocalU64 *stats;
    uint32_t probes, runs, cover;

    __device__ void rd(int p, double &r, double &d) const {
        const double2 v = rdp[p];
        r = v.x;
        d = v.y;
    }
    __device__ PrefixRecord prefix(int p) const { return pre[p]; }
    __device__ double w(int p) const { return wp[p]; }
    __device__ double z(int p) const { return zp[p]; }
    __device__ double dw(int p) const { return dwp[p]; }
    __device__ void add(int bin, double we, double xi, double rp, double rt, double z,
                        uint32_t n) {
        LocalBin *b = (LocalBin *)((LocalByte *)hist +
                                   __umul24((uint32_t)bin, (uint32_t)sizeof(LdsBin)));
        __hip_atomic_fetch_add(&b->we, we, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(&b->xi, xi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(&b->rp, rp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(&b->rt, rt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(&b->z, z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(&b->nb, n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    // Adds the runs that the lanes of the wave closed in one pass (all 64 lanes call it;
    // `on` false: no run).  Neighbouring lanes (consecutive pixels of delta1 walking in
    // lockstep) mostly close runs of the same bin in the same pass: a segmented sum over
    // each group of WALK_FLUSH_GROUP lanes merges the lanes of equal bin, and the last lane
    // of each segment makes the six LDS atomics, so that a bin takes one atomic per group
    // instead of one per lane.
    __device__ void flush(bool on, const RunSums &run) {
        static_assert(WALK_FLUSH_GROUP >= 1 && WALK_FLUSH_GROUP <= 16 &&
                      (WALK_FLUSH_GROUP & (WALK_FLUSH_GROUP - 1)) == 0, "group of 1..16 lanes");
        // segment key: bin + 1 for a lane with a run, 0 without (and outside the row)
        const int key = on ? run.bin + 1 : 0;
        const int gl = __lane_id() & (WALK_FLUSH_GROUP - 1);
        int head = (key != dpp_int<DPP_ROW_SHR1>(key)) || gl == 0;
        double we = run.we, xi = run.xi, rp = run.rp, rt = run.rt, z = run.z;
        uint32_t n = run.n;
        if constexpr (WALK_FLUSH_GROUP > 1) segment_step<DPP_ROW_SHR1>(head, we, xi, rp, rt, z, n);
        if constexpr (WALK_FLUSH_GROUP > 2) segment_step<DPP_ROW_SHR2>(head, we, xi, rp, rt, z, n);
        if constexpr (WALK_FLUSH_GROUP > 4) segment_step<DPP_ROW_SHR4>(head, we, xi, rp, rt, z, n);
        if constexpr (WALK_FLUSH_GROUP > 8) segment_step<DPP_ROW_SHR8>(head, we, xi, rp, rt, z, n);
        const bool tail = (key != dpp_int<DPP_ROW_SHL1>(key)) || gl == WALK_FLUSH_GROUP - 1;
        if (on && tail) add(run.bin, we, xi, rp, rt, z, n);
    }
    // One step of the inclusive segmented sum: a lane whose segment head lies within the
    // last CTRL lanes keeps its value, the others add the value of lane l - CTRL.
    template <int CTRL>
    __device__ static void segment_step(int &head, double &we, double &xi, double &rp,
                                        double &rt, double &z, uint32_t &n) {
        const double g = head ? 0.0 : 1.0;
        const int h = dpp_int<CTRL>(head);
        we = fma(dpp_f64<CTRL>(we), g, we);
        xi = fma(dpp_f64<CTRL>(xi), g, xi);
        rp = fma(dpp_f64<CTRL>(rp), g, rp);
        rt = fma(dpp_f64<CTRL>(rt), g, rt);
        z = fma(dpp_f64<CTRL>(z), g, z);
        const uint32_t m = (uint32_t)dpp_int<CTRL>((int)n);
        n += head ? 0u : m;
        head |= h;
    }
    __device__ bool any(bool v) const { return __ballot(v) != 0; }
    // Keeps the run's sums where the walk computes them, at the top of the pass: left to
    // itself the compiler moves them to the flush, after the prefix record of the run's
    // start has been replaced, and keeps two copies of that record around the pass.
    __device__ static void pin(RunSums &r) {
        asm volatile("" : "+v"(r.we), "+v"(r.xi), "+v"(r.rp), "+v"(r.rt), "+v"(r.z), "+v"(r.n),
                     "+v"(r.bin));
    }
    __device__ void count_edge() {
        __hip_atomic_fetch_add(&stats[XI_STAT_EDGE], 1ull, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __device__ void count_overflow() {
        __hip_atomic_fetch_add(&stats[XI_STAT_OVERFLOW], 1ull, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __device__ void count_probe() { ++probes; }
    __device__ void count_run(uint32_t n) {
        ++runs;
        cover += n;
    }
};

// One task: pixels [64*chunk, 64*chunk + 64) of delta1 against the pruned window of delta2.
//   GENERAL = false: standard binning, auto-correlation, r_par_min == 0, no z / zerr /
//   same-half-plate cuts (the cuts of the reference that cannot trigger are compiled out).
//   LDS_HIST: the histogram is the padded LDS layout (row stride nt + 1), otherwise the
//   output row in HBM.
template <int MODE, bool XCORR, bool GENERAL, bool LDS_HIST, bool WALK, typename Bin>
__device__ void xi_task(const XiConst &C, const XiConst &C_lds, const ForestView &F,
                        const PairView &P, int64_t p, int chunk, int lane, Bin *hist,
                        TilePixel *tile, TileExact *tile_x, LaneCounters &cnt,
                        LocalU64 *lds_stats, uint64_t &n_steps, uint64_t &n_walk) {
    const int a = uniform_int(P.a[p]);
    const int b = uniform_int(P.b[p]);
    const int64_t oa = uniform_i64(F.off[a]);
    const int la = (int)(uniform_i64(F.off[a + 1]) - oa);
    const int64_t ob = uniform_i64(F.off[b]);
    const int lb = (int)(uniform_i64(F.off[b + 1]) - ob);
    const double ang = uniform_f64(P.ang[p]);
    const double cos_half = uniform_f64(P.cs[2 * p]);
    const double sin_half = uniform_f64(P.cs[2 * p + 1]);
    bool same_half_plate = false;
    bool zerr_on = false;
    double z_qso_1 = 0, z_qso_2 = 0;
    if (GENERAL) {
        same_half_plate = (uniform_int(P.flags[p]) & PAIR_SAME_HALF_PLATE) != 0;
        zerr_on = (C.flags & XI_HAS_ZERR) && (ang < C.zerr_ang);
        z_qso_1 = uniform_f64(F.zq[a]);
        z_qso_2 = uniform_f64(F.zq[b]);
    }

    const int i = chunk * 64 + lane;
    bool active = i < la;
    double r1 = 0, d1 = 0, zh1 = 0, w1 = 0, dw1 = 0;
    if (active) {
        r1 = F.r[oa + i];
        d1 = F.d[oa + i];
        zh1 = F.z[oa + i] * 0.5;
        w1 = F.w[oa + i];
        dw1 = F.dw[oa + i];
        active = w1 != 0;
    }
    if (GENERAL && zerr_on && active) {
        const double z1 = zh1 * 2.0;
        double z_qf = 0.5 * (z1 + z_qso_2);
        double dv_kms = fabs(z1 - z_qso_2) / (1 + z_qf);
        dv_kms *= SPEED_LIGHT_KMS;
        if (dv_kms < C.zerr_kms) active = false;
    }
    if (__ballot(active) == 0) return;
    // Drain the loads of this lane's pixel now: the pixel loop keeps tile loads in
    // flight, and a value still pending here would otherwise be waited for with
    // vmcnt(0) inside the loop.  (gfx9 s_waitcnt: vmcnt 0, expcnt 7, lgkmcnt 15.)
    __builtin_amdgcn_s_waitcnt(0x0F70);

    const double *__restrict__ Br = F.r + ob;
    const double *__restrict__ Bd = F.d + ob;
    const double *__restrict__ Bz = F.z + ob;
    const double *__restrict__ Bw = F.w + ob;
    const double *__restrict__ Bdw = F.dw + ob;

    int j_lo = 0, j_hi = lb;
    bool pruned = false;
    if (MODE == MODE_STANDARD && (C.flags & XI_PRUNE) && cos_half > 0 &&
        (uniform_int(F.mono[b]) & 3) == 3) {
        pruned = true;
        const double r_min = wave_min(active ? r1 : INFINITY);
        const double r_max = wave_max(active ? r1 : -INFINITY);
        const double d_min = wave_min(active ? d1 : INFINITY);
        // r_par = (r1 - r2) cos_half must reach [r_par_min - tol, r_par_max + tol)
        const double slack = 4.0 * C.edge_tol / cos_half + 1e-6;
        double lo_r = r_min - C.rpmax / cos_half;
        double hi_r = XCORR ? r_max - C.rpmin / cos_half : r_max + C.rpmax / cos_half;
        lo_r -= slack + 1e-12 * fabs(lo_r);
        hi_r += slack + 1e-12 * fabs(hi_r);
        j_lo = wave_search<true>(Br, lb, lo_r, lane);
        j_hi = wave_search<false>(Br, lb, hi_r, lane);
        if (sin_half > 0) {
            double d_lim = (C.rtmax + 4.0 * C.edge_tol) / sin_half - d_min;
            d_lim += 1e-6 + 1e-12 * fabs(d_lim);
            const int j_d = wave_search<false>(Bd, lb, d_lim, lane);
            if (j_d < j_hi) j_hi = j_d;
        }
    }
    j_lo = uniform_int(j_lo);
    j_hi = uniform_int(j_hi);
    // an empty window (about half the tasks of a survey): no pair of this chunk is in range
    if (j_lo >= j_hi) return;

    if constexpr (!GENERAL && LDS_HIST) {
        if (WALK && (C.flags & XI_WALK) && pruned) {
            // Run walk (xi_walk.h) for the tasks walk_task admits; otherwise the pair loop
            // below.
            const double spc_r = uniform_f64(F.spc[b].x);
            const double spc_d = uniform_f64(F.spc[b].y);
            const double cp = cos_half * C.inv_p;
            const double st = sin_half * C.inv_t;
            double q0 = 1.0;
            if (j_hi - j_lo >= 2) {
                const double span = (uniform_f64(F.rd[ob + j_hi - 1].x) -
                                     uniform_f64(F.rd[ob + j_lo].x)) * cp;
                if (span > 0) q0 = (double)(j_hi - 1 - j_lo) / span;
            }
            if (walk_task(C, spc_r, spc_d, cp, st, q0)) {
                WalkTask T;
                T.cos_half = cos_half;
                T.sin_half = sin_half;
                T.cp = cp;
                T.st = st;
                T.rc = uniform_f64(F.rd[ob + (lb >> 1)].x);
                T.dc = uniform_f64(F.rd[ob + (lb >> 1)].y);
                T.j_lo = j_lo;
                T.j_hi = j_hi;
                T.q0 = q0;
                DeviceWalkMem mem{F.rd + ob, F.pre + (ob + b), F.w + ob, F.z + ob, F.dw + ob,
                                  hist, lds_stats, 0u, 0u, 0u};
                const WalkLane L{r1, d1, zh1, w1, dw1};
                walk_lane(C, T, L, mem, active);
                cnt.visit += mem.cover;
                cnt.runs += mem.runs;
                cnt.probes += mem.probes;
                ++n_walk;
                // wave passes of the walk: the lanes' loops run until the longest is done
                n_steps += (uint64_t)uniform_int(wave_max_int((int)mem.probes));
                return;
            }
        }
        // Pixel j of delta2 is wave-uniform: its record comes through the scalar cache into
        // SGPRs and the loop touches LDS only to flush runs.  Two pixels per iteration, the
        // next two already loading; the store is padded, so reading up to three records
        // past j_hi stays in bounds.
        //
        // Every pair of an active lane takes a bin of the padded histogram: a pair whose
        // r_par or r_trans lies beyond the range clamps to row np or column nt, which
        // collect the rejected pairs and are never read.  Inactive lanes are moved far
        // beyond r_par_max so that their (zero-weight) pairs always land there.
        const PixelRecord *rec = F.pix + ob;
        const double cp = cos_half * C.inv_p;  // bins per Mpc/h of |r1 - r2|
        const double st = sin_half * C.inv_t;  // bins per Mpc/h of d1 + d2
        const double cap_p = C.np_d;  // clamps of the bin coordinates: row np, column nt
        const double cap_t = C.nt_d;
        const int trash = (C.np + 1) * C.ntp - 1;
        const double r1f = active ? r1 : 1e300;
        uint32_t steps = 0;
        FastRun run;
        run.bin = trash;
        run.count = 0.0;
        run.we = run.xi = run.rp = run.rt = run.z = 0.0;
        PixelRecord c0 = load_record(rec + j_lo), c1 = load_record(rec + j_lo + 1);
        // consume the first two records here, so that their loads are waited for before
        // the loop and the loop only ever waits for records it prefetched an iteration
        // earlier
        asm volatile("" ::"s"(c0.w), "s"(c0.z), "s"(c1.w), "s"(c1.z));
        auto step = [&](const PixelRecord &q) {
            // the reference skips w2 == 0 (either sign) before any arithmetic
            if (uniform_zero_weight(q.w)) return;
            ++steps;
            const double diff = r1f - q.r;
            const double sum = d1 + q.d;
            // |r_par| and r_trans in bin units.  Away from an edge the bin is the integer
            // part; within the rounding margin of an edge (or of the range limits, which
            // are edges) the pair is redone with the reference's arithmetic.  r_par == 0
            // exactly (equal r_comov) is bin 0 and never an edge pair.
            const double tp = fabs(diff) * cp;
            const double tt = sum * st;
            const double fp = __builtin_amdgcn_fract(tp);
            const double ft = __builtin_amdgcn_fract(tt);
            const bool near = ((fabs(fp - 0.5) > C.half_p) & (diff != 0)) |
                              (fabs(ft - 0.5) > C.half_t);
            int bin = (int)__umul24((unsigned)C.ntp, (unsigned)(int)min_f64(tp, cap_p)) +
                      (int)min_f64(tt, cap_t);
            if (active & near) {
                // exact re-evaluation (py/picca/cf.py:355-376), constants from the LDS copy
                bin = exact_fast(C_lds, r1, d1, q.r, q.d, cos_half, sin_half, trash, lds_stats);
            }
            const double w12 = w1 * q.w;
            const double dw12 = dw1 * q.dw;
            const double z = fma(q.z, 0.5, zh1);
            // a new bin adds the lane's run to the histogram and restarts it: the restart is
            // a multiply by 0 (every input is finite), which costs no branch of its own
            const bool change = bin != run.bin;
            if (change) flush_fast(run, hist, cos_half, sin_half);
            const double keep = change ? 0.0 : 1.0;
            run.bin = bin;
            run.count = fma(run.count, keep, 1.0);
            run.we = fma(run.we, keep, w12);
            run.xi = fma(run.xi, keep, dw12);
            run.rp = fma(fabs(diff), w12, run.rp * keep);
            run.rt = fma(sum, w12, run.rt * keep);
            run.z = fma(z, w12, run.z * keep);
        };
        for (int j = j_lo; j < j_hi; j += 2) {
            const PixelRecord n0 = load_record(rec + j + 2), n1 = load_record(rec + j + 3);
            step(c0);
            if (j + 1 < j_hi) step(c1);
            c0 = n0;
            c1 = n1;
        }
        flush_fast(run, hist, cos_half, sin_half);
        n_steps += steps;
        if (active) cnt.visit += steps;
        return;
    } else {
        LaneRun run;
        run.bin = -1;
        run.count = 0;
        run.we = run.xi = run.rp = run.rt = run.z = 0.0;
        uint32_t visit = 0;
        uint32_t edge = 0;
        uint32_t overflow = 0;
        const int tile_len = C.tile_len;

        double pr = 0, pd = 0, pz = 0, pw = 0, pdw = 0;
        bool pvalid = false;
        {
            const int j = j_lo + lane;
            pvalid = lane < tile_len && j < j_hi;
            if (pvalid) {
                pr = Br[j];
                pd = Bd[j];
                pz = Bz[j];
                pw = Bw[j];
                pdw = Bdw[j];
            }
        }
        for (int j0 = j_lo; j0 < j_hi; j0 += tile_len) {
            // stage the prefetched tile: drop what the reference skips, compact the rest
            bool keep = pvalid && pw != 0;
            if (GENERAL && zerr_on && keep) {
                double z_qf = 0.5 * (pz + z_qso_1);
                double dv_kms = fabs(pz - z_qso_1) / (1 + z_qf);
                dv_kms *= SPEED_LIGHT_KMS;
                if (dv_kms < C.zerr_kms) keep = false;
            }
            const uint64_t mask = __ballot(keep);
            const int n_tile = uniform_int(__popcll(mask));
            const int pos = __builtin_amdgcn_mbcnt_hi(
                (uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
            if (keep) {
                TilePixel *t = tile + pos;
                t->cr = cos_half * pr;
                t->sd = sin_half * pd;
                t->w = pw;
                t->dw = pdw;
                t->zh = pz * 0.5;
                tile_x[pos].r = pr;
                tile_x[pos].d = pd;
            }
            __builtin_amdgcn_wave_barrier();
            {
                const int j = j0 + tile_len + lane;
                pvalid = lane < tile_len && j < j_hi;
                if (pvalid) {
                    pr = Br[j];
                    pd = Bd[j];
                    pz = Bz[j];
                    pw = Bw[j];
                    pdw = Bdw[j];
                }
            }
            n_steps += n_tile;
            for (int u = 0; u < n_tile; ++u) {
                const double2 wdw = *reinterpret_cast<const double2 *>(&tile[u].w);
                const double zh2 = tile[u].zh;
                const double2 rd = *reinterpret_cast<const double2 *>(&tile_x[u].r);
                const double z = zh1 + zh2;
                if (!active) continue;
                if (!((C.flags & XI_HAS_ZMIN) && z < C.zmin) &&
                    !((C.flags & XI_HAS_ZMAX) && z > C.zmax))
                    ++visit;
                int bin, pbin;
                double r_par, r_trans;
                if (exact_pair<MODE, XCORR, GENERAL>(C, r1, d1, rd.x, rd.y, z, ang, cos_half,
                                                     sin_half, same_half_plate, bin, pbin,
                                                     r_par, r_trans, edge, overflow))
                    add_pair(run, hist, cnt, LDS_HIST ? pbin : bin, w1 * wdw.x, dw1 * wdw.y,
                             r_par, r_trans, z);
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (run.bin >= 0) {
            flush_run(run, hist);
            cnt.accept += run.count;
        }
        cnt.visit += visit;
        cnt.edge += edge;
        cnt.overflow += overflow;
    }
}

// Adds the real bins of the padded LDS histogram to output row `row` in HBM and clears
// them; returns the pairs this thread moved (the accepted-pair statistic).
__device__ __forceinline__ uint64_t flush_row(const XiConst &C, LocalBin *hist, XiBin *out,
                                              int tid) {
    uint64_t moved = 0;
    for (int k = tid; k < C.nb; k += XI_BLOCK) {
        const int kp = k + k / C.nt;  // padded index: row k / nt, column k % nt
        const uint32_t n = hist[kp].nb;
        if (n) {
            unsafeAtomicAdd(&out[k].we, hist[kp].we);
            unsafeAtomicAdd(&out[k].xi, hist[kp].xi);
            unsafeAtomicAdd(&out[k].rp, hist[kp].rp);
            unsafeAtomicAdd(&out[k].rt, hist[kp].rt);
            unsafeAtomicAdd(&out[k].z, hist[kp].z);
            atomicAdd(&out[k].nb, (unsigned long long)n);
            hist[kp].we = hist[kp].xi = hist[kp].rp = hist[kp].rt = hist[kp].z = 0.0;
            hist[kp].nb = 0;
            moved += n;
        }
    }
    return moved;
}

template <int MODE, bool XCORR, bool GENERAL, bool LDS_HIST, bool WALK>
__global__ void __launch_bounds__(XI_BLOCK)
xi_kernel(XiConst C, ForestView F, PairView P, GroupView G, XiOut O) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int nbp = (C.np + 1) * C.ntp;  // padded LDS histogram
    LocalBin *hist = (LocalBin *)reinterpret_cast<LdsBin *>(smem);
    const size_t hist_bytes = LDS_HIST ? (size_t)nbp * sizeof(LdsBin) : 0;
    // per wave: (tile_len + 1) staged pixels, then (tile_len + 1) exact coordinates
    const int wave = threadIdx.x >> 6;
    unsigned char *wave_tiles = smem + hist_bytes +
                                (size_t)wave * (C.tile_len + 1) * XI_TILE_PIXEL_BYTES;
    TilePixel *tile = reinterpret_cast<TilePixel *>(wave_tiles);
    TileExact *tile_x = reinterpret_cast<TileExact *>(
        wave_tiles + (size_t)(C.tile_len + 1) * sizeof(TilePixel));

    __shared__ int64_t s_group;
    __shared__ int s_task;
    __shared__ int s_prefix[XI_GROUP_PAIRS + 1];
    __shared__ int s_first[XI_GROUP_PAIRS];   // first chunk of each pair of the group
    __shared__ unsigned long long s_stats[XI_NSTATS];
    __shared__ XiConst s_const;
    LocalU64 *lds_stats = (LocalU64 *)s_stats;

    const int tid = threadIdx.x;
    const int lane = tid & 63;

    if (LDS_HIST) {
        for (int k = tid; k < nbp; k += XI_BLOCK) {
            hist[k].we = hist[k].xi = hist[k].rp = hist[k].rt = hist[k].z = 0.0;
            hist[k].nb = 0;
        }
    }
    if (tid < XI_NSTATS) s_stats[tid] = 0;
    if (tid == 0) s_const = C;  // published by the first __syncthreads of the group loop

    LaneCounters cnt = {0, 0, 0, 0, 0, 0};
    uint64_t accepted = 0;  // LDS_HIST: pairs moved from the histogram to HBM
    uint64_t n_steps = 0, n_tasks = 0, n_empty_tasks = 0, n_walk = 0;  // wave-uniform
    int current_row = -1;

    for (;;) {
        __syncthreads();
        if (tid == 0) s_group = (int64_t)atomicAdd(O.queue, 1ull);
        __syncthreads();
        const int64_t g = s_group;
        if (g >= G.count) break;
        const int row = G.row[g];
        if (LDS_HIST && row != current_row && current_row >= 0)
            accepted += flush_row(C, hist, O.bins + (int64_t)current_row * C.nb, tid);
        current_row = row;
        const int64_t p0 = G.begin[g];
        const int npairs = (int)(G.begin[g + 1] - p0);
        // chunk prefix over the pairs of the group (one wave)
        if (tid < 64) {
            int chunks = 0;
            if (tid < npairs) {
                // the chunks of delta1 that can reach delta2 (xi_chunk_kernel)
                const int first = P.chunks[2 * (p0 + tid)];
                chunks = P.chunks[2 * (p0 + tid) + 1] - first;
                s_first[tid] = first;
            }
            int incl = chunks;
            for (int off = 1; off < 64; off <<= 1) {
                int v = __shfl_up(incl, off, 64);
                if (lane >= off) incl += v;
            }
            if (tid < npairs) s_prefix[tid + 1] = incl;
            if (tid == 0) {
                s_prefix[0] = 0;
                s_task = 0;
            }
        }
        __syncthreads();
        const int total = s_prefix[npairs];
        for (;;) {
            int t = 0;
            if (lane == 0) t = atomicAdd(&s_task, 1);
            t = __shfl(t, 0, 64);
            t = uniform_int(t);
            if (t >= total) break;
            // largest q with s_prefix[q] <= t
            int lo = 0, hi = npairs - 1;
            while (lo < hi) {
                int mid = (lo + hi + 1) >> 1;
                if (s_prefix[mid] <= t) lo = mid;
                else hi = mid - 1;
            }
            const int q = uniform_int(lo);
            const int chunk = uniform_int(s_first[q] + t - s_prefix[q]);
            const uint64_t steps_before = n_steps;
            if (LDS_HIST) {
                xi_task<MODE, XCORR, GENERAL, true, WALK>(C, s_const, F, P, p0 + q, chunk, lane, hist,
                                                    tile, tile_x, cnt, lds_stats, n_steps, n_walk);
            } else {
                xi_task<MODE, XCORR, GENERAL, false, false>(C, s_const, F, P, p0 + q, chunk, lane,
                                                     O.bins + (int64_t)row * C.nb, tile, tile_x,
                                                     cnt, lds_stats, n_steps, n_walk);
            }
            ++n_tasks;
            if (n_steps == steps_before) ++n_empty_tasks;
        }
    }
    // final flush (all waves are past the last __syncthreads of the loop)
    if (LDS_HIST && current_row >= 0)
        accepted += flush_row(C, hist, O.bins + (int64_t)current_row * C.nb, tid);
    if (!LDS_HIST) accepted = cnt.accept;
    // statistics: lane values -> wave sums -> workgroup -> HBM
    const uint64_t v_visit = wave_sum_u64(cnt.visit);
    const uint64_t v_edge = wave_sum_u64(cnt.edge);
    const uint64_t v_over = wave_sum_u64(cnt.overflow);
    const uint64_t v_accept = wave_sum_u64(accepted);
    const uint64_t v_runs = wave_sum_u64(cnt.runs);
    const uint64_t v_probes = wave_sum_u64(cnt.probes);
    if (lane == 0) {
        atomicAdd(&s_stats[XI_STAT_VISITED], (unsigned long long)v_visit);
        atomicAdd(&s_stats[XI_STAT_EDGE], (unsigned long long)v_edge);
        atomicAdd(&s_stats[XI_STAT_OVERFLOW], (unsigned long long)v_over);
        atomicAdd(&s_stats[XI_STAT_STEPS], (unsigned long long)n_steps);
        atomicAdd(&s_stats[XI_STAT_ACCEPTED], (unsigned long long)v_accept);
        atomicAdd(&s_stats[XI_STAT_TASKS], (unsigned long long)n_tasks);
        atomicAdd(&s_stats[XI_STAT_EMPTY_TASKS], (unsigned long long)n_empty_tasks);
        atomicAdd(&s_stats[XI_STAT_RUNS], (unsigned long long)v_runs);
        atomicAdd(&s_stats[XI_STAT_PROBES], (unsigned long long)v_probes);
        atomicAdd(&s_stats[XI_STAT_WALK_TASKS], (unsigned long long)n_walk);
    }
    __syncthreads();
    if (tid < XI_NSTATS) atomicAdd(O.stats + tid, s_stats[tid]);
}

// Work-queue head and counters of picca_xi_execute_noalloc: one copy per GPU, held until
// the process ends.  Host code reaches it only through the address hipGetSymbolAddress
// returns (xi_scratch_address), never through this name.
static __device__ __align__(256) unsigned long long scratch[XI_SCRATCH_WORDS];

hipError_t xi_scratch_address(unsigned long long **ptr) {
    void *p = nullptr;
    hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    *ptr = err == hipSuccess ? (unsigned long long *)p : nullptr;
    return err;
}

// Per pair, the 64-pixel chunks of delta1 [first, end) whose windows of delta2 can be
// non-empty: with both forests monotone, delta1's pixels must lie within r_par_max / cos of
// delta2's r_comov span and have dist_m below (r_trans_max / sin) - dist_m of delta2's first
// pixel, each bound widened by twice the pair kernel's window slack, so every skipped chunk
// is one whose pruned window the pair kernel would find empty.  Other pairs keep every chunk.
__global__ void xi_chunk_kernel(XiConst C, ForestView F, PairView P, int64_t num_pairs,
                                bool standard) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= num_pairs) return;
    const int a = P.a[p], b = P.b[p];
    const int64_t oa = F.off[a], ob = F.off[b];
    const int la = (int)(F.off[a + 1] - oa), lb = (int)(F.off[b + 1] - ob);
    int first = 0, end = (la + 63) >> 6;
    const double cos_half = P.cs[2 * p], sin_half = P.cs[2 * p + 1];
    const int ma = F.mono[a], mb = F.mono[b];
    if (standard && (C.flags & XI_PRUNE) && cos_half > 0 && (mb & 3) == 3 && (ma & 1) &&
        la > 0 && lb > 0) {
        const double *__restrict__ R1 = F.r + oa;
        const double *__restrict__ D1 = F.d + oa;
        const double reach = fmax(C.rpmax, fabs(C.rpmin)) / cos_half;
        const double slack = 8.0 * C.edge_tol / cos_half + 2e-6;
        double lo = F.r[ob] - reach, hi = F.r[ob + lb - 1] + reach;
        lo -= slack + 2e-12 * fabs(lo);
        hi += slack + 2e-12 * fabs(hi);
        // first pixel with r1 >= lo, first with r1 > hi
        int i_lo = 0, i_hi = la;
        for (int n = la; n > 0;) {
            const int h = n >> 1;
            if (R1[i_lo + h] < lo) { i_lo += h + 1; n -= h + 1; } else { n = h; }
        }
        int base = 0;
        for (int n = la; n > 0;) {
            const int h = n >> 1;
            if (R1[base + h] <= hi) { base += h + 1; n -= h + 1; } else { n = h; }
        }
        i_hi = base;
        if (sin_half > 0 && (ma & 2)) {
            double d_lim = (C.rtmax + 8.0 * C.edge_tol) / sin_half - F.d[ob];
            d_lim += 2e-6 + 2e-12 * fabs(d_lim);
            int k = 0;
            for (int n = la; n > 0;) {
                const int h = n >> 1;
                if (D1[k + h] <= d_lim) { k += h + 1; n -= h + 1; } else { n = h; }
            }
            if (k < i_hi) i_hi = k;
        }
        if (i_lo >= i_hi) {
            first = end = 0;
        } else {
            first = i_lo >> 6;
            end = (i_hi + 63) >> 6;
        }
    }
    P.chunks[2 * p] = first;
    P.chunks[2 * p + 1] = end;
}

hipError_t launch_xi_chunks(int mode, const XiConst &C, const ForestView &F, const PairView &P,
                            int64_t num_pairs, hipStream_t stream) {
    if (num_pairs <= 0) return hipSuccess;
    const int64_t blocks = (num_pairs + 255) / 256;
    xi_chunk_kernel<<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(C, F, P, num_pairs,
                                                                       mode == MODE_STANDARD);
    return hipGetLastError();
}

__global__ void zero_u64_kernel(unsigned long long *__restrict__ a, int64_t n) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; k < n; k += stride) a[k] = 0ull;
}

__global__ __launch_bounds__(256) void copy16_kernel(const uint4 *__restrict__ src,
                                                    uint4 *__restrict__ dst, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
        dst[k] = src[k];
}

hipError_t launch_copy16(const void *src, void *dst, int64_t n16, hipStream_t stream) {
    if (n16 <= 0) return hipSuccess;
    int64_t blocks = (n16 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    copy16_kernel<<<dim3((unsigned)blocks), dim3(256), 0, stream>>>((const uint4 *)src,
                                                                    (uint4 *)dst, n16);
    return hipGetLastError();
}

hipError_t launch_zero_u64(unsigned long long *ptr, int64_t n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    zero_u64_kernel<<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(ptr, n);
    return hipGetLastError();
}

template <int MODE, bool XCORR, bool GENERAL, bool LDS_HIST, bool WALK>
static hipError_t launch_xi_t(const XiConst &C, const ForestView &F, const PairView &P,
                              const GroupView &G, const XiOut &O, int grid, size_t lds,
                              hipStream_t stream) {
    auto kern = xi_kernel<MODE, XCORR, GENERAL, LDS_HIST, WALK>;
    hipError_t err = hipFuncSetAttribute((const void *)kern,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (err != hipSuccess) return err;
    kern<<<dim3(grid), dim3(XI_BLOCK), lds, stream>>>(C, F, P, G, O);
    return hipGetLastError();
}

template <int MODE, bool XCORR, bool GENERAL>
static hipError_t launch_xi_h(bool lds_hist, const XiConst &C, const ForestView &F,
                              const PairView &P, const GroupView &G, const XiOut &O, int grid,
                              size_t lds, hipStream_t stream) {
    return lds_hist
               ? launch_xi_t<MODE, XCORR, GENERAL, true, false>(C, F, P, G, O, grid, lds, stream)
               : launch_xi_t<MODE, XCORR, GENERAL, false, false>(C, F, P, G, O, grid, lds, stream);
}

// general == false is only valid for MODE_STANDARD, auto-correlation, r_par_min == 0 and
// no z / zerr / same-half-plate cuts (the host decides).  The run walk (XI_WALK) has a
// kernel of its own, so that its registers do not weigh on the pair loop's.
hipError_t launch_xi(int mode, bool xcorr, bool general, bool lds_hist, const XiConst &C,
                     const ForestView &F, const PairView &P, const GroupView &G,
                     const XiOut &O, int grid, size_t lds, hipStream_t stream) {
    if (mode == MODE_STANDARD) {
        if (!general && !xcorr) {
            if (lds_hist && (C.flags & XI_WALK))
                return launch_xi_t<MODE_STANDARD, false, false, true, true>(C, F, P, G, O, grid,
                                                                            lds, stream);
            return launch_xi_h<MODE_STANDARD, false, false>(lds_hist, C, F, P, G, O, grid, lds, stream);
        }
        return xcorr ? launch_xi_h<MODE_STANDARD, true, true>(lds_hist, C, F, P, G, O, grid, lds, stream)
                     : launch_xi_h<MODE_STANDARD, false, true>(lds_hist, C, F, P, G, O, grid, lds, stream);
    }
    if (mode == MODE_RMU)
        return xcorr ? launch_xi_h<MODE_RMU, true, true>(lds_hist, C, F, P, G, O, grid, lds, stream)
                     : launch_xi_h<MODE_RMU, false, true>(lds_hist, C, F, P, G, O, grid, lds, stream);
    return xcorr ? launch_xi_h<MODE_ANGULAR, true, true>(lds_hist, C, F, P, G, O, grid, lds, stream)
                 : launch_xi_h<MODE_ANGULAR, false, true>(lds_hist, C, F, P, G, O, grid, lds, stream);
}

}  // namespace picca
