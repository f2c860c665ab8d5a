// Pair kernel of the 3D correlation function on gfx950 (MI355X).
//
// GPU restatement of compute_xi_forest_pairs_fast (py/picca/cf.py:250-387) for a whole
// batch of forest pairs, grouped by output row (the healpix of delta1: picca_cf calls
// compute_xi once per healpix, py/picca/bin/picca_cf.py:452-458).
//
// Work decomposition
//   * one persistent workgroup of 1024 threads (16 wave64) per CU keeps the histogram of
//     its current output row in LDS, 48 B per bin (5 FP64 sums + u32 count, 120 KB at
//     50x50 bins), and adds it to the row in HBM with FP64 atomics when the row changes.
//     When np*nt does not fit in LDS the runs are added straight to the row in HBM;
//   * workgroups pull "groups" (<= 64 consecutive pairs of one row) from an HBM queue;
//   * inside a group, waves pull tasks = (pair, chunk of 64 pixels of delta1) from an
//     LDS counter.  Lane l owns pixel i = 64*chunk + l of delta1 in registers; the wave
//     walks the pixels j of delta2 in lockstep;
//   * pixels j are staged per wave through LDS: lane l loads pixel j0 + l (coalesced),
//     pixels the reference skips before any arithmetic (w2 == 0, zerr cut) are dropped,
//     the rest compacted, and the next tile is already loading while the wave walks the
//     current one with wave-uniform (broadcast) LDS reads;
//   * for the standard (r_par, r_trans) binning the j range is pruned to the window where
//     |r_par| < r_par_max and r_trans < r_trans_max can hold for some lane (r_comov and
//     dist_m are non-decreasing along a forest; forests where they are not are walked in
//     full).  The skipped pixel pairs are provably rejected by the reference's range test;
//   * a lane accumulates consecutive accepted pairs that land in the same bin in
//     registers (r_par and r_trans move monotonically with j) and adds the run to the
//     histogram with six LDS atomics when its bin changes.
//
// Arithmetic: r_par, r_trans and the bin index use the reference's FP64 operations in the
// reference's order (built with -ffp-contract=off).  The bin index is first taken with a
// reciprocal multiply; whenever the fractional part lies within the edge tolerance (plus
// a rounding margin) of an integer it is recomputed with the reference's
// divide-then-multiply, so every pair gets the reference's bin.  The mean pair redshift
// (z1 + z2) / 2 is formed as z1/2 + z2/2, which is the same double.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_util.h"
#include "engine.h"
#include "pair_common.h"

namespace picca {

// Everything compute_xi_forest_pairs_fast does with one pixel pair after the weight and
// zerr skips, with the reference's arithmetic (py/picca/cf.py:336-387): pair-redshift cut,
// separations, range test, bins (reciprocal multiply, exact divide-then-multiply near an
// edge), same-half-plate cut, overflow guard.  Returns true and sets bin / r_par / r_trans
// for an accepted pair.
template <int MODE, bool XCORR, bool GENERAL>
__device__ __forceinline__ bool exact_pair(const XiConst &C, double r1, double d1, double r2,
                                           double d2, double z, double ang, double cos_half,
                                           double sin_half, bool same_half_plate, int &bin,
                                           double &r_par, double &r_trans, uint32_t &edge,
                                           uint32_t &overflow) {
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
        // pairs just outside the range are edge pairs too
        if ((r_par >= C.rpmax && r_par - C.rpmax < C.edge_tol) ||
            (r_trans >= C.rtmax && r_trans - C.rtmax < C.edge_tol) ||
            (r_par < C.rpmin && C.rpmin - r_par < C.edge_tol))
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
    return true;
}

// One task: pixels [64*chunk, 64*chunk + 64) of delta1 against the pruned window of delta2.
//   GENERAL = false: standard binning, auto-correlation, r_par_min == 0, no z / zerr /
//   same-half-plate cuts (the cuts of the reference that cannot trigger are compiled out).
template <int MODE, bool XCORR, bool GENERAL, typename Bin>
__device__ void xi_task(const XiConst &C, const XiConst &C_lds, const ForestView &F,
                        const PairView &P, int64_t p, int chunk, int lane, Bin *hist,
                        TilePixel *tile, TileExact *tile_x, LaneCounters &cnt,
                        uint64_t &n_steps) {
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
    if (MODE == MODE_STANDARD && (C.flags & XI_PRUNE) && cos_half > 0 &&
        (uniform_int(F.mono[b]) & 3) == 3) {
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

    LaneRun run;
    run.bin = -1;
    run.count = 0;
    run.we = run.xi = run.rp = run.rt = run.z = 0.0;
    uint32_t visit = 0;
    uint32_t edge = 0;
    uint32_t overflow = 0;
    const int tile_len = C.tile_len;

    if (!GENERAL) {
        // Pixel j of delta2 is wave-uniform: its record comes through the scalar cache into
        // SGPRs and the loop touches LDS only to flush runs.  Two pixels per iteration, the
        // next two already loading; the store is padded, so reading up to three records
        // past j_hi stays in bounds.
        const PixelRecord *rec = F.pix + ob;
        uint32_t steps = 0;
        PixelRecord c0 = load_record(rec + j_lo), c1 = load_record(rec + j_lo + 1);
        auto step = [&](const PixelRecord &q) {
            // the reference skips w2 == 0 (either sign) before any arithmetic
            if ((__double_as_longlong(q.w) << 1) == 0) return;
            ++steps;
            double r_par = (r1 - q.r) * cos_half;
            const double r_trans = (d1 + q.d) * sin_half;
            const double t_par = fabs(r_par) * C.inv_p;
            const double t_trans = r_trans * C.inv_t;
            const double f_par = floor(t_par);
            const double f_trans = floor(t_trans);
            // r_par == 0 exactly (equal r_comov) is bin 0 and never an edge pair
            const bool near = ((fabs(t_par - f_par - 0.5) > C.half_p) & (r_par != 0)) |
                              (fabs(t_trans - f_trans - 0.5) > C.half_t);
            bool accept = active & !near & (t_par < C.np_d) & (t_trans < C.nt_d);
            int bin = (int)__umul24((unsigned)C.nt, (unsigned)(int)f_par) + (int)f_trans;
            double rp = r_par, rt = r_trans;
            if (active & near) {
                // exact re-evaluation (py/picca/cf.py:355-376)
                // (constants from the LDS copy: this rare path keeps no SGPRs busy)
                accept = exact_pair<MODE, XCORR, false>(C_lds, r1, d1, q.r, q.d, 0.0, ang,
                                                        cos_half, sin_half, false, bin, rp, rt,
                                                        edge, overflow);
            }
            if (accept)
                add_pair(run, hist, cnt, bin, w1 * q.w, dw1 * q.dw, fabs(rp), rt,
                         fma(q.z, 0.5, zh1));
        };
        for (int j = j_lo; j < j_hi; j += 2) {
            const PixelRecord n0 = load_record(rec + j + 2), n1 = load_record(rec + j + 3);
            step(c0);
            if (j + 1 < j_hi) step(c1);
            c0 = n0;
            c1 = n1;
        }
        n_steps += steps;
        if (active) visit += steps;
        if (run.bin >= 0) {
            flush_run(run, hist);
            cnt.accept += run.count;
        }
        cnt.visit += visit;
        cnt.edge += edge;
        cnt.overflow += overflow;
        return;
    }

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
        const int pos = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
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
        {
            for (int u = 0; u < n_tile; ++u) {
                const double2 wdw = *reinterpret_cast<const double2 *>(&tile[u].w);
                const double zh2 = tile[u].zh;
                const double2 rd = *reinterpret_cast<const double2 *>(&tile_x[u].r);
                const double z = zh1 + zh2;
                if (!active) continue;
                if (!((C.flags & XI_HAS_ZMIN) && z < C.zmin) &&
                    !((C.flags & XI_HAS_ZMAX) && z > C.zmax))
                    ++visit;
                int bin;
                double r_par, r_trans;
                if (exact_pair<MODE, XCORR, GENERAL>(C, r1, d1, rd.x, rd.y, z, ang, cos_half,
                                                     sin_half, same_half_plate, bin, r_par,
                                                     r_trans, edge, overflow))
                    add_pair(run, hist, cnt, bin, w1 * wdw.x, dw1 * wdw.y, r_par, r_trans, z);
            }
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

template <int MODE, bool XCORR, bool GENERAL, bool LDS_HIST>
__global__ void __launch_bounds__(XI_BLOCK)
xi_kernel(XiConst C, ForestView F, PairView P, GroupView G, XiOut O) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int nb = C.nb;
    LocalBin *hist = (LocalBin *)reinterpret_cast<LdsBin *>(smem);
    const size_t hist_bytes = LDS_HIST ? (size_t)nb * sizeof(LdsBin) : 0;
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
    __shared__ unsigned long long s_stats[XI_NSTATS];
    __shared__ XiConst s_const;

    const int tid = threadIdx.x;
    const int lane = tid & 63;

    if (LDS_HIST) {
        for (int k = tid; k < nb; k += XI_BLOCK) {
            hist[k].we = hist[k].xi = hist[k].rp = hist[k].rt = hist[k].z = 0.0;
            hist[k].nb = 0;
        }
    }
    if (tid < XI_NSTATS) s_stats[tid] = 0;
    if (tid == 0) s_const = C;  // published by the first __syncthreads of the group loop

    LaneCounters cnt = {0, 0, 0, 0};
    uint64_t n_steps = 0, n_tasks = 0, n_empty_tasks = 0;  // wave-uniform
    int current_row = -1;

    for (;;) {
        __syncthreads();
        if (tid == 0) s_group = (int64_t)atomicAdd(O.queue, 1ull);
        __syncthreads();
        const int64_t g = s_group;
        if (g >= G.count) break;
        const int row = G.row[g];
        if (LDS_HIST && row != current_row) {
            if (current_row >= 0) {
                XiBin *out = O.bins + (int64_t)current_row * nb;
                for (int k = tid; k < nb; k += XI_BLOCK) {
                    const uint32_t n = hist[k].nb;
                    if (n) {
                        unsafeAtomicAdd(&out[k].we, hist[k].we);
                        unsafeAtomicAdd(&out[k].xi, hist[k].xi);
                        unsafeAtomicAdd(&out[k].rp, hist[k].rp);
                        unsafeAtomicAdd(&out[k].rt, hist[k].rt);
                        unsafeAtomicAdd(&out[k].z, hist[k].z);
                        atomicAdd(&out[k].nb, (unsigned long long)n);
                        hist[k].we = hist[k].xi = hist[k].rp = hist[k].rt = hist[k].z = 0.0;
                        hist[k].nb = 0;
                    }
                }
            }
        }
        current_row = row;
        const int64_t p0 = G.begin[g];
        const int npairs = (int)(G.begin[g + 1] - p0);
        // chunk prefix over the pairs of the group (one wave)
        if (tid < 64) {
            int chunks = 0;
            if (tid < npairs) {
                const int a = P.a[p0 + tid];
                const int64_t la = F.off[a + 1] - F.off[a];
                chunks = (int)((la + 63) >> 6);
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
            const int chunk = uniform_int(t - s_prefix[q]);
            const uint64_t steps_before = n_steps;
            if (LDS_HIST) {
                xi_task<MODE, XCORR, GENERAL>(C, s_const, F, P, p0 + q, chunk, lane, hist, tile,
                                              tile_x, cnt, n_steps);
            } else {
                xi_task<MODE, XCORR, GENERAL>(C, s_const, F, P, p0 + q, chunk, lane,
                                              O.bins + (int64_t)row * nb, tile, tile_x, cnt,
                                              n_steps);
            }
            ++n_tasks;
            if (n_steps == steps_before) ++n_empty_tasks;
        }
    }
    // final flush (all waves are past the last __syncthreads of the loop)
    if (LDS_HIST && current_row >= 0) {
        XiBin *out = O.bins + (int64_t)current_row * nb;
        for (int k = tid; k < nb; k += XI_BLOCK) {
            const uint32_t n = hist[k].nb;
            if (n) {
                unsafeAtomicAdd(&out[k].we, hist[k].we);
                unsafeAtomicAdd(&out[k].xi, hist[k].xi);
                unsafeAtomicAdd(&out[k].rp, hist[k].rp);
                unsafeAtomicAdd(&out[k].rt, hist[k].rt);
                unsafeAtomicAdd(&out[k].z, hist[k].z);
                atomicAdd(&out[k].nb, (unsigned long long)n);
            }
        }
    }
    // statistics: lane values -> wave sums -> workgroup -> HBM
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
        atomicAdd(&s_stats[XI_STAT_EMPTY_TASKS], (unsigned long long)n_empty_tasks);
    }
    __syncthreads();
    if (tid < XI_NSTATS) atomicAdd(O.stats + tid, s_stats[tid]);
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

template <int MODE, bool XCORR, bool GENERAL, bool LDS_HIST>
static hipError_t launch_xi_t(const XiConst &C, const ForestView &F, const PairView &P,
                              const GroupView &G, const XiOut &O, int grid, size_t lds,
                              hipStream_t stream) {
    auto kern = xi_kernel<MODE, XCORR, GENERAL, LDS_HIST>;
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
    return lds_hist ? launch_xi_t<MODE, XCORR, GENERAL, true>(C, F, P, G, O, grid, lds, stream)
                    : launch_xi_t<MODE, XCORR, GENERAL, false>(C, F, P, G, O, grid, lds, stream);
}

// general == false is only valid for MODE_STANDARD, auto-correlation, r_par_min == 0 and
// no z / zerr / same-half-plate cuts (the host decides).
hipError_t launch_xi(int mode, bool xcorr, bool general, bool lds_hist, const XiConst &C,
                     const ForestView &F, const PairView &P, const GroupView &G,
                     const XiOut &O, int grid, size_t lds, hipStream_t stream) {
    if (mode == MODE_STANDARD) {
        if (!general && !xcorr)
            return launch_xi_h<MODE_STANDARD, false, false>(lds_hist, C, F, P, G, O, grid, lds, stream);
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
