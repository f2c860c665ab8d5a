// Pair kernel of the 3D correlation function on gfx950 (MI355X).
//
// GPU restatement of compute_xi_forest_pairs_fast (py/picca/cf.py:250-387) for a whole
// batch of forest pairs, grouped by output row (healpix of delta1, as compute_xi is called
// per healpix in py/picca/bin/picca_cf.py:452-458).
//
// Work decomposition
//   * one persistent workgroup of 1024 threads (16 wave64) per CU; each keeps a private
//     histogram of the current output row in LDS: 5 FP64 sums + 1 u32 count per bin
//     (44 B/bin, 110 KB at 50x50 bins) and flushes it to HBM with FP64 atomics when the
//     row changes;
//   * workgroups pull "groups" (<= 64 consecutive pairs of one row) from an HBM queue;
//   * inside a group, waves pull tasks = (pair, chunk of 64 pixels of delta1) from an LDS
//     counter.  Lane l of the wave owns pixel i = 64*chunk + l of delta1 and keeps it in
//     registers; the wave walks pixel j of delta2 in lockstep, so the values of pixel j
//     are wave-uniform (scalar loads);
//   * for the standard (r_par, r_trans) binning the j range is pruned to the window
//     where |r_par| < r_par_max and r_trans < r_trans_max can hold for some lane
//     (r_comov and dist_m are non-decreasing along a forest; forests where they are not
//     are walked in full).  The skipped pixel pairs are provably rejected by the
//     reference's range test;
//   * a lane accumulates consecutive accepted pairs that land in the same bin in
//     registers (r_par and r_trans move monotonically with j) and adds the run to the
//     LDS histogram with FP64 / u32 LDS atomics when its bin changes.
//
// Arithmetic: r_par, r_trans and the bin index use the reference's FP64 operations in
// the reference's order (no contraction: built with -ffp-contract=off).  The bin index is
// first taken with a reciprocal multiply; whenever the fractional part lies within the
// edge tolerance (plus a rounding margin) of an integer it is recomputed with the
// reference's divide-then-multiply, so every pair gets the reference's bin.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine.h"

namespace picca {

__device__ __forceinline__ int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ int64_t uniform_i64(int64_t v) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ double uniform_f64(double v) {
    uint64_t bits = __double_as_longlong(v);
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)bits);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(bits >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__device__ __forceinline__ double wave_min(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// First index in [0, n) with arr[idx] >= x (n if none); arr must be non-decreasing.
// All 64 lanes probe one level at a time, so the search takes ceil(log64 n) rounds.
__device__ int wave_lower_bound(const double *__restrict__ arr, int n, double x, int lane) {
    int lo = 0, len = n;
    while (len > 0) {
        int stride = (len + 63) >> 6;
        int probe = lane * stride;
        bool below = false;
        if (probe < len) below = arr[lo + probe] < x;
        int cnt = __popcll(__ballot(below));
        if (cnt == 0) break;
        int start = lo + (cnt - 1) * stride + 1;
        int stop = lo + cnt * stride;
        if (stop > lo + len) stop = lo + len;
        lo = start;
        len = stop - start;
    }
    return lo;
}

// First index in [0, n) with arr[idx] > x (n if none); arr must be non-decreasing.
__device__ int wave_upper_bound(const double *__restrict__ arr, int n, double x, int lane) {
    int lo = 0, len = n;
    while (len > 0) {
        int stride = (len + 63) >> 6;
        int probe = lane * stride;
        bool below = false;
        if (probe < len) below = arr[lo + probe] <= x;
        int cnt = __popcll(__ballot(below));
        if (cnt == 0) break;
        int start = lo + (cnt - 1) * stride + 1;
        int stop = lo + cnt * stride;
        if (stop > lo + len) stop = lo + len;
        lo = start;
        len = stop - start;
    }
    return lo;
}

__device__ __forceinline__ double edge_distance(double t, double width) {
    double frac = t - floor(t);
    double lo = frac * width;
    double hi = (1.0 - frac) * width;
    return lo < hi ? lo : hi;
}

struct LaneAccumulator {
    uint32_t bin;
    uint32_t count;
    double we, xi, rp, rt, z;
};

__device__ __forceinline__ void flush_run(const LaneAccumulator &acc, double *h_we, double *h_xi,
                                          double *h_rp, double *h_rt, double *h_z,
                                          uint32_t *h_nb) {
    __hip_atomic_fetch_add(h_we + acc.bin, acc.we, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(h_xi + acc.bin, acc.xi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(h_rp + acc.bin, acc.rp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(h_rt + acc.bin, acc.rt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(h_z + acc.bin, acc.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(h_nb + acc.bin, acc.count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// One task: pixels [64*chunk, 64*chunk + 64) of delta1 against the pruned window of delta2.
template <int MODE, bool XCORR>
__device__ void xi_task(const XiConst &C, const ForestView &F, int64_t p, int chunk, int lane,
                        const PairView &P, double *h_we, double *h_xi, double *h_rp,
                        double *h_rt, double *h_z, uint32_t *h_nb, uint64_t &n_visit,
                        uint32_t &n_edge, uint32_t &n_overflow, uint64_t &n_steps,
                        uint64_t &n_accept, double *tile, int tile_len) {
    const int a = uniform_int(P.a[p]);
    const int b = uniform_int(P.b[p]);
    const int64_t oa = uniform_i64(F.off[a]);
    const int la = (int)(uniform_i64(F.off[a + 1]) - oa);
    const int64_t ob = uniform_i64(F.off[b]);
    const int lb = (int)(uniform_i64(F.off[b + 1]) - ob);
    const double ang = uniform_f64(P.ang[p]);
    const double cos_half = uniform_f64(P.cs[2 * p]);
    const double sin_half = uniform_f64(P.cs[2 * p + 1]);
    const bool same_half_plate = (uniform_int(P.flags[p]) & PAIR_SAME_HALF_PLATE) != 0;
    const double z_qso_1 = uniform_f64(F.zq[a]);
    const double z_qso_2 = uniform_f64(F.zq[b]);
    const bool zerr_on = (C.flags & XI_HAS_ZERR) && (ang < C.zerr_ang);

    const int i = chunk * 64 + lane;
    bool active = i < la;
    double r1 = 0, d1 = 0, z1 = 0, w1 = 0, dw1 = 0;
    if (active) {
        r1 = F.r[oa + i];
        d1 = F.d[oa + i];
        z1 = F.z[oa + i];
        w1 = F.w[oa + i];
        dw1 = F.dw[oa + i];
        active = w1 != 0;
    }
    if (zerr_on && active) {
        double z_qf = 0.5 * (z1 + z_qso_2);
        double dv_kms = fabs(z1 - z_qso_2) / (1 + z_qf);
        dv_kms *= SPEED_LIGHT_KMS;
        if (dv_kms < C.zerr_kms) active = false;
    }
    if (__ballot(active) == 0) return;

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
        double lo_r, hi_r;
        if (XCORR) {
            lo_r = r_min - C.rpmax / cos_half;
            hi_r = r_max - C.rpmin / cos_half;
        } else {
            lo_r = r_min - C.rpmax / cos_half;
            hi_r = r_max + C.rpmax / cos_half;
        }
        lo_r -= slack + 1e-12 * fabs(lo_r);
        hi_r += slack + 1e-12 * fabs(hi_r);
        j_lo = wave_lower_bound(Br, lb, lo_r, lane);
        j_hi = wave_upper_bound(Br, lb, hi_r, lane);
        if (sin_half > 0) {
            double d_lim = (C.rtmax + 4.0 * C.edge_tol) / sin_half - d_min;
            d_lim += 1e-6 + 1e-12 * fabs(d_lim);
            int j_d = wave_upper_bound(Bd, lb, d_lim, lane);
            if (j_d < j_hi) j_hi = j_d;
        }
    }
    j_lo = uniform_int(j_lo);
    j_hi = uniform_int(j_hi);

    LaneAccumulator acc;
    acc.bin = 0xffffffffu;
    acc.count = 0;
    acc.we = acc.xi = acc.rp = acc.rt = acc.z = 0.0;
    uint32_t visit = 0;
    uint32_t steps = 0;

    // Pixel j of delta2 is staged through this wave's LDS tile: lane l loads pixel
    // j0 + l (coalesced), pixels with w2 == 0 or failing the zerr cut are dropped
    // (the reference `continue`s on them before anything else), the survivors are
    // compacted, and the next tile's loads are already in flight while the wave
    // walks the current one with wave-uniform LDS reads.
    double *t_r = tile;
    double *t_d = tile + tile_len;
    double *t_z = tile + 2 * tile_len;
    double *t_w = tile + 3 * tile_len;
    double *t_dw = tile + 4 * tile_len;
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
        bool keep = pvalid && pw != 0;
        if (zerr_on && keep) {
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
            t_r[pos] = pr;
            t_d[pos] = pd;
            t_z[pos] = pz;
            t_w[pos] = pw;
            t_dw[pos] = pdw;
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
        steps += n_tile;
    for (int u = 0; u < n_tile; ++u) {
        const double r2 = t_r[u];
        const double d2 = t_d[u];
        const double z2 = t_z[u];
        const double w2 = t_w[u];
        const double dw2 = t_dw[u];

        bool ok = active;
        const double z = (z1 + z2) / 2;
        if ((C.flags & XI_HAS_ZMIN) && z < C.zmin) ok = false;
        if ((C.flags & XI_HAS_ZMAX) && z > C.zmax) ok = false;

        double r_par, r_trans;
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
        visit += ok ? 1u : 0u;
        const bool out = (r_par >= C.rpmax) | (r_trans >= C.rtmax) | (r_par < C.rpmin);
        if (ok && out) {
            if ((r_par >= C.rpmax && r_par - C.rpmax < C.edge_tol) ||
                (r_trans >= C.rtmax && r_trans - C.rtmax < C.edge_tol) ||
                (r_par < C.rpmin && C.rpmin - r_par < C.edge_tol))
                ++n_edge;
        }
        if (ok && !out) {
            double t_par = (r_par - C.rpmin) * C.inv_p;
            double t_trans = r_trans * C.inv_t;
            double f_par = floor(t_par);
            double f_trans = floor(t_trans);
            const double frac_p = t_par - f_par;
            const double frac_t = t_trans - f_trans;
            if (frac_p < C.fast_p || frac_p > 1.0 - C.fast_p || frac_t < C.fast_t ||
                frac_t > 1.0 - C.fast_t) {
                t_par = (r_par - C.rpmin) / C.span_p * C.np_d;
                t_trans = r_trans / C.rtmax * C.nt_d;
                f_par = floor(t_par);
                f_trans = floor(t_trans);
                // an exact zero (equal r_comov, or ang = 0) cannot move under rounding
                if ((r_par != 0 && edge_distance(t_par, C.width_p) < C.edge_tol) ||
                    (r_trans != 0 && edge_distance(t_trans, C.width_t) < C.edge_tol))
                    ++n_edge;
            }
            bool keep = true;
            if ((C.flags & XI_REMOVE_SHP) && same_half_plate && fabs(r_par) < C.width_p)
                keep = false;
            const double bin_f = f_trans + C.nt_d * f_par;
            if (keep && !(bin_f >= 0.0 && bin_f < C.nb_d)) {
                ++n_overflow;
                keep = false;
            }
            if (keep) {
                const uint32_t bin = (uint32_t)bin_f;
                const double w12 = w1 * w2;
                const double dw12 = dw1 * dw2;
                if (bin != acc.bin) {
                    if (acc.bin != 0xffffffffu) {
                        flush_run(acc, h_we, h_xi, h_rp, h_rt, h_z, h_nb);
                        n_accept += acc.count;
                    }
                    acc.bin = bin;
                    acc.count = 1;
                    acc.we = w12;
                    acc.xi = dw12;
                    acc.rp = r_par * w12;
                    acc.rt = r_trans * w12;
                    acc.z = z * w12;
                } else {
                    acc.count += 1;
                    acc.we += w12;
                    acc.xi += dw12;
                    acc.rp += r_par * w12;
                    acc.rt += r_trans * w12;
                    acc.z += z * w12;
                }
            }
        }
    }
        __builtin_amdgcn_wave_barrier();
    }
    if (acc.bin != 0xffffffffu) {
        flush_run(acc, h_we, h_xi, h_rp, h_rt, h_z, h_nb);
        n_accept += acc.count;
    }
    n_visit += visit;
    n_steps += steps;
}

template <int MODE, bool XCORR>
__global__ void __launch_bounds__(XI_BLOCK)
xi_kernel(XiConst C, ForestView F, PairView P, GroupView G, XiOut O) {
    extern __shared__ double smem[];
    const int nb = C.nb;
    double *h_we = smem;
    double *h_xi = h_we + nb;
    double *h_rp = h_xi + nb;
    double *h_rt = h_rp + nb;
    double *h_z = h_rt + nb;
    uint32_t *h_nb = reinterpret_cast<uint32_t *>(h_z + nb);
    // per-wave staging tiles after the histogram (8-byte aligned)
    double *tile = h_z + nb + (nb + 1) / 2 + (threadIdx.x >> 6) * 5 * C.tile_len;

    __shared__ int64_t s_group;
    __shared__ int s_task;
    __shared__ int s_prefix[XI_GROUP_PAIRS + 1];
    __shared__ unsigned long long s_stats[XI_NSTATS];

    const int tid = threadIdx.x;
    const int lane = tid & 63;

    for (int k = tid; k < nb; k += XI_BLOCK) {
        h_we[k] = h_xi[k] = h_rp[k] = h_rt[k] = h_z[k] = 0.0;
        h_nb[k] = 0;
    }
    if (tid < XI_NSTATS) s_stats[tid] = 0;

    uint64_t n_visit = 0, n_steps = 0, n_accept = 0;
    uint32_t n_edge = 0, n_overflow = 0;
    int current_row = -1;

    for (;;) {
        __syncthreads();
        if (tid == 0) s_group = (int64_t)atomicAdd(O.queue, 1ull);
        __syncthreads();
        const int64_t g = s_group;
        if (g >= G.count) break;
        const int row = G.row[g];
        if (row != current_row) {
            if (current_row >= 0) {
                const int64_t base = (int64_t)current_row * nb;
                for (int k = tid; k < nb; k += XI_BLOCK) {
                    const uint32_t cnt = h_nb[k];
                    if (cnt) {
                        unsafeAtomicAdd(O.we + base + k, h_we[k]);
                        unsafeAtomicAdd(O.xi + base + k, h_xi[k]);
                        unsafeAtomicAdd(O.rp + base + k, h_rp[k]);
                        unsafeAtomicAdd(O.rt + base + k, h_rt[k]);
                        unsafeAtomicAdd(O.z + base + k, h_z[k]);
                        atomicAdd(O.nb + base + k, (unsigned long long)cnt);
                        h_we[k] = h_xi[k] = h_rp[k] = h_rt[k] = h_z[k] = 0.0;
                        h_nb[k] = 0;
                    }
                }
            }
            current_row = row;
        }
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
            xi_task<MODE, XCORR>(C, F, p0 + q, chunk, lane, P, h_we, h_xi, h_rp, h_rt, h_z,
                                 h_nb, n_visit, n_edge, n_overflow, n_steps, n_accept, tile,
                                 C.tile_len);
        }
    }
    // final flush (all waves are past the last __syncthreads of the loop)
    if (current_row >= 0) {
        const int64_t base = (int64_t)current_row * nb;
        for (int k = tid; k < nb; k += XI_BLOCK) {
            const uint32_t cnt = h_nb[k];
            if (cnt) {
                unsafeAtomicAdd(O.we + base + k, h_we[k]);
                unsafeAtomicAdd(O.xi + base + k, h_xi[k]);
                unsafeAtomicAdd(O.rp + base + k, h_rp[k]);
                unsafeAtomicAdd(O.rt + base + k, h_rt[k]);
                unsafeAtomicAdd(O.z + base + k, h_z[k]);
                atomicAdd(O.nb + base + k, (unsigned long long)cnt);
            }
        }
    }
    // statistics: lane values -> wave sums -> workgroup -> HBM
    const uint64_t v_visit = wave_sum_u64(n_visit);
    const uint64_t v_edge = wave_sum_u64(n_edge);
    const uint64_t v_over = wave_sum_u64(n_overflow);
    const uint64_t v_accept = wave_sum_u64(n_accept);
    if (lane == 0) {
        atomicAdd(&s_stats[XI_STAT_VISITED], (unsigned long long)v_visit);
        atomicAdd(&s_stats[XI_STAT_EDGE], (unsigned long long)v_edge);
        atomicAdd(&s_stats[XI_STAT_OVERFLOW], (unsigned long long)v_over);
        atomicAdd(&s_stats[XI_STAT_STEPS], (unsigned long long)n_steps);
        atomicAdd(&s_stats[XI_STAT_ACCEPTED], (unsigned long long)v_accept);
    }
    __syncthreads();
    if (tid < XI_NSTATS) atomicAdd(O.stats + tid, s_stats[tid]);
}

__global__ void zero_kernel(double *__restrict__ a, int64_t n) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; k < n; k += stride) a[k] = 0.0;
}

__global__ void zero_u64_kernel(unsigned long long *__restrict__ a, int64_t n) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; k < n; k += stride) a[k] = 0ull;
}

hipError_t launch_zero(double *ptr, int64_t n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    zero_kernel<<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(ptr, n);
    return hipGetLastError();
}

hipError_t launch_zero_u64(unsigned long long *ptr, int64_t n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    zero_u64_kernel<<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(ptr, n);
    return hipGetLastError();
}

template <int MODE, bool XCORR>
static hipError_t launch_xi_t(const XiConst &C, const ForestView &F, const PairView &P,
                              const GroupView &G, const XiOut &O, int grid, size_t lds,
                              hipStream_t stream) {
    auto kern = xi_kernel<MODE, XCORR>;
    hipError_t err = hipFuncSetAttribute((const void *)kern,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (err != hipSuccess) return err;
    kern<<<dim3(grid), dim3(XI_BLOCK), lds, stream>>>(C, F, P, G, O);
    return hipGetLastError();
}

hipError_t launch_xi(int mode, bool xcorr, const XiConst &C, const ForestView &F,
                     const PairView &P, const GroupView &G, const XiOut &O, int grid,
                     size_t lds, hipStream_t stream) {
    if (mode == MODE_STANDARD)
        return xcorr ? launch_xi_t<MODE_STANDARD, true>(C, F, P, G, O, grid, lds, stream)
                     : launch_xi_t<MODE_STANDARD, false>(C, F, P, G, O, grid, lds, stream);
    if (mode == MODE_RMU)
        return xcorr ? launch_xi_t<MODE_RMU, true>(C, F, P, G, O, grid, lds, stream)
                     : launch_xi_t<MODE_RMU, false>(C, F, P, G, O, grid, lds, stream);
    return xcorr ? launch_xi_t<MODE_ANGULAR, true>(C, F, P, G, O, grid, lds, stream)
                 : launch_xi_t<MODE_ANGULAR, false>(C, F, P, G, O, grid, lds, stream);
}

}  // namespace picca
