// Metal distortion matrices on gfx950 (MI355X): compute_metal_dmat of picca's auto/cross
// correlation (py/picca/cf.py:890-1232) and of the forest x quasar correlation
// (py/picca/xcf.py:677-800), one pass (one absorber on delta1, one on delta2) over a list of
// kept pairs.
//
// For every pixel pair the reference takes the Lya coordinates for the data bin A and the
// absorbers' coordinates for the model bin M, and adds
//     weights_dmat[A] += w1 w2,
//     dmat[A, M] += w1 w2 f,  r_par_eff[M] += r_par' w1 w2 f,  r_trans_eff[M] += ...,
//     z_eff[M] += (z1' + z2') / 2 w1 w2 f,  weight_eff[M] += w1 w2 f,
// where f is the pair's redshift-evolution factor, A must be in range and M in range for
// the second group.  It is a histogram with no continuum-projection terms, so the kernel is
// a pair loop like the correlation's: one wave per (pair, 64 pixels of delta1), lanes over
// the pixels i, a scalar walk over the pixels j of delta2 inside the window that can give an
// in-range data bin.  Pairs with the same (A, M) merge into runs that travel along the
// diagonal (i + 1, j + 1) from lane to lane (a DPP rotation of the wave); a run adds to weights_dmat and the four
// effective sums in LDS (privatised per workgroup when they fit) and to dmat in HBM.  All
// arithmetic is FP64; the bins are the reference's divide-then-multiply ones (a reciprocal
// multiply away from bin edges); pixel weights are 0 where the absorber's redshift is not
// below the quasar's (the reference drops those pixels).
#include <hip/hip_runtime.h>
#include <math.h>

#include "device_util.h"
#include "engine.h"

namespace picca {
namespace {

struct MetalRun {
    int a, m;   // data bin, model bin (-1: out of range) of the run; a < 0: no run
    double w, wz, rp, rt, z;
};

__device__ __forceinline__ void lds_add(double *p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <bool LDS_HIST>
__device__ __forceinline__ void metal_flush(const MetalRun &run, const MetalConst &C,
                                            double *h_wdm, double *h_eff, const MetalOut &O) {
    if (run.a < 0) return;
    if (run.w != 0) {
        if (LDS_HIST)
            lds_add(h_wdm + run.a, run.w);
        else
            unsafeAtomicAdd(O.weights_dmat + run.a, run.w);
    }
    if (run.m < 0 || run.wz == 0) return;
    unsafeAtomicAdd(O.dmat + (int64_t)run.a * C.nm + run.m, run.wz);
    if (LDS_HIST) {
        lds_add(h_eff + run.m, run.rp);
        lds_add(h_eff + C.nm + run.m, run.rt);
        lds_add(h_eff + 2 * C.nm + run.m, run.z);
        lds_add(h_eff + 3 * C.nm + run.m, run.wz);
    } else {
        unsafeAtomicAdd(O.r_par_eff + run.m, run.rp);
        unsafeAtomicAdd(O.r_trans_eff + run.m, run.rt);
        unsafeAtomicAdd(O.z_eff + run.m, run.z);
        unsafeAtomicAdd(O.weight_eff + run.m, run.wz);
    }
}

// Flat bin of (r_par, r_trans) with n_par x n_trans bins, or -1 when out of range
// (py/picca/cf.py:1036-1052 / py/picca/xcf.py:733-742): floor of the r_par coordinate,
// truncation of the r_trans coordinate (numpy's astype(int)).  The coordinates come from a
// multiply by bins per Mpc/h; when one lies within the margin (0.5 - half) of an integer,
// both are recomputed with the reference's divide-then-multiply, so the bin is the
// reference's.  METAL_XCF counts a bin that rounds up to n_par (the reference's bincount
// would fail there) in `over`.
template <int MODE>
__device__ __forceinline__ int metal_bin(double r_par, double r_trans, const MetalConst &C,
                                         int n_par, int n_trans, double inv_p, double inv_t,
                                         double half_p, double half_t, uint32_t &over) {
    if (MODE == METAL_XCF && !(r_par > C.rpmin && r_par < C.rpmax && r_trans < C.rtmax))
        return -1;
    const double x = r_par - C.rpmin;
    double tp = x * inv_p, tt = r_trans * inv_t;
    double fp = floor(tp), ft = floor(tt);
    if ((fabs(tp - fp - 0.5) > half_p) | (fabs(tt - ft - 0.5) > half_t)) {
        tp = x / C.span * n_par;
        tt = r_trans / C.rtmax * n_trans;
        fp = floor(tp);
    }
    ft = trunc(tt);
    if (MODE == METAL_XCF) {
        const int bp = (int)fp;
        const int bt = (int)ft;
        const int bin = bt + n_trans * bp;
        if (bin >= n_par * n_trans) {
            ++over;
            return -1;
        }
        return bt >= 0 ? bin : -1;   // a negative r_trans (never for physical inputs)
    } else {
        if (!(fp < n_par && ft < n_trans && fp >= 0 && ft >= 0)) return -1;
        return (int)ft + n_trans * (int)fp;
    }
}

// Reads a wave-uniform value through the constant address space (a scalar load).
__device__ __forceinline__ double scalar_load(const double *p) {
    return *(const __attribute__((address_space(4))) double *)p;
}

// Lane l receives lane l - 1's value, lane 0 lane 63's (DPP wave_ror:1 over the full wave).
// Called with every lane of the wave active.
__device__ __forceinline__ int rotate_lanes(int v) {
    return __builtin_amdgcn_mov_dpp(v, 0x13c, 0xf, 0xf, false);
}

__device__ __forceinline__ double rotate_lanes(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)rotate_lanes((int)(uint32_t)b);
    const uint32_t hi = (uint32_t)rotate_lanes((int)(uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__device__ __forceinline__ MetalRun rotate_run(const MetalRun &r) {
    return MetalRun{rotate_lanes(r.a), rotate_lanes(r.m), rotate_lanes(r.w),
                    rotate_lanes(r.wz), rotate_lanes(r.rp), rotate_lanes(r.rt),
                    rotate_lanes(r.z)};
}

// Record of a wave-uniform pixel through the constant address space (scalar loads).
__device__ __forceinline__ MetalRecord load_metal_record(const MetalRecord *p) {
    typedef const __attribute__((address_space(4))) MetalRecord ConstRecord;
    const ConstRecord *q = (const ConstRecord *)p;
    MetalRecord out;
    out.r = q->r;
    out.d = q->d;
    out.w = q->w;
    out.za = q->za;
    out.ra = q->ra;
    out.da = q->da;
    out.f = q->f;
    out.pad = 0;
    return out;
}

// x == 0 (either sign) from the bits, so that a uniform value is tested on the scalar unit
__device__ __forceinline__ bool is_zero(double x) {
    return ((uint64_t)__double_as_longlong(x) << 1) == 0;
}

}  // namespace

template <int MODE, bool LDS_HIST>
__global__ __launch_bounds__(METAL_BLOCK) void metal_kernel(
    MetalConst C, const int64_t *__restrict__ off, const uint8_t *__restrict__ mono,
    const double *__restrict__ Fr, const double *__restrict__ Fd, MetalTable TA,
    MetalTable TB, MetalPairs P, const int32_t *__restrict__ task_pair, MetalOut O) {
    extern __shared__ __align__(16) double hist[];   // [nb] weights_dmat, [4][nm] eff
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    double *h_wdm = hist;
    double *h_eff = hist + C.nb;
    if (LDS_HIST) {
        for (int k = tid; k < C.nb + 4 * C.nm; k += METAL_BLOCK) hist[k] = 0.0;
        __syncthreads();
    }
    uint64_t n_data = 0, n_model = 0, n_over = 0, n_visit = 0;
    uint32_t over = 0;
    for (;;) {
        int64_t t = 0;
        if (lane == 0) t = (int64_t)atomicAdd(O.queue, 1ull);
        t = uniform_i64(t);
        if (t >= P.tasks) break;
        const int s = uniform_int(task_pair[t]);
        const int chunk = (int)(t - uniform_i64(P.task_begin[s]));
        const int a = uniform_int(P.a[s]);
        const int b = uniform_int(P.b[s]);
        const int64_t oa = uniform_i64(off[a]);
        const int la = (int)(uniform_i64(off[a + 1]) - oa);
        const int64_t ob = uniform_i64(off[b]);
        const int lb = (int)(uniform_i64(off[b + 1]) - ob);
        const double cth = uniform_f64(P.cs[2 * s]);
        const double sth = uniform_f64(P.cs[2 * s + 1]);
        const bool shp = MODE != METAL_XCF && C.shp && (P.flags[s] & PAIR_SAME_HALF_PLATE);

        const int i = chunk * 64 + lane;
        double r1 = 0, d1 = 0, w1 = 0, z1a = 0, r1a = 0, d1a = 0, f1a = 0;
        if (i < la) {
            const int64_t p = oa + i;
            r1 = Fr[p];
            d1 = Fd[p];
            w1 = TA.w[p];
            z1a = TA.z[p];
            r1a = TA.r[p];
            d1a = TA.d[p];
            f1a = TA.f[p];
        }
        const bool live = i < la && w1 != 0;
        if (__ballot(live) == 0) continue;

        // pixels j of delta2 that can give an in-range data bin with a live lane
        int j_lo = 0, j_hi = lb;
        const int mb = uniform_int((int)mono[b]);
        if (lb > 1 && (mb & 1)) {
            const double rmin = wave_min(live ? r1 : INFINITY);
            const double rmax = wave_max(live ? r1 : -INFINITY);
            double lo_r = rmin - C.rpmax / cth;
            double hi_r = MODE == METAL_AUTO ? rmax + C.rpmax / cth : rmax - C.rpmin / cth;
            lo_r -= 1e-6 + 1e-12 * fabs(lo_r);
            hi_r += 1e-6 + 1e-12 * fabs(hi_r);
            j_lo = uniform_int(wave_search<true>(Fr + ob, lb, lo_r, lane));
            j_hi = uniform_int(wave_search<false>(Fr + ob, lb, hi_r, lane));
            if ((mb & 2) && sth > 0) {
                const double dmin = wave_min(live ? d1 : INFINITY);
                double d_lim = C.rtmax / sth - dmin;
                d_lim += 1e-6 + 1e-12 * fabs(d_lim);
                j_hi = min(j_hi, uniform_int(wave_search<false>(Fd + ob, lb, d_lim, lane)));
            }
        }

        // Accumulators travel along the diagonal: after each pixel j every lane hands its
        // run to the next lane (lane 63 to lane 0), so lane l + 1 meets at pixel j + 1 the
        // run that lane l left at pixel j.  On a common wavelength grid (i + 1, j + 1) is
        // almost the separation of (i, j), so a run stays in one (A, M) cell for many
        // pixels and the matrix sees one atomic per run instead of one per lane and cell.
        // A lane flushes the run it receives when its own in-range pair falls in another
        // cell, and passes it on unchanged when it has no in-range pair; every run is
        // flushed at the end of the task.
        MetalRun run{-1, -1, 0, 0, 0, 0, 0};
        uint32_t visit = 0, data = 0, model = 0;   // this task's counts (j < 2^31)
        const double f1s = f1a * C.inv_den;
        // pixel j's record is in SGPRs; the next one loads while this one is evaluated (the
        // table's padding keeps the load of record j_hi in bounds)
        const MetalRecord *rec = TB.rec + ob;
        MetalRecord cur = load_metal_record(rec + j_lo);
        // wait for the first record here, so that the loop only ever waits for a record it
        // requested an iteration earlier
        asm volatile("" ::"s"(cur.r), "s"(cur.w), "s"(cur.ra), "s"(cur.f));
        for (int j = j_lo; j < j_hi; ++j) {
            const MetalRecord nxt = load_metal_record(rec + j + 1);
            if (!is_zero(cur.w)) {
                // branch-free pair: every lane evaluates both bins, the conditions select
                double r_par = (r1 - cur.r) * cth;
                if (MODE == METAL_AUTO) r_par = fabs(r_par);
                const double r_trans = (d1 + cur.d) * sth;
                uint32_t over_a = 0;
                int A = metal_bin<MODE>(r_par, r_trans, C, C.np, C.nt, C.inv_p, C.inv_t,
                                        C.half_p, C.half_t, over_a);
                double r_par_a = (r1a - cur.ra) * cth;
                if (MODE == METAL_AUTO) r_par_a = fabs(r_par_a);
                const double r_trans_a = (d1a + cur.da) * sth;
                uint32_t over_m = 0;
                int M = metal_bin<MODE>(r_par_a, r_trans_a, C, C.npm, C.ntm, C.inv_pm,
                                        C.inv_tm, C.half_pm, C.half_tm, over_m);
                A = live ? A : -1;
                const bool in_a = A >= 0;
                if (!in_a) M = -1;
                const bool in_m = M >= 0;
                // a data-bin overflow counts for a live lane, a model-bin one for a data pair
                over += (live ? over_a : 0u) + (in_a ? over_m : 0u);
                visit += live ? 1u : 0u;
                data += in_a ? 1u : 0u;
                model += in_m ? 1u : 0u;
                double w12 = w1 * cur.w;
                if (shp && fabs(r_par) < C.width_p) w12 = 0.0;
                const double c_w = in_a ? w12 : 0.0;
                const double wzw = in_m ? w12 * (f1s * cur.f) : 0.0;
                if (in_a && (A != run.a || M != run.m)) {
                    metal_flush<LDS_HIST>(run, C, h_wdm, h_eff, O);
                    run = MetalRun{A, M, 0, 0, 0, 0, 0};
                }
                // a pair out of range adds zeros to the run it passes on
                run.w += c_w;
                run.wz += wzw;
                run.rp += r_par_a * wzw;
                run.rt += r_trans_a * wzw;
                run.z += (z1a + cur.za) / 2 * wzw;
                run = rotate_run(run);
            }
            cur = nxt;
        }
        metal_flush<LDS_HIST>(run, C, h_wdm, h_eff, O);
        n_visit += visit;
        n_data += data;
        n_model += model;
    }
    n_over = over;

    const uint64_t v_data = wave_sum_u64(n_data);
    const uint64_t v_model = wave_sum_u64(n_model);
    const uint64_t v_over = wave_sum_u64(n_over);
    const uint64_t v_visit = wave_sum_u64(n_visit);
    if (lane == 0) {
        atomicAdd(O.stats + METAL_STAT_DATA, (unsigned long long)v_data);
        atomicAdd(O.stats + METAL_STAT_MODEL, (unsigned long long)v_model);
        atomicAdd(O.stats + METAL_STAT_OVERFLOW, (unsigned long long)v_over);
        atomicAdd(O.stats + METAL_STAT_VISITED, (unsigned long long)v_visit);
    }
    if (LDS_HIST) {
        __syncthreads();
        for (int k = tid; k < C.nb; k += METAL_BLOCK)
            if (h_wdm[k] != 0) unsafeAtomicAdd(O.weights_dmat + k, h_wdm[k]);
        double *eff[4] = {O.r_par_eff, O.r_trans_eff, O.z_eff, O.weight_eff};
        for (int k = tid; k < 4 * C.nm; k += METAL_BLOCK) {
            const double v = h_eff[k];
            if (v != 0) unsafeAtomicAdd(eff[k / C.nm] + (k % C.nm), v);
        }
    }
}

__global__ __launch_bounds__(256) void metal_pack_kernel(const double *__restrict__ r,
                                                         const double *__restrict__ d,
                                                         MetalTable T, int64_t num_pixels,
                                                         MetalRecord *__restrict__ rec) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= num_pixels + METAL_REC_PAD) return;
    MetalRecord out{0, 0, 0, 0, 0, 0, 0, 0};
    if (p < num_pixels) {
        out.r = r[p];
        out.d = d[p];
        out.w = T.w[p];
        out.za = T.z[p];
        out.ra = T.r[p];
        out.da = T.d[p];
        out.f = T.f[p];
    }
    rec[p] = out;
}

hipError_t launch_metal_pack(const double *r, const double *d, const MetalTable &T,
                             int64_t num_pixels, MetalRecord *rec, hipStream_t stream) {
    const int64_t n = num_pixels + METAL_REC_PAD;
    hipLaunchKernelGGL(metal_pack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       stream, r, d, T, num_pixels, rec);
    return hipGetLastError();
}

size_t metal_lds_bytes(const MetalConst &C) {
    return sizeof(double) * ((size_t)C.nb + 4 * (size_t)C.nm);
}

template <int MODE, bool LDS_HIST>
static hipError_t launch_metal_t(const MetalConst &C, const int64_t *off, const uint8_t *mono,
                                 const double *r, const double *d, const MetalTable &A,
                                 const MetalTable &B, const MetalPairs &P,
                                 const int32_t *task_pair, const MetalOut &O, int grid,
                                 hipStream_t stream) {
    auto kern = metal_kernel<MODE, LDS_HIST>;
    const size_t lds = LDS_HIST ? metal_lds_bytes(C) : 0;
    if (LDS_HIST) {
        hipError_t err = hipFuncSetAttribute((const void *)kern,
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)lds);
        if (err != hipSuccess) return err;
    }
    hipLaunchKernelGGL(kern, dim3(grid), dim3(METAL_BLOCK), lds, stream, C, off, mono, r, d, A,
                       B, P, task_pair, O);
    return hipGetLastError();
}

hipError_t launch_metal(const MetalConst &C, const int64_t *off, const uint8_t *mono,
                        const double *r, const double *d, const MetalTable &A,
                        const MetalTable &B, const MetalPairs &P, const int32_t *task_pair,
                        const MetalOut &O, bool lds_hist, int grid, hipStream_t stream) {
    if (P.tasks <= 0) return hipSuccess;
    switch (C.mode * 2 + (lds_hist ? 1 : 0)) {
        case METAL_AUTO * 2:
            return launch_metal_t<METAL_AUTO, false>(C, off, mono, r, d, A, B, P, task_pair, O,
                                                     grid, stream);
        case METAL_AUTO * 2 + 1:
            return launch_metal_t<METAL_AUTO, true>(C, off, mono, r, d, A, B, P, task_pair, O,
                                                    grid, stream);
        case METAL_CROSS * 2:
            return launch_metal_t<METAL_CROSS, false>(C, off, mono, r, d, A, B, P, task_pair,
                                                      O, grid, stream);
        case METAL_CROSS * 2 + 1:
            return launch_metal_t<METAL_CROSS, true>(C, off, mono, r, d, A, B, P, task_pair,
                                                     O, grid, stream);
        case METAL_XCF * 2:
            return launch_metal_t<METAL_XCF, false>(C, off, mono, r, d, A, B, P, task_pair, O,
                                                    grid, stream);
        case METAL_XCF * 2 + 1:
            return launch_metal_t<METAL_XCF, true>(C, off, mono, r, d, A, B, P, task_pair, O,
                                                   grid, stream);
        default:
            return hipErrorInvalidValue;
    }
}

}  // namespace picca
