// Wick covariance terms T1, T2, T3 of the auto-correlation on gfx950 (MI355X):
// compute_wickT123_pairs of py/picca/cf.py:1486-1602 for a batch of forest pairs.
//
// For each pair the reference lists its in-range pixel pairs (i, j) with bin p, adds
// w1w2 to weights_wick[p], 1 to num_pairs_wick[p] and w1w2 zev1 zev2 to T1[p, p], and then,
// for every two listed pairs a < b, adds one product to [p_a, p_b] and [p_b, p_a]:
//     i_a == i_b:  Wx2[j_a, j_b] w1[i] zev1[i]                        -> T2
//     j_a == j_b:  Wx1[i_a, i_b] w2[j] zev2[j]                        -> T2
//     otherwise:   Wx1[i_a, i_b] Wx2[j_a, j_b]                        -> T3
// (Wx the forests' weighted 1-D correlations).  That is T[P, Q] = sum over a in P, b in Q,
// b != a, of the product, so:
//   wick_list_kernel  one workgroup per pair: the in-range pixel pairs sorted by bin
//                     (LDS counting sort) into a scratch list, the per-bin sums and T1,
//                     and one task per non-empty bin;
//   wick_row_kernel   one workgroup per task (pair, bin P): the pixel pairs a of P on the
//                     lanes; the partners b come as segments (bin Q, j_b, i_lo..i_hi):
//                     the pixels of delta1 that pixel j_b pairs with in Q, consecutive.
//                     Over a segment the products share Wx2[j_a, j_b] (or w2 zev2), so
//                     their sum needs only sum_{i_b = i_lo..i_hi} Wx1[i_a, i_b], one
//                     difference of the row prefix sums c1 of Wx1 (minus the i_b == i_a
//                     term, which is a T2 product or a itself).  The segments (sorted by
//                     bin) are staged through LDS and walked in step by the whole wave;
//                     the lanes' sums are reduced when Q changes and added to an LDS row
//                     of T2 / T3, which goes to HBM at the end.
// FP64 throughout; bins by truncation after divide-then-multiply, as the reference.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_util.h"
#include "engine.h"

namespace picca {

namespace {

__device__ __forceinline__ void lds_add(double *p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// bin of pixel pair (i, j) or -1 when out of range; r_trans from the comoving distances
__device__ __forceinline__ int wick_bin(const WickConst &C, double r1, double r2, double c,
                                        double s) {
    double r_par = (r1 - r2) * c;
    if (!C.xcorr) r_par = fabs(r_par);
    const double r_trans = (r1 + r2) * s;
    if (!(r_par < C.rpmax && r_trans < C.rtmax && r_par >= C.rpmin)) return -1;
    const int bp = (int)((r_par - C.rpmin) / C.span * C.np);
    const int bt = (int)(r_trans / C.rtmax * C.nt);
    const int p = bt + C.nt * bp;
    return p < C.nb ? p : -2;   // -2: rounds up to np (the reference writes out of bounds)
}

}  // namespace

__global__ __launch_bounds__(WICK_BLOCK) void wick_list_kernel(WickConst C, WickInputs I,
                                                               WickPairs P, WickWork W,
                                                               WickOut O) {
    extern __shared__ __align__(16) unsigned char smem[];
    double *hw = (double *)smem;          // [nb] sum w1 w2
    double *ht = hw + C.nb;               // [nb] sum w1 w2 zev1 zev2
    int *cnt = (int *)(ht + C.nb);        // [nb] pairs per bin, then fill positions
    int *scnt = cnt + C.nb;               // [nb] segments per bin, then fill positions
    __shared__ int s_part[WICK_BLOCK];
    __shared__ int s_spart[WICK_BLOCK];
    __shared__ unsigned long long s_over;
    const int s = blockIdx.x;
    const int tid = threadIdx.x;
    const int a = P.a[s], b = P.b[s];
    const int64_t oa = I.off[a], ob = I.off[b];
    const int n1 = (int)(I.off[a + 1] - oa), n2 = (int)(I.off[b + 1] - ob);
    const double c = P.cs[2 * s], sn = P.cs[2 * s + 1];
    for (int k = tid; k < C.nb; k += WICK_BLOCK) {
        hw[k] = 0.0;
        ht[k] = 0.0;
        cnt[k] = 0;
        scnt[k] = 0;
    }
    if (tid == 0) s_over = 0;
    __syncthreads();
    const int64_t total = (int64_t)n1 * n2;
    for (int64_t idx = tid; idx < total; idx += WICK_BLOCK) {
        const int j = (int)(idx / n1), i = (int)(idx - (int64_t)j * n1);
        const int p = wick_bin(C, I.r[oa + i], I.r[ob + j], c, sn);
        if (p == -2) atomicAdd(&s_over, 1ull);
        if (p < 0) continue;
        const double w12 = I.w[oa + i] * I.w[ob + j];
        atomicAdd(&cnt[p], 1);
        lds_add(hw + p, w12);
        lds_add(ht + p, w12 * (I.zev1[oa + i] * I.zev2[ob + j]));
        // a segment starts where pixel i - 1 is not in bin p with pixel j
        if (i == 0 || wick_bin(C, I.r[oa + i - 1], I.r[ob + j], c, sn) != p)
            atomicAdd(&scnt[p], 1);
    }
    __syncthreads();
    // per-bin sums, T1 and the tasks; then exclusive prefix sum of the counts
    for (int k = tid; k < C.nb; k += WICK_BLOCK) {
        if (cnt[k] == 0) continue;
        unsafeAtomicAdd(O.weights_wick + k, hw[k]);
        atomicAdd(O.num_pairs_wick + k, (unsigned long long)cnt[k]);
        unsafeAtomicAdd(O.t1 + (int64_t)k * C.nb + k, ht[k]);
    }
    const int per = (C.nb + WICK_BLOCK - 1) / WICK_BLOCK;
    const int k0 = tid * per, k1 = min(C.nb, k0 + per);
    int part = 0, spart = 0;
    for (int k = k0; k < k1; ++k) {
        part += cnt[k];
        spart += scnt[k];
    }
    s_part[tid] = part;
    s_spart[tid] = spart;
    __syncthreads();
    if (tid == 0) {
        int run = 0, srun = 0;
        for (int t = 0; t < WICK_BLOCK; ++t) {
            const int v = s_part[t], sv = s_spart[t];
            s_part[t] = run;
            s_spart[t] = srun;
            run += v;
            srun += sv;
        }
        W.count[s] = run;
        W.seg_count[s] = srun;
        if (s_over) atomicAdd(O.stats + WICK_STAT_OVERFLOW, s_over);
        atomicAdd(O.stats + WICK_STAT_IN_RANGE, (unsigned long long)run);
    }
    __syncthreads();
    int run = s_part[tid];
    int srun = s_spart[tid];
    for (int k = k0; k < k1; ++k) {
        const int sv = scnt[k];
        scnt[k] = srun;   // segment fill position
        srun += sv;
        const int v = cnt[k];
        if (v > 0) {
            const unsigned long long t = atomicAdd(W.num_tasks, 1ull);
            if (t < (unsigned long long)W.max_tasks) {
                W.task_pair[t] = s;
                W.task_bin[t] = k;
                W.task_start[t] = run;
                W.task_count[t] = v;
            }
        }
        cnt[k] = run;   // fill position
        run += v;
    }
    __syncthreads();
    int64_t *list = W.list + W.base[s];
    int4 *seg = W.seg + W.base[s];
    // i-major, so that a bin's list keeps the pairs of one pixel i together: the lanes of
    // the row kernel then share few rows of the prefix sums
    for (int64_t idx = tid; idx < total; idx += WICK_BLOCK) {
        const int i = (int)(idx / n2), j = (int)(idx - (int64_t)i * n2);
        const double rj = I.r[ob + j];
        const int p = wick_bin(C, I.r[oa + i], rj, c, sn);
        if (p < 0) continue;
        const int pos = atomicAdd(&cnt[p], 1);
        // packed entry: bin (bits 40..63), j (20..39), i (0..19)
        list[pos] = ((int64_t)p << 40) | ((int64_t)j << 20) | (int64_t)i;
        if (i == 0 || wick_bin(C, I.r[oa + i - 1], rj, c, sn) != p) {
            int hi = i;
            while (hi + 1 < n1 && wick_bin(C, I.r[oa + hi + 1], rj, c, sn) == p) ++hi;
            seg[atomicAdd(&scnt[p], 1)] = make_int4(p, j, i, hi);
        }
    }
}

__global__ __launch_bounds__(WICK_BLOCK) void wick_row_kernel(WickConst C, WickInputs I,
                                                              WickPairs P, WickWork W,
                                                              WickOut O, int split) {
    extern __shared__ __align__(16) unsigned char smem[];
    double *r2 = (double *)smem;                  // [nb] row of T2
    double *r3 = r2 + C.nb;                       // [nb] row of T3
    int4 *sb = (int4 *)(r3 + C.nb);               // [WICK_STAGE] staged segments
    __shared__ unsigned long long s_products;
    // workgroup g of `split` for task blockIdx / split takes part g of the pair's segments
    const int task = blockIdx.x / split;
    const int g = blockIdx.x - task * split;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int s = W.task_pair[task];
    const int a_start = W.task_start[task], a_count = W.task_count[task];
    const int n_all = W.seg_count[s];
    const int part = (n_all + split - 1) / split;
    const int lo = min(n_all, g * part), hi = min(n_all, lo + part);
    const int fa = P.a[s], fb = P.b[s];
    const int64_t oa = I.off[fa], ob = I.off[fb];
    const int n1 = (int)(I.off[fa + 1] - oa), n2 = (int)(I.off[fb + 1] - ob);
    const double *wx1 = I.wx + I.wx_off[fa];
    const double *wx2 = I.wx + I.wx_off[fb];
    const double *c1 = I.c1 + I.c1_off[fa];
    const int64_t *list = W.list + W.base[s];
    const int4 *segs = W.seg + W.base[s];
    for (int k = tid; k < C.nb; k += WICK_BLOCK) {
        r2[k] = 0.0;
        r3[k] = 0.0;
    }
    if (tid == 0) s_products = 0;
    // work units: (chunk of 64 pixel pairs a of bin P on the lanes, part of the segments)
    const int achunks = (a_count + 63) >> 6;
    const int bsplit = max(1, (WICK_BLOCK / 64) / achunks);
    const int units = achunks * bsplit;
    uint64_t products = 0;
    for (int s0 = lo; s0 < hi; s0 += WICK_STAGE) {
        const int ns = min(WICK_STAGE, hi - s0);
        __syncthreads();
        for (int k = tid; k < ns; k += WICK_BLOCK) sb[k] = segs[s0 + k];
        __syncthreads();
        for (int u = wave; u < units; u += WICK_BLOCK / 64) {
            const int c = u / bsplit, q = u - c * bsplit;
            const int ai = c * 64 + lane;
            const bool live = ai < a_count;
            int ia = 0, ja = 0;
            double w1z1 = 0.0, w2z2 = 0.0, d1 = 0.0;
            const double *c1row = c1;
            if (live) {
                const int64_t e = list[a_start + ai];
                ia = (int)(e & 0xFFFFF);
                ja = (int)((e >> 20) & 0xFFFFF);
                w1z1 = I.w[oa + ia] * I.zev1[oa + ia];
                w2z2 = I.w[ob + ja] * I.zev2[ob + ja];
                d1 = wx1[(int64_t)ia * n1 + ia];
                c1row = c1 + (int64_t)ia * (n1 + 1);
            }
            const int per = (ns + bsplit - 1) / bsplit;
            const int k0 = min(ns, q * per), k1 = min(ns, k0 + per);
            int cur = -1;
            double acc2 = 0.0, acc3 = 0.0;
            auto flush = [&]() {
                const double t2v = wave_sum_f64(acc2), t3v = wave_sum_f64(acc3);
                if (lane == 0) {
                    if (t2v != 0.0) lds_add(r2 + cur, t2v);
                    if (t3v != 0.0) lds_add(r3 + cur, t3v);
                }
                acc2 = acc3 = 0.0;
            };
            for (int k = k0; k < k1; ++k) {
                const int4 e = sb[k];   // (bin, j_b, i_lo, i_hi)
                if (e.x != cur) {
                    if (cur >= 0) flush();
                    cur = e.x;
                }
                if (!live) continue;
                // sum of Wx1[i_a, i_b] over the segment without i_b == i_a
                const bool in = ia >= e.z && ia <= e.w;
                double s1 = c1row[e.w + 1] - c1row[e.z];
                if (in) s1 -= d1;
                products += (uint64_t)(e.w - e.z + 1) - ((in && e.y == ja) ? 1 : 0);
                if (e.y == ja) {
                    acc2 += s1 * w2z2;   // same j: Wx1[i_a, i_b] w2 zev2
                } else {
                    const double x2 = wx2[(int64_t)e.y * n2 + ja];
                    acc3 += x2 * s1;                // i_b != i_a: Wx1 Wx2
                    if (in) acc2 += x2 * w1z1;      // i_b == i_a: Wx2[j_a, j_b] w1 zev1
                }
            }
            if (cur >= 0) flush();
        }
    }
    const uint64_t wp = wave_sum_u64(products);
    if (lane == 0) atomicAdd(&s_products, (unsigned long long)wp);
    __syncthreads();
    const int pbin = W.task_bin[task];
    const int64_t row = (int64_t)pbin * C.nb;
    for (int k = tid; k < C.nb; k += WICK_BLOCK) {
        if (r2[k] != 0.0) unsafeAtomicAdd(O.t2 + row + k, r2[k]);
        if (r3[k] != 0.0) unsafeAtomicAdd(O.t3 + row + k, r3[k]);
    }
    if (tid == 0) atomicAdd(O.stats + WICK_STAT_PRODUCTS, s_products);
}

// c1[i][k] = sum_{m < k} Wx[i][m] for the rows of one listed forest per workgroup
__global__ __launch_bounds__(256) void wick_prefix_kernel(const double *__restrict__ wx,
                                                          const int64_t *__restrict__ wx_off,
                                                          const int64_t *__restrict__ off,
                                                          const int32_t *__restrict__ forests,
                                                          double *__restrict__ c1,
                                                          const int64_t *__restrict__ c1_off) {
    const int f = forests[blockIdx.x];
    const int n = (int)(off[f + 1] - off[f]);
    const double *m = wx + wx_off[f];
    double *out = c1 + c1_off[f];
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double *row = m + (int64_t)i * n;
        double *dst = out + (int64_t)i * (n + 1);
        double acc = 0.0;
        dst[0] = 0.0;
        for (int k = 0; k < n; ++k) {
            acc += row[k];
            dst[k + 1] = acc;
        }
    }
}

hipError_t launch_wick_prefix(const double *wx, const int64_t *wx_off, const int64_t *off,
                              const int32_t *forests, int32_t num_forests, double *c1,
                              const int64_t *c1_off, hipStream_t stream) {
    if (num_forests <= 0) return hipSuccess;
    hipLaunchKernelGGL(wick_prefix_kernel, dim3(num_forests), dim3(256), 0, stream, wx, wx_off,
                       off, forests, c1, c1_off);
    return hipGetLastError();
}

size_t wick_list_lds(const WickConst &C) { return (size_t)C.nb * (8 + 8 + 4 + 4); }
size_t wick_row_lds(const WickConst &C) {
    return (size_t)C.nb * 16 + (size_t)WICK_STAGE * sizeof(int4);
}

hipError_t launch_wick_list(const WickConst &C, const WickInputs &I, const WickPairs &P,
                            const WickWork &W, const WickOut &O, hipStream_t stream) {
    if (P.count <= 0) return hipSuccess;
    const size_t lds = wick_list_lds(C);
    hipError_t err = hipFuncSetAttribute((const void *)wick_list_kernel,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(wick_list_kernel, dim3(P.count), dim3(WICK_BLOCK), lds, stream, C, I, P,
                       W, O);
    return hipGetLastError();
}

hipError_t launch_wick_rows(const WickConst &C, const WickInputs &I, const WickPairs &P,
                            const WickWork &W, const WickOut &O, int64_t tasks,
                            hipStream_t stream) {
    if (tasks <= 0) return hipSuccess;
    const size_t lds = wick_row_lds(C);
    hipError_t err = hipFuncSetAttribute((const void *)wick_row_kernel,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (err != hipSuccess) return err;
    // enough workgroups for about eight waves per SIMD: each task's list split in parts
    const int64_t want = 8 * 1024 / (WICK_BLOCK / 64);
    const int split = (int)std::max<int64_t>(1, std::min<int64_t>(64, (want + tasks - 1) / tasks));
    hipLaunchKernelGGL(wick_row_kernel, dim3((unsigned)(tasks * split)), dim3(WICK_BLOCK), lds,
                       stream, C, I, P, W, O, split);
    return hipGetLastError();
}

}  // namespace picca
