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
//   wick_row_kernel   one workgroup per task (pair, bin P) and part of the pair's partner
//                     segments b (bin Q, j_b, i_lo..i_hi), one segment per lane:
//                     the pixels of delta1 that pixel j_b pairs with in Q, consecutive.
//                     Over a segment the products share Wx2[j_a, j_b] (or w2 zev2), so
//                     their sum needs only sum_{i_b = i_lo..i_hi} Wx1[i_a, i_b], one
//                     difference of the row prefix sums c1 of Wx1 (minus the i_b == i_a
//                     term, which is a T2 product or a itself).  The pixel pairs a of P
//                     are staged through LDS and walked in step by the wave, so each step
//                     reads one row of c1 (i_a) and one row of the symmetric Wx2 (j_a);
//                     each lane's sums go to an LDS row of T2 / T3 (at its segment's bin),
//                     which goes to HBM at the end.
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
    // i-major, so that a bin's list keeps the pairs of one pixel i together: consecutive
    // steps of the row kernel then read the same row of the prefix sums
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

// pixel pair a of the task's bin as the row kernel stages it
struct WickA {
    int32_t ia, ja;
    double w1z1, w2z2;   // w1 zev1 at i_a, w2 zev2 at j_a
    double d1;           // Wx1[i_a, i_a]
};
static_assert(sizeof(WickA) == 32, "WickA layout");
constexpr int WICK_ASTAGE = 512;   // pixel pairs a staged in LDS at a time

__global__ __launch_bounds__(WICK_BLOCK) void wick_row_kernel(WickConst C, WickInputs I,
                                                              WickPairs P, WickWork W,
                                                              WickOut O, int split) {
    extern __shared__ __align__(16) unsigned char smem[];
    double *r2 = (double *)smem;                  // [nb] row of T2
    double *r3 = r2 + C.nb;                       // [nb] row of T3
    WickA *sa = (WickA *)(r3 + C.nb);             // [WICK_ASTAGE] staged pixel pairs a
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
    // Lanes own segments b; the wave walks the staged pixel pairs a in step, so every read
    // of a step is in one row: row i_a of the prefix sums and row j_a of the symmetric Wx2.
    uint64_t products = 0;
    for (int a0 = 0; a0 < a_count; a0 += WICK_ASTAGE) {
        const int na = min(WICK_ASTAGE, a_count - a0);
        __syncthreads();
        for (int k = tid; k < na; k += WICK_BLOCK) {
            const int64_t e = list[a_start + a0 + k];
            WickA A;
            A.ia = (int)(e & 0xFFFFF);
            A.ja = (int)((e >> 20) & 0xFFFFF);
            A.w1z1 = I.w[oa + A.ia] * I.zev1[oa + A.ia];
            A.w2z2 = I.w[ob + A.ja] * I.zev2[ob + A.ja];
            A.d1 = wx1[(int64_t)A.ia * n1 + A.ia];
            sa[k] = A;
        }
        __syncthreads();
        for (int c0 = lo + wave * 64; c0 < hi; c0 += WICK_BLOCK) {
            const int sidx = c0 + lane;
            const bool live = sidx < hi;
            // (bin, j_b, i_lo, i_hi); a dead lane's empty segment reads c1[i_a][0] twice
            const int4 e = live ? segs[sidx] : make_int4(0, -1, 0, -1);
            double acc2 = 0.0, acc3 = 0.0;
            uint32_t cnt = 0;
#pragma unroll 8
            for (int k = 0; k < na; ++k) {
                const WickA A = sa[k];
                const double *crow = c1 + (int64_t)A.ia * (n1 + 1);
                // sum of Wx1[i_a, i_b] over the segment without i_b == i_a
                const bool in = A.ia >= e.z && A.ia <= e.w;
                const bool same_j = e.y == A.ja;
                // Wx2[j_a, j_b], 0 for j_b == j_a (those products are T2) and dead lanes
                double x2 = 0.0;
                if (live && !same_j) x2 = wx2[(int64_t)A.ja * n2 + e.y];
                double s1 = crow[e.w + 1] - crow[e.z];
                if (in) s1 -= A.d1;
                cnt += (uint32_t)(e.w - e.z + 1) - ((in && same_j) ? 1u : 0u);
                acc3 += x2 * s1;   // i_b != i_a, j_b != j_a: Wx1 Wx2
                // same j: Wx1[i_a, i_b] w2 zev2; i_b == i_a: Wx2[j_a, j_b] w1 zev1
                acc2 += same_j ? s1 * A.w2z2 : (in ? x2 * A.w1z1 : 0.0);
            }
            products += cnt;
            if (live) {
                if (acc2 != 0.0) lds_add(r2 + e.x, acc2);
                if (acc3 != 0.0) lds_add(r3 + e.x, acc3);
            }
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
    return (size_t)C.nb * 16 + (size_t)WICK_ASTAGE * sizeof(WickA);
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
