// Pieces shared by the pair kernels of the correlation functions: the LDS histogram
// record, the per-lane run of consecutive pairs in one bin and its flush, the scalar-cache
// pixel read.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine.h"

namespace picca {

__device__ __forceinline__ double edge_distance(double t, double width) {
    double frac = t - floor(t);
    double lo = frac * width;
    double hi = (1.0 - frac) * width;
    return lo < hi ? lo : hi;
}

// histogram record of one bin in LDS
struct LdsBin {
    double we, xi, rp, rt, z;
    uint32_t nb;
    uint32_t pad;
};
static_assert(sizeof(LdsBin) == 48, "LdsBin layout");
static_assert(sizeof(XiBin) == 48, "XiBin layout");

// Staged pixel j of delta2.  The fast loop reads the 48-byte record (pre-scaled
// coordinates, weights, z/2); the exact coordinates sit in a parallel 16-byte record that
// only the exact re-evaluation near bin edges reads.
struct TilePixel {
    double cr, sd;  // cos(ang/2) * r_comov, sin(ang/2) * dist_m
    double w, dw;   // weight, delta * weight
    double zh;      // z / 2
    double pad;
};
static_assert(sizeof(TilePixel) == 48, "TilePixel layout");

struct TileExact {
    double r, d;    // r_comov, dist_m
};
static_assert(sizeof(TileExact) == 16, "TileExact layout");

struct LaneRun {
    int32_t bin;
    uint32_t count;
    double we, xi, rp, rt, z;
};

// LDS-qualified view of the histogram: 32-bit addressing (a generic pointer costs a 64-bit
// multiply-add per flush)
typedef __attribute__((address_space(3))) LdsBin LocalBin;

typedef __attribute__((address_space(3))) unsigned char LocalByte;

__device__ __forceinline__ void flush_run(const LaneRun &run, LocalBin *hist) {
    // bin < 2^24: one 24-bit multiply-add forms the byte address
    LocalBin *b = (LocalBin *)((LocalByte *)hist +
                               __umul24((uint32_t)run.bin, (uint32_t)sizeof(LdsBin)));
    __hip_atomic_fetch_add(&b->we, run.we, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&b->xi, run.xi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&b->rp, run.rp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&b->rt, run.rt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&b->z, run.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&b->nb, run.count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void flush_run(const LaneRun &run, XiBin *hist) {
    XiBin *b = hist + run.bin;
    unsafeAtomicAdd(&b->we, run.we);
    unsafeAtomicAdd(&b->xi, run.xi);
    unsafeAtomicAdd(&b->rp, run.rp);
    unsafeAtomicAdd(&b->rt, run.rt);
    unsafeAtomicAdd(&b->z, run.z);
    atomicAdd(&b->nb, (unsigned long long)run.count);
}

// Reads one pixel record through the constant address space, so a wave-uniform address
// becomes scalar loads.
__device__ __forceinline__ PixelRecord load_record(const PixelRecord *p) {
    typedef const __attribute__((address_space(4))) double ConstDouble;
    const ConstDouble *q = (const ConstDouble *)p;
    PixelRecord out;
    out.r = q[0];
    out.d = q[1];
    out.z = q[2];
    out.w = q[3];
    out.dw = q[4];
    return out;
}

struct LaneCounters {
    uint64_t visit;
    uint64_t accept;
    uint32_t edge;
    uint32_t overflow;
    uint64_t runs;     // run walk: runs and probes of the lane
    uint64_t probes;
};

// Adds one accepted pair to the lane's run, flushing the run when the bin changes.  The
// restart is a multiply by 0 instead of a branch that zeroes registers (the engine only
// accepts finite inputs, so the sums are finite).
template <typename Bin>
__device__ __forceinline__ void add_pair(LaneRun &run, Bin *hist, LaneCounters &cnt, int bin,
                                         double w12, double dw12, double r_par, double r_trans,
                                         double z) {
    const bool change = bin != run.bin;
    if (change & (run.bin >= 0)) {
        flush_run(run, hist);
        cnt.accept += run.count;
    }
    const double keep = change ? 0.0 : 1.0;
    run.bin = bin;
    run.count = (change ? 0u : run.count) + 1u;
    run.we = fma(run.we, keep, w12);
    run.xi = fma(run.xi, keep, dw12);
    run.rp = fma(r_par, w12, run.rp * keep);
    run.rt = fma(r_trans, w12, run.rt * keep);
    run.z = fma(z, w12, run.z * keep);
}

}  // namespace picca
