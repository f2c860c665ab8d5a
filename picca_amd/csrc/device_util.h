// Wave-level helpers shared by the kernels (wave64, gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

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

__device__ __forceinline__ int wave_min_int(int v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ int wave_max_int(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ double wave_sum_f64(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// First index in [0, n) with arr[idx] >= x (n if none) when `STRICT`, or with
// arr[idx] > x otherwise; arr must be non-decreasing and x wave-uniform.  All 64 lanes
// probe one level at a time, so the search takes ceil(log64 n) rounds.
template <bool STRICT>
__device__ int wave_search(const double *__restrict__ arr, int n, double x, int lane) {
    int lo = 0, len = n;
    while (len > 0) {
        const int stride = (len + 63) >> 6;
        const int probe = lane * stride;
        bool below = false;
        if (probe < len) below = STRICT ? (arr[lo + probe] < x) : (arr[lo + probe] <= x);
        const int cnt = __popcll(__ballot(below));
        if (cnt == 0) break;
        const int start = lo + (cnt - 1) * stride + 1;
        int stop = lo + cnt * stride;
        if (stop > lo + len) stop = lo + len;
        lo = start;
        len = stop - start;
    }
    return lo;
}

// DPP row moves (gfx9 dpp_ctrl): lane l of a row of 16 reads lane l - n (ROW_SHR) or
// l + n (ROW_SHL) of the same row, and 0 where that lane lies outside the row (bound_ctrl).
// Every lane of the wave must be active.
constexpr int DPP_ROW_SHL1 = 0x101;
constexpr int DPP_ROW_SHR1 = 0x111;
constexpr int DPP_ROW_SHR2 = 0x112;
constexpr int DPP_ROW_SHR4 = 0x114;
constexpr int DPP_ROW_SHR8 = 0x118;

template <int CTRL>
__device__ __forceinline__ int dpp_int(int v) {
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, true);
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)dpp_int<CTRL>((int)(uint32_t)b);
    const uint32_t hi = (uint32_t)dpp_int<CTRL>((int)(uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

}  // namespace picca
