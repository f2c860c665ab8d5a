// Internal declarations shared by the kernels and the C-ABI host code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace picca {

constexpr int XI_BLOCK = 1024;       // threads per workgroup of the pair kernel
constexpr int XI_GROUP_PAIRS = 64;   // max forest pairs per work item of the queue
constexpr double SPEED_LIGHT_KMS = 299792458.0 / 1000.0;
constexpr size_t PIXEL_PAD = 8;      // zero pixels after the last forest in device memory

enum : int { MODE_STANDARD = 0, MODE_RMU = 1, MODE_ANGULAR = 2 };

enum : int {
    XI_HAS_ZMIN = 1,
    XI_HAS_ZMAX = 2,
    XI_HAS_ZERR = 4,
    XI_REMOVE_SHP = 8,
    XI_PRUNE = 16,
};

enum : int { PAIR_SAME_HALF_PLATE = 1 };

enum : int {
    XI_STAT_VISITED = 0,
    XI_STAT_EDGE = 1,
    XI_STAT_OVERFLOW = 2,
    XI_STAT_STEPS = 3,
    XI_STAT_ACCEPTED = 4,
    XI_STAT_TASKS = 5,
    XI_STAT_EMPTY_TASKS = 6,
    XI_NSTATS = 7,
};

// Launch constants derived on the host from picca_cf_params.
struct XiConst {
    int32_t np, nt, nb, flags;
    int32_t tile_len;       // delta2 pixels staged per wave and tile (<= 64)
    int32_t pad_;
    double rpmin, rpmax, rtmax;
    double span_p;          // r_par_max - r_par_min
    double inv_p, inv_t;    // np / span_p, nt / r_trans_max
    double width_p;         // span_p / np   (also the same-half-plate cut)
    double width_t;         // r_trans_max / nt
    double fast_p, fast_t;  // fractional-part margins that trigger the exact re-check
    double half_p, half_t;  // 0.5 - fast_p, 0.5 - fast_t
    double edge_tol;
    double zmin, zmax;
    double zerr_ang;        // zerr_cut_deg * pi / 180
    double zerr_kms;
    double np_d, nt_d, nb_d;
};

// One pixel as the pair kernel streams it through the scalar cache.
struct PixelRecord {
    double r;   // r_comov
    double d;   // dist_m
    double z;
    double w;   // weight
    double dw;  // delta * weight
};

struct ForestView {
    const PixelRecord *__restrict__ pix;  // the same pixels, one 40-byte record each
    const double *__restrict__ z;
    const double *__restrict__ r;
    const double *__restrict__ d;
    const double *__restrict__ w;
    const double *__restrict__ dw;   // delta * weight, per pixel
    const double *__restrict__ ll;   // log10 lambda (dmat only)
    const int64_t *__restrict__ off; // [n + 1]
    const double *__restrict__ zq;
    const int32_t *__restrict__ order;
    const uint8_t *__restrict__ mono; // bit0 r_comov non-decreasing, bit1 dist_m non-decreasing
};

struct PairView {
    const int32_t *__restrict__ a;
    const int32_t *__restrict__ b;
    const double *__restrict__ ang;
    const double *__restrict__ cs;   // cos(ang/2), sin(ang/2) interleaved
    const uint8_t *__restrict__ flags;
};

struct GroupView {
    const int64_t *__restrict__ begin;  // [count + 1] pair offsets
    const int32_t *__restrict__ row;    // [count]
    int64_t count;
};

// One output bin of one row in HBM: the five FP64 sums of compute_xi and the pair count.
struct XiBin {
    double we, xi, rp, rt, z;
    unsigned long long nb;
};

struct XiOut {
    XiBin *bins;                 // [num_rows * nb]
    unsigned long long *queue;   // work-queue head (zeroed before each launch)
    unsigned long long *stats;   // [XI_NSTATS]
};

// LDS bytes of the pair kernel: histogram (48 B per bin, when it fits) + per-wave tiles
// of (tile_len + 1) staged pixels of 48 B + 16 B.
constexpr size_t XI_HIST_BIN_BYTES = 48;
constexpr size_t XI_TILE_PIXEL_BYTES = 64;

hipError_t launch_xi(int mode, bool xcorr, bool general, bool lds_hist, const XiConst &C,
                     const ForestView &F, const PairView &P, const GroupView &G,
                     const XiOut &O, int grid, size_t lds, hipStream_t stream);
hipError_t launch_zero_u64(unsigned long long *ptr, int64_t n, hipStream_t stream);

}  // namespace picca
