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
    XI_WALK = 32,   // run walk for the eligible tasks of the specialised kernel
    XI_WALK_AUTO = 64,  // ... only where the window holds XI_WALK_MIN_DENSITY pixels per bin
};

// The automatic schedule walks a task when its delta2 window holds at least this many
// pixels per r_par bin: on the 0.8 A grid (about 7.5 per bin) the walk measured 29.0 s on
// the 100k survey against the pair loop's 31.1 s; on the log10-lambda grid (about 2 per bin)
// the pair loop is faster.
constexpr double XI_WALK_MIN_DENSITY = 4.0;

enum : int { PAIR_SAME_HALF_PLATE = 1, PAIR_XCF_COUNT = 2 };

enum : int {
    XI_STAT_VISITED = 0,
    XI_STAT_EDGE = 1,
    XI_STAT_OVERFLOW = 2,
    XI_STAT_STEPS = 3,
    XI_STAT_ACCEPTED = 4,
    XI_STAT_TASKS = 5,
    XI_STAT_EMPTY_TASKS = 6,
    XI_STAT_RUNS = 7,
    XI_STAT_PROBES = 8,
    XI_STAT_WALK_TASKS = 9,
    XI_NSTATS = 10,
};

// Launch constants derived on the host from picca_cf_params.
struct XiConst {
    int32_t np, nt, nb, flags;
    int32_t tile_len;       // delta2 pixels staged per wave and tile (<= 64)
    int32_t ntp;            // nt + 1: row stride of the LDS histogram of the xi kernel,
                            // whose last column and last row collect rejected pairs
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
    double ntp_d;           // nt + 1
    // r_trans row window of the cross-correlation kernel's LDS histogram: rows
    // [win_t0, win_t0 + win_rows) of the output, padded to (np + 2) x (win_rows + 2) bins
    int32_t win_t0, win_rows, win_last, win_pad_;
};

// One pixel as the pair kernel streams it through the scalar cache.
struct PixelRecord {
    double r;   // r_comov
    double d;   // dist_m
    double z;
    double w;   // weight
    double dw;  // delta * weight
};

struct PrefixRecord;  // xi_walk.h

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
    const double2 *__restrict__ rd;   // (r_comov, dist_m) per pixel
    const PrefixRecord *__restrict__ pre;       // run-walk prefix sums, forest b at off[b] + b
    const double2 *__restrict__ spc;  // per forest: smallest step of r_comov and of dist_m
                                      // between consecutive pixels, -1 unless both increase
                                      // strictly
};

struct PairView {
    const int32_t *__restrict__ a;
    const int32_t *__restrict__ b;
    const double *__restrict__ ang;
    const double *__restrict__ cs;   // cos(ang/2), sin(ang/2) interleaved
    const uint8_t *__restrict__ flags;
    int32_t *chunks;                 // per pair: first and end 64-pixel chunk of delta1 that
                                     // can reach delta2 (xi_chunk_kernel; the xi kernels read)
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

// 64-bit words of the no-allocation variant's device array: queue head + counters
constexpr int XI_SCRATCH_WORDS = 32;
static_assert(1 + XI_NSTATS <= XI_SCRATCH_WORDS, "xi scratch too small");

struct XiOut {
    XiBin *bins;                 // [num_rows * nb]
    unsigned long long *queue;   // work-queue head (zeroed before each launch)
    unsigned long long *stats;   // [XI_NSTATS]
};

// LDS bytes of the pair kernel: histogram (48 B per bin, when it fits) + per-wave tiles
// of (tile_len + 1) staged pixels of 48 B + 16 B.
constexpr size_t XI_HIST_BIN_BYTES = 48;
constexpr size_t XI_TILE_PIXEL_BYTES = 64;

// ---------------------------------------------------------------------------------------
// distortion matrix
// ---------------------------------------------------------------------------------------

constexpr int DMAT_PASS_BLOCK = 256;   // threads of the pair-pass kernel (lane per pixel i)
constexpr int DMAT_PASS_ROWS = 6;      // r_trans rows per LDS band of the pair pass
constexpr int DMAT_FILL_BLOCK = 256;   // threads of the fill kernel (4 waves)
constexpr int DMAT_FILL_RP = 4;        // 16-bin blocks of data r_par bins per chunk (one per wave)
constexpr int DMAT_FILL_MRP = 7;       // 16-bin blocks of model r_par bins per chunk
constexpr int DMAT_CHUNK_RP = 16 * DMAT_FILL_RP;
constexpr int DMAT_CHUNK_MRP = 16 * DMAT_FILL_MRP;

enum : int {
    DMAT_REDSHIFT_EVOL = 32,
    DMAT_XCORR = 64,
    DMAT_RMU = 128,
};

enum : int {
    DMAT_STAT_IN_RANGE = 0,     // in-range pairs of the main pass
    DMAT_STAT_SELECTED = 1,     // selected pairs with a valid model bin
    DMAT_STAT_BEYOND = 2,       // in-range pairs beyond the count-pass size
    DMAT_STAT_OVERFLOW = 3,     // selected pairs whose data bin is >= np*nt
    DMAT_STAT_NEGATIVE = 4,     // selected pairs whose data bin is < 0
    DMAT_STAT_MFMA = 5,         // FP64 matrix-core instructions of the fill (16x16x4 each)
    DMAT_NSTATS = 6,
};

struct DmatConst {
    int32_t np, nt, nb;        // data bins
    int32_t npm, ntm, nm;      // model bins
    int32_t flags;             // XI_HAS_* | XI_REMOVE_SHP | DMAT_*
    int32_t words_u;           // 32-bit words of a model-bin bitmap
    int32_t tb_max, ub_max;    // r_trans bins of the LDS bands of the pass kernel
    double rpmin, rpmax, rtmax, span_p, width_p;
    double zmin, zmax, zerr_ang, zerr_kms;
    double np_d, nt_d, npm_d, ntm_d;
    double inv_span_p, inv_rtmax;  // 1 / span_p, 1 / rtmax (bin coordinates by multiplies)
};

// Per-pixel and per-forest inputs of the distortion matrix (store order).
struct DmatForests {
    const double *__restrict__ r;
    const double *__restrict__ d;
    const double *__restrict__ z;
    const double *__restrict__ w;
    const double *__restrict__ dll;   // log_lambda - weighted mean log_lambda of the forest
    const double *__restrict__ zfa;   // ((1+z)/(1+z_ref))^(alpha-1)
    const double *__restrict__ zfb;   // ((1+z)/(1+z_ref))^(alpha2-1)
    const int64_t *__restrict__ off;
    const double *__restrict__ zq;
    const double *__restrict__ sum_w;   // per forest: sum of weights
    const double *__restrict__ sw_dll;  // per forest: sum of w * dll^2
    const int32_t *__restrict__ order;
    const uint8_t *__restrict__ mono;
};

// One batch of kept forest pairs.
struct DmatPairs {
    const int32_t *__restrict__ a;
    const int32_t *__restrict__ b;
    const double *__restrict__ ang;
    const double *__restrict__ cs;     // cos(ang/2), sin(ang/2)
    const uint8_t *__restrict__ flags;
    int32_t count;
};

// Per-slot workspace of a batch (slot = position of the pair in the batch).
// Data r_trans bin t / model r_trans bin u below are the remainders A % nt and M % ntm of
// the reference's flat bin indices A and M (so a bin index that rounds to nt lands where
// the reference writes it).
struct DmatWork {
    double *S;          // [slot][4][nt][np]   S0, Sj, Si, Sij per data bin (band only)
    double *eta;        // [slot][4][ntm][npm] eta5..eta8 per model bin, masked by U (band)
    uint32_t *U;        // [slot][words_u]     unique model bins of the fill loop
    int32_t *band;      // [slot][4]           t_lo, t_hi, u_lo, u_hi (empty: lo > hi)
    // pixel ranges per slot, [slot][bins][2] = (first, last) pixel index, first > last
    // when empty: pixels i of delta1 with a selected pair in data bin row t (it) / an
    // in-range pair in model bin row u (iu); the same for pixels j of delta2 (jt, ju)
    int32_t *rng_it, *rng_iu, *rng_jt, *rng_ju;
    int32_t *cnt_i;     // [pix_off[slot] + i] in-range pairs of pixel i (same-half-plate)
    const int64_t *pix_off;
    uint32_t *touched;  // [(nt*ntm + 31) / 32] (t, u) blocks some slot covers
    unsigned long long *stats;  // [DMAT_NSTATS]
};

struct DmatOut {
    double *weights_dmat;   // [nb]
    double *dmat;           // [nb][nm]
    double *r_par_eff, *r_trans_eff, *z_eff, *weight_eff;  // [nm]
};

hipError_t launch_dmat_pass(const DmatConst &C, const DmatForests &F, const DmatPairs &P,
                            const DmatWork &W, const DmatOut &O, size_t lds,
                            hipStream_t stream);
// quasar: every delta2 of the batch has order < 0 (the forest x quasar form)
hipError_t launch_dmat_fill(const DmatConst &C, const DmatForests &F, const DmatPairs &P,
                            const DmatWork &W, const DmatOut &O, int groups, bool quasar,
                            hipStream_t stream);
size_t dmat_fill_lds(const DmatConst &C);

// ---------------------------------------------------------------------------------------
// metal distortion matrix
// ---------------------------------------------------------------------------------------

constexpr int METAL_BLOCK = 1024;   // threads per workgroup (one per CU, persistent)

enum : int {
    METAL_AUTO = 0,    // |r_par|, bins by floor / truncation, in-range by bin index
    METAL_CROSS = 1,   // signed r_par, otherwise as METAL_AUTO
    METAL_XCF = 2,     // forest x quasar: r_min < r_par < r_max, r_trans < r_t_max
};

enum : int {
    METAL_STAT_DATA = 0,      // pairs added to weights_dmat
    METAL_STAT_MODEL = 1,     // pairs added to dmat
    METAL_STAT_OVERFLOW = 2,  // METAL_XCF pairs whose bin rounds to np (skipped)
    METAL_STAT_VISITED = 3,   // pixel pairs evaluated after pruning
    METAL_NSTATS = 4,
};

// One absorber's view of every pixel of the store: its redshift, comoving distance,
// transverse distance, the pixel weight (0 where the absorber's redshift is not below the
// quasar's) and the redshift-evolution factor.
struct MetalTable {
    const double *__restrict__ z;
    const double *__restrict__ r;
    const double *__restrict__ d;
    const double *__restrict__ w;
    const double *__restrict__ f;
    const struct MetalRecord *__restrict__ rec;   // the same pixels as one record each
};

// One pixel of an absorber table with its forest's Lya coordinates, read as one 64-byte
// scalar load by the pixel loop (the table holds METAL_REC_PAD zero records past the end,
// so that the loop can load the next record unconditionally).
struct MetalRecord {
    double r, d;    // Lya r_comov, dist_m
    double w;       // pixel weight (0 where the absorber's redshift is not below the quasar's)
    double za, ra, da;   // absorber redshift, r_comov, dist_m
    double f;       // redshift-evolution factor
    double pad;
};
static_assert(sizeof(MetalRecord) == 64, "MetalRecord layout");
constexpr int METAL_REC_PAD = 4;

hipError_t launch_metal_pack(const double *r, const double *d, const MetalTable &T,
                             int64_t num_pixels, MetalRecord *rec, hipStream_t stream);

struct MetalConst {
    int32_t np, nt, nb;
    int32_t npm, ntm, nm;
    int32_t mode;       // METAL_*
    int32_t shp;        // remove same-half-plate close pairs
    double rpmin, rpmax, rtmax, span, width_p;
    double inv_den;     // 1 / the evolution factor's denominator
    // bins per Mpc/h (data r_par, r_trans, model r_par, r_trans) for the reciprocal-multiply
    // bin coordinates, and 0.5 - the margin around an integer inside which a coordinate is
    // recomputed with the reference's divide-then-multiply
    double inv_p, inv_t, inv_pm, inv_tm;
    double half_p, half_t, half_pm, half_tm;
};

struct MetalPairs {
    const int32_t *__restrict__ a;
    const int32_t *__restrict__ b;
    const double *__restrict__ cs;          // cos(ang/2), sin(ang/2)
    const uint8_t *__restrict__ flags;      // PAIR_SAME_HALF_PLATE
    const int64_t *__restrict__ task_begin; // [count] first 64-pixel task of each pair
    int64_t count;
    int64_t tasks;                          // task_pair[t]: the pair of task t
};

struct MetalOut {
    double *weights_dmat;   // [nb]
    double *dmat;           // [nb][nm]
    double *r_par_eff, *r_trans_eff, *z_eff, *weight_eff;  // [nm]
    unsigned long long *queue;
    unsigned long long *stats;   // [METAL_NSTATS]
};

// lds_hist: weights_dmat and the four effective-coordinate sums are privatised in LDS
hipError_t launch_metal(const MetalConst &C, const int64_t *off, const uint8_t *mono,
                        const double *r, const double *d, const MetalTable &A,
                        const MetalTable &B, const MetalPairs &P, const int32_t *task_pair,
                        const MetalOut &O, bool lds_hist, int grid, hipStream_t stream);
size_t metal_lds_bytes(const MetalConst &C);

// ---------------------------------------------------------------------------------------
// subsampling covariance
// ---------------------------------------------------------------------------------------

// S[num_bins] and mean[num_bins] are workspace; cov is [num_bins][num_bins]
hipError_t launch_cov_prep(const double *xi, const double *w, int64_t num_samples,
                           int num_bins, double *S, double *mean, hipStream_t stream);
hipError_t launch_cov_syrk(const double *xi, const double *w, const double *S,
                           const double *mean, int64_t num_samples, int num_bins, double *cov,
                           hipStream_t stream);

// ---------------------------------------------------------------------------------------
// Wick covariance terms T1-T3
// ---------------------------------------------------------------------------------------

constexpr int WICK_BLOCK = 256;

enum : int {
    WICK_STAT_IN_RANGE = 0,    // listed pixel pairs
    WICK_STAT_PRODUCTS = 1,    // (a, b) products evaluated
    WICK_STAT_OVERFLOW = 2,    // in-range pairs whose bin rounds up to np (skipped)
    WICK_NSTATS = 3,
};

struct WickConst {
    int32_t np, nt, nb;
    int32_t xcorr;
    double rpmin, rpmax, rtmax, span;
};

struct WickInputs {
    const int64_t *__restrict__ off;     // store offsets
    const double *__restrict__ r;        // r_comov
    const double *__restrict__ w;        // weights
    const double *__restrict__ zev1;     // ((1+z)/(1+z_ref))^(alpha-1)
    const double *__restrict__ zev2;     // ((1+z)/(1+z_ref))^(alpha2-1)
    const double *__restrict__ wx;       // weighted 1-D correlations, n x n per forest
    const int64_t *__restrict__ wx_off;  // [forest] offset of its matrix in wx
    const double *__restrict__ c1;       // row prefix sums of wx, n x (n + 1) per forest
    const int64_t *__restrict__ c1_off;  // [forest] offset of its prefix sums in c1
};

struct WickPairs {
    const int32_t *__restrict__ a;
    const int32_t *__restrict__ b;
    const double *__restrict__ cs;       // cos(ang/2), sin(ang/2)
    int32_t count;
};

struct WickWork {
    int64_t *list;             // per pair, its in-range pixel pairs sorted by bin
    const int64_t *base;       // [pair] offset of its list (capacity n1 n2)
    int32_t *count;            // [pair] listed pixel pairs
    int4 *seg;                 // per pair (same offsets as list), its segments sorted by bin:
                               // (bin, j, i_lo, i_hi), pixels i_lo..i_hi of delta1 that
                               // pixel j of delta2 pairs with in that bin
    int32_t *seg_count;        // [pair] segments
    int32_t *task_pair, *task_bin, *task_start, *task_count;
    unsigned long long *num_tasks;
    int64_t max_tasks;
};

struct WickOut {
    double *weights_wick;                  // [nb]
    unsigned long long *num_pairs_wick;    // [nb]
    double *t1, *t2, *t3;                  // [nb][nb]
    unsigned long long *stats;             // [WICK_NSTATS]
};

hipError_t launch_wick_list(const WickConst &C, const WickInputs &I, const WickPairs &P,
                            const WickWork &W, const WickOut &O, hipStream_t stream);
hipError_t launch_wick_rows(const WickConst &C, const WickInputs &I, const WickPairs &P,
                            const WickWork &W, const WickOut &O, int64_t tasks,
                            hipStream_t stream);
size_t wick_list_lds(const WickConst &C);
size_t wick_row_lds(const WickConst &C);
// row prefix sums c1 of the weighted 1-D correlation of each listed forest
hipError_t launch_wick_prefix(const double *wx, const int64_t *wx_off, const int64_t *off,
                              const int32_t *forests, int32_t num_forests, double *c1,
                              const int64_t *c1_off, hipStream_t stream);

// ---------------------------------------------------------------------------------------
// continuum fit
// ---------------------------------------------------------------------------------------

struct ContConst {
    int32_t max_iter;
    int32_t use_constant_weight;
    double tol;      // stop when g^T H^-1 g / 2 < tol * max(1, |chi2|)
    double tol_abs;  // or when g^T H^-1 g / 2 < tol_abs
};

struct ContIn {
    const int64_t *__restrict__ off;   // [n + 1]
    const double *__restrict__ flux;
    const double *__restrict__ ivar;
    const double *__restrict__ mean_cont;   // mean continuum x transmission correction
    const double *__restrict__ x;           // (log_lambda - Lmin) / (Lmax - Lmin)
    const double *__restrict__ eta;
    const double *__restrict__ var_lss;
    const double *__restrict__ fudge;
    const int32_t *__restrict__ order;      // per forest
    const double *__restrict__ a0;          // per forest start
    const double *__restrict__ b0;
};

struct ContOut {
    double *a, *b, *chi2;
    int32_t *ndata, *status, *evals;
};

hipError_t launch_cont_fit(const ContConst &K, const ContIn &I, const ContOut &O,
                           int64_t n_forests, hipStream_t stream);
// DR16 iteration: weighted F / C stack in rest-frame bins over the forests whose last fit
// is valid (status 0); interpolation of the new mean continuum to every pixel
hipError_t launch_cont_stack(const ContIn &I, const double *a, const double *b,
                             const int32_t *status, const int32_t *bins, int32_t nbins,
                             double *sum_ratio, double *sum_weight, int64_t n_forests,
                             int num_cus, hipStream_t stream);
hipError_t launch_cont_interp(const double *ll_rest, const double *trans, const double *xs,
                              const double *ys, int32_t n, double *mean_cont, int64_t npix,
                              int num_cus, hipStream_t stream);

// ---------------------------------------------------------------------------------------
// seed network of the continuum fit
// ---------------------------------------------------------------------------------------

constexpr int MLP_MAX_IN = 256;        // inputs (multiple of 32)
constexpr int MLP_MAX_HIDDEN = 128;    // hidden units (multiple of 16)
constexpr int MLP_MAX_TILES_PER_WAVE = (MLP_MAX_IN / 16) * (MLP_MAX_HIDDEN / 16) / 4;
constexpr int MLP_BW_ROWS = 256;       // batch rows per block of the dW1 product

struct MlpView {
    int32_t n_in, n_hidden;
    float *w1t;          // [n_hidden][n_in] FP32 master (W1 transposed)
    uint16_t *w1t_bf16;  // [n_hidden][n_in] BF16 copy read by the matrix cores
    float *b1;           // [n_hidden]
    float *w2;           // [n_hidden][2]
    float *b2;           // [2]
};

struct MlpGrads {
    float *w1t, *b1, *w2, *b2;
    float *loss;
};

struct MlpAdam {
    float *m, *v;        // [n_hidden * n_in + 3 * n_hidden + 2], parameter order w1t b1 w2 b2
};

hipError_t launch_mlp_forward(const MlpView &M, const float *x, const int32_t *rows, int64_t n,
                              float *h, float *y, hipStream_t stream);
hipError_t launch_mlp_backward(const MlpView &M, const MlpGrads &G, const float *x,
                               const float *target, const int32_t *rows, int64_t n,
                               const float *h, const float *y, float *dh, hipStream_t stream);
hipError_t launch_mlp_adam(const MlpView &M, const MlpGrads &G, const MlpAdam &A, float lr,
                           float beta1, float beta2, float eps, float corr1, float corr2,
                           hipStream_t stream);
hipError_t launch_mlp_pack(const MlpView &M, hipStream_t stream);

// ---------------------------------------------------------------------------------------
// forest x quasar cross-correlation
// ---------------------------------------------------------------------------------------

// One (forest, quasar) pair: the quasar's r_comov, dist_m, z, weight and the pair's
// cos(ang/2), sin(ang/2).  In the angular mode r is the quasar's wavelength and c the
// angle.
struct XcfPairRecord {
    double r, d, z, w, c, s;
};

struct XcfView {
    const XcfPairRecord *__restrict__ rec;  // [num_pairs], grouped by forest
    const int64_t *__restrict__ idx;        // group-local pair k -> rec[idx[begin + k]]
                                            // (nullptr: rec[begin + k])
    const int32_t *__restrict__ forest;     // [count] forest of each group
    const int64_t *__restrict__ begin;      // [count + 1] pair range of each group
    const int32_t *__restrict__ row;        // [count] output row of each group
    const int64_t *__restrict__ chunk;      // [nchunks + 1] first group of each claim
    int64_t count;
    int64_t nchunks;
};

hipError_t launch_xcf(int mode, bool lds_hist, bool fast, const XiConst &C,
                      const ForestView &F, const XcfView &X, const XiOut &O, int grid,
                      size_t lds, hipStream_t stream);

hipError_t launch_xi(int mode, bool xcorr, bool general, bool lds_hist, const XiConst &C,
                     const ForestView &F, const PairView &P, const GroupView &G,
                     const XiOut &O, int grid, size_t lds, hipStream_t stream);
hipError_t launch_zero_u64(unsigned long long *ptr, int64_t n, hipStream_t stream);
// per pair, the chunks of delta1 whose pixels can pair with delta2 (P.chunks)
hipError_t launch_xi_chunks(int mode, const XiConst &C, const ForestView &F, const PairView &P,
                            int64_t num_pairs, hipStream_t stream);
// device address of the current GPU's scratch array (XI_SCRATCH_WORDS words)
hipError_t xi_scratch_address(unsigned long long **ptr);
hipError_t launch_copy16(const void *src, void *dst, int64_t n16, hipStream_t stream);

}  // namespace picca
