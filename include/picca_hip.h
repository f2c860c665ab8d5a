/*
 * picca_hip.h -- C-ABI of the MI355X (gfx950) engine for picca's 3D Lya correlation path.
 *
 * Plain C types only: POD structs, raw host pointers and sizes.  Every function returns
 * PICCA_OK (0) or a negative PICCA_E_* code; picca_last_error() gives the message of the
 * last failure on the calling thread.  One context owns one device's buffers and one
 * HIP stream; calls on one context must not overlap.
 *
 * Reference interfaces replaced (paths relative to the picca repository):
 *   picca_xi_*         compute_xi            py/picca/cf.py:138-247, batched over healpix
 *                      rows; its per-pair kernel compute_xi_forest_pairs_fast
 *                      py/picca/cf.py:250-387 runs on the GPU for every listed pair
 *   picca_dmat_*       compute_dmat          py/picca/cf.py:390-517 and its per-pair
 *                      kernel compute_dmat_forest_pairs_fast py/picca/cf.py:520-887
 *   picca_xcf_*        xcf.compute_xi        py/picca/xcf.py:131-220 and its per-forest
 *                      kernel py/picca/xcf.py:223-322
 *   picca_cont_*       compute_continuum     py/picca/delta_extraction/expected_fluxes/
 *                      utils.py:12-136 (the Migrad fit of its chi2), batched over forests
 *   picca_metal_*      compute_metal_dmat    py/picca/cf.py:890-1232 and
 *                      py/picca/xcf.py:677-800, one absorber pass over the kept pairs
 *   picca_cov_subsample compute_cov       py/picca/utils.py:100-127
 *   picca_wick_t123    compute_wickT123_pairs py/picca/cf.py:1486-1602 over a pair list
 *   picca_neighbour_candidates  the healpy query_disc superset of fill_neighs
 *                      py/picca/cf.py:82-135 and py/picca/xcf.py:71-128
 *   picca_cf_params    the module globals    py/picca/cf.py:28-79 (None -> has_* = 0)
 */
#ifndef PICCA_HIP_H
#define PICCA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PICCA_HIP_ABI_VERSION 4

#define PICCA_OK 0
#define PICCA_E_VALUE (-1)   /* invalid argument: picca raises ValueError          */
#define PICCA_E_RUNTIME (-2) /* unsupported / inconsistent state: RuntimeError     */
#define PICCA_E_INDEX (-3)   /* negative dmat data bin: IndexError (cf.py:855-856) */
#define PICCA_E_HIP (-4)     /* HIP runtime failure                                */
#define PICCA_E_NOMEM (-5)   /* device or host allocation failed                   */

/* pair flag bits (picca_xi_set_pairs / picca_dmat_run: pair_flags) */
#define PICCA_PAIR_SAME_HALF_PLATE 1u
/* picca_dmat_run: a forest x quasar pair, whose count pass rejects r_par <= r_par_min
   (py/picca/xcf.py:464-466) */
#define PICCA_PAIR_XCF_COUNT 2u

/*
 * Configuration of one correlation run: the cf module globals (py/picca/cf.py:28-79)
 * as set by picca_cf.py (py/picca/bin/picca_cf.py:343-368) and picca_dmat.py
 * (py/picca/bin/picca_dmat.py:353-394).  A global that is None in picca has its has_*
 * flag at 0.  edge_tol is this engine's edge-pair report threshold (Mpc/h).
 */
typedef struct picca_cf_params {
    int32_t num_bins_r_par;
    int32_t num_bins_r_trans;
    int32_t num_model_bins_r_par;
    int32_t num_model_bins_r_trans;
    double r_par_min;
    double r_par_max;
    double r_trans_max;
    double z_min_pairs;
    double z_max_pairs;
    double zerr_cut_deg;
    double zerr_cut_kms;
    double z_ref;
    double alpha;
    double alpha2;
    double edge_tol;
    int32_t has_z_min_pairs;
    int32_t has_z_max_pairs;
    int32_t has_zerr_cut;
    int32_t x_correlation;
    int32_t rmu_binning;
    int32_t ang_correlation;
    int32_t remove_same_half_plate_close_pairs;
    int32_t redshift_evolution;
} picca_cf_params;

/* Counters of one xi execution (no counterpart in picca, which reports none). */
typedef struct picca_xi_report {
    int64_t accepted_pairs;  /* sum of num_pairs over all rows and bins             */
    int64_t edge_pairs;      /* pairs within edge_tol of a bin edge, or outside the range
                                but within edge_tol of it in both coordinates        */
    int64_t overflow_pairs;  /* bin index >= np*nt (picca writes out of bounds)    */
    int64_t visited_pairs;   /* pixel pairs the kernel evaluated after pruning     */
    int64_t wave_steps;      /* 64-lane pair evaluations issued                     */
    int64_t tasks;           /* (pair, 64-pixel chunk of delta1) work items         */
    int64_t empty_tasks;     /* work items whose pruned pixel window was empty      */
    double kernel_ms;        /* device time of the pair kernel (HIP events)         */
    int64_t launches;        /* pair-kernel launches inside kernel_ms               */
    int64_t runs;            /* run walk: runs of one bin summed from prefix sums    */
    int64_t probes;          /* run walk: pixels probed to find the runs' ends       */
    int64_t walk_tasks;      /* work items taken by the run walk                     */
} picca_xi_report;

typedef struct picca_ctx picca_ctx;

/* version / device ---------------------------------------------------------- */
int picca_abi_version(void);
const char *picca_last_error(void);
int picca_device_count(int32_t *count);
/* name (NUL-terminated, truncated to name_len), compute units, LDS bytes per workgroup */
int picca_device_info(int32_t device, char *name, int32_t name_len, int32_t *num_cus,
                      int64_t *lds_per_block, int64_t *global_mem);

/* context ------------------------------------------------------------------- */
int picca_create(int32_t device, picca_ctx **out);
int picca_destroy(picca_ctx *ctx);

/*
 * Forest store: num_forests forests, forest k owns pixels [offsets[k], offsets[k+1]).
 * Arrays are the Delta attributes after read_deltas (py/picca/io.py:492-510):
 * z, r_comov, dist_m, weights, delta (FP64, one value per pixel), z_qso (per forest).
 * log_lambda and order are needed by the distortion matrix only (may be NULL).
 * The store is copied to device memory and stays resident until replaced.
 */
int picca_set_forests(picca_ctx *ctx, int64_t num_forests, const int64_t *offsets,
                      const double *z, const double *r_comov, const double *dist_m,
                      const double *weights, const double *delta, const double *log_lambda,
                      const double *z_qso, const int32_t *order);

/*
 * xi pair list (the (delta1, delta2) loops of compute_xi): pair k correlates forest
 * pair_a[k] (delta1) with pair_b[k] (delta2) at angle pair_ang[k] (radians, as
 * QSO.get_angle_between, py/picca/data.py:106-162) and accumulates into output row
 * pair_row[k] (0 <= row < num_rows; picca_cf uses one row per delta1 healpix).  Pairs
 * of one row must be contiguous.  cos(ang/2) and sin(ang/2) are taken with the host
 * C library, as the reference's compiled loop does.
 */
int picca_xi_set_pairs(picca_ctx *ctx, int64_t num_pairs, const int32_t *pair_a,
                       const int32_t *pair_b, const double *pair_ang,
                       const uint8_t *pair_flags, const int32_t *pair_row, int32_t num_rows);

/*
 * Run the pair kernel over the resident pairs; results stay on the device.  The launch
 * sequence (clearing kernels and the pair kernel) is captured in a HIP graph and replayed
 * while its launch arguments stay the same; picca_xi_use_graph(ctx, 0) launches it
 * directly instead (no counterpart in picca).
 */
int picca_xi_execute(picca_ctx *ctx, const picca_cf_params *params, picca_xi_report *report);
int picca_xi_use_graph(picca_ctx *ctx, int32_t on);

/*
 * picca_xi_execute for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc); no counterpart in picca.  The
 * sums go to the caller's device buffer out_bins: num_rows * np*nt records of 48 bytes
 * (sum w1 w2, sum w1 d1 w2 d2, sum r_par w1 w2, sum r_trans w1 w2, sum z w1 w2 as FP64, the
 * pair count as uint64), out_capacity records long; a smaller buffer is rejected
 * (PICCA_E_VALUE).  The work-queue head and the counters live in a fixed device array of
 * 32 words that the library holds per GPU for the life of the process, whatever the input
 * size; it is cleared at every call, and calls that overlap on one GPU race on it.  The
 * forests and pairs must already be resident (picca_set_forests, picca_xi_set_pairs).
 * Calls no hipMalloc or hipFree; the result is not kept for picca_xi_download.
 */
int picca_xi_execute_noalloc(picca_ctx *ctx, const picca_cf_params *params, void *out_bins,
                             int64_t out_capacity, picca_xi_report *report);

/*
 * Schedule of the pair kernel's specialised path (standard binning, auto-correlation,
 * r_par_min == 0, no cuts).  RUNS walks each pixel of delta1 through runs of pixels of
 * delta2 that share one bin, summed from per-forest prefix sums, wherever delta2's pixels
 * increase strictly and are spaced well above the rounding margins; PAIRS evaluates every
 * pixel pair.  AUTO (the default) walks the tasks whose window of delta2 holds at least
 * four pixels per r_par bin and takes PAIRS for the others, and for the whole execution
 * when the store's median pixel step is longer than a quarter bin.  Results are identical
 * up to FP64 summation order; no counterpart in picca.
 */
#define PICCA_XI_SCHEDULE_AUTO 0
#define PICCA_XI_SCHEDULE_PAIRS 1
#define PICCA_XI_SCHEDULE_RUNS 2
int picca_xi_set_schedule(picca_ctx *ctx, int32_t schedule);

/*
 * Copy the un-normalised sums of the last execution to the host, each [num_rows * np*nt]:
 * sum w1 w2, sum w1 d1 w2 d2, sum r_par w1 w2, sum r_trans w1 w2, sum z w1 w2, pair
 * counts.  Dividing by the weights where they are > 0 gives compute_xi's outputs
 * (py/picca/cf.py:242-246).
 */
int picca_xi_download(picca_ctx *ctx, double *weights, double *xi, double *r_par,
                      double *r_trans, double *z, int64_t *num_pairs);

/*
 * Forest x quasar cross-correlation (compute_xi of py/picca/xcf.py:131-220 with its
 * per-pair kernel :223-322).  Pair k correlates forest pair_forest[k] of the resident
 * store (the deltas) with a quasar at r_comov obj_r[k], dist_m obj_d[k], redshift obj_z[k]
 * and weight obj_w[k], at angle pair_ang[k], into output row pair_row[k]; the pairs of one
 * forest must be contiguous and belong to one row, rows contiguous.  In the angular
 * mode obj_r is the quasar's wavelength (obj_d unused).  Results are read with
 * picca_xi_download.  cos/sin(ang/2) are taken with the host C library.
 */
int picca_xcf_set_pairs(picca_ctx *ctx, int64_t num_pairs, const int32_t *pair_forest,
                        const double *obj_r, const double *obj_d, const double *obj_z,
                        const double *obj_w, const double *pair_ang, const int32_t *pair_row,
                        int32_t num_rows, int32_t angular);
int picca_xcf_execute(picca_ctx *ctx, const picca_cf_params *params, picca_xi_report *report);

/*
 * Distortion matrix (compute_dmat, py/picca/cf.py:390-517) over an explicit list of
 * kept forest pairs (the rejection draws of py/picca/cf.py:444 are made by the caller,
 * in reference order).  Needs a forest store with log_lambda and order.
 * Outputs are accumulated into the caller's arrays:
 *   weights_dmat[np*nt], dmat[np*nt * nm] (nm = model np*nt), r_par_eff[nm],
 *   r_trans_eff[nm], z_eff[nm], weight_eff[nm].
 * A forest with order < 0 in the store is not projected: the cross-correlation
 * distortion matrix (compute_dmat of py/picca/xcf.py:324-658) stores each kept quasar
 * as a one-pixel forest (r_comov, dist_m, z_qso, weight) with order -1 and lists the
 * (delta, quasar) pairs with PICCA_PAIR_XCF_COUNT and x_correlation = 1; alpha2 is then
 * the quasars' redshift-evolution exponent.
 * stats (may be NULL) receives: [0] in-range pairs, [1] selected pairs,
 * [2] in-range pairs beyond the count-pass size (see py/picca/cf.py:619,702-703),
 * [3] selected pairs whose data bin reaches np*nt by rounding (the reference writes those
 * past the end of weights_dmat; here they are counted and dropped).  stats has 4 entries.
 */
/*
 * Device timing and work of the last picca_dmat_run on ctx (no counterpart in picca):
 * pass_ms / fill_ms the summed HIP-event times of the pair-pass and fill kernels over the
 * batches, mfma_flops the FP64 operations the fill issued on the matrix cores (2048 per
 * 16x16x4 instruction).
 */
typedef struct picca_dmat_report {
    double pass_ms;
    double fill_ms;
    int64_t batches;
    int64_t mfma_flops;
} picca_dmat_report;
int picca_dmat_last_report(picca_ctx *ctx, picca_dmat_report *report);

int picca_dmat_run(picca_ctx *ctx, const picca_cf_params *params, int64_t num_pairs,
                   const int32_t *pair_a, const int32_t *pair_b, const double *pair_ang,
                   const uint8_t *pair_flags, double *weights_dmat, double *dmat,
                   double *r_par_eff, double *r_trans_eff, double *z_eff,
                   double *weight_eff, int64_t *stats);

/*
 * Quasar-continuum fit (compute_continuum,
 * py/picca/delta_extraction/expected_fluxes/utils.py:12-136, with the objective of
 * LeastsSquaresContModel, py/picca/delta_extraction/least_squares/
 * least_squares_cont_model.py:65-197), batched over forests.
 *
 * picca_cont_set_forests makes the fit inputs resident: forest k owns pixels
 * [offsets[k], offsets[k+1]); per pixel the flux, ivar, the mean continuum times the
 * transmission correction (mean_cont), x = (log_lambda - Lmin) / (Lmax - Lmin) with
 * Lmin/Lmax the rest-frame grid ends shifted to the forest's redshift (:59-78), and the
 * variance terms eta, var_lss, fudge evaluated at the pixel (:80-86); per forest the
 * polynomial order (0 fixes the slope at 0).
 *
 * picca_cont_fit minimises chi2(a, b) for every forest from the start (a0, b0) given per
 * forest (compute_continuum starts at a0 = sum f ivar / sum ivar, b0 = 0) and returns a
 * (zero point), b (slope), the minimum chi2, ndata and a status: 0 fitted, 1 not
 * converged within max_iter, 2 negative continuum, 3 zero continuum, 4 non-finite chi2
 * (the reference's bad_continuum_reason cases, :111-120), and the number of objective
 * evaluations.
 */
typedef struct picca_cont_params {
    int32_t max_iter;             /* Levenberg-Marquardt iterations (e.g. 100)          */
    int32_t use_constant_weight;  /* 1: all weights 1 (chi2 = sum of squares)           */
    double tol;                   /* stop when g^T H^-1 g / 2 < tol * max(1, |chi2|)    */
    double tol_abs;               /* or when g^T H^-1 g / 2 < tol_abs (0: off); Migrad's
                                     own rule is 2e-4 at its default tolerance          */
} picca_cont_params;

int picca_cont_set_forests(picca_ctx *ctx, int64_t num_forests, const int64_t *offsets,
                           const double *flux, const double *ivar, const double *mean_cont,
                           const double *x, const double *eta, const double *var_lss,
                           const double *fudge, const int32_t *order);
int picca_cont_fit(picca_ctx *ctx, const picca_cont_params *params, const double *a0,
                   const double *b0, double *a, double *b, double *chi2, int32_t *ndata,
                   int32_t *status, int32_t *evaluations, double *kernel_ms);

/*
 * The mean-continuum update of the DR16 iteration (Dr16ExpectedFlux.compute_expected_flux,
 * py/picca/delta_extraction/expected_fluxes/dr16_expected_flux.py:380-464).
 *
 * picca_cont_set_rest_frame sets, for the resident forests, each pixel's rest-frame
 * log lambda, its transmission correction and its bin of the rest-frame grid (num_bins
 * bins, find_bins_log / find_bins_lin of py/picca/delta_extraction/utils.py:139-186).
 *
 * picca_cont_stack stacks the last fit: over the pixels of the forests with status 0,
 * sum_ratio[bin] = sum F / (C + 1e-16) w and sum_weight[bin] = sum w, with C the fitted
 * continuum and w the forest weight 1 / (eta var_pipe + var_lss + fudge / var_pipe),
 * var_pipe = 1 / ivar / C^2, 0 where ivar <= 0 (ExpectedFlux._compute_mean_cont,
 * py/picca/delta_extraction/expected_flux.py:256-308, with compute_forest_weights,
 * dr16_expected_flux.py:466-488).  The caller normalises the stack and multiplies it into
 * the previous mean continuum (a few hundred bins).
 *
 * picca_cont_set_mean_cont sets every pixel's mean continuum to the linear interpolation
 * of the nodes (x, y) (n >= 2, x increasing; the end intervals extend beyond them, as
 * scipy's interp1d with fill_value="extrapolate") at its rest-frame log lambda, times its
 * transmission correction; the next picca_cont_fit uses it.
 */
#define PICCA_CONT_MAX_BINS 8192
int picca_cont_set_rest_frame(picca_ctx *ctx, const double *log_lambda_rest,
                              const double *transmission, const int32_t *bins,
                              int32_t num_bins);
int picca_cont_stack(picca_ctx *ctx, double *sum_ratio, double *sum_weight,
                     double *kernel_ms);
int picca_cont_set_mean_cont(picca_ctx *ctx, int32_t n, const double *x, const double *y,
                             double *kernel_ms);

/*
 * Seed network of the continuum fit (no counterpart in picca: it predicts the start
 * (a, b) of the fit above from the binned spectrum).  A two-layer perceptron
 * y = relu(x W1 + b1) W2 + b2 with n_in inputs (multiple of 32, <= 256), n_hidden units
 * (multiple of 16, <= 128) and 2 outputs; W1 is [n_in][n_hidden] row-major, W2
 * [n_hidden][2].  The n_in x n_hidden products run in BF16 on the matrix cores with FP32
 * accumulation; parameters, gradients and Adam state are FP32.
 *
 * picca_mlp_init sets the parameters (and resets Adam); picca_mlp_get reads them back;
 * picca_mlp_forward evaluates n rows of x [n][n_in] into y [n][2] (kernel_ms, may be
 * NULL, receives the device time of the forward kernel); picca_mlp_gradients
 * returns the gradients of loss = sum (y - t)^2 / (2 n) over one batch without updating;
 * picca_mlp_set_data makes a training set resident and picca_mlp_train runs `steps`
 * Adam steps (beta1 0.9, beta2 0.999, eps 1e-8) on minibatches of `batch` rows drawn
 * without replacement per epoch from a seeded permutation, writing each step's loss.
 */
int picca_mlp_init(picca_ctx *ctx, int32_t n_in, int32_t n_hidden, const float *w1,
                   const float *b1, const float *w2, const float *b2);
int picca_mlp_get(picca_ctx *ctx, float *w1, float *b1, float *w2, float *b2);
int picca_mlp_forward(picca_ctx *ctx, int64_t n, const float *x, float *y, double *kernel_ms);
int picca_mlp_gradients(picca_ctx *ctx, int64_t n, const float *x, const float *t,
                        float *g_w1, float *g_b1, float *g_w2, float *g_b2, float *loss);
int picca_mlp_set_data(picca_ctx *ctx, int64_t n, const float *x, const float *t);
int picca_mlp_train(picca_ctx *ctx, int64_t steps, int32_t batch, float lr, uint64_t seed,
                    float *losses);

/*
 * Diagnostics (no counterpart in picca): runs a plain device-to-device copy kernel over
 * `bytes` bytes (a multiple of 16) between two buffers of the context and returns its
 * device time.  Used to calibrate hardware byte counters (FETCH_SIZE / WRITE_SIZE) against
 * a kernel of known traffic.
 */
int picca_diag_copy(picca_ctx *ctx, int64_t bytes, double *kernel_ms);

/*
 * Neighbour candidates for fill_neighs (py/picca/cf.py:82-135), host code.  Replaces the
 * healpy.query_disc(inclusive=True) superset (py/picca/cf.py:93-98): returns, per source,
 * the targets (ascending index) with x2*x1 + y2*y1 + z2*z1 >= cos_min or inside the
 * 2-arcsec box of the small-angle formula, excluding equal ids and, in auto mode, targets
 * with ra >= the source's ra.  The caller applies the exact angle test.  Two passes:
 * offsets == NULL fills counts[n_src]; otherwise the candidates are written at
 * [offsets[s], offsets[s] + counts[s]) of out_target / out_cos.
 */
int picca_neighbour_candidates(int64_t n_src, const double *sx, const double *sy,
                               const double *sz, const double *sra, const double *sdec,
                               const int64_t *sid, int64_t n_tgt, const double *tx,
                               const double *ty, const double *tz, const double *tra,
                               const double *tdec, const int64_t *tid, double cos_min,
                               double chord_max, double small_angle, int32_t auto_mode,
                               int64_t *counts, const int64_t *offsets, int64_t *out_target,
                               double *out_cos);

/*
 * Metal distortion matrices (compute_metal_dmat of py/picca/cf.py:890-1232 and of
 * py/picca/xcf.py:677-800).
 *
 * picca_metal_set_absorber stores one absorber's view of every pixel of the forest store
 * (store order, the store's pixel count): its redshift z = 10^log_lambda / lambda_abs - 1,
 * r_comov(z), dist_m(z), the pixel weight with 0 where z >= the forest's z_qso (the
 * reference drops those pixels, py/picca/cf.py:1010-1019, :1038-1047, xcf.py:722-731), and
 * the redshift-evolution factor of the pixel: (1+z)^(alpha_abs-1) for the auto/cross forms
 * (py/picca/cf.py:1062-1066), ((1+z)/(1+z_ref))^(alpha_abs-1) for the forest x quasar form
 * (py/picca/xcf.py:747-749).  A quasar stored as a one-pixel forest has its own z_qso,
 * r_comov, dist_m, weight and factor 1.  slot is 0..3.
 */
int picca_metal_set_absorber(picca_ctx *ctx, int32_t slot, const double *z,
                             const double *r_comov, const double *dist_m,
                             const double *weights, const double *zfac);

#define PICCA_METAL_AUTO 0  /* |r_par|; in range when the floor / truncated bins are   */
#define PICCA_METAL_CROSS 1 /* signed r_par, bins as PICCA_METAL_AUTO                   */
#define PICCA_METAL_XCF 2   /* forest x quasar: r_min < r_par < r_max, r_t < r_t_max   */

/* Counters of one metal pass (no counterpart in picca). */
typedef struct picca_metal_report {
    int64_t data_pairs;      /* pairs added to weights_dmat                            */
    int64_t model_pairs;     /* pairs added to dmat                                    */
    int64_t overflow_pairs;  /* PICCA_METAL_XCF bins rounded up to np (skipped; picca's
                                bincount would fail there)                             */
    int64_t visited_pairs;   /* pixel pairs evaluated after pruning                    */
    double kernel_ms;        /* device time of the pass kernel (HIP events)           */
} picca_metal_report;

/*
 * One pass of compute_metal_dmat over an explicit list of kept pairs (the rejection
 * draws are made by the caller in reference order): delta1's pixels as the absorber of
 * slot_a, delta2's as slot_b, data bins from the store's Lya r_comov / dist_m, model bins
 * from the absorbers'.  The pair's evolution factor is zfac_a * zfac_b / zw_den
 * ((1+z_ref)^(alpha_a+alpha_b-2) for the auto/cross forms, 1 for the forest x quasar
 * form).  Uses num_bins_* / num_model_bins_*, r_par_min/max, r_trans_max and
 * remove_same_half_plate_close_pairs of params (with PICCA_PAIR_SAME_HALF_PLATE flags).
 * Outputs are accumulated into the caller's arrays as in picca_dmat_run (the swapped
 * second pass of py/picca/cf.py:1092-1200 is a second call with the slots exchanged).
 * report may be NULL.
 */
int picca_metal_pass(picca_ctx *ctx, const picca_cf_params *params, int32_t mode,
                     int32_t slot_a, int32_t slot_b, double zw_den, int64_t num_pairs,
                     const int32_t *pair_a, const int32_t *pair_b, const double *pair_ang,
                     const uint8_t *pair_flags, double *weights_dmat, double *dmat,
                     double *r_par_eff, double *r_trans_eff, double *z_eff,
                     double *weight_eff, picca_metal_report *report);

/*
 * Covariance of the per-healpix correlation rows by subsampling (compute_cov of
 * py/picca/utils.py:100-127, called by picca_export, py/picca/bin/picca_export.py:276):
 * xi and weights are [num_samples][num_bins] row-major (the COR rows' DA and WE),
 * cov receives [num_bins][num_bins].  kernel_ms (may be NULL) receives two device times
 * in ms: [0] the whole computation, [1] the matrix-core product Y^T Y alone.
 */
int picca_cov_subsample(picca_ctx *ctx, int64_t num_samples, int64_t num_bins,
                        const double *xi, const double *weights, double *cov,
                        double *kernel_ms);

/* Counters of one Wick call (no counterpart in picca). */
typedef struct picca_wick_report {
    int64_t in_range_pairs;  /* listed pixel pairs (sum of num_pairs_wick)           */
    int64_t products;        /* unordered pixel-pair products {a, b}, a != b: the
                                iterations of the reference's index1 < index2 loop     */
    int64_t overflow_pairs;  /* in-range pairs whose bin rounds up to np (skipped;
                                numba writes them out of bounds)                       */
    int64_t batches;         /* batches the pair list was split into                   */
    double kernel_ms;        /* device time of the kernels (HIP events)                */
} picca_wick_report;

/*
 * Wick covariance terms T1, T2, T3 (compute_wickT123_pairs, py/picca/cf.py:1486-1602) over
 * an explicit list of forest pairs of the resident store (compute_wick_terms,
 * py/picca/cf.py:1328-1424, draws one uniform per delta of a healpix and takes every
 * neighbour of a kept delta: the caller makes those draws).  Uses the store's r_comov and
 * weights; zev1 / zev2 [store pixels] are ((1+z)/(1+z_ref))^(alpha-1) and ^(alpha2-1);
 * wx + wx_offsets[f] is forest f's n x n weighted 1-D correlation
 * w_k w_l xi_1d(|log_lambda_k - log_lambda_l|) sqrt(v_k v_l) (py/picca/cf.py:1392-1398),
 * symmetric (its rows are read as its columns), needed for every forest of the pair list.
 * Device memory per call: the row prefix sums of those matrices (n (n+1) doubles each) and
 * per batch up to 24 bytes per (pixel, pixel) of its pairs.  Outputs (accumulated):
 * weights_wick[nb], num_pairs_wick[nb], t1, t2, t3 [nb][nb].  report may be NULL.
 */
int picca_wick_t123(picca_ctx *ctx, const picca_cf_params *params, int64_t num_pairs,
                    const int32_t *pair_a, const int32_t *pair_b, const double *pair_ang,
                    const double *zev1, const double *zev2, const int64_t *wx_offsets,
                    int64_t wx_size, const double *wx, double *weights_wick,
                    int64_t *num_pairs_wick, double *t1, double *t2, double *t3,
                    picca_wick_report *report);

#ifdef __cplusplus
}
#endif

#endif /* PICCA_HIP_H */
