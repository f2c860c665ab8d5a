/*
 * CPU oracle for the picca correlation hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is a plain-C, FP64, scalar restatement of the two numba kernels of
 * picca's cf module.  It is the checker the HIP engine is compared against and the
 * CPU baseline timed beside it; nothing in the product links or calls it.
 *
 *   oracle_xi_pair    <- compute_xi_forest_pairs_fast, py/picca/cf.py:250-387
 *   oracle_dmat_pair  <- compute_dmat_forest_pairs_fast, py/picca/cf.py:520-887
 *   oracle_xi_rows    <- the per-healpix loop of compute_xi, py/picca/cf.py:162-231,
 *                        run for many healpix rows at once (one OpenMP task per row,
 *                        as picca_cf runs one pool task per healpix,
 *                        py/picca/bin/picca_cf.py:452-458)
 *   oracle_dmat_rows  <- the loop of compute_dmat over kept pairs,
 *                        py/picca/cf.py:425-502
 *   oracle_xcf_pair   <- compute_xi_forest_pairs_fast of xcf, py/picca/xcf.py:223-322
 *   oracle_xdmat_forest <- compute_dmat_forest_pairs_fast of xcf,
 *                        py/picca/xcf.py:427-658 (one delta with its kept quasars)
 *   oracle_metal_pair <- one pass of compute_metal_dmat for one pair,
 *                        py/picca/cf.py:890-1232 and py/picca/xcf.py:677-800
 *   oracle_wick_t123  <- compute_wickT123_pairs over a list of pairs,
 *                        py/picca/cf.py:1486-1602
 *
 * Loop order, branch order and every floating-point operation follow the reference
 * (divide-then-multiply bin assignment, cos/sin of ang/2 from libm, sequential sums
 * as numba's ndarray.sum).  Build with -O2 -ffp-contract=off and without -ffast-math.
 *
 * Additions that the reference does not have, reported through `stats`:
 *   stats[0] accepted pairs (== sum of num_pairs increments)
 *   stats[1] edge pairs: pairs inside the (r_par, r_trans) range whose non-zero r_par
 *            or r_trans lies within params->edge_tol of a bin edge, plus pairs outside
 *            the range whose r_par and r_trans both lie within edge_tol of it (a pair
 *            far outside in one coordinate cannot enter the range whatever the other)
 *   stats[2] overflow pairs: bin index >= np*nt (numba writes out of bounds there;
 *            the oracle skips and counts them)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define SPEED_LIGHT_KMS (299792458.0 / 1000.0)

typedef struct {
    int32_t num_bins_r_par;
    int32_t num_bins_r_trans;
    int32_t num_model_bins_r_par;
    int32_t num_model_bins_r_trans;
    double r_par_min;
    double r_par_max;
    double r_trans_max;
    double z_min_pairs;   /* used when has_z_min_pairs */
    double z_max_pairs;   /* used when has_z_max_pairs */
    double zerr_cut_deg;  /* used when has_zerr_cut */
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
} oracle_params;

/* distance of t to the nearest integer, measured in units of `width` */
static double edge_distance(double t, double width) {
    double frac = t - floor(t);
    double lo = frac * width;
    double hi = (1.0 - frac) * width;
    return lo < hi ? lo : hi;
}

void oracle_xi_pair(const oracle_params *p,
                    const double *z1, const double *r_comov1, const double *dist_m1,
                    const double *weights1, const double *delta1, int64_t n1, double z_qso_1,
                    const double *z2, const double *r_comov2, const double *dist_m2,
                    const double *weights2, const double *delta2, int64_t n2, double z_qso_2,
                    double ang, int32_t same_half_plate,
                    double *rebin_weight, double *rebin_xi, double *rebin_r_par,
                    double *rebin_r_trans, double *rebin_z, int64_t *rebin_num_pairs,
                    int64_t *stats) {
    const int64_t nbins = (int64_t)p->num_bins_r_par * p->num_bins_r_trans;
    const int zerr_on = p->has_zerr_cut && (ang < p->zerr_cut_deg * M_PI / 180.0);
    const double width_p = (p->r_par_max - p->r_par_min) / p->num_bins_r_par;
    const double width_t = p->r_trans_max / p->num_bins_r_trans;
    for (int64_t i = 0; i < n1; ++i) {
        if (weights1[i] == 0) continue;
        if (zerr_on) {
            double z_qf = 0.5 * (z1[i] + z_qso_2);
            double dv_kms = fabs(z1[i] - z_qso_2) / (1 + z_qf);
            dv_kms *= SPEED_LIGHT_KMS;
            if (dv_kms < p->zerr_cut_kms) continue;
        }
        for (int64_t j = 0; j < n2; ++j) {
            if (weights2[j] == 0) continue;
            double z = (z1[i] + z2[j]) / 2;
            if ((p->has_z_min_pairs && z < p->z_min_pairs) ||
                (p->has_z_max_pairs && z > p->z_max_pairs))
                continue;
            if (zerr_on) {
                double z_qf = 0.5 * (z2[j] + z_qso_1);
                double dv_kms = fabs(z2[j] - z_qso_1) / (1 + z_qf);
                dv_kms *= SPEED_LIGHT_KMS;
                if (dv_kms < p->zerr_cut_kms) continue;
            }
            double r_par, r_trans;
            if (p->ang_correlation) {
                r_par = r_comov1[i] / r_comov2[j];
                if (!p->x_correlation && r_par < 1.0) r_par = 1.0 / r_par;
                r_trans = ang;
            } else {
                r_par = (r_comov1[i] - r_comov2[j]) * cos(ang / 2);
                r_trans = (dist_m1[i] + dist_m2[j]) * sin(ang / 2);
                if (p->rmu_binning) {
                    r_trans = sqrt(r_trans * r_trans + r_par * r_par);
                    r_par /= r_trans;
                }
                if (!p->x_correlation) r_par = fabs(r_par);
            }
            if (r_par >= p->r_par_max || r_trans >= p->r_trans_max || r_par < p->r_par_min) {
                if (r_par - p->r_par_max < p->edge_tol && p->r_par_min - r_par < p->edge_tol &&
                    r_trans - p->r_trans_max < p->edge_tol)
                    stats[1] += 1;
                continue;
            }
            double delta_times_weight1 = delta1[i] * weights1[i];
            double delta_times_weight2 = delta2[j] * weights2[j];
            double delta_times_weight12 = delta_times_weight1 * delta_times_weight2;
            double weights12 = weights1[i] * weights2[j];
            double t_par = (r_par - p->r_par_min) / (p->r_par_max - p->r_par_min) * p->num_bins_r_par;
            double t_trans = r_trans / p->r_trans_max * p->num_bins_r_trans;
            double bins_r_par = floor(t_par);
            double bins_r_trans = floor(t_trans);
            int64_t bins = (int64_t)(bins_r_trans + p->num_bins_r_trans * bins_r_par);
            /* an exact zero (equal r_comov, or ang = 0) cannot move under rounding */
            if ((r_par != 0 && edge_distance(t_par, width_p) < p->edge_tol) ||
                (r_trans != 0 && edge_distance(t_trans, width_t) < p->edge_tol))
                stats[1] += 1;
            if (p->remove_same_half_plate_close_pairs && same_half_plate) {
                if (fabs(r_par) < (p->r_par_max - p->r_par_min) / p->num_bins_r_par) continue;
            }
            if (bins < 0 || bins >= nbins) {
                stats[2] += 1;
                continue;
            }
            rebin_xi[bins] += delta_times_weight12;
            rebin_weight[bins] += weights12;
            rebin_r_par[bins] += r_par * weights12;
            rebin_r_trans[bins] += r_trans * weights12;
            rebin_z[bins] += z * weights12;
            rebin_num_pairs[bins] += 1;
            stats[0] += 1;
        }
    }
}

/*
 * compute_xi_forest_pairs_fast of the forest x quasar cross-correlation
 * (py/picca/xcf.py:227-322): forest pixels i against quasars j, each quasar with its own
 * angle.  r_par keeps its sign; the lower range limit is exclusive (r_par <= r_par_min is
 * rejected).  Bin indices >= np*nt (rounding at the range limit) are counted in stats[2]
 * and left out; stats[0] counts accepted pairs, stats[1] edge pairs.
 */
void oracle_xcf_pair(const oracle_params *p,
                     const double *z1, const double *r_comov1, const double *dist_m1,
                     const double *weights1, const double *delta1, int64_t n1,
                     const double *z2, const double *r_comov2, const double *dist_m2,
                     const double *weights2, const double *ang, int64_t n2,
                     double *rebin_weight, double *rebin_xi, double *rebin_r_par,
                     double *rebin_r_trans, double *rebin_z, int64_t *rebin_num_pairs,
                     int64_t *stats) {
    const int64_t nbins = (int64_t)p->num_bins_r_par * p->num_bins_r_trans;
    const double width_p = (p->r_par_max - p->r_par_min) / p->num_bins_r_par;
    const double width_t = p->r_trans_max / p->num_bins_r_trans;
    for (int64_t i = 0; i < n1; ++i) {
        if (weights1[i] == 0) continue;
        for (int64_t j = 0; j < n2; ++j) {
            if (weights2[j] == 0) continue;
            double z = (z1[i] + z2[j]) / 2;
            if ((p->has_z_min_pairs && z < p->z_min_pairs) ||
                (p->has_z_max_pairs && z > p->z_max_pairs))
                continue;
            double r_par, r_trans;
            if (p->ang_correlation) {
                r_par = r_comov1[i] / r_comov2[j];
                r_trans = ang[j];
            } else {
                r_par = (r_comov1[i] - r_comov2[j]) * cos(ang[j] / 2);
                r_trans = (dist_m1[i] + dist_m2[j]) * sin(ang[j] / 2);
            }
            if (p->rmu_binning) {
                r_trans = sqrt(r_trans * r_trans + r_par * r_par);
                r_par /= r_trans;
            }
            if (r_par >= p->r_par_max || r_trans >= p->r_trans_max || r_par <= p->r_par_min) {
                if (r_par - p->r_par_max < p->edge_tol && p->r_par_min - r_par < p->edge_tol &&
                    r_trans - p->r_trans_max < p->edge_tol)
                    stats[1] += 1;
                continue;
            }
            double delta_times_weight = delta1[i] * weights1[i] * weights2[j];
            double weights12 = weights1[i] * weights2[j];
            double t_par = (r_par - p->r_par_min) / (p->r_par_max - p->r_par_min) * p->num_bins_r_par;
            double t_trans = r_trans / p->r_trans_max * p->num_bins_r_trans;
            double bins_r_par = floor(t_par);
            double bins_r_trans = floor(t_trans);
            int64_t bins = (int64_t)(bins_r_trans + p->num_bins_r_trans * bins_r_par);
            if ((r_par != p->r_par_min && edge_distance(t_par, width_p) < p->edge_tol) ||
                (r_trans != 0 && edge_distance(t_trans, width_t) < p->edge_tol))
                stats[1] += 1;
            if (bins < 0 || bins >= nbins) {
                stats[2] += 1;
                continue;
            }
            rebin_xi[bins] += delta_times_weight;
            rebin_weight[bins] += weights12;
            rebin_r_par[bins] += r_par * weights12;
            rebin_r_trans[bins] += r_trans * weights12;
            rebin_z[bins] += z * weights12;
            rebin_num_pairs[bins] += 1;
            stats[0] += 1;
        }
    }
}

/*
 * Many healpix rows at once.  Pairs are grouped by row (pair_row non-decreasing is NOT
 * required: the pairs of each row are visited in their input order).  Forest k owns
 * pixels [offsets[k], offsets[k+1]) of the pixel arrays.  Outputs are [num_rows * nb].
 */
void oracle_xi_rows(const oracle_params *p, const int64_t *offsets,
                    const double *z, const double *r_comov, const double *dist_m,
                    const double *weights, const double *delta, const double *z_qso,
                    int64_t num_pairs, const int32_t *pair_a, const int32_t *pair_b,
                    const double *pair_ang, const uint8_t *pair_shp,
                    const int64_t *row_start, int32_t num_rows,
                    double *out_we, double *out_xi, double *out_rp, double *out_rt,
                    double *out_z, int64_t *out_nb, int64_t *stats, int32_t num_threads) {
    const int64_t nb = (int64_t)p->num_bins_r_par * p->num_bins_r_trans;
    (void)num_pairs;
    int64_t s0 = 0, s1 = 0, s2 = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(num_threads) reduction(+ : s0, s1, s2)
    for (int32_t row = 0; row < num_rows; ++row) {
        int64_t local[3] = {0, 0, 0};
        for (int64_t k = row_start[row]; k < row_start[row + 1]; ++k) {
            int32_t a = pair_a[k], b = pair_b[k];
            oracle_xi_pair(p, z + offsets[a], r_comov + offsets[a], dist_m + offsets[a],
                           weights + offsets[a], delta + offsets[a], offsets[a + 1] - offsets[a],
                           z_qso[a], z + offsets[b], r_comov + offsets[b], dist_m + offsets[b],
                           weights + offsets[b], delta + offsets[b], offsets[b + 1] - offsets[b],
                           z_qso[b], pair_ang[k], pair_shp[k], out_we + row * nb,
                           out_xi + row * nb, out_rp + row * nb, out_rt + row * nb,
                           out_z + row * nb, out_nb + row * nb, local);
        }
        s0 += local[0];
        s1 += local[1];
        s2 += local[2];
    }
    stats[0] += s0;
    stats[1] += s1;
    stats[2] += s2;
}

/* ------------------------------------------------------------------------------ */
/* distortion matrix                                                               */
/* ------------------------------------------------------------------------------ */

static int cmp_i32(const void *x, const void *y) {
    int32_t a = *(const int32_t *)x, b = *(const int32_t *)y;
    return (a > b) - (a < b);
}

/*
 * Returns 0, or -1 when a selected pair has a negative data bin (the reference raises
 * IndexError there, py/picca/cf.py:855-856), or -2 on allocation failure.
 * stats[0] in-range pairs of the main pass, stats[1] selected pairs,
 * stats[3] selected pairs whose data bin is >= np*nt (left out),
 * stats[2] in-range pairs beyond the count-pass size (the reference writes those
 * model bins past the end of all_model_bins; np.unique then sees only the first
 * count-pass-many, which is what this function reproduces).
 */
int oracle_dmat_pair(const oracle_params *p,
                     const double *log_lambda1, const double *log_lambda2,
                     const double *r_comov1, const double *r_comov2,
                     const double *dist_m1, const double *dist_m2,
                     const double *z1, const double *z2,
                     const double *weights1, const double *weights2,
                     int64_t n1, int64_t n2, double z_qso_1, double z_qso_2, double ang,
                     double *weights_dmat, double *dmat, double *r_par_eff,
                     double *r_trans_eff, double *z_eff, double *weight_eff,
                     int32_t same_half_plate, int32_t order1, int32_t order2,
                     int64_t *stats) {
    const double cos_half = cos(ang / 2);
    const double sin_half = sin(ang / 2);
    const int64_t n_model = (int64_t)p->num_model_bins_r_par * p->num_model_bins_r_trans;

    /* (i) count pass */
    int64_t num_pairs = 0;
    for (int64_t i = 0; i < n1; ++i) {
        if (weights1[i] == 0) continue;
        for (int64_t j = 0; j < n2; ++j) {
            if (weights2[j] == 0) continue;
            double r_par = (r_comov1[i] - r_comov2[j]) * cos_half;
            double r_trans = (dist_m1[i] + dist_m2[j]) * sin_half;
            if (p->rmu_binning) {
                r_trans = sqrt(r_trans * r_trans + r_par * r_par);
                r_par /= r_trans;
            }
            if (!p->x_correlation) r_par = fabs(r_par);
            if (r_par >= p->r_par_max || r_trans >= p->r_trans_max || r_par < p->r_par_min) continue;
            if (p->remove_same_half_plate_close_pairs && same_half_plate) {
                if (fabs(r_par) < (p->r_par_max - p->r_par_min) / p->num_bins_r_par) continue;
            }
            num_pairs += 1;
        }
    }
    if (num_pairs == 0) return 0;

    /* (ii) per-forest sums, sequential like numba's ndarray.sum */
    double sum_weights1 = 0, sum_weights2 = 0;
    for (int64_t i = 0; i < n1; ++i) sum_weights1 += weights1[i];
    for (int64_t j = 0; j < n2; ++j) sum_weights2 += weights2[j];
    double acc1 = 0, acc2 = 0;
    for (int64_t i = 0; i < n1; ++i) acc1 += log_lambda1[i] * weights1[i];
    for (int64_t j = 0; j < n2; ++j) acc2 += log_lambda2[j] * weights2[j];
    double mean_log_lambda1 = acc1 / sum_weights1;
    double mean_log_lambda2 = acc2 / sum_weights2;
    double *dll1 = (double *)malloc(sizeof(double) * (size_t)(n1 > 0 ? n1 : 1));
    double *dll2 = (double *)malloc(sizeof(double) * (size_t)(n2 > 0 ? n2 : 1));
    double *eta1 = (double *)calloc((size_t)(n_model * n1), sizeof(double));
    double *eta2 = (double *)calloc((size_t)(n_model * n2), sizeof(double));
    double *eta3 = (double *)calloc((size_t)(n_model * n1), sizeof(double));
    double *eta4 = (double *)calloc((size_t)(n_model * n2), sizeof(double));
    double *eta5 = (double *)calloc((size_t)n_model, sizeof(double));
    double *eta6 = (double *)calloc((size_t)n_model, sizeof(double));
    double *eta7 = (double *)calloc((size_t)n_model, sizeof(double));
    double *eta8 = (double *)calloc((size_t)n_model, sizeof(double));
    int32_t *sel_data = (int32_t *)malloc(sizeof(int32_t) * (size_t)num_pairs);
    int32_t *sel_model = (int32_t *)calloc((size_t)num_pairs, sizeof(int32_t));
    int32_t *sel_i = (int32_t *)calloc((size_t)num_pairs, sizeof(int32_t));
    int32_t *sel_j = (int32_t *)calloc((size_t)num_pairs, sizeof(int32_t));
    int32_t *all_model = (int32_t *)calloc((size_t)num_pairs, sizeof(int32_t));
    int status = 0;
    if (!dll1 || !dll2 || !eta1 || !eta2 || !eta3 || !eta4 || !eta5 || !eta6 || !eta7 ||
        !eta8 || !sel_data || !sel_model || !sel_i || !sel_j || !all_model) {
        status = -2;
        goto done;
    }
    for (int64_t k = 0; k < num_pairs; ++k) sel_data[k] = -1;
    for (int64_t i = 0; i < n1; ++i) dll1[i] = log_lambda1[i] - mean_log_lambda1;
    for (int64_t j = 0; j < n2; ++j) dll2[j] = log_lambda2[j] - mean_log_lambda2;
    double sw_dll1 = 0, sw_dll2 = 0;
    for (int64_t i = 0; i < n1; ++i) sw_dll1 += weights1[i] * (dll1[i] * dll1[i]);
    for (int64_t j = 0; j < n2; ++j) sw_dll2 += weights2[j] * (dll2[j] * dll2[j]);

    /* (iii) main pass */
    int64_t counter_selected = 0, counter_pairs = 0;
    const int zerr_on = p->has_zerr_cut && (ang < p->zerr_cut_deg * M_PI / 180.0);
    for (int64_t i = 0; i < n1; ++i) {
        if (weights1[i] == 0) continue;
        int i_selected = 1;
        if (zerr_on) {
            double z_qf = 0.5 * (z1[i] + z_qso_2);
            double dv_kms = fabs(z1[i] - z_qso_2) / (1 + z_qf);
            dv_kms *= SPEED_LIGHT_KMS;
            if (dv_kms < p->zerr_cut_kms) i_selected = 0;
        }
        for (int64_t j = 0; j < n2; ++j) {
            if (weights2[j] == 0) continue;
            double z = (z1[i] + z2[j]) / 2;
            int j_selected = 1;
            if ((p->has_z_min_pairs && z < p->z_min_pairs) ||
                (p->has_z_max_pairs && z > p->z_max_pairs))
                j_selected = 0;
            if (zerr_on) {
                double z_qf = 0.5 * (z2[j] + z_qso_1);
                double dv_kms = fabs(z2[j] - z_qso_1) / (1 + z_qf);
                dv_kms *= SPEED_LIGHT_KMS;
                if (dv_kms < p->zerr_cut_kms) j_selected = 0;
            }
            double r_par = (r_comov1[i] - r_comov2[j]) * cos_half;
            double r_trans = (dist_m1[i] + dist_m2[j]) * sin_half;
            if (p->rmu_binning) {
                r_trans = sqrt(r_trans * r_trans + r_par * r_par);
                r_par /= r_trans;
            }
            if (!p->x_correlation) r_par = fabs(r_par);
            if (r_par >= p->r_par_max || r_trans >= p->r_trans_max || r_par < p->r_par_min) continue;
            if (p->remove_same_half_plate_close_pairs && same_half_plate) {
                if (fabs(r_par) < (p->r_par_max - p->r_par_min) / p->num_bins_r_par) j_selected = 0;
            }
            double weights12 = weights1[i] * weights2[j];
            double zfac;
            if (p->redshift_evolution) {
                zfac = pow((1 + z1[i]) / (1 + p->z_ref), p->alpha - 1) *
                       pow((1 + z2[j]) / (1 + p->z_ref), p->alpha2 - 1);
            } else {
                zfac = 1.0;
            }
            double bins_r_par = floor((r_par - p->r_par_min) / (p->r_par_max - p->r_par_min) * p->num_bins_r_par);
            double bins_r_trans = floor(r_trans / p->r_trans_max * p->num_bins_r_trans);
            int32_t bins = (int32_t)(bins_r_trans + p->num_bins_r_trans * bins_r_par);
            double model_bins_r_par = floor((r_par - p->r_par_min) / (p->r_par_max - p->r_par_min) * p->num_model_bins_r_par);
            double model_bins_r_trans = floor(r_trans / p->r_trans_max * p->num_model_bins_r_trans);
            int32_t model_bins = (int32_t)(model_bins_r_trans + p->num_model_bins_r_trans * model_bins_r_par);

            if (counter_pairs < num_pairs) all_model[counter_pairs] = model_bins;
            else stats[2] += 1;
            counter_pairs += 1;
            stats[0] += 1;

            if (model_bins < 0 || model_bins >= n_model) {
                /* numba would index out of bounds here; keep the pair out of eta */
                continue;
            }
            if (i_selected && j_selected) {
                sel_model[counter_selected] = model_bins;
                sel_data[counter_selected] = bins;
                sel_i[counter_selected] = (int32_t)i;
                sel_j[counter_selected] = (int32_t)j;
                counter_selected += 1;
                stats[1] += 1;
                r_par_eff[model_bins] += weights12 * r_par;
                r_trans_eff[model_bins] += weights12 * r_trans;
                z_eff[model_bins] += weights12 * z;
                weight_eff[model_bins] += weights12;
                /* numba would write out of bounds for bins >= np*nt (bin rounding at the
                 * range limit): such pairs are counted and left out, as in the fill */
                if (bins < (int64_t)p->num_bins_r_par * p->num_bins_r_trans) {
                    if (bins >= 0) weights_dmat[bins] += weights12;
                } else {
                    stats[3] += 1;
                }
            }
            eta1[i + n1 * model_bins] += zfac * weights2[j] / sum_weights2;
            eta2[j + n2 * model_bins] += zfac * weights1[i] / sum_weights1;
            eta5[model_bins] += zfac * weights12 / sum_weights1 / sum_weights2;
            if (order2 == 1) {
                eta3[i + n1 * model_bins] += zfac * weights2[j] * dll2[j] / sw_dll2;
                eta6[model_bins] += zfac * weights1[i] / sum_weights1 * (weights2[j] * dll2[j] / sw_dll2);
            }
            if (order1 == 1) {
                eta4[j + n2 * model_bins] += zfac * weights1[i] * dll1[i] / sw_dll1;
                eta7[model_bins] += zfac * weights2[j] / sum_weights2 * (weights1[i] * dll1[i] / sw_dll1);
                if (order2 == 1) {
                    eta8[model_bins] += zfac * weights1[i] * dll1[i] * weights2[j] * dll2[j] / sw_dll1 / sw_dll2;
                }
            }
        }
    }

    /* (iv) fill over unique model bins of the first num_pairs in-range pairs */
    {
        int64_t n_all = counter_pairs < num_pairs ? counter_pairs : num_pairs;
        int32_t *uniq = (int32_t *)malloc(sizeof(int32_t) * (size_t)(num_pairs > 0 ? num_pairs : 1));
        if (!uniq) {
            status = -2;
            goto done;
        }
        /* entries never written keep their zero initialisation, as in the reference */
        memcpy(uniq, all_model, sizeof(int32_t) * (size_t)num_pairs);
        (void)n_all;
        qsort(uniq, (size_t)num_pairs, sizeof(int32_t), cmp_i32);
        int64_t n_uniq = 0;
        for (int64_t k = 0; k < num_pairs; ++k)
            if (k == 0 || uniq[k] != uniq[n_uniq - 1]) uniq[n_uniq++] = uniq[k];
        for (int64_t pair = 0; pair < counter_selected; ++pair) {
            int32_t i = sel_i[pair], j = sel_j[pair];
            int32_t bins = sel_data[pair];
            if (bins < 0) {
                status = -1;
                free(uniq);
                goto done;
            }
            int32_t model_bins = sel_model[pair];
            double weights12 = weights1[i] * weights2[j];
            int64_t dmat_bin = model_bins + n_model * (int64_t)bins;
            double zfac;
            if (p->redshift_evolution) {
                zfac = pow((1 + z1[i]) / (1 + p->z_ref), p->alpha - 1) *
                       pow((1 + z2[j]) / (1 + p->z_ref), p->alpha2 - 1);
            } else {
                zfac = 1;
            }
            int64_t nb = (int64_t)p->num_bins_r_par * p->num_bins_r_trans;
            if (bins >= nb) continue; /* out-of-range data bin: counted by caller */
            dmat[dmat_bin] += weights12 * zfac;
            for (int64_t u = 0; u < n_uniq; ++u) {
                int32_t k = uniq[u];
                if (k < 0 || k >= n_model) continue;
                dmat_bin = k + n_model * (int64_t)bins;
                dmat[dmat_bin] += weights12 * (eta5[k] + eta6[k] * dll2[j] + eta7[k] * dll1[i] +
                                               eta8[k] * dll1[i] * dll2[j] - eta1[i + n1 * k] -
                                               eta2[j + n2 * k] - eta3[i + n1 * k] * dll2[j] -
                                               eta4[j + n2 * k] * dll1[i]);
            }
        }
        free(uniq);
    }
done:
    free(dll1);
    free(dll2);
    free(eta1);
    free(eta2);
    free(eta3);
    free(eta4);
    free(eta5);
    free(eta6);
    free(eta7);
    free(eta8);
    free(sel_data);
    free(sel_model);
    free(sel_i);
    free(sel_j);
    free(all_model);
    return status;
}

/*
 * The kept pairs of a dmat run, in reference order, accumulated into one set of
 * outputs (compute_dmat sums everything of a process list into one matrix).
 */
int oracle_dmat_rows(const oracle_params *p, const int64_t *offsets,
                     const double *log_lambda, const double *z, const double *r_comov,
                     const double *dist_m, const double *weights, const double *z_qso,
                     const int32_t *order, int64_t num_pairs, const int32_t *pair_a,
                     const int32_t *pair_b, const double *pair_ang, const uint8_t *pair_shp,
                     double *weights_dmat, double *dmat, double *r_par_eff,
                     double *r_trans_eff, double *z_eff, double *weight_eff,
                     int64_t *stats) {
    for (int64_t k = 0; k < num_pairs; ++k) {
        int32_t a = pair_a[k], b = pair_b[k];
        int status = oracle_dmat_pair(
            p, log_lambda + offsets[a], log_lambda + offsets[b], r_comov + offsets[a],
            r_comov + offsets[b], dist_m + offsets[a], dist_m + offsets[b], z + offsets[a],
            z + offsets[b], weights + offsets[a], weights + offsets[b],
            offsets[a + 1] - offsets[a], offsets[b + 1] - offsets[b], z_qso[a], z_qso[b],
            pair_ang[k], weights_dmat, dmat, r_par_eff, r_trans_eff, z_eff, weight_eff,
            pair_shp[k], order[a], order[b], stats);
        if (status != 0) return status;
    }
    return 0;
}


/*
 * The cross-correlation distortion matrix of one delta against its kept quasars
 * (compute_dmat_forest_pairs_fast of py/picca/xcf.py:427-658).  Quasar j has r_comov2[j],
 * dist_m2[j], z2[j] (its redshift), weights2[j] and angle ang[j].  Only the delta's
 * continuum fit distorts the field: eta2 (w1 / sum w1) and, for order1 == 1, eta4
 * (w1 dll1 / sum w1 dll1^2) per quasar and model bin.  The redshift-evolution factor uses
 * alpha for the pixel and alpha2 (the quasars' exponent) for the quasar.
 * Returns 0, -1 for a selected pair with a negative data bin (IndexError in the
 * reference), -2 on allocation failure.  stats as oracle_dmat_pair: [0] in-range pairs of
 * the main pass, [1] selected pairs, [2] in-range pairs beyond the count-pass size (the
 * count pass rejects r_par <= r_par_min, the main pass only r_par < r_par_min,
 * py/picca/xcf.py:464-466, 542-544), [3] selected pairs with a data bin >= np*nt.
 */
int oracle_xdmat_forest(const oracle_params *p, const double *log_lambda1,
                        const double *r_comov1, const double *dist_m1, const double *z1,
                        const double *weights1, int64_t n1, const double *r_comov2,
                        const double *dist_m2, const double *z2, const double *weights2,
                        const double *ang, int64_t n2, double *weights_dmat, double *dmat,
                        double *r_par_eff, double *r_trans_eff, double *z_eff,
                        double *weight_eff, int32_t order1, int64_t *stats) {
    const int64_t n_model = (int64_t)p->num_model_bins_r_par * p->num_model_bins_r_trans;
    const int64_t nb = (int64_t)p->num_bins_r_par * p->num_bins_r_trans;

    /* (i) count pass */
    int64_t num_pairs = 0;
    for (int64_t i = 0; i < n1; ++i) {
        if (weights1[i] == 0) continue;
        for (int64_t j = 0; j < n2; ++j) {
            if (weights2[j] == 0) continue;
            double r_par = (r_comov1[i] - r_comov2[j]) * cos(ang[j] / 2);
            double r_trans = (dist_m1[i] + dist_m2[j]) * sin(ang[j] / 2);
            if (p->rmu_binning) {
                r_trans = sqrt(r_trans * r_trans + r_par * r_par);
                r_par /= r_trans;
            }
            if (r_par >= p->r_par_max || r_trans >= p->r_trans_max || r_par <= p->r_par_min)
                continue;
            num_pairs += 1;
        }
    }
    if (num_pairs == 0) return 0;

    /* (ii) per-forest sums, sequential like numba's ndarray.sum */
    double sum_weights1 = 0;
    for (int64_t i = 0; i < n1; ++i) sum_weights1 += weights1[i];
    double acc1 = 0;
    for (int64_t i = 0; i < n1; ++i) acc1 += log_lambda1[i] * weights1[i];
    const double mean_log_lambda1 = acc1 / sum_weights1;
    double *dll1 = (double *)malloc(sizeof(double) * (size_t)(n1 > 0 ? n1 : 1));
    double *eta2 = (double *)calloc((size_t)(n_model * n2), sizeof(double));
    double *eta4 = (double *)calloc((size_t)(n_model * n2), sizeof(double));
    /* the main pass can find more in-range pairs than the count pass: size the records
       by an upper bound and keep the reference's count-pass-many model bins for unique */
    const int64_t cap = n1 * n2 > 0 ? n1 * n2 : 1;
    int32_t *sel_data = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
    int32_t *sel_model = (int32_t *)calloc((size_t)cap, sizeof(int32_t));
    int32_t *sel_i = (int32_t *)calloc((size_t)cap, sizeof(int32_t));
    int32_t *sel_j = (int32_t *)calloc((size_t)cap, sizeof(int32_t));
    int32_t *all_model = (int32_t *)calloc((size_t)num_pairs, sizeof(int32_t));
    int32_t *uniq = (int32_t *)malloc(sizeof(int32_t) * (size_t)num_pairs);
    int status = 0;
    if (!dll1 || !eta2 || !eta4 || !sel_data || !sel_model || !sel_i || !sel_j ||
        !all_model || !uniq) {
        status = -2;
        goto done;
    }
    for (int64_t i = 0; i < n1; ++i) dll1[i] = log_lambda1[i] - mean_log_lambda1;
    double sw_dll1 = 0;
    for (int64_t i = 0; i < n1; ++i) sw_dll1 += weights1[i] * (dll1[i] * dll1[i]);

    /* (iii) main pass */
    int64_t counter_selected = 0, counter_pairs = 0;
    for (int64_t i = 0; i < n1; ++i) {
        if (weights1[i] == 0) continue;
        for (int64_t j = 0; j < n2; ++j) {
            if (weights2[j] == 0) continue;
            double z = (z1[i] + z2[j]) / 2;
            int j_selected = 1;
            if ((p->has_z_min_pairs && z < p->z_min_pairs) ||
                (p->has_z_max_pairs && z > p->z_max_pairs))
                j_selected = 0;
            double r_par = (r_comov1[i] - r_comov2[j]) * cos(ang[j] / 2);
            double r_trans = (dist_m1[i] + dist_m2[j]) * sin(ang[j] / 2);
            if (p->rmu_binning) {
                r_trans = sqrt(r_trans * r_trans + r_par * r_par);
                r_par /= r_trans;
            }
            if (r_par >= p->r_par_max || r_trans >= p->r_trans_max || r_par < p->r_par_min)
                continue;
            double weights12 = weights1[i] * weights2[j];
            double zfac;
            if (p->redshift_evolution) {
                zfac = pow((1 + z1[i]) / (1 + p->z_ref), p->alpha - 1) *
                       pow((1 + z2[j]) / (1 + p->z_ref), p->alpha2 - 1);
            } else {
                zfac = 1;
            }
            double bins_r_par = floor((r_par - p->r_par_min) / (p->r_par_max - p->r_par_min) * p->num_bins_r_par);
            double bins_r_trans = floor(r_trans / p->r_trans_max * p->num_bins_r_trans);
            int32_t bins = (int32_t)(bins_r_trans + p->num_bins_r_trans * bins_r_par);
            double model_bins_r_par = floor((r_par - p->r_par_min) / (p->r_par_max - p->r_par_min) * p->num_model_bins_r_par);
            double model_bins_r_trans = floor(r_trans / p->r_trans_max * p->num_model_bins_r_trans);
            int32_t model_bins = (int32_t)(model_bins_r_trans + p->num_model_bins_r_trans * model_bins_r_par);

            if (counter_pairs < num_pairs) all_model[counter_pairs] = model_bins;
            else stats[2] += 1;
            counter_pairs += 1;
            stats[0] += 1;

            if (model_bins < 0 || model_bins >= n_model) continue;  /* out of bounds in numba */
            if (j_selected) {
                sel_model[counter_selected] = model_bins;
                sel_data[counter_selected] = bins;
                sel_i[counter_selected] = (int32_t)i;
                sel_j[counter_selected] = (int32_t)j;
                counter_selected += 1;
                stats[1] += 1;
                r_par_eff[model_bins] += weights12 * r_par;
                r_trans_eff[model_bins] += weights12 * r_trans;
                z_eff[model_bins] += weights12 * z;
                weight_eff[model_bins] += weights12;
                if (bins < nb) {
                    if (bins >= 0) weights_dmat[bins] += weights12;
                } else {
                    stats[3] += 1;
                }
            }
            eta2[j + n2 * model_bins] += zfac * weights1[i] / sum_weights1;
            if (order1 == 1)
                eta4[j + n2 * model_bins] += zfac * (weights1[i] * dll1[i] / sw_dll1);
        }
    }

    /* (iv) fill over the unique model bins of the first num_pairs in-range pairs */
    {
        memcpy(uniq, all_model, sizeof(int32_t) * (size_t)num_pairs);
        qsort(uniq, (size_t)num_pairs, sizeof(int32_t), cmp_i32);
        int64_t n_uniq = 0;
        for (int64_t k = 0; k < num_pairs; ++k)
            if (k == 0 || uniq[k] != uniq[n_uniq - 1]) uniq[n_uniq++] = uniq[k];
        for (int64_t pair = 0; pair < counter_selected; ++pair) {
            int32_t i = sel_i[pair], j = sel_j[pair];
            int32_t bins = sel_data[pair];
            if (bins < 0) {
                status = -1;
                goto done;
            }
            int32_t model_bins = sel_model[pair];
            double weights12 = weights1[i] * weights2[j];
            double zfac;
            if (p->redshift_evolution) {
                zfac = pow((1 + z1[i]) / (1 + p->z_ref), p->alpha - 1) *
                       pow((1 + z2[j]) / (1 + p->z_ref), p->alpha2 - 1);
            } else {
                zfac = 1;
            }
            if (bins >= nb) continue;
            dmat[model_bins + n_model * (int64_t)bins] += zfac * weights12;
            for (int64_t u = 0; u < n_uniq; ++u) {
                int32_t k = uniq[u];
                if (k < 0 || k >= n_model) continue;
                dmat[k + n_model * (int64_t)bins] +=
                    weights12 * (-eta2[j + n2 * k] - eta4[j + n2 * k] * dll1[i]);
            }
        }
    }
done:
    free(dll1);
    free(eta2);
    free(eta4);
    free(sel_data);
    free(sel_model);
    free(sel_i);
    free(sel_j);
    free(all_model);
    free(uniq);
    return status;
}


/*
 * One pass of the metal distortion matrix for one kept pair.
 *
 *   mode 0: compute_metal_dmat of the auto-correlation (py/picca/cf.py:994-1090 for the
 *           first pass, :1092-1200 for the swapped one), |r_par|;
 *   mode 1: the same with x_correlation (signed r_par);
 *   mode 2: compute_metal_dmat of xcf (py/picca/xcf.py:710-790): delta x quasar.
 *
 * Pixel i of the first forest has its Lya coordinates (r1, d1), its weight w1 (0 where
 * the absorber's redshift z1a is not below the quasar's: the reference drops those
 * pixels, which changes no sum), the absorber's redshift, coordinates (z1a, r1a, d1a)
 * and the redshift-evolution factor f1a; the same for the second forest (a quasar in
 * mode 2: its own redshift and coordinates, f = 1).  The factor of a pair is
 * f1a * f2a / zw_den, as the reference's
 * (1+z1)^(a1-1) * (1+z2)^(a2-1) / (1+z_ref)^(a1+a2-2) (modes 0, 1) or
 * ((1+z1)/(1+z_ref))^(a-1) (mode 2, with f2a = zw_den = 1).
 *
 * Data bins use the Lya coordinates, model bins the absorbers'.  Modes 0 and 1 keep a pair
 * whose bins fall in range (floor of r_par, truncation of r_trans); mode 2 keeps
 * r_min < r_par < r_max, r_trans < r_t_max for both sets of coordinates, then truncates.
 * Same-half-plate close pairs get weight 0 (modes 0, 1).  A data bin beyond np*nt in
 * mode 2 (rounding up at r_par_max; the reference's bincount would then fail) is skipped
 * and counted.  stats: [0] pairs added to weights_dmat, [1] pairs added to dmat,
 * [2] skipped overflow pairs.
 */
void oracle_metal_pair(const oracle_params *p, int32_t mode, double zw_den,
                       const double *r1, const double *d1, const double *w1,
                       const double *z1a, const double *r1a, const double *d1a,
                       const double *f1a, int64_t n1, const double *r2, const double *d2,
                       const double *w2, const double *z2a, const double *r2a,
                       const double *d2a, const double *f2a, int64_t n2, double ang,
                       int32_t same_half_plate, double *weights_dmat, double *dmat,
                       double *r_par_eff, double *r_trans_eff, double *z_eff,
                       double *weight_eff, int64_t *stats) {
    const int np_ = p->num_bins_r_par, nt = p->num_bins_r_trans;
    const int npm = p->num_model_bins_r_par, ntm = p->num_model_bins_r_trans;
    const int64_t nb = (int64_t)np_ * nt, nm = (int64_t)npm * ntm;
    const double span = p->r_par_max - p->r_par_min;
    const double c = cos(ang / 2), s = sin(ang / 2);
    const int shp = p->remove_same_half_plate_close_pairs && same_half_plate && mode != 2;
    for (int64_t i = 0; i < n1; ++i) {
        if (w1[i] == 0) continue;
        for (int64_t j = 0; j < n2; ++j) {
            if (w2[j] == 0) continue;
            double r_par = (r1[i] - r2[j]) * c;
            if (mode == 0) r_par = fabs(r_par);
            const double r_trans = (d1[i] + d2[j]) * s;
            double w12 = w1[i] * w2[j];
            int64_t bin;
            if (mode == 2) {
                if (!(r_par > p->r_par_min && r_par < p->r_par_max && r_trans < p->r_trans_max))
                    continue;
                const int64_t bp = (int64_t)((r_par - p->r_par_min) / span * np_);
                const int64_t bt = (int64_t)(r_trans / p->r_trans_max * nt);
                bin = bt + (int64_t)nt * bp;
                if (bin >= nb) {
                    stats[2] += 1;
                    continue;
                }
                if (bt < 0) continue;   /* negative r_trans: not a physical input */
            } else {
                const double bp = floor((r_par - p->r_par_min) / span * np_);
                const double bt = trunc(r_trans / p->r_trans_max * nt);
                if (!(bp < np_ && bt < nt && bp >= 0 && bt >= 0)) continue;
                if (shp && fabs(r_par) < span / np_) w12 = 0.0;
                bin = (int64_t)bt + (int64_t)nt * (int64_t)bp;
            }
            weights_dmat[bin] += w12;
            stats[0] += 1;

            double r_par_a = (r1a[i] - r2a[j]) * c;
            if (mode == 0) r_par_a = fabs(r_par_a);
            const double r_trans_a = (d1a[i] + d2a[j]) * s;
            int64_t model;
            if (mode == 2) {
                if (!(r_par_a > p->r_par_min && r_par_a < p->r_par_max &&
                      r_trans_a < p->r_trans_max))
                    continue;
                const int64_t mp = (int64_t)((r_par_a - p->r_par_min) / span * npm);
                const int64_t mt = (int64_t)(r_trans_a / p->r_trans_max * ntm);
                model = mt + (int64_t)ntm * mp;
                if (model >= nm) {
                    stats[2] += 1;
                    continue;
                }
                if (mt < 0) continue;
            } else {
                const double mp = floor((r_par_a - p->r_par_min) / span * npm);
                const double mt = trunc(r_trans_a / p->r_trans_max * ntm);
                if (!(mp < npm && mt < ntm && mp >= 0 && mt >= 0)) continue;
                model = (int64_t)mt + (int64_t)ntm * (int64_t)mp;
            }
            const double zw = f1a[i] * f2a[j] / zw_den;
            const double wz = w12 * zw;
            dmat[model + nm * bin] += wz;
            r_par_eff[model] += r_par_a * w12 * zw;
            r_trans_eff[model] += r_trans_a * w12 * zw;
            z_eff[model] += (z1a[i] + z2a[j]) / 2 * w12 * zw;
            weight_eff[model] += wz;
            stats[1] += 1;
        }
    }
}


/*
 * Wick terms T1, T2, T3 of the auto-correlation covariance for a list of forest pairs
 * (compute_wickT123_pairs, py/picca/cf.py:1486-1602, called per pair by
 * compute_wick_terms, :1380-1424).  Forest f has pixels offsets[f] .. offsets[f+1]-1 with
 * r_comov, weights and the evolution factors zev1 = ((1+z)/(1+z_ref))^(alpha-1),
 * zev2 = ((1+z)/(1+z_ref))^(alpha2-1); wx + wx_offsets[f] is its n x n weighted 1-D
 * correlation (w_k w_l xi_1d(|log_lambda_k - log_lambda_l|) sqrt(v_k v_l)).
 * r_trans is (r1 + r2) sin(ang/2) with the comoving distances, as the reference.
 * The pairs are split over OpenMP threads with private accumulators, summed at the end.
 */
static void wick_pair(const oracle_params *p, const double *r1, const double *w1,
                      const double *za1, int64_t n1, const double *r2, const double *w2,
                      const double *zb2, int64_t n2, double ang, const double *wx1,
                      const double *wx2, double *weights_wick, int64_t *num_pairs_wick,
                      double *t1, double *t2, double *t3, int64_t nb) {
    const double c = cos(ang / 2), s = sin(ang / 2);
    const double span = p->r_par_max - p->r_par_min;
    int64_t count = 0;
    for (int64_t j = 0; j < n2; ++j)
        for (int64_t i = 0; i < n1; ++i) {
            double r_par = (r1[i] - r2[j]) * c;
            if (!p->x_correlation) r_par = fabs(r_par);
            const double r_trans = (r1[i] + r2[j]) * s;
            if (r_par < p->r_par_max && r_trans < p->r_trans_max && r_par >= p->r_par_min)
                ++count;
        }
    if (count == 0) return;
    int64_t *pi = (int64_t *)malloc(sizeof(int64_t) * (size_t)count);
    int64_t *pj = (int64_t *)malloc(sizeof(int64_t) * (size_t)count);
    int64_t *pb = (int64_t *)malloc(sizeof(int64_t) * (size_t)count);
    int64_t k = 0;
    for (int64_t j = 0; j < n2; ++j)
        for (int64_t i = 0; i < n1; ++i) {
            double r_par = (r1[i] - r2[j]) * c;
            if (!p->x_correlation) r_par = fabs(r_par);
            const double r_trans = (r1[i] + r2[j]) * s;
            if (!(r_par < p->r_par_max && r_trans < p->r_trans_max && r_par >= p->r_par_min))
                continue;
            const int64_t bp = (int64_t)((r_par - p->r_par_min) / span * p->num_bins_r_par);
            const int64_t bt = (int64_t)(r_trans / p->r_trans_max * p->num_bins_r_trans);
            pi[k] = i;
            pj[k] = j;
            pb[k] = bt + (int64_t)p->num_bins_r_trans * bp;
            ++k;
        }
    for (int64_t a = 0; a < count; ++a) {
        const int64_t p1 = pb[a], i1 = pi[a], j1 = pj[a];
        if (p1 < 0 || p1 >= nb) continue;   /* rounding up to np: numba writes out of bounds */
        const double w12 = w1[i1] * w2[j1];
        weights_wick[p1] += w12;
        num_pairs_wick[p1] += 1;
        t1[p1 * nb + p1] += w12 * (za1[i1] * zb2[j1]);
        for (int64_t b = a + 1; b < count; ++b) {
            const int64_t p2 = pb[b], i2 = pi[b], j2 = pj[b];
            if (p2 < 0 || p2 >= nb) continue;
            double prod;
            if (i1 == i2) {
                prod = wx2[j1 * n2 + j2] * w1[i1] * za1[i1];
                t2[p1 * nb + p2] += prod;
                t2[p2 * nb + p1] += prod;
            } else if (j1 == j2) {
                prod = wx1[i1 * n1 + i2] * w2[j2] * zb2[j1];
                t2[p1 * nb + p2] += prod;
                t2[p2 * nb + p1] += prod;
            } else {
                prod = wx1[i1 * n1 + i2] * wx2[j1 * n2 + j2];
                t3[p1 * nb + p2] += prod;
                t3[p2 * nb + p1] += prod;
            }
        }
    }
    free(pi);
    free(pj);
    free(pb);
}

void oracle_wick_t123(const oracle_params *p, int64_t num_pairs, const int32_t *pair_a,
                      const int32_t *pair_b, const double *pair_ang, const int64_t *offsets,
                      const double *r_comov, const double *weights, const double *zev1,
                      const double *zev2, const int64_t *wx_offsets, const double *wx,
                      double *weights_wick, int64_t *num_pairs_wick, double *t1, double *t2,
                      double *t3) {
    const int64_t nb = (int64_t)p->num_bins_r_par * p->num_bins_r_trans;
#pragma omp parallel
    {
        double *lw = (double *)calloc((size_t)nb, sizeof(double));
        int64_t *ln = (int64_t *)calloc((size_t)nb, sizeof(int64_t));
        double *l1 = (double *)calloc((size_t)(nb * nb), sizeof(double));
        double *l2 = (double *)calloc((size_t)(nb * nb), sizeof(double));
        double *l3 = (double *)calloc((size_t)(nb * nb), sizeof(double));
#pragma omp for schedule(dynamic, 1)
        for (int64_t k = 0; k < num_pairs; ++k) {
            const int64_t a = pair_a[k], b = pair_b[k];
            const int64_t oa = offsets[a], ob = offsets[b];
            wick_pair(p, r_comov + oa, weights + oa, zev1 + oa, offsets[a + 1] - oa,
                      r_comov + ob, weights + ob, zev2 + ob, offsets[b + 1] - ob, pair_ang[k],
                      wx + wx_offsets[a], wx + wx_offsets[b], lw, ln, l1, l2, l3, nb);
        }
#pragma omp critical
        {
            for (int64_t q = 0; q < nb; ++q) {
                weights_wick[q] += lw[q];
                num_pairs_wick[q] += ln[q];
            }
            for (int64_t q = 0; q < nb * nb; ++q) {
                t1[q] += l1[q];
                t2[q] += l2[q];
                t3[q] += l3[q];
            }
        }
        free(lw);
        free(ln);
        free(l1);
        free(l2);
        free(l3);
    }
}
