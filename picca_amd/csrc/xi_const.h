// Host side: launch constants of the pair kernels from picca_cf_params (the module globals
// of py/picca/cf.py:28-79), shared by the library and the CPU harness of the tests.
#pragma once

#include <cmath>

#include "../../include/picca_hip.h"
#include "engine.h"

namespace picca {

// Fills C and the binning mode; returns nullptr, or the message of the ValueError picca's
// configuration would deserve.
inline const char *fill_xi_const(const picca_cf_params *p, XiConst &C, int &mode) {
    if (!p) return "params is NULL";
    if (p->num_bins_r_par <= 0 || p->num_bins_r_trans <= 0)
        return "num_bins_r_par and num_bins_r_trans must be positive";
    if (!(p->r_par_max > p->r_par_min))
        return "r_par_max must be larger than r_par_min";
    if (!(p->r_trans_max > 0))
        return "r_trans_max must be positive";
    if (!(p->edge_tol >= 0)) return "edge_tol must be >= 0";
    if (p->has_zerr_cut && !(p->zerr_cut_kms == p->zerr_cut_kms))
        return "zerr_cut_kms must be given with zerr_cut_deg";
    C = XiConst{};
    C.np = p->num_bins_r_par;
    C.nt = p->num_bins_r_trans;
    C.nb = C.np * C.nt;
    C.ntp = C.nt + 1;
    C.rpmin = p->r_par_min;
    C.rpmax = p->r_par_max;
    C.rtmax = p->r_trans_max;
    C.span_p = p->r_par_max - p->r_par_min;
    C.np_d = (double)C.np;
    C.nt_d = (double)C.nt;
    C.nb_d = (double)C.nb;
    C.ntp_d = (double)C.ntp;
    C.inv_p = C.np_d / C.span_p;
    C.inv_t = C.nt_d / C.rtmax;
    C.width_p = C.span_p / C.np;
    C.width_t = C.rtmax / C.nt;
    C.edge_tol = p->edge_tol;
    // |t_fast - t_exact| <= 4 ulp(t) << 1e-12 for t < 2^20; the margin also absorbs the
    // rounding of frac * width in the edge test.
    C.fast_p = p->edge_tol / C.width_p * (1.0 + 1e-6) + 1e-12;
    C.fast_t = p->edge_tol / C.width_t * (1.0 + 1e-6) + 1e-12;
    C.half_p = 0.5 - C.fast_p;
    C.half_t = 0.5 - C.fast_t;
    C.zmin = p->z_min_pairs;
    C.zmax = p->z_max_pairs;
    C.zerr_ang = p->zerr_cut_deg * M_PI / 180.0;
    C.zerr_kms = p->zerr_cut_kms;
    int flags = 0;
    if (p->has_z_min_pairs) flags |= XI_HAS_ZMIN;
    if (p->has_z_max_pairs) flags |= XI_HAS_ZMAX;
    if (p->has_zerr_cut) flags |= XI_HAS_ZERR;
    if (p->remove_same_half_plate_close_pairs) flags |= XI_REMOVE_SHP;
    mode = p->ang_correlation ? MODE_ANGULAR
                              : (p->rmu_binning ? MODE_RMU : MODE_STANDARD);
    if (mode == MODE_STANDARD) flags |= XI_PRUNE;
    C.flags = flags;
    return nullptr;
}

// The fast loops take the bin coordinates as |r1 - r2| * (cos * np / span) and
// (d1 + d2) * (sin * nt / r_trans_max) instead of the reference's
// ((r1 - r2) * cos - r_par_min) / span * np and (d1 + d2) * sin / r_trans_max * nt; both
// differ from the exact value by a few ulp of the coordinate.  The margins around every bin
// edge and range limit are widened by a bound of that difference (in bin units), so that
// every pair whose bin or range test could differ is re-evaluated exactly.  max_abs_r and
// max_abs_d bound |r_comov| and |dist_m| over the store.
inline const char *widen_xi_margins(XiConst &C, double max_abs_r, double max_abs_d) {
    const double u = std::ldexp(1.0, -53);
    const double err_p = 16.0 * u * max_abs_r * C.inv_p + 8.0 * u * (C.np_d + 1.0);
    const double err_t = 16.0 * u * max_abs_d * C.inv_t + 8.0 * u * (C.nt_d + 1.0);
    C.fast_p += err_p;
    C.fast_t += err_t;
    C.half_p = 0.5 - C.fast_p;
    C.half_t = 0.5 - C.fast_t;
    if (!(C.half_p > 0.25 && C.half_t > 0.25))
        return "edge_tol is too large compared with the bin width";
    return nullptr;
}

// Schedule of the specialised kernel (PICCA_XI_SCHEDULE_*): RUNS walks every eligible task,
// AUTO the eligible tasks whose window holds at least XI_WALK_MIN_DENSITY pixels per r_par
// bin (walk_task in xi_walk.h), PAIRS none.
inline void apply_xi_schedule(XiConst &C, int schedule) {
    C.flags &= ~(XI_WALK | XI_WALK_AUTO);
    if (schedule == PICCA_XI_SCHEDULE_RUNS) C.flags |= XI_WALK;
    else if (schedule == PICCA_XI_SCHEDULE_AUTO) C.flags |= XI_WALK | XI_WALK_AUTO;
}

// The specialised kernel (and its run walk) covers standard binning, auto-correlation,
// r_par_min == 0 and no z / zerr / same-half-plate cuts.
inline bool xi_specialised(const picca_cf_params *p, int mode) {
    return mode == MODE_STANDARD && !p->x_correlation && p->r_par_min == 0.0 &&
           !p->has_z_min_pairs && !p->has_z_max_pairs && !p->has_zerr_cut &&
           !p->remove_same_half_plate_close_pairs;
}

}  // namespace picca
