"""Absorber tables of the metal distortion matrices.

compute_metal_dmat (``py/picca/cf.py:890-1232``, ``py/picca/xcf.py:677-800``) re-reads each
forest's pixels as another absorber: z = 10^log_lambda / lambda_abs - 1, its comoving and
transverse distances, the pixels with z >= z_qso dropped, and a redshift-evolution factor.
Here that view is built once per absorber for every pixel of a forest store (weight 0
stands for a dropped pixel) and handed to the engine (``picca_metal_set_absorber``).
"""
import numpy as np


def absorber_table(store, lambda_abs, alpha_abs, cosmo, z_ref=None):
    """(z, r_comov, dist_m, weights, zfac) of every pixel of ``store`` seen as absorber
    ``lambda_abs``.  zfac is (1+z)^(alpha_abs-1) (``py/picca/cf.py:1062-1066``) or, with
    z_ref, ((1+z)/(1+z_ref))^(alpha_abs-1) (``py/picca/xcf.py:747-749``)."""
    if store.log_lambda is None:
        raise RuntimeError("the metal distortion matrix needs the deltas' log_lambda")
    z = 10**store.log_lambda / lambda_abs - 1
    r = np.asarray(cosmo.get_r_comov(z), dtype=np.float64)
    d = np.asarray(cosmo.get_dist_m(z), dtype=np.float64)
    lengths = np.diff(store.offsets)
    z_qso = np.repeat(store.z_qso, lengths)
    w = np.where(z < z_qso, store.weights, 0.0)
    if z_ref is None:
        f = (1 + z) ** (alpha_abs - 1)
    else:
        f = ((1.0 + z) / (1.0 + z_ref)) ** (alpha_abs - 1.0)
    return z, r, d, w, f


def quasar_table(z_qso, r_comov, dist_m, weights):
    """Quasars stored as one-pixel forests: their own redshift and distances, weight, and
    evolution factor 1 (the quasar weight already carries its evolution)."""
    n = len(z_qso)
    return (np.asarray(z_qso, dtype=np.float64), np.asarray(r_comov, dtype=np.float64),
            np.asarray(dist_m, dtype=np.float64), np.asarray(weights, dtype=np.float64),
            np.ones(n))


def outputs(num_bins, num_model_bins):
    """Zeroed (weights_dmat, dmat (flat), r_par_eff, r_trans_eff, z_eff, weight_eff)."""
    return [np.zeros(num_bins), np.zeros(num_bins * num_model_bins)] + \
        [np.zeros(num_model_bins) for _ in range(4)]
