"""CPU oracle of the covariance export -- TEST INFRASTRUCTURE ONLY.

Only tests and the CPU-baseline leg of ``bench.py`` use this module, as the checker.

* ``compute_cov`` restates the subsampling covariance of ``py/picca/utils.py:100-127``
  (numpy, the same operations in the same order; the product ``Y^T Y`` goes through
  numpy's BLAS).
* ``smooth_cov`` restates ``py/picca/utils.py:177-259`` with the reference's loops and
  dictionaries (average correlation per rounded (|delta r_par|, |delta r_trans|) or, per
  r_par, per (r_par bin, |delta r_par|, |delta r_trans|)).
* ``export_columns`` restates the default path of ``py/picca/bin/picca_export.py:176-404``
  (no multipoles, no blinding): subsampling covariance, smoothing, weighted mean of the
  rows, the distortion matrix and its coordinates.
"""
import numpy as np


def compute_cov(xi, weights):
    mean_xi = (xi * weights).sum(axis=0)
    sum_weights = weights.sum(axis=0)
    w = sum_weights > 0.
    mean_xi[w] /= sum_weights[w]
    meanless_xi_times_weight = weights * (xi - mean_xi)
    covariance = meanless_xi_times_weight.T.dot(meanless_xi_times_weight)
    sum_weights_squared = sum_weights * sum_weights[:, None]
    w = sum_weights_squared > 0.
    covariance[w] /= sum_weights_squared[w]
    return covariance


def smooth_cov(covariance, r_par, r_trans, delta_r_trans=4.0, delta_r_par=4.0,
               per_r_par=False):
    num_bins = covariance.shape[1]
    var = np.diagonal(covariance)
    if np.any(var == 0.):
        return covariance
    correlation = covariance / np.sqrt(var * var[:, None])
    correlation_smooth = np.zeros([num_bins, num_bins])
    sums = {}
    counts = {}
    for i in range(num_bins):
        ip = int(r_par[i] / delta_r_par) if per_r_par else None
        for j in range(i + 1, num_bins):
            key = (round(abs(r_par[j] - r_par[i]) / delta_r_par),
                   round(abs(r_trans[i] - r_trans[j]) / delta_r_trans))
            if per_r_par:
                key = (ip,) + key
            sums[key] = sums.get(key, 0.) + correlation[i, j]
            counts[key] = counts.get(key, 0) + 1
    for i in range(num_bins):
        correlation_smooth[i, i] = 1.
        ip = int(r_par[i] / delta_r_par) if per_r_par else None
        for j in range(i + 1, num_bins):
            key = (round(abs(r_par[j] - r_par[i]) / delta_r_par),
                   round(abs(r_trans[i] - r_trans[j]) / delta_r_trans))
            if per_r_par:
                key = (ip,) + key
            correlation_smooth[i, j] = sums[key] / counts[key]
            correlation_smooth[j, i] = correlation_smooth[i, j]
    return correlation_smooth * np.sqrt(var * var[:, None])


def export_columns(cf_hdus, dmat_hdus=None, smooth=True, per_r_par=False):
    """cf_hdus: (ATTRI columns, COR columns, ATTRI header) of a cf file; dmat_hdus: (DMAT
    columns, ATTRI columns) or None.  Returns the COR and DMATTRI columns."""
    attri, cor, head = cf_hdus
    r_par = np.array(attri["RP"])
    r_trans = np.array(attri["RT"])
    z = np.array(attri["Z"])
    num_pairs = np.array(attri["NB"])
    weights = np.array(cor["WE"])
    xi = np.array(cor["DA"])
    covariance = compute_cov(xi, weights)
    if smooth:
        delta_r_par = (head["RPMAX"] - head["RPMIN"]) / head["NP"]
        delta_r_trans = (head["RTMAX"] - 0.0) / head["NT"]
        covariance = smooth_cov(covariance, r_par, r_trans, delta_r_trans=delta_r_trans,
                                delta_r_par=delta_r_par, per_r_par=per_r_par)
    xi = (xi * weights).sum(axis=0)
    weights = weights.sum(axis=0)
    w = weights > 0
    xi[w] /= weights[w]
    if dmat_hdus is not None:
        dmat = np.array(dmat_hdus[0]["DM"])
        r_par_dmat = np.array(dmat_hdus[1]["RP"])
        r_trans_dmat = np.array(dmat_hdus[1]["RT"])
        z_dmat = np.array(dmat_hdus[1]["Z"])
        if dmat.shape == (xi.size, xi.size):
            r_par_dmat, r_trans_dmat, z_dmat = r_par.copy(), r_trans.copy(), z.copy()
    else:
        dmat = np.eye(len(xi))
        r_par_dmat, r_trans_dmat, z_dmat = r_par.copy(), r_trans.copy(), z.copy()
    return {"DA": xi, "RP": r_par, "RT": r_trans, "Z": z, "CO": covariance, "DM": dmat,
            "NB": num_pairs, "DMRP": r_par_dmat, "DMRT": r_trans_dmat, "DMZ": z_dmat}
