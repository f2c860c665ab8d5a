"""Small helpers on the correlation path.

* ``userprint`` mirrors ``py/picca/utils.py:20-28`` (print and flush);
* ``compute_ang_max`` mirrors ``py/picca/utils.py:419-450``: the largest angle at which a
  pair can still have r_trans < r_trans_max, 2 arcsin(r_trans_max / (D_M(z1) + D_M(z2)))
  at the lowest pixel redshifts, or pi when that argument exceeds one;
* ``shuffle_distrib_forests`` mirrors ``py/picca/utils.py:453-500``: reassigns sky
  positions among forests with a legacy-seeded permutation;
* ``compute_cov`` (the GPU subsampling covariance), ``compute_cov_boot`` and
  ``smooth_cov`` mirror ``py/picca/utils.py:100-259`` for picca_export.
"""
import sys

import numpy as np


def userprint(*args, **kwds):
    print(*args, **kwds)
    sys.stdout.flush()


def compute_ang_max(cosmo, r_trans_max, z_min, z_min2=None):
    if z_min2 is None:
        z_min2 = z_min
    r_min = cosmo.get_dist_m(z_min)
    r_min2 = cosmo.get_dist_m(z_min2)
    if r_min + r_min2 < r_trans_max:
        return np.pi
    return 2.0 * np.arcsin(r_trans_max / (r_min + r_min2))


def shuffle_distrib_forests(data, seed):
    """Give every forest the sky position of another one (seeded, legacy MT19937)."""
    userprint("INFO: Shuffling the forests angular position with seed {}".format(seed))
    params = ["ra", "dec", "x_cart", "y_cart", "z_cart", "cos_dec", "thingid"]
    info = {p: [] for p in params}
    info["healpix"] = []
    for healpix, deltas in data.items():
        for delta in deltas:
            info["healpix"].append(healpix)
            for p in params:
                info[p].append(getattr(delta, p))
    np.random.seed(seed)
    order = np.arange(len(info["ra"]))
    np.random.shuffle(order)
    shuffled = {}
    index = 0
    for deltas in data.values():
        for delta in deltas:
            src = order[index]
            for p in params:
                setattr(delta, p, info[p][src])
            shuffled.setdefault(info["healpix"][src], []).append(delta)
            index += 1
    return shuffled


def compute_cov(xi, weights, device=0):
    """Covariance of the rows of ``xi`` (one per healpix) by subsampling
    (``py/picca/utils.py:100-127``), on the GPU (``picca_cov_subsample``).  ``xi`` and
    ``weights`` are [num_samples, num_bins]; returns [num_bins, num_bins]."""
    from .engine import get_engine
    userprint("Computing cov...")
    return get_engine(device).cov_subsample(xi, weights)[0]


def compute_cov_boot(xi, weights, nboots=10000, seed=121567):
    """Bootstrap covariance (``py/picca/utils.py:130-151``): the same generator, the same
    draws and numpy's covariance of the bootstrap means."""
    nhpx, ndata = xi.shape
    boot_xis = np.empty((nboots, ndata))
    rnst = np.random.default_rng(seed)
    for i in range(nboots):
        idx = rnst.choice(nhpx, size=nhpx)
        wei = weights[idx]
        boot_xis[i] = np.sum(wei * xi[idx], axis=0) / wei.sum(0)
    return np.cov(boot_xis, rowvar=False)


def smooth_cov(xi, weights, r_par, r_trans, delta_r_trans=4.0, delta_r_par=4.0,
               covariance=None, per_r_par=False):
    """Smoothed covariance (``py/picca/utils.py:154-259``): the correlation coefficient of
    every pair of bins replaced by its mean over the pairs with the same rounded
    (|delta r_par| / delta_r_par, |delta r_trans| / delta_r_trans) -- and the same
    int(r_par / delta_r_par) of the first bin when ``per_r_par`` -- the variances kept.
    The means are accumulated over the upper triangle in row order, as the reference's
    loops do, in one bincount."""
    if covariance is None:
        covariance = compute_cov(xi, weights)
    num_bins = covariance.shape[1]
    var = np.diagonal(covariance)
    if np.any(var == 0.):
        userprint('WARNING: data has some empty bins, impossible to smooth')
        userprint('WARNING: returning the unsmoothed covariance')
        return covariance
    correlation = covariance / np.sqrt(var * var[:, None])
    r_par = np.asarray(r_par, dtype=np.float64)
    r_trans = np.asarray(r_trans, dtype=np.float64)
    i_up, j_up = np.triu_indices(num_bins, k=1)
    # round() of the reference rounds half to even, as np.rint
    key_par = np.rint(np.abs(r_par[j_up] - r_par[i_up]) / delta_r_par).astype(np.int64)
    key_trans = np.rint(np.abs(r_trans[i_up] - r_trans[j_up]) / delta_r_trans).astype(np.int64)
    keys = [key_par, key_trans]
    if per_r_par:
        keys.insert(0, (r_par / delta_r_par).astype(np.int64)[i_up])
    _, inverse = np.unique(np.stack(keys), axis=1, return_inverse=True)
    inverse = inverse.ravel()
    sums = np.bincount(inverse, weights=correlation[i_up, j_up])
    counts = np.bincount(inverse)
    correlation_smooth = np.eye(num_bins)
    values = sums[inverse] / counts[inverse]
    correlation_smooth[i_up, j_up] = values
    correlation_smooth[j_up, i_up] = values
    return correlation_smooth * np.sqrt(var * var[:, None])
