"""Host preprocessing of the cross-correlation drivers: removal of the mean delta as a
function of observed wavelength.

``remove_mean_lambda_obs`` restates the block of py/picca/bin/picca_xcf.py:436-468: the
observed-frame grid step is the smallest log-lambda step of any forest, the grid spans
log10((1 + z) lambda_abs) over the deltas' redshift range widened by half a step, the
stack is ``stack(data, stack_from_deltas=True)`` (py/picca/prep_del.py:13-60: weighted
mean delta per grid bin, bins = int((log_lambda - min) / step + 0.5)), and each delta
has the stack of its bins subtracted.
"""
import numpy as np


def stack_deltas(data, log_lambda_min, log_lambda_max, delta_log_lambda):
    """(stack_log_lambda, stack_delta, stack_weight) of the deltas in ``data``."""
    num_bins = int((log_lambda_max - log_lambda_min) / delta_log_lambda) + 1
    stack_log_lambda = log_lambda_min + np.arange(num_bins) * delta_log_lambda
    stack_delta = np.zeros(num_bins)
    stack_weight = np.zeros(num_bins)
    for healpix in sorted(data):
        for forest in data[healpix]:
            bins = ((forest.log_lambda - log_lambda_min) / delta_log_lambda + 0.5).astype(int)
            rebin = np.bincount(bins, weights=forest.delta * forest.weights)
            stack_delta[:len(rebin)] += rebin
            rebin = np.bincount(bins, weights=forest.weights)
            stack_weight[:len(rebin)] += rebin
    w = stack_weight > 0
    stack_delta[w] /= stack_weight[w]
    return stack_log_lambda, stack_delta, stack_weight


def remove_mean_lambda_obs(data, z_min, z_max, lambda_abs):
    """Subtract the observed-frame mean delta from every delta of ``data`` in place."""
    delta_log_lambda = None
    for healpix in data:
        for delta in data[healpix]:
            step = np.diff(delta.log_lambda).min()
            delta_log_lambda = step if delta_log_lambda is None else min(step, delta_log_lambda)
    log_lambda_min = np.log10((z_min + 1.0) * lambda_abs) - delta_log_lambda / 2.0
    log_lambda_max = np.log10((z_max + 1.0) * lambda_abs) + delta_log_lambda / 2.0
    _, mean_delta, _ = stack_deltas(data, log_lambda_min, log_lambda_max, delta_log_lambda)
    for healpix in data:
        for delta in data[healpix]:
            bins = ((delta.log_lambda - log_lambda_min) / delta_log_lambda + 0.5).astype(int)
            delta.delta -= mean_delta[bins]
