"""CPU oracle of the quasar-continuum fit -- TEST INFRASTRUCTURE ONLY.

Only tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` use this
module, and only as the checker; the engine never imports it.

FP64 numpy restatement of picca's continuum fit for one forest:

* ``continuum_model``  -- ``LeastsSquaresContModel.get_continuum_model``
  (py/picca/delta_extraction/least_squares/least_squares_cont_model.py:97-139):
  C = (b (L - Lmin) / (Lmax - Lmin) + a) * mean_cont;
* ``continuum_weights`` -- ``get_continuum_weights`` (same file :141-192): 0 where
  ivar == 0, 1 with use_constant_weight, else 1 / C^2 / (eta var_pipe + var_lss +
  fudge / var_pipe) with var_pipe = 1 / ivar / C^2;
* ``chi2`` -- ``LeastsSquaresContModel.__call__`` (:65-95):
  sum (f - C)^2 w - sum_{w > 0} log w;
* ``fit`` -- the minimisation of ``compute_continuum``
  (py/picca/delta_extraction/expected_fluxes/utils.py:92-120): start a0 = sum f ivar /
  sum ivar, b0 = 0, slope fixed when order == 0, then the reference's validity checks
  (negative or all-zero continuum).  The reference minimises with iminuit's Migrad, which
  is not installed here; this oracle uses scipy's Nelder-Mead simplex to full precision
  followed by Newton polishing on finite-difference derivatives, so it finds the same
  minimum without sharing code with the GPU fit.  Parity is claimed on the objective:
  the minimum chi2 and the (a, b) that reach it.
* ``stack_mean_cont`` -- ``ExpectedFlux._compute_mean_cont``
  (py/picca/delta_extraction/expected_flux.py:256-308) with the weights of
  ``Dr16ExpectedFlux.compute_forest_weights``
  (py/picca/delta_extraction/expected_fluxes/dr16_expected_flux.py:466-488) and the
  bins of ``find_bins_log`` / ``find_bins_lin`` (py/picca/delta_extraction/utils.py:139-186).
* ``rest_frame_grid_log`` -- ``Forest.log_lambda_rest_frame_grid`` of the log wavelength
  solution (py/picca/delta_extraction/astronomical_objects/forest.py:729-732).
* ``iterate`` -- the loop of ``Dr16ExpectedFlux.compute_expected_flux``
  (py/picca/delta_extraction/expected_fluxes/dr16_expected_flux.py:380-464): fit every
  forest with the current mean continuum, then, except after the last iteration, re-stack
  the mean continuum and replace the interpolator (scipy ``interp1d`` with
  ``fill_value="extrapolate"``, as ``_compute_mean_cont`` does).  eta, var_lss and fudge
  stay as given: the reference skips ``compute_var_stats`` with ``use constant weight`` or
  ``use ivar as weight``, and the Dr16Fixed* classes hold them.
"""
import numpy as np
from scipy.optimize import minimize


def continuum_model(a, b, log_lambda, log_lambda_min, log_lambda_max, mean_cont):
    line = b * (log_lambda - log_lambda_min) / (log_lambda_max - log_lambda_min) + a
    return line * mean_cont


def continuum_weights(ivar, cont, eta, var_lss, fudge, use_constant_weight):
    w = ivar > 0
    weights = np.empty_like(cont)
    weights[~w] = 0
    if use_constant_weight:
        weights[w] = 1
    else:
        var_pipe = 1.0 / ivar[w] / cont[w] ** 2
        variance = eta[w] * var_pipe + var_lss[w] + fudge[w] / var_pipe
        weights[w] = 1.0 / cont[w] ** 2 / variance
    return weights


def chi2(a, b, spec, use_constant_weight=False):
    cont = continuum_model(a, b, spec["log_lambda"], spec["log_lambda_min"],
                           spec["log_lambda_max"], spec["mean_cont"])
    weights = continuum_weights(spec["ivar"], cont, spec["eta"], spec["var_lss"],
                                spec["fudge"], use_constant_weight)
    w = weights > 0
    return float(((spec["flux"] - cont) ** 2 * weights).sum() - np.log(weights[w]).sum())


def start_point(spec):
    """Migrad's starting point in compute_continuum."""
    a0 = (spec["flux"] * spec["ivar"]).sum() / spec["ivar"].sum()
    return a0, 0.0


def _polish(fun, x, free):
    """Newton steps on central finite differences (only when they lower the objective)."""
    x = np.array(x, dtype=float)
    best = fun(x)
    for _ in range(20):
        h = 1e-6 * np.maximum(1.0, np.abs(x))
        g = np.zeros(2)
        H = np.eye(2)
        for k in range(2):
            if not free[k]:
                continue
            e = np.zeros(2)
            e[k] = h[k]
            fp, fm = fun(x + e), fun(x - e)
            g[k] = (fp - fm) / (2 * h[k])
            H[k, k] = (fp - 2 * best + fm) / h[k] ** 2
        if free[0] and free[1]:
            e0 = np.array([h[0], 0.0])
            e1 = np.array([0.0, h[1]])
            H[0, 1] = H[1, 0] = (fun(x + e0 + e1) - fun(x + e0 - e1) - fun(x - e0 + e1)
                                 + fun(x - e0 - e1)) / (4 * h[0] * h[1])
        try:
            step = -np.linalg.solve(H, g)
        except np.linalg.LinAlgError:
            break
        step[~np.asarray(free)] = 0
        trial = x + step
        val = fun(trial)
        if not val < best:
            break
        x, best = trial, val
    return x, best


def fit(spec, order=1, use_constant_weight=False):
    """Returns (a, b, chi2_min, ndata, valid) of one forest."""
    a0, b0 = start_point(spec)
    free = (True, order != 0)

    def fun(p):
        return chi2(p[0], p[1] if free[1] else 0.0, spec, use_constant_weight)

    scale = max(abs(a0) / 2, 1e-6)
    if free[1]:
        simplex = np.array([[a0, b0], [a0 + scale, b0], [a0, b0 + scale]])
        res = minimize(fun, [a0, b0], method="Nelder-Mead",
                       options={"initial_simplex": simplex, "xatol": 1e-13, "fatol": 1e-13,
                                "maxiter": 20000, "maxfev": 40000})
        x = res.x
    else:
        res = minimize(lambda p: fun([p[0], 0.0]), [a0], method="Nelder-Mead",
                       options={"initial_simplex": np.array([[a0], [a0 + scale]]),
                                "xatol": 1e-13, "fatol": 1e-13, "maxiter": 20000})
        x = np.array([res.x[0], 0.0])
    x, best = _polish(fun, x, free)
    a, b = float(x[0]), float(x[1]) if free[1] else 0.0
    cont = continuum_model(a, b, spec["log_lambda"], spec["log_lambda_min"],
                           spec["log_lambda_max"], spec["mean_cont"])
    ndata = int((continuum_weights(spec["ivar"], cont, spec["eta"], spec["var_lss"],
                                   spec["fudge"], use_constant_weight) > 0).sum())
    valid = bool(np.all(cont >= 0) and not np.all(cont == 0) and np.isfinite(best))
    return a, b, float(best), ndata, valid


def ols(spec, order=1):
    """use_constant_weight fit in closed form: linear least squares of f on (m, x m) over
    the pixels with ivar > 0 (the objective is then sum (f - C)^2)."""
    w = spec["ivar"] > 0
    m = spec["mean_cont"][w]
    x = (spec["log_lambda"][w] - spec["log_lambda_min"]) / (
        spec["log_lambda_max"] - spec["log_lambda_min"])
    if order == 0:
        A = m[:, None]
    else:
        A = np.column_stack([m, x * m])
    coef, *_ = np.linalg.lstsq(A, spec["flux"][w], rcond=None)
    a = float(coef[0])
    b = float(coef[1]) if order != 0 else 0.0
    return a, b


def find_bins_log(log_lambda_rest, grid):
    step = grid[1] - grid[0]
    return ((log_lambda_rest - grid[0]) / step + 0.5).astype(np.int64)


def find_bins_lin(log_lambda_rest, grid):
    aux = 10 ** grid[0]
    step = 10 ** grid[1] - aux
    return ((10 ** log_lambda_rest - aux) / step + 0.5).astype(np.int64)


def forest_weights(ivar, cont, eta, var_lss, fudge):
    """Dr16ExpectedFlux.compute_forest_weights with var_lss_mod = 1."""
    w = ivar > 0
    weights = np.zeros_like(cont)
    var_pipe = 1.0 / ivar[w] / cont[w] ** 2
    weights[w] = 1.0 / (eta[w] * var_pipe + var_lss[w] + fudge[w] / var_pipe)
    return weights


def stack_mean_cont(specs, conts, grid, bins_of, previous):
    """One update of the mean continuum: returns (log_lambda grid points kept, new mean
    continuum there, stacked weights there).  ``previous`` is the old mean continuum
    evaluated on the grid; forests with ``conts[k] is None`` (bad fits) are skipped."""
    mean_cont = np.zeros_like(grid)
    mean_weight = np.zeros_like(grid)
    for spec, cont in zip(specs, conts):
        if cont is None:
            continue
        bins = bins_of(spec["log_lambda"] - np.log10(1 + spec["z_qso"]), grid)
        weights = forest_weights(spec["ivar"], cont, spec["eta"], spec["var_lss"],
                                 spec["fudge"])
        mean_cont += np.bincount(bins, weights=spec["flux"] / (cont + 1e-16) * weights,
                                 minlength=grid.size)
        mean_weight += np.bincount(bins, weights=weights, minlength=grid.size)
    w = mean_weight > 0
    mean_cont[w] /= mean_weight[w]
    mean_cont /= mean_cont[w].mean()
    return grid[w], previous[w] * mean_cont[w], mean_weight[w]


def rest_frame_grid_log(lambda_min, lambda_max, step):
    return np.arange(np.log10(lambda_min), np.log10(lambda_max) + step / 2, step)


def mean_cont_interpolator(x, y):
    from scipy.interpolate import interp1d
    return interp1d(x, y, fill_value="extrapolate")


def iterate(specs, grid, num_iterations, order=1, use_constant_weight=False,
            bins_of=find_bins_log):
    """Returns (fits of the last iteration [(a, b, chi2, ndata, valid)], kept grid points,
    mean continuum there, history of the mean continuum after each re-stack).  The specs'
    ``mean_cont`` entries are overwritten with the interpolated mean continuum (times
    ``transmission`` when a spec holds one)."""
    x_keep, y_keep = grid, np.ones_like(grid)
    history = []
    fits = []
    for iteration in range(num_iterations):
        get_mean_cont = mean_cont_interpolator(x_keep, y_keep)
        for spec in specs:
            ll_rest = spec["log_lambda"] - np.log10(1 + spec["z_qso"])
            spec["mean_cont"] = get_mean_cont(ll_rest) * spec.get("transmission", 1.0)
        fits = [fit(spec, order=order, use_constant_weight=use_constant_weight)
                for spec in specs]
        if iteration < num_iterations - 1:
            conts = []
            for spec, (a, b, _, _, valid) in zip(specs, fits):
                conts.append(continuum_model(a, b, spec["log_lambda"], spec["log_lambda_min"],
                                             spec["log_lambda_max"], spec["mean_cont"])
                             if valid else None)
            x_keep, y_keep, _ = stack_mean_cont(specs, conts, grid, bins_of,
                                                get_mean_cont(grid))
            history.append((x_keep, y_keep))
    return fits, x_keep, y_keep, history
