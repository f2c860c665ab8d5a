"""Quasar continua on the GPU: the continuum-fitting step of picca's delta extraction.

``compute_continuum`` keeps the signature and return values of picca's
``compute_continuum`` (py/picca/delta_extraction/expected_fluxes/utils.py:12-136), with
the rest-frame wavelength grid passed in (picca reads it from the ``Forest`` class);
``compute_continua`` does the same for a list of forests in one batched GPU fit, which is
how ``Dr16ExpectedFlux.compute_expected_flux`` uses it
(py/picca/delta_extraction/expected_fluxes/dr16_expected_flux.py:380-464).

The fit itself (``picca_cont_fit``) minimises the reference's chi2 by damped Newton
iterations on the GPU, one wave per forest; it starts where Migrad starts (a0 = weighted
mean flux, b0 = 0) unless start values are given (for example the learned seeds of
``picca_amd.seed_mlp``).

``compute_expected_flux`` runs the iteration of ``Dr16ExpectedFlux.compute_expected_flux``
(py/picca/delta_extraction/expected_fluxes/dr16_expected_flux.py:380-464) on the GPU:
fit every forest with the current mean continuum, then (except after the last iteration)
stack F / C in rest-frame bins with the forest weights of ``compute_forest_weights``
(:466-488), normalise the stack, multiply it into the previous mean continuum
(py/picca/delta_extraction/expected_flux.py:256-308) and interpolate the result back to
every pixel.  eta, var_lss and fudge stay as given, as in the reference with ``use
constant weight`` or ``use ivar as weight`` (which skip ``compute_var_stats``) and in the
Dr16Fixed* classes.
"""
import numpy as np

from . import engine as _engine

BAD_REASONS = {1: "minuit didn't converge", 2: "negative continuum", 3: "zero continuum",
               4: "minuit didn't converge"}

device = 0


def start_values(offsets, flux, ivar):
    """Migrad's start of compute_continuum: a0 = sum f ivar / sum ivar, b0 = 0."""
    seg = np.repeat(np.arange(len(offsets) - 1), np.diff(offsets))
    n = len(offsets) - 1
    a0 = np.bincount(seg, weights=flux * ivar, minlength=n) / np.bincount(
        seg, weights=ivar, minlength=n)
    return a0, np.zeros(n)


def fit_arrays(inputs, order=1, use_constant_weight=False, a0=None, b0=None, max_iter=100,
               tol=1e-12, tol_abs=0.0):
    """Batched fit over flat per-pixel inputs (see ``synth_spectra.forest_inputs``).

    Returns the dict of ``Engine.cont_fit`` (a, b, chi2, ndata, status, evaluations,
    kernel_ms)."""
    offsets = np.asarray(inputs["offsets"], dtype=np.int64)
    n = len(offsets) - 1
    orders = np.broadcast_to(np.asarray(order, dtype=np.int32), (n,)).copy()
    eng = _engine.get_engine(device)
    eng.cont_set_forests(offsets, inputs["flux"], inputs["ivar"], inputs["mean_cont"],
                         inputs["x"], inputs["eta"], inputs["var_lss"], inputs["fudge"], orders)
    if a0 is None or b0 is None:
        s0, t0 = start_values(offsets, inputs["flux"], inputs["ivar"])
        a0 = s0 if a0 is None else a0
        b0 = t0 if b0 is None else b0
    return eng.cont_fit(a0, b0, max_iter=max_iter, tol=tol,
                        use_constant_weight=use_constant_weight, tol_abs=tol_abs)


def _forest_arrays(forest, get_mean_cont, get_eta, get_var_lss, get_fudge, rest_frame_grid):
    """Per-pixel inputs of one forest (py/picca/delta_extraction/expected_fluxes/
    utils.py:59-86)."""
    log_lambda = np.asarray(forest.log_lambda, dtype=np.float64)
    mean_cont = np.asarray(get_mean_cont(log_lambda - np.log10(1 + forest.z)), dtype=np.float64)
    mean_cont = mean_cont * forest.transmission_correction
    ll_max = rest_frame_grid[-1] + np.log10(1 + forest.z)
    ll_min = rest_frame_grid[0] + np.log10(1 + forest.z)
    x = (log_lambda - ll_min) / (ll_max - ll_min)
    return (log_lambda, np.asarray(forest.flux, dtype=np.float64),
            np.asarray(forest.ivar, dtype=np.float64), mean_cont, x,
            np.broadcast_to(get_eta(log_lambda), log_lambda.shape).astype(np.float64),
            np.broadcast_to(get_var_lss(log_lambda), log_lambda.shape).astype(np.float64),
            np.broadcast_to(get_fudge(log_lambda), log_lambda.shape).astype(np.float64))


def compute_continua(forests, get_mean_cont, get_eta, get_var_lss, get_fudge,
                     use_constant_weight, order, rest_frame_grid, seeds=None):
    """compute_continuum for every forest of the list in one GPU fit.

    Returns a list of (cont_model, bad_continuum_reason, continuum_fit_parameters)."""
    parts = [_forest_arrays(f, get_mean_cont, get_eta, get_var_lss, get_fudge, rest_frame_grid)
             for f in forests]
    lengths = np.array([len(p[0]) for p in parts], dtype=np.int64)
    offsets = np.zeros(len(parts) + 1, dtype=np.int64)
    np.cumsum(lengths, out=offsets[1:])

    def cat(k):
        return np.concatenate([p[k] for p in parts]) if parts else np.zeros(0)

    inputs = {"offsets": offsets, "flux": cat(1), "ivar": cat(2), "mean_cont": cat(3),
              "x": cat(4), "eta": cat(5), "var_lss": cat(6), "fudge": cat(7)}
    a0 = b0 = None
    if seeds is not None:
        a0, b0 = seeds
    res = fit_arrays(inputs, order=order, use_constant_weight=use_constant_weight, a0=a0, b0=b0)
    results = []
    for k, p in enumerate(parts):
        a, b = res["a"][k], res["b"][k]
        cont_model = (b * p[4] + a) * p[3]
        reason = BAD_REASONS.get(int(res["status"][k]))
        if reason is None:
            results.append((cont_model, None, (a, b, res["chi2"][k], int(res["ndata"][k]))))
        else:
            results.append((None, reason, (np.nan, np.nan, np.nan, int(res["ndata"][k]))))
    return results


def compute_continuum(forest, get_mean_cont, get_eta, get_var_lss, get_fudge,
                      use_constant_weight, order, rest_frame_grid):
    """picca's compute_continuum for one forest (a batch of one on the GPU)."""
    return compute_continua([forest], get_mean_cont, get_eta, get_var_lss, get_fudge,
                            use_constant_weight, order, rest_frame_grid)[0]


def rest_frame_grid_log(lambda_min, lambda_max, step):
    """Forest.log_lambda_rest_frame_grid of the log wavelength solution
    (py/picca/delta_extraction/astronomical_objects/forest.py:729-732)."""
    return np.arange(np.log10(lambda_min), np.log10(lambda_max) + step / 2, step)


def find_bins_log(log_lambda_rest, grid):
    """find_bins_log (py/picca/delta_extraction/utils.py:162-186)."""
    step = grid[1] - grid[0]
    return ((log_lambda_rest - grid[0]) / step + 0.5).astype(np.int64)


def compute_expected_flux(inputs, grid, num_iterations, order=1, use_constant_weight=False,
                          transmission=None, bins_of=find_bins_log, max_iter=100, tol=1e-12,
                          tol_abs=0.0, seeds=None):
    """The DR16 iteration over the flat forest inputs of ``synth_spectra.forest_inputs``
    (per pixel: log_lambda, flux, ivar, x, eta, var_lss, fudge; per forest: z_qso).

    Returns a dict: ``fit`` (the last iteration's fit, as ``fit_arrays``), ``x`` / ``y``
    (the mean continuum at the kept rest-frame grid points), ``history`` (the mean
    continuum after each re-stack) and ``kernel_ms`` (device time of every fit, stack and
    interpolation).  ``seeds(iteration, inputs)`` may return start values (a0, b0) for an
    iteration's fit."""
    offsets = np.asarray(inputs["offsets"], dtype=np.int64)
    n = len(offsets) - 1
    seg = np.repeat(np.arange(n), np.diff(offsets))
    ll_rest = np.asarray(inputs["log_lambda"], dtype=np.float64) - np.log10(
        1 + np.asarray(inputs["z_qso"], dtype=np.float64))[seg]
    bins = np.asarray(bins_of(ll_rest, grid), dtype=np.int64)
    if bins.size and (bins.min() < 0 or bins.max() >= grid.size):
        raise ValueError("pixels outside the rest-frame grid")
    trans = np.ones_like(ll_rest) if transmission is None else np.asarray(transmission, float)
    orders = np.broadcast_to(np.asarray(order, dtype=np.int32), (n,)).copy()
    eng = _engine.get_engine(device)
    eng.cont_set_forests(offsets, inputs["flux"], inputs["ivar"], np.ones_like(ll_rest),
                         inputs["x"], inputs["eta"], inputs["var_lss"], inputs["fudge"], orders)
    eng.cont_set_rest_frame(ll_rest, trans, bins, grid.size)
    x_keep, y_keep = np.asarray(grid, dtype=np.float64), np.ones(grid.size)
    history, kernel_ms = [], 0.0
    res = None
    plain_start = None
    for iteration in range(num_iterations):
        kernel_ms += eng.cont_set_mean_cont(x_keep, y_keep)
        seed = seeds(iteration, inputs) if seeds is not None else None
        if seed is not None:
            a0, b0 = seed
        else:
            # picca's start values depend on flux and ivar only: the same every iteration
            if plain_start is None:
                plain_start = start_values(offsets, inputs["flux"], inputs["ivar"])
            a0, b0 = plain_start
        res = eng.cont_fit(a0, b0, max_iter=max_iter, tol=tol,
                           use_constant_weight=use_constant_weight, tol_abs=tol_abs)
        kernel_ms += res["kernel_ms"]
        if iteration < num_iterations - 1:
            total, weight, ms = eng.cont_stack(grid.size)
            kernel_ms += ms
            keep = weight > 0
            stack = np.zeros(grid.size)
            stack[keep] = total[keep] / weight[keep]
            stack /= stack[keep].mean()
            previous = _interp1d(x_keep, y_keep, grid)
            x_keep, y_keep = grid[keep], previous[keep] * stack[keep]
            history.append((x_keep, y_keep))
    return {"fit": res, "x": x_keep, "y": y_keep, "history": history, "kernel_ms": kernel_ms}


def _interp1d(x, y, v):
    """scipy's interp1d(x, y, fill_value="extrapolate") at v (its _call_linear)."""
    idx = np.clip(np.searchsorted(x, v), 1, len(x) - 1)
    lo, hi = idx - 1, idx
    slope = (y[hi] - y[lo]) / (x[hi] - x[lo])
    return slope * (v - x[lo]) + y[lo]
