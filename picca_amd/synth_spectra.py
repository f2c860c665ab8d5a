"""Synthetic quasar spectra for the continuum fit and its learned seed.

Each spectrum is observed on the SDSS log10-lambda grid (step 1e-4 from 3600 A) over the
rest frame 1040-1600 A:

    continuum = A (lambda_rest / 1200)^alpha (1 + sum_l A_l exp(-(lambda_rest - lambda_l)^2
                                                              / (2 sigma_l^2)))
    flux      = continuum exp(-tau) + noise

with alpha ~ N(-1.5, 0.3), A log-uniform in [0.5, 5], Gaussian emission lines at Ly-beta
1025.72, OVI 1031.9 / 1037.6, Ly-alpha 1215.67, NV 1238.8 / 1242.8 and SiIV 1393.8 /
1402.8 A (the absorber table of py/picca/constants.py:266-298) with random amplitudes
and widths of 1500-4000 km/s, a log-normal Ly-alpha optical depth blueward of Ly-alpha
around tau_eff(z) = 0.0023 (1 + z)^3.64, and Gaussian noise with the signal-to-noise of
the spectrum drawn from U(1, 10) (sigma = continuum / SNR).

``Spectra`` holds the flat arrays; ``forest_inputs`` cuts the Ly-alpha forest
(1040-1200 A rest) and builds the per-pixel inputs of the continuum fit for picca's
first iteration (mean continuum 1, eta 1, var_lss 0.1, fudge 0,
py/picca/delta_extraction/expected_fluxes/dr16_expected_flux.py:37-39);
``binned_inputs`` gives the 256-bin rest-frame spectrum the seed network reads.
"""
import numpy as np

SPEED_LIGHT_KMS = 299792.458
LINES = (1025.72, 1031.9, 1037.6, 1215.67, 1238.8, 1242.8, 1393.8, 1402.8)
LYA = 1215.67
LOG_STEP = 1e-4
LOG_LAMBDA0 = np.log10(3600.0)
REST_MIN, REST_MAX = 1040.0, 1600.0
FOREST_MIN, FOREST_MAX = 1040.0, 1200.0
NUM_BINS = 256


class Spectra:
    def __init__(self, offsets, log_lambda, flux, ivar, z_qso, truth):
        self.offsets = offsets
        self.log_lambda = log_lambda
        self.flux = flux
        self.ivar = ivar
        self.z_qso = z_qso
        self.truth = truth

    @property
    def num_spectra(self):
        return len(self.offsets) - 1


def make_spectra(n, seed=0):
    rng = np.random.default_rng(seed)
    z_qso = rng.uniform(2.1, 3.5, n)
    alpha = rng.normal(-1.5, 0.3, n)
    amp = np.exp(rng.uniform(np.log(0.5), np.log(5.0), n))
    snr = rng.uniform(1.0, 10.0, n)
    line_amp = rng.uniform(0.05, 0.6, (n, len(LINES)))
    line_amp[:, 3] = rng.uniform(1.0, 3.0, n)                  # Ly-alpha
    line_sigma_v = rng.uniform(1500.0, 4000.0, (n, len(LINES)))

    k_lo = np.ceil((np.log10(REST_MIN * (1 + z_qso)) - LOG_LAMBDA0) / LOG_STEP).astype(np.int64)
    k_hi = np.floor((np.log10(REST_MAX * (1 + z_qso)) - LOG_LAMBDA0) / LOG_STEP).astype(np.int64)
    lengths = k_hi - k_lo + 1
    offsets = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lengths, out=offsets[1:])
    npix = int(offsets[-1])
    seg = np.repeat(np.arange(n), lengths)
    k = np.repeat(k_lo, lengths) + (np.arange(npix) - np.repeat(offsets[:-1], lengths))
    log_lambda = LOG_LAMBDA0 + k * LOG_STEP
    lam = 10 ** log_lambda
    lam_rest = lam / (1 + z_qso[seg])

    shape = np.ones(npix)
    for l, centre in enumerate(LINES):
        sigma = centre * line_sigma_v[seg, l] / SPEED_LIGHT_KMS
        shape += line_amp[seg, l] * np.exp(-0.5 * ((lam_rest - centre) / sigma) ** 2)
    continuum = amp[seg] * (lam_rest / 1200.0) ** alpha[seg] * shape

    z_abs = lam / LYA - 1
    tau_eff = 0.0023 * (1 + z_abs) ** 3.64
    sig_g = 0.8
    tau = tau_eff * np.exp(rng.normal(0.0, sig_g, npix) - 0.5 * sig_g ** 2)
    trans = np.where(lam_rest < LYA, np.exp(-tau), 1.0)
    sigma = continuum / snr[seg]
    flux = continuum * trans + sigma * rng.normal(size=npix)
    ivar = 1.0 / sigma ** 2
    truth = {"alpha": alpha, "amp": amp, "snr": snr, "continuum": continuum}
    return Spectra(offsets, log_lambda, flux, ivar, z_qso, truth)


def forest_inputs(spectra):
    """Per-pixel inputs of the continuum fit over the Ly-alpha forest of each spectrum.

    Returns dict(offsets, log_lambda, flux, ivar, mean_cont, x, eta, var_lss, fudge,
    log_lambda_min, log_lambda_max, z_qso)."""
    n = spectra.num_spectra
    seg = np.repeat(np.arange(n), np.diff(spectra.offsets))
    lam_rest = 10 ** spectra.log_lambda / (1 + spectra.z_qso[seg])
    keep = (lam_rest >= FOREST_MIN) & (lam_rest <= FOREST_MAX)
    counts = np.bincount(seg[keep], minlength=n)
    offsets = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=offsets[1:])
    ll = spectra.log_lambda[keep]
    # rest-frame grid ends shifted to the quasar redshift
    # (py/picca/delta_extraction/expected_fluxes/utils.py:73-78)
    ll_min = np.log10(FOREST_MIN) + np.log10(1 + spectra.z_qso)
    ll_max = np.log10(FOREST_MAX) + np.log10(1 + spectra.z_qso)
    s = seg[keep]
    x = (ll - ll_min[s]) / (ll_max[s] - ll_min[s])
    npix = int(offsets[-1])
    return {
        "offsets": offsets, "log_lambda": ll, "flux": spectra.flux[keep],
        "ivar": spectra.ivar[keep], "mean_cont": np.ones(npix), "x": x,
        "eta": np.ones(npix), "var_lss": np.full(npix, 0.1), "fudge": np.zeros(npix),
        "log_lambda_min": ll_min, "log_lambda_max": ll_max, "z_qso": spectra.z_qso,
    }


def forest_spec(inputs, k):
    """Forest k of ``forest_inputs`` as the per-forest dict the oracle reads."""
    o, e = inputs["offsets"][k], inputs["offsets"][k + 1]
    spec = {name: inputs[name][o:e] for name in ("log_lambda", "flux", "ivar", "mean_cont",
                                                  "eta", "var_lss", "fudge")}
    spec["log_lambda_min"] = inputs["log_lambda_min"][k]
    spec["log_lambda_max"] = inputs["log_lambda_max"][k]
    spec["z_qso"] = inputs["z_qso"][k]
    return spec


def binned_inputs(spectra):
    """256 log-spaced rest-frame bins over 1040-1600 A: ivar-weighted mean flux per bin
    divided by the spectrum's ivar-weighted mean flux (scale).  Returns (X [n, 256]
    float32, scale [n])."""
    n = spectra.num_spectra
    seg = np.repeat(np.arange(n), np.diff(spectra.offsets))
    ll_rest = spectra.log_lambda - np.log10(1 + spectra.z_qso[seg])
    lo, hi = np.log10(REST_MIN), np.log10(REST_MAX)
    b = np.clip(((ll_rest - lo) / (hi - lo) * NUM_BINS).astype(np.int64), 0, NUM_BINS - 1)
    idx = seg * NUM_BINS + b
    wsum = np.bincount(idx, weights=spectra.ivar, minlength=n * NUM_BINS)
    fsum = np.bincount(idx, weights=spectra.ivar * spectra.flux, minlength=n * NUM_BINS)
    binned = np.where(wsum > 0, fsum / np.where(wsum > 0, wsum, 1.0), 0.0).reshape(n, NUM_BINS)
    tot_w = np.bincount(seg, weights=spectra.ivar, minlength=n)
    tot_f = np.bincount(seg, weights=spectra.ivar * spectra.flux, minlength=n)
    scale = tot_f / tot_w
    scale = np.where(np.abs(scale) > 1e-12, scale, 1.0)
    return (binned / scale[:, None]).astype(np.float32), scale
