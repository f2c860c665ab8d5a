"""Synthetic delta catalogues of the benchmark workloads.

Forests are generated directly as flat arrays (a ``ForestStore``), then transformed the
way ``read_deltas`` transforms real deltas (``py/picca/io.py:492-510``): pixel redshift
z = lambda / 1215.67 - 1, r_comov(z) and D_M(z) from the fiducial ``Cosmo``, weights times
((1+z)/(1+z_ref))^(alpha-1) and the order-1 projection of ``Delta.project``
(``py/picca/data.py:622-655``, evaluated with segment sums).

Workloads (all FP64, numpy ``default_rng(seed)``):

* ``survey_small``   1000 forests, RA/Dec ~ U(0, 10 deg), z_qso ~ U(2.1, 3.5), log10 lambda
                     grid with step 3e-4 from 3600 A over rest frame 1040-1200 A,
                     delta ~ N(0, 0.3), w = 1/(sigma^2 + 0.05) with sigma ~ U(0.1, 1),
                     Omega_m = 0.315, nside 16;
* ``survey_100k``    100k forests over 1000 deg^2 (RA/Dec ~ U(0, 31.62 deg)), linear grid
                     lambda_k = 3600 + 0.8 k A, L ~ U{50..1500} pixels ending at the last grid
                     pixel below 1200 (1 + z_qso) (clipped at 3600 A), same delta and w,
                     nside 64;
* ``survey_300k``    as survey_100k with 300k forests and L log-uniform in [50, 1500];
* ``survey_100k_heavy`` as survey_100k with delta = 0.3 t(nu=3) and w log-uniform in
                     [1e-3, 1e2] (precision study).

``scale`` shrinks the sky area and the forest count together (same density), for tests.
"""
import numpy as np

from .constants import ABSORBER_IGM, Cosmo
from .healpix import ang2pix
from .store import ForestStore

LYA = ABSORBER_IGM["LYA"]

WORKLOADS = {
    "survey_small": dict(n=1000, side_deg=10.0, grid="log", nside=16, length="rest",
                         noise="gauss"),
    "survey_100k": dict(n=100_000, side_deg=31.62, grid="lin", nside=64, length="uniform",
                        noise="gauss"),
    "survey_300k": dict(n=300_000, side_deg=54.77, grid="lin", nside=64, length="loguniform",
                        noise="gauss"),
    "survey_100k_heavy": dict(n=100_000, side_deg=31.62, grid="lin", nside=64,
                              length="uniform", noise="heavy"),
}


class Survey:
    """A synthetic catalogue: forest store + per-forest healpix + global pixel z_min."""

    def __init__(self, store, healpix, z_min, nside, name):
        self.store = store
        self.healpix = healpix
        self.z_min = z_min
        self.nside = nside
        self.name = name

    def rows(self):
        """Forest indices grouped by healpix (ascending), each group in store order."""
        order = np.argsort(self.healpix, kind="stable")
        hp_sorted = self.healpix[order]
        uniq, starts = np.unique(hp_sorted, return_index=True)
        bounds = list(starts) + [len(order)]
        return uniq, [order[bounds[k]:bounds[k + 1]] for k in range(len(uniq))]


def _segment_sum(values, offsets):
    csum = np.concatenate([[0.0], np.cumsum(values)])
    return csum[offsets[1:]] - csum[offsets[:-1]]


def make_survey(name="survey_small", seed=0, scale=1.0, z_ref=2.25, alpha=2.9, om=0.315,
                project=True):
    spec = WORKLOADS[name]
    rng = np.random.default_rng(seed)
    n = max(2, int(round(spec["n"] * scale)))
    side = np.radians(spec["side_deg"] * np.sqrt(scale))
    ra = rng.uniform(0.0, side, n)
    dec = rng.uniform(0.0, side, n)
    z_qso = rng.uniform(2.1, 3.5, n)

    if spec["grid"] == "log":
        step = 3e-4
        ll0 = np.log10(3600.0)
        k_lo = np.ceil((np.log10(1040.0 * (1 + z_qso)) - ll0) / step).astype(np.int64)
        k_hi = np.floor((np.log10(1200.0 * (1 + z_qso)) - ll0) / step).astype(np.int64)
        k_lo = np.maximum(k_lo, 0)
        lengths = np.maximum(k_hi - k_lo + 1, 1)
        offsets = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(lengths, out=offsets[1:])
        k = np.repeat(k_lo, lengths) + (np.arange(offsets[-1]) - np.repeat(offsets[:-1], lengths))
        log_lambda = ll0 + k * step
        wave = 10 ** log_lambda
    else:
        dwave = 0.8
        k_end = np.floor((1200.0 * (1 + z_qso) - 3600.0) / dwave).astype(np.int64)
        if spec["length"] == "uniform":
            want = rng.integers(50, 1501, n)
        else:
            want = np.round(np.exp(rng.uniform(np.log(50), np.log(1500), n))).astype(np.int64)
        k_start = np.maximum(k_end - want + 1, 0)
        lengths = k_end - k_start + 1
        offsets = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(lengths, out=offsets[1:])
        k = np.repeat(k_start, lengths) + (np.arange(offsets[-1]) - np.repeat(offsets[:-1], lengths))
        wave = 3600.0 + dwave * k
        log_lambda = np.log10(wave)

    npix = int(offsets[-1])
    if spec["noise"] == "gauss":
        delta = rng.normal(0.0, 0.3, npix)
        sigma = rng.uniform(0.1, 1.0, npix)
        weights = 1.0 / (sigma ** 2 + 0.05)
    else:
        delta = 0.3 * rng.standard_t(3, npix)
        weights = np.exp(rng.uniform(np.log(1e-3), np.log(1e2), npix))

    # read_deltas transforms
    cosmo = Cosmo(Om=om, verbose=False)
    z = 10 ** log_lambda / LYA - 1.0
    r_comov = cosmo.get_r_comov(z)
    dist_m = cosmo.get_dist_m(z)
    weights = weights * ((1 + z) / (1 + z_ref)) ** (alpha - 1)
    if project:
        sw = _segment_sum(weights, offsets)
        seg = np.repeat(np.arange(n), lengths)
        mean_delta = _segment_sum(weights * delta, offsets) / sw
        mean_ll = _segment_sum(weights * log_lambda, offsets) / sw
        dll = log_lambda - mean_ll[seg]
        num = _segment_sum(weights * delta * dll, offsets)
        den = _segment_sum(weights * dll ** 2, offsets)
        slope = np.where(den > 0, num / np.where(den > 0, den, 1.0), 0.0)
        delta = delta - (mean_delta[seg] + slope[seg] * dll)

    store = ForestStore(offsets, z, r_comov, dist_m, weights, delta, z_qso, ra, dec,
                        np.arange(n, dtype=np.int64), log_lambda=log_lambda,
                        plate=np.arange(n, dtype=np.int64), fiberid=np.arange(n, dtype=np.int64),
                        order=np.ones(n, dtype=np.int32))
    healpix = ang2pix(spec["nside"], np.pi / 2 - dec, ra)
    return Survey(store, healpix, float(z.min()), spec["nside"], name)
