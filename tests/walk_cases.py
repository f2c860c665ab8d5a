"""Inputs of the run-walk tests, shared by the CPU harness tests and the GPU tests."""
import numpy as np

import cor_cases
from oracle import cf_oracle as O
from picca_amd import batch, synth
from picca_amd.constants import Cosmo
from picca_amd.store import ForestStore
from picca_amd.utils import compute_ang_max


class Globals:
    def __init__(self, **kw):
        self.num_bins_r_par = 50
        self.num_bins_r_trans = 50
        self.num_model_bins_r_par = 0
        self.num_model_bins_r_trans = 0
        self.r_par_min = 0.0
        self.r_par_max = 200.0
        self.r_trans_max = 200.0
        self.z_min_pairs = None
        self.z_max_pairs = None
        self.zerr_cut_deg = None
        self.zerr_cut_kms = None
        self.z_ref = 2.25
        self.alpha = 2.9
        self.alpha2 = 2.9
        self.x_correlation = False
        self.rmu_binning = False
        self.ang_correlation = False
        self.remove_same_half_plate_close_pairs = False
        self.redshift_evolution_in_distortion_matrix = True
        for k, v in kw.items():
            setattr(self, k, v)


class Case:
    """Globals, forest store and pair list (a, b, ang, row, num_rows) of one test input."""

    def __init__(self, g, store, pair_a, pair_b, pair_ang, pair_row, num_rows):
        self.g = g
        self.store = store
        self.pairs = batch.XiPairs(np.asarray(pair_a), np.asarray(pair_b),
                                   np.asarray(pair_ang, dtype=float),
                                   np.zeros(len(pair_a), dtype=np.uint8),
                                   np.asarray(pair_row, dtype=np.int32),
                                   np.arange(num_rows))


def survey_case(name, scale, r_max=200.0):
    s = synth.make_survey(name, seed=0, scale=scale)
    ang_max = compute_ang_max(Cosmo(0.315, verbose=False), r_max, s.z_min)
    p = batch.xi_pairs(s.store, s.healpix, ang_max)
    return Case(Globals(), s.store, p.pair_a, p.pair_b, p.pair_ang, p.pair_row, p.num_rows)


def golden_case(case):
    """picca's own test deltas with the fixture binning (15 x 15 bins to 60 Mpc/h) and the
    same-half-plate cut dropped, so that the specialised kernel takes every pair."""
    argv = [a for a in cor_cases.CF_CASES[case] if a != "--remove-same-half-plate-close-pairs"]
    setup = cor_cases.Setup(argv)
    hps = sorted(setup.data)
    deltas = [d for hp in hps for d in setup.data[hp]]
    index = {id(d): k for k, d in enumerate(deltas)}
    store = ForestStore.from_deltas(deltas)
    neigh = O.brute_force_neighbours(setup.data, None, hps, setup.ang_max)
    a, b, ang, row = [], [], [], []
    for r, hp in enumerate(hps):
        for d1 in setup.data[hp]:
            for d2 in neigh[id(d1)]:
                a.append(index[id(d1)])
                b.append(index[id(d2)])
                ang.append(O.angle(d1, d2))
                row.append(r)
    return Case(setup, store, a, b, ang, row, len(hps))


def _store(forests):
    lengths = [len(f["r"]) for f in forests]
    offsets = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    cat = lambda key: np.concatenate([np.asarray(f[key], dtype=float) for f in forests])
    n = len(forests)
    return ForestStore(offsets, cat("z"), cat("r"), cat("d"), cat("w"), cat("delta"),
                       np.full(n, 3.0), np.linspace(0.01, 0.0101, n), np.full(n, 0.01),
                       np.arange(n))


def _all_pairs(n, ang):
    """Every ordered pair of n forests, spread over three output rows (contiguous, as the
    engine requires)."""
    a, b = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    keep = a != b
    pa, pb = a[keep], b[keep]
    row = (pa % 3).astype(np.int32)
    order = np.argsort(row, kind="stable")
    return pa[order], pb[order], np.full(len(pa), ang), row[order]


def lattice_case():
    """Pixels on an integer r_comov lattice with cos(ang/2) == 1: r_par is an exact multiple
    of the bin width for many pairs, so run ends fall on exact edges (edge pairs, r_par == 0
    pairs, both sides of r_comov1), and pixels 1e-10 Mpc/h off the lattice are edge pairs on
    either side of an edge."""
    rng = np.random.default_rng(3)
    ang = 2e-8
    assert np.cos(ang / 2) == 1.0
    forests = []
    for n, r0, step in ((400, 3800.0, 1.0), (300, 3850.0, 1.0), (350, 3801.0, 2.0),
                        (200, 3900.5, 0.5), (250, 3820.0 + 1e-10, 1.0),
                        (260, 3820.0 - 1e-10, 1.0)):
        r = r0 + step * np.arange(n)
        forests.append({"r": r, "d": r.copy(), "z": 2.0 + r / 1e4,
                        "w": rng.uniform(0.5, 2, n), "delta": rng.normal(0, 0.3, n)})
    pa, pb, pang, prow = _all_pairs(len(forests), ang)
    return Case(Globals(), _store(forests), pa, pb, pang, prow, 3)


def irregular_case():
    """Zero weights (both signs, at run ends and in runs), forests of 1 and 2 pixels, an
    all-zero forest, gaps, and forests the kernel may not walk (non-monotone, a repeated
    r_comov), mixed in one pair list."""
    rng = np.random.default_rng(7)

    def forest(n, r0=3800.0, dr=0.5):
        r = r0 + np.cumsum(rng.uniform(0.2 * dr, 1.8 * dr, n))
        return {"r": r, "d": r - 3.0, "z": 2.0 + r / 1e4, "w": rng.uniform(0.1, 3, n),
                "delta": rng.normal(0, 0.3, n)}

    f_zero = forest(500)
    f_zero["w"][::3] = 0.0
    f_zero["w"][1::7] = -0.0
    f_one = forest(1, r0=3850.0)
    f_two = forest(2, r0=3860.0)
    f_allzero = forest(80)
    f_allzero["w"][:] = 0.0
    f_gaps = forest(600)
    keep = np.ones(600, bool)
    keep[100:180] = False
    keep[400:404] = False
    f_gaps = {k: v[keep] for k, v in f_gaps.items()}
    f_nonmono = forest(150)
    f_nonmono["r"] = f_nonmono["r"][rng.permutation(150)]
    f_repeat = forest(200)
    f_repeat["r"][50] = f_repeat["r"][49]
    f_repeat["d"][50] = f_repeat["d"][49]
    forests = [f_zero, f_one, f_two, f_allzero, f_gaps, f_nonmono, f_repeat, forest(700),
               forest(300, dr=2.0)]
    pa, pb, pang, prow = _all_pairs(len(forests), 0.004)
    return Case(Globals(), _store(forests), pa, pb, pang, prow, 3)


def random_case(seed):
    """Random binnings, ranges, pixel spacings (0.05-5 Mpc/h), angles up to 0.2 rad and
    weights spanning 1e-3..1e2."""
    rng = np.random.default_rng(100 + seed)
    g = Globals(num_bins_r_par=int(rng.integers(1, 61)), num_bins_r_trans=int(rng.integers(1, 61)),
                r_par_max=float(rng.uniform(5, 250)), r_trans_max=float(rng.uniform(5, 250)))
    forests = []
    for _ in range(int(rng.integers(3, 8))):
        n = int(rng.integers(1, 400))
        step = float(np.exp(rng.uniform(np.log(0.05), np.log(5.0))))
        r = float(rng.uniform(3600, 4200)) + np.cumsum(rng.uniform(0.5 * step, 1.5 * step, n))
        w = np.exp(rng.uniform(np.log(1e-3), np.log(1e2), n))
        w[rng.random(n) < 0.1] = 0.0
        forests.append({"r": r, "d": r * float(rng.uniform(0.9, 1.0)), "z": 2.0 + r / 1e4,
                        "w": w, "delta": 0.3 * rng.standard_t(3, n)})
    pa, pb, _, prow = _all_pairs(len(forests), 0.0)
    pang = rng.uniform(0, 0.2, len(pa))
    pang[rng.random(len(pa)) < 0.2] = rng.uniform(0, 1e-4)
    return Case(g, _store(forests), pa, pb, pang, prow, 3)


def oracle(case):
    """Oracle sums [rows, nb] and stats (accepted, edge, overflow) of a case."""
    p = case.pairs
    return O.xi_rows(O.params_from_module(case.g), case.store, p.pair_a, p.pair_b, p.pair_ang,
                     p.pair_flags, p.pair_row, p.num_rows, num_threads=8)


def assert_matches(out, edge, overflow, ref, ref_stats, rtol):
    """num_pairs bit-exact, edge / overflow counts equal, sums within rtol after
    normalisation (xi against the scale of the row's values: its sums cancel)."""
    assert np.array_equal(out[5], ref[5])
    assert edge == ref_stats[1]
    assert overflow == ref_stats[2]
    we = ref[0]
    np.testing.assert_allclose(out[0], we, rtol=rtol, atol=1e-300)
    mask = we > 0
    for k in range(1, 5):
        mine = out[k][mask] / out[0][mask]
        theirs = ref[k][mask] / we[mask]
        atol = 1e-12 * np.abs(theirs).max() if k == 1 and theirs.size else 0.0
        np.testing.assert_allclose(mine, theirs, rtol=rtol, atol=atol)
