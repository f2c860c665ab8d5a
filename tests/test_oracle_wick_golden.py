"""Pins the Wick-covariance oracle against picca's golden wick file.

Replays test_wick (py/picca/tests/test_3_cor.py:412-441: picca_wick on Delta_LYA with the
golden cf1d, rp/rt < 60, 15 x 15 bins, rej 0.99, one process list) with the oracle:
brute-force neighbours, the legacy stream seeded with the first healpix
(py/picca/bin/picca_wick.py:20-37), one draw per delta of each healpix, the C pair loop of
compute_wickT123_pairs, and the normalisation of py/picca/bin/picca_wick.py:541-560.
"""
import numpy as np

import cor_cases
from oracle import cf_oracle as O

WICK_CASE = ["--in-dir", cor_cases.DELTA_LYA, "--rp-min", "+0.0", "--rp-max", "+60.0",
             "--rt-max", "+60.0", "--np", "15", "--nt", "15", "--rej", "0.99", "--nproc", "1",
             "--in-attributes", cor_cases.ATTRIBUTES]


def normalise(weights_wick, num_pairs_wick, num_pairs, num_pairs_used, t1, t2, t3):
    """py/picca/bin/picca_wick.py:541-560 (diagrams 4-6 are zero at max_diagram 3)."""
    t = [t1.copy(), t2.copy(), t3.copy()]
    weights = weights_wick * weights_wick[:, None]
    w = weights > 0.0
    for x in t:
        x[w] /= weights[w]
        x *= 1.0 * num_pairs_used / num_pairs
    zero = np.zeros_like(t1)
    return {"CO": t[0] + t[1] + t[2] + zero + zero + zero, "WALL": weights_wick,
            "NB": num_pairs_wick, "T1": t[0], "T2": t[1], "T3": t[2], "T4": zero, "T5": zero,
            "T6": zero}


def oracle_wick():
    setup = cor_cases.Setup(WICK_CASE)
    setup.alpha2 = setup.alpha     # picca_wick sets alpha2 = --z-evol (picca_wick.py:333)
    params = O.params_from_module(setup)
    healpixs = sorted(setup.data)
    neigh = O.brute_force_neighbours(setup.data, None, healpixs, setup.ang_max)
    variance, xi = O.read_cf1d(cor_cases.golden_path("cf1d"))
    np.random.seed(healpixs[0])
    return setup, O.compute_wick_terms(params, setup.data, healpixs, neigh, setup.reject,
                                       variance, xi)


def test_oracle_reproduces_golden_wick():
    _, out = oracle_wick()
    cols = normalise(*out)
    golden = cor_cases.FITS(cor_cases.golden_path("wick"))
    header = golden["COV"].read_header()
    assert out[2] == header["NPALL"] and out[3] == header["NPUSED"]
    report = cor_cases.compare_to_golden(cols, cor_cases.golden_columns("wick"))
    assert {"CO", "T1", "T2", "T3", "WALL", "NB"} <= set(report)
    # beyond picca's own tolerance: every covariance entry within 1e-6 of the largest
    ref = cor_cases.golden_columns("wick")
    for name in ("CO", "T1", "T2", "T3"):
        scale = np.abs(ref[name]).max()
        assert scale > 0
        assert np.abs(cols[name] - ref[name]).max() < 1e-6 * scale, name
