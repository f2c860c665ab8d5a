"""Small helpers on the correlation path.

* ``userprint`` mirrors ``py/picca/utils.py:20-28`` (print and flush);
* ``compute_ang_max`` mirrors ``py/picca/utils.py:419-450``: the largest angle at which a
  pair can still have r_trans < r_trans_max, 2 arcsin(r_trans_max / (D_M(z1) + D_M(z2)))
  at the lowest pixel redshifts, or pi when that argument exceeds one;
* ``shuffle_distrib_forests`` mirrors ``py/picca/utils.py:453-500``: reassigns sky
  positions among forests with a legacy-seeded permutation.
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
