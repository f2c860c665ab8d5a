"""Array-level front end: whole catalogues in one launch, without Delta objects.

``picca_cf`` runs ``fill_neighs([hp]); compute_xi([hp])`` once per healpix
(``py/picca/bin/picca_cf.py:21-37,452-458``).  ``xi_pairs`` builds the equivalent pair
list for every healpix of a ``ForestStore`` at once (rows = healpix ascending, delta1 in
store order, neighbours in the reference's order), and ``run_xi`` evaluates it on the
engine.
"""
import numpy as np

from . import _native
from . import engine as _engine
from .neighbours import find_neighbours


class XiPairs:
    def __init__(self, pair_a, pair_b, pair_ang, pair_flags, pair_row, healpixs):
        self.pair_a = pair_a
        self.pair_b = pair_b
        self.pair_ang = pair_ang
        self.pair_flags = pair_flags
        self.pair_row = pair_row
        self.healpixs = healpixs

    @property
    def num_pairs(self):
        return len(self.pair_a)

    @property
    def num_rows(self):
        return len(self.healpixs)

    def evaluated_pixel_pairs(self, store):
        """Pixel pairs with w1 != 0 and w2 != 0 over all forest pairs: the iterations of
        the reference's j loop body (py/picca/cf.py:330-387)."""
        nz = store.nonzero_weight_counts()
        return int(np.dot(nz[self.pair_a].astype(np.float64), nz[self.pair_b]))


def xi_pairs(store, healpix, ang_max, target_store_index=None, target_healpix=None,
             auto=True, remove_same_half_plate_close_pairs=False):
    """Pair list of an auto (or, with targets, cross) correlation over all healpixs."""
    healpix = np.asarray(healpix)
    order = np.argsort(healpix, kind="stable")
    uniq, starts, counts = np.unique(healpix[order], return_index=True, return_counts=True)
    row_of_src = np.repeat(np.arange(len(uniq), dtype=np.int32), counts)
    src = order.astype(np.int64)
    if target_store_index is None:
        tgt = src
    else:
        t_order = np.argsort(np.asarray(target_healpix), kind="stable")
        tgt = np.asarray(target_store_index, dtype=np.int64)[t_order]
    nl = find_neighbours(store, src, tgt, ang_max, auto=auto)
    s_of = nl.source_index()
    pair_a = src[s_of]
    pair_b = tgt[nl.target]
    if remove_same_half_plate_close_pairs:
        pa, pb = store.plate[pair_a], store.plate[pair_b]
        fa, fb = store.fiberid[pair_a], store.fiberid[pair_b]
        flags = ((pa == pb) & (((fa <= 500) & (fb <= 500)) | ((fa > 500) & (fb > 500))))
        flags = flags.astype(np.uint8)
    else:
        flags = np.zeros(len(pair_a), dtype=np.uint8)
    return XiPairs(pair_a, pair_b, nl.angle, flags, row_of_src[s_of], uniq)


def run_xi(store, pairs, params, device=0, store_key=None, pairs_key=None, schedule=None):
    """Upload (if needed) and run; returns an engine.XiResult with [rows, nb] sums.

    schedule: None keeps the engine's (AUTO unless set), else one of
    _native.XI_SCHEDULE_AUTO / _PAIRS / _RUNS."""
    eng = _engine.get_engine(device)
    eng.set_forests(store, key=store_key)
    eng.set_xi_pairs(pairs.pair_a, pairs.pair_b, pairs.pair_ang, pairs.pair_flags,
                     pairs.pair_row, pairs.num_rows, key=pairs_key)
    if schedule is None:
        return eng.xi(params)
    eng.set_xi_schedule(schedule)
    try:
        return eng.xi(params)
    finally:
        eng.set_xi_schedule(_native.XI_SCHEDULE_AUTO)
