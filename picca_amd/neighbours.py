"""Neighbour search of ``fill_neighs`` (``py/picca/cf.py:82-135``) on flat arrays.

For every source forest the reference keeps, in this order, the forests of the healpix
pixels returned (ascending) by ``healpy.query_disc(nside, xyz, ang_max, inclusive=True)``
that are in the target catalogue, minus the ones with the same thingid, minus those with
``get_angle_between >= ang_max``, and in auto-correlation mode minus those with
``ra >= ra_source``.  Because the disc query returns a superset of the pixels that hold a
forest within ang_max, the neighbour list is fixed by the exact angle test alone; the
order is (target healpix ascending, position in that healpix list).

Here the candidate superset comes from the engine's native grid search
(``picca_neighbour_candidates``), which also evaluates the dot product with the
reference's operation order.  The arccos, the clamping and the 2-arcsec flat-sky
replacement are then applied with numpy exactly as ``QSO.get_angle_between``
(``py/picca/data.py:118-141``), so the angles are the ones the reference computes on the
same machine.
"""
import numpy as np

from . import _native
from .constants import SMALL_ANGLE_CUT_OFF


class NeighbourList:
    """CSR neighbour lists: targets of source s are ``target[offsets[s]:offsets[s+1]]``."""

    def __init__(self, offsets, target, angle):
        self.offsets = offsets
        self.target = target
        self.angle = angle

    @property
    def num_pairs(self):
        return int(self.offsets[-1])

    def source_index(self):
        return np.repeat(np.arange(len(self.offsets) - 1, dtype=np.int64),
                         np.diff(self.offsets))


def _catalog(store, index):
    index = np.ascontiguousarray(index, dtype=np.int64)
    return {
        "x": np.ascontiguousarray(store.x_cart[index]),
        "y": np.ascontiguousarray(store.y_cart[index]),
        "z": np.ascontiguousarray(store.z_cart[index]),
        "ra": np.ascontiguousarray(store.ra[index]),
        "dec": np.ascontiguousarray(store.dec[index]),
        "cos_dec": np.ascontiguousarray(store.cos_dec[index]),
        "id": np.ascontiguousarray(store.thingid[index]),
    }


def find_neighbours(store, source_index, target_index, ang_max, auto, target_store=None):
    """Neighbour lists of ``source_index`` among ``target_index`` (both into ``store``).

    ``target_index`` must already be in the reference's neighbour order (healpix
    ascending, then file order).  Returns a NeighbourList whose ``target`` entries are
    positions in ``target_index``.  ``target_store`` (any object with the same position
    and id arrays, e.g. a quasar catalogue) holds the targets when they are not forests of
    ``store``.
    """
    lib = _native.load()
    src = _catalog(store, source_index)
    tgt = _catalog(store if target_store is None else target_store, target_index)
    n_src = len(src["x"])
    n_tgt = len(tgt["x"])
    ang_max = float(ang_max)
    # conservative dot-product threshold: arccos is monotone, so every pair with
    # arccos(cos) < ang_max has cos > cos(ang_max) minus a few ulp
    cos_min = np.cos(ang_max) - 1e-12
    chord = np.sqrt(max(2.0 - 2.0 * cos_min, 0.0)) + 1e-9
    chord = min(chord, 2.0)
    counts = np.zeros(n_src, dtype=np.int64)

    def call(offsets, out_target, out_cos):
        status = lib.picca_neighbour_candidates(
            n_src, _native.ptr(src["x"]), _native.ptr(src["y"]), _native.ptr(src["z"]),
            _native.ptr(src["ra"]), _native.ptr(src["dec"]), _native.ptr(src["id"]),
            n_tgt, _native.ptr(tgt["x"]), _native.ptr(tgt["y"]), _native.ptr(tgt["z"]),
            _native.ptr(tgt["ra"]), _native.ptr(tgt["dec"]), _native.ptr(tgt["id"]),
            float(cos_min), float(chord), float(SMALL_ANGLE_CUT_OFF), int(bool(auto)),
            _native.ptr(counts), _native.ptr(offsets), _native.ptr(out_target),
            _native.ptr(out_cos))
        _native.check(status, "picca_neighbour_candidates")

    call(None, None, None)
    offsets = np.zeros(n_src + 1, dtype=np.int64)
    np.cumsum(counts, out=offsets[1:])
    total = int(offsets[-1])
    cand = np.zeros(total, dtype=np.int64)
    cosv = np.zeros(total, dtype=np.float64)
    call(offsets, cand, cosv)

    # QSO.get_angle_between on the candidate lists
    src_of = np.repeat(np.arange(n_src, dtype=np.int64), counts)
    cosv[cosv >= 1.0] = 1.0
    cosv[cosv <= -1.0] = -1.0
    angle = np.arccos(cosv)
    ra1 = src["ra"][src_of]
    dec1 = src["dec"][src_of]
    ra2 = tgt["ra"][cand]
    dec2 = tgt["dec"][cand]
    close = ((np.absolute(ra2 - ra1) < SMALL_ANGLE_CUT_OFF)
             & (np.absolute(dec2 - dec1) < SMALL_ANGLE_CUT_OFF))
    if close.any():
        angle[close] = np.sqrt((dec2[close] - dec1[close]) ** 2
                               + (src["cos_dec"][src_of[close]] * (ra2[close] - ra1[close])) ** 2)
    keep = angle < ang_max
    kept_counts = np.bincount(src_of[keep], minlength=n_src).astype(np.int64)
    out_offsets = np.zeros(n_src + 1, dtype=np.int64)
    np.cumsum(kept_counts, out=out_offsets[1:])
    return NeighbourList(out_offsets, cand[keep], angle[keep])
