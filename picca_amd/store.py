"""Structure-of-arrays forest store: the layout the engine keeps resident in HBM.

picca holds every forest as a ``Delta`` object with its own small numpy arrays
(``py/picca/data.py:238-374``).  The engine instead concatenates the per-pixel arrays of
all forests (z, r_comov, dist_m, weights, delta, log_lambda; FP64) with an ``offsets``
index (forest k owns pixels [offsets[k], offsets[k+1])), next to per-forest columns
(z_qso, ra, dec, unit vector, ids, continuum order).  ``ForestStore.from_deltas`` packs
Delta objects; synthetic surveys build the arrays directly.
"""
import numpy as np


class ForestStore:
    """Concatenated forests.  All per-pixel arrays share ``offsets``."""

    PIXEL_FIELDS = ("z", "r_comov", "dist_m", "weights", "delta", "log_lambda")
    FOREST_FIELDS = ("z_qso", "ra", "dec", "x_cart", "y_cart", "z_cart", "cos_dec",
                     "thingid", "plate", "fiberid", "order")

    def __init__(self, offsets, z, r_comov, dist_m, weights, delta, z_qso, ra, dec,
                 thingid, log_lambda=None, plate=None, fiberid=None, order=None,
                 x_cart=None, y_cart=None, z_cart=None, cos_dec=None):
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        n = len(self.offsets) - 1
        self.z = np.ascontiguousarray(z, dtype=np.float64)
        self.r_comov = np.ascontiguousarray(r_comov, dtype=np.float64)
        self.dist_m = np.ascontiguousarray(dist_m, dtype=np.float64)
        self.weights = np.ascontiguousarray(weights, dtype=np.float64)
        self.delta = np.ascontiguousarray(delta, dtype=np.float64)
        self.log_lambda = (None if log_lambda is None
                           else np.ascontiguousarray(log_lambda, dtype=np.float64))
        self.z_qso = np.ascontiguousarray(z_qso, dtype=np.float64)
        self.ra = np.ascontiguousarray(ra, dtype=np.float64)
        self.dec = np.ascontiguousarray(dec, dtype=np.float64)
        # unit vectors exactly as QSO.__init__ (py/picca/data.py:86-89)
        self.x_cart = (np.cos(self.ra) * np.cos(self.dec) if x_cart is None
                       else np.ascontiguousarray(x_cart, dtype=np.float64))
        self.y_cart = (np.sin(self.ra) * np.cos(self.dec) if y_cart is None
                       else np.ascontiguousarray(y_cart, dtype=np.float64))
        self.z_cart = (np.sin(self.dec) if z_cart is None
                       else np.ascontiguousarray(z_cart, dtype=np.float64))
        self.cos_dec = (np.cos(self.dec) if cos_dec is None
                        else np.ascontiguousarray(cos_dec, dtype=np.float64))
        self.thingid = np.ascontiguousarray(thingid, dtype=np.int64)
        self.plate = (np.zeros(n, dtype=np.int64) if plate is None
                      else np.ascontiguousarray(plate, dtype=np.int64))
        self.fiberid = (np.zeros(n, dtype=np.int64) if fiberid is None
                        else np.ascontiguousarray(fiberid, dtype=np.int64))
        self.order = (np.full(n, -1, dtype=np.int32) if order is None
                      else np.ascontiguousarray(order, dtype=np.int32))
        self.version = 0
        for name in self.PIXEL_FIELDS:
            arr = getattr(self, name)
            if arr is not None and arr.shape != (self.offsets[-1],):
                raise ValueError("pixel array %s has shape %s, expected (%d,)"
                                 % (name, arr.shape, self.offsets[-1]))

    @property
    def num_forests(self):
        return len(self.offsets) - 1

    @property
    def num_pixels(self):
        return int(self.offsets[-1])

    def lengths(self):
        return np.diff(self.offsets)

    def nonzero_weight_counts(self):
        """Pixels with w != 0 per forest (the reference skips the others)."""
        nz = (self.weights != 0).astype(np.int64)
        csum = np.concatenate([[0], np.cumsum(nz)])
        return csum[self.offsets[1:]] - csum[self.offsets[:-1]]

    def forest(self, k):
        lo, hi = self.offsets[k], self.offsets[k + 1]
        return {name: getattr(self, name)[lo:hi] for name in self.PIXEL_FIELDS
                if getattr(self, name) is not None}

    @classmethod
    def from_deltas(cls, deltas, angular=False):
        """Pack a sequence of Delta objects (after read_deltas) in the given order.

        With ``angular`` the separation coordinates r_comov and dist_m are both the
        wavelength 10**log_lambda, which is what compute_xi hands its kernel for the
        angular correlation (py/picca/cf.py:186-191)."""
        deltas = list(deltas)
        lengths = np.array([len(d.z) for d in deltas], dtype=np.int64)
        offsets = np.zeros(len(deltas) + 1, dtype=np.int64)
        np.cumsum(lengths, out=offsets[1:])

        def cat(attr):
            if not deltas:
                return np.zeros(0)
            return np.concatenate([np.asarray(getattr(d, attr), dtype=np.float64)
                                   for d in deltas])

        def ids(attr):
            out = []
            for d in deltas:
                value = getattr(d, attr, 0)
                try:
                    out.append(int(value))
                except (TypeError, ValueError):
                    out.append(-1)
            return np.array(out, dtype=np.int64)

        orders = np.array([-1 if d.order is None else int(d.order) for d in deltas],
                          dtype=np.int32)
        has_ll = all(d.log_lambda is not None for d in deltas)
        if angular:
            wave = (np.concatenate([10.0 ** np.asarray(d.log_lambda, dtype=np.float64)
                                    for d in deltas]) if deltas else np.zeros(0))
            r_comov, dist_m = wave, wave
        else:
            r_comov, dist_m = cat("r_comov"), cat("dist_m")
        return cls(offsets, cat("z"), r_comov, dist_m, cat("weights"),
                   cat("delta"),
                   np.array([d.z_qso for d in deltas], dtype=np.float64),
                   np.array([d.ra for d in deltas], dtype=np.float64),
                   np.array([d.dec for d in deltas], dtype=np.float64),
                   ids("thingid"),
                   log_lambda=cat("log_lambda") if has_ll else None,
                   plate=ids("plate"), fiberid=ids("fiberid"), order=orders,
                   x_cart=np.array([d.x_cart for d in deltas], dtype=np.float64),
                   y_cart=np.array([d.y_cart for d in deltas], dtype=np.float64),
                   z_cart=np.array([d.z_cart for d in deltas], dtype=np.float64),
                   cos_dec=np.array([d.cos_dec for d in deltas], dtype=np.float64))
