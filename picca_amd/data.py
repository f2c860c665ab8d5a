"""Line-of-sight records: ``QSO`` and ``Delta``.

Host-side mirror of ``py/picca/data.py``:

* ``QSO.__init__`` (:60-104) stores the unit vector (cos ra cos dec, sin ra cos dec,
  sin dec) and ``cos_dec``;
* ``QSO.get_angle_between`` (:106-162) is the angular separation
  arccos(clamp(x1 x2 + y1 y2 + z1 z2)), replaced by the flat-sky form
  sqrt(d_dec^2 + (cos dec1 d_ra)^2) when both |d_ra| and |d_dec| are below 2 arcsec;
* ``Delta.from_fitsio`` (:375-474) and ``Delta.from_image`` (:519-620) decode the two
  delta file layouts (per-forest BINTABLE HDUs, or DESI image HDUs with w <= 0 pixels
  dropped at read time);
* ``Delta.project`` (:622-655) removes the weighted mean and, for order 1, the weighted
  linear trend in log lambda (eqs. 5-6 of du Mas des Bourboux et al. 2020);
* ``Delta.rebin`` (:657-686) merges ``factor`` consecutive linear-wavelength pixels.

The arithmetic is written with the same numpy operations in the same order, because the
angle and the pixel distances feed bin assignment, where one ulp can move a pair.
"""
import numpy as np

from . import constants
from .utils import userprint


class QSO:
    """A quasar sight line: position on the sky, redshift and identifiers."""

    def __init__(self, los_id, ra, dec, z_qso, plate, mjd, fiberid):
        self.ra = ra
        self.dec = dec
        self.plate = plate
        self.mjd = mjd
        self.fiberid = fiberid
        self.x_cart = np.cos(ra) * np.cos(dec)
        self.y_cart = np.sin(ra) * np.cos(dec)
        self.z_cart = np.sin(dec)
        self.cos_dec = np.cos(dec)
        self.z_qso = z_qso
        self.los_id = los_id
        self.thingid = los_id
        self.weight = None
        self.r_comov = None
        self.dist_m = None
        self.log_lambda = None

    def get_angle_between(self, data):
        """Angular separation (radians) to one QSO or to a sequence of them."""
        if isinstance(data, QSO):
            return angle_between(self.x_cart, self.y_cart, self.z_cart, self.ra, self.dec,
                                 self.cos_dec, data.x_cart, data.y_cart, data.z_cart,
                                 data.ra, data.dec, scalar=True)
        others = list(data)
        x_cart = np.array([d.x_cart for d in others])
        y_cart = np.array([d.y_cart for d in others])
        z_cart = np.array([d.z_cart for d in others])
        ra = np.array([d.ra for d in others])
        dec = np.array([d.dec for d in others])
        return angle_between(self.x_cart, self.y_cart, self.z_cart, self.ra, self.dec,
                             self.cos_dec, x_cart, y_cart, z_cart, ra, dec)


def angle_between(x1, y1, z1, ra1, dec1, cos_dec1, x2, y2, z2, ra2, dec2, scalar=False):
    """Angle between sight line 1 and sight line(s) 2 (``py/picca/data.py:118-161``).

    The dot product is evaluated as ``x2*x1 + y2*y1 + z2*z1`` left to right; values
    outside [-1, 1] are clamped (with the reference's warning); pairs closer than
    ``SMALL_ANGLE_CUT_OFF`` in both ra and dec use the flat-sky formula.
    """
    if scalar:
        cos = x2 * x1 + y2 * y1 + z2 * z1
        if cos >= 1.0:
            userprint("WARNING: 1 pair has cosinus>=1.")
            cos = 1.0
        elif cos <= -1.0:
            userprint("WARNING: 1 pair has cosinus<=-1.")
            cos = -1.0
        angle = np.arccos(cos)
        if (np.absolute(ra2 - ra1) < constants.SMALL_ANGLE_CUT_OFF) & (
                np.absolute(dec2 - dec1) < constants.SMALL_ANGLE_CUT_OFF):
            angle = np.sqrt((dec2 - dec1) ** 2 + (cos_dec1 * (ra2 - ra1)) ** 2)
        return angle
    cos = x2 * x1 + y2 * y1 + z2 * z1
    high = cos >= 1.0
    if high.sum() != 0:
        userprint("WARNING: {} pairs have cos>=1.".format(high.sum()))
        cos[high] = 1.0
    low = cos <= -1.0
    if low.sum() != 0:
        userprint("WARNING: {} pairs have cos<=-1.".format(low.sum()))
        cos[low] = -1.0
    angle = np.arccos(cos)
    close = ((np.absolute(ra2 - ra1) < constants.SMALL_ANGLE_CUT_OFF)
             & (np.absolute(dec2 - dec1) < constants.SMALL_ANGLE_CUT_OFF))
    if close.sum() != 0:
        angle[close] = np.sqrt((dec2[close] - dec1) ** 2
                               + (cos_dec1 * (ra2[close] - ra1)) ** 2)
    return angle


class Delta(QSO):
    """The transmitted-flux fluctuation field of one sight line."""

    def __init__(self, los_id, ra, dec, z_qso, plate, mjd, fiberid, log_lambda,
                 weights, cont, delta, order, ivar=None, exposures_diff=None,
                 mean_snr=None, mean_reso=None, mean_z=None, resolution_matrix=None,
                 mean_resolution_matrix=None, mean_reso_pix=None):
        QSO.__init__(self, los_id, ra, dec, z_qso, plate, mjd, fiberid)
        self.log_lambda = log_lambda
        self.weights = weights
        self.cont = cont
        self.delta = delta
        self.order = order
        self.ivar = ivar
        self.exposures_diff = exposures_diff
        self.mean_snr = mean_snr
        self.mean_reso = mean_reso
        self.mean_z = mean_z
        self.resolution_matrix = resolution_matrix
        self.mean_resolution_matrix = mean_resolution_matrix
        self.mean_reso_pix = mean_reso_pix
        self.z = None
        self.r_comov = None
        self.dist_m = None
        self.neighbours = None
        self.fname = None

    @classmethod
    def from_fitsio(cls, hdu, pk1d_type=False, order=None):
        """One forest from a BINTABLE HDU (``py/picca/data.py:375-474``)."""
        if pk1d_type:
            raise ValueError("Pk1D delta files are outside the correlation path")
        header = hdu.read_header()
        blinding = header["BLINDING"] if "BLINDING" in header else "none"
        delta_name = "DELTA_BLIND" if blinding != "none" else "DELTA"
        delta = hdu[delta_name][:].astype(float)
        colnames = hdu.get_colnames()
        if "LOGLAM" in colnames:
            log_lambda = hdu["LOGLAM"][:].astype(float)
        elif "LAMBDA" in colnames:
            log_lambda = np.log10(hdu["LAMBDA"][:].astype(float))
        else:
            raise KeyError("Did not find LOGLAM or LAMBDA in delta file")
        weights = hdu["WEIGHT"][:].astype(float)
        cont = hdu["CONT"][:].astype(float)
        if "THING_ID" in header:
            los_id = header["THING_ID"]
            plate = header["PLATE"]
            mjd = header["MJD"]
            fiberid = header["FIBERID"]
        elif "LOS_ID" in header:
            los_id = header["LOS_ID"]
            plate = mjd = fiberid = los_id
        else:
            raise Exception("Could not find THING_ID or LOS_ID")
        return cls(los_id, header["RA"], header["DEC"], header["Z"], plate, mjd, fiberid,
                   log_lambda, weights, cont, delta, order)

    @classmethod
    def from_image(cls, hdul, pk1d_type=False, z_min_qso=0, z_max_qso=10, order=None):
        """All forests of a DESI image-format file (``py/picca/data.py:519-620``)."""
        if pk1d_type:
            raise ValueError("ImageHDU format not implemented for Pk1D forests.")
        meta = hdul["METADATA"]
        header = meta.read_header()
        blinding = header["BLINDING"] if "BLINDING" in header else "none"
        delta_name = "DELTA_BLIND" if blinding != "none" else "DELTA"
        delta = hdul[delta_name].read().astype(float)
        if "LOGLAM" in hdul:
            log_lambda = hdul["LOGLAM"].read().astype(float)
        elif "LAMBDA" in hdul:
            log_lambda = np.log10(hdul["LAMBDA"].read().astype(float))
        else:
            raise KeyError("Did not find LOGLAM or LAMBDA in delta file")
        weights = hdul["WEIGHT"].read().astype(float)
        cont = hdul["CONT"].read().astype(float)
        keep = weights > 0
        colnames = meta.get_colnames()
        if "THING_ID" in colnames:
            los_id = meta["THING_ID"][:]
            plate = meta["PLATE"][:]
            mjd = meta["MJD"][:]
            fiberid = meta["FIBERID"][:]
        elif "LOS_ID" in colnames:
            los_id = meta["LOS_ID"][:]
            plate = mjd = fiberid = los_id
        else:
            raise Exception("Could not find THING_ID or LOS_ID")
        ra = meta["RA"][:]
        dec = meta["DEC"][:]
        z_qso = meta["Z"][:]
        deltas = []
        for k in range(len(ra)):
            if z_min_qso <= z_qso[k] <= z_max_qso:
                sel = keep[k]
                deltas.append(cls(los_id[k], ra[k], dec[k], z_qso[k], plate[k], mjd[k],
                                  fiberid[k], log_lambda[sel], weights[k][sel],
                                  cont[k][sel], delta[k][sel], order))
        return deltas

    def project(self):
        """Remove the continuum-fit modes from delta (``py/picca/data.py:622-655``)."""
        if self.order is None:
            raise RuntimeError(
                "Trying to project but order is not defined for the deltas. "
                "Check previous warning to solve this issue")
        sum_weights = np.sum(self.weights)
        if not sum_weights > 0.0:
            return
        mean_delta = np.average(self.delta, weights=self.weights)
        res = 0
        if self.order == 1 and self.delta.shape[0] > 1:
            mean_log_lambda = np.average(self.log_lambda, weights=self.weights)
            meanless_log_lambda = self.log_lambda - mean_log_lambda
            slope = (np.sum(self.weights * self.delta * meanless_log_lambda)
                     / np.sum(self.weights * meanless_log_lambda ** 2))
            res = slope * meanless_log_lambda
        elif self.order == 1:
            res = self.delta
        self.delta -= mean_delta + res

    def rebin(self, factor, dwave=0.8):
        """Merge ``factor`` linear-wavelength pixels (``py/picca/data.py:657-686``)."""
        wave = 10 ** np.array(self.log_lambda)
        start = wave.min() - dwave / 2
        num_bins = np.ceil(((wave[-1] - wave[0]) / dwave + 1) / factor)
        edges = np.arange(num_bins) * dwave * factor + start
        index = np.searchsorted(edges, wave)
        binned_delta = np.bincount(index, weights=self.delta * self.weights,
                                   minlength=edges.size + 1)[1:-1]
        binned_weight = np.bincount(index, weights=self.weights,
                                    minlength=edges.size + 1)[1:-1]
        mask = binned_weight != 0
        binned_delta[mask] /= binned_weight[mask]
        new_wave = (edges[1:] + edges[:-1]) / 2
        self.log_lambda = np.log10(new_wave[mask])
        self.delta = binned_delta[mask]
        self.weights = binned_weight[mask]
