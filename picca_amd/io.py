"""Delta-file reading on the correlation path.

Host-side mirror of ``py/picca/io.py``:

* ``find_order`` (:31-112): the continuum polynomial order, from the ``FITORDER`` key of
  ``FIT_METADATA`` in the delta-attributes file (falling back to ``STACK_DELTAS`` and to
  the ``.config.ini`` of the delta run);
* ``read_blinding`` (:297-330);
* ``read_delta_file`` (:333-380): both file layouts, the z_qso window, optional rebinning;
* ``read_deltas`` (:383-512): files in sorted order, ``--nspec`` truncation, healpix
  assignment ``ang2pix(nside, pi/2 - dec, ra)``, then per forest
  z = 10^loglam / lambda_abs - 1, r_comov(z), D_M(z), w *= ((1+z)/(1+z_ref))^(alpha-1)
  and the projection of ``Delta.project``.

The reference decodes files in a process pool; decoding is cheap next to the pair
kernels, so this reads them in order in one process (the result does not depend on it).
"""
import glob
import os
from configparser import ConfigParser

import numpy as np

from .data import Delta
from .fitsio_lite import FITS
from .healpix import ang2pix
from .utils import userprint


def _config_order(in_dir):
    config = ConfigParser()
    config.read(in_dir + "/../.config.ini")
    if "expected flux" in config and "order" in config["expected flux"]:
        order = config["expected flux"].getint("order")
        userprint("WARNING: Found `order` in delta config file, continuing the analysis")
        return order
    userprint("WARNING: `order` not found in delta config file")
    return None


def find_order(in_dir, delta_attributes=None):
    """Order (0, 1 or None) of the log-lambda polynomial of the continuum fit."""
    if delta_attributes is None:
        delta_attributes = in_dir + "/../Log/delta_attributes.fits.gz"
        userprint(f"WARNING: delta_attributes file not given, setting to {delta_attributes}")
    userprint(f"Reading delta attributes from {delta_attributes}")
    try:
        hdul = FITS(delta_attributes)
    except OSError as err:
        userprint(f"WARNING: OSError encountered: {err}")
        order = _config_order(in_dir)
        userprint(f"Setting order={order} for the polynomial used for the continuum fitting")
        return order
    for extname in ("FIT_METADATA", "STACK_DELTAS"):
        if extname in hdul and "FITORDER" in hdul[extname].read_header():
            order = hdul[extname].read_header()["FITORDER"]
            userprint(f"Setting order={order} for the polynomial used for the continuum fitting")
            return order
        userprint(f"WARNING: FITORDER not found in {extname}")
    order = _config_order(in_dir)
    userprint(f"Setting order={order} for the polynomial used for the continuum fitting")
    return order


def _list_files(in_dir):
    in_dir = os.path.expandvars(in_dir)
    if len(in_dir) > 8 and in_dir[-8:] == ".fits.gz":
        files = glob.glob(in_dir)
    elif len(in_dir) > 5 and in_dir[-5:] == ".fits":
        files = glob.glob(in_dir)
    else:
        files = glob.glob(in_dir + "/*.fits") + glob.glob(in_dir + "/*.fits.gz")
    return sorted(files)


def read_blinding(in_dir):
    """Blinding keyword of the first delta file ("none" when absent)."""
    filename = _list_files(in_dir)[0]
    hdul = FITS(filename)
    if "LAMBDA" in hdul:
        return hdul["METADATA"].read_header()["BLINDING"]
    header = hdul[1].read_header()
    return header["BLINDING"] if "BLINDING" in header else "none"


def read_delta_file(filename, z_min_qso=0, z_max_qso=10, rebin_factor=None, order=None):
    """All deltas of one file within the z_qso window."""
    hdul = FITS(filename)
    if "LAMBDA" in hdul:
        deltas = Delta.from_image(hdul, z_min_qso=z_min_qso, z_max_qso=z_max_qso,
                                  order=order)
    else:
        deltas = [Delta.from_fitsio(hdu, order=order) for hdu in hdul[1:]
                  if z_min_qso < hdu.read_header()["Z"] < z_max_qso]
    if rebin_factor is not None:
        card = "LAMBDA" if "LAMBDA" in hdul else 1
        header = hdul[card].read_header()
        if header["WAVE_SOLUTION"] != "lin":
            raise ValueError("Delta rebinning only implemented for linear lambda bins")
        dwave = header["DELTA_LAMBDA"]
        for delta in deltas:
            delta.rebin(rebin_factor, dwave=dwave)
    return deltas


def read_deltas(in_dir, nside, lambda_abs, alpha, z_ref, cosmo, max_num_spec=None,
                no_project=False, nproc=None, rebin_factor=None, z_min_qso=0,
                z_max_qso=10, delta_attributes=None):
    """Read, place and transform all deltas.  Returns (data, num_data, z_min, z_max)."""
    files = _list_files(in_dir)
    if rebin_factor is not None:
        userprint(f"Rebinning deltas by a factor of {rebin_factor}\n")
    order = find_order(in_dir, delta_attributes)

    deltas = []
    num_data = 0
    for filename in files:
        deltas += read_delta_file(filename, z_min_qso, z_max_qso, rebin_factor, order)
        num_data = len(deltas)
        if max_num_spec is not None and num_data > max_num_spec:
            break
    if max_num_spec is not None:
        deltas = deltas[:max_num_spec]
        num_data = len(deltas)

    phi = [delta.ra for delta in deltas]
    theta = [np.pi / 2.0 - delta.dec for delta in deltas]
    healpixs = ang2pix(nside, theta, phi) if deltas else np.zeros(0, dtype=np.int64)
    if healpixs.size == 0:
        raise AssertionError("ERROR: No data in {}".format(in_dir))

    data = {}
    z_min = 10 ** deltas[0].log_lambda[0] / lambda_abs - 1.0
    z_max = 0.0
    for delta, healpix in zip(deltas, healpixs):
        z = 10 ** delta.log_lambda / lambda_abs - 1.0
        z_min = min(z_min, z.min())
        z_max = max(z_max, z.max())
        delta.z = z
        if cosmo is not None:
            delta.r_comov = cosmo.get_r_comov(z)
            delta.dist_m = cosmo.get_dist_m(z)
        delta.weights *= ((1 + z) / (1 + z_ref)) ** (alpha - 1)
        if not no_project:
            delta.project()
        data.setdefault(int(healpix), []).append(delta)
    return data, num_data, z_min, z_max


def read_drq(drq_filename, z_min=0, z_max=10.0, keep_bal=False, bi_max=None, mode="sdss"):
    """Quasars of a DRQ catalogue (``read_drq``, py/picca/io.py:160-294): the selection
    id > 0, ra != dec, ra != 0, dec != 0, z_min <= z < z_max, the BAL cuts, and RA/DEC
    converted to radians.  Returns a dict of column arrays."""
    userprint("Reading catalog from ", drq_filename)
    hdu = FITS(drq_filename)[1]
    names = hdu.get_colnames()
    catalog = {name: np.asarray(hdu[name][:]) for name in names}
    if "desi" in mode and "TARGETID" in catalog:
        obj_id_name = "TARGETID"
        if "TARGET_RA" in catalog:
            catalog["RA"] = catalog.pop("TARGET_RA")
            catalog["DEC"] = catalog.pop("TARGET_DEC")
    else:
        obj_id_name = "THING_ID"
    if "Z" not in catalog:
        if "Z_VI" in catalog:
            catalog["Z"] = catalog.pop("Z_VI")
            userprint("Z not found (new DRQ >= DRQ14 style), using Z_VI (DRQ <= DRQ12)")
        else:
            userprint("ERROR: No valid column for redshift found in ", drq_filename)
            return None
    if mode == "desi_mocks":
        for key in ("RA", "DEC"):
            catalog[key] = catalog[key].astype("float64")
    w = np.ones(len(catalog["Z"]), dtype=bool)
    w &= catalog[obj_id_name] > 0
    w &= catalog["RA"] != catalog["DEC"]
    w &= catalog["RA"] != 0.0
    w &= catalog["DEC"] != 0.0
    w &= catalog["Z"] >= z_min
    w &= catalog["Z"] < z_max
    if not keep_bal and bi_max is None:
        if "BAL_FLAG_VI" in catalog:
            w &= catalog["BAL_FLAG_VI"] == 0
        else:
            userprint("WARNING: BAL_FLAG_VI not found")
    if bi_max is not None:
        if "BI_CIV" not in catalog:
            raise ValueError("--bi-max set but no BI_CIV field in HDU")
        w &= catalog["BI_CIV"] <= bi_max
    index = np.where(w)[0]
    out = {name: values[index] for name, values in catalog.items()}
    out["RA"] = np.radians(out["RA"])
    out["DEC"] = np.radians(out["DEC"])
    out["_id_name"] = obj_id_name
    return out


def read_objects(filename, nside, z_min, z_max, alpha, z_ref, cosmo, mode="sdss",
                 keep_bal=True):
    """Quasar objects by healpix (``read_objects``, py/picca/io.py:515-612): QSO
    records with weights ((1+z)/(1+z_ref))^(alpha-1) and, with a cosmology, r_comov and
    D_M at z_qso.  Returns (objs, z_min of the catalogue)."""
    from .data import QSO
    catalog = read_drq(filename, z_min=z_min, z_max=z_max, keep_bal=keep_bal, mode=mode)
    phi = catalog["RA"]
    theta = np.pi / 2.0 - catalog["DEC"]
    healpixs = ang2pix(nside, theta, phi)
    if healpixs.size == 0:
        raise AssertionError()
    userprint("Reading objects ")
    objs = {}
    for healpix in np.unique(healpixs):
        index = np.where(healpixs == healpix)[0]
        if "TARGETID" in catalog:
            fiber = catalog["FIBER"] if "FIBER" in catalog else catalog["TARGETID"]
            night = None
            for col in ("LAST_NIGHT", "NIGHT"):
                if col in catalog:
                    night = catalog[col]
                    break
            if night is None or mode == "desi_mocks":
                night = catalog["TARGETID"]
            objs[int(healpix)] = [
                QSO(int(catalog["TARGETID"][k]), float(catalog["RA"][k]),
                    float(catalog["DEC"][k]), float(catalog["Z"][k]),
                    int(catalog["TARGETID"][k]), int(night[k]), int(fiber[k]))
                for k in index]
        else:
            objs[int(healpix)] = [
                QSO(int(catalog["THING_ID"][k]), float(catalog["RA"][k]),
                    float(catalog["DEC"][k]), float(catalog["Z"][k]), int(catalog["PLATE"][k]),
                    int(catalog["MJD"][k]), int(catalog["FIBERID"][k]))
                for k in index]
        for obj in objs[int(healpix)]:
            obj.weights = ((1.0 + obj.z_qso) / (1.0 + z_ref)) ** (alpha - 1.0)
            if cosmo is not None:
                obj.r_comov = float(cosmo.get_r_comov(obj.z_qso))
                obj.dist_m = float(cosmo.get_dist_m(obj.z_qso))
    return objs, catalog["Z"].min()
