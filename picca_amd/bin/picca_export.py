#!/usr/bin/env python3
"""picca_export on the MI355X engine: a correlation with its covariance and distortion
matrix, written as picca's exported COR + DMATTRI file.

Command line and product as ``py/picca/bin/picca_export.py`` for the correlation path.  The
subsampling covariance runs on the GPU (``utils.compute_cov``); ``--cov``, ``--cor``,
bootstrap covariances, smoothing and ``--remove-shuffled-correlation`` are supported.
Multipoles and DESI's blinding templates (site files) are not part of this engine and are
refused.
"""
import sys

import numpy as np
import scipy.linalg

from picca_amd import driver as D
from picca_amd.fitsio_lite import FITS
from picca_amd.utils import compute_cov, compute_cov_boot, smooth_cov

# strategies whose correlations go out unblinded (py/picca/bin/picca_export.py:26)
UNBLINDED = ("none", "desi_m2", "desi_y1", "desi_y3")

# keywords copied from the correlation's ATTRI header when present
PASSED_ON = ("ZMIN", "ZMAX", "OMEGAM", "OMEGAR", "OMEGAK", "WL")


class Correlation:
    """Rows and attributes of a picca_cf / picca_xcf file."""

    def __init__(self, path):
        hdus = FITS(path)
        attri, cor = hdus[1], hdus[2]
        self.head = attri.read_header()
        for col in ("RP", "RT", "Z", "NB"):
            setattr(self, col.lower(), np.array(attri[col][:]))
        self.column = "DA_BLIND" if "DA_BLIND" in cor.get_colnames() else "DA"
        self.xi = np.array(cor[self.column][:])
        self.we = np.array(cor["WE"][:])
        blinding = self.head["BLINDING"] if "BLINDING" in self.head else "none"
        # the retired "minimal" strategy is read as corr_yshift
        self.blinding = "corr_yshift" if blinding == "minimal" else blinding

    def bin_widths(self):
        h = self.head
        return (h["RPMAX"] - h["RPMIN"]) / h["NP"], h["RTMAX"] / h["NT"]

    def subtract_shuffled(self, path):
        """Subtracts, from every row, the weighted mean over bins of the same row of a
        correlation of shuffled forests."""
        cor = FITS(path)["COR"]
        we = cor["WE"][:]
        per_row = (cor[self.column][:] * we).sum(axis=1)
        total = we.sum(axis=1)
        per_row[total > 0] /= total[total > 0]
        self.xi -= per_row[:, None]

    def mean(self):
        total = self.we.sum(axis=0)
        xi = (self.xi * self.we).sum(axis=0)
        xi[total > 0] /= total[total > 0]
        return xi


def covariance_of(opts, corr):
    if opts.cov is not None:
        D.say("covariance from", opts.cov)
        return FITS(opts.cov)[1]["CO"][:]
    if opts.cor is not None:
        D.say("correlation matrix from", opts.cor)
        rho = FITS(opts.cor)[1]["CO"][:]
        diag = np.diagonal(rho)
        if rho.min() < -1 or rho.max() > 1 or np.any(diag != 1.0):
            D.say("warning: the correlation matrix holds values outside [-1, 1] or a "
                  "diagonal that is not 1")
        rho = rho / np.sqrt(diag * diag[:, None])
        var = np.diagonal(compute_cov(corr.xi, corr.we))
        return rho * np.sqrt(var * var[:, None])
    if opts.num_boot_cov > 0:
        D.say("bootstrap covariance, %d realisations" % opts.num_boot_cov)
        cov = compute_cov_boot(corr.xi, corr.we, nboots=opts.num_boot_cov)
    else:
        cov = compute_cov(corr.xi, corr.we)
    if opts.do_not_smooth_cov:
        return cov
    d_par, d_trans = corr.bin_widths()
    return smooth_cov(None, None, corr.rp, corr.rt, delta_r_trans=d_trans, delta_r_par=d_par,
                      covariance=cov, per_r_par=opts.smooth_per_r_par)


def matrix_of(opts, corr, size):
    """(matrix, its column name, model RP, RT, Z): the file's matrix, matching the
    correlation's blinding, or the identity when there is none."""
    if opts.dmat is None:
        return np.eye(size), "DM_EMPTY", corr.rp.copy(), corr.rt.copy(), corr.z.copy()
    hdus = FITS(opts.dmat)
    blind_file = "DM_BLIND" in hdus[1].get_colnames()
    if blind_file != (corr.column == "DA_BLIND"):
        D.say("the correlation and the distortion matrix differ in blinding; they cannot "
              "be exported together")
        sys.exit(1)
    name = "DM_BLIND" if blind_file else "DM"
    dmat = np.array(hdus[1][name][:])
    try:
        model = [np.array(hdus[2][c][:]) for c in ("RP", "RT", "Z")]
    except (IndexError, KeyError):
        model = [corr.rp.copy(), corr.rt.copy(), corr.z.copy()]
    if dmat.shape == (size, size):
        model = [corr.rp.copy(), corr.rt.copy(), corr.z.copy()]
    return (dmat, name, *model)


def main(cmdargs=None):
    opts = D.parse("picca_export", cmdargs)
    if opts.multipoles:
        raise NotImplementedError("--multipoles is not part of this engine")
    corr = Correlation(opts.data)
    rows = corr.xi.shape[0]
    if opts.remove_shuffled_correlation is not None:
        corr.subtract_shuffled(opts.remove_shuffled_correlation)
    cov = covariance_of(opts, corr)
    xi = corr.mean()
    try:
        scipy.linalg.cholesky(cov)
    except scipy.linalg.LinAlgError:
        D.say("warning: the covariance is not positive definite")
    dmat, dmat_name, m_rp, m_rt, m_z = matrix_of(opts, corr, xi.size)

    if corr.blinding not in UNBLINDED:
        raise NotImplementedError("blinding '%s' needs DESI's blinding templates, which are "
                                  "not part of this engine" % corr.blinding)
    if corr.blinding != "none":
        D.say("'%s' correlations go out unblinded" % corr.blinding)
    # unblinded output names its columns DA and DM whatever the inputs were called,
    # the identity stand-in for a missing matrix included
    data_name, dmat_name = "DA", "DM"

    h = corr.head
    cards = D.header(BLINDING="none", RPMIN=h["RPMIN"], RPMAX=h["RPMAX"], RTMAX=h["RTMAX"],
                     NP=h["NP"], NT=h["NT"])
    cards += D.header(**{k: h[k] for k in PASSED_ON if k in h})
    cards += D.header(RMU_BIN=bool("RMU_BIN" in h and h["RMU_BIN"]), NSAMPLES=rows,
                      MLTPOLE=False)
    cor = D.Product("COR", cards)
    cor.add(data_name, xi, like="DA").add("RP", corr.rp).add("RT", corr.rt).add("Z", corr.z)
    cor.add("CO", cov).add(dmat_name, dmat, like="DM").add("NB", corr.nb)
    model = D.Product("DMATTRI").add("DMRP", m_rp).add("DMRT", m_rt).add("DMZ", m_z)
    D.write_products(opts.out, [cor, model])


if __name__ == "__main__":
    D.main_entry(main)
