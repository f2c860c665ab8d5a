"""The file-format edges of the correlation drivers: rebinned image deltas and blinding.

* picca's test_cf_image_data_rebin (py/picca/tests/test_3_cor.py:288-316): Delta_LYA_image
  with --rebin-factor 3 against cf_image_rebinned.fits.gz, with picca's own comparison
  (exact NB and HEALPID, allclose rtol 1e-5 / atol 1e-8 elsewhere).  The fixture's LAMBDA
  header holds neither WAVE_SOLUTION nor DELTA_LAMBDA, which the current reader needs
  (py/picca/io.py:369-373; picca itself stops there with a KeyError on these files), so
  the test supplies the fixture's 0.8 A grid to Delta.rebin (py/picca/data.py:657-686)
  exactly where the reader would have read it.
* Blinding: the Delta_LYA files rewritten by this package's FITS writer with the delta
  column named DELTA_BLIND and a BLINDING card (py/picca/data.py from_fitsio, blinded
  branch) go through picca_cf and picca_dmat; the products carry DA_BLIND / DM_BLIND
  (py/picca/bin/picca_cf.py:571-573, py/picca/bin/picca_dmat.py) and the BLINDING card, and
  their values equal those of the unblinded run.
"""
import os

import numpy as np
import pytest

import cor_cases
from picca_amd.fitsio_lite import FITS, FITSWriter

pytestmark = pytest.mark.gpu


def products(path):
    out, headers = {}, {}
    f = FITS(path)
    for hdu in f.hdus[1:]:
        headers[hdu.get_extname()] = hdu.read_header()
        for col in hdu.get_colnames():
            out[col] = hdu[col][:]
    return out, headers


def test_rebinned_image_deltas_against_golden(tmp_path, monkeypatch):
    from picca_amd import io as pio
    from picca_amd.bin import picca_cf
    original = pio.read_delta_file

    def read_on_the_fixture_grid(filename, z_min_qso=0, z_max_qso=10, rebin_factor=None,
                                 order=None):
        deltas = original(filename, z_min_qso, z_max_qso, None, order)
        if rebin_factor is not None:
            header = FITS(filename)["LAMBDA"].read_header()
            assert "DELTA_LAMBDA" not in header and "WAVE_SOLUTION" not in header
            for delta in deltas:
                delta.rebin(rebin_factor, dwave=0.8)
        return deltas

    monkeypatch.setattr(pio, "read_delta_file", read_on_the_fixture_grid)
    out = str(tmp_path / "cf_image_rebinned.fits.gz")
    picca_cf.main(cor_cases.CF_CASES["cf_image"] + ["--rebin-factor", "3", "--out", out])
    mine, _ = products(out)
    cor_cases.compare_to_golden(mine, cor_cases.golden_columns("cf_image_rebinned"))


def blind_copy(src_dir, dst_dir, blinding="desi_y1"):
    """Every BinTable delta file of src_dir rewritten with DELTA -> DELTA_BLIND and a
    BLINDING card."""
    os.makedirs(dst_dir, exist_ok=True)
    for name in sorted(os.listdir(src_dir)):
        if not name.endswith(".fits.gz"):
            continue
        src = FITS(os.path.join(src_dir, name))
        with FITSWriter(os.path.join(dst_dir, name)) as out:
            for hdu in src.hdus[1:]:
                header = hdu.read_header()
                cols = hdu.get_colnames()
                names = ["DELTA_BLIND" if c == "DELTA" else c for c in cols]
                cards = [{"name": k, "value": header[k]} for k in
                         ("RA", "DEC", "Z", "PMF", "THING_ID", "PLATE", "MJD", "FIBERID",
                          "ORDER") if k in header]
                cards.append({"name": "BLINDING", "value": blinding})
                out.write([hdu[c][:] for c in cols], names, header=cards,
                          extname=hdu.get_extname())


def test_blinded_round_trip(tmp_path):
    from picca_amd.bin import picca_cf, picca_dmat
    blind_dir = str(tmp_path / "Delta_LYA_blind")
    blind_copy(cor_cases.DELTA_LYA, blind_dir)
    for program, case, column in ((picca_cf, cor_cases.CF_CASES["cf"], "DA"),
                                  (picca_dmat, cor_cases.DMAT_CASES["dmat"], "DM")):
        plain_out = str(tmp_path / ("plain_%s.fits.gz" % column))
        blind_out = str(tmp_path / ("blind_%s.fits.gz" % column))
        program.main(case + ["--out", plain_out])
        blind_args = list(case)
        blind_args[blind_args.index("--in-dir") + 1] = blind_dir
        program.main(blind_args + ["--out", blind_out])
        plain, _ = products(plain_out)
        blind, headers = products(blind_out)
        assert column in plain and column not in blind
        assert column + "_BLIND" in blind
        assert any(h.get("BLINDING") == "desi_y1" for h in headers.values())
        np.testing.assert_allclose(blind[column + "_BLIND"], plain[column], rtol=1e-12,
                                   atol=1e-12 * np.abs(plain[column]).max())
        for col in set(plain) - {column}:
            np.testing.assert_allclose(blind[col], plain[col], rtol=1e-12, atol=0, err_msg=col)
