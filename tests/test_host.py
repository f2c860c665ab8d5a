"""CPU tests of the host side: C-ABI library, FITS I/O, HEALPix, cosmology, neighbours."""
import ctypes
import os
import re

import numpy as np
import pytest
from scipy import integrate

import cor_cases
from oracle import cf_oracle as O
from picca_amd import _native, healpix, io, neighbours
from picca_amd.constants import SPEED_LIGHT, Cosmo
from picca_amd.fitsio_lite import FITS, FITSWriter
from picca_amd.store import ForestStore

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "picca_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(picca_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _native.load()
    names = declared_functions()
    assert "picca_xi_execute" in names and "picca_dmat_run" in names
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(names) == sorted(_native.exported_symbols())
    assert lib.picca_abi_version() == _native.ABI_VERSION == 4


def test_library_reports_errors_without_device_work():
    lib = _native.load()
    count = ctypes.c_int32(-1)
    lib.picca_device_count(ctypes.byref(count))  # 0 here without a GPU
    # a NULL context is a value error with a message
    status = lib.picca_xi_execute(None, None, None)
    assert status == _native.PICCA_E_VALUE
    with pytest.raises(ValueError):
        _native.check(status)


def test_fits_roundtrip(tmp_path):
    path = str(tmp_path / "t.fits.gz")
    rng = np.random.default_rng(0)
    a = rng.normal(size=7)
    b = rng.integers(-5, 5, 7).astype(np.int64)
    m = rng.normal(size=(7, 9))
    w = FITSWriter(path)
    w.write([a, b], names=["A", "B"], header=[{"name": "RPMIN", "value": 0.5, "comment": "x"},
                                             {"name": "ZMIN", "value": None, "comment": ""},
                                             {"name": "BLINDING", "value": "none", "comment": ""},
                                             {"name": "RMU_BIN", "value": False, "comment": ""}],
            extname="ATTRI", units=["h^-1 Mpc", ""])
    w.write([b, m], names=["HEALPID", "WE"], extname="COR")
    w.close()
    f = FITS(path)
    assert len(f) == 3
    assert np.array_equal(f["ATTRI"]["A"][:], a)
    assert np.array_equal(f["ATTRI"]["B"][:], b)
    assert np.array_equal(f["COR"]["WE"][:], m)
    hdr = f["ATTRI"].read_header()
    assert hdr["RPMIN"] == 0.5 and hdr["ZMIN"] is None and hdr["BLINDING"] == "none"
    assert hdr["RMU_BIN"] is False


def test_fits_reads_golden_layout():
    f = FITS(cor_cases.golden_path("cf"))
    assert [h.get_extname() for h in f.hdus[1:]] == ["ATTRI", "COR"]
    assert f["COR"]["WE"][:].shape == (14, 225)
    assert f["ATTRI"].read_header()["NSIDE"] == 16


def test_healpix_pixel_centres_roundtrip():
    for nside in (1, 2, 16, 64):
        pix = np.arange(healpix.nside2npix(nside))
        theta, phi = healpix.pix2ang(nside, pix)
        assert np.array_equal(healpix.ang2pix(nside, theta, phi), pix)


def test_healpix_known_values():
    # published RING values: north pole -> 0 .. 3, south pole -> last ring, equator
    assert healpix.ang2pix(1, 0.0, 0.0) == 0
    assert healpix.ang2pix(1, np.pi, 0.0) == 8
    assert healpix.ang2pix(1, np.pi / 2, 0.0) == 4
    assert healpix.ang2pix(16, np.pi, 0.0) == healpix.nside2npix(16) - 4
    # every pixel gets the same number of uniformly drawn points within shot noise
    rng = np.random.default_rng(0)
    n = 200000
    pix = healpix.ang2pix(4, np.arccos(rng.uniform(-1, 1, n)), rng.uniform(0, 2 * np.pi, n))
    counts = np.bincount(pix, minlength=192)
    assert counts.min() > n / 192 * 0.85 and counts.max() < n / 192 * 1.15


def test_healpix_matches_golden_healpids():
    assert list(cor_cases.golden_columns("cf")["HEALPID"]) == [
        1184, 1185, 1248, 1249, 1312, 1313, 1314, 1376, 1377, 1440, 1441, 1442, 1504, 1505]
    cosmo = Cosmo(0.315, verbose=False)
    data, _, _, _ = io.read_deltas(cor_cases.DELTA_LYA, 16, 1215.67, 2.9, 2.25, cosmo,
                                   delta_attributes=cor_cases.ATTRIBUTES)
    assert sorted(data) == list(cor_cases.golden_columns("cf")["HEALPID"])


def test_cosmo_against_quadrature():
    cosmo = Cosmo(0.315, verbose=False)
    for z in (0.5, 1.96, 2.5, 3.5):
        exact = integrate.quad(lambda x: SPEED_LIGHT / (100 * np.sqrt(0.685 + 0.315 * (1 + x) ** 3)),
                               0, z, epsabs=0, epsrel=1e-12)[0]
        assert abs(cosmo.get_r_comov(z) / exact - 1) < 1e-6
        assert cosmo.get_dist_m(z) == cosmo.get_r_comov(z)
    # curved: D_M = sinh(sqrt(Ok) H0 r / c) c / (sqrt(Ok) H0)
    curved = Cosmo(0.3, Ok=0.05, verbose=False)
    r = curved.get_r_comov(2.0)
    k = 100 * np.sqrt(0.05) / SPEED_LIGHT
    assert abs(curved.get_dist_m(2.0) - np.sinh(k * r) / k) < 1e-6 * r


def test_cosmo_table_is_sequential_trapezoid():
    cosmo = Cosmo(0.315, verbose=False)
    z = cosmo.z_table
    h = 100 * np.sqrt(0.685 * (1 + z) ** 0.0 + 0.0 * (1 + z) ** 2 + 0.315 * (1 + z) ** 3
                      + 0.0 * (1 + z) ** 4)
    r = 0.0
    for k in range(1, 200):
        r = SPEED_LIGHT * (1.0 / h[k - 1] + 1.0 / h[k]) / 2.0 * 1e-3 + r
        assert cosmo.r_comov_table[k] == r


@pytest.mark.parametrize("auto", [True, False])
def test_native_neighbours_match_brute_force(auto):
    cosmo = Cosmo(0.315, verbose=False)
    data, _, z_min, _ = io.read_deltas(cor_cases.DELTA_LYA, 16, 1215.67, 2.9, 2.25, cosmo,
                                       delta_attributes=cor_cases.ATTRIBUTES)
    from picca_amd.utils import compute_ang_max
    ang_max = compute_ang_max(cosmo, 60.0, z_min)
    deltas = [d for hp in sorted(data) for d in data[hp]]
    store = ForestStore.from_deltas(deltas)
    idx = np.arange(len(deltas))
    nl = neighbours.find_neighbours(store, idx, idx, ang_max, auto=auto)
    ref = O.brute_force_neighbours(data, None if auto else data, sorted(data), ang_max)
    pos = {id(d): k for k, d in enumerate(deltas)}
    for s, d in enumerate(deltas):
        got = list(nl.target[nl.offsets[s]:nl.offsets[s + 1]])
        want = [pos[id(x)] for x in ref[id(d)]]
        assert got == want
    if auto:
        assert nl.num_pairs == 11845  # NPALL of the golden dmat.fits.gz
    else:
        assert nl.num_pairs == 23690  # NPALL of the golden dmat_cross.fits.gz


def test_neighbours_small_angle_branch_and_ra_wrap():
    """Pairs closer than 2 arcsec use the flat-sky angle; pairs across ra = 0 are found."""
    ra = np.array([0.001, 0.001 + 5e-6, 2 * np.pi - 0.001, 1.0])
    dec = np.array([0.2, 0.2 + 3e-6, 0.2, 0.2])
    n = len(ra)
    store = ForestStore(np.arange(n + 1), np.zeros(n), np.zeros(n), np.zeros(n), np.ones(n),
                        np.zeros(n), np.ones(n), ra, dec, np.arange(n))
    idx = np.arange(n)
    nl = neighbours.find_neighbours(store, idx, idx, 0.01, auto=False)
    got = {s: list(nl.target[nl.offsets[s]:nl.offsets[s + 1]]) for s in range(n)}
    assert got[0] == [1, 2] and got[3] == []
    ang01 = nl.angle[nl.offsets[0]]
    assert ang01 == np.sqrt((dec[1] - dec[0]) ** 2 + (np.cos(dec[0]) * (ra[1] - ra[0])) ** 2)
