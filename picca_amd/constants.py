"""Physical constants and the tabulated fiducial cosmology.

Mirrors ``py/picca/constants.py``: ``SMALL_ANGLE_CUT_OFF`` (:16), ``SPEED_LIGHT`` (:18),
``ABSORBER_IGM`` (:240-312, a subset) and
``Cosmo`` (:159-233, without the camb distance template, which is off the path).

``Cosmo`` integrates c/H(z) with the trapezoid rule on 10 000 steps of dz = 1e-3 from
z = 0 to 10 with H0 = 100 km/s/Mpc (so distances come out in Mpc/h) and interpolates
linearly with ``scipy.interpolate.interp1d``.  The cumulative sum is carried out
left to right, one step at a time, so the table is bit-for-bit the one the sequential
loop at ``py/picca/constants.py:205-209`` produces.
"""
import numpy as np
from scipy import interpolate
from scipy.constants import speed_of_light as _speed_of_light

SMALL_ANGLE_CUT_OFF = 2.0 / 3600.0 * np.pi / 180.0  # 2 arcsec, in radians
SPEED_LIGHT = _speed_of_light / 1000.0  # km/s

# rest-frame wavelengths in Angstrom; a subset of the table at
# py/picca/constants.py:240-312 (the absorbers a correlation run can name with
# --lambda-abs/--lambda-abs2, and the emission lines of the synthetic spectra)
ABSORBER_IGM = {
    "NV(1243)": 1242.804,
    "NV(1239)": 1238.821,
    "SiIV(1403)": 1402.77291,
    "SiIV(1394)": 1393.76018,
    "CIV(eff)": 1549.06,
    "SiII(1527)": 1526.70698,
    "SiII(1304)": 1304.3702,
    "SiII(1260)": 1260.4221,
    "LYA": 1215.67,
    "SiIII(1207)": 1206.500,
    "SiII(1193)": 1193.2897,
    "SiII(1190)": 1190.4158,
    "OVI(1038)": 1037.613,
    "OVI(1032)": 1031.912,
    "LYB": 1025.72,
}


class Cosmo:
    """Fiducial flat/curved wCDM distance tables (Mpc/h).

    Attributes ``get_r_comov``, ``get_dist_m``, ``get_hubble``, ``get_dist_hubble``,
    ``get_dist_v`` and ``distance_to_redshift`` are interp1d callables, as in
    ``py/picca/constants.py:211-233``.
    """

    NUM_BINS = 10000
    Z_MAX = 10.0

    def __init__(self, Om, Ok=0.0, Or=0.0, wl=-1.0, blinding=False, verbose=True,
                 template=None):
        if template is not None:
            raise NotImplementedError(
                "distance templates need camb, which is not part of this engine")
        hubble0 = 100.0
        omega_lambda = 1.0 - Ok - Om - Or
        delta_z = self.Z_MAX / self.NUM_BINS
        z = np.arange(self.NUM_BINS, dtype=float) * delta_z
        hubble = hubble0 * np.sqrt(omega_lambda * (1.0 + z) ** (3.0 * (1.0 + wl))
                                   + Ok * (1.0 + z) ** 2
                                   + Om * (1.0 + z) ** 3
                                   + Or * (1.0 + z) ** 4)
        # trapezoid steps, then a strictly sequential running sum
        steps = SPEED_LIGHT * (1.0 / hubble[:-1] + 1.0 / hubble[1:]) / 2.0 * delta_z
        r_comov = np.zeros(self.NUM_BINS)
        r_comov[1:] = np.add.accumulate(steps)
        if Ok == 0.0:
            dist_m = r_comov
        elif Ok < 0.0:
            k = hubble0 * np.sqrt(-Ok) / SPEED_LIGHT
            dist_m = np.sin(k * r_comov) / k
        else:
            k = hubble0 * np.sqrt(Ok) / SPEED_LIGHT
            dist_m = np.sinh(k * r_comov) / k

        self.z_table = z
        self.r_comov_table = r_comov
        self.dist_m_table = dist_m
        self.get_r_comov = interpolate.interp1d(z, r_comov)
        self.get_dist_m = interpolate.interp1d(z, dist_m)
        self.get_hubble = interpolate.interp1d(z, hubble)
        self.distance_to_redshift = interpolate.interp1d(r_comov, z)
        self.get_dist_hubble = interpolate.interp1d(z, SPEED_LIGHT / hubble)
        dist_v = np.power(z * self.get_dist_m(z) ** 2 * self.get_dist_hubble(z), 1.0 / 3.0)
        self.get_dist_v = interpolate.interp1d(z, dist_v)
