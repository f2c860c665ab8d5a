"""HEALPix RING-scheme pixel index (Górski et al. 2005, ApJ 622, 759, section 4).

picca assigns every forest to ``healpy.ang2pix(nside, pi/2 - dec, ra)``
(``py/picca/io.py:486-488``) and uses the result as the unit of work and as the
``HEALPID`` column of the output.  healpy is not part of this stack; this is a
vectorised restatement of the published RING indexing: equatorial belt for
|cos theta| <= 2/3, polar caps above, with the sin(theta) form of the cap radius
near the poles (|cos theta| > 0.99) as in the reference implementation of the paper.
"""
import numpy as np

_TWO_THIRDS = 2.0 / 3.0


def nside2npix(nside):
    return 12 * int(nside) * int(nside)


def ang2pix(nside, theta, phi):
    """RING pixel index of colatitude ``theta`` and longitude ``phi`` (radians)."""
    nside = int(nside)
    theta = np.asarray(theta, dtype=float)
    phi = np.asarray(phi, dtype=float)
    scalar = theta.ndim == 0 and phi.ndim == 0
    theta, phi = np.broadcast_arrays(np.atleast_1d(theta), np.atleast_1d(phi))
    if np.any((theta < 0) | (theta > np.pi)):
        raise ValueError("THETA is out of range [0,pi]")

    z = np.cos(theta)
    sin_theta = np.sin(theta)
    za = np.abs(z)
    # phi folded into [0, 2 pi) and expressed in units of pi/2
    tt = np.mod(phi, 2.0 * np.pi) * (2.0 / np.pi)
    tt = np.where(tt >= 4.0, tt - 4.0, tt)

    npix = nside2npix(nside)
    ncap = 2 * nside * (nside - 1)
    nl4 = 4 * nside
    pix = np.empty(z.shape, dtype=np.int64)

    belt = za <= _TWO_THIRDS
    if np.any(belt):
        zb = z[belt]
        t1 = nside * (0.5 + tt[belt])
        t2 = nside * (zb * 0.75)
        jp = (t1 - t2).astype(np.int64)  # index of the ascending edge line
        jm = (t1 + t2).astype(np.int64)  # index of the descending edge line
        ring = nside + 1 + jp - jm  # 1 .. 2 nside + 1, counted in the belt
        kshift = 1 - (ring & 1)
        ip = (jp + jm - nside + kshift + 1) // 2
        ip = np.mod(ip, nl4)
        pix[belt] = ncap + (ring - 1) * nl4 + ip

    cap = ~belt
    if np.any(cap):
        zc = z[cap]
        ttc = tt[cap]
        tp = ttc - np.floor(ttc)
        near_pole = np.abs(zc) > 0.99
        tmp = np.where(near_pole,
                       nside * sin_theta[cap] / np.sqrt((1.0 + np.abs(zc)) / 3.0),
                       nside * np.sqrt(3.0 * (1.0 - np.abs(zc))))
        jp = (tp * tmp).astype(np.int64)
        jm = ((1.0 - tp) * tmp).astype(np.int64)
        ring = jp + jm + 1  # ring number counted from the nearest pole
        ip = (ttc * ring).astype(np.int64)
        ip = np.mod(ip, 4 * ring)
        north = zc > 0
        pix[cap] = np.where(north, 2 * ring * (ring - 1) + ip,
                            npix - 2 * ring * (ring + 1) + ip)
    if scalar:
        return int(pix[0])
    return pix


def pix2ang(nside, pix):
    """Centre (theta, phi) of RING pixels; used by tests as the inverse of ang2pix."""
    nside = int(nside)
    pix = np.atleast_1d(np.asarray(pix, dtype=np.int64))
    npix = nside2npix(nside)
    ncap = 2 * nside * (nside - 1)
    z = np.empty(pix.shape)
    phi = np.empty(pix.shape)

    north = pix < ncap
    if np.any(north):
        p = pix[north]
        ring = ((1 + np.sqrt(1 + 2 * p)) // 2).astype(np.int64)
        # guard the float square root
        ring = np.where(2 * ring * (ring - 1) > p, ring - 1, ring)
        ring = np.where(2 * (ring + 1) * ring <= p, ring + 1, ring)
        iphi = p + 1 - 2 * ring * (ring - 1)
        z[north] = 1.0 - ring * ring / (3.0 * nside * nside)
        phi[north] = (iphi - 0.5) * np.pi / (2.0 * ring)

    south = pix >= npix - ncap
    if np.any(south):
        p = npix - pix[south]
        ring = ((1 + np.sqrt(2 * p - 1)) // 2).astype(np.int64)
        ring = np.where(2 * ring * (ring - 1) >= p, ring - 1, ring)
        ring = np.where(2 * (ring + 1) * ring < p, ring + 1, ring)
        iphi = 4 * ring + 1 - (p - 2 * ring * (ring - 1))
        z[south] = -1.0 + ring * ring / (3.0 * nside * nside)
        phi[south] = (iphi - 0.5) * np.pi / (2.0 * ring)

    belt = ~(north | south)
    if np.any(belt):
        p = pix[belt] - ncap
        nl4 = 4 * nside
        ring = p // nl4 + nside  # ring index from the north pole
        iphi = p % nl4 + 1
        fodd = np.where(((ring + nside) & 1) == 1, 1.0, 0.5)
        z[belt] = (2 * nside - ring) * 2.0 / (3.0 * nside)
        phi[belt] = (iphi - fodd) * np.pi / (2.0 * nside)

    return np.arccos(z), phi
