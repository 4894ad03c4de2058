"""HEALPix RING-scheme pixelisation (Gorski et al. 2005, ApJ 622, 759).

Replaces the two healpy calls of the reference: ``ang2pix(nside, theta, phi)``
(forest_class.py:24) and the candidate-pixel list of ``query_disc(...,
inclusive=True)`` (forest_class.py:89).  The pair list only needs a candidate
*superset*, because the final cut is the exact dot-product test.
"""
import numpy as np

_TWO_PI = 2.0 * np.pi


def ang2pix_ring(nside, theta, phi):
    """RING pixel index of colatitude ``theta`` and longitude ``phi`` (radians).

    Closed form of Gorski et al. (2005) sec. 4: equatorial belt |z| <= 2/3 is
    indexed along the two families of edge lines, polar caps by rings of
    growing length.  Accepts scalars or arrays; returns int64.
    """
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)
    scalar = theta.ndim == 0 and phi.ndim == 0
    theta, phi = np.broadcast_arrays(np.atleast_1d(theta), np.atleast_1d(phi))
    n = np.int64(nside)
    z = np.cos(theta)
    za = np.abs(z)
    tt = np.mod(phi, _TWO_PI) * (2.0 / np.pi)
    tt = np.where(tt >= 4.0, tt - 4.0, tt)
    pix = np.empty(theta.shape, dtype=np.int64)

    eq = za <= 2.0 / 3.0
    if np.any(eq):
        t1 = n * (0.5 + tt[eq])
        t2 = n * z[eq] * 0.75
        jp = np.floor(t1 - t2).astype(np.int64)
        jm = np.floor(t1 + t2).astype(np.int64)
        ir = n + 1 + jp - jm
        kshift = 1 - (ir & 1)
        ip = (jp + jm - n + kshift + 1) // 2
        ip = np.mod(ip, 4 * n)
        pix[eq] = 2 * n * (n - 1) + (ir - 1) * 4 * n + ip

    pol = ~eq
    if np.any(pol):
        ttp = tt[pol]
        tp = ttp - np.floor(ttp)
        tmp = n * np.sqrt(3.0 * (1.0 - za[pol]))
        jp = np.floor(tp * tmp).astype(np.int64)
        jm = np.floor((1.0 - tp) * tmp).astype(np.int64)
        ir = jp + jm + 1
        ip = np.floor(ttp * ir).astype(np.int64)
        ip = np.mod(ip, 4 * ir)
        north = z[pol] > 0
        pix[pol] = np.where(north, 2 * ir * (ir - 1) + ip, 12 * n * n - 2 * ir * (ir + 1) + ip)
    return int(pix[0]) if scalar else pix


def pix2ang_ring(nside, pix):
    """Colatitude and longitude of RING pixel centres (inverse of the layout above)."""
    pix = np.atleast_1d(np.asarray(pix, dtype=np.int64))
    n = np.int64(nside)
    npix = 12 * n * n
    ncap = 2 * n * (n - 1)
    z = np.empty(pix.shape)
    phi = np.empty(pix.shape)

    north = pix < ncap
    if np.any(north):
        p = pix[north]
        iring = (1 + np.floor(np.sqrt(1.0 + 2.0 * p)).astype(np.int64)) >> 1
        # guard the float sqrt against off-by-one at perfect squares
        iring = np.where(2 * iring * (iring - 1) > p, iring - 1, iring)
        iring = np.where(2 * (iring + 1) * iring <= p, iring + 1, iring)
        iphi = p + 1 - 2 * iring * (iring - 1)
        z[north] = 1.0 - iring * iring * 4.0 / npix
        phi[north] = (iphi - 0.5) * np.pi / (2.0 * iring)

    belt = (pix >= ncap) & (pix < npix - ncap)
    if np.any(belt):
        p = pix[belt] - ncap
        iring = p // (4 * n) + n
        iphi = p % (4 * n) + 1
        fodd = np.where(((iring + n) & 1) == 1, 1.0, 0.5)
        z[belt] = (2 * n - iring) * 2.0 / (3.0 * n)
        phi[belt] = (iphi - fodd) * np.pi / (2.0 * n)

    south = pix >= npix - ncap
    if np.any(south):
        p = npix - pix[south]
        iring = (1 + np.floor(np.sqrt(2.0 * p - 1.0)).astype(np.int64)) >> 1
        iring = np.where(2 * iring * (iring - 1) >= p, iring - 1, iring)
        iring = np.where(2 * (iring + 1) * iring < p, iring + 1, iring)
        iphi = 4 * iring + 1 - (p - 2 * iring * (iring - 1))
        z[south] = -1.0 + iring * iring * 4.0 / npix
        phi[south] = (iphi - 0.5) * np.pi / (2.0 * iring)

    return np.arccos(np.clip(z, -1.0, 1.0)), phi


def max_pixel_radius(nside):
    """Upper bound on the angular distance from a pixel centre to any point of
    the pixel.  1.5x the mean pixel size is above the true maximum for every
    nside (the worst pixels, at the cap/belt transition, reach about 1.1x)."""
    return 1.5 * np.sqrt(np.pi / 3.0) / nside


def disc_candidates(nside, vec, radius, pixels):
    """Pixels of the iterable ``pixels`` that can hold a point within ``radius``
    of unit vector ``vec``, in ascending order (superset of healpy
    ``query_disc(..., inclusive=True)`` restricted to ``pixels``)."""
    keys = np.array(sorted(int(p) for p in pixels), dtype=np.int64)
    if keys.size == 0:
        return []
    th, ph = pix2ang_ring(nside, keys)
    st = np.sin(th)
    cosd = st * np.cos(ph) * vec[0] + st * np.sin(ph) * vec[1] + np.cos(th) * vec[2]
    lim = np.cos(min(np.pi, radius + max_pixel_radius(nside)))
    return [int(k) for k in keys[cosd >= lim]]
