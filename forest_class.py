"""Per-forest record stored in the pickled ``data#.npy`` files.

The class name, module name and attribute names are the input contract of the
reference drivers (reference: forest_class.py:9-46): ``data#.npy`` is
``np.save`` of a dict ``{healpix_pixel: [quasar, ...]}``, so unpickling needs
a class ``forest_class.quasar`` with these attributes.  This version computes
the HEALPix pixel with the package's own RING ``ang2pix`` instead of healpy.

Attributes
----------
ra, dec : float             radians
fib, pl, name               identifiers; ``name`` breaks self-pairs
pix : int                   RING pixel at ``parameters.nside``
x, y, z : float             unit vector of the line of sight
dc, we, dw : float64[n]     comoving distance (Mpc/h), weight, weight*delta (projected)
rx, ry, rz : float64[n]     Cartesian pixel positions (unused by the pair loop)
delta_lambda : float64[n]   loglam minus its weighted mean
lambda_average, omega, omega_delta_lambda2 : float   projection moments
"""
import numpy as np

from parameters import nside, halfpi
from lya2pcf_amd import healpix


class quasar(object):
    """One Lyman-alpha forest (reference: forest_class.py:9-46)."""

    def __init__(self, fib, pl, name, ra, dec, lenght):
        self.ra = ra
        self.fib = fib
        self.pl = pl
        self.name = name
        self.dec = dec
        self.lenght = lenght
        self.pix = int(healpix.ang2pix_ring(nside, halfpi - dec, ra))

        self.dw = np.zeros(lenght, dtype=np.float64)
        self.we = np.zeros(lenght, dtype=np.float64)
        self.dc = np.zeros(lenght, dtype=np.float64)

        self.x = np.cos(ra) * np.cos(dec)
        self.y = np.sin(ra) * np.cos(dec)
        self.z = np.sin(dec)

        self.rx = np.zeros(lenght, dtype=np.float64)
        self.ry = np.zeros(lenght, dtype=np.float64)
        self.rz = np.zeros(lenght, dtype=np.float64)

        self.lambda_average = 0
        self.delta_lambda = np.zeros(lenght, dtype=np.float64)
        self.omega_delta_lambda2 = 0
        self.omega = 0

    def fill_dw(self, de, loglam, project):
        """Set ``dw`` from deltas; with ``project`` remove the weighted mean and
        the weighted slope in loglam (reference: forest_class.py:48-68)."""
        if project:
            we = self.we
            mean_delta = np.average(de, weights=we)
            self.lambda_average = np.average(loglam, weights=we)
            self.delta_lambda = loglam - self.lambda_average
            self.omega_delta_lambda2 = np.sum(self.delta_lambda ** 2 * we)
            slope_num = np.average(self.delta_lambda * de, weights=we)
            slope_den = np.average(self.delta_lambda * self.delta_lambda, weights=we)
            if slope_den == 0:
                print("fill_dw: zero loglam variance in forest", self.name)
            self.dw = we * (de - mean_delta - self.delta_lambda * slope_num / slope_den)
            self.omega = np.sum(we)
        else:
            self.dw = de * self.we

    def dot_product(self, other):
        """cos(angle) between two lines of sight, evaluated left to right
        without fused multiply-add (reference: forest_class.py:70-76)."""
        return self.x * other.x + self.y * other.y + self.z * other.z

    def neighborhood(self, data, angmax):
        """Forests of ``data`` that pair with this one: cos(angle) > cos(angmax),
        different name and larger ra (reference: forest_class.py:78-98).

        The reference takes candidates from healpy ``query_disc(inclusive=True)``;
        any candidate superset gives the same list, so this scans the pixels whose
        centres lie within angmax plus the pixel radius.  Order: ascending pixel,
        then list order.
        """
        mumin = np.cos(angmax)
        out = []
        for pix2 in healpix.disc_candidates(nside, (self.x, self.y, self.z), angmax, data.keys()):
            for forest2 in data[pix2]:
                mu = self.dot_product(forest2)
                if mu > mumin and self.name != forest2.name and self.ra < forest2.ra:
                    out.append(forest2)
        return out
