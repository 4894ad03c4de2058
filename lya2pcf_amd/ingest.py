"""Delta-file ingest: PICCA/DESI and eBOSS delta FITS files to forests.

Two file layouts, as the reference readers take them:

* DESI (reference delta_reader.py:25-62): one file holds many forests in
  image HDUs on a common wavelength grid -- ``LAMBDA`` [n_lambda],
  ``<delta_key>`` and ``WEIGHT`` [n_forest, n_lambda] -- and a ``METADATA``
  table with LOS_ID, TARGETID, RA, DEC (radians).  Pixels where delta is not
  finite are dropped; z = lambda / lambda_lya - 1, loglam = log10(lambda).
* eBOSS (reference delta_reader_eboss.py:25-56): every extension HDU is one
  forest, a table with LOGLAM, DELTA, WEIGHT and header cards FIBERID, PLATE,
  THING_ID, RA, DEC; z = 10^(loglam - log10(lambda_lya)) - 1.

In both, the weight is rescaled by ((1 + z) / (1 + z_ref))^(gamma / 2), delta
is projected (weighted mean and loglam slope removed, forest_class.py:48-68)
and dc = d_c(z).  ``records_*`` return the per-forest arrays; the drop-in
readers turn them into ``forest_class.quasar`` objects pickled in data#.npy,
and :func:`forest_set` packs them straight into the kernels' SoA, skipping the
pickle.  A forest without any finite delta is skipped with a warning (the
reference stops there with a division by zero in ``np.average``).
"""
import warnings
from dataclasses import dataclass

import numpy as np

from . import fits, healpix
from .forests import ForestSet
from .synth import project_forests

LAMBDA_LYA = 1215.67


@dataclass
class Settings:
    """Ingest constants (reference parameters.py:11,31-40,44-54)."""
    delta_key: str = "DELTA"
    lambda_lya: float = LAMBDA_LYA
    z_ref: float = 2.25
    gamma: float = 3.8
    nside: int = 32
    omega_m: float = 0.3153
    zmin: float = 0.0
    zmax: float = 10.0
    nz: int = 10000

    @classmethod
    def from_parameters(cls, P):
        return cls(delta_key=P.delta_key, lambda_lya=float(P.lambdaa), z_ref=float(P.z_ref), gamma=float(P.gamma),
                   nside=int(P.nside), omega_m=float(P.Omm), zmin=float(P.zmin), zmax=float(P.zmax), nz=int(P.nz))

    def distance(self):
        from .cosmology import ComovingDistance
        return ComovingDistance(self.omega_m, self.zmin, self.zmax, self.nz)


@dataclass
class DeltaRecord:
    """One forest as read from a delta file (weights already rescaled)."""
    fib: object
    pl: object
    name: object
    ra: float
    dec: float
    lenght: int           # the reference's record length (all grid points for DESI files)
    loglam: np.ndarray
    z: np.ndarray
    delta: np.ndarray
    we: np.ndarray


def _rescale(w, z, st):
    return w * np.power((z + 1.) / (1. + st.z_ref), st.gamma / 2)


def records_desi(path, st=None):
    """Forests of a DESI-layout delta file (reference delta_reader.py:25-62)."""
    st = st or Settings()
    hdus = fits.read(path)
    meta = fits.by_name(hdus, "METADATA").data
    delta = fits.by_name(hdus, st.delta_key).data
    weight = fits.by_name(hdus, "WEIGHT").data
    lam = np.asarray(fits.by_name(hdus, "LAMBDA").data, dtype=np.float64).ravel()
    n_forest, n_lambda = delta.shape
    out = []
    for i in range(n_forest):
        d = np.asarray(delta[i], dtype=np.float64)
        mask = np.isfinite(d)
        if not mask.any():
            warnings.warn("%s: forest %d has no finite delta; skipped" % (path, i))
            continue
        l = lam[mask]
        z = l / st.lambda_lya - 1
        out.append(DeltaRecord(meta["LOS_ID"][i], meta["LOS_ID"][i], meta["TARGETID"][i], float(meta["RA"][i]),
                               float(meta["DEC"][i]), n_lambda, np.log10(l), z, d[mask],
                               _rescale(np.asarray(weight[i], dtype=np.float64)[mask], z, st)))
    return out


def records_eboss(path, st=None):
    """Forests of an eBOSS-layout delta file, one HDU each (reference
    delta_reader_eboss.py:25-56)."""
    st = st or Settings()
    out = []
    for h in fits.read(path)[1:]:
        t = h.data
        loglam = np.asarray(t["LOGLAM"], dtype=np.float64)
        z = np.power(10, loglam - np.log10(st.lambda_lya)) - 1.
        w = np.asarray(t["WEIGHT"], dtype=np.float64)
        if len(w) == 0:
            warnings.warn("%s: empty forest HDU skipped" % path)
            continue
        hd = h.header
        out.append(DeltaRecord(hd["FIBERID"], hd["PLATE"], hd["THING_ID"], float(hd["RA"]), float(hd["DEC"]),
                               len(w), loglam, z, np.asarray(t["DELTA"], dtype=np.float64), _rescale(w, z, st)))
    return out


def pixel_of(records, nside):
    """RING HEALPix pixel of every record (theta = pi/2 - dec, phi = ra)."""
    ra = np.array([r.ra for r in records], dtype=np.float64)
    dec = np.array([r.dec for r in records], dtype=np.float64)
    return np.asarray(healpix.ang2pix_ring(nside, np.pi / 2 - dec, ra), dtype=np.int64)


def quasars(records, quasar_cls, dist):
    """``quasar_cls`` objects filled as the reference readers fill them."""
    out = []
    for r in records:
        q = quasar_cls(r.fib, r.pl, r.name, r.ra, r.dec, r.lenght)
        q.we = r.we
        q.fill_dw(r.delta, r.loglam, True)
        dc = np.asarray(dist(r.z), dtype=np.float64)
        q.dc = dc
        q.rx, q.ry, q.rz = q.x * dc, q.y * dc, q.z * dc
        out.append(q)
    return out


def forest_set(records, st=None, dist=None, pixel_order="sorted"):
    """Pack records straight into a :class:`ForestSet` (no quasar objects, no
    pickle).  Forests are grouped by HEALPix pixel -- ascending pixel
    (``"sorted"``, as delta_reader.py:116-117 stores them) or in order of
    first appearance (``"first"``, as delta_reader_eboss.py:80-85) -- keeping
    file order within a pixel.  The projection is evaluated vectorised; it
    agrees with the per-forest ``fill_dw`` to rounding."""
    st = st or Settings()
    dist = dist or st.distance()
    if not records:
        raise NameError("Empty data: no forests in the delta files")
    pix = pixel_of(records, st.nside)
    if pixel_order == "sorted":
        keys = np.unique(pix)
    elif pixel_order == "first":
        _, first = np.unique(pix, return_index=True)
        keys = pix[np.sort(first)]
    else:
        raise ValueError("pixel_order must be 'sorted' or 'first'")
    rank = {int(k): s for s, k in enumerate(keys)}
    slot = np.array([rank[int(p)] for p in pix], dtype=np.int32)
    order = np.argsort(slot, kind="stable")
    recs = [records[k] for k in order]
    slot = slot[order]
    n = np.array([len(r.we) for r in recs], dtype=np.int64)
    off = np.zeros(len(recs) + 1, dtype=np.int64)
    np.cumsum(n, out=off[1:])
    cat = lambda a: np.concatenate([getattr(r, a) for r in recs])
    loglam, z, delta, we = cat("loglam"), cat("z"), cat("delta"), cat("we")
    dw, dl, omega, odl2, _ = project_forests(off, loglam, delta, we)
    ra = np.array([r.ra for r in recs], dtype=np.float64)
    dec = np.array([r.dec for r in recs], dtype=np.float64)
    names = np.array([int(r.name) for r in recs], dtype=np.int64)
    fs = ForestSet(off, [int(k) for k in keys], ux=np.cos(ra) * np.cos(dec), uy=np.sin(ra) * np.cos(dec),
                   uz=np.sin(dec), ra=ra, dec=dec, name=names, slot=slot, omega=omega, odl2=odl2,
                   dc=np.asarray(dist(z), dtype=np.float64), we=we, dw=dw, dl=dl)
    return fs


def neighbour_names(data, rtmax):
    """For every forest of ``data`` ({pixel: [quasar]}), the names and pixels
    of its neighbours under its own cut angle 2 asin(rtmax / (2 dc[0])), in
    ascending pixel then list order (reference delta_reader.py:119-128 with
    forest_class.py:100-121).  Sets ``neigh_names`` and ``neigh_pixels`` on
    each forest and returns the neighbour counts in sorted-pixel order."""
    from . import xi
    from .distortion import neighbour_order
    pixels = sorted(data.keys())
    fs = ForestSet.from_data(data, pixels)
    if fs.nf == 0:
        return []
    first = fs.dc[fs.pix_off[:-1]]
    own = 2 * np.arcsin(0.5 * rtmax / first)
    off, idx = xi.build_pairs(fs, float(np.max(own)))   # superset: the widest cut
    rank = neighbour_order(fs)
    counts = []
    for i, q in enumerate(fs.objects):
        nb = idx[off[i]:off[i + 1]]
        mu = (fs.ux[i] * fs.ux[nb] + fs.uy[i] * fs.uy[nb]) + fs.uz[i] * fs.uz[nb]
        nb = nb[mu > np.cos(own[i])]
        nb = nb[np.argsort(rank[nb], kind="stable")]
        q.neigh_names = [fs.objects[j].name for j in nb]
        q.neigh_pixels = [fs.slots[fs.slot[j]] for j in nb]
        counts.append(len(nb))
    return counts
