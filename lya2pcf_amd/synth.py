"""Seeded synthetic Lyman-alpha surveys in the layout the pair kernels consume.

The reference reads PICCA delta files (reference: delta_reader.py:25-62); no
data ship with it, so every workload here is generated.  A survey is built
directly as ragged structure-of-arrays (one row per forest, one entry per
spectral pixel, ``pix_off`` giving each forest's slice), and can be written as
the reference's pickled ``data#.npy`` (``{healpix: [quasar]}``) for the
drop-in drivers.

Recipe (shared by every configuration):
  * quasar positions uniform on a patch RA in [0, W), sin(dec) in [-sin h, sin h];
    optional per-HEALPix-pixel density factor 1 + c*U(0,1) by rejection;
  * pixels on the loglam grid k * step, restricted to the rest-frame window
    (lambda_rest in [1040, 1200) A) and the observed window [3600, 5500) A, or,
    with ``length_range``, n ~ U{lo..hi} pixels from the blue edge (clipped at 5500 A);
  * z = lambda / 1215.67 - 1, dc = d_c(z) in flat LCDM;
  * delta = N(0, 0.2^2) + 0.05 cos(k . x), |k| = 2 pi / 100 h/Mpc along a seeded
    direction, x = dc * (unit vector of the quasar);
  * w = 1 / (eta * sigma2_pip + sigma2_lss(z)), sigma2_pip ~ U(0, 0.2), eta = 1,
    sigma2_lss = 0.05 at 10 redshift nodes in [2, 3.5] (linear interpolation);
    this is w = 1 / (0.05 + U(0, 0.2));
  * dw = w * (delta - weighted mean - delta_lambda * weighted slope), the
    projection of reference forest_class.py:54-66.
"""
from dataclasses import dataclass, field
import numpy as np

from . import healpix
from .cosmology import ComovingDistance

LYA = 1215.67
DEG = np.pi / 180.0


@dataclass
class SurveyConfig:
    name: str
    n_forests: int
    area_deg2: float
    dec_half_deg: float
    zq_range: tuple
    seed: int
    nside: int
    loglam_step: float = 3e-4
    rest_window: tuple = (1040.0, 1200.0)
    obs_window: tuple = (3600.0, 5500.0)
    length_range: tuple = None
    density_contrast: float = 0.0
    omega_m: float = 0.315
    delta_sigma: float = 0.2
    mode_amp: float = 0.05
    mode_wavelength: float = 100.0
    eta: float = 1.0
    sigma2_lss: float = 0.05
    sigma2_pip_max: float = 0.2
    z_nodes: tuple = field(default_factory=lambda: tuple(np.linspace(2.0, 3.5, 10)))
    project: bool = True


# The five workloads of the benchmark plan.
CONFIGS = {
    "patch2k": SurveyConfig("2k forests, 10x10 deg, zq U[2.1,3.0]", 2000, 99.7, 5.0, (2.1, 3.0), 1000, 32),
    "sky100k": SurveyConfig("100k forests, 2500 deg2, nside 16, zq U[2.1,3.5]", 100_000, 2500.0, 25.0,
                       (2.1, 3.5), 1001, 16),
    "sky300k": SurveyConfig("300k forests, 8000 deg2, nside 8, density 1+9U, n U{50..600}", 300_000, 8000.0,
                       40.0, (2.1, 3.5), 1003, 8, length_range=(50, 600), density_contrast=9.0),
    "dist20k": SurveyConfig("20k forests, 1000 deg2, zq U[2.1,3.0]", 20_000, 1000.0, 15.0, (2.1, 3.0), 1004, 32),
}


def scaled(cfg, n_forests, area_deg2=None, seed=None, **kw):
    """Same recipe at another size (density kept unless an area is given)."""
    area = cfg.area_deg2 * n_forests / cfg.n_forests if area_deg2 is None else area_deg2
    d = dict(cfg.__dict__)
    d.update(n_forests=int(n_forests), area_deg2=float(area), **kw)
    if seed is not None:
        d["seed"] = seed
    # keep the patch roughly square in degrees for small areas
    side = np.sqrt(area)
    d["dec_half_deg"] = min(cfg.dec_half_deg, max(0.5, 0.5 * side))
    return SurveyConfig(**d)


class Survey:
    """Ragged SoA survey.  Forest arrays have length ``nf``; pixel arrays have
    length ``pix_off[-1]``.  Forests are grouped by HEALPix pixel in ascending
    pixel order (the order in which the data dict is written)."""

    FOREST_FIELDS = ("ra", "dec", "zq", "name", "hpix", "ux", "uy", "uz",
                     "omega", "omega_dl2", "lambda_average")
    PIXEL_FIELDS = ("loglam", "z", "dc", "delta", "sigma2_pip", "we", "dw", "delta_lambda")

    def __init__(self, **arrays):
        for k, v in arrays.items():
            setattr(self, k, v)

    @property
    def nf(self):
        return len(self.pix_off) - 1

    @property
    def n(self):
        return np.diff(self.pix_off)

    def subset(self, idx):
        """Survey restricted to forests ``idx`` (kept in the given order)."""
        idx = np.asarray(idx, dtype=np.int64)
        n = self.n[idx]
        off = np.zeros(len(idx) + 1, dtype=np.int64)
        np.cumsum(n, out=off[1:])
        sel = np.concatenate([np.arange(self.pix_off[i], self.pix_off[i + 1]) for i in idx]) if len(idx) else \
            np.zeros(0, dtype=np.int64)
        out = {"pix_off": off, "nside": self.nside, "config": getattr(self, "config", None)}
        for f in self.FOREST_FIELDS:
            out[f] = getattr(self, f)[idx]
        for f in self.PIXEL_FIELDS:
            out[f] = getattr(self, f)[sel]
        return Survey(**out)


def _sigma2_lss(z, cfg):
    nodes = np.asarray(cfg.z_nodes)
    vals = np.full(len(nodes), cfg.sigma2_lss)
    return np.interp(z, nodes, vals)


def project_forests(pix_off, loglam, delta, we):
    """Weighted mean and slope removal per forest (reference forest_class.py:54-66),
    vectorised over forests.  Returns dw, delta_lambda, omega, omega_dl2, lambda_average."""
    starts = pix_off[:-1]
    seg = lambda a: np.add.reduceat(a, starts)
    omega = seg(we)
    mean_delta = seg(delta * we) / omega
    lam_avg = seg(loglam * we) / omega
    counts = np.diff(pix_off)
    dl = loglam - np.repeat(lam_avg, counts)
    odl2 = seg(dl ** 2 * we)
    slope_num = seg(dl * delta * we) / omega
    slope_den = seg(dl * dl * we) / omega
    dw = we * (delta - np.repeat(mean_delta, counts) - dl * np.repeat(slope_num / slope_den, counts))
    return dw, dl, omega, odl2, lam_avg


def generate(cfg, dist=None):
    """Build the survey described by ``cfg`` (deterministic in ``cfg.seed``)."""
    ss = np.random.SeedSequence(cfg.seed)
    r_pos, r_zq, r_len, r_k, r_delta, r_pip, r_dens = [np.random.default_rng(s) for s in ss.spawn(7)]
    dist = dist or ComovingDistance(cfg.omega_m)
    N = cfg.n_forests

    h = cfg.dec_half_deg * DEG
    width = (cfg.area_deg2 * DEG * DEG) / (2.0 * np.sin(h))
    if width > 2 * np.pi:
        raise ValueError("patch wider than the sky")

    # positions (with optional density modulation per HEALPix pixel)
    if cfg.density_contrast > 0:
        dens = 1.0 + cfg.density_contrast * r_dens.random(12 * cfg.nside * cfg.nside)
        ra_l, dec_l, got = [], [], 0
        while got < N:
            m = 2 * (N - got) + 1024
            ra_c = width * r_pos.random(m)
            dec_c = np.arcsin(np.sin(h) * (2.0 * r_pos.random(m) - 1.0))
            p = healpix.ang2pix_ring(cfg.nside, np.pi / 2 - dec_c, ra_c)
            keep = r_pos.random(m) * (1.0 + cfg.density_contrast) < dens[p]
            ra_l.append(ra_c[keep]); dec_l.append(dec_c[keep]); got += int(keep.sum())
        ra = np.concatenate(ra_l)[:N]
        dec = np.concatenate(dec_l)[:N]
    else:
        ra = width * r_pos.random(N)
        dec = np.arcsin(np.sin(h) * (2.0 * r_pos.random(N) - 1.0))
    zq = r_zq.uniform(cfg.zq_range[0], cfg.zq_range[1], N)
    name = np.arange(N, dtype=np.int64) + 1_000_000_000
    hp = healpix.ang2pix_ring(cfg.nside, np.pi / 2 - dec, ra)

    # pixel grids
    step = cfg.loglam_step
    lo = np.maximum(cfg.rest_window[0] * (1.0 + zq), cfg.obs_window[0])
    k_lo = np.ceil(np.log10(lo) / step).astype(np.int64)
    k_obs_hi = np.ceil(np.log10(cfg.obs_window[1]) / step).astype(np.int64) - 1
    if cfg.length_range is not None:
        nlen = r_len.integers(cfg.length_range[0], cfg.length_range[1] + 1, N)
        k_hi = np.minimum(k_lo + nlen - 1, k_obs_hi)
    else:
        hi = np.minimum(cfg.rest_window[1] * (1.0 + zq), cfg.obs_window[1])
        k_hi = np.ceil(np.log10(hi) / step).astype(np.int64) - 1
    n = k_hi - k_lo + 1
    keep = n >= 2
    if not np.all(keep):
        ra, dec, zq, name, hp, k_lo, n = ra[keep], dec[keep], zq[keep], name[keep], hp[keep], k_lo[keep], n[keep]

    # group by HEALPix pixel (stable, so generation order is kept inside a pixel)
    order = np.argsort(hp, kind="stable")
    ra, dec, zq, name, hp, k_lo, n = ra[order], dec[order], zq[order], name[order], hp[order], k_lo[order], n[order]
    nf = len(ra)
    pix_off = np.zeros(nf + 1, dtype=np.int64)
    np.cumsum(n, out=pix_off[1:])
    P = int(pix_off[-1])
    fid = np.repeat(np.arange(nf), n)
    kk = np.repeat(k_lo, n) + (np.arange(P) - np.repeat(pix_off[:-1], n))
    loglam = kk * step
    z = 10.0 ** loglam / LYA - 1.0
    dc = np.asarray(dist(z), dtype=np.float64)

    ux = np.cos(ra) * np.cos(dec)
    uy = np.sin(ra) * np.cos(dec)
    uz = np.sin(dec)

    kdir = r_k.normal(size=3)
    kdir /= np.linalg.norm(kdir)
    kvec = kdir * (2.0 * np.pi / cfg.mode_wavelength)
    phase = dc * (kvec[0] * ux[fid] + kvec[1] * uy[fid] + kvec[2] * uz[fid])
    delta = r_delta.normal(0.0, cfg.delta_sigma, P) + cfg.mode_amp * np.cos(phase)
    s2pip = r_pip.uniform(0.0, cfg.sigma2_pip_max, P)
    we = 1.0 / (cfg.eta * s2pip + _sigma2_lss(z, cfg))

    if cfg.project:
        dw, dl, omega, odl2, lam_avg = project_forests(pix_off, loglam, delta, we)
    else:
        dw = delta * we
        dl = np.zeros(P)
        omega = np.add.reduceat(we, pix_off[:-1])
        odl2 = np.zeros(nf)
        lam_avg = np.zeros(nf)

    return Survey(ra=ra, dec=dec, zq=zq, name=name, hpix=hp, ux=ux, uy=uy, uz=uz,
                  omega=omega, omega_dl2=odl2, lambda_average=lam_avg,
                  pix_off=pix_off, loglam=loglam, z=z, dc=dc, delta=delta, sigma2_pip=s2pip,
                  we=we, dw=dw, delta_lambda=dl, nside=cfg.nside, config=cfg, kvec=kvec)


def to_data_dict(survey):
    """``{healpix: [forest_class.quasar]}`` carrying exactly the survey's values."""
    from forest_class import quasar
    data = {}
    off = survey.pix_off
    for f in range(survey.nf):
        a, b = int(off[f]), int(off[f + 1])
        q = quasar.__new__(quasar)
        q.ra = float(survey.ra[f]); q.dec = float(survey.dec[f])
        q.fib = int(survey.name[f]); q.pl = int(survey.name[f]); q.name = int(survey.name[f])
        q.lenght = b - a
        q.pix = int(survey.hpix[f])
        q.x = np.float64(survey.ux[f]); q.y = np.float64(survey.uy[f]); q.z = np.float64(survey.uz[f])
        q.dc = survey.dc[a:b].copy(); q.we = survey.we[a:b].copy(); q.dw = survey.dw[a:b].copy()
        q.rx = q.x * q.dc; q.ry = q.y * q.dc; q.rz = q.z * q.dc
        q.delta_lambda = survey.delta_lambda[a:b].copy()
        q.lambda_average = float(survey.lambda_average[f])
        q.omega = float(survey.omega[f]); q.omega_delta_lambda2 = float(survey.omega_dl2[f])
        data.setdefault(q.pix, []).append(q)
    return data


def write_data_npy(survey, path):
    """Write the survey as the reference's pickled ``data#.npy``
    (reference: delta_reader.py:139-146)."""
    np.save(path, to_data_dict(survey), allow_pickle=True)


def adversarial(n_pairs, seed=1002, n_pix=8, eps=1e-6, rmax=200.0, nbins=50, omega_m=0.315, nside=16,
                theta=(5e-4, 0.04)):
    """Forest pairs whose pixel pairs sit within ``eps`` Mpc/h of bin edges.

    Each pair (f1, f2) gets a random small separation theta; forest 1 has
    ``n_pix`` pixels spaced ~2 Mpc/h; every forest-2 pixel is placed so that
    its pair with one chosen forest-1 pixel has r_par (even pixels) or r_perp
    (odd pixels) equal to an edge 4k + e, e ~ U(-eps, eps) (one pixel in four
    with e = 0, i.e. on the edge up to FP64 rounding).  The half-angle factors
    are computed from the stored unit vectors with the FP64 formula the
    kernels use, so the targets are edges of the evaluated coordinates.
    Pairs sit at random sky positions, far apart from each other, with
    separations theta ~ U(theta[0], theta[1]) rad.  r_perp edges at small theta
    need r1 + r2 = edge / s: the default range reaches ~8e5 Mpc/h; (0.03, 0.055)
    keeps every distance below 10^4 Mpc/h, as in real data.
    """
    rng = np.random.default_rng(seed)
    bsz = rmax / nbins
    nf = 2 * n_pairs
    ra = np.empty(nf); dec = np.empty(nf)
    dcs = []
    for k in range(n_pairs):
        ra1 = rng.uniform(0.2, 2 * np.pi - 0.2)
        dec1 = np.arcsin(rng.uniform(-0.9, 0.9))
        th = rng.uniform(theta[0], theta[1])
        # forest 2 at angle th, displaced mostly in ra so that ra2 > ra1
        phi = rng.uniform(-0.6, 0.6)
        dec2 = dec1 + th * np.sin(phi)
        ra2 = ra1 + th * np.cos(phi) / np.cos(dec1)
        ra[2 * k], dec[2 * k], ra[2 * k + 1], dec[2 * k + 1] = ra1, dec1, ra2, dec2
    ux = np.cos(ra) * np.cos(dec); uy = np.sin(ra) * np.cos(dec); uz = np.sin(dec)
    for k in range(n_pairs):
        a, b = 2 * k, 2 * k + 1
        cos12 = (ux[a] * ux[b] + uy[a] * uy[b]) + uz[a] * uz[b]
        cos12 = min(1.0, max(-1.0, cos12))
        c, s = np.sqrt(0.5 * (1.0 + cos12)), np.sqrt(0.5 * (1.0 - cos12))
        r0 = rng.uniform(3700.0, 4800.0)
        r1 = r0 + np.sort(rng.uniform(0.0, 2.0 * n_pix, n_pix))
        r2 = np.empty(n_pix)
        for j in range(n_pix):
            i = rng.integers(n_pix)
            e = 0.0 if rng.random() < 0.25 else rng.uniform(-eps, eps)
            if j % 2 == 0 or s * 2 * r1[i] > rmax - 2 * bsz:
                kk = rng.integers(1, nbins)
                d = (kk * bsz + e) / c
                r2[j] = r1[i] + d if rng.random() < 0.5 else r1[i] - d
            else:
                lo = int(np.ceil(s * 2 * r1[i] / bsz)) + 1
                kk = rng.integers(lo, nbins) if lo < nbins else nbins - 1
                r2[j] = (kk * bsz + e) / s - r1[i]
        dcs.append(r1)
        dcs.append(np.sort(r2))
    n = np.array([len(x) for x in dcs])
    pix_off = np.zeros(nf + 1, dtype=np.int64)
    np.cumsum(n, out=pix_off[1:])
    dc = np.concatenate(dcs)
    P = len(dc)
    we = 1.0 / (0.05 + rng.uniform(0.0, 0.2, P))
    dw = we * rng.normal(0.0, 0.2, P)
    hp = healpix.ang2pix_ring(nside, np.pi / 2 - dec, ra)
    order = np.argsort(hp, kind="stable")
    sel = np.concatenate([np.arange(pix_off[f], pix_off[f + 1]) for f in order])
    n2 = n[order]
    off2 = np.zeros(nf + 1, dtype=np.int64)
    np.cumsum(n2, out=off2[1:])
    zero_f = np.zeros(nf)
    return Survey(ra=ra[order], dec=dec[order], zq=zero_f, name=np.arange(nf, dtype=np.int64)[order] + 7_000_000,
                  hpix=hp[order], ux=ux[order], uy=uy[order], uz=uz[order], omega=zero_f.copy(),
                  omega_dl2=zero_f.copy(), lambda_average=zero_f.copy(), pix_off=off2, loglam=np.zeros(P),
                  z=np.zeros(P), dc=dc[sel], delta=np.zeros(P), sigma2_pip=np.zeros(P), we=we[sel], dw=dw[sel],
                  delta_lambda=np.zeros(P), nside=nside, config=None)
