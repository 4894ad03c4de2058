"""Pinning the oracle and the data model without the reference (which cannot run
here: numba, healpy, mpi4py and pycuda are absent).  Each check is chosen so
that a plausible mistake (dropped term, wrong sign or index, transposed
operand, wrong rounding direction) fails it."""
import numpy as np
import pytest

import oracle
from lya2pcf_amd import cosmology, forests, healpix, synth, xi


# ------------------------------------------------------------------ healpix
@pytest.mark.parametrize("nside", [1, 2, 4, 8, 16, 32])
def test_healpix_centres_round_trip(nside):
    pix = np.arange(12 * nside * nside)
    th, ph = healpix.pix2ang_ring(nside, pix)
    assert np.array_equal(healpix.ang2pix_ring(nside, th, ph), pix)


def test_healpix_known_values_and_equal_area():
    # nside=1: the north cap holds pixels 0-3 by quadrant, the south cap 8-11
    assert healpix.ang2pix_ring(1, 0.1, 0.1) == 0
    assert healpix.ang2pix_ring(1, 0.1, np.pi / 2 + 0.1) == 1
    assert healpix.ang2pix_ring(1, np.pi - 0.1, 0.1) == 8
    assert healpix.ang2pix_ring(1, np.pi / 2, 0.0) == 4   # equator at phi=0 is pixel 4's corner region
    # uniform points fill the 3072 pixels of nside 16 equally (chi^2 test)
    rng = np.random.default_rng(0)
    n = 3_000_000
    z = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    cnt = np.bincount(healpix.ang2pix_ring(16, np.arccos(z), ph), minlength=3072)
    e = n / 3072
    chi2 = np.sum((cnt - e) ** 2 / e)
    assert chi2 < 3072 + 6 * np.sqrt(2 * 3072)


def test_disc_candidates_superset():
    rng = np.random.default_rng(1)
    nside = 16
    pts_th = np.arccos(rng.uniform(-1, 1, 20000))
    pts_ph = rng.uniform(0, 2 * np.pi, 20000)
    pix = healpix.ang2pix_ring(nside, pts_th, pts_ph)
    v = np.stack([np.sin(pts_th) * np.cos(pts_ph), np.sin(pts_th) * np.sin(pts_ph), np.cos(pts_th)], 1)
    for k in range(20):
        c = v[k]
        rad = 0.06
        inside = v @ c > np.cos(rad)
        cand = set(healpix.disc_candidates(nside, c, rad, set(pix.tolist())))
        assert set(pix[inside].tolist()) <= cand


# ------------------------------------------------------------------ cosmology
def test_comoving_distance_table():
    from scipy import integrate
    d = cosmology.ComovingDistance(0.3153)
    for z in (0.5, 2.1, 2.7, 3.5, 9.0):
        q = 2997.92458 * integrate.quad(lambda x: 1 / np.sqrt(0.3153 * (1 + x) ** 3 + 0.6847), 0, z,
                                        epsabs=0, epsrel=1e-13)[0]
        assert abs(d(z) - q) < 1e-9 * q
        assert abs(d.exact(z) - q) < 1e-12 * q
    # Einstein-de Sitter closed form: d = 2 c/H0 (1 - 1/sqrt(1+z))
    eds = cosmology.ComovingDistance(1.0)
    for z in (1.0, 3.0):
        assert abs(eds(z) - 2 * 2997.92458 * (1 - 1 / np.sqrt(1 + z))) < 1e-9 * eds(z)


# ------------------------------------------------------------------ pair loop
def _two_forests(dc1, dc2, theta, we1=None, we2=None, dw1=None, dw2=None):
    dc1, dc2 = np.asarray(dc1, float), np.asarray(dc2, float)
    n1, n2 = len(dc1), len(dc2)
    ux = np.array([1.0, np.cos(theta)])
    uy = np.array([0.0, np.sin(theta)])
    return forests.ForestSet(np.array([0, n1, n1 + n2]), [0], ux=ux, uy=uy, uz=np.zeros(2),
                             ra=np.array([0.0, theta]), dec=np.zeros(2), name=np.array([1, 2]),
                             slot=np.zeros(2, np.int32), omega=np.ones(2), odl2=np.ones(2),
                             dc=np.concatenate([dc1, dc2]),
                             we=np.concatenate([np.ones(n1) if we1 is None else we1, np.ones(n2) if we2 is None else we2]),
                             dw=np.concatenate([np.ones(n1) if dw1 is None else dw1, np.ones(n2) if dw2 is None else dw2]),
                             dl=np.zeros(n1 + n2))


def test_line_of_sight_pair_known_answer():
    """theta = 0: rt = 0 and rp = |dr| exactly, so each pair lands in bin (floor(|dr|/4), 0)."""
    fs = _two_forests([4000.0, 4010.0], [4003.0, 4099.5, 4199.0], 0.0)
    off = np.array([0, 1, 1]); idx = np.array([1], np.int32)
    cnt, w, d = oracle.xi_gpu_form(fs, off, idx)
    exp = np.zeros((50, 50), np.uint64)
    for a in (4000.0, 4010.0):
        for b in (4003.0, 4099.5, 4199.0):
            k = int(abs(a - b) // 4)
            if k < 50:
                exp[k, 0] += 1
    assert np.array_equal(cnt[0], exp)
    # rp == 200 exactly (4199 - 3999 style) is out of range; 4199.0-4000.0 = 199 -> bin 49
    assert cnt[0, 49, 0] == 1


def test_edge_goes_to_upper_bin_and_rmax_excluded():
    fs = _two_forests([4000.0], [4008.0, 4200.0, 4199.999999], 0.0)
    cnt, _, _ = oracle.xi_gpu_form(fs, np.array([0, 1, 1]), np.array([1], np.int32))
    assert cnt[0, 2, 0] == 1      # rp = 8 exactly -> bin 2 (floor)
    assert cnt[0].sum() == 2      # rp = 200 excluded, rp = 199.999999 kept in bin 49
    assert cnt[0, 49, 0] == 1


def test_perpendicular_closed_form():
    """Equal distances: rp = 0, rt = 2 r sin(theta/2)."""
    th = 0.01
    r = 4000.0
    fs = _two_forests([r], [r], th)
    cnt, _, _ = oracle.xi_gpu_form(fs, np.array([0, 1, 1]), np.array([1], np.int32))
    k = int(np.floor(2 * r * np.sin(th / 2) / 4))
    assert cnt[0, 0, k] == 1 and cnt.sum() == 1


def test_oracle_vs_numpy_brute_force(small_set, small_pairs):
    am, (off, idx) = small_pairs
    # restrict numpy to the first 40 forests as forest 1 to keep it fast
    lim = off[40]
    off2 = np.minimum(off, lim)
    c2, w2, d2 = oracle.numpy_xi_gpu_form(small_set, off2, idx[:lim])
    f1 = np.arange(40, dtype=np.int32)
    c3, w3, d3 = oracle.xi_gpu_form(small_set, off2, idx[:lim], f1_list=f1)
    assert np.array_equal(c2, c3)
    np.testing.assert_allclose(w2, w3, rtol=1e-12)
    _, _, a3 = oracle.xi_gpu_form(small_set, off2, idx[:lim], f1_list=f1, dw=np.abs(small_set.dw))
    assert np.all(np.abs(d2 - d3) <= 1e-12 * a3 + 1e-300)


def test_invariants(small_set, small_pairs):
    am, (off, idx) = small_pairs
    fs = small_set
    c, w, d = oracle.xi_gpu_form(fs, off, idx)
    # global weight scale: xi invariant, W scales by k^2
    k = 3.0
    c2, w2, d2 = oracle.xi_gpu_form(fs, off, idx, we=fs.we * k, dw=fs.dw * k)
    assert np.array_equal(c, c2)
    np.testing.assert_allclose(w2, w * k * k, rtol=1e-13)
    ok = w > 0
    np.testing.assert_allclose(d2[ok] / w2[ok], d[ok] / w[ok], rtol=1e-10, atol=1e-16)
    # constant dw = w (delta = 1, unprojected): xi = 1 in every populated bin
    c3, w3, d3 = oracle.xi_gpu_form(fs, off, idx, dw=fs.we)
    np.testing.assert_allclose(d3[ok] / w3[ok], 1.0, rtol=1e-12)


def test_swapping_roles_keeps_bins(small_set, small_pairs):
    """Pair (f1, f2) and (f2, f1) give the same bins (|r1-r2|, r1+r2 symmetric)."""
    am, (off, idx) = small_pairs
    fs = small_set
    toff, tidx = xi.transpose_pairs(off, idx, fs.nf)
    one = forests.ForestSet(fs.pix_off, [0], **{k: getattr(fs, k) for k in fs.FOREST + fs.PIXEL})
    one.slot = np.zeros(fs.nf, np.int32)
    c1, w1, _ = oracle.xi_gpu_form(one, off, idx)
    c2, w2, _ = oracle.xi_gpu_form(one, toff, tidx)
    assert np.array_equal(c1, c2)
    np.testing.assert_allclose(w1, w2, rtol=1e-12)


def test_superset_pair_list_same_histograms(small_set, small_pairs):
    """Pairs beyond angmax have rt >= rtmax: a larger cut adds no binned pair."""
    am, (off, idx) = small_pairs
    c1, _, _ = oracle.xi_gpu_form(small_set, off, idx)
    off2, idx2 = xi.build_pairs(small_set, am * 1.3)
    assert len(idx2) > len(idx)
    c2, _, _ = oracle.xi_gpu_form(small_set, off2, idx2)
    assert np.array_equal(c1, c2)


def test_cpu_form_agrees_with_gpu_form(small_set, small_pairs):
    """The two reference FP64 formulas (arccos of ra/dec vs sqrt of the dot
    product) may differ only for pixel pairs within ~1e-9 of a bin edge."""
    am, (off, idx) = small_pairs
    c1, w1, _ = oracle.xi_gpu_form(small_set, off, idx)
    c2, w2, _ = oracle.xi_cpu_form(small_set, off, idx, am, 2. / 3600. * 3.14159 / 180.)
    diff = np.abs(c1.astype(np.int64) - c2.astype(np.int64)).sum()
    assert diff <= 2e-8 * c1.sum()
    c3, _, _ = oracle.xi_cpu_form(small_set, off, idx, am, 2. / 3600. * 3.14159 / 180., threads=4)
    assert np.array_equal(c2, c3)


def test_golden_fixture():
    """Committed inputs and oracle outputs (tests/golden/xi_tiny.npz)."""
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "xi_tiny.npz")
    z = np.load(path)
    fs = forests.ForestSet(z["pix_off"], list(z["slots"]), ux=z["ux"], uy=z["uy"], uz=z["uz"], ra=z["ra"],
                           dec=z["dec"], name=z["name"], slot=z["slot"], omega=z["omega"], odl2=z["odl2"],
                           dc=z["dc"], we=z["we"], dw=z["dw"], dl=z["dl"])
    off, idx = xi.build_pairs(fs, float(z["angmax"]))
    assert np.array_equal(off, z["nb_off"]) and np.array_equal(idx, z["nb_idx"])
    c, w, d = oracle.xi_gpu_form(fs, off, idx)
    assert np.array_equal(c, z["cnt"])
    np.testing.assert_allclose(w, z["w"], rtol=1e-13)
    np.testing.assert_allclose(d, z["dw_sum"], rtol=1e-9, atol=1e-9)


def test_multicore_gpu_form_and_subset_rows(small_set, small_pairs):
    """The OpenMP GPU form equals the sequential one (counts exactly, sums to 1e-12 of
    the absolute sums), and pairs_of gives the O(N^2) rows of the forests asked for."""
    am, (off, idx) = small_pairs
    c0, w0, d0 = oracle.xi_gpu_form(small_set, off, idx)
    _, aw, ad = oracle.xi_gpu_form(small_set, off, idx, we=np.abs(small_set.we), dw=np.abs(small_set.dw))
    c1, w1, d1 = oracle.xi_gpu_form(small_set, off, idx, threads=4)
    assert np.array_equal(c0, c1)
    assert np.all(np.abs(w1 - w0) <= 1e-12 * aw) and np.all(np.abs(d1 - d0) <= 1e-12 * ad)
    keep = np.nonzero(small_set.slot == small_set.slot[len(small_set.slot) // 2])[0]
    soff, sidx = oracle.pairs_of(small_set, am, keep)
    for f in range(small_set.nf):
        row = sidx[soff[f]:soff[f + 1]]
        assert np.array_equal(row, idx[off[f]:off[f + 1]] if f in set(keep) else row[:0]), f
    c2, _, _ = oracle.xi_gpu_form(small_set, soff, sidx, threads=3)
    k = int(small_set.slot[keep[0]])
    assert np.array_equal(c2[k], c0[k]) and c2.sum() == c0[k].sum()
