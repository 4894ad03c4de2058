"""Post-processing consumer (reference post_processing.py:30-169) reading the
histogram files the drop-in driver writes."""
import numpy as np
import pytest

import oracle
from lya2pcf_amd import postproc


def _loops_cov(xi, weight):
    n = xi.shape[1]
    m = np.array([np.sum(xi[:, i] * weight[:, i]) / np.sum(weight[:, i]) for i in range(n)])
    C = np.zeros((n, n))
    for a in range(n):
        for b in range(n):
            C[a, b] = np.sum(weight[:, a] * (xi[:, a] - m[a]) * weight[:, b] * (xi[:, b] - m[b])) / (
                np.sum(weight[:, a]) * np.sum(weight[:, b]))
    return C, m


def test_cov_small_against_loops_and_invariants():
    rng = np.random.default_rng(0)
    xi = rng.normal(size=(7, 9))
    w = rng.uniform(0.5, 2.0, size=(7, 9))
    C, m = postproc.cov(xi, w)
    C2, m2 = _loops_cov(xi, w)
    np.testing.assert_allclose(C, C2, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(m, m2, rtol=1e-13)
    # identical subsamples: zero covariance, mean = the common value
    C3, m3 = postproc.cov(np.tile(xi[0], (5, 1)), w[:5])
    assert np.max(np.abs(C3)) < 1e-28
    np.testing.assert_allclose(m3, xi[0], rtol=1e-14)


def test_cov_smooth_keeps_step_stationary_matrix():
    rp, rt = postproc.bin_coordinates(3, 3)
    n = 9
    sig = np.linspace(1, 2, n)
    rho = lambda dp, dt: 0.5 ** (dp + 2 * dt)
    corr = np.array([[rho(abs(rp[a] - rp[b]), abs(rt[a] - rt[b])) for b in range(n)] for a in range(n)])
    C = corr * np.outer(sig, sig)
    S = postproc.cov_smooth(np.zeros((4, n)), np.ones((4, n)), rp, rt, C)
    np.testing.assert_allclose(S, C, rtol=1e-13)


def test_diag_error_closed_form():
    da = np.array([[1.0, 2.0], [3.0, 2.0]])
    we = np.array([[1.0, 1.0], [1.0, 3.0]])
    err, m = postproc.diag_error(da, we)
    np.testing.assert_allclose(m, [2.0, 2.0])
    np.testing.assert_allclose(err, [np.sqrt(2.0) / 2.0, 0.0])


def test_mean_xi_equals_summed_histograms(tmp_path, small_set, small_pairs):
    """mean over pixels of w-weighted xi_p equals sum(dw)/sum(w) of all pixels."""
    am, (off, idx) = small_pairs
    c, w, d = oracle.xi_gpu_form(small_set, off, idx)
    for k, pix in enumerate(small_set.slots):
        np.save(tmp_path / ("2d_histogram_pixel_%d" % pix), (w[k], d[k]))
        np.save(tmp_path / ("2d_paircounts_pixel_%d" % pix), c[k])
    da, we, shape = postproc.read_histograms(str(tmp_path))
    assert da.shape[0] == len(small_set.slots) and shape == (50, 50)
    _, mean = postproc.cov(da, we)
    W, D = w.sum(0).ravel(), d.sum(0).ravel()
    ok = W > 0
    np.testing.assert_allclose(mean[ok], D[ok] / W[ok], rtol=1e-10, atol=1e-16)
    corr, err = postproc.main(["--corr-dir", str(tmp_path), "--diagonal-error"])
    assert corr.shape == (2500,) and (tmp_path / "correlation_2d.npy").exists()


def test_fits_table_with_distortion(tmp_path, small_set, small_pairs):
    """correlation.out.gz: the COR table of post_processing.py:186-216, read back with
    the package's FITS reader (pinned separately by a hand-built FITS file)."""
    from lya2pcf_amd import fits
    am, (off, idx) = small_pairs
    c, w, d = oracle.xi_gpu_form(small_set, off, idx)
    for k, pix in enumerate(small_set.slots):
        np.save(tmp_path / ("2d_histogram_pixel_%d" % pix), (w[k], d[k]))
    dist = np.random.default_rng(3).normal(size=(2500, 2500))
    np.save(tmp_path / "distortion", dist)
    corr, err = postproc.main(["--corr-dir", str(tmp_path)])
    hdus = fits.read(tmp_path / "correlation.out.gz")
    t = fits.by_name(hdus, "COR")
    assert t.header["NP"] == 50 and t.header["NT"] == 50 and t.header["RPMAX"] == 200.0 and t.header["WL"] == -1
    assert np.array_equal(t.data["DA"], corr.ravel())
    np.testing.assert_array_equal(t.data["DM"], dist)
    np.testing.assert_array_equal(t.data["CO"], np.load(tmp_path / "covariance.npy"))
    assert t.data["RP"][0] == 2.0 and t.data["RT"][1] == 6.0 and t.data["RP"][50] == 6.0
    # without the distortion file: a warning and no FITS file
    (tmp_path / "correlation.out.gz").unlink()
    (tmp_path / "distortion.npy").unlink()
    with pytest.warns(UserWarning, match="distortion"):
        postproc.main(["--corr-dir", str(tmp_path), "--diagonal-error"])
    assert not (tmp_path / "correlation.out.gz").exists()
