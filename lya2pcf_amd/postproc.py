"""Consumer of the per-pixel histograms: mean xi, subsample covariance,
smoothing and diagonal errors (reference post_processing.py:13-109, 122-169),
vectorised with numpy (the reference uses numba loops).  When
``distortion.npy`` is present, the combined FITS table ``correlation.out.gz``
(HDU ``COR``: DA, RP, RT, CO, DM and the binning/cosmology header, reference
post_processing.py:186-216) is written with the package's FITS writer.

    python -m lya2pcf_amd.postproc [--corr-dir DIR] [--write-coordinates] [--diagonal-error]

Reads every ``2d_histogram_pixel_*`` file of the directory (the glob of
post_processing.py:122), so the counts files written by ``2pla.py --counts``
(``2d_paircounts_pixel_*``) are not picked up.
"""
import argparse
import glob
import os
import warnings

import numpy as np


def bin_coordinates(numpix_r_p, numpix_r_t):
    """Integer bin labels 1..n, r_p major (post_processing.py:13-28)."""
    rp, rt = np.meshgrid(np.arange(1, numpix_r_p + 1), np.arange(1, numpix_r_t + 1), indexing="ij")
    return rp.ravel(), rt.ravel()


def read_histograms(corr_dir):
    """(xi per pixel [n_pix, nb], weight per pixel [n_pix, nb], bin shape)."""
    files = sorted(glob.glob(os.path.join(corr_dir, "2d_histogram_pixel_*")))
    if not files:
        raise FileNotFoundError("no 2d_histogram_pixel_* files in " + corr_dir)
    we, da, shape = [], [], None
    for f in files:
        w_hist, dw_hist = np.load(f)
        shape = shape or w_hist.shape
        w = w_hist.ravel()
        d = dw_hist.ravel()
        we.append(w)
        da.append(np.divide(d, w, out=np.zeros_like(d), where=w != 0))
    return np.array(da), np.array(we), shape


def cov(xi, weight):
    """Weighted mean per bin and covariance of the pixel subsamples
    (post_processing.py:30-52): C_AB = sum_p w_pA (xi_pA - m_A) w_pB (xi_pB - m_B) / (W_A W_B)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        W = weight.sum(0)
        means = (xi * weight).sum(0) / W
        X = weight * (xi - means)
        C = (X.T @ X) / np.outer(W, W)
    return C, means


def cov_smooth(xi, weight, r_p, r_t, covariance):
    """Average the correlation matrix over equal (|d r_p|, |d r_t|) steps
    (post_processing.py:54-91)."""
    variance = np.diagonal(covariance)
    if np.any(variance == 0.):
        warnings.warn("Data has some empty bins. Returning the unsmoothed covariance")
        return covariance
    norm = np.sqrt(np.outer(variance, variance))
    corr = covariance / norm
    n = variance.shape[0]
    nbins = int(np.sqrt(xi.shape[1]))
    a, b = np.triu_indices(n, 1)
    srp = np.rint(np.abs(r_p[b] - r_p[a])).astype(np.int64)
    srt = np.rint(np.abs(r_t[b] - r_t[a])).astype(np.int64)
    key = srp * nbins + srt
    s = np.bincount(key, weights=corr[a, b], minlength=nbins * nbins)
    c = np.bincount(key, minlength=nbins * nbins)
    with np.errstate(divide="ignore", invalid="ignore"):
        avg = s / c
    sm = np.empty_like(corr)
    sm[a, b] = avg[key]
    sm[b, a] = avg[key]
    np.fill_diagonal(sm, np.diagonal(corr))
    return sm * norm


def diag_error(da, we):
    """Weighted mean and its diagonal error (post_processing.py:94-109)."""
    mda = (da * we).sum(axis=0)
    swe = we.sum(axis=0)
    w = swe > 0.
    mda[w] /= swe[w]
    wda = we * (da - mda)
    error = np.sqrt((wda ** 2).sum(axis=0))
    error[w] /= swe[w]
    return error, mda


def write_fits(path, correlation, covariance, distortion, np_rp, np_rt, rpmax, rtmax, omega_m):
    """The ``COR`` table of post_processing.py:186-216: one row per bin (r_p major)
    with DA (xi), RP and RT (bin centres), CO (covariance row) and DM (distortion
    row)."""
    from . import fits
    n = np_rp * np_rt
    i, j = np.meshgrid(np.arange(np_rp), np.arange(np_rt), indexing="ij")
    cols = {"DA": np.asarray(correlation, dtype=np.float64).ravel(),
            "RP": ((i + 0.5) * rpmax / np_rp).ravel(), "RT": ((j + 0.5) * rtmax / np_rt).ravel(),
            "CO": np.asarray(covariance, dtype=np.float64).reshape(n, n),
            "DM": np.asarray(distortion, dtype=np.float64).reshape(n, n)}
    header = {"RPMIN": 0, "RPMAX": float(rpmax), "RTMAX": float(rtmax), "NP": int(np_rp), "NT": int(np_rt),
              "OMEGAM": float(omega_m), "OMEGAR": 0, "OMEGAK": 0, "WL": -1}
    fits.write(path, [fits.table_bytes(cols, "COR", header)])


def main(argv=None):
    try:
        import parameters as P
        corr_default, np_rp, np_rt, rpmax, rtmax = P.corr_dir, P.numpix_rp, P.numpix_rt, P.rpmax, P.rtmax
        omega_m = float(P.Omm)
    except Exception:
        corr_default, np_rp, np_rt, rpmax, rtmax = "./outputs/test/", 50, 50, 200, 200
        omega_m = 0.3153
    ap = argparse.ArgumentParser(description="Correlation and covariance from the histogram files.")
    ap.add_argument("--corr-dir", default=corr_default)
    ap.add_argument("--write-coordinates", action="store_true")
    ap.add_argument("--diagonal-error", action="store_true")
    a = ap.parse_args(argv)
    da, we, shape = read_histograms(a.corr_dir)
    if a.diagonal_error:
        error, correlation = diag_error(da, we)
    else:
        cnot, correlation = cov(da, we)
        rp, rt = bin_coordinates(np_rp, np_rt)
        covariance = cov_smooth(da, we, rp, rt, cnot)
        error = np.sqrt(np.diagonal(covariance))
        np.save(os.path.join(a.corr_dir, "covariance"), covariance)
        np.save(os.path.join(a.corr_dir, "covariance_not_smooth"), cnot)
    np.save(os.path.join(a.corr_dir, "correlation_2d"), correlation.reshape(shape))
    np.save(os.path.join(a.corr_dir, "error_2d"), error.reshape(shape))
    if a.write_coordinates:
        i, j = np.meshgrid(np.arange(np_rp), np.arange(np_rt), indexing="ij")
        np.save(os.path.join(a.corr_dir, "rp"), (i + 0.5) * rpmax / np_rp)
        np.save(os.path.join(a.corr_dir, "rt"), (j + 0.5) * rtmax / np_rt)
    dpath = os.path.join(a.corr_dir, "distortion.npy")
    if not os.path.exists(dpath):
        warnings.warn("No distortion.npy found. Compute it first if you want to obtain a fits.gz file with all "
                      "the outputs.")
        return correlation, error
    # with --diagonal-error the table carries the diagonal covariance error**2 (the
    # reference has no covariance there and stops with a NameError)
    cov_out = covariance if not a.diagonal_error else np.diag(error ** 2)
    write_fits(os.path.join(a.corr_dir, "correlation.out.gz"), correlation, cov_out, np.load(dpath),
               np_rp, np_rt, rpmax, rtmax, omega_m)
    return correlation, error


if __name__ == "__main__":
    main()
