"""Drop-in contract: data1.npy in, per-pixel .npy files out, read the way
post-processing reads them (reference 2pla.py:52-140, distortion.py:48-125,
post_processing.py:130-147)."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import oracle
from lya2pcf_amd import forests, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _workspace(tmp_path, n=150, seed=7, params=(), **kw):
    """tmp dir with the drivers, a parameters.py pointing at tmp paths (with the lines
    ``params`` replaced: (old, new) pairs), and data1.npy."""
    for f in ("2pla.py", "distortion.py", "correlation_procedures_hip.py", "distortion_procedures_hip.py",
              "forest_class.py"):
        shutil.copy(os.path.join(ROOT, f), tmp_path / f)
    par = open(os.path.join(ROOT, "parameters.py")).read()
    par = par.replace("data_dir = './deltas_lya2pcf/'", "data_dir = '%s/deltas/'" % tmp_path)
    par = par.replace("corr_dir = './outputs/test/'", "corr_dir = '%s/out/'" % tmp_path)
    for a, b in params:
        assert a in par, a
        par = par.replace(a, b)
    (tmp_path / "parameters.py").write_text(par)
    (tmp_path / "deltas").mkdir()
    s = synth.generate(synth.scaled(synth.CONFIGS["patch2k"], n, seed=seed, **kw))   # nside 32, as parameters.py
    synth.write_data_npy(s, str(tmp_path / "deltas" / "data1.npy"))
    return s


def _run(tmp_path, *args):
    env = dict(os.environ, PYTHONPATH=ROOT)
    r = subprocess.run([sys.executable, str(tmp_path / args[0]), *args[1:]], cwd=tmp_path, env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


def test_data_npy_round_trip(tmp_path):
    """data1.npy (pickled {pixel: [quasar]}) packs to the same arrays as the survey."""
    s = synth.generate(synth.scaled(synth.CONFIGS["patch2k"], 40, seed=3))
    path = tmp_path / "data1.npy"
    synth.write_data_npy(s, str(path))
    data = np.load(path, allow_pickle=True).item()
    assert sorted(data.keys()) == list(data.keys())
    fa = forests.ForestSet.from_data(data)
    fb = forests.ForestSet.from_survey(s)
    for k in ("pix_off", "dc", "we", "dw", "ux", "uy", "uz", "ra", "name"):
        assert np.array_equal(getattr(fa, k), getattr(fb, k)), k
    q = data[next(iter(data))][0]
    for attr in ("ra", "dec", "name", "pix", "x", "y", "z", "dc", "we", "dw", "rx", "delta_lambda", "omega",
                 "omega_delta_lambda2", "lenght", "fib", "pl"):
        assert hasattr(q, attr), attr
    from forest_class import quasar
    from parameters import nside
    from lya2pcf_amd import healpix
    assert isinstance(q, quasar)
    assert q.pix == healpix.ang2pix_ring(nside, np.pi / 2 - q.dec, q.ra)


def test_quasar_neighborhood_matches_pair_list():
    """forest_class.quasar.neighborhood (healpy-free) gives the native pair list."""
    from lya2pcf_amd import xi
    s = synth.generate(synth.scaled(synth.CONFIGS["patch2k"], 120, seed=5))
    data = synth.to_data_dict(s)
    fs = forests.ForestSet.from_data(data)
    am = forests.angmax_from(fs, 200.0)
    off, idx = xi.build_pairs(fs, am)
    pos = {id(q): k for k, q in enumerate(fs.objects)}
    for k, q in enumerate(fs.objects):
        got = sorted(pos[id(f)] for f in q.neighborhood(data, am))
        assert got == list(idx[off[k]:off[k + 1]]), k


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_2pla_gpu_outputs(tmp_path, precision):
    s = _workspace(tmp_path)
    _run(tmp_path, "2pla.py", "--gpu", "--counts", "--precision", precision)
    out = tmp_path / "out"
    fs = forests.ForestSet.from_survey(s)
    am = forests.angmax_from(fs, 200.0)
    off, idx = oracle.pairs(fs, am)
    c0, w0, d0 = oracle.xi_gpu_form(fs, off, idx)
    files = sorted(out.glob("2d_histogram_pixel_*"))
    assert len(files) == len(fs.slots)
    for k, pix in enumerate(fs.slots):
        h = np.load(out / ("2d_histogram_pixel_%d.npy" % pix))
        assert h.shape == (2, 50, 50) and h.dtype == np.float64
        # post_processing.py:139 unpacks exactly two arrays
        w_hist, dw_hist = h
        cnt = np.load(out / ("2d_paircounts_pixel_%d.npy" % pix))
        assert np.array_equal(cnt, c0[k])
        np.testing.assert_allclose(w_hist, w0[k], rtol=1e-12)
        np.testing.assert_allclose(dw_hist, d0[k], rtol=1e-9, atol=1e-9 * np.abs(w0[k]).max())
    assert (out / "thread_0_of_1.log").exists()


@pytest.mark.gpu
@pytest.mark.parametrize("bin_size,rtmax", [(2, 200), (2, 120)])
def test_2pla_gpu_large_binnings(tmp_path, bin_size, rtmax):
    """A parameters.py with 2 Mpc/h bins (100 x 100, and 100 x 60 with rtmax = 120):
    more bins than the LDS histogram holds; 2pla.py --gpu runs them in bands.  Counts
    bit-exact, sums as the default binning."""
    s = _workspace(tmp_path, params=(("bin_size_r = 4", "bin_size_r = %d" % bin_size),
                                     ("rtmax = rmax", "rtmax = %d" % rtmax)))
    np_rp, np_rt = 200 // bin_size, rtmax // bin_size
    for precision in ("fp64", "mixed"):
        _run(tmp_path, "2pla.py", "--gpu", "--counts", "--precision", precision)
        out = tmp_path / "out"
        fs = forests.ForestSet.from_survey(s)
        am = forests.angmax_from(fs, float(rtmax))
        off, idx = oracle.pairs(fs, am)
        c0, w0, d0 = oracle.xi_gpu_form(fs, off, idx, np_rp, np_rt, 200.0, float(rtmax))
        assert c0.sum() > 0
        for k, pix in enumerate(fs.slots):
            w_hist, dw_hist = np.load(out / ("2d_histogram_pixel_%d.npy" % pix))
            assert w_hist.shape == (np_rp, np_rt)
            assert np.array_equal(np.load(out / ("2d_paircounts_pixel_%d.npy" % pix)), c0[k]), (precision, pix)
            np.testing.assert_allclose(w_hist, w0[k], rtol=1e-12)
            np.testing.assert_allclose(dw_hist, d0[k], rtol=1e-9, atol=1e-9 * np.abs(w0[k]).max())


@pytest.mark.gpu
def test_two_point_per_pixel_any_pixel(tmp_path):
    """A pixel outside init's compute_pixels is computed when asked for (no silent
    zeros); a pixel not in the data raises KeyError, as the reference's data[pixel]."""
    import correlation_procedures_hip as cp
    s = _workspace(tmp_path, n=120, seed=9)
    data = np.load(tmp_path / "deltas" / "data1.npy", allow_pickle=True).item()
    pixels = list(data.keys())
    fs = forests.ForestSet.from_survey(s)
    am = forests.angmax_from(fs, 200.0)
    off, idx = oracle.pairs(fs, am)
    c0, w0, d0 = oracle.xi_gpu_form(fs, off, idx)
    cp.init(data, None, (50, 50), am, compute_pixels=pixels[:1], counts=True)
    for k in [0, len(pixels) - 1, 1]:
        w, d = cp.two_point_per_pixel(pixels[k])
        assert np.array_equal(cp.pair_counts(pixels[k]), c0[k]), k
        np.testing.assert_allclose(w, w0[k], rtol=1e-12)
    assert c0[len(pixels) - 1].sum() > 0
    with pytest.raises(KeyError):
        cp.two_point_per_pixel(-5)


@pytest.mark.gpu
def test_distortion_per_pixel_api(tmp_path):
    """The reference's own distortion entry point (distortion.py:101-125): init, then
    distortion_per_pixel(data[pixel]) for every pixel in dict order, summed in FP64 and
    divided by the summed weights.  Within 1e-10 of the oracle's η algebra."""
    import distortion_procedures_hip as dp
    from lya2pcf_amd import distortion as dist
    s = _workspace(tmp_path, n=60, seed=8, length_range=(30, 60))
    data = np.load(tmp_path / "deltas" / "data1.npy", allow_pickle=True).item()
    fs = forests.ForestSet.from_survey(s)
    am = forests.angmax_from(fs, 200.0)
    # a fresh subsample stream, as the reference's random.seed(1) at import
    import random
    random.seed(1)
    dp.init(data, None, (50, 50), am, 0.8)
    D = np.zeros((2500, 2500))
    wA = np.zeros(2500)
    for pixel in data:
        d, w = dp.distortion_per_pixel(data[pixel])
        assert d.shape == (2500, 2500) and w.shape == (2500,)
        D += d
        wA += w
    off, idx = oracle.pairs(fs, am)
    pairs = dist.subsample(fs, off, idx, range(fs.nf), 0.8, rng=random.Random(1))
    D0, w0 = oracle.distortion(fs, pairs)
    with np.errstate(divide="ignore", invalid="ignore"):
        E, G = D0 / w0[:, None], D / wA[:, None]
    ok = w0 > 0
    assert np.array_equal(np.isfinite(G).all(1), ok)
    assert np.max(np.abs(G[ok] - E[ok])) <= 1e-10 * np.abs(E[ok]).max()


@pytest.mark.gpu
def test_distortion_driver_large_binning(tmp_path):
    """distortion.py with 2 Mpc/h bins: D is 10,000 x 10,000; the LDS is budgeted by
    each pair's active bins."""
    import random
    from lya2pcf_amd import distortion as dist
    s = _workspace(tmp_path, n=30, seed=10, length_range=(20, 40), params=(("bin_size_r = 4", "bin_size_r = 2"),))
    _run(tmp_path, "distortion.py", "--excluded", "0.5")
    D = np.load(tmp_path / "out" / "distortion.npy", mmap_mode="r")
    assert D.shape == (10000, 10000) and D.dtype == np.float64
    fs = forests.ForestSet.from_survey(s)
    am = forests.angmax_from(fs, 200.0)
    off, idx = oracle.pairs(fs, am)
    pairs = dist.subsample(fs, off, idx, range(fs.nf), 0.5, rng=random.Random(1))
    D0, w0 = oracle.distortion(fs, pairs, 100, 100)
    ok = np.nonzero(w0 > 0)[0]
    assert len(ok) > 0
    E = D0[ok] / w0[ok, None]
    G = np.asarray(D[ok])
    assert np.isfinite(G).all()
    assert np.max(np.abs(G - E)) <= 1e-10 * np.abs(E).max()
    del D0
    assert not np.isfinite(np.asarray(D[np.nonzero(w0 == 0)[0][:50]])).any()


@pytest.mark.gpu
def test_distortion_driver(tmp_path):
    import random
    from lya2pcf_amd import distortion as dist
    s = _workspace(tmp_path, n=60, seed=8, length_range=(30, 60))
    _run(tmp_path, "distortion.py", "--excluded", "0.8")
    D = np.load(tmp_path / "out" / "distortion.npy")
    assert D.shape == (2500, 2500) and D.dtype == np.float64
    fs = forests.ForestSet.from_survey(s)
    am = forests.angmax_from(fs, 200.0)
    off, idx = oracle.pairs(fs, am)
    pairs = dist.subsample(fs, off, idx, range(fs.nf), 0.8, rng=random.Random(1))
    D0, w0 = oracle.distortion(fs, pairs)
    with np.errstate(divide="ignore", invalid="ignore"):
        E = D0 / w0[:, None]
    ok = w0 > 0
    assert np.array_equal(np.isfinite(D).all(1), ok)
    scale = np.abs(E[ok]).max()
    assert np.max(np.abs(D[ok] - E[ok])) <= 1e-10 * scale


# ------------------------------------------------------------ multi-file drivers
def _two_files(tmp_path):
    """data1.npy and data2.npy: two surveys over the same patch (shared pixels)."""
    s1 = _workspace(tmp_path, n=80, seed=21, length_range=(30, 80))
    s2 = synth.generate(synth.scaled(synth.CONFIGS["patch2k"], 70, seed=22, length_range=(30, 80)))
    synth.write_data_npy(s2, str(tmp_path / "deltas" / "data2.npy"))
    for f in ("2pla_multiple_data.py", "distortion_multiple_data.py"):
        shutil.copy(os.path.join(ROOT, f), tmp_path / f)
    return [s1, s2]


def test_driver_helpers(tmp_path):
    from lya2pcf_amd import drivers
    for name in ("data2.npy", "data10.npy", "data1.npy", "other.npy"):
        (tmp_path / name).write_bytes(b"")
    files = drivers.data_files(str(tmp_path))
    assert [os.path.basename(f) for f in files] == ["data1.npy", "data10.npy", "data2.npy"]
    parts = [drivers.files_of_rank(str(tmp_path), r, 2) for r in range(2)]
    assert parts[0] + parts[1] == files and len(parts[0]) == 2
    with pytest.raises(FileNotFoundError):
        drivers.data_files(str(tmp_path / "nothing"))
    s = synth.generate(synth.scaled(synth.CONFIGS["patch2k"], 30, seed=4))
    dmin, am = drivers.cut_angle(synth.to_data_dict(s), 200.0)
    fs = forests.ForestSet.from_survey(s)
    assert dmin == fs.dc[fs.pix_off[:-1]].min() and am == forests.angmax_from(fs, 200.0)
    with pytest.raises(NameError):
        drivers.cut_angle({}, 200.0)


@pytest.mark.gpu
def test_2pla_multiple_data_outputs(tmp_path):
    surveys = _two_files(tmp_path)
    _run(tmp_path, "2pla_multiple_data.py", "--gpu", "--counts")
    out = tmp_path / "out"
    expect = {}
    for s in surveys:  # sorted file order: the later file wins a shared pixel
        fs = forests.ForestSet.from_survey(s)
        off, idx = oracle.pairs(fs, forests.angmax_from(fs, 200.0))  # pairs within the file only
        c, w, d = oracle.xi_gpu_form(fs, off, idx)
        for k, pix in enumerate(fs.slots):
            expect[pix] = (c[k], w[k], d[k])
    shared = set(forests.ForestSet.from_survey(surveys[0]).slots) & set(forests.ForestSet.from_survey(surveys[1]).slots)
    assert shared, "the two files should share pixels"
    files = sorted(out.glob("2d_histogram_pixel_*"))
    assert len(files) == len(expect)
    for pix, (c0, w0, d0) in expect.items():
        w_hist, dw_hist = np.load(out / ("2d_histogram_pixel_%d.npy" % pix))
        assert np.array_equal(np.load(out / ("2d_paircounts_pixel_%d.npy" % pix)), c0)
        np.testing.assert_allclose(w_hist, w0, rtol=1e-12)
        np.testing.assert_allclose(dw_hist, d0, rtol=1e-9, atol=1e-9 * np.abs(w0).max())


@pytest.mark.gpu
def test_distortion_multiple_data(tmp_path):
    import random
    from lya2pcf_amd import distortion as dist
    surveys = _two_files(tmp_path)
    _run(tmp_path, "distortion_multiple_data.py", "--excluded", "0.8")
    D = np.load(tmp_path / "out" / "distortion.npy")
    rng = random.Random(1)  # one stream, continued from file to file
    D0 = np.zeros((2500, 2500))
    w0 = np.zeros(2500)
    for s in surveys:
        fs = forests.ForestSet.from_survey(s)
        off, idx = oracle.pairs(fs, forests.angmax_from(fs, 200.0))
        d, w = oracle.distortion(fs, dist.subsample(fs, off, idx, range(fs.nf), 0.8, rng=rng))
        D0 += d
        w0 += w
    with np.errstate(divide="ignore", invalid="ignore"):
        E = D0 / w0[:, None]
    ok = w0 > 0
    assert np.array_equal(np.isfinite(D).all(1), ok)
    assert np.max(np.abs(D[ok] - E[ok])) <= 1e-10 * np.abs(E[ok]).max()


@pytest.mark.gpu
def test_postprocessing_of_gpu_outputs(tmp_path):
    """End to end: 2pla.py --gpu and distortion.py write their files, the
    post-processing consumer (reference post_processing.py:122-217) reads them, and
    xi, the smoothed covariance and the COR table equal the same consumer run on
    oracle histograms and the oracle distortion."""
    import random
    from lya2pcf_amd import distortion as dist, fits, postproc
    s = _workspace(tmp_path, n=120, seed=12, length_range=(40, 120))
    _run(tmp_path, "2pla.py", "--gpu")
    _run(tmp_path, "distortion.py", "--excluded", "0.8")
    out = tmp_path / "out"
    corr, err = postproc.main(["--corr-dir", str(out)])
    # the same consumer on oracle outputs
    ref = tmp_path / "ref"
    ref.mkdir()
    fs = forests.ForestSet.from_survey(s)
    am = forests.angmax_from(fs, 200.0)
    off, idx = oracle.pairs(fs, am)
    c0, w0, d0 = oracle.xi_gpu_form(fs, off, idx)
    for k, pix in enumerate(fs.slots):
        np.save(ref / ("2d_histogram_pixel_%d" % pix), (w0[k], d0[k]))
    pairs = dist.subsample(fs, off, idx, range(fs.nf), 0.8, rng=random.Random(1))
    D0, wA0 = oracle.distortion(fs, pairs)
    with np.errstate(divide="ignore", invalid="ignore"):
        np.save(ref / "distortion", D0 / wA0[:, None])
    corr0, err0 = postproc.main(["--corr-dir", str(ref)])
    x, x0 = np.load(out / "correlation_2d.npy"), np.load(ref / "correlation_2d.npy")
    assert x.shape == (50, 50) and np.abs(x0).max() > 0
    np.testing.assert_allclose(x, x0, rtol=1e-9, atol=1e-10 * np.abs(x0).max())
    C, C0 = np.load(out / "covariance.npy"), np.load(ref / "covariance.npy")
    np.testing.assert_allclose(C, C0, rtol=1e-8, atol=1e-9 * np.nanmax(np.abs(C0)))
    t = fits.by_name(fits.read(out / "correlation.out.gz"), "COR")
    t0 = fits.by_name(fits.read(ref / "correlation.out.gz"), "COR")
    assert t.header["NP"] == t0.header["NP"] == 50
    np.testing.assert_allclose(t.data["DA"], t0.data["DA"], rtol=1e-9, atol=1e-10 * np.abs(x0).max())
    np.testing.assert_allclose(t.data["CO"], t0.data["CO"], rtol=1e-8, atol=1e-9 * np.nanmax(np.abs(C0)))
    dm, dm0 = t.data["DM"], t0.data["DM"]
    ok = np.isfinite(dm0).all(1)
    assert np.array_equal(np.isfinite(dm).all(1), ok) and ok.any()
    assert np.max(np.abs(dm[ok] - dm0[ok])) <= 1e-10 * np.abs(dm0[ok]).max()
    assert np.array_equal(t.data["RP"], t0.data["RP"]) and np.array_equal(t.data["RT"], t0.data["RT"])
