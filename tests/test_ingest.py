"""Delta-file ingest (reference delta_reader.py:25-150, delta_reader_eboss.py:25-85)
and the minimal FITS reader/writer it uses.

fitsio and astropy are not installed, so no reference-written FITS file can be
read here: the reader is pinned by a FITS byte string assembled card by card
in this file (independent of the writer), then the writer by round trips.
The ingest is checked against the raw arrays the synthetic files were written
from, and the neighbour attributes against a brute-force per-forest rule.
"""
import gzip
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import oracle
from lya2pcf_amd import fits, forests, ingest, synth, xi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ------------------------------------------------------------------- FITS
def _cards(*cards):
    text = "".join(c.ljust(80) for c in cards) + "END".ljust(80)
    return (text + " " * ((-len(text)) % 2880)).encode("ascii")


def _pad(b):
    return b + b"\0" * ((-len(b)) % 2880)


def test_reader_on_hand_built_file():
    img = (np.arange(6.0).reshape(2, 3) * 1.5).astype(">f8")
    rows = np.zeros(2, dtype=[("A", ">i4"), ("B", ">f8", (2,)), ("C", "S4"), ("D", "i1")])
    rows["A"] = [7, -3]
    rows["B"] = [[1.0, 2.0], [3.0, 4.5]]
    rows["C"] = [b"ab", b"wxyz"]
    rows["D"] = [ord("T"), ord("F")]
    buf = (_cards("SIMPLE  =                    T", "BITPIX  =                    8", "NAXIS   =                    0",
                  "EXTEND  =                    T")
           + _cards("XTENSION= 'IMAGE   '", "BITPIX  =                  -64", "NAXIS   =                    2",
                    "NAXIS1  =                    3", "NAXIS2  =                    2", "PCOUNT  =                    0",
                    "GCOUNT  =                    1", "EXTNAME = 'FLUX    '           / a comment",
                    "RA      =   1.2500000000000D+00", "NOTE    = 'it''s'")
           + _pad(img.tobytes())
           + _cards("XTENSION= 'BINTABLE'", "BITPIX  =                    8", "NAXIS   =                    2",
                    "NAXIS1  =                   25", "NAXIS2  =                    2", "PCOUNT  =                    0",
                    "GCOUNT  =                    1", "TFIELDS =                    4", "TTYPE1  = 'A       '",
                    "TFORM1  = 'J       '", "TTYPE2  = 'B       '", "TFORM2  = '2D      '", "TTYPE3  = 'C       '",
                    "TFORM3  = '4A      '", "TTYPE4  = 'D       '", "TFORM4  = 'L       '", "EXTNAME = 'TAB'")
           + _pad(rows.tobytes()))
    for b in (buf, gzip.compress(buf)):
        h = fits.read_bytes(b)
        assert [x.name for x in h] == ["PRIMARY", "FLUX", "TAB"]
        assert np.array_equal(h[1].data, img) and h[1].data.dtype == np.float64
        assert h[1].header["RA"] == 1.25 and h[1].header["NOTE"] == "it's"
        t = h[2].data
        assert list(t["A"]) == [7, -3] and np.array_equal(t["B"], rows["B"])
        assert list(t["C"]) == [b"ab", b"wxyz"] and list(t["D"]) == [True, False]


def test_writer_round_trip(tmp_path):
    img = np.random.default_rng(0).normal(size=(4, 5))
    cols = {"LOGLAM": np.linspace(3.5, 3.7, 6), "N": np.arange(6, dtype=np.int64),
            "I": np.arange(6, dtype=np.int32) - 3, "F": np.ones(6, np.float32), "V": np.ones((6, 3)),
            "S": np.array(["a", "bb", "", "dddd", "e", "f"]), "L": np.arange(6) % 2 == 0}
    p = tmp_path / "x.fits.gz"
    fits.write(p, [fits.image_bytes(img, "DELTA"), fits.table_bytes(cols, "META", {"RA": 0.5, "THING_ID": 42})])
    assert open(p, "rb").read(2) == b"\x1f\x8b"
    h = fits.read(p)
    assert np.array_equal(fits.by_name(h, "delta").data, img)
    t = fits.by_name(h, "META")
    assert t.header["RA"] == 0.5 and t.header["THING_ID"] == 42
    for k, v in cols.items():
        got = t.data[k]
        if k == "S":
            got = got.astype(str)
        assert np.array_equal(got, v), k
    with pytest.raises(ValueError):
        fits.read_bytes(b"X" * 2880)


# ------------------------------------------------------------ delta files
def _survey(n=40, seed=3):
    return synth.generate(synth.scaled(synth.CONFIGS["patch2k"], n, seed=seed, length_range=(30, 90)))


def _raw(s, f):
    """loglam, delta, weight of forest f as the synthetic survey made them."""
    a, b = s.pix_off[f], s.pix_off[f + 1]
    return s.loglam[a:b], s.delta[a:b], s.we[a:b]


def write_desi(s, path, forest_ids, st):
    """DESI layout: common loglam grid, NaN delta outside each forest."""
    grid = np.unique(np.concatenate([_raw(s, f)[0] for f in forest_ids]))
    lam = 10 ** grid
    D = np.full((len(forest_ids), len(grid)), np.nan)
    W = np.zeros_like(D)
    for r, f in enumerate(forest_ids):
        ll, d, w = _raw(s, f)
        k = np.searchsorted(grid, ll)
        D[r, k] = d
        W[r, k] = w
    meta = {"LOS_ID": s.name[forest_ids].astype(np.int64) + 10 ** 6, "TARGETID": s.name[forest_ids].astype(np.int64),
            "RA": s.ra[forest_ids], "DEC": s.dec[forest_ids]}
    fits.write(path, [fits.image_bytes(lam, "LAMBDA"), fits.table_bytes(meta, "METADATA"),
                      fits.image_bytes(D, st.delta_key), fits.image_bytes(W, "WEIGHT")])
    return lam


def write_eboss(s, path, forest_ids):
    parts = []
    for f in forest_ids:
        ll, d, w = _raw(s, f)
        parts.append(fits.table_bytes({"LOGLAM": ll, "DELTA": d, "WEIGHT": w, "CONT": np.ones_like(d)}, None,
                                      {"RA": float(s.ra[f]), "DEC": float(s.dec[f]), "THING_ID": int(s.name[f]),
                                       "PLATE": 4000 + int(f), "FIBERID": 7 + int(f)}))
    fits.write(path, parts)


def _expect(s, f, st, dist):
    ll, d, w = _raw(s, f)
    z = 10 ** ll / st.lambda_lya - 1
    we = w * ((1 + z) / (1 + st.z_ref)) ** (st.gamma / 2)
    m = np.sum(d * we) / np.sum(we)
    lm = np.sum(ll * we) / np.sum(we)
    dl = ll - lm
    slope = np.sum(dl * d * we) / np.sum(dl * dl * we)
    return we, we * (d - m - dl * slope), dist(z)


def test_records_desi_and_eboss(tmp_path):
    st = ingest.Settings()
    dist = st.distance()
    s = _survey()
    ids = np.arange(s.nf)
    write_desi(s, tmp_path / "d.fits.gz", ids, st)
    write_eboss(s, tmp_path / "e.fits.gz", ids)
    rd = ingest.records_desi(tmp_path / "d.fits.gz", st)
    re_ = ingest.records_eboss(tmp_path / "e.fits.gz", st)
    assert len(rd) == len(re_) == s.nf
    from forest_class import quasar
    for recs in (rd, re_):
        qs = ingest.quasars(recs, quasar, dist)
        for f, q in enumerate(qs):
            we, dw, dc = _expect(s, f, st, dist)
            np.testing.assert_allclose(q.we, we, rtol=1e-13)
            np.testing.assert_allclose(q.dw, dw, rtol=1e-10, atol=1e-13 * np.abs(dw).max())
            np.testing.assert_allclose(q.dc, dc, rtol=1e-13)
            assert q.ra == s.ra[f] and q.dec == s.dec[f] and q.pix == s.hpix[f]
    assert rd[0].lenght == len(ingest.fits.by_name(fits.read(tmp_path / "d.fits.gz"), "LAMBDA").data)
    assert rd[0].name == s.name[0] and rd[0].fib == s.name[0] + 10 ** 6
    assert re_[3].name == s.name[3] and re_[3].pl == 4003 and re_[3].fib == 10 and re_[3].lenght == s.pix_off[4] - s.pix_off[3]


def test_forest_set_direct_matches_pickle_path(tmp_path):
    """FITS -> SoA directly equals FITS -> quasar objects -> ForestSet.from_data."""
    st = ingest.Settings()
    dist = st.distance()
    s = _survey(60, seed=5)
    write_desi(s, tmp_path / "d.fits.gz", np.arange(s.nf)[::-1], st)   # file order != pixel order
    recs = ingest.records_desi(tmp_path / "d.fits.gz", st)
    from forest_class import quasar
    data = {}
    for q in ingest.quasars(recs, quasar, dist):
        data.setdefault(q.pix, []).append(q)
    for order, keys in (("sorted", sorted(data)), ("first", list(data))):
        a = ingest.forest_set(recs, st, dist, pixel_order=order)
        b = forests.ForestSet.from_data(data, keys)
        assert a.slots == b.slots
        for k in ("pix_off", "ux", "uy", "uz", "ra", "name", "slot", "dc", "we"):
            assert np.array_equal(getattr(a, k), getattr(b, k)), k
        for k in ("dw", "dl", "omega", "odl2"):
            np.testing.assert_allclose(getattr(a, k), getattr(b, k), rtol=1e-10, atol=1e-12)
    # the histograms of the two packings agree (oracle, CPU)
    fs = ingest.forest_set(recs, st, dist)
    am = forests.angmax_from(fs, 200.0)
    off, idx = oracle.pairs(fs, am)
    c, w, d = oracle.xi_gpu_form(fs, off, idx)
    fs2 = forests.ForestSet.from_data(data, sorted(data))
    c2, w2, d2 = oracle.xi_gpu_form(fs2, off, idx)
    assert np.array_equal(c, c2)
    np.testing.assert_allclose(w, w2, rtol=1e-14)


def _workspace(tmp_path):
    for f in ("delta_reader.py", "delta_reader_eboss.py", "forest_class.py", "parameters.py"):
        shutil.copy(os.path.join(ROOT, f), tmp_path / f)
    (tmp_path / "deltas").mkdir()


def _run(tmp_path, *args):
    env = dict(os.environ, PYTHONPATH=ROOT)
    r = subprocess.run([sys.executable, *args], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


def test_delta_reader_dropin(tmp_path):
    _workspace(tmp_path)
    st = ingest.Settings()
    s = _survey(50, seed=9)
    ids = np.arange(s.nf)
    write_desi(s, tmp_path / "deltas" / "delta-1.fits.gz", ids[:30], st)
    write_desi(s, tmp_path / "deltas" / "delta-2.fits.gz", ids[30:], st)
    _run(tmp_path, "delta_reader.py", "--delta-dir", "deltas", "--data-dir", "out", "--split-number", "2")
    parts = [np.load(tmp_path / "out" / ("data%d.npy" % i), allow_pickle=True).item() for i in (1, 2)]
    keys = [list(p.keys()) for p in parts]
    allk = keys[0] + keys[1]
    assert allk == sorted(allk) and set(allk) == set(s.hpix.tolist())
    data = {k: v for p in parts for k, v in p.items()}
    qs = [q for k in allk for q in data[k]]
    assert len(qs) == s.nf
    # neighbour names: brute force with each forest's own cut angle, ascending pixel then list order
    for q in qs:
        amax = 2 * np.arcsin(0.5 * 200.0 / q.dc[0])
        exp = [(p.name, k) for k in allk for p in data[k]
               if q.x * p.x + q.y * p.y + q.z * p.z > np.cos(amax) and q.name != p.name and q.ra < p.ra]
        assert list(zip(q.neigh_names, q.neigh_pixels)) == exp
    sizes = np.loadtxt(tmp_path / "sizes")
    assert sorted(sizes.astype(int)) == sorted(np.diff(s.pix_off).tolist())
    assert np.loadtxt(tmp_path / "neighbors").astype(int).tolist() == [len(q.neigh_names) for q in qs]


def test_delta_reader_eboss_dropin(tmp_path):
    _workspace(tmp_path)
    s = _survey(30, seed=10)
    order = np.random.default_rng(1).permutation(s.nf)
    write_eboss(s, tmp_path / "deltas" / "a.fits.gz", order)
    _run(tmp_path, "delta_reader_eboss.py", "--delta-dir", "deltas", "--data-dir", "out")
    data = np.load(tmp_path / "out" / "data1.npy", allow_pickle=True).item()
    first = []
    for f in order:
        if s.hpix[f] not in first:
            first.append(int(s.hpix[f]))
    assert list(data.keys()) == first
    got = [q.name for k in data for q in data[k]]
    assert sorted(got) == sorted(s.name.tolist())
    for k in data:  # file order within a pixel
        names = [q.name for q in data[k]]
        assert names == [s.name[f] for f in order if s.hpix[f] == k]


@pytest.mark.gpu
def test_fits_to_gpu_histograms(tmp_path):
    """Delta files straight to the device kernels: identical to the oracle."""
    st = ingest.Settings()
    s = _survey(80, seed=12)
    write_desi(s, tmp_path / "d.fits.gz", np.arange(s.nf), st)
    fs = ingest.forest_set(ingest.records_desi(tmp_path / "d.fits.gz", st), st)
    am = forests.angmax_from(fs, 200.0)
    print("fits to gpu", fs.nf, flush=True)
    eng = xi.XiEngine(fs, am)
    eng.run(xi.default_config("fp64"))
    c, w, d = eng.results()
    off, idx = oracle.pairs(fs, am)
    c0, w0, d0 = oracle.xi_gpu_form(fs, off, idx)
    assert np.array_equal(c, c0)
    np.testing.assert_allclose(w, w0, rtol=1e-12)
