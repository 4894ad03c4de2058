"""GPU parity of the diagonal-walk kernel (walk 1, FP64) against the oracle.

Same bar as the other FP64 kernels: per-bin pair counts bit-exact, sums within
1e-12 of the per-bin absolute sums.  Cases cover both workgroup sizes, counts
on and off, power-of-two and other binners, unsorted forests (full diagonal
range, no r_perp cut), pixel pairs on bin edges, items split over several
neighbour batches, long forests (batches cut by the sub-group capacity) and
the drop-in procedures.
"""
import numpy as np
import pytest

import oracle
from lya2pcf_amd import forests, synth, xi

pytestmark = pytest.mark.gpu


def _abs_sums(fs, off, idx, **kw):
    _, aw, ad = oracle.xi_gpu_form(fs, off, idx, we=np.abs(fs.we), dw=np.abs(fs.dw), **kw)
    return aw, ad


def _check(fs, off, idx, got, label, tol=1e-12, **kw):
    c0, w0, d0 = oracle.xi_gpu_form(fs, off, idx, **kw)
    aw, ad = _abs_sums(fs, off, idx, **kw)
    c, w, d = got
    print(label, "pairs", int(c0.sum()), flush=True)
    assert np.array_equal(c, c0), "%s: counts differ in %d bins" % (label, int((c != c0).sum()))
    assert np.all(np.abs(w - w0) <= tol * aw + 1e-300), label
    assert np.all(np.abs(d - d0) <= tol * ad + 1e-300), label


@pytest.mark.parametrize("waves", [8, 16])
def test_walk_matches_oracle(small_set, small_pairs, waves):
    am, (off, idx) = small_pairs
    eng = xi.XiEngine(small_set, am, pairs=(off, idx))
    eng.run(xi.default_config("fp64", walk=1, waves=waves))
    _check(small_set, off, idx, eng.results(), "walk w%d" % waves)


def test_walk_counts_off_and_other_binner(small_set, small_pairs):
    am, (off, idx) = small_pairs
    eng = xi.XiEngine(small_set, am, pairs=(off, idx))
    eng.run(xi.default_config("fp64", walk=1))
    _, w0, d0 = eng.results()
    eng.run(xi.default_config("fp64", walk=1, counts=False))
    c, w, d = eng.results()
    assert c.sum() == 0
    np.testing.assert_allclose(w, w0, rtol=1e-12)
    # 40 x 30 bins over (200, 150): binners 0.2 and 0.2 are not powers of two
    eng2 = xi.XiEngine(small_set, am, np_rp=40, np_rt=30, rpmax=200.0, rtmax=150.0, pairs=(off, idx))
    eng2.run(xi.default_config("fp64", walk=1))
    _check(small_set, off, idx, eng2.results(), "walk 40x30", np_rp=40, np_rt=30, rpmax=200.0, rtmax=150.0)


def test_walk_unsorted_forests(small_set, small_pairs):
    """Forests whose dc is not monotone take the full diagonal range."""
    am, (off, idx) = small_pairs
    fs = small_set
    args = {k: getattr(fs, k).copy() for k in fs.FOREST + fs.PIXEL}
    for f in range(0, fs.nf, 3):  # reverse every third forest
        sl = fs.forest_slice(f)
        for k in fs.PIXEL:
            args[k][sl] = args[k][sl][::-1]
    fr = forests.ForestSet(fs.pix_off, fs.slots, **args)
    assert fr.sorted.sum() < fs.nf
    eng = xi.XiEngine(fr, am, pairs=(off, idx))
    eng.run(xi.default_config("fp64", walk=1))
    _check(fr, off, idx, eng.results(), "walk unsorted")


def test_walk_edge_pairs_bit_exact():
    s = synth.adversarial(2000, seed=1002)
    fs = forests.ForestSet.from_survey(s, float32_inputs=True)
    am = 0.06
    off, idx = xi.build_pairs(fs, am)
    c0, _, _ = oracle.xi_gpu_form(fs, off, idx)
    eng = xi.XiEngine(fs, am, pairs=(off, idx))
    eng.run(xi.default_config("fp64", walk=1))
    c, _, _ = eng.results()
    assert np.array_equal(c, c0)


def test_walk_batches_and_long_forests():
    """Items of > 64 neighbours (several batches) and 600-pixel forests (a batch
    cut by the sub-group capacity) against the tile kernel."""
    s = synth.generate(synth.scaled(synth.CONFIGS["sky300k"], 4000, seed=31))
    fs = forests.ForestSet.from_survey(s, float32_inputs=True)
    am = forests.angmax_from(fs, 200.0)
    print("long forests", fs.nf, int(np.max(fs.n)), flush=True)
    eng = xi.XiEngine(fs, am, item_cost=1 << 40)   # one item per forest 1
    assert np.max(np.diff(eng.nb_off)) > 64
    eng.run(xi.default_config("fp64"))
    c0, w0, d0 = eng.results()
    eng.run(xi.default_config("fp64", walk=1))
    c, w, d = eng.results()
    assert np.array_equal(c, c0)
    np.testing.assert_allclose(w, w0, rtol=1e-11)


def test_walk_graph_replays(small_set, small_pairs):
    am, (off, idx) = small_pairs
    eng = xi.XiEngine(small_set, am, pairs=(off, idx))
    g = eng.graph(xi.default_config("fp64", walk=1))
    c0, w0, _ = oracle.xi_gpu_form(small_set, off, idx)
    for _ in range(3):
        g.launch()
        c, w, _ = eng.results()
        assert np.array_equal(c, c0)
        np.testing.assert_allclose(w, w0, rtol=1e-12)
    g.close()
