"""Binnings larger than the LDS histogram (reference parameters.py:14-21 lets
bin_size_r, rpmax and rtmax change; the reference bins any numpix_rp x numpix_rt
with global atomics, cuda_kernels.cpp:94-99,152-156).

The xi tile kernels then run in bands of bins; the distortion budgets its LDS by
the active bins of each pair; the adjoint reads its table from global memory when
the table exceeds a workgroup's LDS.  Bar as for the default binning: counts
bit-exact against the FP64 oracle, sums within 1e-12 of the per-bin absolute sums,
the distortion within 1e-10 of max|D|.
"""
import random

import numpy as np
import pytest

import oracle
from lya2pcf_amd import forests, synth, xi

pytestmark = pytest.mark.gpu

# (np_rp, np_rt, rpmax, rtmax): 2 Mpc/h bins (5 column bands); a non-square binning
# with non-power-of-two binners; 1 Mpc/h bins (19 column bands); a row-banded binning
# (more rows than one column per row fits: 5 blocks of rows)
BINNINGS = [(100, 100, 200.0, 200.0), (100, 37, 200.0, 111.0), (200, 200, 200.0, 200.0), (1300, 7, 200.0, 70.0)]


@pytest.fixture(scope="module")
def f32_set(small_set):
    """The small set with w, dw rounded to FP32: the same inputs for both modes."""
    fs = forests.ForestSet(small_set.pix_off, small_set.slots,
                           **{k: getattr(small_set, k) for k in small_set.FOREST + small_set.PIXEL})
    fs.we = fs.we32.astype(np.float64)
    fs.dw = fs.dw32.astype(np.float64)
    return fs


@pytest.mark.parametrize("np_rp,np_rt,rpmax,rtmax", BINNINGS)
def test_banded_xi_matches_oracle(f32_set, small_pairs, np_rp, np_rt, rpmax, rtmax):
    am, (off, idx) = small_pairs
    fs = f32_set
    print("binning", np_rp, np_rt, rpmax, rtmax, flush=True)
    c0, w0, d0 = oracle.xi_gpu_form(fs, off, idx, np_rp, np_rt, rpmax, rtmax)
    _, aw, ad = oracle.xi_gpu_form(fs, off, idx, np_rp, np_rt, rpmax, rtmax, we=np.abs(fs.we), dw=np.abs(fs.dw))
    assert c0.sum() > 0
    eng = xi.XiEngine(fs, am, np_rp=np_rp, np_rt=np_rt, rpmax=rpmax, rtmax=rtmax, pairs=(off, idx))
    for mode in ("fp64", "mixed"):
        for ppl, merge in ((1, 0), (1, 1), (2, 1)):
            eng.run(xi.default_config(mode, px_per_lane=ppl, merge=merge))
            c, w, d = eng.results()
            label = (mode, ppl, merge, np_rp, np_rt)
            assert np.array_equal(c, c0), (label, int((c != c0).sum()))
            assert np.all(np.abs(w - w0) <= 1e-12 * aw + 1e-300), label
            assert np.all(np.abs(d - d0) <= 1e-12 * ad + 1e-300), label
            # counts off: the same sums, counts left zero
            eng.run(xi.default_config(mode, px_per_lane=ppl, merge=merge, counts=False))
            c2, w2, d2 = eng.results()
            assert c2.sum() == 0, label
            assert np.all(np.abs(w2 - w0) <= 1e-12 * aw + 1e-300), label
            assert np.all(np.abs(d2 - d0) <= 1e-12 * ad + 1e-300), label


def test_banded_graph_and_edges():
    """Adversarial edge pairs at 2 Mpc/h bins (every band edge an r_perp edge of some
    pair) through a captured graph, replayed with eager work between replays."""
    import torch
    s = synth.adversarial(2_000, seed=1005, theta=(0.03, 0.055), nbins=100)
    fs = forests.ForestSet.from_survey(s, float32_inputs=True)
    off, idx = xi.build_pairs(fs, 0.06)
    c0, _, _ = oracle.xi_gpu_form(fs, off, idx, 100, 100, 200.0, 200.0)
    eng = xi.XiEngine(fs, 0.06, np_rp=100, np_rt=100, pairs=(off, idx))
    g = eng.graph(xi.default_config("mixed", merge=1))
    for rep in range(3):
        g.launch()
        torch.cuda.synchronize()
        c, _, _ = eng.results()
        assert np.array_equal(c, c0), rep
        eng.run(xi.default_config("fp64", px_per_lane=2, merge=1))
        assert np.array_equal(eng.results()[0], c0), rep
    g.close()


def test_adjoint_global_table(f32_set, small_pairs):
    """150 x 100 bins: the (alpha, beta) table (240 kB) exceeds a workgroup's LDS and is
    read from global memory; both sides of every pair (the list and its transpose)
    equal the oracle adjoint to 1e-12, and the forward sums run banded."""
    import torch
    from lya2pcf_amd import adjoint
    am, (off, idx) = small_pairs
    fs = f32_set
    np_rp, np_rt = 150, 100
    nb = np_rp * np_rt
    rng = np.random.default_rng(5)
    alpha, beta = rng.normal(size=nb), rng.normal(size=nb)
    g0d, g0w = oracle.xi_adjoint(fs, off, idx, alpha, beta, np_rp, np_rt, 200.0, 200.0)
    ga, gb = oracle.xi_adjoint(fs, off, idx, np.abs(alpha), np.abs(beta), np_rp, np_rt, 200.0, 200.0,
                               we=np.abs(fs.we), dw=np.abs(fs.dw))
    _, w0, d0 = oracle.xi_gpu_form(fs, off, idx, np_rp, np_rt, 200.0, 200.0)
    for mode in ("fp64", "mixed"):
        eng = adjoint.AdjointEngine(fs, am, mode=mode, np_rp=np_rp, np_rt=np_rt, pairs=(off, idx))
        gd, gw = eng.adjoint(torch.from_numpy(alpha).cuda(), torch.from_numpy(beta).cuda())
        torch.cuda.synchronize()
        assert np.all(np.abs(gd.cpu().numpy() - g0d) <= 1e-12 * ga + 1e-300), mode
        assert np.all(np.abs(gw.cpu().numpy() - g0w) <= 1e-12 * gb + 1e-300), mode
        W, DW = eng.sums()
        np.testing.assert_allclose(W.cpu().numpy(), w0.sum(0).ravel(), rtol=1e-12, atol=1e-12 * np.abs(w0).sum())


def test_distortion_nonsquare_large():
    """120 x 70 bins (8,400): more than two rows of every bin fit the LDS no longer;
    the LDS is budgeted by each pair's active bins.  D within 1e-11 of max|D|."""
    from lya2pcf_amd import distortion as dist
    s = synth.generate(synth.scaled(synth.CONFIGS["patch2k"], 40, seed=13, length_range=(20, 50)))
    fs = forests.ForestSet.from_survey(s)
    am = forests.angmax_from(fs, 200.0)
    off, idx = oracle.pairs(fs, am)
    pairs = dist.subsample(fs, off, idx, range(fs.nf), 0.5, rng=random.Random(3))
    np_rp, np_rt, rpmax, rtmax = 120, 70, 180.0, 140.0
    D0, w0 = oracle.distortion(fs, pairs, np_rp, np_rt, rpmax, rtmax)
    eng = dist.DistortionEngine(fs, am, np_rp, np_rt, rpmax, rtmax, pairs=(off, idx))
    eng.accumulate(pairs)
    D, w = eng.result()
    assert eng.skipped() == (0, 0)
    assert np.abs(D0).max() > 0
    assert np.max(np.abs(D - D0)) <= 1e-11 * np.abs(D0).max()
    assert np.allclose(w, w0, rtol=1e-12, atol=0)
