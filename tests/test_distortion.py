"""Distortion matrix: the oracle's eta algebra (reference cuda_kernels.cpp:165-278)
against the definition D_AB = sum_{(i,j) in A} w_i w_j sum_{(k,l) in B} eta_ik eta_jl,
and the GPU kernel (FP64 accumulation) against the oracle."""
import random

import numpy as np
import pytest

import oracle
from lya2pcf_amd import distortion as dist
from lya2pcf_amd import forests, synth, xi


def _short_set(n, lo, hi, seed):
    cfg = synth.scaled(synth.CONFIGS["patch2k"], n, seed=seed, length_range=(lo, hi))
    return forests.ForestSet.from_survey(synth.generate(cfg))


def test_eta_algebra_equals_definition():
    fs = _short_set(40, 6, 14, 21)
    am = forests.angmax_from(fs, 200.0)
    off, idx = xi.build_pairs(fs, am)
    f1 = np.repeat(np.arange(fs.nf), np.diff(off))
    pairs = np.stack([f1, idx], 1)[:60]
    assert len(pairs) > 10
    D1, w1 = oracle.distortion(fs, pairs, form="eta")
    D2, w2 = oracle.distortion(fs, pairs, form="direct")
    assert np.array_equal(w1 > 0, w2 > 0)
    np.testing.assert_allclose(w1, w2, rtol=1e-13)
    scale = np.abs(D2).max()
    assert np.max(np.abs(D1 - D2)) <= 1e-12 * scale


def test_projection_removes_mean_and_slope():
    """The projection (forest_class.py:54-66) maps delta = a + b*dl to zero."""
    fs = _short_set(30, 8, 12, 5)
    # a projected field of a+b*dl is zero: dw computed by the projection of synth
    d = synth.project_forests(fs.pix_off, fs.dl + 0.0, 3.0 + 7.0 * fs.dl, fs.we)[0]
    assert np.max(np.abs(d)) < 1e-9


def test_subsample_matches_reference_semantics():
    """ceil(n*(1-reject)) after random.shuffle with the global seed-1 stream."""
    fs = _short_set(60, 20, 30, 9)
    am = forests.angmax_from(fs, 200.0)
    off, idx = xi.build_pairs(fs, am)
    rng = random.Random(1)
    got = dist.subsample(fs, off, idx, range(fs.nf), 0.95, rng=rng)
    rng = random.Random(1)
    exp = []
    for f in range(fs.nf):
        nb = list(idx[off[f]:off[f + 1]])
        keep = int(np.ceil(len(nb) * (1. - 0.95)))
        rng.shuffle(nb)
        exp += [(f, g) for g in nb[:keep]]
    assert [tuple(p) for p in got] == exp


@pytest.mark.gpu
def test_distortion_gpu_matches_oracle():
    fs = _short_set(120, 30, 60, 13)
    am = forests.angmax_from(fs, 200.0)
    off, idx = xi.build_pairs(fs, am)
    f1 = np.repeat(np.arange(fs.nf), np.diff(off))
    pairs = np.stack([f1, idx], 1)[::7][:400].astype(np.int32)
    D0, w0 = oracle.distortion(fs, pairs, form="eta")
    eng = dist.DistortionEngine(fs, am, pairs=(off, idx))
    eng.accumulate(pairs)
    D, w = eng.result()
    np.testing.assert_allclose(w, w0, rtol=1e-12, atol=0)
    scale = np.abs(D0).max()
    assert np.max(np.abs(D - D0)) <= 1e-11 * scale, np.max(np.abs(D - D0)) / scale


@pytest.mark.gpu
def test_distortion_gpu_long_forests():
    """Forests of 300-450 pixels (more pixel pairs than one LDS bin table would hold)
    and pairs of very unequal length, against the oracle."""
    fs = _short_set(40, 300, 450, 17)
    am = forests.angmax_from(fs, 200.0)
    off, idx = xi.build_pairs(fs, am)
    f1 = np.repeat(np.arange(fs.nf), np.diff(off))
    pairs = np.stack([f1, idx], 1)[::3][:60].astype(np.int32)
    assert int(np.max(fs.n)) > 300 and len(pairs) > 10
    D0, w0 = oracle.distortion(fs, pairs, form="eta")
    eng = dist.DistortionEngine(fs, am, pairs=(off, idx))
    eng.accumulate(pairs)
    D, w = eng.result()
    np.testing.assert_allclose(w, w0, rtol=1e-12, atol=0)
    scale = np.abs(D0).max()
    assert np.max(np.abs(D - D0)) <= 1e-11 * scale, np.max(np.abs(D - D0)) / scale
    assert np.array_equal(D, D.T)


def _active_bins(fs, f1, f2, np_rp, np_rt, rpmax, rtmax):
    r1 = fs.dc[fs.pix_off[f1]:fs.pix_off[f1 + 1]]
    r2 = fs.dc[fs.pix_off[f2]:fs.pix_off[f2 + 1]]
    cos = fs.ux[f1] * fs.ux[f2] + fs.uy[f1] * fs.uy[f2]
    cos = min(1.0, max(-1.0, cos + fs.uz[f1] * fs.uz[f2]))
    c, s = np.sqrt(0.5 * (1 + cos)), np.sqrt(0.5 * (1 - cos))
    bp = np.floor(np.abs(r1[:, None] - r2[None, :]) * c * (np_rp / rpmax))
    bt = np.floor((r1[:, None] + r2[None, :]) * s * (np_rt / rtmax))
    ok = (bp < np_rp) & (bt < np_rt)
    return len(np.unique((bp * np_rt + bt)[ok]))


@pytest.mark.gpu
def test_distortion_gpu_many_active_bins():
    """100 x 40 bins: forest pairs with more than 240 active bins take several passes
    over their Z rows (the 16x16 blocks of one pass do not cover D_pair)."""
    np_rp, np_rt, rmax = 100, 40, 200.0
    fs = _short_set(60, 150, 208, 23)
    am = forests.angmax_from(fs, 200.0)
    off, idx = xi.build_pairs(fs, am)
    f1 = np.repeat(np.arange(fs.nf), np.diff(off))
    allp = np.stack([f1, idx], 1).astype(np.int32)
    nact = np.array([_active_bins(fs, a, b, np_rp, np_rt, rmax, rmax) for a, b in allp])
    pairs = allp[np.argsort(-nact)[:24]]
    assert nact.max() > 240
    D0, w0 = oracle.distortion(fs, pairs, np_rp, np_rt, rmax, rmax, form="eta")
    eng = dist.DistortionEngine(fs, am, np_rp, np_rt, rmax, rmax, pairs=(off, idx))
    eng.accumulate(pairs)
    D, w = eng.result()
    np.testing.assert_allclose(w, w0, rtol=1e-12, atol=0)
    scale = np.abs(D0).max()
    assert np.max(np.abs(D - D0)) <= 1e-11 * scale, np.max(np.abs(D - D0)) / scale


@pytest.mark.gpu
def test_distortion_gpu_short_and_flat_forests():
    """Forests of 2-6 pixels, every fifth with dl = 0 (L = sum w dl^2 = 0, the zero-slope
    convention of both the oracle and the kernel), and pairs without any in-range pixel pair."""
    cfg = synth.scaled(synth.CONFIGS["patch2k"], 200, seed=31, length_range=(1, 6))
    fs = forests.ForestSet.from_survey(synth.generate(cfg))
    for f in range(0, fs.nf, 5):
        fs.dl[fs.pix_off[f]:fs.pix_off[f + 1]] = 0.0
        fs.odl2[f] = 0.0
    am = forests.angmax_from(fs, 200.0)
    off, idx = xi.build_pairs(fs, am)
    f1 = np.repeat(np.arange(fs.nf), np.diff(off))
    pairs = np.stack([f1, idx], 1)[::11].astype(np.int32)
    nact = np.array([_active_bins(fs, a, b, 50, 50, 200.0, 200.0) for a, b in pairs])
    assert (nact == 0).any() and (fs.odl2 == 0).sum() >= 30
    D0, w0 = oracle.distortion(fs, pairs, form="eta")
    eng = dist.DistortionEngine(fs, am, pairs=(off, idx))
    eng.accumulate(pairs)
    D, w = eng.result()
    np.testing.assert_allclose(w, w0, rtol=1e-12, atol=0)
    scale = np.abs(D0).max()
    assert np.max(np.abs(D - D0)) <= 1e-11 * scale, np.max(np.abs(D - D0)) / scale
