"""Property tests (hypothesis): random geometries against the oracle's brute force.

CPU: the neighbour rule (three builders), the FP64 pair formula (C oracle vs numpy,
role swap, weight scale) and the distortion eta algebra (vs its direct definition).
GPU: random sky sets, kernels and tuned shapes, bit-exact counts against the oracle.
"""
import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings, strategies as st

import oracle
from lya2pcf_amd import autotune, forests, synth, xi

CPU = settings(max_examples=40, deadline=None, derandomize=True,
               suppress_health_check=[HealthCheck.too_slow, HealthCheck.data_too_large])


def _forest_set(rng, n, lengths, spread_rad, n_names, ra_step):
    """n forests around (1, 0, 0) within spread_rad, dc increasing from 3540 Mpc/h;
    names from a small set and ra on a coarse grid, so ties of both occur."""
    ra = np.round(rng.uniform(-spread_rad, spread_rad, n) / ra_step) * ra_step
    dec = rng.uniform(-spread_rad, spread_rad, n)
    ux, uy, uz = np.cos(dec) * np.cos(ra), np.cos(dec) * np.sin(ra), np.sin(dec)
    off = np.zeros(n + 1, np.int64)
    np.cumsum(lengths, out=off[1:])
    dc = np.concatenate([3540.0 + np.sort(rng.uniform(0, 600, k)) for k in lengths]) if n else np.zeros(0)
    npix = int(off[-1])
    return forests.ForestSet(off, [0], ux=ux, uy=uy, uz=uz, ra=ra, dec=dec,
                             name=rng.integers(0, n_names, n).astype(np.int64), slot=np.zeros(n, np.int32),
                             omega=np.ones(n), odl2=np.ones(n), dc=dc, we=rng.uniform(5, 20, npix),
                             dw=rng.normal(0, 2, npix), dl=rng.normal(0, 1e-3, npix))


@CPU
@given(seed=st.integers(0, 2**31 - 1), n=st.integers(0, 30), spread=st.floats(0.01, 0.2),
       n_names=st.integers(1, 40), ra_step=st.sampled_from([1e-9, 1e-3, 1e-2]), angmax=st.floats(0.005, 0.3))
def test_neighbour_rule_three_builders(seed, n, spread, n_names, ra_step, angmax):
    """forest_class.py:96 (dot > cos angmax, different names, ra1 < ra2): the C oracle,
    the numpy brute force and the OpenMP host builder give the same CSR."""
    rng = np.random.default_rng(seed)
    fs = _forest_set(rng, n, np.ones(n, np.int64), spread, n_names, ra_step)
    o_off, o_idx = oracle.pairs(fs, angmax)
    n_off, n_idx = oracle.numpy_pairs(fs, angmax)
    h_off, h_idx = xi.build_pairs(fs, angmax)
    assert np.array_equal(o_off, n_off) and np.array_equal(o_idx, n_idx)
    assert np.array_equal(np.asarray(h_off), o_off) and np.array_equal(np.asarray(h_idx), o_idx)


def _pair(rng, n1, n2, theta, r0, span):
    dc1 = r0 + np.sort(rng.uniform(0, span, n1))
    dc2 = r0 + np.sort(rng.uniform(0, span, n2))
    ux = np.array([1.0, np.cos(theta)]); uy = np.array([0.0, np.sin(theta)])
    return forests.ForestSet(np.array([0, n1, n1 + n2]), [0], ux=ux, uy=uy, uz=np.zeros(2),
                             ra=np.array([0.0, theta]), dec=np.zeros(2), name=np.array([1, 2]),
                             slot=np.zeros(2, np.int32), omega=np.ones(2), odl2=np.ones(2),
                             dc=np.concatenate([dc1, dc2]), we=rng.uniform(1, 20, n1 + n2),
                             dw=rng.normal(0, 3, n1 + n2), dl=np.zeros(n1 + n2))


@CPU
@given(seed=st.integers(0, 2**31 - 1), n1=st.integers(1, 60), n2=st.integers(1, 60),
       theta=st.floats(0.0, 0.06), r0=st.floats(3500.0, 6000.0), span=st.floats(1.0, 400.0),
       bins=st.sampled_from([(50, 50, 200.0, 200.0), (40, 30, 200.0, 150.0), (7, 13, 61.0, 97.0)]))
def test_pair_formula_brute_force(seed, n1, n2, theta, r0, span, bins):
    """The C oracle's FP64 bins equal numpy's floor(|dr| c np/rmax), floor((r1+r2) s np/rmax);
    swapping the two forests keeps every bin; a weight scale k scales W by k^2."""
    np_rp, np_rt, rpmax, rtmax = bins
    fs = _pair(np.random.default_rng(seed), n1, n2, theta, r0, span)
    off, idx = np.array([0, 1, 1]), np.array([1], np.int32)
    c, w, d = oracle.xi_gpu_form(fs, off, idx, np_rp, np_rt, rpmax, rtmax)
    cn, wn, dn = oracle.numpy_xi_gpu_form(fs, off, idx, np_rp, np_rt, rpmax, rtmax)
    assert np.array_equal(c, cn)
    _, aw, ad = oracle.xi_gpu_form(fs, off, idx, np_rp, np_rt, rpmax, rtmax, we=np.abs(fs.we), dw=np.abs(fs.dw))
    assert np.all(np.abs(w - wn) <= 1e-13 * aw) and np.all(np.abs(d - dn) <= 1e-13 * ad)
    cs, _, _ = oracle.xi_gpu_form(fs, np.array([0, 0, 1]), np.array([0], np.int32), np_rp, np_rt, rpmax, rtmax)
    assert np.array_equal(c, cs)
    c2, w2, _ = oracle.xi_gpu_form(fs, off, idx, np_rp, np_rt, rpmax, rtmax, we=fs.we * 4.0)
    assert np.array_equal(c, c2) and np.allclose(w2, 16.0 * w, rtol=1e-14, atol=0)


@settings(max_examples=15, deadline=None, derandomize=True, suppress_health_check=[HealthCheck.too_slow])
@given(seed=st.integers(0, 2**31 - 1), n=st.integers(2, 6), max_len=st.integers(1, 9),
       bins=st.sampled_from([(5, 5, 200.0, 200.0), (10, 4, 200.0, 200.0)]))
def test_distortion_eta_algebra_vs_definition(seed, n, max_len, bins):
    """D_AB = sum over pairs in A of w_i w_j sum over pairs in B of eta_ik eta_jl, from the
    expanded eta terms and from the definition: equal to 1e-12 of max|D|."""
    rng = np.random.default_rng(seed)
    lengths = rng.integers(1, max_len + 1, n)
    fs = _forest_set(rng, n, lengths, 0.01, 1000, 1e-9)
    fs.name = np.arange(n, dtype=np.int64)
    # per-forest normalisations of the projection, as forest_class.py:54-66 defines them
    for f in range(n):
        a = fs.forest_slice(f)
        fs.omega[f] = fs.we[a].sum()
        fs.odl2[f] = (fs.we[a] * fs.dl[a] ** 2).sum()
    pr = np.array([(i, j) for i in range(n) for j in range(n) if i != j], np.int32)
    np_rp, np_rt, rpmax, rtmax = bins
    D1, w1 = oracle.distortion(fs, pr, np_rp, np_rt, rpmax, rtmax, form="eta")
    D2, w2 = oracle.distortion(fs, pr, np_rp, np_rt, rpmax, rtmax, form="direct")
    scale = max(np.abs(D2).max(), 1e-300)
    assert np.all(np.abs(D1 - D2) <= 1e-12 * scale)
    np.testing.assert_allclose(w1, w2, rtol=1e-13, atol=0)


# ---------------------------------------------------------------------- GPU
_SHAPES = {m: list(autotune.candidates(m)) for m in ("mixed", "fp64")}


@pytest.mark.gpu
@settings(max_examples=12, deadline=None, derandomize=True, suppress_health_check=[HealthCheck.too_slow])
@given(seed=st.integers(0, 2**31 - 1), recipe=st.sampled_from(["patch2k", "sky100k", "sky300k"]),
       n=st.integers(20, 120), mode=st.sampled_from(["mixed", "fp64"]), pick=st.integers(0, 10**6),
       counts=st.booleans())
def test_kernels_on_random_skies(seed, recipe, n, mode, pick, counts):
    """Random synthetic skies of the three recipes, every kernel family and tuned shape:
    per-bin counts bit-exact against the oracle, sums within 1e-12 of the absolute sums."""
    fs = forests.ForestSet.from_survey(synth.generate(synth.scaled(synth.CONFIGS[recipe], n, seed=seed)),
                                       float32_inputs=(mode == "mixed"))
    am = forests.angmax_from(fs, 200.0)
    off, idx = xi.build_pairs(fs, am)
    shape = _SHAPES[mode][pick % len(_SHAPES[mode])]
    print("random sky", recipe, n, seed, mode, shape, flush=True)
    eng = xi.XiEngine(fs, am, pairs=(off, idx))
    eng.run(xi.default_config(mode, counts=True, **shape))
    c, w, d = eng.results()
    c0, w0, d0 = oracle.xi_gpu_form(fs, off, idx)
    _, aw, ad = oracle.xi_gpu_form(fs, off, idx, we=np.abs(fs.we), dw=np.abs(fs.dw))
    assert np.array_equal(c, c0)
    assert np.all(np.abs(w - w0) <= 1e-12 * aw) and np.all(np.abs(d - d0) <= 1e-12 * ad)
    if not counts:
        eng.run(xi.default_config(mode, counts=False, **shape))
        c2, w2, d2 = eng.results()
        assert c2.sum() == 0
        assert np.all(np.abs(w2 - w0) <= 1e-12 * aw) and np.all(np.abs(d2 - d0) <= 1e-12 * ad)
