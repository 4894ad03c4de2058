"""GPU parity of the pair-binning kernels against the CPU oracle.

Bar: per-bin pair counts bit-exact in both precision modes; FP64 sums within
1e-12 of the per-bin absolute sums; mixed sums within 1e-6 of them (FP32
products and FP32 LDS partials, flushed to FP64 every work item).
"""
import itertools

import numpy as np
import pytest

import oracle
from lya2pcf_amd import forests, synth, xi

pytestmark = pytest.mark.gpu


def _abs_sums(fs, off, idx):
    """Per-bin sums of |w1 w2| and |dw1 dw2| (error scale of the sums)."""
    _, aw, ad = oracle.xi_gpu_form(fs, off, idx, we=np.abs(fs.we), dw=np.abs(fs.dw))
    return aw, ad


def _check(fs, off, idx, got, tol, label):
    c0, w0, d0 = oracle.xi_gpu_form(fs, off, idx)
    aw, ad = _abs_sums(fs, off, idx)
    c, w, d = got
    print(label, "pairs", int(c0.sum()), flush=True)
    assert np.array_equal(c, c0), "%s: counts differ in %d bins" % (label, int((c != c0).sum()))
    assert np.all(np.abs(w - w0) <= tol * aw + 1e-300), label
    assert np.all(np.abs(d - d0) <= tol * ad + 1e-300), label


@pytest.mark.parametrize("mode,merge", [("fp64", 0), ("fp64", 1), ("mixed", 0), ("mixed", 1), ("mixed32", 0)])
def test_xi_matches_oracle(small_set, small_pairs, mode, merge):
    """fp64 and mixed (exact FP64 products, FP64 LDS sums): 1e-12 of the absolute
    sums; mixed32 (FP32 products and FP32 LDS partials): 1e-5."""
    am, (off, idx) = small_pairs
    fs = small_set
    if mode != "fp64":
        # identical input values for both: w, dw rounded to float32
        fs = forests.ForestSet(fs.pix_off, fs.slots, **{k: getattr(fs, k) for k in fs.FOREST + fs.PIXEL})
        fs.we = fs.we32.astype(np.float64)
        fs.dw = fs.dw32.astype(np.float64)
    eng = xi.XiEngine(fs, am, pairs=(off, idx))
    eng.run(xi.default_config(mode, merge=merge))
    _check(fs, off, idx, eng.results(), 1e-5 if mode == "mixed32" else 1e-12, "%s merge %d" % (mode, merge))


def test_kernel_shapes(small_set, small_pairs):
    """Every tile size / workgroup / LDS layout gives the same counts, and the same sums
    with counts off (those shapes are separate kernels, several built for 8 waves per
    SIMD)."""
    am, (off, idx) = small_pairs
    c0, _, _ = oracle.xi_gpu_form(small_set, off, idx)
    eng = xi.XiEngine(small_set, am, pairs=(off, idx))
    for mode, ppl, (waves, rep), pad, lmap, merge in itertools.product(["fp64", "mixed", "mixed32"], [1, 2, 4],
                                                                        [(4, 1), (8, 1), (8, 2), (16, 1)], [0, 1],
                                                                        [0, 1, 2], [0, 1]):
        if merge and mode == "mixed32":
            continue
        shape = dict(px_per_lane=ppl, waves=waves, replicas=rep, row_pad=pad, lane_map=lmap, merge=merge)
        eng.run(xi.default_config(mode, **shape))
        c, w, d = eng.results()
        assert np.array_equal(c, c0), (mode, shape)
        eng.run(xi.default_config(mode, counts=False, **shape))
        c2, w2, d2 = eng.results()
        tol = 1e-5 if mode == "mixed32" else 1e-12
        assert c2.sum() == 0, (mode, shape)
        assert np.all(np.abs(w2 - w) <= tol * np.abs(w) + 1e-300), (mode, shape)
        assert np.all(np.abs(d2 - d) <= tol * np.abs(w) + 1e-300), (mode, shape)


def test_without_counts(small_set, small_pairs):
    """Counts off: identical sums, out_cnt left zero."""
    am, (off, idx) = small_pairs
    eng = xi.XiEngine(small_set, am, pairs=(off, idx))
    for mode, merge in (("fp64", 0), ("fp64", 1), ("mixed", 0), ("mixed", 1), ("mixed32", 0)):
        eng.run(xi.default_config(mode, merge=merge))
        _, w0, d0 = eng.results()
        eng.run(xi.default_config(mode, counts=False, merge=merge))
        c, w, d = eng.results()
        assert c.sum() == 0
        tol = 1e-5 if mode == "mixed32" else 1e-12
        np.testing.assert_allclose(w, w0, rtol=tol)
        assert np.all(np.abs(d - d0) <= tol * np.abs(w0) + 1e-300)


@pytest.mark.parametrize("seed,theta", [(1002, (5e-4, 0.04)), (1003, (0.03, 0.055))])
def test_adversarial_edges_bit_exact(seed, theta):
    """10,000 seeded forest pairs whose pixel pairs sit within 1e-6 Mpc/h of bin edges
    (80,000 targeted pixel pairs): mixed bins equal FP64 bins, in every kernel.  The
    first set puts r_perp edges at small angles (dc up to ~8e5 Mpc/h); the second keeps
    every dc below 10^4 Mpc/h, as real forests are."""
    s = synth.adversarial(10_000, seed=seed, theta=theta)
    print("adversarial seed", seed, "theta", theta, flush=True)
    fs = forests.ForestSet.from_survey(s, float32_inputs=True)
    am = 0.06
    off, idx = xi.build_pairs(fs, am)
    c0, _, _ = oracle.xi_gpu_form(fs, off, idx)
    eng = xi.XiEngine(fs, am, pairs=(off, idx))
    for mode in ("mixed", "mixed32", "fp64"):
        for ppl, merge in ((1, 0), (2, 0), (1, 1), (2, 1)):
            if merge and mode == "mixed32":
                continue
            eng.run(xi.default_config(mode, px_per_lane=ppl, merge=merge))
            c, _, _ = eng.results()
            assert np.array_equal(c, c0), (mode, ppl, int((c != c0).sum()))


def test_graph_replays(small_set, small_pairs):
    import torch
    am, (off, idx) = small_pairs
    c0, _, _ = oracle.xi_gpu_form(small_set, off, idx)
    eng = xi.XiEngine(small_set, am, pairs=(off, idx))
    g = eng.graph(xi.default_config("mixed"))
    ref = None
    for rep in range(3):
        g.launch()
        torch.cuda.synchronize()
        c, w, d = eng.results()
        assert np.array_equal(c, c0), rep
        if ref is None:
            ref = (w, d)
        # eager work between replays
        eng.run(xi.default_config("fp64"))
        torch.cuda.synchronize()
    g.close()


def test_halfangles_bit_exact():
    """Device sqrt(0.5(1 +- cos12)) equals the correctly rounded host value."""
    import torch
    from lya2pcf_amd import _lib
    rng = np.random.default_rng(3)
    x = np.concatenate([1.0 - rng.uniform(0, 2e-3, 200000), rng.uniform(-1, 1, 50000), [1.0, -1.0, 0.0]])
    xt = torch.from_numpy(x).cuda()
    c = torch.empty_like(xt)
    s = torch.empty_like(xt)
    _lib.check(_lib.lib().lya_debug_halfangles(_lib.ptr(xt), _lib.ptr(c), _lib.ptr(s), len(x), _lib.stream_ptr()))
    torch.cuda.synchronize()
    assert np.array_equal(c.cpu().numpy(), np.sqrt(0.5 * (1.0 + x)))
    assert np.array_equal(s.cpu().numpy(), np.sqrt(0.5 * (1.0 - x)))


def test_slot_subset(small_set, small_pairs):
    """Computing only some slots gives those slots' histograms unchanged."""
    am, (off, idx) = small_pairs
    c0, w0, _ = oracle.xi_gpu_form(small_set, off, idx)
    keep = [0, len(small_set.slots) - 1]
    eng = xi.XiEngine(small_set, am, pairs=(off, idx), slots=keep)
    eng.run(xi.default_config("fp64"))
    c, w, _ = eng.results()
    for k in range(len(small_set.slots)):
        if k in keep:
            assert np.array_equal(c[k], c0[k])
        else:
            assert c[k].sum() == 0


def test_empty_problem(small_set):
    """No pairs (angmax tiny): zero histograms, no launch error."""
    eng = xi.XiEngine(small_set, 1e-9)
    assert len(eng.nb_idx) == 0
    eng.run(xi.default_config("mixed"))
    c, w, d = eng.results()
    assert c.sum() == 0 and not w.any() and not d.any()


@pytest.mark.parametrize("np_rp,np_rt,rpmax,rtmax", [(1, 1, 200.0, 200.0), (3, 5, 60.0, 100.0),
                                                     (7, 13, 61.0, 97.0)])
def test_small_binnings_merged(small_set, small_pairs, np_rp, np_rt, rpmax, rtmax):
    """Merged runs on binnings of 1 to 7 rows: out-of-range runs of many lanes share the
    few trash entries of the padding column (all 64 lanes one entry at one row), and
    non-power-of-two binners take the unfolded FP64 path; counts bit-exact, sums 1e-12."""
    am, (off, idx) = small_pairs
    fs = forests.ForestSet(small_set.pix_off, small_set.slots,
                           **{k: getattr(small_set, k) for k in small_set.FOREST + small_set.PIXEL})
    fs.we = fs.we32.astype(np.float64)
    fs.dw = fs.dw32.astype(np.float64)
    eng = xi.XiEngine(fs, am, np_rp=np_rp, np_rt=np_rt, rpmax=rpmax, rtmax=rtmax, pairs=(off, idx))
    c0, w0, d0 = oracle.xi_gpu_form(fs, off, idx, np_rp, np_rt, rpmax, rtmax)
    _, aw, ad = oracle.xi_gpu_form(fs, off, idx, np_rp, np_rt, rpmax, rtmax, we=np.abs(fs.we), dw=np.abs(fs.dw))
    for mode in ("fp64", "mixed"):
        for ppl in (1, 2):
            eng.run(xi.default_config(mode, px_per_lane=ppl, merge=1))
            c, w, d = eng.results()
            assert np.array_equal(c, c0), (mode, ppl, np_rp, np_rt)
            assert np.all(np.abs(w - w0) <= 1e-12 * aw + 1e-300), (mode, ppl)
            assert np.all(np.abs(d - d0) <= 1e-12 * ad + 1e-300), (mode, ppl)


@pytest.mark.parametrize("xcd_groups", [1, 3, 16])
def test_tile_group_counts(small_set, small_pairs, xcd_groups):
    """One queue (what callers get with n_tile_groups = 0), and tile groups of any count,
    including groups left empty: every tile is taken exactly once (counts bit-exact)."""
    am, (off, idx) = small_pairs
    c0, _, _ = oracle.xi_gpu_form(small_set, off, idx)
    eng = xi.XiEngine(small_set, am, pairs=(off, idx), xcd_groups=xcd_groups, item_cost=20_000)
    for mode, shape in (("fp64", dict(px_per_lane=2, merge=1)), ("mixed", dict(merge=1)),
                        ("fp64", dict(walk=1, waves=16))):
        eng.run(xi.default_config(mode, **shape))
        c, _, _ = eng.results()
        assert np.array_equal(c, c0), (xcd_groups, mode, shape)


def _shifted_set(n, seed):
    """Forests of 50-600 pixels (the 300k-forest recipe at small size) with every forest's
    distances shifted by a seeded U(-600, 600) Mpc/h: many forest pairs have no pixel pair
    within rmax, and many pairs only over part of forest 1, so the trimmed ranges start
    and end inside the forests with every remainder modulo 64."""
    s = synth.generate(synth.scaled(synth.CONFIGS["sky300k"], n, seed=seed))
    fs = forests.ForestSet.from_survey(s, float32_inputs=True)
    rng = np.random.default_rng(seed)
    shift = rng.uniform(-600.0, 600.0, fs.nf)
    dc = fs.dc.copy()
    for f in range(fs.nf):
        dc[fs.pix_off[f]:fs.pix_off[f + 1]] += shift[f]
    return forests.ForestSet(fs.pix_off, fs.slots, **{**{k: getattr(fs, k) for k in fs.FOREST + fs.PIXEL},
                                                      "dc": dc})


@pytest.mark.parametrize("seed", [21, 22])
def test_trimmed_units_and_tails(seed):
    """Each forest pair is trimmed to the forest-1 pixels that can pair before its slabs
    are cut, and the tail is walked against the forest-2 pixels that can pair with it:
    every tail threshold (off, 16, 32, 48, 63 pixels), both slab widths, both precisions,
    merged runs on and off give the oracle's counts bit for bit and its sums to 1e-12."""
    fs = _shifted_set(150, seed)
    am = forests.angmax_from(fs, 200.0)
    off, idx = xi.build_pairs(fs, am)
    c0, w0, d0 = oracle.xi_gpu_form(fs, off, idx)
    aw, ad = _abs_sums(fs, off, idx)
    print("seed", seed, "forest pairs", len(idx), "pixel pairs binned", int(c0.sum()), flush=True)
    eng = xi.XiEngine(fs, am, pairs=(off, idx))
    stats = eng.work_stats(xi.default_config("fp64", tail_max=-1))
    assert stats["tasks"] > 0
    for mode, ppl, merge, tail in itertools.product(["fp64", "mixed"], [1, 2], [0, 1], [-1, 16, 32, 48, 63]):
        eng.run(xi.default_config(mode, px_per_lane=ppl, merge=merge, tail_max=tail))
        c, w, d = eng.results()
        label = (mode, ppl, merge, tail)
        assert np.array_equal(c, c0), "%s: counts differ in %d bins" % (label, int((c != c0).sum()))
        assert np.all(np.abs(w - w0) <= 1e-12 * aw + 1e-300), label
        assert np.all(np.abs(d - d0) <= 1e-12 * ad + 1e-300), label
