"""Forest-pair list and work decomposition: host (OpenMP) and GPU builders.

The neighbour rule is forest_class.py:96 (reference quasar.neighborhood,
forest_class.py:78-98); the list is checked against the oracle's brute force.
The item and tile cuts are pinned by the invariants that define them: items
partition each forest's neighbour range, each is greedy-maximal under the cost
bound; tiles partition the items into runs of one slot, greedy-maximal, in
stably decreasing cost order.
"""
import types

import numpy as np
import pytest

import oracle
from lya2pcf_amd import synth, forests, xi


def _item_costs(fs, nb_off, nb_idx, items):
    n = np.diff(fs.pix_off)
    return np.array([n[f] * n[nb_idx[b:e]].sum() for f, b, e, _ in items], dtype=np.int64)


def _check_items(fs, nb_off, nb_idx, f1_list, items, costs, item_cost):
    n = np.diff(fs.pix_off)
    k = 0
    for f in f1_list:
        b, e = nb_off[f], nb_off[f + 1]
        pos = b
        while pos < e:
            f_, ib, ie, slot = items[k]
            assert (f_, ib, slot) == (f, pos, fs.slot[f]), k
            assert ie > ib
            assert costs[k] == n[f] * n[nb_idx[ib:ie]].sum()
            assert costs[k] <= item_cost or ie - ib == 1
            if ie < e:  # greedy: one more neighbour would exceed the bound
                assert costs[k] + n[f] * n[nb_idx[ie]] > item_cost
            pos = ie
            k += 1
    assert k == len(items)


def _check_tiles(items, costs, tiles, tile_cost):
    tc = []
    nxt = 0
    for b, e in sorted(map(tuple, tiles)):
        assert b == nxt and e > b
        assert len(set(items[b:e, 3])) == 1
        c = costs[b:e].sum()
        assert c <= tile_cost or e - b == 1
        if e < len(items) and items[e, 3] == items[b, 3]:
            assert c + costs[e] > tile_cost
        nxt = e
    assert nxt == len(items)
    tc = np.array([costs[b:e].sum() for b, e in tiles])
    assert np.all(np.diff(tc) <= 0)
    for a in range(len(tiles) - 1):  # stable: equal costs keep item order
        if tc[a] == tc[a + 1]:
            assert tiles[a][0] < tiles[a + 1][0]


@pytest.mark.parametrize("item_cost,tile_cost", [(20_000, 60_000), (1, 1), (10 ** 12, 10 ** 12)])
def test_host_work_invariants(small_set, small_pairs, item_cost, tile_cost):
    fs = small_set
    _, (off, idx) = small_pairs
    f1_list = np.argsort(fs.slot, kind="stable").astype(np.int32)
    items, tiles, total = xi.build_work(fs, off, idx, f1_list, item_cost, tile_cost)
    costs = _item_costs(fs, off, idx, items)
    _check_items(fs, off, idx, f1_list, items, costs, item_cost)
    _check_tiles(items, costs, tiles, tile_cost)
    n = np.diff(fs.pix_off)
    assert total == float(sum(n[f] * n[idx[off[f]:off[f + 1]]].sum() for f in range(fs.nf)))
    # the tile step alone, from the items and their costs
    assert np.array_equal(xi.build_tiles(items, costs, tile_cost), tiles)


def test_build_tiles_empty_and_errors():
    assert xi.build_tiles(np.zeros((0, 4), np.int32), np.zeros(0, np.int64), 10).shape == (0, 2)
    with pytest.raises(RuntimeError):
        xi.build_tiles(np.zeros((1, 4), np.int32), np.ones(1, np.int64), 0)


# ----------------------------------------------------------------- GPU builders
def _raw_set(n, seed, angle_deg=4.0, ra_ties=False, name_dups=False):
    """Unit vectors in a small cap, with optional ra ties and duplicated names."""
    rng = np.random.default_rng(seed)
    ra = np.deg2rad(rng.uniform(0, angle_deg, n))
    dec = np.arcsin(rng.uniform(-np.sin(np.deg2rad(angle_deg / 2)), np.sin(np.deg2rad(angle_deg / 2)), n))
    if ra_ties:
        ra[rng.integers(0, n, n // 4)] = ra[0]
    name = np.arange(n, dtype=np.int64) + 1000
    if name_dups:
        name[rng.integers(0, n, n // 5)] = name[1]
    return types.SimpleNamespace(nf=n, ra=ra, name=name, ux=np.cos(dec) * np.cos(ra), uy=np.cos(dec) * np.sin(ra),
                                 uz=np.sin(dec))


def _device_pairs(s, angmax):
    import torch
    t = lambda a, d: torch.from_numpy(np.ascontiguousarray(a, dtype=d)).cuda()
    off, idx, n = xi.build_pairs_device(t(s.ux, np.float64), t(s.uy, np.float64), t(s.uz, np.float64),
                                        t(s.ra, np.float64), t(s.name, np.int64), angmax)
    return off.cpu().numpy(), idx.cpu().numpy(), n


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["plain", "ties", "full_sky", "tiny_angle", "one", "empty"])
def test_device_pairs_match_oracle(case):
    kw = {"plain": dict(n=1500, seed=1), "ties": dict(n=1500, seed=2, ra_ties=True, name_dups=True),
          "full_sky": dict(n=700, seed=3, angle_deg=170.0), "tiny_angle": dict(n=3000, seed=4, angle_deg=0.5),
          "one": dict(n=1, seed=5), "empty": dict(n=0, seed=6)}[case]
    s = _raw_set(**kw)
    angmax = {"full_sky": np.pi, "tiny_angle": 2e-4}.get(case, 0.0565)
    print("device pairs", case, s.nf, flush=True)
    off0, idx0 = oracle.pairs(s, angmax)
    off, idx, n = _device_pairs(s, angmax)
    assert n == len(idx0)
    assert np.array_equal(off, off0)
    assert np.array_equal(idx, idx0)


@pytest.mark.gpu
def test_device_pairs_and_items_match_host():
    import torch
    s = synth.generate(synth.scaled(synth.CONFIGS["sky100k"], 20_000, seed=11))
    fs = forests.ForestSet.from_survey(s)
    am = forests.angmax_from(fs, 200.0)
    print("device vs host pairs", fs.nf, flush=True)
    off_h, idx_h = xi.build_pairs(fs, am)
    off, idx, n = _device_pairs(fs, am)
    assert np.array_equal(off, off_h) and np.array_equal(idx, idx_h)
    f1_list = np.argsort(fs.slot, kind="stable").astype(np.int32)
    t = lambda a, d: torch.from_numpy(np.ascontiguousarray(a, dtype=d)).cuda()
    items_d, costs_d, total_d = xi.build_items_device(t(f1_list, np.int32), t(fs.slot[f1_list], np.int32),
                                                      t(fs.pix_off, np.int64), t(off_h, np.int64),
                                                      t(idx_h, np.int32), 300_000)
    items_h, tiles_h, total_h = xi.build_work(fs, off_h, idx_h, f1_list, 300_000, 3_000_000)
    assert np.array_equal(items_d.cpu().numpy(), items_h)
    assert total_d == total_h
    costs = costs_d.cpu().numpy()
    assert np.array_equal(costs, _item_costs(fs, off_h, idx_h, items_h))
    assert np.array_equal(xi.build_tiles(items_h, costs, 3_000_000), tiles_h)


@pytest.mark.gpu
def test_device_pairs_row_limit():
    """A forest with more than 16384 neighbours is refused with an error (no fill launched)."""
    rng = np.random.default_rng(9)
    n = 16500
    ra = np.sort(rng.uniform(0, 1e-3, n))
    dec = rng.uniform(-1e-3, 1e-3, n)
    s = types.SimpleNamespace(nf=n, ra=ra, name=np.arange(n, dtype=np.int64),
                              ux=np.cos(dec) * np.cos(ra), uy=np.cos(dec) * np.sin(ra), uz=np.sin(dec))
    with pytest.raises(RuntimeError, match="neighbours"):
        _device_pairs(s, 0.05)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_engine_device_builder_same_results(small_set, mode):
    fs = small_set
    am = forests.angmax_from(fs, 200.0)
    print("engine device builder", mode, flush=True)
    e_h = xi.XiEngine(fs, am, item_cost=40_000, pair_builder="host")
    e_d = xi.XiEngine(fs, am, item_cost=40_000, pair_builder="device")
    assert e_d.pixel_pairs == e_h.pixel_pairs and e_d.n_pairs == e_h.n_pairs
    assert np.array_equal(e_d.nb_off, e_h.nb_off) and np.array_equal(e_d.nb_idx, e_h.nb_idx)
    cfg = xi.default_config(mode)
    e_h.run(cfg)
    e_d.run(cfg)
    c0, w0, d0 = e_h.results()
    c1, w1, d1 = e_d.results()
    assert np.array_equal(c0, c1)
    np.testing.assert_allclose(w1, w0, rtol=1e-12, atol=0)
    np.testing.assert_allclose(d1, d0, rtol=1e-9, atol=1e-12 * np.abs(w0).max())


@pytest.mark.gpu
def test_engine_falls_back_to_host_pair_list():
    """A forest with more than 16384 neighbours: the engine warns and builds the list
    on the host; the histograms still match the oracle."""
    rng = np.random.default_rng(12)
    n = 16500
    ra = np.sort(rng.uniform(0, 1e-3, n))
    dec = rng.uniform(-1e-3, 1e-3, n)
    npx = 2
    off = np.arange(n + 1, dtype=np.int64) * npx
    dc = np.repeat(rng.uniform(3600.0, 3700.0, n), npx) + np.tile(np.arange(npx) * 2.0, n)
    fs = forests.ForestSet(off, [0], ux=np.cos(dec) * np.cos(ra), uy=np.cos(dec) * np.sin(ra), uz=np.sin(dec),
                           ra=ra, dec=dec, name=np.arange(n, dtype=np.int64), slot=np.zeros(n, np.int32),
                           omega=np.ones(n), odl2=np.ones(n), dc=dc, we=rng.uniform(0.5, 2.0, n * npx),
                           dw=rng.normal(size=n * npx), dl=np.zeros(n * npx))
    am = forests.angmax_from(fs, 200.0)
    print("fallback", n, flush=True)
    with pytest.warns(UserWarning, match="host builder"):
        eng = xi.XiEngine(fs, am)
    assert eng.pair_builder == "host"
    eng.run(xi.default_config("fp64"))
    c, w, _ = eng.results()
    off0, idx0 = oracle.pairs(fs, am)
    assert np.array_equal(eng.nb_off, off0) and np.array_equal(eng.nb_idx, idx0)
    c0, w0, _ = oracle.xi_gpu_form(fs, off0, idx0)
    assert np.array_equal(c, c0)
    np.testing.assert_allclose(w, w0, rtol=1e-10)
