"""C-ABI library: loads, exports every declared symbol, host-side entry points."""
import os
import re

import numpy as np

import oracle
from lya2pcf_amd import _lib, forests, xi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_match_header():
    hdr = open(os.path.join(ROOT, "include", "lya2pcf.h")).read()
    declared = set(re.findall(r"^\s*(?:const char\*|int64_t|int)\s+(lya_\w+)\s*\(", hdr, re.M))
    assert declared == set(_lib.EXPORTS), declared ^ set(_lib.EXPORTS)
    L = _lib.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert L.lya_version() >= 1


def test_pairs_match_bruteforce(small_set, small_pairs):
    am, (off, idx) = small_pairs
    off2, idx2 = oracle.pairs(small_set, am)
    off3, idx3 = oracle.numpy_pairs(small_set, am)
    assert np.array_equal(off, off2) and np.array_equal(idx, idx2)
    assert np.array_equal(off, off3) and np.array_equal(idx, idx3)


def test_pair_rule_details():
    """name equality and ra ties drop a pair (forest_class.py:96)."""
    fs = forests.ForestSet(np.array([0, 2, 4, 6, 8]), [0], ux=np.array([1.0, 1.0, 1.0, 1.0]),
                           uy=np.zeros(4), uz=np.zeros(4), ra=np.array([0.1, 0.1, 0.2, 0.3]),
                           dec=np.zeros(4), name=np.array([5, 6, 7, 7]), slot=np.zeros(4, np.int32),
                           omega=np.ones(4), odl2=np.ones(4), dc=np.arange(8.0) + 4000, we=np.ones(8),
                           dw=np.ones(8), dl=np.zeros(8))
    off, idx = xi.build_pairs(fs, 0.01)
    lists = [list(idx[off[i]:off[i + 1]]) for i in range(4)]
    # 0 and 1 tie in ra -> no pair; 2 and 3 share a name -> no pair
    assert lists == [[2, 3], [2, 3], [], []]


def test_build_work_covers_pairs(small_set, small_pairs):
    am, (off, idx) = small_pairs
    f1 = np.arange(small_set.nf, dtype=np.int32)
    items, tiles, tot = xi.build_work(small_set, off, idx, f1, 300_000, 3_000_000)
    # every neighbour range covered once, items inside one slot per tile
    cover = np.zeros(len(idx), dtype=np.int64)
    for f, b, e, s in items:
        assert off[f] <= b < e <= off[f + 1]
        assert s == small_set.slot[f]
        cover[b:e] += 1
    assert np.all(cover == 1)
    seen = np.zeros(len(items), dtype=np.int64)
    for b, e in tiles:
        assert len(set(items[b:e, 3])) == 1
        seen[b:e] += 1
    assert np.all(seen == 1)
    n = np.diff(small_set.pix_off)
    f1r = np.repeat(np.arange(small_set.nf), np.diff(off))
    assert tot == float((n[f1r] * n[idx]).sum())


def test_tau_bound_small():
    """The FP64 recheck band is a few 1e-5 bins for these magnitudes."""
    import ctypes as C
    p = _lib.XiProblem()
    p.np_rp = p.np_rt = 50
    t = C.c_double()
    _lib.check(_lib.lib().lya_xi_tau(C.byref(p), 210.0, 6000.0, 0.25, 0.25, C.byref(t)))
    assert 1e-6 < t.value < 1e-4
    # the FP64 term grows with the largest distance; the FP32 terms dominate
    t2 = C.c_double()
    _lib.check(_lib.lib().lya_xi_tau(C.byref(p), 210.0, 1e9, 0.25, 0.25, C.byref(t2)))
    assert t2.value > t.value and t2.value < 1.01 * t.value + 1.25 * 8 * 2.0 ** -53 * 2e9 * 0.25
    assert _lib.lib().lya_xi_tau(C.byref(p), 210.0, -1.0, 0.25, 0.25, C.byref(t2)) == _lib.LYA_ERR_ARG


def test_missing_library_fails_loudly(tmp_path):
    """Without the HIP library the product path raises; there is no fallback."""
    import subprocess
    import sys
    code = ("from lya2pcf_amd import _lib, xi\n"
            "try:\n"
            "    _lib.lib()\n"
            "except _lib.LyaError as e:\n"
            "    print('raised', e)\n")
    env = dict(os.environ, LYA2PCF_LIB=str(tmp_path / "absent" / "liblya2pcf.so"))
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.startswith("raised") and "not built" in out.stdout


def test_histogram_capacity_refused_before_any_device_call():
    """Binnings beyond the LDS histogram run in bands with the fp64 and mixed tile
    kernels (8 waves, one replica); other shapes, mixed32 and the diagonal walk refuse
    them while the configuration is resolved, before the library touches a device or
    the problem's buffers."""
    import ctypes as C
    L = _lib.lib()

    def run(np_rp, np_rt, mode="fp64", **cfg):
        p = _lib.XiProblem()  # binning only: every buffer pointer stays NULL
        p.np_rp, p.np_rt, p.rpmax, p.rtmax = np_rp, np_rt, 4.0 * np_rp, 4.0 * np_rt
        return L.lya_xi_run(C.byref(p), C.byref(xi.default_config(mode, counts=False, **cfg)), None)

    accepted = (_lib.LYA_ERR_ARG, _lib.LYA_ERR_HIP)   # NULL buffers, or no device here
    # one histogram: np_rp * (np_rt + 1) entries, at most LYA_MAX_BINS = 50 * 51
    assert _lib.LYA_MAX_BINS == 50 * 51
    assert run(50, 50) in accepted
    # larger binnings: banded tile kernels
    for np_rp, np_rt in ((51, 50), (100, 100), (100, 37), (200, 200), (1400, 30), (3000, 3)):
        for mode in ("fp64", "mixed"):
            for merge in (0, 1):
                assert run(np_rp, np_rt, mode, merge=merge) in accepted, (np_rp, np_rt, mode, merge)
    for bad in (dict(mode="mixed32"), dict(waves=4), dict(waves=16), dict(replicas=2)):
        assert run(100, 100, **bad) == _lib.LYA_ERR_UNSUPPORTED, bad
        assert b"bands" in L.lya_last_error()
        assert run(50, 50, **bad) in accepted, bad
    # merged runs always take the padded stride: 50 x 51 bins fit unpadded only
    assert run(50, 51, row_pad=0) in accepted
    assert run(50, 51, "mixed32", row_pad=0) in accepted
    assert run(50, 51, "mixed32", row_pad=1) == _lib.LYA_ERR_UNSUPPORTED
    # diagonal walk: (np_rp + 1) * (np_rt + 1) LDS bins, at most 2688
    assert run(52, 51, walk=1, waves=16) == _lib.LYA_ERR_UNSUPPORTED


def test_mixed_edge_band_flags_every_misbinned_pair():
    """Mixed mode's FP32 arithmetic, emulated in numpy (fma through an exact FP64
    product), against the FP64 reference bins: the FP32 error stays within the bound
    behind tau, every pair whose FP32 bin differs is flagged (fract < 2 tau), and pairs
    at r_par = 0 exactly (pixels on a common wavelength grid) are binned without a flag."""
    import ctypes as C
    rng = np.random.default_rng(11)
    max_off, binner, n = 300.0, 0.25, 400_000
    p = _lib.XiProblem()
    p.np_rp = p.np_rt = 50
    t = C.c_double()
    _lib.check(_lib.lib().lya_xi_tau(C.byref(p), max_off, 5600.0 + max_off, binner, binner, C.byref(t)))
    tau = np.float32(t.value)
    if float(tau) < t.value:
        tau = np.nextafter(tau, np.float32(1))
    f32 = np.float32
    fma32 = lambda a, b, c: (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(f32)
    base2 = rng.integers(3600, 5600, n).astype(np.float64)
    off2 = rng.uniform(-max_off, max_off, n)
    r2 = base2 + off2
    c = np.cos(rng.uniform(0, 0.03, n))          # half-angle cosines of the 3.2 deg cut
    s = np.sqrt(1 - c * c)
    # forest-1 distances: random, within a few tau of an r_par edge, or equal to r2
    rp_bins = rng.uniform(0, 53, n)
    kind = rng.integers(0, 3, n)
    near = kind == 1
    rp_bins[near] = np.floor(rp_bins[near]) + rng.uniform(-3, 3, near.sum()) * float(tau)
    sign = np.where(rng.random(n) < 0.5, 1.0, -1.0)
    r1 = r2 + sign * np.abs(rp_bins) / (c * binner)
    same = kind == 2
    r1[same] = r2[same]
    cb, sb = c * binner, s * binner
    o2 = (r2 - base2).astype(f32)
    # kernel: P1 = (r1 - base2) c b, T1 = (r1 + base2) s b + tau, both rounded once to FP32
    P1 = ((r1 - base2) * cb).astype(f32)
    T1 = ((r1 + base2) * sb + float(tau)).astype(f32)
    qv = (np.abs(fma32(o2, -cb.astype(f32), P1)) + tau).astype(f32)
    qp = np.maximum(qv, f32(0.5))                  # below half a bin: bin 0 either way
    qt = fma32(o2, sb.astype(f32), T1)
    # reference FP64 coordinates (binner a power of two: folding it in is exact)
    qp64 = (np.abs(r1 - r2) * c) * binner
    qt64 = ((r1 + r2) * s) * binner
    rel = (qp64 < 53) & (qt64 < 53)
    bound = t.value / 1.25
    assert np.abs(qv[rel] - (qp64[rel] + float(tau))).max() <= bound
    assert np.abs(qt[rel] - (qt64[rel] + float(tau))).max() <= bound
    flagged = np.minimum(qp - np.floor(qp), qt - np.floor(qt)) < 2 * tau
    wrong = (np.floor(qp) != np.floor(qp64)) | (np.floor(qt) != np.floor(qt64))
    assert wrong[rel].sum() > 0                    # the near-edge half does hit wrong FP32 bins
    assert not np.any(wrong & rel & ~flagged)      # and every one of them is rechecked
    # r_par = 0 exactly: bin 0, and its zero edge calls no recheck (r_perp edges still may)
    assert np.all(qp64[same] == 0) and not np.any(wrong[same])
    assert np.all(qp[same] - np.floor(qp[same]) >= 2 * tau)


def test_group_tiles_partition():
    """XCD tile groups: a permutation of the tiles, contiguous slot ranges per group,
    costs balanced, each group in decreasing cost; one group keeps the tiles as built."""
    rng = np.random.default_rng(5)
    n_items = 400
    slots = np.sort(rng.integers(0, 60, n_items))
    items = np.stack([np.arange(n_items), np.zeros(n_items), np.ones(n_items), slots], 1).astype(np.int32)
    costs = rng.integers(1, 1000, n_items)
    tiles = xi.build_tiles(items, costs, 3000)
    t1, o1 = xi.group_tiles(items, tiles, costs, 1)
    assert np.array_equal(t1, tiles) and list(o1) == [0, len(tiles)]
    t8, o8 = xi.group_tiles(items, tiles, costs, 8)
    assert len(o8) == 9 and o8[0] == 0 and o8[-1] == len(tiles) and np.all(np.diff(o8) >= 0)
    key = lambda t: sorted(map(tuple, t.tolist()))
    assert key(t8) == key(tiles)
    cum = np.concatenate([[0], np.cumsum(costs)])
    tc = cum[t8[:, 1]] - cum[t8[:, 0]]
    sl = items[t8[:, 0], 3]
    g_cost = [tc[o8[g]:o8[g + 1]].sum() for g in range(8)]
    assert max(g_cost) <= sum(g_cost) / 8 + tc.max()
    for g in range(7):
        a, b = sl[o8[g]:o8[g + 1]], sl[o8[g + 1]:o8[g + 2]]
        if len(a) and len(b):
            assert a.max() <= b.min()
    for g in range(8):
        assert np.all(np.diff(tc[o8[g]:o8[g + 1]]) <= 0)
