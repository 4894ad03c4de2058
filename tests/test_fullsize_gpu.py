"""The benchmarked workload itself against the oracle: 100k forests over 2,500 deg^2,
HEALPix nside 16, 50 x 50 bins (reference cuda_kernels.cpp:90-162 at full size).

Three seeded output slots (HEALPix pixels of forest 1) are computed by the tuned
fp64 and mixed kernels on the full survey (device pair list, all forests as forest
2) and by the OpenMP form of the oracle over the same slots, whose neighbour rows
come from the O(N) per-forest rule in numpy.  Counts bit-exact, sums within 1e-12
of the per-bin absolute sums; every other slot stays zero.
"""
import os

import numpy as np
import pytest

import oracle
from lya2pcf_amd import autotune, forests, synth, xi

pytestmark = pytest.mark.gpu


def test_sky100k_slots_match_oracle():
    s = synth.generate(synth.CONFIGS["sky100k"])
    fs = forests.ForestSet.from_survey(s, float32_inputs=True)
    del s
    am = forests.angmax_from(fs, 200.0)
    rng = np.random.default_rng(2024)
    slots = np.sort(rng.choice(len(fs.slots), 3, replace=False))
    f1 = np.nonzero(np.isin(fs.slot, slots))[0].astype(np.int32)
    print("slots", slots.tolist(), "forests", len(f1), flush=True)
    off, idx = oracle.pairs_of(fs, am, f1)
    threads = min(16, os.cpu_count() or 1)
    c0, w0, d0 = oracle.xi_gpu_form(fs, off, idx, f1_list=f1, threads=threads)
    _, aw, ad = oracle.xi_gpu_form(fs, off, idx, f1_list=f1, we=np.abs(fs.we), dw=np.abs(fs.dw), threads=threads)
    print("oracle pairs", int(c0.sum()), flush=True)
    assert c0[slots].sum() > 1e9
    eng = xi.XiEngine(fs, am, slots=slots.tolist())
    assert len(eng.f1_list) == len(f1)
    others = np.setdiff1d(np.arange(len(fs.slots)), slots)
    for mode in ("fp64", "mixed"):
        shape = dict(autotune.load_best(mode) or {})
        eng.run(xi.default_config(mode, **shape))
        c, w, d = eng.results()
        assert np.array_equal(c[slots], c0[slots]), mode
        assert not c[others].any() and not w[others].any(), mode
        assert np.all(np.abs(w - w0) <= 1e-12 * aw + 1e-300), mode
        assert np.all(np.abs(d - d0) <= 1e-12 * ad + 1e-300), mode
