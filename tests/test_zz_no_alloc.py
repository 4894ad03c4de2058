"""The allocation-free distortion entry point (lya_distortion_run_noalloc, a built-in
device scratch buffer) against the entry point that takes a caller's scratch
(lya_distortion_run), through the C ABI.  Two calls in a row, the second on a smaller
input: the built-in buffer keeps the first call's contents and must not leak them.
Bar as the distortion's: D within 1e-11 of max|D|, wA within 1e-12.
"""
import ctypes as C
import random

import numpy as np
import pytest

from lya2pcf_amd import _lib, forests, synth
from lya2pcf_amd import distortion as dist

pytestmark = pytest.mark.gpu


def _noalloc(eng, pairs, skipped, groups=None):
    import torch
    d = eng.d
    D = torch.zeros((eng.nb, eng.nb), dtype=torch.float64, device="cuda")
    wA = torch.zeros(eng.nb, dtype=torch.float64, device="cuda")
    pt = torch.from_numpy(np.ascontiguousarray(pairs, dtype=np.int32)).cuda()
    L = _lib.lib()
    rc = L.lya_distortion_run_noalloc(eng.fs.nf, _lib.ptr(d["pix_off"]), _lib.ptr(d["dc"]), _lib.ptr(d["we"]),
                                      _lib.ptr(d["dl"]), _lib.ptr(d["ux"]), _lib.ptr(d["uy"]), _lib.ptr(d["uz"]),
                                      _lib.ptr(d["omega"]), _lib.ptr(d["odl2"]), _lib.ptr(pt), len(pairs),
                                      eng.max_len, eng.np_rp, eng.np_rt, eng.rpmax, eng.rtmax, _lib.ptr(D),
                                      _lib.ptr(wA), _lib.ptr(skipped), groups or eng.groups, _lib.stream_ptr())
    if rc:
        return rc, None, None
    _lib.check(L.lya_distortion_mirror(_lib.ptr(D), eng.nb, _lib.stream_ptr()))
    torch.cuda.synchronize()
    return rc, D.cpu().numpy(), wA.cpu().numpy()


def test_noalloc_matches_scratch_entry_point():
    import torch
    s = synth.generate(synth.scaled(synth.CONFIGS["patch2k"], 90, seed=31, length_range=(30, 120)))
    fs = forests.ForestSet.from_survey(s)
    am = forests.angmax_from(fs, 200.0)
    eng = dist.DistortionEngine(fs, am)
    pairs = dist.subsample(fs, eng.nb_off, eng.nb_idx, range(fs.nf), 0.5, rng=random.Random(2))
    assert len(pairs) > 30
    skipped = torch.zeros(2, dtype=torch.int32, device="cuda")
    for k, sub in enumerate((pairs, pairs[: len(pairs) // 3])):
        print("call", k, "pairs", len(sub), flush=True)
        eng.zero()
        eng.accumulate(sub)
        D0, w0 = eng.result()
        rc, D, w = _noalloc(eng, sub, skipped if k == 0 else None)
        assert rc == 0, _lib.lib().lya_last_error()
        scale = np.abs(D0).max()
        assert scale > 0
        assert np.max(np.abs(D - D0)) <= 1e-11 * scale, k
        assert np.all(np.abs(w - w0) <= 1e-12 * np.abs(w0)), k
    assert skipped.cpu().numpy().tolist() == [0, 0]


def test_noalloc_refuses_more_than_its_buffer():
    """More workgroups than the built-in buffer holds: LYA_ERR_CAPACITY with a message,
    before any launch."""
    import torch
    s = synth.generate(synth.scaled(synth.CONFIGS["patch2k"], 20, seed=32, length_range=(30, 60)))
    fs = forests.ForestSet.from_survey(s)
    eng = dist.DistortionEngine(fs, forests.angmax_from(fs, 200.0))
    need = _lib.lib().lya_distortion_scratch_bytes(1000, eng.max_len, eng.nb)
    assert need > _lib.LYA_DIST_NOALLOC_BYTES
    rc, _, _ = _noalloc(eng, np.zeros((1, 2), np.int32), None, groups=1000)
    assert rc == _lib.LYA_ERR_CAPACITY
    assert b"built-in buffer" in _lib.lib().lya_last_error()
    torch.cuda.synchronize()
