"""Distortion matrix on the GPU (FP64 accumulation) and its drop-in module API.

Subsampling follows the reference exactly (distortion_procedures_pycuda.py:11,
:220-225): Python's ``random`` is seeded with 1 when the module is imported,
and for every forest 1, in pixel order and list order, its neighbour list is
shuffled with ``random.shuffle`` and the first ceil(n * (1 - reject)) kept.
The neighbour list is ordered as the reference's ``neighborhood`` returns it:
ascending HEALPix pixel, then the pixel's list order.

Drop-in API (reference distortion_procedures_pycuda.py:19, :195):
    init(data, log_file, shape_hist, angmax, reject_fraction, pixel_list=None)
    distortion_per_pixel(forest_list, **kargs) -> (D f64 [nb, nb], weight f64 [nb])
The driver divides rows of sum(D) by sum(weight) (distortion.py:125).
"""
import ctypes as C
import math
import random

import numpy as np

from . import _lib
from . import forests as _forests
from . import xi as _xi
from ._lib import check, lib, ptr

random.seed(1)


def neighbour_order(fs):
    """Rank of every forest in the reference's neighbour order (pixel value, list position)."""
    pix_val = np.asarray(fs.slots, dtype=np.int64)[fs.slot]
    return np.lexsort((np.arange(fs.nf), pix_val)).argsort()


def subsample(fs, nb_off, nb_idx, f1_forests, reject, rng=random):
    """[(f1, f2)] pairs kept for the forests ``f1_forests`` (in that order)."""
    rank = neighbour_order(fs)
    out = []
    frac = 1. - reject
    for f1 in f1_forests:
        nb = nb_idx[nb_off[f1]:nb_off[f1 + 1]]
        nb = list(nb[np.argsort(rank[nb], kind="stable")])
        keep = int(np.ceil(len(nb) * frac))
        rng.shuffle(nb)
        out.extend((int(f1), int(f2)) for f2 in nb[:keep])
    return np.array(out, dtype=np.int32).reshape(-1, 2)


class DistortionEngine:
    def __init__(self, fs, angmax, np_rp=50, np_rt=50, rpmax=200.0, rtmax=200.0, pairs=None, groups=0):
        import torch
        self.torch = torch
        self.fs = fs
        self.np_rp, self.np_rt = int(np_rp), int(np_rt)
        self.rpmax, self.rtmax = float(rpmax), float(rtmax)
        self.nb = self.np_rp * self.np_rt
        dev = torch.device("cuda")
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        self.d = {k: t(getattr(fs, k)) for k in ("pix_off", "dc", "we", "dl", "ux", "uy", "uz", "omega", "odl2")}
        if pairs is None:
            # the pair list is built on the GPU; the subsampling walks a host copy
            off, idx, _ = _xi.build_pairs_device(self.d["ux"], self.d["uy"], self.d["uz"],
                                                 t(np.asarray(fs.ra, dtype=np.float64)),
                                                 t(np.asarray(fs.name, dtype=np.int64)), angmax)
            pairs = (off.cpu().numpy(), idx.cpu().numpy())
        self.nb_off, self.nb_idx = pairs
        self.max_len = int(np.max(fs.n)) if fs.nf else 1
        props = torch.cuda.get_device_properties(0)
        self.groups = groups or props.multi_processor_count
        nbytes = lib().lya_distortion_scratch_bytes(self.groups, self.max_len, self.nb)
        if nbytes < 0:
            raise _lib.LyaError("bad distortion sizes")
        # zeroed once: words 1 and 2 count skipped pairs over every run
        self.scratch = torch.zeros(int(nbytes), dtype=torch.uint8, device=dev)
        self.D = torch.zeros((self.nb, self.nb), dtype=torch.float64, device=dev)
        self.wA = torch.zeros(self.nb, dtype=torch.float64, device=dev)

    def zero(self):
        self.D.zero_()
        self.wA.zero_()

    def accumulate(self, pairs, stream=None):
        """Add the contributions of forest pairs [(f1, f2)] to D and wA (on device)."""
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        if len(pairs) == 0:
            return
        pt = self.torch.from_numpy(pairs).cuda()
        d = self.d
        check(lib().lya_distortion_run(self.fs.nf, ptr(d["pix_off"]), ptr(d["dc"]), ptr(d["we"]), ptr(d["dl"]),
                                       ptr(d["ux"]), ptr(d["uy"]), ptr(d["uz"]), ptr(d["omega"]), ptr(d["odl2"]),
                                       ptr(pt), len(pairs), self.max_len, self.np_rp, self.np_rt, self.rpmax,
                                       self.rtmax, ptr(self.D), ptr(self.wA), ptr(self.scratch),
                                       self.scratch.numel(), self.groups, _lib.stream_ptr(stream)))
        self._keep = pt

    def skipped(self):
        """(pairs with a forest longer than max_len, pairs with more active bins than the
        LDS holds) skipped since the engine was made; both are zero unless a limit was hit."""
        w = self.scratch[:16].cpu().numpy().view(np.uint32)
        return int(w[_lib.LYA_DIST_SKIP_LONG]), int(w[_lib.LYA_DIST_SKIP_BINS])

    def result(self):
        """(D, wA) on the host; D is completed from its accumulated upper triangle.
        Raises LyaError when a pair was skipped (the sums would be incomplete)."""
        check(lib().lya_distortion_mirror(ptr(self.D), self.nb, _lib.stream_ptr(None)))
        self.torch.cuda.synchronize()
        long_, bins = self.skipped()
        if long_ or bins:
            raise _lib.LyaError("distortion: %d forest pairs skipped (%d with a forest longer than %d pixels, "
                                "%d with more active bins than the LDS holds at %d bins)"
                                % (long_ + bins, long_, self.max_len, bins, self.nb), _lib.LYA_ERR_UNSUPPORTED)
        return self.D.cpu().numpy(), self.wA.cpu().numpy()


# --------------------------------------------------------------------- drop-in
_S = {}


def init(data_aux, log_file_aux, shape_hist_aux, angmax_aux, reject_aux, pixel_list=None):
    try:
        import parameters as P
        rpmax, rtmax = float(P.rpmax), float(P.rtmax)
    except Exception:
        rpmax = rtmax = 200.0
    fs = _forests.ForestSet.from_data(data_aux, pixel_list)
    _S.clear()
    _S["fs"] = fs
    _S["data"] = data_aux
    _S["index"] = {id(q): k for k, q in enumerate(fs.objects)}
    _S["reject"] = float(reject_aux)
    np_rp, np_rt = (int(x) for x in shape_hist_aux)
    _S["eng"] = DistortionEngine(fs, float(angmax_aux), np_rp, np_rt, rpmax, rtmax)


def distortion_per_pixel(forest_list, **kargs):
    eng = _S["eng"]
    f1 = [_S["index"][id(q)] for q in forest_list]
    pairs = subsample(_S["fs"], eng.nb_off, eng.nb_idx, f1, _S["reject"])
    eng.zero()
    eng.accumulate(pairs)
    return eng.result()


def distortion_of_pixels(pixels, log_file=None):
    """Sum of ``distortion_per_pixel`` over the forests of ``pixels`` (in that
    order, so the subsample draws the same random sequence), accumulated on the
    device in FP64: (D f64 [nb, nb], weight f64 [nb])."""
    eng = _S["eng"]
    data = _S["data"]
    eng.zero()
    for k, pixel in enumerate(pixels):
        if log_file is not None:
            log_file.write("\nComputing pixel %s, completed %d%%" % (pixel, int(k / max(1, len(pixels)) * 100)))
            log_file.flush()
        f1 = [_S["index"][id(q)] for q in data[pixel]]
        eng.accumulate(subsample(_S["fs"], eng.nb_off, eng.nb_idx, f1, _S["reject"]))
    return eng.result()
