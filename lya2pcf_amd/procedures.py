"""Drop-in xi procedures: the module API of the reference's
correlation_procedures_pycuda (init :16-111, two_point_per_pixel :114-171).

    init(data, log_file, shape_hist, angmax, pixel_list=None)
    two_point_per_pixel(pixel, **kargs) -> (w_hist, dw_hist)   float64 (np_rp, np_rt)

The reference uploads the padded arrays in ``init`` and then, per HEALPix
pixel, searches neighbours on the host and launches one small kernel per
forest 1 with a device synchronisation after each.  Here ``init`` packs the
forests once (ragged SoA), builds the whole pair list natively, and the first
``two_point_per_pixel`` call computes every requested pixel in one replay of
the captured persistent kernel; later calls return the cached histograms.

Extra keyword arguments of ``init`` (all optional):
  compute_pixels  pixels computed by the first replay (default: every pixel of data);
                  a pixel of the data outside this list is computed when it is first
                  asked for (one more launch over every pixel not computed yet)
  precision       'fp64' (reference formula, default) or 'mixed'
  config          an XiConfig (default: the tuned shape of the precision)
  counts          also count pixel pairs per bin (default False, as the reference,
                  whose outputs are the two sums; on costs one LDS atomic per run)
  rpmax, rtmax    default from the ``parameters`` module
Pair counts per bin are available from ``pair_counts(pixel)``.

A pixel that is not in the uploaded data raises ``KeyError``, as the
reference's ``data[pixel]`` does (correlation_procedures_pycuda.py:141).
"""
import numpy as np

from . import forests as _forests
from . import xi as _xi

_S = {}


def _params():
    try:
        import parameters as P
        return float(P.rpmax), float(P.rtmax)
    except Exception:
        return 200.0, 200.0


def init(data_aux, log_file_aux, shape_hist_aux, angmax_aux, pixel_list=None, compute_pixels=None,
         precision="fp64", rpmax=None, rtmax=None, config=None, counts=False):
    dr, dt = _params()
    _S.clear()
    _S["log"] = log_file_aux
    fs = _forests.ForestSet.from_data(data_aux, pixel_list)
    slot_of = {p: k for k, p in enumerate(fs.slots)}
    want = None
    if compute_pixels is not None:
        want = [slot_of[p] for p in compute_pixels if p in slot_of]
    np_rp, np_rt = (int(x) for x in shape_hist_aux)
    rp = rpmax if rpmax is not None else dr
    rt = rtmax if rtmax is not None else dt
    want = list(range(len(fs.slots))) if want is None else sorted(set(want))
    _S.update(fs=fs, slot_of=slot_of, shape=(np_rp, np_rt), precision=precision, angmax=float(angmax_aux),
              binning=(rp, rt), engine=_xi.XiEngine(fs, float(angmax_aux), np_rp=np_rp, np_rt=np_rt,
                                                    rpmax=rp, rtmax=rt, slots=want),
              pending=want, computed=set(), config=config, counts=bool(counts))
    if log_file_aux is not None:
        log_file_aux.write("\nHIP pair list: %d forests, %d forest pairs, %.4g pixel pairs"
                           % (fs.nf, _S["engine"].n_pairs, _S["engine"].pixel_pairs))
        log_file_aux.flush()


def _config():
    cfg = _S["config"]
    if cfg is None:
        # the tuned kernel shape of this precision, if a tuning table is present
        from . import autotune
        shape = autotune.load_best(_S["precision"]) or {}
        np_rp, np_rt = _S["shape"]
        if np_rp * (np_rt + 1) > _xi.LYA_MAX_BINS and (shape.get("walk") or shape.get("waves", 8) != 8
                                                       or shape.get("replicas", 1) != 1):
            # binnings beyond one LDS histogram run banded, in 8-wave tile kernels
            shape = {"merge": 1, "px_per_lane": shape.get("px_per_lane", 1) if not shape.get("walk") else 1}
        cfg = _xi.default_config(_S["precision"], counts=_S["counts"], **shape)
    return cfg


def _compute(slot=None):
    """Run the pending slots (those of ``compute_pixels``) in one graph replay; a later
    request for a slot outside them runs every slot not computed yet in one more launch."""
    if slot is not None and slot in _S["computed"]:
        return
    if not _S["pending"] and slot is not None:
        nf = len(_S["fs"].slots)
        _S["pending"] = [k for k in range(nf) if k not in _S["computed"]]
        np_rp, np_rt = _S["shape"]
        rp, rt = _S["binning"]
        _S["engine"] = _xi.XiEngine(_S["fs"], _S["angmax"], np_rp=np_rp, np_rt=np_rt, rpmax=rp, rtmax=rt,
                                    slots=_S["pending"])
    if not _S["pending"]:
        return
    eng = _S["engine"]
    g = eng.graph(_config())
    g.launch()
    cnt, w, dw = eng.results()
    g.close()
    if "w" not in _S:
        _S["cnt"], _S["w"], _S["dw"] = cnt, w, dw
    else:
        k = np.asarray(_S["pending"], dtype=np.int64)
        _S["cnt"][k], _S["w"][k], _S["dw"][k] = cnt[k], w[k], dw[k]
    _S["computed"].update(_S["pending"])
    _S["pending"] = []


def _slot(pixel):
    if "engine" not in _S:
        raise RuntimeError("init() must be called first")
    k = _S["slot_of"].get(pixel)
    if k is None:
        raise KeyError(pixel)
    _compute(k)
    return k


def two_point_per_pixel(pixel, **kargs):
    """(w_hist, dw_hist) of the pairs whose forest 1 lies in HEALPix ``pixel``."""
    k = _slot(pixel)
    return (_S["w"][k].copy(), _S["dw"][k].copy())


def pair_counts(pixel):
    """Number of pixel pairs per bin (uint64) for HEALPix ``pixel``."""
    if not _S.get("counts", False):
        raise RuntimeError("pair counts were not requested (init(..., counts=False))")
    k = _slot(pixel)
    return _S["cnt"][k].copy()


def pixel_pairs():
    """Sum of n1*n2 over the forest-pair list of the pixels computed by the first replay."""
    return _S["engine"].pixel_pairs
