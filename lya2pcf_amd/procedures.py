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
  compute_pixels  pixels that will be asked for (default: every pixel of data)
  precision       'fp64' (reference formula, default) or 'mixed'
  config          an XiConfig (default: the tuned shape of the precision)
  counts          also count pixel pairs per bin (default False, as the reference,
                  whose outputs are the two sums; on costs one LDS atomic per run)
  rpmax, rtmax    default from the ``parameters`` module
Pair counts per bin are available from ``pair_counts(pixel)``.
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
    _S.update(fs=fs, slot_of=slot_of, shape=(np_rp, np_rt), precision=precision,
              engine=_xi.XiEngine(fs, float(angmax_aux), np_rp=np_rp, np_rt=np_rt,
                                  rpmax=rpmax if rpmax is not None else dr,
                                  rtmax=rtmax if rtmax is not None else dt, slots=want),
              config=config, counts=bool(counts), done=False)
    if log_file_aux is not None:
        log_file_aux.write("\nHIP pair list: %d forests, %d forest pairs, %.4g pixel pairs"
                           % (fs.nf, _S["engine"].n_pairs, _S["engine"].pixel_pairs))
        log_file_aux.flush()


def _compute():
    if _S.get("done"):
        return
    eng = _S["engine"]
    cfg = _S["config"]
    if cfg is None:
        # the tuned kernel shape of this precision, if a tuning table is present
        from . import autotune
        shape = autotune.load_best(_S["precision"]) or {}
        cfg = _xi.default_config(_S["precision"], counts=_S["counts"], **shape)
    g = eng.graph(cfg)
    g.launch()
    _S["cnt"], _S["w"], _S["dw"] = eng.results()
    g.close()
    _S["done"] = True


def two_point_per_pixel(pixel, **kargs):
    """(w_hist, dw_hist) of the pairs whose forest 1 lies in HEALPix ``pixel``."""
    if "engine" not in _S:
        raise RuntimeError("init() must be called first")
    _compute()
    k = _S["slot_of"].get(pixel)
    if k is None:
        z = np.zeros(_S["shape"])
        return (z, z.copy())
    return (_S["w"][k].copy(), _S["dw"][k].copy())


def pair_counts(pixel):
    """Number of pixel pairs per bin (uint64) for HEALPix ``pixel``."""
    if not _S.get("counts", False):
        raise RuntimeError("pair counts were not requested (init(..., counts=False))")
    _compute()
    k = _S["slot_of"].get(pixel)
    if k is None:
        return np.zeros(_S["shape"], dtype=np.uint64)
    return _S["cnt"][k].copy()


def pixel_pairs():
    """Sum of n1*n2 over the forest-pair list of the computed pixels."""
    return _S["engine"].pixel_pairs
