"""ctypes binding of liblya2pcf.so (the C ABI declared in include/lya2pcf.h).

The library is built in-tree by ``lya2pcf_amd.build`` into
``lya2pcf_amd/_build/liblya2pcf.so``.  Loading fails loudly when it is
missing: there is no CPU fallback behind these entry points.
"""
import ctypes as C
import os

LIB_PATH = os.environ.get("LYA2PCF_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build",
                                                         "liblya2pcf.so")

LYA_MODE_FP64 = 0
LYA_MODE_MIXED = 1
LYA_MODE_MIXED_F32 = 2
LYA_MAX_BINS = 2550
LYA_DIST_SKIP_LONG = 1   # uint32 words of the distortion scratch counting skipped pairs
LYA_DIST_SKIP_BINS = 2
LYA_DIST_NOALLOC_BYTES = 32 << 20   # built-in scratch of lya_distortion_run_noalloc

_vp = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64
_f64 = C.c_double


class XiProblem(C.Structure):
    """Mirror of ``lya_xi_problem``."""
    _fields_ = [
        ("n_forest", _i32), ("n_slots", _i32),
        ("pix_off", _vp), ("ux", _vp), ("uy", _vp), ("uz", _vp), ("base", _vp), ("sorted", _vp),
        ("dc", _vp), ("we", _vp), ("dw", _vp), ("off32", _vp), ("we32", _vp), ("dw32", _vp),
        ("nb_idx", _vp), ("items", _vp), ("tiles", _vp),
        ("n_items", _i32), ("n_tiles", _i32), ("np_rp", _i32), ("np_rt", _i32),
        ("rpmax", _f64), ("rtmax", _f64), ("tau", C.c_float), ("lane_map", _i32), ("max_len", _i32),
        ("pad", _i32),
        ("out_cnt", _vp), ("out_w", _vp), ("out_dw", _vp), ("queue", _vp),
        ("tile_groups", _vp), ("n_tile_groups", _i32),
    ]


class XiConfig(C.Structure):
    """Mirror of ``lya_xi_config`` (``lane_map`` is a Python-side field copied into
    the problem's ``lane_map``)."""
    _fields_ = [("mode", _i32), ("px_per_lane", _i32), ("waves", _i32), ("replicas", _i32),
                ("row_pad", _i32), ("groups", _i32), ("no_counts", _i32), ("walk", _i32),
                ("merge", _i32), ("tail_max", _i32)]
    lane_map = 1

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


LYA_ERR_ARG, LYA_ERR_HIP, LYA_ERR_CAPACITY, LYA_ERR_UNSUPPORTED = -1, -2, -3, -4


class LyaError(RuntimeError):
    """A nonzero LYA_ERR_* status; ``code`` holds it."""

    def __init__(self, msg, code=0):
        super().__init__(msg)
        self.code = code



_lib = None


def lib():
    """Load the library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LyaError("liblya2pcf.so not built (%s); run `python -m lya2pcf_amd.build`" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    L.lya_last_error.restype = C.c_char_p
    L.lya_version.restype = C.c_int
    L.lya_build_pairs.argtypes = [_i32, _vp, _vp, _vp, _vp, _vp, _f64, _vp, _vp, _i64, _vp]
    L.lya_build_work.argtypes = [_i32, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp]
    L.lya_xi_tau.argtypes = [C.POINTER(XiProblem), _f64, _f64, _f64, _f64, C.POINTER(_f64)]
    L.lya_xi_run.argtypes = [C.POINTER(XiProblem), C.POINTER(XiConfig), _vp]
    L.lya_xi_graph_create.argtypes = [C.POINTER(XiProblem), C.POINTER(XiConfig), _vp, C.POINTER(_vp)]
    L.lya_graph_launch.argtypes = [_vp, _vp]
    L.lya_graph_destroy.argtypes = [_vp]
    L.lya_xi_grid.argtypes = [C.POINTER(XiConfig), C.POINTER(_i32)]
    L.lya_xi_work_stats.argtypes = [C.POINTER(XiProblem), C.POINTER(XiConfig), _vp, _vp]
    L.lya_xi_adjoint.argtypes = [C.POINTER(XiProblem), C.POINTER(XiConfig), _vp, _vp, _vp, _vp, _vp]
    L.lya_distortion_scratch_bytes.argtypes = [_i32, _i32, _i32]
    L.lya_distortion_scratch_bytes.restype = _i64
    L.lya_distortion_run.argtypes = [_i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _i32,
                                     _i32, _i32, _f64, _f64, _vp, _vp, _vp, _i64, _i32, _vp]
    L.lya_distortion_run_noalloc.argtypes = [_i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _i32,
                                             _i32, _i32, _f64, _f64, _vp, _vp, _vp, _i32, _vp]
    L.lya_distortion_mirror.argtypes = [_vp, _i32, _vp]
    L.lya_segment_sum.argtypes = [_vp, _vp, _i32, _vp, _vp]
    L.lya_debug_halfangles.argtypes = [_vp, _vp, _vp, _i32, _vp]
    L.lya_copy_bytes.argtypes = [_vp, _vp, _i64, _vp]
    L.lya_debug_lds_atomic.argtypes = [_i32, _i32, _i32, _vp, _vp]
    L.lya_build_tiles.argtypes = [_i64, _vp, _vp, _i64, _vp, _vp]
    L.lya_pairs_device_scratch_bytes.argtypes = [_i32, _f64, _vp]
    L.lya_build_pairs_device.argtypes = [_i32, _vp, _vp, _vp, _vp, _vp, _f64, _vp, _vp, _i64, _vp, _i64, _vp, _vp]
    L.lya_items_device_scratch_bytes.argtypes = [_i32, _vp]
    L.lya_build_items_device.argtypes = [_i32, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _i64, _vp, _i64, _vp, _vp,
                                         _vp]
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise LyaError("lya2pcf error %d: %s" % (rc, lib().lya_last_error().decode(errors="replace")), rc)
    return rc


# Every symbol include/lya2pcf.h declares.
EXPORTS = ("lya_last_error", "lya_version", "lya_build_pairs", "lya_build_work", "lya_xi_tau", "lya_xi_run",
           "lya_xi_graph_create", "lya_graph_launch", "lya_graph_destroy", "lya_xi_grid", "lya_xi_work_stats",
           "lya_xi_adjoint",
           "lya_distortion_scratch_bytes", "lya_distortion_run", "lya_distortion_run_noalloc",
           "lya_distortion_mirror", "lya_segment_sum", "lya_debug_halfangles", "lya_copy_bytes",
           "lya_debug_lds_atomic",
           "lya_build_tiles", "lya_pairs_device_scratch_bytes", "lya_build_pairs_device",
           "lya_items_device_scratch_bytes", "lya_build_items_device")


def ptr(t):
    """Raw address of a torch tensor or numpy array (None -> NULL)."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream
