"""Python side of the CPU oracle.  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker (or the timed CPU baseline); the product path never
does.  It wraps oracle/_build/liboracle.so (FP64 C restatements, see
xi_oracle.c for the reference file:line each follows) and adds numpy
brute-force forms for tiny inputs:

  numpy_xi_gpu_form   vectorised O(n1 n2) per forest pair with np.floor
                      (cuda_kernels.cpp:131-156)
  numpy_pairs         O(N^2) neighbour rule (forest_class.py:96)
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

_vp, _i32, _i64, _f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_double
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.oracle_pairs.restype = _i64
        L.oracle_pairs.argtypes = [_i32, _vp, _vp, _vp, _vp, _vp, _f64, _vp, _vp]
        L.oracle_xi_gpu_form.argtypes = [_vp] * 11 + [_i32, _i32, _i32, _f64, _f64, _vp, _vp, _vp]
        L.oracle_xi_gpu_form_mt.argtypes = [_vp] * 11 + [_i32, _i32, _i32, _f64, _f64, _i32, _i32, _vp, _vp, _vp]
        L.oracle_xi_cpu_form.argtypes = [_vp] * 10 + [_i32, _i32, _i32, _f64, _f64, _f64, _f64, _vp, _vp, _vp]
        L.oracle_xi_cpu_form_mt.argtypes = [_vp] * 10 + [_i32, _i32, _i32, _f64, _f64, _f64, _f64, _i32, _i32,
                                                         _vp, _vp, _vp]
        L.oracle_distortion_eta.argtypes = [_vp] * 10 + [_i32, _i32, _i32, _f64, _f64, _vp, _vp]
        L.oracle_distortion_direct.argtypes = [_vp] * 10 + [_i32, _i32, _i32, _f64, _f64, _vp, _vp]
        L.oracle_xi_adjoint.argtypes = [_vp] * 10 + [_i32, _i32, _i32, _f64, _f64, _vp, _vp, _vp, _vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data if a is not None else None


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def pairs(fs, angmax):
    """Brute-force neighbour CSR (nb_off, nb_idx) in ascending f2 order."""
    L = lib()
    ux, uy, uz, ra = (_c(getattr(fs, k), np.float64) for k in ("ux", "uy", "uz", "ra"))
    name = _c(fs.name, np.int64)
    off = np.zeros(fs.nf + 1, dtype=np.int64)
    cm = float(np.cos(angmax))
    n = L.oracle_pairs(fs.nf, _p(ux), _p(uy), _p(uz), _p(ra), _p(name), cm, _p(off), None)
    idx = np.zeros(max(1, n), dtype=np.int32)
    L.oracle_pairs(fs.nf, _p(ux), _p(uy), _p(uz), _p(ra), _p(name), cm, _p(off), _p(idx))
    return off, idx[:n]


def _f1(fs, f1_list):
    if f1_list is None:
        f1_list = np.arange(fs.nf, dtype=np.int32)
    f1_list = _c(f1_list, np.int32)
    return f1_list, _c(fs.slot[f1_list], np.int32)


def xi_gpu_form(fs, nb_off, nb_idx, np_rp=50, np_rt=50, rpmax=200.0, rtmax=200.0, f1_list=None, we=None, dw=None,
                threads=0):
    """(cnt, w, dw) of shape [n_slots, np_rp, np_rt] with the GPU-form FP64 formula
    (reference cuda_kernels.cpp:90-162).  threads > 0: the same loop over forest 1
    with OpenMP (same bins and counts; sums in another order)."""
    L = lib()
    f1_list, f1_slot = _f1(fs, f1_list)
    ns = len(fs.slots)
    cnt = np.zeros((ns, np_rp, np_rt), dtype=np.uint64)
    w = np.zeros((ns, np_rp, np_rt))
    d = np.zeros((ns, np_rp, np_rt))
    we = _c(fs.we if we is None else we, np.float64)
    dwa = _c(fs.dw if dw is None else dw, np.float64)
    nbi = _c(nb_idx if len(nb_idx) else np.zeros(1), np.int32)
    args = [_p(_c(fs.pix_off, np.int64)), _p(_c(fs.dc, np.float64)), _p(we), _p(dwa),
            _p(_c(fs.ux, np.float64)), _p(_c(fs.uy, np.float64)), _p(_c(fs.uz, np.float64)),
            _p(_c(nb_off, np.int64)), _p(nbi), _p(f1_list), _p(f1_slot), len(f1_list), np_rp, np_rt, rpmax, rtmax]
    if threads:
        L.oracle_xi_gpu_form_mt(*args, ns, int(threads), _p(cnt), _p(w), _p(d))
    else:
        L.oracle_xi_gpu_form(*args, _p(cnt), _p(w), _p(d))
    return cnt, w, d


def pairs_of(fs, angmax, f1_list):
    """Neighbour CSR (nb_off [nf+1], nb_idx) with rows only for the forests of
    ``f1_list`` (others empty): the rule of reference forest_class.py:96 evaluated in
    numpy per forest 1, (x1 x2 + y1 y2) + z1 z2 > cos(angmax), name1 != name2,
    ra1 < ra2; ascending f2.  For checking a few output slots of a large survey."""
    cm = np.cos(angmax)
    rows = {}
    for f1 in np.asarray(f1_list, dtype=np.int64):
        mu = (fs.ux[f1] * fs.ux + fs.uy[f1] * fs.uy) + fs.uz[f1] * fs.uz
        rows[int(f1)] = np.nonzero((mu > cm) & (fs.name != fs.name[f1]) & (fs.ra[f1] < fs.ra))[0].astype(np.int32)
    counts = np.zeros(fs.nf, dtype=np.int64)
    for f1, r in rows.items():
        counts[f1] = len(r)
    off = np.zeros(fs.nf + 1, dtype=np.int64)
    np.cumsum(counts, out=off[1:])
    idx = np.concatenate([rows[f] for f in sorted(rows)]) if rows else np.zeros(0, np.int32)
    return off, idx.astype(np.int32)


def xi_cpu_form(fs, nb_off, nb_idx, angmax, chiquito, np_rp=50, np_rt=50, rpmax=200.0, rtmax=200.0, f1_list=None,
                threads=0):
    """CPU-form FP64 (angles from ra/dec).  threads=0: sequential oracle;
    threads>0: the multi-core baseline (same bins, sums in another order)."""
    L = lib()
    f1_list, f1_slot = _f1(fs, f1_list)
    ns = len(fs.slots)
    cnt = np.zeros((ns, np_rp, np_rt), dtype=np.uint64)
    w = np.zeros((ns, np_rp, np_rt))
    d = np.zeros((ns, np_rp, np_rt))
    nbi = _c(nb_idx if len(nb_idx) else np.zeros(1), np.int32)
    args = [_p(_c(fs.pix_off, np.int64)), _p(_c(fs.dc, np.float64)), _p(_c(fs.we, np.float64)),
            _p(_c(fs.dw, np.float64)), _p(_c(fs.ra, np.float64)), _p(_c(fs.dec, np.float64)),
            _p(_c(nb_off, np.int64)), _p(nbi), _p(f1_list), _p(f1_slot), len(f1_list), np_rp, np_rt,
            rpmax, rtmax, float(angmax), float(chiquito)]
    if threads:
        L.oracle_xi_cpu_form_mt(*args, ns, int(threads), _p(cnt), _p(w), _p(d))
    else:
        L.oracle_xi_cpu_form(*args, _p(cnt), _p(w), _p(d))
    return cnt, w, d


def distortion(fs, pairs_arr, np_rp=50, np_rt=50, rpmax=200.0, rtmax=200.0, form="eta"):
    """(D [nb, nb], wA [nb]) unnormalised, FP64; form 'eta' or 'direct'."""
    L = lib()
    nb = np_rp * np_rt
    D = np.zeros((nb, nb))
    wA = np.zeros(nb)
    pr = _c(pairs_arr, np.int32).reshape(-1, 2)
    fn = L.oracle_distortion_eta if form == "eta" else L.oracle_distortion_direct
    fn(_p(_c(fs.pix_off, np.int64)), _p(_c(fs.dc, np.float64)), _p(_c(fs.we, np.float64)),
       _p(_c(fs.dl, np.float64)), _p(_c(fs.ux, np.float64)), _p(_c(fs.uy, np.float64)), _p(_c(fs.uz, np.float64)),
       _p(_c(fs.omega, np.float64)), _p(_c(fs.odl2, np.float64)), _p(pr if len(pr) else np.zeros((1, 2), np.int32)),
       len(pr), np_rp, np_rt, rpmax, rtmax, _p(D), _p(wA))
    return D, wA


def xi_adjoint(fs, nb_off, nb_idx, alpha, beta, np_rp=50, np_rt=50, rpmax=200.0, rtmax=200.0, f1_list=None,
               we=None, dw=None):
    L = lib()
    f1_list, _ = _f1(fs, f1_list)
    g_dw = np.zeros(fs.n_pix)
    g_w = np.zeros(fs.n_pix)
    nbi = _c(nb_idx if len(nb_idx) else np.zeros(1), np.int32)
    L.oracle_xi_adjoint(_p(_c(fs.pix_off, np.int64)), _p(_c(fs.dc, np.float64)),
                        _p(_c(fs.we if we is None else we, np.float64)), _p(_c(fs.dw if dw is None else dw, np.float64)),
                        _p(_c(fs.ux, np.float64)), _p(_c(fs.uy, np.float64)), _p(_c(fs.uz, np.float64)),
                        _p(_c(nb_off, np.int64)), _p(nbi), _p(f1_list), len(f1_list), np_rp, np_rt, rpmax, rtmax,
                        _p(_c(alpha, np.float64)), _p(_c(beta, np.float64)), _p(g_dw), _p(g_w))
    return g_dw, g_w


# ---------------------------------------------------------------- numpy forms
def numpy_pairs(fs, angmax):
    mu = (fs.ux[:, None] * fs.ux[None, :] + fs.uy[:, None] * fs.uy[None, :]) + fs.uz[:, None] * fs.uz[None, :]
    ok = (mu > np.cos(angmax)) & (fs.name[:, None] != fs.name[None, :]) & (fs.ra[:, None] < fs.ra[None, :])
    off = np.zeros(fs.nf + 1, dtype=np.int64)
    np.cumsum(ok.sum(1), out=off[1:])
    return off, np.nonzero(ok)[1].astype(np.int32)


def numpy_xi_gpu_form(fs, nb_off, nb_idx, np_rp=50, np_rt=50, rpmax=200.0, rtmax=200.0):
    """Vectorised per forest pair; counts exact, sums in numpy's order."""
    ns = len(fs.slots)
    cnt = np.zeros((ns, np_rp * np_rt), dtype=np.uint64)
    w = np.zeros((ns, np_rp * np_rt))
    d = np.zeros((ns, np_rp * np_rt))
    bp, bt = np_rp / rpmax, np_rt / rtmax
    for f1 in range(fs.nf):
        a = fs.forest_slice(f1)
        for f2 in nb_idx[nb_off[f1]:nb_off[f1 + 1]]:
            b = fs.forest_slice(f2)
            cos12 = (fs.ux[f1] * fs.ux[f2] + fs.uy[f1] * fs.uy[f2]) + fs.uz[f1] * fs.uz[f2]
            cos12 = min(1.0, max(-1.0, cos12))
            c, s = np.sqrt(0.5 * (1.0 + cos12)), np.sqrt(0.5 * (1.0 - cos12))
            r1, r2 = fs.dc[a][:, None], fs.dc[b][None, :]
            qp = np.floor((np.abs(r1 - r2) * c) * bp)
            qt = np.floor(((r1 + r2) * s) * bt)
            ok = (qp < np_rp) & (qt < np_rt)
            idx = (qp[ok] * np_rt + qt[ok]).astype(np.int64)
            sl = fs.slot[f1]
            np.add.at(cnt[sl], idx, 1)
            np.add.at(w[sl], idx, (fs.we[a][:, None] * fs.we[b][None, :])[ok])
            np.add.at(d[sl], idx, (fs.dw[a][:, None] * fs.dw[b][None, :])[ok])
    sh = (ns, np_rp, np_rt)
    return cnt.reshape(sh), w.reshape(sh), d.reshape(sh)
