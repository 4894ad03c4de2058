"""Differentiable xi: gradients of a chi^2 against a model xi with respect to
delta, w and the weight-model parameters.

    W_A  = sum_{pairs in A} w_i w_j,   DW_A = sum_{pairs in A} dw_i dw_j
    xi_A = DW_A / W_A                  (all output slots summed)

``PairSums`` (torch.autograd.Function) maps (dw, w) -> (W, DW) with the
persistent forward kernel and, in backward, calls the adjoint kernel
``lya_xi_adjoint`` twice (pair list and its transpose) to form
    dL/d dw_i = sum_j  alpha_A(i,j) dw_j,   dL/d w_i = sum_j beta_A(i,j) w_j,
alpha = dL/dDW, beta = dL/dW, re-walking the pairs without storing them.
The rest of the map is ordinary torch code:
  weight model  w = 1 / (eta * sigma2_pip + sigma2_lss(z)),  sigma2_lss linear
                in z between 10 nodes (the reference's weights are PICCA's,
                delta_reader.py:49-51; this parametric model is new here);
  projection    dw = w (delta - <delta>_w - dl <dl delta>_w / <dl^2>_w) with
                dl = loglam - <loglam>_w (forest_class.py:54-66);
  chi^2         sum_A (xi_A - m_A)^2 / sigma_A^2 over bins with W_A > 0.
The reference has no gradient path; this whole module is an addition.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib, ptr
from . import xi as _xi


class AdjointEngine:
    """Forward + adjoint kernels on one forest set (all pixels as forest 1)."""

    def __init__(self, fs, angmax, mode="fp64", np_rp=50, np_rt=50, rpmax=200.0, rtmax=200.0, pairs=None):
        self.fwd = _xi.XiEngine(fs, angmax, np_rp, np_rt, rpmax, rtmax, pairs=pairs)
        self.bwd = _xi.XiEngine(fs, angmax, np_rp, np_rt, rpmax, rtmax,
                                pairs=(self.fwd.nb_off, self.fwd.nb_idx), transposed=True)
        # the transposed engine reads the forward engine's pixel buffers
        for k in ("we", "dw", "we32", "dw32"):
            self.bwd.d[k] = self.fwd.d[k]
        self.mode = mode
        # adjoint walks: mode and workgroup size; forward sums: the tuned tile shape of the
        # mode, without per-bin counts (the gradient needs only W and DW)
        self.cfg = _xi.default_config(mode)
        from . import autotune
        shape = dict(autotune.load_best(mode, walk=0) or {})
        shape.pop("walk", None)
        self.cfg_fwd = _xi.default_config(mode, counts=False, **shape)
        self.nb = np_rp * np_rt

    def set_inputs(self, dw, w):
        d = self.fwd.d
        d["dw"].copy_(dw.detach().to(d["dw"].dtype))
        d["we"].copy_(w.detach().to(d["we"].dtype))
        d["dw32"].copy_(dw.detach().float())
        d["we32"].copy_(w.detach().float())

    def sums(self):
        self.fwd.run(self.cfg_fwd)
        return self.fwd.out_w.sum(0), self.fwd.out_dw.sum(0)

    def adjoint(self, alpha, beta):
        torch = self.fwd.torch
        n = self.fwd.forests.n_pix
        g_dw = torch.zeros(n, dtype=torch.float64, device=self.fwd.device)
        g_w = torch.zeros(n, dtype=torch.float64, device=self.fwd.device)
        a = alpha.detach().to(torch.float64).contiguous()
        b = beta.detach().to(torch.float64).contiguous()
        for eng in (self.fwd, self.bwd):
            p = eng.problem(self.cfg)
            check(lib().lya_xi_adjoint(C.byref(p), C.byref(self.cfg), ptr(a), ptr(b), ptr(g_dw), ptr(g_w),
                                       _lib.stream_ptr()))
        return g_dw, g_w


def _pair_sums_fn():
    import torch

    class PairSums(torch.autograd.Function):
        @staticmethod
        def forward(ctx, dw, w, eng):
            eng.set_inputs(dw, w)
            W, DW = eng.sums()
            ctx.eng = eng
            ctx.dtypes = (dw.dtype, w.dtype)
            return W.clone(), DW.clone()

        @staticmethod
        def backward(ctx, gW, gDW):
            eng = ctx.eng
            if gW is None:
                gW = torch.zeros(eng.nb, dtype=torch.float64, device=eng.fwd.device)
            if gDW is None:
                gDW = torch.zeros(eng.nb, dtype=torch.float64, device=eng.fwd.device)
            g_dw, g_w = eng.adjoint(gDW, gW)
            return g_dw.to(ctx.dtypes[0]), g_w.to(ctx.dtypes[1]), None

    return PairSums


_PairSums = None


def pair_sums(dw, w, eng):
    global _PairSums
    if _PairSums is None:
        _PairSums = _pair_sums_fn()
    return _PairSums.apply(dw, w, eng)


def _interp_fn():
    import torch

    class NodeInterp(torch.autograd.Function):
        """s_i = (1 - f_i) v[k_i] + f_i v[k_i + 1]; backward sums per node over the
        pixels pre-sorted by k (no scatter of millions of values into a few nodes)."""

        @staticmethod
        def forward(ctx, v, wm):
            ctx.wm = wm
            return (1.0 - wm.f) * v[wm.k0] + wm.f * v[wm.k0 + 1]

        @staticmethod
        def backward(ctx, g):
            wm = ctx.wm
            a = (g * (1.0 - wm.f))[wm.perm]
            b = (g * wm.f)[wm.perm]
            parts = []
            for k in range(wm.n_nodes):
                t = g.new_zeros(())
                if k < wm.n_nodes - 1 and wm.offs[k + 1] > wm.offs[k]:
                    t = t + a[wm.offs[k]:wm.offs[k + 1]].sum()
                if k > 0 and wm.offs[k] > wm.offs[k - 1]:
                    t = t + b[wm.offs[k - 1]:wm.offs[k]].sum()
                parts.append(t)
            return torch.stack(parts), None

    return NodeInterp


_INTERP = None


class WeightModel:
    """w = 1 / (eta * sigma2_pip + sigma2_lss(z)), sigma2_lss linear between nodes."""

    def __init__(self, z, sigma2_pip, nodes, device):
        import torch
        global _INTERP
        if _INTERP is None:
            _INTERP = _interp_fn()
        nodes = np.asarray(nodes, dtype=np.float64)
        zc = np.clip(z, nodes[0], nodes[-1])
        k0 = np.clip(np.searchsorted(nodes, zc, side="right") - 1, 0, len(nodes) - 2)
        f = (zc - nodes[k0]) / (nodes[k0 + 1] - nodes[k0])
        self.n_nodes = len(nodes)
        perm = np.argsort(k0, kind="stable")
        # pixels of interval k (between nodes k and k+1) are perm[offs[k]:offs[k+1]]
        self.offs = [int(x) for x in np.searchsorted(k0[perm], np.arange(len(nodes)), side="left")]
        self.perm = torch.from_numpy(perm.astype(np.int64)).to(device)
        self.k0 = torch.from_numpy(k0.astype(np.int64)).to(device)
        self.f = torch.from_numpy(f).to(device)
        self.s2pip = torch.from_numpy(np.asarray(sigma2_pip, dtype=np.float64)).to(device)

    def __call__(self, eta, s2lss):
        s = _INTERP.apply(s2lss, self)
        return 1.0 / (eta * self.s2pip + s)


def _segment_fns():
    import torch

    def seg_sum(x, off, nf):
        if not (x.is_cuda and off.is_cuda):
            raise ValueError("segment sum: device tensors expected")
        x = x.contiguous()
        out = torch.empty(nf, dtype=torch.float64, device=x.device)
        check(lib().lya_segment_sum(ptr(x), ptr(off), nf, ptr(out), _lib.stream_ptr()))
        return out

    class SegmentSum(torch.autograd.Function):
        """Per-forest sums of a pixel vector (HIP kernel); backward broadcasts."""

        @staticmethod
        def forward(ctx, x, proj):
            ctx.proj = proj
            return seg_sum(x.to(torch.float64), proj.off, proj.nf)

        @staticmethod
        def backward(ctx, g):
            return g[ctx.proj.seg], None

    class Broadcast(torch.autograd.Function):
        """Per-forest values onto their pixels; backward sums per forest (HIP kernel)."""

        @staticmethod
        def forward(ctx, v, proj):
            ctx.proj = proj
            return v[proj.seg]

        @staticmethod
        def backward(ctx, g):
            return seg_sum(g.to(torch.float64), ctx.proj.off, ctx.proj.nf), None

    return SegmentSum, Broadcast


_SEG = None


class Projection:
    """dw = w (delta - <delta>_w - dl <dl delta>_w / <dl^2>_w), per forest."""

    def __init__(self, pix_off, loglam, device):
        import torch
        global _SEG
        if _SEG is None:
            _SEG = _segment_fns()
        n = np.diff(pix_off)
        self.seg = torch.from_numpy(np.repeat(np.arange(len(n)), n)).to(device)
        self.off = torch.from_numpy(np.asarray(pix_off, dtype=np.int64)).to(device)
        self.nf = len(n)
        self.loglam = torch.from_numpy(np.asarray(loglam, dtype=np.float64)).to(device)

    def _sum(self, x):
        if self.off.is_cuda:
            return _SEG[0].apply(x, self)
        # host tensors (CPU tests of the projection algebra): torch's own segment sum
        import torch
        return torch.zeros(self.nf, dtype=x.dtype).index_add_(0, self.seg, x)

    def _bcast(self, v):
        return _SEG[1].apply(v, self) if self.off.is_cuda else v[self.seg]

    def __call__(self, delta, w):
        om = self._sum(w)
        avg = lambda x: self._bcast(self._sum(w * x) / om)
        dl = self.loglam - avg(self.loglam)
        slope = avg(dl * delta) / avg(dl * dl)
        return w * (delta - avg(delta) - dl * slope)


def model_xi(np_rp=50, np_rt=50, rpmax=200.0, rtmax=200.0, amp=1e-3, scale=50.0):
    """m_A = amp * exp(-r_A / scale) at bin centres r_A = |(rp_c, rt_c)|."""
    rp = (np.arange(np_rp) + 0.5) * rpmax / np_rp
    rt = (np.arange(np_rt) + 0.5) * rtmax / np_rt
    r = np.sqrt(rp[:, None] ** 2 + rt[None, :] ** 2).ravel()
    return amp * np.exp(-r / scale)


def chi2(W, DW, m, sigma):
    import torch
    ok = W > 0
    xi = torch.where(ok, DW / torch.where(ok, W, torch.ones_like(W)), torch.zeros_like(W))
    return torch.sum(torch.where(ok, (xi - m) ** 2 / sigma ** 2, torch.zeros_like(W)))
