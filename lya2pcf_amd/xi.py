"""xi(r_par, r_perp) pair histograms on the GPU.

``XiEngine`` owns one problem: the packed forests in HBM, the forest-pair
list, the work decomposition (items and tiles), the per-slot outputs and the
work-queue counter.  One call of :meth:`run` (or one replay of
:meth:`graph`) computes, for every output slot (HEALPix pixel of forest 1),
  cnt[slot, bp, bt]  number of pixel pairs in the bin,
  w[slot, bp, bt]    sum of w1 w2,
  dw[slot, bp, bt]   sum of dw1 dw2,
with the binning of reference cuda_kernels.cpp:90-162 and the pair list of
forest_class.py:78-98.  Device buffers are torch tensors (plumbing only).
"""
import ctypes as C
import time
import warnings

import numpy as np

from . import _lib
from ._lib import XiProblem, XiConfig, LYA_MODE_FP64, LYA_MODE_MIXED, LYA_MODE_MIXED_F32, LYA_MAX_BINS, check, lib, ptr

MODES = {"fp64": LYA_MODE_FP64, "mixed": LYA_MODE_MIXED, "mixed32": LYA_MODE_MIXED_F32}
# dc of the padding around the pixel arrays: any pair with it falls outside the bins
DC_PAD = 1e300


def build_pairs(forests, angmax):
    """CSR neighbour list: (nb_off int64[nf+1], nb_idx int32[n_pairs])."""
    L = lib()
    nf = forests.nf
    nb_off = np.zeros(nf + 1, dtype=np.int64)
    npairs = C.c_int64(0)
    ux, uy, uz = (np.ascontiguousarray(a, dtype=np.float64) for a in (forests.ux, forests.uy, forests.uz))
    ra = np.ascontiguousarray(forests.ra, dtype=np.float64)
    name = np.ascontiguousarray(forests.name, dtype=np.int64)
    cosm = float(np.cos(angmax))
    check(L.lya_build_pairs(nf, ptr(ux), ptr(uy), ptr(uz), ptr(ra), ptr(name), cosm, ptr(nb_off), None, 0,
                            C.byref(npairs)))
    nb_idx = np.zeros(max(1, npairs.value), dtype=np.int32)
    check(L.lya_build_pairs(nf, ptr(ux), ptr(uy), ptr(uz), ptr(ra), ptr(name), cosm, ptr(nb_off), ptr(nb_idx),
                            len(nb_idx), C.byref(npairs)))
    return nb_off, nb_idx[:npairs.value]


def transpose_pairs(nb_off, nb_idx, nf):
    """Pair list seen from forest 2: CSR over f2 of the f1 that list it."""
    counts = np.bincount(nb_idx, minlength=nf).astype(np.int64)
    off = np.zeros(nf + 1, dtype=np.int64)
    np.cumsum(counts, out=off[1:])
    f1 = np.repeat(np.arange(nf, dtype=np.int32), np.diff(nb_off))
    order = np.argsort(nb_idx, kind="stable")
    return off, f1[order].astype(np.int32)


def build_work(forests, nb_off, nb_idx, f1_list, item_cost, tile_cost):
    L = lib()
    f1_list = np.ascontiguousarray(f1_list, dtype=np.int32)
    f1_slot = np.ascontiguousarray(forests.slot[f1_list], dtype=np.int32)
    ni, nt = C.c_int32(0), C.c_int32(0)
    tot = C.c_double(0.0)
    nbi = nb_idx if len(nb_idx) else np.zeros(1, dtype=np.int32)
    args = (len(f1_list), ptr(f1_list), ptr(f1_slot), ptr(forests.pix_off), ptr(nb_off), ptr(nbi),
            int(item_cost), int(tile_cost))
    check(L.lya_build_work(*args, None, None, C.byref(ni), C.byref(nt), C.byref(tot)))
    items = np.zeros((max(1, ni.value), 4), dtype=np.int32)
    tiles = np.zeros((max(1, nt.value), 2), dtype=np.int32)
    check(L.lya_build_work(*args, ptr(items), ptr(tiles), C.byref(ni), C.byref(nt), C.byref(tot)))
    return items[:ni.value], tiles[:nt.value], tot.value


def build_tiles(items, costs, tile_cost):
    """Tiles {item_begin, item_end} (host int32 [n_tiles, 2]) from items and their costs."""
    L = lib()
    items = np.ascontiguousarray(items, dtype=np.int32).reshape(-1, 4)
    costs = np.ascontiguousarray(costs, dtype=np.int64)
    if len(costs) != len(items):
        raise ValueError("build_tiles: %d items but %d costs" % (len(items), len(costs)))
    nt = C.c_int32(0)
    check(L.lya_build_tiles(len(items), ptr(items), ptr(costs), int(tile_cost), None, C.byref(nt)))
    tiles = np.zeros((max(1, nt.value), 2), dtype=np.int32)
    check(L.lya_build_tiles(len(items), ptr(items), ptr(costs), int(tile_cost), ptr(tiles), C.byref(nt)))
    return tiles[:nt.value]


XCDS = 8  # MI355X: 8 accelerator dies (XCDs), each with its own L2


def group_tiles(items, tiles, item_costs, n_groups):
    """Tiles regrouped for XCD-local work: tiles ordered by slot (neighbouring HEALPix
    pixels, hence neighbouring forests), cut into ``n_groups`` contiguous runs of about
    equal cost, each run ordered by decreasing cost.  Returns (tiles, offsets) with
    offsets of length n_groups + 1; the kernels' workgroup b serves group b % n_groups
    first, and consecutive workgroups sit on different XCDs."""
    tiles = np.asarray(tiles, dtype=np.int32).reshape(-1, 2)
    nt = len(tiles)
    if n_groups <= 1 or nt == 0:
        return tiles, np.array([0, nt], dtype=np.int32)
    cum = np.concatenate([[0], np.cumsum(np.asarray(item_costs, dtype=np.float64))])
    tcost = cum[tiles[:, 1]] - cum[tiles[:, 0]]
    slot = np.asarray(items, dtype=np.int32).reshape(-1, 4)[tiles[:, 0], 3]
    order = np.lexsort((np.arange(nt), slot))
    c = np.cumsum(tcost[order])
    total = c[-1] if nt else 0.0
    cuts = np.searchsorted(c, total * np.arange(1, n_groups) / n_groups, side="left")
    offs = np.concatenate([[0], np.minimum(cuts + 1, nt), [nt]]).astype(np.int64)
    offs = np.maximum.accumulate(offs)
    out = []
    for g in range(n_groups):
        sel = order[offs[g]:offs[g + 1]]
        out.append(sel[np.argsort(-tcost[sel], kind="stable")])
    return tiles[np.concatenate(out)], offs.astype(np.int32)


def _device_vec(t, dtype, n, what):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == dtype and t.is_contiguous() and t.numel() == n):
        raise ValueError("%s must be a contiguous %s device tensor of %d elements" % (what, dtype, n))
    return t


def build_pairs_device(ux, uy, uz, ra, name, angmax, stream=None):
    """The neighbour list of :func:`build_pairs`, built on the GPU from device tensors
    (ux, uy, uz, ra float64; name int64).  Returns (nb_off int64 [nf+1], nb_idx int32
    [n_pairs]) device tensors and n_pairs; the arrays equal the host builder's."""
    import torch
    L = lib()
    n = ux.numel()
    for t, what in ((ux, "ux"), (uy, "uy"), (uz, "uz"), (ra, "ra")):
        _device_vec(t, torch.float64, n, what)
    _device_vec(name, torch.int64, n, "name")
    cosm = float(np.cos(angmax))
    nbytes = C.c_int64(0)
    check(L.lya_pairs_device_scratch_bytes(n, cosm, C.byref(nbytes)))
    dev = ux.device
    scratch = torch.empty(max(1, nbytes.value), dtype=torch.uint8, device=dev)
    nb_off = torch.empty(n + 1, dtype=torch.int64, device=dev)
    npairs = C.c_int64(0)
    sp = _lib.stream_ptr(stream)
    args = (n, ptr(ux), ptr(uy), ptr(uz), ptr(ra), ptr(name), cosm, ptr(nb_off))
    check(L.lya_build_pairs_device(*args, None, 0, ptr(scratch), nbytes.value, C.byref(npairs), sp))
    nb_idx = torch.empty(max(1, npairs.value), dtype=torch.int32, device=dev)
    check(L.lya_build_pairs_device(*args, ptr(nb_idx), nb_idx.numel(), ptr(scratch), nbytes.value, C.byref(npairs),
                                   sp))
    return nb_off, nb_idx[:npairs.value], npairs.value


def build_items_device(f1_list, f1_slot, pix_off, nb_off, nb_idx, item_cost, stream=None):
    """The item cut of :func:`build_work` on the GPU.  Device tensors in (f1_list,
    f1_slot int32; pix_off, nb_off int64; nb_idx int32); returns (items int32 [n, 4],
    costs int64 [n]) device tensors and the total pixel-pair count."""
    import torch
    L = lib()
    n = f1_list.numel()
    _device_vec(f1_list, torch.int32, n, "f1_list")
    _device_vec(f1_slot, torch.int32, n, "f1_slot")
    nbytes = C.c_int64(0)
    check(L.lya_items_device_scratch_bytes(n, C.byref(nbytes)))
    dev = f1_list.device
    scratch = torch.empty(max(1, nbytes.value), dtype=torch.uint8, device=dev)
    ni = C.c_int64(0)
    tot = C.c_double(0.0)
    sp = _lib.stream_ptr(stream)
    nbi = nb_idx if nb_idx.numel() else torch.zeros(1, dtype=torch.int32, device=dev)
    args = (n, ptr(f1_list), ptr(f1_slot), ptr(pix_off), ptr(nb_off), ptr(nbi), int(item_cost))
    check(L.lya_build_items_device(*args, None, None, 0, ptr(scratch), nbytes.value, C.byref(ni), C.byref(tot), sp))
    items = torch.empty((max(1, ni.value), 4), dtype=torch.int32, device=dev)
    costs = torch.empty(max(1, ni.value), dtype=torch.int64, device=dev)
    check(L.lya_build_items_device(*args, ptr(items), ptr(costs), ni.value, ptr(scratch), nbytes.value, C.byref(ni),
                                   C.byref(tot), sp))
    return items[:ni.value], costs[:ni.value], tot.value


def default_config(mode="mixed", **kw):
    cfg = XiConfig()
    cfg.mode = MODES[mode] if isinstance(mode, str) else int(mode)
    cfg.px_per_lane = kw.get("px_per_lane", 1)
    cfg.waves = kw.get("waves", 8)
    cfg.replicas = kw.get("replicas", 1)
    cfg.row_pad = kw.get("row_pad", 1)
    cfg.groups = kw.get("groups", 0)
    cfg.no_counts = 0 if kw.get("counts", True) else 1
    cfg.lane_map = kw.get("lane_map", 1)
    cfg.walk = kw.get("walk", 0)
    cfg.merge = kw.get("merge", 0)
    cfg.tail_max = kw.get("tail_max", 0)
    return cfg


class XiEngine:
    """One xi problem resident on the GPU.

    forests   ForestSet (host); uploaded once.
    angmax    neighbour cut (radians).
    slots     slot indices whose forests act as forest 1 (default: all).
    pair_builder  "device" (default): the forest-pair list and the work items are
              built on the GPU and stay resident; "host": OpenMP builders.  Both
              give identical arrays.  A given (pairs=) or transposed list is host.
    """

    def __init__(self, forests, angmax, np_rp=50, np_rt=50, rpmax=200.0, rtmax=200.0, slots=None,
                 item_cost=2_000_000, tiles_per_group=6, device="cuda", pairs=None, transposed=False,
                 xcd_groups=XCDS,
                 pair_builder="device"):
        import torch
        if pair_builder not in ("host", "device"):
            raise ValueError("pair_builder must be 'host' or 'device'")
        self.torch = torch
        self.forests = forests
        self.angmax = float(angmax)
        self.np_rp, self.np_rt = int(np_rp), int(np_rt)
        self.rpmax, self.rtmax = float(rpmax), float(rtmax)
        self.device = torch.device(device)
        # a given or transposed list is a host list
        self.pair_builder = pair_builder if (pairs is None and not transposed) else "host"
        self.n_slots = len(forests.slots)
        if slots is None:
            f1_list = np.arange(forests.nf, dtype=np.int32)
        else:
            want = np.zeros(self.n_slots, dtype=bool)
            want[np.asarray(list(slots), dtype=np.int64)] = True
            f1_list = np.nonzero(want[forests.slot])[0].astype(np.int32)
        self.f1_list = f1_list
        self.item_cost = int(item_cost)
        self.tiles_per_group = tiles_per_group
        self.xcd_groups = int(xcd_groups)
        self._work = {}
        self._items_host = None
        self.setup_seconds = {}
        clock = time.perf_counter
        if self.pair_builder == "device":
            try:
                self._upload()
                self._build_on_device(forests, angmax, f1_list, clock)
            except _lib.LyaError as e:
                if e.code != _lib.LYA_ERR_UNSUPPORTED:
                    raise
                # a forest with more neighbours than the device row sort holds
                warnings.warn("device pair list unavailable (%s); using the host builder" % e)
                self.pair_builder = "host"
        if self.pair_builder == "host":
            t0 = clock()
            nb_off, nb_idx = pairs if pairs is not None else build_pairs(forests, angmax)
            if transposed:
                nb_off, nb_idx = transpose_pairs(nb_off, nb_idx, forests.nf)
            self._nb_host = (nb_off, nb_idx)
            self.n_pairs = len(nb_idx)
            t1 = clock()
            # total pixel pairs over the list (the metric's numerator)
            _, _, self.pixel_pairs = build_work(forests, nb_off, nb_idx, f1_list, self.item_cost, 1 << 62)
            self.setup_seconds = {"pairs": t1 - t0, "items": clock() - t1}
            self._upload()
        self._graphs = []

    def _build_on_device(self, forests, angmax, f1_list, clock):
        torch = self.torch
        self._nb_host = None
        d = self.d
        torch.cuda.synchronize()
        t0 = clock()
        self.d_nb_off, nb_idx_t, self.n_pairs = build_pairs_device(d["ux"], d["uy"], d["uz"], d["ra"], d["name"],
                                                                   angmax)
        d["nb_idx"] = nb_idx_t if self.n_pairs else torch.zeros(1, dtype=torch.int32, device=self.device)
        torch.cuda.synchronize()
        t1 = clock()
        f1_slot = np.ascontiguousarray(forests.slot[f1_list], dtype=np.int32)
        items, costs, self.pixel_pairs = build_items_device(self._t(f1_list), self._t(f1_slot), d["pix_off"],
                                                            self.d_nb_off, nb_idx_t, self.item_cost)
        self._items_host = (items.cpu().numpy(), costs.cpu().numpy())
        self.setup_seconds = {"pairs": t1 - t0, "items": clock() - t1}

    def _host_pairs(self):
        if self._nb_host is None:
            self._nb_host = (self.d_nb_off.cpu().numpy(), self.d["nb_idx"][:self.n_pairs].cpu().numpy())
        return self._nb_host

    @property
    def nb_off(self):
        """CSR offsets of the forest-pair list (host copy)."""
        return self._host_pairs()[0]

    @property
    def nb_idx(self):
        """Neighbour forest indices of the forest-pair list (host copy)."""
        return self._host_pairs()[1]

    # ------------------------------------------------------------------
    def _t(self, a, dtype=None):
        t = self.torch.from_numpy(np.ascontiguousarray(a))
        if dtype is not None:
            t = t.to(dtype)
        return t.to(self.device)

    def _padded(self, a, fill=0.0):
        """Device copy of a pixel array with ``self.pad`` entries of ``fill`` on both
        sides (the diagonal walk reads up to one forest length outside the forests);
        returns the view of the data, whose first element is pixel 0."""
        buf = self.torch.full((len(a) + 2 * self.pad,), fill, dtype=self.torch.float64, device=self.device)
        buf[self.pad:self.pad + len(a)] = self.torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64))
        return buf[self.pad:self.pad + len(a)]

    def _upload(self):
        torch = self.torch
        f = self.forests
        self.max_len = int(np.max(f.n)) if f.nf else 0
        self.pad = max(1, self.max_len)
        self.d = {
            "pix_off": self._t(f.pix_off), "ux": self._t(f.ux), "uy": self._t(f.uy), "uz": self._t(f.uz),
            "base": self._t(f.base), "sorted": self._t(f.sorted),
            "dc": self._padded(f.dc, DC_PAD), "we": self._padded(f.we), "dw": self._padded(f.dw),
            "off32": self._t(f.off32), "we32": self._t(f.we32), "dw32": self._t(f.dw32),
        }
        if self.pair_builder == "host":
            self.d["nb_idx"] = self._t(self.nb_idx if len(self.nb_idx) else np.zeros(1, np.int32))
        else:
            self.d["ra"] = self._t(np.ascontiguousarray(f.ra, dtype=np.float64))
            self.d["name"] = self._t(np.ascontiguousarray(f.name, dtype=np.int64))
        nbins = self.np_rp * self.np_rt
        self.out_cnt = torch.zeros((self.n_slots, nbins), dtype=torch.int64, device=self.device)
        self.out_w = torch.zeros((self.n_slots, nbins), dtype=torch.float64, device=self.device)
        self.out_dw = torch.zeros((self.n_slots, nbins), dtype=torch.float64, device=self.device)
        self.queue = torch.zeros(16, dtype=torch.int32, device=self.device)
        self.tau = self._tau()

    def _tau(self):
        p = XiProblem()
        p.np_rp, p.np_rt = self.np_rp, self.np_rt
        t = C.c_double(0.0)
        max_dc = float(np.max(self.forests.dc)) if len(self.forests.dc) else 0.0
        check(lib().lya_xi_tau(C.byref(p), self.forests.max_abs_off, max_dc, self.np_rp / self.rpmax,
                               self.np_rt / self.rtmax, C.byref(t)))
        # round up to the next float32 so the kernel band is never narrower than the bound
        tf = np.float32(t.value)
        if float(tf) < t.value:
            tf = np.nextafter(tf, np.float32(1.0))
        return float(tf)

    def work(self, cfg):
        groups = C.c_int32(0)
        check(lib().lya_xi_grid(C.byref(cfg), C.byref(groups)))
        # tile kernels: XCD tile groups (fewer L2 misses, same time); the walk, with a
        # quarter of the workgroups, keeps one queue in global cost order (groups cost it
        # 3.5 % at 300k forests)
        n_groups = 1 if getattr(cfg, "walk", 0) else self.xcd_groups
        key = (groups.value, n_groups)
        if key not in self._work:
            tile_cost = max(self.item_cost, int(self.pixel_pairs / max(1, groups.value * self.tiles_per_group)))
            if self._items_host is not None:
                items, costs = self._items_host
                tiles = build_tiles(items, costs, tile_cost)
            else:
                items, tiles, _ = build_work(self.forests, self.nb_off, self.nb_idx, self.f1_list, self.item_cost,
                                             tile_cost)
                costs = self._item_costs(items)
            tiles, offs = group_tiles(items, tiles, costs, n_groups)
            self._work[key] = (self._t(items if len(items) else np.zeros((1, 4), np.int32)),
                               self._t(tiles if len(tiles) else np.zeros((1, 2), np.int32)),
                               len(items), len(tiles), self._t(offs), len(offs) - 1)
        return self._work[key]

    def _item_costs(self, items):
        """n1 * sum of n2 over each item's neighbours (host arrays)."""
        items = np.asarray(items, dtype=np.int64).reshape(-1, 4)
        if len(items) == 0:
            return np.zeros(0, dtype=np.int64)
        nlen = np.diff(np.asarray(self.forests.pix_off, dtype=np.int64))
        nb = np.asarray(self.nb_idx, dtype=np.int64)
        cum = np.concatenate([[0], np.cumsum(nlen[nb])])
        return nlen[items[:, 0]] * (cum[items[:, 2]] - cum[items[:, 1]])

    def problem(self, cfg):
        items, tiles, ni, nt, offs, ng = self.work(cfg)
        d = self.d
        p = XiProblem()
        p.n_forest, p.n_slots = self.forests.nf, self.n_slots
        for k in ("pix_off", "ux", "uy", "uz", "base", "sorted", "dc", "we", "dw", "off32", "we32", "dw32", "nb_idx"):
            setattr(p, k, ptr(d[k]))
        p.items, p.tiles, p.n_items, p.n_tiles = ptr(items), ptr(tiles), ni, nt
        p.np_rp, p.np_rt, p.rpmax, p.rtmax = self.np_rp, self.np_rt, self.rpmax, self.rtmax
        p.tau = self.tau
        p.lane_map = getattr(cfg, "lane_map", 1)
        p.max_len = self.max_len
        p.pad = self.pad
        p.out_cnt, p.out_w, p.out_dw, p.queue = ptr(self.out_cnt), ptr(self.out_w), ptr(self.out_dw), ptr(self.queue)
        p.tile_groups, p.n_tile_groups = ptr(offs), ng
        self._keep = (items, tiles, offs)
        return p

    def run(self, cfg=None, stream=None):
        cfg = cfg or default_config()
        p = self.problem(cfg)
        check(lib().lya_xi_run(C.byref(p), C.byref(cfg), _lib.stream_ptr(stream)))

    def work_stats(self, cfg=None):
        """Work of the tile kernels for cfg (one band): {'evaluated_pairs': lane slots
        stepped (pixel pairs evaluated, idle lanes included), 'evaluated_valid': the same
        over lanes holding a pixel, 'tasks', 'tasks_nonempty'}."""
        cfg = cfg or default_config()
        p = self.problem(cfg)
        out = self.torch.zeros(4, dtype=self.torch.int64, device=self.device)
        check(lib().lya_xi_work_stats(C.byref(p), C.byref(cfg), ptr(out), _lib.stream_ptr(None)))
        v = out.cpu().numpy().view(np.uint64)
        return {"evaluated_pairs": float(v[0]), "evaluated_valid": float(v[1]), "tasks": int(v[2]),
                "tasks_nonempty": int(v[3])}

    def graph(self, cfg=None):
        cfg = cfg or default_config()
        return XiGraph(self, cfg)

    def results(self):
        """(cnt uint64, w f64, dw f64) host arrays of shape [n_slots, np_rp, np_rt]."""
        self.torch.cuda.synchronize()
        sh = (self.n_slots, self.np_rp, self.np_rt)
        cnt = self.out_cnt.cpu().numpy().view(np.uint64).reshape(sh)
        return cnt, self.out_w.cpu().numpy().reshape(sh), self.out_dw.cpu().numpy().reshape(sh)


class XiGraph:
    """The zeroing + persistent-kernel sequence captured once as a hipGraph."""

    def __init__(self, eng, cfg):
        self.eng = eng
        self.cfg = cfg
        self.p = eng.problem(cfg)
        h = C.c_void_p(0)
        check(lib().lya_xi_graph_create(C.byref(self.p), C.byref(cfg), None, C.byref(h)))
        self.h = h

    def launch(self, stream=None):
        check(lib().lya_graph_launch(self.h, _lib.stream_ptr(stream)))

    def close(self):
        if self.h:
            lib().lya_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
