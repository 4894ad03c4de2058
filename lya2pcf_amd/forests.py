"""Packed forest set: the ragged structure-of-arrays the kernels read.

Layout in HBM (all arrays flat, forests contiguous in packing order):
  per forest  pix_off int64[nf+1], ux/uy/uz f64, base f64 (integer-valued
              origin of the FP32 offsets), sorted u8 (dc non-decreasing),
              plus ra, dec, name, slot on the host;
  per pixel   dc f64, we f64, dw f64 (FP64 path) and off32 = dc - base, we32,
              dw32 as f32 (mixed path; 12 bytes per pixel).
Packing order follows the data dict: pixels (slots) in the given order, then
each pixel's forest list (reference: correlation_procedures_pycuda.py:73-87,
which pads every forest to max_lenght instead).
"""
import numpy as np


class ForestSet:
    FOREST = ("ux", "uy", "uz", "ra", "dec", "name", "slot", "omega", "odl2")
    PIXEL = ("dc", "we", "dw", "dl")

    def __init__(self, pix_off, slots, **a):
        self.pix_off = np.ascontiguousarray(pix_off, dtype=np.int64)
        self.slots = list(slots)            # slot index -> HEALPix pixel key
        for k in self.FOREST + self.PIXEL:
            v = a.get(k)
            setattr(self, k, None if v is None else np.ascontiguousarray(v))
        self._derive()

    # ------------------------------------------------------------------
    @classmethod
    def from_survey(cls, s, float32_inputs=False):
        keys, slot = np.unique(s.hpix, return_inverse=True)
        if np.any(np.diff(slot) < 0):
            raise ValueError("survey forests must be grouped by HEALPix pixel")
        we, dw = s.we, s.dw
        if float32_inputs:
            we = we.astype(np.float32).astype(np.float64)
            dw = dw.astype(np.float32).astype(np.float64)
        return cls(s.pix_off, [int(k) for k in keys], ux=s.ux, uy=s.uy, uz=s.uz, ra=s.ra, dec=s.dec,
                   name=s.name.astype(np.int64), slot=slot.astype(np.int32), omega=s.omega, odl2=s.omega_dl2,
                   dc=s.dc, we=we, dw=dw, dl=s.delta_lambda)

    @classmethod
    def from_data(cls, data, pixel_list=None):
        """Pack ``{pixel: [quasar]}`` (the unpickled data#.npy)."""
        pixels = list(data.keys()) if not pixel_list else list(pixel_list)
        forests, slot = [], []
        for k, pix in enumerate(pixels):
            for f in data[pix]:
                forests.append(f)
                slot.append(k)
        n = np.array([len(f.dc) for f in forests], dtype=np.int64)
        off = np.zeros(len(forests) + 1, dtype=np.int64)
        np.cumsum(n, out=off[1:])
        cat = lambda attr: (np.concatenate([np.asarray(getattr(f, attr), dtype=np.float64) for f in forests])
                            if forests else np.zeros(0))
        scal = lambda attr, dt=np.float64: np.array([getattr(f, attr) for f in forests], dtype=dt)
        names = np.array([hash(f.name) if not isinstance(f.name, (int, np.integer)) else int(f.name)
                          for f in forests], dtype=np.int64)
        dl = [np.asarray(getattr(f, "delta_lambda", np.zeros(len(f.dc))), dtype=np.float64) for f in forests]
        dl = [d if d.shape == (len(f.dc),) else np.zeros(len(f.dc)) for d, f in zip(dl, forests)]
        fs = cls(off, pixels, ux=scal("x"), uy=scal("y"), uz=scal("z"), ra=scal("ra"), dec=scal("dec"),
                 name=names, slot=np.array(slot, dtype=np.int32),
                 omega=scal("omega"), odl2=scal("omega_delta_lambda2"),
                 dc=cat("dc"), we=cat("we"), dw=cat("dw"), dl=np.concatenate(dl) if dl else np.zeros(0))
        fs.objects = forests
        return fs

    # ------------------------------------------------------------------
    def _derive(self):
        nf = self.nf
        n = self.n
        dc = self.dc
        first = self.pix_off[:-1]
        last = np.maximum(self.pix_off[1:] - 1, first)
        if nf and np.any(n <= 0):
            raise ValueError("empty forest in the set")
        if nf:
            d = np.diff(dc)
            bad = np.zeros(len(dc), dtype=bool)
            bad[:-1] = d < 0
            # a decrease inside forest f sits at pixel index p with p+1 in the same forest
            within = np.ones(len(dc), dtype=bool)
            within[self.pix_off[1:-1] - 1] = False
            within[-1] = False
            viol = np.add.reduceat((bad & within).astype(np.int64), first)
            self.sorted = (viol == 0).astype(np.uint8)
            lo = np.minimum.reduceat(dc, first)
            hi = np.maximum.reduceat(dc, first)
            self.base = np.round(0.5 * (lo + hi))
            off = dc - np.repeat(self.base, n)
            self.off32 = off.astype(np.float32)
            self.max_abs_off = float(np.max(np.abs(off)))
        else:
            self.sorted = np.zeros(0, dtype=np.uint8)
            self.base = np.zeros(0)
            self.off32 = np.zeros(0, dtype=np.float32)
            self.max_abs_off = 0.0
        self.we32 = self.we.astype(np.float32)
        self.dw32 = self.dw.astype(np.float32)
        del first, last

    @property
    def nf(self):
        return len(self.pix_off) - 1

    @property
    def n(self):
        return np.diff(self.pix_off)

    @property
    def n_pix(self):
        return int(self.pix_off[-1])

    def forest_slice(self, f):
        return slice(int(self.pix_off[f]), int(self.pix_off[f + 1]))

    def min_dc_first(self):
        """min over forests of dc[0] (reference 2pla.py:64-69)."""
        if self.nf == 0:
            raise NameError("Empty data. Did you write correctly the input directory?")
        return float(np.min(self.dc[self.pix_off[:-1]]))


def angmax_from(forests, rtmax):
    """angmax = 2 asin(rtmax / (2 dmin)), reference 2pla.py:73."""
    return 2 * np.arcsin(0.5 * rtmax / forests.min_dc_first())
