"""Minimal FITS reader and writer (numpy only), for the delta-file ingest.

Covers what delta files use (FITS standard 4.0): a primary HDU and IMAGE and
BINTABLE extensions, big-endian data, 2880-byte blocks of 80-character header
cards, gzip-compressed files (``*.fits.gz``, detected by their magic bytes).
Table columns of type L, B, I, J, K, E, D and A with a repeat count are
supported; BSCALE/BZERO and TSCALn/TZEROn are applied when present.  Variable
length arrays, bit and complex columns, CONTINUE cards and random groups are
not (they raise ``ValueError``).  The reference reads these files with fitsio
(delta_reader.py:33-49, delta_reader_eboss.py:33-44).
"""
import gzip
import io

import numpy as np

BLOCK = 2880
CARD = 80

_BITPIX = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}
_TFORM = {"L": ("i1", 1), "B": (">u1", 1), "I": (">i2", 2), "J": (">i4", 4), "K": (">i8", 8), "E": (">f4", 4),
          "D": (">f8", 8), "A": ("S", 1)}


class HDU:
    """One header-data unit: ``header`` (dict, card order kept), ``name``
    (EXTNAME, or 'PRIMARY'), ``kind`` ('image' or 'table') and ``data``
    (ndarray for images, numpy structured array for tables, None if empty)."""

    def __init__(self, header, data, kind):
        self.header = header
        self.data = data
        self.kind = kind
        self.name = str(header.get("EXTNAME", "PRIMARY" if "SIMPLE" in header else "")).strip().upper()

    def __getitem__(self, key):
        return self.data[key]

    def __repr__(self):
        shape = None if self.data is None else (self.data.shape if self.kind == "image" else len(self.data))
        return "HDU(%s, %s, %s)" % (self.name, self.kind, shape)


# ---------------------------------------------------------------------- read
def _value(text):
    t = text.strip()
    if not t:
        return None
    if t[0] == "'":
        out, k = [], 1
        while k < len(t):
            if t[k] == "'":
                if k + 1 < len(t) and t[k + 1] == "'":
                    out.append("'")
                    k += 2
                    continue
                break
            out.append(t[k])
            k += 1
        return "".join(out).rstrip()
    v = t.split("/", 1)[0].strip()
    if v == "T":
        return True
    if v == "F":
        return False
    try:
        return int(v)
    except ValueError:
        pass
    try:
        return float(v.replace("D", "E").replace("d", "e"))
    except ValueError:
        return v


def _parse_header(buf, pos):
    header = {}
    while True:
        if pos + BLOCK > len(buf):
            raise ValueError("truncated FITS header")
        block = buf[pos:pos + BLOCK].decode("ascii")
        pos += BLOCK
        for c in range(0, BLOCK, CARD):
            card = block[c:c + CARD]
            key = card[:8].strip()
            if key == "END":
                return header, pos
            if key == "CONTINUE":
                raise ValueError("CONTINUE cards are not supported")
            if card[8:10] == "= " and key:
                header[key] = _value(card[10:])


def _table_dtype(h):
    fields = []
    for i in range(1, int(h["TFIELDS"]) + 1):
        form = str(h["TFORM%d" % i]).strip()
        k = 0
        while k < len(form) and form[k].isdigit():
            k += 1
        rep = int(form[:k]) if k else 1
        code = form[k:k + 1]
        if code not in _TFORM:
            raise ValueError("unsupported table column type %r" % form)
        name = str(h.get("TTYPE%d" % i, "COL%d" % i)).strip()
        base = _TFORM[code][0]
        if code == "A":
            fields.append((name, "S%d" % rep))
        else:
            fields.append((name, base) if rep == 1 else (name, base, (rep,)))
    return np.dtype(fields)


def _read_data(h, buf, pos, kind):
    bitpix = int(h["BITPIX"])
    naxis = [int(h["NAXIS%d" % i]) for i in range(1, int(h.get("NAXIS", 0)) + 1)]
    if int(h.get("GCOUNT", 1)) != 1:
        raise ValueError("random groups are not supported")
    nbytes = abs(bitpix) // 8 * (int(h.get("PCOUNT", 0)) + (int(np.prod(naxis)) if naxis else 0))
    data = None
    if kind == "image" and naxis:
        data = np.frombuffer(buf, dtype=_BITPIX[bitpix], count=int(np.prod(naxis)), offset=pos)
        data = data.reshape(naxis[::-1])
        if "BSCALE" in h or "BZERO" in h:
            data = data * float(h.get("BSCALE", 1.0)) + float(h.get("BZERO", 0.0))
        else:
            data = data.astype(data.dtype.newbyteorder("="))
    elif kind == "table":
        dt = _table_dtype(h)
        if dt.itemsize != naxis[0]:
            raise ValueError("table row of %d bytes, columns give %d" % (naxis[0], dt.itemsize))
        raw = np.frombuffer(buf, dtype=dt, count=naxis[1], offset=pos)
        out = {}
        for i, name in enumerate(dt.names, start=1):
            col = raw[name]
            if col.dtype.kind == "S":
                out[name] = np.char.rstrip(col)
                continue
            col = col.astype(col.dtype.newbyteorder("="))
            if _TFORM[str(h["TFORM%d" % i]).strip().lstrip("0123456789")[0]][0] == "i1":
                col = col == ord("T")
            if "TSCAL%d" % i in h or "TZERO%d" % i in h:
                col = col * float(h.get("TSCAL%d" % i, 1.0)) + float(h.get("TZERO%d" % i, 0.0))
            out[name] = col
        if dt.names:
            fields = [(n, out[n].dtype, out[n].shape[1:]) for n in dt.names]
            data = np.zeros(naxis[1], dtype=fields).view(np.recarray)
            for n in dt.names:
                data[n] = out[n]
    pos += (nbytes + BLOCK - 1) // BLOCK * BLOCK
    return data, pos


def read_bytes(buf):
    """All HDUs of a FITS byte string."""
    if buf[:2] == b"\x1f\x8b":
        buf = gzip.decompress(buf)
    hdus, pos = [], 0
    while pos + BLOCK <= len(buf) and buf[pos:pos + BLOCK].strip(b" \x00"):
        h, pos = _parse_header(buf, pos)
        if not hdus:
            if h.get("SIMPLE") is not True:
                raise ValueError("not a FITS file (no SIMPLE = T)")
            kind = "image"
        else:
            x = str(h.get("XTENSION", "")).strip().upper()
            if x not in ("IMAGE", "BINTABLE"):
                raise ValueError("unsupported extension %r" % x)
            kind = "image" if x == "IMAGE" else "table"
        data, pos = _read_data(h, buf, pos, kind)
        hdus.append(HDU(h, data, kind))
    return hdus


def read(path):
    """All HDUs of a FITS (or gzip-compressed FITS) file."""
    with open(path, "rb") as f:
        return read_bytes(f.read())


def by_name(hdus, name):
    for h in hdus:
        if h.name == name.upper():
            return h
    raise KeyError("no HDU named %s" % name)


# --------------------------------------------------------------------- write
def _card(key, value=None, comment=None):
    if value is None:
        s = "%-8s" % key
    else:
        if isinstance(value, bool):
            v = "%20s" % ("T" if value else "F")
        elif isinstance(value, (int, np.integer)):
            v = "%20d" % int(value)
        elif isinstance(value, (float, np.floating)):
            v = "%20s" % repr(float(value)).upper()
        else:
            v = "'%-8s'" % str(value).replace("'", "''")
        s = "%-8s= %s" % (key, v)
        if comment:
            s += " / " + comment
    if len(s) > CARD:
        raise ValueError("card too long: %s" % s)
    return s.ljust(CARD)


def _header_bytes(cards):
    text = "".join(cards) + "END".ljust(CARD)
    text += " " * ((-len(text)) % BLOCK)
    return text.encode("ascii")


def _pad(b, fill=b"\x00"):
    return b + fill * ((-len(b)) % BLOCK)


def _extra(cards, header):
    for k, v in (header or {}).items():
        cards.append(_card(k.upper(), v))


def primary_bytes(header=None):
    cards = [_card("SIMPLE", True), _card("BITPIX", 8), _card("NAXIS", 0), _card("EXTEND", True)]
    _extra(cards, header)
    return _header_bytes(cards)


def image_bytes(data, name=None, header=None):
    data = np.asarray(data)
    code = {np.dtype(v).newbyteorder("=").str: k for k, v in _BITPIX.items()}.get(data.dtype.newbyteorder("=").str)
    if code is None:
        raise ValueError("unsupported image dtype %s" % data.dtype)
    cards = [_card("XTENSION", "IMAGE"), _card("BITPIX", code), _card("NAXIS", data.ndim)]
    cards += [_card("NAXIS%d" % (i + 1), n) for i, n in enumerate(data.shape[::-1])]
    cards += [_card("PCOUNT", 0), _card("GCOUNT", 1)]
    if name:
        cards.append(_card("EXTNAME", name))
    _extra(cards, header)
    return _header_bytes(cards) + _pad(np.ascontiguousarray(data, dtype=_BITPIX[code]).tobytes())


def table_bytes(columns, name=None, header=None):
    """BINTABLE from {name: 1-D or 2-D array}, in the given column order."""
    codes = {"f8": "D", "f4": "E", "i8": "K", "i4": "J", "i2": "I", "u1": "B", "b1": "L"}
    cols, fields, forms = [], [], []
    nrow = None
    for cname, a in columns.items():
        a = np.asarray(a)
        nrow = len(a) if nrow is None else nrow
        if len(a) != nrow:
            raise ValueError("columns of different lengths")
        rep = int(np.prod(a.shape[1:])) if a.ndim > 1 else 1
        if a.dtype.kind in "SU":
            a = a.astype("S")
            w = max(1, a.dtype.itemsize)
            fields.append((cname, "S%d" % w))
            forms.append("%dA" % w)
            cols.append(a)
            continue
        key = a.dtype.newbyteorder("=").str[1:] if a.dtype.kind != "b" else "b1"
        if key not in codes:
            raise ValueError("unsupported column dtype %s" % a.dtype)
        code = codes[key]
        if code == "L":
            a = np.where(a, ord("T"), ord("F")).astype("i1")
            base = "i1"
        else:
            base = ">" + key
        fields.append((cname, base) if rep == 1 else (cname, base, a.shape[1:]))
        forms.append(("%d" % rep if rep != 1 else "") + code)
        cols.append(a)
    nrow = nrow or 0
    rec = np.zeros(nrow, dtype=np.dtype(fields))
    for (cname, *_), a in zip(fields, cols):
        rec[cname] = a
    cards = [_card("XTENSION", "BINTABLE"), _card("BITPIX", 8), _card("NAXIS", 2),
             _card("NAXIS1", rec.dtype.itemsize), _card("NAXIS2", nrow), _card("PCOUNT", 0), _card("GCOUNT", 1),
             _card("TFIELDS", len(fields))]
    for i, ((cname, *_), form) in enumerate(zip(fields, forms), start=1):
        cards += [_card("TTYPE%d" % i, cname), _card("TFORM%d" % i, form)]
    if name:
        cards.append(_card("EXTNAME", name))
    _extra(cards, header)
    return _header_bytes(cards) + _pad(rec.tobytes())


def write(path, parts, compress=None):
    """Write a primary HDU followed by the extension byte strings ``parts``
    (from image_bytes / table_bytes); gzip when ``compress`` or path ends .gz."""
    buf = io.BytesIO()
    buf.write(primary_bytes())
    for p in parts:
        buf.write(p)
    b = buf.getvalue()
    if compress or (compress is None and str(path).endswith(".gz")):
        b = gzip.compress(b, mtime=0)
    with open(path, "wb") as f:
        f.write(b)
