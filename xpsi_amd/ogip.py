"""FITS files in the OGIP layouts the reference reads (response matrices, an
effective area, event lists and spectra), without astropy.

read_fits(path) returns the file's header-data units in order, each a dict
with ``header`` (keyword -> value, strings unquoted, numbers converted,
logicals as bool), ``name`` (EXTNAME, or PRIMARY) and, for binary tables,
``columns`` (TTYPE -> numpy array; variable-length array columns as a list
of per-row arrays).  The format follows the FITS standard 4.0: 2880-byte
records, 80-character header cards, big-endian binary tables whose
variable-length arrays live in the heap after the table (THEAP).

data_checksum(path, index) is the standard's 32-bit ones' complement sum of
an HDU's data unit (the value DATASUM records), hdu_checksum the same over
header and data (0xFFFFFFFF when CHECKSUM is right).
"""
import re

import numpy as np

BLOCK = 2880

_CODES = {"L": ("u1", 1), "B": ("u1", 1), "I": (">i2", 2), "J": (">i4", 4), "K": (">i8", 8),
          "E": (">f4", 4), "D": (">f8", 8), "A": ("S1", 1), "X": ("u1", None),
          "C": (">c8", 8), "M": (">c16", 16)}


def _value(raw):
    raw = raw.strip()
    if raw.startswith("'"):
        m = re.match(r"'((?:[^']|'')*)'", raw)
        return m.group(1).replace("''", "'").rstrip() if m else raw.strip("'")
    v = raw.split("/", 1)[0].strip()
    if v == "T":
        return True
    if v == "F":
        return False
    for cast in (int, float):
        try:
            return cast(v.replace("D", "E")) if cast is float else cast(v)
        except ValueError:
            pass
    return v


def _read_header(buf, pos):
    header, start = {}, pos
    while True:
        if pos + BLOCK > len(buf):
            raise ValueError("truncated FITS header")
        block = buf[pos:pos + BLOCK]
        pos += BLOCK
        done = False
        for k in range(36):
            card = block[80 * k:80 * (k + 1)].decode("ascii", "replace")
            key = card[:8].strip()
            if key == "END":
                done = True
                break
            if card[8:10] == "= " and key:
                header[key] = _value(card[10:])
        if done:
            return header, start, pos


def _parse_tform(tform):
    m = re.match(r"\s*(\d*)([PQ]?)([LXBIJKAEDCM])(.*)", tform)
    if not m:
        raise ValueError("unsupported TFORM %r" % tform)
    rep = int(m.group(1)) if m.group(1) else 1
    return rep, m.group(2), m.group(3)


def _columns(h, data, heap):
    nrow, rowlen = h["NAXIS2"], h["NAXIS1"]
    cols, off = {}, 0
    rows = np.frombuffer(data, dtype=np.uint8, count=nrow * rowlen).reshape(nrow, rowlen) if nrow else None
    for k in range(1, h["TFIELDS"] + 1):
        name = h.get("TTYPE%d" % k, "COL%d" % k)
        rep, var, code = _parse_tform(h["TFORM%d" % k])
        if var:
            dsz = 4 if var == "P" else 8
            width = 2 * dsz
            desc = rows[:, off:off + width].copy().view(">i4" if var == "P" else ">i8").reshape(nrow, 2)
            dt, size = _CODES[code]
            vals = []
            for n, o in desc:
                vals.append(np.frombuffer(heap, dtype=dt, count=int(n), offset=int(o)).astype(dt[1:] if dt[0] == ">" else dt))
            cols[name] = vals
        else:
            dt, size = _CODES[code]
            if code == "X":
                width = (rep + 7) // 8
                cols[name] = rows[:, off:off + width].copy()
            elif code == "A":
                width = rep
                cols[name] = np.array([bytes(r).decode("ascii", "replace").rstrip("\x00 ")
                                       for r in rows[:, off:off + width]])
            else:
                width = rep * size
                a = rows[:, off:off + width].copy().view(dt).reshape(nrow, rep) if nrow else np.zeros((0, rep))
                a = a.astype(dt[1:] if dt[0] == ">" else dt)
                scale, zero = h.get("TSCAL%d" % k), h.get("TZERO%d" % k)
                if scale is not None or zero is not None:
                    a = a * (scale if scale is not None else 1.0) + (zero if zero is not None else 0.0)
                cols[name] = a[:, 0] if rep == 1 else a
        off += width
    return cols


def read_fits(path):
    buf = open(path, "rb").read()
    hdus, pos = [], 0
    while pos + BLOCK <= len(buf):
        h, start, pos = _read_header(buf, pos)
        naxis = h.get("NAXIS", 0)
        nbytes = 0
        if naxis:
            nbytes = abs(h.get("BITPIX", 8)) // 8
            for k in range(1, naxis + 1):
                nbytes *= h["NAXIS%d" % k]
        nbytes = h.get("GCOUNT", 1) * (nbytes + h.get("PCOUNT", 0)) if naxis else 0
        data = buf[pos:pos + nbytes]
        hdu = dict(header=h, name=h.get("EXTNAME", "PRIMARY"), hdr_start=start, data_start=pos,
                   data_bytes=nbytes)
        if h.get("XTENSION") == "BINTABLE":
            main = h["NAXIS1"] * h["NAXIS2"]
            theap = h.get("THEAP", main)
            hdu["columns"] = _columns(h, data, data[theap:])
        hdus.append(hdu)
        pos += ((nbytes + BLOCK - 1) // BLOCK) * BLOCK
    return hdus


def hdu(hdus, name):
    for x in hdus:
        if x["name"] == name:
            return x
    raise KeyError(name)


def _ones_sum(b):
    w = np.frombuffer(b, dtype=">u4").astype(np.uint64)
    s = int(w.sum())
    while s >> 32:
        s = (s & 0xFFFFFFFF) + (s >> 32)
    return s


def data_checksum(path, index):
    buf = open(path, "rb").read()
    h = read_fits(path)[index]
    padded = ((h["data_bytes"] + BLOCK - 1) // BLOCK) * BLOCK
    return _ones_sum(buf[h["data_start"]:h["data_start"] + padded])


def hdu_checksum(path, index):
    buf = open(path, "rb").read()
    h = read_fits(path)[index]
    padded = ((h["data_bytes"] + BLOCK - 1) // BLOCK) * BLOCK
    return _ones_sum(buf[h["hdr_start"]:h["data_start"] + padded])
