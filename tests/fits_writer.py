"""A small FITS writer for test files (binary tables, variable-length array
columns in the heap), used to build synthetic OGIP files for the reader."""
import numpy as np

BLOCK = 2880
_FMT = {np.dtype(">f4"): "E", np.dtype(">f8"): "D", np.dtype(">i2"): "I", np.dtype(">i4"): "J",
        np.dtype(">i8"): "K", np.dtype("u1"): "B"}


def _card(key, value, comment=""):
    if isinstance(value, bool):
        v = "%20s" % ("T" if value else "F")
    elif isinstance(value, (int, np.integer)):
        v = "%20d" % value
    elif isinstance(value, (float, np.floating)):
        v = "%20s" % repr(float(value)).upper()
    else:
        v = "'%-8s'" % str(value).replace("'", "''")
    c = "%-8s= %s" % (key, v)
    if comment:
        c += " / " + comment
    return c[:80].ljust(80)


def _header(cards):
    s = "".join(cards) + "END".ljust(80)
    pad = (-len(s)) % BLOCK
    return (s + " " * pad).encode("ascii")


def _pad(b):
    return b + b"\0" * ((-len(b)) % BLOCK)


def primary():
    return _header([_card("SIMPLE", True), _card("BITPIX", 8), _card("NAXIS", 0), _card("EXTEND", True)])


def bintable(name, columns, extra=()):
    """columns: list of (name, dtype, values) with values a 1-D array (fixed
    column) or a list of 1-D arrays (variable-length, descriptor P)."""
    nrow = None
    row_parts, forms, heap = [], [], b""
    for cname, dt, vals in columns:
        dt = np.dtype(dt).newbyteorder(">") if np.dtype(dt).itemsize > 1 else np.dtype(dt)
        if isinstance(vals, list):
            desc = np.zeros((len(vals), 2), dtype=">i4")
            for k, v in enumerate(vals):
                arr = np.asarray(v, dtype=dt)
                desc[k] = (len(arr), len(heap))
                heap += arr.tobytes()
            row_parts.append(desc.view(np.uint8).reshape(len(vals), 8))
            forms.append("P%s(%d)" % (_FMT[dt], max((len(v) for v in vals), default=0)))
            n = len(vals)
        else:
            arr = np.asarray(vals, dtype=dt)
            row_parts.append(arr.reshape(len(arr), 1).view(np.uint8).reshape(len(arr), dt.itemsize))
            forms.append("1%s" % _FMT[dt])
            n = len(arr)
        nrow = n if nrow is None else nrow
    rows = np.concatenate(row_parts, axis=1)
    cards = [_card("XTENSION", "BINTABLE"), _card("BITPIX", 8), _card("NAXIS", 2),
             _card("NAXIS1", rows.shape[1]), _card("NAXIS2", nrow), _card("PCOUNT", len(heap)),
             _card("GCOUNT", 1), _card("TFIELDS", len(columns))]
    for k, ((cname, _, _), f) in enumerate(zip(columns, forms), 1):
        cards += [_card("TTYPE%d" % k, cname), _card("TFORM%d" % k, f)]
    cards.append(_card("EXTNAME", name))
    for key, val in extra:
        cards.append(_card(key, val))
    return _header(cards) + _pad(rows.tobytes() + heap)


def write(path, hdus):
    with open(path, "wb") as f:
        f.write(primary())
        for h in hdus:
            f.write(h)
