"""Minimal FITS reader/writer (numpy + gzip) for the files on the correlation path.

picca reads and writes FITS through ``fitsio`` (``py/picca/io.py:19``,
``py/picca/bin/picca_cf.py:14``).  That library is not part of this stack, so this
module implements the subset of the FITS standard (Pence et al. 2010, A&A 524, A42)
that the correlation path touches:

* primary HDUs with no data, IMAGE extensions (BITPIX 8/16/32/64/-32/-64 with
  BZERO/BSCALE), BINTABLE extensions with fixed-width columns (L, B, I, J, K, E, D, A,
  with repeat counts, TDIM, TZERO/TSCAL);
* gzip-compressed (``.fits.gz``) and plain files;
* a BINTABLE writer for the ``ATTRI`` / ``COR`` / ``DMAT`` outputs of
  ``py/picca/bin/picca_cf.py:486-580`` and ``py/picca/bin/picca_dmat.py:521-627``.

The API is a small look-alike of fitsio's: ``FITS(path)[ext]`` gives an HDU with
``read_header()``, ``read()``, ``get_colnames()``, ``get_nrows()`` and column access
``hdu["NAME"][:]``.
"""
from __future__ import annotations

import gzip
import os
import re

import numpy as np

BLOCK = 2880

_TFORM_RE = re.compile(r"^\s*(\d*)([LXBIJKAEDCMPQ])(.*)$")
_TFORM_DTYPE = {
    "L": ("i1", 1),
    "B": ("u1", 1),
    "I": (">i2", 2),
    "J": (">i4", 4),
    "K": (">i8", 8),
    "E": (">f4", 4),
    "D": (">f8", 8),
    "C": (">c8", 8),
    "M": (">c16", 16),
}
_BITPIX_DTYPE = {8: "u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}


class FITSHeader(dict):
    """Ordered keyword → value mapping that keeps the card comments."""

    def __init__(self):
        super().__init__()
        self.comments = {}
        self.order = []

    def records(self):
        """Return a list of ``{"name", "value", "comment"}`` dicts, like fitsio."""
        return [{"name": k, "value": self[k], "comment": self.comments.get(k, "")}
                for k in self.order]


def _parse_value(raw: str):
    raw = raw.strip()
    if raw.startswith("'"):
        # string with '' escapes, terminated by a lone quote
        out = []
        idx = 1
        while idx < len(raw):
            ch = raw[idx]
            if ch == "'":
                if idx + 1 < len(raw) and raw[idx + 1] == "'":
                    out.append("'")
                    idx += 2
                    continue
                idx += 1
                break
            out.append(ch)
            idx += 1
        rest = raw[idx:]
        comment = rest.split("/", 1)[1].strip() if "/" in rest else ""
        return "".join(out).rstrip(), comment
    value, _, comment = raw.partition("/")
    value = value.strip()
    comment = comment.strip()
    if value == "T":
        return True, comment
    if value == "F":
        return False, comment
    if value == "":
        return None, comment
    try:
        return int(value), comment
    except ValueError:
        pass
    try:
        return float(value.replace("D", "E")), comment
    except ValueError:
        return value, comment


def _parse_header(buf: bytes, offset: int):
    header = FITSHeader()
    pos = offset
    done = False
    while not done:
        block = buf[pos:pos + BLOCK]
        if len(block) < BLOCK:
            raise OSError("truncated FITS header")
        pos += BLOCK
        text = block.decode("ascii", errors="replace")
        for k in range(0, BLOCK, 80):
            card = text[k:k + 80]
            key = card[:8].strip()
            if key == "END":
                done = True
                break
            if key in ("", "COMMENT", "HISTORY"):
                continue
            if card[8:10] != "= ":
                continue
            value, comment = _parse_value(card[10:])
            if key == "CONTINUE":
                continue
            if key not in header:
                header.order.append(key)
            header[key] = value
            header.comments[key] = comment
    return header, pos


def _data_size(header) -> int:
    naxis = int(header.get("NAXIS", 0))
    if naxis == 0:
        return 0
    size = abs(int(header["BITPIX"])) // 8
    for k in range(1, naxis + 1):
        size *= int(header["NAXIS%d" % k])
    size = int(header.get("GCOUNT", 1)) * (int(header.get("PCOUNT", 0)) + size)
    return size


class _Column:
    def __init__(self, hdu, name):
        self._hdu = hdu
        self._name = name

    def __getitem__(self, item):
        return self._hdu.read_column(self._name)[item]

    def read(self):
        return self._hdu.read_column(self._name)


class HDU:
    """One header-data unit of a FITS file."""

    def __init__(self, header, data: bytes):
        self._header = header
        self._data = data
        self._cols = None
        self._cache = {}

    # -- fitsio-like API -------------------------------------------------
    def read_header(self):
        return self._header

    @property
    def is_table(self):
        return self._header.get("XTENSION", "").strip() in ("BINTABLE",)

    def get_extname(self):
        return str(self._header.get("EXTNAME", "")).strip()

    def get_nrows(self):
        if self.is_table:
            return int(self._header["NAXIS2"])
        naxis = int(self._header.get("NAXIS", 0))
        return int(self._header["NAXIS%d" % naxis]) if naxis else 0

    def get_colnames(self):
        return [c["name"] for c in self._columns()]

    def __contains__(self, name):
        return name in self.get_colnames()

    def __getitem__(self, name):
        if name not in self.get_colnames():
            raise KeyError(name)
        return _Column(self, name)

    def read(self, columns=None):
        if not self.is_table:
            return self._read_image()
        names = self.get_colnames() if columns is None else list(columns)
        return {n: self.read_column(n) for n in names}

    # -- internals ---------------------------------------------------------
    def _read_image(self):
        header = self._header
        naxis = int(header.get("NAXIS", 0))
        if naxis == 0:
            return None
        dtype = np.dtype(_BITPIX_DTYPE[int(header["BITPIX"])])
        shape = tuple(int(header["NAXIS%d" % k]) for k in range(naxis, 0, -1))
        count = int(np.prod(shape))
        arr = np.frombuffer(self._data, dtype=dtype, count=count).reshape(shape)
        arr = arr.astype(dtype.newbyteorder("="))
        bscale = header.get("BSCALE", 1)
        bzero = header.get("BZERO", 0)
        if bscale != 1 or bzero != 0:
            if int(header["BITPIX"]) > 0 and bscale == 1 and float(bzero).is_integer():
                # unsigned-integer convention
                if int(header["BITPIX"]) == 8:
                    arr = (arr.astype(np.int16) + int(bzero)).astype(np.int8)
                else:
                    big = {16: np.uint16, 32: np.uint32, 64: np.uint64}[int(header["BITPIX"])]
                    arr = (arr.astype(np.int64) + int(bzero)).astype(big) if bzero > 0 else arr + int(bzero)
            else:
                arr = arr * bscale + bzero
        return arr

    def _columns(self):
        if self._cols is not None:
            return self._cols
        if not self.is_table:
            self._cols = []
            return self._cols
        header = self._header
        nfields = int(header["TFIELDS"])
        cols = []
        offset = 0
        for k in range(1, nfields + 1):
            tform = str(header["TFORM%d" % k]).strip()
            match = _TFORM_RE.match(tform)
            if match is None:
                raise ValueError("unsupported TFORM %r" % tform)
            repeat = int(match.group(1)) if match.group(1) else 1
            code = match.group(2)
            if code == "A":
                width = repeat
                dtype = "S%d" % max(repeat, 1)
            elif code in ("P", "Q", "X"):
                raise ValueError("variable-length/bit columns are not supported (%s)" % tform)
            else:
                base, size = _TFORM_DTYPE[code]
                width = repeat * size
                dtype = base
            name = str(header.get("TTYPE%d" % k, "COL%d" % k)).strip()
            tdim = header.get("TDIM%d" % k)
            cols.append(dict(name=name, code=code, repeat=repeat, offset=offset,
                             width=width, dtype=dtype,
                             tzero=header.get("TZERO%d" % k),
                             tscal=header.get("TSCAL%d" % k), tdim=tdim))
            offset += width
        self._cols = cols
        return cols

    def read_column(self, name):
        if name in self._cache:
            return self._cache[name]
        col = next(c for c in self._columns() if c["name"] == name)
        nrows = int(self._header["NAXIS2"])
        rowlen = int(self._header["NAXIS1"])
        raw = np.frombuffer(self._data, dtype=np.uint8, count=rowlen * nrows).reshape(nrows, rowlen)
        chunk = np.ascontiguousarray(raw[:, col["offset"]:col["offset"] + col["width"]])
        if col["code"] == "A":
            out = chunk.view(col["dtype"]).reshape(nrows)
            out = np.char.rstrip(out.astype(str))
        else:
            dtype = np.dtype(col["dtype"])
            out = chunk.view(dtype).reshape((nrows, col["repeat"]) if col["repeat"] != 1 else (nrows,))
            out = out.astype(dtype.newbyteorder("="))
            if col["code"] == "L":
                out = out == ord("T")
            if col["tscal"] not in (None, 1) or col["tzero"] not in (None, 0):
                tzero = col["tzero"] or 0
                tscal = col["tscal"] if col["tscal"] is not None else 1
                if col["code"] == "K" and tzero == 2 ** 63 and tscal == 1:
                    out = (out.view(np.uint64) ^ np.uint64(1 << 63)).astype(np.uint64)
                elif col["code"] in ("I", "J", "B") and tscal == 1 and float(tzero).is_integer():
                    out = out.astype(np.int64) + int(tzero)
                else:
                    out = out * tscal + tzero
            if col["tdim"] and col["repeat"] > 1:
                dims = [int(x) for x in str(col["tdim"]).strip("() ").split(",")]
                out = out.reshape((nrows,) + tuple(reversed(dims)))
        self._cache[name] = out
        return out


def _read_bytes(path):
    with open(path, "rb") as handle:
        magic = handle.read(2)
    if magic == b"\x1f\x8b":
        with gzip.open(path, "rb") as handle:
            return handle.read()
    with open(path, "rb") as handle:
        return handle.read()


class FITS:
    """Read-only FITS file, indexable by HDU number or EXTNAME (case-insensitive)."""

    def __init__(self, path):
        self.path = os.fspath(path)
        buf = _read_bytes(self.path)
        self.hdus = []
        pos = 0
        while pos + BLOCK <= len(buf):
            if not buf[pos:pos + 8].strip(b" \x00"):
                break
            header, pos = _parse_header(buf, pos)
            size = _data_size(header)
            data = buf[pos:pos + size]
            pos += ((size + BLOCK - 1) // BLOCK) * BLOCK
            self.hdus.append(HDU(header, data))

    def __len__(self):
        return len(self.hdus)

    def __iter__(self):
        return iter(self.hdus)

    def _index(self, key):
        if isinstance(key, (int, np.integer)):
            return int(key)
        for idx, hdu in enumerate(self.hdus):
            if hdu.get_extname().upper() == str(key).upper():
                return idx
        raise KeyError(key)

    def __contains__(self, key):
        try:
            self._index(key)
            return True
        except KeyError:
            return False

    def __getitem__(self, key):
        if isinstance(key, slice):
            return self.hdus[key]
        return self.hdus[self._index(key)]

    def close(self):
        self.hdus = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ---------------------------------------------------------------------------
# writer
# ---------------------------------------------------------------------------

def _format_value(value):
    if isinstance(value, (bool, np.bool_)):
        return "%20s" % ("T" if value else "F")
    if isinstance(value, (int, np.integer)):
        return "%20d" % int(value)
    if isinstance(value, (float, np.floating)):
        text = repr(float(value)).upper()
        if "E" not in text and "." not in text and "N" not in text and "I" not in text:
            text += "."
        return "%20s" % text
    if value is None:
        return " " * 20
    text = "'%-8s'" % str(value).replace("'", "''")
    return "%-20s" % text


def _card(key, value, comment=""):
    body = "%-8s= %s" % (key[:8], _format_value(value))
    if comment:
        body += " / " + comment
    return body[:80].ljust(80)


def _pad(data: bytes, fill=b"\x00"):
    rem = (-len(data)) % BLOCK
    return data + fill * rem


def _header_bytes(cards):
    text = "".join(cards) + "END".ljust(80)
    return _pad(text.encode("ascii"), b" ")


def _column_spec(arr):
    arr = np.asarray(arr)
    if arr.dtype.kind == "U" or arr.dtype.kind == "S":
        width = max(1, int(arr.dtype.itemsize // (4 if arr.dtype.kind == "U" else 1)))
        return "%dA" % width, arr.astype("S%d" % width), width
    kind = {
        ("b", 1): ("L", "i1"),
        ("i", 2): ("I", ">i2"),
        ("i", 4): ("J", ">i4"),
        ("i", 8): ("K", ">i8"),
        ("u", 1): ("B", "u1"),
        ("f", 4): ("E", ">f4"),
        ("f", 8): ("D", ">f8"),
    }
    key = (arr.dtype.kind, arr.dtype.itemsize)
    if arr.dtype.kind == "u" and arr.dtype.itemsize > 1:
        key = ("i", 8)
    code, dtype = kind[key]
    repeat = int(np.prod(arr.shape[1:])) if arr.ndim > 1 else 1
    if code == "L":
        data = np.where(arr, ord("T"), ord("F")).astype("i1")
    else:
        data = arr.astype(dtype)
    return ("%d%s" % (repeat, code) if repeat != 1 else code), data, repeat * np.dtype(dtype).itemsize


class FITSWriter:
    """Writes a primary HDU followed by BINTABLE extensions (gzip when the name ends in .gz)."""

    def __init__(self, path):
        self.path = os.fspath(path)
        self._chunks = [_header_bytes([
            _card("SIMPLE", True, "file does conform to FITS standard"),
            _card("BITPIX", 16, "number of bits per data pixel"),
            _card("NAXIS", 0, "number of data axes"),
            _card("EXTEND", True, "FITS dataset may contain extensions"),
        ])]

    def write(self, columns, names, header=None, extname=None, comment=None, units=None):
        """Append a BINTABLE; ``header`` is a list of {"name","value","comment"} dicts."""
        specs = [_column_spec(c) for c in columns]
        nrows = len(np.asarray(columns[0]))
        rowlen = sum(s[2] for s in specs)
        cards = [
            _card("XTENSION", "BINTABLE", "binary table extension"),
            _card("BITPIX", 8, "8-bit bytes"),
            _card("NAXIS", 2, "2-dimensional binary table"),
            _card("NAXIS1", rowlen, "width of table in bytes"),
            _card("NAXIS2", nrows, "number of rows in table"),
            _card("PCOUNT", 0, "size of special data area"),
            _card("GCOUNT", 1, "one data group (required keyword)"),
            _card("TFIELDS", len(columns), "number of fields in each row"),
        ]
        for k, (name, spec) in enumerate(zip(names, specs), start=1):
            cards.append(_card("TTYPE%d" % k, name, comment[k - 1] if comment else ""))
            cards.append(_card("TFORM%d" % k, spec[0], "data format of field"))
            if units and units[k - 1]:
                cards.append(_card("TUNIT%d" % k, units[k - 1], ""))
        if extname:
            cards.append(_card("EXTNAME", extname, "name of this binary table extension"))
        for record in header or []:
            value = record["value"]
            cards.append(_card(record["name"], value, record.get("comment", "")))
        rows = np.zeros(nrows, dtype=[("c%d" % k, s[1].dtype, s[1].shape[1:])
                                      for k, s in enumerate(specs)])
        for k, spec in enumerate(specs):
            rows["c%d" % k] = spec[1]
        payload = rows.tobytes()
        assert len(payload) == rowlen * nrows
        self._chunks.append(_header_bytes(cards))
        self._chunks.append(_pad(payload))

    def close(self):
        blob = b"".join(self._chunks)
        if self.path.endswith(".gz"):
            with gzip.open(self.path, "wb", compresslevel=6) as handle:
                handle.write(blob)
        else:
            with open(self.path, "wb") as handle:
                handle.write(blob)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
