"""Minimal FITS reader for the reference's fixture files (oracle only).

astropy/healpy are absent.  The fixtures the oracle reads are plain FITS:
an empty primary HDU followed by BINTABLE extensions whose columns are
``TFORM='D'`` (big-endian float64, one value per row) as written by
``hp.write_map`` / ``hp.write_cl``, or IMAGE extensions with
``BITPIX=-64``.  This covers ``hp.read_map`` (mask, maps) and
``hp.read_cl`` (beam) as the reference calls them at
``src/cosmoforge.cosmocore/cosmocore/in_out.py:174-206,439-590`` and
``beam.py:202``.
"""

from __future__ import annotations

import numpy as np

_BLOCK = 2880


def _read_header(buf: bytes, off: int):
    cards = {}
    order = []
    while True:
        block = buf[off : off + _BLOCK]
        if len(block) < _BLOCK:
            raise ValueError("truncated FITS header")
        off += _BLOCK
        done = False
        for i in range(36):
            card = block[i * 80 : (i + 1) * 80].decode("ascii", "replace")
            key = card[:8].strip()
            if key == "END":
                done = True
                break
            if card[8:10] == "= ":
                val = card[10:].split("/")[0].strip()
                if val.startswith("'"):
                    val = val.strip("'").strip()
                cards[key] = val
                order.append(key)
        if done:
            return cards, off


def _tform_dtype(tform: str):
    tform = tform.strip()
    rep = ""
    while tform and tform[0].isdigit():
        rep += tform[0]
        tform = tform[1:]
    rep = int(rep) if rep else 1
    code = tform[0]
    table = {"D": ">f8", "E": ">f4", "J": ">i4", "K": ">i8", "I": ">i2", "B": "u1"}
    if code not in table:
        raise ValueError(f"unsupported TFORM {tform!r}")
    return np.dtype(table[code]), rep


def read_hdus(path: str):
    """Return a list of (header dict, data) for every HDU in ``path``."""
    with open(path, "rb") as fh:
        buf = fh.read()
    off = 0
    hdus = []
    while off < len(buf):
        hdr, off = _read_header(buf, off)
        bitpix = int(hdr.get("BITPIX", "8"))
        naxis = int(hdr.get("NAXIS", "0"))
        dims = [int(hdr[f"NAXIS{i}"]) for i in range(1, naxis + 1)]
        nbytes = 0
        if naxis:
            nbytes = abs(bitpix) // 8
            for d in dims:
                nbytes *= d
            nbytes += int(hdr.get("PCOUNT", "0"))
        raw = buf[off : off + nbytes]
        off += ((nbytes + _BLOCK - 1) // _BLOCK) * _BLOCK
        data = None
        xt = hdr.get("XTENSION", "")
        if xt == "BINTABLE":
            nrow = dims[1]
            ncol = int(hdr["TFIELDS"])
            fields = []
            for c in range(1, ncol + 1):
                dt, rep = _tform_dtype(hdr[f"TFORM{c}"])
                name = hdr.get(f"TTYPE{c}", f"COL{c}")
                fields.append((name, dt, (rep,) if rep > 1 else ()))
            rec = np.frombuffer(raw[: dims[0] * nrow], dtype=np.dtype(fields))
            data = {name: np.asarray(rec[name], dtype=np.float64).ravel() for name, _, _ in fields}
            data = [data[name] for name, _, _ in fields]
        elif naxis and bitpix in (-64, -32, 8, 16, 32, 64):
            code = {-64: ">f8", -32: ">f4", 8: "u1", 16: ">i2", 32: ">i4", 64: ">i8"}[bitpix]
            arr = np.frombuffer(raw[: np.prod(dims) * abs(bitpix) // 8], dtype=code)
            data = arr.reshape(dims[::-1]).astype(np.float64)
        hdus.append((hdr, data))
    return hdus


def read_table_columns(path: str, ncols: int | None = None) -> np.ndarray:
    """Columns of the first BINTABLE as a (ncols, nrows) float64 array."""
    for hdr, data in read_hdus(path):
        if hdr.get("XTENSION", "") == "BINTABLE":
            cols = data if ncols is None else data[:ncols]
            return np.array(cols, dtype=np.float64)
    raise ValueError(f"no BINTABLE in {path}")


def read_image_hdus(path: str) -> list[np.ndarray]:
    """Data arrays of every IMAGE extension (e.g. ``SIM_000`` maps)."""
    return [d for h, d in read_hdus(path) if h.get("XTENSION", "") == "IMAGE"]


def ring_from_nest_map(m: np.ndarray, nside: int) -> np.ndarray:
    """Unused helper guard: the fixtures are all RING-ordered."""
    raise NotImplementedError("NESTED map files are not used by the fixtures")
