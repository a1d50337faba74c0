"""File inputs and outputs of the pixel path (host side).

The reference reads its inputs through ``src/cosmoforge.cosmocore/cosmocore/
in_out.py`` and ``beam.py`` (healpy / astropy); neither library is used
here.  This module reads and writes the same files:

* ``read_covmat`` -- full-sky float64 row-major ``(nmaps npix)^2`` binary,
  reduced to the active rows and columns (``in_out.py:59-108``).  The file is
  memory-mapped and gathered one row block at a time, so a full-sky matrix
  larger than host memory (nside 64 TQU: 340 GB) can still be reduced;
* ``read_covmat_reduced`` / ``write_covmat_reduced`` (``in_out.py:111-171``);
* ``read_mask`` / ``read_maps`` / ``read_cl_fits`` -- HEALPix FITS maps and
  window functions (``in_out.py:174-206,439-590``, ``beam.py:202``), with a
  FITS reader of its own (binary tables and images, implicit and explicit
  pixel indexing, RING <-> NESTED reordering);
* ``readcl`` / ``convert_spectra_normalization`` (``in_out.py:255-381``);
* ``output_geometry`` / ``writecl`` / ``write_out_matrix``
  (``in_out.py:209-252,384-436``).

Output helpers are no-ops on a falsy path (``None`` or blank), as in the
reference.
"""

from __future__ import annotations

import os
from pathlib import Path

import numpy as np

UNSEEN = -1.6375e30            # HEALPix "no data" value
_BLOCK = 2880                  # FITS record length
_GATHER_BYTES = 256 << 20      # host memory per row block of read_covmat


def _path(p) -> str:
    return os.fspath(p).strip()


def _falsy(p) -> bool:
    return not (os.fspath(p) if p is not None else "").strip()


def _make_parent(p: str) -> None:
    Path(p).parent.mkdir(parents=True, exist_ok=True)


# ---------------------------------------------------------------------------
# noise covariance binaries
# ---------------------------------------------------------------------------

def read_covmat(covmatfile, npix: int, nmaps: int, active, C: np.ndarray | None = None) -> np.ndarray:
    """Reduce a full-sky covariance file to the active pixels.

    The file holds ``(nmaps npix)^2`` float64 values in row-major order;
    ``active`` is the position of every reduced row in the flattened
    ``(nmaps npix,)`` vector.  Returns ``C[i, j] = full[active[i], active[j]]``
    (written into ``C`` when given).  Follows ``in_out.py:59-108``.
    """
    path = _path(covmatfile)
    full = int(npix) * int(nmaps)
    size = os.path.getsize(path)
    if size != full * full * 8:
        raise ValueError(f"covariance file '{path}' holds {size // 8} float64 values, expected "
                         f"{full * full} for a ({full}, {full}) full-sky matrix")
    idx = np.asarray(active, dtype=np.int64).ravel()
    if idx.size and (idx.min() < 0 or idx.max() >= full):
        raise ValueError(f"active index out of range for a {full}-row covariance")
    n = idx.size
    if C is None:
        C = np.empty((n, n), dtype=np.float64)
    if C.shape != (n, n):
        raise ValueError(f"output has shape {C.shape}, expected ({n}, {n})")
    if n == 0:
        return C
    mm = np.memmap(path, dtype=np.float64, mode="r", shape=(full, full))
    rows = max(1, _GATHER_BYTES // (8 * full))
    # rows are read in file order (sorted) so the map is swept once
    order = np.argsort(idx, kind="stable")
    for s in range(0, n, rows):
        sel = order[s: s + rows]
        C[sel, :] = mm[idx[sel]][:, idx]
    del mm
    return C


def read_covmat_reduced(covmatfile, C: np.ndarray) -> np.ndarray:
    """A covariance already reduced to the active pixels (``in_out.py:111-144``)."""
    path = _path(covmatfile)
    ntot = C.shape[0]
    data = np.fromfile(path, dtype=np.float64)
    if data.size != ntot * ntot:
        raise ValueError(
            f"Invalid reduced covariance matrix size in '{path}': expected {ntot * ntot} float64 values "
            f"({ntot * ntot * 8} bytes) for shape ({ntot}, {ntot}), but found {data.size} values "
            f"({data.size * 8} bytes).")
    C[:] = data.reshape(ntot, ntot)
    return C


def write_covmat_reduced(outcovmatfile, C) -> None:
    """Row-major float64 binary of ``C`` (a host array or a device tensor)."""
    if _falsy(outcovmatfile):
        return
    path = _path(outcovmatfile)
    _make_parent(path)
    if hasattr(C, "detach"):
        C = C.detach().cpu().numpy()
    np.ascontiguousarray(C, dtype=np.float64).tofile(path)


# ---------------------------------------------------------------------------
# FITS
# ---------------------------------------------------------------------------

_TFORM = {"D": ">f8", "E": ">f4", "K": ">i8", "J": ">i4", "I": ">i2", "B": "u1", "L": "u1"}
_BITPIX = {-64: ">f8", -32: ">f4", 64: ">i8", 32: ">i4", 16: ">i2", 8: "u1"}


def _card_value(text: str):
    text = text.strip()
    if text.startswith("'"):
        end = text.find("'", 1)
        while end > 0 and end + 1 < len(text) and text[end + 1] == "'":   # '' escapes a quote
            end = text.find("'", end + 2)
        return text[1:end if end > 0 else None].replace("''", "'").rstrip()
    text = text.split("/", 1)[0].strip()
    if text in ("T", "F"):
        return text == "T"
    try:
        return int(text)
    except ValueError:
        pass
    try:
        return float(text.replace("D", "E"))
    except ValueError:
        return text


class _Hdu:
    """One header-data unit: its keywords and the byte range of its data."""

    def __init__(self, header: dict, start: int, nbytes: int):
        self.header = header
        self.start = start
        self.nbytes = nbytes

    @property
    def kind(self) -> str:
        return str(self.header.get("XTENSION", "PRIMARY")).strip().upper()

    @property
    def name(self) -> str:
        return str(self.header.get("EXTNAME", "")).strip().upper()


def _fits_hdus(buf) -> list:
    hdus, off, end = [], 0, len(buf)
    if bytes(buf[:8]).rstrip() != b"SIMPLE":
        raise ValueError("not a FITS file (no SIMPLE = T card)")
    while off + _BLOCK <= end:
        header, done = {}, False
        while not done:
            if off + _BLOCK > end:
                raise ValueError("truncated FITS header")
            rec = bytes(buf[off: off + _BLOCK]).decode("ascii", "replace")
            off += _BLOCK
            for i in range(0, _BLOCK, 80):
                card = rec[i: i + 80]
                key = card[:8].strip()
                if key == "END":
                    done = True
                    break
                if card[8:10] == "= " and key not in header:
                    header[key] = _card_value(card[10:])
        if not hdus and header.get("SIMPLE") is not True:
            raise ValueError("not a FITS file (no SIMPLE = T card)")
        naxis = int(header.get("NAXIS", 0))
        nbytes = 0
        if naxis:
            nbytes = abs(int(header.get("BITPIX", 8))) // 8
            for a in range(1, naxis + 1):
                nbytes *= int(header[f"NAXIS{a}"])
            nbytes = (nbytes + int(header.get("PCOUNT", 0))) * int(header.get("GCOUNT", 1))
        if off + nbytes > end:
            raise ValueError("truncated FITS data unit")
        hdus.append(_Hdu(header, off, nbytes))
        off += -(-nbytes // _BLOCK) * _BLOCK
    return hdus


class FitsFile:
    """Read-only view of a FITS file (memory-mapped)."""

    def __init__(self, path):
        self.path = _path(path)
        self.buf = np.memmap(self.path, dtype=np.uint8, mode="r")
        self.hdus = _fits_hdus(self.buf)

    def __contains__(self, name: str) -> bool:
        return any(h.name == name.upper() for h in self.hdus)

    def hdu(self, key) -> _Hdu:
        if isinstance(key, int):
            return self.hdus[key]
        for h in self.hdus:
            if h.name == str(key).upper():
                return h
        raise KeyError(f"no HDU named {key!r} in {self.path}")

    def first_table(self) -> _Hdu:
        for h in self.hdus[1:]:
            if h.kind == "BINTABLE":
                return h
        raise ValueError(f"no binary table in {self.path}")

    def columns(self, h: _Hdu) -> list:
        """Columns of a binary table, each flattened to float64 (a vector
        column of repeat r over R rows gives R r values in row order)."""
        hd = h.header
        if h.kind != "BINTABLE":
            raise ValueError(f"HDU {h.name or '?'} of {self.path} is not a binary table")
        width, nrow, ncol = int(hd["NAXIS1"]), int(hd["NAXIS2"]), int(hd["TFIELDS"])
        fields, offset = [], 0
        for c in range(1, ncol + 1):
            tf = str(hd[f"TFORM{c}"]).strip().upper()
            k = 0
            while k < len(tf) and tf[k].isdigit():
                k += 1
            rep, code = (int(tf[:k]) if k else 1), tf[k: k + 1]
            if code not in _TFORM:
                raise ValueError(f"unsupported TFORM{c} = {tf!r} in {self.path}")
            dt = np.dtype(_TFORM[code])
            fields.append((c, dt, rep, offset))
            offset += dt.itemsize * rep
        if offset != width:
            raise ValueError(f"binary table row of {self.path} is {width} bytes, columns add up to {offset}")
        raw = self.buf[h.start: h.start + width * nrow]
        out = []
        for c, dt, rep, o in fields:
            rec = np.ndarray((nrow, rep), dtype=dt, buffer=raw, offset=o, strides=(width, dt.itemsize))
            col = rec.astype(np.float64).ravel()
            scale, zero = hd.get(f"TSCAL{c}", 1), hd.get(f"TZERO{c}", 0)
            if scale != 1 or zero != 0:
                col = col * float(scale) + float(zero)
            out.append(col)
        return out

    def column_names(self, h: _Hdu) -> list:
        return [str(h.header.get(f"TTYPE{c}", f"COL{c}")).strip().upper()
                for c in range(1, int(h.header["TFIELDS"]) + 1)]

    def image(self, h: _Hdu) -> np.ndarray:
        """An image unit as a float64 array of shape (NAXISn, ..., NAXIS1)."""
        hd = h.header
        naxis = int(hd.get("NAXIS", 0))
        dims = [int(hd[f"NAXIS{a}"]) for a in range(naxis, 0, -1)]
        dt = np.dtype(_BITPIX[int(hd["BITPIX"])])
        count = int(np.prod(dims)) if dims else 0
        arr = np.frombuffer(self.buf[h.start: h.start + count * dt.itemsize], dtype=dt).astype(np.float64)
        scale, zero = hd.get("BSCALE", 1), hd.get("BZERO", 0)
        if scale != 1 or zero != 0:
            arr = arr * float(scale) + float(zero)
        return arr.reshape(dims)


# ---------------------------------------------------------------------------
# HEALPix pixel orderings
# ---------------------------------------------------------------------------

_JRLL = np.array([2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4], dtype=np.int64)
_JPLL = np.array([1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7], dtype=np.int64)


def _compress_bits(v: np.ndarray) -> np.ndarray:
    """Even bits of ``v`` packed together (the x of a Morton code)."""
    v = v & 0x5555555555555555
    v = (v | (v >> 1)) & 0x3333333333333333
    v = (v | (v >> 2)) & 0x0F0F0F0F0F0F0F0F
    v = (v | (v >> 4)) & 0x00FF00FF00FF00FF
    v = (v | (v >> 8)) & 0x0000FFFF0000FFFF
    v = (v | (v >> 16)) & 0x00000000FFFFFFFF
    return v


def nest2ring(nside: int, ipix) -> np.ndarray:
    """RING index of every NESTED pixel (Gorski et al. 2005, ApJ 622, 759:
    face number and in-face Morton code -> ring number and position)."""
    if nside < 1 or nside & (nside - 1):
        raise ValueError(f"NESTED ordering needs a power-of-two nside, got {nside}")
    ipix = np.asarray(ipix, dtype=np.int64)
    npface = nside * nside
    ncap = 2 * nside * (nside - 1)
    npix = 12 * npface
    face = ipix // npface
    ipf = ipix % npface
    ix = _compress_bits(ipf)
    iy = _compress_bits(ipf >> 1)
    jr = _JRLL[face] * nside - ix - iy - 1
    nr = np.where(jr < nside, jr, np.where(jr > 3 * nside, 4 * nside - jr, nside))
    before = np.where(jr < nside, 2 * nr * (nr - 1),
                      np.where(jr > 3 * nside, npix - 2 * (nr + 1) * nr, ncap + (jr - nside) * 4 * nside))
    kshift = np.where((jr >= nside) & (jr <= 3 * nside), (jr - nside) & 1, 0)
    jp = (_JPLL[face] * nr + ix - iy + 1 + kshift) // 2
    jp = np.where(jp > 4 * nr, jp - 4 * nr, jp)
    jp = np.where(jp < 1, jp + 4 * nr, jp)
    return before + jp - 1


def reorder(m: np.ndarray, nside: int, to_nest: bool) -> np.ndarray:
    """Reorder the last axis of ``m`` between RING and NESTED."""
    n2r = nest2ring(nside, np.arange(12 * nside * nside))
    return m[..., n2r] if to_nest else _scatter_last(m, n2r)


def _scatter_last(m: np.ndarray, idx: np.ndarray) -> np.ndarray:
    out = np.empty_like(m)
    out[..., idx] = m
    return out


# ---------------------------------------------------------------------------
# maps, masks and window functions
# ---------------------------------------------------------------------------

def read_healpix(path, field=None, nest: bool = False, hdu=1) -> np.ndarray:
    """HEALPix map columns of a binary table as ``(nfield, npix)`` float64,
    in the requested ordering.

    ``field`` selects value columns (an int, a sequence, or ``None`` for
    all); an explicitly indexed (partial-sky) table is expanded to the full
    sky with ``UNSEEN`` where it holds no pixel.  Covers ``hp.read_map`` as
    ``in_out.py:205,563`` calls it.
    """
    f = FitsFile(path)
    h = f.hdu(hdu)
    hd = h.header
    cols = f.columns(h)
    names = f.column_names(h)
    explicit = str(hd.get("INDXSCHM", "IMPLICIT")).strip().upper() == "EXPLICIT"
    pix = None
    if explicit:
        k = names.index("PIXEL") if "PIXEL" in names else 0
        pix = cols[k].astype(np.int64)
        cols = cols[:k] + cols[k + 1:]
    if field is None:
        sel = list(range(len(cols)))
    elif isinstance(field, (int, np.integer)):
        sel = [int(field)]
    else:
        sel = [int(i) for i in field]
    for i in sel:
        if not 0 <= i < len(cols):
            raise ValueError(f"{path}: field {i} not present (the table holds {len(cols)} map column(s))")
    nside = int(hd["NSIDE"]) if "NSIDE" in hd else None
    if explicit:
        if nside is None:
            raise ValueError(f"{path}: explicitly indexed map without NSIDE")
        npix = 12 * nside * nside
        out = np.full((len(sel), npix), UNSEEN)
        for r, i in enumerate(sel):
            out[r, pix] = cols[i][: pix.size]
    else:
        out = np.array([cols[i] for i in sel], dtype=np.float64)
        npix = out.shape[1]
        if nside is None:
            nside = int(round(np.sqrt(npix / 12.0)))
        if 12 * nside * nside != npix:
            raise ValueError(f"{path}: {npix} values per column is not a HEALPix map")
    ordering = str(hd.get("ORDERING", "RING")).strip().upper()
    if ordering not in ("RING", "NESTED"):
        raise ValueError(f"{path}: unknown ORDERING {ordering!r}")
    if (ordering == "NESTED") != bool(nest):
        out = reorder(out, nside, to_nest=bool(nest))
    return out


def read_mask(maskfile, mask, nest: bool = False) -> np.ndarray:
    """``(npix, nmaps)`` mask; ``mask.shape[0]`` gives the field count
    (``in_out.py:174-206``)."""
    nmaps = mask.shape[0]
    return read_healpix(maskfile, field=range(nmaps), nest=nest).T.copy()


def read_cl_fits(path) -> np.ndarray:
    """Every column of the first binary table as ``(ncol, nrow)`` float64
    (``hp.read_cl`` as ``beam.py:202`` calls it)."""
    f = FitsFile(path)
    return np.array(f.columns(f.first_table()), dtype=np.float64)


def get_field_index(header: dict, field_name: str) -> int:
    """Position of ``field_name`` in a ``FIELDS`` keyword, either
    comma-separated (``"T,Q,U"``) or concatenated (``"TQU"``)
    (``in_out.py:592-636``)."""
    s = str(header.get("FIELDS", ""))
    avail = [x.strip() for x in s.split(",")] if "," in s else list(s.strip())
    if field_name not in avail:
        raise ValueError(f"Field '{field_name}' not found in {avail}")
    return avail.index(field_name)


_SIMPLE_FIELDS = {"T": 0, "Q": 1, "U": 2, "E": 1, "B": 2}


def read_maps(maps: np.ndarray, filename, pixact, field_labels) -> np.ndarray:
    """Fill ``maps`` (n_active, nsims) from a ``.npy`` stack, a multi-simulation
    FITS file (``SIM_XXX`` image units with a ``FIELDS`` keyword) or a plain
    HEALPix FITS map (one simulation), following ``in_out.py:439-590``.
    Plain maps are read in RING order, multi-simulation images as stored."""
    if maps.shape[0] != sum(len(p) for p in pixact):
        raise ValueError("maps array has incorrect shape")
    if len(field_labels) != len(pixact):
        raise ValueError(f"Mismatch between pixact length ({len(pixact)}) and field labels "
                         f"({len(field_labels)}): {field_labels}")
    nsims = maps.shape[1]
    path = _path(filename)
    pix = [np.asarray(p, dtype=np.int64) for p in pixact]
    offs = np.concatenate([[0], np.cumsum([p.size for p in pix])])
    if path.endswith(".npy"):
        data = np.load(path, mmap_mode="r")
        if data.ndim != 3:
            raise ValueError(f"Numpy map file must be 3D (n_fields, n_pix, n_sims), got shape {data.shape}")
        if data.shape[0] < len(pix):
            raise ValueError(f"Numpy map file has {data.shape[0]} fields, need at least {len(pix)}")
        if data.shape[2] < nsims:
            raise ValueError(f"Numpy map file has {data.shape[2]} sims, need at least {nsims}")
        for k, p in enumerate(pix):
            maps[offs[k]: offs[k + 1], :] = data[k][p, :nsims]
        return maps
    f = FitsFile(path)
    if "SIM_000" in f:
        first = f.hdu("SIM_000").header
        where = [get_field_index(first, lab) for lab in field_labels]
        for s in range(nsims):
            img = f.image(f.hdu(f"SIM_{s:03d}"))
            img = img.reshape(1, -1) if img.ndim == 1 else img
            for k, p in enumerate(pix):
                maps[offs[k]: offs[k + 1], s] = img[where[k]][p]
        return maps
    if nsims != 1:
        raise ValueError(f"Simple HEALPix format only supports nsims=1, got nsims={nsims}")
    allm = read_healpix(path, field=None, nest=False)
    for k, lab in enumerate(field_labels):
        if lab not in _SIMPLE_FIELDS:
            raise ValueError(f"Unknown field label '{lab}' for simple HEALPix format")
        i = _SIMPLE_FIELDS[lab]
        if i >= allm.shape[0]:
            raise ValueError(f"Field '{lab}' (index {i}) not available in file with {allm.shape[0]} fields")
        maps[offs[k]: offs[k + 1], 0] = allm[i][pix[k]]
    return maps


# ---------------------------------------------------------------------------
# power spectra text files
# ---------------------------------------------------------------------------

def _parse_convention(value) -> str:
    key = str(value).strip().lower()
    if key not in ("cl", "dl"):
        raise ValueError(f"Unknown spectra convention '{value}'. Must be 'Cl' or 'Dl'.")
    return "Cl" if key == "cl" else "Dl"


def convert_spectra_normalization(cls_dict: dict, from_norm, to_norm) -> dict:
    """In-place C_l <-> D_l = l(l+1)/(2 pi) C_l on l-indexed arrays; the
    D_l -> C_l factor at l = 0 is 0 (``in_out.py:264-302``)."""
    a, b = _parse_convention(from_norm), _parse_convention(to_norm)
    if a == b:
        return cls_dict
    n = next(iter(cls_dict.values())).shape[0]
    ell = np.arange(n, dtype=np.float64)
    if a == "Dl":
        fac = np.zeros(n)
        fac[1:] = 2.0 * np.pi / (ell[1:] * (ell[1:] + 1.0))
    else:
        fac = ell * (ell + 1.0) / (2.0 * np.pi)
    for k in cls_dict:
        cls_dict[k] = cls_dict[k] * fac
    return cls_dict


def readcl(inputclfile, Params, logger=None, lmax: int | None = None) -> dict:
    """``{label: C_l}`` (l-indexed, length lmax + 1) from a text file whose
    first line is ``# label label ...``.  An ``ell`` column places each row;
    without one the rows start at l = 2.  Rows beyond ``lmax`` are dropped,
    missing multipoles are zero, and ``Params.input_convention == "Dl"``
    converts to C_l (``in_out.py:305-381``)."""
    lmax = Params.lmax if lmax is None else lmax
    path = _path(inputclfile)
    with open(path) as fh:
        header = fh.readline()
        if not header.lstrip().startswith("#"):
            raise ValueError("First line must be a header starting with '#'")
        labels = header.strip().lstrip("#").split()
        arr = np.loadtxt(fh, dtype=np.float64, ndmin=2)
    if logger is not None:
        logger.log_with_feedback(f"Read Cls: {arr.shape} {labels}", level=4)
    low = [lab.lower() for lab in labels]
    ell_col = low.index("ell") if "ell" in low else None
    n_out = lmax + 1
    if ell_col is not None:
        ells = arr[:, ell_col].astype(int)
        keep = (ells >= 0) & (ells <= lmax)
    else:
        nfit = max(0, min(arr.shape[0], lmax - 1))
    out = {}
    for i, lab in enumerate(labels):
        if i == ell_col:
            continue
        col = np.zeros(n_out)
        if ell_col is not None:
            col[ells[keep]] = arr[keep, i]
        elif nfit > 0:
            col[2: 2 + nfit] = arr[:nfit, i]
        out[lab] = col
    conv = _parse_convention(getattr(Params, "input_convention", "Cl"))
    if conv != "Cl":
        convert_spectra_normalization(out, conv, "Cl")
    return out


def writecl(filename, power_spectra) -> None:
    if _falsy(filename):
        return
    path = _path(filename)
    _make_parent(path)
    np.savetxt(path, np.asarray(power_spectra), fmt="%.16e")


def write_out_matrix(outfilematrix, matrix) -> None:
    """One row per line, ``%24.16E`` per entry (``in_out.py:408-436``)."""
    if _falsy(outfilematrix):
        return
    path = _path(outfilematrix)
    _make_parent(path)
    m = np.asarray(matrix, dtype=np.float64)
    with open(path, "w") as fh:
        for row in m:
            fh.write("".join(f"{x:24.16E}" for x in row) + "\n")


def output_geometry(filegeometry, npixs, point_vectors, active) -> None:
    """Per field a header ``field, n, 0, 0`` (``%6d`` each), then per active
    pixel its index (``%6d``) and unit vector (``%24.16e`` x 3)
    (``in_out.py:209-252``)."""
    if _falsy(filegeometry):
        return
    nmaps = len(npixs)
    if len(active) != nmaps:
        raise ValueError(f"active has {len(active)} entries but npixs has {nmaps}: both must be per field. "
                         "Did you pass the per-component active pixels (pixact)?")
    path = _path(filegeometry)
    _make_parent(path)
    with open(path, "w") as fh:
        for k in range(nmaps):
            n = int(npixs[k])
            fh.write(f"{k + 1:6d}{n:6d}{0:6d}{0:6d}\n")
            vec = np.asarray(point_vectors[k], dtype=np.float64)
            act = np.asarray(active[k])
            fh.write("".join(f"{int(act[i]):6d}{vec[i, 0]:24.16e}{vec[i, 1]:24.16e}{vec[i, 2]:24.16e}\n"
                             for i in range(n)))
