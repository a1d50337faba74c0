"""Oracle signal and derivative matrices (FP64, CPU; test infrastructure only).

Assembly follows the reference:

* ``compute_signal_matrix`` -- ``src/cosmoforge.cosmocore/cosmocore/
  signal_kernels.py:390-502`` (field-pair walk, lower blocks, mirror);
* ``PixelBasis.get_binned_derivative_direct`` -- ``src/cosmoforge.cosmocore/
  cosmocore/basis/pixel.py:802-950`` (one recurrence pass per bin with the
  in-bin weights ``w_l * b_l^2``);
* field / spectrum bookkeeping -- ``fields.py:495-690`` and
  ``spectra_io.py:155-190`` (auto spectra per field, then cross spectra).

The pair loops themselves are the C restatement in ``csrc/oracle_kernels.c``
(built by ``build_oracle()``); ``legendre_*_py`` below are pure-Python
restatements of ``basics/legendre.py:10-149`` used to pin the C code on
small inputs.
"""

from __future__ import annotations

import ctypes
import math
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")
_SO = os.path.join(_BUILD, "liboracle_kernels.so")
_SRC = os.path.join(_HERE, "csrc", "oracle_kernels.c")


def build_oracle(force: bool = False) -> str:
    """Compile the C restatement with gcc + OpenMP into ``oracle/_build``."""
    os.makedirs(_BUILD, exist_ok=True)
    if force or not os.path.exists(_SO) or os.path.getmtime(_SO) < os.path.getmtime(_SRC):
        tmp = _SO + ".tmp"
        subprocess.check_call(
            ["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
             "-o", tmp, _SRC, "-lm"]
        )
        os.replace(tmp, _SO)
    return _SO


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(build_oracle())
        dp = ctypes.POINTER(ctypes.c_double)
        i64 = ctypes.c_int64
        L.oracle_compute_00.argtypes = [dp, ctypes.c_int, ctypes.c_void_p, i64, dp, dp, i64, ctypes.c_int]
        L.oracle_compute_22.argtypes = [dp, dp, dp, ctypes.c_int, ctypes.c_void_p, i64, dp, dp, i64]
        L.oracle_compute_02.argtypes = [dp, dp, ctypes.c_int, ctypes.c_void_p, i64, dp, dp, i64, i64]
        L.oracle_copy_lower_to_upper.argtypes = [ctypes.c_void_p, i64, i64]
        L.oracle_matrix_trace.argtypes = [dp, dp, i64]
        L.oracle_matrix_trace.restype = ctypes.c_double
        L.oracle_rotation_angle.argtypes = [dp, dp, dp, dp]
        L.oracle_legendre_00.argtypes = [ctypes.c_double, dp, ctypes.c_int]
        L.oracle_legendre_22.argtypes = [ctypes.c_double, dp, dp, dp, ctypes.c_int]
        L.oracle_legendre_02.argtypes = [ctypes.c_double, dp, ctypes.c_int]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


# ---------------------------------------------------------------------------
# Pure-Python restatements (small inputs only)
# ---------------------------------------------------------------------------


def legendre_00_py(x: float, lmax: int) -> np.ndarray:
    """(2l+1)/(4pi) P_l(x), l=0..lmax  (legendre.py:10-47)."""
    leg = np.zeros(lmax + 1)
    leg[0] = 1.0
    if lmax >= 1:
        leg[1] = x
    if lmax >= 2:
        leg[2] = 1.5 * x * x - 0.5
    for ell in range(3, lmax + 1):
        leg[ell] = ((2 * ell - 1) * x * leg[ell - 1] - (ell - 1) * leg[ell - 2]) / ell
    for ell in range(lmax + 1):
        leg[ell] *= (2 * ell + 1) / (4 * math.pi)
    return leg


def legendre_22_py(x: float, lmax: int):
    """Normalised P^22_l, F1_l, F2_l  (legendre.py:51-107)."""
    leg = np.zeros(lmax + 1)
    f1 = np.zeros(lmax + 1)
    f2 = np.zeros(lmax + 1)
    if lmax >= 2:
        leg[2] = 3.0
        f1[2] = 6.0 * (1.0 + x * x)
        f2[2] = -12.0 * x
    for ell in range(3, lmax + 1):
        leg[ell] = (x * (2 * ell - 1) * leg[ell - 1] - (ell + 1) * leg[ell - 2]) / (ell - 2)
        f1[ell] = (-(2 * ell - 8 + ell * (ell - 1) * (1.0 - x * x)) * leg[ell]
                   + (2 * ell + 4) * x * leg[ell - 1])
        f2[ell] = 4.0 * (-(ell - 1) * x * leg[ell] + (ell + 2) * leg[ell - 1])
    for ell in range(2, lmax + 1):
        factor2 = 1.0 / ((ell + 2) * (ell + 1) * ell * (ell - 1))
        norm = (2 * ell + 1) / (4 * math.pi) * factor2
        leg[ell] *= norm
        f1[ell] *= norm
        f2[ell] *= norm
    return leg, f1, f2


def legendre_02_py(x: float, lmax: int) -> np.ndarray:
    """Normalised P^02_l  (legendre.py:111-149)."""
    leg = np.zeros(lmax + 1)
    if lmax >= 2:
        leg[2] = 3.0 * (1.0 - x * x)
    for ell in range(3, lmax + 1):
        leg[ell] = (x * (2 * ell - 1) * leg[ell - 1] - (ell + 1) * leg[ell - 2]) / (ell - 2)
    for ell in range(2, lmax + 1):
        leg[ell] *= (2 * ell + 1) / (4 * math.pi) * math.sqrt(
            1.0 / ((ell + 2) * (ell + 1) * ell * (ell - 1)))
    return leg


def legendre_c(kind: str, x: float, lmax: int):
    """The C restatement's scalar recurrences (for pinning against Python)."""
    n = lmax + 1
    if kind == "00":
        out = np.zeros(n)
        lib().oracle_legendre_00(x, _dp(out), n)
        return out
    if kind == "22":
        a, b, c = np.zeros(n), np.zeros(n), np.zeros(n)
        lib().oracle_legendre_22(x, _dp(a), _dp(b), _dp(c), n)
        return a, b, c
    out = np.zeros(n)
    lib().oracle_legendre_02(x, _dp(out), n)
    return out


def rotation_angle(r1, r2):
    """(a12, a21) of ``geometry.py:49-110``."""
    r1 = np.ascontiguousarray(r1, dtype=np.float64)
    r2 = np.ascontiguousarray(r2, dtype=np.float64)
    a = np.zeros(1)
    b = np.zeros(1)
    lib().oracle_rotation_angle(_dp(r1), _dp(r2), _dp(a), _dp(b))
    return float(a[0]), float(b[0])


# ---------------------------------------------------------------------------
# Field layout and spectrum bookkeeping
# ---------------------------------------------------------------------------


@dataclass
class Field:
    spin: int
    labels: list  # ["T"] or ["E", "B"]
    pixels: np.ndarray  # active pixel indices (ascending)
    vec: np.ndarray = None  # (n, 3) unit vectors

    @property
    def n(self) -> int:
        return int(self.pixels.size)

    @property
    def rows(self) -> int:
        return 2 * self.n if self.spin == 2 else self.n

    def auto_labels(self):
        if self.spin == 0:
            a = self.labels[0].upper()
            return [a + a]
        e, b = self.labels[0].upper(), self.labels[1].upper()
        return [e + e, b + b, e + b]

    def cross_labels(self, other: "Field"):
        if self.spin == 0 and other.spin == 0:
            return [self.labels[0].upper() + other.labels[0].upper()]
        if self.spin == 0 and other.spin == 2:
            t = self.labels[0].upper()
            return [t + other.labels[0].upper(), t + other.labels[1].upper()]
        if self.spin == 2 and other.spin == 0:
            t = other.labels[0].upper()
            return [t + self.labels[0].upper(), t + self.labels[1].upper()]
        return [a.upper() + b.upper() for a in self.labels for b in other.labels]


@dataclass
class Layout:
    fields: list
    spectra_map: dict = field(default_factory=dict)  # (i, j, mode) -> label

    def __post_init__(self):
        self.spectra_map = {}
        for i, f in enumerate(self.fields):
            for mode, lab in enumerate(f.auto_labels()):
                self.spectra_map[(i, i, mode)] = lab
        for i in range(len(self.fields)):
            for j in range(i + 1, len(self.fields)):
                for mode, lab in enumerate(self.fields[i].cross_labels(self.fields[j])):
                    self.spectra_map[(i, j, mode)] = lab

    @property
    def labels(self):
        return list(self.spectra_map.values())

    @property
    def n(self) -> int:
        return sum(f.rows for f in self.fields)

    def offsets(self):
        off, out = 0, []
        for f in self.fields:
            out.append(off)
            off += f.rows
        return out

    def spectra_list(self):
        """(fi, fj, mode, label) in the Fisher order; SYMMETRIC drops spin-2
        cross-field CG entries (``spectra_io.py:365-425``)."""
        out = []
        for (fi, fj, mode), lab in self.spectra_map.items():
            if fi != fj and self.fields[fi].spin == 2 and self.fields[fj].spin == 2 and mode == 2:
                continue
            out.append((fi, fj, mode, lab))
        return out


def layout_from_mask(mask: np.ndarray, spins, labels, nside: int, nest: bool = False,
                     pointings=None) -> Layout:
    """Fields from an (npix, ncomp) mask, as ``Core.setup_fields`` builds them
    (``core.py:474-519``: a field's pixels come from its last mask column)."""
    from . import healpix

    mask = np.asarray(mask, dtype=np.float64)
    if mask.ndim == 1:
        mask = mask[:, None]
    fields, counter = [], 0
    for spin in spins:
        if spin == 0:
            labs = [labels[counter]]
            counter += 1
        else:
            labs = [labels[counter], labels[counter + 1]]
            counter += 2
        pix = np.nonzero(mask[:, counter - 1] > 0.5)[0]
        f = Field(spin, labs, pix)
        f.vec = pointings(f.pixels) if pointings is not None else healpix.pointings(nside, pix, nest)
        fields.append(f)
    return Layout(fields)


def _get_cls(layout: Layout, cls: dict, i: int, j: int, mode: int) -> np.ndarray:
    lab = layout.spectra_map[(i, j, mode)]
    return np.ascontiguousarray(cls[lab], dtype=np.float64)


# ---------------------------------------------------------------------------
# Signal matrix and binned derivative
# ---------------------------------------------------------------------------


def _block_ptr(S: np.ndarray, r0: int, c0: int) -> int:
    return S.ctypes.data + 8 * (r0 + c0 * S.shape[0])


def signal_matrix(layout: Layout, cls: dict, lmax: int) -> np.ndarray:
    """S over all field pairs, restating ``compute_signal_matrix``.

    ``cls`` maps labels to l-indexed, beam-smoothed C_l of length >= lmax+1.
    """
    L = lib()
    n = layout.n
    S = np.zeros((n, n), order="F")
    ld = n
    offs = layout.offsets()
    n_ell = lmax + 1
    for i, fi in enumerate(layout.fields):
        for j, fj in enumerate(layout.fields):
            if j < i:
                continue
            ri, rj = offs[i], offs[j]
            if fi.spin == 0 and fj.spin == 0:
                cl = _get_cls(layout, cls, i, j, 0)[:n_ell].copy()
                if i == j:
                    L.oracle_compute_00(_dp(cl), n_ell, _block_ptr(S, ri, ri), ld,
                                        _dp(fi.vec), _dp(fj.vec), fi.n, 0)
                else:
                    if fi.n != fj.n:
                        raise ValueError("cross spin-0 fields must share the pixel set")
                    L.oracle_compute_00(_dp(cl), n_ell, _block_ptr(S, rj, ri), ld,
                                        _dp(fj.vec), _dp(fi.vec), fi.n, 1)
            elif fi.spin == 2 and fj.spin == 2:
                if i != j:
                    raise NotImplementedError("Cross-correlation for spin-2 fields not implemented yet.")
                c11 = _get_cls(layout, cls, i, i, 0)[:n_ell].copy()
                c22 = _get_cls(layout, cls, i, i, 1)[:n_ell].copy()
                c12 = _get_cls(layout, cls, i, i, 2)[:n_ell].copy()
                L.oracle_compute_22(_dp(c11), _dp(c22), _dp(c12), n_ell, _block_ptr(S, rj, ri), ld,
                                    _dp(fi.vec), _dp(fj.vec), fi.n)
            else:
                c12 = _get_cls(layout, cls, min(i, j), max(i, j), 0)[:n_ell].copy()
                c13 = _get_cls(layout, cls, min(i, j), max(i, j), 1)[:n_ell].copy()
                if fi.spin == 0:
                    L.oracle_compute_02(_dp(c12), _dp(c13), n_ell, _block_ptr(S, rj, ri), ld,
                                        _dp(fi.vec), _dp(fj.vec), fi.n, fj.n)
                else:
                    L.oracle_compute_02(_dp(c13), _dp(c12), n_ell, _block_ptr(S, ri, rj), ld,
                                        _dp(fj.vec), _dp(fi.vec), fj.n, fi.n)
    L.oracle_copy_lower_to_upper(S.ctypes.data, n, ld)
    return S


def key_mode(layout: Layout, fi: int, fj: int, mode: int) -> int:
    """Legacy derivative mode of a spectrum (``spectrum_key.py:145``): the
    auto spin-2 modes are EE=0, BB=1, EB=2; spin-0 x spin-2 TE=0, TB=1."""
    return mode


def binned_derivative(layout: Layout, fi: int, fj: int, mode: int, weights: np.ndarray) -> np.ndarray:
    """dC_b for one spectrum with l-indexed in-bin weights (``pixel.py:802-950``)."""
    L = lib()
    n = layout.n
    dS = np.zeros((n, n), order="F")
    offs = layout.offsets()
    f_i, f_j = layout.fields[fi], layout.fields[fj]
    ri, rj = offs[fi], offs[fj]
    w = np.ascontiguousarray(weights, dtype=np.float64)
    n_ell = w.size
    zeros = np.zeros(n_ell)
    if f_i.spin == 0 and f_j.spin == 0:
        sub = 0 if fi == fj else 1
        L.oracle_compute_00(_dp(w), n_ell, _block_ptr(dS, rj, ri), n, _dp(f_i.vec), _dp(f_j.vec),
                            f_i.n, sub)
        blk = dS[rj:rj + f_j.rows, ri:ri + f_i.rows]
        if sub == 0:
            blk[:] = np.tril(blk) + np.tril(blk, -1).T
        else:
            dS[ri:ri + f_i.rows, rj:rj + f_j.rows] = blk.T
    elif f_i.spin == 2 and f_j.spin == 2 and fi == fj:
        c11, c22, c12 = [(w if mode == m else zeros) for m in (0, 1, 2)]
        L.oracle_compute_22(_dp(c11), _dp(c22), _dp(c12), n_ell, _block_ptr(dS, rj, ri), n,
                            _dp(f_i.vec), _dp(f_j.vec), f_i.n)
        blk = dS[rj:rj + f_j.rows, ri:ri + f_i.rows]
        blk[:] = np.tril(blk) + np.tril(blk, -1).T
    elif {f_i.spin, f_j.spin} == {0, 2}:
        c12 = w if mode == 0 else zeros
        c13 = w if mode == 1 else zeros
        if f_i.spin == 0:
            L.oracle_compute_02(_dp(c12), _dp(c13), n_ell, _block_ptr(dS, rj, ri), n,
                                _dp(f_i.vec), _dp(f_j.vec), f_i.n, f_j.n)
            dS[ri:ri + f_i.rows, rj:rj + f_j.rows] = dS[rj:rj + f_j.rows, ri:ri + f_i.rows].T
        else:
            L.oracle_compute_02(_dp(c12), _dp(c13), n_ell, _block_ptr(dS, ri, rj), n,
                                _dp(f_j.vec), _dp(f_i.vec), f_j.n, f_i.n)
            dS[rj:rj + f_j.rows, ri:ri + f_i.rows] = dS[ri:ri + f_i.rows, rj:rj + f_j.rows].T
    else:
        raise NotImplementedError("spin-2 x spin-2 cross-field derivative")
    return dS


def matrix_trace(A: np.ndarray, B: np.ndarray) -> float:
    """Tr(AB) with fixed-order block partials (``linalg.py:66-98``)."""
    A = np.ascontiguousarray(A)
    B = np.ascontiguousarray(B)
    return float(lib().oracle_matrix_trace(_dp(A), _dp(B), A.shape[0]))
