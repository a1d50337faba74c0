"""Shared input handling of Fisher / Spectra / PICSLike (host side).

Mirrors the reference's ``Core`` seams for in-memory inputs
(``src/cosmoforge.cosmocore/cosmocore/core.py:224-766``, ADR-0017):

* ``params``: dict, YAML path or an object with attributes; defaults follow
  ``settings.py:178-253`` (file paths included); the multipole chain
  ``max(lmin_signal) <= lmin <= lmax <= lmax_signal`` clamps ``lmax``
  (``core.py:378-401``);
* ``mask`` (npix,) or (npix, ncomp), active = > 0.5, Q/U columns must agree
  (``core.py:403-472``);
* ``noise_cov1``: reduced (n, n) float64 in active-row order, used as-is
  (a host array, or a device tensor that stays where it is);
* ``cls_data``: {label: physical C_l}, l-indexed, every spectrum label
  required (``spectra_io.py:257-326``);
* ``beam``: (>= 3, >= lmax_signal + 1) T/E/B windows; wins over
  ``smoothing_type`` (``beam.py:128-221``).

An input given as an array wins; without one the file named by ``params``
is read (``maskfile``, ``covmatfile1`` with ``load_reduced``,
``inputclfile`` with ``input_convention``, ``beam_file`` when
``smoothing_type == "file"``, ``inputmapfile1``) through ``in_out``, as the
reference's file adapters do (``core.py:403-472,638-714``, ``fields.py:915-940``).
"""

from __future__ import annotations

import math
import os
from types import SimpleNamespace

import numpy as np

from . import in_out
from .layout import Layout, build_layout

DEFAULTS = dict(
    nside=16, spins=[0, 2], labels=["T", "E", "B"], physical_labels=None, feedback=1,
    do_cross=False, smoothing_type="cosine_legacy", apply_pixwin=True, fwhmarcmin=440.0, lmax=64,
    delta_ell=1, bin_lmins=None, bin_lmaxs=None, ordering="RING", nsims=None, remove_nb=True,
    input_convention="Cl", output_convention="Cl", lmin_signal=2, lmax_signal=None, lmin=2,
    parameters={}, fiducialfile=None,
    maskfile="inputs/mask.fits", covmatfile1="inputs/NCVM1.bin", covmatfile2="inputs/NCVM2.bin",
    inputclfile="inputs/cls.dat", beam_file="inputs/beam.fits", inputmapfile1="", inputmapfile2="",
    load_reduced=False, output_geometry_file=None, outnoisecovmat1=None, outnoisecovmat2=None,
    outinvcovmatfile1=None, outinvcovmatfile2=None, outfilefisher=None, outcovmatfile=None, outerrfile=None,
    root_dir="inputs", root_filename="theory_spectra",
)


def read_params(params) -> SimpleNamespace:
    """Resolve ``params`` (dict / YAML path / object) over the defaults."""
    values = dict(DEFAULTS)
    if params is None:
        pass
    elif isinstance(params, dict):
        values.update(params)
    elif isinstance(params, (str, os.PathLike)):
        import yaml

        with open(params) as fh:
            values.update(yaml.safe_load(fh) or {})
    elif hasattr(params, "__dict__"):
        values.update({k: v for k, v in vars(params).items() if not k.startswith("_")})
    else:
        raise TypeError("params must be an instance of InputParams, a string with the path to a "
                        "parameter file, or a dictionary with parameters.")
    p = SimpleNamespace(**values)
    if isinstance(p.labels, str):
        p.labels = list(p.labels)
    if p.physical_labels is None:
        p.physical_labels = list(p.labels)
    elif isinstance(p.physical_labels, str):
        p.physical_labels = list(p.physical_labels)
    p.nfields = sum(1 if s == 0 else 2 for s in p.spins)
    ls = p.lmin_signal
    ls = [ls] * len(p.spins) if isinstance(ls, int) else list(ls)
    if len(ls) != len(p.spins):
        raise ValueError(f"len(lmin_signal)={len(ls)} != len(spins)={len(p.spins)}")
    for i, (lm, s) in enumerate(zip(ls, p.spins)):
        if lm < abs(s):
            raise ValueError(f"lmin_signal[{i}]={lm} invalid for spin-{s} component at index {i}; "
                             f"minimum allowed is {abs(s)}")
    p.lmin_signal = ls
    lmax_signal = p.lmax_signal if p.lmax_signal is not None else 4 * p.nside
    if p.lmin < max(ls):
        raise ValueError(f"lmin={p.lmin} < max(lmin_signal)={max(ls)}")
    if p.lmax is not None and p.lmax > lmax_signal:
        p.lmax = lmax_signal
    if p.lmax is not None and p.lmax < p.lmin:
        raise ValueError(f"lmax={p.lmax} < lmin={p.lmin}")
    p.do_cross = bool(p.do_cross)
    p.smoothing_type = normalize_smoothing_type(p.smoothing_type)
    return p


SMOOTHING_TYPES = ("none", "gaussian", "cosine_legacy", "cosine_npipe", "file")
# legacy integer codes of ``smoothing_type`` (``settings.py:353-380``)
SMOOTHING_CODES = {0: "none", 1: "gaussian", 2: "cosine_legacy", 3: "file"}


def normalize_smoothing_type(value) -> str:
    """A canonical ``smoothing_type``: the names above in any case, the
    deprecated alias "cosine" (= "cosine_legacy") and the deprecated integer
    codes 0-3, each alias with a DeprecationWarning (``settings.py:353-380``)."""
    import warnings

    if isinstance(value, bool):
        raise ValueError(f"Unknown smoothing_type {value!r}")
    if isinstance(value, (int, float, np.integer, np.floating)):
        code = int(value)
        if code != value or code not in SMOOTHING_CODES:
            raise ValueError(f"Unknown smoothing_type={value!r}; integer codes are "
                             + ", ".join(f"{k}={v}" for k, v in SMOOTHING_CODES.items()))
        warnings.warn(f"integer smoothing_type={code} is deprecated; use '{SMOOTHING_CODES[code]}'",
                      DeprecationWarning, stacklevel=3)
        return SMOOTHING_CODES[code]
    if isinstance(value, str):
        key = value.strip().lower()
        if key == "cosine":
            warnings.warn("smoothing_type='cosine' is deprecated; use 'cosine_legacy' or 'cosine_npipe'",
                          DeprecationWarning, stacklevel=3)
            return "cosine_legacy"
        if key in SMOOTHING_TYPES:
            return key
    raise ValueError(f"Unknown smoothing_type {value!r}; expected one of {', '.join(SMOOTHING_TYPES)}")


def resolve_basis(basis):
    """``basis=`` semantics (ADR-0018).  ``None``/``"auto"``/``False``/
    ``"pixel"`` select the pixel-space estimator; ``"harmonic"`` the same
    estimator with its traces contracted over harmonic modes
    (``harmonic.HarmonicContraction``), uncompressed."""
    if basis is None or basis is False:
        return "pixel"
    if isinstance(basis, str):
        b = basis.lower()
        if b in ("auto", "pixel"):
            return "pixel"
        if b == "harmonic":
            return "harmonic"
        raise ValueError(f"unknown basis {basis!r}")
    if isinstance(basis, dict):
        if basis.get("epsilon") or basis.get("mode_fraction") or basis.get("compression_target"):
            raise NotImplementedError("compressed bases are not part of this build")
        return resolve_basis(basis.get("method", "pixel"))
    raise TypeError(f"basis must be None, False, str or dict; got {type(basis).__name__}")


def resolve_mask(p, mask) -> np.ndarray:
    if mask is None:
        nest = str(p.ordering).upper() == "NESTED"
        mask = in_out.read_mask(p.maskfile, np.empty((p.nfields, 0)), nest=nest)
    mask = np.asarray(mask, dtype=np.float64)
    if mask.ndim == 1:
        mask = mask[:, None]
    npix = 12 * p.nside**2
    if mask.shape[0] != npix:
        raise ValueError(f"mask has npix={mask.shape[0]}, expected {npix} for nside={p.nside}")
    if mask.shape[1] != p.nfields:
        raise ValueError(f"mask has {mask.shape[1]} column(s), expected nfields={p.nfields}")
    col = 0
    for s in p.spins:
        if s != 0:
            if not np.array_equal(mask[:, col] > 0.5, mask[:, col + 1] > 0.5):
                raise ValueError(f"mask columns {col} and {col + 1} differ, but they are the two "
                                 "components of one spin-2 field and must share a single pixel set")
            col += 2
        else:
            col += 1
    return mask


def coswinbeam(nside: int, ell1=None, ell2=None) -> np.ndarray:
    """Raised-cosine window b(l) (``beam.py:27-74``)."""
    ell1 = nside if ell1 is None else ell1
    ell2 = 3 * nside if ell2 is None else ell2
    if not (0 <= ell1 < ell2 <= 4 * nside):
        raise ValueError("coswinbeam transition multipoles must satisfy 0 <= ell1 < ell2 <= 4*nside")
    b = np.zeros(4 * nside + 1)
    b[: ell1 + 1] = 1.0
    ell = np.arange(ell1 + 1, ell2 + 1)
    b[ell1 + 1: ell2 + 1] = 0.5 * (1.0 + np.cos((ell - ell1) * np.pi / (ell2 - ell1)))
    return b


def gaussian_beam(fwhm_rad: float, lmax: int) -> np.ndarray:
    """(3, lmax+1) Gaussian T/E/B windows, E and B with exp(2 sigma^2)."""
    sigma = fwhm_rad / math.sqrt(8.0 * math.log(2.0))
    ell = np.arange(lmax + 1)
    g = np.exp(-0.5 * ell * (ell + 1) * sigma**2)
    pol = math.exp(2.0 * sigma**2)
    return np.vstack([g, g * pol, g * pol])


def resolve_beam(p, beam, lmax: int) -> np.ndarray:
    if beam is not None:
        b = np.asarray(beam, dtype=np.float64)
        if b.shape[0] < 3:
            raise ValueError(f"injected beam must have at least 3 rows (T, E, B), got {b.shape[0]}")
        if b.shape[1] < lmax + 1:
            raise ValueError(f"injected beam has {b.shape[1]} multipoles, need lmax_signal+1={lmax + 1}")
        return b[:3, : lmax + 1].copy()
    st = p.smoothing_type
    if st == "none":
        return np.ones((3, lmax + 1))
    if st == "gaussian":
        return gaussian_beam(math.radians(p.fwhmarcmin / 60.0), lmax)
    if st in ("cosine_legacy", "cosine_npipe"):
        b = coswinbeam(p.nside, ell1=p.nside if st == "cosine_legacy" else 1, ell2=3 * p.nside)
        if b.size < lmax + 1:
            raise ValueError("cosine beam shorter than lmax_signal + 1")
        b = b[: lmax + 1]
        return np.vstack([b, b, b])
    if st == "file":
        bls = in_out.read_cl_fits(p.beam_file)
        if bls.shape[0] < 3:
            raise ValueError(f"Beam file must have at least 3 columns (T, E, B), got {bls.shape[0]}")
        if bls.shape[1] < lmax + 1:
            raise ValueError(f"Beam shape mismatch: expected (3, {lmax + 1}), got (3, {bls.shape[1]})")
        return bls[:3, : lmax + 1].copy()
    raise ValueError(f"Unknown smoothtype='{st}'")


def beam_by_label(layout: Layout, beam3: np.ndarray) -> dict:
    """Field-map label -> window (``beam.py:261-286``)."""
    named = {"T": 0, "E": 1, "B": 2}
    out = {}
    for f in layout.fields:
        if f.spin == 0:
            out[f.labels[0]] = beam3[named.get(f.labels[0], 0)]
        else:
            out[f.labels[0]] = beam3[1]
            out[f.labels[1]] = beam3[2]
    return out


def smoothing_factors(layout: Layout, beams: dict, lmax: int) -> dict:
    out = {}
    for (i, j, _), lab in layout.spectra_map.items():
        a, b = layout.split_label(lab, i, j)
        if beams[a].size < lmax + 1 or beams[b].size < lmax + 1:
            raise ValueError(f"Beam for spectrum '{lab}' is too short: expected at least {lmax + 1} multipoles")
        out[lab] = beams[a][: lmax + 1] * beams[b][: lmax + 1]
    return out


def resolve_cls(layout: Layout, cls_data, lmax: int, p=None) -> dict:
    if cls_data is None:
        if p is None:
            raise ValueError("cls_data= is required without params")
        cls_data = in_out.readcl(p.inputclfile, p, lmax=lmax)
    if isinstance(cls_data, np.ndarray):
        if cls_data.ndim != 2 or cls_data.shape[1] != len(layout.labels):
            raise ValueError(f"Expected {len(layout.labels)} spectra columns, got {cls_data.shape}")
        cls_data = {lab: cls_data[:, k] for k, lab in enumerate(layout.labels)}
    out = {}
    for lab in layout.labels:
        if lab not in cls_data:
            raise ValueError(f"Missing power spectrum for {lab}")
        arr = np.asarray(cls_data[lab], dtype=np.float64)[: lmax + 1]
        if arr.size < lmax + 1:
            raise ValueError(f"power spectrum {lab} has {arr.size} multipoles, need {lmax + 1}")
        out[lab] = arr.copy()
    return out


def resolve_noise(n: int, noise_cov, p=None, layout: Layout | None = None, which: int = 1):
    """The (n, n) noise covariance: a host array, or a torch tensor already on
    the device (kept there, so no host copy or PCIe transfer per solve).
    Without one, ``p.covmatfile{which}`` is read: already reduced when
    ``p.load_reduced``, else reduced from the full-sky file
    (``core.py:638-680``)."""
    name = f"noise_cov{which}"
    if noise_cov is None:
        if p is None or layout is None:
            raise ValueError(f"{name}= is required without params")
        path = getattr(p, f"covmatfile{which}")
        C = np.empty((n, n), dtype=np.float64)
        if p.load_reduced:
            return in_out.read_covmat_reduced(path, C)
        npix = 12 * p.nside**2
        return in_out.read_covmat(path, npix, p.nfields, layout.active_index(npix), C)
    if _is_device_tensor(noise_cov):
        N = noise_cov
        if tuple(N.shape) != (n, n):
            raise ValueError(f"{name} has shape {tuple(N.shape)}, expected ({n}, {n}) for the active-pixel count")
        return N
    N = np.asarray(noise_cov)
    if N.shape != (n, n):
        raise ValueError(f"{name} has shape {N.shape}, expected ({n}, {n}) for the active-pixel count")
    return N


def _is_device_tensor(x) -> bool:
    try:
        import torch
    except ImportError:
        return False
    return isinstance(x, torch.Tensor) and x.is_cuda


def resolve_maps(n: int, nsims, maps, p=None, layout: Layout | None = None, which: int = 1) -> np.ndarray:
    """(n, nsims) maps: the array given, or ``p.inputmapfile{which}`` read for
    the per-component active pixels and ``p.physical_labels``
    (``core.py:682-714``)."""
    if maps is None:
        if p is None or layout is None:
            raise ValueError(f"maps{which}= is required without params")
        if not nsims:
            raise ValueError(f"nsims must be set to read maps from inputmapfile{which}")
        m = np.empty((n, int(nsims)), dtype=np.float64)
        pixact = [f.pixels for f in layout.fields for _ in range(2 if f.spin == 2 else 1)]
        return in_out.read_maps(m, getattr(p, f"inputmapfile{which}"), pixact, p.physical_labels)
    m = np.asarray(maps, dtype=np.float64)
    if m.ndim == 1:
        m = m[:, None]
    if nsims is None:
        nsims = m.shape[1]
    if m.shape != (n, nsims):
        raise ValueError(f"injected maps have shape {m.shape}, expected ({n}, {nsims}) for the "
                         "active-pixel count and nsims")
    return m


def bins_from_params(p):
    """[(l_min, l_max)] of the bins (``bins.py:80-140``): explicit
    ``bin_lmins``/``bin_lmaxs`` (bins below l = 2 dropped, sorted, no
    overlap), else uniform ``delta_ell`` bins from ``lmin`` (the floor)."""
    if p.bin_lmins is not None and p.bin_lmaxs is not None:
        if len(p.bin_lmins) != len(p.bin_lmaxs):
            raise ValueError(f"lmins and lmaxs must have the same length, got {len(p.bin_lmins)} and "
                             f"{len(p.bin_lmaxs)}")
        lo, hi = np.asarray(p.bin_lmins, int), np.asarray(p.bin_lmaxs, int)
        floor = 2
    else:
        nb = (p.lmax - p.lmin + 1) // p.delta_ell
        lo = p.lmin + np.arange(nb) * p.delta_ell
        hi = lo + p.delta_ell - 1
        floor = p.lmin
    keep = (hi >= floor) & (lo >= floor)
    lo, hi = lo[keep], hi[keep]
    if lo.size == 0:
        raise ValueError(f"No valid bins (all bins below lmin_floor={floor})")
    for i, (a, b) in enumerate(zip(lo, hi)):
        if a > b:
            raise ValueError(f"Bin {i}: lmin={a} > lmax={b}")
    order = np.argsort(lo, kind="stable")
    lo, hi = lo[order], hi[order]
    for i in range(lo.size - 1):
        if lo[i + 1] <= hi[i]:
            raise ValueError(f"Bins {i} and {i + 1} overlap: [{lo[i]}, {hi[i]}] and [{lo[i + 1]}, {hi[i + 1]}]")
    return list(zip(lo.tolist(), hi.tolist()))


def shape_weights(lmax: int, convention: str) -> np.ndarray:
    key = str(convention).strip().lower()
    if key not in ("cl", "dl"):
        raise ValueError(f"Unknown shape convention '{convention}'. Must be 'Cl' or 'Dl'.")
    if key == "cl":
        return np.ones(lmax + 1)
    ell = np.arange(lmax + 1, dtype=np.float64)
    w = np.zeros(lmax + 1)
    w[1:] = 2 * np.pi / (ell[1:] * (ell[1:] + 1))
    return w


class Inputs:
    """Resolved in-memory inputs shared by the three estimators."""

    def __init__(self, params, mask, noise_cov1, cls_data, beam, noise_cov2=None):
        self.params = read_params(params)
        p = self.params
        self.lmax_signal = p.lmax_signal if p.lmax_signal is not None else 4 * p.nside
        self.mask = resolve_mask(p, mask)
        self.layout = build_layout(self.mask, p.spins, p.labels)
        self.n = self.layout.n
        self.noise = resolve_noise(self.n, noise_cov1, p, self.layout)
        in_out.write_covmat_reduced(p.outnoisecovmat1, self.noise)
        # the second noise covariance is read only for a cross analysis (core.py:614-622)
        self.noise2 = None
        if p.do_cross:
            self.noise2 = resolve_noise(self.n, noise_cov2, p, self.layout, which=2)
            in_out.write_covmat_reduced(p.outnoisecovmat2, self.noise2)
        self.cls = resolve_cls(self.layout, cls_data, self.lmax_signal, p)
        self.beam3 = resolve_beam(p, beam, self.lmax_signal)
        self.beams = beam_by_label(self.layout, self.beam3)
        self.smoothing = smoothing_factors(self.layout, self.beams, self.lmax_signal)

    def smoothed_cls(self, cls=None) -> dict:
        cls = self.cls if cls is None else cls
        return {lab: cls[lab][: self.lmax_signal + 1] * self.smoothing[lab] for lab in self.layout.labels}
