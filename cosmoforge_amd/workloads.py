"""Synthetic inputs of the benchmark configurations (host side, NumPy).

* ``config(0)``: nside 16, T, lmax 47, |z| >= sin 20 deg mask, TT = 2 pi 1e3 /
  (l(l+1)) uK^2, Gaussian 8 deg beam, white noise 1 uK^2.
* ``config(1)``: nside 16 TQU, lmax 47; EE = 0.05 TT + 0.02 exp(-(l-4)^2/8),
  TE = 0.3 sqrt(TT EE), BB = 0.01 EE, EB = TB = 0; noise 1 (T), 2 (Q/U).
* ``config(2)``: as config 1 at nside 32, lmax 95, 4 deg beam, noise 0.5 / 1.
* ``config(3)``: nside 16 QU discs (centres uniform on the sphere, radius
  U[25, 60] deg) with (tau, r) points: EE(tau) = 0.05 TT + 0.02 (tau/0.06)^2
  exp(-(l-4)^2/8), BB = r 2 pi 5e-3 / (l(l+1)), EB = 0, noise 1.

Seeds are ``numpy.random.default_rng(20261203 + config)``.  lmax_signal is
the reference default 4 nside.  C_0 = C_1 = 0.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .core import gaussian_beam

SEED0 = 20261203


def pixel_vectors(nside: int) -> np.ndarray:
    """Unit vectors of every RING pixel (platform libm, host)."""
    lib = _lib.load()
    pix = np.arange(12 * nside * nside, dtype=np.int64)
    vec = np.empty((pix.size, 3))
    rc = lib.cfmi_pointings_host(nside, 0, pix.ctypes.data, pix.size, vec.ctypes.data)
    if rc != 0:
        raise ValueError("pointings_host failed")
    return vec


def galactic_cut(nside: int, deg: float = 20.0) -> np.ndarray:
    """1 where |z| >= sin(deg), 0 elsewhere."""
    z = pixel_vectors(nside)[:, 2]
    return (np.abs(z) >= math.sin(math.radians(deg))).astype(np.float64)


def tt_spectrum(lmax: int) -> np.ndarray:
    ell = np.arange(lmax + 1, dtype=np.float64)
    tt = np.zeros(lmax + 1)
    tt[2:] = 2 * np.pi * 1e3 / (ell[2:] * (ell[2:] + 1))
    return tt


def pol_spectra(lmax: int, tau: float = 0.06, r: float | None = None) -> dict:
    ell = np.arange(lmax + 1, dtype=np.float64)
    tt = tt_spectrum(lmax)
    ee = 0.05 * tt + 0.02 * (tau / 0.06) ** 2 * np.exp(-((ell - 4.0) ** 2) / 8.0)
    ee[:2] = 0.0
    if r is None:
        bb = 0.01 * ee
    else:
        bb = np.zeros(lmax + 1)
        bb[2:] = r * 2 * np.pi * 5e-3 / (ell[2:] * (ell[2:] + 1))
    te = 0.3 * np.sqrt(tt * ee)
    z = np.zeros(lmax + 1)
    return {"TT": tt, "EE": ee, "BB": bb, "TE": te, "TB": z.copy(), "EB": z.copy()}


@dataclass
class Config:
    index: int
    name: str
    params: dict
    mask: np.ndarray
    cls: dict
    beam: np.ndarray
    noise_diag: np.ndarray
    extra: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.noise_diag.size)

    @property
    def noise(self) -> np.ndarray:
        return np.diag(self.noise_diag)


def config(index: int) -> Config:
    if index in (0, 1, 2):
        nside = 32 if index == 2 else 16
        lmax = 95 if index == 2 else 47
        L = 4 * nside
        fwhm = 4.0 if index == 2 else 8.0
        cut = galactic_cut(nside)
        beam = gaussian_beam(math.radians(fwhm), L)
        if index == 0:
            spins, labels = [0], ["T"]
            mask = cut[:, None]
            cls = {"TT": tt_spectrum(L)}
            nd = np.ones(int(cut.sum()))
        else:
            spins, labels = [0, 2], ["T", "E", "B"]
            mask = np.stack([cut, cut, cut], axis=1)
            cls = pol_spectra(L)
            nt, nq = (0.5, 1.0) if index == 2 else (1.0, 2.0)
            k = int(cut.sum())
            nd = np.concatenate([np.full(k, nt), np.full(2 * k, nq)])
        params = dict(nside=nside, spins=spins, labels=labels, lmin=2, lmax=lmax, lmax_signal=L, nsims=1,
                      ordering="RING")
        names = {0: "nside16_T_lmax47", 1: "nside16_TQU_lmax47", 2: "nside32_TQU_lmax95"}
        return Config(index, names[index], params, mask, cls, beam, nd)
    if index == 3:
        nside, lmax = 16, 47
        L = 4 * nside
        rng = np.random.default_rng(SEED0 + 3)
        vec = pixel_vectors(nside)
        npatch = 64
        u = rng.uniform(-1.0, 1.0, npatch)
        ph = rng.uniform(0.0, 2 * np.pi, npatch)
        centres = np.stack([np.sqrt(1 - u * u) * np.cos(ph), np.sqrt(1 - u * u) * np.sin(ph), u], axis=1)
        radii = rng.uniform(25.0, 60.0, npatch)
        masks = [(vec @ c >= math.cos(math.radians(rd))).astype(np.float64) for c, rd in zip(centres, radii)]
        npts = 256
        tau = rng.uniform(0.03, 0.09, (npatch, npts))
        r = rng.uniform(0.0, 0.1, (npatch, npts))
        beam = np.ones((3, L + 1))
        params = dict(nside=nside, spins=[2], labels=["E", "B"], lmin=2, lmax=lmax, lmax_signal=L, nsims=1,
                      ordering="RING", smoothing_type="none")
        return Config(3, "nside16_QU_64discs_256pts", params, None, pol_spectra(L), beam, np.zeros(0),
                      extra=dict(masks=masks, tau=tau, r=r, centres=centres, radii=radii))
    raise ValueError(f"unknown config {index}")


def qu_cls(lmax: int, tau: float, r: float) -> dict:
    """EE/BB/EB spectra of one config-3 parameter point."""
    s = pol_spectra(lmax, tau=tau, r=r)
    return {"EE": s["EE"], "BB": s["BB"], "EB": s["EB"]}
