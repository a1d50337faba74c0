"""HEALPix pixel centres, RING and NESTED (oracle restatement).

healpy is not available, so the oracle owns ``pix2ang``.  The formulas are
those of Gorski et al. 2005 (ApJ 622, 759), in the form used by the HEALPix
C++ library: polar-cap rings use ``z = 1 - i^2 * 4/npix`` with
``theta = atan2(sqrt(t(2-t)), z)`` near the poles, equatorial rings use
``z = (2 nside - i) * 2 nside * 4/npix``.

The reference calls ``hp.pix2ang`` at ``src/cosmoforge.cosmocore/cosmocore/
signal_kernels.py:137`` and turns the angles into unit vectors with a scalar
sin/cos loop at ``signal_kernels.py:144-161``; ``unit_vectors`` below keeps
that scalar loop (Python's ``math`` is the platform libm, as Numba's scalar
path is; squares are products, as Numba emits for ``x**2``).

Pinned by the reference's ``geometry.dat`` fixtures (RING, nside 8) and the
NESTED golden signal matrix (nside 4).
"""

from __future__ import annotations

import math

import numpy as np

_JRLL = (2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4)
_JPLL = (1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7)


def _isqrt(v: int) -> int:
    return math.isqrt(v)


def _compress_bits(v: int) -> int:
    """Keep the even bits of ``v`` and pack them (NESTED bit de-interleave)."""
    out = 0
    bit = 0
    while v:
        out |= (v & 1) << bit
        v >>= 2
        bit += 1
    return out


def pix2zphi(nside: int, pix: int, nest: bool) -> tuple[float, float, float | None]:
    """Return (z, phi, sin_theta or None) of the centre of ``pix``."""
    npix = 12 * nside * nside
    fact2 = 4.0 / npix
    fact1 = (nside << 1) * fact2
    sth = None
    if not nest:
        ncap = 2 * nside * (nside - 1)
        if pix < ncap:
            iring = (1 + _isqrt(1 + 2 * pix)) >> 1
            iphi = (pix + 1) - 2 * iring * (iring - 1)
            tmp = (iring * iring) * fact2
            z = 1.0 - tmp
            if z > 0.99:
                sth = math.sqrt(tmp * (2.0 - tmp))
            phi = (iphi - 0.5) * (0.5 * math.pi) / iring
        elif pix < npix - ncap:
            nl4 = 4 * nside
            ip = pix - ncap
            tmp = ip // nl4
            iring = tmp + nside
            iphi = ip - nl4 * tmp + 1
            fodd = 1.0 if ((iring + nside) & 1) else 0.5
            z = (2 * nside - iring) * fact1
            phi = (iphi - fodd) * math.pi * 0.75 * fact1
        else:
            ip = npix - pix
            iring = (1 + _isqrt(2 * ip - 1)) >> 1
            iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1))
            tmp = (iring * iring) * fact2
            z = tmp - 1.0
            if z < -0.99:
                sth = math.sqrt(tmp * (2.0 - tmp))
            phi = (iphi - 0.5) * (0.5 * math.pi) / iring
        return z, phi, sth
    order = nside.bit_length() - 1
    npface = nside * nside
    face = pix // npface
    ipf = pix & (npface - 1)
    ix = _compress_bits(ipf)
    iy = _compress_bits(ipf >> 1)
    jr = (_JRLL[face] << order) - ix - iy - 1
    if jr < nside:
        nr = jr
        tmp = (nr * nr) * fact2
        z = 1.0 - tmp
        if z > 0.99:
            sth = math.sqrt(tmp * (2.0 - tmp))
    elif jr > 3 * nside:
        nr = 4 * nside - jr
        tmp = (nr * nr) * fact2
        z = tmp - 1.0
        if z < -0.99:
            sth = math.sqrt(tmp * (2.0 - tmp))
    else:
        nr = nside
        z = (2 * nside - jr) * fact1
    t = _JPLL[face] * nr + ix - iy
    if t < 0:
        t += 8 * nr
    if nr == nside:
        phi = 0.75 * (0.5 * math.pi) * t * fact1
    else:
        phi = (0.5 * (0.5 * math.pi) * t) / nr
    return z, phi, sth


def pix2ang(nside: int, pixels, nest: bool = False):
    """(theta, phi) arrays of pixel centres, as ``hp.pix2ang`` returns them."""
    pixels = np.asarray(pixels, dtype=np.int64).ravel()
    theta = np.empty(pixels.size)
    phi = np.empty(pixels.size)
    for k, p in enumerate(pixels.tolist()):
        z, ph, sth = pix2zphi(nside, int(p), nest)
        theta[k] = math.atan2(sth, z) if sth is not None else math.acos(z)
        phi[k] = ph
    return theta, phi


def unit_vectors(theta, phi):
    """Scalar-libm unit vectors, restating ``signal_kernels.py:144-161``."""
    out = np.empty((len(theta), 3))
    for i in range(len(theta)):
        st = math.sin(theta[i])
        x = st * math.cos(phi[i])
        y = st * math.sin(phi[i])
        z = math.cos(theta[i])
        # Numba lowers ``x**2`` with a literal integer exponent to x*x; CPython
        # would call pow(x, 2.0), which can differ by an ulp.
        norm = math.sqrt(x * x + y * y + z * z)
        out[i, 0] = x / norm
        out[i, 1] = y / norm
        out[i, 2] = z / norm
    return out


def pointings(nside: int, pixels, nest: bool = False) -> np.ndarray:
    """(n, 3) unit vectors of active pixels (``compute_pointings``, :113-141)."""
    theta, phi = pix2ang(nside, pixels, nest)
    return unit_vectors(theta, phi)
