"""Real spin-0 / spin-2 harmonic modes on pixel centres (oracle, test
infrastructure only).

Restates the mode operator V of ``src/cosmoforge.cosmocore/cosmocore/basis/
harmonic_basis.py:19-117`` (``_fill_V_spin0``, ``_fill_V_spin2``) with the
Wigner small-d convention of ``basics/wigner.py:12-160`` (d^l_{m,s} with
d^m_{m,s} = +sqrt((2m)! / ((m+s)! (m-s)!)) cos^{m+s}(t/2) sin^{m-s}(t/2)),
evaluated here through Jacobi polynomials (Varshalovich et al. 1988, 4.3.4;
checked against the finite factorial sum at low l) rather than the
l-recurrence of the device kernel, so that it checks one.

Modes are ordered l-major: for each l in [lmin, lmax], m = 0, then for
m = 1..l the cos and the sin mode.  For spin 2 every mode has an E and a B
column over the (Q, U) rows, negated relative to the reference's V so that
the T x polarisation blocks carry the pixel-space sign
(Q + iU = -sum (a_E + i a_B) 2Y).  By the addition theorems
sum_m V_lm V_lm^T is the pixel-space derivative dS/dC_l of the TT, EE, BB,
TE, TB and EB blocks.
"""

from __future__ import annotations

import math

import numpy as np


def wigner_d_sum(ell: int, m: int, s: int, theta: np.ndarray) -> np.ndarray:
    """d^l_{m,s}(theta) from the finite factorial sum (exact in principle; its
    alternating terms cancel, so it is used for l <= 12 only)."""
    theta = np.asarray(theta, dtype=np.float64)
    if ell < abs(m) or ell < abs(s):
        return np.zeros_like(theta)
    # the standard d^j_{a b} sum with a = s, b = m
    a, b, j = s, m, ell
    c, sn = np.cos(0.5 * theta), np.sin(0.5 * theta)
    pref = math.sqrt(math.factorial(j + a) * math.factorial(j - a) * math.factorial(j + b) * math.factorial(j - b))
    out = np.zeros_like(theta)
    for k in range(max(0, b - a), min(j + b, j - a) + 1):
        den = (math.factorial(j + b - k) * math.factorial(k) * math.factorial(j - k - a)
               * math.factorial(k - b + a))
        out += ((-1.0) ** (k - b + a)) * (pref / den) * c ** (2 * j - 2 * k + b - a) * sn ** (2 * k - b + a)
    return out


def wigner_d(ell: int, m: int, s: int, theta: np.ndarray) -> np.ndarray:
    """d^l_{m,s}(theta) through the Jacobi polynomial P_k^{(mu, nu)} (scipy's
    stable evaluation): with mu = |m - s|, nu = |m + s|, k = l - (mu + nu)/2,
    d = xi sqrt(k! (k + mu + nu)! / ((k + mu)! (k + nu)!)) sin^mu(t/2)
    cos^nu(t/2) P_k^{(mu, nu)}(cos t), xi = 1 for m >= s, else (-1)^{m - s}."""
    from scipy.special import eval_jacobi, gammaln

    theta = np.asarray(theta, dtype=np.float64)
    if ell < abs(m) or ell < abs(s):
        return np.zeros_like(theta)
    mu, nu = abs(m - s), abs(m + s)
    k = ell - (mu + nu) // 2
    xi = 1.0 if m >= s else (-1.0) ** (m - s)
    lg = 0.5 * (gammaln(k + 1) + gammaln(k + mu + nu + 1) - gammaln(k + mu + 1) - gammaln(k + nu + 1))
    return (xi * math.exp(lg) * np.sin(0.5 * theta) ** mu * np.cos(0.5 * theta) ** nu
            * eval_jacobi(k, mu, nu, np.cos(theta)))


def n_modes(lmin: int, lmax: int) -> int:
    return (lmax + 1) ** 2 - lmin ** 2


def mode_ells(lmin: int, lmax: int) -> np.ndarray:
    return np.concatenate([np.full(2 * ell + 1, ell) for ell in range(lmin, lmax + 1)])


def _angles(vec: np.ndarray):
    vec = np.asarray(vec, dtype=np.float64)
    theta = np.arccos(np.clip(vec[:, 2], -1.0, 1.0))
    phi = np.arctan2(vec[:, 1], vec[:, 0])
    return theta, phi


def modes_spin0(vec: np.ndarray, lmin: int, lmax: int) -> np.ndarray:
    """(npix, n_modes) real spherical harmonics."""
    theta, phi = _angles(vec)
    cols = []
    for ell in range(lmin, lmax + 1):
        sc = math.sqrt((2 * ell + 1) / (4 * math.pi))
        cols.append(sc * wigner_d(ell, 0, 0, theta))
        for m in range(1, ell + 1):
            base = math.sqrt(2.0) * sc * wigner_d(ell, m, 0, theta)
            cols.append(base * np.cos(m * phi))
            cols.append(base * np.sin(m * phi))
    return np.column_stack(cols)


def modes_spin2(vec: np.ndarray, lmin: int, lmax: int):
    """(E, B): each (2 npix, n_modes) over rows [Q | U]; l >= 2."""
    if lmin < 2:
        raise ValueError("spin-2 modes start at l = 2")
    theta, phi = _angles(vec)
    npix = theta.size
    E, B = [], []
    for ell in range(lmin, lmax + 1):
        sc = math.sqrt((2 * ell + 1) / (4 * math.pi))
        for m in range(0, ell + 1):
            dp = wigner_d(ell, m, 2, theta)
            dm = wigner_d(ell, m, -2, theta)
            Dp, Dm = dm + dp, dm - dp
            if m == 0:
                h = -0.5 * sc
                z = np.zeros(npix)
                E.append(np.r_[h * Dp, z])
                B.append(np.r_[z, h * Dp])
                continue
            h = -0.5 * math.sqrt(2.0) * sc
            cm, sm = np.cos(m * phi), np.sin(m * phi)
            E.append(np.r_[h * Dp * cm, h * Dm * sm])
            B.append(np.r_[-h * Dm * sm, h * Dp * cm])
            E.append(np.r_[h * Dp * sm, -h * Dm * cm])
            B.append(np.r_[h * Dm * cm, h * Dp * sm])
    return np.column_stack(E), np.column_stack(B)


def mode_groups(layout, lmin: int, lmax: int) -> dict:
    """Field-map label -> (n, n_modes) mode matrix over all rows of the layout."""
    n = layout.n
    out = {}
    for f, off in zip(layout.fields, layout.offsets()):
        if f.spin == 0:
            G = np.zeros((n, n_modes(lmin, lmax)))
            G[off: off + f.n] = modes_spin0(f.vec, lmin, lmax)
            out[f.labels[0].upper()] = G
        else:
            E, B = modes_spin2(f.vec, lmin, lmax)
            for lab, X in zip(f.labels, (E, B)):
                G = np.zeros((n, n_modes(lmin, lmax)))
                G[off: off + 2 * f.n] = X
                out[lab.upper()] = G
    return out


def _split(groups: dict, label: str):
    hits = [(a, b) for a in groups for b in groups if a + b == label.upper()]
    if len(hits) != 1:
        raise ValueError(f"label {label!r} does not split into the mode groups {sorted(groups)}")
    return hits[0]


def mode_derivative(groups: dict, label: str, w: np.ndarray, lmin: int, lmax: int) -> np.ndarray:
    """dC = sum_k w_{l_k} (x_k y_k^T [+ y_k x_k^T]) of spectrum ``label``."""
    a, b = _split(groups, label)
    ells = mode_ells(lmin, lmax)
    ww = np.zeros(ells.size)
    ok = ells < w.size
    ww[ok] = w[ells[ok]]
    X, Y = groups[a], groups[b]
    D = (X * ww) @ Y.T
    return D if a == b else D + D.T


def fisher_modes(layout, Cinv: np.ndarray, params, lmin: int, lmax: int) -> np.ndarray:
    """F_ij = 1/2 Tr(C^-1 dC_i C^-1 dC_j) with the mode derivatives, in the
    reference formulation (A_i = C^-1 dC_i, ``fisher.py:506-603``)."""
    from . import kernels as K

    groups = mode_groups(layout, lmin, lmax)
    A = [Cinv @ mode_derivative(groups, lab, w, lmin, lmax) for _, _, _, lab, w in params]
    F = np.zeros((len(A), len(A)))
    for i in range(len(A)):
        for j in range(i, len(A)):
            F[i, j] = F[j, i] = 0.5 * K.matrix_trace(A[i], A[j])
    return F


def qml_modes(layout, Cinv: np.ndarray, N: np.ndarray, maps: np.ndarray, params, lmin: int, lmax: int,
              remove_nb: bool = True):
    """(q, nb) with the mode derivatives (``spectra.py:1124-1186``)."""
    groups = mode_groups(layout, lmin, lmax)
    y = Cinv @ np.asarray(maps, dtype=np.float64)
    G = Cinv @ (N @ Cinv)
    q = np.zeros((y.shape[1], len(params)))
    nb = np.zeros(len(params))
    for p, (_, _, _, lab, w) in enumerate(params):
        dC = mode_derivative(groups, lab, w, lmin, lmax)
        nb[p] = 0.5 * float(np.sum(G * dC))
        q[:, p] = 0.5 * np.sum(y * (dC @ y), axis=0) - (nb[p] if remove_nb else 0.0)
    return q, nb
