"""FP64 Fisher / QML / PICSLike oracle (test infrastructure only).

Restates the reference's traditional pixel-space path:

* smoothing -- ``src/cosmoforge.cosmocore/cosmocore/beam.py:223-414`` and
  ``spectra_io.py:446-507`` (C_l <- C_l * B1(l) B2(l)) of a given (3, L+1)
  beam array (every oracle call passes the beam explicitly);
* Fisher -- ``src/cosmoforge.qube/qube/fisher.py:311-344`` (C = N + S, C^-1),
  ``:506-531`` (A_i = C^-1 dC_i) and ``:557-603`` (F_ij = 1/2 Tr(A_i A_j)),
  with the beam-smoothing rows of ``:790-811``;
* QML -- ``src/cosmoforge.qube/qube/spectra.py:1124-1186`` (y = C^-1 x,
  G = C^-1 N C^-1, q_b = 1/2 y^T dC_b y - 1/2 Tr(G dC_b)),
  ``:607-661`` (F^-1 via the Jacobi-scaled inverse), ``:1340-1356`` and
  ``:1713-1733`` (C_hat = q F^-1, noise bias = nb F^-1);
* PICSLike -- ``src/cosmoforge.picslike/picslike/picslike.py:264-319,643-664``
  (chi2 = d^T C^-1 d, logdet by a second Cholesky, lnL = -(chi2+logdet)/2);
* pixel filters -- ``src/cosmoforge.cosmocore/cosmocore/basics/linalg.py:477-528``
  (Sigma W^T X W Sigma, Sigma W^T d) at the seams of ``core.py:1070-1140,1455-1469``
  (N, S, every binned derivative and the maps restricted, the estimator run in r dimensions);
* cross analysis (``do_cross``) -- ``fisher.py:522-531,584-590``
  (A_i = C2^-1 dC_i C1^-1, F_ij = 1/2 Tr(dC_j A_i)) and ``spectra.py:1146-1166``
  (q_b = 1/2 sum y2 * (dC_b y1), no noise bias).
"""

from __future__ import annotations

import numpy as np

from . import kernels as K
from .linalg import matrix_inverse_symm, matrix_slogdet_symm


def beam_by_label(layout: K.Layout, beam3: np.ndarray, lmax: int) -> dict:
    """Field-map label -> beam window (``beam.py:223-309``)."""
    out = {}
    named = {"T": 0, "E": 1, "B": 2}
    for f in layout.fields:
        if f.spin == 0:
            row = named.get(f.labels[0], 0)
            out[f.labels[0]] = np.asarray(beam3[row][: lmax + 1], dtype=np.float64)
        else:
            out[f.labels[0]] = np.asarray(beam3[1][: lmax + 1], dtype=np.float64)
            out[f.labels[1]] = np.asarray(beam3[2][: lmax + 1], dtype=np.float64)
    return out


def split_label(layout: K.Layout, label: str, i: int, j: int):
    cands = list(layout.fields[i].labels) + list(layout.fields[j].labels)
    s = {(a, b) for a in cands for b in cands if (a + b).upper() == label}
    if len(s) != 1:
        raise ValueError(f"label {label} does not split uniquely")
    return s.pop()


def smoothing_factors(layout: K.Layout, beams: dict, lmax: int) -> dict:
    """label -> B1(l) B2(l), l = 0..lmax (``spectra_io.py:446-507``)."""
    out = {}
    for (i, j, _), lab in layout.spectra_map.items():
        a, b = split_label(layout, lab, i, j)
        out[lab] = beams[a][: lmax + 1] * beams[b][: lmax + 1]
    return out


def smoothed_cls(layout: K.Layout, cls: dict, beams: dict, lmax: int) -> dict:
    sm = smoothing_factors(layout, beams, lmax)
    out = {}
    for lab in layout.labels:
        if lab not in cls:
            raise ValueError(f"Missing power spectrum for {lab}")
        out[lab] = np.asarray(cls[lab], dtype=np.float64)[: lmax + 1].copy() * sm[lab]
    return out


def bins_fromdeltal(lmin: int, lmax: int, delta: int = 1):
    nb = (lmax - lmin + 1) // delta
    lmins = lmin + np.arange(nb) * delta
    return list(zip(lmins.tolist(), (lmins + delta - 1).tolist()))


def shape_weights(lmax_bins: int, convention: str = "Cl") -> np.ndarray:
    if convention.lower() == "cl":
        return np.ones(lmax_bins + 1)
    ell = np.arange(lmax_bins + 1, dtype=np.float64)
    w = np.zeros(lmax_bins + 1)
    w[1:] = 2 * np.pi / (ell[1:] * (ell[1:] + 1))
    return w


def parameter_weights(layout: K.Layout, beams: dict, lmin: int, lmax: int, lmax_signal: int,
                      bins=None, convention: str = "Cl"):
    """For each Fisher parameter: (fi, fj, mode, label, l-indexed weights).

    Parameter index = spectrum * nbins + bin (``fisher.py:506-519``); the
    weight is w_l * b1(l) b2(l) inside the bin (``core.py:1361-1409``)."""
    bins = bins if bins is not None else bins_fromdeltal(lmin, lmax)
    sw = shape_weights(max(b[1] for b in bins), convention)
    sm = smoothing_factors(layout, beams, lmax)
    out = []
    for fi, fj, mode, lab in layout.spectra_list():
        for lo, hi in bins:
            w = np.zeros(lmax_signal + 1)
            for ell in range(lo, hi + 1):
                if ell > lmax_signal:
                    continue
                w[ell] = sw[ell] * sm[lab][ell]
            out.append((fi, fj, mode, lab, w))
    return out


def fisher(layout: K.Layout, cls_s: dict, lmax_signal: int, N: np.ndarray, params, S=None):
    """(F, C^-1, S) in the reference formulation (stores every A_i)."""
    if S is None:
        S = K.signal_matrix(layout, cls_s, lmax_signal)
    Cinv = matrix_inverse_symm(N + S)
    A = []
    for fi, fj, mode, _, w in params:
        dC = K.binned_derivative(layout, fi, fj, mode, w)
        A.append(Cinv @ dC)
    npar = len(A)
    F = np.zeros((npar, npar))
    for i in range(npar):
        for j in range(i, npar):
            v = 0.5 * K.matrix_trace(A[i], A[j])
            F[i, j] = v
            F[j, i] = v
    return F, Cinv, S


def qml(layout: K.Layout, Cinv: np.ndarray, N: np.ndarray, maps: np.ndarray, F: np.ndarray,
        params, remove_nb: bool = True):
    """(q, nb, C_hat, nb_hat); maps is (n, nsims)."""
    maps = np.asarray(maps, dtype=np.float64)
    if maps.ndim == 1:
        maps = maps[:, None]
    y = Cinv @ maps
    G = Cinv @ (N @ Cinv)
    npar = len(params)
    q = np.zeros((maps.shape[1], npar))
    nb = np.zeros(npar)
    for p, (fi, fj, mode, _, w) in enumerate(params):
        dC = K.binned_derivative(layout, fi, fj, mode, w)
        nb[p] = 0.5 * K.matrix_trace(G, dC)
        qv = 0.5 * np.sum(y * (dC @ y), axis=0)
        if remove_nb:
            qv = qv - nb[p]
        q[:, p] = qv
    invF = matrix_inverse_symm(F)
    return q, nb, q @ invF, nb @ invF


def picslike_point(layout: K.Layout, cls_s: dict, lmax_signal: int, N: np.ndarray,
                   maps: np.ndarray):
    """(chi2, logdet, lnL) per simulation for one parameter point."""
    maps = np.asarray(maps, dtype=np.float64)
    if maps.ndim == 1:
        maps = maps[:, None]
    S = K.signal_matrix(layout, cls_s, lmax_signal)
    C = S + N
    Cinv = matrix_inverse_symm(C)
    chi2 = np.einsum("in,ij,jn->n", maps, Cinv, maps)
    logdet = matrix_slogdet_symm(C)[1]
    return chi2, logdet, -0.5 * (chi2 + logdet)


def fisher_cross(layout: K.Layout, cls_s: dict, lmax_signal: int, N1: np.ndarray, N2: np.ndarray, params,
                 S=None):
    """(F, C1^-1, C2^-1) of a cross analysis in the reference formulation."""
    if S is None:
        S = K.signal_matrix(layout, cls_s, lmax_signal)
    Cinv1 = matrix_inverse_symm(N1 + S)
    Cinv2 = matrix_inverse_symm(N2 + S)
    dCs, A = [], []
    for fi, fj, mode, _, w in params:
        dC = K.binned_derivative(layout, fi, fj, mode, w)
        dCs.append(dC)
        A.append(Cinv2 @ (dC @ Cinv1))
    npar = len(A)
    F = np.zeros((npar, npar))
    for i in range(npar):
        for j in range(i, npar):
            v = 0.5 * K.matrix_trace(dCs[j], A[i])
            F[i, j] = v
            F[j, i] = v
    return F, Cinv1, Cinv2


def qml_cross(layout: K.Layout, Cinv1: np.ndarray, Cinv2: np.ndarray, maps1: np.ndarray, maps2: np.ndarray,
              F: np.ndarray, params):
    """(q, C_hat) of a cross analysis; maps are (n, nsims)."""
    y1 = Cinv1 @ np.asarray(maps1, dtype=np.float64)
    y2 = Cinv2 @ np.asarray(maps2, dtype=np.float64)
    q = np.zeros((y1.shape[1], len(params)))
    for p, (fi, fj, mode, _, w) in enumerate(params):
        dC = K.binned_derivative(layout, fi, fj, mode, w)
        q[:, p] = 0.5 * np.sum(y2 * (dC @ y1), axis=0)
    return q, q @ matrix_inverse_symm(F)


def restrict(X: np.ndarray, W: np.ndarray, sigma=None) -> np.ndarray:
    """Sigma W^T X W Sigma, symmetrised (``linalg.py:477-503``)."""
    M = W.T @ X @ W
    if sigma is not None:
        M = M * sigma[:, None] * sigma[None, :]
    return 0.5 * (M + M.T)


def restrict_data(d: np.ndarray, W: np.ndarray, sigma=None) -> np.ndarray:
    """Sigma W^T d (``linalg.py:506-528``)."""
    out = W.T @ d
    if sigma is not None:
        out = out * (sigma if out.ndim == 1 else sigma[:, None])
    return out


def fisher_filtered(layout: K.Layout, cls_s: dict, lmax_signal: int, N: np.ndarray, params, W: np.ndarray,
                    sigma=None, S=None):
    """(F, C'^-1, N') of the restricted problem: C' = S' + N', A_i = C'^-1 dC'_i,
    F_ij = 1/2 Tr(A_i A_j)."""
    if S is None:
        S = K.signal_matrix(layout, cls_s, lmax_signal)
    Nr = restrict(N, W, sigma)
    Cinv = matrix_inverse_symm(restrict(S, W, sigma) + Nr)
    A = []
    for fi, fj, mode, _, w in params:
        A.append(Cinv @ restrict(K.binned_derivative(layout, fi, fj, mode, w), W, sigma))
    npar = len(A)
    F = np.zeros((npar, npar))
    for i in range(npar):
        for j in range(i, npar):
            F[i, j] = F[j, i] = 0.5 * K.matrix_trace(A[i], A[j])
    return F, Cinv, Nr


def qml_filtered(layout: K.Layout, Cinv: np.ndarray, Nr: np.ndarray, maps: np.ndarray, F: np.ndarray, params,
                 W: np.ndarray, sigma=None, remove_nb: bool = True):
    """(q, nb, C_hat) of the restricted problem with d' = Sigma W^T d."""
    y = Cinv @ restrict_data(np.asarray(maps, dtype=np.float64), W, sigma)
    G = Cinv @ (Nr @ Cinv)
    q = np.zeros((y.shape[1], len(params)))
    nb = np.zeros(len(params))
    for p, (fi, fj, mode, _, w) in enumerate(params):
        dC = restrict(K.binned_derivative(layout, fi, fj, mode, w), W, sigma)
        nb[p] = 0.5 * K.matrix_trace(G, dC)
        q[:, p] = 0.5 * np.sum(y * (dC @ y), axis=0) - (nb[p] if remove_nb else 0.0)
    return q, nb, q @ matrix_inverse_symm(F)
