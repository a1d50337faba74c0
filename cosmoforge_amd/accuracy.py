"""Differences between two Fisher + QML solves (e.g. mixed against FP64),
in the units the estimator's users read them in.

* F: |dF_ij| / sqrt(F_ii F_jj), each entry against the geometric mean of
  its two diagonal entries (the diagonal itself: the plain relative error).
* C_l: |dC| / max(|C|, sigma) with sigma = sqrt((F^-1)_ll): relative where
  the estimate exceeds its error bar, in units of the error bar where it
  does not (EB, TB and noise-dominated bins scatter around zero).
* noise bias and q: |d| / sqrt(F_pp), the shift in units of the standard
  deviation of q_p.
"""

from __future__ import annotations

import numpy as np


def errors(res, ref) -> dict:
    """``res`` / ``ref``: (F, C_hat, noise bias, q, sigma) of two solves
    (sigma of ``ref`` is used; ``res`` may omit it)."""
    F, C, nb, q, sig = ref[:5]
    Fm, Cm, nbm, qm = res[:4]
    d = np.sqrt(np.diag(F))
    E = np.abs(Fm - F) / np.outer(d, d)
    C, Cm, sig = np.ravel(C), np.ravel(Cm), np.ravel(sig)
    return {"F_entry_max": float(E.max()),
            "F_diag_rel_max": float(np.max(np.abs(np.diag(Fm) - np.diag(F)) / np.diag(F))),
            "Cl_rel_max": float(np.max(np.abs(Cm - C) / np.maximum(np.abs(C), sig))),
            "Cl_over_sigma_max": float(np.max(np.abs(Cm - C) / sig)),
            "noise_bias_over_sigma_q_max": float(np.max(np.abs(np.ravel(nbm) - np.ravel(nb)) / d)),
            "q_over_sigma_q_max": float(np.max(np.abs(np.atleast_2d(qm) - np.atleast_2d(q)) / d[None, :]))}


def per_bin(res, ref, labels, ells) -> list:
    """One row per parameter (spectrum label, multipole): the largest F
    error of its row, its C_l error, and its noise-bias and q shifts, in the
    units of ``errors``."""
    F, C, nb, q, sig = ref[:5]
    Fm, Cm, nbm, qm = res[:4]
    d = np.sqrt(np.diag(F))
    E = np.abs(Fm - F) / np.outer(d, d)
    C, Cm, sig = np.ravel(C), np.ravel(Cm), np.ravel(sig)
    ec = np.abs(Cm - C) / np.maximum(np.abs(C), sig)
    en = np.abs(np.ravel(nbm) - np.ravel(nb)) / d
    eq = np.max(np.abs(np.atleast_2d(qm) - np.atleast_2d(q)) / d[None, :], axis=0)
    rows = []
    p = 0
    for lab in labels:
        for ell in ells:
            rows.append((lab, int(ell), float(E[p].max()), float(ec[p]), float(en[p]), float(eq[p])))
            p += 1
    return rows


def summarize(rows, ell_edges) -> list:
    """Maxima of ``per_bin`` rows per (spectrum, multipole range)."""
    out = []
    labels = list(dict.fromkeys(r[0] for r in rows))
    for lab in labels:
        for lo, hi in zip(ell_edges[:-1], ell_edges[1:]):
            sel = [r for r in rows if r[0] == lab and lo <= r[1] < hi]
            if sel:
                out.append((lab, lo, hi - 1) + tuple(max(r[k] for r in sel) for k in (2, 3, 4, 5)))
    return out
