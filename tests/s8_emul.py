"""Host restatement of the int8-slice arithmetic of ``cfmi_slice8`` /
``cfmi_gemm_s8`` (test infrastructure; no reference counterpart).

Every step uses the same IEEE operations as the device code (correctly
rounded products, differences, rint, no fused multiply-add), so the device
results are reproduced bit for bit:

* a row x is scaled by mu = max |x| (1 for a zero row), t = (x / mu) 127,
  slice_1 = rint(t), t <- (t - slice_1) 254, slice_2 = rint(t), ...
* a product sums the slice pairs with p + q < s exactly (integers), then
  v = acc_{s-1}; v = acc_d + v / 254 (d = s-2 .. 0, the division as a
  product with the rounded constant 1/254), v <- (v (1/16129)) ((alpha mu_m) nu_n).
"""

from __future__ import annotations

import numpy as np

INV254 = 1.0 / 254.0
INV127SQ = 1.0 / 16129.0


def slices(X: np.ndarray, s: int):
    """(mu (rows,), int64 slices (s, rows, k)) of the rows of X."""
    X = np.asarray(X, dtype=np.float64)
    mu = np.max(np.abs(X), axis=1)
    mu = np.where(mu == 0.0, 1.0, mu)
    inv = 1.0 / mu
    t = (X * inv[:, None]) * 127.0
    out = np.empty((s,) + X.shape, dtype=np.int64)
    for p in range(s):
        sp = np.rint(t)
        out[p] = sp.astype(np.int64)
        t = (t - sp) * 254.0
    return mu, out


def unslice(mu: np.ndarray, sl: np.ndarray) -> np.ndarray:
    s = sl.shape[0]
    v = sl[s - 1].astype(np.float64)
    for p in range(s - 2, -1, -1):
        v = sl[p].astype(np.float64) + v * INV254
    return (v * (1.0 / 127.0)) * mu[:, None]


def kmax(s: int) -> int:
    """Largest k of one exact int32 pass (``cfmi_s8_kmax``)."""
    return (0x7FFFFFFF // (16129 * s)) // 64 * 64


def gemm(A: np.ndarray, B: np.ndarray, s: int, alpha: float = 1.0, beta: float = 0.0, C=None,
         scale=None) -> np.ndarray:
    """alpha A B^T (+ beta C) with A (m, k) and B (n, k) sliced by rows; k
    beyond ``kmax(s)`` in chunks added in FP64, as the device splits it.
    ``scale``: elementwise factors applied to each chunk's product before the
    addition (the packed band's diagonal factor)."""
    muA, sa = slices(A, s)
    muB, sb = slices(B, s)
    k = A.shape[1]
    step = kmax(s)
    out = None
    b = beta
    for k0 in range(0, k, step):
        k1 = min(k, k0 + step)
        acc = []
        for d in range(s):
            # integer sums below 2^31 are exact in float64 products (BLAS)
            t = np.zeros((A.shape[0], B.shape[0]))
            for p in range(d + 1):
                t += sa[p][:, k0:k1].astype(np.float64) @ sb[d - p][:, k0:k1].T.astype(np.float64)
            assert np.max(np.abs(t), initial=0.0) < 2.0**31
            acc.append(t)
        v = acc[s - 1].astype(np.float64)
        for d in range(s - 2, -1, -1):
            v = acc[d].astype(np.float64) + v * INV254
        sc = (alpha * muA)[:, None] * muB[None, :]
        part = (v * INV127SQ) * sc
        if scale is not None:
            part = part * scale
        prev = out if out is not None else C
        out = part if b == 0.0 else part + b * prev
        b = 1.0
    return out


def bound(A: np.ndarray, B: np.ndarray, s: int) -> np.ndarray:
    """Elementwise bound on |sliced product - A B^T|: each operand is held to
    mu / (2 127 254^(s-1)) per entry, and the dropped levels p + q >= s add
    at most (s + 2) 254^-s mu nu per k; the FP64 epilogue a few ulps."""
    muA = np.max(np.abs(A), axis=1)
    muB = np.max(np.abs(B), axis=1)
    e = 0.5 / (127.0 * 254.0 ** (s - 1))
    k = A.shape[1]
    rep = e * (muA[:, None] * np.abs(B).sum(axis=1)[None, :] + np.abs(A).sum(axis=1)[:, None] * muB[None, :])
    drop = k * (s + 2) * 254.0 ** (-s) * muA[:, None] * muB[None, :]
    return rep + drop + 8e-16 * (np.abs(A) @ np.abs(B).T) + 1e-300      # + FP64 rounding of the epilogue
