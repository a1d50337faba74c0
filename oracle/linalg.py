"""Dense FP64 linear algebra of the oracle (test infrastructure only).

Restates ``src/cosmoforge.cosmocore/cosmocore/basics/linalg.py``:

* ``matrix_inverse_symm`` -- :152-219: Jacobi scaling D M D with
  D = diag(1/sqrt|M_ii|), LAPACK dpotrf + dpotri, mirror, unscale;
  raises ``ValueError("dpotrf failed with info=k")``.
* ``matrix_slogdet_symm`` -- :430-474: a second dpotrf, 2 * sum(log L_ii).
* ``cholesky_factor`` -- :362-406: ``np.linalg.LinAlgError`` on non-PD.
"""

from __future__ import annotations

import numpy as np
from scipy.linalg import lapack


def matrix_inverse_symm(M: np.ndarray) -> np.ndarray:
    if M.shape[0] != M.shape[1]:
        raise ValueError("Matrix must be square")
    M = np.asfortranarray(M.copy())
    diag = np.diag(M).copy()
    abs_diag = np.abs(diag)
    abs_diag[abs_diag == 0.0] = 1.0
    d_inv = 1.0 / np.sqrt(abs_diag)
    np.multiply(M, d_inv[None, :], out=M)
    np.multiply(M, d_inv[:, None], out=M)
    L, info = lapack.dpotrf(M, lower=True, overwrite_a=True, clean=True)
    if info != 0:
        raise ValueError(f"dpotrf failed with info={info}")
    inv, info_i = lapack.dpotri(L, lower=True, overwrite_c=True)
    if info_i != 0:
        raise ValueError(f"dpotri failed with info={info_i}")
    inv = np.tril(inv) + np.tril(inv, -1).T
    inv = np.asfortranarray(inv)
    np.multiply(inv, d_inv[None, :], out=inv)
    np.multiply(inv, d_inv[:, None], out=inv)
    return inv


def matrix_slogdet_symm(M: np.ndarray):
    if M.shape[0] != M.shape[1]:
        raise ValueError("Matrix must be square")
    L, info = lapack.dpotrf(M, lower=True, overwrite_a=False, clean=True)
    if info != 0:
        raise ValueError(f"dpotrf failed with info={info}")
    d = np.diag(L)
    if np.any(d <= 0):
        raise ValueError("Non-positive diagonal element in Cholesky factor")
    return 1.0, 2.0 * float(np.sum(np.log(d)))


def cholesky_factor(M: np.ndarray):
    L, info = lapack.dpotrf(M, lower=True, overwrite_a=False, clean=True)
    if info != 0:
        raise np.linalg.LinAlgError(
            f"dpotrf failed: matrix is not positive definite (info={info})")
    return L, True
