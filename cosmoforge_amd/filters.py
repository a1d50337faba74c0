"""Pixel filters as restriction to the filter's range (host setup + device
restriction).

Follows ``src/cosmoforge.cosmocore/cosmocore/filters.py:44-490`` (the
``Filter`` record and its constructors) and the seams of ``core.py:1070-1140,
1455-1469`` (ADR-0020): a linear filter ``F = U Sigma W^T`` truncated at rank
r is a subspace, and the analysis runs inside it.  Raw pixel-space inputs map
as ``Sigma W^T x`` and matrices as ``Sigma W^T X W Sigma``; inputs that
already carry the filter map as ``U^T x`` and ``U^T X U``.  For a projector
``U`` is ``W`` and ``Sigma = I``.

The constructors factor F (or the template stack) once on the host.  The
restrictions run on the device: ``DeviceFilter`` holds W, U and Sigma in HBM
and restricts N, S, every binned derivative block stack (two batched GEMMs
per block, all bins of a spectrum per launch) and the maps.
"""

from __future__ import annotations

import warnings
from hashlib import blake2b

import numpy as np
import torch

from . import engine
from .engine import P

MIN_RANGE_EPSILON = float(np.sqrt(np.finfo(np.float64).eps))
GAP_WARN_RATIO = 0.1          # sigma_{k+1}/sigma_k above this is "no spectral gap"
_PROJECTOR_ATOL = 1e-10


def active_pixel_index(mask) -> np.ndarray:
    """Row positions of the reduced arrays in the flattened (ncomp npix)
    vector (``geometry.py:94-134``)."""
    mask = np.asarray(mask, dtype=np.float64)
    if mask.ndim == 1:
        mask = mask[:, None]
    npix = mask.shape[0]
    return np.concatenate([np.nonzero(mask[:, c] > 0.5)[0] + c * npix for c in range(mask.shape[1])]).astype(np.intp)


def mask_fingerprint(mask) -> bytes:
    """16-byte blake2b digest of the active-pixel ordering (``filters.py:44-60``)."""
    return blake2b(active_pixel_index(mask).tobytes(), digest_size=16).digest()


def _complement(A: np.ndarray, epsilon: float):
    """(kept, removed): orthonormal bases of the complement of range(A) and of
    its directions above the relative cut ``epsilon`` (``filters.py:63-79``)."""
    U, s, _ = np.linalg.svd(A, full_matrices=True)
    nrem = int(np.count_nonzero(s > epsilon * s[0])) if s.size else 0
    if nrem >= A.shape[0]:
        raise ValueError(f"the {A.shape[1]} columns handed in span all {A.shape[0]} pixels above "
                         f"epsilon={epsilon:g}, leaving a rank-0 filter; the geometry cannot support this "
                         "many independent templates")
    return np.ascontiguousarray(U[:, nrem:]), np.ascontiguousarray(U[:, :nrem])


class Filter:
    """A pixel-space filter as its truncated thin SVD ``F = U Sigma W^T``
    (``filters.py:82-170``).  ``W``, ``U``: (n_pixels, rank) orthonormal;
    ``sigma``: kept singular values; a projector has ``U is W``."""

    __slots__ = ("W", "U", "sigma", "rank", "is_projector", "range_epsilon", "range_rank", "Q_removed",
                 "fingerprint", "n_pixels")

    def __init__(self, W, U, sigma, rank, is_projector, range_epsilon, range_rank, Q_removed, fingerprint,
                 n_pixels):
        if W.ndim != 2:
            raise ValueError(f"W must be 2-D, got shape {W.shape}")
        if W.shape != (n_pixels, rank):
            raise ValueError(f"W has shape {W.shape}, expected ({n_pixels}, {rank})")
        if U.shape != W.shape:
            raise ValueError(f"U has shape {U.shape}, expected {W.shape} (same as W)")
        if sigma.shape != (rank,):
            raise ValueError(f"sigma has shape {sigma.shape}, expected ({rank},)")
        if is_projector and U is not W:
            raise ValueError("a projector's U must be the same object as W, so that raw and pre-filtered "
                             "inputs land in identical coordinates")
        for k, v in zip(self.__slots__, (W, U, sigma, rank, is_projector, range_epsilon, range_rank, Q_removed,
                                         fingerprint, n_pixels)):
            object.__setattr__(self, k, v)

    def __setattr__(self, key, value):
        raise AttributeError("Filter is immutable")

    def __repr__(self):
        kind = "projector" if self.is_projector else "graded"
        knob = (f"range_rank={self.range_rank}" if self.range_rank is not None
                else f"range_epsilon={self.range_epsilon:g}")
        return f"Filter({kind}, n_pixels={self.n_pixels}, rank={self.rank}, {knob})"

    @property
    def scaling(self):
        """``sigma``, or None for a projector."""
        return None if self.is_projector else self.sigma

    @classmethod
    def from_operator(cls, F, *, range_epsilon=None, range_rank=None, mask=None) -> "Filter":
        """Thin SVD of a dense (n, n) operator, truncated by ``range_epsilon``
        (sigma_i > eps sigma_1, clipped to MIN_RANGE_EPSILON) or ``range_rank``
        (``filters.py:209-311``)."""
        if range_epsilon is None and range_rank is None:
            raise ValueError("a rank threshold is required: pass range_epsilon or range_rank. There is no "
                             "default because the right cut depends on the operator's spectrum, and the "
                             "restricted covariance carries Σ²")
        if range_epsilon is not None and range_rank is not None:
            raise ValueError("range_epsilon and range_rank are mutually exclusive; pass one")
        F = np.asarray(F, dtype=np.float64)
        if F.ndim != 2 or F.shape[0] != F.shape[1]:
            raise ValueError(f"F must be square (n_pixels, n_pixels), got shape {F.shape}")
        n = F.shape[0]
        Uf, s, Wt = np.linalg.svd(F)
        eff = None
        if range_rank is not None:
            k = int(range_rank)
            if not 1 <= k <= s.size:
                raise ValueError(f"range_rank={k} outside 1..{s.size} for this operator")
            if s[k - 1] <= MIN_RANGE_EPSILON * s[0]:
                raise ValueError(f"range_rank={k} keeps σ_{k}/σ₁ = {s[k - 1] / s[0]:.3g}, at or below "
                                 f"MIN_RANGE_EPSILON={MIN_RANGE_EPSILON:g}: the restricted covariance would be "
                                 "singular. Lower range_rank")
        else:
            eps = float(range_epsilon)
            if eps < MIN_RANGE_EPSILON:
                warnings.warn(f"range_epsilon={eps:g} is below MIN_RANGE_EPSILON={MIN_RANGE_EPSILON:g} and has "
                              "been clipped: the restricted covariance carries Σ², so a smaller cut puts it "
                              "beyond double precision", stacklevel=2)
                eps = MIN_RANGE_EPSILON
            eff = eps
            k = int(np.count_nonzero(s > eps * s[0]))
            if k == 0:
                raise ValueError(f"no singular value of F exceeds {eps:g} × σ₁; the filter would have rank 0")
        _warn_on_truncation(s, k)
        sigma = np.ascontiguousarray(s[:k])
        W = np.ascontiguousarray(Wt[:k].T)
        proj = bool(np.allclose(sigma, 1.0, rtol=0.0, atol=_PROJECTOR_ATOL)
                    and np.allclose(Uf[:, :k], W, rtol=0.0, atol=1e-8))
        if proj:
            U, Qr = W, np.ascontiguousarray(Wt[k:].T)
        else:
            U, Qr = np.ascontiguousarray(Uf[:, :k]), None
        return cls(W, U, sigma, k, proj, eff, None if range_rank is None else int(range_rank), Qr,
                   None if mask is None else mask_fingerprint(mask), n)

    @classmethod
    def from_deprojection(cls, T, *, template_epsilon=1e-3, mask=None) -> "Filter":
        """The projector annihilating the template stack T (n, n_templates)
        (``filters.py:313-356``)."""
        T = np.asarray(T, dtype=np.float64)
        if T.ndim != 2:
            raise ValueError(f"T must be 2-D (n_pixels, n_templates), got {T.shape}")
        W, Qr = _complement(T, float(template_epsilon))
        return cls(W, W, np.ones(W.shape[1]), W.shape[1], True, None, W.shape[1], Qr,
                   None if mask is None else mask_fingerprint(mask), T.shape[0])

    @classmethod
    def from_subspace(cls, W, *, mask=None) -> "Filter":
        """The projector onto an orthonormal (n, r) basis (``filters.py:358-400``)."""
        W = np.ascontiguousarray(W, dtype=np.float64)
        if W.ndim != 2:
            raise ValueError(f"W must be 2-D (n_pixels, rank), got {W.shape}")
        if W.shape[1] > W.shape[0]:
            raise ValueError(f"W has {W.shape[1]} columns for {W.shape[0]} pixels; a subspace cannot exceed "
                             "the space")
        norms = np.linalg.norm(W, axis=0)
        if not np.allclose(norms, 1.0, atol=1e-8):
            raise ValueError(f"W's columns must be orthonormal; column norms range over "
                             f"[{norms.min():g}, {norms.max():g}]")
        return cls(W, W, np.ones(W.shape[1]), W.shape[1], True, None, W.shape[1], None,
                   None if mask is None else mask_fingerprint(mask), W.shape[0])

    def __and__(self, other):
        return intersect(self, other)


def _warn_on_truncation(s: np.ndarray, k: int) -> None:
    if k < s.size:
        gap = s[k] / s[k - 1]
        if gap > GAP_WARN_RATIO:
            with np.errstate(invalid="ignore", divide="ignore"):
                largest = int(np.nanargmin(s[1:] / s[:-1])) + 1
            warnings.warn(f"filter truncated at rank {k} without a spectral gap: σ_{k + 1}/σ_{k} = {gap:.3g}. "
                          f"The largest gap in the spectrum is at rank {largest} "
                          f"(ratio {s[largest] / s[largest - 1]:.3g})", stacklevel=3)
    else:
        cond = s[0] / s[-1]
        warnings.warn(f"filter kept every singular value (rank {k}); it is invertible and restriction leaves "
                      f"the estimator unchanged, but the restricted covariance carries condition {cond**2:.3g}",
                      stacklevel=3)


def intersect(f1: Filter, f2: Filter) -> Filter:
    """Projector onto range(W1) ∩ range(W2), as the complement of the union of
    the removed subspaces (``filters.py:431-490``)."""
    for f in (f1, f2):
        if not f.is_projector:
            raise ValueError("intersection is defined for projectors only; a graded filter is a transfer "
                             "function, not a subspace")
        if f.Q_removed is None:
            raise ValueError("intersection needs Q_removed, which from_subspace does not form; build the "
                             "operands with from_deprojection")
    if f1.n_pixels != f2.n_pixels:
        raise ValueError(f"pixel-space dimensions differ: {f1.n_pixels} and {f2.n_pixels}")
    if f1.fingerprint is not None and f2.fingerprint is not None and f1.fingerprint != f2.fingerprint:
        raise ValueError("the two filters were built against different active-pixel orderings")
    W, Qr = _complement(np.column_stack([f1.Q_removed, f2.Q_removed]), MIN_RANGE_EPSILON)
    return Filter(W, W, np.ones(W.shape[1]), W.shape[1], True, None, W.shape[1], Qr,
                  f1.fingerprint or f2.fingerprint, f1.n_pixels)


class DeviceFilter:
    """W, U and Sigma in HBM (column-major) and the restrictions of the
    analysis objects; every product runs in libcfmi (FP64)."""

    def __init__(self, ctx: engine.Context, f: Filter, noise_prefiltered: bool = False,
                 maps_prefiltered: bool = False):
        self.ctx, self.f = ctx, f
        self.n, self.r = f.n_pixels, f.rank
        self.W = ctx.upload_cm(f.W)
        self.U = self.W if f.U is f.W else ctx.upload_cm(f.U)
        self.sigma = None if f.is_projector else ctx.upload(np.asarray(f.sigma, dtype=np.float64))
        self.noise_prefiltered = bool(noise_prefiltered)
        self.maps_prefiltered = bool(maps_prefiltered)
        self._tiled = {}

    def _basis(self, prefiltered: bool):
        return (self.U, None) if prefiltered else (self.W, self.sigma)

    def _sigma_tiled(self, k: int):
        if k not in self._tiled:
            self._tiled[k] = self.ctx.upload(np.tile(np.asarray(self.f.sigma, dtype=np.float64), k))
        return self._tiled[k]

    def restrict_matrix(self, X: torch.Tensor, prefiltered: bool = False) -> torch.Tensor:
        """Sigma B^T X B Sigma (r x r, column-major) for an n x n column-major
        X, with B = W (or U when ``prefiltered``, unscaled)."""
        ctx, n, r = self.ctx, self.n, self.r
        B, sig = self._basis(prefiltered)
        T = ctx.empty(n, r)
        engine.gemm(ctx, False, False, n, r, n, 1.0, X, 0, n, B, 0, n, 0.0, T, 0, n)
        M = ctx.empty(r, r)
        # flag 64: lower tiles computed and mirrored, so M is exactly symmetric
        engine.gemm(ctx, True, False, r, r, n, 1.0, B, 0, n, T, 0, n, 0.0, M, 0, r, 64)
        del T
        if sig is not None:
            ctx.check(ctx.lib.cfmi_scale(ctx.h, 0, r, r, P(M), r, P(sig)), "scale")
            ctx.check(ctx.lib.cfmi_scale(ctx.h, 1, r, r, P(M), r, P(sig)), "scale")
        return M

    def restrict_columns(self, X: torch.Tensor, k: int) -> torch.Tensor:
        """Sigma W^T X for an n x k column-major device X (r x k, column-major):
        pixel-space mode vectors into the range (``basis/base.py``'s
        V' = V W Sigma)."""
        ctx, n, r = self.ctx, self.n, self.r
        Y = torch.empty((k, r), dtype=torch.float64, device=ctx.device)
        engine.gemm(ctx, True, False, r, k, n, 1.0, self.W, 0, n, X, 0, n, 0.0, Y, 0, r)
        if self.sigma is not None:
            ctx.check(ctx.lib.cfmi_scale(ctx.h, 1, r, k, P(Y), r, P(self.sigma)), "scale")
        return Y

    def restrict_data(self, maps: np.ndarray, prefiltered: bool = False) -> np.ndarray:
        """Sigma B^T d for (n, nsims) maps (host in, host out)."""
        ctx, n, r = self.ctx, self.n, self.r
        maps = np.asarray(maps, dtype=np.float64)
        if maps.ndim == 1:
            maps = maps[:, None]
        if maps.shape[0] != n:
            raise ValueError(f"maps have {maps.shape[0]} rows, the filter expects {n}")
        B, sig = self._basis(prefiltered)
        ns = maps.shape[1]
        X = ctx.upload_cm(maps)
        Y = ctx.empty(r, ns)
        engine.gemm(ctx, True, False, r, ns, n, 1.0, B, 0, n, X, 0, n, 0.0, Y, 0, r)
        if sig is not None:
            ctx.check(ctx.lib.cfmi_scale(ctx.h, 1, r, ns, P(Y), r, P(sig)), "scale")
        return engine.to_numpy_cm(Y)

    def restrict_stack(self, blk: engine.Block, out: torch.Tensor | None = None) -> torch.Tensor:
        """Sigma W^T dC_b W Sigma for every bin of a block stack, dC_b =
        e_a D_b e_b^T (+ transpose): X_b = W_a^T D_b W_b, then X_b + X_b^T
        (or X_b alone for a diagonal block).  Returns (nbins, r, r)."""
        ctx, n, r = self.ctx, self.n, self.r
        D = blk.stack
        nb, ma, mb = D.shape[0], blk.ma, blk.mb
        W = self.W
        T = torch.empty((nb, r, ma), dtype=torch.float64, device=ctx.device)
        # T_b = D_b W[rb:rb+mb, :]   (ma x r per bin, column-major, ld ma)
        engine.gemm_batched(ctx, False, False, ma, r, mb, 1.0, D, 0, ma, ma * mb, W, blk.rb, n, 0,
                            0.0, T, 0, ma, ma * r, nb)
        R = out if out is not None else torch.empty((nb, r, r), dtype=torch.float64, device=ctx.device)
        # R_b = W[ra:ra+ma, :]^T T_b (symmetric for a diagonal block: flag 64 mirrors it)
        engine.gemm_batched(ctx, True, False, r, r, ma, 1.0, W, blk.ra, n, 0, T, 0, ma, ma * r,
                            0.0, R, 0, r, r * r, nb, 64 if blk.sym else 0)
        if not blk.sym:
            # + T_b^T W[ra:ra+ma, :] = (W_a^T D_b W_b)^T
            engine.gemm_batched(ctx, True, False, r, r, ma, 1.0, T, 0, ma, ma * r, W, blk.ra, n, 0,
                                1.0, R, 0, r, r * r, nb)
        del T
        if self.sigma is not None:
            # the stack is an r x (nb r) column-major matrix: rows by sigma,
            # columns by sigma tiled over the bins
            ctx.check(ctx.lib.cfmi_scale(ctx.h, 1, r, nb * r, P(R), r, P(self.sigma)), "scale")
            ctx.check(ctx.lib.cfmi_scale(ctx.h, 0, r, nb * r, P(R), r, P(self._sigma_tiled(nb))), "scale")
        return R
