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

# Cuts on singular values are relative to the largest one.  Below sqrt(eps)
# the restricted covariance, which carries sigma^2, loses every digit.
MIN_RANGE_EPSILON = float(np.sqrt(np.finfo(np.float64).eps))
# A cut is "clean" when the first dropped singular value is at most this
# fraction of the last kept one.
GAP_WARN_RATIO = 0.1
# Tolerance on |sigma - 1| and on the agreement of the two singular bases
# below which a factored operator counts as an orthogonal projector.
PROJECTOR_TOL = 1e-8


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


def _fingerprint_or_none(mask):
    return None if mask is None else mask_fingerprint(mask)


def _split_space(A: np.ndarray, rel_cut: float):
    """Orthonormal bases (span, complement) of the column space of A: the left
    singular vectors whose singular value exceeds ``rel_cut`` times the
    largest, and the remaining ones of the full pixel space."""
    n = A.shape[0]
    left, sv, _ = np.linalg.svd(A, full_matrices=True)
    keep = 0 if sv.size == 0 or sv[0] == 0.0 else int(np.sum(sv > rel_cut * sv[0]))
    return np.ascontiguousarray(left[:, :keep]), np.ascontiguousarray(left[:, keep:n])


def _pick_rank(sv: np.ndarray, range_epsilon, range_rank):
    """(rank, effective epsilon) of a truncated SVD; raises ValueError on an
    unusable knob.  ``range_epsilon`` keeps sigma_i > eps sigma_1 (eps clipped
    up to MIN_RANGE_EPSILON with a warning), ``range_rank`` keeps the leading
    singular values, which must stay above MIN_RANGE_EPSILON sigma_1."""
    if (range_epsilon is None) == (range_rank is None):
        raise ValueError("from_operator needs exactly one truncation knob: range_epsilon or range_rank "
                         "(both were %s)" % ("missing" if range_epsilon is None else "given"))
    top = sv[0] if sv.size else 0.0
    if range_rank is not None:
        k = int(range_rank)
        if k < 1 or k > sv.size:
            raise ValueError(f"range_rank={k} is not in 1..{sv.size}")
        if sv[k - 1] <= MIN_RANGE_EPSILON * top:
            raise ValueError(f"range_rank={k} reaches sigma_k / sigma_1 = {sv[k - 1] / top:.3g}, not above "
                             f"MIN_RANGE_EPSILON = {MIN_RANGE_EPSILON:.3g}")
        return k, None
    eps = float(range_epsilon)
    if eps < MIN_RANGE_EPSILON:
        warnings.warn(f"range_epsilon={eps:g} raised (clipped) to MIN_RANGE_EPSILON = {MIN_RANGE_EPSILON:.3g}",
                      stacklevel=3)
        eps = MIN_RANGE_EPSILON
    k = int(np.sum(sv > eps * top))
    if k == 0:
        raise ValueError(f"range_epsilon={eps:g} keeps no singular value")
    return k, eps


def _gap_warning(sv: np.ndarray, k: int) -> None:
    """Warn when the cut at rank k is not at a spectral gap, or when nothing
    was cut (an invertible filter leaves the estimator unchanged)."""
    if k == sv.size:
        warnings.warn(f"every one of the {k} singular values was kept: the filter is invertible "
                      f"(restricted condition number {(sv[0] / sv[-1]) ** 2:.3g})", stacklevel=3)
        return
    ratio = sv[k] / sv[k - 1]
    if ratio > GAP_WARN_RATIO:
        with np.errstate(divide="ignore", invalid="ignore"):
            steps = sv[1:] / sv[:-1]
        best = int(np.nanargmin(steps)) + 1
        warnings.warn(f"rank-{k} truncation is not at a spectral gap (sigma_{k + 1} / sigma_{k} = {ratio:.3g}); "
                      f"the sharpest drop is after rank {best} ({steps[best - 1]:.3g})", stacklevel=3)


class Filter:
    """A pixel filter F = U diag(sigma) W^T, truncated at rank r (the public
    record of ``filters.py:82-170``): ``W`` and ``U`` are (n_pixels, r) with
    orthonormal columns, ``sigma`` the kept singular values.  For an
    orthogonal projector U and W are one array and sigma is all ones;
    ``Q_removed`` spans the removed directions when they were formed."""

    _FIELDS = ("W", "U", "sigma", "rank", "is_projector", "range_epsilon", "range_rank", "Q_removed",
               "fingerprint", "n_pixels")
    __slots__ = _FIELDS

    def __init__(self, W, U, sigma, rank, is_projector, range_epsilon, range_rank, Q_removed, fingerprint,
                 n_pixels):
        shape = (int(n_pixels), int(rank))
        if np.ndim(W) != 2 or W.shape != shape:
            raise ValueError(f"W must have shape {shape}; got {np.shape(W)}")
        if np.shape(U) != shape:
            raise ValueError(f"U must have shape {shape} like W; got {np.shape(U)}")
        if np.shape(sigma) != (shape[1],):
            raise ValueError(f"sigma must have shape ({shape[1]},); got {np.shape(sigma)}")
        if is_projector and U is not W:
            raise ValueError("a projector keeps one basis: pass the same array as U and W")
        for name, value in zip(self._FIELDS, (W, U, sigma, int(rank), bool(is_projector), range_epsilon,
                                              range_rank, Q_removed, fingerprint, int(n_pixels))):
            object.__setattr__(self, name, value)

    def __setattr__(self, key, value):
        raise AttributeError(f"cannot set {key!r}: Filter is immutable")

    def __repr__(self):
        cut = f"range_epsilon={self.range_epsilon:g}" if self.range_epsilon is not None else \
            f"range_rank={self.range_rank}"
        return (f"Filter({'projector' if self.is_projector else 'graded'}, n_pixels={self.n_pixels}, "
                f"rank={self.rank}, {cut})")

    @property
    def scaling(self):
        """The kept singular values of a graded filter; None for a projector."""
        return None if self.is_projector else self.sigma

    @classmethod
    def _projector(cls, basis, removed, range_rank, mask):
        """The orthogonal projector onto span(basis)."""
        r = basis.shape[1]
        return cls(basis, basis, np.ones(r), r, True, None, range_rank, removed, _fingerprint_or_none(mask),
                   basis.shape[0])

    @classmethod
    def from_operator(cls, F, *, range_epsilon=None, range_rank=None, mask=None) -> "Filter":
        """Factor a dense (n, n) operator by its SVD and truncate it
        (``filters.py:209-311``).  An operator whose kept part is an
        orthogonal projector (unit singular values, equal left and right
        bases) is stored as one, with its removed directions."""
        F = np.asarray(F, dtype=np.float64)
        if F.ndim != 2 or F.shape[0] != F.shape[1]:
            raise ValueError(f"F must be a square (n_pixels, n_pixels) array; got shape {F.shape}")
        left, sv, right_t = np.linalg.svd(F)
        k, eff = _pick_rank(sv, range_epsilon, range_rank)
        _gap_warning(sv, k)
        W = np.ascontiguousarray(right_t[:k].T)
        U = np.ascontiguousarray(left[:, :k])
        rr = None if range_rank is None else int(range_rank)
        fp = _fingerprint_or_none(mask)
        # projector: F restricted to its kept range acts as the identity there
        unit = np.max(np.abs(sv[:k] - 1.0)) <= PROJECTOR_TOL
        if unit and np.max(np.abs(U.T @ W - np.eye(k))) <= PROJECTOR_TOL:
            return cls(W, W, np.ones(k), k, True, eff, rr, np.ascontiguousarray(right_t[k:].T), fp, F.shape[0])
        return cls(W, U, np.ascontiguousarray(sv[:k]), k, False, eff, rr, None, fp, F.shape[0])

    @classmethod
    def from_deprojection(cls, T, *, template_epsilon=1e-3, mask=None) -> "Filter":
        """The orthogonal projector that removes the span of the template
        columns T (n, n_templates), counting only directions above
        ``template_epsilon`` times the strongest (``filters.py:313-356``)."""
        T = np.asarray(T, dtype=np.float64)
        if T.ndim != 2:
            raise ValueError(f"T must be a 2-D (n_pixels, n_templates) array; got shape {T.shape}")
        removed, kept = _split_space(T, float(template_epsilon))
        if kept.shape[1] == 0:
            raise ValueError(f"the templates span all {T.shape[0]} pixels: deprojecting them leaves a rank-0 "
                             "filter")
        return cls._projector(kept, removed, kept.shape[1], mask)

    @classmethod
    def from_subspace(cls, W, *, mask=None) -> "Filter":
        """The orthogonal projector onto the span of an orthonormal (n, r)
        basis (``filters.py:358-400``)."""
        W = np.ascontiguousarray(W, dtype=np.float64)
        if W.ndim != 2:
            raise ValueError(f"W must be a 2-D (n_pixels, rank) array; got shape {W.shape}")
        n, r = W.shape
        if r > n:
            raise ValueError(f"a subspace of {n} pixels cannot have {r} basis vectors")
        if np.max(np.abs(W.T @ W - np.eye(r)), initial=0.0) > PROJECTOR_TOL:
            raise ValueError("W must have orthonormal columns (W^T W = I)")
        return cls._projector(W, None, r, mask)

    def __and__(self, other):
        return intersect(self, other)


def intersect(f1: Filter, f2: Filter) -> Filter:
    """The projector onto range(f1) ∩ range(f2): the orthogonal complement of
    the union of the two removed subspaces (``filters.py:431-490``)."""
    for f in (f1, f2):
        if not f.is_projector:
            raise ValueError("only projectors can be intersected (a graded filter is not a subspace)")
        if f.Q_removed is None:
            raise ValueError("intersect needs each filter's removed directions (Q_removed); from_subspace "
                             "does not form them, from_deprojection and from_operator do")
    if f1.n_pixels != f2.n_pixels:
        raise ValueError(f"the filters act on {f1.n_pixels} and {f2.n_pixels} pixels")
    if None not in (f1.fingerprint, f2.fingerprint) and f1.fingerprint != f2.fingerprint:
        raise ValueError("the filters were built for different active-pixel orderings")
    removed, kept = _split_space(np.hstack([f1.Q_removed, f2.Q_removed]), MIN_RANGE_EPSILON)
    if kept.shape[1] == 0:
        raise ValueError("the two removed subspaces together span every pixel: the intersection is empty")
    fp = f1.fingerprint if f1.fingerprint is not None else f2.fingerprint
    r = kept.shape[1]
    return Filter(kept, kept, np.ones(r), r, True, None, r, removed, fp, f1.n_pixels)


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
