"""Harmonic-mode contraction of the Fisher and QML traces (``basis="harmonic"``).

The reference's harmonic route (``src/cosmoforge.qube/qube/fisher.py:416-500``,
``cosmocore/basis/harmonic_basis.py:19-182``) writes every binned derivative
as a sum of rank-one terms over real harmonic modes,
dC_b = sum_{l in b} w_l sum_m (x_lm y_lm^T [+ y_lm x_lm^T]), so that

* F_pq = 1/2 sum_{modes of p, q} w w' sum_pairs M[v, u'] M[u, v'] with
  M = V^T C^-1 V = (W V)^T (W V), W the whitener of C = S + N;
* q_p = 1/2 sum w sum_pairs (u^T y)(v^T y) with y = C^-1 x, and
  nb_p = 1/2 sum w sum_pairs u^T G v with G = C^-1 N C^-1, both from
  R = C^-1 V = W^T Z (u^T y = R_u^T x, u^T G v = R_u^T N R_v).

C and its factor are the pixel path's (S from the pixel recurrences).  The
device work is three FP64 GEMMs (Z = W V, M = Z^T Z, R = W^T Z) and
HBM-bound mode reductions (``cfmi_harmonic_fisher``, ``cfmi_harmonic_pairsum``), instead of
the pixel path's n^2 x n products per parameter: about 30x less work at
nside 16 and 60x at nside 32.

``precision="mixed"``: the three products run on the int8 matrix cores with
sliced operands (``engine.gemm_s8``: exact integer sums, operands held to
2^-32 / 2^-40 of their row maxima), and without the whitener: Z = L^-1 D V
and R = D L^-T Z as blocked triangular solves with L, the Cholesky factor of
D C D (FP64, as are the diagonal-block solves, ln|C| and y = C^-1 x).  The
factor's rows are flat where the whitener's are dominated by the diagonal,
which is what lets few slices reach ~1e-9 on F.

The modes span l in [first bin's l_min, last bin's l_max]; the derivative of
the pixel path carries the reference's acos-based rotation angles, which
agree with the exact spin-2 harmonics to ~1e-8 relative, so the two routes
agree to that level on polarisation (to rounding on temperature).
"""

from __future__ import annotations

import numpy as np
import torch

from . import engine, pipeline
from .engine import P


class HarmonicContraction:
    """Mode operator V (n x N_modes, column-major, in HBM) and the tables of
    one (Covariance, bins) pair."""

    # int8 slices of the mixed products: Z = L^-1 D V, M = Z^T Z, R = D L^-T Z
    MIXED_SLICES = (4, 5, 4)
    TRSM_BLOCK = 512          # diagonal blocks of the FP64 solves
    # mixed: 2048-row diagonal blocks, their solves (products with the
    # block inverses, K = 2048) on 5 int8 slices; the factorisation's
    # trailing updates on 5 slices (Covariance.build(slices=...))
    MIXED_TRSM_BLOCK = 2048
    TRSM_OUTER = 2048         # rows per outer block of the blocked solves (engine.trsm), FP64
    MIXED_TRSM_OUTER = 4096   # mixed: each right-hand-side pass covers more of L (measured 3% faster)
    DIAG_SLICES = 5
    FACTOR_SLICES = 5

    def __init__(self, cov: pipeline.Covariance, bins, convention: str = "Cl",
                 cov2: pipeline.Covariance | None = None, precision: str = "fp64", slices=None):
        """``cov2``: C2 = S + N2 of a cross analysis (``do_cross``);
        ``precision``: "fp64" or "mixed" (auto analyses without a filter)."""
        self.cov, self.bins = cov, [(int(a), int(b)) for a, b in bins]
        self.cov2 = cov2
        if precision not in pipeline.PRECISIONS:
            raise ValueError(f"precision must be one of {pipeline.PRECISIONS}; got {precision!r}")
        if precision == "mixed" and (cov2 is not None or getattr(cov, "pixel_filter", None) is not None):
            raise ValueError("the mixed harmonic contraction covers auto analyses without a pixel filter")
        self.precision = precision
        self.slices = tuple(slices) if slices is not None else self.MIXED_SLICES
        if precision == "mixed" and getattr(cov.factor, "nb", 0) == 0:
            raise ValueError("the mixed harmonic contraction needs the factor's diagonal-block inverses: "
                             "build the covariance with diag_blocks")
        ctx, inp = cov.ctx, cov.inp
        lay = inp.layout
        self.lmin = min(a for a, _ in self.bins)
        self.lmax = max(b for _, b in self.bins)
        if self.lmin < 2:
            raise ValueError("harmonic modes start at l = 2")
        if self.lmax > inp.lmax_signal:
            raise ValueError(f"bins reach l = {self.lmax} above lmax_signal = {inp.lmax_signal}")
        self.ng = (self.lmax + 1) ** 2 - self.lmin ** 2
        # one mode group per field-map label: T, or E and B of a spin-2 field
        self.group = {}
        for f in lay.fields:
            for lab in f.labels:
                self.group[lab.upper()] = len(self.group)
        self.nt = self.ng * len(self.group)
        # HBM: V (n x N, twice while a filter restricts it), Z and M (N x N),
        # then R and N R; C's factor and W are already resident
        need = 8 * (3 * inp.n * self.nt + self.nt * self.nt) * (2 if cov2 is not None else 1)
        free = pipeline._free_bytes(ctx)
        if need > free:
            raise MemoryError(f"the harmonic contraction needs {need / 1e9:.1f} GB of device memory for "
                              f"{self.nt} modes over {inp.n} rows, {free / 1e9:.1f} GB are free; use basis=False "
                              "(the pixel route bands its products)")
        # modes on the active pixels; under a pixel filter they are then
        # restricted to its range, V' = Sigma W^T V (r x N)
        flt = getattr(cov, "pixel_filter", None)
        n = inp.n
        self.V = torch.zeros((self.nt, n), dtype=torch.float64, device=ctx.device)   # column-major n x N
        for fi, f in enumerate(lay.fields):
            vec = cov.vecs[fi]
            if f.spin == 0:
                g = self.group[f.labels[0].upper()]
                rc = ctx.lib.cfmi_harmonic_modes(ctx.h, 0, f.npix, P(vec), self.lmin, self.lmax, P(self.V), n,
                                                 self.nt, f.offset, 0, g * self.ng, 0)
            else:
                ge, gb = (self.group[x.upper()] for x in f.labels)
                rc = ctx.lib.cfmi_harmonic_modes(ctx.h, 2, f.npix, P(vec), self.lmin, self.lmax, P(self.V), n,
                                                 self.nt, f.offset, f.offset + f.npix, ge * self.ng, gb * self.ng)
            ctx.check(rc, "harmonic_modes")
        if flt is not None:
            self.V = flt.restrict_columns(self.V, self.nt)
        specs = lay.spectra_list()
        pairs = np.full((len(specs), 4), -1, dtype=np.int32)
        for k, sp in enumerate(specs):
            a, b = lay.split_label(sp.label, sp.fi, sp.fj)
            ga, gb = self.group[a.upper()], self.group[b.upper()]
            pairs[k, :2] = (ga, ga) if ga == gb else (ga, gb)
            if ga != gb:
                pairs[k, 2:] = (gb, ga)
        wts = pipeline.parameter_weights(inp, self.bins, convention)
        self.lw = inp.lmax_signal + 1
        w = np.stack([wts[sp.label][: self.lw] for sp in specs])
        self.nspec, self.nbins = len(specs), len(self.bins)
        self.npar = self.nspec * self.nbins
        self._pairs = ctx.upload(pairs, dtype=torch.int32)
        self._goff = ctx.upload(np.arange(len(self.group), dtype=np.int64) * self.ng, dtype=torch.int64)
        self._w = ctx.upload(np.ascontiguousarray(w, dtype=np.float64))
        self._lo = ctx.upload(np.array([a for a, _ in self.bins], dtype=np.int32), dtype=torch.int32)
        self._hi = ctx.upload(np.array([b for _, b in self.bins], dtype=np.int32), dtype=torch.int32)
        self._Z = {}

    @property
    def flops(self) -> float:
        """FP64 flops of Z = W V and R = W^T Z (W triangular) and M = Z^T Z
        (its lower half)."""
        n, nt = self.cov.n, self.nt
        return 2.0 * n * n * nt + float(n) * nt * (nt + 1)

    def _tables(self):
        return (self.lmin, self.nbins, P(self._lo), P(self._hi), self.nspec, P(self._pairs), P(self._goff),
                P(self._w), self.lw)

    def _z(self, cov) -> torch.Tensor:
        """Z = W V = L^-1 D V (n x N, column-major) of ``cov``, kept for the
        QML forms."""
        key = id(cov)
        if key not in self._Z:
            ctx, n, nt = cov.ctx, cov.n, self.nt
            Z = torch.empty((nt, n), dtype=torch.float64, device=ctx.device)
            if self._blocked(cov):
                B = self.V.clone()
                ctx.check(ctx.lib.cfmi_scale(ctx.h, 1, n, nt, P(B), n, P(cov.factor.d)), "scale")
                engine.trsm(ctx, cov.factor, n, B, Z, nt, slices=self._s(0), outer=self._outer(),
                            diag_slices=self._ds())
                del B
            else:
                engine.gemm(ctx, False, False, n, nt, n, 1.0, cov.factor.W, 0, n, self.V, 0, n, 0.0, Z, 0, n, 1)
            self._Z[key] = Z
        return self._Z[key]

    def _m(self, cov) -> torch.Tensor:
        ctx, n, nt = cov.ctx, cov.n, self.nt
        Z = self._z(cov)
        M = torch.empty((nt, nt), dtype=torch.float64, device=ctx.device)
        if self.precision == "mixed":
            Zs = engine.slice8(ctx, self.slices[1], 0, nt, n, Z, 0, n)
            engine.gemm_s8(ctx, nt, nt, n, 1.0, Zs, 0, Zs, 0, 0.0, M, 0, nt, flags=1)
            del Zs
        else:
            engine.gram(ctx, nt, nt, n, Z, Z, out=M)
        return M

    def _r(self, cov) -> torch.Tensor:
        """R = C^-1 V = W^T Z = D L^-T Z (column-major n x N); releases Z."""
        ctx, n, nt = cov.ctx, cov.n, self.nt
        Z = self._z(cov)
        R = torch.empty((nt, n), dtype=torch.float64, device=ctx.device)
        if self._blocked(cov):
            del self._Z[id(cov)]
            engine.trsm(ctx, cov.factor, n, Z, R, nt, transpose=True, slices=self._s(2), outer=self._outer(),
                        diag_slices=self._ds())
            del Z
            ctx.check(ctx.lib.cfmi_scale(ctx.h, 1, n, nt, P(R), n, P(cov.factor.d)), "scale")
            return R
        engine.gemm(ctx, True, False, n, nt, n, 1.0, cov.factor.W, 0, n, Z, 0, n, 0.0, R, 0, n, 2)
        del self._Z[id(cov)]
        return R

    def _blocked(self, cov) -> bool:
        """Z and R by blocked solves with L (the factor keeps L's diagonal
        blocks inverted and no whitener), else by products with W."""
        return cov.factor.W is None and getattr(cov.factor, "nb", 0) > 0

    def _s(self, i: int) -> int:
        """Slices of product i (0 when the product runs in FP64)."""
        return self.slices[i] if self.precision == "mixed" else 0

    def _outer(self) -> int:
        return self.MIXED_TRSM_OUTER if self.precision == "mixed" else self.TRSM_OUTER

    def _ds(self) -> int:
        """Slices of the blocked solves' diagonal products (long K only)."""
        nb = getattr(self.cov.factor, "nb", 0)
        return self.DIAG_SLICES if self.precision == "mixed" and nb >= self.MIXED_TRSM_BLOCK else 0

    @staticmethod
    def _apply_inv(cov, X: torch.Tensor, ns: int) -> torch.Tensor:
        """D L^-T L^-1 D X in FP64 (X is consumed)."""
        ctx, n = cov.ctx, cov.n
        ctx.check(ctx.lib.cfmi_scale(ctx.h, 1, n, ns, P(X), n, P(cov.factor.d)), "scale")
        T = ctx.empty(n, ns)
        engine.trsm(ctx, cov.factor, n, X, T, ns)
        Y = ctx.empty(n, ns)
        engine.trsm(ctx, cov.factor, n, T, Y, ns, transpose=True)
        ctx.check(ctx.lib.cfmi_scale(ctx.h, 1, n, ns, P(Y), n, P(cov.factor.d)), "scale")
        return Y

    def _solve(self, cov, maps: np.ndarray) -> torch.Tensor:
        """y = C^-1 x (column-major n x nsims), FP64 solves with the factor.
        A factor with sliced trailing updates reproduces C only to ~2^-(8 s)
        relative, so y then takes one step of FP64 refinement against
        C = S + N: y += C~^-1 (x - S y - N y)."""
        ctx, n = cov.ctx, cov.n
        ns = maps.shape[1]
        Y = self._apply_inv(cov, ctx.upload_cm(maps), ns)
        if cov.factor.slices:
            Rr = ctx.upload_cm(maps)
            engine.gemm(ctx, False, False, n, ns, n, -1.0, cov.S, 0, n, Y, 0, n, 1.0, Rr, 0, n)
            if cov.noise_is_diagonal():
                # N y for a diagonal N: y's rows scaled by N's diagonal
                NY = Y.clone()
                d = torch.empty(n, dtype=torch.float64, device=ctx.device)
                ctx.check(ctx.lib.cfmi_copy(ctx.h, engine.F64, 1, n, P(cov.N), n + 1, P(d), 1), "copy")
                ctx.check(ctx.lib.cfmi_scale(ctx.h, 1, n, ns, P(NY), n, P(d)), "scale")
                ctx.check(ctx.lib.cfmi_axpby(ctx.h, n, ns, -1.0, P(NY), n, 1.0, P(Rr), n), "axpby")
                del NY, d
            else:
                engine.gemm(ctx, False, False, n, ns, n, -1.0, cov.N, 0, n, Y, 0, n, 1.0, Rr, 0, n)
            D = self._apply_inv(cov, Rr, ns)
            ctx.check(ctx.lib.cfmi_axpby(ctx.h, n, ns, 1.0, P(D), n, 1.0, P(Y), n), "axpby")
        return Y

    def _ut(self, R: torch.Tensor, maps: np.ndarray) -> torch.Tensor:
        """u^T C^-1 x = R^T x for every mode and simulation (column-major ns x N)."""
        ctx, n, nt = self.cov.ctx, self.cov.n, self.nt
        ns = maps.shape[1]
        X = ctx.upload_cm(maps)
        Ut = torch.empty((nt, ns), dtype=torch.float64, device=ctx.device)
        engine.gemm(ctx, True, False, ns, nt, n, 1.0, X, 0, n, R, 0, n, 0.0, Ut, 0, ns)
        return Ut

    def fisher(self) -> np.ndarray:
        """F = 1/2 Tr[C^-1 dC_p C^-1 dC_q], or 1/2 Tr[C1^-1 dC_p C2^-1 dC_q]
        with ``cov2`` (``fisher.py:522-531,584-590``)."""
        ctx = self.cov.ctx
        M = self._m(self.cov)
        M2 = self._m(self.cov2) if self.cov2 is not None else None
        F = torch.empty((self.npar, self.npar), dtype=torch.float64, device=ctx.device)
        ctx.check(ctx.lib.cfmi_harmonic_fisher(ctx.h, P(M), P(M2) if M2 is not None else None, self.nt,
                                               *self._tables(), P(F)), "harmonic_fisher")
        del M, M2
        return engine.to_numpy_cm(F)

    @staticmethod
    def _maps(maps, n):
        maps = np.asarray(maps, dtype=np.float64)
        if maps.ndim == 1:
            maps = maps[:, None]
        if maps.shape[0] != n:
            raise ValueError(f"maps have {maps.shape[0]} rows, expected {n}")
        return maps

    def qml(self, maps: np.ndarray, remove_nb: bool = True, maps2: np.ndarray | None = None):
        """(q (nsims, n_p), nb (n_p,)) as ``pipeline.qml``.  With R = C^-1 V =
        W^T Z: u^T y = R_u^T x and u^T G v = R_u^T N R_v, so neither y nor G is
        formed; a diagonal N scales the rows of R.  With ``cov2`` and
        ``maps2``: q = 1/2 y2^T dC y1 and no noise bias (``spectra.py:1146-1166``)."""
        ctx, n, nt = self.cov.ctx, self.cov.n, self.nt
        maps = self._maps(maps, n)
        ns = maps.shape[1]
        q = torch.empty((ns, self.npar), dtype=torch.float64, device=ctx.device)  # column-major npar x ns
        if self.cov2 is not None:
            if maps2 is None:
                raise ValueError("a cross analysis needs maps2")
            maps2 = self._maps(maps2, n)
            if maps2.shape != maps.shape:
                raise ValueError(f"maps1 {maps.shape} and maps2 {maps2.shape} differ")
            Ut1 = self._ut(self._r(self.cov), maps)
            Ut2 = self._ut(self._r(self.cov2), maps2)
            ctx.check(ctx.lib.cfmi_harmonic_pairsum(ctx.h, 1, P(Ut2), ns, 1, P(Ut1), ns, 1, ns, *self._tables(),
                                                    0.5, P(q), self.npar), "harmonic_pairsum")
            return q.cpu().numpy(), np.zeros(self.npar)
        R = self._r(self.cov)
        if self._blocked(self.cov):
            # u^T y from y = C^-1 x solved in FP64 (rather than from a sliced R)
            Y = self._solve(self.cov, maps)
            Ut = torch.empty((nt, ns), dtype=torch.float64, device=ctx.device)
            engine.gemm(ctx, True, False, ns, nt, n, 1.0, Y, 0, n, self.V, 0, n, 0.0, Ut, 0, ns)
            del Y
        else:
            Ut = self._ut(R, maps)
        ctx.check(ctx.lib.cfmi_harmonic_pairsum(ctx.h, 1, P(Ut), ns, 1, P(Ut), ns, 1, ns, *self._tables(), 0.5,
                                                P(q), self.npar), "harmonic_pairsum")
        # nb_p = 1/2 sum w sum_pairs R_u^T N R_v: a diagonal N weights the
        # rows of R (no N R formed), a dense one is applied by a product
        nb = torch.empty(self.npar, dtype=torch.float64, device=ctx.device)
        if self.cov.noise_is_diagonal():
            d = torch.empty(n, dtype=torch.float64, device=ctx.device)
            ctx.check(ctx.lib.cfmi_copy(ctx.h, engine.F64, 1, n, P(self.cov.N), n + 1, P(d), 1), "copy")
            X, wrow = R, d
        else:
            X = torch.empty_like(R)
            engine.gemm(ctx, False, False, n, nt, n, 1.0, self.cov.N, 0, n, R, 0, n, 0.0, X, 0, n)
            wrow = None
        ctx.check(ctx.lib.cfmi_harmonic_coldots(ctx.h, n, P(X), n, P(R), n, P(wrow), len(self.group), self.ng,
                                                *self._tables(), 0.5, P(nb)), "harmonic_coldots")
        del R, X
        qh = q.cpu().numpy()              # (ns, npar): row s holds simulation s
        nbh = nb.cpu().numpy()
        if remove_nb:
            qh = qh - nbh[None, :]
        return qh, nbh
