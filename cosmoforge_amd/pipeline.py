"""Device orchestration of the estimators (one GPU).

Every step below is a libcfmi call (``engine``); this module only sizes
buffers and walks the spectra / bins in the reference's order:

* ``Covariance.build``   S (``signal_kernels.py:390-537``), C = N + S and its
                         Jacobi-scaled Cholesky (``linalg.py:152-219``)
* ``fisher``             F_ij = 1/2 Tr[C^-1 dC_i C^-1 dC_j] (``fisher.py:362-620``)
                         as the Gram matrix of the whitened derivatives
                         W dC_i W^T, with C^-1 = W^T W
* ``qml``                q_b, noise bias (``spectra.py:1124-1186``) and the
                         F^-1 normalisation (``spectra.py:607-661,1340-1356``)
* ``likelihood``         chi2 and ln|C| (``picslike.py:568-664``)
"""

from __future__ import annotations

import numpy as np
import torch

from . import engine
from .core import Inputs, shape_weights
from .engine import P


class Covariance:
    """Device-resident geometry, noise and (on demand) S, C and its factor."""

    def __init__(self, ctx: engine.Context, inputs: Inputs):
        self.ctx = ctx
        self.inp = inputs
        p = inputs.params
        nest = str(p.ordering).upper() == "NESTED"
        self.vecs = [engine.pointings(ctx, p.nside, f.pixels, nest) for f in inputs.layout.fields]
        self.n = inputs.n
        self.N = ctx.upload_cm(inputs.noise)
        self.S = None
        self.factor = None

    def signal(self, cls=None) -> torch.Tensor:
        ctx = self.ctx
        cs = self.inp.smoothed_cls(cls)
        dev = {k: ctx.upload(v) for k, v in cs.items()}
        self.S = engine.signal_matrix(ctx, self.inp.layout, self.vecs, dev, self.inp.lmax_signal, out=self.S)
        return self.S

    def build(self, cls=None, need_whitener: bool = True) -> engine.Factor:
        """S, then C = S + N factorised in a separate buffer (S is kept)."""
        ctx, n = self.ctx, self.n
        S = self.signal(cls)
        C = ctx.empty(n, n)
        ctx.check(ctx.lib.cfmi_axpby(ctx.h, n, n, 1.0, P(S), n, 0.0, P(C), n), "axpby")
        ctx.check(ctx.lib.cfmi_axpby(ctx.h, n, n, 1.0, P(self.N), n, 1.0, P(C), n), "axpby")
        self.factor = engine.factor(ctx, C, n, need_whitener=need_whitener)
        return self.factor

    def inverse(self) -> torch.Tensor:
        return engine.inverse_from_whitener(self.ctx, self.factor.W, self.n)


def parameter_weights(inp: Inputs, bins, convention: str) -> dict:
    """label -> l-indexed w_l B1(l) B2(l) inside the bins (zero elsewhere)."""
    L = inp.lmax_signal
    sw = shape_weights(max(b[1] for b in bins), convention)
    out = {}
    for lab in inp.layout.labels:
        w = np.zeros(L + 1)
        for lo, hi in bins:
            for ell in range(lo, min(hi, L) + 1):
                w[ell] = sw[ell] * inp.smoothing[lab][ell]
        out[lab] = w
    return out


class StageTimer:
    """HIP events (via torch, on the context's stream) around named stages."""

    def __init__(self):
        self.events = {}

    def mark(self, name: str):
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        self.events.setdefault(name, []).append(ev)

    def elapsed_ms(self, start: str, end: str) -> float:
        return sum(a.elapsed_time(b) for a, b in zip(self.events[start], self.events[end]))


def whitening_flops(n: int, blocks, nbins: int) -> float:
    """Multiply-add flops (x2) of the whitened-derivative GEMMs, counting the
    products the kernels need: the nonzero (lower-triangular) part of W and,
    for the output, its lower triangle only."""
    tot = 0.0
    for blk in blocks:
        ra, ma, rb, mb = blk.ra, blk.ma, blk.rb, blk.mb
        i = np.arange(n - ra)
        z = 2.0 * mb * np.minimum(i + 1, ma).sum()
        j = np.arange(n - rb)
        kj = np.minimum(j + 1, mb)
        if ra == rb and ma == mb:
            x = 2.0 * ((n - ra - j) * kj).sum()
        else:
            x = 2.0 * (n - ra) * kj.sum()
        tot += (z + x) * nbins
    return float(tot)


def fisher(cov: Covariance, bins, convention: str = "Cl", timer: StageTimer | None = None):
    """(F, V): F is (n_p, n_p) on the host, V the device (n_p, n^2) stack of
    whitened derivatives W dC_p W^T (kept for QML)."""
    ctx, inp, n = cov.ctx, cov.inp, cov.n
    lay = inp.layout
    for lo, hi in bins:
        if hi > inp.lmax_signal:
            raise ValueError(f"bin [{lo}, {hi}] extends beyond lmax_signal={inp.lmax_signal}")
    w = parameter_weights(inp, bins, convention)
    dev_w = {k: ctx.upload(v) for k, v in w.items()}
    if timer:
        timer.mark("deriv0")
    blocks = engine.deriv_stacks(ctx, lay, cov.vecs, dev_w, inp.lmax_signal, bins)
    nb = len(bins)
    npar = len(blocks) * nb
    K = engine.packed_size(n)
    V = torch.zeros((npar, K), dtype=torch.float64, device=ctx.device)
    if timer:
        timer.mark("white0")
        timer.flops = whitening_flops(n, blocks, nb)
    engine.whitened_derivatives(ctx, cov.factor.W, n, blocks, nb, V)
    if timer:
        timer.mark("white1")
    del blocks
    G = engine.gram(ctx, npar, npar, K, V, V)
    if timer:
        timer.mark("gram1")
    F = 0.5 * (G + G.T)
    return F, V


def whitened_maps(cov: Covariance, maps: np.ndarray) -> torch.Tensor:
    """z = W x for every simulation (n x nsims, column-major)."""
    ctx, n = cov.ctx, cov.n
    X = ctx.upload_cm(maps)
    ns = maps.shape[1]
    Z = ctx.empty(n, ns)
    engine.gemm(ctx, False, False, n, ns, n, 1.0, cov.factor.W, 0, n, X, 0, n, 0.0, Z, 0, n, 1)
    return Z


def qml(cov: Covariance, V: torch.Tensor, maps: np.ndarray, remove_nb: bool = True):
    """(q, nb): raw QML estimates (nsims, n_p) and noise bias (n_p,)."""
    ctx, n = cov.ctx, cov.n
    W = cov.factor.W
    ns = maps.shape[1]
    npar = V.shape[0]
    Z = whitened_maps(cov, maps)
    K = engine.packed_size(n)
    # Packed like V (diagonal / sqrt 2), so <pack(A), V_b> = 1/2 <A, W dC_b W^T>.
    Q = torch.empty((ns + 1, K), dtype=torch.float64, device=ctx.device)
    for s in range(ns):
        engine.gemm_packed(ctx, False, True, n, n, 1, 1.0, Z, s * n, n, Z, s * n, n, 0.0, Q[s], n, 0, 0)
    T = ctx.empty(n, n)
    engine.gemm(ctx, False, False, n, n, n, 1.0, W, 0, n, cov.N, 0, n, 0.0, T, 0, n, 1)
    engine.gemm_packed(ctx, False, True, n, n, n, 1.0, T, 0, n, W, 0, n, 0.0, Q[ns], n, 0, 0, flags=4)
    del T
    G = engine.gram(ctx, ns + 1, npar, K, Q, V)
    nb = G[ns]
    q = G[:ns]
    if remove_nb:
        q = q - nb[None, :]
    return q, nb


def inverse_symm(ctx: engine.Context, M: np.ndarray) -> np.ndarray:
    """Jacobi-scaled Cholesky inverse of a small SPD matrix on the device
    (``matrix_inverse_symm``); raises ValueError("dpotrf failed with info=k")."""
    m = M.shape[0]
    if M.shape != (m, m):
        raise ValueError("Matrix must be square")
    dM = ctx.upload_cm(np.asarray(M, dtype=np.float64))
    f = engine.factor(ctx, dM, m)
    return engine.to_numpy_cm(engine.inverse_from_whitener(ctx, f.W, m))


def likelihood(cov: Covariance, maps: np.ndarray, cls=None):
    """(chi2 (nsims,), logdet) of one parameter point."""
    f = cov.build(cls)
    Z = whitened_maps(cov, maps)
    ns = maps.shape[1]
    G = engine.gram(cov.ctx, ns, ns, cov.n, Z, Z)
    return np.diag(G).copy(), f.logdet
