"""Device pipeline of the pixel-space path on one MI355X.

PyTorch-ROCm tensors are only device buffers here; every computation is a
libcfmi kernel called through the C ABI.  Matrices are column-major: an
``m x n`` matrix is a torch tensor of shape ``(n, m)`` (its row-major view).

Pipeline (reference functions in parentheses):

1. pointings of active pixels (``compute_pointings``)
2. S from the pair recurrences (``compute_signal_matrix``)
3. C = S + N, D C D Jacobi scaling, Cholesky L L^T, U = L^-1, and the
   whitener W = U D with C^-1 = W^T W (``matrix_inverse_symm``)
4. per spectrum, all bins of dC_b in one recurrence pass
   (``get_binned_derivative_direct``)
5. Fisher: F_ij = 1/2 Tr[C^-1 dC_i C^-1 dC_j] = 1/2 <W dC_i W^T, W dC_j W^T>
   (``fisher.py:506-603``, reassociated through the Cholesky factor)
6. QML: with z = W x and Nw = W N W^T,
   q_b = 1/2 z^T (W dC_b W^T) z - 1/2 <Nw, W dC_b W^T>
   (``spectra.py:1124-1186``: y = C^-1 x, 1/2 y^T dC_b y - 1/2 Tr[C^-1 N C^-1 dC_b])
7. PICSLike: chi2 = |W x|^2, ln|C| from diag(L) (``picslike.py:643-664``).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .layout import Layout

F64 = _lib.CFMI_F64
F32 = _lib.CFMI_F32


class LinAlgError(ValueError):
    """Cholesky failure; carries the LAPACK-style info."""

    def __init__(self, info: int):
        super().__init__(f"dpotrf failed with info={info}")
        self.info = info


class Context:
    """A libcfmi context bound to torch's current stream on one device."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        if not torch.cuda.is_available():
            raise RuntimeError("cosmoforge_amd needs a HIP device (torch.cuda.is_available() is False)")
        self.device = torch.device("cuda", device)
        with torch.cuda.device(self.device):
            self.stream = torch.cuda.current_stream(self.device)
        h = ctypes.c_void_p()
        rc = self.lib.cfmi_create(device, ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(h))
        if rc != 0:
            raise _lib.CfmiError(f"cfmi_create failed with code {rc}")
        self.h = h

    def check(self, rc: int, what: str) -> int:
        if rc < 0:
            msg = self.lib.cfmi_last_error(self.h)
            raise _lib.CfmiError(f"{what}: {msg.decode() if msg else rc}")
        return rc

    def sync(self):
        self.check(self.lib.cfmi_synchronize(self.h), "synchronize")

    def close(self):
        if getattr(self, "h", None):
            self.lib.cfmi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- allocation helpers --------------------------------------------------
    def empty(self, m: int, n: int, dtype=torch.float64) -> torch.Tensor:
        return torch.empty((n, m), dtype=dtype, device=self.device)

    def zeros(self, m: int, n: int, dtype=torch.float64) -> torch.Tensor:
        return torch.zeros((n, m), dtype=dtype, device=self.device)

    def upload(self, a, dtype=torch.float64) -> torch.Tensor:
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(self.device)

    def upload_cm(self, a: np.ndarray, dtype=torch.float64) -> torch.Tensor:
        """Host (m, n) array -> column-major device matrix."""
        a = np.asarray(a)
        return torch.as_tensor(np.ascontiguousarray(a.T), dtype=dtype).to(self.device)


def P(t: torch.Tensor, off: int = 0) -> ctypes.c_void_p:
    """Device address of element ``off`` of a tensor."""
    if t is None:
        return ctypes.c_void_p(0)
    return ctypes.c_void_p(t.data_ptr() + off * t.element_size())


def to_numpy_cm(t: torch.Tensor) -> np.ndarray:
    """Column-major device matrix -> host (m, n) array."""
    return t.detach().cpu().numpy().T.copy()


def dtype_code(t: torch.Tensor) -> int:
    if t.dtype == torch.float64:
        return F64
    if t.dtype == torch.float32:
        return F32
    raise TypeError(f"unsupported dtype {t.dtype}")


# ---------------------------------------------------------------------------
# A1 - A9: geometry and signal
# ---------------------------------------------------------------------------

def pointings(ctx: Context, nside: int, pixels: np.ndarray, nest: bool,
              libm_exact: bool = True) -> torch.Tensor:
    """(n, 3) unit vectors.  ``libm_exact`` evaluates them with the platform
    libm on the host (bit-compatible with the reference); otherwise on the
    device (within 2 ulp)."""
    pixels = np.asarray(pixels, dtype=np.int64)
    if nside < 1 or (nest and nside & (nside - 1)):
        raise ValueError(f"invalid nside {nside}")
    if pixels.size and (pixels.min() < 0 or pixels.max() >= 12 * nside * nside):
        raise ValueError("pixel index out of range for nside")
    if libm_exact:
        host = np.empty((pixels.size, 3))
        pix_h = np.ascontiguousarray(pixels)
        ctx.check(ctx.lib.cfmi_pointings_host(nside, 1 if nest else 0, pix_h.ctypes.data, pixels.size,
                                              host.ctypes.data), "pointings_host")
        return ctx.upload(host)
    pix = ctx.upload(pixels, dtype=torch.int64)
    vec = torch.empty((pix.numel(), 3), dtype=torch.float64, device=ctx.device)
    ctx.check(ctx.lib.cfmi_pointings(ctx.h, nside, 1 if nest else 0, P(pix), pix.numel(), P(vec)),
              "pointings")
    return vec


def signal_matrix(ctx: Context, layout: Layout, vecs: list, cls: dict, lmax: int,
                  out: torch.Tensor | None = None) -> torch.Tensor:
    """S (n x n, full symmetric).  ``cls``: label -> device C_l (lmax + 1)."""
    n = layout.n
    S = out if out is not None else ctx.empty(n, n)
    if out is not None:
        ctx.check(ctx.lib.cfmi_fill(ctx.h, F64, n, n, P(S), n, 0.0), "fill")
    lib, h = ctx.lib, ctx.h
    fl = layout.fields
    for i, fi in enumerate(fl):
        for j in range(i, len(fl)):
            fj = fl[j]
            ri, rj = fi.offset, fj.offset
            lab = lambda m, a=i, b=j: layout.spectra_map[(a, b, m)]  # noqa: E731
            if fi.spin == 0 and fj.spin == 0:
                cl = cls[lab(0)]
                if i == j:
                    ctx.check(lib.cfmi_signal_00(h, P(cl), lmax, P(vecs[i]), fi.npix, P(vecs[i]), fi.npix, 1,
                                                 P(S, ri + ri * n), n, None, 0), "signal_00")
                else:
                    ctx.check(lib.cfmi_signal_00(h, P(cl), lmax, P(vecs[j]), fj.npix, P(vecs[i]), fi.npix, 0,
                                                 P(S, rj + ri * n), n, P(S, ri + rj * n), n), "signal_00")
            elif fi.spin == 2 and fj.spin == 2:
                if i != j:
                    raise NotImplementedError("Cross-correlation for spin-2 fields not implemented yet.")
                ctx.check(lib.cfmi_signal_22(h, P(cls[lab(0)]), P(cls[lab(1)]), P(cls[lab(2)]), lmax,
                                             P(vecs[i]), fi.npix, P(S, ri + ri * n), n), "signal_22")
            else:
                if fi.spin == 2:
                    raise NotImplementedError(
                        "a spin-2 field listed before a spin-0 field is not supported; list spin-0 first")
                ctx.check(lib.cfmi_signal_02(h, P(cls[lab(0)]), P(cls[lab(1)]), lmax, P(vecs[i]), fi.npix,
                                             P(vecs[j]), fj.npix, P(S, rj + ri * n), n, P(S, ri + rj * n), n),
                          "signal_02")
    return S


# ---------------------------------------------------------------------------
# A11: derivative stacks
# ---------------------------------------------------------------------------

@dataclass
class Block:
    """One spectrum's derivative support: rows [ra, ra+ma), cols [rb, rb+mb);
    dC = e_a D e_b^T (+ its transpose when a != b)."""
    ra: int
    ma: int
    rb: int
    mb: int
    stack: torch.Tensor  # (nbins, mb, ma) col-major blocks


def deriv_stacks(ctx: Context, layout: Layout, vecs: list, weights: dict, lmax: int,
                 bins: list) -> list:
    """For every spectrum of ``layout.spectra_list()``, every bin's D block.

    ``weights``: label -> device l-indexed w_l B1 B2 (lmax + 1).
    Returns a list of Block, one per spectrum, in spectra-list order."""
    lo = ctx.upload(np.array([b[0] for b in bins], dtype=np.int32), dtype=torch.int32)
    hi = ctx.upload(np.array([b[1] for b in bins], dtype=np.int32), dtype=torch.int32)
    nb = len(bins)
    lib, h = ctx.lib, ctx.h
    out = []
    cache22 = {}
    cache02 = {}
    for sp in layout.spectra_list():
        fi, fj = layout.fields[sp.fi], layout.fields[sp.fj]
        if fi.spin == 0 and fj.spin == 0:
            if sp.fi == sp.fj:
                st = torch.empty((nb, fi.rows, fi.rows), dtype=torch.float64, device=ctx.device)
                ctx.check(lib.cfmi_deriv_stack_00(h, P(weights[sp.label]), lmax, P(lo), P(hi), nb,
                                                  P(vecs[sp.fi]), fi.npix, P(vecs[sp.fi]), fi.npix, 1, P(st),
                                                  fi.rows, fi.rows * fi.rows, None, 0, 0), "deriv_stack_00")
                out.append(Block(fi.offset, fi.rows, fi.offset, fi.rows, st))
            else:
                st = torch.empty((nb, fi.rows, fj.rows), dtype=torch.float64, device=ctx.device)
                ctx.check(lib.cfmi_deriv_stack_00(h, P(weights[sp.label]), lmax, P(lo), P(hi), nb,
                                                  P(vecs[sp.fj]), fj.npix, P(vecs[sp.fi]), fi.npix, 0, P(st),
                                                  fj.rows, fi.rows * fj.rows, None, 0, 0), "deriv_stack_00")
                out.append(Block(fj.offset, fj.rows, fi.offset, fi.rows, st))
        elif fi.spin == 2 and fj.spin == 2:
            if sp.fi != sp.fj:
                raise NotImplementedError("spin-2 x spin-2 cross-field derivative")
            key = sp.fi
            if key not in cache22:
                labs = [layout.spectra_map[(sp.fi, sp.fi, m)] for m in range(3)]
                sts = [torch.empty((nb, fi.rows, fi.rows), dtype=torch.float64, device=ctx.device)
                       for _ in range(3)]
                ctx.check(lib.cfmi_deriv_stack_22(h, P(weights[labs[0]]), P(weights[labs[1]]),
                                                  P(weights[labs[2]]), lmax, P(lo), P(hi), nb, P(vecs[key]),
                                                  fi.npix, P(sts[0]), P(sts[1]), P(sts[2]), fi.rows,
                                                  fi.rows * fi.rows), "deriv_stack_22")
                cache22[key] = sts
            out.append(Block(fi.offset, fi.rows, fi.offset, fi.rows, cache22[key][sp.mode]))
        else:
            if fi.spin == 2:
                raise NotImplementedError("spin-2 field before spin-0 field")
            key = (sp.fi, sp.fj)
            if key not in cache02:
                labs = [layout.spectra_map[(sp.fi, sp.fj, m)] for m in range(2)]
                sts = [torch.empty((nb, fi.rows, fj.rows), dtype=torch.float64, device=ctx.device)
                       for _ in range(2)]
                ctx.check(lib.cfmi_deriv_stack_02(h, P(weights[labs[0]]), P(weights[labs[1]]), lmax, P(lo),
                                                  P(hi), nb, P(vecs[sp.fi]), fi.npix, P(vecs[sp.fj]), fj.npix,
                                                  P(sts[0]), P(sts[1]), fj.rows, fi.rows * fj.rows, None, None,
                                                  0, 0), "deriv_stack_02")
                cache02[key] = sts
            out.append(Block(fj.offset, fj.rows, fi.offset, fi.rows, cache02[key][sp.mode]))
    return out


# ---------------------------------------------------------------------------
# A12 - A13: factorisation
# ---------------------------------------------------------------------------

@dataclass
class Factor:
    L: torch.Tensor      # Cholesky factor of D C D (lower, col-major)
    W: torch.Tensor      # whitener U D with C^-1 = W^T W (lower, col-major)
    d: torch.Tensor      # Jacobi scaling 1/sqrt|C_ii|
    logdet: float        # ln|C|


def factor(ctx: Context, C: torch.Tensor, n: int, need_whitener: bool = True) -> Factor:
    """Jacobi-scaled Cholesky of C (overwritten by L) and the whitener."""
    d = torch.empty(n, dtype=torch.float64, device=ctx.device)
    ctx.check(ctx.lib.cfmi_jacobi_scale(ctx.h, n, P(C), n, P(d)), "jacobi_scale")
    U = ctx.empty(n, n) if need_whitener else None
    info = ctypes.c_int(0)
    rc = ctx.lib.cfmi_potrf(ctx.h, F64, n, P(C), n, P(U), n, ctypes.byref(info))
    ctx.check(rc, "potrf")
    if info.value:
        raise LinAlgError(info.value)
    ld = ctypes.c_double(0.0)
    ctx.check(ctx.lib.cfmi_logdet_chol(ctx.h, n, P(C), n, P(d), ctypes.byref(ld)), "logdet")
    if need_whitener:
        ctx.check(ctx.lib.cfmi_trtri(ctx.h, F64, n, P(C), n, P(U), n), "trtri")
        ctx.check(ctx.lib.cfmi_scale(ctx.h, 0, n, n, P(U), n, P(d)), "scale")
    return Factor(C, U, d, float(ld.value))


def gemm(ctx: Context, ta: bool, tb: bool, m: int, n: int, k: int, alpha: float, A, offa: int, lda: int,
         B, offb: int, ldb: int, beta: float, C, offc: int, ldc: int, flags: int = 0):
    ctx.check(ctx.lib.cfmi_gemm(ctx.h, dtype_code(C), 1 if ta else 0, 1 if tb else 0, m, n, k, alpha,
                                P(A, offa), lda, P(B, offb), ldb, beta, P(C, offc), ldc, flags), "gemm")


def inverse_from_whitener(ctx: Context, W: torch.Tensor, n: int) -> torch.Tensor:
    """C^-1 = W^T W (full symmetric)."""
    Ci = ctx.empty(n, n)
    # op(A) = W^T is upper (A[m,k] = 0 for k < m): flag bit 1.
    gemm(ctx, True, False, n, n, n, 1.0, W, 0, n, W, 0, n, 0.0, Ci, 0, n, 2)
    return Ci


# ---------------------------------------------------------------------------
# A14 - A17: Fisher and QML through the whitened derivatives
# ---------------------------------------------------------------------------

SQRT_HALF = 0.7071067811865476


def packed_size(n: int) -> int:
    return n * (n + 1) // 2


def gemm_packed(ctx: Context, ta: bool, tb: bool, m: int, n: int, k: int, alpha: float, A, offa: int, lda: int,
                B, offb: int, ldb: int, beta: float, Pk: torch.Tensor, pn: int, row0: int, col0: int,
                pdiag: float = SQRT_HALF, poff: float = 1.0, flags: int = 0):
    ctx.check(ctx.lib.cfmi_gemm_packed(ctx.h, dtype_code(Pk), 1 if ta else 0, 1 if tb else 0, m, n, k, alpha,
                                       P(A, offa), lda, P(B, offb), ldb, beta, P(Pk), pn, row0, col0, pdiag,
                                       poff, flags), "gemm_packed")


def gemm_batched(ctx: Context, ta: bool, tb: bool, m: int, n: int, k: int, alpha: float, A, offa: int, lda: int,
                 sA: int, B, offb: int, ldb: int, sB: int, beta: float, C, offc: int, ldc: int, sC: int, batch: int,
                 flags: int = 0):
    ctx.check(ctx.lib.cfmi_gemm_batched(ctx.h, dtype_code(C), 1 if ta else 0, 1 if tb else 0, m, n, k, alpha,
                                        P(A, offa), lda, sA, P(B, offb), ldb, sB, beta, P(C, offc), ldc, sC, batch,
                                        flags), "gemm_batched")


def gemm_packed_batched(ctx: Context, ta: bool, tb: bool, m: int, n: int, k: int, alpha: float, A, offa: int,
                        lda: int, sA: int, B, offb: int, ldb: int, sB: int, beta: float, Pk: torch.Tensor, sP: int,
                        pn: int, row0: int, col0: int, batch: int, pdiag: float = SQRT_HALF, poff: float = 1.0,
                        flags: int = 0):
    ctx.check(ctx.lib.cfmi_gemm_packed_batched(ctx.h, dtype_code(Pk), 1 if ta else 0, 1 if tb else 0, m, n, k,
                                               alpha, P(A, offa), lda, sA, P(B, offb), ldb, sB, beta, P(Pk), sP, pn,
                                               row0, col0, pdiag, poff, batch, flags), "gemm_packed_batched")


def whitened_derivatives(ctx: Context, W: torch.Tensor, n: int, blocks: list, nbins: int,
                         out: torch.Tensor, zbuf: torch.Tensor | None = None):
    """out[p] = packed lower triangle of W dC_p W^T, diagonal scaled by
    1/sqrt(2), for every parameter p = spectrum * nbins + bin.

    With that scaling <out[i], out[j]> = 1/2 Tr[W dC_i W^T W dC_j W^T] =
    1/2 Tr[C^-1 dC_i C^-1 dC_j] = F_ij.  ``out`` is (n_params, n(n+1)/2),
    zero-filled.  All bins of a spectrum go through one batched launch per
    product."""
    K = packed_size(n)
    for s, blk in enumerate(blocks):
        ra, ma, rb, mb = blk.ra, blk.ma, blk.rb, blk.mb
        zs = (n - ra) * mb
        need = zs * nbins
        if zbuf is None or zbuf.numel() < need:
            zbuf = None
            zbuf = torch.empty(need, dtype=torch.float64, device=ctx.device)
        Vs = out[s * nbins:(s + 1) * nbins]
        # Z_b = W[ra:, ra:ra+ma] D_b      ((n-ra) x mb each; W block lower)
        gemm_batched(ctx, False, False, n - ra, mb, ma, 1.0, W, ra + ra * n, n, 0, blk.stack, 0, ma, ma * mb, 0.0,
                     zbuf, 0, n - ra, zs, nbins, 1)
        # X_b = Z_b W[rb:, rb:rb+mb]^T at rows ra.., cols rb.. (lower part)
        gemm_packed_batched(ctx, False, True, n - ra, n - rb, mb, 1.0, zbuf, 0, n - ra, zs, W, rb + rb * n, n, 0,
                            0.0, Vs, K, n, ra, rb, nbins, flags=4)
        if ra != rb or ma != mb:
            # + X_b^T = W[rb:, rb:rb+mb] Z_b^T at rows rb.., cols ra..
            gemm_packed_batched(ctx, False, True, n - rb, n - ra, mb, 1.0, W, rb + rb * n, n, 0, zbuf, 0, n - ra,
                                zs, 1.0, Vs, K, n, rb, ra, nbins, flags=1)
    return zbuf


def gram(ctx: Context, m: int, nn: int, K: int, X: torch.Tensor, Y: torch.Tensor) -> np.ndarray:
    G = torch.empty((nn, m), dtype=torch.float64, device=ctx.device)
    ctx.check(ctx.lib.cfmi_gram(ctx.h, dtype_code(X), m, nn, K, P(X), K, P(Y), K, P(G), m), "gram")
    return to_numpy_cm(G)
