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
BF16X3 = "bf16x3"   # stack dtype marker: three BF16 planes x = x0 + x1 + x2 (int16 storage)


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
        """Host (m, n) array -> column-major device matrix (the transpose is
        made on the device, not by a host copy)."""
        a = np.asarray(a)
        if a.ndim == 2 and a.flags.c_contiguous:
            return torch.as_tensor(a).to(self.device, dtype=dtype).T.contiguous()
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
            elif fi.spin == 2:
                # A spin-2 field listed before a spin-0 field: the reference
                # writes this block above the diagonal and its final
                # lower-to-upper copy then overwrites it with the (unwritten)
                # block below, so its S carries no T x P correlation
                # (signal_kernels.py:489-499 with linalg.py:101-120).  Kept
                # for parity: both blocks are zero.
                ctx.check(lib.cfmi_fill(h, F64, fj.rows, fi.rows, P(S, rj + ri * n), n, 0.0), "fill")
                ctx.check(lib.cfmi_fill(h, F64, fi.rows, fj.rows, P(S, ri + rj * n), n, 0.0), "fill")
            else:
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
    stack: torch.Tensor | None = None   # (nbins, mb, ma) col-major blocks; BF16X3: (3, nbins, mb, ma)
    mirror: torch.Tensor | None = None  # BF16X3 only: (3, nbins, ma, mb), the transposed blocks

    @property
    def sym(self) -> bool:
        return self.ra == self.rb and self.ma == self.mb

    @property
    def nbins(self) -> int:
        return self.stack.shape[1] if self.stack.dim() == 4 else self.stack.shape[0]


def block_geometry(layout: Layout) -> list:
    """Block (without stack) of every spectrum of ``layout.spectra_list()``."""
    out = []
    for sp in layout.spectra_list():
        fi, fj = layout.fields[sp.fi], layout.fields[sp.fj]
        if fi.spin == 2 and fj.spin == 2 and sp.fi != sp.fj:
            raise NotImplementedError("spin-2 x spin-2 cross-field derivative")
        if sp.fi == sp.fj:
            out.append(Block(fi.offset, fi.rows, fi.offset, fi.rows))
        elif fi.spin == 2 and fj.spin == 0:
            # P rows x T columns, as for a spin-0 field listed first
            out.append(Block(fi.offset, fi.rows, fj.offset, fj.rows))
        else:
            out.append(Block(fj.offset, fj.rows, fi.offset, fi.rows))
    return out


def deriv_state(ctx: Context, layout: Layout, group_spectrum: int) -> torch.Tensor:
    """Recurrence-state buffer for the recurrence pass of one spectrum group."""
    sp = layout.spectra_list()[group_spectrum]
    fi, fj = layout.fields[sp.fi], layout.fields[sp.fj]
    if sp.fi == sp.fj:
        k = ctx.lib.cfmi_deriv_state_doubles(fi.npix, fi.npix, 1)
    elif fi.spin == 0:
        k = ctx.lib.cfmi_deriv_state_doubles(fj.npix, fi.npix, 0)
    else:   # a spin-2 field listed before the spin-0 one: T first, as the 0 x 2 pass runs
        k = ctx.lib.cfmi_deriv_state_doubles(fi.npix, fj.npix, 0)
    return torch.empty(int(k), dtype=torch.float64, device=ctx.device)


def deriv_stacks(ctx: Context, layout: Layout, vecs: list, weights: dict, lmax: int, bins: list,
                 dtype=torch.float64, only=None, bin_range=None, state: torch.Tensor | None = None,
                 lresume: int = 0, pool: torch.Tensor | None = None, mirrors: bool = False) -> list:
    """For every spectrum of ``layout.spectra_list()``, every bin's D block.

    ``weights``: label -> device l-indexed w_l B1 B2 (lmax + 1).  ``only``:
    indices of the spectra to produce (the others get ``stack=None``);
    ``bin_range``: (b0, b1), produce bins b0 .. b1-1 only.  Stacks are FP64
    sums of the FP64 recurrences stored as ``dtype``.  The three spin-2
    spectra of a field share one recurrence pass, and TE/TB another.
    ``state`` (``deriv_state``, one recurrence group per call): the pass
    saves its final recurrence state there and, with ``lresume`` > 0,
    resumes at l = lresume from it.  ``pool``: a flat ``dtype`` buffer the
    stacks are carved from (instead of fresh allocations).

    ``dtype=BF16X3``: every block as three BF16 planes (``split_planes``),
    shape (3, nbins, mb, ma) on an int16 pool; with ``mirrors`` the blocks
    of a != b spectra also get their transposes (``Block.mirror``), which
    the banded three-piece whitening multiplies by."""
    b0, b1 = bin_range if bin_range is not None else (0, len(bins))
    sel = bins[b0:b1]
    lo = ctx.upload(np.array([b[0] for b in sel], dtype=np.int32), dtype=torch.int32)
    hi = ctx.upload(np.array([b[1] for b in sel], dtype=np.int32), dtype=torch.int32)
    nb = len(sel)
    b3 = isinstance(dtype, str)
    if b3 and dtype != BF16X3:
        raise ValueError(f"unknown stack dtype {dtype!r}")
    dcode = _lib.CFMI_BF16X3 if b3 else (F64 if dtype == torch.float64 else F32)
    tdt = torch.int16 if b3 else dtype
    lib, h = ctx.lib, ctx.h
    specs = layout.spectra_list()
    geo = block_geometry(layout)
    want = set(range(len(specs)) if only is None else only)
    out = [Block(g.ra, g.ma, g.rb, g.mb) for g in geo]
    done = set()
    used = [0]

    def alloc(shape):
        if b3:
            shape = (3,) + tuple(shape)
        if pool is None:
            return torch.empty(shape, dtype=tdt, device=ctx.device)
        k = int(np.prod(shape))
        if used[0] + k > pool.numel():
            raise MemoryError("derivative stack chunk exceeds its pool")
        t = pool[used[0]:used[0] + k].view(shape)
        used[0] += k
        return t
    for k, sp in enumerate(specs):
        if k not in want or k in done:
            continue
        fi, fj = layout.fields[sp.fi], layout.fields[sp.fj]
        blk = out[k]
        if fi.spin == 0 and fj.spin == 0:
            st = alloc((nb, blk.mb, blk.ma))
            if sp.fi == sp.fj:
                ctx.check(lib.cfmi_deriv_stack_00(h, dcode, P(weights[sp.label]), lmax, P(lo), P(hi), nb,
                                                  P(vecs[sp.fi]), fi.npix, P(vecs[sp.fi]), fi.npix, 1, P(st),
                                                  fi.rows, fi.rows * fi.rows, None, 0, 0, P(state), lresume),
                          "deriv_stack_00")
            else:
                mr = alloc((nb, blk.ma, blk.mb)) if (b3 and mirrors) else None
                ctx.check(lib.cfmi_deriv_stack_00(h, dcode, P(weights[sp.label]), lmax, P(lo), P(hi), nb,
                                                  P(vecs[sp.fj]), fj.npix, P(vecs[sp.fi]), fi.npix, 0, P(st),
                                                  fj.rows, fi.rows * fj.rows, P(mr), blk.mb if mr is not None else 0,
                                                  blk.ma * blk.mb if mr is not None else 0, P(state), lresume),
                          "deriv_stack_00")
                blk.mirror = mr
            blk.stack = st
            done.add(k)
        elif fi.spin == 2 and fj.spin == 2:
            # EE, BB, EB of this field: one pass, only the requested planes.
            members = [m for m, s2 in enumerate(specs) if s2.fi == sp.fi and s2.fj == sp.fj]
            sts = [None, None, None]
            for m in members:
                if m in want:
                    sts[specs[m].mode] = alloc((nb, fi.rows, fi.rows))
            labs = [layout.spectra_map[(sp.fi, sp.fi, md)] for md in range(3)]
            ctx.check(lib.cfmi_deriv_stack_22(h, dcode, P(weights[labs[0]]), P(weights[labs[1]]),
                                              P(weights[labs[2]]), lmax, P(lo), P(hi), nb, P(vecs[sp.fi]),
                                              fi.npix, P(sts[0]), P(sts[1]), P(sts[2]), fi.rows,
                                              fi.rows * fi.rows, P(state), lresume), "deriv_stack_22")
            for m in members:
                if m in want:
                    out[m].stack = sts[specs[m].mode]
                    done.add(m)
        else:
            # TE, TB: the spin-0 field (listed first or second) and the spin-2 one
            it, ip = (sp.fi, sp.fj) if fi.spin == 0 else (sp.fj, sp.fi)
            ft, fp = layout.fields[it], layout.fields[ip]
            members = [m for m, s2 in enumerate(specs) if s2.fi == sp.fi and s2.fj == sp.fj]
            sts = [None, None]
            mrs = [None, None]
            for m in members:
                if m in want:
                    sts[specs[m].mode] = alloc((nb, ft.rows, fp.rows))
                    if b3 and mirrors:
                        mrs[specs[m].mode] = alloc((nb, fp.rows, ft.rows))
            labs = [layout.spectra_map[(sp.fi, sp.fj, md)] for md in range(2)]
            has_m = mrs[0] is not None or mrs[1] is not None
            ctx.check(lib.cfmi_deriv_stack_02(h, dcode, P(weights[labs[0]]), P(weights[labs[1]]), lmax, P(lo),
                                              P(hi), nb, P(vecs[it]), ft.npix, P(vecs[ip]), fp.npix,
                                              P(sts[0]), P(sts[1]), fp.rows, ft.rows * fp.rows, P(mrs[0]),
                                              P(mrs[1]), ft.rows if has_m else 0,
                                              ft.rows * fp.rows if has_m else 0, P(state), lresume),
                      "deriv_stack_02")
            for m in members:
                if m in want:
                    out[m].stack = sts[specs[m].mode]
                    out[m].mirror = mrs[specs[m].mode]
                    done.add(m)
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
    U: torch.Tensor | None = None   # inverses of L's diagonal nb-blocks (when W is not formed)
    nb: int = 0
    slices: int = 0                 # > 0: trailing updates on int8 slices (L L^T = D C D to ~2^-(8 s))


def factor(ctx: Context, C: torch.Tensor, n: int, need_whitener: bool = True, diag_blocks: int = 0,
           slices: int = 0) -> Factor:
    """Jacobi-scaled Cholesky of C (overwritten by L) and the whitener.

    ``diag_blocks`` = nb > 0 (instead of the whitener): ``Factor.U`` holds
    the inverses of L's nb x nb diagonal blocks at their places (the rest of
    U is unused), the diagonal solves of the blocked ``trsm``.  ``slices``
    = s > 0: the trailing updates of the factorisation on the int8 matrix
    cores, operands as s int8 slices (``cfmi_potrf_s8``)."""
    d = torch.empty(n, dtype=torch.float64, device=ctx.device)
    ctx.check(ctx.lib.cfmi_jacobi_scale(ctx.h, n, P(C), n, P(d)), "jacobi_scale")
    U = ctx.empty(n, n) if (need_whitener or diag_blocks) else None
    info = ctypes.c_int(0)
    if slices:
        rc = ctx.lib.cfmi_potrf_s8(ctx.h, slices, n, P(C), n, P(U), n, ctypes.byref(info))
    else:
        rc = ctx.lib.cfmi_potrf(ctx.h, F64, n, P(C), n, P(U), n, ctypes.byref(info))
    ctx.check(rc, "potrf")
    if info.value:
        raise LinAlgError(info.value)
    ld = ctypes.c_double(0.0)
    ctx.check(ctx.lib.cfmi_logdet_chol(ctx.h, n, P(C), n, P(d), ctypes.byref(ld)), "logdet")
    if need_whitener:
        ctx.check(ctx.lib.cfmi_trtri(ctx.h, F64, n, P(C), n, P(U), n), "trtri")
        ctx.check(ctx.lib.cfmi_scale(ctx.h, 0, n, n, P(U), n, P(d)), "scale")
        return Factor(C, U, d, float(ld.value), slices=slices)
    if diag_blocks:
        if diag_blocks % 64 or (diag_blocks // 64) & (diag_blocks // 64 - 1):
            raise ValueError("diagonal blocks must be 64 times a power of two rows")
        ctx.check(ctx.lib.cfmi_trtri_blocks(ctx.h, F64, n, diag_blocks, P(C), n, P(U), n), "trtri_blocks")
        f = Factor(C, None, d, float(ld.value))
        f.U, f.nb, f.slices = U, diag_blocks, slices
        return f
    return Factor(C, None, d, float(ld.value), slices=slices)


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


def gemm_ex(ctx: Context, ta: bool, tb: bool, m: int, n: int, k: int, alpha: float, A, offa: int, lda: int,
            B, offb: int, ldb: int, beta: float, C, offc: int = 0, ldc: int = 1, *, sA: int = 0, sB: int = 0,
            sC: int = 0, batch: int = 1, flags: int = 0, tri_a: int = 0, tri_b: int = 0, out_mode: int = 0,
            row0: int = 0, col0: int = 0, pbase: int = 0, pdiag: float = 1.0, poff: float = 1.0):
    """``cfmi_gemm_ex``: input dtype from A, output dtype from C."""
    if m <= 0 or n <= 0:
        return
    ctx.check(ctx.lib.cfmi_gemm_ex(ctx.h, dtype_code(A), dtype_code(C), 1 if ta else 0, 1 if tb else 0, m, n, k,
                                   alpha, P(A, offa), lda, sA, tri_a, P(B, offb), ldb, sB, tri_b, beta,
                                   P(C, offc), ldc, sC, batch, flags, out_mode, row0, col0, pbase, pdiag, poff),
              "gemm_ex")


def _gemm_ex(*a, **k):
    return gemm_ex(*a, **k)


def band_size(r0: int, r1: int) -> int:
    """Entries of lower-triangle rows r0 .. r1-1."""
    return (r1 * (r1 + 1) - r0 * (r0 + 1)) // 2


def whiten_band(ctx: Context, W: torch.Tensor, n: int, blk: Block, X: torch.Tensor, r0: int, r1: int,
                zbuf: torch.Tensor | None = None, flags: int = 0, zcap: int | None = None,
                counter: list | None = None) -> torch.Tensor | None:
    """X[b] += rows r0..r1-1 of the lower triangle of W dC_b W^T, row-major
    packed (element (i, c <= i) at i(i+1)/2 + c - r0(r0+1)/2), diagonal
    scaled by 1/sqrt(2), for every bin b of ``blk.stack``.

    With that scaling the sum over bands of <X_i, X_j> is
    1/2 Tr[W dC_i W^T W dC_j W^T] = 1/2 Tr[C^-1 dC_i C^-1 dC_j].
    W is lower triangular (its zero parts are skipped); ``X`` is
    (nbins, band_size(r0, r1)) and must hold zeros (or earlier terms).
    ``zcap``: at most this many elements of intermediate products at once
    (bins are then taken in sub-batches)."""
    D = blk.stack
    if zcap is not None:
        rows = r1 - min(max(r0, blk.ra), max(r0, blk.rb))
        per = max(rows, 1) * max(blk.ma, blk.mb)
        step = max(1, zcap // per)
        if step < D.shape[0]:
            for b0 in range(0, D.shape[0], step):
                sub = Block(blk.ra, blk.ma, blk.rb, blk.mb, D[b0:b0 + step])
                zbuf = whiten_band(ctx, W, n, sub, X[b0:b0 + step], r0, r1, zbuf, flags, counter=counter)
            return zbuf
    nbb = D.shape[0]
    ra, ma, rb, mb = blk.ra, blk.ma, blk.rb, blk.mb
    sX = X.shape[1]
    pbase = r0 * (r0 + 1) // 2
    full = r0 == 0 and r1 == n
    fl = flags

    def gemm_ex(*a, **k):  # noqa: F811  (counts the launches of this call)
        if counter is not None:
            counter.append(1)
        _gemm_ex(*a, **k)

    def need(k):
        nonlocal zbuf
        if zbuf is None or zbuf.numel() < k or zbuf.dtype != W.dtype:
            zbuf = None
            zbuf = torch.empty(k, dtype=W.dtype, device=ctx.device)
        return zbuf

    # term 1: X[i, c] += sum_k (W[i, a] D)[k] W[c, rb + k],   i >= ra, c >= rb
    i0 = max(r0, ra)
    rr = r1 - i0
    if rr > 0:
        Z = need(rr * mb * nbb)
        gemm_ex(ctx, False, False, rr, mb, ma, 1.0, W, i0 + ra * n, n, D, 0, ma, 0.0, Z, 0, rr, sB=ma * mb,
                sC=rr * mb, batch=nbb, flags=1 | fl, tri_a=i0 - ra)
        gemm_ex(ctx, False, True, rr, r1 - rb, mb, 1.0, Z, 0, rr, W, rb + rb * n, n, 1.0, X, 0, sA=rr * mb,
                sC=sX, batch=nbb, flags=4 | fl, out_mode=2, row0=i0, col0=rb, pbase=pbase, pdiag=SQRT_HALF)
    if blk.sym:
        return zbuf
    # term 2: the transpose, X[i, c] += (W dC^T W^T)[i, c] = sum_k' (W[i, b] D^T)[k'] W[c, ra + k']
    if full and rr > 0:
        # Z above holds all rows of W[:, a] D: reuse it as op(B) = Z^T.
        gemm_ex(ctx, False, True, n - rb, n - ra, mb, 1.0, W, rb + rb * n, n, Z, 0, rr, 1.0, X, 0, sB=rr * mb,
                sC=sX, batch=nbb, flags=1 | fl, out_mode=2, row0=rb, col0=ra, pbase=pbase, pdiag=SQRT_HALF)
        return zbuf
    j0 = max(r0, rb)
    rj = r1 - j0
    if rj <= 0 or r1 <= ra:
        return zbuf
    Z2 = need(rj * ma * nbb)
    gemm_ex(ctx, False, True, rj, ma, mb, 1.0, W, j0 + rb * n, n, D, 0, ma, 0.0, Z2, 0, rj, sB=ma * mb,
            sC=rj * ma, batch=nbb, flags=1 | fl, tri_a=j0 - rb)
    gemm_ex(ctx, False, True, rj, r1 - ra, ma, 1.0, Z2, 0, rj, W, ra + ra * n, n, 1.0, X, 0, sA=rj * ma, sC=sX,
            batch=nbb, flags=4 | fl, out_mode=2, row0=j0, col0=ra, pbase=pbase, pdiag=SQRT_HALF)
    return zbuf


# ---------------------------------------------------------------------------
# Three-piece BF16 operands: FP32-accurate products on the BF16 MFMA
# ---------------------------------------------------------------------------

def planes_ld(cols: int) -> int:
    """Row pitch of three-piece planes (a multiple of 8 elements, so rows
    start on 16-byte boundaries)."""
    return (cols + 7) // 8 * 8


def split_planes(ctx: Context, A: torch.Tensor, rows: int, cols: int, lda: int) -> torch.Tensor:
    """Column-major FP64/FP32 (rows x cols, ld lda) -> (3, rows, planes_ld(cols))
    int16 tensor of BF16 pieces, row-major: [p, i, j] is piece p of A[i, j]."""
    ld = planes_ld(cols)
    out = torch.empty((3, rows, ld), dtype=torch.int16, device=ctx.device)
    ctx.check(ctx.lib.cfmi_split3(ctx.h, dtype_code(A), rows, cols, P(A), lda, P(out), ld, rows * ld), "split3")
    return out


def gemm_b3(ctx: Context, mode: int, m: int, n: int, k: int, alpha: float, A, offa: int, lda: int, sA: int,
            pA: int, B, offb: int, ldb: int, sB: int, pB: int, beta: float, C, offc: int = 0, ldc: int = 0,
            sC: int = 0, pC: int = 0, *, batch: int = 1, flags: int = 0, tri_a: int = 0, tri_b: int = 0,
            row0: int = 0, col0: int = 0, pbase: int = 0, pdiag: float = 1.0):
    """``cfmi_gemm_b3``: C[m, n] = alpha sum_k A[m][k] B[n][k] (+ beta C) on
    k-contiguous three-piece operands (offsets and strides in elements,
    planes pA / pB apart).  mode 0: FP32 column-major C; 1: three-piece
    row-major C (planes pC apart); 2: FP32 packed lower row band."""
    if m <= 0 or n <= 0:
        return
    ctx.check(ctx.lib.cfmi_gemm_b3(ctx.h, mode, m, n, k, alpha, P(A, offa), lda, sA, pA, tri_a, P(B, offb), ldb,
                                   sB, pB, tri_b, beta, P(C, offc), ldc, sC, pC, batch, flags, row0, col0, pbase,
                                   pdiag), "gemm_b3")


def _gemm_b3(*a, **k):
    return gemm_b3(*a, **k)


# Extra cfmi_gemm_b3 flags for the whitening products (256 / 512 force the
# 128- / 256-tile configuration; 0 keeps the library's choice).
B3_WHITEN_FLAGS = 0


def _plane_buffer(ctx: Context, k: int) -> torch.Tensor:
    """Flat storage for k three-piece plane entries (int16 BF16 bits)."""
    return torch.empty(k, dtype=torch.int16, device=ctx.device)


def whiten_band_b3(ctx: Context, Wp: torch.Tensor, n: int, blk: Block, X: torch.Tensor, r0: int, r1: int,
                   zbuf: torch.Tensor | None = None, zcap: int | None = None,
                   counter: list | None = None) -> torch.Tensor | None:
    """``whiten_band`` on three-piece BF16 operands (FP32 accuracy at the
    BF16 MFMA rate): ``Wp`` = ``split_planes`` of W (3, n, ldw);
    ``blk.stack`` (3, nbins, mb, ma) planes of the column-major D blocks,
    which are k-contiguous for W D as they stand; ``blk.mirror`` holds
    D^T for a != b when the band is not all of [0, n) (W D^T has D^T as
    its k-contiguous operand).  The products W D (W D^T) are stored as
    three-piece row-major Z, the operand of the band product, and ``X``
    (FP32) receives the packed band as in ``whiten_band``."""
    D = blk.stack
    nbb = D.shape[1]
    ra, ma, rb, mb = blk.ra, blk.ma, blk.rb, blk.mb
    if zcap is not None:
        rows = r1 - min(max(r0, ra), max(r0, rb))
        step = max(1, zcap // (max(rows, 1) * max(ma, mb)))
        if step < nbb:
            for b0 in range(0, nbb, step):
                Mi = blk.mirror[:, b0:b0 + step] if blk.mirror is not None else None
                sub = Block(ra, ma, rb, mb, D[:, b0:b0 + step], Mi)
                zbuf = whiten_band_b3(ctx, Wp, n, sub, X[b0:b0 + step], r0, r1, zbuf, counter=counter)
            return zbuf
    ldw, pW = Wp.shape[2], Wp.stride(0)
    sX = X.stride(0)
    pbase = r0 * (r0 + 1) // 2
    full = r0 == 0 and r1 == n

    def g(*a, **k):
        if counter is not None:
            counter.append(1)
        k["flags"] = k.get("flags", 0) | B3_WHITEN_FLAGS
        _gemm_b3(ctx, *a, **k)

    def need(k):
        nonlocal zbuf
        if zbuf is None or zbuf.numel() < 3 * k:
            zbuf = None
            zbuf = _plane_buffer(ctx, 3 * k)
        return zbuf

    # term 1: Z = W[i0:r1, ra:ra+ma] D (rr x mb), X[i, c] += sum_k Z[i, k] W[c, rb + k]
    i0 = max(r0, ra)
    rr = r1 - i0
    if rr > 0:
        Z = need(rr * mb * nbb)
        pZ = rr * mb * nbb
        g(1, rr, mb, ma, 1.0, Wp, i0 * ldw + ra, ldw, 0, pW, D, 0, ma, D.stride(1), D.stride(0), 0.0, Z, 0, mb,
          rr * mb, pZ, batch=nbb, flags=1, tri_a=i0 - ra)
        g(2, rr, r1 - rb, mb, 1.0, Z, 0, mb, rr * mb, pZ, Wp, rb * ldw + rb, ldw, 0, pW, 1.0, X, 0, 0, sX, 0,
          batch=nbb, flags=4, row0=i0, col0=rb, pbase=pbase, pdiag=SQRT_HALF)
    if blk.sym:
        return zbuf
    # term 2: X[i, c] += sum_k' (W[i, rb:] D^T)[k'] W[c, ra + k']
    if full and rr > 0:
        # Z holds all rows of W[:, a] D: X[i, c] += sum_k W[i, rb + k] Z[c - ra, k]
        g(2, n - rb, n - ra, mb, 1.0, Wp, rb * ldw + rb, ldw, 0, pW, Z, 0, mb, rr * mb, pZ, 1.0, X, 0, 0, sX, 0,
          batch=nbb, flags=1, row0=rb, col0=ra, pbase=pbase, pdiag=SQRT_HALF)
        return zbuf
    j0 = max(r0, rb)
    rj = r1 - j0
    if rj <= 0 or r1 <= ra:
        return zbuf
    Dt = blk.mirror
    if Dt is None:
        raise ValueError("whiten_band_b3: a != b block over a partial band needs the transposed stack (mirrors)")
    Z2 = need(rj * ma * nbb)
    pZ2 = rj * ma * nbb
    g(1, rj, ma, mb, 1.0, Wp, j0 * ldw + rb, ldw, 0, pW, Dt, 0, mb, Dt.stride(1), Dt.stride(0), 0.0, Z2, 0, ma,
      rj * ma, pZ2, batch=nbb, flags=1, tri_a=j0 - rb)
    g(2, rj, r1 - ra, ma, 1.0, Z2, 0, ma, rj * ma, pZ2, Wp, ra * ldw + ra, ldw, 0, pW, 1.0, X, 0, 0, sX, 0,
      batch=nbb, flags=4, row0=j0, col0=ra, pbase=pbase, pdiag=SQRT_HALF)
    return zbuf


def gram(ctx: Context, m: int, nn: int, K: int, X: torch.Tensor, Y: torch.Tensor,
         out: torch.Tensor | None = None, beta: float = 0.0):
    """G = X^T Y (m x nn, FP64) over K-long columns; with ``out`` the device
    G is accumulated (beta) and returned, otherwise a host copy."""
    G = out if out is not None else torch.empty((nn, m), dtype=torch.float64, device=ctx.device)
    sym = 64 if (X.data_ptr() == Y.data_ptr() and m == nn) else 0
    gemm_ex(ctx, True, False, m, nn, K, 1.0, X, 0, K, Y, 0, K, beta, G, 0, m, flags=sym | _lib.CFMI_GEMM_ACC64)
    return G if out is not None else to_numpy_cm(G)


# ---------------------------------------------------------------------------
# FP64 products on the int8 matrix cores: sliced operands (cfmi_gemm_s8)
# ---------------------------------------------------------------------------

class Slices:
    """int8 slices of up to ``rows`` operand rows over ``k`` values:
    ``data`` (s, rows, ldk) int8 (k contiguous, ldk a multiple of 64),
    ``mu`` (rows,) FP64 row scales.  A buffer can be refilled with fewer
    rows; the plane stride stays rows * ldk."""

    def __init__(self, ctx: Context, s: int, rows: int, k: int):
        self.s, self.rows, self.k = int(s), int(rows), int(k)
        self.ldk = (self.k + 63) // 64 * 64
        self.data = torch.empty((self.s, max(self.rows, 1), self.ldk), dtype=torch.int8, device=ctx.device)
        self.mu = torch.empty(max(self.rows, 1), dtype=torch.float64, device=ctx.device)

    @property
    def ps(self) -> int:
        return self.data.shape[1] * self.ldk


def slice8(ctx: Context, s: int, layout: int, rows: int, k: int, src: torch.Tensor, off: int, lds: int,
           out: Slices | None = None) -> Slices:
    """Slice ``rows`` operand rows of ``k`` FP64 values: layout 0, row r at
    src[off + r lds] (k contiguous); layout 1, element (r, k) at
    src[off + r + k lds]."""
    if out is None:
        out = Slices(ctx, s, rows, k)
    if out.s != s or rows > out.data.shape[1] or k > out.ldk:
        raise ValueError("slice buffer too small")
    ctx.check(ctx.lib.cfmi_slice8(ctx.h, s, layout, rows, k, P(src, off), lds, P(out.data), out.ldk, out.ps,
                                  P(out.mu)), "slice8")
    out.k = k
    return out


class S8Recorder:
    """The context's sliced-GEMM launch record while open
    (``cfmi_s8_stats_*``): HIP events on the context's stream around every
    gemm_s8 kernel launch, those issued from here and those the library
    issues itself (the sliced Cholesky updates), with each launch's
    algorithmic flops (2 k per entry computed: 2 m n k, about m n k for a
    lower or symmetric result)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        ctx.check(ctx.lib.cfmi_s8_stats_enable(ctx.h, 1), "s8_stats_enable")

    def total(self):
        """(sum of algorithmic flops, sum of launch times in ms, launches,
        int8 products issued per algorithmic multiply-add, averaged)."""
        n, ms, fl, prod = ctypes.c_int64(0), ctypes.c_double(0), ctypes.c_double(0), ctypes.c_double(0)
        self.ctx.check(self.ctx.lib.cfmi_s8_stats_read(self.ctx.h, ctypes.byref(n), ctypes.byref(ms),
                                                       ctypes.byref(fl), ctypes.byref(prod)), "s8_stats_read")
        return fl.value, ms.value, int(n.value), prod.value / max(fl.value, 1.0)

    def close(self):
        self.ctx.check(self.ctx.lib.cfmi_s8_stats_enable(self.ctx.h, 0), "s8_stats_enable")


def gemm_s8(ctx: Context, m: int, n: int, k: int, alpha: float, A: Slices, a_row0: int, B: Slices, b_row0: int,
            beta: float, C: torch.Tensor, offc: int, ldc: int, flags: int = 0, *, tri_a: int | None = None,
            tri_b: int | None = None, packed=None):
    """C[m x n] (column-major FP64, at offc) = alpha A B^T + beta C over rows
    a_row0.. of A and b_row0.. of B (``cfmi_gemm_s8``).  ``tri_a`` / ``tri_b``:
    A[m][k] = 0 for k > m + tri_a (B likewise), which shortens K per tile;
    ``packed`` = (row0, col0, pbase, pdiag): C (at offc) is the packed lower
    row band (ldc unused)."""
    if m <= 0 or n <= 0:
        return
    if A.s != B.s:
        raise ValueError("operands sliced with different slice counts")
    if tri_a is not None:
        flags |= 4
    if tri_b is not None:
        flags |= 8
    row0 = col0 = pbase = 0
    pdiag = 1.0
    if packed is not None:
        flags |= 16
        row0, col0, pbase, pdiag = packed
    ctx.check(ctx.lib.cfmi_gemm_s8(ctx.h, A.s, m, n, k, alpha, P(A.data, a_row0 * A.ldk), A.ldk, A.ps,
                                   P(A.mu, a_row0), tri_a or 0, P(B.data, b_row0 * B.ldk), B.ldk, B.ps,
                                   P(B.mu, b_row0), tri_b or 0, beta, P(C, offc), ldc, flags, row0, col0, pbase,
                                   pdiag), "gemm_s8")


def transpose(ctx: Context, A: torch.Tensor, m: int, n: int, lda: int) -> torch.Tensor:
    """A^T (n x m, column-major) of a column-major m x n FP64 matrix."""
    T = ctx.empty(n, m)
    ctx.check(ctx.lib.cfmi_transpose(ctx.h, m, n, P(A), lda, P(T), n), "transpose")
    return T


def trsm(ctx: Context, f: Factor, n: int, B: torch.Tensor, X: torch.Tensor, nrhs: int, transpose: bool = False,
         slices: int = 0, outer: int = 2048, diag_slices: int = 0):
    """X = L^-1 B (or L^-T B with ``transpose``) for the factor's L, n x nrhs
    column-major with leading dimension n; B is consumed (it holds the
    running right-hand side).

    Two-level blocking: inside an ``outer``-row block, nb-row steps (nb the
    factor's diagonal-block size) whose diagonal solve is an FP64 product
    with the block inverse (``factor(..., diag_blocks=nb)``) and whose
    update reaches the rows of the same outer block; after each outer block
    one update of every remaining row with K = ``outer``, so the right-hand
    side is read and written n / outer times instead of n / nb.  Updates run
    in FP64, or on the int8 matrix cores with ``slices`` = s: the rows of L's
    panel (rows of L^T, kept once per factor) and the solved block's columns
    are sliced per update, each row scaled over the panel's width.  The
    diagonal solves likewise with ``diag_slices`` (the block inverses sliced
    once per factor; worth it for large nb, where K = nb is long)."""
    nb = f.nb
    if not nb or f.U is None:
        raise ValueError("trsm needs the factor's diagonal-block inverses (factor(..., diag_blocks=nb))")
    outer = max(nb, outer // nb * nb)
    L, U = f.L, f.U
    if slices and not transpose and getattr(f, "Lt", None) is None:
        f.Lt = transpose_tri(ctx, L, n)
    As = Bs = None

    def update(r0: int, r1: int, k0: int, k1: int):
        """B[r0:r1] -= L[r0:r1, k0:k1] X[k0:k1] (forward) or
        B[r0:r1] -= L[k0:k1, r0:r1]^T X[k0:k1] (transpose)."""
        nonlocal As, Bs
        m, w = r1 - r0, k1 - k0
        if m <= 0:
            return
        if not slices:
            if not transpose:
                gemm(ctx, False, False, m, nrhs, w, -1.0, L, r0 + k0 * n, n, X, k0, n, 1.0, B, r0, n)
            else:
                gemm(ctx, True, False, m, nrhs, w, -1.0, L, k0 + r0 * n, n, X, k0, n, 1.0, B, r0, n)
            return
        if As is None:
            As, Bs = Slices(ctx, slices, n, outer), Slices(ctx, slices, nrhs, outer)
        # rows r of the A operand: L[r, k0:k1] is column r of L^T; L[k0:k1, r] column r of L
        src = f.Lt if not transpose else L
        slice8(ctx, slices, 0, m, w, src, k0 + r0 * n, n, As)
        slice8(ctx, slices, 0, nrhs, w, X, k0, n, Bs)
        gemm_s8(ctx, m, nrhs, w, -1.0, As, 0, Bs, 0, 1.0, B, r0, n)

    def diag_solve(j0: int, j1: int):
        """X[j0:j1] = U_jj B[j0:j1] (forward) or U_jj^T B[j0:j1] (transpose)."""
        nonlocal Ds
        w = j1 - j0
        o = j0 + j0 * n
        if not diag_slices:
            gemm(ctx, transpose, False, w, nrhs, w, 1.0, U, o, n, B, j0, n, 0.0, X, j0, n, 2 if transpose else 1)
            return
        # rows of U_jj (forward) or of U_jj^T = columns of U_jj (transpose)
        if getattr(f, "u_slices", None) is None:
            f.u_slices = {}
        key = (diag_slices, transpose, j0)
        if key not in f.u_slices:
            f.u_slices[key] = (slice8(ctx, diag_slices, 0, w, w, U, o, n) if transpose
                               else slice8(ctx, diag_slices, 1, w, w, U, o, n))
        if Ds is None:
            Ds = Slices(ctx, diag_slices, nrhs, nb)
        slice8(ctx, diag_slices, 0, nrhs, w, B, j0, n, Ds)
        gemm_s8(ctx, w, nrhs, w, 1.0, f.u_slices[key], 0, Ds, 0, 0.0, X, j0, n, tri_a=None if transpose else 0)

    Ds = None

    outers = [(o0, min(n, o0 + outer)) for o0 in range(0, n, outer)]
    if transpose:
        outers = outers[::-1]
    for o0, o1 in outers:
        inner = [(j0, min(o1, j0 + nb)) for j0 in range(o0, o1, nb)]
        if transpose:
            inner = inner[::-1]
        for j0, j1 in inner:
            diag_solve(j0, j1)
            if not transpose:
                update(j1, o1, j0, j1)
            else:
                update(o0, j0, j0, j1)
        if not transpose:
            update(o1, n, o0, o1)
        else:
            update(0, o0, o0, o1)
    return X


def transpose_tri(ctx: Context, L: torch.Tensor, n: int) -> torch.Tensor:
    """L^T of a column-major n x n matrix (its rows become contiguous)."""
    return transpose(ctx, L, n, n, n)


class S8Whitener:
    """The whitener W (lower triangular, column-major n x n FP64) as int8
    slices per field column block: for a block of columns [r0, r0 + m),
    rows r0..n-1 of W[:, block] (the rows above r0 are zero there), each row
    scaled by its largest entry within the block."""

    def __init__(self, ctx: Context, W: torch.Tensor, n: int, geometry, s: int = 5):
        self.s, self.n = int(s), int(n)
        self.blocks = {}
        for g in geometry:
            for r0, m in ((g.ra, g.ma), (g.rb, g.mb)):
                if (r0, m) not in self.blocks:
                    self.blocks[(r0, m)] = slice8(ctx, self.s, 1, n - r0, m, W, r0 + r0 * n, n)

    def __getitem__(self, key) -> Slices:
        return self.blocks[key]

    @property
    def nbytes(self) -> int:
        return sum(v.data.numel() + 8 * v.mu.numel() for v in self.blocks.values())


class S8Work:
    """Reused buffers of ``whiten_band_s8``: one bin's FP64 intermediate
    product and the slices of it and of the derivative block."""

    def __init__(self, ctx: Context, s: int):
        self.ctx, self.s = ctx, s
        self.z = None
        self.sl = {}

    def fp64(self, count: int) -> torch.Tensor:
        if self.z is None or self.z.numel() < count:
            self.z = None
            self.z = torch.empty(count, dtype=torch.float64, device=self.ctx.device)
        return self.z

    def slices(self, name: str, rows: int, k: int) -> Slices:
        b = self.sl.get(name)
        if b is None or b.data.shape[1] < rows or b.ldk < k:
            self.sl[name] = None
            b = self.sl[name] = Slices(self.ctx, self.s, rows, k)
        return b

    @staticmethod
    def bytes_for(s: int, rows: int, width: int) -> int:
        """Device bytes one bin needs for ``rows`` intermediate rows of a
        block ``width`` wide (FP64 product, its slices, the block's slices)."""
        ld = (width + 63) // 64 * 64
        return rows * width * 8 + s * rows * ld + 2 * s * width * ld + 16 * (rows + width)


def whiten_band_s8(ctx: Context, Ws: S8Whitener, n: int, blk: Block, X: torch.Tensor, r0: int, r1: int,
                   work: S8Work | None = None, counter: list | None = None) -> S8Work:
    """``whiten_band`` with every product on the int8 matrix cores at FP64
    accuracy: X[b] (FP64, packed rows r0..r1-1) += the lower triangle of
    W dC_b W^T, diagonal scaled by 1/sqrt(2), bin by bin.

    Per bin, Z^T = D^T W[i, a]^T is formed in FP64 from the sliced columns
    of D and the sliced rows of W (their zeros right of the diagonal trimmed
    per tile), re-sliced by rows of Z, and multiplied by W[c, b]^T into the
    packed band; the transposed term of an off-diagonal block reuses Z for a
    full band, or forms (W[j, b] D^T)^T from the sliced rows of D for a
    partial one."""
    s = Ws.s
    if work is None:
        work = S8Work(ctx, s)
    D = blk.stack
    nbb = D.shape[0]
    ra, ma, rb, mb = blk.ra, blk.ma, blk.rb, blk.mb
    Wa, Wb = Ws[(ra, ma)], Ws[(rb, mb)]
    sX = X.shape[1]
    pbase = r0 * (r0 + 1) // 2
    full = r0 == 0 and r1 == n
    i0, j0 = max(r0, ra), max(r0, rb)
    rr, rj = r1 - i0, r1 - j0

    def gemm(*a, **k):
        if counter is not None:
            counter.append(1)
        gemm_s8(*a, **k)

    for b in range(nbb):
        off = b * ma * mb
        if rr > 0:
            # term 1: X[i, c] += sum_k (W[i, a] D)[k] W[c, rb + k],  i >= ra, c >= rb
            Dc = slice8(ctx, s, 0, mb, ma, D, off, ma, out=work.slices("dcol", mb, ma))
            ZT = work.fp64(rr * mb)
            gemm(ctx, mb, rr, ma, 1.0, Dc, 0, Wa, i0 - ra, 0.0, ZT, 0, mb, tri_b=i0 - ra)
            Zs = slice8(ctx, s, 0, rr, mb, ZT, 0, mb, out=work.slices("z", rr, mb))
            if r1 > rb:
                gemm(ctx, rr, r1 - rb, mb, 1.0, Zs, 0, Wb, 0, 1.0, X, b * sX, 0, tri_b=0,
                     packed=(i0, rb, pbase, SQRT_HALF))
        if blk.sym:
            continue
        # term 2, the transpose: X[i, c] += sum_k W[i, rb + k] (W[c, a] D)[k],  i >= rb, c >= ra
        if full and rr > 0:
            gemm(ctx, n - rb, n - ra, mb, 1.0, Wb, 0, Zs, 0, 1.0, X, b * sX, 0, tri_a=0,
                 packed=(rb, ra, 0, SQRT_HALF))
            continue
        if rj <= 0 or r1 <= ra:
            continue
        # partial band: Z2^T = D (W[j, b])^T, then X[j, c] += sum_q Z2[j, q] W[c, ra + q]
        Dr = slice8(ctx, s, 1, ma, mb, D, off, ma, out=work.slices("drow", ma, mb))
        Z2T = work.fp64(rj * ma)
        gemm(ctx, ma, rj, mb, 1.0, Dr, 0, Wb, j0 - rb, 0.0, Z2T, 0, ma, tri_b=j0 - rb)
        Z2s = slice8(ctx, s, 0, rj, ma, Z2T, 0, ma, out=work.slices("z2", rj, ma))
        gemm(ctx, rj, r1 - ra, ma, 1.0, Z2s, 0, Wa, 0, 1.0, X, b * sX, 0, tri_b=0,
             packed=(j0, ra, pbase, SQRT_HALF))
    return work
