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

import ctypes

import numpy as np
import torch

from . import engine, in_out
from .core import Inputs, shape_weights
from . import _lib
from .engine import P


class Covariance:
    """Device-resident geometry, noise and (on demand) S, C and its factor."""

    def __init__(self, ctx: engine.Context, inputs: Inputs, noise=None, vecs=None, pixel_filter=None):
        """``noise``: another noise covariance over the same pixels (the
        second one of a cross analysis), ``vecs``: pointings to share,
        ``pixel_filter``: a ``filters.DeviceFilter``; N, S and C then live in
        its r-dimensional range (``core.py:628-634``, ``signal_kernels.py``'s
        ``pixel_filter=`` seam) and ``n`` is r."""
        self.ctx = ctx
        self.inp = inputs
        p = inputs.params
        if vecs is None:
            nest = str(p.ordering).upper() == "NESTED"
            vecs = [engine.pointings(ctx, p.nside, f.pixels, nest) for f in inputs.layout.fields]
            if not in_out._falsy(p.output_geometry_file):
                fields = inputs.layout.fields
                in_out.output_geometry(p.output_geometry_file, [f.pixels.size for f in fields],
                                       [v.cpu().numpy() for v in vecs], [f.pixels for f in fields])
        self.vecs = vecs
        self.n = inputs.n
        N = inputs.noise if noise is None else noise
        if isinstance(N, torch.Tensor):
            # (n, n) row-major device tensor -> column-major copy on the device
            self.N = N.to(device=ctx.device, dtype=torch.float64).T.contiguous()
        else:
            self.N = ctx.upload_cm(N)
        self.pixel_filter = pixel_filter
        if pixel_filter is not None:
            if pixel_filter.n != self.n:
                raise ValueError(f"pixel_filter was built for {pixel_filter.n} pixels but this run has "
                                 f"{self.n} active pixels; rebuild it against this mask")
            self.N = pixel_filter.restrict_matrix(self.N, prefiltered=pixel_filter.noise_prefiltered)
            self.n = pixel_filter.r
        self.S = None
        self.factor = None
        self._w32 = None
        self._wp = None
        self._noise_diag = None

    def noise_is_diagonal(self) -> bool:
        """True when N has no nonzero off-diagonal entry (counted once, on
        the device)."""
        if self._noise_diag is None:
            cnt = ctypes.c_int64()
            self.ctx.check(self.ctx.lib.cfmi_offdiag_nnz(self.ctx.h, self.n, P(self.N), self.n, ctypes.byref(cnt)),
                           "offdiag_nnz")
            self._noise_diag = cnt.value == 0
        return self._noise_diag

    def signal(self, cls=None) -> torch.Tensor:
        ctx = self.ctx
        cs = self.inp.smoothed_cls(cls)
        dev = {k: ctx.upload(v) for k, v in cs.items()}
        if self.pixel_filter is not None:
            full = engine.signal_matrix(ctx, self.inp.layout, self.vecs, dev, self.inp.lmax_signal)
            self.S = self.pixel_filter.restrict_matrix(full)
            return self.S
        self.S = engine.signal_matrix(ctx, self.inp.layout, self.vecs, dev, self.inp.lmax_signal, out=self.S)
        return self.S

    def build(self, cls=None, need_whitener: bool = True, diag_blocks: int = 0, slices: int = 0) -> engine.Factor:
        """S, then C = S + N factorised in a separate buffer (S is kept).
        ``diag_blocks``: keep the inverses of L's diagonal blocks (for
        ``engine.trsm``) instead of forming the whitener; ``slices``: the
        factorisation's trailing updates on int8 slices (``engine.factor``)."""
        ctx, n = self.ctx, self.n
        S = self.signal(cls)
        C = ctx.empty(n, n)
        ctx.check(ctx.lib.cfmi_axpby(ctx.h, n, n, 1.0, P(S), n, 0.0, P(C), n), "axpby")
        if self.noise_is_diagonal():
            # only the diagonal of N: an n-vector with stride n + 1 on both sides
            ctx.check(ctx.lib.cfmi_axpby(ctx.h, 1, n, 1.0, P(self.N), n + 1, 1.0, P(C), n + 1), "axpby")
        else:
            ctx.check(ctx.lib.cfmi_axpby(ctx.h, n, n, 1.0, P(self.N), n, 1.0, P(C), n), "axpby")
        self.factor = engine.factor(ctx, C, n, need_whitener=need_whitener and not diag_blocks,
                                    diag_blocks=diag_blocks, slices=slices)
        self._w32 = None
        self._wp = None
        return self.factor

    def inverse(self) -> torch.Tensor:
        f = self.factor
        if f.W is not None:
            return engine.inverse_from_whitener(self.ctx, f.W, self.n)
        # C^-1 = D L^-T L^-1 D from the blocked solves (no whitener formed)
        ctx, n = self.ctx, self.n
        X = ctx.empty(n, n)
        ctx.check(ctx.lib.cfmi_fill(ctx.h, engine.F64, n, n, P(X), n, 0.0), "fill")
        ctx.check(ctx.lib.cfmi_copy(ctx.h, engine.F64, 1, n, P(f.d), 1, P(X), n + 1), "copy")
        T = ctx.empty(n, n)
        engine.trsm(ctx, f, n, X, T, n)
        engine.trsm(ctx, f, n, T, X, n, transpose=True)
        ctx.check(ctx.lib.cfmi_scale(ctx.h, 1, n, n, P(X), n, P(f.d)), "scale")
        return X


def make_source(cov: Covariance, bins, convention: str, dtype, budget: int | None = None):
    """The derivative source of ``cov``: pixel-space block stacks, or their
    restrictions to the filter's range when ``cov`` carries a pixel filter."""
    if getattr(cov, "pixel_filter", None) is not None:
        return RestrictedSource(cov, bins, convention, dtype, budget)
    return DerivativeSource(cov, bins, convention, dtype, budget)


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


def whitening_flops(n: int, blocks, nbins: int, bands=None) -> float:
    """Multiply-add flops (x2) of the whitened-derivative GEMMs, counting the
    products the kernels need: the nonzero (lower-triangular) part of W and,
    for the output, its lower triangle only.  ``bands``: the row bands the
    work is split into (one band: the whole matrix)."""
    bands = bands or [(0, n)]
    tot = 0.0
    for blk in blocks:
        ra, ma, rb, mb = blk.ra, blk.ma, blk.rb, blk.mb
        for r0, r1 in bands:
            i = np.arange(max(r0, ra), r1)
            tot += 2.0 * mb * np.minimum(i - ra + 1, ma).sum() * nbins      # Z = W[i, a] D
            tot += 2.0 * _tri_sum(i, rb, mb) * nbins                        # Z W[c, b]^T
            if blk.sym:
                continue
            if r0 == 0 and r1 == n:                                         # W[i, b] Z^T
                j = np.arange(rb, n)
                tot += 2.0 * (np.minimum(j - rb + 1, mb) * np.maximum(j - ra + 1, 0)).sum() * nbins
            else:                                                           # Z2 = W[j, b] D^T, Z2 W[c, a]^T
                j = np.arange(max(r0, rb), r1)
                tot += 2.0 * ma * np.minimum(j - rb + 1, mb).sum() * nbins
                tot += 2.0 * _tri_sum(j, ra, ma) * nbins
    return float(tot)


def _tri_sum(rows: np.ndarray, c0: int, kmax: int) -> float:
    """sum over rows i and columns c in [c0, i] of min(c - c0 + 1, kmax)."""
    m = np.maximum(rows - c0 + 1, 0)            # columns per row
    # sum_{t=1..m} min(t, kmax)
    full = np.minimum(m, kmax)
    return float((full * (full + 1) / 2 + np.maximum(m - kmax, 0) * kmax).sum())


def plan_bands(n: int, npar: int, esize: int, budget: int, align: int = 128) -> list:
    """Row bands [r0, r1) of the lower triangle such that one band's packed
    whitened derivatives for all npar parameters fit in ``budget`` bytes."""
    if npar * esize * band_size(0, n) <= budget:
        return [(0, n)]
    bands = []
    r0 = 0
    cap = budget // (npar * esize)
    while r0 < n:
        lo, hi = 1, n - r0
        while lo < hi:                                   # largest r with band_size(r0, r0 + r) <= cap
            mid = (lo + hi + 1) // 2
            if band_size(r0, r0 + mid) <= cap:
                lo = mid
            else:
                hi = mid - 1
        r = lo
        if r0 + r < n and r > align:
            r = (r // align) * align
        if band_size(r0, r0 + r) > cap:
            raise MemoryError(f"one row of {n} packed whitened derivatives for {npar} parameters exceeds "
                              f"the {budget / 2**30:.1f} GiB budget")
        bands.append((r0, r0 + r))
        r0 += r
    return bands


def band_size(r0: int, r1: int) -> int:
    return (r1 * (r1 + 1) - r0 * (r0 + 1)) // 2


PRECISIONS = ("fp64", "mixed")


def _dtype(precision: str):
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {PRECISIONS}; got {precision!r}")
    return torch.float64 if precision == "fp64" else torch.float32


def whitener(cov: Covariance, precision: str) -> torch.Tensor:
    """W in the arithmetic type of ``precision`` (FP32: W rounded once)."""
    W = cov.factor.W
    if precision == "fp64":
        return W
    if getattr(cov, "_w32", None) is None:
        n = cov.n
        w32 = torch.empty((n, n), dtype=torch.float32, device=cov.ctx.device)
        cov.ctx.check(cov.ctx.lib.cfmi_convert(cov.ctx.h, 0, n, n, P(W), n, P(w32), n), "convert")
        cov._w32 = w32
    return cov._w32


def whitener_planes(cov: Covariance) -> torch.Tensor:
    """W as three-piece BF16 row-major planes (3, n, ldw), split once from
    FP64 (each entry to 2^-26 relative)."""
    if getattr(cov, "_wp", None) is None:
        cov._wp = engine.split_planes(cov.ctx, cov.factor.W, cov.n, cov.n, cov.n)
    return cov._wp


MIXED_GEMMS = ("s8", "bf16x3", "fp32")

# int8 slices per operand of the sliced whitening products (each entry held
# to 2^-39 of its row's largest entry; 15 int8 products per multiply-add)
S8_WHITEN_SLICES = 5


def s8_whitener(cov: Covariance, geometry) -> engine.S8Whitener:
    """W as int8 slices per field column block, made once per factor."""
    key = tuple((g.ra, g.ma, g.rb, g.mb) for g in geometry)
    cached = getattr(cov, "_ws8", None)
    if cached is None or cached[0] is not cov.factor or cached[1] != key:
        cov._ws8 = None
        cov._ws8 = (cov.factor, key, engine.S8Whitener(cov.ctx, cov.factor.W, cov.n, geometry,
                                                         S8_WHITEN_SLICES))
    return cov._ws8[2]


def _stack_bytes(dtype) -> int:
    if dtype == engine.BF16X3:
        return 6
    return 4 if dtype == torch.float32 else 8


class DerivativeSource:
    """The binned derivative stacks of one (Covariance, bins) pair: kept on
    the device when they fit, regenerated spectrum by spectrum (and in bin
    chunks) otherwise."""

    def __init__(self, cov: Covariance, bins, convention: str, dtype, budget: int | None = None,
                 mirrors: bool = False):
        self.cov, self.bins, self.dtype, self.convention = cov, bins, dtype, convention
        inp = cov.inp
        for lo, hi in bins:
            if hi > inp.lmax_signal:
                raise ValueError(f"bin [{lo}, {hi}] extends beyond lmax_signal={inp.lmax_signal}")
        w = parameter_weights(inp, bins, convention)
        self.weights = {k: cov.ctx.upload(v) for k, v in w.items()}
        self.geometry = engine.block_geometry(inp.layout)
        self.resident = None
        self.budget = budget
        self.set_mirrors(mirrors)

    def set_mirrors(self, on: bool):
        """Three-piece stacks: also produce D^T of the a != b blocks (the
        banded whitening's operand); counted in the byte totals."""
        self.mirrors = bool(on) and self.dtype == engine.BF16X3
        self.resident = None
        es = _stack_bytes(self.dtype)
        self.bytes_per_bin = [g.ma * g.mb * es * (2 if (self.mirrors and not g.sym) else 1)
                              for g in self.geometry]
        self.total_bytes = sum(self.bytes_per_bin) * len(self.bins)

    def fits(self, budget: int) -> bool:
        return self.total_bytes <= budget

    def make_resident(self):
        if self.resident is None:
            self.resident = self._make(None, None)
        return self.resident

    def _make(self, only, bin_range, state=None, lresume=0, pool=None):
        inp = self.cov.inp
        return engine.deriv_stacks(self.cov.ctx, inp.layout, self.cov.vecs, self.weights, inp.lmax_signal,
                                   self.bins, dtype=self.dtype, only=only, bin_range=bin_range, state=state,
                                   lresume=lresume, pool=pool, mirrors=self.mirrors)

    def chunks(self, chunk_budget: int):
        """Yield (spectrum index, first bin, Block with stack) over all
        spectra and bins; sibling spectra sharing a recurrence pass are
        produced together."""
        if self.resident is not None:
            for s, blk in enumerate(self.resident):
                yield s, 0, blk
            return
        specs = self.cov.inp.layout.spectra_list()
        nb = len(self.bins)
        groups = []
        seen = set()
        for s, sp in enumerate(specs):
            if s in seen:
                continue
            g = [m for m, s2 in enumerate(specs) if (s2.fi, s2.fj) == (sp.fi, sp.fj)]
            seen.update(g)
            groups.append(g)
        plan = []
        for g in groups:
            per_bin = sum(self.bytes_per_bin[m] for m in g)
            plan.append((g, max(1, min(nb, chunk_budget // max(per_bin, 1))), per_bin))
        # One pool for every chunk (no per-chunk allocations of tens of GB).
        pdt = torch.int16 if self.dtype == engine.BF16X3 else self.dtype
        pes = torch.empty((), dtype=pdt).element_size()
        pool = torch.empty(max(cb * pb for _, cb, pb in plan) // pes, dtype=pdt, device=self.cov.ctx.device)
        for g, cb, _ in plan:
            # Consecutive bin chunks continue one recurrence pass.
            state = engine.deriv_state(self.cov.ctx, self.cov.inp.layout, g[0]) if cb < nb else None
            for b0 in range(0, nb, cb):
                lres = self.bins[b0 - 1][1] + 1 if b0 > 0 else 0
                blocks = self._make(g, (b0, min(nb, b0 + cb)), state, lres, pool)
                for m in g:
                    yield m, b0, blocks[m]
                del blocks
            del state
        del pool


class RestrictedSource:
    """Binned derivatives restricted to a pixel filter's range,
    Sigma W^T dC_b W Sigma (``core.py:1455-1469``): dense symmetric r x r
    blocks (one ``Block(0, r, 0, r)`` per spectrum), made from the
    pixel-space block stacks chunk by chunk (``DeviceFilter.restrict_stack``)."""

    def __init__(self, cov: Covariance, bins, convention: str, dtype, budget: int | None = None):
        if dtype != torch.float64:
            raise ValueError("a pixel filter runs in fp64")
        self.cov, self.bins, self.dtype, self.convention = cov, bins, dtype, convention
        self.base = DerivativeSource(cov, bins, convention, torch.float64, budget)
        r = cov.n
        self.geometry = [engine.Block(0, r, 0, r) for _ in self.base.geometry]
        self.weights = self.base.weights
        self.mirrors = False
        self.bytes_per_bin = [r * r * 8 for _ in self.geometry]
        self.total_bytes = sum(self.bytes_per_bin) * len(bins)
        self.resident = None
        self.budget = budget

    def set_mirrors(self, on: bool):
        self.mirrors = False

    def fits(self, budget: int) -> bool:
        return self.total_bytes <= budget

    def _restricted(self, chunk_budget: int):
        r = self.cov.n
        grow = max(r * r / max(g.ma * g.mb, 1) for g in self.base.geometry)
        for s, b0, blk in self.base.chunks(max(1, int(chunk_budget / (1.0 + grow)))):
            R = self.cov.pixel_filter.restrict_stack(blk)
            yield s, b0, engine.Block(0, r, 0, r, stack=R)
            del blk

    def make_resident(self):
        if self.resident is None:
            r, nb = self.cov.n, len(self.bins)
            out = [engine.Block(0, r, 0, r, stack=torch.empty((nb, r, r), dtype=torch.float64,
                                                             device=self.cov.ctx.device))
                   for _ in self.geometry]
            for s, b0, blk in self._restricted(int(_free_bytes(self.cov.ctx) * 0.25)):
                k = blk.stack.shape[0]
                out[s].stack[b0: b0 + k].copy_(blk.stack)
                del blk
            self.resident = out
        return self.resident

    def chunks(self, chunk_budget: int):
        if self.resident is not None:
            for s, blk in enumerate(self.resident):
                yield s, 0, blk
            return
        yield from self._restricted(chunk_budget)


def _free_bytes(ctx) -> int:
    """Device memory available to new buffers: free memory plus what torch's
    caching allocator holds but does not use."""
    free, _ = torch.cuda.mem_get_info(ctx.device)
    cached = torch.cuda.memory_reserved(ctx.device) - torch.cuda.memory_allocated(ctx.device)
    return int(free + max(cached, 0))


class LaunchTimer:
    """HIP events around individual kernel launches on the context's stream
    (the stream the library enqueues on)."""

    def __init__(self):
        self.pairs = []
        self._open = None
        self.launches = []

    def start(self):
        self._open = torch.cuda.Event(enable_timing=True)
        self._open.record()

    def stop(self):
        b = torch.cuda.Event(enable_timing=True)
        b.record()
        self.pairs.append((self._open, b))
        self._open = None

    def total_ms(self) -> float:
        torch.cuda.synchronize()
        return sum(a.elapsed_time(b) for a, b in self.pairs)


def plan_memory(n: int, npar: int, es: int, geometry, nbins: int, stack_bytes: int, free: int,
                es_z: int | None = None):
    """(bands, stacks_resident, chunk_budget, z_budget) for the banded Fisher.

    One band holds npar packed row bands of es-byte entries; the derivative
    stacks stay resident when they fit beside it, otherwise they are
    regenerated per band in chunks of at most ``chunk_budget`` bytes.  The
    intermediate products W D take es_z bytes per entry (default es)."""
    es_z = es if es_z is None else es_z
    maxw = max(max(g.ma, g.mb) for g in geometry)
    zfull = max((n - min(g.ra, g.rb)) * maxw for g in geometry) * nbins * es_z
    if stack_bytes + npar * es * band_size(0, n) + zfull <= free:
        return [(0, n)], True, stack_bytes, zfull
    resident = stack_bytes <= free // 4
    chunk = stack_bytes if resident else free // 4
    zres = free // 16
    bands = plan_bands(n, npar, es, free - chunk - zres)
    h = max(r1 - r0 for r0, r1 in bands)
    if h * maxw * es_z > zres:        # one bin of the tallest band must fit
        zres = h * maxw * es_z
        bands = plan_bands(n, npar, es, free - chunk - zres)
    return bands, resident, chunk, zres


def fisher(cov: Covariance, bins, convention: str = "Cl", timer: StageTimer | None = None,
           precision: str = "fp64", keep_stacks: bool = True, mem_fraction: float = 0.8,
           launch_timer: LaunchTimer | None = None, mem_budget: int | None = None, progress=None,
           mixed_gemm: str = "s8"):
    """(F, derivative source): F is (n_p, n_p) on the host.

    F_ij = 1/2 <W dC_i W^T, W dC_j W^T>, accumulated over row bands of the
    packed lower triangles (one band when everything fits on the device).
    ``precision="mixed"`` runs the whitening products on the low-precision
    matrix cores; ``mixed_gemm`` picks how:

    * "s8" (default): FP64-accurate products from int8 slices of W, D and
      the intermediate W D (``cfmi_gemm_s8``, exact int32 sums); the
      whitened derivatives, their Gram matrix and everything else stay FP64;
    * "bf16x3": FP32-accurate products from three-piece BF16 operands
      (``cfmi_gemm_b3``), FP32 whitened derivatives, their Gram matrix from
      FP32 products accumulated in FP64 across 256-term chunks;
    * "fp32": W and D rounded to FP32, FP32 MFMA, FP32 storage as above.

    The FP32 variants lose accuracy in proportion to C's conditioning (the
    error of W D W^T is cancelled in F only to that level); the sliced
    products do not."""
    ctx, n = cov.ctx, cov.n
    if mixed_gemm not in MIXED_GEMMS:
        raise ValueError(f"mixed_gemm must be one of {MIXED_GEMMS}; got {mixed_gemm!r}")
    s8 = precision == "mixed" and mixed_gemm == "s8"
    T = torch.float64 if s8 else _dtype(precision)
    b3 = precision == "mixed" and mixed_gemm == "bf16x3"
    es = 4 if T == torch.float32 else 8
    sdt = engine.BF16X3 if b3 else T
    src = make_source(cov, bins, convention, sdt)
    if s8:
        W = s8_whitener(cov, src.geometry)
        es_z = 8 + 3 * W.s
    else:
        W = whitener_planes(cov) if b3 else whitener(cov, precision)
        es_z = _stack_bytes(sdt)
    nb = len(bins)
    npar = len(src.geometry) * nb
    # Whitening products (K <= a few 10^4, terms of both signs) accumulate in
    # FP32 inside the MFMA (FP64 chunk accumulation there was measured to
    # change nothing: the FP32 error of F comes from rounding W and D); the
    # Gram matrix (K ~ n^2 / 2, positive diagonal terms) accumulates in FP64
    # across 256-term chunks.
    flags = 0
    free = int(mem_budget) if mem_budget is not None else int(_free_bytes(ctx) * mem_fraction)
    # the sliced whitening takes one bin at a time
    bands, resident, chunk_budget, zbytes = plan_memory(n, npar, es, src.geometry, 1 if s8 else nb,
                                                        src.total_bytes, free, es_z)
    if b3 and len(bands) > 1 and not all(g.sym for g in src.geometry):
        src.set_mirrors(True)             # partial bands multiply the a != b blocks by W D^T
        bands, resident, chunk_budget, zbytes = plan_memory(n, npar, es, src.geometry, nb, src.total_bytes,
                                                            free, es_z)
    if timer:
        timer.mark("deriv0")
    if resident:
        src.make_resident()
    if timer:
        timer.flops = whitening_flops(n, src.geometry, nb, bands)
        timer.bands = bands
    G = torch.zeros((npar, npar), dtype=torch.float64, device=ctx.device)
    zbuf = None
    xbuf = torch.empty(npar * max(band_size(r0, r1) for r0, r1 in bands), dtype=T, device=ctx.device)
    for bi, (r0, r1) in enumerate(bands):
        X = xbuf[:npar * band_size(r0, r1)].view(npar, band_size(r0, r1))
        ctx.check(ctx.lib.cfmi_fill(ctx.h, engine.dtype_code(X), X.numel(), 1, P(X), X.numel(), 0.0), "fill")
        if timer:
            timer.mark("white0")
        for s, b0, blk in src.chunks(chunk_budget):
            k = blk.nbins
            Xs = X[s * nb + b0: s * nb + b0 + k]
            if launch_timer is not None:
                launch_timer.start()
            cnt = launch_timer.launches if launch_timer is not None else None
            if s8:
                zbuf = engine.whiten_band_s8(ctx, W, n, blk, Xs, r0, r1, zbuf, counter=cnt)
            elif b3:
                zbuf = engine.whiten_band_b3(ctx, W, n, blk, Xs, r0, r1, zbuf, zbytes // es_z, counter=cnt)
            else:
                zbuf = engine.whiten_band(ctx, W, n, blk, Xs, r0, r1, zbuf, flags, zbytes // es_z, counter=cnt)
            if launch_timer is not None:
                launch_timer.stop()
            del blk, Xs                  # the next chunk is allocated before the loop variable is rebound
        if timer:
            timer.mark("white1")
        engine.gram(ctx, npar, npar, X.shape[1], X, X, out=G, beta=0.0 if bi == 0 else 1.0)
        if timer:
            timer.mark("gram1")
        del X
        if progress is not None:
            progress(bi + 1, len(bands))
    del zbuf, xbuf
    if not keep_stacks:
        src.resident = None
    Gh = engine.to_numpy_cm(G)
    F = 0.5 * (Gh + Gh.T)
    return F, src


def fisher_cross(cov1: Covariance, cov2: Covariance, bins, convention: str = "Cl",
                 mem_budget: int | None = None, mem_fraction: float = 0.8):
    """(F, derivative source) of a cross analysis (``do_cross``), FP64.

    F_ij = 1/2 Tr[C1^-1 dC_i C2^-1 dC_j] (``fisher.py:522-531,584-590``).  With
    C_k^-1 = W_k^T W_k and E_i = W1 dC_i W2^T this is 1/2 <E_i, E_j>, a Gram
    matrix of full (not symmetric) n x n products.  E_i is formed per
    spectrum and all bins at once: T = D W2[:, b]^T, E = W1[:, a] T, plus the
    transposed block's W1[:, b] (D^T W2[:, a]^T) when a != b.  The E of as
    many spectra as fit twice in the budget are kept; spectrum groups that do
    not fit together are regenerated per Gram block."""
    ctx, n = cov1.ctx, cov1.n
    src = make_source(cov1, bins, convention, torch.float64)
    nb = len(bins)
    geo = src.geometry
    npar = len(geo) * nb
    free = int(mem_budget) if mem_budget is not None else int(_free_bytes(ctx) * mem_fraction)
    tmax = nb * max(max(g.ma, g.mb) for g in geo) * n * 8
    if src.total_bytes + tmax > free // 2:
        raise ValueError(f"cross Fisher needs its derivative stacks resident: {src.total_bytes / 1e9:.1f} GB "
                         f"plus {tmax / 1e9:.1f} GB of staging exceed half the {free / 1e9:.1f} GB budget")
    blocks = src.make_resident()
    per = nb * n * n * 8                     # E of one spectrum
    g = max(1, min(len(geo), (free - src.total_bytes - tmax) // (2 * per)))
    groups = [list(range(s0, min(len(geo), s0 + g))) for s0 in range(0, len(geo), g)]
    W1, W2 = cov1.factor.W, cov2.factor.W
    T = torch.empty(tmax // 8, dtype=torch.float64, device=ctx.device)

    def whitened(specs, out):
        for k, s in enumerate(specs):
            blk = blocks[s]
            D, ma, mb = blk.stack, blk.ma, blk.mb
            E = out[k * nb: (k + 1) * nb]
            # T = D W2[:, rb:rb+mb]^T (ma x n per bin), E = W1[:, ra:ra+ma] T
            engine.gemm_batched(ctx, False, True, ma, n, mb, 1.0, D, 0, ma, ma * mb, W2, blk.rb * n, n, 0,
                                0.0, T, 0, ma, ma * n, nb)
            engine.gemm_batched(ctx, False, False, n, n, ma, 1.0, W1, blk.ra * n, n, 0, T, 0, ma, ma * n,
                                0.0, E, 0, n, n * n, nb)
            if not blk.sym:
                engine.gemm_batched(ctx, True, True, mb, n, ma, 1.0, D, 0, ma, ma * mb, W2, blk.ra * n, n, 0,
                                    0.0, T, 0, mb, mb * n, nb)
                engine.gemm_batched(ctx, False, False, n, n, mb, 1.0, W1, blk.rb * n, n, 0, T, 0, mb, mb * n,
                                    1.0, E, 0, n, n * n, nb)
        return out[: len(specs) * nb]

    F = np.zeros((npar, npar))
    bufA = torch.empty((g * nb, n * n), dtype=torch.float64, device=ctx.device)
    bufB = torch.empty((g * nb, n * n), dtype=torch.float64, device=ctx.device) if len(groups) > 1 else None
    for gi, I in enumerate(groups):
        EI = whitened(I, bufA)
        pi = slice(I[0] * nb, (I[-1] + 1) * nb)
        F[pi, pi] = engine.gram(ctx, EI.shape[0], EI.shape[0], n * n, EI, EI)
        for J in groups[gi + 1:]:
            EJ = whitened(J, bufB)
            pj = slice(J[0] * nb, (J[-1] + 1) * nb)
            Gij = engine.gram(ctx, EI.shape[0], EJ.shape[0], n * n, EI, EJ)      # (|I|, |J|)
            F[pi, pj] = Gij
            F[pj, pi] = Gij.T
    del T, bufA, bufB
    F = 0.25 * (F + F.T)              # 1/2 G, symmetrised
    return F, src


def qml_cross(cov1: Covariance, cov2: Covariance, src: DerivativeSource, maps1: np.ndarray, maps2: np.ndarray,
              chunk_budget: int | None = None):
    """(q, nb) of a cross analysis: q_b = 1/2 y2^T dC_b y1 with y_k = C_k^-1 x_k
    and no noise bias (``spectra.py:1146-1166``).  For dC_b = e_a D e_b^T
    (+ transpose) the contraction is <D, y2_a y1_b^T + y1_a y2_b^T>
    (the second term when a != b), one ``cfmi_dot_batched`` pass per spectrum."""
    ctx, n = cov1.ctx, cov1.n
    if maps1.shape != maps2.shape:
        raise ValueError(f"maps1 {maps1.shape} and maps2 {maps2.shape} differ")
    ns = maps1.shape[1]
    Y1 = solve(cov1, maps1, "fp64")
    Y2 = solve(cov2, maps2, "fp64")
    if src.dtype != torch.float64:
        src = make_source(cov1, src.bins, src.convention, torch.float64)
    nbb = len(src.bins)
    npar = len(src.geometry) * nbb
    out = np.zeros((ns, npar))
    if chunk_budget is None:
        chunk_budget = int(_free_bytes(ctx) * 0.25)
    cache = {}
    for s, b0, blk in src.chunks(chunk_budget):
        D = blk.stack
        k = D.shape[0]
        key = (blk.ra, blk.ma, blk.rb, blk.mb)
        if key not in cache:
            cache.clear()
            ma, mb = blk.ma, blk.mb
            M = torch.empty((ns, mb, ma), dtype=torch.float64, device=ctx.device)
            for q in range(ns):
                engine.gemm_ex(ctx, False, True, ma, mb, 1, 1.0, Y2, blk.ra + q * n, ma, Y1, blk.rb + q * n, mb,
                               0.0, M, q * ma * mb, ma)
                if not blk.sym:
                    engine.gemm_ex(ctx, False, True, ma, mb, 1, 1.0, Y1, blk.ra + q * n, ma, Y2, blk.rb + q * n,
                                   mb, 1.0, M, q * ma * mb, ma)
            cache[key] = M
        M = cache[key]
        R = torch.empty((ns, k), dtype=torch.float64, device=ctx.device)
        K = blk.ma * blk.mb
        ctx.check(ctx.lib.cfmi_dot_batched(ctx.h, K, k, P(D), K, ns, P(M), K, P(R), k), "dot_batched")
        out[:, s * nbb + b0: s * nbb + b0 + k] = 0.5 * R.cpu().numpy()
        del blk, D
    return out, np.zeros(npar)


def whitened_maps(cov: Covariance, maps: np.ndarray) -> torch.Tensor:
    """z = W x for every simulation (n x nsims, column-major)."""
    ctx, n = cov.ctx, cov.n
    X = ctx.upload_cm(maps)
    ns = maps.shape[1]
    Z = ctx.empty(n, ns)
    engine.gemm(ctx, False, False, n, ns, n, 1.0, cov.factor.W, 0, n, X, 0, n, 0.0, Z, 0, n, 1)
    return Z


def solve(cov: Covariance, maps: np.ndarray, precision: str = "fp64", iters: int = 3) -> torch.Tensor:
    """y = C^-1 x (n x nsims, FP64, column-major).

    fp64: y = W^T (W x).  mixed: FP32 solves through W (rounded to FP32)
    refined with FP64 residuals r = x - (S + N) y."""
    ctx, n = cov.ctx, cov.n
    ns = maps.shape[1]
    X = ctx.upload_cm(maps)
    if precision == "fp64":
        W = cov.factor.W
        Z = ctx.empty(n, ns)
        engine.gemm(ctx, False, False, n, ns, n, 1.0, W, 0, n, X, 0, n, 0.0, Z, 0, n, 1)
        Y = ctx.empty(n, ns)
        engine.gemm(ctx, True, False, n, ns, n, 1.0, W, 0, n, Z, 0, n, 0.0, Y, 0, n, 2)
        return Y
    W32 = whitener(cov, precision)

    def apply_inv32(R64):
        r32 = torch.empty((ns, n), dtype=torch.float32, device=ctx.device)
        ctx.check(ctx.lib.cfmi_convert(ctx.h, 0, n, ns, P(R64), n, P(r32), n), "convert")
        z32 = torch.empty_like(r32)
        engine.gemm(ctx, False, False, n, ns, n, 1.0, W32, 0, n, r32, 0, n, 0.0, z32, 0, n, 1)
        y32 = torch.empty_like(r32)
        engine.gemm(ctx, True, False, n, ns, n, 1.0, W32, 0, n, z32, 0, n, 0.0, y32, 0, n, 2)
        y64 = ctx.empty(n, ns)
        ctx.check(ctx.lib.cfmi_convert(ctx.h, 1, n, ns, P(y32), n, P(y64), n), "convert")
        return y64

    Y = apply_inv32(X)
    R = ctx.empty(n, ns)
    for _ in range(iters):
        # R = X - S Y - N Y   (FP64)
        ctx.check(ctx.lib.cfmi_axpby(ctx.h, n, ns, 1.0, P(X), n, 0.0, P(R), n), "axpby")
        engine.gemm(ctx, False, False, n, ns, n, -1.0, cov.S, 0, n, Y, 0, n, 1.0, R, 0, n)
        engine.gemm(ctx, False, False, n, ns, n, -1.0, cov.N, 0, n, Y, 0, n, 1.0, R, 0, n)
        dY = apply_inv32(R)
        ctx.check(ctx.lib.cfmi_axpby(ctx.h, n, ns, 1.0, P(dY), n, 1.0, P(Y), n), "axpby")
    return Y


def noise_weighted_inverse(cov: Covariance, precision: str = "fp64") -> torch.Tensor:
    """G = C^-1 N C^-1 (n x n) in the arithmetic type of ``precision``."""
    ctx, n = cov.ctx, cov.n
    T = _dtype(precision)
    W = whitener(cov, precision)
    Ci = torch.empty((n, n), dtype=T, device=ctx.device)
    engine.gemm(ctx, True, False, n, n, n, 1.0, W, 0, n, W, 0, n, 0.0, Ci, 0, n, 2)
    if T == torch.float64 and cov.noise_is_diagonal():
        # G = (C^-1 N) C^-1 with N diagonal: scale the columns of a copy of
        # C^-1 by diag(N), then one product, lower tiles only (G symmetric).
        d = torch.empty(n, dtype=T, device=ctx.device)
        ctx.check(ctx.lib.cfmi_copy(ctx.h, engine.F64, 1, n, P(cov.N), n + 1, P(d), 1), "copy")
        T1 = torch.empty_like(Ci)
        ctx.check(ctx.lib.cfmi_copy(ctx.h, engine.F64, n, n, P(Ci), n, P(T1), n), "copy")
        ctx.check(ctx.lib.cfmi_scale(ctx.h, 0, n, n, P(T1), n, P(d)), "scale")
        G = torch.empty_like(Ci)
        engine.gemm(ctx, False, False, n, n, n, 1.0, T1, 0, n, Ci, 0, n, 0.0, G, 0, n, 64)
        return G
    if T == torch.float64:
        Nt = cov.N
    else:
        Nt = torch.empty((n, n), dtype=T, device=ctx.device)
        ctx.check(ctx.lib.cfmi_convert(ctx.h, 0, n, n, P(cov.N), n, P(Nt), n), "convert")
    T1 = torch.empty_like(Ci)
    engine.gemm(ctx, False, False, n, n, n, 1.0, Ci, 0, n, Nt, 0, n, 0.0, T1, 0, n)
    del Nt
    G = torch.empty_like(Ci)
    engine.gemm(ctx, False, False, n, n, n, 1.0, T1, 0, n, Ci, 0, n, 0.0, G, 0, n)
    return G


def qml(cov: Covariance, src: DerivativeSource, maps: np.ndarray, remove_nb: bool = True,
        precision: str = "fp64", chunk_budget: int | None = None):
    """(q, nb): raw QML estimates (nsims, n_p) and noise bias (n_p,).

    q_b + nb_b = 1/2 y^T dC_b y with y = C^-1 x, and nb_b = 1/2 Tr[G dC_b]
    with G = C^-1 N C^-1 (``spectra.py:1124-1186``).  With dC_b = e_a D_b e_b^T
    (+ transpose) both are <M_ab, D_b> (x2 when a != b) for M = y y^T and G:
    one HBM-bound FP64 pass per spectrum over the flattened D_b
    (``cfmi_dot_batched``).

    The contractions sum n^2 terms into results about n times smaller than
    the terms, so FP32 rounding of G or D would leave ~n * 6e-8 relative
    error (3e-5 measured at n = 768): they run in FP64 in both precisions.
    ``precision`` selects how y is solved (``solve``)."""
    ctx, n = cov.ctx, cov.n
    ns = maps.shape[1]
    Y = solve(cov, maps, precision)
    if src.dtype != torch.float64:
        src = make_source(cov, src.bins, src.convention, torch.float64)
    G = noise_weighted_inverse(cov, "fp64")
    nb_bins = len(src.bins)
    npar = len(src.geometry) * nb_bins
    out = np.zeros((ns + 1, npar))
    if chunk_budget is None:
        chunk_budget = int(_free_bytes(ctx) * 0.25)
    cache = {}
    for s, b0, blk in src.chunks(chunk_budget):
        D = blk.stack
        k = D.shape[0]
        key = (blk.ra, blk.ma, blk.rb, blk.mb)
        if key not in cache:
            cache.clear()
            ma, mb = blk.ma, blk.mb
            M = torch.empty((ns + 1, mb, ma), dtype=torch.float64, device=ctx.device)
            # G[a, b] block, column-major ma x mb
            ctx.check(ctx.lib.cfmi_copy(ctx.h, engine.F64, ma, mb, P(G, blk.ra + blk.rb * n), n, P(M), ma), "copy")
            for q in range(ns):
                # y_a y_b^T (K = 1 product)
                engine.gemm_ex(ctx, False, True, ma, mb, 1, 1.0, Y, blk.ra + q * n, ma, Y, blk.rb + q * n, mb,
                               0.0, M, (q + 1) * ma * mb, ma)
            cache[key] = M
        M = cache[key]
        R = torch.empty((ns + 1, k), dtype=torch.float64, device=ctx.device)
        K = blk.ma * blk.mb
        ctx.check(ctx.lib.cfmi_dot_batched(ctx.h, K, k, P(D), K, ns + 1, P(M), K, P(R), k), "dot_batched")
        mult = 0.5 * (1.0 if blk.sym else 2.0)
        out[:, s * nb_bins + b0: s * nb_bins + b0 + k] = mult * R.cpu().numpy()
        del blk, D
    del cache, G
    nb = out[0]
    q = out[1:]
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
