"""Batched exact pixel likelihood on one MI355X (no reference counterpart:
the reference evaluates ``PICSLike._compute_likelihood_point``,
``picslike/picslike.py:568-664``, one point at a time).

For covariances linear in a few spectrum amplitudes,
C(theta) = sum_k c_k(theta) M_k (S is linear in C_l: ``signal_kernels.py:390-537``),
every patch's component matrices M_k are built once by the FP64 recurrence
kernels, and all (patch, parameter point) jobs are factorised and solved by
the persistent kernels of ``csrc/batched.hip`` (``cfmi_batch_*``), captured
as a hipGraph and replayed for new parameter sets.  The panel width and the
number of persistent workgroups per CU are picked by ``autotune``.

-2 lnL = chi2 + ln|C| per job, as the reference's lnL = -(chi2 + logdet)/2.
"""

from __future__ import annotations

import ctypes
import json
import math
import os
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import engine
from .engine import P

NB_CHOICES = (16, 32, 64)   # 16, 32: one persistent workgroup per matrix; 64: step-synchronous launches
WG_CHOICES = (1, 2, 4)


@dataclass
class Patch:
    """One sky patch: QU rows of the active pixels of a disc."""
    pixels: np.ndarray
    vec: torch.Tensor          # (npix, 3) device
    comps: torch.Tensor        # (ncomp, n, n) device FP64 (each col-major)
    x: torch.Tensor            # (n,) device FP64 data vector

    @property
    def n(self) -> int:
        return int(self.comps.shape[1])


def qu_signal(ctx: engine.Context, vec: torch.Tensor, npix: int, cls: dict, lmax: int) -> torch.Tensor:
    """The 2 npix x 2 npix QU signal block of the (beam-smoothed) spectra."""
    n = 2 * npix
    S = ctx.empty(n, n)
    dev = {k: ctx.upload(np.asarray(cls[k], dtype=np.float64)[: lmax + 1]) for k in ("EE", "BB", "EB")}
    ctx.check(ctx.lib.cfmi_signal_22(ctx.h, P(dev["EE"]), P(dev["BB"]), P(dev["EB"]), lmax, P(vec), npix, P(S), n),
              "signal_22")
    return S


def qu_patches(ctx: engine.Context, nside: int, masks: list, basis_cls: list, noise_var: float, lmax: int,
               truth: np.ndarray, seed: int) -> list:
    """Patches with components [N, S(basis_cls[0]), S(basis_cls[1]), ...]
    and a data vector x = chol(sum_k truth_k M_k) z (seeded, FP64 on the
    device)."""
    rng = np.random.default_rng(seed)
    out = []
    for m in masks:
        pix = np.nonzero(np.asarray(m) > 0.5)[0].astype(np.int64)
        vec = engine.pointings(ctx, nside, pix, nest=False)
        npix = pix.size
        n = 2 * npix
        comps = torch.empty((1 + len(basis_cls), n, n), dtype=torch.float64, device=ctx.device)
        ctx.check(ctx.lib.cfmi_fill(ctx.h, engine.F64, n, n, P(comps[0]), n, 0.0), "fill")
        eye = torch.full((n,), float(noise_var), dtype=torch.float64, device=ctx.device)
        comps[0].view(-1)[:: n + 1].copy_(eye)
        for k, cl in enumerate(basis_cls):
            comps[1 + k].copy_(qu_signal(ctx, vec, npix, cl, lmax))
        C = ctx.empty(n, n)
        ctx.check(ctx.lib.cfmi_fill(ctx.h, engine.F64, n, n, P(C), n, 0.0), "fill")
        for k, t in enumerate(truth):
            ctx.check(ctx.lib.cfmi_axpby(ctx.h, n, n, float(t), P(comps[k]), n, 1.0, P(C), n), "axpby")
        f = engine.factor(ctx, C, n, need_whitener=False)
        z = ctx.upload_cm(rng.standard_normal((n, 1)))
        xs = ctx.empty(n, 1)
        engine.gemm(ctx, False, False, n, 1, n, 1.0, f.L, 0, n, z, 0, n, 0.0, xs, 0, n, 1)
        x = (xs.view(-1) / f.d).contiguous()
        out.append(Patch(pix, vec, comps, x))
        del C, f
    return out


MODES = {"mixed": 1, "mixed_f64_storage": 2, "fp64": 3}


class BatchedLikelihood:
    """chi2 and ln|C| for jobs (patch j_p, coefficients c_j) in one captured
    sequence: ``nb`` = 64 runs the step-synchronous launches (every job's
    step k at once), 16 or 32 one persistent workgroup per matrix.
    ``precision``: "mixed" (factor stored in FP32, FP32 MFMA trailing
    updates; FP64 diagonal blocks, panels, ln|C| and refinement),
    "mixed_f64_storage" (factor stored in FP64, trailing products in FP32
    MFMA) or "fp64"."""

    def __init__(self, ctx: engine.Context, patches: list, job_patch: np.ndarray, precision: str = "mixed",
                 nb: int = 64, wg_per_cu: int = 2, refine_iters: int = 1):
        if precision not in MODES:
            raise ValueError(f"precision must be one of {list(MODES)}; got {precision!r}")
        self.ctx = ctx
        self.patches = patches
        self.precision = precision
        self.nb, self.wg_per_cu, self.refine_iters = nb, wg_per_cu, refine_iters
        ncomp = {int(p.comps.shape[0]) for p in patches}
        if len(ncomp) != 1:
            raise ValueError("every patch must have the same number of component matrices")
        self.ncomp = ncomp.pop()
        ns = np.array([p.n for p in patches], dtype=np.int64)
        # one flat FP64 buffer for all components and all data vectors
        coff = np.zeros(len(patches), dtype=np.int64)
        xoff = np.zeros(len(patches), dtype=np.int64)
        tot_c = 0
        tot_x = 0
        for i, p in enumerate(patches):
            coff[i] = tot_c
            xoff[i] = tot_x
            tot_c += self.ncomp * p.n * p.n
            tot_x += p.n
        self.comps = torch.empty(tot_c, dtype=torch.float64, device=ctx.device)
        self.xs = torch.empty(tot_x, dtype=torch.float64, device=ctx.device)
        for i, p in enumerate(patches):
            self.comps[coff[i]: coff[i] + self.ncomp * p.n * p.n].copy_(p.comps.reshape(-1))
            self.xs[xoff[i]: xoff[i] + p.n].copy_(p.x)
        self.job_patch = np.ascontiguousarray(job_patch, dtype=np.int32)
        self.njobs = int(self.job_patch.size)
        self._ns, self._coff, self._xoff = ns, coff, xoff
        h = ctypes.c_void_p()
        ctx.check(ctx.lib.cfmi_batch_create(ctx.h, MODES[precision], len(patches), ns.ctypes.data, self.ncomp, P(self.comps),
                                            coff.ctypes.data, P(self.xs), xoff.ctypes.data, self.njobs,
                                            self.job_patch.ctypes.data, nb, wg_per_cu, refine_iters,
                                            ctypes.byref(h)), "batch_create")
        self.h = h
        self.coef = torch.zeros((self.njobs, self.ncomp), dtype=torch.float64, device=ctx.device)
        self.chi2 = torch.empty(self.njobs, dtype=torch.float64, device=ctx.device)
        self.logdet = torch.empty(self.njobs, dtype=torch.float64, device=ctx.device)
        self.info = torch.empty(self.njobs, dtype=torch.int32, device=ctx.device)
        self._captured = False

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.cfmi_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _set_coef(self, coef):
        c = torch.as_tensor(np.asarray(coef, dtype=np.float64)).reshape(self.njobs, self.ncomp)
        self.coef.copy_(c)

    def run(self, coef, graph: bool = False):
        """Enqueue every job; returns device (chi2, logdet, info)."""
        self._set_coef(coef)
        lib = self.ctx.lib
        if graph:
            if not self._captured:
                self.ctx.check(lib.cfmi_batch_capture(self.h, P(self.coef), P(self.chi2), P(self.logdet),
                                                      P(self.info)), "batch_capture")
                self._captured = True
            self.ctx.check(lib.cfmi_batch_replay(self.h), "batch_replay")
        else:
            self.ctx.check(lib.cfmi_batch_run(self.h, P(self.coef), P(self.chi2), P(self.logdet), P(self.info)),
                           "batch_run")
        return self.chi2, self.logdet, self.info

    def minus2lnl(self, coef, graph: bool = False) -> np.ndarray:
        chi2, logdet, info = self.run(coef, graph)
        torch.cuda.synchronize()
        bad = info.cpu().numpy()
        if np.any(bad):
            j = int(np.nonzero(bad)[0][0])
            raise engine.LinAlgError(int(bad[j]))
        return (chi2 + logdet).cpu().numpy()


def _tuning_path() -> str:
    base = os.environ.get("XDG_CACHE_HOME", os.path.join(os.path.expanduser("~"), ".cache"))
    return os.environ.get("COSMOFORGE_AMD_TUNING", os.path.join(base, "cosmoforge_amd", "batched_tuning.json"))


def _signature(ns: np.ndarray, precision: str) -> str:
    dev = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "cpu"
    hist = np.histogram(ns, bins=[0, 256, 512, 768, 1024, 1536, 2049])[0]
    return f"{dev}|{precision}|" + ",".join(str(int(v)) for v in hist)


def autotune(ctx: engine.Context, patches: list, job_patch: np.ndarray, coef: np.ndarray, precision: str = "mixed",
             sample: int = 2048, repeats: int = 2, use_cache: bool = True, log=None) -> dict:
    """Pick (nb, wg_per_cu) for this batch shape by timing a job sample that
    keeps the patch-size mix; the choice is cached per device, precision and
    size histogram."""
    ns = np.array([patches[p].n for p in job_patch])
    key = _signature(ns, precision)
    path = _tuning_path()
    cache = {}
    if use_cache and os.path.exists(path):
        try:
            with open(path) as fh:
                cache = json.load(fh)
        except (OSError, ValueError):
            cache = {}
        if key in cache:
            return cache[key]
    rng = np.random.default_rng(0)
    sel = np.sort(rng.choice(len(job_patch), size=min(sample, len(job_patch)), replace=False))
    jp, cf = job_patch[sel], np.asarray(coef)[sel]
    results = []
    for nb in NB_CHOICES:
        for wg in WG_CHOICES:
            b = BatchedLikelihood(ctx, patches, jp, precision, nb=nb, wg_per_cu=wg)
            b.run(cf)
            torch.cuda.synchronize()
            best = math.inf
            for _ in range(repeats):
                t0 = time.perf_counter()
                b.run(cf)
                torch.cuda.synchronize()
                best = min(best, time.perf_counter() - t0)
            b.close()
            results.append({"nb": nb, "wg_per_cu": wg, "s": best})
            if log:
                log(f"autotune nb={nb} wg_per_cu={wg}: {best * 1e3:.1f} ms for {len(sel)} jobs")
    choice = min(results, key=lambda r: r["s"])
    out = {"nb": choice["nb"], "wg_per_cu": choice["wg_per_cu"], "sample_jobs": int(len(sel)),
           "timings": results}
    cache[key] = out
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as fh:
            json.dump(cache, fh, indent=1)
    except OSError:
        pass
    return out


# ---------------------------------------------------------------------------
# The QU disc workload: 64 patches, (tau, r[, A]) points
# ---------------------------------------------------------------------------

def disc_basis(lmax: int):
    """Basis spectra with C(tau, r, A) = N + e^A (S_fix + (tau/0.06)^2 S_bump + r S_BB)."""
    from .workloads import pol_spectra

    ell = np.arange(lmax + 1, dtype=np.float64)
    base = pol_spectra(lmax, tau=0.06)
    z = np.zeros(lmax + 1)
    ee_fix = 0.05 * base["TT"]
    bump = 0.02 * np.exp(-((ell - 4.0) ** 2) / 8.0)
    bump[:2] = 0.0
    bb = np.zeros(lmax + 1)
    bb[2:] = 2 * np.pi * 5e-3 / (ell[2:] * (ell[2:] + 1))
    return [{"EE": ee_fix, "BB": z, "EB": z}, {"EE": bump, "BB": z, "EB": z}, {"EE": z, "BB": bb, "EB": z}]


def disc_coefficients(tau, r, logamp=None) -> np.ndarray:
    tau = np.asarray(tau, dtype=np.float64).ravel()
    r = np.asarray(r, dtype=np.float64).ravel()
    A = np.zeros_like(tau) if logamp is None else np.asarray(logamp, dtype=np.float64).ravel()
    e = np.exp(A)
    return np.stack([np.ones_like(tau), e, e * (tau / 0.06) ** 2, e * r], axis=1)


TRUTH = (0.06, 0.05, 0.0)


class JointLikelihood:
    """-2 ln L(theta) summed over independent patches, theta = (tau, r, A)
    rows; evaluated with ``BatchedLikelihood`` in chunks of at most
    ``max_jobs`` (patch, point) jobs, one plan per chunk size."""

    def __init__(self, ctx: engine.Context, patches: list, precision: str = "mixed", nb: int = 64,
                 wg_per_cu: int = 4, max_jobs: int = 16384):
        self.ctx, self.patches, self.precision = ctx, patches, precision
        self.nb, self.wg_per_cu = nb, wg_per_cu
        self.per_call = max(1, max_jobs // len(patches))
        self._plans = {}
        self.evaluations = 0

    def _plan(self, npts: int) -> BatchedLikelihood:
        if npts not in self._plans:
            if len(self._plans) > 2:
                for b in self._plans.values():
                    b.close()
                self._plans = {}
            jp = np.tile(np.arange(len(self.patches)), npts)
            self._plans[npts] = BatchedLikelihood(self.ctx, self.patches, jp, self.precision, nb=self.nb,
                                                  wg_per_cu=self.wg_per_cu)
        return self._plans[npts]

    def __call__(self, theta: np.ndarray, graph: bool = False) -> np.ndarray:
        theta = np.atleast_2d(np.asarray(theta, dtype=np.float64))
        out = np.empty(len(theta))
        npatch = len(self.patches)
        for s in range(0, len(theta), self.per_call):
            th = theta[s:s + self.per_call]
            coef = disc_coefficients(np.repeat(th[:, 0], npatch), np.repeat(th[:, 1], npatch),
                                     np.repeat(th[:, 2], npatch))
            v = self._plan(len(th)).minus2lnl(coef, graph=graph)
            out[s:s + len(th)] = v.reshape(len(th), npatch).sum(axis=1)
        self.evaluations += len(theta)
        return out

    def close(self):
        for b in self._plans.values():
            b.close()
        self._plans = {}


BOUNDS = np.array([[0.03, 0.09], [0.0, 0.1], [-0.1, 0.1]])


def draw_points(n: int, seed: int, bounds=None) -> np.ndarray:
    """Uniform (tau, r, A) inside ``bounds`` (default BOUNDS)."""
    b = BOUNDS if bounds is None else np.asarray(bounds, dtype=np.float64)
    rng = np.random.default_rng(seed)
    return rng.uniform(b[:, 0], b[:, 1], (n, 3))
