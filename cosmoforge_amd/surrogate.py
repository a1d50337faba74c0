"""Learned surrogate of the exact pixel log-likelihood (no reference
counterpart).

An ensemble of E MLPs (3 -> 256 -> 256 -> 256 -> 1, SiLU, BF16 weights, FP32
accumulation) maps parameters theta = (tau, r, A) to -2 ln L.  It is trained
on the device by the hand-written kernels of ``csrc/mlp.hip``
(``cfmi_mlp_forward`` / ``cfmi_mlp_backward`` / ``cfmi_mlp_adam``) on exact
evaluations produced by ``batched.BatchedLikelihood``.  Screening: every
candidate whose ensemble spread exceeds ``spread_threshold`` (-2 ln L units)
falls back to the exact likelihood, and so does every candidate predicted
within ``keep_delta`` of the best prediction; the rest are rejected on the
surrogate's word.  ``screen`` reports what it did.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import engine
from .engine import P


def _init_params(rng: np.random.Generator, P_: int) -> np.ndarray:
    """U(-1/sqrt(fan_in), 1/sqrt(fan_in)) per layer, in the kernels' layout."""
    H, NIN = 256, 3
    out = np.zeros(P_)
    off = 0
    for fo, fi in ((H, NIN), (H, H), (H, H), (1, H)):
        b = 1.0 / np.sqrt(fi)
        out[off:off + fo * fi] = rng.uniform(-b, b, fo * fi)
        off += fo * fi
        out[off:off + fo] = rng.uniform(-b, b, fo)
        off += fo
    assert off == P_
    return out


class EnsembleMLP:
    """E members trained together (one launch per layer for all members)."""

    def __init__(self, ctx: engine.Context, members: int = 4, seed: int = 0, bounds=None):
        self.ctx = ctx
        self.E = members
        self.P = int(ctx.lib.cfmi_mlp_param_count())
        rng = np.random.default_rng(seed)
        w = np.stack([_init_params(np.random.default_rng(rng.integers(2**62)), self.P) for _ in range(members)])
        dev = ctx.device
        self.w32 = torch.tensor(w, dtype=torch.float32, device=dev).contiguous()
        self.wb = torch.empty((members, self.P), dtype=torch.int16, device=dev)
        self.m = torch.zeros_like(self.w32)
        self.v = torch.zeros_like(self.w32)
        self.grad = torch.zeros_like(self.w32)
        ctx.check(ctx.lib.cfmi_mlp_to_bf16(ctx.h, self.w32.numel(), P(self.w32), P(self.wb)), "mlp_to_bf16")
        self.step = 0
        self.bounds = np.asarray(bounds if bounds is not None else [[0.0, 1.0]] * 3, dtype=np.float64)
        self.t_mean, self.t_std = 0.0, 1.0
        self._ws = {}
        self.history = []

    # -- normalisation --------------------------------------------------------
    def _xn(self, theta: np.ndarray) -> np.ndarray:
        lo, hi = self.bounds[:, 0], self.bounds[:, 1]
        return (2.0 * (np.asarray(theta, dtype=np.float64) - lo) / (hi - lo) - 1.0).astype(np.float32)

    def _workspace(self, B: int) -> torch.Tensor:
        if B not in self._ws:
            nbytes = int(self.ctx.lib.cfmi_mlp_workspace_bytes(self.E, B))
            self._ws = {B: torch.empty(nbytes, dtype=torch.uint8, device=self.ctx.device)}
        return self._ws[B]

    # -- raw kernels ------------------------------------------------------------
    def forward_raw(self, Xd: torch.Tensor, idx: torch.Tensor | None, B: int) -> torch.Tensor:
        y = torch.empty((self.E, B), dtype=torch.float32, device=self.ctx.device)
        ws = self._workspace(B)
        self.ctx.check(self.ctx.lib.cfmi_mlp_forward(self.ctx.h, self.E, B, P(Xd), P(idx), P(self.w32), P(self.wb),
                                                     P(y), P(ws)), "mlp_forward")
        return y

    def backward_raw(self, Xd, idx, B: int, dy: torch.Tensor, chunk: int = 256):
        ws = self._workspace(B)
        nsplit = (B + chunk - 1) // chunk
        part = torch.empty((self.E, nsplit, 256, 256), dtype=torch.float32, device=self.ctx.device)
        self.ctx.check(self.ctx.lib.cfmi_mlp_backward(self.ctx.h, self.E, B, P(Xd), P(idx), P(self.wb), P(dy),
                                                      P(ws), P(self.grad), P(part), chunk), "mlp_backward")
        return self.grad

    # -- training ---------------------------------------------------------------
    def fit(self, theta: np.ndarray, target: np.ndarray, epochs: int = 200, batch: int = 1024, lr: float = 2e-3,
            lr_final: float = 1e-4, seed: int = 1, log=None):
        """Adam on the mean squared error of the standardised target; each
        member sees its own shuffles of the data (full batches only)."""
        ctx = self.ctx
        N = len(target)
        if N < batch:
            batch = max(32, (N // 32) * 32)
        self.t_mean = float(np.mean(target))
        self.t_std = float(np.std(target)) or 1.0
        Xd = torch.tensor(self._xn(theta), device=ctx.device).contiguous()
        Td = torch.tensor(((np.asarray(target) - self.t_mean) / self.t_std).astype(np.float32),
                          device=ctx.device).contiguous()
        rng = np.random.default_rng(seed)
        steps_per_epoch = N // batch
        total = epochs * steps_per_epoch
        dy = torch.empty((self.E, batch), dtype=torch.float32, device=ctx.device)
        loss = torch.empty(self.E, dtype=torch.float32, device=ctx.device)
        k = 0
        for ep in range(epochs):
            perms = np.stack([rng.permutation(N) for _ in range(self.E)]).astype(np.int32)
            perms_d = torch.tensor(perms, device=ctx.device)
            for s in range(steps_per_epoch):
                idx = perms_d[:, s * batch:(s + 1) * batch].contiguous()
                y = self.forward_raw(Xd, idx, batch)
                ctx.check(ctx.lib.cfmi_mlp_loss(ctx.h, self.E, batch, P(y), P(Td), P(idx), P(dy), P(loss)),
                          "mlp_loss")
                self.backward_raw(Xd, idx, batch, dy)
                self.step += 1
                frac = k / max(total - 1, 1)
                lr_t = lr * (lr_final / lr) ** frac          # geometric decay
                ctx.check(ctx.lib.cfmi_mlp_adam(ctx.h, self.w32.numel(), P(self.w32), P(self.grad), P(self.m),
                                                P(self.v), lr_t, 0.9, 0.999, 1e-8, self.step, P(self.wb)), "mlp_adam")
                k += 1
            if log and (ep % max(1, epochs // 10) == 0 or ep == epochs - 1):
                lv = loss.cpu().numpy()
                self.history.append((ep, lv.copy()))
                log(f"epoch {ep}: mse (standardised) per member {np.array2string(lv, precision=5)}")
        return self

    # -- inference ----------------------------------------------------------------
    def predict(self, theta: np.ndarray, return_members: bool = False):
        """(mean, spread) of -2 ln L over the members (spread: population
        standard deviation), for theta (N, 3)."""
        ctx = self.ctx
        theta = np.asarray(theta, dtype=np.float64)
        B = theta.shape[0]
        Xd = torch.tensor(self._xn(theta), device=ctx.device).contiguous()
        y = self.forward_raw(Xd, None, B)
        mean = torch.empty(B, dtype=torch.float32, device=ctx.device)
        sd = torch.empty(B, dtype=torch.float32, device=ctx.device)
        ctx.check(ctx.lib.cfmi_mlp_ensemble_stats(ctx.h, self.E, B, P(y), P(mean), P(sd)), "mlp_stats")
        mu = mean.cpu().numpy().astype(np.float64) * self.t_std + self.t_mean
        sg = sd.cpu().numpy().astype(np.float64) * self.t_std
        if return_members:
            return mu, sg, y.cpu().numpy().astype(np.float64) * self.t_std + self.t_mean
        return mu, sg


@dataclass
class ScreenReport:
    n: int
    exact: np.ndarray          # indices evaluated exactly
    fallback: np.ndarray       # of which: spread above threshold
    rejected: np.ndarray       # indices decided by the surrogate alone
    value: np.ndarray          # -2 ln L (exact where evaluated, surrogate elsewhere)
    spread: np.ndarray


def screen(model: EnsembleMLP, theta: np.ndarray, exact_fn, spread_threshold: float = 0.5,
           keep_delta: float = 25.0) -> ScreenReport:
    """Surrogate-screened evaluation of -2 ln L at the candidates theta.

    exact: spread > spread_threshold, or predicted within keep_delta of the
    best confident prediction; the exact values come from ``exact_fn``."""
    mu, sg = model.predict(theta)
    confident = sg <= spread_threshold
    best = np.min(mu[confident]) if np.any(confident) else np.min(mu)
    fallback = np.nonzero(~confident)[0]
    near = np.nonzero(confident & (mu - best <= keep_delta))[0]
    exact = np.union1d(fallback, near)
    value = mu.copy()
    if exact.size:
        value[exact] = exact_fn(theta[exact])
    rejected = np.setdiff1d(np.arange(len(theta)), exact)
    return ScreenReport(len(theta), exact, fallback, rejected, value, sg)
