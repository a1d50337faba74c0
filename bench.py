#!/usr/bin/env python3
"""Benchmark: Fisher + QML time-to-solution on one MI355X, mixed vs FP64.

Default workload: nside 32 TQU, lmax 95, |b| > 20 deg (n = 24192 rows, 564
parameters), beam 4 deg, white noise 0.5 / 1 uK^2, one seeded map of S + N.
One step = one complete Fisher + QML solve with the inputs resident in HBM:
pointings, S, C = S + N and its Cholesky factor, the harmonic-mode
contraction of the traces (F from M = V^T C^-1 V, q from y = C^-1 x, the
noise bias from C^-1 V), and the F^-1 normalisation.  The mixed precision
runs the large products on the int8 matrix cores with sliced operands
(exact integer sums): the Cholesky's trailing updates, the blocked solves
for C^-1 V (updates and 2048-row diagonal products) and M; the Cholesky
panels, y = C^-1 x (refined once against S + N) and the recurrences stay
FP64.  The FP64 precision runs the same algorithm with FP64 MFMA products
(blocked solves with 512-row diagonal blocks, its fastest measured setting).
Both are timed over K steps after W warm-ups; ``value`` is the mixed
seconds per solve.

Outside the timed loops: the pixel-space route of the reference's
formulation (per-parameter whitened derivatives, FP64) once, as the accuracy
anchor, and its mixed variant once; the batched PICSLike likelihood of the
disc workload; the learned surrogate (trained on exact evaluations of that
likelihood, then screening); and the CPU oracle on a bounded sample of the
same workload, extrapolated (labelled as such).

    python bench.py [--steps K] [--warmup W] [--config 0|1|2|3|4] [--no-pixel] [--no-picslike]
                    [--no-surrogate] [--no-cpu]
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Fisher+QML time-to-solution & picslike evals/s, mixed vs FP64 at stated error"
INT8_MFMA_PEAK_TOPS = 5000.0       # MI355X dense int8 MFMA
FP64_MFMA_PEAK_TFLOPS = 78.6
# HBM bytes per gemm_s8 launch of the config-2 mixed solve (mean over its 150
# launches), from rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE over one
# solve, each scaled by what the counter reads for a plain 1 GiB device copy
# (FETCH x 2.0, WRITE x 1.0); not measured live.  Keyed by config.
S8_TRAFFIC_BYTES_PER_LAUNCH = {2: 4.31e9}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_map(ctx, cfg, seed):
    """x = C^(1/2) z, z ~ N(0, I) seeded, through the device factorisation."""
    from cosmoforge_amd import engine, pipeline
    from cosmoforge_amd.core import Inputs

    inp = Inputs(cfg.params, cfg.mask, cfg.noise, cfg.cls, cfg.beam)
    cov = pipeline.Covariance(ctx, inp)
    f = cov.build(need_whitener=False)
    n = inp.n
    z = np.random.default_rng(seed).standard_normal((n, 1))
    dz = ctx.upload_cm(z)
    x = ctx.empty(n, 1)
    engine.gemm(ctx, False, False, n, 1, n, 1.0, f.L, 0, n, dz, 0, n, 0.0, x, 0, n, 1)
    return engine.to_numpy_cm(x) / f.d.cpu().numpy()[:, None]


def cpu_baseline(cfg, stride: int, k_params: int = 2):
    """The oracle in the reference formulation (FP64, LAPACK + OpenMP C
    recurrences) on the host cores, on a bounded sample of the workload:
    every ``stride``-th active pixel of the same mask, spectra, beam and
    noise (n_s = n / stride rows), timing S, C^-1, k parameters' dC_b and
    C^-1 dC_b, one trace, and the QML's G = C^-1 N C^-1, y and per-parameter
    contractions; each term is scaled to n by its operation count (pairs
    n^2, dense products n^3, traces n^2) and to every parameter."""
    from threadpoolctl import threadpool_info

    from oracle import kernels as K
    from oracle import linalg as LA
    from oracle import pipeline as OP

    threads = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    p = cfg.params
    mask = np.array(cfg.mask, dtype=np.float64)
    if stride > 1:
        act = np.nonzero(mask[:, 0] > 0.5)[0]
        keep = np.zeros(mask.shape[0], dtype=bool)
        keep[act[::stride]] = True
        mask[~keep] = 0.0
    lay = K.layout_from_mask(mask, p["spins"], p["labels"], p["nside"])
    n_s, n = lay.n, cfg.n
    L = p["lmax_signal"]
    beams = OP.beam_by_label(lay, cfg.beam, L)
    cls_s = OP.smoothed_cls(lay, cfg.cls, beams, L)
    prm = OP.parameter_weights(lay, beams, p["lmin"], p["lmax"], L)
    # the workload's white noise levels (T first, then Q/U rows)
    nd = np.concatenate([np.full(f.rows, cfg.noise_diag[0] if f.spin == 0 else cfg.noise_diag[-1])
                         for f in lay.fields])
    N = np.diag(nd)
    x = np.random.default_rng(0).standard_normal((n_s, 1))
    t0 = time.perf_counter()
    S = K.signal_matrix(lay, cls_s, L)
    t_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    Cinv = LA.matrix_inverse_symm(N + S)
    t_inv = time.perf_counter() - t0
    picks = np.linspace(0, len(prm) - 1, k_params).astype(int)
    A, t_dc, t_prod = [], 0.0, 0.0
    for i in picks:
        fi, fj, mode, _, w = prm[i]
        t0 = time.perf_counter()
        dC = K.binned_derivative(lay, fi, fj, mode, w)
        t1 = time.perf_counter()
        A.append(Cinv @ dC)
        t2 = time.perf_counter()
        t_dc += t1 - t0
        t_prod += t2 - t1
    t_dc /= len(picks)
    t_prod /= len(picks)
    t0 = time.perf_counter()
    K.matrix_trace(A[0], A[-1])
    t_tr = time.perf_counter() - t0
    t0 = time.perf_counter()
    G = Cinv @ (N @ Cinv)
    y = Cinv @ x
    t_g = time.perf_counter() - t0
    t0 = time.perf_counter()
    for i in picks:     # QML per parameter: dC_b y and Tr(G dC_b) (dC_b rebuilt, cache_derivatives=False)
        fi, fj, mode, _, w = prm[i]
        dC = K.binned_derivative(lay, fi, fj, mode, w)
        0.5 * float(y[:, 0] @ (dC @ y[:, 0]))
        K.matrix_trace(G, dC)
    t_qp = (time.perf_counter() - t0) / len(picks)
    del G, y, A
    npar = len(prm)
    r2, r3 = (n / n_s) ** 2, (n / n_s) ** 3
    fisher = t_s * r2 + t_inv * r3 + npar * (t_dc * r2 + t_prod * r3) + npar * (npar + 1) / 2 * t_tr * r2
    qml = t_g * r3 + npar * t_qp * r2
    blas = ",".join(sorted({d.get("internal_api", "?") for d in threadpool_info()})) or "unknown"
    return {
        "value": round(fisher + qml, 1), "unit": "s per Fisher+QML solve (extrapolated)", "cores": threads,
        "kind": "port",
        "sample": (f"oracle on every {stride}-th active pixel (n_s = {n_s} of {n}): S {t_s:.2f}s (x{r2:.0f}), "
                   f"C^-1 {t_inv:.2f}s (x{r3:.0f}), {len(picks)} of {npar} parameters (dC_b {t_dc:.2f}s x{r2:.0f}, "
                   f"C^-1 dC_b {t_prod:.2f}s x{r3:.0f}), one trace {t_tr:.4f}s (x{r2:.0f}, "
                   f"x{npar * (npar + 1) // 2} pairs), QML G and y {t_g:.2f}s (x{r3:.0f}), per-parameter QML "
                   f"{t_qp:.3f}s (x{r2:.0f}, x{npar}); BLAS {blas}"),
    }


def picslike_leg(ctx):
    """Config-3 batched exact likelihood: evals/s of the batched engine the
    autotuner picks (step-synchronous launches or one persistent workgroup
    per matrix; mixed and FP64, eager and graph replay) against the plain
    per-point GPU path (the drop-in PICSLike point by point, FP64)."""
    import torch

    from cosmoforge_amd import batched, pipeline, workloads
    from cosmoforge_amd.core import Inputs

    cfg = workloads.config(3)
    L = cfg.params["lmax_signal"]
    truth = batched.disc_coefficients([batched.TRUTH[0]], [batched.TRUTH[1]], [batched.TRUTH[2]])[0]
    patches = batched.qu_patches(ctx, 16, cfg.extra["masks"], batched.disc_basis(L), 1.0, L, truth,
                                 seed=workloads.SEED0 + 3)
    ns = np.array([p.n for p in patches])
    tau, r = cfg.extra["tau"], cfg.extra["r"]
    job_patch = np.repeat(np.arange(len(patches)), tau.shape[1])
    coef = batched.disc_coefficients(tau.ravel(), r.ravel())
    tun = batched.autotune(ctx, patches, job_patch, coef, "mixed", use_cache=False)
    out = {"workload": cfg.name, "jobs": int(len(job_patch)), "n_min": int(ns.min()), "n_max": int(ns.max()),
           "n_mean": round(float(ns.mean()), 1), "autotune": {"nb": tun["nb"], "wg_per_cu": tun["wg_per_cu"]}}
    vals = {}
    for prec in ("mixed", "fp64"):
        b = batched.BatchedLikelihood(ctx, patches, job_patch, prec, nb=tun["nb"], wg_per_cu=tun["wg_per_cu"])
        vals[prec] = b.minus2lnl(coef)
        for graph in (False, True):
            b.run(coef, graph=graph)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            reps = 3
            for _ in range(reps):
                b.run(coef, graph=graph)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / reps
            out[f"{prec}{'_graph' if graph else ''}_evals_per_s"] = round(b.njobs / dt, 1)
        b.close()
        log(f"[bench] picslike batched {prec} done")
    sel = np.linspace(0, len(job_patch) - 1, 24).astype(int)
    plain = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for j in sel:
        p = job_patch[j]
        mask = np.stack([cfg.extra["masks"][p]] * 2, axis=1)
        inp = Inputs(dict(cfg.params), mask, np.eye(patches[p].n), workloads.qu_cls(L, tau.ravel()[j], r.ravel()[j]),
                     cfg.beam)
        cov = pipeline.Covariance(ctx, inp)
        chi2, logdet = pipeline.likelihood(cov, patches[p].x.cpu().numpy()[:, None], None)
        plain.append(chi2[0] + logdet)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / len(sel)
    out["plain_per_point_evals_per_s"] = round(1.0 / dt, 1)
    out["speedup_batched_mixed_vs_plain"] = round(out["mixed_graph_evals_per_s"] / out["plain_per_point_evals_per_s"], 1)
    out["mixed_vs_fp64_max_abs_m2lnL"] = float(np.max(np.abs(vals["mixed"] - vals["fp64"])))
    out["plain_vs_fp64_batched_max_abs_m2lnL"] = float(np.max(np.abs(np.array(plain) - vals["fp64"][sel])))
    return out, (patches, tun)


def surrogate_leg(ctx, patches, tun, plain_rate, n_train=20000, n_cand=65536):
    """Config 4: 4-member BF16 MLP ensemble trained on exact joint -2 ln L of
    the config-3 patches; 65,536 candidates screened (spread > 0.5 or within
    25 of the best prediction -> exact)."""
    import torch

    from cosmoforge_amd import batched
    from cosmoforge_amd.surrogate import EnsembleMLP, screen

    like = batched.JointLikelihood(ctx, patches, "mixed", nb=tun["nb"], wg_per_cu=tun["wg_per_cu"])
    train = batched.draw_points(n_train, seed=20261207)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    target = like(train, graph=True)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    log(f"[bench] surrogate training set: {n_train} exact joint evaluations in {t_data:.1f} s")
    model = EnsembleMLP(ctx, members=4, seed=7, bounds=batched.BOUNDS)
    t0 = time.perf_counter()
    model.fit(train, target, epochs=120, batch=512, lr=3e-3, lr_final=5e-5)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    log(f"[bench] surrogate trained in {t_train:.1f} s")
    cand = batched.draw_points(n_cand, seed=20261208)
    model.predict(cand[:1024])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mu, sg = model.predict(cand)
    torch.cuda.synchronize()
    t_inf = time.perf_counter() - t0
    like.evaluations = 0
    t0 = time.perf_counter()
    rep = screen(model, cand, lambda th: like(th, graph=True), spread_threshold=0.5, keep_delta=25.0)
    torch.cuda.synchronize()
    t_screen = time.perf_counter() - t0
    # the surrogate's error, measured with the exact likelihood on a sample of the rejected candidates
    rng = np.random.default_rng(3)
    rej = rep.rejected
    chk = rng.choice(rej, size=min(2048, rej.size), replace=False) if rej.size else rej
    exact_chk = like(cand[chk], graph=True) if chk.size else np.zeros(0)
    ex = rep.exact
    like.close()
    return {"train_points": n_train, "train_data_s": round(t_data, 2), "train_s": round(t_train, 2),
            "candidates": n_cand, "inference_s": round(t_inf, 4), "screen_s": round(t_screen, 2),
            "screened_points_per_s": round(n_cand / t_screen, 1),
            "exhaustive_batched_points_per_s": round(n_train / t_data, 1),
            "plain_per_point_points_per_s": round(plain_rate / len(patches), 2),
            "speedup_screened_vs_plain_per_point": round(n_cand / t_screen / (plain_rate / len(patches)), 1),
            "exact_fraction": round(len(ex) / n_cand, 4), "fallback_by_spread": int(len(rep.fallback)),
            "rejected": int(rej.size),
            "surrogate_error_on_rejected_sample": {
                "n": int(chk.size),
                "max_abs": float(np.max(np.abs(mu[chk] - exact_chk))) if chk.size else None,
                "rms": float(np.sqrt(np.mean((mu[chk] - exact_chk) ** 2))) if chk.size else None,
                "min_exact_minus_best_kept": float(np.min(exact_chk) - np.min(rep.value[ex])) if chk.size else None}}


def stages_leg(ctx, cfg, reps):
    """S (pair recurrences), the FP64 Cholesky and the sliced one (5 int8
    slices in the trailing updates) of the workload's C, ``reps`` times
    each, with the FP64 VALU flops of S and n^3/3 per factorisation."""
    import ctypes

    import torch

    from cosmoforge_amd import engine, pipeline
    from cosmoforge_amd.core import Inputs

    inp = Inputs(cfg.params, cfg.mask, cfg.noise, cfg.cls, cfg.beam)
    cov = pipeline.Covariance(ctx, inp)
    n = inp.n
    out = {"workload": cfg.name, "n": n, "reps": reps}
    ts = []
    for _ in range(reps):
        cov.S = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        S = cov.signal()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    out["signal_ms"] = round(1e3 * min(ts), 2)
    for name, s in (("potrf_fp64", 0), ("potrf_s8", 5)):
        ts = []
        for _ in range(reps):
            C = ctx.empty(n, n)
            ctx.check(ctx.lib.cfmi_axpby(ctx.h, n, n, 1.0, engine.P(S), n, 0.0, engine.P(C), n), "axpby")
            ctx.check(ctx.lib.cfmi_axpby(ctx.h, n, n, 1.0, engine.P(cov.N), n, 1.0, engine.P(C), n), "axpby")
            d = torch.empty(n, dtype=torch.float64, device=ctx.device)
            ctx.check(ctx.lib.cfmi_jacobi_scale(ctx.h, n, engine.P(C), n, engine.P(d)), "jacobi_scale")
            info = ctypes.c_int(0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if s:
                rc = ctx.lib.cfmi_potrf_s8(ctx.h, s, n, engine.P(C), n, None, n, ctypes.byref(info))
            else:
                rc = ctx.lib.cfmi_potrf(ctx.h, engine.F64, n, engine.P(C), n, None, n, ctypes.byref(info))
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
            ctx.check(rc, name)
            del C, d
            torch.cuda.empty_cache()
        t = min(ts)
        out[name + "_ms"] = round(1e3 * t, 2)
        out[name + "_tflops_n3_over_3"] = round(n ** 3 / 3 / t / 1e12, 2)
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2, choices=[0, 1, 2, 3, 4],
                    help="0-2: Fisher+QML workloads; 3: batched likelihood; 4: surrogate-screened likelihood")
    ap.add_argument("--no-pixel", action="store_true", help="skip the pixel-route anchor solves")
    ap.add_argument("--no-picslike", action="store_true", help="skip the batched-likelihood leg")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--cpu-stride", type=int, default=0, help="CPU sample: every k-th active pixel (0: auto)")
    ap.add_argument("--stages", action="store_true",
                    help="time the covariance stages alone (S, the FP64 and the sliced Cholesky) and exit; "
                         "the command the kernel profiles of those stages run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import torch

    from cosmoforge_amd import Fisher, Spectra, engine, workloads
    from cosmoforge_amd.accuracy import errors

    ctx = engine.Context(0)

    if args.stages:
        print(json.dumps(stages_leg(ctx, workloads.config(args.config), args.steps)))
        return

    if args.config in (3, 4):
        pl, (patches, tun) = picslike_leg(ctx)
        out = {"metric": METRIC, "n_gpus": 1, "steps": 3, "warmup": 1, "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "data": "synthetic (seeded discs, maps and parameter draws)"}
        if args.config == 3:
            out.update({"value": pl["mixed_graph_evals_per_s"], "unit": "likelihood evaluations/s",
                        "ms_per_step": round(1e3 * pl["jobs"] / pl["mixed_graph_evals_per_s"], 2),
                        "dtype": "f32 MFMA (FP64 diagonal blocks, panels, ln|C|, refinement)",
                        "config": {"workload": pl["workload"], "jobs": pl["jobs"]}, "picslike": pl})
        else:
            sl = surrogate_leg(ctx, patches, tun, pl["plain_per_point_evals_per_s"])
            out.update({"value": sl["screened_points_per_s"], "unit": "parameter points/s (joint over 64 patches)",
                        "ms_per_step": round(1e3 * sl["screen_s"], 2), "dtype": "bf16 MFMA surrogate + f32 MFMA exact",
                        "config": {"workload": "nside16_QU_64discs_surrogate", "candidates": sl["candidates"]},
                        "surrogate": sl, "picslike": pl})
        print(json.dumps(out))
        return

    cfg = workloads.config(args.config)
    maps = make_map(ctx, cfg, workloads.SEED0 + args.config)
    log(f"[bench] config {args.config} ({cfg.name}): n={cfg.n}")
    # Inputs resident in HBM before the timed region: the (n, n) noise
    # covariance is uploaded once (a host array would add n^2 8 B of PCIe
    # per solve).
    noise_dev = torch.as_tensor(cfg.noise).to(ctx.device)

    def solve(prec, basis="harmonic", solver="trsm"):
        f = Fisher(cfg.params, basis=basis, mask=cfg.mask, noise_cov1=noise_dev, cls_data=cfg.cls, beam=cfg.beam,
                   context=ctx, precision=prec, harmonic_solver=solver)
        f.run()
        s = Spectra(fisher=f, maps1=maps)
        s.run()
        out = (f.get_fisher_matrix(), s.get_power_spectra(), s.qml_noise_bias, s.qml_results, f.get_error_bars())
        if basis == "harmonic":
            out = out + (f.derivs.nt, f.derivs.flops)
        f.release()
        del f, s
        return out

    def timed(prec, record=False):
        for w in range(args.warmup):
            solve(prec)
            log(f"[bench] {prec} warmup {w + 1}/{args.warmup} done")
        torch.cuda.synchronize()
        rec = engine.S8Recorder(ctx) if record else None
        t0 = time.perf_counter()
        for k in range(args.steps):
            res = solve(prec)
            log(f"[bench] {prec} step {k + 1}/{args.steps} done")
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        return dt, res, rec

    t_mixed, res_m, rec = timed("mixed", record=True)
    t_fp64, res_f, _ = timed("fp64")
    n_modes, flops = res_m[5], res_m[6]
    fl, ms, launches, prod = rec.total()
    rec.close()
    achieved = fl / (ms * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": round(t_mixed, 4),
        "unit": "s per Fisher+QML solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_mixed * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8 MFMA on sliced FP64 operands (exact int32 sums) for the Cholesky trailing updates, C^-1 V "
                 "and V^T C^-1 V; FP64 Cholesky panels, recurrences, y = C^-1 x",
        "data": "synthetic (seeded map of S+N; reference-recipe spectra, mask, beam, noise)",
        "config": {"workload": cfg.name, "n_rows": cfg.n, "n_params": int(res_m[0].shape[0]),
                   "lmax_signal": cfg.params["lmax_signal"], "nsims": 1, "n_modes": int(n_modes),
                   "route": "harmonic-mode contraction (basis='harmonic'), blocked triangular solves",
                   "precision": "mixed"},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": INT8_MFMA_PEAK_TOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / INT8_MFMA_PEAK_TOPS, 4),
                     "traffic": S8_TRAFFIC_BYTES_PER_LAUNCH.get(args.config),
                     "kernel": "gemm_s8_kernel (FP64-accurate products on v_mfma_i32_16x16x64_i8: sliced operands, "
                               "s(s+1)/2 int8 products per multiply-add, exact int32 sums)",
                     "launches_per_step": launches / args.steps, "flops_per_launch": fl / max(launches, 1),
                     "ms_per_launch": round(ms / max(launches, 1), 3),
                     "int8_products_per_flop": round(prod, 2),
                     "mfma_issue_frac": round(prod * achieved / INT8_MFMA_PEAK_TOPS, 4),
                     "vs_fp64_mfma_peak": round(achieved / FP64_MFMA_PEAK_TFLOPS, 3)},
        "mixed_s": round(t_mixed, 4),
        "fp64_s": round(t_fp64, 4),
        "speedup_mixed_vs_fp64": round(t_fp64 / t_mixed, 3),
        "mixed_error_vs_fp64": errors(res_m, res_f),
        "harmonic_gemm_flops_per_solve": flops,
    }
    log("[bench] harmonic legs done")
    # hand the cached device memory back before each leg (the library's own
    # allocations, e.g. the batched likelihood's workspace, do not see it)
    torch.cuda.empty_cache()
    if not args.no_pixel:
        # the reference's formulation (pixel-space whitened derivatives), once
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res_p = solve("fp64", basis=False)
        torch.cuda.synchronize()
        t_p = time.perf_counter() - t0
        log(f"[bench] pixel-route FP64 solve done ({t_p:.1f} s)")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res_pm = solve("mixed", basis=False)
        torch.cuda.synchronize()
        t_pm = time.perf_counter() - t0
        log(f"[bench] pixel-route mixed solve done ({t_pm:.1f} s)")
        out["pixel_route"] = {
            "fp64_s": round(t_p, 3), "mixed_s": round(t_pm, 3), "speedup_mixed_vs_fp64": round(t_p / t_pm, 3),
            "mixed_error_vs_pixel_fp64": errors(res_pm, res_p),
            "harmonic_fp64_error_vs_pixel_fp64": errors(res_f, res_p),
            "harmonic_mixed_error_vs_pixel_fp64": errors(res_m, res_p),
            "note": "one solve each, outside the timed loops; the pixel route carries the reference's acos rotation "
                    "angles, so the routes agree to that level"}
        del res_p, res_pm
        torch.cuda.empty_cache()
    if not args.no_picslike:
        pl, (patches, tun) = picslike_leg(ctx)
        out["picslike"] = pl
        if not args.no_surrogate:
            out["surrogate"] = surrogate_leg(ctx, patches, tun, pl["plain_per_point_evals_per_s"])
            log("[bench] surrogate leg done")
    if not args.no_cpu:
        stride = args.cpu_stride or {0: 1, 1: 1, 2: 4}[args.config]
        out["cpu_baseline"] = cpu_baseline(cfg, stride)
        log("[bench] cpu baseline done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
