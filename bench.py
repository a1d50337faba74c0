#!/usr/bin/env python3
"""Benchmark: Fisher + QML time-to-solution on one MI355X, mixed vs FP64.

One step = one complete Fisher + QML solve of the nside-16 TQU, lmax-47
configuration (n = 5952 rows, 276 parameters): pointings, S, C = S + N and
its Cholesky, the 6 derivative stacks, the whitened derivatives, the Fisher
Gram matrix, the QML estimates, noise bias and their F^-1 normalisation.
Each precision (FP64, mixed) is timed over K steps after W warm-ups; the
line reports the mixed-precision time as ``value``, the FP64 time, the
speedup and the mixed-vs-FP64 errors.  Inputs (mask, spectra, beam, noise,
one seeded map) are resident before the timed region.

    python bench.py [--steps K] [--warmup W] [--config 0|1|2] [--precisions fp64,mixed] [--no-cpu]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_MFMA_PEAK_TFLOPS = 78.6
FP32_MFMA_PEAK_TFLOPS = 157.3
BF16_MFMA_PEAK_TFLOPS = 2500.0
# HBM bytes per whitening launch, measured with rocprofv3 --pmc FETCH_SIZE and
# --pmc WRITE_SIZE on this workload (each scaled by what the counter reads for
# a 1 GiB copy: x2.0 and x1.0); not measured live.  Keyed by (config, kernel).
WHITENING_TRAFFIC_BYTES_PER_LAUNCH = {(1, "f64"): 87.5e9, (1, "f32"): 44.0e9, (1, "bf16x3"): 41.8e9}
# The three-piece kernel issues six BF16 MFMA products per FP32-accurate one.
B3_PRODUCTS = 6


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


def cpu_baseline(cfg, maps, k_params: int):
    """The oracle (reference formulation, FP64) on the host cores, on a
    bounded sample: S, C^-1 and k parameters' C^-1 dC_b products and traces,
    extrapolated to every parameter of the configuration."""
    from threadpoolctl import threadpool_info

    from oracle import kernels as K
    from oracle import linalg as LA
    from oracle import pipeline as OP

    threads = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    lay = K.layout_from_mask(cfg.mask, cfg.params["spins"], cfg.params["labels"], cfg.params["nside"])
    L = cfg.params["lmax_signal"]
    beams = OP.beam_by_label(lay, cfg.beam, L)
    cls_s = OP.smoothed_cls(lay, cfg.cls, beams, L)
    prm = OP.parameter_weights(lay, beams, cfg.params["lmin"], cfg.params["lmax"], L)
    N = cfg.noise
    t0 = time.perf_counter()
    S = K.signal_matrix(lay, cls_s, L)
    t_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    Cinv = LA.matrix_inverse_symm(N + S)
    t_inv = time.perf_counter() - t0
    picks = np.linspace(0, len(prm) - 1, k_params).astype(int)
    A = []
    t0 = time.perf_counter()
    for i in picks:
        fi, fj, mode, _, w = prm[i]
        dC = K.binned_derivative(lay, fi, fj, mode, w)
        A.append(Cinv @ dC)
    t_par = (time.perf_counter() - t0) / len(picks)
    t0 = time.perf_counter()
    K.matrix_trace(A[0], A[-1])
    t_tr = time.perf_counter() - t0
    t0 = time.perf_counter()
    G = Cinv @ (N @ Cinv)
    y = Cinv @ maps
    t_g = time.perf_counter() - t0
    del G, y
    npar = len(prm)
    total = t_s + t_inv + npar * t_par + npar * (npar + 1) / 2 * t_tr + t_g + npar * (t_par + t_tr)
    blas = ",".join(sorted({d.get("internal_api", "?") for d in threadpool_info()})) or "unknown"
    return {
        "value": round(total, 2), "unit": "s (extrapolated)", "cores": threads, "kind": "port",
        "sample": (f"oracle S ({t_s:.2f}s) + C^-1 ({t_inv:.2f}s) + {len(picks)} of {npar} parameters "
                   f"(dC_b and C^-1 dC_b {t_par:.2f}s each, one trace {t_tr:.3f}s) + QML G and y "
                   f"({t_g:.2f}s); Fisher extrapolated x{npar}, traces x{npar * (npar + 1) // 2}, "
                   f"QML x{npar}; BLAS {blas}"),
    }


def picslike_leg(ctx, log):
    """Config-3 batched exact likelihood: evals/s of the persistent batched
    kernel (mixed and FP64, eager and graph replay) against the plain
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


def surrogate_leg(ctx, patches, tun, log, n_train=20000, n_cand=65536):
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
    jobs_rate = len(patches) * len(ex) / max(t_screen - t_inf, 1e-9)
    return {"train_points": n_train, "train_data_s": round(t_data, 2), "train_s": round(t_train, 2),
            "candidates": n_cand, "inference_s": round(t_inf, 4), "screen_s": round(t_screen, 2),
            "screened_points_per_s": round(n_cand / t_screen, 1),
            "exhaustive_batched_points_per_s": round(n_train / t_data, 1),
            "exact_fraction": round(len(ex) / n_cand, 4), "fallback_by_spread": int(len(rep.fallback)),
            "rejected": int(rej.size), "patch_jobs_per_s_in_screen": round(jobs_rate, 1),
            "surrogate_error_on_rejected_sample": {
                "n": int(chk.size),
                "max_abs": float(np.max(np.abs(mu[chk] - exact_chk))) if chk.size else None,
                "rms": float(np.sqrt(np.mean((mu[chk] - exact_chk) ** 2))) if chk.size else None,
                "min_exact_minus_best_kept": float(np.min(exact_chk) - np.min(rep.value[ex])) if chk.size else None}}


def rel(a, b) -> float:
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=1, choices=[0, 1, 2, 3, 4],
                    help="0-2: Fisher+QML workloads; 3: batched likelihood; 4: surrogate-screened likelihood")
    ap.add_argument("--no-picslike", action="store_true", help="skip the batched-likelihood leg")
    ap.add_argument("--precisions", default="fp64,mixed", help="comma list of fp64 / mixed")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--no-harmonic", action="store_true", help="skip the harmonic-mode (basis='harmonic') leg")
    ap.add_argument("--cpu-params", type=int, default=2)
    ap.add_argument("--mixed-gemm", default="bf16x3", choices=["bf16x3", "fp32"],
                    help="whitening arithmetic of the mixed precision")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import torch

    from cosmoforge_amd import Fisher, Spectra, engine, pipeline, workloads

    ctx = engine.Context(0)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    if args.config in (3, 4):
        pl, (patches, tun) = picslike_leg(ctx, log)
        out = {"metric": "Fisher+QML time-to-solution & picslike evals/s, mixed vs FP64 at stated error",
               "n_gpus": 1, "steps": 3, "warmup": 1, "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "data": "synthetic (seeded discs, maps and parameter draws)"}
        if args.config == 3:
            out.update({"value": pl["mixed_graph_evals_per_s"], "unit": "likelihood evaluations/s",
                        "ms_per_step": round(1e3 * pl["jobs"] / pl["mixed_graph_evals_per_s"], 2),
                        "dtype": "f32 MFMA (FP64 diagonal blocks, panels, ln|C|, refinement)",
                        "config": {"workload": pl["workload"], "jobs": pl["jobs"]}, "picslike": pl})
        else:
            sl = surrogate_leg(ctx, patches, tun, log)
            out.update({"value": sl["screened_points_per_s"], "unit": "parameter points/s (joint over 64 patches)",
                        "ms_per_step": round(1e3 * sl["screen_s"], 2), "dtype": "bf16 MFMA surrogate + f32 MFMA exact",
                        "config": {"workload": "nside16_QU_64discs_surrogate", "candidates": sl["candidates"]},
                        "surrogate": sl, "picslike": pl,
                        "speedup_screened_vs_plain_per_point": round(
                            sl["screened_points_per_s"] / (pl["plain_per_point_evals_per_s"] / 64.0), 1)})
        print(json.dumps(out))
        return

    cfg = workloads.config(args.config)
    maps = make_map(ctx, cfg, workloads.SEED0 + args.config)
    print(f"[bench] config {args.config} ({cfg.name}): n={cfg.n}", file=sys.stderr, flush=True)

    # Inputs resident in HBM before the timed region: the (n, n) noise
    # covariance is uploaded once (the PCIe-inclusive figure adds ~n^2 8 B
    # per solve, ~35 ms at n = 5952).
    noise_dev = torch.as_tensor(cfg.noise).to(ctx.device)

    def step(prec, timer=None, lt=None, basis=False):
        f = Fisher(cfg.params, basis=basis, mask=cfg.mask, noise_cov1=noise_dev, cls_data=cfg.cls,
                   beam=cfg.beam, context=ctx, timer=timer, launch_timer=lt, precision=prec,
                   mixed_gemm=args.mixed_gemm)
        f.run()
        s = Spectra(fisher=f, maps1=maps)
        s.run()
        out = (f.get_fisher_matrix(), s.get_power_spectra(), s.qml_noise_bias, s.qml_results,
               f.get_error_bars())
        if basis == "harmonic":
            out = out + (f.derivs.nt, f.derivs.flops)
        f.release()
        del f, s
        return out

    runs = {}
    for prec in args.precisions.split(","):
        for w in range(args.warmup):
            step(prec)
            print(f"[bench] {prec} warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        timer = pipeline.StageTimer()
        lt = pipeline.LaunchTimer()
        t0 = time.perf_counter()
        for k in range(args.steps):
            res = step(prec, timer, lt)
            print(f"[bench] {prec} step {k + 1}/{args.steps} done", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        per_step = (time.perf_counter() - t0) / args.steps
        runs[prec] = dict(s=per_step, res=res, timer=timer, lt=lt)

    main_prec = "mixed" if "mixed" in runs else "fp64"
    r = runs[main_prec]
    timer, lt = r["timer"], r["lt"]
    white_ms = lt.total_ms() / args.steps
    launches = len(lt.launches) / args.steps
    achieved = timer.flops / (white_ms * 1e-3) / 1e12
    b3 = main_prec == "mixed" and args.mixed_gemm == "bf16x3"
    dt = "bf16x3" if b3 else ("f32" if main_prec == "mixed" else "f64")
    peak = {"bf16x3": BF16_MFMA_PEAK_TFLOPS, "f32": FP32_MFMA_PEAK_TFLOPS, "f64": FP64_MFMA_PEAK_TFLOPS}[dt]
    kernel = {"bf16x3": "gemm_b3_kernel (three-piece BF16 operands, six BF16 MFMA products per FP32-accurate "
                        "product; whitened derivatives W dC_b W^T, batched over bins)",
              "f32": "gemm_kernel<float> (whitened derivatives W dC_b W^T, batched over bins)",
              "f64": "gemm_kernel<double> (whitened derivatives W dC_b W^T, batched over bins)"}[dt]
    F = r["res"][0]
    out = {
        "metric": "Fisher+QML time-to-solution & picslike evals/s, mixed vs FP64 at stated error",
        "value": round(r["s"], 4),
        "unit": "s per Fisher+QML solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["s"] * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("bf16x3 MFMA at f32 accuracy" if b3 else "f32 MFMA") +
                 " (FP64 recurrences, factorisation, Gram accumulation, QML)" if main_prec == "mixed" else "f64",
        "data": "synthetic (seeded map of S+N; reference-recipe spectra, mask, beam, noise)",
        "config": {"workload": cfg.name, "n_rows": cfg.n, "n_params": int(F.shape[0]),
                   "lmax_signal": cfg.params["lmax_signal"], "nsims": 1, "precision": main_prec,
                   "mixed_gemm": args.mixed_gemm if main_prec == "mixed" else None,
                   "row_bands": len(timer.bands)},
        "stages_ms": {"deriv_stacks": round(timer.elapsed_ms("deriv0", "white0") / args.steps, 2),
                      "whitened_derivatives": round(timer.elapsed_ms("white0", "white1") / args.steps, 2),
                      "fisher_gram": round(timer.elapsed_ms("white1", "gram1") / args.steps, 2)},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4),
                     "traffic": WHITENING_TRAFFIC_BYTES_PER_LAUNCH.get((args.config, dt)),
                     "kernel": kernel,
                     "launches_per_step": launches, "flops_per_launch": timer.flops / max(launches, 1),
                     "ms_per_launch": round(white_ms / max(launches, 1), 3)},
    }
    if b3:
        # achieved counts the FP32-equivalent (algorithmic) flops; the matrix
        # cores execute six BF16 products for each of them.
        out["roofline"]["mfma_products_per_flop"] = B3_PRODUCTS
        out["roofline"]["mfma_issue_frac"] = round(B3_PRODUCTS * achieved / peak, 4)
    for prec, rr in runs.items():
        out[f"{prec}_s"] = round(rr["s"], 4)
    if "fp64" in runs and "mixed" in runs:
        (F, C, nb, q, sig), (Fm, Cm, nbm, qm, _) = runs["fp64"]["res"], runs["mixed"]["res"]
        out["speedup_mixed_vs_fp64"] = round(runs["fp64"]["s"] / runs["mixed"]["s"], 3)
        out["mixed_error_vs_fp64"] = {
            "F_maxrel": rel(Fm, F), "F_diag_rel_max": float(np.max(np.abs(np.diag(Fm) - np.diag(F)) / np.abs(np.diag(F)))),
            "Cl_maxrel": rel(Cm, C), "Cl_over_sigma_max": float(np.max(np.abs(Cm - C) / sig)),
            "noise_bias_maxrel": rel(nbm, nb), "qml_raw_maxrel": rel(qm, q), "cond_F": float(np.linalg.cond(F))}
    if not args.no_harmonic and args.config < 3:
        # the same estimator with its traces contracted over harmonic modes
        # (FP64): M = (W V)^T (W V) and mode reductions instead of the
        # per-parameter whitening products
        step("fp64", basis="harmonic")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            hres = step("fp64", basis="harmonic")
        torch.cuda.synchronize()
        hs = (time.perf_counter() - t0) / args.steps
        print("[bench] harmonic steps done", file=sys.stderr, flush=True)
        Fh, Ch, nbh, qh, _, nmodes, hflops = hres
        hleg = {"s": round(hs, 4), "modes": int(nmodes), "gemm_flops_per_solve": hflops,
                "speedup_vs_pixel": {p: round(rr["s"] / hs, 2) for p, rr in runs.items()}}
        if "fp64" in runs:
            F, C, nb, q, sig = runs["fp64"]["res"]
            hleg["error_vs_pixel_fp64"] = {
                "F_maxrel": rel(Fh, F),
                "F_diag_rel_max": float(np.max(np.abs(np.diag(Fh) - np.diag(F)) / np.abs(np.diag(F)))),
                "Cl_over_sigma_max": float(np.max(np.abs(Ch - C) / sig)), "noise_bias_maxrel": rel(nbh, nb)}
        out["harmonic_fp64"] = hleg
    if not args.no_picslike:
        out["picslike"] = picslike_leg(ctx, log)[0]
    if not args.no_cpu and args.config < 2:
        out["cpu_baseline"] = cpu_baseline(cfg, maps, args.cpu_params)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
