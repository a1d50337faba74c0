"""Surrogate MLP kernels (GPU) against the NumPy oracle with the same BF16
roundings, Adam against its closed form, training, and the screening gain
over exhaustive exact evaluation."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from cosmoforge_amd import engine

    return engine.Context(0)


def _model(ctx, E=2, seed=4):
    from cosmoforge_amd.surrogate import EnsembleMLP

    return EnsembleMLP(ctx, members=E, seed=seed, bounds=[[-1, 1]] * 3)


def test_forward_backward_match_oracle(ctx):
    import torch
    from oracle import mlp as M

    m = _model(ctx)
    rng = np.random.default_rng(8)
    B = 96
    X = rng.uniform(-1, 1, (B, 3)).astype(np.float32)
    T = rng.standard_normal(B).astype(np.float32)
    Xd = torch.tensor(X, device=ctx.device)
    Td = torch.tensor(T, device=ctx.device)
    y = m.forward_raw(Xd, None, B)
    dy = torch.empty_like(y)
    loss = torch.empty(m.E, dtype=torch.float32, device=ctx.device)
    ctx.check(ctx.lib.cfmi_mlp_loss(ctx.h, m.E, B, _p(y), _p(Td), None, _p(dy), _p(loss)), "loss")
    g = m.backward_raw(Xd, None, B, dy, chunk=32).cpu().numpy()
    w = m.w32.cpu().numpy().astype(np.float64)
    yg = y.cpu().numpy()
    for e in range(m.E):
        yo, cache = M.forward(w[e], X.astype(np.float64), bf16=True)
        assert np.max(np.abs(yg[e] - yo)) <= 2e-3 * max(1.0, np.max(np.abs(yo)))
        dyo = 2.0 * (yo - T) / B
        go = M.backward(cache, dyo, bf16=True)
        for name, (a, b) in {"W1": (0, 768), "b1": (768, 1024), "W2": (1024, 66560), "b2": (66560, 66816),
                             "W3": (66816, 132352), "b3": (132352, 132608), "W4": (132608, 132864),
                             "b4": (132864, 132865)}.items():
            err = np.linalg.norm(g[e, a:b] - go[a:b]) / np.linalg.norm(go[a:b])
            print(f"member {e} {name}: relative gradient error {err:.2e}")
            assert err < 2e-2, (e, name, err)


def _p(t):
    from cosmoforge_amd.engine import P

    return P(t)


def test_adam_matches_closed_form(ctx):
    import torch

    rng = np.random.default_rng(2)
    n = 1000
    w = rng.standard_normal(n).astype(np.float32)
    wd = torch.tensor(w, device=ctx.device)
    md = torch.zeros_like(wd)
    vd = torch.zeros_like(wd)
    wb = torch.empty(n, dtype=torch.int16, device=ctx.device)
    m = np.zeros(n)
    v = np.zeros(n)
    ww = w.astype(np.float64)
    for step in range(1, 4):
        g = rng.standard_normal(n).astype(np.float32)
        gd = torch.tensor(g, device=ctx.device)
        ctx.check(ctx.lib.cfmi_mlp_adam(ctx.h, n, _p(wd), _p(gd), _p(md), _p(vd), 1e-2, 0.9, 0.999, 1e-8, step,
                                        _p(wb)), "adam")
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g.astype(np.float64) ** 2
        ww = ww - 1e-2 * (m / (1 - 0.9**step)) / (np.sqrt(v / (1 - 0.999**step)) + 1e-8)
    np.testing.assert_allclose(wd.cpu().numpy(), ww, rtol=1e-5, atol=1e-6)
    from oracle.mlp import bf16_round

    bits = wb.cpu().numpy().astype(np.uint16).astype(np.uint32) << 16
    np.testing.assert_array_equal(bits.view(np.float32), bf16_round(wd.cpu().numpy()).astype(np.float32))


def test_training_fits_smooth_function(ctx):
    m = _model(ctx, E=4, seed=11)
    rng = np.random.default_rng(5)
    th = rng.uniform(-1, 1, (4096, 3))
    f = lambda t: 50 * (t[:, 0] - 0.2) ** 2 + 30 * (t[:, 1] + 0.1) ** 2 + 10 * t[:, 0] * t[:, 2] + 5 * t[:, 2]  # noqa
    m.fit(th, f(th), epochs=60, batch=512, lr=3e-3, lr_final=3e-4)
    test = rng.uniform(-0.9, 0.9, (2048, 3))
    mu, sg = m.predict(test)
    err = np.abs(mu - f(test))
    print(f"fit: max abs err {err.max():.3f}, rms {np.sqrt(np.mean(err**2)):.3f}, spread median {np.median(sg):.3f}")
    assert np.sqrt(np.mean(err**2)) < 0.02 * np.std(f(test))


def test_screening_gain_over_exhaustive_exact(ctx):
    """Trained on exact joint -2 ln L of small QU discs, the surrogate screens
    4096 candidates: the exact best point and every candidate within 5 of the
    exact minimum are kept for exact evaluation, with far fewer exact
    evaluations than the exhaustive method needs."""
    import math

    from cosmoforge_amd import batched, workloads
    from cosmoforge_amd.surrogate import EnsembleMLP, screen

    nside, L = 8, 32
    vec = workloads.pixel_vectors(nside)
    rng = np.random.default_rng(31)
    masks = []
    for radius in (50.0, 60.0, 70.0, 80.0, 65.0, 55.0):
        u, ph = rng.uniform(-1, 1), rng.uniform(0, 2 * np.pi)
        c = np.array([math.sqrt(1 - u * u) * math.cos(ph), math.sqrt(1 - u * u) * math.sin(ph), u])
        masks.append((vec @ c >= math.cos(math.radians(radius))).astype(np.float64))
    truth = batched.disc_coefficients([0.06], [0.05], [0.0])[0]
    patches = batched.qu_patches(ctx, nside, masks, batched.disc_basis(L), 1.0, L, truth, seed=9)
    like = batched.JointLikelihood(ctx, patches, "mixed")
    bounds = [[0.03, 0.09], [0.0, 0.1], [-1.0, 1.0]]      # a log-amplitude range the data constrain
    train = batched.draw_points(3000, seed=1, bounds=bounds)
    model = EnsembleMLP(ctx, members=4, seed=3, bounds=bounds)
    model.fit(train, like(train), epochs=400, batch=256, lr=3e-3, lr_final=5e-5)
    cand = batched.draw_points(4096, seed=2, bounds=bounds)
    exact_all = like(cand)
    like.evaluations = 0
    rep = screen(model, cand, like, spread_threshold=0.5, keep_delta=25.0)
    kept = set(rep.exact.tolist())
    good = np.nonzero(exact_all - exact_all.min() <= 5.0)[0]
    print(f"screen: {len(rep.exact)} exact of {rep.n} ({len(rep.fallback)} by spread), "
          f"{len(good)} candidates within 5 of the minimum; exact range {np.ptp(exact_all):.1f}; "
          f"surrogate max |err| {np.max(np.abs(model.predict(cand)[0] - exact_all)):.2f}")
    assert int(np.argmin(exact_all)) in kept
    assert set(good.tolist()) <= kept
    assert len(rep.exact) < 0.5 * rep.n
    np.testing.assert_allclose(rep.value[rep.exact], exact_all[rep.exact], rtol=0, atol=1e-9)
    like.close()
