"""cfmi_mlp_backward_noalloc (static scratch, no allocation) against
cfmi_mlp_backward (workspace argument): bit-identical gradients on two calls
in a row, the second with a smaller batch, and a clear error beyond the
scratch size."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _grads(ctx, model, X, idx, B, dy, chunk, noalloc):
    import torch
    from cosmoforge_amd.engine import P

    ws = model._workspace(B)
    grad = torch.full_like(model.w32, float("nan"))
    if noalloc:
        rc = ctx.lib.cfmi_mlp_backward_noalloc(ctx.h, model.E, B, P(X), P(idx), P(model.wb), P(dy), P(ws), P(grad),
                                               chunk)
    else:
        nsplit = (B + chunk - 1) // chunk
        part = torch.empty((model.E, nsplit, 256, 256), dtype=torch.float32, device=ctx.device)
        rc = ctx.lib.cfmi_mlp_backward(ctx.h, model.E, B, P(X), P(idx), P(model.wb), P(dy), P(ws), P(grad),
                                       P(part), chunk)
    ctx.check(rc, "mlp_backward")
    torch.cuda.synchronize()
    return grad.cpu().numpy()


def test_noalloc_backward_matches_workspace_backward():
    import torch
    from cosmoforge_amd import engine
    from cosmoforge_amd.surrogate import EnsembleMLP

    ctx = engine.Context(0)
    model = EnsembleMLP(ctx, members=4, seed=3, bounds=[[-1, 1]] * 3)
    rng = np.random.default_rng(11)
    for B, chunk in ((1024, 256), (160, 32)):       # the second call smaller
        X = torch.tensor(rng.uniform(-1, 1, (B + 7, 3)).astype(np.float32), device=ctx.device)
        idx = torch.tensor(np.stack([rng.permutation(B + 7)[:B] for _ in range(4)]).astype(np.int32),
                           device=ctx.device)
        y = model.forward_raw(X, idx, B)
        dy = torch.tensor(rng.standard_normal((4, B)).astype(np.float32), device=ctx.device)
        g_ws = _grads(ctx, model, X, idx, B, dy, chunk, noalloc=False)
        model.forward_raw(X, idx, B)                  # the same activations again
        g_na = _grads(ctx, model, X, idx, B, dy, chunk, noalloc=True)
        assert np.isfinite(g_ws).all()
        np.testing.assert_array_equal(g_na, g_ws)
        del y
    # beyond the static scratch: 4 members x ceil(8192 / 32) = 1024 > 128 chunks
    Bbig = 8192
    rc = ctx.lib.cfmi_mlp_backward_noalloc(ctx.h, 4, Bbig, None, None, None, None, None, None, 32)
    assert rc < 0 and b"static scratch" in ctx.lib.cfmi_last_error(ctx.h)
