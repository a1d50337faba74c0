"""Batched exact likelihood (``cfmi_batch_*``) against the FP64 oracle (GPU).

Small QU discs at nside 8 (n = 2 n_pix between ~60 and ~400 rows), random
(tau, r, log-amplitude) points.  The FP64 batch must match the oracle's
``picslike_point`` (Jacobi-scaled dpotrf/dpotri chi2 and slogdet) to 1e-10
relative; the mixed batch (FP32 MFMA trailing updates, FP64 diagonal
blocks, ln|C| and refinement) to 1e-3 absolute in -2 lnL.  Graph replays
must reproduce eager runs bit for bit, and a non-positive-definite job must
report its pivot without disturbing the others.
"""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NSIDE, LMAX = 8, 32


@pytest.fixture(scope="module")
def setup():
    from cosmoforge_amd import batched, engine, workloads

    ctx = engine.Context(0)
    vec = workloads.pixel_vectors(NSIDE)
    rng = np.random.default_rng(17)
    masks = []
    for radius in (20.0, 35.0, 55.0, 75.0):
        u, ph = rng.uniform(-1, 1), rng.uniform(0, 2 * np.pi)
        c = np.array([math.sqrt(1 - u * u) * math.cos(ph), math.sqrt(1 - u * u) * math.sin(ph), u])
        masks.append((vec @ c >= math.cos(math.radians(radius))).astype(np.float64))
    basis = batched.disc_basis(LMAX)
    truth = batched.disc_coefficients([0.06], [0.05], [0.0])[0]
    patches = batched.qu_patches(ctx, NSIDE, masks, basis, 1.0, LMAX, truth, seed=5)
    npts = 6
    tau = rng.uniform(0.03, 0.09, (len(patches), npts))
    r = rng.uniform(0.0, 0.1, (len(patches), npts))
    A = rng.uniform(-0.1, 0.1, (len(patches), npts))
    job_patch = np.repeat(np.arange(len(patches)), npts)
    coef = batched.disc_coefficients(tau.ravel(), r.ravel(), A.ravel())
    return ctx, masks, basis, patches, job_patch, coef, (tau.ravel(), r.ravel(), A.ravel())


def _oracle(masks, basis, patches, job_patch, params):
    from oracle import kernels as K
    from oracle import pipeline as OP

    out = []
    tau, r, A = params
    for j, p in enumerate(job_patch):
        lay = K.layout_from_mask(np.stack([masks[p], masks[p]], axis=1), [2], ["E", "B"], NSIDE)
        e = math.exp(A[j])
        cls = {k: e * (basis[0][k] + (tau[j] / 0.06) ** 2 * basis[1][k] + r[j] * basis[2][k])
               for k in ("EE", "BB", "EB")}
        beams = OP.beam_by_label(lay, np.ones((3, LMAX + 1)), LMAX)
        cls_s = OP.smoothed_cls(lay, cls, beams, LMAX)
        n = patches[p].n
        x = patches[p].x.cpu().numpy()
        chi2, logdet, _ = OP.picslike_point(lay, cls_s, LMAX, np.eye(n), x)
        out.append((float(chi2[0]), float(logdet)))
    return np.array(out)


def test_batched_fp64_matches_oracle(setup):
    from cosmoforge_amd import batched

    ctx, masks, basis, patches, job_patch, coef, params = setup
    ref = _oracle(masks, basis, patches, job_patch, params)
    for nb in batched.NB_CHOICES:
        b = batched.BatchedLikelihood(ctx, patches, job_patch, "fp64", nb=nb, wg_per_cu=2)
        chi2, logdet, info = b.run(coef)
        chi2, logdet = chi2.cpu().numpy(), logdet.cpu().numpy()
        assert not np.any(info.cpu().numpy())
        np.testing.assert_allclose(chi2, ref[:, 0], rtol=1e-10)
        np.testing.assert_allclose(logdet, ref[:, 1], rtol=1e-10)
        b.close()


def test_batched_mixed_within_1e3(setup):
    from cosmoforge_amd import batched

    ctx, masks, basis, patches, job_patch, coef, params = setup
    exact = batched.BatchedLikelihood(ctx, patches, job_patch, "fp64").minus2lnl(coef)
    for mode in ("mixed", "mixed_f64_storage"):
        for nb in batched.NB_CHOICES:
            m = batched.BatchedLikelihood(ctx, patches, job_patch, mode, nb=nb).minus2lnl(coef)
            err = np.max(np.abs(m - exact))
            print(f"{mode} nb={nb}: max |d(-2lnL)| = {err:.3e}")
            assert err < 1e-3


@pytest.mark.parametrize("nb", [32, 64])
@pytest.mark.parametrize("iters", [0, 2])
def test_batched_refinement_steps(setup, iters, nb):
    """No refinement and two refinement steps (per-patch residual products
    between the factor and refine kernels): FP64 still at 1e-10 of the
    oracle, mixed within 1e-3 of FP64."""
    from cosmoforge_amd import batched

    ctx, masks, basis, patches, job_patch, coef, params = setup
    ref = _oracle(masks, basis, patches, job_patch, params)
    b = batched.BatchedLikelihood(ctx, patches, job_patch, "fp64", refine_iters=iters, nb=nb)
    chi2, logdet, info = b.run(coef)
    assert not np.any(info.cpu().numpy())
    np.testing.assert_allclose(chi2.cpu().numpy(), ref[:, 0], rtol=1e-10)
    np.testing.assert_allclose(logdet.cpu().numpy(), ref[:, 1], rtol=1e-10)
    b.close()
    exact = ref[:, 0] + ref[:, 1]
    m = batched.BatchedLikelihood(ctx, patches, job_patch, "mixed", refine_iters=iters, nb=nb).minus2lnl(coef)
    err = np.max(np.abs(m - exact))
    print(f"mixed, {iters} refinement steps: max |d(-2lnL)| = {err:.3e}")
    assert err < (1e-3 if iters else 1.0)


@pytest.mark.parametrize("nb", [32, 64])
def test_graph_replay_matches_eager(setup, nb):
    from cosmoforge_amd import batched

    ctx, masks, basis, patches, job_patch, coef, params = setup
    b = batched.BatchedLikelihood(ctx, patches, job_patch, "mixed", refine_iters=2, nb=nb)
    eager1 = b.minus2lnl(coef)
    coef2 = coef.copy()
    coef2[:, 3] *= 0.5
    eager2 = b.minus2lnl(coef2)
    for rep in range(3):
        g1 = b.minus2lnl(coef, graph=True)
        _ = (ctx.upload(np.arange(10.0)) * 2).sum().item()  # eager work between replays
        g2 = b.minus2lnl(coef2, graph=True)
        assert np.array_equal(g1, eager1), rep
        assert np.array_equal(g2, eager2), rep


@pytest.mark.parametrize("nb", [32, 64])
def test_non_positive_definite_job_reported(setup, nb):
    import torch
    from cosmoforge_amd import batched

    ctx, masks, basis, patches, job_patch, coef, params = setup
    bad = coef.copy()
    bad[2] = [-1.0, 0.0, 0.0, 0.0]          # C = -N
    b = batched.BatchedLikelihood(ctx, patches, job_patch, "mixed", nb=nb)
    chi2, logdet, info = b.run(bad)
    torch.cuda.synchronize()
    info = info.cpu().numpy()
    assert info[2] == 1 and np.isnan(chi2[2].item()) and np.isnan(logdet[2].item())
    assert np.count_nonzero(info) == 1
    good = batched.BatchedLikelihood(ctx, patches, job_patch, "mixed", nb=nb).minus2lnl(coef)
    mask = np.arange(len(coef)) != 2
    np.testing.assert_array_equal((chi2 + logdet).cpu().numpy()[mask], good[mask])
    with pytest.raises(Exception):
        b.minus2lnl(bad)
