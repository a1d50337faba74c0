"""Kernel-level parity of libcfmi against the oracle (GPU).

Each test calls the library through the C ABI (``cosmoforge_amd.engine``)
and compares with the FP64 oracle or NumPy on the same seeded inputs:
GEMM and the MFMA operand layout (all transposes, ragged sizes, triangular
flags), pointings (RING ``geometry.dat`` and NESTED), signal blocks against
the reference's golden matrices, derivative stacks against the oracle's
binned derivative, and Cholesky / inverse / log-determinant.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from cosmoforge_amd import engine

    return engine.Context(0)


def _cm(ctx, a, dtype=None):
    import torch

    return ctx.upload_cm(a, dtype=dtype or torch.float64)


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_layouts(ctx, dt, ta, tb):
    import torch
    from cosmoforge_amd import engine

    rng = np.random.default_rng(11)
    tdt = torch.float64 if dt == "f64" else torch.float32
    tol = 1e-12 if dt == "f64" else 2e-5
    for (M, N, K) in [(37, 53, 29), (130, 129, 200), (257, 64, 1000)]:
        A = rng.standard_normal((K, M) if ta else (M, K))
        B = rng.standard_normal((N, K) if tb else (K, N))
        C0 = rng.standard_normal((M, N))
        dA, dB, dC = _cm(ctx, A, tdt), _cm(ctx, B, tdt), _cm(ctx, C0, tdt)
        engine.gemm(ctx, bool(ta), bool(tb), M, N, K, 0.5, dA, 0, A.shape[0], dB, 0, B.shape[0], -2.0, dC, 0, M)
        got = engine.to_numpy_cm(dC).astype(np.float64)
        ref = 0.5 * (A.T if ta else A) @ (B.T if tb else B) - 2.0 * C0
        assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < tol, (M, N, K)


def test_gemm_triangular_flags(ctx):
    from cosmoforge_amd import engine

    rng = np.random.default_rng(5)
    n = 300
    Lw = np.tril(rng.standard_normal((n, n)))
    B = rng.standard_normal((n, 77))
    dL, dB = _cm(ctx, Lw), _cm(ctx, B)
    dC = ctx.empty(n, 77)
    engine.gemm(ctx, False, False, n, 77, n, 1.0, dL, 0, n, dB, 0, n, 0.0, dC, 0, n, 1)
    assert np.allclose(engine.to_numpy_cm(dC), Lw @ B, rtol=0, atol=1e-11)
    # lower-only output of L L^T
    dS = ctx.zeros(n, n)
    engine.gemm(ctx, False, True, n, n, n, 1.0, dL, 0, n, dL, 0, n, 0.0, dS, 0, n, 4 | 16)
    got = engine.to_numpy_cm(dS)
    ref = Lw @ Lw.T
    assert np.allclose(np.tril(got), np.tril(ref), atol=1e-10)
    assert np.all(np.triu(got, 1) == 0)


def test_pointings(ctx):
    from cosmoforge_amd import engine
    from oracle import healpix
    from tests import cases

    g = cases.case_fortran_b().expect["geometry"]
    v = engine.pointings(ctx, 8, g[:, 0].astype(np.int64), nest=False).cpu().numpy()
    np.testing.assert_array_equal(v, g[:, 1:4])
    for nside in (1, 4, 16):
        pix = np.arange(12 * nside * nside)
        for nest in (False, True):
            r = healpix.pointings(nside, pix, nest)
            v = engine.pointings(ctx, nside, pix, nest).cpu().numpy()
            np.testing.assert_array_equal(v, r)
            # device evaluation: same formulas, device libm (within 2 ulp)
            v = engine.pointings(ctx, nside, pix, nest, libm_exact=False).cpu().numpy()
            assert np.max(np.abs(v - r)) <= 4.5e-16, (nside, nest)


def _device_signal(ctx, case):
    import torch
    from cosmoforge_amd import engine, layout as LY
    from tests import oracle_run as R

    lay = LY.build_layout(case.mask, case.spins, case.labels)
    vecs = [engine.pointings(ctx, case.nside, f.pixels, case.nest) for f in lay.fields]
    olay = R.layout(case)
    cls_s, _ = R.smoothed(case, olay)
    dcls = {k: ctx.upload(v) for k, v in cls_s.items()}
    S = engine.signal_matrix(ctx, lay, vecs, dcls, case.lmax_signal)
    torch.cuda.synchronize()
    return engine.to_numpy_cm(S), lay, vecs, olay, cls_s


@pytest.mark.parametrize("name", ["signal_tqu", "signal_t1t2"])
def test_signal_golden(ctx, name):
    from tests import cases

    case = cases.case_signal_tqu() if name == "signal_tqu" else cases.case_signal_t1t2()
    S, *_ = _device_signal(ctx, case)
    np.testing.assert_allclose(S, case.expect["S"], rtol=1e-12, atol=1e-20)


def test_signal_vs_oracle_with_eb_tb(ctx):
    """Nonzero EB/TB (no reference fixture has them): GPU vs oracle."""
    import torch
    from cosmoforge_amd import engine, layout as LY
    from oracle import kernels as K
    from tests import cases
    from tests import oracle_run as R

    case = cases.case_fisher_nside4("TQU")
    case.nside = 4
    olay = R.layout(case)
    cls_s, _ = R.smoothed(case, olay)
    rng = np.random.default_rng(9)
    cls_s["EB"] = 0.1 * cls_s["EE"] * rng.random(cls_s["EE"].size)
    cls_s["TB"] = 0.1 * cls_s["TE"] * rng.random(cls_s["TE"].size)
    So = K.signal_matrix(olay, cls_s, case.lmax_signal)
    lay = LY.build_layout(case.mask, case.spins, case.labels)
    vecs = [engine.pointings(ctx, 4, f.pixels, False) for f in lay.fields]
    S = engine.signal_matrix(ctx, lay, vecs, {k: ctx.upload(v) for k, v in cls_s.items()}, case.lmax_signal)
    torch.cuda.synchronize()
    S = engine.to_numpy_cm(S)
    assert np.max(np.abs(S - So)) <= 1e-12 * np.max(np.abs(So))


def test_deriv_stacks_vs_oracle(ctx):
    import torch
    from cosmoforge_amd import engine, layout as LY
    from oracle import kernels as K
    from oracle import pipeline as OP
    from tests import cases
    from tests import oracle_run as R

    case = cases.case_fisher_nside4("TQU")
    olay = R.layout(case)
    _, beams = R.smoothed(case, olay)
    prm = R.params(case, olay, beams)
    lay = LY.build_layout(case.mask, case.spins, case.labels)
    vecs = [engine.pointings(ctx, 4, f.pixels, False) for f in lay.fields]
    bins = OP.bins_fromdeltal(case.lmin, case.lmax)
    nb = len(bins)
    sm = OP.smoothing_factors(olay, beams, case.lmax_signal)
    wts = {lab: ctx.upload(sm[lab]) for lab in lay.labels}
    blocks = engine.deriv_stacks(ctx, lay, vecs, wts, case.lmax_signal, bins)
    torch.cuda.synchronize()
    n = lay.n
    for s, blk in enumerate(blocks):
        for b in range(nb):
            fi, fj, mode, _, w = prm[s * nb + b]
            ref = K.binned_derivative(olay, fi, fj, mode, w)
            D = blk.stack[b].cpu().numpy().T
            full = np.zeros((n, n))
            full[blk.ra:blk.ra + blk.ma, blk.rb:blk.rb + blk.mb] = D
            if blk.ra != blk.rb:
                full[blk.rb:blk.rb + blk.mb, blk.ra:blk.ra + blk.ma] = D.T
            scale = np.max(np.abs(ref)) + 1e-300
            assert np.max(np.abs(full - ref)) / scale < 1e-12, (s, b)


@pytest.mark.parametrize("n,nb", [(1000, 256), (2500, 512), (4096, 2048), (300, 64), (130, 1024)])
def test_diagonal_block_inverses(ctx, n, nb):
    """cfmi_trtri_blocks: every nb x nb diagonal block of U (the last one
    short) is the inverse of L's block, L from the same factorisation."""
    from cosmoforge_amd import engine

    rng = np.random.default_rng(n + nb)
    A = rng.standard_normal((n, n))
    C = A @ A.T / n + np.diag(rng.uniform(0.5, 2.0, n))
    dC = _cm(ctx, C)
    f = engine.factor(ctx, dC, n, need_whitener=False, diag_blocks=nb)
    L, U = engine.to_numpy_cm(f.L), engine.to_numpy_cm(f.U)
    for j0 in range(0, n, nb):
        j1 = min(n, j0 + nb)
        Lb, Ub = L[j0:j1, j0:j1], U[j0:j1, j0:j1]
        assert np.array_equal(np.triu(Ub, 1), np.zeros_like(Ub))
        ref = np.linalg.inv(Lb)
        assert np.max(np.abs(Ub - ref)) / np.max(np.abs(ref)) < 1e-12, (j0, j1)
    with pytest.raises(ValueError):
        engine.factor(ctx, _cm(ctx, C), n, need_whitener=False, diag_blocks=192)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 300, 1000, 2500])
def test_factor_and_inverse(ctx, n):
    from cosmoforge_amd import engine
    from oracle import linalg as LA

    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    C = A @ A.T + n * np.diag(rng.uniform(0.5, 2.0, n)) * 10 ** rng.uniform(-3, 3, n)
    C = 0.5 * (C + C.T)
    dC = _cm(ctx, C)
    f = engine.factor(ctx, dC, n)
    Ci = engine.to_numpy_cm(engine.inverse_from_whitener(ctx, f.W, n))
    ref = LA.matrix_inverse_symm(C)
    assert np.max(np.abs(Ci - ref)) / np.max(np.abs(ref)) < 1e-12
    assert f.logdet == pytest.approx(LA.matrix_slogdet_symm(C)[1], rel=1e-12, abs=1e-10)


@pytest.mark.parametrize("n", [700, 2500])
def test_potrf_fp32(ctx, n):
    """cfmi_potrf in FP32 (one-wave FP32 diagonal factor and inverse, FP32
    panel solves and updates): the factor equals the FP64 Cholesky factor to
    FP32 rounding times the growth of the factorisation, the diagonal-block
    inverses are the inverses of its 64 x 64 blocks, the strict upper triangle
    is zero."""
    import ctypes

    import torch
    from cosmoforge_amd import engine

    rng = np.random.default_rng(n + 3)
    X = rng.standard_normal((n, n // 2))
    A = X @ X.T / n + np.eye(n)
    d = 1.0 / np.sqrt(np.diag(A))
    A = d[:, None] * A * d[None, :]
    C = ctx.upload_cm(A, dtype=torch.float32)
    U = torch.zeros((n, n), dtype=torch.float32, device=ctx.device)
    info = ctypes.c_int(-1)
    rc = ctx.lib.cfmi_potrf(ctx.h, engine.F32, n, engine.P(C), n, engine.P(U), n, ctypes.byref(info))
    assert rc == 0 and info.value == 0
    L = C.cpu().numpy().T.astype(np.float64)
    Lref = np.linalg.cholesky(A)
    assert not np.triu(L, 1).any()
    assert np.max(np.abs(L - Lref)) < 2e-5
    Uh = U.cpu().numpy().T.astype(np.float64)
    for k0 in (0, 64 * ((n // 64) - 1), 64 * (n // 64)):
        kb = min(64, n - k0)
        if kb <= 0:
            continue
        blk = L[k0:k0 + kb, k0:k0 + kb]
        np.testing.assert_allclose(Uh[k0:k0 + kb, k0:k0 + kb] @ blk, np.eye(kb), atol=1e-5)


@pytest.mark.parametrize("n", [6500, 8200])
def test_potrf_lookahead(ctx, n):
    """Factorisations past two 2048-column panels update the next panel's
    columns on the context's stream and the rest of the trailing matrix on a
    second stream while the next panel is factored.  The FP64 factor equals
    the host Cholesky factor, the FP64 and sliced factors are the same bits
    on a repeat (the ordering between the streams leaves no race), and the
    sliced factor equals the FP64 one to the slicing level.  n = 8200 ends
    with an 8-column panel."""
    import ctypes

    import torch
    from cosmoforge_amd import engine

    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, n // 4))
    A = X @ X.T / n + np.eye(n)
    d = 1.0 / np.sqrt(np.diag(A))
    A = d[:, None] * A * d[None, :]
    dA = ctx.upload_cm(A)
    runs = {}
    for mode in ("fp64", "fp64", "s8", "s8"):
        C = dA.clone()
        info = ctypes.c_int(-1)
        if mode == "fp64":
            rc = ctx.lib.cfmi_potrf(ctx.h, engine.F64, n, engine.P(C), n, None, n, ctypes.byref(info))
        else:
            rc = ctx.lib.cfmi_potrf_s8(ctx.h, 5, n, engine.P(C), n, None, n, ctypes.byref(info))
        assert rc == 0 and info.value == 0
        runs.setdefault(mode, []).append(C.cpu().numpy().T.copy())
        del C
        torch.cuda.empty_cache()
    for mode, (a, b) in runs.items():
        np.testing.assert_array_equal(a, b, err_msg=mode)
    L = runs["fp64"][0]
    assert not np.triu(L, 1).any()
    Lref = np.linalg.cholesky(A)
    assert np.max(np.abs(L - Lref)) < 1e-12
    assert np.max(np.abs(runs["s8"][0] - L)) < 1e-8


def test_factor_not_pd_raises(ctx):
    from cosmoforge_amd import engine

    n = 200
    C = np.eye(n)
    C[150, 150] = -1.0
    with pytest.raises(ValueError, match="dpotrf failed with info=151"):
        engine.factor(ctx, _cm(ctx, C), n)


def test_gemm_batched_and_packed(ctx):
    import torch
    from cosmoforge_amd import engine

    rng = np.random.default_rng(21)
    M, N, K, nb = 150, 131, 77, 5
    A = rng.standard_normal((nb, M, K))
    B = rng.standard_normal((K, N))
    dA = torch.as_tensor(np.ascontiguousarray(A.transpose(0, 2, 1))).cuda()  # col-major blocks
    dB = ctx.upload_cm(B)
    dC = torch.zeros((nb, N, M), dtype=torch.float64, device="cuda")
    engine.gemm_batched(ctx, False, False, M, N, K, 1.0, dA, 0, M, M * K, dB, 0, K, 0, 0.0, dC, 0, M, M * N, nb)
    got = dC.cpu().numpy().transpose(0, 2, 1)
    assert np.max(np.abs(got - A @ B)) < 1e-11
    # packed lower triangle of a symmetric product, offset inside a larger matrix
    n, r0 = 200, 37
    X = rng.standard_normal((n - r0, 40))
    dX = ctx.upload_cm(X)
    Kp = engine.packed_size(n)
    Pk = torch.zeros((2, Kp), dtype=torch.float64, device="cuda")
    engine.gemm_packed_batched(ctx, False, True, n - r0, n - r0, 40, 1.0, dX, 0, n - r0, 0, dX, 0, n - r0, 0, 0.0,
                               Pk, Kp, n, r0, r0, 2)
    full = np.zeros((n, n))
    full[r0:, r0:] = X @ X.T
    i, j = np.tril_indices(n)
    ref = full[i, j] * np.where(i == j, engine.SQRT_HALF, 1.0)
    order = np.lexsort((i, j))  # column-major packing
    for b in range(2):
        assert np.max(np.abs(Pk[b].cpu().numpy() - ref[order])) < 1e-11


def test_gemm_ex_band_packing_and_offsets(ctx):
    """cfmi_gemm_ex: row-major packed lower band output and shifted
    triangular trims, FP64 and FP32, batched."""
    import torch
    from cosmoforge_amd import engine

    rng = np.random.default_rng(21)
    n = 300
    W = np.tril(rng.standard_normal((n, n)))
    r0, r1, ra = 96, 233, 40
    m, K, N = r1 - r0, 150, 90
    for dt, tol in [(torch.float64, 1e-12), (torch.float32, 3e-5)]:
        Wd = ctx.upload_cm(W, dtype=dt)
        Bm = rng.standard_normal((2, K, N))
        Bd = torch.tensor(np.ascontiguousarray(np.transpose(Bm, (0, 2, 1))), dtype=dt, device=ctx.device)
        # Z_b = W[r0:r1, ra:ra+K] B_b with A-lower trim offset r0 - ra
        Z = torch.zeros((2, N, m), dtype=dt, device=ctx.device)
        engine.gemm_ex(ctx, False, False, m, N, K, 1.0, Wd, r0 + ra * n, n, Bd, 0, K, 0.0, Z, 0, m, sB=K * N,
                       sC=m * N, batch=2, flags=1, tri_a=r0 - ra)
        for b in range(2):
            want = W[r0:r1, ra:ra + K] @ Bm[b]
            got = Z[b].cpu().numpy().T
            assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < tol
        # band output of Z_b W[0:r1, 0:N]^T (B-upper trim) into rows r0..r1
        pb = r0 * (r0 + 1) // 2
        X = torch.zeros((2, engine.band_size(r0, r1)), dtype=dt, device=ctx.device)
        engine.gemm_ex(ctx, False, True, m, r1, N, 1.0, Z, 0, m, Wd, 0, n, 0.0, X, 0, sA=m * N,
                       sC=X.shape[1], batch=2, flags=4, out_mode=2, row0=r0, col0=0, pbase=pb, pdiag=0.5)
        for b in range(2):
            P = Z[b].cpu().double().numpy().T @ W[:r1, :N].T
            want = np.concatenate([P[i - r0, :i + 1] * np.r_[np.ones(i), 0.5] for i in range(r0, r1)])
            got = X[b].cpu().double().numpy()
            assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < tol


def test_gram_fp32_accumulates_in_fp64(ctx):
    """FP32 inputs over a long K: FP32 products, FP64 accumulation across
    chunks, so the error stays at FP32 rounding of the inputs, not of a
    10^7-term FP32 sum."""
    import torch
    from cosmoforge_amd import engine

    rng = np.random.default_rng(3)
    K, m = 4_000_000, 5
    X = rng.random((m, K)).astype(np.float32)          # positive terms: worst case for a plain FP32 sum
    Xd = torch.tensor(X, device=ctx.device)
    G = engine.gram(ctx, m, m, K, Xd, Xd)
    want = X.astype(np.float64) @ X.astype(np.float64).T
    assert np.max(np.abs(G - want) / want) < 1e-6
    Xd64 = Xd.double()
    G64 = engine.gram(ctx, m, m, K, Xd64, Xd64)
    assert np.max(np.abs(G64 - want) / want) < 1e-12


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("flags", [0, 1, 4 | 16])
def test_batched_gemm_tile_orders_agree_bitwise(ctx, dt, flags):
    """Batched products: the XCD-grouped tile order (default for batch > 1)
    and the plain order over the batch dimension (flag 1024) run the same K
    loop per tile, so they agree bit for bit, over several groups of tiles,
    with the A-trim (1) and the B-trim with a lower-triangle output (4 | 16);
    both match FP64 NumPy."""
    import torch
    from cosmoforge_amd import engine

    rng = np.random.default_rng(31 + flags)
    tdt = torch.float64 if dt == "f64" else torch.float32
    nb, m, n, k = 3, 1100, 1000, 900
    A = np.tril(rng.standard_normal((nb, m, k)), 0)
    B = np.triu(rng.standard_normal((nb, k, n)), 0)
    dA = torch.tensor(A.transpose(0, 2, 1).copy(), dtype=tdt, device=ctx.device)   # column-major m x k per batch
    dB = torch.tensor(B.transpose(0, 2, 1).copy(), dtype=tdt, device=ctx.device)   # column-major k x n per batch
    outs = []
    for order in (0, 1024):
        C = torch.zeros((nb, n, m), dtype=tdt, device=ctx.device)
        engine.gemm_ex(ctx, False, False, m, n, k, 1.0, dA, 0, m, dB, 0, k, 0.0, C, 0, m, sA=m * k, sB=k * n,
                       sC=m * n, batch=nb, flags=flags | order)
        outs.append(C.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])
    tol = 1e-12 if dt == "f64" else 2e-5
    for b in range(nb):
        ref = A[b] @ B[b]
        got = outs[0][b].T.astype(np.float64)
        if flags & 16:
            ref, got = np.tril(ref), np.tril(got)
        assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < tol, b


def test_deriv_stacks_chunked_resume_and_fp32(ctx):
    """Bin chunks that resume the recurrence from its saved state give the
    same stacks as one pass (bit for bit); FP32 stacks are the FP64 sums
    rounded once."""
    import torch
    from cosmoforge_amd import engine
    from tests import cases

    case = cases.case_fisher_nside4("TEB")
    from cosmoforge_amd.core import Inputs, bins_from_params
    from cosmoforge_amd import pipeline

    p = dict(nside=case.nside, spins=case.spins, labels=case.labels, lmin=case.lmin, lmax=case.lmax,
             lmax_signal=case.lmax_signal, ordering="NESTED" if case.nest else "RING")
    inp = Inputs(p, case.mask, case.noise, case.cls, case.beam)
    cov = pipeline.Covariance(ctx, inp)
    bins = bins_from_params(inp.params)
    full = pipeline.DerivativeSource(cov, bins, "Cl", torch.float64)
    ref = full.make_resident()
    got = {}
    for s, b0, blk in full.__class__(cov, bins, "Cl", torch.float64).chunks(1):   # one bin per chunk
        got[(s, b0)] = blk.stack.clone()
    for s, blk in enumerate(ref):
        for b in range(len(bins)):
            assert torch.equal(got[(s, b)][0], blk.stack[b]), (s, b)
    f32 = pipeline.DerivativeSource(cov, bins, "Cl", torch.float32).make_resident()
    for a, b in zip(ref, f32):
        assert torch.equal(a.stack.float(), b.stack)


def test_deriv_stacks_bf16x3_and_mirrors(ctx):
    """Three-piece BF16 stacks: the joined pieces equal the FP64 stacks to
    2^-25 relative (three BF16 pieces carry 24 bits; above an FP64-noise
    floor), and the transposed
    stacks of the a != b blocks (spin-0 cross fields in TEB, TE/TB in TQU)
    hold the same pieces transposed."""
    import torch
    from cosmoforge_amd import engine, pipeline
    from cosmoforge_amd.core import Inputs, bins_from_params
    from tests import cases

    for name in ("TEB", "TQU"):
        case = cases.case_fisher_nside4(name)
        p = dict(nside=case.nside, spins=case.spins, labels=case.labels, lmin=case.lmin, lmax=case.lmax,
                 lmax_signal=case.lmax_signal, ordering="NESTED" if case.nest else "RING")
        inp = Inputs(p, case.mask, case.noise, case.cls, case.beam)
        cov = pipeline.Covariance(ctx, inp)
        bins = bins_from_params(inp.params)
        ref = pipeline.DerivativeSource(cov, bins, "Cl", torch.float64).make_resident()
        b3 = pipeline.DerivativeSource(cov, bins, "Cl", engine.BF16X3, mirrors=True).make_resident()
        nmirror = 0
        for a, b in zip(ref, b3):
            nb, mb, ma = a.stack.shape
            assert b.stack.shape == (3, nb, mb, ma)
            cnt = nb * mb * ma
            j = torch.empty(cnt, dtype=torch.float64, device=ctx.device)
            ctx.check(ctx.lib.cfmi_join3(ctx.h, cnt, engine.P(b.stack), b.stack.stride(0), engine.P(j)), "join3")
            x = a.stack.reshape(-1)
            # The FP64 and three-piece kernels are separate instantiations: FMA
            # contraction may differ at FP64 rounding (entries that cancel in the
            # spin-2 rotation), hence the 2^-40 max|x| floor.
            err = (j - x).abs()
            bound = 2.0 ** -25 * x.abs() + 2.0 ** -40 * x.abs().max()
            bad = err > bound
            if bad.any():
                print(name, "worst:", [(float(x[i]), float(j[i])) for i in torch.nonzero(bad)[:5, 0]])
            assert not bad.any(), name
            if a.sym:
                assert b.mirror is None
            else:
                nmirror += 1
                assert b.mirror.shape == (3, nb, ma, mb)
                assert torch.equal(b.mirror, b.stack.transpose(2, 3)), name
        assert nmirror == sum(1 for g in engine.block_geometry(inp.layout) if not g.sym) > 0


@pytest.mark.parametrize("K,na,nb", [(1_000_003, 37, 3), (4096, 1, 1), (5000, 5, 6), (17, 3, 2), (0, 2, 2)])
def test_dot_batched(ctx, K, na, nb):
    """cfmi_dot_batched against NumPy FP64 (ragged K, more than four B
    vectors, K = 0), and bit-for-bit repeatable."""
    import torch
    from cosmoforge_amd.engine import P

    rng = np.random.default_rng(K + na)
    A = rng.standard_normal((na, K))
    B = rng.standard_normal((nb, K))
    Ad = torch.tensor(A, device=ctx.device)
    Bd = torch.tensor(B, device=ctx.device)
    outs = []
    for _ in range(2):
        R = torch.full((nb, na), np.nan, dtype=torch.float64, device=ctx.device)
        ctx.check(ctx.lib.cfmi_dot_batched(ctx.h, K, na, P(Ad), K, nb, P(Bd), K, P(R), na), "dot_batched")
        outs.append(R.cpu().numpy())
    want = B @ A.T
    if K == 0:
        assert np.array_equal(outs[0], np.zeros((nb, na)))
        return
    scale = np.sqrt(K) * max(1.0, np.max(np.abs(A)) * np.max(np.abs(B)))
    assert np.max(np.abs(outs[0] - want)) <= 1e-13 * scale
    assert np.array_equal(outs[0], outs[1])


def test_offdiag_nnz_and_diagonal_noise_G(ctx):
    """cfmi_offdiag_nnz counts off-diagonal nonzeros; the QML's G for a
    diagonal N (one symmetric product) equals the general three-product
    form to FP64 rounding."""
    import ctypes

    import torch
    from cosmoforge_amd import pipeline
    from cosmoforge_amd.core import Inputs
    from cosmoforge_amd.engine import P
    from tests import cases

    rng = np.random.default_rng(2)
    A = np.diag(rng.random(300) + 1.0)
    A[5, 7] = 1e-300
    A[299, 0] = -2.0
    Ad = ctx.upload_cm(A)
    cnt = ctypes.c_int64()
    ctx.check(ctx.lib.cfmi_offdiag_nnz(ctx.h, 300, P(Ad), 300, ctypes.byref(cnt)), "offdiag_nnz")
    assert cnt.value == 2

    case = cases.case_fisher_nside4("TQU")
    p = dict(nside=case.nside, spins=case.spins, labels=case.labels, lmin=case.lmin, lmax=case.lmax,
             lmax_signal=case.lmax_signal, ordering="NESTED" if case.nest else "RING")
    cov = pipeline.Covariance(ctx, Inputs(p, case.mask, case.noise, case.cls, case.beam))
    cov.build()
    assert cov.noise_is_diagonal()
    G1 = pipeline.noise_weighted_inverse(cov, "fp64").cpu().numpy()
    cov._noise_diag = False                       # force the general three-product form
    G3 = pipeline.noise_weighted_inverse(cov, "fp64").cpu().numpy()
    assert np.max(np.abs(G1 - G3)) <= 1e-12 * np.max(np.abs(G3))
    assert np.array_equal(G1, G1.T)


@pytest.mark.parametrize("lmin,lmax", [(2, 12), (5, 40)])
def test_harmonic_modes_vs_oracle(ctx, lmin, lmax):
    """cfmi_harmonic_modes (Wigner recurrences on the device) against the
    oracle's closed-form Wigner sums for a T + QU layout, NESTED pixels in
    both hemispheres; out-of-range arguments are refused."""
    import torch

    from cosmoforge_amd import engine
    from oracle import harmonic as OH
    from oracle import healpix
    from oracle import kernels as OK

    nside = 8
    rng = np.random.default_rng(1)
    mask = np.zeros((768, 3))
    mask[np.sort(rng.choice(768, 300, replace=False)), 0] = 1
    qu = np.sort(rng.choice(768, 250, replace=False))
    mask[qu, 1:] = 1
    lay = OK.layout_from_mask(mask, [0, 2], ["T", "E", "B"], nside, nest=True)
    groups = OH.mode_groups(lay, lmin, lmax)
    ng = OH.n_modes(lmin, lmax)
    n = lay.n
    V = torch.zeros((3 * ng, n), dtype=torch.float64, device=ctx.device)
    offs = lay.offsets()
    for fi, f in enumerate(lay.fields):
        vec = ctx.upload(np.ascontiguousarray(healpix.pointings(nside, f.pixels, nest=True)))
        if f.spin == 0:
            rc = ctx.lib.cfmi_harmonic_modes(ctx.h, 0, f.n, engine.P(vec), lmin, lmax, engine.P(V), n, 3 * ng,
                                             offs[fi], 0, 0, 0)
        else:
            rc = ctx.lib.cfmi_harmonic_modes(ctx.h, 2, f.n, engine.P(vec), lmin, lmax, engine.P(V), n, 3 * ng,
                                             offs[fi], offs[fi] + f.n, ng, 2 * ng)
        ctx.check(rc, "harmonic_modes")
    Vh = V.cpu().numpy().T            # (n, 3 ng)
    for g, lab in enumerate(["T", "E", "B"]):
        ref = groups[lab]
        got = Vh[:, g * ng:(g + 1) * ng]
        assert np.abs(got - ref).max() < 1e-12 * max(1.0, np.abs(ref).max()), lab
    vec = ctx.upload(np.zeros((4, 3)))
    assert ctx.lib.cfmi_harmonic_modes(ctx.h, 2, 4, engine.P(vec), 1, 8, engine.P(V), n, 3 * ng, 0, 0, 0, 0) < 0
    assert ctx.lib.cfmi_harmonic_modes(ctx.h, 0, 4, engine.P(vec), lmin, lmax, engine.P(V), n, ng - 1, 0, 0, 0,
                                       0) < 0


def test_harmonic_coldots_equals_pairsum(ctx):
    """The column-dot noise-bias reduction against the per-parameter pair
    sums over an explicit N R: diagonal row weights and a dense X."""
    import torch

    from cosmoforge_amd import engine

    rng = np.random.default_rng(12)
    n, lmin, lmax, G = 700, 2, 9, 3
    ng = (lmax + 1) ** 2 - lmin ** 2
    nt = G * ng
    bins = [(2, 3), (4, 6), (7, 9)]
    pairs = np.array([[0, 0, -1, -1], [1, 1, -1, -1], [2, 2, -1, -1], [1, 2, 2, 1], [0, 1, 1, 0], [0, 2, 2, 0]],
                     dtype=np.int32)
    lw = lmax + 1
    w = rng.uniform(0.5, 2.0, (len(pairs), lw))
    tabs = (lmin, len(bins), ctx.upload(np.array([a for a, _ in bins], dtype=np.int32), dtype=torch.int32),
            ctx.upload(np.array([b for _, b in bins], dtype=np.int32), dtype=torch.int32), len(pairs),
            ctx.upload(pairs, dtype=torch.int32), ctx.upload(np.arange(G, dtype=np.int64) * ng, dtype=torch.int64),
            ctx.upload(w), lw)
    targs = (tabs[0], tabs[1], engine.P(tabs[2]), engine.P(tabs[3]), tabs[4], engine.P(tabs[5]), engine.P(tabs[6]),
             engine.P(tabs[7]), tabs[8])
    npar = len(pairs) * len(bins)
    R = rng.standard_normal((n, nt))
    d = rng.uniform(0.5, 2.0, n)
    dR, dd = ctx.upload_cm(R), ctx.upload(d)
    want = torch.empty(npar, dtype=torch.float64, device=ctx.device)
    NR = ctx.upload_cm(d[:, None] * R)
    ctx.check(ctx.lib.cfmi_harmonic_pairsum(ctx.h, n, engine.P(NR), n, 0, engine.P(dR), n, 0, 1, *targs, 0.5,
                                            engine.P(want), npar), "pairsum")
    got = torch.empty(npar, dtype=torch.float64, device=ctx.device)
    ctx.check(ctx.lib.cfmi_harmonic_coldots(ctx.h, n, engine.P(dR), n, engine.P(dR), n, engine.P(dd), G, ng, *targs,
                                            0.5, engine.P(got)), "coldots")
    np.testing.assert_allclose(got.cpu().numpy(), want.cpu().numpy(), rtol=1e-12, atol=1e-12)
    got2 = torch.empty(npar, dtype=torch.float64, device=ctx.device)
    ctx.check(ctx.lib.cfmi_harmonic_coldots(ctx.h, n, engine.P(NR), n, engine.P(dR), n, None, G, ng, *targs, 0.5,
                                            engine.P(got2)), "coldots")
    np.testing.assert_allclose(got2.cpu().numpy(), want.cpu().numpy(), rtol=1e-12, atol=1e-12)
    # host check of one parameter: TE bin 1 (pairs (0, 1) and (1, 0))
    p = 4 * len(bins) + 1
    ks = range(4 * 4 - 4, 7 * 7 - 4)
    ell = lambda k: int(np.floor(np.sqrt(k + lmin * lmin) + 1e-9))  # noqa: E731
    ref = 0.5 * sum(w[4, ell(k)] * (np.sum(d * R[:, k] * R[:, ng + k]) + np.sum(d * R[:, ng + k] * R[:, k]))
                    for k in ks)
    assert got.cpu().numpy()[p] == pytest.approx(ref, rel=1e-12)
