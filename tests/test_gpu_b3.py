"""Parity of the three-piece BF16 GEMM (``cfmi_gemm_b3``) with FP64 NumPy.

Operands are split into three BF16 planes on the device (``cfmi_split3``);
the product of the joined pieces must match the FP64 product to FP32
accuracy: |error| <= 2e-6 max_ij (|A| |B|^T)_ij over K <= 1000, which is a
few units of FP32 rounding (2^-24 = 6e-8) times sqrt(K) and leaves a wide
margin below a dropped cross term (a0 b1 alone is ~2^-9 relative).
Covers the three outputs (FP32 column-major with beta, BF16 planes, packed
lower row band with diagonal scale), ragged sizes, unaligned leading
dimensions (element-wise staging), triangular K-trims and batches, each with
both block configurations (flags 256: 128 x 128 tiles, 512: 256 x 256).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 2e-6


@pytest.fixture(scope="module")
def ctx():
    from cosmoforge_amd import engine

    return engine.Context(0)


def _planes(ctx, X, ld):
    """(rows x cols) FP64 host array -> three row-major BF16 planes on the
    device, leading dimension ld, planes rows*ld apart."""
    import torch
    from cosmoforge_amd.engine import P

    rows, cols = X.shape
    src = ctx.upload_cm(X, dtype=torch.float64)
    ps = rows * ld
    dst = torch.zeros(3 * ps, dtype=torch.int16, device=ctx.device)
    ctx.check(ctx.lib.cfmi_split3(ctx.h, 0, rows, cols, P(src), rows, P(dst), ld, ps), "split3")
    return dst, ps


def _join(ctx, planes, count, ps):
    import torch
    from cosmoforge_amd.engine import P

    out = torch.empty(count, dtype=torch.float64, device=ctx.device)
    ctx.check(ctx.lib.cfmi_join3(ctx.h, count, P(planes), ps, P(out)), "join3")
    return out.cpu().numpy()


def _scale(A, B):
    return np.max(np.abs(A) @ np.abs(B).T)


def test_split_join_round_trip(ctx):
    rng = np.random.default_rng(0)
    X = rng.standard_normal((37, 53)) * 10.0 ** rng.integers(-6, 6, (37, 53))
    pl, ps = _planes(ctx, X, 56)
    got = _join(ctx, pl, 37 * 56, ps).reshape(37, 56)[:, :53]
    assert np.max(np.abs(got - X) / np.abs(X)) < 2.0 ** -25


TILES = [256, 512]


@pytest.mark.parametrize("tile", TILES)
@pytest.mark.parametrize("M,N,K,lda,ldb", [(300, 190, 333, 336, 336), (128, 128, 256, 256, 256),
                                           (77, 45, 1000, 1001, 1003), (1, 3, 5, 5, 5), (520, 600, 96, 96, 104)])
def test_gemm_b3_fp32_out(ctx, M, N, K, lda, ldb, tile):
    import torch
    from cosmoforge_amd.engine import P

    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((N, K))
    C0 = rng.standard_normal((M, N))
    Ap, pA = _planes(ctx, A, lda)
    Bp, pB = _planes(ctx, B, ldb)
    C = ctx.upload_cm(C0, dtype=torch.float32)
    ctx.check(ctx.lib.cfmi_gemm_b3(ctx.h, 0, M, N, K, 0.75, P(Ap), lda, 0, pA, 0, P(Bp), ldb, 0, pB, 0, 0.5,
                                   P(C), M, 0, 0, 1, tile, 0, 0, 0, 1.0), "gemm_b3")
    got = C.cpu().double().numpy().reshape(N, M).T
    want = 0.75 * A @ B.T + 0.5 * C0.astype(np.float32)
    assert np.max(np.abs(got - want)) <= TOL * (_scale(A, B) + np.max(np.abs(C0)))


@pytest.mark.parametrize("tile", TILES)
def test_gemm_b3_planes_out_batched_and_trimmed(ctx, tile):
    """Mode 1 (BF16 planes of the product), batch of 2, A lower trapezoid
    with the K-trim flag: the trim must change nothing but the work done."""
    import torch
    from cosmoforge_amd.engine import P

    rng = np.random.default_rng(5)
    M, N, K, off = 260, 200, 400, 40
    lda = ldb = 400
    ldc = 200
    As = [np.tril(rng.standard_normal((M, K)), off) for _ in range(2)]
    Bs = [rng.standard_normal((N, K)) for _ in range(2)]
    Ap, pA = _planes(ctx, np.vstack(As), lda)        # batch stride M*lda inside one plane set
    Bp, pB = _planes(ctx, np.vstack(Bs), ldb)
    pC = 2 * M * ldc
    for flags in (0, 1):
        Cp = torch.zeros(3 * pC, dtype=torch.int16, device=ctx.device)
        ctx.check(ctx.lib.cfmi_gemm_b3(ctx.h, 1, M, N, K, 1.0, P(Ap), lda, M * lda, pA, off, P(Bp), ldb, N * ldb, pB,
                                       0, 0.0, P(Cp), ldc, M * ldc, pC, 2, flags | tile, 0, 0, 0, 1.0), "gemm_b3")
        got = _join(ctx, Cp, 2 * M * ldc, pC).reshape(2, M, ldc)
        for b in range(2):
            want = As[b] @ Bs[b].T
            # the BF16x3 output adds its own representation error (~2^-26 relative)
            assert np.max(np.abs(got[b] - want)) <= TOL * _scale(As[b], Bs[b])


@pytest.mark.parametrize("tile", TILES)
def test_gemm_b3_packed_band(ctx, tile):
    """Mode 2: rows r0..r1 of the lower triangle of alpha Z W^T, row-major
    packed from pbase, diagonal times pdiag; B upper-trimmed (W lower)."""
    import torch
    from cosmoforge_amd.engine import P

    rng = np.random.default_rng(9)
    n, r0, r1 = 300, 96, 233
    m, K = r1 - r0, 208
    W = np.tril(rng.standard_normal((n, n)))[:r1, :K]     # B[n][k] = W[n, k]: zero for k > n
    Z = rng.standard_normal((m, K))
    Zp, pZ = _planes(ctx, Z, K)
    Wp, pW = _planes(ctx, W, K)
    pb = r0 * (r0 + 1) // 2
    size = r1 * (r1 + 1) // 2 - pb
    X = torch.zeros(size, dtype=torch.float32, device=ctx.device)
    ctx.check(ctx.lib.cfmi_gemm_b3(ctx.h, 2, m, r1, K, 1.0, P(Zp), K, 0, pZ, 0, P(Wp), K, 0, pW, 0, 0.0, P(X), 0, 0,
                                   0, 1, 4 | tile, r0, 0, pb, 0.5), "gemm_b3")
    Pm = Z @ W.T
    want = np.concatenate([Pm[i - r0, :i + 1] * np.r_[np.ones(i), 0.5] for i in range(r0, r1)])
    got = X.cpu().double().numpy()
    assert np.max(np.abs(got - want)) <= TOL * _scale(Z, W)


@pytest.mark.parametrize("tile", TILES)
@pytest.mark.parametrize("mode,flags", [(1, 1), (2, 4), (0, 0)])
def test_gemm_b3_tile_orders_agree_bitwise(ctx, tile, mode, flags):
    """The XCD-grouped tile order (default) and the column-by-column order
    (flag 1024) compute every tile with the same K loop, so a batched,
    trimmed product over several groups of tiles is bit-identical in both,
    and within tolerance of FP64."""
    import torch
    from cosmoforge_amd.engine import P

    rng = np.random.default_rng(11 + mode)
    nb, n, r0 = 3, 1100, 0
    A = [np.tril(rng.standard_normal((n, n))) for _ in range(nb)]
    B = [np.tril(rng.standard_normal((n, n))) for _ in range(nb)]
    Ap, pA = _planes(ctx, np.vstack(A), n)
    Bp, pB = _planes(ctx, np.vstack(B), n)
    outs = []
    for order in (0, 1024):
        if mode == 1:
            pC = nb * n * n
            C = torch.zeros(3 * pC, dtype=torch.int16, device=ctx.device)
            args = (C, n, n * n, pC)
        elif mode == 2:
            C = torch.zeros(nb * (n * (n + 1) // 2), dtype=torch.float32, device=ctx.device)
            args = (C, 0, n * (n + 1) // 2, 0)
        else:
            C = torch.zeros(nb * n * n, dtype=torch.float32, device=ctx.device)
            args = (C, n, n * n, 0)
        ctx.check(ctx.lib.cfmi_gemm_b3(ctx.h, mode, n, n, n, 1.0, P(Ap), n, n * n, pA, 0, P(Bp), n, n * n, pB, 0,
                                       0.0, P(args[0]), args[1], args[2], args[3], nb, flags | tile | order, r0, 0,
                                       0, 1.0), "gemm_b3")
        outs.append(C.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])
    if mode == 0:
        got = outs[0].astype(np.float64).reshape(nb, n, n)
        for b in range(nb):
            assert np.max(np.abs(got[b].T - A[b] @ B[b].T)) <= TOL * _scale(A[b], B[b])


def test_gemm_b3_rejects_bad_arguments(ctx):
    from cosmoforge_amd.engine import P

    lib = ctx.lib
    assert lib.cfmi_gemm_b3(ctx.h, 3, 4, 4, 4, 1.0, None, 4, 0, 16, 0, None, 4, 0, 16, 0, 0.0, None, 4, 0, 0, 1, 0,
                            0, 0, 0, 1.0) != 0
    assert lib.cfmi_gemm_b3(ctx.h, 0, 4, 4, 8, 1.0, None, 4, 0, 16, 0, None, 8, 0, 16, 0, 0.0, None, 4, 0, 0, 1, 0,
                            0, 0, 0, 1.0) != 0
    assert lib.cfmi_gemm_b3(ctx.h, 1, 4, 4, 4, 1.0, None, 4, 0, 16, 0, None, 4, 0, 16, 0, 0.0, None, 4, 0, 8, 1, 0,
                            0, 0, 0, 1.0) != 0
