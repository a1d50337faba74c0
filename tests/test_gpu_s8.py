"""The int8-sliced FP64 products (``cfmi_slice8`` / ``cfmi_gemm_s8``) and the
blocked triangular solve built on them (``engine.trsm``).

* Slicing and products equal the host restatement (``tests/s8_emul.py``) bit
  for bit: same slices, same exact integer level sums, same FP64 epilogue,
  for 3, 4 and 5 slices, both source layouts, ragged sizes, alpha / beta,
  the symmetric (lower tiles + mirror) output and a K split beyond the exact
  int32 range.
* The products stay within the slicing bound of the FP64 product, which a
  dropped level or a wrong scale would exceed.
* The blocked solves (one and two levels of blocking) equal SciPy's
  triangular solves to the bound of their arithmetic (FP64: 1e-12; 4 or 5
  slices: the slicing error times the growth of the solve).
"""

import numpy as np
import pytest

from tests import s8_emul as E

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from cosmoforge_amd import engine

    return engine.Context(0)


def _mat(rng, m, k, spread=3):
    """Rows of different magnitudes, some entries zero, one zero row."""
    X = rng.standard_normal((m, k)) * 10.0 ** rng.uniform(-spread, spread, (m, 1))
    X[rng.random((m, k)) < 0.05] = 0.0
    X[m // 3] = 0.0
    return X


def _dev_slices(ctx, X, s, layout):
    """Slice the rows of host X on the device; layout 1 reads X column-major
    (rows of a column-major matrix), layout 0 reads it as k-contiguous rows."""
    import torch
    from cosmoforge_amd import engine

    m, k = X.shape
    if layout == 1:
        src = ctx.upload_cm(X)                     # element (r, k) at r + k m
        return engine.slice8(ctx, s, 1, m, k, src, 0, m)
    src = torch.as_tensor(np.ascontiguousarray(X)).to(ctx.device)   # row r at r k
    return engine.slice8(ctx, s, 0, m, k, src, 0, k)


@pytest.mark.parametrize("s", [3, 4, 5])
@pytest.mark.parametrize("layout", [0, 1])
def test_slices_bit_exact(ctx, s, layout):
    rng = np.random.default_rng(10 + s + 7 * layout)
    X = _mat(rng, 203, 331)
    sl = _dev_slices(ctx, X, s, layout)
    mu, ref = E.slices(X, s)
    got = sl.data.cpu().numpy()[:, :203, :]
    np.testing.assert_array_equal(sl.mu.cpu().numpy()[:203], mu)
    np.testing.assert_array_equal(got[:, :, :331].astype(np.int64), ref)
    assert not got[:, :, 331:].any()                          # zero padding up to ldk
    # the represented values stay within the slicing bound
    err = np.abs(E.unslice(mu, ref) - X)
    assert np.all(err <= (0.5 / (127.0 * 254.0 ** (s - 1)) + 4e-16) * mu[:, None])   # + FP64 rounding


@pytest.mark.parametrize("s", [3, 4, 5])
@pytest.mark.parametrize("shape", [(130, 97, 200), (257, 385, 1000), (64, 64, 64)])
def test_gemm_s8_bit_exact_and_bound(ctx, s, shape):
    from cosmoforge_amd import engine

    m, n, k = shape
    rng = np.random.default_rng(s * 100 + m)
    A, B = _mat(rng, m, k), _mat(rng, n, k)
    C0 = rng.standard_normal((m, n))
    Sa, Sb = _dev_slices(ctx, A, s, 1), _dev_slices(ctx, B, s, 0)
    C = ctx.upload_cm(C0)
    engine.gemm_s8(ctx, m, n, k, -0.75, Sa, 0, Sb, 0, 0.5, C, 0, m)
    got = engine.to_numpy_cm(C)
    want = E.gemm(A, B, s, alpha=-0.75, beta=0.5, C=C0)
    np.testing.assert_array_equal(got, want)
    exact = -0.75 * (A @ B.T) + 0.5 * C0
    assert np.all(np.abs(got - exact) <= 0.75 * E.bound(A, B, s) + 1e-15 * np.abs(C0))


@pytest.mark.parametrize("s", [3, 5])
def test_gemm_s8_single_stage(ctx, s):
    """Flag bit 1 (one LDS stage staged through registers) gives the same
    result as the two asynchronous stages, bit for bit."""
    from cosmoforge_amd import engine

    m, n, k = 257, 190, 700
    rng = np.random.default_rng(70 + s)
    A, B = _mat(rng, m, k), _mat(rng, n, k)
    C0 = rng.standard_normal((m, n))
    Sa, Sb = _dev_slices(ctx, A, s, 1), _dev_slices(ctx, B, s, 0)
    got = {}
    for flags in (0, 2):
        C = ctx.upload_cm(C0)
        engine.gemm_s8(ctx, m, n, k, 1.25, Sa, 0, Sb, 0, -0.5, C, 0, m, flags=flags)
        got[flags] = engine.to_numpy_cm(C)
    np.testing.assert_array_equal(got[2], got[0])
    np.testing.assert_array_equal(got[0], E.gemm(A, B, s, alpha=1.25, beta=-0.5, C=C0))


def test_gemm_s8_symmetric_and_offsets(ctx):
    """Lower tiles + mirror equal the full product; row offsets into the
    slice buffers address the right rows."""
    import torch
    from cosmoforge_amd import engine

    rng = np.random.default_rng(5)
    Z = rng.standard_normal((300, 777))
    S = _dev_slices(ctx, Z, 4, 0)
    M = torch.zeros((300, 300), dtype=torch.float64, device=ctx.device)
    engine.gemm_s8(ctx, 300, 300, 777, 1.0, S, 0, S, 0, 0.0, M, 0, 300, flags=1)
    got = engine.to_numpy_cm(M)
    want = E.gemm(Z, Z, 4)
    low = np.tril(want)
    np.testing.assert_array_equal(np.tril(got), low)
    np.testing.assert_array_equal(got, low + np.tril(low, -1).T)
    # rows 40.. of A against rows 7.. of B
    C = torch.zeros((50, 100), dtype=torch.float64, device=ctx.device)   # column-major 100 x 50
    engine.gemm_s8(ctx, 100, 50, 777, 1.0, S, 40, S, 7, 0.0, C, 0, 100)
    full = want
    # the slices of a row do not depend on the other rows, so the block of
    # the full product is the product of the blocks
    np.testing.assert_array_equal(engine.to_numpy_cm(C), full[40:140, 7:57])


def test_gemm_s8_lower_trapezoid(ctx):
    """Flag bit 5 with n < m (the sliced Cholesky's in-panel updates): the
    entries on or below the diagonal equal the full product's, bit for bit,
    the strict upper part keeps C's values; n > m is rejected."""
    import torch
    from cosmoforge_amd import engine

    rng = np.random.default_rng(6)
    P = rng.standard_normal((300, 320))
    S = _dev_slices(ctx, P, 5, 0)
    C0 = rng.standard_normal((300, 140))
    C = ctx.upload_cm(C0)
    engine.gemm_s8(ctx, 300, 140, 320, -1.0, S, 0, S, 0, 1.0, C, 0, 300, flags=32)
    got = engine.to_numpy_cm(C)
    want = C0 - E.gemm(P, P[:140], 5)
    low = np.tril(np.ones((300, 140), dtype=bool))
    np.testing.assert_allclose(got[low], want[low], rtol=0, atol=1e-13 * np.max(np.abs(want)))
    np.testing.assert_array_equal(got[~low], C0[~low])
    W = torch.zeros((300, 140), dtype=torch.float64, device=ctx.device)
    bad = ctx.lib.cfmi_gemm_s8(ctx.h, 5, 140, 300, 320, 1.0, engine.P(S.data), S.ldk, S.ps, engine.P(S.mu), 0,
                               engine.P(S.data), S.ldk, S.ps, engine.P(S.mu), 0, 0.0, engine.P(W), 140, 32, 0, 0, 0,
                               1.0)
    assert bad < 0 and b"n <= m" in ctx.lib.cfmi_last_error(ctx.h)


def test_gemm_s8_k_split(ctx):
    """k beyond the exact int32 range of one pass: chunks added in FP64."""
    from cosmoforge_amd import engine

    s = 5
    k = E.kmax(s) + 1000
    assert ctx.lib.cfmi_s8_kmax(s) == E.kmax(s)
    rng = np.random.default_rng(9)
    A, B = rng.standard_normal((70, k)), rng.standard_normal((33, k))
    Sa, Sb = _dev_slices(ctx, A, s, 0), _dev_slices(ctx, B, s, 0)
    C = ctx.empty(70, 33)
    engine.gemm_s8(ctx, 70, 33, k, 1.0, Sa, 0, Sb, 0, 0.0, C, 0, 70)
    got = engine.to_numpy_cm(C)
    np.testing.assert_array_equal(got, E.gemm(A, B, s))
    assert np.all(np.abs(got - A @ B.T) <= E.bound(A, B, s))


@pytest.mark.parametrize("k_extra", [0, 1])
def test_gemm_s8_trims_and_packed_band(ctx, k_extra):
    """Triangular K-trims on A and B (they skip only zeros, so the result is
    the full product bit for bit) into a packed lower row band with a scaled
    diagonal, accumulated onto its contents; once more with K beyond the
    exact int32 range of one pass (the trims are then relative to each
    pass)."""
    import torch
    from cosmoforge_amd import engine

    s = 4
    rng = np.random.default_rng(31 + k_extra)
    m, n = 300, 200
    k = (E.kmax(s) + 700) if k_extra else 900
    offa, offb = (E.kmax(s) - 150, E.kmax(s) + 90) if k_extra else (37, 290)
    A = rng.standard_normal((m, k))
    B = rng.standard_normal((n, k))
    A[np.arange(k)[None, :] > np.arange(m)[:, None] + offa] = 0.0
    B[np.arange(k)[None, :] > np.arange(n)[:, None] + offb] = 0.0
    Sa, Sb = _dev_slices(ctx, A, s, 0), _dev_slices(ctx, B, s, 0)
    row0, col0 = 250, 180                     # band rows 250..549, columns 180..379
    pbase = row0 * (row0 + 1) // 2
    size = (row0 + m) * (row0 + m + 1) // 2 - pbase
    X0 = rng.standard_normal(size)
    X = torch.tensor(X0, device=ctx.device)
    engine.gemm_s8(ctx, m, n, k, 1.0, Sa, 0, Sb, 0, 1.0, X, 0, 0, tri_a=offa, tri_b=offb,
                   packed=(row0, col0, pbase, 0.7))
    ii = row0 + np.arange(m)[:, None]
    cc = col0 + np.arange(n)[None, :]
    ok = cc <= ii
    idx = np.where(ok, ii * (ii + 1) // 2 + cc - pbase, 0)
    C0 = np.where(ok, X0[idx], 0.0)
    full = E.gemm(A, B, s, beta=1.0, C=C0, scale=np.where(cc == ii, 0.7, 1.0))
    want = X0.copy()
    want[idx[ok]] = full[ok]
    np.testing.assert_array_equal(X.cpu().numpy(), want)


def test_gemm_s8_rejects_bad_arguments(ctx):
    from cosmoforge_amd import engine

    S = engine.Slices(ctx, 4, 10, 64)
    C = ctx.empty(10, 10)
    bad = ctx.lib.cfmi_gemm_s8(ctx.h, 6, 10, 10, 64, 1.0, engine.P(S.data), S.ldk, S.ps, engine.P(S.mu), 0,
                               engine.P(S.data), S.ldk, S.ps, engine.P(S.mu), 0, 0.0, engine.P(C), 10, 0, 0, 0, 0, 1.0)
    assert bad < 0 and b"slices" in ctx.lib.cfmi_last_error(ctx.h)
    bad = ctx.lib.cfmi_gemm_s8(ctx.h, 4, 10, 10, 200, 1.0, engine.P(S.data), S.ldk, S.ps, engine.P(S.mu), 0,
                               engine.P(S.data), S.ldk, S.ps, engine.P(S.mu), 0, 0.0, engine.P(C), 10, 0, 0, 0, 0,
                               1.0)
    assert bad < 0   # row pitch 64 below k = 200


@pytest.mark.parametrize("slices,outer,diag", [(0, 2048, 0), (4, 2048, 0), (5, 2048, 0), (0, 384, 0), (4, 384, 0),
                                               (4, 2048, 5), (0, 384, 5)])
def test_blocked_trsm(ctx, slices, outer, diag):
    """Blocked L^-1 B and L^-T B against the host solves; updates and
    diagonal products in FP64 or on int8 slices."""
    import torch
    from scipy.linalg import solve_triangular

    from cosmoforge_amd import engine

    rng = np.random.default_rng(3 + slices)
    n, nrhs = 900, 130
    G = rng.standard_normal((n, n))
    C = G @ G.T / n + np.diag(rng.uniform(0.5, 2.0, n))
    dC = ctx.upload_cm(C)
    f = engine.factor(ctx, dC, n, need_whitener=False, diag_blocks=128)
    Lh = np.tril(engine.to_numpy_cm(f.L))
    Bh = rng.standard_normal((n, nrhs))
    X = torch.zeros((nrhs, n), dtype=torch.float64, device=ctx.device)
    engine.trsm(ctx, f, n, ctx.upload_cm(Bh), X, nrhs, slices=slices, outer=outer, diag_slices=diag)
    want = solve_triangular(Lh, Bh, lower=True)
    err = np.max(np.abs(engine.to_numpy_cm(X) - want)) / np.max(np.abs(want))
    Y = torch.zeros((nrhs, n), dtype=torch.float64, device=ctx.device)
    engine.trsm(ctx, f, n, X.clone(), Y, nrhs, transpose=True, slices=slices, outer=outer, diag_slices=diag)
    want2 = solve_triangular(Lh.T, want, lower=False)
    err2 = np.max(np.abs(engine.to_numpy_cm(Y) - want2)) / np.max(np.abs(want2))
    tol = {0: 1e-12, 4: 1e-7, 5: 1e-9}[slices] if not diag else {0: 1e-9, 4: 1e-7}[slices]
    assert err < tol and err2 < tol, (err, err2)


@pytest.mark.parametrize("s", [4, 5])
def test_potrf_s8_matches_fp64_factor(ctx, s):
    """The Cholesky factor with its rank-2048 trailing updates and its
    in-panel updates of K >= 256 on int8 slices equals the FP64 factor to
    the slicing level (n = 5000: two sliced trailing updates), keeps the
    strict upper triangle zero, and fills the diagonal-block inverses as
    cfmi_potrf does."""
    import ctypes

    import torch
    from cosmoforge_amd import engine

    n = 5000
    rng = np.random.default_rng(40 + s)
    X = rng.standard_normal((n, n // 2))
    A = X @ X.T / n + np.eye(n)                     # SPD, cond ~ 10
    d = 1.0 / np.sqrt(np.diag(A))
    A = d[:, None] * A * d[None, :]
    out = {}
    for mode in ("fp64", "s8"):
        C = ctx.upload_cm(A)
        U = ctx.empty(n, n)
        info = ctypes.c_int(0)
        if mode == "fp64":
            rc = ctx.lib.cfmi_potrf(ctx.h, engine.F64, n, engine.P(C), n, engine.P(U), n, ctypes.byref(info))
        else:
            rc = ctx.lib.cfmi_potrf_s8(ctx.h, s, n, engine.P(C), n, engine.P(U), n, ctypes.byref(info))
        assert rc == 0 and info.value == 0
        out[mode] = (engine.to_numpy_cm(C), engine.to_numpy_cm(U))
        del C, U
        torch.cuda.empty_cache()
    (L, U), (Ls, Us) = out["fp64"], out["s8"]
    assert not np.triu(Ls, 1).any()
    err = np.max(np.abs(Ls - L))
    print(f"s={s}: max |L_s8 - L_fp64| = {err:.2e}")
    assert err < (1e-6 if s == 4 else 1e-8)
    np.testing.assert_allclose(Us[:64, :64], np.linalg.inv(Ls[:64, :64]), atol=1e-12)
    bad = ctx.lib.cfmi_potrf_s8(ctx.h, 6, 10, engine.P(ctx.empty(10, 10)), 10, None, 10, None)
    assert bad < 0 and b"slices" in ctx.lib.cfmi_last_error(ctx.h)
