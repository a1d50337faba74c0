"""Host-side logic of the banded whitening (CPU only).

``engine.whiten_band`` is run with a NumPy emulator in place of
``cfmi_gemm_ex``; the emulator asserts that every triangular K-trim the
band code requests skips only zeros, and applies the packed row-band output
mapping.  The result must equal the packed lower triangle of W dC W^T
computed directly, for one band and for several."""

import math

import numpy as np
import pytest
import torch

from cosmoforge_amd import engine, pipeline


def _view(flat, off, ld, rows, cols):
    """Column-major rows x cols matrix at flat[off:] with leading dim ld."""
    idx = off + np.arange(rows)[:, None] + ld * np.arange(cols)[None, :]
    return flat[idx]


def fake_gemm_ex(ctx, ta, tb, m, n, k, alpha, A, offa, lda, B, offb, ldb, beta, C, offc=0, ldc=1, *, sA=0, sB=0,
                 sC=0, batch=1, flags=0, tri_a=0, tri_b=0, out_mode=0, row0=0, col0=0, pbase=0, pdiag=1.0,
                 poff=1.0):
    if m <= 0 or n <= 0:
        return
    a_flat = A.detach().reshape(-1).numpy()
    b_flat = B.detach().reshape(-1).numpy()
    c_flat = C.detach().reshape(-1).numpy()  # shares memory with C
    for z in range(batch):
        opA = _view(a_flat, offa + z * sA, lda, k, m).T if ta else _view(a_flat, offa + z * sA, lda, m, k)
        opB = _view(b_flat, offb + z * sB, ldb, n, k).T if tb else _view(b_flat, offb + z * sB, ldb, k, n)
        mm = np.arange(m)[:, None]
        kk = np.arange(k)[None, :]
        if flags & 1:
            assert np.all(opA[kk > mm + tri_a] == 0), "A-lower trim skips nonzeros"
        if flags & 2:
            assert np.all(opA[kk < mm + tri_a] == 0), "A-upper trim skips nonzeros"
        kk2 = np.arange(k)[:, None]
        nn = np.arange(n)[None, :]
        if flags & 4:
            assert np.all(opB[kk2 > nn + tri_b] == 0), "B-upper trim skips nonzeros"
        if flags & 8:
            assert np.all(opB[kk2 < nn + tri_b] == 0), "B-lower trim skips nonzeros"
        prod = alpha * (opA @ opB)
        if out_mode == 0:
            for j in range(n):
                base = offc + z * sC + j * ldc
                c_flat[base:base + m] = prod[:, j] + (beta * c_flat[base:base + m] if beta else 0.0)
        else:
            assert out_mode == 2
            for r in range(m):
                i = row0 + r
                for cidx in range(n):
                    c = col0 + cidx
                    if c > i:
                        continue
                    v = prod[r, cidx] * (pdiag if i == c else poff)
                    p = offc + z * sC + i * (i + 1) // 2 + c - pbase
                    c_flat[p] = v + (beta * c_flat[p] if beta else 0.0)


class FakeCtx:
    device = torch.device("cpu")


def _problem(rng, nf, nb):
    n = 3 * nf
    L = np.tril(rng.standard_normal((n, n))) + 4 * np.eye(n)
    W = np.linalg.inv(L)                                  # lower triangular
    W[np.triu_indices(n, 1)] = 0.0
    blocks = []
    for ra, ma, rb, mb in [(0, nf, 0, nf), (nf, 2 * nf, nf, 2 * nf), (nf, 2 * nf, 0, nf)]:
        D = rng.standard_normal((nb, mb, ma))             # torch (nb, mb, ma) = col-major ma x mb
        if ra == rb:
            D = D + np.transpose(D, (0, 2, 1))
        blocks.append((ra, ma, rb, mb, D))
    return n, W, blocks


def _direct(n, W, ra, ma, rb, mb, Db):
    dC = np.zeros((n, n))
    Dm = Db.T                                              # ma x mb
    dC[ra:ra + ma, rb:rb + mb] += Dm
    if not (ra == rb and ma == mb):
        dC[rb:rb + mb, ra:ra + ma] += Dm.T
    V = W @ dC @ W.T
    out = []
    for i in range(n):
        for c in range(i + 1):
            out.append(V[i, c] * (engine.SQRT_HALF if i == c else 1.0))
    return np.array(out)


@pytest.mark.parametrize("bands,zcap", [([(0, 12)], None), ([(0, 5), (5, 12)], None),
                                        ([(0, 3), (3, 4), (4, 9), (9, 12)], None), ([(0, 5), (5, 12)], 40)])
def test_whiten_band_matches_direct(monkeypatch, bands, zcap):
    rng = np.random.default_rng(5)
    n, W, blocks = _problem(rng, 4, 3)
    monkeypatch.setattr(engine, "gemm_ex", fake_gemm_ex)
    Wt = torch.tensor(np.ascontiguousarray(W.T))          # col-major n x n
    for ra, ma, rb, mb, D in blocks:
        blk = engine.Block(ra, ma, rb, mb, torch.tensor(D))
        got = []
        for r0, r1 in bands:
            X = torch.zeros((D.shape[0], engine.band_size(r0, r1)), dtype=torch.float64)
            engine.whiten_band(FakeCtx(), Wt, n, blk, X, r0, r1, zcap=zcap)
            got.append(X.numpy())
        got = np.concatenate(got, axis=1)
        for b in range(D.shape[0]):
            want = _direct(n, W, ra, ma, rb, mb, D[b])
            np.testing.assert_allclose(got[b], want, rtol=1e-12, atol=1e-12)


def test_band_gram_is_half_trace():
    """sum over bands of <X_i, X_j> = 1/2 Tr[V_i V_j] for symmetric V."""
    rng = np.random.default_rng(9)
    n = 7
    Vs = [rng.standard_normal((n, n)) for _ in range(2)]
    Vs = [v + v.T for v in Vs]
    packs = [np.array([v[i, c] * (engine.SQRT_HALF if i == c else 1.0) for i in range(n) for c in range(i + 1)])
             for v in Vs]
    assert math.isclose(packs[0] @ packs[1], 0.5 * np.trace(Vs[0] @ Vs[1]), rel_tol=1e-13)


def test_band_size_and_plan():
    n = 1000
    assert sum(engine.band_size(a, b) for a, b in [(0, 10), (10, 500), (500, 1000)]) == n * (n + 1) // 2
    for budget in [10**6, 4 * 10**6, 10**9]:
        bands = pipeline.plan_bands(n, 3, 8, budget, align=128)
        assert bands[0][0] == 0 and bands[-1][1] == n
        assert all(b0[1] == b1[0] for b0, b1 in zip(bands, bands[1:]))
        assert all(3 * 8 * engine.band_size(a, b) <= budget for a, b in bands)
        assert all((b - a) % 128 == 0 for a, b in bands[:-1] if b - a > 128)
    assert pipeline.plan_bands(n, 3, 8, 10**9) == [(0, n)]
    with pytest.raises(MemoryError):
        pipeline.plan_bands(n, 3, 8, 3 * 8 * 500)


def _brute_flops(n, blk, nb, bands):
    ra, ma, rb, mb = blk.ra, blk.ma, blk.rb, blk.mb
    tot = 0
    for r0, r1 in bands:
        full = r0 == 0 and r1 == n
        for i in range(max(r0, ra), r1):
            tot += 2 * mb * min(i - ra + 1, ma)
            for c in range(rb, i + 1):
                tot += 2 * min(c - rb + 1, mb)
        if blk.sym:
            continue
        if full:
            for i in range(rb, n):
                for c in range(ra, i + 1):
                    tot += 2 * min(i - rb + 1, mb)
        else:
            for j in range(max(r0, rb), r1):
                tot += 2 * ma * min(j - rb + 1, mb)
                for c in range(ra, j + 1):
                    tot += 2 * min(c - ra + 1, ma)
    return tot * nb


@pytest.mark.parametrize("bands", [None, [(0, 7), (7, 20), (20, 30)]])
def test_whitening_flops_brute_force(bands):
    n, nf = 30, 10
    blocks = [engine.Block(0, nf, 0, nf), engine.Block(nf, 2 * nf, nf, 2 * nf), engine.Block(nf, 2 * nf, 0, nf)]
    want = sum(_brute_flops(n, b, 3, bands or [(0, n)]) for b in blocks)
    assert pipeline.whitening_flops(n, blocks, 3, bands) == pytest.approx(want, rel=1e-12)


# --- three-piece BF16 whitening (engine.whiten_band_b3) ----------------------

FRAC = (0.5, 0.3, 0.2)   # the emulated "pieces" of a value: they sum to it


def _rows(flat, off, ld, ps, rows, cols):
    """k-contiguous rows x cols operand: sum over the three planes of
    flat[off + p ps + r ld + c]."""
    idx = off + np.arange(rows)[:, None] * ld + np.arange(cols)[None, :]
    return flat[idx] + flat[idx + ps] + flat[idx + 2 * ps]


def _flat(t):
    """The storage from t's first element on, as the device pointer P(t) sees it
    (strided views included)."""
    t = t.detach()
    total = t.untyped_storage().nbytes() // t.element_size() - t.storage_offset()
    return torch.as_strided(t, (total,), (1,)).numpy()


def fake_gemm_b3(ctx, mode, m, n, k, alpha, A, offa, lda, sA, pA, B, offb, ldb, sB, pB, beta, C, offc=0, ldc=0,
                 sC=0, pC=0, *, batch=1, flags=0, tri_a=0, tri_b=0, row0=0, col0=0, pbase=0, pdiag=1.0):
    """NumPy stand-in for cfmi_gemm_b3 on float64 'planes'; asserts that the
    K-trims skip only zeros."""
    if m <= 0 or n <= 0:
        return
    a_flat, b_flat, c_flat = _flat(A), _flat(B), _flat(C)
    assert lda >= k and ldb >= k
    for z in range(batch):
        Am = _rows(a_flat, offa + z * sA, lda, pA, m, k)
        Bm = _rows(b_flat, offb + z * sB, ldb, pB, n, k)
        if flags & 1:
            assert np.all(Am[np.arange(k)[None, :] > np.arange(m)[:, None] + tri_a] == 0), "A trim skips nonzeros"
        if flags & 4:
            assert np.all(Bm[np.arange(k)[None, :] > np.arange(n)[:, None] + tri_b] == 0), "B trim skips nonzeros"
        prod = alpha * (Am @ Bm.T)
        if mode == 0:
            for j in range(n):
                base = offc + z * sC + j * ldc
                c_flat[base:base + m] = prod[:, j] + (beta * c_flat[base:base + m] if beta else 0.0)
        elif mode == 1:
            assert ldc >= n and pC >= m * ldc
            for r in range(m):
                base = offc + z * sC + r * ldc
                for p in range(3):
                    c_flat[base + p * pC: base + p * pC + n] = FRAC[p] * prod[r]
        else:
            for r in range(m):
                i = row0 + r
                for cidx in range(n):
                    c = col0 + cidx
                    if c > i:
                        continue
                    v = prod[r, cidx] * (pdiag if i == c else 1.0)
                    q = offc + z * sC + i * (i + 1) // 2 + c - pbase
                    c_flat[q] = v + (beta * c_flat[q] if beta else 0.0)


def _as_planes(a):
    return torch.tensor(np.stack([f * a for f in FRAC]))


@pytest.mark.parametrize("bands,zcap", [([(0, 12)], None), ([(0, 5), (5, 12)], None),
                                        ([(0, 3), (3, 4), (4, 9), (9, 12)], None), ([(0, 5), (5, 12)], 40)])
def test_whiten_band_b3_matches_direct(monkeypatch, bands, zcap):
    """Same bands as the FP64 layout test, on the three-piece operand
    layout: row-major W planes (padded pitch), (3, nbins, mb, ma) stacks,
    their transposes for the a != b blocks, three-piece Z."""
    rng = np.random.default_rng(5)
    n, W, blocks = _problem(rng, 4, 3)
    monkeypatch.setattr(engine, "_gemm_b3", fake_gemm_b3)
    monkeypatch.setattr(engine, "_plane_buffer", lambda ctx, k: torch.full((k,), np.nan, dtype=torch.float64))
    ldw = engine.planes_ld(n)
    Wpad = np.zeros((n, ldw))
    Wpad[:, :n] = W
    Wp = _as_planes(Wpad)                                 # (3, n, ldw)
    for ra, ma, rb, mb, D in blocks:
        mirror = _as_planes(np.ascontiguousarray(np.transpose(D, (0, 2, 1))))
        blk = engine.Block(ra, ma, rb, mb, _as_planes(D), None if (ra == rb and ma == mb) else mirror)
        got = []
        for r0, r1 in bands:
            X = torch.zeros((D.shape[0], engine.band_size(r0, r1)), dtype=torch.float64)
            engine.whiten_band_b3(FakeCtx(), Wp, n, blk, X, r0, r1, zcap=zcap)
            got.append(X.numpy())
        got = np.concatenate(got, axis=1)
        for b in range(D.shape[0]):
            want = _direct(n, W, ra, ma, rb, mb, D[b])
            np.testing.assert_allclose(got[b], want, rtol=1e-12, atol=1e-12)


def test_whiten_band_b3_needs_mirror_for_partial_bands(monkeypatch):
    rng = np.random.default_rng(1)
    n, W, blocks = _problem(rng, 4, 2)
    monkeypatch.setattr(engine, "_gemm_b3", fake_gemm_b3)
    monkeypatch.setattr(engine, "_plane_buffer", lambda ctx, k: torch.zeros(k, dtype=torch.float64))
    Wp = _as_planes(np.pad(W, ((0, 0), (0, engine.planes_ld(n) - n))))
    ra, ma, rb, mb, D = blocks[2]
    blk = engine.Block(ra, ma, rb, mb, _as_planes(D))
    X = torch.zeros((2, engine.band_size(5, 12)), dtype=torch.float64)
    with pytest.raises(ValueError, match="transposed stack"):
        engine.whiten_band_b3(FakeCtx(), Wp, n, blk, X, 5, 12)
    X = torch.zeros((2, engine.band_size(0, 12)), dtype=torch.float64)
    engine.whiten_band_b3(FakeCtx(), Wp, n, blk, X, 0, 12)   # one full band reuses W D: no mirror needed


def test_bins_from_params_rules():
    """Bin construction as ``cosmocore/bins.py:80-140``: uniform delta_ell
    bins from lmin (remainder dropped), explicit bins sorted with l < 2
    dropped, and the overlap / order / empty errors."""
    from types import SimpleNamespace as NS

    from cosmoforge_amd.core import bins_from_params

    def P(**k):
        base = dict(bin_lmins=None, bin_lmaxs=None, lmin=2, lmax=10, delta_ell=3)
        base.update(k)
        return NS(**base)

    assert bins_from_params(P()) == [(2, 4), (5, 7), (8, 10)]
    assert bins_from_params(P(lmax=11)) == [(2, 4), (5, 7), (8, 10)]
    assert bins_from_params(P(delta_ell=1, lmax=4)) == [(2, 2), (3, 3), (4, 4)]
    assert bins_from_params(P(bin_lmins=[10, 0, 2], bin_lmaxs=[12, 1, 9])) == [(2, 9), (10, 12)]
    with pytest.raises(ValueError, match="overlap"):
        bins_from_params(P(bin_lmins=[2, 5], bin_lmaxs=[6, 9]))
    with pytest.raises(ValueError, match="lmin=7 > lmax=5"):
        bins_from_params(P(bin_lmins=[7], bin_lmaxs=[5]))
    with pytest.raises(ValueError, match="No valid bins"):
        bins_from_params(P(bin_lmins=[0], bin_lmaxs=[1]))
    with pytest.raises(ValueError, match="same length"):
        bins_from_params(P(bin_lmins=[2, 3], bin_lmaxs=[4]))


def _group_tile_plan(m, n, batch, big, tri, gm=2, gn=8):
    """Host mirror of the grouped launches of cfmi_gemm_b3 (2 x 8 groups)
    and of batched cfmi_gemm_ex (1 x 16): sizes as set in the entry points,
    mapping as group_tile."""
    bt = 256 if big else 128
    tm, tn = -(-m // bt), -(-n // bt)
    gtm, gtn = -(-tm // gm), -(-tn // gn)
    gtmp = gtm | 1
    ng = gtmp * gtn * batch
    nwg = -(-ng // 8) * 8 * gm * gn

    def tile(x):
        G = gm * gn
        slot = x >> 3
        gg = (slot // G) * 8 + (x & 7)
        if gg >= ng:
            return None
        w, gi, rest = slot % G, gg % gtmp, gg // gtmp
        if gi >= gtm:
            return None
        ti, tj = gi * gm + w % gm, (rest % gtn) * gn + w // gm
        if ti >= tm or tj >= tn:
            return None
        return rest // gtn, (tm - 1 - ti if tri & 1 else ti), (tn - 1 - tj if tri & 4 else tj)

    return tm, tn, nwg, tile


@pytest.mark.parametrize("m,n,batch", [(5952, 5952, 46), (3968, 3968, 46), (1984, 5952, 46), (16128, 16128, 3),
                                       (1, 1, 1), (129, 7000, 5), (1100, 1100, 3), (8064, 257, 7)])
@pytest.mark.parametrize("gm,gn", [(2, 8), (1, 16)])
def test_grouped_tile_order_reaches_every_tile_once(m, n, batch, gm, gn):
    """The XCD-grouped order is a bijection onto (batch, row tile, column
    tile) for both tile sizes and trims, and keeps the eight XCDs' shares of
    the square row-trimmed whitening products (work ~ row index + 1) within
    6% of each other."""
    for big in (0, 1):
        for tri in (0, 1, 4, 5):
            tm, tn, nwg, tile = _group_tile_plan(m, n, batch, big, tri, gm, gn)
            seen, load = set(), [0] * 8
            for x in range(nwg):
                t = tile(x)
                if t is None:
                    continue
                z, ti, tj = t
                assert 0 <= z < batch and 0 <= ti < tm and 0 <= tj < tn
                assert t not in seen
                seen.add(t)
                load[x & 7] += ti + 1
            assert len(seen) == tm * tn * batch
            if tri & 1 and m == n and m >= 3968:     # the trimmed whitening products
                assert max(load) / min(load) < 1.06


def test_smoothing_type_normalisation():
    """Names in any case, the deprecated 'cosine' alias and the legacy
    integer codes 0-3 (``settings.py:353-380``); anything else raises."""
    import warnings

    from cosmoforge_amd.core import read_params

    cases = {"Gaussian": "gaussian", "NONE": "none", " cosine_npipe ": "cosine_npipe", "file": "file"}
    for v, want in cases.items():
        assert read_params({"smoothing_type": v}).smoothing_type == want
    for v, want in {"cosine": "cosine_legacy", 0: "none", 1: "gaussian", 2: "cosine_legacy", 3: "file"}.items():
        with warnings.catch_warnings(record=True) as w:
            warnings.simplefilter("always")
            assert read_params({"smoothing_type": v}).smoothing_type == want
        assert any(issubclass(x.category, DeprecationWarning) for x in w)
    for bad in (4, -1, 1.5, "cos", "gauss", True):
        with pytest.raises(ValueError):
            read_params({"smoothing_type": bad})


def test_default_beams_follow_smoothing_type():
    """Without beam=: 'none' is unity, 'cosine_legacy' / 'cosine_npipe' the
    raised cosine from nside / 1 to 3 nside in all three rows (the window of
    the reference's signal fixture, coswinbeam(4, 4, 12)), 'gaussian' the
    healpy convention exp(-l(l+1) s^2 / 2) with exp(2 s^2) on E and B, and
    an integer code selects the same beam as its name."""
    import math
    import warnings

    from cosmoforge_amd.core import read_params, resolve_beam

    L = 16
    one = resolve_beam(read_params({"nside": 4, "smoothing_type": "none"}), None, L)
    assert np.array_equal(one, np.ones((3, L + 1)))
    cl = resolve_beam(read_params({"nside": 4, "smoothing_type": "cosine_legacy"}), None, L)
    assert np.array_equal(cl[0], cl[1]) and np.array_equal(cl[0], cl[2])
    assert np.all(cl[0, :5] == 1.0) and np.all(cl[0, 12:] == 0.0)
    np.testing.assert_allclose(cl[0, 8], 0.5, atol=1e-15)                 # half-way between 4 and 12
    assert np.all(np.diff(cl[0, 4:13]) < 0)
    cn = resolve_beam(read_params({"nside": 4, "smoothing_type": "cosine_npipe"}), None, L)
    assert cn[0, 1] == 1.0 and cn[0, 2] < 1.0 and cn[0, 12] == 0.0
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        assert np.array_equal(resolve_beam(read_params({"nside": 4, "smoothing_type": 2}), None, L), cl)
    g = resolve_beam(read_params({"nside": 4, "smoothing_type": "gaussian", "fwhmarcmin": 60.0}), None, L)
    sigma = math.radians(1.0) / math.sqrt(8.0 * math.log(2.0))
    ell = np.arange(L + 1)
    np.testing.assert_allclose(np.log(g[0]), -0.5 * ell * (ell + 1) * sigma**2, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(g[1] / g[0], math.exp(2 * sigma**2), rtol=1e-14)
    np.testing.assert_array_equal(g[1], g[2])
    # an injected beam wins over smoothing_type
    inj = np.full((4, L + 3), 0.5)
    assert np.array_equal(resolve_beam(read_params({"smoothing_type": "gaussian"}), inj, L), inj[:3, : L + 1])
