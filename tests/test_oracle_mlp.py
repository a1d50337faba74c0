"""The MLP oracle (CPU): gradient formulas against central finite
differences in float64, BF16 rounding, parameter layout."""

import numpy as np

from oracle import mlp as M


def test_bf16_round_is_nearest_even():
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-8, -2.5, 3.14159, 1e-30, 65504.0], dtype=np.float32)
    r = M.bf16_round(x)
    bits = r.astype(np.float32).view(np.uint32)
    assert np.all(bits & 0xFFFF == 0)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 4 * 2**-8     # ties to even
    ulp = 2.0 ** (np.floor(np.log2(np.abs(x.astype(np.float64)))) - 7)
    assert np.all(np.abs(r - x) <= ulp / 2 * (1 + 1e-12) + 1e-45)


def test_pack_unpack_roundtrip():
    p = M.init_params(np.random.default_rng(0))
    assert p.size == M.NPAR == 132865
    np.testing.assert_array_equal(M.pack(M.unpack(p)), p)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = M.init_params(rng)
    X = rng.uniform(-1, 1, (7, 3))
    T = rng.standard_normal(7)
    _, dy, cache = M.mse(p, X, T, bf16=False)
    g = M.backward(cache, dy, bf16=False)
    picks = np.concatenate([rng.choice(np.arange(a, b), min(6, b - a), replace=False) for a, b in
                            [(0, 768), (768, 1024), (1024, 66560), (66560, 66816), (66816, 132352),
                             (132352, 132608), (132608, 132864), (132864, 132865)]])
    h = 1e-6
    for i in picks:
        e = np.zeros_like(p)
        e[i] = h
        fd = (M.mse(p + e, X, T)[0] - M.mse(p - e, X, T)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])
