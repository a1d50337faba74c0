"""The ``Filter`` record and its constructors (CPU; ``cosmocore/filters.py``
semantics): truncation knobs, projector detection, deprojection rank,
intersection as complement of the union, and the guards."""

import warnings

import numpy as np
import pytest

from cosmoforge_amd.filters import MIN_RANGE_EPSILON, Filter, active_pixel_index, mask_fingerprint


def _rng():
    return np.random.default_rng(2)


def test_from_operator_graded_and_rank_knobs():
    rng = _rng()
    n = 40
    Q1, _ = np.linalg.qr(rng.standard_normal((n, n)))
    Q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.r_[np.linspace(3.0, 1.0, 30), np.full(10, 1e-9)]
    F = Q1 @ np.diag(s) @ Q2.T
    f = Filter.from_operator(F, range_epsilon=1e-6)
    assert f.rank == 30 and not f.is_projector and f.scaling is not None
    # F = U Sigma W^T on the kept part
    np.testing.assert_allclose(f.U @ np.diag(f.sigma) @ f.W.T, Q1[:, :30] @ np.diag(s[:30]) @ Q2[:, :30].T,
                               atol=1e-12)
    assert Filter.from_operator(F, range_rank=30).rank == 30
    with pytest.raises(ValueError, match="exactly one truncation knob"):
        Filter.from_operator(F)
    with pytest.raises(ValueError, match="exactly one truncation knob"):
        Filter.from_operator(F, range_epsilon=1e-3, range_rank=3)
    with pytest.raises(ValueError, match="MIN_RANGE_EPSILON"):
        Filter.from_operator(F, range_rank=35)
    with pytest.warns(UserWarning, match="clipped"):
        Filter.from_operator(F, range_epsilon=1e-12)
    with pytest.raises(ValueError, match="square"):
        Filter.from_operator(F[:, :5], range_rank=2)


def test_projector_detection_and_deprojection():
    rng = _rng()
    n = 30
    T = rng.standard_normal((n, 3))
    f = Filter.from_deprojection(T)
    assert f.is_projector and f.rank == n - 3 and f.U is f.W and f.scaling is None
    assert np.abs(f.W.T @ T).max() < 1e-12                    # templates annihilated
    np.testing.assert_allclose(f.W.T @ f.W, np.eye(n - 3), atol=1e-12)
    Pm = f.W @ f.W.T
    g = Filter.from_operator(Pm, range_epsilon=1e-6)
    assert g.is_projector and g.rank == n - 3 and g.U is g.W
    # a rotation has unit singular values but is not a projector
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        Qr, _ = np.linalg.qr(rng.standard_normal((n, n)))
        assert not Filter.from_operator(Qr, range_rank=n).is_projector
    with pytest.raises(ValueError, match="rank-0"):
        Filter.from_deprojection(rng.standard_normal((n, n + 2)))


def test_from_subspace_and_intersection():
    rng = _rng()
    n = 25
    W, _ = np.linalg.qr(rng.standard_normal((n, 7)))
    f = Filter.from_subspace(W)
    assert f.rank == 7 and f.is_projector
    with pytest.raises(ValueError, match="orthonormal"):
        Filter.from_subspace(2 * W)
    with pytest.raises(ValueError, match="cannot have"):
        Filter.from_subspace(rng.standard_normal((3, 5)))
    T1, T2 = rng.standard_normal((n, 2)), rng.standard_normal((n, 3))
    f1, f2 = Filter.from_deprojection(T1), Filter.from_deprojection(T2)
    g = f1 & f2
    assert g.rank == n - 5
    assert np.abs(g.W.T @ np.column_stack([T1, T2])).max() < 1e-12
    with pytest.raises(ValueError, match="Q_removed"):
        f1 & f
    with pytest.raises(AttributeError):
        f1.rank = 3


def test_fingerprint_follows_active_ordering():
    m = np.zeros((48, 3))
    m[[1, 5, 9], 0] = 1
    m[[2, 3], 1:] = 1
    idx = active_pixel_index(m)
    assert list(idx) == [1, 5, 9, 48 + 2, 48 + 3, 96 + 2, 96 + 3]
    m2 = m.copy()
    m2[7, 0] = 1
    assert mask_fingerprint(m) != mask_fingerprint(m2)
    assert mask_fingerprint(m) == mask_fingerprint(m.copy())
