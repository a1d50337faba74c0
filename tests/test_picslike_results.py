"""``ParameterGrid`` / ``LikelihoodResult`` host logic (CPU;
``picslike/likelihood_result.py:33-420``, ``parameter_grid.py:287-505``):
marginals and highest-density intervals of a separable Gaussian grid
against their closed forms, best fit, save/load round trip and errors."""

from types import SimpleNamespace

import numpy as np
import pytest

from cosmoforge_amd.picslike import LikelihoodResult, ParameterGrid


def _grid():
    a = np.linspace(-3.0, 3.0, 121)
    b = np.linspace(0.0, 1.0, 5)
    spec = {p: {} for p in __import__("itertools").product(a, b)}
    return ParameterGrid(SimpleNamespace(nside=4), {"a": a, "b": b}, theoretical_spectra=spec), a, b


def test_grid_helpers():
    g, a, b = _grid()
    assert len(g) == g.get_total_points() == a.size * b.size
    assert g.get_grid_shape() == (a.size, b.size)
    pt = (a[7], b[3])
    assert g.get_parameter_index(pt) == 7 * b.size + 3      # Cartesian product order, last fastest
    assert pt in g and list(iter(g))[0] == (a[0], b[0])
    assert g.get_parameter_dict(pt) == {"a": a[7], "b": b[3]}
    with pytest.raises(ValueError, match="not found in grid"):
        g.get_parameter_index((9.0, 9.0))
    with pytest.raises(ValueError, match="expected 2"):
        g.get_parameter_dict((1.0,))
    with pytest.raises(KeyError, match="No theoretical spectrum"):
        ParameterGrid(SimpleNamespace(nside=4), {"a": a}, theoretical_spectra={}, _require_spectra=False) \
            .get_spectrum((a[0],))
    with pytest.raises(ValueError, match="Missing theoretical spectra"):
        ParameterGrid(SimpleNamespace(nside=4), {"a": a}, theoretical_spectra={})


def test_marginals_intervals_and_best_fit(tmp_path):
    g, a, b = _grid()
    sig, mu = 0.5, 0.4
    pts = np.array(g.grid_points)
    chi2 = ((pts[:, 0] - mu) / sig) ** 2 + 0.1 * pts[:, 1]
    r = LikelihoodResult(g, chi2, -0.5 * chi2)
    assert r.get_best_fit() == {"a": a[np.argmin(np.abs(a - mu))], "b": 0.0}
    assert r.get_chi_squared_minimum() == pytest.approx(chi2.min())
    assert r.get_maximum_likelihood() == 1.0
    m = r.get_marginalized_likelihood("a")
    ref = np.exp(-0.5 * ((a - mu) / sig) ** 2)
    np.testing.assert_allclose(m, ref / ref.sum(), rtol=1e-12)
    assert m.sum() == pytest.approx(1.0)
    lo, hi = r.get_confidence_intervals(0.68)["a"]
    da = a[1] - a[0]
    assert abs(lo - (mu - 0.994 * sig)) <= da and abs(hi - (mu + 0.994 * sig)) <= da
    lo95, hi95 = r.get_confidence_intervals(0.95)["a"]
    assert abs(hi95 - lo95 - 2 * 1.96 * sig) <= 2 * da
    s = r.get_summary_statistics()
    assert s["total_parameter_points"] == len(g) and s["parameter_names"] == ["a", "b"]
    with pytest.raises(ValueError, match="not found in grid"):
        r.get_marginalized_likelihood("c")
    path = tmp_path / "res" / "mean.npz"
    r.save(path)
    r2 = LikelihoodResult.load(path)
    np.testing.assert_array_equal(r2.chi_squared_values, r.chi_squared_values)
    np.testing.assert_array_equal(r2.log_likelihood_values, r.log_likelihood_values)
    assert r2.get_best_fit() == r.get_best_fit()
    assert r2.get_confidence_intervals(0.68) == r.get_confidence_intervals(0.68)
    with pytest.raises(FileNotFoundError):
        LikelihoodResult.load(tmp_path / "missing.npz")
    with pytest.raises(ValueError, match="doesn't match grid size"):
        LikelihoodResult(g, chi2[:-1], -0.5 * chi2)
