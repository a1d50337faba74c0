"""``PICSLike``: drop-in for ``picslike.PICSLike`` (exact pixel likelihood).

Follows ``src/cosmoforge.picslike/picslike/``: the constructor
(``picslike.py:160-175``), ``compute`` (:479-546), the per-point likelihood
(:568-664: chi2 = d^T C^-1 d, ln|C|, lnL = -(chi2 + ln|C|)/2), the getters
(:671-810), ``ParameterGrid`` (``parameter_grid.py:85-167``, Cartesian
product order) and ``LikelihoodResult`` (``likelihood_result.py:84-171``).

Per point the device path builds S from that point's spectra, factors
C = S + N once (the reference factors it twice: potrf+potri, then potrf for
the log-determinant) and takes chi2 = |W d|^2 with C^-1 = W^T W.
"""

from __future__ import annotations

import itertools
from copy import deepcopy

import numpy as np

from . import engine, pipeline
from .core import Inputs, resolve_basis, resolve_maps


class ParameterGrid:
    def __init__(self, core_params, parameter_ranges: dict, theoretical_spectra: dict | None = None,
                 root_dir=None, root_filename=None, lmax_signal=None):
        if theoretical_spectra is None:
            raise NotImplementedError("file-backed theory spectra are not part of this build; pass "
                                      "theoretical_spectra={point: {label: C_l}}")
        self.core_params = core_params
        self.lmax_signal = lmax_signal if lmax_signal is not None else 4 * core_params.nside
        self.parameter_names = list(parameter_ranges.keys())
        self.parameter_ranges = deepcopy(parameter_ranges)
        vals = [self.parameter_ranges[k] for k in self.parameter_names]
        self.grid_points = list(itertools.product(*vals)) if vals else []
        self.theoretical_spectra = deepcopy(theoretical_spectra)
        missing = [p for p in self.grid_points if p not in self.theoretical_spectra]
        if missing:
            raise ValueError(f"no theoretical spectra for grid points {missing[:3]}")

    def get_spectrum(self, point):
        return self.theoretical_spectra[point]

    def get_parameter_dict(self, point):
        return dict(zip(self.parameter_names, point))

    @property
    def n_points(self):
        return len(self.grid_points)


class LikelihoodResult:
    def __init__(self, parameter_grid: ParameterGrid, chi_squared_values, log_likelihood_values):
        self.parameter_grid = parameter_grid
        self.chi_squared_values = np.asarray(chi_squared_values)
        self.log_likelihood_values = np.asarray(log_likelihood_values)

    @property
    def likelihood_values(self):
        return np.exp(self.log_likelihood_values - np.max(self.log_likelihood_values))

    def get_best_fit(self):
        i = int(np.argmax(self.likelihood_values))
        return self.parameter_grid.get_parameter_dict(self.parameter_grid.grid_points[i])

    def get_chi_squared_minimum(self) -> float:
        return float(np.min(self.chi_squared_values))


class PICSLike:
    def __init__(self, params=None, basis=None, mask=None, noise_cov1=None, noise_cov2=None, cls_data=None,
                 fiducial_cls=None, beam=None, maps1=None, maps2=None, *, params_file=None,
                 parameter_grid: ParameterGrid | None = None, theoretical_spectra: dict | None = None,
                 device: int = 0, context: engine.Context | None = None, **kwargs):
        if params is None and params_file is not None:
            params = params_file
        if noise_cov2 is not None or maps2 is not None:
            raise NotImplementedError("do_cross is not part of this build")
        self.basis = resolve_basis(basis)
        if cls_data is None and theoretical_spectra:
            cls_data = next(iter(theoretical_spectra.values()))
        self.inputs = Inputs(params, mask, noise_cov1, cls_data if cls_data is not None else fiducial_cls, beam)
        self.params = self.inputs.params
        self.maps1 = resolve_maps(self.inputs.n, self.params.nsims, maps1)
        self.params.nsims = self.maps1.shape[1]
        if parameter_grid is None and theoretical_spectra is not None:
            pts = list(theoretical_spectra)
            parameter_grid = ParameterGrid(self.params, {"point": [p[0] if isinstance(p, tuple) else p
                                                                  for p in pts]},
                                           {(p[0] if isinstance(p, tuple) else p,): v
                                            for p, v in theoretical_spectra.items()},
                                           lmax_signal=self.inputs.lmax_signal)
        self.parameter_grid = parameter_grid
        self.ctx = context if context is not None else engine.Context(device)
        self.cov = pipeline.Covariance(self.ctx, self.inputs)
        self.likelihood_result = None
        self.simulation_results = None

    @property
    def lmax_signal(self):
        return self.inputs.lmax_signal

    def _compute_likelihood_point(self, param_point):
        cls = self.inputs.cls if param_point is None else self.parameter_grid.get_spectrum(param_point)
        from .core import resolve_cls

        cls = resolve_cls(self.inputs.layout, cls, self.inputs.lmax_signal)
        chi2, logdet = pipeline.likelihood(self.cov, self.maps1, cls)
        return chi2, -0.5 * (chi2 + logdet)

    def compute(self):
        if self.parameter_grid is None:
            raise ValueError("PICSLike needs parameter_grid= or theoretical_spectra=")
        pts = self.parameter_grid.grid_points
        ns = self.maps1.shape[1]
        chi2 = np.zeros((len(pts), ns))
        lnl = np.zeros((len(pts), ns))
        for i, p in enumerate(pts):
            chi2[i], lnl[i] = self._compute_likelihood_point(p)
        self.simulation_results = [LikelihoodResult(self.parameter_grid, chi2[:, s], lnl[:, s])
                                   for s in range(ns)]
        self.likelihood_result = LikelihoodResult(self.parameter_grid, chi2.mean(axis=1), lnl.mean(axis=1))

    run = compute

    def _need(self):
        if self.likelihood_result is None:
            raise RuntimeError("Likelihood not computed. Call compute_likelihood_grid() first.")

    def get_chi_squared(self):
        self._need()
        return self.likelihood_result.chi_squared_values

    def get_log_likelihood(self):
        self._need()
        return self.likelihood_result.log_likelihood_values

    def get_best_fit(self):
        self._need()
        return self.likelihood_result.get_best_fit()

    def get_simulation_results(self):
        if self.simulation_results is None:
            raise RuntimeError("Simulation results not available. Call compute() first.")
        return self.simulation_results

    def get_mean_likelihood_result(self):
        self._need()
        return self.likelihood_result
