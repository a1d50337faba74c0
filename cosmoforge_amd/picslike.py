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

from . import engine, in_out, pipeline
from .core import Inputs, resolve_basis, resolve_maps


def _blend(param: dict, fid: dict, lmin: int, lmax: int) -> dict:
    """Fiducial spectrum outside [lmin, lmax], the point's inside
    (``parameter_grid.py:225-260``)."""
    out = {}
    for key, arr in param.items():
        if key not in fid:
            out[key] = arr.copy()
            continue
        n = min(len(arr), len(fid[key]))
        res = fid[key][:n].copy()
        lo, hi = max(0, lmin), min(n - 1, lmax)
        if lo <= hi:
            res[lo: hi + 1] = arr[lo: hi + 1]
        out[key] = res
    return out


class ParameterGrid:
    """Grid points (Cartesian product of ``parameter_ranges``) and their
    spectra: given in memory, or read per point from
    ``{root_dir}/{root_filename}_{name}{value:.6g}[_...].txt`` and blended
    with ``core_params.fiducialfile`` outside [lmin, lmax]
    (``parameter_grid.py:85-262``)."""

    def __init__(self, core_params, parameter_ranges: dict, theoretical_spectra: dict | None = None,
                 root_dir=None, root_filename=None, lmax_signal=None, _require_spectra: bool = True):
        self.core_params = core_params
        self.lmax_signal = lmax_signal if lmax_signal is not None else 4 * core_params.nside
        self.parameter_names = list(parameter_ranges.keys())
        self.parameter_ranges = deepcopy(parameter_ranges)
        vals = [self.parameter_ranges[k] for k in self.parameter_names]
        self.grid_points = list(itertools.product(*vals)) if vals else []
        if theoretical_spectra is not None:
            self.theoretical_spectra = deepcopy(theoretical_spectra)
        else:
            if root_dir is None or root_filename is None:
                raise ValueError("Either theoretical_spectra or both root_dir and root_filename must be provided")
            self.root_dir, self.root_filename = root_dir, root_filename
            self.read_theoretical_spectra()
        missing = [p for p in self.grid_points if p not in self.theoretical_spectra]
        if missing and _require_spectra:
            raise ValueError(f"Missing theoretical spectra for {len(missing)} grid points")

    def spectrum_file(self, point) -> str:
        tag = "_".join(f"{self.parameter_names[i]}{v:.6g}" for i, v in enumerate(point))
        return f"{self.root_dir}/{self.root_filename}_{tag}.txt"

    def read_theoretical_spectra(self) -> None:
        p = self.core_params
        lmin, lmax = getattr(p, "lmin", None), getattr(p, "lmax", None)
        fid = None
        if getattr(p, "fiducialfile", None) is not None and lmin is not None and lmax is not None:
            fid = in_out.readcl(p.fiducialfile, p, lmax=self.lmax_signal)
        self.theoretical_spectra = {}
        for point in self.grid_points:
            spec = in_out.readcl(self.spectrum_file(point), p, lmax=self.lmax_signal)
            self.theoretical_spectra[point] = spec if fid is None else _blend(spec, fid, lmin, lmax)

    def get_spectrum(self, point):
        try:
            return self.theoretical_spectra[point]
        except KeyError as exc:
            raise KeyError(f"No theoretical spectrum found for parameters: {point}") from exc

    def get_parameter_dict(self, point):
        if len(point) != len(self.parameter_names):
            raise ValueError(f"Parameter point has {len(point)} values, expected {len(self.parameter_names)}")
        return dict(zip(self.parameter_names, point))

    def get_total_points(self) -> int:
        return len(self.grid_points)

    @property
    def n_points(self):
        return len(self.grid_points)

    def get_grid_shape(self) -> tuple:
        return tuple(len(self.parameter_ranges[k]) for k in self.parameter_names)

    def get_parameter_index(self, point) -> int:
        try:
            return self.grid_points.index(point)
        except ValueError as exc:
            raise ValueError(f"Parameter point {point} not found in grid") from exc

    def add_spectrum(self, point, spectrum) -> None:
        self.theoretical_spectra[point] = deepcopy(spectrum)

    def __len__(self) -> int:
        return len(self.grid_points)

    def __iter__(self):
        return iter(self.grid_points)

    def __contains__(self, point) -> bool:
        return point in self.grid_points


class LikelihoodResult:
    """chi2 and ln L over a parameter grid (``likelihood_result.py:33-420``)."""

    def __init__(self, parameter_grid: ParameterGrid, chi_squared_values, log_likelihood_values):
        npts = len(parameter_grid.grid_points)
        if len(chi_squared_values) != npts:
            raise ValueError(f"Chi-squared array length {len(chi_squared_values)} doesn't match grid size {npts}")
        if len(log_likelihood_values) != npts:
            raise ValueError(f"Log-likelihood array length {len(log_likelihood_values)} doesn't match grid size "
                             f"{npts}")
        self.parameter_grid = parameter_grid
        self.chi_squared_values = np.array(chi_squared_values, dtype=np.float64)
        self.log_likelihood_values = np.array(log_likelihood_values, dtype=np.float64)
        # exp(lnL - max lnL), the grid's relative likelihood
        self.likelihood_values = np.exp(self.log_likelihood_values - np.max(self.log_likelihood_values))

    def get_best_fit(self):
        i = int(np.argmax(self.likelihood_values))
        return self.parameter_grid.get_parameter_dict(self.parameter_grid.grid_points[i])

    def get_chi_squared_minimum(self) -> float:
        return float(np.min(self.chi_squared_values))

    def get_maximum_likelihood(self) -> float:
        return float(np.max(self.likelihood_values))

    def get_marginalized_likelihood(self, parameter_name: str) -> np.ndarray:
        """Likelihood summed over the other parameters, normalised to unit sum."""
        grid = self.parameter_grid
        if parameter_name not in grid.parameter_names:
            raise ValueError(f"Parameter '{parameter_name}' not found in grid")
        k = grid.parameter_names.index(parameter_name)
        values = np.asarray(grid.parameter_ranges[parameter_name])
        col = np.array([pt[k] for pt in grid.grid_points])
        marg = np.array([self.likelihood_values[col == v].sum() for v in values])
        tot = marg.sum()
        return marg / tot if tot > 0 else marg

    @staticmethod
    def _interval(values: np.ndarray, like: np.ndarray, level: float):
        """Highest-density set of grid values holding ``level`` of the mass;
        its hull (``likelihood_result.py:270-328``)."""
        tot = like.sum()
        if tot == 0:
            return float(values[0]), float(values[-1])
        order = np.argsort(like)[::-1]
        mass = np.cumsum(like[order]) / tot
        keep = values[order[: int(np.searchsorted(mass, level)) + 1]]
        return float(keep.min()), float(keep.max())

    def get_confidence_intervals(self, confidence_level: float = 0.68) -> dict:
        grid = self.parameter_grid
        return {k: self._interval(np.asarray(grid.parameter_ranges[k]), self.get_marginalized_likelihood(k),
                                  confidence_level) for k in grid.parameter_names}

    def get_summary_statistics(self) -> dict:
        return {"best_fit_parameters": self.get_best_fit(), "chi_squared_minimum": self.get_chi_squared_minimum(),
                "maximum_likelihood": self.get_maximum_likelihood(),
                "confidence_intervals_68": self.get_confidence_intervals(0.68),
                "confidence_intervals_95": self.get_confidence_intervals(0.95),
                "total_parameter_points": len(self.parameter_grid),
                "parameter_names": list(self.parameter_grid.parameter_names)}

    def save(self, output_path) -> None:
        """The grid and the results as a NumPy archive (no pickled objects;
        the reference pickles the object)."""
        from pathlib import Path

        path = Path(output_path)
        path.parent.mkdir(parents=True, exist_ok=True)
        grid = self.parameter_grid
        with open(path, "wb") as fh:
            np.savez(fh, names=np.array(grid.parameter_names, dtype=str),
                     points=np.array(grid.grid_points, dtype=np.float64).reshape(len(grid.grid_points), -1),
                     chi2=self.chi_squared_values, lnl=self.log_likelihood_values,
                     **{f"range_{i}": np.asarray(grid.parameter_ranges[k], dtype=np.float64)
                        for i, k in enumerate(grid.parameter_names)})

    @classmethod
    def load(cls, input_path) -> "LikelihoodResult":
        """A result written by ``save`` (the grid without its spectra)."""
        from pathlib import Path
        from types import SimpleNamespace

        path = Path(input_path)
        if not path.exists():
            raise FileNotFoundError(f"Results file not found: {path}")
        with np.load(path, allow_pickle=False) as d:
            names = [str(x) for x in d["names"]]
            ranges = {k: d[f"range_{i}"] for i, k in enumerate(names)}
            grid = ParameterGrid(SimpleNamespace(nside=0), ranges, theoretical_spectra={}, lmax_signal=0,
                                 _require_spectra=False)
            return cls(grid, d["chi2"], d["lnl"])

    def __repr__(self) -> str:
        return (f"LikelihoodResult(parameters={len(self.parameter_grid.parameter_names)}, "
                f"grid_points={len(self.parameter_grid)}, χ²_min={self.get_chi_squared_minimum():.3f})")


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
        self.maps1 = resolve_maps(self.inputs.n, self.params.nsims, maps1, self.params, self.inputs.layout)
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
        # pixel filter (ADR-0020): N, S and the maps restricted to its range
        self.pixel_filter = kwargs.get("pixel_filter")
        flt = None
        if self.pixel_filter is not None:
            from .fisher import _check_filter
            from .filters import DeviceFilter

            _check_filter(self.pixel_filter, self.inputs)
            flt = DeviceFilter(self.ctx, self.pixel_filter, kwargs.get("noise_prefiltered", False),
                               kwargs.get("maps_prefiltered", False))
            self.maps1 = flt.restrict_data(self.maps1, prefiltered=flt.maps_prefiltered)
        self.cov = pipeline.Covariance(self.ctx, self.inputs, pixel_filter=flt)
        self.likelihood_result = None
        self.simulation_results = None
        self.simulation_index = 0

    @property
    def lmax_signal(self):
        return self.inputs.lmax_signal

    def _compute_likelihood_point(self, param_point):
        cls = self.inputs.cls if param_point is None else self.parameter_grid.get_spectrum(param_point)
        from .core import resolve_cls

        cls = resolve_cls(self.inputs.layout, cls, self.inputs.lmax_signal)
        chi2, logdet = pipeline.likelihood(self.cov, self.maps1, cls)
        return chi2, -0.5 * (chi2 + logdet)

    def setup_parameter_grid(self) -> None:
        """Grid from ``params.parameters`` ({name: [min, max, n]} -> linspace)
        with spectra read from ``root_dir`` / ``root_filename``
        (``picslike.py:409-431``)."""
        names = list(self.params.parameters.keys())
        ranges = {k: np.linspace(*self.params.parameters[k]) for k in names}
        self.parameter_grid = ParameterGrid(self.params, ranges, root_dir=self.params.root_dir,
                                            root_filename=self.params.root_filename,
                                            lmax_signal=self.lmax_signal)

    def compute(self):
        if self.parameter_grid is None and self.params.parameters:
            self.setup_parameter_grid()
        if self.parameter_grid is None:
            raise ValueError("PICSLike needs parameter_grid=, theoretical_spectra= or params.parameters")
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
        if self.likelihood_result is None:
            raise RuntimeError("Likelihood not computed. Call compute() first.")
        return self.likelihood_result

    def set_simulation_index(self, sim_idx: int) -> None:
        """Select the simulation an analysis reports on (``picslike.py:377-407``)."""
        if self.maps1 is not None:
            top = self.maps1.shape[1] - 1
            if sim_idx < 0 or sim_idx > top:
                raise ValueError(f"Simulation index {sim_idx} out of range [0, {top}]")
        self.simulation_index = sim_idx

    def save_results(self, output_path: str) -> None:
        """The mean result at ``output_path`` and every simulation's at
        ``{stem}_sim_XX{suffix}`` beside it (``picslike.py:771-810``)."""
        from pathlib import Path

        if self.likelihood_result is None:
            raise RuntimeError("No results to save. Run compute() first.")
        self.likelihood_result.save(output_path)
        base = Path(output_path)
        for i, r in enumerate(self.simulation_results or []):
            r.save(str(base.parent / f"{base.stem}_sim_{i:02d}{base.suffix}"))
