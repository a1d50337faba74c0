"""``Spectra``: drop-in for ``qube.Spectra`` (QML power spectra) on the
pixel-space path.

Follows ``src/cosmoforge.qube/qube/spectra.py``: constructor :205-334
(``mask``, ``noise_cov*``, ``cls_data``, ``fiducial_cls`` and ``beam`` are
refused together with ``fisher=``), ``run`` :1220-1290, the traditional
estimator :1124-1186 (q_b = 1/2 y^T dC_b y - nb_b, nb_b = 1/2 Tr[C^-1 N C^-1
dC_b], y = C^-1 x), the F^-1 normalisation :607-661,1340-1356 and the
getters :1358-1733.  The Fisher instance's ``precision`` applies: in
"mixed", y comes from FP32 solves refined against the FP64 C, and
C^-1 N C^-1 and the contractions run on FP32 inputs with FP64 accumulation.
"""

from __future__ import annotations

import numpy as np

from . import engine, in_out, pipeline
from .core import resolve_maps
from .fisher import Fisher


class Spectra:
    def __init__(self, params=None, fisher: Fisher | None = None, basis=None, mask=None, noise_cov1=None,
                 noise_cov2=None, cls_data=None, fiducial_cls=None, beam=None, maps1=None, maps2=None, *,
                 params_file=None, device: int = 0, **kwargs):
        if params is None and params_file is not None:
            params = params_file
        if fisher is not None:
            clash = [k for k, v in dict(mask=mask, noise_cov1=noise_cov1, noise_cov2=noise_cov2,
                                         cls_data=cls_data, fiducial_cls=fiducial_cls, beam=beam).items()
                     if v is not None]
            if clash:
                raise ValueError(f"{', '.join(clash)} cannot be injected together with fisher=: the Fisher "
                                 "instance already owns these inputs")
            self.fisher_instance = fisher
        else:
            self.fisher_instance = Fisher(params, basis=basis, mask=mask, noise_cov1=noise_cov1,
                                          cls_data=cls_data, fiducial_cls=fiducial_cls, beam=beam,
                                          device=device, **kwargs)
        self.params = self.fisher_instance.params
        if params is not None and fisher is not None:
            from .core import read_params

            extra = read_params(params)
            for key in ("nsims", "remove_nb", "inputmapfile1", "inputmapfile2", "outcovmatfile", "outerrfile"):
                setattr(self.params, key, getattr(extra, key))
        self._maps_in = maps1
        self._maps2_in = maps2
        self.maps1 = None
        self.maps2 = None
        self.qml_results = None
        self.qml_noise_bias = None
        self.invfisher = None
        self.fisher_normalized = None
        self.inv_fisher_sqrt = None

    @property
    def ctx(self):
        return self.fisher_instance.ctx

    def setup_maps(self):
        inp = self.fisher_instance.inputs
        self.maps1 = resolve_maps(inp.n, self.params.nsims, self._maps_in, self.params, inp.layout)
        if self.params.do_cross:
            self.maps2 = resolve_maps(inp.n, self.params.nsims, self._maps2_in, self.params, inp.layout, which=2)
        # filter seam for d (ADR-0020): restricted to the filter's range
        flt = self.fisher_instance.device_filter()
        if flt is not None:
            self.maps1 = flt.restrict_data(self.maps1, prefiltered=flt.maps_prefiltered)
            if self.maps2 is not None:
                self.maps2 = flt.restrict_data(self.maps2, prefiltered=flt.maps_prefiltered)

    def setup_fisher_inversion(self):
        F = self.fisher_instance.get_fisher_matrix()
        if F is None:
            raise ValueError("Fisher matrix not available")
        self.fisher_normalized = F.copy()
        self.invfisher = pipeline.inverse_symm(self.ctx, F)
        # F^-1 and the error bars as (nbins, nspectra) columns (spectra.py:651-661)
        in_out.write_out_matrix(self.params.outcovmatfile, self.invfisher)
        if not in_out._falsy(self.params.outerrfile):
            err = np.sqrt(np.diag(self.invfisher))
            nb = len(self.fisher_instance.bins)
            in_out.writecl(self.params.outerrfile, err.reshape(-1, nb).T)

    def run(self):
        f = self.fisher_instance
        if f.fisher is None:
            f.run()
        src = f.derivs
        if f.basis == "harmonic":
            if src is None:
                from .harmonic import HarmonicContraction

                src = f.derivs = HarmonicContraction(f.cov, f.bins, self.params.output_convention,
                                                     cov2=f.cov2 if self.params.do_cross else None,
                                                     precision=f.precision)
            self.setup_maps()
            self.setup_fisher_inversion()
            self.qml_results, self.qml_noise_bias = src.qml(self.maps1, remove_nb=bool(self.params.remove_nb),
                                                            maps2=self.maps2)
            return
        if src is None:
            src = pipeline.make_source(f.cov, f.bins, self.params.output_convention,
                                            pipeline._dtype(f.precision))
        self.setup_maps()
        self.setup_fisher_inversion()
        if self.params.do_cross:
            q, nb = pipeline.qml_cross(f.cov, f.cov2, src, self.maps1, self.maps2)
        else:
            q, nb = pipeline.qml(f.cov, src, self.maps1, remove_nb=bool(self.params.remove_nb),
                                 precision=f.precision)
        self.qml_results = q
        self.qml_noise_bias = nb

    compute = run

    # -- getters ------------------------------------------------------------
    _MODES = ("deconvolved", "decorrelated", "convolved")

    def _check_mode(self, mode):
        if mode not in self._MODES:
            raise ValueError(f"mode must be one of {self._MODES}, got '{mode}'")

    def _normalize_spectra(self, spectra: np.ndarray) -> np.ndarray:
        if self.invfisher is None:
            raise ValueError("Fisher inversion must be set up first.")
        return spectra @ self.invfisher

    def _inv_fisher_sqrt(self) -> np.ndarray:
        """F^-1/2 from the eigendecomposition, eigenvalues at or below 1e-12 of
        the largest dropped (``spectra.py:663-740``)."""
        if self.inv_fisher_sqrt is None:
            w, v = np.linalg.eigh(self.fisher_normalized)
            inv = np.zeros_like(w)
            ok = w > w.max() * 1e-12
            inv[ok] = 1.0 / np.sqrt(w[ok])
            self.inv_fisher_sqrt = (v * inv) @ v.T
        return self.inv_fisher_sqrt

    def _as_dict(self, arr):
        f = self.fisher_instance
        nb = len(f.bins)
        return {lab: arr[..., k * nb:(k + 1) * nb] for k, lab in enumerate(f.spectra_list)}

    def _make_convolve_theory(self, W: np.ndarray):
        """Callable taking a flat (n_params,) binned theory or a label-keyed
        dict of per-bin arrays; returns W @ theory (``spectra.py:1573-1614``)."""
        labs = self.fisher_instance.spectra_list

        def convolve_theory(cl_theory):
            if isinstance(cl_theory, dict):
                missing = [lab for lab in labs if lab not in cl_theory]
                if missing:
                    raise KeyError(f"convolve_theory dict missing entries for: {missing}. "
                                   f"Expected keys: {labs}")
                cl_theory = np.concatenate([np.asarray(cl_theory[lab]) for lab in labs])
            return W @ np.asarray(cl_theory)

        return convolve_theory

    def get_power_spectra(self, mode: str = "deconvolved", *, as_dict: bool = False):
        """deconvolved y F^-1, decorrelated y F^-1/2, or convolved
        (y, W, convolve_theory) (``spectra.py:1358-1553``)."""
        self._check_mode(mode)
        if self.qml_results is None:
            return None
        if mode == "deconvolved":
            out = self._normalize_spectra(self.qml_results)
        elif mode == "decorrelated":
            out = self.qml_results @ self._inv_fisher_sqrt()
        else:
            W = self.fisher_instance.get_window_matrix()
            if W is None:
                raise ValueError("Window matrix not available from Fisher instance.")
            y = self.qml_results
            return ((self._as_dict(y) if as_dict else y), W, self._make_convolve_theory(W))
        return self._as_dict(out) if as_dict else out

    def get_noise_bias(self):
        if self.qml_noise_bias is None:
            return None
        return self._normalize_spectra(self.qml_noise_bias)

    def get_covariance(self, mode: str = "deconvolved"):
        """F^-1, the identity, or F for the three modes (``spectra.py:1880-1971``)."""
        self._check_mode(mode)
        if mode == "convolved":
            return None if self.fisher_normalized is None else self.fisher_normalized.copy()
        if self.invfisher is None:
            return None
        return self.invfisher.copy() if mode == "deconvolved" else np.eye(self.invfisher.shape[0])

    def get_error_bars(self, mode: str = "deconvolved", *, as_dict: bool = False):
        cov = self.get_covariance(mode)
        if cov is None:
            return None
        err = np.sqrt(np.diag(cov))
        return self._as_dict(err) if as_dict else err

    def convolve_theory(self, cl_theory):
        """W @ binned theory, the expected raw estimate (``spectra.py:2039-2097``)."""
        W = self.fisher_instance.get_window_matrix()
        return None if W is None else W @ np.asarray(cl_theory)

    def convolve_theory_for_inference(self, cl_theory):
        """Expected deconvolved bandpowers of a per-l theory through the
        bandpower window (``spectra.py:1735-1878``): a flat spectrum-major
        vector of nspectra n_ell values, a label-keyed dict, or for one
        spectrum an (n_ell,) or (>= lmax+1,) l-indexed array."""
        f = self.fisher_instance
        if getattr(f, "fisher", None) is None:
            raise ValueError("Fisher matrix has not been computed. Call fisher.run() "
                             "(and spectra.run()) before convolve_theory_for_inference().")
        lmin, lmax = self.params.lmin, self.params.lmax
        n_ell = lmax - lmin + 1
        labs = f.spectra_list

        def perell(arr, what):
            if arr.ndim != 1:
                raise ValueError(f"{what} must be 1D, got shape {arr.shape}")
            if arr.size == n_ell:
                return arr
            if arr.size >= lmax + 1:
                return arr[lmin:lmax + 1]
            raise ValueError(f"{what} length {arr.size} does not match expected "
                             f"{n_ell} (ℓ=lmin..lmax) or {lmax + 1} (ℓ=0..lmax)")

        if isinstance(cl_theory, dict):
            missing = [lab for lab in labs if lab not in cl_theory]
            if missing:
                raise KeyError(f"cl_theory dict missing entries for: {missing}. Expected keys: {labs}")
            vec = np.concatenate([perell(np.asarray(cl_theory[lab], dtype=np.float64), lab) for lab in labs])
        else:
            cl = np.asarray(cl_theory, dtype=np.float64)
            if cl.ndim != 1:
                raise ValueError(f"cl_theory must be 1D, got shape {cl.shape}")
            if len(labs) > 1:
                if cl.size != len(labs) * n_ell:
                    raise ValueError(f"cl_theory length {cl.size} does not match expected "
                                     f"{len(labs) * n_ell} (nspectra·n_ell, spectrum-major) for a "
                                     f"{len(labs)}-spectrum run; or pass a label-keyed dict.")
                vec = cl
            else:
                vec = perell(cl, "cl_theory")
        return f.get_bandpower_window_function() @ vec

    def write_power_spectra(self, mode: str = "deconvolved", filename: str | None = None,
                            include_errors: bool = True) -> None:
        """Mean over simulations as (nbins, nspectra) columns, with the error
        bars in ``{base}_errors.{ext}``; convolved mode writes the mean raw y
        and the window ``{base}_window.{ext}`` instead
        (``spectra.py:2099-2239``).  The default name is
        ``{outclfile base}_{mode}.{ext}``, else ``spectra_{mode}.dat``."""
        self._check_mode(mode)
        if filename is None:
            out = getattr(self.params, "outclfile", None)
            if not in_out._falsy(out):
                base, _, ext = str(out).rpartition(".") if "." in str(out) else (str(out), "", "dat")
                filename = f"{base}_{mode}.{ext}"
            else:
                filename = f"spectra_{mode}.dat"
        filename = str(filename)
        stem, dot, ext = filename.rpartition(".")

        def sibling(tag):
            return f"{stem}_{tag}.{ext}" if dot else f"{filename}_{tag}"

        nb = len(self.fisher_instance.bins)
        if mode == "convolved":
            res = self.get_power_spectra(mode="convolved")
            if res is None:
                return
            y, W, _ = res
            in_out._make_parent(in_out._path(filename))
            np.savetxt(filename, np.mean(y, axis=0), header="Raw QML estimates (y vector, mean across simulations)")
            np.savetxt(sibling("window"), W, header="Window matrix W: <y> = W @ C_true")
            return
        spectra = self.get_power_spectra(mode=mode)
        if spectra is None:
            return
        # spectrum-major vector -> (nbins, nspectra) columns (spectra_io.py:72-97)
        in_out.writecl(filename, np.mean(spectra, axis=0).reshape(-1, nb).T)
        if include_errors:
            err = self.get_error_bars(mode=mode)
            if err is not None:
                in_out.writecl(sibling("errors"), err.reshape(-1, nb).T)

    def get_effective_ells(self, use_midpoint: bool = False):
        """Bin midpoints (``use_midpoint``), else the window centroids of
        ``Fisher.get_effective_ells`` (``spectra.py:1686-1711``)."""
        bins = self.fisher_instance.bins
        if use_midpoint:
            return np.array([0.5 * (a + b) for a, b in bins])
        return self.fisher_instance.get_effective_ells()
