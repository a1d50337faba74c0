"""``Fisher``: drop-in for ``qube.Fisher`` on the pixel-space path.

Constructor kwargs and getters follow ``src/cosmoforge.qube/qube/fisher.py``
(:141-266 constructor, :622-820 ``run``, :826-1232 getters).  The matrix is
F_ij = 1/2 Tr[C^-1 dC_i C^-1 dC_j] over parameters ordered spectrum-major
(``spectra_list`` order, then bins), beam-smoothed as the reference's.
All device work runs in libcfmi on one GPU.  ``precision="mixed"`` (no
reference counterpart) runs the large products on the int8 matrix cores at
FP64 accuracy (operands as int8 slices, exact int32 sums): the whitening
products of the pixel route (``mixed_gemm="s8"``, the default; "bf16x3" and
"fp32" keep the FP32-accurate variants) and the solves and contractions of
the harmonic route; recurrences, factorisations and F^-1 stay FP64.
"""

from __future__ import annotations

import numpy as np

from . import engine, in_out, pipeline
from .core import Inputs, bins_from_params, resolve_basis


def _check_filter(f, inputs) -> None:
    """The filter against this run's active pixels (``core.py:1090-1111``)."""
    from .filters import mask_fingerprint

    if f.n_pixels != inputs.n:
        raise ValueError(f"pixel_filter was built for {f.n_pixels} pixels but this run has {inputs.n} active "
                         "pixels; rebuild it against this mask")
    if f.fingerprint is not None and f.fingerprint != mask_fingerprint(inputs.mask):
        raise ValueError("pixel_filter was built against a different active-pixel ordering than this run's "
                         "mask; its rows would be permuted. Rebuild it with active_pixel_index(mask) of this mask")


class Fisher:
    def __init__(self, params=None, basis=None, cache_derivatives: bool = False, symmetry_mode=None,
                 mask=None, noise_cov1=None, noise_cov2=None, cls_data=None, fiducial_cls=None, beam=None,
                 *, params_file=None, device: int = 0, context: engine.Context | None = None,
                 precision: str = "fp64", **kwargs):
        if params is None and params_file is not None:
            params = params_file
        if symmetry_mode is not None:
            name = getattr(symmetry_mode, "name", symmetry_mode)
            if not isinstance(name, str):
                raise TypeError(f"symmetry_mode must be SymmetryMode | str | None; got "
                                f"{type(symmetry_mode).__name__}")
            if name.upper() not in ("SYMMETRIC", "DIRECTIONAL"):
                raise ValueError(f"unknown symmetry_mode {symmetry_mode!r}; expected one of "
                                 "['SYMMETRIC', 'DIRECTIONAL']")
        self.basis = resolve_basis(basis)
        self._cache_derivatives = cache_derivatives
        self.inputs = Inputs(params, mask, noise_cov1, cls_data, beam, noise_cov2)
        self.params = self.inputs.params
        self.fiducial_cls = fiducial_cls  # inert on the pixel-space path, as in the reference
        self.ctx = context if context is not None else engine.Context(device)
        self.cov: pipeline.Covariance | None = None
        self.cov2: pipeline.Covariance | None = None   # C2 = S + N2 of a cross analysis
        self.fisher = None
        self.beam_smoothing = None
        self.spectra_list = None
        self.bins = None
        if precision not in pipeline.PRECISIONS:
            raise ValueError(f"precision must be one of {pipeline.PRECISIONS}; got {precision!r}")
        self.precision = precision
        if self.basis == "harmonic" and precision != "fp64" and (self.params.do_cross or
                                                                  kwargs.get("pixel_filter") is not None):
            raise NotImplementedError("basis='harmonic' with precision='mixed' covers auto analyses without "
                                      "a pixel filter")
        if self.params.do_cross and precision != "fp64":
            raise ValueError("precision='mixed' covers the auto-spectrum path; do_cross runs in fp64")
        self.derivs = None  # device derivative stacks (reused by Spectra when resident)
        self.timings = {}
        self.timer = kwargs.get("timer")
        self.launch_timer = kwargs.get("launch_timer")
        # pixel filter (ADR-0020): the analysis runs in the filter's range
        self.pixel_filter = kwargs.get("pixel_filter")
        self.maps_prefiltered = bool(kwargs.get("maps_prefiltered", False))
        self.noise_prefiltered = bool(kwargs.get("noise_prefiltered", False))
        self._device_filter = None
        if self.pixel_filter is not None:
            _check_filter(self.pixel_filter, self.inputs)
            if precision != "fp64":
                raise ValueError("precision='mixed' covers the unfiltered path; a pixel filter runs in fp64")
        # harmonic route in FP64: "whitener" (products with W = L^-1 D) or
        # "trsm" (blocked solves with L, as the mixed precision does)
        self.harmonic_solver = kwargs.get("harmonic_solver", "whitener")
        if self.harmonic_solver not in ("whitener", "trsm"):
            raise ValueError(f"harmonic_solver must be 'whitener' or 'trsm'; got {self.harmonic_solver!r}")
        self.mixed_gemm = kwargs.get("mixed_gemm", "s8")
        if self.mixed_gemm not in pipeline.MIXED_GEMMS:
            raise ValueError(f"mixed_gemm must be one of {pipeline.MIXED_GEMMS}; got {self.mixed_gemm!r}")

    # -- pipeline -----------------------------------------------------------
    @property
    def lmax_signal(self) -> int:
        return self.inputs.lmax_signal

    @property
    def reduced_noise_cov1(self) -> np.ndarray:
        N = self.inputs.noise
        return N.detach().cpu().numpy() if hasattr(N, "detach") else N

    @property
    def reduced_noise_cov2(self) -> np.ndarray | None:
        N = self.inputs.noise2
        return N.detach().cpu().numpy() if hasattr(N, "detach") else N

    def set_binning(self, bins):
        """bins: list of (lmin_b, lmax_b) or an object with lmins/lmaxs."""
        if hasattr(bins, "lmins"):
            bins = list(zip(np.asarray(bins.lmins).tolist(), np.asarray(bins.lmaxs).tolist()))
        bins = [(int(a), int(b)) for a, b in bins]
        if bins[0][0] != self.params.lmin:
            raise ValueError(f"bins.lmin={bins[0][0]} != params.lmin={self.params.lmin}")
        if max(b for _, b in bins) > self.params.lmax:
            raise ValueError("binning cannot extend above the inference ceiling.")
        self.bins = bins

    def device_filter(self):
        """The pixel filter's device copy (None without a filter)."""
        if self.pixel_filter is not None and self._device_filter is None:
            from .filters import DeviceFilter

            self._device_filter = DeviceFilter(self.ctx, self.pixel_filter, self.noise_prefiltered,
                                               self.maps_prefiltered)
        return self._device_filter

    def prepare_covariance_matrices(self):
        if self.cov is None:
            self.cov = pipeline.Covariance(self.ctx, self.inputs, pixel_filter=self.device_filter())
        if self.basis == "harmonic" and (self.precision == "mixed" or self.harmonic_solver == "trsm"):
            from .harmonic import HarmonicContraction

            mixed = self.precision == "mixed"
            self.cov.build(diag_blocks=HarmonicContraction.MIXED_TRSM_BLOCK if mixed else HarmonicContraction.TRSM_BLOCK,
                           slices=HarmonicContraction.FACTOR_SLICES if mixed else 0)
        else:
            self.cov.build()
        if not in_out._falsy(self.params.outinvcovmatfile1):
            in_out.write_covmat_reduced(self.params.outinvcovmatfile1, engine.to_numpy_cm(self.cov.inverse()))
        if self.params.do_cross:
            if self.cov2 is None:
                self.cov2 = pipeline.Covariance(self.ctx, self.inputs, noise=self.inputs.noise2, vecs=self.cov.vecs,
                                                pixel_filter=self.device_filter())
            self.cov2.build()
            if not in_out._falsy(self.params.outinvcovmatfile2):
                in_out.write_covmat_reduced(self.params.outinvcovmatfile2,
                                            engine.to_numpy_cm(self.cov2.inverse()))

    def run(self) -> None:
        import time

        t0 = time.perf_counter()
        self.prepare_covariance_matrices()
        if self.bins is None:
            self.bins = bins_from_params(self.params)
        lay = self.inputs.layout
        self.spectra_list = [s.label for s in lay.spectra_list()]
        lmin, lmax = self.params.lmin, self.params.lmax
        sm = self.inputs.smoothing
        self.beam_smoothing = np.concatenate([sm[lab][lmin: lmax + 1] for lab in self.spectra_list])
        self.n_params = len(self.spectra_list) * len(self.bins)
        self.ctx.sync()
        t1 = time.perf_counter()
        self._bandpower_window = None
        if self.basis == "harmonic":
            from .harmonic import HarmonicContraction

            self.derivs = HarmonicContraction(self.cov, self.bins, self.params.output_convention,
                                              cov2=self.cov2 if self.params.do_cross else None,
                                              precision=self.precision)
            self.fisher = self.derivs.fisher()
        elif self.params.do_cross:
            self.fisher, self.derivs = pipeline.fisher_cross(self.cov, self.cov2, self.bins,
                                                             self.params.output_convention)
        else:
            self.fisher, self.derivs = pipeline.fisher(self.cov, self.bins, self.params.output_convention,
                                                       timer=self.timer, precision=self.precision,
                                                       launch_timer=self.launch_timer, mixed_gemm=self.mixed_gemm)
        self.ctx.sync()
        self.timings = {"covariance_s": t1 - t0, "fisher_s": time.perf_counter() - t1}
        in_out.write_out_matrix(self.params.outfilefisher, self.fisher)

    compute = run

    def release(self):
        """Drop the device derivative stacks (Spectra regenerates them)."""
        self.derivs = None

    # -- getters ------------------------------------------------------------
    def get_fisher_matrix(self) -> np.ndarray | None:
        return self.fisher

    @property
    def signal_matrix(self) -> np.ndarray | None:
        if self.cov is None or self.cov.S is None:
            return None
        return engine.to_numpy_cm(self.cov.S)

    @property
    def inverse_covariance(self) -> np.ndarray:
        """C^-1 (the reference keeps it in ``noise_cov1`` after ``run``)."""
        return engine.to_numpy_cm(self.cov.inverse())

    @property
    def inverse_covariance2(self) -> np.ndarray | None:
        """C2^-1 = (S + N2)^-1 of a cross analysis (``noise_cov2`` after ``run``)."""
        return None if self.cov2 is None else engine.to_numpy_cm(self.cov2.inverse())

    def get_error_bars(self, *, as_dict: bool = False):
        if self.fisher is None:
            return None
        cov = pipeline.inverse_symm(self.ctx, self.fisher)
        err = np.sqrt(np.diag(cov))
        if not as_dict:
            return err
        nb = len(self.bins)
        return {lab: err[k * nb:(k + 1) * nb] for k, lab in enumerate(self.spectra_list)}

    def get_bandpower_slices(self):
        if self.spectra_list is None:
            return None
        nb = len(self.bins)
        return {lab: slice(k * nb, (k + 1) * nb) for k, lab in enumerate(self.spectra_list)}

    def get_fisher_block(self, label_i: str, label_j: str | None = None):
        F, sl = self.get_fisher_matrix(), self.get_bandpower_slices()
        if F is None or sl is None:
            return None
        label_j = label_i if label_j is None else label_j
        for lab in (label_i, label_j):
            if lab not in sl:
                raise KeyError(f"unknown spectrum label {lab!r}; legal labels: {list(sl)}")
        return F[sl[label_i], sl[label_j]]

    def get_window_matrix(self):
        """The beam-smoothed Fisher matrix, as the reference defines the
        window of the convolved mode (``fisher.py:1000-1049``)."""
        return self.get_fisher_matrix()

    def get_bandpower_window_function(self, per_ell_fisher: np.ndarray | None = None) -> np.ndarray:
        """W = F_b^-1 Q F_perell (``fisher.py:1051-1184``): maps a per-l
        (unbeamed, output-convention) theory to the expected binned
        deconvolved bandpowers, <C_hat_b> = W C_l.  Q sums the l of each bin
        (block-diagonal over spectra); F_perell is the delta_ell = 1 Fisher of
        the same covariance, computed on the device when not given, and the
        result is cached.  Shape (nspectra nbins, nspectra n_ell), blocks
        spectrum-major as ``get_bandpower_slices``."""
        if self.bins is None or self.fisher is None:
            raise ValueError("Bandpower window requires binning. Call set_binning() before run().")
        cached = getattr(self, "_bandpower_window", None)
        if per_ell_fisher is None and cached is not None:
            return cached
        lmin, lmax = self.params.lmin, self.params.lmax
        n_ell = lmax - lmin + 1
        if per_ell_fisher is None:
            per_ell = [(ell, ell) for ell in range(lmin, lmax + 1)]
            if self.basis == "harmonic":
                from .harmonic import HarmonicContraction

                per_ell_fisher = HarmonicContraction(self.cov, per_ell, self.params.output_convention,
                                                     cov2=self.cov2 if self.params.do_cross else None,
                                                     precision=self.precision).fisher()
            else:
                per_ell_fisher, _ = pipeline.fisher(self.cov, per_ell, self.params.output_convention,
                                                    precision=self.precision, keep_stacks=False,
                                                    mixed_gemm=self.mixed_gemm)
        per_ell_fisher = np.asarray(per_ell_fisher, dtype=np.float64)
        nspec = len(self.spectra_list)
        if per_ell_fisher.shape != (nspec * n_ell, nspec * n_ell):
            raise ValueError(f"per_ell_fisher has shape {per_ell_fisher.shape}, expected "
                             f"({nspec * n_ell}, {nspec * n_ell})")
        Qb = np.zeros((len(self.bins), n_ell))
        for b, (lo, hi) in enumerate(self.bins):
            Qb[b, lo - lmin: hi - lmin + 1] = 1.0
        Q = np.kron(np.eye(nspec), Qb)
        # F_b^-1 from the device's Jacobi-scaled Cholesky inverse, then two
        # parameter-sized products
        W = pipeline.inverse_symm(self.ctx, self.fisher) @ (Q @ per_ell_fisher)
        self._bandpower_window = W
        return W

    def get_effective_ells(self) -> np.ndarray:
        """Per bin, the window-weighted centroid sum_l W[b, l] l / sum_l W[b, l]
        over the bin's own spectrum block (``fisher.py:1186-1232``)."""
        if self.fisher is None:
            raise ValueError("get_effective_ells requires a completed Fisher run. Call run() first.")
        W = self.get_bandpower_window_function()
        lmin = self.params.lmin
        n_ell = self.params.lmax - lmin + 1
        nb = len(self.bins)
        ells = np.arange(lmin, lmin + n_ell, dtype=np.float64)
        eff = np.empty(len(self.spectra_list) * nb)
        for s in range(len(self.spectra_list)):
            Wb = W[s * nb:(s + 1) * nb, s * n_ell:(s + 1) * n_ell]
            eff[s * nb:(s + 1) * nb] = (Wb @ ells) / Wb.sum(axis=1)
        return eff
