"""End-to-end parity of the drop-in classes (GPU) with the oracle and the
reference's fixtures.

Tolerances: F and the deconvolved C_l against the FP64 oracle at 1e-10
relative (max-abs difference over max-abs value, and elementwise on the
diagonal); the fixtures at the reference tests' own tolerances.
"""

import numpy as np
import pytest

from tests import cases
from tests import oracle_run as R

pytestmark = pytest.mark.gpu


def _params(case, **extra):
    p = dict(nside=case.nside, spins=case.spins, labels=case.labels, lmin=case.lmin, lmax=case.lmax,
             lmax_signal=case.lmax_signal, ordering="NESTED" if case.nest else "RING",
             output_convention=case.convention, nsims=None if case.maps is None else case.maps.shape[1])
    p.update(extra)
    return p


def _gpu_fisher(case):
    from cosmoforge_amd import Fisher

    f = Fisher(_params(case), basis=False, mask=case.mask, noise_cov1=case.noise, cls_data=case.cls,
               beam=case.beam)
    f.run()
    return f


def _oracle_fisher(case):
    from oracle import pipeline as OP

    lay = R.layout(case)
    cls_s, beams = R.smoothed(case, lay)
    prm = R.params(case, lay, beams)
    F, Cinv, S = OP.fisher(lay, cls_s, case.lmax_signal, case.noise, prm)
    return F, Cinv, S, lay, prm, beams


def _rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("name", ["T", "QU", "TQU", "TEB"])
def test_fisher_nside4(name):
    case = cases.case_fisher_nside4(name)
    f = _gpu_fisher(case)
    F = f.get_fisher_matrix()
    Fo, *_ = _oracle_fisher(case)
    assert _rel(F, Fo) < 1e-10
    d, do = np.diag(F), np.diag(Fo)
    assert np.max(np.abs(d - do) / np.abs(do)) < 1e-10
    bs = f.beam_smoothing
    np.testing.assert_allclose(F / np.outer(bs, bs), case.expect["fisher"], atol=1e-3, rtol=1e-5)


def test_fortran_suite_gpu():
    from cosmoforge_amd import Spectra

    case = cases.case_fortran_b()
    f = _gpu_fisher(case)
    S = f.signal_matrix
    ref = case.expect["S"]
    rel = np.where(np.abs(ref) > 1e-20, np.abs((S - ref) / ref), 0)
    assert rel.max() < 1e-10
    dinv = np.diag(f.inverse_covariance)
    d = case.expect["invcov_diag"]
    assert np.max(np.abs((dinv - d) / d)) < 1e-12
    F = f.get_fisher_matrix()
    Fo, Cinv, _, lay, prm, beams = _oracle_fisher(case)
    assert _rel(F, Fo) < 1e-10
    e = np.arange(case.lmin, case.lmax + 1.0)
    dd = e * (e + 1) / (2 * np.pi)
    fr = F / np.outer(f.beam_smoothing, f.beam_smoothing) * np.outer(dd, dd)
    fd = np.diag(case.expect["fisher"])
    assert np.max(np.abs((np.diag(fr) - fd) / fd)) < 1e-6
    sp = Spectra(fisher=f, maps1=case.maps)
    sp.run()
    ch = sp.get_power_spectra()[0]
    ref_sp = case.expect["spectra"]
    m = min(ref_sp.size, ch.size)
    assert np.max(np.abs((ch[:m] - ref_sp[:m]) / ref_sp[:m])) < 1e-10
    nbh = sp.get_noise_bias()
    nr = case.expect["noise_bias"]
    m = min(nr.size, nbh.size)
    assert np.max(np.abs((nbh[:m] - nr[:m]) / nr[:m])) < 1e-6
    # against the oracle's QML on the same inputs
    from oracle import pipeline as OP

    q, nb, cho, nbho = OP.qml(lay, Cinv, case.noise, case.maps, Fo, prm)
    assert _rel(sp.qml_results, q) < 1e-10
    assert _rel(sp.qml_noise_bias, nb) < 1e-10
    assert _rel(sp.get_power_spectra(), cho) < 1e-10


def test_qml_tqu_with_eb_tb_vs_oracle():
    """TQU with nonzero EB/TB spectra and a seeded map: every spectrum's
    Fisher row and QML estimate against the oracle."""
    from cosmoforge_amd import Fisher, Spectra
    from oracle import pipeline as OP

    case = cases.case_fisher_nside4("TQU")
    rng = np.random.default_rng(20261203)
    cls = dict(case.cls)
    cls["EB"] = 0.2 * np.sqrt(cls["EE"] * cls["BB"])
    cls["TB"] = 0.1 * np.sqrt(cls["TT"] * cls["BB"])
    case.cls = cls
    lay = R.layout(case)
    cls_s, beams = R.smoothed(case, lay)
    prm = R.params(case, lay, beams)
    Fo, Cinv, S, = OP.fisher(lay, cls_s, case.lmax_signal, case.noise, prm)
    Lc = np.linalg.cholesky(S + case.noise)
    maps = Lc @ rng.standard_normal((lay.n, 2))
    q, nb, cho, nbho = OP.qml(lay, Cinv, case.noise, maps, Fo, prm)
    f = Fisher(_params(case, nsims=2), basis=False, mask=case.mask, noise_cov1=case.noise, cls_data=cls,
               beam=case.beam)
    f.run()
    assert _rel(f.get_fisher_matrix(), Fo) < 1e-10
    sp = Spectra(fisher=f, maps1=maps)
    sp.run()
    assert _rel(sp.qml_results, q) < 1e-10
    assert _rel(sp.qml_noise_bias, nb) < 1e-10
    assert _rel(sp.get_power_spectra(), cho) < 1e-10
    # The F^-1-normalised noise bias inherits the conditioning of F (TQU at
    # nside 4 mixes cosmic-variance- and noise-limited bins): measured 2e-10.
    assert _rel(sp.get_noise_bias(), nbho) < 1e-9


def test_spin2_field_listed_first_vs_oracle():
    """spins=[2, 0] (Q/U rows before T): S, Fisher rows, QML estimates and
    noise bias against the oracle's restatement of the reference's assembly
    (``signal_kernels.py:479-499``), whose T x P block of S ends at zero; the
    TE/TB derivatives are those of the [0, 2] order with the blocks moved;
    the harmonic route gives the same F on the same covariance."""
    from cosmoforge_amd import Fisher, Spectra
    from oracle import pipeline as OP

    base = cases.case_fisher_nside4("TQU")
    npix = base.mask.shape[0]
    case = cases.Case("tqu_p_first", base.nside, [2, 0], ["E", "B", "T"], base.mask[:, [1, 2, 0]], dict(base.cls),
                      base.beam, base.lmax_signal, lmin=base.lmin, lmax=base.lmax, nest=base.nest)
    # noise in the Q, U, T row order
    nt = np.diag(base.noise)[: npix]
    nq = np.diag(base.noise)[npix:]
    case.noise = np.diag(np.r_[nq, nt])
    lay = R.layout(case)
    cls_s, beams = R.smoothed(case, lay)
    prm = R.params(case, lay, beams)
    Fo, Cinv, So = OP.fisher(lay, cls_s, case.lmax_signal, case.noise, prm)
    maps = np.linalg.cholesky(So + case.noise) @ np.random.default_rng(4).standard_normal((lay.n, 2))
    q, nb, cho, _ = OP.qml(lay, Cinv, case.noise, maps, Fo, prm)
    f = Fisher(_params(case, nsims=2), basis=False, mask=case.mask, noise_cov1=case.noise, cls_data=case.cls,
               beam=case.beam)
    f.run()
    assert f.spectra_list == [lab for *_, lab in lay.spectra_list()]
    S = f.signal_matrix
    assert np.max(np.abs(S - So)) <= 1e-12 * np.max(np.abs(So))
    assert not S[2 * npix:, : 2 * npix].any()
    assert _rel(f.get_fisher_matrix(), Fo) < 1e-10
    sp = Spectra(fisher=f, maps1=maps)
    sp.run()
    assert _rel(sp.qml_results, q) < 1e-10
    assert _rel(sp.qml_noise_bias, nb) < 1e-10
    assert _rel(sp.get_power_spectra(), cho) < 1e-9
    fh = Fisher(_params(case, nsims=2), basis="harmonic", mask=case.mask, noise_cov1=case.noise,
                cls_data=case.cls, beam=case.beam, context=f.ctx)
    fh.run()
    assert _rel(fh.get_fisher_matrix(), Fo) < 1e-7       # the pixel route's acos angles (polarisation)


def test_v110_pins_gpu():
    from cosmoforge_amd import PICSLike

    case = cases.case_v110()
    pl = PICSLike(_params(case, nsims=1), basis=False, mask=case.mask, noise_cov1=case.noise,
                  cls_data=case.cls, beam=case.beam, maps1=case.maps)
    chi2, lnl = pl._compute_likelihood_point(None)
    logdet = -2.0 * lnl[0] - chi2[0]
    assert logdet == pytest.approx(case.expect["logdet"], rel=1e-10)
    assert chi2[0] == pytest.approx(case.expect["quadform"], rel=1e-10)
    S = pl.cov.S.cpu().numpy().T
    C = S + case.noise
    assert np.trace(C) == pytest.approx(case.expect["cov_trace"], rel=1e-10)
    assert np.linalg.norm(C) == pytest.approx(case.expect["cov_fro"], rel=1e-10)


def test_default_cosine_beam_reproduces_signal_golden():
    """smoothing_type='cosine_legacy' and no beam= (the reference default,
    ``settings.py:219``; ``beam.py:27,192-194``): S equals ref_TQU_signal.dat,
    which the reference's test generated with coswinbeam(4, 4, 12); the
    legacy integer code 2 selects the same beam."""
    import warnings

    from cosmoforge_amd import engine, pipeline
    from cosmoforge_amd.core import Inputs

    case = cases.case_signal_tqu()
    ctx = engine.Context(0)
    for st in ("cosine_legacy", 2):
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", DeprecationWarning)
            inp = Inputs(_params(case, smoothing_type=st), case.mask, np.eye(576), case.cls, None)
        S = engine.to_numpy_cm(pipeline.Covariance(ctx, inp).signal())
        np.testing.assert_allclose(S, case.expect["S"], rtol=1e-12, atol=1e-20)


def test_v110_pins_with_default_gaussian_beam():
    """The v1.1.0 pins (``tests/test_core_pipeline.py:276-308``) reached
    through smoothing_type='gaussian', fwhmarcmin=60 and no beam=: pins the
    T column of the Gaussian window (the field is T only; the E/B factor
    exp(2 sigma^2) has no reference fixture)."""
    from cosmoforge_amd import PICSLike

    case = cases.case_v110()
    pl = PICSLike(_params(case, nsims=1, smoothing_type="gaussian", fwhmarcmin=60.0), basis=False, mask=case.mask,
                  noise_cov1=case.noise, cls_data=case.cls, maps1=case.maps)
    chi2, lnl = pl._compute_likelihood_point(None)
    logdet = -2.0 * lnl[0] - chi2[0]
    assert logdet == pytest.approx(case.expect["logdet"], rel=1e-10)
    assert chi2[0] == pytest.approx(case.expect["quadform"], rel=1e-10)
    C = pl.cov.S.cpu().numpy().T + case.noise
    assert np.trace(C) == pytest.approx(case.expect["cov_trace"], rel=1e-10)
    assert np.linalg.norm(C) == pytest.approx(case.expect["cov_fro"], rel=1e-10)


def test_picslike_fortran_gpu():
    from cosmoforge_amd import PICSLike

    case, points, res = cases.case_picslike_b()
    theory = {(k,): cl for k, cl in enumerate(points)}
    pl = PICSLike(_params(case, nsims=1), basis=False, mask=case.mask, noise_cov1=case.noise,
                  beam=case.beam, maps1=case.maps, theoretical_spectra=theory)
    pl.compute()
    chi2 = pl.get_chi_squared()
    lnl = pl.get_log_likelihood()
    assert np.max(np.abs(chi2 / res[:, 3] - 1)) < 1e-9
    assert np.max(np.abs(lnl - res[:, 2])) / np.abs(res[:, 2]).mean() < 1e-9
    assert pl.get_best_fit() == {"point": int(np.argmax(lnl))}


def test_fisher_bands_equal_single_band():
    """The banded Fisher (several row bands, stacks regenerated per band)
    equals the single-band Fisher at FP64 rounding level."""
    from cosmoforge_amd import Fisher, pipeline

    case = cases.case_fisher_nside4("TQU")
    f1 = _gpu_fisher(case)
    F1 = f1.get_fisher_matrix()
    f2 = Fisher(_params(case), basis=False, mask=case.mask, noise_cov1=case.noise, cls_data=case.cls,
                beam=case.beam)
    f2.prepare_covariance_matrices()
    from cosmoforge_amd.core import bins_from_params

    bins = bins_from_params(f2.params)
    n, npar = f2.cov.n, 6 * len(bins)
    budget = 3 * npar * 8 * (n * (n + 1) // 2) // 10   # forces bands; stacks no longer resident
    timer = pipeline.StageTimer()
    F2, src = pipeline.fisher(f2.cov, bins, f2.params.output_convention, timer=timer, mem_budget=budget)
    assert len(timer.bands) > 1 and src.resident is None
    assert _rel(F2, F1) < 1e-12


def test_mixed_precision_fisher_qml_close_to_fp64():
    """precision="mixed" (int8-slice whitening on the pixel route) against
    the FP64 path, per entry: |dF_ij| / sqrt(F_ii F_jj), the noise bias and
    q in units of sigma(q_p), and |dC_l| / max(|C_l|, sigma_l), on the
    nside-8 Fortran case and on nside-4 TEB.  The nside-4 TEB F is singular
    (cond ~ 1e16: its EB/TB bins are unconstrained), so its C_l = q F^-1 is
    not a well-defined quantity to compare; F, q and the noise bias are."""
    from cosmoforge_amd import Fisher, Spectra
    from cosmoforge_amd.accuracy import errors

    for case in [cases.case_fortran_b(), cases.case_fisher_nside4("TEB")]:
        out = {}
        for prec in ("fp64", "mixed"):
            f = Fisher(_params(case), basis=False, mask=case.mask, noise_cov1=case.noise, cls_data=case.cls,
                       beam=case.beam, precision=prec)
            f.run()
            x = case.maps if case.maps is not None else np.random.default_rng(1).standard_normal((f.cov.n, 1))
            s = Spectra(fisher=f, maps1=x)
            s.run()
            sig = np.sqrt(np.abs(np.diag(np.linalg.pinv(f.get_fisher_matrix()))))
            out[prec] = (f.get_fisher_matrix(), s.get_power_spectra(), s.qml_noise_bias, s.qml_results, sig)
        e = errors(out["mixed"], out["fp64"])
        condF = np.linalg.cond(out["fp64"][0])
        print(f"mixed vs fp64 ({case.name}, cond F {condF:.1e}):", e)
        assert e["F_entry_max"] < 1e-9
        assert e["noise_bias_over_sigma_q_max"] < 1e-9 and e["q_over_sigma_q_max"] < 1e-9
        if case.name != "fisher_nside4_TEB":
            assert e["Cl_rel_max"] < 1e-6


def test_mixed_whitening_engines_and_bands():
    """Mixed Fisher on nside-4 TQU (TE/TB blocks): the int8-slice whitening
    in one band and in several (partial bands take the transposed term from
    the rows of D; stacks regenerated per band) within 1e-9 of FP64 per
    entry; the FP32-accurate variants (three-piece BF16, one band and
    several; FP32 MFMA) within 3e-5 per entry (measured 7e-6: FP32 rounding
    of W D W^T, which F cancels only to C's conditioning), the banded
    three-piece result within 1e-5 of its single band (FP32 sums in
    another order)."""
    from cosmoforge_amd import Fisher, pipeline
    from cosmoforge_amd.core import bins_from_params

    case = cases.case_fisher_nside4("TQU")
    f = Fisher(_params(case), basis=False, mask=case.mask, noise_cov1=case.noise, cls_data=case.cls,
               beam=case.beam)
    f.prepare_covariance_matrices()
    bins = bins_from_params(f.params)
    conv = f.params.output_convention
    n, npar = f.cov.n, 6 * len(bins)
    F64, _ = pipeline.fisher(f.cov, bins, conv)
    d = np.sqrt(np.diag(F64))

    def err(F):
        return float(np.max(np.abs(F - F64) / np.outer(d, d)))

    Fs, _ = pipeline.fisher(f.cov, bins, conv, precision="mixed")
    timer = pipeline.StageTimer()
    Fsb, src = pipeline.fisher(f.cov, bins, conv, precision="mixed", timer=timer,
                               mem_budget=3 * npar * 8 * (n * (n + 1) // 2) // 10)
    assert len(timer.bands) > 1 and src.resident is None
    Fb, _ = pipeline.fisher(f.cov, bins, conv, precision="mixed", mixed_gemm="bf16x3")
    Ff, _ = pipeline.fisher(f.cov, bins, conv, precision="mixed", mixed_gemm="fp32")
    timer = pipeline.StageTimer()
    Fbb, src = pipeline.fisher(f.cov, bins, conv, precision="mixed", mixed_gemm="bf16x3", timer=timer,
                               mem_budget=3 * npar * 4 * (n * (n + 1) // 2) // 10)
    assert len(timer.bands) > 1 and src.mirrors
    errs = dict(s8=err(Fs), s8_bands=err(Fsb), b3=err(Fb), fp32=err(Ff), b3_bands=err(Fbb))
    print("mixed whitening vs fp64 (per entry):", errs)
    assert errs["s8"] < 1e-9 and errs["s8_bands"] < 1e-9
    assert errs["b3"] < 3e-5 and errs["fp32"] < 3e-5 and errs["b3_bands"] < 3e-5
    assert np.max(np.abs(Fbb - Fb) / np.outer(d, d)) < 1e-5
    with pytest.raises(ValueError, match="mixed_gemm"):
        pipeline.fisher(f.cov, bins, conv, precision="mixed", mixed_gemm="fp16")


def test_device_resident_noise_matches_host():
    """noise_cov1 given as a device tensor (kept in HBM) gives the same F,
    bit for bit, as the host array, and the getter returns host data."""
    import torch
    from cosmoforge_amd import Fisher

    case = cases.case_fisher_nside4("TQU")
    f1 = Fisher(_params(case), basis=False, mask=case.mask, noise_cov1=case.noise, cls_data=case.cls,
                beam=case.beam)
    f1.run()
    Nd = torch.as_tensor(np.ascontiguousarray(case.noise)).to("cuda")
    f2 = Fisher(_params(case), basis=False, mask=case.mask, noise_cov1=Nd, cls_data=case.cls, beam=case.beam,
                context=f1.ctx)
    f2.run()
    assert np.array_equal(f1.get_fisher_matrix(), f2.get_fisher_matrix())
    assert np.array_equal(f2.reduced_noise_cov1, np.asarray(case.noise))


def test_qml_correlated_noise_vs_oracle():
    """A non-diagonal (correlated) noise covariance takes the general
    G = C^-1 N C^-1 path: Fisher, QML estimates and noise bias against the
    oracle."""
    from cosmoforge_amd import Fisher, Spectra
    from oracle import pipeline as OP

    case = cases.case_fisher_nside4("TQU")
    rng = np.random.default_rng(11)
    lay = R.layout(case)
    Q = rng.standard_normal((lay.n, 20))
    case.noise = np.asarray(case.noise) + 0.05 * (Q @ Q.T) / 20.0
    cls_s, beams = R.smoothed(case, lay)
    prm = R.params(case, lay, beams)
    Fo, Cinv, S, = OP.fisher(lay, cls_s, case.lmax_signal, case.noise, prm)
    maps = np.linalg.cholesky(S + case.noise) @ rng.standard_normal((lay.n, 1))
    q, nb, cho, nbho = OP.qml(lay, Cinv, case.noise, maps, Fo, prm)
    f = Fisher(_params(case), basis=False, mask=case.mask, noise_cov1=case.noise, cls_data=case.cls,
               beam=case.beam)
    f.run()
    assert not f.cov.noise_is_diagonal()
    assert _rel(f.get_fisher_matrix(), Fo) < 1e-10
    sp = Spectra(fisher=f, maps1=maps)
    sp.run()
    assert _rel(sp.qml_results, q) < 1e-10
    assert _rel(sp.qml_noise_bias, nb) < 1e-10


@pytest.mark.parametrize("name", ["T", "TQU"])
def test_bandpower_window_and_effective_ells_vs_oracle(name):
    """Fisher.get_bandpower_window_function with delta_ell = 2 bins on
    nside-4 T and TQU: W = F_b^-1 Q F_perell against the oracle's F_b and
    F_perell (tolerance scaled by cond(F_b));
    each spectrum's rows sum to 1 over its own l (W Q^T = I); effective
    multipoles as the oracle's window gives; Spectra delegates to Fisher."""
    import scipy.linalg
    from cosmoforge_amd import Fisher, Spectra
    from oracle import pipeline as OP

    case = cases.case_fisher_nside4(name)
    lay = R.layout(case)
    cls_s, beams = R.smoothed(case, lay)
    lmin, lmax = case.lmin, case.lmax
    bins2 = [(lo, min(lo + 1, lmax)) for lo in range(lmin, lmax + 1, 2)]
    per_ell = [(ell, ell) for ell in range(lmin, lmax + 1)]
    pw = lambda b: OP.parameter_weights(lay, beams, lmin, lmax, case.lmax_signal, bins=b,  # noqa: E731
                                        convention=case.convention)
    Fb, _, S = OP.fisher(lay, cls_s, case.lmax_signal, case.noise, pw(bins2))
    Fl, _, _ = OP.fisher(lay, cls_s, case.lmax_signal, case.noise, pw(per_ell), S=S)
    nspec, n_ell = len(lay.spectra_list()), lmax - lmin + 1
    Qb = np.zeros((len(bins2), n_ell))
    for b, (lo, hi) in enumerate(bins2):
        Qb[b, lo - lmin: hi - lmin + 1] = 1.0
    Q = np.kron(np.eye(nspec), Qb)
    Wo = scipy.linalg.cho_solve(scipy.linalg.cho_factor(Fb, lower=True), Q @ Fl)

    f = Fisher(_params(case), basis=False, mask=case.mask, noise_cov1=case.noise, cls_data=case.cls,
               beam=case.beam)
    f.set_binning(bins2)
    f.run()
    W = f.get_bandpower_window_function()
    tol = 1e-10 * np.linalg.cond(Fb)
    assert np.max(np.abs(W - Wo)) <= tol * np.max(np.abs(Wo))
    if np.linalg.cond(Fb) < 1e8:       # W Q^T = F_b^-1 (Q F_perell Q^T) = I up to conditioning
        assert np.allclose(W @ Q.T, np.eye(W.shape[0]), atol=1e-8)
    eff = f.get_effective_ells()
    nb = len(bins2)
    ells = np.arange(lmin, lmax + 1, dtype=np.float64)
    eff_o = np.concatenate([(Wo[s * nb:(s + 1) * nb, s * n_ell:(s + 1) * n_ell] @ ells)
                            / Wo[s * nb:(s + 1) * nb, s * n_ell:(s + 1) * n_ell].sum(axis=1) for s in range(nspec)])
    assert np.max(np.abs(eff - eff_o)) <= tol * np.max(np.abs(eff_o))
    sp = Spectra(fisher=f, maps1=np.random.default_rng(3).standard_normal((f.cov.n, 1)))
    assert np.array_equal(sp.get_effective_ells(), eff)
    assert f.get_bandpower_window_function() is W          # cached
    with pytest.raises(ValueError, match="per_ell_fisher has shape"):
        f.get_bandpower_window_function(per_ell_fisher=np.eye(3))


def _b_config(tmp_path):
    """The reference's nside-8 B-only PICSLike configuration
    (``picslike/tests/data/nside8/B/fortran_reference/config.yaml``) over the
    committed copies of its data files; the full-sky noise binary and the
    theory-spectra directory are rebuilt from the fixtures."""
    import os
    import shutil

    files = os.path.join(cases.GOLDEN, "files")
    d = cases.load("picslike_nside8_B.npz")
    cases.untril(768, d["ncov_tril"]).tofile(str(tmp_path / "ncov_B_only.bin"))
    th = tmp_path / "theory_spectra"
    th.mkdir()
    for tag in ("0", "1.25e-05", "2.5e-05"):
        shutil.copyfile(os.path.join(files, f"B_theory_r{tag}.txt"), str(th / f"fine_dls_r_likelihood_r{tag}.txt"))
    return dict(
        nside=8, spins=[0], labels=["B"], do_cross=False, ordering="RING",
        maskfile=os.path.join(files, "B_mask.fits"), output_geometry_file=str(tmp_path / "out/geometry.dat"),
        inputclfile=os.path.join(files, "B_input_dls.txt"), input_convention="Dl",
        covmatfile1=str(tmp_path / "ncov_B_only.bin"), lmax_signal=32, lmin=2, lmax=16,
        fiducialfile=os.path.join(files, "B_fiducial_dls.txt"), smoothing_type="file",
        beam_file=os.path.join(files, "B_beam.fits"), apply_pixwin=False,
        outnoisecovmat1=str(tmp_path / "out/output_ncov.bin"), outinvcovmatfile1=str(tmp_path / "out/invcov.bin"),
        outfilefisher=str(tmp_path / "out/fisher.dat"), nsims=1,
        inputmapfile1=os.path.join(files, "B_map.fits"),
        outcovmatfile=str(tmp_path / "out/cov_matrix.dat"), outerrfile=str(tmp_path / "out/errors.dat"),
        remove_nb=True, root_dir=str(th), root_filename="fine_dls_r_likelihood",
        parameters={"r": [0.0, 2.5e-05, 3]})


def test_picslike_from_yaml_files_matches_fortran_results(tmp_path):
    """PICSLike run from a YAML file of paths alone (mask, beam, noise, map,
    C_l and per-point theory files): the same chi2 / lnL as the in-memory
    run and the reference's Fortran results."""
    import yaml

    from cosmoforge_amd import PICSLike

    cfg = _b_config(tmp_path)
    path = tmp_path / "config.yaml"
    path.write_text(yaml.safe_dump(cfg))
    print("picslike from files", path, flush=True)
    pl = PICSLike(str(path), basis=False)
    pl.compute()
    _, _, res = cases.case_picslike_b()
    chi2, lnl = pl.get_chi_squared(), pl.get_log_likelihood()
    assert np.max(np.abs(chi2 / res[:, 3] - 1)) < 1e-9
    assert np.max(np.abs(lnl - res[:, 2])) / np.abs(res[:, 2]).mean() < 1e-9
    assert pl.get_best_fit() == {"r": pl.parameter_grid.grid_points[int(np.argmax(lnl))][0]}
    g = np.loadtxt(cfg["output_geometry_file"], skiprows=1)
    ref = cases.case_fortran_b().expect["geometry"]
    assert np.array_equal(g[:, 0], ref[:, 0])
    assert np.max(np.abs(g[:, 1:] - ref[:, 1:])) < 1e-15
    red = np.fromfile(cfg["outnoisecovmat1"]).reshape(768, 768)
    assert np.array_equal(red, pl.inputs.noise)


def test_fisher_spectra_from_files_match_in_memory(tmp_path):
    """Fisher / Spectra driven by file paths give bit-identical F and C_l to
    the same inputs passed in memory, and write F, F^-1, the error bars and
    C^-1 where the configuration names them."""
    from cosmoforge_amd import Fisher, Spectra

    cfg = _b_config(tmp_path)
    print("fisher from files", flush=True)
    f = Fisher(cfg, basis=False)
    f.run()
    sp = Spectra(fisher=f)
    sp.run()
    case, _, _ = cases.case_picslike_b()
    f2 = Fisher(_params(case, nsims=1), basis=False, mask=case.mask, noise_cov1=case.noise,
                cls_data=f.inputs.cls, beam=case.beam)
    f2.run()
    sp2 = Spectra(fisher=f2, maps1=case.maps)
    sp2.run()
    assert np.array_equal(f.get_fisher_matrix(), f2.get_fisher_matrix())
    assert np.array_equal(sp.get_power_spectra(), sp2.get_power_spectra())
    F = f.get_fisher_matrix()
    assert np.array_equal(np.loadtxt(cfg["outfilefisher"]), F)
    assert np.array_equal(np.loadtxt(cfg["outcovmatfile"]), sp.get_covariance())
    err = np.loadtxt(cfg["outerrfile"])
    assert np.array_equal(err.reshape(-1), sp.get_error_bars())
    Ci = np.fromfile(cfg["outinvcovmatfile1"]).reshape(768, 768)
    assert np.array_equal(Ci, f.inverse_covariance)
    sp.params.outclfile = str(tmp_path / "out" / "cl.dat")
    sp.write_power_spectra()
    cl = np.loadtxt(tmp_path / "out" / "cl_deconvolved.dat").reshape(len(f.bins), -1)
    np.testing.assert_allclose(cl.T.reshape(-1), sp.get_power_spectra().mean(0), rtol=1e-15)
    sp.write_power_spectra(mode="convolved")
    np.testing.assert_allclose(np.loadtxt(tmp_path / "out" / "cl_convolved_window.dat"), F, rtol=1e-15)


@pytest.mark.parametrize("name", ["T", "TQU"])
def test_cross_fisher_and_qml_vs_oracle(name):
    """do_cross: two noise covariances (the second correlated), two map sets
    drawn from C1 and C2; F, the raw q and the deconvolved C_l against the
    oracle's restatement of the reference at 1e-10; the Fisher computed in
    several spectrum groups equals the one-group result."""
    from cosmoforge_amd import Fisher, Spectra, pipeline
    from oracle import kernels as OK
    from oracle import pipeline as OP

    case = cases.case_fisher_nside4(name)
    n = case.noise.shape[0]
    rng = np.random.default_rng(7)
    A = rng.standard_normal((n, n)) / np.sqrt(n)
    N2 = np.diag(np.diag(case.noise)) * 1.7 + 0.05 * np.mean(np.diag(case.noise)) * (A @ A.T)
    lay = R.layout(case)
    cls_s, beams = R.smoothed(case, lay)
    prm = R.params(case, lay, beams)
    S = OK.signal_matrix(lay, cls_s, case.lmax_signal)
    x1 = np.linalg.cholesky(S + case.noise) @ rng.standard_normal((n, 2))
    x2 = np.linalg.cholesky(S + N2) @ rng.standard_normal((n, 2))
    Fo, C1, C2 = OP.fisher_cross(lay, cls_s, case.lmax_signal, case.noise, N2, prm, S=S)
    qo, cho = OP.qml_cross(lay, C1, C2, x1, x2, Fo, prm)
    print("cross", name, n, flush=True)
    f = Fisher(_params(case, do_cross=True, nsims=2), basis=False, mask=case.mask, noise_cov1=case.noise,
               noise_cov2=N2, cls_data=case.cls, beam=case.beam)
    f.run()
    sp = Spectra(fisher=f, maps1=x1, maps2=x2)
    sp.run()
    F = f.get_fisher_matrix()
    assert _rel(F, Fo) < 1e-10
    assert np.max(np.abs(np.diag(F) / np.diag(Fo) - 1)) < 1e-10
    assert _rel(sp.qml_results, qo) < 1e-10
    assert _rel(sp.get_power_spectra(), cho) < 1e-10
    assert np.all(sp.qml_noise_bias == 0)
    assert np.max(np.abs(f.inverse_covariance2 - C2)) / np.abs(C2).max() < 1e-10
    ngeo = len(f.derivs.geometry)
    if ngeo > 1:
        per = len(f.bins) * n * n * 8
        Fg, _ = pipeline.fisher_cross(f.cov, f.cov2, f.bins, f.params.output_convention,
                                      mem_budget=2 * (f.derivs.total_bytes + 4 * per))
        assert _rel(Fg, F) < 1e-12


def _filter_case(name):
    case = cases.case_fisher_nside4(name)
    case.lmax = 6
    lay = R.layout(case)
    cls_s, beams = R.smoothed(case, lay)
    prm = R.params(case, lay, beams)
    return case, lay, cls_s, prm


@pytest.mark.parametrize("kind", ["deprojection_TQU", "graded_T"])
def test_filtered_fisher_qml_picslike_vs_oracle(kind):
    """A pixel filter (ADR-0020): N, S, the binned derivatives and the maps
    restricted to its range on the device; F, q, the noise bias, C_hat and
    the PICSLike chi2 / ln|C'| against the oracle's restricted problem at
    1e-10."""
    from cosmoforge_amd import Fisher, PICSLike, Spectra
    from cosmoforge_amd.filters import Filter
    from oracle import healpix
    from oracle import kernels as OK
    from oracle import pipeline as OP
    from oracle.linalg import matrix_slogdet_symm

    name = kind.split("_")[1]
    case, lay, cls_s, prm = _filter_case(name)
    n = case.noise.shape[0]
    rng = np.random.default_rng(13)
    if kind.startswith("deprojection"):
        # monopole and dipole of T, and a Q/U template, removed
        act = np.nonzero(case.mask[:, 0] > 0.5)[0]
        vec = healpix.pointings(4, act, nest=False)
        T = np.zeros((n, 5))
        nt = act.size
        T[:nt, 0] = 1.0
        T[:nt, 1:4] = vec
        T[nt:, 4] = rng.standard_normal(n - nt)
        flt = Filter.from_deprojection(T, mask=case.mask)
        W, sig = flt.W, None
    else:
        Q1, _ = np.linalg.qr(rng.standard_normal((n, n)))
        Q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
        s = np.r_[np.linspace(2.0, 0.5, n - 20), np.full(20, 1e-12)]
        flt = Filter.from_operator(Q1 @ np.diag(s) @ Q2.T, range_epsilon=1e-6)
        W, sig = flt.W, flt.sigma
    S = OK.signal_matrix(lay, cls_s, case.lmax_signal)
    x = np.linalg.cholesky(S + case.noise) @ rng.standard_normal((n, 2))
    Fo, Cf, Nr = OP.fisher_filtered(lay, cls_s, case.lmax_signal, case.noise, prm, W, sig, S=S)
    qo, nbo, cho = OP.qml_filtered(lay, Cf, Nr, x, Fo, prm, W, sig)
    print("filter", kind, n, flt.rank, flush=True)
    f = Fisher(_params(case, nsims=2), basis=False, mask=case.mask, noise_cov1=case.noise, cls_data=case.cls,
               beam=case.beam, pixel_filter=flt)
    f.run()
    sp = Spectra(fisher=f, maps1=x)
    sp.run()
    F = f.get_fisher_matrix()
    assert _rel(F, Fo) < 1e-10
    assert np.max(np.abs(np.diag(F) / np.diag(Fo) - 1)) < 1e-10
    # q: cond(C) is 1.1e8 here, and two FP64 formulations of the oracle
    # itself (explicit C^-1 against solves) differ by 2.3e-10 in q and
    # 1.2e-10 in q + nb, 2e-11 in nb: q and C_hat are held to 1e-9
    assert _rel(sp.qml_noise_bias, nbo) < 1e-10
    assert _rel(sp.qml_results + sp.qml_noise_bias, qo + nbo) < 1e-9
    assert _rel(sp.qml_results, qo) < 1e-9
    assert _rel(sp.get_power_spectra(), cho) < 1e-9
    pl = PICSLike(_params(case, nsims=2), basis=False, mask=case.mask, noise_cov1=case.noise, beam=case.beam,
                  maps1=x, theoretical_spectra={(0,): case.cls}, pixel_filter=flt)
    pl.compute()
    Cr = OP.restrict(S, W, sig) + Nr
    xr = OP.restrict_data(x, W, sig)
    chi2 = np.einsum("in,ij,jn->n", xr, np.linalg.inv(Cr), xr)
    lnl = -0.5 * (chi2 + matrix_slogdet_symm(Cr)[1])
    assert np.max(np.abs(pl.simulation_results[0].chi_squared_values / chi2[0] - 1)) < 1e-9
    assert np.max(np.abs(pl.get_log_likelihood() - lnl.mean())) < 1e-9 * np.abs(lnl).max()


def test_invertible_filter_leaves_the_estimator_unchanged():
    """Full-rank graded filter: F, q and the noise bias equal the unfiltered
    device results (Gaussian Fisher invariance), to the conditioning the
    filter adds."""
    import warnings

    from cosmoforge_amd import Fisher, Spectra
    from cosmoforge_amd.filters import Filter

    case, lay, cls_s, prm = _filter_case("TQU")
    n = case.noise.shape[0]
    rng = np.random.default_rng(3)
    Q1, _ = np.linalg.qr(rng.standard_normal((n, n)))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        flt = Filter.from_operator(Q1 @ np.diag(np.linspace(1.5, 0.7, n)), range_rank=n)
    x = rng.standard_normal((n, 1)) * np.sqrt(np.diag(case.noise))[:, None]
    out = []
    for pf in (None, flt):
        f = Fisher(_params(case, nsims=1), basis=False, mask=case.mask, noise_cov1=case.noise, cls_data=case.cls,
                   beam=case.beam, pixel_filter=pf)
        f.run()
        sp = Spectra(fisher=f, maps1=x)
        sp.run()
        out.append((f.get_fisher_matrix(), sp.qml_results, sp.qml_noise_bias))
    for a, b in zip(out[0], out[1]):
        assert _rel(b, a) < 1e-9


@pytest.mark.parametrize("name", ["T", "TQU", "TEB"])
def test_harmonic_contraction_vs_oracle_and_pixel_path(name):
    """basis="harmonic": modes built on the device, M = (W V)^T (W V), F and
    the QML forms reduced over the modes; against the oracle's mode
    restatement at 1e-10 and against the pixel path on the device (1e-10 for
    temperature, 1e-8 with polarisation: the pixel path's acos rotation
    angles)."""
    from cosmoforge_amd import Fisher, Spectra
    from oracle import harmonic as OH
    from oracle import kernels as OK
    from oracle import pipeline as OP

    case = cases.case_fisher_nside4(name)
    lay = R.layout(case)
    cls_s, beams = R.smoothed(case, lay)
    prm = R.params(case, lay, beams)
    S = OK.signal_matrix(lay, cls_s, case.lmax_signal)
    n = case.noise.shape[0]
    x = np.linalg.cholesky(S + case.noise) @ np.random.default_rng(9).standard_normal((n, 2))
    Cinv = OP.matrix_inverse_symm(S + case.noise)
    Fo = OH.fisher_modes(lay, Cinv, prm, case.lmin, case.lmax)
    qo, nbo = OH.qml_modes(lay, Cinv, case.noise, x, prm, case.lmin, case.lmax)
    print("harmonic", name, n, flush=True)
    out = {}
    for basis in ("harmonic", False):
        f = Fisher(_params(case, nsims=2), basis=basis, mask=case.mask, noise_cov1=case.noise, cls_data=case.cls,
                   beam=case.beam)
        f.run()
        sp = Spectra(fisher=f, maps1=x)
        sp.run()
        out[basis] = (f.get_fisher_matrix(), sp.qml_results, sp.qml_noise_bias, sp.get_power_spectra())
    F, q, nb, ch = out["harmonic"]
    assert _rel(F, Fo) < 1e-10
    assert np.max(np.abs(np.diag(F) / np.diag(Fo) - 1)) < 1e-10
    # nb = 1/2 <G, dC_b> is a cancelling sum (terms ~1e3 x nb): 1e-8
    assert _rel(nb, nbo) < 1e-8
    assert _rel(q + nb, qo + nbo) < 1e-9
    cho = (qo) @ OP.matrix_inverse_symm(Fo)
    assert _rel(ch, cho) < 1e-9
    # against the pixel route: F, nb and q + nb; C_hat = q F^-1 would carry
    # the routes' 1e-8 difference in F times cond(F)
    tol = 1e-10 if name == "T" else 1e-8
    Fp, qp, nbp, chp = out[False]
    assert _rel(F, Fp) < tol
    assert _rel(nb, nbp) < 1e-8
    assert _rel(q + nb, qp + nbp) < 10 * tol


def test_harmonic_contraction_with_wide_and_gapped_bins():
    """basis="harmonic" with delta_ell = 2 bins from lmin = 3 and with
    explicit bins leaving a gap: the same F and noise bias as the pixel route
    on the device (TQU, 1e-8), and the mode range follows the bins."""
    from cosmoforge_amd import Fisher, Spectra

    case = cases.case_fisher_nside4("TQU")
    n = case.noise.shape[0]
    x = np.random.default_rng(2).standard_normal((n, 1)) * np.sqrt(np.diag(case.noise))[:, None]
    for extra in (dict(lmin=3, lmax=8, delta_ell=2), dict(bin_lmins=[2, 6, 9], bin_lmaxs=[3, 7, 12], lmax=12)):
        print("harmonic bins", extra, flush=True)
        out = []
        for basis in ("harmonic", False):
            p = _params(case, nsims=1, **extra)
            f = Fisher(p, basis=basis, mask=case.mask, noise_cov1=case.noise, cls_data=case.cls, beam=case.beam)
            f.run()
            sp = Spectra(fisher=f, maps1=x)
            sp.run()
            out.append((f.get_fisher_matrix(), sp.qml_noise_bias, f))
        (F, nb, fh), (Fp, nbp, _) = out
        assert F.shape == Fp.shape
        assert _rel(F, Fp) < 1e-8
        assert _rel(nb, nbp) < 1e-8
        lo, hi = min(a for a, _ in fh.bins), max(b for _, b in fh.bins)
        assert fh.derivs is not None and (fh.derivs.lmin, fh.derivs.lmax) == (lo, hi)


def test_prefiltered_inputs_match_raw_inputs():
    """noise_prefiltered / maps_prefiltered: a noise covariance F N F^T and
    maps F d that already carry a graded filter F = U Sigma W^T, restricted
    with U^T, give the F, q and noise bias of the raw N and d restricted with
    Sigma W^T (exactly so for the rank-r operator)."""
    from cosmoforge_amd import Fisher, Spectra
    from cosmoforge_amd.filters import Filter

    case, lay, cls_s, prm = _filter_case("T")
    n = case.noise.shape[0]
    rng = np.random.default_rng(21)
    Q1, _ = np.linalg.qr(rng.standard_normal((n, n)))
    Q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.r_[np.linspace(1.8, 0.6, n - 12), np.zeros(12)]
    Fop = Q1 @ np.diag(s) @ Q2.T
    flt = Filter.from_operator(Fop, range_rank=n - 12)
    Fr = flt.U @ np.diag(flt.sigma) @ flt.W.T
    x = rng.standard_normal((n, 1)) * np.sqrt(np.diag(case.noise))[:, None]
    out = []
    for pref in (False, True):
        N = Fr @ case.noise @ Fr.T if pref else case.noise
        d = Fr @ x if pref else x
        f = Fisher(_params(case, nsims=1), basis=False, mask=case.mask, noise_cov1=N, cls_data=case.cls,
                   beam=case.beam, pixel_filter=flt, noise_prefiltered=pref, maps_prefiltered=pref)
        f.run()
        sp = Spectra(fisher=f, maps1=d)
        sp.run()
        out.append((f.get_fisher_matrix(), sp.qml_results + sp.qml_noise_bias, sp.qml_noise_bias))
    for a, b in zip(out[0], out[1]):
        assert _rel(b, a) < 1e-9


@pytest.mark.parametrize("name", ["T", "TQU"])
def test_harmonic_cross_analysis_vs_oracle(name):
    """basis="harmonic" with do_cross: F = 1/2 Tr[C1^-1 dC C2^-1 dC'] from
    M1 and M2 and q = 1/2 y2^T dC y1, against the oracle's cross restatement
    of the reference (1e-10 for temperature, 1e-8 with polarisation: the
    pixel derivative's acos rotation angles)."""
    from cosmoforge_amd import Fisher, Spectra
    from oracle import kernels as OK
    from oracle import pipeline as OP

    case = cases.case_fisher_nside4(name)
    n = case.noise.shape[0]
    rng = np.random.default_rng(17)
    A = rng.standard_normal((n, n)) / np.sqrt(n)
    N2 = np.diag(np.diag(case.noise)) * 1.3 + 0.05 * np.mean(np.diag(case.noise)) * (A @ A.T)
    lay = R.layout(case)
    cls_s, beams = R.smoothed(case, lay)
    prm = R.params(case, lay, beams)
    S = OK.signal_matrix(lay, cls_s, case.lmax_signal)
    x1 = np.linalg.cholesky(S + case.noise) @ rng.standard_normal((n, 2))
    x2 = np.linalg.cholesky(S + N2) @ rng.standard_normal((n, 2))
    Fo, C1, C2 = OP.fisher_cross(lay, cls_s, case.lmax_signal, case.noise, N2, prm, S=S)
    qo, _ = OP.qml_cross(lay, C1, C2, x1, x2, Fo, prm)
    print("harmonic cross", name, n, flush=True)
    f = Fisher(_params(case, do_cross=True, nsims=2), basis="harmonic", mask=case.mask, noise_cov1=case.noise,
               noise_cov2=N2, cls_data=case.cls, beam=case.beam)
    f.run()
    sp = Spectra(fisher=f, maps1=x1, maps2=x2)
    sp.run()
    tol = 1e-10 if name == "T" else 1e-8
    assert _rel(f.get_fisher_matrix(), Fo) < tol
    assert _rel(sp.qml_results, qo) < 10 * tol
    assert np.all(sp.qml_noise_bias == 0)


@pytest.mark.parametrize("name", ["T", "TQU"])
def test_harmonic_contraction_under_pixel_filter(name):
    """basis="harmonic" with a pixel filter: the modes restricted to the
    filter's range (V' = Sigma W^T V) give the filtered pixel route's F, q + nb
    and nb (1e-10 for temperature, 1e-8 with polarisation)."""
    from cosmoforge_amd import Fisher, Spectra
    from cosmoforge_amd.filters import Filter

    case, lay, cls_s, prm = _filter_case(name)
    n = case.noise.shape[0]
    rng = np.random.default_rng(29)
    T = rng.standard_normal((n, 3))
    flt = Filter.from_deprojection(T, mask=case.mask)
    x = rng.standard_normal((n, 1)) * np.sqrt(np.diag(case.noise))[:, None]
    out = []
    for basis in ("harmonic", False):
        f = Fisher(_params(case, nsims=1), basis=basis, mask=case.mask, noise_cov1=case.noise, cls_data=case.cls,
                   beam=case.beam, pixel_filter=flt)
        f.run()
        sp = Spectra(fisher=f, maps1=x)
        sp.run()
        out.append((f.get_fisher_matrix(), sp.qml_results + sp.qml_noise_bias, sp.qml_noise_bias))
    tol = 1e-10 if name == "T" else 1e-8
    for k, (a, b) in enumerate(zip(out[0], out[1])):
        assert _rel(a, b) < (tol if k == 0 else 100 * tol), k
