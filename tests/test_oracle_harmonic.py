"""The oracle's harmonic modes (CPU): the Wigner small-d sum against
closed forms, and sum_m x_lm y_lm^T against the oracle's pixel-space
derivative of every spectrum (spin-0 exact by the addition theorem; spin-2
to the ~1e-8 that the reference's acos-based rotation angles carry)."""

import math

import numpy as np
import pytest
from scipy.special import lpmv

from oracle import harmonic as H
from oracle import healpix
from oracle import kernels as K


def test_wigner_d_closed_forms():
    th = np.linspace(0.05, 3.1, 17)
    for ell in range(0, 9):
        for m in range(0, ell + 1):
            # d^l_{m,0} = sqrt((l-m)!/(l+m)!) P_l^m(cos) up to the Condon-Shortley sign
            ref = math.sqrt(math.factorial(ell - m) / math.factorial(ell + m)) * lpmv(m, ell, np.cos(th))
            np.testing.assert_allclose(np.abs(H.wigner_d(ell, m, 0, th)), np.abs(ref), atol=1e-13)
    # d^2_{2,2} = cos^4(t/2), d^2_{2,-2} = sin^4(t/2), d^l_{m,s} = (-1)^{m-s} d^l_{s,m}
    np.testing.assert_allclose(H.wigner_d(2, 2, 2, th), np.cos(th / 2) ** 4, atol=1e-14)
    np.testing.assert_allclose(H.wigner_d(2, 2, -2, th), np.sin(th / 2) ** 4, atol=1e-14)
    for ell, m, s in ((5, 1, 2), (6, 0, -2), (7, 3, 2)):
        np.testing.assert_allclose(H.wigner_d(ell, m, s, th), (-1) ** (m - s) * H.wigner_d(ell, s, m, th),
                                   atol=1e-13)
    # the Jacobi-polynomial form against the finite factorial sum
    for ell in range(0, 11):
        for m in range(-ell, ell + 1):
            for s in (-2, 0, 2):
                np.testing.assert_allclose(H.wigner_d(ell, m, s, th), H.wigner_d_sum(ell, m, s, th), atol=1e-12)
    # unitarity: sum_m d^l_{m,s}^2 = 1, also at high l
    for ell in (2, 5, 8, 40, 95):
        for s in (0, 2, -2):
            tot = sum(H.wigner_d(ell, m, s, th) ** 2 for m in range(-ell, ell + 1))
            np.testing.assert_allclose(tot, 1.0, atol=1e-12)


@pytest.mark.parametrize("ell", [2, 3, 7])
def test_modes_reproduce_pixel_derivatives(ell):
    nside = 4
    pix = np.arange(0, 192, 5)
    mask = np.zeros((192, 3))
    mask[pix, :] = 1
    lay = K.layout_from_mask(mask, [0, 2], ["T", "E", "B"], nside)
    vec = healpix.pointings(nside, pix)
    lmin, lmax = 2, 8
    VT = H.modes_spin0(vec, lmin, lmax)
    VE, VB = H.modes_spin2(vec, lmin, lmax)
    nf = pix.size
    n = 3 * nf
    sel = H.mode_ells(lmin, lmax) == ell
    assert sel.sum() == 2 * ell + 1
    G = {"T": np.zeros((n, sel.sum())), "E": np.zeros((n, sel.sum())), "B": np.zeros((n, sel.sum()))}
    G["T"][:nf] = VT[:, sel]
    G["E"][nf:] = VE[:, sel]
    G["B"][nf:] = VB[:, sel]
    for fi, fj, mode, lab in lay.spectra_list():
        w = np.zeros(lmax + 1)
        w[ell] = 1.0
        D = K.binned_derivative(lay, fi, fj, mode, w)
        x, y = lab[0], lab[1]
        A = G[x] @ G[y].T
        if x != y:
            A = A + A.T
        tol = 1e-13 if lab == "TT" else 1e-7
        assert np.abs(A - D).max() < tol * np.abs(D).max(), lab


@pytest.mark.parametrize("name", ["T", "TQU"])
def test_mode_fisher_and_qml_equal_pixel_formulation(name):
    """The harmonic-mode restatement of F, q and nb against the oracle's
    pixel-space formulation on the nside-4 reference cases."""
    from oracle import pipeline as P
    from tests import cases
    from tests import oracle_run as R

    c = cases.case_fisher_nside4(name)
    c.lmax = 6
    lay = R.layout(c)
    cls_s, beams = R.smoothed(c, lay)
    prm = R.params(c, lay, beams)
    F, Cinv, S = P.fisher(lay, cls_s, c.lmax_signal, c.noise, prm)
    Fh = H.fisher_modes(lay, Cinv, prm, c.lmin, c.lmax)
    tol = 1e-12 if name == "T" else 1e-9
    assert np.abs(Fh - F).max() < tol * np.abs(F).max()
    x = np.linalg.cholesky(S + c.noise) @ np.random.default_rng(4).standard_normal((c.noise.shape[0], 2))
    q, nb, _, _ = P.qml(lay, Cinv, c.noise, x, F, prm)
    qh, nbh = H.qml_modes(lay, Cinv, c.noise, x, prm, c.lmin, c.lmax)
    # nb = 1/2 <G, dC_b> cancels: its terms are ~1e3 times nb for T here, so
    # the two summation orders agree to ~1e-12 of the terms
    assert np.abs(nbh - nb).max() < 1e-8 * np.abs(nb).max()
    assert np.abs(qh - q).max() < 10 * tol * np.abs(q).max()
