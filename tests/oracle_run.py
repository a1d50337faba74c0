"""Run the FP64 oracle on a ``tests.cases.Case`` (tests only)."""

from __future__ import annotations

import numpy as np

from oracle import kernels as K
from oracle import pipeline as P


def layout(case):
    return K.layout_from_mask(case.mask, case.spins, case.labels, case.nside, nest=case.nest)


def smoothed(case, lay=None, cls=None):
    lay = lay or layout(case)
    beams = P.beam_by_label(lay, case.beam, case.lmax_signal)
    return P.smoothed_cls(lay, cls if cls is not None else case.cls, beams, case.lmax_signal), beams


def params(case, lay, beams):
    return P.parameter_weights(lay, beams, case.lmin, case.lmax, case.lmax_signal,
                               convention=case.convention)


def beam_smoothing(case, lay, beams):
    sm = P.smoothing_factors(lay, beams, case.lmax)
    return np.concatenate([sm[lab][case.lmin: case.lmax + 1] for (*_, lab) in lay.spectra_list()])
