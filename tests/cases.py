"""Input builders for the golden-fixture cases (shared by oracle and GPU tests).

Each builder turns a committed fixture under ``tests/golden`` into the
reference's in-memory inputs: mask, l-indexed physical C_l, (3, L+1) beam,
reduced noise covariance, reduced maps.  Reading conventions follow the
reference: ``readcl`` maps columns by header label and converts D_l -> C_l
with 2 pi / (l (l+1)) (``cosmocore/in_out.py:264-381``); simple HEALPix map
files pick the column T=0, Q/E=1, U/B=2 (``in_out.py:566-590``).
"""

from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name: str):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def untril(n: int, t: np.ndarray) -> np.ndarray:
    a = np.zeros((n, n))
    a[np.tril_indices(n)] = t
    return a + np.tril(a, -1).T


def readcl_arrays(labels, arr, lmax: int, dl: bool) -> dict:
    labels = [str(x) for x in labels]
    ell_col = [i for i, lab in enumerate(labels) if lab.lower() == "ell"][0]
    ells = arr[:, ell_col].astype(int)
    keep = (ells >= 0) & (ells <= lmax)
    out = {}
    for i, lab in enumerate(labels):
        if i == ell_col:
            continue
        o = np.zeros(lmax + 1)
        o[ells[keep]] = arr[keep, i]
        out[lab] = o
    if dl:
        ell = np.arange(lmax + 1.0)
        fac = np.zeros(lmax + 1)
        fac[1:] = 2 * np.pi / (ell[1:] * (ell[1:] + 1))
        out = {k: v * fac for k, v in out.items()}
    return out


@dataclass
class Case:
    name: str
    nside: int
    spins: list
    labels: list
    mask: np.ndarray  # (npix, ncomp)
    cls: dict
    beam: np.ndarray  # (3, >= L+1)
    lmax_signal: int
    lmin: int = 2
    lmax: int = 2
    nest: bool = False
    noise: np.ndarray = None
    maps: np.ndarray = None
    convention: str = "Cl"
    expect: dict = field(default_factory=dict)


def case_signal_tqu() -> Case:
    nside, lmax = 4, 8
    ell = np.arange(2, lmax + 1)

    def z(a):
        return np.array([float("%.16e" % v) for v in np.r_[0.0, 0.0, a]])

    cls = {"TT": z(1e-4 / ell**2), "EE": z(0.5e-4 / ell**2), "BB": z(0.1e-4 / ell**2),
           "TE": z(0.3e-4 / ell**2), "TB": np.zeros(lmax + 1), "EB": np.zeros(lmax + 1)}
    b = np.zeros(4 * nside + 1)
    b[: nside + 1] = 1.0
    el = np.arange(nside + 1, 3 * nside + 1)
    b[nside + 1: 3 * nside + 1] = 0.5 * (1.0 + np.cos((el - nside) * np.pi / (2 * nside)))
    d = load("signal_tqu_nside4_nest.npz")
    return Case("signal_tqu", nside, [0, 2], ["T", "E", "B"], np.ones((192, 3)), cls,
                np.vstack([b, b, b])[:, : lmax + 1], lmax, nest=True,
                expect={"S": untril(int(d["n"]), d["tril"])})


def case_signal_t1t2() -> Case:
    nside, lmax = 4, 8
    ell = np.arange(2, lmax + 1)

    def z(a):
        return np.array([float("%.16e" % v) for v in np.r_[0.0, 0.0, a]])

    cls = {"T1T1": z(1e-4 / ell**2), "T2T2": z(0.8e-4 / ell**2), "T1T2": z(0.5e-4 / ell**2)}
    d = load("signal_t1t2_nside4_nest.npz")
    return Case("signal_t1t2", nside, [0, 0], ["T1", "T2"], np.ones((192, 2)), cls,
                np.ones((3, lmax + 1)), lmax, nest=True,
                expect={"S": untril(int(d["n"]), d["tril"])})


def case_fortran_b() -> Case:
    d = load("fortran_nside8_B.npz")
    Ls = 32
    mask = d["mask"][:, None]
    act = np.nonzero(mask[:, 0] > 0.5)[0]
    cls = readcl_arrays(d["dls_labels"], d["dls"], Ls, dl=True)
    n = act.size
    return Case("fortran_B", 8, [0], ["B"], mask, cls, d["beam"], Ls, lmin=2, lmax=24,
                noise=np.diag(d["ncov_diag"][act]), maps=d["map"][2][act][:, None],
                convention="Dl",
                expect={"S": untril(n, d["signal_tril"]), "invcov_diag": d["invcov_diag"],
                        "fisher": d["fisher"], "spectra": d["spectra"][:, 1],
                        "noise_bias": d["noise_bias"][:, 1], "geometry": d["geometry"]})


_NSIDE4 = {"T": ([0], ["T"]), "QU": ([2], ["E", "B"]), "TQU": ([0, 2], ["T", "E", "B"]),
           "TEB": ([0, 0, 0], ["T", "E", "B"])}


def case_fisher_nside4(name: str) -> Case:
    spins, labels = _NSIDE4[name]
    d = load(f"fisher_nside4_{name}.npz")
    mask = d["mask"].T
    Ls = 16
    cls = readcl_arrays(d["dls_labels"], d["dls"], Ls, dl=True)
    idx = np.concatenate([np.nonzero(mask[:, c] > 0.5)[0] + c * 192 for c in range(mask.shape[1])])
    return Case(f"fisher_nside4_{name}", 4, spins, labels, mask, cls, d["beam"], Ls, lmin=2, lmax=8,
                noise=np.diag(d["ncov_diag_full"][idx]), expect={"fisher": d["ref_fisher"]})


def case_v110() -> Case:
    nside, Ls = 8, 16
    npix = 12 * nside**2
    mask = np.ones((npix, 1))
    mask[: npix // 2] = 0.0
    ell = np.arange(0, 4 * nside + 1)
    tt = 1000.0 * np.exp(-((ell / 50.0) ** 2)) + 1.0
    tt = np.array([float("%.6e" % v) for v in tt])[: Ls + 1]
    sigma = math.radians(60.0 / 60.0) / math.sqrt(8.0 * math.log(2.0))
    el = np.arange(Ls + 1)
    g = np.exp(-0.5 * el * (el + 1) * sigma**2)
    pol = math.exp(2.0 * sigma**2)
    n = npix // 2
    return Case("v110", nside, [0], ["T"], mask, {"TT": tt}, np.vstack([g, g * pol, g * pol]), Ls,
                lmin=2, lmax=10, noise=0.1 * np.eye(n),
                maps=np.random.default_rng(20260908).standard_normal(n)[:, None],
                expect={"logdet": 1612.9143724664957, "quadform": 1581.7900899163615,
                        "cov_trace": 8321344.223152173, "cov_fro": 673367.4529629507})


def case_picslike_b():
    """(Case, list of blended per-point C_l dicts, expected results rows)."""
    d = load("picslike_nside8_B.npz")
    Ls, lmin, lmax = 32, 2, 16
    mask = d["mask"][:, None]
    act = np.nonzero(mask[:, 0] > 0.5)[0]
    nc = untril(768, d["ncov_tril"])[np.ix_(act, act)]
    fid = readcl_arrays(d["fid_labels"], d["fid"], Ls, dl=True)
    points = []
    for r in np.linspace(0.0, 0.005, 401)[:3]:
        ps = readcl_arrays(d["theory_labels"], d["theory_%s" % ("%.6g" % r)], Ls, dl=True)
        bl = {}
        for k, v in ps.items():
            if k not in fid:
                bl[k] = v.copy()
                continue
            x = fid[k].copy()
            x[lmin: lmax + 1] = v[lmin: lmax + 1]
            bl[k] = x
        points.append(bl)
    case = Case("picslike_B", 8, [0], ["B"], mask, fid, d["beam"], Ls, lmin=lmin, lmax=lmax,
                noise=nc, maps=d["map"][2][act][:, None])
    return case, points, d["results"]
