"""Host-side bookkeeping of fields, rows and spectra (no device work).

Semantics follow the reference:

* fields are built from ``spins``/``labels`` and the mask column of each
  field (``src/cosmoforge.cosmocore/cosmocore/core.py:474-519``); active
  pixels are ``mask > 0.5`` in ascending pixel order
  (``cosmocore/geometry.py:49-134``);
* rows are ``[field 0 | field 1 | ...]``, a spin-2 field contributing its
  Q rows then its U rows;
* spectra: auto spectra per field (``TT``; ``EE, BB, EB``), then cross
  spectra per field pair (``fields.py:495-690``, ``spectra_io.py:155-190``);
  under the SYMMETRIC mode spin-2 x spin-2 cross pairs drop their ``CG``
  entry (``spectra_io.py:365-425``).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class FieldInfo:
    spin: int
    labels: list
    pixels: np.ndarray
    offset: int = 0

    @property
    def npix(self) -> int:
        return int(self.pixels.size)

    @property
    def rows(self) -> int:
        return 2 * self.npix if self.spin == 2 else self.npix

    def auto_labels(self):
        if self.spin == 0:
            a = self.labels[0].upper()
            return [a + a]
        e, b = (x.upper() for x in self.labels)
        return [e + e, b + b, e + b]

    def cross_labels(self, other: "FieldInfo"):
        if self.spin == 0 and other.spin == 0:
            return [self.labels[0].upper() + other.labels[0].upper()]
        if self.spin == 0:
            t = self.labels[0].upper()
            return [t + other.labels[0].upper(), t + other.labels[1].upper()]
        if other.spin == 0:
            t = other.labels[0].upper()
            return [t + self.labels[0].upper(), t + self.labels[1].upper()]
        return [a.upper() + b.upper() for a in self.labels for b in other.labels]


@dataclass
class Spectrum:
    fi: int
    fj: int
    mode: int
    label: str


@dataclass
class Layout:
    fields: list
    spectra_map: dict = field(default_factory=dict)

    def __post_init__(self):
        off = 0
        for f in self.fields:
            f.offset = off
            off += f.rows
        self.n = off
        self.spectra_map = {}
        for i, f in enumerate(self.fields):
            for m, lab in enumerate(f.auto_labels()):
                self.spectra_map[(i, i, m)] = lab
        for i in range(len(self.fields)):
            for j in range(i + 1, len(self.fields)):
                for m, lab in enumerate(self.fields[i].cross_labels(self.fields[j])):
                    self.spectra_map[(i, j, m)] = lab

    @property
    def labels(self):
        return list(self.spectra_map.values())

    def spectra_list(self):
        out = []
        for (fi, fj, m), lab in self.spectra_map.items():
            if fi != fj and self.fields[fi].spin == 2 and self.fields[fj].spin == 2 and m == 2:
                continue
            out.append(Spectrum(fi, fj, m, lab))
        return out

    def split_label(self, label: str, i: int, j: int):
        cands = list(self.fields[i].labels) + list(self.fields[j].labels)
        s = {(a, b) for a in cands for b in cands if (a + b).upper() == label}
        if len(s) != 1:
            raise ValueError(
                f"spectrum label {label!r} does not split uniquely into the field-map "
                f"labels {cands}; got {sorted(s)}")
        return s.pop()

    def active_index(self, npix: int) -> np.ndarray:
        """Positions of the rows in the flattened (ncomponents * npix) vector."""
        out, comp = [], 0
        for f in self.fields:
            for _ in range(2 if f.spin == 2 else 1):
                out.append(f.pixels + comp * npix)
                comp += 1
        return np.concatenate(out) if out else np.zeros(0, dtype=np.int64)


def build_layout(mask: np.ndarray, spins, labels) -> Layout:
    mask = np.asarray(mask, dtype=np.float64)
    if mask.ndim == 1:
        mask = mask[:, None]
    fields, counter = [], 0
    for spin in spins:
        if spin == 0:
            labs = [labels[counter]]
            counter += 1
        elif spin == 2:
            labs = [labels[counter], labels[counter + 1]]
            counter += 2
        else:
            raise ValueError(f"unsupported spin {spin}")
        pix = np.nonzero(mask[:, counter - 1] > 0.5)[0].astype(np.int64)
        fields.append(FieldInfo(spin, labs, pix))
    return Layout(fields)
