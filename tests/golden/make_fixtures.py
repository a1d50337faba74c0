"""Write the golden fixtures under tests/golden/ from the reference's data files.

Run in a container where the reference checkout is mounted (default
``/root/reference``; override with ``COSMOFORGE_REFERENCE``).  Only data
files that the reference's own tests hold are read -- no reference code is
imported or run.  Symmetric matrices are stored as their lower triangle,
diagonal noise matrices as their diagonal.

Fixtures written (all float64, ``np.savez_compressed``):

* ``signal_tqu_nside4_nest.npz`` -- ``cosmocore/tests/data/ref_TQU_signal.dat``
  (``tests/test_signal.py:30-119``: nside 4 NESTED, TQU, lmax 8, cosine beam).
* ``signal_t1t2_nside4_nest.npz`` -- ``cosmocore/tests/data/ref_T1T2_signal.dat``
  (``tests/test_signal.py:122-212``).
* ``fortran_nside8_B.npz`` -- ``qube/tests/data/nside8/B/fortran_reference``
  (``qube/tests/test_fortran_validation.py``).
* ``fisher_nside4_{T,QU,TQU,TEB}.npz`` -- ``qube/tests/data/nside4/*``
  (``qube/tests/test_fisher.py:83-129``).
* ``picslike_nside8_B.npz`` -- ``picslike/tests/data/nside8/B/fortran_reference``
  (``picslike/tests/test_fortran_validation.py:162-426``).
* ``files/`` -- data files copied byte for byte (text files cut to their first
  ``CL_ROWS`` multipole rows, the multi-simulation FITS file to its first two
  ``SIM_XXX`` units) for the file-reader tests:
  ``nside4_TQU_mask.fits`` and ``nside4_TQU_geometry.dat``
  (``picslike/tests/data/nside4/TQU``: the reference's ``output_geometry`` of
  that full-sky RING geometry), ``nside4_all_sims_TQU_2.fits``
  (``nside4/maps/all_sims_TQU.fits``), and from ``nside8/B/fortran_reference``
  ``B_mask.fits``, ``B_beam.fits``, ``B_map.fits``, ``B_input_dls.txt``,
  ``B_fiducial_dls.txt`` and ``B_theory_r{0,1.25e-05,2.5e-05}.txt``.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import fitsio  # noqa: E402

REF = os.environ.get("COSMOFORGE_REFERENCE", "/root/reference")
CC = os.path.join(REF, "src/cosmoforge.cosmocore/tests/data")
QB = os.path.join(REF, "src/cosmoforge.qube")
PL = os.path.join(REF, "src/cosmoforge.picslike/tests/data")


def tril(a):
    return a[np.tril_indices(a.shape[0])]


def read_cl_text(path):
    """Header labels + numeric columns of a reference C_l / D_l text file."""
    with open(path) as fh:
        hdr = fh.readline().lstrip("#").split()
    return np.array(hdr), np.loadtxt(path)


CL_ROWS = 63            # multipoles 2..64 of a C_l text file


def copy_files(out: str) -> None:
    import shutil

    dst = os.path.join(out, "files")
    os.makedirs(dst, exist_ok=True)
    t4 = os.path.join(PL, "nside4/TQU")
    fr = os.path.join(PL, "nside8/B/fortran_reference")
    for src, name in ((os.path.join(t4, "mask.fits"), "nside4_TQU_mask.fits"),
                      (os.path.join(t4, "geometry.dat"), "nside4_TQU_geometry.dat"),
                      (os.path.join(fr, "mask.fits"), "B_mask.fits"),
                      (os.path.join(fr, "beam.fits"), "B_beam.fits"),
                      (os.path.join(fr, "map.fits"), "B_map.fits")):
        shutil.copyfile(src, os.path.join(dst, name))
    texts = [(os.path.join(fr, "input_dls.txt"), "B_input_dls.txt"),
             (os.path.join(fr, "fiducial_dls.txt"), "B_fiducial_dls.txt")]
    for tag in ("0", "1.25e-05", "2.5e-05"):
        texts.append((os.path.join(fr, f"theory_spectra/fine_dls_r_likelihood_r{tag}.txt"), f"B_theory_r{tag}.txt"))
    for src, name in texts:
        with open(src) as fh:
            lines = fh.readlines()[: 1 + CL_ROWS]
        with open(os.path.join(dst, name), "w") as fh:
            fh.writelines(lines)
    # primary unit + SIM_000 + SIM_001: each unit is a whole number of 2880-byte records
    raw = open(os.path.join(PL, "nside4/maps/all_sims_TQU.fits"), "rb").read()
    off, units = 0, 0
    while units < 3:
        nax, bitpix, dims = 0, 8, {}
        while True:
            rec = raw[off: off + 2880].decode("ascii")
            off += 2880
            cards = [rec[i: i + 80] for i in range(0, 2880, 80)]
            for c in cards:
                key = c[:8].strip()
                if key == "BITPIX":
                    bitpix = int(c[10:].split("/")[0])
                elif key == "NAXIS":
                    nax = int(c[10:].split("/")[0])
                elif key.startswith("NAXIS"):
                    dims[key] = int(c[10:].split("/")[0])
            if any(c[:8].strip() == "END" for c in cards):
                break
        nbytes = abs(bitpix) // 8 * int(np.prod([dims[f"NAXIS{a}"] for a in range(1, nax + 1)])) if nax else 0
        off += -(-nbytes // 2880) * 2880
        units += 1
    with open(os.path.join(dst, "nside4_all_sims_TQU_2.fits"), "wb") as fh:
        fh.write(raw[:off])


def main():
    out = HERE
    copy_files(out)
    a = np.loadtxt(os.path.join(CC, "ref_TQU_signal.dat"))
    np.savez_compressed(os.path.join(out, "signal_tqu_nside4_nest.npz"), n=a.shape[0], tril=tril(a))
    a = np.loadtxt(os.path.join(CC, "ref_T1T2_signal.dat"))
    np.savez_compressed(os.path.join(out, "signal_t1t2_nside4_nest.npz"), n=a.shape[0], tril=tril(a))

    fr = os.path.join(QB, "tests/data/nside8/B/fortran_reference")
    hdr, dls = read_cl_text(os.path.join(fr, "input_dls.txt"))
    ncov = np.fromfile(os.path.join(fr, "input_ncov.bin")).reshape(768, 768)
    assert not np.any(ncov - np.diag(np.diag(ncov)))
    sig = np.fromfile(os.path.join(fr, "signal.bin")).reshape(768, 768)
    inv = np.fromfile(os.path.join(fr, "invcov.bin")).reshape(768, 768)
    np.savez_compressed(
        os.path.join(out, "fortran_nside8_B.npz"),
        mask=fitsio.read_table_columns(os.path.join(fr, "mask.fits"), 1)[0],
        dls_labels=hdr, dls=dls,
        beam=fitsio.read_table_columns(os.path.join(fr, "beam.fits"), 3),
        ncov_diag=np.diag(ncov).copy(),
        map=fitsio.read_table_columns(os.path.join(fr, "map.fits"), 3),
        signal_tril=tril(sig), invcov_diag=np.diag(inv).copy(),
        fisher=np.loadtxt(os.path.join(fr, "fisher.dat")),
        spectra=np.loadtxt(os.path.join(fr, "spectra.dat")),
        noise_bias=np.loadtxt(os.path.join(fr, "noise_bias.dat")),
        geometry=np.loadtxt(os.path.join(fr, "geometry.dat"), skiprows=1),
    )

    beam = fitsio.read_table_columns(os.path.join(QB, "scripts/beam_440TP_pixwin16.fits"), 3)
    for name, ncomp in (("T", 1), ("QU", 2), ("TQU", 3), ("TEB", 3)):
        d = os.path.join(QB, "tests/data/nside4", name)
        hdr, dls = read_cl_text(os.path.join(d, "dls.txt"))
        nc = np.fromfile(os.path.join(d, "NCVM1.bin")).reshape(ncomp * 192, ncomp * 192)
        assert not np.any(nc - np.diag(np.diag(nc)))
        np.savez_compressed(
            os.path.join(out, f"fisher_nside4_{name}.npz"),
            mask=fitsio.read_table_columns(os.path.join(d, "mask.fits"), ncomp),
            dls_labels=hdr, dls=dls, beam=beam, ncov_diag_full=np.diag(nc).copy(),
            ref_fisher=np.loadtxt(os.path.join(d, "ref_fisher.dat")),
        )

    pr = os.path.join(PL, "nside8/B/fortran_reference")
    nc = np.fromfile(os.path.join(pr, "ncov_B_only.bin")).reshape(768, 768)
    theory = {}
    for r in np.linspace(0.0, 0.005, 401)[:3]:
        tag = "%.6g" % r
        h, arr = read_cl_text(os.path.join(pr, f"theory_spectra/fine_dls_r_likelihood_r{tag}.txt"))
        theory[f"theory_{tag}"] = arr
        theory["theory_labels"] = h
    fh, fid = read_cl_text(os.path.join(pr, "fiducial_dls.txt"))
    res = np.loadtxt(os.path.join(pr, "results.txt"))
    np.savez_compressed(
        os.path.join(out, "picslike_nside8_B.npz"),
        mask=fitsio.read_table_columns(os.path.join(pr, "mask.fits"), 1)[0],
        beam=fitsio.read_table_columns(os.path.join(pr, "beam.fits"), 3),
        ncov_tril=tril(nc), map=fitsio.read_table_columns(os.path.join(pr, "map.fits"), 3),
        fid_labels=fh, fid=fid, results=res[res[:, 1] < 3], **theory,
    )
    print("fixtures written to", out)


if __name__ == "__main__":
    main()
