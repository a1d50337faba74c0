"""File readers and writers of ``cosmoforge_amd.in_out`` (CPU only).

Pinned against the reference's own data files under ``tests/golden/files``
(copied by ``tests/golden/make_fixtures.py``): its FITS masks, maps, beams
and multi-simulation images (read independently by ``oracle.fitsio``), its
``geometry.dat`` (written by the reference's ``output_geometry``,
``cosmocore/in_out.py:209-252``) and its D_l text files; the D_l -> C_l and
reduction rules are pinned by closed forms and by index-encoding matrices
whose every entry names its own row and column.
"""

import os

import numpy as np
import pytest

from cosmoforge_amd import in_out
from cosmoforge_amd.core import Inputs
from oracle import fitsio, healpix

from .cases import GOLDEN, case_picslike_b, load, untril

FILES = os.path.join(GOLDEN, "files")


def f(name):
    return os.path.join(FILES, name)


# --------------------------------------------------------------------------
# FITS
# --------------------------------------------------------------------------

@pytest.mark.parametrize("name", ["nside4_TQU_mask.fits", "B_mask.fits", "B_map.fits"])
def test_read_healpix_matches_oracle_reader(name):
    ours = in_out.read_healpix(f(name))
    ref = fitsio.read_table_columns(f(name))
    assert ours.shape == ref.shape
    assert np.array_equal(ours, ref)


def test_read_mask_shape_and_fixture_values():
    m = in_out.read_mask(f("nside4_TQU_mask.fits"), np.empty((3, 0)))
    assert m.shape == (192, 3)
    assert np.array_equal(m.T, load("fisher_nside4_TQU.npz")["mask"])
    m1 = in_out.read_mask(f("B_mask.fits"), np.empty((1, 0)))
    assert np.array_equal(m1[:, 0], load("picslike_nside8_B.npz")["mask"])


def test_read_cl_fits_beam_matches_fixture():
    b = in_out.read_cl_fits(f("B_beam.fits"))
    assert b.shape[0] >= 3
    assert np.array_equal(b[:3], load("picslike_nside8_B.npz")["beam"][:3])


def test_read_maps_multi_simulation_images():
    pix = [np.arange(192), np.arange(0, 192, 3), np.arange(5, 150)]
    n = sum(p.size for p in pix)
    maps = np.zeros((n, 2))
    in_out.read_maps(maps, f("nside4_all_sims_TQU_2.fits"), pix, ["T", "Q", "U"])
    imgs = fitsio.read_image_hdus(f("nside4_all_sims_TQU_2.fits"))
    exp = np.zeros((n, 2))
    for s in range(2):
        exp[:, s] = np.concatenate([imgs[s][k][p] for k, p in enumerate(pix)])
    assert np.array_equal(maps, exp)
    with pytest.raises(ValueError, match="not found"):
        in_out.read_maps(np.zeros((192, 1)), f("nside4_all_sims_TQU_2.fits"), [np.arange(192)], ["E"])


def test_read_maps_simple_healpix_picks_column_by_label():
    pix = [np.arange(0, 768, 2)]
    maps = np.zeros((pix[0].size, 1))
    in_out.read_maps(maps, f("B_map.fits"), pix, ["B"])
    assert np.array_equal(maps[:, 0], load("picslike_nside8_B.npz")["map"][2][pix[0]])
    with pytest.raises(ValueError, match="only supports nsims=1"):
        in_out.read_maps(np.zeros((384, 2)), f("B_map.fits"), pix, ["B"])
    with pytest.raises(ValueError, match="Unknown field label"):
        in_out.read_maps(np.zeros((384, 1)), f("B_map.fits"), pix, ["X"])


def test_read_maps_numpy_stack(tmp_path):
    rng = np.random.default_rng(3)
    stack = rng.standard_normal((3, 48, 4))
    path = str(tmp_path / "maps.npy")
    np.save(path, stack)
    pix = [np.array([0, 5, 7]), np.array([1, 2]), np.array([47])]
    maps = np.zeros((6, 3))
    in_out.read_maps(maps, path, pix, ["T", "Q", "U"])
    exp = np.concatenate([stack[k][p, :3] for k, p in enumerate(pix)])
    assert np.array_equal(maps, exp)


def _card(key, value):
    if isinstance(value, str):
        v = f"'{value:<8}'"
        return f"{key:<8}= {v:<70}"
    if isinstance(value, bool):
        return f"{key:<8}= {'T' if value else 'F':>20}" + " " * 50
    return f"{key:<8}= {value:>20}" + " " * 50


def _write_table(path, cols, names, extra):
    """A BINTABLE written per the FITS standard (4.0, section 7.3): big-endian
    float64 columns of one value per row."""
    prim = [_card("SIMPLE", True), _card("BITPIX", 8), _card("NAXIS", 0), _card("EXTEND", True)]
    nrow = cols[0].size
    hdr = [_card("XTENSION", "BINTABLE"), _card("BITPIX", 8), _card("NAXIS", 2),
           _card("NAXIS1", 8 * len(cols)), _card("NAXIS2", nrow), _card("PCOUNT", 0), _card("GCOUNT", 1),
           _card("TFIELDS", len(cols))]
    for i, (nm, c) in enumerate(zip(names, cols), 1):
        hdr += [_card(f"TTYPE{i}", nm), _card(f"TFORM{i}", "K" if c.dtype.kind == "i" else "D")]
    hdr += [_card(k, v) for k, v in extra.items()]

    def block(cards):
        s = "".join(cards) + "END".ljust(80)
        s += " " * (-len(s) % 2880)
        return s.encode("ascii")

    rows = np.empty(nrow, dtype=[(f"c{i}", ">i8" if c.dtype.kind == "i" else ">f8") for i, c in enumerate(cols)])
    for i, c in enumerate(cols):
        rows[f"c{i}"] = c
    data = rows.tobytes()
    data += b"\0" * (-len(data) % 2880)
    with open(path, "wb") as fh:
        fh.write(block(prim) + block(hdr) + data)


def test_partial_sky_and_nested_files(tmp_path):
    nside = 4
    npix = 192
    rng = np.random.default_rng(5)
    full = rng.standard_normal((2, npix))
    # NESTED full-sky file: values stored in NESTED order
    n2r = in_out.nest2ring(nside, np.arange(npix))
    nest_vals = full[:, n2r]
    path = str(tmp_path / "nest.fits")
    _write_table(path, list(nest_vals), ["T", "Q"], {"ORDERING": "NESTED", "NSIDE": nside,
                                                     "INDXSCHM": "IMPLICIT"})
    assert np.array_equal(in_out.read_healpix(path, nest=False), full)
    assert np.array_equal(in_out.read_healpix(path, nest=True), nest_vals)
    assert np.array_equal(in_out.read_healpix(path, field=1), full[1:2])
    # explicit (partial-sky) RING file
    pix = np.sort(rng.choice(npix, 40, replace=False)).astype(np.int64)
    path = str(tmp_path / "partial.fits")
    _write_table(path, [pix, full[0, pix]], ["PIXEL", "SIGNAL"], {"ORDERING": "RING", "NSIDE": nside,
                                                                 "INDXSCHM": "EXPLICIT"})
    m = in_out.read_healpix(path)
    assert m.shape == (1, npix)
    assert np.array_equal(m[0, pix], full[0, pix])
    others = np.setdiff1d(np.arange(npix), pix)
    assert np.all(m[0, others] == in_out.UNSEEN)
    with pytest.raises(ValueError, match="not present"):
        in_out.read_healpix(path, field=1)


def test_nest2ring_is_the_healpix_pixel_map():
    for nside in (1, 2, 4, 8):
        npix = 12 * nside * nside
        r = in_out.nest2ring(nside, np.arange(npix))
        assert np.array_equal(np.sort(r), np.arange(npix))
        zr = np.array([healpix.pix2zphi(nside, int(x), False)[:2] for x in r])
        zn = np.array([healpix.pix2zphi(nside, int(x), True)[:2] for x in range(npix)])
        assert np.allclose(zr, zn, rtol=0, atol=1e-13)
    with pytest.raises(ValueError, match="power-of-two"):
        in_out.nest2ring(3, [0])


def test_not_a_fits_file(tmp_path):
    p = tmp_path / "x.fits"
    p.write_bytes(b" " * 2880)
    with pytest.raises(ValueError, match="not a FITS"):
        in_out.read_healpix(str(p))


# --------------------------------------------------------------------------
# covariance binaries
# --------------------------------------------------------------------------

def test_read_covmat_reduces_rows_and_columns(tmp_path):
    npix, nmaps = 12, 3
    full = npix * nmaps
    # entry (r, c) encodes its own row and column; not symmetric, so a
    # transposed read fails
    enc = (np.arange(full)[:, None] * 1000 + np.arange(full)[None, :]).astype(np.float64)
    path = str(tmp_path / "ncvm.bin")
    enc.tofile(path)
    active = np.array([0, 3, 4, 11, 12 + 2, 12 + 5, 24 + 0, 24 + 11])
    C = in_out.read_covmat(path, npix, nmaps, active)
    for i, a in enumerate(active):
        for j, b in enumerate(active):
            assert C[i, j] == a * 1000 + b
    # unsorted active order and several row blocks
    old = in_out._GATHER_BYTES
    try:
        in_out._GATHER_BYTES = 2 * 8 * full
        perm = active[[3, 0, 7, 1, 6, 2, 5, 4]]
        C2 = in_out.read_covmat(path, npix, nmaps, perm, np.empty((8, 8)))
    finally:
        in_out._GATHER_BYTES = old
    assert np.array_equal(C2, perm[:, None] * 1000.0 + perm[None, :])
    assert in_out.read_covmat(path, npix, nmaps, np.zeros(0, int)).shape == (0, 0)
    with pytest.raises(ValueError, match="expected"):
        in_out.read_covmat(path, npix + 1, nmaps, active)
    with pytest.raises(ValueError, match="out of range"):
        in_out.read_covmat(path, npix, nmaps, np.array([full]))


def test_reduced_covmat_round_trip_and_size_error(tmp_path):
    rng = np.random.default_rng(1)
    A = rng.standard_normal((7, 7))
    path = str(tmp_path / "sub" / "red.bin")
    in_out.write_covmat_reduced(path, A)
    assert os.path.getsize(path) == 7 * 7 * 8
    B = in_out.read_covmat_reduced(path, np.empty((7, 7)))
    assert np.array_equal(A, B)
    with pytest.raises(ValueError, match="Invalid reduced covariance matrix size"):
        in_out.read_covmat_reduced(path, np.empty((6, 6)))


# --------------------------------------------------------------------------
# spectra text files
# --------------------------------------------------------------------------

def test_convert_spectra_normalization_closed_forms():
    ell = np.arange(6.0)
    x = np.array([0.0, 7.0, 1.0, 2.0, 3.0, 4.0])
    d = in_out.convert_spectra_normalization({"TT": x.copy()}, "Cl", "Dl")
    assert np.allclose(d["TT"], x * ell * (ell + 1) / (2 * np.pi), rtol=1e-15, atol=0)
    c = in_out.convert_spectra_normalization({"TT": d["TT"].copy()}, "dl", "CL")
    assert c["TT"][0] == 0.0
    assert np.allclose(c["TT"][1:], x[1:], rtol=1e-14, atol=0)
    same = {"TT": x.copy()}
    assert in_out.convert_spectra_normalization(same, "Dl", "dl") is same
    with pytest.raises(ValueError, match="Unknown spectra convention"):
        in_out.convert_spectra_normalization({"TT": x}, "Cl", "XX")


class _P:
    lmax = 4
    input_convention = "Cl"


def test_readcl_ell_column_and_default_start(tmp_path):
    p1 = tmp_path / "a.dat"
    p1.write_text("# ell TT EE\n2 1.0 10\n3 2.0 20\n4 3.0 30\n5 4.0 40\n")
    out = in_out.readcl(str(p1), _P)
    assert np.array_equal(out["TT"], [0, 0, 1, 2, 3])
    assert np.array_equal(out["EE"], [0, 0, 10, 20, 30])
    assert "ell" not in out
    # rows placed by their ell value, holes zero
    p2 = tmp_path / "b.dat"
    p2.write_text("#ELL TT\n0 9\n3 2\n1 5\n")
    assert np.array_equal(in_out.readcl(str(p2), _P)["TT"], [9, 5, 0, 2, 0])
    # no ell column: rows start at l = 2
    p3 = tmp_path / "c.dat"
    p3.write_text("# TT\n1\n2\n3\n4\n")
    assert np.array_equal(in_out.readcl(str(p3), _P, lmax=3)["TT"], [0, 0, 1, 2])
    assert np.array_equal(in_out.readcl(str(p3), _P, lmax=8)["TT"], [0, 0, 1, 2, 3, 4, 0, 0, 0])
    p4 = tmp_path / "d.dat"
    p4.write_text("TT EE\n1 2\n")
    with pytest.raises(ValueError, match="header starting with '#'"):
        in_out.readcl(str(p4), _P)


def test_readcl_dl_file_matches_reference_conversion():
    class P:
        lmax = 32
        input_convention = "Dl"

    cls = in_out.readcl(f("B_input_dls.txt"), P)
    raw = np.loadtxt(f("B_input_dls.txt"))
    ell = raw[:, 0].astype(int)
    keep = ell <= 32
    for k, lab in enumerate(["TT", "EE", "BB", "TE"], 1):
        exp = np.zeros(33)
        # the factor first, then the product (in_out.py:289-298)
        exp[ell[keep]] = raw[keep, k] * (2 * np.pi / (ell[keep] * (ell[keep] + 1.0)))
        assert np.array_equal(cls[lab], exp)
    assert np.array_equal(raw, load("picslike_nside8_B.npz")["fid"][: raw.shape[0]])


# --------------------------------------------------------------------------
# writers
# --------------------------------------------------------------------------

def test_output_geometry_reproduces_reference_file(tmp_path):
    pix = np.arange(192)
    vec = healpix.pointings(4, pix, nest=False)
    out = str(tmp_path / "g" / "geometry.dat")
    in_out.output_geometry(out, [192, 192], [vec, vec], [pix, pix])
    # byte for byte: the oracle's pointings are the platform libm's, as healpy's
    assert open(out).read() == open(f("nside4_TQU_geometry.dat")).read()
    with pytest.raises(ValueError, match="per field"):
        in_out.output_geometry(out, [192, 192], [vec, vec], [pix, pix, pix])


@pytest.mark.parametrize("path", [None, "", "   "])
def test_writers_noop_on_falsy_path(path, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    in_out.write_covmat_reduced(path, np.eye(3))
    in_out.write_out_matrix(path, np.eye(3))
    in_out.writecl(path, np.arange(6.0).reshape(3, 2))
    in_out.output_geometry(path, [2], [np.zeros((2, 3))], [np.array([0, 1])])
    assert list(tmp_path.rglob("*")) == []


def test_write_out_matrix_and_writecl_formats(tmp_path):
    M = np.array([[1.0, -2.5e-10], [3.0, 4.0]])
    p = str(tmp_path / "m.dat")
    in_out.write_out_matrix(p, M)
    lines = open(p).read().splitlines()
    assert lines[0] == f"{1.0:24.16E}{-2.5e-10:24.16E}"
    assert np.array_equal(np.loadtxt(p), M)
    q = str(tmp_path / "e.dat")
    in_out.writecl(q, M)
    assert open(q).read().splitlines()[1] == "3.0000000000000000e+00 4.0000000000000000e+00"


# --------------------------------------------------------------------------
# Inputs driven by file paths (no device work)
# --------------------------------------------------------------------------

def _b_files(tmp_path):
    d = load("picslike_nside8_B.npz")
    ncov = untril(768, d["ncov_tril"])
    ncov.tofile(str(tmp_path / "ncov.bin"))
    return dict(nside=8, spins=[0], labels=["B"], maskfile=f("B_mask.fits"), ordering="RING",
                inputclfile=f("B_input_dls.txt"), input_convention="Dl",
                covmatfile1=str(tmp_path / "ncov.bin"), lmax_signal=32, lmin=2, lmax=16,
                smoothing_type="file", beam_file=f("B_beam.fits"), apply_pixwin=False, nsims=1,
                inputmapfile1=f("B_map.fits"), outnoisecovmat1=str(tmp_path / "out" / "red.bin"))


def test_inputs_from_files_match_in_memory_inputs(tmp_path):
    params = _b_files(tmp_path)
    case, _, _ = case_picslike_b()
    inp = Inputs(params, None, None, None, None)
    assert np.array_equal(inp.mask, case.mask)
    assert np.array_equal(inp.noise, case.noise)
    assert np.array_equal(inp.beam3, case.beam[:3, :33])
    for lab in inp.layout.labels:
        assert np.allclose(inp.cls[lab], case.cls[lab][:33], rtol=1e-15, atol=0)
    red = np.fromfile(params["outnoisecovmat1"]).reshape(inp.n, inp.n)
    assert np.array_equal(red, case.noise)
    from cosmoforge_amd.core import resolve_maps

    m = resolve_maps(inp.n, 1, None, inp.params, inp.layout)
    assert np.array_equal(m, case.maps)
    # load_reduced reads the reduced file back
    p2 = dict(params, covmatfile1=params["outnoisecovmat1"], load_reduced=True, outnoisecovmat1=None)
    assert np.array_equal(Inputs(p2, None, None, None, None).noise, case.noise)


def test_cross_inputs_from_files(tmp_path):
    """do_cross: the second noise covariance from covmatfile2 (reduced to the
    active pixels, written back to outnoisecovmat2) and the second map set
    from inputmapfile2, as the first ones."""
    from cosmoforge_amd.core import resolve_maps

    params = _b_files(tmp_path)
    d = load("picslike_nside8_B.npz")
    n2 = 2.0 * untril(768, d["ncov_tril"])
    n2.tofile(str(tmp_path / "ncov2.bin"))
    params.update(do_cross=True, covmatfile2=str(tmp_path / "ncov2.bin"), inputmapfile2=f("B_map.fits"),
                  outnoisecovmat2=str(tmp_path / "out" / "red2.bin"))
    inp = Inputs(params, None, None, None, None)
    case, _, _ = case_picslike_b()
    assert np.array_equal(inp.noise2, 2.0 * case.noise)
    assert np.array_equal(np.fromfile(params["outnoisecovmat2"]).reshape(inp.n, inp.n), 2.0 * case.noise)
    m2 = resolve_maps(inp.n, 1, None, inp.params, inp.layout, which=2)
    assert np.array_equal(m2, case.maps)
    # without do_cross the second covariance is not read
    params.update(do_cross=False, covmatfile2=str(tmp_path / "missing.bin"))
    assert Inputs(params, None, None, None, None).noise2 is None
