"""Output files: the tdave checkpoint (Fortran unformatted) and the FITS
grids, SED hypercubes and images of the reference's output modules.

What pins them:
* the unformatted layout against a Fortran runtime: tests/fortran/unf_grid.f90
  writes a grid with the reference's WRITE statements (the file must be
  byte-identical to ours) and reads ours with the reference's READ
  statements (every value must come back exactly);
* the FITS files against astropy.io.fits (an independent FITS
  implementation, in the image's Python 3.9), which must accept them under
  its standard-conformance check and see the same headers and values; our
  reader is checked against astropy on FITS files written by other tools;
* the SED-binning constants against the bin edges they describe.
"""
import json
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from src3_mpi_cfitsio_amd import model as M
from src3_mpi_cfitsio_amd import output as O

HERE = os.path.dirname(os.path.abspath(__file__))
FLANG = "/opt/rocm/lib/llvm/bin/flang"
PY39 = "/opt/conda/bin/python3.9"
OTHER_FITS = ["/usr/local/lib/python3.10/dist-packages/numpy/_core/tests/data/recarray_from_file.fits",
              "/opt/conda/lib/python3.9/site-packages/astropy/wcs/tests/data/dist_lookup.fits.gz"]


@pytest.fixture(scope="module")
def small():
    # distinct sizes on every axis, so that a transposed axis shows
    return M.build("out", nr=7, nt=6, nphi=3, npeel=2, nx=5, nfreq=4, nmu=3, nlam=10, nT=10, nnu=9,
                   bands=((1.0, 2.0), (3.0, 4.0)), thete_deg=[30.0, 75.5], phie_deg=45.0)


@pytest.fixture(scope="module")
def unf_grid(tmp_path_factory):
    if not os.path.exists(FLANG):
        pytest.skip("flang not present")
    exe = str(tmp_path_factory.mktemp("fortran") / "unf_grid")
    subprocess.run([FLANG, "-O1", os.path.join(HERE, "fortran", "unf_grid.f90"), "-o", exe], check=True,
                   cwd=os.path.dirname(exe))
    return exe


def _astropy(tmp_path, paths):
    if not os.path.exists(PY39):
        pytest.skip("no Python 3.9 with astropy in this image")
    out = str(tmp_path / "astropy_dump.npz")
    env = {k: v for k, v in os.environ.items() if not k.startswith("PYTHON")}
    r = subprocess.run([PY39, "-s", os.path.join(HERE, "fits_astropy_dump.py"), out] + list(paths),
                       capture_output=True, text=True, env=env, timeout=300)
    if r.returncode != 0 and "No module named 'astropy'" in r.stderr:
        pytest.skip("astropy not importable")
    assert r.returncode == 0, r.stderr[-3000:]
    z = np.load(out, allow_pickle=False)
    res = []
    for i in range(len(paths)):
        hdus = []
        for j in range(int(z[f"f{i}_nhdu"])):
            head = json.loads(bytes(z[f"f{i}_h{j}_header"]).decode())
            cols = {k.split("_col_", 1)[1]: z[k] for k in z.files if k.startswith(f"f{i}_h{j}_col_")}
            data = z[f"f{i}_h{j}_data"] if f"f{i}_h{j}_data" in z.files else (cols or None)
            hdus.append((head, data))
        res.append(hdus)
    return res


# ---------------------------------------------------------------------------
# Fortran unformatted
# ---------------------------------------------------------------------------
def test_unformatted_bytes_equal_fortran_writer(tmp_path, unf_grid):
    nr, nt, npc = 5, 4, 3
    f_path = str(tmp_path / "fortran.unf")
    subprocess.run([unf_grid, "write", f_path, str(nr), str(nt), str(npc)], check=True)
    ir, it, ip = (np.arange(1, n + 1, dtype=np.float64) for n in (nr, nt, npc))
    a = 1000.0 * ir[None, None, :] + 10.0 * it[None, :, None] + 0.25 * ip[:, None, None] + 0.125
    p_path = str(tmp_path / "python.unf")
    O.write_unformatted(p_path, [[np.array([nr, nt, npc, 1], np.int32)], [1.5 * ir + 0.0625], [0.01 * it],
                                 [0.5 * ip], [a]])
    assert open(f_path, "rb").read() == open(p_path, "rb").read()
    got, (c1, c2, c3) = O.read_grid(f_path)
    assert np.array_equal(got, a) and np.array_equal(c1, 1.5 * ir + 0.0625) and np.array_equal(c3, 0.5 * ip)


def test_tdave_checkpoint_reads_back_in_fortran(tmp_path, unf_grid, small):
    rng = np.random.default_rng(5)
    t_ev = rng.uniform(1.0, 1500.0, small.cell_shape) / M.TK
    path = O.write_tdave(str(tmp_path), small, t_ev, iteration=3)
    assert os.path.basename(path) == "tdave_003.unf"          # initarr_2_new.f90:58-60
    nr, nt, npc = small.nrg - 1, small.ntg - 1, small.npg - 1
    out = subprocess.run([unf_grid, "read", path, str(nr), str(nt), str(npc)], check=True, capture_output=True,
                         text=True).stdout.split("\n")
    assert [int(x) for x in out[0].split()] == [nr, nt, npc, 1]
    vals = np.array([float(x) for x in out[1:] if x.strip()])
    rave, tave, pave = O.cell_centres(small)
    n0 = nr + nt + npc
    assert np.array_equal(vals[:nr], rave / O.autors(small))
    assert np.array_equal(vals[nr:nr + nt], tave) and np.array_equal(vals[nr + nt:n0], pave)
    # the checkpoint holds kelvin (tfinal_5_test1.f90:418-420), cell order ir fastest
    assert np.array_equal(vals[n0:], (t_ev * M.TK).ravel())
    back = O.read_tdave(path, small)
    assert np.array_equal(back, (t_ev * M.TK) / M.TK)
    np.testing.assert_allclose(back, t_ev, rtol=2e-16, atol=0)


def test_grid_centres_are_wall_midpoints(small):
    rave, tave, pave = O.cell_centres(small)
    for c, w in ((rave, small.rarr), (tave, small.thetarr), (pave, small.phiarr)):
        assert len(c) == len(w) - 1 and np.all(c > w[:-1]) and np.all(c < w[1:])
        np.testing.assert_allclose(c - w[:-1], w[1:] - c, rtol=1e-12, atol=1e-15)
    # setup.f90:45: autors = 214.94/rstar stellar radii per AU (1 AU = 214.94 Rsun to 5 digits)
    assert abs(O.autors(small) * small.rstar - M.AU / M.RSUN) < 0.01


def test_unformatted_rejects_damaged_files(tmp_path):
    p = str(tmp_path / "bad.unf")
    O.write_unformatted(p, [[np.arange(3, dtype=np.int32)]])
    raw = bytearray(open(p, "rb").read())
    raw[-1] ^= 1
    open(p, "wb").write(bytes(raw))
    with pytest.raises(ValueError):
        O.read_unformatted(p)
    open(p, "wb").write(bytes(raw[:-2]))
    with pytest.raises(ValueError):
        O.read_unformatted(p)


# ---------------------------------------------------------------------------
# FITS
# ---------------------------------------------------------------------------
def test_reader_matches_astropy_on_foreign_files(tmp_path):
    paths = [p for p in OTHER_FITS if os.path.exists(p)]
    if not paths:
        pytest.skip("no FITS files from other writers in this image")
    ref = _astropy(tmp_path, paths)
    for path, hdus in zip(paths, ref):
        ours = O.read_fits(path)
        assert len(ours) == len(hdus), path
        for h, (head, data) in zip(ours, hdus):
            for k in ("BITPIX", "NAXIS", "NAXIS1", "NAXIS2", "EXTNAME", "TFIELDS", "CRVAL1", "CDELT1"):
                if k in head:
                    assert h.header[k] == head[k], (path, k)
            if isinstance(data, dict):
                for name, col in data.items():
                    mine = h.data[name]
                    if mine.dtype.kind == "S":
                        assert [x.decode().rstrip() for x in mine] == [str(x).rstrip() for x in col], (path, name)
                    else:
                        assert np.array_equal(mine, col), (path, name)
            elif data is not None:
                assert np.array_equal(h.data, data), path


def test_grid_fits_as_astropy_reads_it(tmp_path, small):
    rng = np.random.default_rng(2)
    t = rng.uniform(10, 900, small.cell_shape)
    nabs = rng.integers(0, 1000, small.cell_shape).astype(np.int32)
    p1 = O.output_grid(str(tmp_path / "tdave_001"), small, t)
    p2 = O.output_grid(str(tmp_path / "nabs"), small, nabs)
    (g1, g2) = _astropy(tmp_path, [p1, p2])
    for path, hdus, vals, bitpix in ((p1, g1, t.astype(np.float32), -32), (p2, g2, nabs, 32)):
        head, data = hdus[0]
        # output_FITS_2_test.f90:41-50: naxes (nrg-1, ntg-1, npg-1)
        assert (head["BITPIX"], head["NAXIS"]) == (bitpix, 3)
        assert (head["NAXIS1"], head["NAXIS2"], head["NAXIS3"]) == (small.nrg - 1, small.ntg - 1, small.npg - 1)
        assert np.array_equal(data, vals)
        names = [h[0]["EXTNAME"] for h in hdus[1:]]
        assert names == ["r walls", "theta walls", "phi walls"]
        np.testing.assert_array_equal(hdus[1][1]["r walls"], (small.rarr * small.rstar * M.RSUN).astype(np.float32))
        np.testing.assert_array_equal(hdus[2][1]["theta walls"], small.thetarr.astype(np.float32))
        assert hdus[1][0]["TUNIT1"] == "cm" and hdus[2][0]["TUNIT1"] == "rad"
        mine, walls = O.read_grid(path)
        assert np.array_equal(mine, vals) and len(walls) == 3


def _synthetic_flux(m, rng):
    NO = 8
    sed = {}
    shp = (NO, m.nap, m.nph, m.nmu, m.nfreq)
    for k in ("si", "sq", "su", "sv", "si2", "sq2", "su2", "sv2"):
        sed[k] = rng.normal(size=shp).astype(np.float32)
    sed["si"] = np.abs(sed["si"]) * (rng.uniform(size=shp) > 0.3)     # zeros where nothing arrived
    pshp = (NO, m.nap, m.npeel, m.nfreq)
    for k in ("ti", "tq", "tu", "tv", "ti2", "tq2", "tu2", "tv2"):
        sed[k] = rng.normal(size=pshp).astype(np.float32)
    sed["ti"] = np.abs(sed["ti"]) * (rng.uniform(size=pshp) > 0.3)
    for s in "iquvdseo":
        sed[f"image_b_{s}"] = rng.normal(size=(m.nband, m.nx, m.nx, m.npeel)).astype(np.float32)
    return sed


def test_sed_binning_constants(small):
    # bin inub spans [numin nurat^((i-1)/nfreq), numin nurat^(i/nfreq)]
    nurat = math.exp(small.lnurat)
    i = np.arange(1, small.nfreq + 1)
    lo, hi = small.numin * nurat ** ((i - 1) / small.nfreq), small.numin * nurat ** (i / small.nfreq)
    centre = np.sqrt(lo * hi)
    # nunorm: the bin width over its (geometric) centre, the same for every bin
    np.testing.assert_allclose((hi - lo) / centre, float(O._nunorm(small)), rtol=1e-6)
    # wavarr: the bin centre in microns, hc = 1.23984 eV um
    hc = 6.62607015e-34 * 2.99792458e8 / 1.602176634e-19 * 1e6
    np.testing.assert_allclose(O.wavelengths(small), hc / centre, rtol=1e-4)
    assert np.all(np.diff(O.wavelengths(small)) < 0)


def test_flux_and_peel_hypercubes_as_astropy_reads_them(tmp_path, small):
    rng = np.random.default_rng(9)
    flux = _synthetic_flux(small, rng)
    files = O.output(str(tmp_path), flux, small, band_names=["2J", "I1"])
    base = {os.path.basename(p) for p in files}
    assert {"flux_hypercube.fits.gz", "peel_hypercube.fits.gz"} <= base
    # e_<filter>_<theta>_<phi>_<S>_img.fits.gz with '(F5.1)' angles
    for name in ("e_2J_30.0_45.0_I_img.fits.gz", "e_I1_75.5_45.0_V_img.fits.gz"):
        assert name in base
    fc, pc, im, imq = _astropy(tmp_path, [os.path.join(str(tmp_path), n) for n in
                                          ("flux_hypercube.fits.gz", "peel_hypercube.fits.gz",
                                           "e_I1_75.5_45.0_I_img.fits.gz", "e_I1_75.5_45.0_Q_img.fits.gz")])
    nun = O._nunorm(small)
    head, cube = fc[0]
    # output_FITS_2_test.f90:394: naxesf = (nfreq, nmu, nph, nap, no, 8)
    assert [head[f"NAXIS{k}"] for k in range(1, 7)] == [small.nfreq, small.nmu, small.nph, small.nap, 8, 8]
    si, pos = flux["si"], flux["si"] > 0
    with np.errstate(divide="ignore", invalid="ignore"):
        expect = [si / nun, flux["si2"] * si / nun, flux["sq"] / si, flux["sq2"], flux["su"] / si, flux["su2"],
                  flux["sv"] / si, flux["sv2"]]
    for k in range(8):
        assert np.array_equal(cube[k][pos], expect[k][pos].astype(np.float32)), k
        assert np.all(cube[k][~pos] == 0.0), k
    assert [h[0]["EXTNAME"] for h in fc[1:]] == ["WAVELENGTH", "APERTURE"]
    np.testing.assert_array_equal(fc[1][1]["WAVELENGTH"], O.wavelengths(small).astype(np.float32))
    np.testing.assert_array_equal(fc[2][1]["APERTURE"],
                                  (np.sqrt(small.aperture2) / O.autors(small)).astype(np.float32))
    head, cube = pc[0]
    assert [head[f"NAXIS{k}"] for k in range(1, 6)] == [small.nfreq, small.npeel, small.nap, 8, 8]
    ti, pos = flux["ti"], flux["ti"] > 0
    assert np.array_equal(cube[0][pos], (ti / nun)[pos]) and np.all(cube[2][~pos] == 0)
    assert [h[0]["EXTNAME"] for h in pc[1:]] == ["WAVELENGTH", "ANGLE", "APERTURE"]
    np.testing.assert_allclose(pc[2][1]["THETA"], [30.0, 75.5], rtol=1e-6)
    np.testing.assert_allclose(pc[2][1]["PHI"], [45.0, 45.0], rtol=1e-6)
    # images: image_b_i(peelid, :, :, ift), naxes (nxhst, nxhst)
    head, img = im[0]
    assert (head["NAXIS1"], head["NAXIS2"]) == (small.nx, small.nx)
    assert np.array_equal(img, flux["image_b_i"][1, :, :, 1])
    assert head["NXHST"] == small.nx and abs(head["THETE"] - 75.5) < 1e-9 and abs(head["PHIE"] - 45.0) < 1e-9
    np.testing.assert_allclose(head["RMAXI"], small.rmaxi * small.rstar * M.RSUN, rtol=1e-10)
    assert "THETE" not in imq[0][0] and np.array_equal(imq[0][1], flux["image_b_q"][1, :, :, 1])
    # our own reader agrees with astropy on our files
    ours = O.read_fits(os.path.join(str(tmp_path), "flux_hypercube.fits.gz"))
    assert np.array_equal(ours[0].data, fc[0][1]) and ours[0].header["NAXIS6"] == 8


def test_angle_labels_follow_f5_1():
    assert O.angle_label(math.radians(60.05)) in ("60.0", "60.1")
    assert O.angle_label(math.radians(5.0)) == "5.0"
    assert O.angle_label(math.radians(120.0)) == "120.0"
    assert O.angle_label(math.radians(1000.0)) == "*****"


def test_fits_cards_are_standard():
    c = O._card("NAXIS1", 12, "x")
    assert len(c) == 80 and c[:10] == b"NAXIS1  = " and c[29:30] == b"2"      # integers end in column 30
    c = O._card("XTENSION", "BINTABLE")
    assert c[10:20] == b"'BINTABLE'"
    c = O._card("EXTNAME", "ab")
    assert c[10:20] == b"'ab      '"                                          # strings padded to 8
    c = O._card("SIMPLE", True)
    assert c[29:30] == b"T"
    with pytest.raises(ValueError):
        O._card("TOOLONGKEY", 1)
