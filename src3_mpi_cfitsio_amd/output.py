"""Output files of a run, in the layouts the reference writes.

Paths are relative to src3_mpi_cfitsio/.  The reference has two
interchangeable output modules with the same interface:

* output_FITS_2_test.f90 (the one its Makefile links) writes every grid and
  SED product as gzipped FITS through cfitsio:
  - output_grid (:41-145): one primary image of the cell array, then three
    binary tables with the r, theta and phi walls;
  - output (:288-560): flux_hypercube.fits.gz, peel_hypercube.fits.gz and
    the broadband images e_<filter>_<theta>_<phi>_{I,Q,U,V}_img.fits.gz.
* output_2_test.f90 writes output_grid as a Fortran unformatted sequential
  file <prefix>.unf (:33-80).  The dust-temperature checkpoint tdave_NNN.unf
  that a resumed run reads back (initarr_2_new.f90:55-75) has this layout.

The FITS writer and reader here are this package's own, written from the
FITS standard (version 4.0: 2880-byte blocks, 80-character header cards,
big-endian IEEE data); cfitsio is a third-party library the reference links
and does not ship.  The unformatted files use the record framing of the
Fortran runtimes on little-endian Linux (gfortran, flang): each record is
its byte count as a 4-byte little-endian integer, the bytes, and the count
again.

Arrays follow the engine's layout: a cell array of shape (np-1, nt-1, nr-1)
in C order holds, byte for byte, the reference's a(ir,it,ip) in Fortran
order, so a file's data section is the array's memory in file byte order.
"""
from __future__ import annotations

import gzip
import math
import os
import struct
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np

from .model import RSUN, TK, Model

__all__ = ["FitsHDU", "write_fits", "read_fits", "write_unformatted", "read_unformatted",
           "autors", "cell_centres", "output_grid", "read_grid", "write_tdave", "read_tdave", "tdave_name",
           "wavelengths", "flux_cube", "peel_cube", "angle_label", "output"]

BLOCK = 2880
CARD = 80
RAD2DEG = 180.0 / math.pi


# ---------------------------------------------------------------------------
# FITS
# ---------------------------------------------------------------------------
class FitsHDU:
    """One header-data unit: `cards` is the ordered list of (keyword, value,
    comment); `data` is an image array (C order, last axis = NAXIS1) or, for
    a binary table, a dict of column name -> 1-D array."""

    def __init__(self, cards: List[Tuple[str, object, str]], data):
        self.cards = cards
        self.data = data

    @property
    def header(self) -> Dict[str, object]:
        return {k: v for k, v, _ in self.cards if k not in ("COMMENT", "HISTORY", "")}

    def __repr__(self):
        h = self.header
        return f"FitsHDU({h.get('XTENSION', 'PRIMARY')}, {h.get('EXTNAME', '')})"


def _fmt_value(v) -> str:
    if isinstance(v, (bool, np.bool_)):
        return f"{'T' if v else 'F':>20}"
    if isinstance(v, (int, np.integer)):
        return f"{int(v):>20}"
    if isinstance(v, (float, np.floating)):
        v = float(v)
        if not math.isfinite(v):
            raise ValueError("FITS header values must be finite")
        return f"{v:>20.10E}"
    if isinstance(v, str):
        if any(ord(c) < 32 or ord(c) > 126 for c in v):
            raise ValueError(f"FITS strings are printable ASCII: {v!r}")
        s = v.replace("'", "''")
        return f"'{s:<8}'"
    raise TypeError(f"no FITS form for {type(v)}")


def _card(key: str, value=None, comment: str = "") -> bytes:
    if len(key) > 8 or key != key.upper():
        raise ValueError(f"bad FITS keyword {key!r}")
    if key == "END":
        text = "END"
    elif key in ("COMMENT", "HISTORY"):
        text = f"{key:<8}{value or ''}"
    else:
        text = f"{key:<8}= {_fmt_value(value)}"
        if comment:
            text += f" / {comment}"
    if len(text) > CARD:
        raise ValueError(f"FITS card longer than {CARD} characters: {text!r}")
    return text.ljust(CARD).encode("ascii")


def _pad(b: bytes, fill: bytes) -> bytes:
    r = len(b) % BLOCK
    return b if r == 0 else b + fill * (BLOCK - r)


_IMAGE_TYPES = {-32: ">f4", -64: ">f8", 32: ">i4", 16: ">i2", 8: "u1", 64: ">i8"}
_TFORM_TYPES = {"E": ">f4", "D": ">f8", "J": ">i4", "I": ">i2", "K": ">i8", "B": "u1"}


def _image_hdu(array: np.ndarray, bitpix: int, primary: bool, extend: bool, keywords) -> bytes:
    a = np.asarray(array)
    naxes = list(reversed(a.shape))
    cards = [_card("SIMPLE", True, "file does conform to FITS standard")] if primary else \
        [_card("XTENSION", "IMAGE", "IMAGE extension")]
    cards += [_card("BITPIX", int(bitpix), "number of bits per data pixel"),
              _card("NAXIS", len(naxes), "number of data axes")]
    cards += [_card(f"NAXIS{i + 1}", int(n), f"length of data axis {i + 1}") for i, n in enumerate(naxes)]
    if primary:
        if extend:
            cards.append(_card("EXTEND", True, "FITS dataset may contain extensions"))
    else:
        cards += [_card("PCOUNT", 0, "required keyword; must = 0"), _card("GCOUNT", 1, "required keyword; must = 1")]
    for k, v, c in keywords:
        cards.append(_card(k, v, c))
    cards.append(_card("END"))
    head = _pad(b"".join(cards), b" ")
    body = np.ascontiguousarray(a, dtype=np.dtype(_IMAGE_TYPES[bitpix])).tobytes() if a.size else b""
    return head + _pad(body, b"\0")


def _table_hdu(columns: Sequence[Tuple[str, str, str, np.ndarray]], extname: str) -> bytes:
    """columns: (name, tform 'rX', unit, values); every column has the same
    number of rows (values of shape (rows,) for r = 1, (rows, r) else)."""
    rows = None
    fields = []
    for name, tform, unit, vals in columns:
        r, code = int(tform[:-1] or 1), tform[-1]
        dt = np.dtype(_TFORM_TYPES[code])
        v = np.asarray(vals).reshape(len(vals), r) if len(np.shape(vals)) else np.asarray([vals]).reshape(1, r)
        if rows is None:
            rows = v.shape[0]
        if v.shape[0] != rows:
            raise ValueError("binary table columns of different lengths")
        fields.append((name, f"{r}{code}", unit, v.astype(dt)))
    row_bytes = sum(f[3].dtype.itemsize * f[3].shape[1] for f in fields)
    cards = [_card("XTENSION", "BINTABLE", "binary table extension"),
             _card("BITPIX", 8, "8-bit bytes"), _card("NAXIS", 2, "2-dimensional binary table"),
             _card("NAXIS1", row_bytes, "width of table in bytes"), _card("NAXIS2", rows or 0, "number of rows in table"),
             _card("PCOUNT", 0, "size of special data area"), _card("GCOUNT", 1, "one data group (required keyword)"),
             _card("TFIELDS", len(fields), "number of fields in each row")]
    for i, (name, tform, unit, _) in enumerate(fields, 1):
        cards.append(_card(f"TTYPE{i}", name, f"label for field {i}"))
        cards.append(_card(f"TFORM{i}", tform, "data format of field"))
        if unit:
            cards.append(_card(f"TUNIT{i}", unit, "physical unit of field"))
    cards.append(_card("EXTNAME", extname, "name of this binary table extension"))
    cards.append(_card("END"))
    head = _pad(b"".join(cards), b" ")
    rec = np.empty(rows or 0, dtype=[(f"f{i}", f[3].dtype, (f[3].shape[1],)) for i, f in enumerate(fields)])
    for i, f in enumerate(fields):
        rec[f"f{i}"] = f[3]
    return head + _pad(rec.tobytes(), b"\0")


def write_fits(path: str, primary: np.ndarray, bitpix: int = -32, keywords=(), tables=(), extend: bool = True):
    """A primary image plus binary-table extensions, gzipped when `path`
    ends in .gz (as cfitsio does for such names).  `keywords`: (key, value,
    comment) cards after the mandatory ones; `tables`: (extname, columns)
    with columns as in _table_hdu."""
    parts = [_image_hdu(primary, bitpix, True, extend or bool(tables), keywords)]
    for extname, cols in tables:
        parts.append(_table_hdu(cols, extname))
    blob = b"".join(parts)
    tmp = path + ".tmp"
    if path.endswith(".gz"):
        with open(tmp, "wb") as raw, gzip.GzipFile(fileobj=raw, mode="wb", mtime=0, filename="") as f:
            f.write(blob)
    else:
        with open(tmp, "wb") as f:
            f.write(blob)
    os.replace(tmp, path)


def _parse_value(s: str):
    s = s.strip()
    if not s:
        return None
    if s.startswith("'"):
        out, i = [], 1
        while i < len(s):
            if s[i] == "'":
                if i + 1 < len(s) and s[i + 1] == "'":
                    out.append("'")
                    i += 2
                    continue
                break
            out.append(s[i])
            i += 1
        return "".join(out).rstrip(), s[i + 1:]
    v, _, rest = s.partition("/")
    v = v.strip()
    if v in ("T", "F"):
        return v == "T", rest
    try:
        return int(v), rest
    except ValueError:
        return float(v.replace("D", "E")), rest


def _read_header(buf: bytes, pos: int):
    cards = []
    while True:
        if pos + BLOCK > len(buf):
            raise ValueError("FITS header runs past the end of the file")
        block = buf[pos:pos + BLOCK]
        pos += BLOCK
        for i in range(0, BLOCK, CARD):
            c = block[i:i + CARD].decode("ascii")
            key = c[:8].strip()
            if key == "END":
                return cards, pos
            if key in ("COMMENT", "HISTORY", ""):
                cards.append((key, c[8:].rstrip(), ""))
            elif c[8:10] == "= ":
                pv = _parse_value(c[10:])
                if isinstance(pv, tuple):
                    val, rest = pv
                    comment = rest.partition("/")[2].strip() if isinstance(val, str) else rest.strip()
                else:
                    val, comment = None, ""
                cards.append((key, val, comment))
            else:
                cards.append((key, None, c[8:].rstrip()))


def read_fits(path: str) -> List[FitsHDU]:
    """Every HDU of a FITS file (gzipped or not) holding images of any BITPIX
    (raw values: BSCALE/BZERO are not applied) and binary tables of
    E/D/J/I/K/B/A columns."""
    with open(path, "rb") as f:
        buf = f.read()
    if buf[:2] == b"\x1f\x8b":
        buf = gzip.decompress(buf)
    hdus, pos = [], 0
    while pos < len(buf):
        cards, pos = _read_header(buf, pos)
        h = {k: v for k, v, _ in cards}
        naxes = [int(h[f"NAXIS{i + 1}"]) for i in range(int(h.get("NAXIS", 0)))]
        bitpix = int(h["BITPIX"])
        if h.get("XTENSION") == "BINTABLE":
            nbytes = naxes[0] * naxes[1] + int(h.get("PCOUNT", 0))
            fields = []
            for i in range(int(h["TFIELDS"])):
                tf = h[f"TFORM{i + 1}"].strip()
                r = int(tf[:-1] or 1)
                if tf[-1] == "A":                    # character column: one string of r bytes
                    fields.append((h[f"TTYPE{i + 1}"], np.dtype(f"S{r}"), 1))
                else:
                    fields.append((h[f"TTYPE{i + 1}"], np.dtype(_TFORM_TYPES[tf[-1]]), r))
            dt = np.dtype([(f"f{i}", d, (r,)) for i, (_, d, r) in enumerate(fields)])
            if dt.itemsize != naxes[0]:
                raise ValueError("binary table row width does not match its columns")
            rec = np.frombuffer(buf, dtype=dt, count=naxes[1], offset=pos)
            data = {}
            for i, (name, d, r) in enumerate(fields):
                col = rec[f"f{i}"].astype(d.newbyteorder("="))
                data[name] = col[:, 0] if r == 1 else col
        else:
            n = int(np.prod(naxes)) if naxes else 0
            nbytes = n * abs(bitpix) // 8
            dt = np.dtype(_IMAGE_TYPES[bitpix])
            data = np.frombuffer(buf, dtype=dt, count=n, offset=pos).astype(dt.newbyteorder("=")).reshape(
                tuple(reversed(naxes))) if n else None
        pos += nbytes + ((-nbytes) % BLOCK)
        hdus.append(FitsHDU(cards, data))
    return hdus


# ---------------------------------------------------------------------------
# Fortran unformatted sequential files
# ---------------------------------------------------------------------------
def write_unformatted(path: str, records: Iterable[Sequence]):
    """Each record is a sequence of numpy arrays / scalars of explicit dtype
    written back to back (a Fortran WRITE statement's item list)."""
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        for items in records:
            body = b"".join(np.ascontiguousarray(x).tobytes() for x in items)
            if len(body) >= 2 ** 31:
                raise ValueError("records of 2 GiB or more need subrecord framing, which is not written")
            n = struct.pack("<i", len(body))
            f.write(n + body + n)
    os.replace(tmp, path)


def read_unformatted(path: str) -> List[bytes]:
    recs = []
    with open(path, "rb") as f:
        buf = f.read()
    pos = 0
    while pos < len(buf):
        if pos + 4 > len(buf):
            raise ValueError("truncated record marker")
        (n,) = struct.unpack_from("<i", buf, pos)
        if n < 0 or pos + 8 + n > len(buf):
            raise ValueError(f"bad record length {n} at byte {pos}")
        body = buf[pos + 4:pos + 4 + n]
        (n2,) = struct.unpack_from("<i", buf, pos + 4 + n)
        if n2 != n:
            raise ValueError(f"record markers differ ({n} vs {n2}) at byte {pos}")
        recs.append(body)
        pos += 8 + n
    return recs


# ---------------------------------------------------------------------------
# grids
# ---------------------------------------------------------------------------
def autors(model: Model) -> float:
    """Stellar radii per AU as setup.f90:45 sets it (214.94 / rstar)."""
    return 214.94 / model.rstar


def cell_centres(model: Model):
    """ravearr, thetavearr, phiavearr: wall midpoints (gridset_5_model13new.f90:525, :745, :780)."""
    r, t, p = (np.asarray(a, np.float64) for a in (model.rarr, model.thetarr, model.phiarr))
    return 0.5 * (r[:-1] + r[1:]), 0.5 * (t[:-1] + t[1:]), 0.5 * (p[:-1] + p[1:])


def _cells(model: Model, array) -> np.ndarray:
    a = np.asarray(array)
    if a.shape != model.cell_shape:
        raise ValueError(f"cell array of shape {a.shape}, the grid has {model.cell_shape}")
    return a


def output_grid(prefix: str, model: Model, array, fmt: str = "fits") -> str:
    """output_grid(prefix, array) of either output module for a cell array
    (booleans are written as 0/1 integers, floats as real(8) in .unf and
    32-bit floats in FITS).  Returns the file written."""
    a = _cells(model, array)
    if a.dtype == np.bool_:
        a = a.astype(np.int32)
    is_int = np.issubdtype(a.dtype, np.integer)
    if fmt == "unf":
        # output_2_test.f90:33-41 (header records) and :64-80 (the array)
        rave, tave, pave = cell_centres(model)
        dims = np.array([model.nrg - 1, model.ntg - 1, model.npg - 1, 1], np.int32)
        path = prefix + ".unf"
        write_unformatted(path, [[dims], [rave / autors(model)], [tave], [pave],
                                 [a.astype(np.int32 if is_int else np.float64)]])
        return path
    if fmt != "fits":
        raise ValueError(f"unknown grid format {fmt!r}")
    # output_FITS_2_test.f90:41-50 (primary), :84-92 (data), :136-145 (walls)
    path = prefix + ".fits.gz"
    walls = [("r walls", [("r walls", "1E", "cm", np.asarray(model.rarr) * model.rstar * RSUN)]),
             ("theta walls", [("theta walls", "1E", "rad", np.asarray(model.thetarr))]),
             ("phi walls", [("phi walls", "1E", "rad", np.asarray(model.phiarr))])]
    write_fits(path, a, bitpix=32 if is_int else -32, tables=walls)
    return path


def read_grid(path: str):
    """Inverse of output_grid for either format: (array, walls_or_centres).
    For .unf the second item is (ravearr/autors, thetavearr, phiavearr);
    for FITS it is the wall lists (r in cm, theta and phi in rad)."""
    if path.endswith(".unf"):
        recs = read_unformatted(path)
        if len(recs) != 5:
            raise ValueError(f"{path}: {len(recs)} records, a grid file has 5")
        nr, nt, npc, last = np.frombuffer(recs[0], "<i4")
        rave, tave, pave = (np.frombuffer(recs[i], "<f8") for i in (1, 2, 3))
        n = nr * nt * npc * last
        dt = "<f8" if len(recs[4]) == 8 * n else "<i4"
        if len(recs[4]) != np.dtype(dt).itemsize * n:
            raise ValueError(f"{path}: data record of {len(recs[4])} bytes for {n} cells")
        shape = (npc, nt, nr) if last == 1 else (last, npc, nt, nr)
        return np.frombuffer(recs[4], dt).reshape(shape).copy(), (rave.copy(), tave.copy(), pave.copy())
    hdus = read_fits(path)
    return hdus[0].data, tuple(next(iter(h.data.values())) for h in hdus[1:4])


def tdave_name(iteration: int) -> str:
    """'tdave_NNN' as tfinal names the checkpoint of Lucy iteration `iteration`
    (suffix written with '("_",I3.3)', initarr_2_new.f90:58)."""
    if not 0 <= iteration <= 999:
        raise ValueError("the reference's checkpoint suffix holds iterations 0..999")
    return f"tdave_{iteration:03d}"


def write_tdave(directory: str, model: Model, tdust, iteration: int, fmt: str = "unf") -> str:
    """The dust temperatures after Lucy iteration `iteration` (engine units,
    K/11605) written as tfinal writes them: in K (tfinal_5_test1.f90:418-420)."""
    return output_grid(os.path.join(directory, tdave_name(iteration)), model,
                       np.asarray(tdust, np.float64) * TK, fmt=fmt)


def read_tdave(path: str, model: Model) -> np.ndarray:
    """Resume: read a tdave checkpoint as initarr_2_new.f90:55-75 does (the
    four integers, three skipped records, the temperatures) and return the
    dust temperatures in engine units (K/11605) for Engine.set_tdust."""
    t, _ = read_grid(path)
    if t.shape != model.cell_shape:
        raise ValueError(f"{path}: checkpoint of shape {t.shape}, the grid has {model.cell_shape}")
    return np.asarray(t, np.float64) / TK


# ---------------------------------------------------------------------------
# SED hypercubes and images (output(), output_FITS_2_test.f90:288-560)
# ---------------------------------------------------------------------------
PLANES = ("I", "I_ERR", "Q/I", "Q_ERR", "U/I", "U_ERR", "V/I", "V_ERR")


def wavelengths(model: Model) -> np.ndarray:
    """wavarr(inub) = 1.2398 / (numin nurat^((inub-0.5)/nfreq)) microns (:325-328)."""
    inub = np.arange(1, model.nfreq + 1, dtype=np.float64)
    return 1.2398 / (model.numin * np.exp(model.lnurat * (inub - 0.5) / model.nfreq))


def _nunorm(model: Model) -> np.float32:
    """nunorm = nurat^(0.5/nfreq) - nurat^(-0.5/nfreq) (setup.f90:42), taken to real."""
    h = 0.5 * model.lnurat / model.nfreq
    return np.float32(math.exp(h) - math.exp(-h))


def _cube(i, i2, q, q2, u, u2, v, v2, nunorm) -> np.ndarray:
    """The eight planes of the reference's where(si > 0.) block (:377-388):
    I/nunorm, I_err*I/nunorm, Q/I, Q_err, U/I, U_err, V/I, V_err; 0 where I <= 0."""
    i = np.asarray(i, np.float32)
    out = np.zeros((8,) + i.shape, np.float32)
    pos = i > 0
    with np.errstate(divide="ignore", invalid="ignore"):
        planes = (i / nunorm, np.asarray(i2, np.float32) * i / nunorm, np.asarray(q, np.float32) / i, q2,
                  np.asarray(u, np.float32) / i, u2, np.asarray(v, np.float32) / i, v2)
        for k, p in enumerate(planes):
            out[k][pos] = np.asarray(p, np.float32)[pos]
    return out


def flux_cube(flux: Dict[str, np.ndarray], model: Model) -> np.ndarray:
    """fluxcube(nfreq,nmu,nph,nap,no,8) of normalized SEDs (diskflux output),
    as a C-order array of shape (8, no, nap, nph, nmu, nfreq)."""
    f = flux
    return _cube(f["si"], f["si2"], f["sq"], f["sq2"], f["su"], f["su2"], f["sv"], f["sv2"], _nunorm(model))


def peel_cube(flux: Dict[str, np.ndarray], model: Model) -> np.ndarray:
    """peelcube(nfreq,npeel,nap,no,8), C-order shape (8, no, nap, npeel, nfreq)."""
    f = flux
    return _cube(f["ti"], f["ti2"], f["tq"], f["tq2"], f["tu"], f["tu2"], f["tv"], f["tv2"], _nunorm(model))


def angle_label(rad: float) -> str:
    """write(c,'(F5.1)') angle*rad2deg, then trim(adjustl(c)) (:474-475)."""
    s = f"{rad * RAD2DEG:5.1f}"
    return "*****" if len(s) > 5 else s.strip()


def _apertures_au(model: Model) -> np.ndarray:
    return np.sqrt(np.asarray(model.aperture2, np.float64)) / autors(model)


def output(directory: str, flux: Dict[str, np.ndarray], model: Model, band_names: Optional[Sequence[str]] = None,
           partial_peeloff: bool = False, nabs=None, tdust=None, iteration: Optional[int] = None) -> List[str]:
    """Write the products of the reference's output() that the packet path
    feeds, from `flux` = engine.diskflux(tallies, model, npackets):

    * flux_hypercube.fits.gz: primary (nfreq,nmu,nph,nap,no,8) cube, then
      WAVELENGTH (microns) and APERTURE (AU) tables (:374-399);
    * peel_hypercube[_partial].fits.gz when peel-off is on: (nfreq,npeel,
      nap,no,8) cube, WAVELENGTH, ANGLE (THETA, PHI in deg), APERTURE (:403-431);
    * e_<band>_<theta>_<phi>_{I,Q,U,V}_img[_partial].fits.gz per peel-off
      direction and band, the I image carrying PHIE/THETE and a WAVELENGTH
      table (:477-519);
    * nabs.fits.gz when `nabs` is given (:616), and the tdave checkpoint of
      `iteration` (FITS and .unf) when `tdust` is given.

    `band_names` label the broadband filters (the reference's filtnames,
    filt_mod.f90:35); default B1, B2, ...  Returns the files written."""
    os.makedirs(directory, exist_ok=True)
    suffix = "_partial" if partial_peeloff else ""
    wav = wavelengths(model)
    ap = _apertures_au(model)
    written = []

    p = os.path.join(directory, "flux_hypercube.fits.gz")
    write_fits(p, flux_cube(flux, model), tables=[("WAVELENGTH", [("WAVELENGTH", "1E", "microns", wav)]),
                                                  ("APERTURE", [("APERTURE", "1E", "AU", ap)])])
    written.append(p)

    if model.ipeel == 1 and model.npeel > 0:
        cube = peel_cube(flux, model)[:, :, :, :model.npeel, :]
        p = os.path.join(directory, f"peel_hypercube{suffix}.fits.gz")
        write_fits(p, cube, tables=[
            ("WAVELENGTH", [("WAVELENGTH", "1E", "microns", wav)]),
            ("ANGLE", [("THETA", "1E", "deg", np.asarray(model.thete[:model.npeel]) * RAD2DEG),
                       ("PHI", "1E", "deg", np.asarray(model.phie[:model.npeel]) * RAD2DEG)]),
            ("APERTURE", [("APERTURE", "1E", "AU", ap)])])
        written.append(p)

        nb = model.nband
        names = list(band_names) if band_names is not None else [f"B{k + 1}" for k in range(nb)]
        if len(names) != nb:
            raise ValueError(f"{len(names)} band names for {nb} bands")
        rmaxi_cm = float(model.rmaxi) * model.rstar * RSUN
        for ip in range(model.npeel):
            th, ph = angle_label(model.thete[ip]), angle_label(model.phie[ip])
            for ib in range(nb):
                beg = os.path.join(directory, f"e_{names[ib].strip()}_{th}_{ph}")
                for s in "IQUV":
                    img = flux[f"image_b_{s.lower()}"][ib, :, :, ip]
                    kw = [("RMAXI", rmaxi_cm, ""), ("NXHST", int(model.nx), "")]
                    tables = ()
                    if s == "I":
                        kw += [("PHIE", float(model.phie[ip]) * RAD2DEG, ""),
                               ("THETE", float(model.thete[ip]) * RAD2DEG, "")]
                        tables = [("WAVELENGTH", [("WAVELENGTH", "1E", "microns", wav)])]
                    p = f"{beg}_{s}_img{suffix}.fits.gz"
                    write_fits(p, img, keywords=kw, tables=tables, extend=False)
                    written.append(p)

    if nabs is not None:
        written.append(output_grid(os.path.join(directory, "nabs"), model, np.asarray(nabs, np.int32)))
    if tdust is not None:
        it = 0 if iteration is None else iteration
        written.append(write_tdave(directory, model, tdust, it, fmt="fits"))
        written.append(write_tdave(directory, model, tdust, it, fmt="unf"))
    return written
