"""The Fortran host shim (src3_mpi_cfitsio_amd/fortran) drives the engine
with reference-layout module arrays and gets the same tallies as the Python
path through the same C ABI."""
import os
import subprocess

import numpy as np
import pytest

from src3_mpi_cfitsio_amd import model as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "src3_mpi_cfitsio_amd")
DRIVER = os.path.join(ROOT, "tests", "fortran", "shim_driver")
FLANG = "/opt/rocm/lib/llvm/bin/flang"


def _reference_layout(m):
    """Cell arrays -> (nrg,ntg,npg[,10]) Fortran arrays as grid_mod holds them."""
    def full(cells, fill=0.0, dtype=np.float64):
        a = np.full((m.nrg, m.ntg, m.npg), fill, dtype=dtype, order="F")
        a[:-1, :-1, :-1] = np.transpose(cells, (2, 1, 0))
        return a
    dens = np.zeros((m.nrg, m.ntg, m.npg, 10), order="F")
    fop = np.ones((m.nrg, m.ntg, m.npg, 10), dtype=np.int32, order="F")
    sp = np.transpose(m.species, (2, 1, 0))
    rho = np.transpose(m.rho, (2, 1, 0))
    for s in range(1, 11):
        dens[:-1, :-1, :-1, s - 1] = np.where(sp == s, rho, 0.0)
    return full(m.vcell), dens, fop, full(m.tdust_initial(), 0.1 / 11605.0)


def _dump(m, path, np_, seed, precision, it):
    vc, dens, fop, td = _reference_layout(m)
    i4 = lambda *v: np.array(v, dtype=np.int32).tobytes()
    f8 = lambda a: np.asarray(a, dtype=np.float64).tobytes(order="F")
    with open(path, "wb") as f:
        f.write(i4(m.nrg, m.ntg, m.npg, m.nopac, m.nlam, m.nT, m.nnu, m.npeel, m.nx, m.nband, m.nwmax,
                   m.nfreq, m.nmu, m.nph, m.nap, precision, it))
        f.write(np.array([np_, seed], dtype=np.int64).tobytes())
        f.write(f8([m.rstar, m.tstar, m.rmin, m.rmaxi, m.numin, m.lnurat]))
        for a in (m.rarr, m.r2arr, m.thetarr, m.costarr, m.tan2arr, m.phiarr, m.aarr, m.barr):
            f.write(f8(a))
        f.write(f8(vc))
        f.write(f8(dens))
        f.write(np.asarray(fop, dtype=np.int32).tobytes(order="F"))
        f.write(f8(td))
        for a in (m.nudust, m.kappad, m.adust, m.gdust, m.pldust):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        for a in (m.kappav, m.lTint, m.Tint, m.nuint, m.kapint, m.kappap):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        f.write(f8(m.thete))
        f.write(f8(m.phie))
        f.write(np.asarray(m.band_nw, dtype=np.int32).tobytes())
        f.write(np.ascontiguousarray(m.band_freq, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(m.band_phi, dtype=np.float64).tobytes())
        f.write(f8(m.aperture2))
        f.write(f8(m.u_mu))


def build_driver():
    mod_dir = os.path.join(ROOT, "_build", "fortran")
    subprocess.run([FLANG, "-O2", "-J", mod_dir, "-I", mod_dir, os.path.join(ROOT, "tests", "fortran", "shim_driver.f90"),
                    "-o", DRIVER, "-L", PKG, "-lmcrt_fortran", "-lmcrt", f"-Wl,-rpath,{PKG}"], check=True)


def test_shim_driver_builds():
    if not os.path.exists(os.path.join(PKG, "libmcrt_fortran.so")):
        pytest.skip("Fortran shim not built")
    build_driver()
    assert os.path.exists(DRIVER)


@pytest.mark.gpu
def test_fortran_disk_matches_python_engine(tmp_path):
    from src3_mpi_cfitsio_amd.engine import Engine
    if not os.path.exists(DRIVER):
        pytest.skip("shim driver not built")
    m = M.small("shim", nr=16, nt=10, nphi=4, pl=0.5)
    n, seed, it = 3000, 99, 2
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _dump(m, inp, n, seed, 0, it)
    print("fortran shim driver run", flush=True)
    r = subprocess.run([DRIVER, inp, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    cs = (m.nrg, m.ntg, m.npg)
    raw = open(out, "rb").read()
    o = 0
    def take(dtype, shape):
        nonlocal o
        cnt = int(np.prod(shape))
        a = np.frombuffer(raw, dtype=dtype, count=cnt, offset=o).reshape(shape, order="F")
        o += cnt * np.dtype(dtype).itemsize
        return a
    dt = take(np.float64, cs)
    na = take(np.int32, cs)
    td = take(np.float64, cs)
    sed = take(np.float32, (m.nfreq, m.nmu, m.nph, m.nap, 8, 10))
    with Engine(m) as e:
        e.zero_tallies()
        e.run_packets(n, seed, first_packet=it * n)
        tal = e.tallies()
        t_py = e.dusttemp(n, update_tdust=False)
    np.testing.assert_array_equal(na[:-1, :-1, :-1], np.transpose(tal["nabs"], (2, 1, 0)))
    np.testing.assert_allclose(dt[:-1, :-1, :-1], np.transpose(tal["dtauabs"], (2, 1, 0)), rtol=1e-12, atol=0)
    np.testing.assert_allclose(td[:-1, :-1, :-1], np.transpose(t_py, (2, 1, 0)), rtol=1e-12, atol=0)
    si_f = sed[..., 0, 0].sum()
    assert abs(si_f - tal["si"][0].sum()) <= 1e-5 * tal["si"][0].sum()
