"""CPU checks of the boundary and the host logic: the C-ABI library loads and
exports every symbol include/mcrt.h declares, fails loudly without a GPU, and
the synthetic model builder honours the reference's grid conventions."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from src3_mpi_cfitsio_amd import model as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mcrt_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from src3_mpi_cfitsio_amd import _lib
    lib = C.CDLL(_lib.LIB_PATH)
    names = _declared("mcrt.h")
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) <= set(_lib.SIGNATURES)


def test_create_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from src3_mpi_cfitsio_amd import _lib
    from src3_mpi_cfitsio_amd.engine import Engine, MCRTError
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.mcrt_create(0, C.byref(h)) == 4          # MCRT_E_NODEVICE
    with pytest.raises(MCRTError):
        Engine(M.small())


def test_grid_conventions():
    m = M.config("c0")
    assert m.ntg % 2 == 1 and m.npg == 2
    ne = (m.ntg + 1) // 2
    assert m.tan2arr[ne - 1] == -1.0 and m.costarr[ne - 1] == 0.0
    assert m.tan2arr[0] == 0.0 and m.tan2arr[-1] == 0.0
    assert m.costarr[0] == 1.0 and m.costarr[-1] == -1.0
    assert m.rarr[0] == 1.0 and np.all(np.diff(m.rarr) > 0)
    np.testing.assert_array_equal(m.r2arr, m.rarr ** 2)
    vol = (4.0 * math.pi / 3.0) * (m.rarr[-1] ** 3 - 1.0) * (m.rstar * M.RSUN) ** 3
    assert abs(m.vcell.sum() / vol - 1.0) < 1e-10
    disk = m.species == 1
    ratio = np.sum(m.rho * m.vcell * disk) / (0.01 * M.MSUN)   # disk cells also hold envelope gas
    assert 1.0 <= ratio < 1.3
    assert np.all(m.rho[:, :, 0] == 0.0)                   # star..Rin gap
    assert np.min(np.abs(m.thete[:, None] - m.thetarr[None, :])) > 0


def test_3d_grid_phi_tables():
    m = M.build("p", nr=10, nt=8, nphi=8, npeel=2, nx=9, nfreq=4, nmu=2, nlam=10, nT=10, nnu=9)
    assert m.npg == 9
    np.testing.assert_allclose(m.aarr, np.sin(m.phiarr))
    np.testing.assert_allclose(m.barr, -np.cos(m.phiarr))
    assert m.phiarr[0] == 0.0 and abs(m.phiarr[-1] - 2 * math.pi) < 1e-15


def test_emission_tables_are_cdfs():
    m = M.small()
    K = m.kapint[0]
    assert np.all(np.diff(K, axis=0) >= 0)
    assert np.all(K[-1] == 1.0) and np.all(K >= 0)
    assert np.all(m.kappap[0] > 0)
    # kappa_P(T) -> kappa_abs at long wavelengths for cold dust, grows with T
    assert m.kappap[0][-1] > m.kappap[0][0]


def test_dust_normalized_at_V():
    d = M.synthetic_dust()
    lam = 1.2398 / d["nudust"]
    k_v = np.interp(0.55, lam, d["kappad"])
    assert abs(k_v - 1.0) < 2e-2
    assert abs(d["kappav"] - 200.0 * 0.55 ** -1.5) / d["kappav"] < 2e-2


def test_fractal_field_statistics():
    f = M.fractal_field((16, 20, 24), 0.5, 2.6, 7)
    lf = np.log(f)
    assert abs(lf.std() - 0.5) < 1e-9
    assert abs(f.mean() - 1.0) < 0.1


def test_ulrich_mu0_solves_cubic():
    rng = np.random.default_rng(0)
    mu = rng.uniform(0, 1, 1000)
    rx = 10 ** rng.uniform(-2, 2, 1000)
    m0 = M.ulrich_mu0(mu, rx)
    res = m0 ** 3 + m0 * (rx - 1) - mu * rx
    assert np.max(np.abs(res) / (1 + rx)) < 1e-10


def test_batch_models_follow_their_ranges():
    b = M.batch_models(24, seed=1002, nfreq=20, nlam=40, nT=100, nnu=65)
    assert b.nmodels == 24
    for p, m in zip(b.params, b.models):
        assert 50 <= p["nr"] <= 400 and 40 <= p["nt"] <= 200 and p["nt"] % 2 == 0
        assert 1e-4 <= p["mdisk"] <= 1e-1 and 1.0 <= p["flare"] <= 1.3
        assert 0.1 <= p["rin_au"] <= 1.0 and 1e-7 <= p["mdot"] <= 1e-5
        assert m.nrg == p["nr"] + 1 and m.ntg == p["nt"] + 1 and m.npg == 2 and m.ipeel == 0 and m.nmu == 10
        # the disk is normalized to the drawn mass; the envelope adds to it
        from src3_mpi_cfitsio_amd.model import MSUN
        assert np.sum(m.rho * m.vcell) >= p["mdisk"] * MSUN * (1.0 - 1e-9)
    # seeded: the same seed gives the same grid
    b2 = M.batch_models(3, seed=1002, nfreq=20, nlam=40, nT=100, nnu=65)
    assert [p for p in b2.params] == b.params[:3]
    cat = b.concat()
    off = b.cell_offsets()
    assert cat["rho"].size == off[-1] == sum(m.ncell for m in b.models)
    k = 5
    assert np.array_equal(cat["rho"][off[k]:off[k + 1]], b.models[k].rho.ravel())
    assert cat["rarr"].size == sum(m.nrg for m in b.models)
