"""Model building: zerod (the Ulrich streamline cubic), densenv's envelope
density and gridset's cell volumes, in the oracle and in the host builder.

* zerod bit for bit against the reference's compiled zerod.f90 (golden
  vectors; the live build when present), including its "2 roots" case.
* densenv against closed forms of the Ulrich (1976) solution: the pole
  (mu0 = 1), the equator inside Rc (mu0 = sqrt(1 - r/Rc)), far outside Rc
  (radial streamlines, mu0 -> mu, rho -> rhoe0 (r/Rc)^-1.5 / sqrt(2) in
  every direction), the cut radii and the ambient floor; the cubic's
  residual at every root.
* Cell volumes: a full sphere shell adds up to 4/3 pi (r2^3 - r1^3).
* The numpy builder (model.zerod / model.densenv / cell_volumes) against the
  oracle.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from src3_mpi_cfitsio_amd import model as M


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz"))


def _orc_zerod(L, mu, z, start=-7.0):
    r = C.c_double(start)
    f = L.orc_zerod(float(mu), float(z), C.byref(r))
    return r.value, f


def test_zerod_matches_reference_bits(gold, oracle_lib):
    L = oracle_lib.oracle()
    for mu, z, want, flag in zip(gold["zd_mu"], gold["zd_z"], gold["zd_out"], gold["zd_flag"]):
        got, f = _orc_zerod(L, mu, z)
        assert got == want, (mu, z)
        assert (f == 1) == (flag == 1), (mu, z)
    assert (gold["zd_flag"] == 1).sum() >= 1          # the "2 roots" case is exercised


def test_zerod_live_reference(oracle_lib, ref_lib):
    if not hasattr(ref_lib, "ref_zerod"):
        pytest.skip("libref.so predates the zerod binding")
    L = oracle_lib.oracle()
    rng = np.random.default_rng(8)
    for mu, z in zip(rng.uniform(0, 1, 2000), 10.0 ** rng.uniform(-5, 5, 2000)):
        a, fa = _orc_zerod(L, mu, z)
        b, cv = C.c_double(-7.0), C.c_int()
        ref_lib.ref_zerod(float(mu), float(z), C.byref(b), C.byref(cv))
        assert a == b.value and (fa == 1) == (cv.value == 1)


def test_zerod_roots_solve_the_cubic(oracle_lib):
    L = oracle_lib.oracle()
    rng = np.random.default_rng(9)
    for mu, z in zip(rng.uniform(1e-6, 1, 3000), 10.0 ** rng.uniform(-4, 4, 3000)):
        x, f = _orc_zerod(L, mu, z)
        assert f == 0 and 0.0 < x <= 1.0 + 1e-12
        assert abs(x ** 3 + (z - 1.0) * x - z * mu) <= 1e-12 * (1.0 + z)


def test_host_zerod_and_densenv_match_the_oracle(oracle_lib):
    L = oracle_lib.oracle()
    rng = np.random.default_rng(10)
    mu, z = rng.uniform(0, 1, 4000), 10.0 ** rng.uniform(-4, 4, 4000)
    x, flag = M.zerod(mu, z)
    ref = np.array([_orc_zerod(L, a, b) for a, b in zip(mu, z)])
    assert np.array_equal(flag, ref[:, 1].astype(np.int32))
    # numpy's vectorized pow/cos/acos may differ from the C library's in the
    # last bit; Cardano's sum factor3 + q/factor3 cancels for r/Rc >> 1 and
    # small mu (root ~ mu, terms ~ (z mu)^(1/3)), which turns one ulp into
    # ~1e-14 absolute on a root in (0, 1]
    np.testing.assert_allclose(x, ref[:, 0], rtol=1e-12, atol=5e-14)
    rad = 10.0 ** rng.uniform(0, 5, 4000)
    cost = rng.uniform(-1, 1, 4000)
    args = (300.0, 2.5e-19, 20.0, 9e4, 1e-23)
    d, fl = M.densenv(rad, cost, *args)
    fo = C.c_int32()
    want = np.array([L.orc_densenv(float(r), float(c), 0.0, *args, C.byref(fo)) for r, c in zip(rad, cost)])
    assert not fl.any()
    np.testing.assert_allclose(d, want, rtol=1e-10, atol=0)


def test_densenv_closed_forms(oracle_lib):
    L = oracle_lib.oracle()
    fo = C.c_int32()
    rc, rho0, rmine, rmax = 500.0, 1.0, 1.0, 1e12

    def dens(r, cost, rhoamb=0.0):
        return L.orc_densenv(r, cost, 0.0, rc, rho0, rmine, rmax, rhoamb, C.byref(fo))

    for r in (20.0, 400.0, 3000.0, 1e6):
        rx = r / rc
        # the pole: mu0 = 1
        np.testing.assert_allclose(dens(r, 1.0), rx ** -1.5 / (1.0 + 2.0 / rx) / math.sqrt(2.0), rtol=1e-12)
        np.testing.assert_allclose(dens(r, -1.0), dens(r, 1.0), rtol=0)        # +z = -z
    for r in (20.0, 100.0, 450.0):
        rx = r / rc
        mu = 1e-9                                                              # next to the equator, inside Rc
        m0 = math.sqrt(1.0 - rx)
        np.testing.assert_allclose(dens(r, mu), rx ** -1.5 / (2.0 * m0 * m0 / rx) / math.sqrt(1.0), rtol=1e-6)
    # far outside Rc the streamlines are radial: mu0 -> mu
    r = 1e9
    for mu in (0.1, 0.5, 0.9):
        rx = r / rc
        np.testing.assert_allclose(dens(r, mu), rx ** -1.5 / math.sqrt(2.0), rtol=1e-5)
    # cut radii and the ambient floor
    assert dens(rmine, 0.5) == 0.0 and dens(0.5 * rmine, 0.5) == 0.0 and dens(rmax, 0.5) == 0.0
    assert dens(1e11, 0.5, rhoamb=1e-3) == 1e-3 and dens(20.0, 0.5, rhoamb=1e-30) == dens(20.0, 0.5)


def test_cell_volumes(oracle_lib):
    L = oracle_lib.oracle()
    r = M.radial_walls(30, 0.1, 100.0, 2.0)
    t = M.theta_walls(20)
    p = M.phi_walls(6)[0]
    rm = 2.0 * M.RSUN
    v = M.cell_volumes(r, t, p, 2.0)
    want = np.array([[[L.orc_cell_volume(r[i], r[i + 1], t[j], t[j + 1], p[k], p[k + 1], rm)
                       for i in range(len(r) - 1)] for j in range(len(t) - 1)] for k in range(len(p) - 1)])
    np.testing.assert_allclose(v, want, rtol=1e-14)
    assert np.all(v > 0)
    shell = v.sum(axis=(0, 1))
    np.testing.assert_allclose(shell, 4.0 / 3.0 * math.pi * (r[1:] ** 3 - r[:-1] ** 3) * rm ** 3, rtol=1e-12)
