"""The oracle's setemiss (dust_5.f90:507-699) and what pins it.

* Bit for bit against the reference's own compiled qsimp/trapzd (nr_2.f),
  Bnu/dBnu/expm1 (mathsubs.f90): golden vectors in tests/golden, and the
  live build when oracle/_ref/libref.so is present.
* Closed forms of the whole table: a grey (constant) opacity gives the
  Planck mean kappa (1 - albedo) and the Rosseland mean kappa, because the
  normalized Planck function and its temperature derivative integrate to 1
  (15/pi^4 * 3! zeta(4) = 1, 15/pi^4 * 4!/4 zeta(4) = 1); an opacity linear
  in frequency gives the Planck mean T * 15/pi^4 * 4! zeta(5); the
  normalized cumulative table is the Planck CDF.
* The smoothing windows of the species above ndg (:612-653), checked on
  unit vectors, and against a running mean in the interior.
* The host model builder's Gauss-Legendre tables agree with qsimp's.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate, special

from src3_mpi_cfitsio_amd import model as M

_dp = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def gold():
    import os
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz"))


def test_qsimp_intervals_match_reference_bits(gold, oracle_lib):
    L = oracle_lib.oracle()
    for row, want, flag in zip(gold["qs_in"], gold["qs_out"], gold["qs_flag"]):
        out = C.c_double()
        f = L.orc_qsimp_seg(int(row[0]), *[float(v) for v in row[1:]], C.byref(out))
        assert out.value == want and f == flag, row
    assert all(L.orc_expm1_as(float(x)) == y for x, y in zip(gold["em1_x"], gold["em1"]))
    assert all(L.orc_dbnu(float(x), 1.0) == y for x, y in zip(gold["dbnu_x"], gold["dbnu"]))


def test_qsimp_live_reference(oracle_lib, ref_lib):
    if not hasattr(ref_lib, "ref_qsimp_seg"):
        pytest.skip("libref.so predates the setemiss bindings")
    L = oracle_lib.oracle()
    rng = np.random.default_rng(77)
    for k in range(500):
        k0, k1 = rng.uniform(1e-3, 10.0, 2)
        T = 10.0 ** rng.uniform(-5.0, 1.5)
        nu0, dnu = rng.uniform(0.0, 32.0) * T, rng.uniform(1e-4, 0.1) * T
        a, b, cv = C.c_double(), C.c_double(), C.c_int()
        fa = L.orc_qsimp_seg(k % 2, k0, k1, nu0, dnu, T, C.byref(a))
        ref_lib.ref_qsimp_seg(k % 2, k0, k1, nu0, dnu, T, C.byref(b), C.byref(cv))
        assert a.value == b.value and fa == cv.value


def _grid(nT=60, nnu=65):
    em = M.emission_tables(M.synthetic_dust(nlam=20), nT=nT, nnu=nnu)
    return em["Tint"], em["nuint"]


def _flat_dust(kappa, albedo, nlam=30, numax=200.0):
    nud = np.linspace(numax, 1e-3, nlam)[None, :]
    return nud, np.full_like(nud, kappa), np.full_like(nud, albedo)


def test_grey_opacity_gives_the_grey_means(oracle_lib):
    Tint, nuint = _grid()
    nud, kap, alb = _flat_dust(3.5, 0.25)
    kapint, kappap, kappar, nfail = oracle_lib.setemiss(nud, kap, alb, Tint, nuint)
    assert nfail == 0
    # Planck mean of kabs = (1 - albedo) kappa; Rosseland mean of chi = kappa
    np.testing.assert_allclose(kappap[0], 0.75 * 3.5, rtol=2e-6)
    np.testing.assert_allclose(kappar[0], 3.5, rtol=2e-6)
    # the normalized cumulative table is the Planck CDF at the grid points
    planck = lambda x: 15.0 / math.pi ** 4 * x ** 3 / math.expm1(x) if x > 0 else 0.0
    for j in (3, 10, 20, 40, 64):
        cdf = integrate.quad(planck, 0.0, nuint[j], epsabs=0, epsrel=1e-12)[0]
        np.testing.assert_allclose(kapint[0, j, :], cdf, rtol=5e-6, atol=1e-12)
    assert np.all(kapint[0, -1] == 1.0) and np.all(np.diff(kapint[0], axis=0) >= 0)


def test_linear_opacity_gives_the_zeta5_planck_mean(oracle_lib):
    """kappa(nu) = nu (exact under the table's linear interpolation): the
    Planck mean at T is T * 15/pi^4 * Gamma(5) zeta(5); a table looked up at
    x instead of nu = x T would lose the factor T."""
    Tint, nuint = _grid()
    sel = Tint * nuint[-1] < 190.0
    nud = np.linspace(200.0, 0.0, 41)[None, :]
    kapint, kappap, kappar, nfail = oracle_lib.setemiss(nud, nud.copy(), np.zeros_like(nud), Tint, nuint)
    # chi(0) = 0: chidB = dB/chi is unbounded on each first interval, where
    # qsimp (like the reference's) stops without converging
    assert nfail == len(Tint)
    want = Tint * 15.0 / math.pi ** 4 * special.gamma(5.0) * special.zeta(5.0)
    np.testing.assert_allclose(kappap[0][sel], want[sel], rtol=2e-6)
    assert sel.sum() > 20


def test_smoothing_windows(oracle_lib):
    """Two passes of kappap1(i) = mean(kappap(lo..hi)) with lo..hi = i-10..i+10
    inside, 1..2i+1 for i <= 10 and 2i-nT..nT for i > nT-10 (1-based)."""
    nT = 40
    Tint, nuint = _grid(nT=nT, nnu=17)
    nud = np.linspace(200.0, 0.0, 41)[None, :]
    kap, alb = nud + 1.0, np.zeros_like(nud)            # kappap grows with T: the windows show
    base = oracle_lib.setemiss(nud, kap, alb, Tint, nuint)[1][0]
    assert np.ptp(base) > 0.5 * base.max()
    sm = oracle_lib.setemiss(nud, kap, alb, Tint, nuint, smooth=[1])[1][0]
    # one pass, as a matrix, from its rows
    W = np.zeros((nT, nT))
    for i in range(1, nT + 1):
        lo, hi = (i - 10, i + 10) if 10 < i <= nT - 10 else ((1, 2 * i + 1) if i <= 10 else (2 * i - nT, nT))
        W[i - 1, lo - 1:hi] = 1.0 / (hi - lo + 1)
    np.testing.assert_allclose(sm, W @ (W @ base), rtol=1e-13)
    # spot rows: i = 1 averages 1..3, i = nT is left alone by a pass
    assert W[0, :3].tolist() == [1 / 3] * 3 and W[-1, -1] == 1.0 and W[-1, :-1].sum() == 0.0
    # interior rows are the 21-point running mean
    run = np.convolve(base, np.ones(21) / 21.0, mode="valid")
    np.testing.assert_allclose((W @ base)[10:nT - 10], run, rtol=1e-13)
    # a smoothed species needs 21 temperatures
    T20, nu20 = _grid(nT=20, nnu=9)
    assert oracle_lib.setemiss(nud, kap, alb, T20, nu20, smooth=[1])[3] == -1


def test_host_tables_agree_with_qsimp(oracle_lib):
    """model.emission_tables (8-node Gauss-Legendre per interval) against the
    reference's qsimp on the synthetic dust."""
    dust = M.synthetic_dust(nlam=40)
    em = M.emission_tables(dust, nT=80, nnu=129)
    alb = dust["adust"]
    kapint, kappap, kappar, nfail = oracle_lib.setemiss(dust["nudust"], dust["kappad"], alb, em["Tint"], em["nuint"])
    assert nfail == 0
    np.testing.assert_allclose(kappap[0], em["kappap"], rtol=1e-5)
    np.testing.assert_allclose(kapint[0], em["kapint"], rtol=1e-5, atol=1e-9)


def test_nlambda_limits_the_table(oracle_lib):
    """Entries past nlambda(id) are not read: garbage there changes nothing."""
    Tint, nuint = _grid(nT=30, nnu=33)
    dust = M.synthetic_dust(nlam=25)
    nud = np.stack([dust["nudust"], dust["nudust"]])
    kap = np.stack([dust["kappad"], dust["kappad"]])
    alb = np.stack([dust["adust"], dust["adust"]])
    nud[1, 20:], kap[1, 20:], alb[1, 20:] = -5.0, 1e9, 7.0
    a = oracle_lib.setemiss(nud[:1, :20], kap[:1, :20], alb[:1, :20], Tint, nuint)
    b = oracle_lib.setemiss(nud, kap, alb, Tint, nuint, nlambda=[25, 20])
    assert np.array_equal(a[0][0], b[0][1]) and np.array_equal(a[1][0], b[1][1])
