"""External illumination: the ISRF tables (isrf_cdf.f90 without
extinction), the packet split (newdisktherm_5_fuv.f90:158-187) and the
oracle's external photons (initpout.f90, rep_isrf.f90).

* Tables: log-log interpolation of the Black (1987) field checked with an
  independent interpolation (np.interp in log space); the CDF against the
  exact integral of the piecewise power law; the reference's first-entry
  values (logcdf = -99, logfreq = 0); the wavelength range 0.05..5000 um.
* Split: the host rule equals the oracle's, including x.5 rounding.
* Photons: in an empty medium every external photon escapes unscattered
  into SED components 1, 6 and 8 (the reference's indx for i_orig = 5);
  its sampled frequencies follow the table CDF (Kolmogorov-Smirnov).
"""
import math

import numpy as np
import pytest

from src3_mpi_cfitsio_amd import model as M


def test_isrf_tables():
    t = M.isrf_tables()
    c = 2.99792458e14
    lnu, lf = np.array(M._ISRF_LNU), np.array(M._ISRF_NUFNU)
    cdf, logcdf, logfreq = t["cdf"], t["logcdf"], t["logfreq"]
    assert cdf[0] == 0.0 and cdf[-1] == 1.0 and np.all(np.diff(cdf) > 0)
    assert logcdf[0] == -99.0 and logfreq[0] == 0.0
    np.testing.assert_allclose(logcdf[1:], np.log10(cdf[1:]), rtol=1e-14)
    # frequency grid: lam(i) = 0.05 (5000/0.05)^(i/200), i = 1..200, as h nu
    # in eV in ascending frequency; entry 1 (lam = 5000 um) has logfreq 0
    hnu = 10.0 ** logfreq[1:]
    lam = 1.2398 / hnu
    np.testing.assert_allclose(lam[-1], 0.05 * (5000.0 / 0.05) ** (1.0 / 200.0), rtol=1e-12)
    np.testing.assert_allclose(lam[0], 0.05 * (5000.0 / 0.05) ** (199.0 / 200.0), rtol=1e-12)
    assert np.allclose(np.diff(np.log(lam)), np.log(0.05 / 5000.0) / 200, rtol=1e-9)
    # pdf = log-log interpolation of the table (independent implementation)
    freq_hz = hnu * c / 1.2398
    want = 10.0 ** np.interp(np.log10(freq_hz), lnu, lf - lnu)
    np.testing.assert_allclose(t["pdf"][1:], want, rtol=1e-10)
    # unnormalized integral: exact integral of the power-law pieces over the
    # grid's range (the trapezoid rule on 200 log-spaced points is within 2%)
    f0, f1 = c / 5000.0, freq_hz[-1]
    grid = np.unique(np.concatenate([[f0, f1], 10.0 ** lnu[(10.0 ** lnu > f0) & (10.0 ** lnu < f1)]]))
    total = 0.0
    for a, b in zip(grid[:-1], grid[1:]):
        pa = 10.0 ** np.interp(math.log10(a), lnu, lf - lnu)
        pb = 10.0 ** np.interp(math.log10(b), lnu, lf - lnu)
        k = math.log(pb / pa) / math.log(b / a)
        total += pa * a / (k + 1.0) * ((b / a) ** (k + 1.0) - 1.0) if abs(k + 1.0) > 1e-12 else pa * a * math.log(b / a)
    assert abs(t["isrf_int"] / total - 1.0) < 0.02


def test_split_rule(oracle_lib):
    cases = [(10, 0.3), (11, 0.3), (1000, 0.25), (7, 0.5), (9, 0.5), (20000, 0.33333), (5, 0.1)]
    for n, f in cases:
        ns = M.isrf_split(n, f)
        m = M.with_isrf(M.small(nr=6, nt=4), f)
        om = oracle_lib.OracleModel(m)
        om.set_run(n)
        assert om.s.isrf_from == ns, (n, f)
        npout = n - ns
        assert abs(npout - f * n) <= 0.5 + 1e-9
    assert M.isrf_split(100, 0.0) == 100


@pytest.fixture(scope="module")
def empty_outside():
    m = M.small("empty", nr=10, nt=8, nphi=4, npeel=2, nx=9)
    m = M.dataclasses.replace(m, rho=np.zeros_like(m.rho), species=np.zeros_like(m.species))
    return M.with_isrf(m, 0.999)


def test_external_photons_cross_an_empty_grid(oracle_lib, empty_outside):
    m = empty_outside
    n = 20000
    r = oracle_lib.run_packets(m, n, seed=3, trace=True)
    ns = M.isrf_split(n, m.isrf_frac)
    tr = r["trace"]
    assert r["counters"]["escaped"] == n and r["counters"]["errors"] == 0
    assert np.all(tr[ns:, 6] == 0)                          # no interaction
    # SED components (propagate_5_fuv.f90:300-313, indx = 1, 1 + i_orig,
    # 5 + i_inter): stellar photons go to 1, 2 (STELLAR) and 6 (direct);
    # external ones (i_orig = 5, i_inter = 3 as for any non-stellar origin)
    # to 1, 6 and 8
    nums = r["nums"].sum(axis=(1, 2, 3, 4))
    assert nums[0] == n and nums[1] == ns and nums[5] == n and nums[7] == n - ns
    assert nums[2] == nums[3] == nums[4] == nums[6] == 0
    # the sampled frequencies follow the table CDF (KS, 99.9%)
    t = m.isrf
    nub = np.sort(tr[ns:, 0])
    x = np.log10(nub)
    cdf_at = np.interp(x, t["logfreq"][1:], t["cdf"][1:])
    emp = np.arange(1, len(x) + 1) / len(x)
    assert np.max(np.abs(emp - cdf_at)) < 1.95 / math.sqrt(len(x))
    assert r["counters"]["sflux"] == 0.0
