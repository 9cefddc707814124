"""Stellar hot spots (initpspot.f90 with ispot = 1, spotset.f90) in the
oracle.

* Geometry: the caps cover the filling factor fspot of the sphere (one
  cap: (1 - cos r)/2 = fspot; two antipodal caps of half that each).
* Photons, in an empty medium (every photon escapes unscattered, so its
  frequency is the emitted one):
  - hot-spot photons: half are X-rays, uniform in h nu over
    1.2398/0.06..1.2398/0.015 eV; the rest follow the shock temperature's
    Planck spectrum, mean h nu = (4 zeta(5)/zeta(4)) k T_shock;
  - stellar photons: a fraction fspot lands inside a cap and takes the
    shock temperature, so the mean of their h nu is the fspot-weighted
    mixture of the two Planck means;
  - with fspot = 1 every stellar photon has the shock temperature.
"""
import math

import numpy as np
import pytest
from scipy import special

from src3_mpi_cfitsio_amd import model as M

PLANCK_MEAN = 4.0 * special.zeta(5.0) / special.zeta(4.0)     # <h nu / kT> of B_nu


@pytest.mark.parametrize("nspot,fspot,lat", [(1, 0.01, 60.0), (1, 0.3, -20.0), (2, 0.2, 75.0), (2, 1.0, 0.0)])
def test_spot_geometry(oracle_lib, nspot, fspot, lat):
    g = oracle_lib.spot_geometry(nspot, fspot, lat, 30.0)
    c1 = np.array([g["scspot1"], g["ssspot1"], g["cspot1"]])
    c2 = np.array([g["scspot2"], g["ssspot2"], g["cspot2"]])
    assert abs(np.linalg.norm(c1) - 1.0) < 1e-14 and np.allclose(c2, -c1, atol=1e-14)
    assert abs(c1[2] - math.sin(math.radians(lat))) < 1e-14
    area = nspot * (1.0 - g["csprad"]) / 2.0
    assert abs(area - fspot) < 1e-12


@pytest.fixture(scope="module")
def empty():
    m = M.small("spots", nr=10, nt=8, nphi=4, npeel=2, nx=9, nfreq=40)
    return M.dataclasses.replace(m, rho=np.zeros_like(m.rho), species=np.zeros_like(m.species))


def _run(oracle_lib, m, n, seed):
    r = oracle_lib.run_packets(m, n, seed=seed, trace=True)
    assert r["counters"]["escaped"] == n and r["counters"]["errors"] == 0
    assert np.all(r["trace"][:, 6] == 0)                      # unscattered
    return r["trace"][:, 0]


def test_hot_spot_and_stellar_photon_frequencies(oracle_lib, empty):
    n = 60000
    tstar = empty.tstar
    m = M.with_spots(empty, nspot=2, fspot=0.3, spotlat=40.0, tshock_k=3.0 * tstar * M.TK, accfrac2=0.4)
    nub = _run(oracle_lib, m, n, seed=7)
    hot_to = int(0.4 * n) - 1                                  # iphot < npaccstar, 1-based
    hot, star = nub[:hot_to], nub[hot_to:]
    lo, hi = 1.2398 / 0.06, 1.2398 / 0.015
    xr = (hot >= lo) & (hot <= hi)
    assert abs(xr.mean() - 0.5) < 4.0 * math.sqrt(0.25 / hot_to)
    assert abs(np.mean((hot[xr] - lo) / (hi - lo)) - 0.5) < 4.0 * math.sqrt(1 / 12.0 / xr.sum())
    planck = hot[~xr]
    want = PLANCK_MEAN * m.spots["tshock"]
    sd = math.sqrt(20.0 * special.zeta(6.0) / special.zeta(4.0) - PLANCK_MEAN ** 2) * m.spots["tshock"]
    assert abs(planck.mean() - want) < 4.0 * sd / math.sqrt(len(planck))
    # stellar photons: a fraction fspot inside the caps at the shock temperature
    mix = PLANCK_MEAN * ((1.0 - 0.3) * tstar + 0.3 * m.spots["tshock"])
    var = (20.0 * special.zeta(6.0) / special.zeta(4.0)) * ((1 - 0.3) * tstar ** 2 + 0.3 * m.spots["tshock"] ** 2) - mix ** 2
    assert abs(star.mean() - mix) < 4.0 * math.sqrt(var / len(star))
    assert np.all(star < lo)                                  # no X-rays among stellar photons


def test_full_coverage_heats_every_stellar_photon(oracle_lib, empty):
    n = 20000
    m = M.with_spots(empty, nspot=1, fspot=1.0, tshock_k=2.0 * empty.tstar * M.TK, accfrac2=0.0)
    nub = _run(oracle_lib, m, n, seed=9)
    sd = math.sqrt(20.0 * special.zeta(6.0) / special.zeta(4.0) - PLANCK_MEAN ** 2) * m.spots["tshock"]
    assert abs(nub.mean() - PLANCK_MEAN * m.spots["tshock"]) < 4.0 * sd / math.sqrt(n)


def test_spots_off_is_the_plain_star(oracle_lib, empty):
    a = oracle_lib.run_packets(empty, 3000, seed=5, trace=True)["trace"]
    b = oracle_lib.run_packets(M.dataclasses.replace(empty, spots=None), 3000, seed=5, trace=True)["trace"]
    assert np.array_equal(a, b)
