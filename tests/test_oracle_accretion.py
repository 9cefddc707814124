"""Disk-accretion photons (initpacc_4.f90) in the oracle and the host tables.

* The per-run split (accretion photons, then hot-spot photons, external
  packets last) against a literal walk of the reference's packet loop
  (newdisktherm_5_fuv.f90:164-190 for the counts, :242, :296, :416 for the
  branch of each iphot).
* accretion_tables against the closed-form integral of the steady-disk
  dissipation, int r^-2 (1 - r^-1/2) dr = -1/r + (2/3) r^-3/2, evaluated
  at the cell walls, summed over the cells before each one; with an inner
  hole no photon is dropped.
* Placement statistics in an optically thin medium, where the only
  absorption of an accretion photon is the one where it is emitted (nabs):
  the radial cell j takes Ldint(j+1) - Ldint(j) (locate's j); heights
  follow a Gaussian of
  scale hdisk(j), so the share of the two theta cells at the midplane is
  erf(c rp / (h sqrt(1 - c^2) sqrt 2)) for their outer wall cosine c;
  azimuths are uniform; a cell without dust is never chosen.
"""
import dataclasses
import math

import numpy as np
import pytest

from src3_mpi_cfitsio_amd import model as M


def _split_by_loop(np_, accfrac, accfrac2, isrf_frac):
    """Kind of every packet (0-based) from the reference's loop conditions."""
    npout = int(np_ * isrf_frac)
    if np_ * isrf_frac - npout > 0.5:
        npout += 1
    ns = int(np_ * (1.0 - accfrac - isrf_frac))
    if np_ * (1.0 - accfrac - isrf_frac) - ns > 0.5:
        ns += 1
    npacc = int(np_ * accfrac)
    npaccstar = int(accfrac2 * np_)
    if ns + npacc + npout != np_:
        ns = ns + np_ - (ns + npacc + npout)
    kinds = []
    for iphot in range(1, np_ + 1):
        if iphot < npacc:
            kinds.append("acc")
        elif iphot <= npacc + ns:
            kinds.append("hot" if iphot < npacc + npaccstar else "star")
        else:
            kinds.append("out")
    return kinds


def _split_by_oracle(om, np_, has_isrf, has_spots):
    om.set_run(np_)
    s = om.s
    kinds = []
    for i in range(np_):
        if i < s.acc_to:
            kinds.append("acc")
        elif has_isrf and s.isrf_from >= 0 and i >= s.isrf_from:
            kinds.append("out")
        elif has_spots and i < s.hot_to:
            kinds.append("hot")
        else:
            kinds.append("star")
    return kinds


@pytest.mark.parametrize("np_", [1, 2, 7, 100, 1001])
@pytest.mark.parametrize("accfrac,accfrac2,isrf", [(0.0, 0.1, 0.0), (0.2, 0.0, 0.0), (0.2, 0.15, 0.0),
                                                   (0.35, 0.1, 0.25), (0.004, 0.3, 0.1)])
def test_split_matches_reference_loop(np_, accfrac, accfrac2, isrf, oracle_lib):
    m = M.small("accsplit", nr=8, nt=6, nphi=1)
    if accfrac:
        m = M.with_accretion(m, accfrac)
    if accfrac2:
        m = M.with_spots(m, accfrac2=accfrac2)
    if isrf:
        m = M.with_isrf(m, isrf)
    om = oracle_lib.OracleModel(m)
    want = _split_by_loop(np_, accfrac, accfrac2, isrf)
    got = _split_by_oracle(om, np_, bool(isrf), bool(accfrac2))
    if not accfrac2:        # without spots the reference's hot range is stellar photons of the star
        want = ["star" if k == "hot" else k for k in want]
    assert got == want


def test_tables_follow_the_steady_disk_integral():
    m = M.small("acctab", nr=60, nt=6, nphi=1)
    rin_au, rout_au = 0.3, 300.0
    t = M.accretion_tables(m, rin_au=rin_au, rout_au=rout_au, h0_au=0.05, flare=1.2)
    ld, hd = t["ldint"], t["hdisk"]
    assert ld.shape == (m.nrg - 1,) and hd.shape == (m.nrg - 1,)
    assert ld[0] == 0.0 and ld[-1] == 1.0 and np.all(np.diff(ld) >= 0.0)
    r = np.asarray(m.rarr)
    autors = M.AU / (m.rstar * M.RSUN)
    rave = 0.5 * (r[:-1] + r[1:])
    np.testing.assert_allclose(hd, 0.05 * (rave / autors) ** 1.2 * autors, rtol=1e-12)
    # the midpoint sum over a cell approximates the integral over it; the
    # cell j is chosen with Ldint(j+1) - Ldint(j): its own luminosity
    inside = (rave / autors >= rin_au) & (rave / autors <= rout_au)
    F = lambda x: -1.0 / x + (2.0 / 3.0) * x ** -1.5
    exact = np.where(inside, F(r[1:]) - F(r[:-1]), 0.0)
    want = np.concatenate(([0.0], np.cumsum(exact[:-1])))
    np.testing.assert_allclose(ld, want / want[-1], atol=0.02)
    first = int(np.argmax(inside))
    assert ld[first] == 0.0 and ld[first + 1] > 0.0       # nothing falls inside the disk's inner edge


def test_default_tables_drop_no_photon(oracle_lib):
    """The synthetic disk has an inner hole; every share lands on dust."""
    m = M.with_accretion(M.small("acchole", nr=20, nt=12, nphi=4, clump_sigma=0.5), 0.5)
    o = oracle_lib.run_packets(m, 3000, seed=41, nthreads=8)
    assert o["counters"]["errors"] == 0


def _thin(m, empty_wedge=None):
    rho = np.full(m.cell_shape, 1e-40)
    sp = np.ones(m.cell_shape, np.uint8)
    if empty_wedge is not None:
        rho[empty_wedge] = 0.0
        sp[empty_wedge] = 0
    return dataclasses.replace(m, rho=rho, species=sp)


def test_placement_statistics(oracle_lib):
    m = _thin(M.small("accplace", nr=24, nt=16, nphi=4))
    cw = np.asarray(m.costarr)
    w = int(np.argmin(np.abs(cw)))
    assert abs(cw[w]) < 1e-12                      # a theta wall at the midplane
    c1 = cw[w - 1]
    np.testing.assert_allclose(-cw[w + 1], c1, rtol=1e-12)
    r = np.asarray(m.rarr)
    rave = 0.5 * (r[:-1] + r[1:])
    t = M.accretion_tables(m)
    hd = c1 * rave / math.sqrt(1.0 - c1 * c1)      # P(midplane cells) = erf(1/sqrt 2)
    m = M.with_accretion(m, 0.9995, dict(ldint=t["ldint"], hdisk=hd))
    n = 20000
    o = oracle_lib.run_packets(m, n, seed=5, nthreads=8)
    nabs = o["nabs"].reshape(m.cell_shape).astype(np.float64)
    nacc = int(0.9995 * n) - 1
    total = nabs.sum()
    assert abs(total - nacc) <= 0.002 * nacc         # one absorption each, none later
    assert o["counters"]["errors"] == 0
    radial = nabs.sum(axis=(0, 1)) / total
    p = np.zeros(m.nrg - 1)
    p[:-1] = np.diff(t["ldint"])                     # cell j takes Ldint(j+1) - Ldint(j)
    tol = 5.0 * np.sqrt(p * (1 - p) / total) + 0.003
    assert np.all(np.abs(radial - p) <= tol), (radial, p)
    mid = nabs[:, w - 1:w + 1, :].sum() / total
    pm = math.erf(1.0 / math.sqrt(2.0))
    assert abs(mid - pm) <= 5.0 * math.sqrt(pm * (1 - pm) / total) + 0.003, mid
    az = nabs.sum(axis=(1, 2)) / total
    assert np.all(np.abs(az - 0.25) <= 5.0 * math.sqrt(0.25 * 0.75 / total)), az


def test_placement_skips_cells_without_dust(oracle_lib):
    m = _thin(M.small("accempty", nr=16, nt=10, nphi=4), empty_wedge=(slice(1, 2), slice(None), slice(None)))
    m = M.with_accretion(m, 0.999)
    n = 6000
    o = oracle_lib.run_packets(m, n, seed=6, nthreads=8)
    nabs = o["nabs"].reshape(m.cell_shape).astype(np.float64)
    assert nabs[1].sum() == 0
    az = nabs.sum(axis=(1, 2)) / nabs.sum()
    assert np.all(np.abs(az[[0, 2, 3]] - 1.0 / 3.0) <= 5.0 * math.sqrt(2.0 / 9.0 / nabs.sum()))
