"""GPU disk-accretion photons (mcrt_set_accretion) against the oracle: whole
packets on the same counter-based draws, accretion photons first in every
run (placed by k_acc_place, absorbed and re-emitted where they start), then
hot-spot photons, external photons last.  Bar as for the stellar packets:
per-packet agreement up to rare last-bit branch flips, outcome counts, the
absorption tallies and images within 2%.  The mixed-precision path is held
to the same statistical bar on the accretion-only tallies."""
import dataclasses

import numpy as np
import pytest

from gpu_util import rel_err
from src3_mpi_cfitsio_amd import model as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("accfrac,spots,isrf", [(0.4, False, 0.0), (0.3, True, 0.0), (0.25, True, 0.2)])
def test_accretion_packets_match_oracle(accfrac, spots, isrf, oracle_lib):
    from src3_mpi_cfitsio_amd.engine import Engine
    m = M.small("acc", nr=20, nt=12, nphi=4, clump_sigma=0.5)
    m = M.with_accretion(m, accfrac)
    if spots:
        m = M.with_spots(m, nspot=1, fspot=0.05, spotlat=50.0, tshock_k=9000.0, accfrac2=0.2)
    if isrf:
        m = M.with_isrf(m, isrf)
    n = 3000
    print(f"accretion parity accfrac={accfrac} spots={spots} isrf={isrf} n={n}", flush=True)
    with Engine(m) as e:
        e.zero_tallies()
        e.enable_trace(n)
        e.run_packets(n, seed=41)
        gt = e.get_trace(n)
        tg = e.tallies()
    o = oracle_lib.run_packets(m, n, seed=41, trace=True)
    same = np.all(rel_err(o["trace"], gt) <= 1e-9, axis=1)
    acc_to = int(accfrac * n) - 1
    print(f"agreement acc {same[:acc_to].mean():.4f}, rest {same[acc_to:].mean():.4f}", flush=True)
    assert same[:acc_to].mean() >= 0.99 and same[acc_to:].mean() >= 0.99
    assert tg["counters"]["errors"] == 0 and o["counters"]["errors"] == 0
    for k in ("escaped", "killed"):
        assert abs(tg["counters"][k] - o["counters"][k]) <= 0.01 * n + 1
    for name in ("nabs", "si", "image_b_i"):
        a, b = tg[name].astype(np.float64), o[name].astype(np.float64)
        assert abs(a.sum() - b.sum()) <= 0.02 * max(abs(b.sum()), 1e-30), name


def test_parked_accretion_photons_match_oracle(oracle_lib):
    """One event per lane and round: a placed accretion photon waits for its
    absorption across a round boundary (parked with i_orig = 0)."""
    from src3_mpi_cfitsio_amd.engine import Engine
    m = M.with_accretion(M.small("accpark", nr=16, nt=10, nphi=2), 0.5)
    n = 1000
    print(f"parked accretion photons n={n}, 256 lanes, 1 event per round", flush=True)
    with Engine(m) as e:
        e.set_launch(lanes=256, queue_per_lane=8, events_per_round=1)
        e.zero_tallies()
        e.enable_trace(n)
        e.run_packets(n, seed=44)
        gt = e.get_trace(n)
        tg = e.tallies()
    o = oracle_lib.run_packets(m, n, seed=44, trace=True)
    same = np.all(rel_err(o["trace"], gt) <= 1e-9, axis=1)
    print(f"agreement {same.mean():.4f}", flush=True)
    assert same.mean() >= 0.99
    assert tg["counters"]["errors"] == 0
    assert abs(tg["nabs"].sum() - o["nabs"].sum()) <= 0.02 * o["nabs"].sum()


def test_mixed_precision_accretion_statistics(oracle_lib):
    from src3_mpi_cfitsio_amd.engine import Engine, MCRT_MIXED
    m = M.small("accmix", nr=20, nt=12, nphi=4)
    rho = np.full(m.cell_shape, 1e-40)
    m = dataclasses.replace(m, rho=rho, species=np.ones(m.cell_shape, np.uint8))
    m = M.with_accretion(m, 0.999)
    n = 40000
    print(f"accretion mixed statistics n={n}", flush=True)
    with Engine(m) as e:
        e.zero_tallies()
        e.run_packets(n, seed=42, precision=MCRT_MIXED)
        tg = e.tallies()
    o = oracle_lib.run_packets(m, n, seed=43, nthreads=8)
    a = tg["nabs"].reshape(m.cell_shape).astype(np.float64)
    b = o["nabs"].reshape(m.cell_shape).astype(np.float64)
    assert tg["counters"]["errors"] == 0
    assert abs(a.sum() - b.sum()) <= 2
    ra, rb = a.sum(axis=(0, 1)) / a.sum(), b.sum(axis=(0, 1)) / b.sum()
    assert np.all(np.abs(ra - rb) <= 5.0 * np.sqrt(np.maximum(rb, 1e-4) * 2.0 / b.sum()) + 1e-3), (ra, rb)
    ta, tb = a.sum(axis=(0, 2)) / a.sum(), b.sum(axis=(0, 2)) / b.sum()
    assert np.all(np.abs(ta - tb) <= 5.0 * np.sqrt(np.maximum(tb, 1e-4) * 2.0 / b.sum()) + 1e-3), (ta, tb)


def test_accretion_rejects_bad_tables():
    from src3_mpi_cfitsio_amd.engine import Engine, MCRTError
    m = M.small("accbad", nr=8, nt=6, nphi=1)
    t = M.accretion_tables(m)
    with Engine(m) as e:
        bad = dict(accfrac=0.2, ldint=t["ldint"][::-1].copy(), hdisk=t["hdisk"])
        with pytest.raises(MCRTError, match="Ldint"):
            e.set_accretion(bad)
        with pytest.raises(MCRTError, match="hdisk"):
            e.set_accretion(dict(accfrac=0.2, ldint=t["ldint"], hdisk=-t["hdisk"]))
        with pytest.raises(MCRTError, match="accfrac"):
            e.set_accretion(dict(accfrac=1.0, ldint=t["ldint"], hdisk=t["hdisk"]))
        e.set_accretion(None)


def test_accretion_at_warm_dust_matches_oracle(oracle_lib):
    """Accretion photons are re-emitted at the dust temperature of the cell
    they start in: with a warm field (150-900 K over the disk, falling with
    radius) re-emission samples interior rows of every CDF, not the f = 0
    row of the 0.1 K start."""
    from src3_mpi_cfitsio_amd.engine import Engine
    m = M.with_accretion(M.small("accw", nr=20, nt=12, nphi=4, clump_sigma=0.5), 0.5)
    rc = 0.5 * (m.rarr[1:] + m.rarr[:-1])
    warm = (900.0 * (rc / rc[0]) ** -0.4)[None, None, :] * np.ones(m.cell_shape)
    tdust = np.where(m.species > 0, np.maximum(warm, 150.0), 0.1) / 11605.0
    n = 3000
    print(f"accretion parity at warm dust, n={n}", flush=True)
    with Engine(m) as e:
        e.set_tdust(tdust)
        e.zero_tallies()
        e.enable_trace(n)
        e.run_packets(n, seed=43)
        gt = e.get_trace(n)
        tg = e.tallies()
    o = oracle_lib.run_packets(m, n, seed=43, tdust=tdust, trace=True)
    same = np.all(rel_err(o["trace"], gt) <= 1e-9, axis=1)
    acc_to = int(0.5 * n) - 1
    print(f"agreement acc {same[:acc_to].mean():.4f}, rest {same[acc_to:].mean():.4f}", flush=True)
    assert same[:acc_to].mean() >= 0.99 and same[acc_to:].mean() >= 0.99
    # re-emitted frequencies of the accretion photons follow the warm rows:
    # far above the 0.1 K row's frequencies
    assert np.median(gt[:acc_to, 0]) > 50.0 * 0.1 / 11605.0
    assert tg["counters"]["errors"] == 0
    for name in ("nabs", "si", "image_b_i"):
        a, b = tg[name].astype(np.float64), o[name].astype(np.float64)
        assert abs(a.sum() - b.sum()) <= 0.02 * max(abs(b.sum()), 1e-30), name
