"""GPU stellar hot spots (mcrt_set_spots) against the oracle: whole
packets on the same counter-based draws, hot-spot photons first in every
run, stellar photons that land in a cap at the shock temperature; with one
and two caps, partial and full coverage, and together with external
illumination.  Bar as for the stellar packets: per-packet agreement up to
rare last-bit branch flips, outcome counts, tallies within 2%."""
import numpy as np
import pytest

from gpu_util import rel_err
from src3_mpi_cfitsio_amd import model as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nspot,fspot,isrf", [(1, 0.05, 0.0), (2, 0.3, 0.0), (1, 1.0, 0.0), (2, 0.1, 0.3)])
def test_spot_packets_match_oracle(nspot, fspot, isrf, oracle_lib):
    from src3_mpi_cfitsio_amd.engine import Engine
    m = M.small("spots", nr=20, nt=12, nphi=4, clump_sigma=0.5)
    m = M.with_spots(m, nspot=nspot, fspot=fspot, spotlat=50.0, tshock_k=9000.0, accfrac2=0.3)
    if isrf:
        m = M.with_isrf(m, isrf)
    n = 3000
    with Engine(m) as e:
        e.zero_tallies()
        e.enable_trace(n)
        e.run_packets(n, seed=31)
        gt = e.get_trace(n)
        tg = e.tallies()
    o = oracle_lib.run_packets(m, n, seed=31, trace=True)
    same = np.all(rel_err(o["trace"], gt) <= 1e-9, axis=1)
    hot_to = int(0.3 * n) - 1
    print(f"nspot={nspot} fspot={fspot} isrf={isrf}: agreement hot {same[:hot_to].mean():.4f}, "
          f"rest {same[hot_to:].mean():.4f}", flush=True)
    assert same[:hot_to].mean() >= 0.99 and same[hot_to:].mean() >= 0.99
    assert tg["counters"]["errors"] == 0
    for k in ("escaped", "killed"):
        assert abs(tg["counters"][k] - o["counters"][k]) <= 0.01 * n + 1
    for name in ("si", "ti", "image_b_i"):
        a, b = tg[name].astype(np.float64), o[name].astype(np.float64)
        assert abs(a.sum() - b.sum()) <= 0.02 * max(abs(b.sum()), 1e-30), name


def test_spots_reject_bad_arguments():
    from src3_mpi_cfitsio_amd.engine import Engine, MCRTError
    m = M.small("spotbad", nr=8, nt=6, nphi=1)
    with Engine(m) as e:
        with pytest.raises(MCRTError, match="nspot"):
            e.set_spots(dict(nspot=3, fspot=0.1, spotlat=0.0, spotlon=0.0, tshock=1.0, accfrac2=0.1))
        with pytest.raises(MCRTError, match="fspot"):
            e.set_spots(dict(nspot=1, fspot=0.0, spotlat=0.0, spotlon=0.0, tshock=1.0, accfrac2=0.1))
