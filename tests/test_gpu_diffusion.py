"""The diffusion approximation (setdiffus_5.f90, emit_common_2.f90:88-378,
tfinal_5_test1.f90:232-233) on the GPU against the oracle.

- setdiffus: the diffusion region of every column, from the same Lucy
  temperatures, equals the oracle's literal restatement of the reference's
  loops (iteration 1 with its 50 K floor and a later iteration).
- a run: no absorption in the region is counted in nabs, every one of them
  is re-emitted once from a surface cell (dbig there = the surface counts),
  photons that never enter the region follow the oracle packet by packet,
  and the surface counts agree with the oracle's within Poisson noise.
  The surface temperature a photon sees depends on the order photons
  arrive in, which differs between the GPU and a serial run, so packets
  re-emitted from a surface part ways with the oracle's.
- the temperatures the run leaves in the region and the Lucy temperatures
  equal the oracle's on the GPU's own counts.
- the mixed path runs the same region without errors.
"""
import math
import os

import numpy as np
import pytest

from src3_mpi_cfitsio_amd import model as M
from gpu_util import rel_err

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def _thick(nphi=1):
    m = M.small(f"diff{nphi}", nr=30, nt=20, nphi=nphi, mdisk=0.1)
    return M.with_diffusion(m)


@pytest.mark.parametrize("nphi", [1, 4])
def test_diffusion_region_run_and_temperatures_match_oracle(nphi, oracle_lib):
    from src3_mpi_cfitsio_amd.engine import Engine
    m = _thick(nphi)
    n = 20_000
    with Engine(m) as e:
        # iteration 1 (setdiffus at the start temperatures, floored at 50 K)
        tal1 = e.disk(n, seed=1, precision=0, iteration=1)
        it1, _ = e.diffusion_state()
        _, _, o_it1 = oracle_lib.setdiffus(m, m.tdust_initial(), 1)
        assert np.array_equal(it1, o_it1)
        # iteration 2 from the first iteration's temperatures
        td = tal1["tdave"]
        ncell = e.setdiffus(2)
        it2, _ = e.diffusion_state()
        dif, ddir, o_it2 = oracle_lib.setdiffus(m, td, 2)
        print(f"nphi {nphi}: diffusion cells {ncell}, columns {np.count_nonzero(it2)}", flush=True)
        assert np.array_equal(it2, o_it2) and ncell == dif.sum() and ncell > 0
        assert np.all(ddir[dif] != 0)
        tdust = np.where(dif | (m.species == 0), 0.1 / 11605.0, td)
        e.zero_tallies()
        e.enable_trace(n)
        e.run_packets(n, seed=2, precision=0, first_packet=10 ** 7)
        gt = e.get_trace(n)
        tg = e.tallies()
        e.enable_trace(0)
        _, cg = e.diffusion_state()
        tgpu = e.dusttemp(n, update_tdust=False)
    cnt = np.zeros(m.cell_shape, np.uint32)
    o = oracle_lib.run_packets(m, n, seed=2, first_packet=10 ** 7, tdust=tdust, nthreads=1, trace=True,
                               diffusion=dict(itmin=o_it2, count=cnt))
    for name, t, c in (("gpu", tg, cg), ("oracle", o, cnt)):
        assert t["nabs"][dif].sum() == 0, name
        assert t["dbig"][dif].sum() == c.sum() > 0, name
        assert t["counters"]["errors"] == 0, name
    # surface counts: Poisson-equal totals and the same surface cells
    a, b = float(cg.sum()), float(cnt.sum())
    assert abs(a - b) <= 5.0 * math.sqrt(a + b) + 2, (a, b)
    assert np.array_equal(cg > 0, cnt > 0) or abs(np.count_nonzero(cg) - np.count_nonzero(cnt)) <= 2
    # packets that never diffuse follow the oracle exactly
    same = np.all(rel_err(o["trace"], gt) <= 1e-9, axis=1)
    frac = b / n
    print(f"nphi {nphi}: surface photons gpu {a:.0f} oracle {b:.0f}, agreement {same.mean():.4f}", flush=True)
    assert same.mean() >= 1.0 - 1.5 * frac - 0.01
    for k in ("escaped", "killed"):
        assert abs(tg["counters"][k] - o["counters"][k]) <= 5.0 * math.sqrt(frac * n) + 0.01 * n, k
    # the region's temperatures and the Lucy temperatures on the GPU's counts
    tcol = oracle_lib.diff_columns(m, it2, cg, tg["nabs"], n, tdust)
    torc = oracle_lib.tfinal2(m, tg["dtauabs"], n, tcol, it2)
    assert np.all(rel_err(tgpu, torc) <= 1e-12)
    assert tgpu[dif].max() > 5.0 * 0.1 / 11605.0


def test_diffusion_mixed_and_lucy_loop():
    from src3_mpi_cfitsio_amd.engine import Engine
    m = _thick(4)
    with Engine(m) as e:
        hist, tal = e.lucy(20_000, seed=9, precision=1, n_iter_min=3, n_iter_max=3)
        itmin, cnt = e.diffusion_state()
        assert len(hist) == 3 and tal["counters"]["errors"] == 0
        assert np.count_nonzero(itmin) > 0 and cnt.sum() > 0
        assert np.all(np.isfinite(tal["tdave"])) and np.all(np.isfinite(tal["si"]))
        dif = np.zeros(m.cell_shape, bool)
        for ir, a in enumerate(itmin):
            if a > 0:
                dif[:, a - 1:m.ntg - a, ir] = True
        assert tal["nabs"][dif].sum() == 0


def test_diffusion_refused_where_it_cannot_apply():
    from src3_mpi_cfitsio_amd.engine import Engine, MCRTError
    m = _thick(1)
    W = np.ones((m.nband, m.nx, m.nx, m.npeel), np.float32)
    with Engine(m) as e:
        with pytest.raises(MCRTError, match="call mcrt_setdiffus first"):
            e.run_packets(100, seed=1)
        with pytest.raises(MCRTError, match="replay cannot reproduce"):
            e.adjoint_image(W, 100, seed=1)
        e.set_diffusion(None)
        e.run_packets(100, seed=1)
