"""GPU tests of the batched model grid (mcrt_set_batch / mcrt_run_batch).

Packet p of model k in a batch runs with seed + k, so every model of a
batch must reproduce a single-model run of that model: integer tallies and
SED packet counts exactly, FP64 sums to their atomic summation order.  The
captured graph must give the eager result on every replay (with other work
in between), and neither the launch shape nor the write-combined tally may
change the results.
"""
import numpy as np
import pytest

from src3_mpi_cfitsio_amd import model as M

pytestmark = pytest.mark.gpu

SEED = 4242
NPK = 20000


@pytest.fixture(scope="module")
def batch():
    return M.batch_models(6, seed=11, nr_range=(12, 40), nt_range=(8, 24), nfreq=20, nlam=40, nT=100, nnu=65)


@pytest.fixture(scope="module")
def beng(batch):
    from src3_mpi_cfitsio_amd.engine import BatchEngine
    e = BatchEngine(batch)
    yield e
    e.close()


def _npk(batch):
    n = np.full(batch.nmodels, NPK, dtype=np.int64)
    n[1] = 0              # an empty model
    n[2] = NPK // 3       # uneven counts
    return n


def _same(a, b, what, rtol=1e-12):
    for k, (x, y) in enumerate(zip(a, b)):
        assert np.array_equal(x["nabs"], y["nabs"]), f"{what}: model {k} nabs"
        assert np.array_equal(x["killed"], y["killed"]), f"{what}: model {k} killed"
        assert np.array_equal(x["nums"], y["nums"]), f"{what}: model {k} SED packet counts"
        np.testing.assert_allclose(x["dtauabs"], y["dtauabs"], rtol=rtol, atol=1e-300, err_msg=f"{what}: model {k} dtauabs")
        np.testing.assert_allclose(x["si"], y["si"], rtol=1e-5, atol=1e-6, err_msg=f"{what}: model {k} SED I")


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_matches_single_model_runs(batch, beng, precision):
    from src3_mpi_cfitsio_amd.engine import Engine
    npk = _npk(batch)
    beng.set_batch_launch(graph=0)
    beng.run_batch(npk, SEED, precision=precision)
    got, work, cnt = beng.batch_tallies()
    ref = []
    for k, m in enumerate(batch.models):
        with Engine(m) as e:
            e.zero_tallies()
            if npk[k]:
                e.run_packets(int(npk[k]), seed=SEED + k, precision=precision)
            ref.append(e.tallies())
    print(f"precision {precision}: work per model {work}, escaped {cnt['escaped']:.0f}, killed {cnt['killed']:.0f}, "
          f"errors {cnt['errors']:.0f}", flush=True)
    assert work[1] == 0 and np.all(work[npk > 0] >= npk[npk > 0])
    assert cnt["escaped"] + cnt["killed"] + cnt["errors"] == npk.sum()
    _same(got, ref, f"batch vs single, precision {precision}")


def test_graph_replays_equal_eager(batch, beng):
    from src3_mpi_cfitsio_amd.engine import Engine
    npk = _npk(batch)
    beng.set_batch_launch(graph=0)
    beng.run_batch(npk, SEED + 1, precision=1)
    eager, _, c0 = beng.batch_tallies()
    beng.set_batch_launch(graph=1)
    other = Engine(batch.models[0])
    try:
        for rep in range(3):
            beng.run_batch(npk, SEED + 1, precision=1)
            got, _, c1 = beng.batch_tallies()
            assert c1 == c0, f"replay {rep}: counters differ"
            _same(got, eager, f"graph replay {rep}")
            # unrelated eager work on the device between replays
            other.zero_tallies()
            other.run_packets(2000, seed=rep, precision=1)
            other.synchronize()
    finally:
        other.close()


def test_launch_shape_and_combining_do_not_change_results(batch, beng):
    npk = _npk(batch)
    res = []
    for shape in [(256, 2, 32, 0), (64, 1, 16, 1), (128, 2, 48, 1), (64, 2, 1, 0)]:
        beng.set_batch_launch(*shape, graph=0)
        beng.run_batch(npk, SEED + 2, precision=0)
        res.append(beng.batch_tallies()[0])
    for r in res[1:]:
        _same(r, res[0], "launch shape", rtol=1e-12)


def test_autotune_picks_a_fitting_shape(batch, beng):
    best, table = beng.autotune(packets_per_model=4000, seed=9, precision=1)
    print("autotune best (block, waves, regroup, combine):", best, "of", len(table), "shapes", flush=True)
    for row in table:
        print("   ", [int(x) for x in row[:4]], f"{row[4]:.3f} ms {row[5]:.4g} packets/s", flush=True)
    assert best[0] in (64, 128, 256) and best[1] in (1, 2, 4) and best[2] in (16, 32, 48) and best[3] in (0, 1)
    assert len(table) >= 6 and np.all(table[:, 4] > 0)
    # the tuned shape gives the same tallies
    npk = _npk(batch)
    beng.set_batch_launch(graph=1)
    beng.run_batch(npk, SEED + 3, precision=0)
    a = beng.batch_tallies()[0]
    beng.set_batch_launch(256, 2, 32, 0, graph=0)
    beng.run_batch(npk, SEED + 3, precision=0)
    b = beng.batch_tallies()[0]
    _same(a, b, "tuned vs default shape")


def test_batch_dusttemp_matches_single_models(batch, beng):
    from src3_mpi_cfitsio_amd.engine import Engine
    npk = np.full(batch.nmodels, NPK, dtype=np.int64)
    beng.set_batch_launch(256, 2, 32, 0, graph=0)
    beng.run_batch(npk, SEED + 4, precision=0)
    T = beng.batch_dusttemp(update_tdust=False)
    for k, m in enumerate(batch.models):
        with Engine(m) as e:
            e.zero_tallies()
            e.run_packets(NPK, seed=SEED + 4 + k, precision=0)
            t = e.dusttemp(NPK, update_tdust=False)
        np.testing.assert_allclose(T[k], t, rtol=1e-9, err_msg=f"model {k}")
