"""The end of a run in k_tail (one wavefront per remaining packet, the
rejection sampling of the scattering angle over 64 lanes at once) against
k_transport on the same RNG counters.

MCRT_TAIL_LANES=0 keeps every round in k_transport; a large value switches
to k_tail as soon as the last packet is handed out, so most of a short run's
interactions go through it.  k_tail draws the same random numbers in the same
order and does the same arithmetic, so:
- FP64: every packet's trace (frequency, direction, Stokes, interactions,
  outcome) is identical, integer tallies and packet counters are equal, and
  floating tallies differ only by the order of their atomic sums;
- mixed precision: the same, with the traces allowed to part for a few
  packets where a trial's weight is formed with another rounding (FMA
  contraction may differ between the two kernels).
"""
import os

import numpy as np
import pytest

from src3_mpi_cfitsio_amd import model as M

pytestmark = pytest.mark.gpu


def _run(e, n, seed, prec, tail):
    os.environ["MCRT_TAIL_LANES"] = str(tail)
    try:
        e.zero_tallies()
        e.enable_trace(n)
        e.run_packets(n, seed=seed, precision=prec, first_packet=0)
        tr = e.get_trace(n)
        tal = e.tallies()
        tal["stats"] = e.run_stats()
        e.enable_trace(0)
    finally:
        del os.environ["MCRT_TAIL_LANES"]
    return tr, tal


def _compare(e, n, seed, prec, what):
    # few lanes and short rounds, so that the run hands out its last packet
    # after a few rounds and goes on for many more
    e.set_launch(lanes=1024, queue_per_lane=16, events_per_round=8)
    try:
        ta, a = _run(e, n, seed, prec, 1 << 30)
        tb, b = _run(e, n, seed, prec, 0)
    finally:
        e.set_launch()
    assert a["stats"]["tail_rounds"] > 0 and b["stats"]["tail_rounds"] == 0, what
    same = np.all(ta == tb, axis=1)
    agree = float(same.mean())
    print(f"{what}: {a['stats']['tail_rounds']:.0f} tail rounds, per-packet agreement {agree:.5f}, "
          f"interactions {ta[:, 6].sum():.0f}", flush=True)
    if prec == 0:
        assert agree == 1.0, what
    else:
        assert agree >= 0.999, what
    if agree == 1.0:
        for k in ("tot", "escaped", "killed", "segments", "nscat", "hit", "htot", "errors", "resets"):
            assert a["counters"][k] == b["counters"][k], (what, k)
        np.testing.assert_array_equal(a["nabs"], b["nabs"])
        np.testing.assert_array_equal(a["dbig"], b["dbig"])
        np.testing.assert_allclose(a["dtauabs"], b["dtauabs"], rtol=1e-9, atol=1e-300)
        for name in ("si", "nums", "ti", "numt", "image_b_i"):
            x, y = a[name].astype(np.float64), b[name].astype(np.float64)
            assert np.abs(x - y).sum() <= 1e-4 * np.abs(y).sum() + 1e-30, (what, name)
    assert a["counters"]["errors"] == 0 and b["counters"]["errors"] == 0


@pytest.mark.parametrize("prec", [0, 1])
def test_tail_kernel_equals_transport_kernel_3d(prec):
    from src3_mpi_cfitsio_amd.engine import Engine
    m = M.small("tail3d", nr=20, nt=12, nphi=8, clump_sigma=0.5, pl=0.5, mdisk=0.05)
    with Engine(m) as e:
        e.disk(20000, seed=3, precision=1)                  # warm dust: re-emission in thick cells
        _compare(e, 4000, 31, prec, f"small 3-D, precision {prec}")


def test_tail_kernel_on_the_200x180x64_grid():
    from src3_mpi_cfitsio_amd.engine import Engine
    m = M.config("c1")
    with Engine(m) as e:
        for k in range(2):
            e.disk(500_000, seed=60 + k, precision=1)
        _compare(e, 20000, 61, 1, "c1 mixed")
        _compare(e, 5000, 62, 0, "c1 fp64")
