"""The mixed path's peel-off walks in three kernels (k_peel_start, k_peel_walk,
k_peel_hits) against the one-kernel k_peel on the same RNG counters.

Every walk does the same arithmetic in both (the reference's first crossing
from the record position, then the same ray steps and the same tally), so:
- the walk counters (peel segments, walks reaching the observer, numt) and
  every packet counter are equal;
- images and peeled SEDs are equal up to the order of their FP32 sums: per
  bin within (2e-5 + 8 sqrt(N) 2^-24) of the bin (of its I for Q, U, V; N
  the walks that reached it) plus 1e-6 of the largest bin;
- no walk error, on a clumpy polarized 3-D model and on the 200x180x64 grid,
  with one chunk per round and with many short rounds (small queues) split
  into several chunks (MCRT_PEEL_CHUNK).
"""
import os

import numpy as np
import pytest

from src3_mpi_cfitsio_amd import model as M

pytestmark = pytest.mark.gpu

NAMES = ("image_b_i", "image_b_q", "image_b_u", "image_b_v", "ti", "tq", "tu", "numt", "si", "nums")


def _run(e, n, seed, split, tdust=None):
    os.environ["MCRT_PEEL_SPLIT"] = "1" if split else "0"
    try:
        if tdust is not None:
            e.set_tdust(tdust)
        e.zero_tallies()
        e.run_packets(n, seed=seed, precision=1, first_packet=0)
        tal = e.tallies()
        tal["stats"] = e.run_stats()
    finally:
        del os.environ["MCRT_PEEL_SPLIT"]
    return tal


def _same(a, b, what):
    for k in ("escaped", "killed", "segments", "peel_segments", "peel_events", "tot", "errors", "resets"):
        assert a["counters"][k] == b["counters"][k], (what, k, a["counters"][k], b["counters"][k])
    assert a["counters"]["errors"] == 0, what
    # A bin's FP32 sums add in another order (the split path sums each
    # direction's peeled SED per block first).  A sum of N terms in FP32
    # rounds by about sqrt(N) u of its size (u = 2^-24); for Q, U and V,
    # whose contributions cancel, the rounding scales with the bin's I
    # (|Q| <= I for every contribution).  A bin or component mixed up moves
    # whole bins and fails by orders of magnitude.
    u = 2.0 ** -24
    scale_of = {"image_b_q": "image_b_i", "image_b_u": "image_b_i", "image_b_v": "image_b_i",
                "tq": "ti", "tu": "ti"}
    for name in NAMES:
        x, y = a[name].astype(np.float64), b[name].astype(np.float64)
        sc = np.abs(b[scale_of.get(name, name)].astype(np.float64))
        n = b["numt"].astype(np.float64) if name in ("ti", "tq", "tu", "numt") else 1.0
        tol = (2e-5 + 8.0 * u * np.sqrt(np.maximum(n, 1.0))) * sc + 1e-6 * max(sc.max(), 1e-30)
        bad = np.abs(x - y) > tol
        rel = np.abs(x - y) / np.maximum(sc, 1e-30)
        assert not bad.any(), (what, name, int(bad.sum()), float(rel.max()))
        print(f"{what}: {name} largest difference {float(rel.max()):.2e} of the bin", flush=True)


def test_split_equals_one_kernel_small_3d():
    from src3_mpi_cfitsio_amd.engine import Engine
    m = M.small("split3d", nr=20, nt=12, nphi=8, clump_sigma=0.5, pl=0.5)
    with Engine(m) as e:
        e.disk(20000, seed=5, precision=1)                   # warm dust: re-emission paths too
        a = _run(e, 30000, 11, split=True)
        b = _run(e, 30000, 11, split=False)
        _same(a, b, "small 3-D")
        # many short rounds, each in chunks of 5000 items
        e.set_launch(lanes=2048, queue_per_lane=16, events_per_round=8)
        os.environ["MCRT_PEEL_CHUNK"] = "5000"
        try:
            a = _run(e, 30000, 12, split=True)
        finally:
            del os.environ["MCRT_PEEL_CHUNK"]
        b = _run(e, 30000, 12, split=False)
        _same(a, b, "small 3-D, short rounds in chunks")
        assert a["stats"]["rounds"] > 10
        e.set_launch()


def test_split_equals_one_kernel_on_the_200x180x64_grid():
    from src3_mpi_cfitsio_amd.engine import Engine
    m = M.config("c1")
    with Engine(m) as e:
        for k in range(2):
            e.disk(500_000, seed=70 + k, precision=1)
        a = _run(e, 200_000, 77, split=True)
        b = _run(e, 200_000, 77, split=False)
        print(f"c1 split: peel {a['stats']['peel_ms']:.1f} ms, one kernel {b['stats']['peel_ms']:.1f} ms; "
              f"transport {a['stats']['transport_ms']:.1f} / {b['stats']['transport_ms']:.1f} ms", flush=True)
        _same(a, b, "c1")
        assert a["stats"]["peel_errors"] == 0 and a["stats"]["other_errors"] == 0
