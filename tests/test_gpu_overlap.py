"""The peel-off kernels of each round on a second stream, overlapping the
next round's transport, against the same runs with both on one stream
(MCRT_OVERLAP=0), and on a caller's stream (mcrt_set_stream).

Round r writes peel-off queue r % 2 and its peel kernels read it while round
r + 1 fills the other queue; round r + 2 waits for them before it reuses
queue r % 2, and the run waits for the last ones.  Every walk is the same
in both schedules, so:
- every packet and walk counter, nabs, dbig and killed are equal;
- dtauabs (FP64 atomics) within 1e-9, images and SEDs within the rounding
  of their FP32 sums in another order (per bin 2e-5 + 8 sqrt(N) 2^-24 of
  the bin, of its I for Q, U, V, plus 1e-6 of the largest bin);
- with many short rounds (small queues), in both precisions.  A queue read
  after it was refilled, or a run that returned before its last peel
  kernels, changes whole bins.
"""
import os

import numpy as np
import pytest

from src3_mpi_cfitsio_amd import model as M

pytestmark = pytest.mark.gpu

NAMES = ("image_b_i", "image_b_q", "image_b_u", "image_b_v", "ti", "tq", "tu", "numt", "si", "nums")
SCALE_OF = {"image_b_q": "image_b_i", "image_b_u": "image_b_i", "image_b_v": "image_b_i", "tq": "ti", "tu": "ti"}


def _run(e, n, seed, prec, overlap):
    # every round in k_transport: where k_tail would take over depends on
    # when each round's queue fills, and in FP32 its scattering trials may
    # round differently (tests/test_gpu_tail.py compares the two kernels)
    os.environ["MCRT_OVERLAP"] = "1" if overlap else "0"
    os.environ["MCRT_TAIL_LANES"] = "0"
    try:
        e.zero_tallies()
        e.run_packets(n, seed=seed, precision=prec, first_packet=0)
        tal = e.tallies()
        tal["stats"] = e.run_stats()
    finally:
        del os.environ["MCRT_OVERLAP"]
        del os.environ["MCRT_TAIL_LANES"]
    return tal


def _same(a, b, what):
    for k in ("tot", "escaped", "killed", "segments", "peel_segments", "peel_events", "nscat", "hit", "htot",
              "errors", "resets"):
        assert a["counters"][k] == b["counters"][k], (what, k, a["counters"][k], b["counters"][k])
    assert a["counters"]["errors"] == 0, what
    for k in ("nabs", "dbig", "killed"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"{what}: {k}")
    np.testing.assert_allclose(a["dtauabs"], b["dtauabs"], rtol=1e-9, atol=1e-300, err_msg=what)
    u = 2.0 ** -24
    for name in NAMES:
        x, y = a[name].astype(np.float64), b[name].astype(np.float64)
        sc = np.abs(b[SCALE_OF.get(name, name)].astype(np.float64))
        cnt = b["numt"] if name in ("ti", "tq", "tu", "numt") else (b["nums"] if name in ("si", "nums") else None)
        nn = cnt.astype(np.float64) if cnt is not None else 1.0
        tol = (2e-5 + 8.0 * u * np.sqrt(np.maximum(nn, 1.0))) * sc + 1e-6 * max(sc.max(), 1e-30)
        bad = np.abs(x - y) > tol
        assert not bad.any(), (what, name, int(bad.sum()), float((np.abs(x - y) / np.maximum(sc, 1e-30)).max()))


@pytest.mark.parametrize("prec", [0, 1])
def test_overlapped_rounds_equal_serial_rounds(prec):
    from src3_mpi_cfitsio_amd.engine import Engine
    m = M.small("overlap3d", nr=20, nt=12, nphi=8, clump_sigma=0.5, pl=0.5, mdisk=0.05)
    with Engine(m) as e:
        e.disk(20000, seed=3, precision=1)                  # warm dust: re-emission paths too
        e.set_launch(lanes=2048, queue_per_lane=16, events_per_round=8)
        try:
            a = _run(e, 20000, 41, prec, overlap=True)
            b = _run(e, 20000, 41, prec, overlap=False)
        finally:
            e.set_launch()
        assert a["stats"]["rounds"] > 20 and a["stats"]["peel_launches"] == a["stats"]["rounds"]
        print(f"precision {prec}: {a['stats']['rounds']:.0f} rounds, peel events {a['counters']['peel_events']:.0f}",
              flush=True)
        _same(a, b, f"overlapped vs serial, precision {prec}")


def test_overlapped_rounds_on_a_caller_stream():
    import ctypes
    from src3_mpi_cfitsio_amd import _lib
    from src3_mpi_cfitsio_amd.engine import Engine
    m = M.small("overlap3d", nr=20, nt=12, nphi=8, clump_sigma=0.5, pl=0.5, mdisk=0.05)
    # a stream of the HIP runtime libmcrt uses (torch carries its own copy),
    # created at the default priority
    _lib.load()
    hip = ctypes.CDLL("libamdhip64.so.7")
    s = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(s)) == 0
    try:
        with Engine(m) as e0, Engine(m, stream=s.value) as e1:
            tdust = e0.disk(20000, seed=3, precision=1)["tdave"]    # warm dust, the same in both
            tal = []
            for e in (e0, e1):
                e.set_tdust(tdust)
                e.set_launch(lanes=2048, queue_per_lane=16, events_per_round=8)
                try:
                    tal.append(_run(e, 20000, 43, 1, overlap=True))
                finally:
                    e.set_launch()
            _same(tal[1], tal[0], "caller stream vs own stream")
    finally:
        hip.hipStreamDestroy(s)
