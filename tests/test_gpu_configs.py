"""GPU tests on the grids of every benchmark configuration, at packet counts
the oracle and the test budget allow.

- 100x90 two-dimensional model, heated by two Lucy iterations: FP64
  per-packet traces against the oracle on the same counters and dust
  temperatures, and the peeled I/Q/U images and peeled SEDs pixel by pixel.
- 200x180x64 clumpy polarized model, heated: the same against the oracle;
  then the mixed path (FP32 ray walker of the peel-off) against the FP64
  path on the same counters in 16 batches: peeled SEDs within 1%, image
  totals and super-pixels within 3 sigma of the FP64 batch scatter.
- the 512-model batch grid: a sample of models against single-model runs.
- 150x120x32 image-gradient model: adjoint in both precisions, its forward
  loss against an ordinary run's image.
- 400x360x128 optically thick model: packet conservation, zero walk
  errors, mixed against FP64 on the same counters.
"""
import os

import numpy as np
import pytest

from src3_mpi_cfitsio_amd import model as M
from gpu_util import rel_err

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def _heat(e, n, iters, seed, precision=1):
    t = None
    for k in range(iters):
        tal = e.disk(n, seed=seed + k, precision=precision)
        t = tal["tdave"]
    return t


def _pixelwise(a, b, agree, what):
    """Tallies of identical packets differ only by the float summation
    order; packets whose paths differ (agreement < 1) may move their own
    contributions, allowed as that share of the total."""
    a, b = a.astype(np.float64), b.astype(np.float64)
    tot = np.abs(b).sum()
    if tot == 0.0:
        assert np.abs(a).sum() == 0.0, what
        return
    d = np.abs(a - b)
    if agree == 1.0:
        tol = 2e-4 * np.abs(b) + 1e-6 * np.abs(b).max()
        bad = d > tol
        assert not bad.any(), f"{what}: {bad.sum()} bins off, worst {d[bad].max():.3g} of {np.abs(b).max():.3g}"
    else:
        assert d.sum() <= (4.0 * (1.0 - agree) + 1e-4) * tot, f"{what}: L1 {d.sum() / tot:.3g}"


def _compare_with_oracle(e, m, tdust, n, seed, first, oracle_lib, label):
    e.zero_tallies()
    e.enable_trace(n)
    e.run_packets(n, seed=seed, precision=0, first_packet=first)
    gt = e.get_trace(n)
    tg = e.tallies()
    e.enable_trace(0)
    print(f"{label}: oracle on {n} packets, {THREADS} threads", flush=True)
    o = oracle_lib.run_packets(m, n, seed=seed, first_packet=first, tdust=tdust, nthreads=THREADS, trace=True)
    same = np.all(rel_err(o["trace"], gt) <= 1e-9, axis=1)
    agree = float(same.mean())
    print(f"{label}: per-packet agreement {agree:.5f}, peel events gpu {tg['counters']['peel_events']:.0f} "
          f"oracle {o['counters']['peel_events']:.0f}", flush=True)
    assert agree >= 0.995
    assert tg["counters"]["errors"] == 0 and o["counters"]["errors"] == 0
    for name in ("image_b_i", "image_b_q", "image_b_u", "ti", "tq", "tu", "numt", "si", "nums"):
        _pixelwise(tg[name], o[name], agree, f"{label} {name}")
    # peel_segments is left out: the transport kernel queues no peel-off
    # record whose walks all end inside their first segment (tau > taumax
    # before any wall), while the oracle walks and counts that segment
    if agree == 1.0:
        for k in ("escaped", "killed", "segments", "peel_events", "tot", "htot"):
            assert tg["counters"][k] == o["counters"][k], k
        np.testing.assert_array_equal(tg["nabs"], o["nabs"])
    return tg


def test_c0_grid_heated_matches_oracle_pixelwise(oracle_lib):
    from src3_mpi_cfitsio_amd.engine import Engine
    m = M.config("c0")
    with Engine(m) as e:
        tdust = _heat(e, 200_000, 2, seed=900)
        assert tdust.max() > 100.0 / 11605.0
        tg = _compare_with_oracle(e, m, tdust, 3000, seed=901, first=10 ** 9, oracle_lib=oracle_lib, label="c0")
        assert tg["image_b_i"].sum() > 0 and tg["counters"]["peel_events"] > 3000


def test_c1_grid_heated_matches_oracle_and_mixed_matches_fp64(oracle_lib):
    from src3_mpi_cfitsio_amd.engine import Engine
    m = M.config("c1")
    with Engine(m) as e:
        tdust = _heat(e, 1_000_000, 2, seed=910)
        tg = _compare_with_oracle(e, m, tdust, 20_000, seed=911, first=10 ** 9, oracle_lib=oracle_lib, label="c1")
        assert np.abs(tg["image_b_q"]).sum() > 0          # polarized light reaches the images
        _mixed_against_fp64_16_batches(e, m, tdust)


def _mixed_against_fp64_16_batches(e, m, tdust, nbatch=16, nper=12_500):
    """Mixed (FP32 ray walker, FP32 Stokes) against FP64 on the same RNG
    counters, in 16 batches of packets so that the Monte Carlo noise of every
    image and SED bin is estimated from 16 independent samples.

    The two precisions follow identical paths for almost every packet; a
    packet whose FP32 optical depth ends a flight in another cell continues
    on an equally likely but different path.  So the images differ by a
    small share of the packets, and that difference must be noise-sized:
    - the peeled SED (ti, tq, tu) within 1% L1 of the FP64 one, and the
      walks reaching the observer within 0.5%;
    - per band and direction, the image's I/Q/U totals within 3 sigma of
      the FP64 batch scatter, and within 1% of the FP64 total I;
    - pixel by pixel on 16x16 super-pixels: |mixed - FP64| <= 3 sigma + 1e-3
      of the direction's total, with sigma the FP64 batch scatter of that
      super-pixel.  A swapped ix/iy, a pixel offset or a Q/U sign error
      moves bright super-pixels by many sigma.
    The full-resolution pixelwise L1 is printed as the stated image error."""
    names = ("image_b_i", "image_b_q", "image_b_u")
    sums = {p: {nm: 0.0 for nm in names + ("ti", "tq", "tu", "numt")} for p in (0, 1)}
    sp = {p: {nm: [] for nm in names} for p in (0, 1)}
    nsub = 16
    for bt in range(nbatch):
        for prec in (0, 1):
            e.set_tdust(tdust)
            e.zero_tallies()
            e.run_packets(nper, seed=912, precision=prec, first_packet=2 * 10 ** 9 + bt * nper)
            t = e.tallies()
            if prec == 1:
                assert t["counters"]["errors"] == 0
            for nm in names:
                img = t[nm].astype(np.float64)                          # [nband][nx][nx][npeel]
                sums[prec][nm] = sums[prec][nm] + img
                nb, nx = img.shape[0], img.shape[1]
                sp[prec][nm].append(img.reshape(nb, nsub, nx // nsub, nsub, nx // nsub, -1).sum(axis=(2, 4)))
            for nm in ("ti", "tq", "tu"):
                sums[prec][nm] = sums[prec][nm] + t[nm][0, 0].astype(np.float64)
            sums[prec]["numt"] += float(t["numt"].sum())
    a, b = sums[0], sums[1]
    ref = np.abs(a["ti"]).sum()
    for nm in ("ti", "tq", "tu"):
        l1 = np.abs(b[nm] - a[nm]).sum() / ref
        print(f"c1 mixed vs fp64: peeled SED {nm} L1 {l1:.2e}", flush=True)
        assert l1 < 0.01, nm
    # walks reaching the observer: FP32 optical depths end walks near the
    # tau = 25 cut differently, and a few packets take other paths
    np.testing.assert_allclose(b["numt"], a["numt"], rtol=5e-3)
    worst = 0.0
    for nm in names:
        s64 = np.array(sp[0][nm])                                       # [batch][band][sub][sub][k]
        sig_sp = s64.std(axis=0, ddof=1) * np.sqrt(nbatch)              # noise of the 16-batch sum
        tot64 = s64.sum(axis=(2, 3))                                    # [batch][band][k]
        sig_tot = tot64.std(axis=0, ddof=1) * np.sqrt(nbatch)
        d_sp = np.array(sp[1][nm]).sum(axis=0) - s64.sum(axis=0)
        for band in range(m.nband):
            for k in range(m.npeel):
                itot = np.abs(a["image_b_i"][band, :, :, k]).sum()
                if itot == 0.0:
                    continue
                x, y = b[nm][band, :, :, k], a[nm][band, :, :, k]
                l1 = np.abs(x - y).sum() / itot
                worst = max(worst, l1)
                dt = x.sum() - y.sum()
                assert abs(dt) <= 3.0 * sig_tot[band, k] + 1e-6 * itot, (nm, band, k, dt, sig_tot[band, k])
                if nm == "image_b_i":
                    assert abs(dt) <= 0.01 * itot, (nm, band, k, dt / itot)
                sg = sig_sp[band, :, :, k]
                # a super-pixel that no FP64 batch reached has no scatter of
                # its own: take the image's median super-pixel scatter
                floor = np.median(sg[sg > 0]) if np.any(sg > 0) else 0.0
                bad = np.abs(d_sp[band, :, :, k]) > 3.0 * np.maximum(sg, floor) + 1e-3 * itot
                assert not bad.any(), (nm, band, k, int(bad.sum()))
    print(f"c1 mixed vs fp64: worst full-resolution pixelwise image L1 {worst:.3e} "
          f"({nbatch * nper} packets)", flush=True)


def test_c2_grid_sample_matches_single_model_runs():
    from src3_mpi_cfitsio_amd.engine import BatchEngine, Engine
    b = M.batch_models(512, seed=1002, device=True)
    n, seed = 2000, 4343
    with BatchEngine(b) as be:
        be.set_batch_launch(graph=1)
        be.run_batch(n, seed, precision=1)
        got, work, cnt = be.batch_tallies()
    assert cnt["errors"] == 0
    for k in (0, 131, 402, 511):
        with Engine(b.models[k]) as e:
            e.zero_tallies()
            e.run_packets(n, seed=seed + k, precision=1)
            ref = e.tallies()
        print(f"c2 model {k}: {b.models[k].nrg - 1}x{b.models[k].ntg - 1}", flush=True)
        assert np.array_equal(got[k]["nabs"], ref["nabs"]), k
        assert np.array_equal(got[k]["nums"], ref["nums"]), k
        np.testing.assert_allclose(got[k]["dtauabs"], ref["dtauabs"], rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(got[k]["si"], ref["si"], rtol=1e-5, atol=1e-6)


def test_c3_grid_adjoint_precisions_and_forward_loss():
    from src3_mpi_cfitsio_amd.engine import Engine
    m = M.config("c3")
    n = 200_000
    W = np.random.default_rng(1003).uniform(0.0, 1.0, size=(m.nband, m.nx, m.nx, m.npeel)).astype(np.float32)
    with Engine(m) as e:
        _heat(e, 200_000, 1, seed=930, precision=0)
        l64, g64 = e.adjoint_image(W, n, seed=931, precision=0, grad=True)
        l32, g32 = e.adjoint_image(W, n, seed=931, precision=1, grad=True)
        e.zero_tallies()
        e.run_packets(n, seed=931, precision=0)
        img = e.tallies()["image_b_i"].astype(np.float64)
    fwd = float(np.sum(W.astype(np.float64) * img))
    print(f"c3: loss fp64 {l64:.6g} mixed {l32:.6g} image {fwd:.6g}", flush=True)
    assert abs(l64 - fwd) <= 1e-3 * abs(fwd)
    assert abs(l32 - l64) <= 1e-3 * abs(l64)
    assert np.all(np.isfinite(g32)) and np.abs(g64).sum() > 0
    assert np.linalg.norm(g32 - g64) <= 0.02 * np.linalg.norm(g64)


def test_c4_grid_conservation_and_mixed_vs_fp64():
    from src3_mpi_cfitsio_amd.engine import Engine
    m = M.config("c4")
    n = 100_000
    with Engine(m) as e:
        res = {}
        for prec in (0, 1):
            e.set_tdust(m.tdust_initial())
            e.zero_tallies()
            e.run_packets(n, seed=940, precision=prec)
            res[prec] = e.tallies()
    for prec, t in res.items():
        c = t["counters"]
        assert c["errors"] == 0, prec
        assert c["escaped"] + c["killed"] == n, prec
        assert t["nums"][0].sum() == c["escaped"]
        assert t["nabs"].sum() == t["dbig"].sum()
    a, b = res[0], res[1]
    for k in ("segments", "peel_segments", "escaped"):
        assert abs(a["counters"][k] - b["counters"][k]) <= 2e-3 * a["counters"][k], k
    assert abs(a["dtauabs"].sum() - b["dtauabs"].sum()) <= 2e-3 * a["dtauabs"].sum()
    assert np.abs(a["si"][0].astype(np.float64) - b["si"][0]).sum() <= 5e-3 * a["si"][0].sum()
