"""Benchmark: photon packets/s of the mixed-precision engine on the 3-D
polarized clumpy disk+envelope workload (200x180x64 cells), against the FP64
engine on the same RNG counters, with the SED / temperature error of the
mixed path stated, the roofline of the packet kernel, and the CPU oracle
timed on the host cores.

A step is one Lucy iteration: zero the tallies, run the step's packets
(emission, cell walk, polarized scattering, 10-direction peel-off into
256x256 images in 3 bands and SEDs, FP64 absorbed-energy tally), then the
FP64 temperature inversion that the next step re-emits with.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--packets N] [--budget S]

The default step is the workload's full Lucy iteration (10^8 packets on the
200x180x64 model). At that size the bulk of the packets sets the step time;
the serial chains of the few longest packets (up to ~10^6 interactions, a
few seconds) overlap with it except at the very end.  The FP64 comparison
runs the first timed step again in FP64, from the same starting dust
temperature on the same RNG counters, timed the same way; the SED and
temperature errors compare the two steps' results.

The run keeps to a wall-clock budget (--budget, default 480 s): W (at most
4) short warm-up Lucy iterations heat the dust, the first timed step, its
FP64 twin and the CPU leg always run, and further timed steps (up to K)
follow while the next one still fits.  The JSON line gives the steps and
warm-ups actually run next to the requested ones.

python bench.py --copy-probe BYTES only copies BYTES of device memory with
torch (one copy kernel): the known-size kernel that scales rocprofv3's
FETCH_SIZE / WRITE_SIZE readings to bytes.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "photon packets/s, mixed-prec vs FP64 GPU path at SED/T err ≤1% of FP64 oracle"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def kernel_bytes(c, st, rec_bytes, lane_bytes, nopac):
    """Algorithmic HBM bytes of the two kinds of kernel of a run.

    transport (k_transport, k_tail): per face crossing the FP32 density
    (4 B) + species (1 B, read only with more than one opacity table) +
    FP64 dtauabs read-modify-write (16 B); per escaped packet its SED row,
    3 components x 10 floats read-modify-write (240 B); per queued peel-off
    record its write (rec_bytes); per parked lane-round the lane state
    written and read back (2 x lane_bytes).
    peel (k_peel_start, k_peel_walk, k_peel_hits): per face crossing the
    density (+ species); per walk that reaches the observer 3 peeled-SED
    components x 9 floats (216 B) + image I,Q,U,V, origin (40 B)
    read-modify-write; per record one read (rec_bytes).  The ray states and
    hits the three kernels pass to each other are not algorithmic bytes:
    they are in the measured traffic."""
    b_rho = 4.0 + (1.0 if nopac > 1 else 0.0)
    transport = (c["segments"] * (b_rho + 16.0) + c["escaped"] * 240.0 + st["records"] * rec_bytes
                 + st["parked"] * 2.0 * lane_bytes)
    peel = c["peel_segments"] * b_rho + c["peel_events"] * 256.0 + st["records"] * rec_bytes
    return transport, peel


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_transport_k_peel_split_c1_hbm_traffic.md")
# the kernels of a mixed-precision run, named as in the traffic summary
TRANSPORT_KERNELS = "k_transport+k_tail<float>"
PEEL_KERNELS = "k_peel_start+k_peel_walk+k_peel_hits<float>"


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE per dispatch over the
    timed step, each scaled by what the counter reads for a copy kernel of
    known size): its line `traffic_bytes_per_launch <kernel> <bytes>`.
    None when the summary holds no figure for that kernel."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                parts = line.strip().strip("`").split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                    return float(parts[2])
    except OSError:
        return None
    return None


def sed_error(a, b):
    """Flux-weighted relative L1 of the total SED over (frequency, inclination)
    bins holding >= 100 packets in the reference run b."""
    fa, fb = a["ti"][0, 0].astype(np.float64), b["ti"][0, 0].astype(np.float64)
    nb = b["numt"][0, 0]
    k = nb >= 100
    return float(np.sum(np.abs(fa[k] - fb[k])) / max(np.sum(np.abs(fb[k])), 1e-300))


def t_error(ta, tb, mass):
    k = (tb > 0.1 / 11605.0 * 1.0001) & (mass > 0)
    rel = np.abs(ta[k] - tb[k]) / tb[k]
    w = mass[k] / mass[k].sum()
    return float(np.sum(w * rel)), float(np.percentile(rel, 99))


def cpu_baseline(m, tdust, seconds, threads):
    """Packets/s of the FP64 C oracle on `threads` host threads: runs of
    growing size until one lasts at least 60% of `seconds` (the work per
    packet is heavy-tailed, so a short probe underestimates it)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    n, first = 2000, 10 ** 12
    for _ in range(6):
        t0 = time.perf_counter()
        oracle_py.run_packets(m, n, seed=777, first_packet=first, tdust=tdust, nthreads=threads)
        dt = time.perf_counter() - t0
        first += n
        if dt >= 0.6 * seconds or n >= 20_000_000:
            break
        n = int(min(20_000_000, max(2 * n, n * 1.2 * seconds / max(dt, 1e-3))))
    return n / dt, n, dt


def run_batch_workload(args):
    """The 512-model grid of axisymmetric disks (uneven grids, SEDs only) in
    one persistent launch: two warm-up Lucy iterations, the autotuner on a
    sample, then a timed run of --packets per model (default 10^6) in mixed
    precision, and FP64 on a quarter of that for the ratio and the SED
    error on identical counters."""
    from src3_mpi_cfitsio_amd import model as M
    from src3_mpi_cfitsio_amd.engine import MCRT_FP64, MCRT_MIXED, BatchEngine
    t0 = time.perf_counter()
    b = M.batch_models(512, seed=1002, device=True)
    log(f"512 models, {sum(x.ncell for x in b.models)} cells, built in {time.perf_counter()-t0:.1f}s "
        f"(cell arrays on the GPU)")
    e = BatchEngine(b)
    for k in range(2):
        e.run_batch(100_000, seed=500 + k, precision=MCRT_MIXED)
        e.batch_dusttemp(update_tdust=True)
        log(f"warm-up Lucy iteration {k}: {e.last_run_ms()/1e3:.2f}s")
    best, table = e.autotune(packets_per_model=2000, seed=3, precision=MCRT_MIXED)
    log(f"autotuner: {len(table)} shapes, best (block, waves, regroup, combine) = {best}")
    e.set_batch_launch(graph=1)
    n = args.packets or 1_000_000
    times = []
    for s in range(args.warmup + args.steps):
        e.synchronize()
        t1 = time.perf_counter()
        e.run_batch(n, seed=900 + s, precision=MCRT_MIXED)
        e.synchronize()
        dt = time.perf_counter() - t1
        log(f"step {s}: {512*n} packets in {dt:.2f}s (kernel {e.last_run_ms()/1e3:.2f}s)")
        if s >= args.warmup:
            times.append(dt)
    ms = 1e3 * float(np.mean(times))
    value = 512 * n / (ms / 1e3)
    extra = {}
    if not args.no_fp64:
        nc = max(1, n // 4)
        res = {}
        for prec in (MCRT_FP64, MCRT_MIXED):
            t1 = time.perf_counter()
            e.run_batch(nc, seed=9999, precision=prec)
            e.synchronize()
            dt = time.perf_counter() - t1
            tal = e.batch_tallies()[0]
            res[prec] = (dt, np.sum([t["si"][0, 0].astype(np.float64) for t in tal], axis=0),
                         np.concatenate([t["si"][0, 0].astype(np.float64).ravel() for t in tal]))
            log(f"{'fp64' if prec == MCRT_FP64 else 'mixed'} comparison: {512*nc} packets in {dt:.2f}s")
        a, f = res[MCRT_MIXED][2], res[MCRT_FP64][2]
        extra = dict(fp64_value=512 * nc / res[MCRT_FP64][0],
                     mixed_over_fp64=res[MCRT_FP64][0] / res[MCRT_MIXED][0],
                     sed_err_vs_fp64=float(np.sum(np.abs(a - f)) / max(np.sum(np.abs(f)), 1e-300)),
                     fp64_comparison_packets=512 * nc)
    e.close()
    out = {"metric": "photon packets/s, batched model grid (one persistent launch)", "value": value, "unit": "packets/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
           "config": {"workload": "c2: 512 axisymmetric disk+envelope models, nr 50-400, ntheta 40-200, SEDs at 10 "
                                  "inclinations, one Lucy iteration per step", "packets_per_model": n,
                      "launch": {"block": best[0], "waves_per_simd": best[1], "regroup": best[2], "combine": best[3]},
                      "parallelism": "1 GPU"}}
    out.update(extra)
    print(json.dumps(out), flush=True)


def run_adjoint_workload(args):
    """Image gradient by path replay on the 150x120x32 model: --packets
    (default 10^7) in mixed precision after two FP64 Lucy iterations, and
    the FP64 gradient on the same counters for the comparison."""
    from src3_mpi_cfitsio_amd import model as M
    from src3_mpi_cfitsio_amd.engine import MCRT_FP64, MCRT_MIXED, Engine
    m = M.config("c3")
    e = Engine(m)
    for k in range(2):
        e.zero_tallies()
        e.run_packets(2_000_000, seed=300 + k, precision=MCRT_FP64)
        e.dusttemp(2_000_000, update_tdust=True)
    log("model heated")
    W = np.random.default_rng(1003).uniform(0.0, 1.0, size=(m.nband, m.nx, m.nx, m.npeel)).astype(np.float32)
    n = args.packets or 10_000_000
    t0 = time.perf_counter()
    l32, g32 = e.adjoint_image(W, n, seed=21, precision=MCRT_MIXED, grad=True)
    dt = time.perf_counter() - t0
    log(f"mixed adjoint: {n} packets in {dt:.2f}s")
    extra = {}
    if not args.no_fp64:
        t0 = time.perf_counter()
        l64, g64 = e.adjoint_image(W, n, seed=21, precision=MCRT_FP64, grad=True)
        dt64 = time.perf_counter() - t0
        log(f"fp64 adjoint: {n} packets in {dt64:.2f}s")
        extra = dict(fp64_value=n / dt64, mixed_over_fp64=dt64 / dt, loss_rel_err_vs_fp64=abs(l32 - l64) / abs(l64),
                     grad_rel_l2_vs_fp64=float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64)))
    e.close()
    out = {"metric": "photon packets/s, path-replay image gradient (forward + replay)", "value": n / dt,
           "unit": "packets/s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": dt * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
           "config": {"workload": "c3: 150x120x32 cells, 128x128 images at 2.2 and 3.6 um, one inclination, "
                                  "d(sum W image)/d(rho) per cell", "packets": n, "parallelism": "1 GPU"}}
    out.update(extra)
    print(json.dumps(out), flush=True)


EMISS_OPS_PER_EVAL = 38     # FP64 operations per integrand evaluation, counted from the source (exp, / as one)
FP64_PEAK_TFLOPS = 78.6


def run_setemiss_workload(args):
    """setemiss (the Planck-mean emissivity tables, qsimp per interval) for
    --species dust species (default 187: a 384 MB kapint table) at nT = 500
    temperatures and nnu = 513 frequency intervals, against the C oracle on
    one host core over the same species."""
    from src3_mpi_cfitsio_amd import model as M
    from src3_mpi_cfitsio_amd.engine import setemiss
    nsp = args.species
    rng = np.random.default_rng(1004)
    sp = [M.synthetic_dust(nlam=100, kappa0=rng.uniform(50, 400), slope=rng.uniform(0.5, 2.0),
                           albedo=(rng.uniform(0.3, 0.7), rng.uniform(0.0, 0.1))) for _ in range(nsp)]
    nud, kap, alb = (np.array([d[k] for d in sp]) for k in ("nudust", "kappad", "adust"))
    em = M.emission_tables(sp[0], nT=500, nnu=513, nodes=2)
    T, nu = em["Tint"], em["nuint"]
    smooth = np.zeros(nsp, np.int32)
    smooth[nsp // 2:] = 1                                  # the species above ndg are smoothed
    setemiss(nud[:2], kap[:2], alb[:2], T, nu, smooth=smooth[:2])      # load the kernels
    walls, kms = [], []
    for s in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        r = setemiss(nud, kap, alb, T, nu, smooth=smooth)
        dt = time.perf_counter() - t0
        log(f"step {s}: {nsp} species in {dt*1e3:.1f} ms (kernels {r['ms']:.2f} ms), "
            f"{r['nevals']/(nsp*500*513):.2f} evaluations per interval, {r['nfail']} qsimp non-converged")
        if s >= args.warmup:
            walls.append(dt)
            kms.append(r["ms"])
    kms_mean = float(np.mean(kms))
    ops = r["nevals"] * EMISS_OPS_PER_EVAL
    achieved = ops / (kms_mean / 1e3) / 1e12
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py
        t0 = time.perf_counter()
        oracle_py.setemiss(nud, kap, alb, T, nu, smooth=smooth)
        dt = time.perf_counter() - t0
        cpu = {"value": nsp / dt, "unit": "species tables/s", "cores": 1, "kind": "port",
               "sample": f"FP64 C oracle (qsimp per interval, as dust_5.f90:507-699), all {nsp} species in "
                         f"{dt:.1f}s on one host thread"}
        log(f"cpu baseline: {nsp} species in {dt:.2f}s")
    out = {"metric": "setemiss species tables/s (500 temperatures x 513 intervals, qsimp per interval)",
           "value": nsp / (kms_mean / 1e3), "unit": "species tables/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": kms_mean, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"setemiss: {nsp} dust species x 500 temperatures x 513 intervals, "
                                  f"{nsp - nsp // 2} smoothed", "wall_ms_per_step_with_copies": 1e3 * float(np.mean(walls)),
                      "parallelism": "1 GPU"},
           "roofline": {"bound": "valu-fp64", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                        "kernel": "k_emiss_seg + k_emiss_scan", "evaluations": r["nevals"],
                        "ops_per_evaluation": EMISS_OPS_PER_EVAL},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def cpu_leg(m, tdust, seconds, threads_list, config):
    """The FP64 C oracle on the host cores at each thread count of
    threads_list (one packet range per thread, as the reference splits its
    packets over MPI ranks), about `seconds` of work each."""
    runs = []
    for th in threads_list:
        rate, n, dt = cpu_baseline(m, tdust, seconds, th)
        runs.append(dict(threads=th, value=rate, packets=n, seconds=dt))
        log(f"cpu baseline: {rate:.4g} packets/s on {th} threads ({n} packets in {dt:.1f}s)")
    best = max(runs, key=lambda r: r["value"])
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return {"value": best["value"], "unit": "packets/s", "cores": best["threads"], "kind": "port",
            "sample": f"FP64 C oracle on the same {config} model and dust temperature, one packet range per "
                      f"host thread, about {seconds:.0f}s of work per thread count; the reference itself cannot "
                      f"be built here (its grid module includes mpif.h and the image has no MPI)",
            "runs": runs, "host_cpus": os.cpu_count(), "cpus_available": avail}


def run_single_workload(args, T0):
    """The headline workload: Lucy iterations of the 200x180x64 model at its
    full 10^8 packets per step, inside a wall-clock budget.

    Order: temperature warm-up (short Lucy iterations), the first timed
    step, the same step again in FP64 on the same counters and starting
    temperature, the CPU leg, then further timed steps while the budget
    allows one more (each estimated at 1.15x the slowest step so far).
    --steps / --warmup are upper limits; the JSON line reports the steps
    actually timed and warmed up."""
    from src3_mpi_cfitsio_amd import model as M
    from src3_mpi_cfitsio_amd.engine import MCRT_FP64, MCRT_MIXED, Engine

    def elapsed():
        return time.perf_counter() - T0

    t0 = time.perf_counter()
    m = M.config(args.config)
    if args.accfrac > 0.0:
        m = M.with_accretion(m, args.accfrac)
    if args.diffusion:
        m = M.with_diffusion(m)
    log(f"model {args.config}: {m.nrg-1}x{m.ntg-1}x{m.npg-1} cells built in {time.perf_counter()-t0:.1f}s")
    mass = m.rho * m.vcell
    eng = Engine(m)
    lucy_iter = [0]

    def setdiffus():
        # gridset calls setdiffus at the start of every Lucy iteration
        if args.diffusion:
            lucy_iter[0] += 1
            nd = eng.setdiffus(lucy_iter[0])
            log(f"diffusion region: {nd} cells (Lucy iteration {lucy_iter[0]})")

    # temperature warm-up: short Lucy iterations so the timed steps transport
    # packets through a heated disk (re-emission at the dust temperature)
    nw = 2_000_000
    nwarm = max(1, min(4, args.warmup)) if args.warmup > 0 else 1
    for k in range(nwarm):
        setdiffus()
        eng.zero_tallies()
        eng.run_packets(nw, seed=1000 + k, precision=MCRT_MIXED, first_packet=0)
        eng.synchronize()
        tw = eng.last_run_ms() / 1e3
        tcur = eng.dusttemp(nw, update_tdust=True)
        log(f"warm-up Lucy iteration {k}: {nw} packets in {tw:.2f}s ({nw/tw:.3g} packets/s), t={elapsed():.0f}s")
    N = args.packets or int(m.npackets)
    log(f"packets per step: {N}, budget {args.budget:.0f}s")

    seed = 20250101
    step_ms, kern_ms, counters, stats = [], [], [], []
    first = 0

    def step(t_in):
        setdiffus()
        eng.synchronize()
        ts = time.perf_counter()
        eng.zero_tallies()
        eng.run_packets(N, seed=seed, precision=MCRT_MIXED, first_packet=first)
        tal_s = eng.tallies()                      # post-reduction tallies to the host
        t_out = eng.dusttemp(N, update_tdust=True)
        eng.synchronize()
        dt = (time.perf_counter() - ts) * 1e3
        km = eng.last_run_ms()
        step_ms.append(dt)
        kern_ms.append(km)
        counters.append(tal_s["counters"])
        stats.append(eng.run_stats())
        sx = stats[-1]
        log(f"timed step {len(step_ms)}: {dt:.0f} ms, kernel {km:.0f} ms, {N/(dt/1e3):.4g} packets/s, "
            f"segments/packet {tal_s['counters']['segments']/N:.0f}, peel segments/packet "
            f"{tal_s['counters']['peel_segments']/N:.0f}, errors {tal_s['counters']['errors']:.0f} "
            f"(walk {sx['walk_errors']:.0f}, peel {sx['peel_errors']:.0f}, other {sx['other_errors']:.0f}), "
            f"t={elapsed():.0f}s")
        return tal_s, t_out

    # the first timed step, then its FP64 twin on the same counters
    t_start, first_step = tcur, first
    tal, tcur = step(t_start)
    first += N
    extra = {}
    if not args.no_fp64 and not args.diffusion:
        t_mixed = eng.dusttemp(N, mode=1, update_tdust=False)     # per-cell T of the mixed step's tallies
        log("fp64 comparison: the first timed step again in FP64 (the engine logs each round to stderr)")
        eng.set_tdust(t_start)
        verbose_before = os.environ.get("MCRT_VERBOSE")
        os.environ["MCRT_VERBOSE"] = "1"             # a progress line per round of this long run
        eng.synchronize()
        ts = time.perf_counter()
        eng.zero_tallies()
        eng.run_packets(N, seed=seed, precision=MCRT_FP64, first_packet=first_step)
        tal64 = eng.tallies()
        eng.dusttemp(N, update_tdust=False)
        eng.synchronize()
        dt64 = time.perf_counter() - ts
        if verbose_before is None:
            del os.environ["MCRT_VERBOSE"]
        else:
            os.environ["MCRT_VERBOSE"] = verbose_before
        t64 = eng.dusttemp(N, mode=1, update_tdust=False)
        log(f"fp64 comparison run: {N} packets in {dt64:.2f}s ({N/dt64:.4g} packets/s); "
            f"mixed step {step_ms[0]/1e3:.2f}s, t={elapsed():.0f}s")
        sed_e = sed_error(tal, tal64)
        t_mean, t_p99 = t_error(t_mixed, t64, mass)
        extra = dict(fp64_value=N / dt64, mixed_over_fp64=dt64 / (step_ms[0] / 1e3),
                     sed_err_vs_fp64=sed_e, t_err_vs_fp64_mass_weighted=t_mean, t_err_vs_fp64_p99=t_p99,
                     fp64_comparison_packets=N)
        del tal64
        eng.set_tdust(tcur)                          # continue the mixed Lucy sequence

    cpu = None
    if not args.no_cpu:
        cap = min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1))
        cpu = cpu_leg(m, t_start, args.cpu_seconds, sorted({cap, 31}), args.config)

    # further timed steps while one more fits in the budget
    while len(step_ms) < max(1, args.steps):
        est = 1.15 * max(step_ms) / 1e3
        if elapsed() + est > args.budget:
            log(f"budget: {elapsed():.0f}s used, next step ~{est:.0f}s would pass {args.budget:.0f}s; "
                f"{len(step_ms)} of {args.steps} requested steps timed")
            break
        tal, tcur = step(tcur)
        first += N
    eng.close()

    ms = float(np.mean(step_ms))
    value = N / (ms / 1e3)
    kmean = float(np.mean(kern_ms))
    c = {k: float(np.mean([cc[k] for cc in counters])) for k in counters[0]}
    st = {k: float(np.mean([ss[k] for ss in stats])) for k in stats[0]}
    b_tr, b_pe = kernel_bytes(c, st, st["record_bytes"], st["lane_bytes"], m.nopac)
    # each kind of kernel's roofline: algorithmic bytes over its summed
    # HIP-event time (every launch of the step bracketed on the run stream);
    # the one that takes longer is the step's "roofline"
    def roof(kname, kms, kB, nl):
        achieved = kB / (kms / 1e3) / 1e9
        return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(kname), "kernel": kname,
                "kernel_ms_per_step": kms, "launches_per_step": nl, "kernel_ms_per_launch": kms / max(nl, 1.0),
                "algorithmic_bytes_per_launch": kB / max(nl, 1.0)}
    r_tr = roof(TRANSPORT_KERNELS, st["transport_ms"], b_tr, st["rounds"])
    r_pe = roof(PEEL_KERNELS, st["peel_ms"], b_pe, st["peel_launches"])
    dom, other = (r_pe, r_tr) if st["peel_ms"] >= st["transport_ms"] else (r_tr, r_pe)
    log(f"kernels per step: transport {st['transport_ms']:.1f} ms ({b_tr/1e9:.2f} GB algorithmic), "
        f"peel {st['peel_ms']:.1f} ms ({b_pe/1e9:.2f} GB), rounds {st['rounds']:.0f}")
    out = {
        "metric": METRIC, "value": value, "unit": "packets/s", "n_gpus": 1, "steps": len(step_ms),
        "warmup": nwarm, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"{args.config}: 3-D polarized clumpy disk+Ulrich envelope, "
                               f"{m.nrg-1}x{m.ntg-1}x{m.npg-1} cells, 10 peel-off inclinations, "
                               f"{m.nx}x{m.nx}x4-Stokes images in {m.nband} bands, dustmat scattering matrix "
                               f"g={m.gdust[0, 0]:g} pl={m.pldust[0, 0]:g}, one Lucy iteration per step",
                   "packets_per_step": N, "precision": "mixed (FP32 opacities, optical depths, sampling, Stokes; "
                                                       "FP64 geometry and dtauabs)",
                   "steps_requested": args.steps, "warmup_requested": args.warmup,
                   "warmup_kind": f"{nwarm} Lucy iterations of {nw} packets (heats the dust; loads the kernels)",
                   "budget_s": args.budget, "wall_s": round(elapsed(), 1),
                   "parallelism": "1 GPU", **({"accfrac": args.accfrac} if args.accfrac > 0.0 else {}),
                   **({"diffusion": True} if args.diffusion else {})},
        "roofline": {**dom, "transport_ms_per_step": st["transport_ms"], "peel_ms_per_step": st["peel_ms"],
                     "run_ms_per_step": kmean, "tail_rounds": st.get("tail_rounds", 0.0)},
        "roofline_other_kernels": other,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


def main():
    T0 = time.perf_counter()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--packets", type=int, default=0, help="packets per step (default: the workload's own)")
    ap.add_argument("--budget", type=float, default=480.0,
                    help="wall-clock seconds the c1 run may take; further timed steps stop short of it")
    ap.add_argument("--config", default="c1")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--accfrac", type=float, default=0.0,
                    help="fraction of each step's packets emitted by disk accretion (c1/c0 workloads)")
    ap.add_argument("--diffusion", action="store_true",
                    help="c1/c0 workloads with the diffusion approximation of thick regions (DIFFUSION = YES); "
                         "no FP64 twin (the diffusion region is set per Lucy iteration)")
    ap.add_argument("--copy-probe", type=int, default=0, help="only copy this many bytes on the GPU")
    ap.add_argument("--workload", default="c1", choices=["c1", "c2", "c3", "setemiss"],
                    help="c1: the headline single-model metric; c2: the 512-model batch; c3: the image gradient; "
                         "setemiss: the emissivity tables")
    ap.add_argument("--species", type=int, default=187, help="dust species of the setemiss workload")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        x = torch.ones(args.copy_probe // 4, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        log(f"copied {x.numel() * 4} bytes")
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path would shard by packet ranges with one tally all-reduce per Lucy iteration)",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "c2":
        return run_batch_workload(args)
    if args.workload == "c3":
        return run_adjoint_workload(args)
    if args.workload == "setemiss":
        return run_setemiss_workload(args)
    return run_single_workload(args, T0)


if __name__ == "__main__":
    main()
