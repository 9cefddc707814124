"""Host-side mirror of the reference's packet-loop interface over libmcrt.so.

The reference drives its packet path with Fortran subroutines on
module-global state (paths relative to src3_mpi_cfitsio/):

  disk(converge, iter, i1)          newdisktherm_5_fuv.f90:1   -> Engine.disk
  propagate / tauint / stokes /
  peeloff inside the packet loop    newdisktherm_5_fuv.f90:220 -> Engine.run_packets
  MPI_ALLREDUCE of every tally      newdisktherm_5_fuv.f90:524 -> Engine.tallies
  tfinal / dusttemp                 tfinal_5_test1.f90:1       -> Engine.dusttemp
  diskflux                          flux_test.f90:1            -> diskflux()

Every call goes through the C ABI (include/mcrt.h); errors raise MCRTError
with the library's message instead of the reference's `stop`.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional

import numpy as np

from . import _lib
from ._lib import MCRT_FP64, MCRT_MIXED, MCRT_T_PERCELL, MCRT_T_REFERENCE, ptr
from .model import LSUN, Batch, Model

COUNTERS = ["tot", "flux", "sflux", "aflux", "abflux", "scount", "nscat", "hit", "htot",
            "escaped", "killed", "segments", "peel_segments", "peel_events", "resets", "errors"]
RUN_STATS = ["rounds", "transport_ms", "peel_ms", "peel_launches", "records", "parked", "lanes", "events0",
             "record_bytes", "lane_bytes", "walk_errors", "peel_errors", "other_errors", "tail_rounds"]
SED_NAMES = ["si", "sq", "su", "sv", "si2", "sq2", "su2", "sv2", "nums", "aveinc"]
PSED_NAMES = ["ti", "tq", "tu", "tv", "ti2", "tq2", "tu2", "tv2", "numt"]
IMAGE_NAMES = ["image_b_i", "image_b_q", "image_b_u", "image_b_v",
               "image_b_d", "image_b_s", "image_b_e", "image_b_o"]
NO = 8


class MCRTError(RuntimeError):
    pass


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Engine:
    """One GPU context holding one model resident in HBM."""

    def __init__(self, model: Model, device: int = 0, stream: Optional[int] = None):
        self.lib = _lib.load()
        self.model = model
        h = C.c_void_p()
        rc = self.lib.mcrt_create(device, C.byref(h))
        if rc != 0:
            raise MCRTError(f"mcrt_create failed with code {rc} (no usable GPU?)")
        self.h = h
        self._keep = []
        if stream is not None:
            self._check(self.lib.mcrt_set_stream(self.h, C.c_void_p(stream)))
        self.upload(model)

    # -- plumbing ------------------------------------------------------------
    def _check(self, rc):
        if rc != 0:
            msg = self.lib.mcrt_last_error(self.h)
            raise MCRTError(f"libmcrt error {rc}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.mcrt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- model upload (the module state gridset/dustinit/setup build) --------
    def upload(self, m: Model):
        L = self.lib
        d = C.c_double
        keep = [_f64(x) for x in (m.rarr, m.r2arr, m.thetarr, m.costarr, m.tan2arr,
                                  m.phiarr, m.aarr, m.barr, m.vcell)]
        self._check(L.mcrt_set_grid(self.h, m.nrg, m.ntg, m.npg, *[ptr(x, d) for x in keep]))
        rho = _f64(m.rho)
        sp = np.ascontiguousarray(m.species, dtype=np.uint8)
        so = np.ascontiguousarray(m.species_opac, dtype=np.int32)
        self._check(L.mcrt_set_medium(self.h, ptr(rho, d), ptr(sp, C.c_uint8), ptr(so, C.c_int32)))
        dust = [_f64(x) for x in (m.nudust, m.kappad, m.adust, m.gdust, m.pldust, m.kappav)]
        tabs = [_f64(x) for x in (m.lTint, m.Tint, m.nuint, m.kapint, m.kappap)]
        self._check(L.mcrt_set_dust(self.h, m.nopac, m.nlam, *[ptr(x, d) for x in dust],
                                    m.nT, m.nnu, *[ptr(x, d) for x in tabs]))
        self._check(L.mcrt_set_source(self.h, m.rstar, m.tstar, m.rmin))
        obs = dict(th=_f64(m.thete), ph=_f64(m.phie), nw=np.ascontiguousarray(m.band_nw, np.int32),
                   bf=_f64(m.band_freq), bp=_f64(m.band_phi), ap=_f64(m.aperture2), mu=_f64(m.u_mu))
        self._check(L.mcrt_set_observer(
            self.h, m.ipeel, m.npeel, ptr(obs["th"], d), ptr(obs["ph"], d), m.nx, m.rmaxi,
            m.nband, m.nwmax, ptr(obs["nw"], C.c_int32), ptr(obs["bf"], d), ptr(obs["bp"], d),
            m.nfreq, m.nmu, m.nph, m.nap, ptr(obs["ap"], d), ptr(obs["mu"], d), m.numin, m.lnurat))
        self.set_tdust(m.tdust_initial())
        self.set_isrf(m.isrf_frac, m.isrf)
        self.set_spots(m.spots)
        self.set_accretion(m.accretion)
        self.set_diffusion(m.diffusion)
        self._check(L.mcrt_zero_tallies(self.h))
        self._iter = 0

    def set_diffusion(self, diffusion=None):
        """The diffusion approximation (dict as model.with_diffusion makes it:
        kappar, hdisk, rmaxd, nr0); None turns it off."""
        if not diffusion:
            self._check(self.lib.mcrt_set_diffusion(self.h, 0, None, None, 0.0, 0))
            self._diffusion = False
            return
        kr, hd = _f64(diffusion["kappar"]), _f64(diffusion["hdisk"])
        d = C.c_double
        self._check(self.lib.mcrt_set_diffusion(self.h, 1, ptr(kr, d), ptr(hd, d), float(diffusion["rmaxd"]),
                                                int(diffusion["nr0"])))
        self._diffusion = True

    def setdiffus(self, iteration: int) -> int:
        """setdiffus for Lucy iteration `iteration` (mcrt_setdiffus):
        flags the diffusion region from the last Lucy temperatures and
        restarts its dust at 0.1 K; returns the number of diffusion cells."""
        n = C.c_int64()
        self._check(self.lib.mcrt_setdiffus(self.h, int(iteration), C.byref(n)))
        return int(n.value)

    def diffusion_state(self):
        """(itmin per radial cell, photons re-emitted per surface cell in cell_shape)."""
        m = self.model
        itmin = np.zeros(m.nrg - 1, np.int32)
        cnt = np.zeros(m.cell_shape, np.uint32)
        self._check(self.lib.mcrt_get_diffusion(self.h, ptr(itmin, C.c_int32), ptr(cnt, C.c_uint32)))
        return itmin, cnt

    def set_accretion(self, accretion=None):
        """Disk-accretion photons (dict as model.with_accretion makes it:
        accfrac, ldint, hdisk); None turns them off."""
        if not accretion or not accretion["accfrac"]:
            self._check(self.lib.mcrt_set_accretion(self.h, None, None, 0.0))
            return
        ld, hd = _f64(accretion["ldint"]), _f64(accretion["hdisk"])
        d = C.c_double
        self._check(self.lib.mcrt_set_accretion(self.h, ptr(ld, d), ptr(hd, d), float(accretion["accfrac"])))

    def set_spots(self, spots=None):
        """Stellar hot spots (dict as model.with_spots makes it); None turns them off."""
        if not spots:
            self._check(self.lib.mcrt_set_spots(self.h, 0, 0.0, 0.0, 0.0, 0.0, 0.0))
            return
        self._check(self.lib.mcrt_set_spots(self.h, int(spots["nspot"]), float(spots["fspot"]),
                                            float(spots["spotlat"]), float(spots["spotlon"]),
                                            float(spots["tshock"]), float(spots["accfrac2"])))

    def set_isrf(self, isrf_frac: float, tables=None):
        """External illumination: the last isrf_frac of every run's packets
        enter from outside with ISRF frequencies (tables as model.isrf_tables
        gives them); 0 turns it off."""
        if not isrf_frac:
            self._check(self.lib.mcrt_set_isrf(self.h, 0, None, None, None, 0.0))
            return
        cdf, lc, lf = (_f64(tables[k]) for k in ("cdf", "logcdf", "logfreq"))
        d = C.c_double
        self._check(self.lib.mcrt_set_isrf(self.h, len(cdf), ptr(cdf, d), ptr(lc, d), ptr(lf, d), float(isrf_frac)))

    def set_density(self, rho):
        """Replace the cell densities (g/cm^3, cell_shape); species unchanged."""
        m = self.model
        r = _f64(np.asarray(rho, dtype=np.float64).reshape(m.cell_shape))
        sp = np.ascontiguousarray(m.species, dtype=np.uint8)
        so = np.ascontiguousarray(m.species_opac, dtype=np.int32)
        self._check(self.lib.mcrt_set_medium(self.h, ptr(r, C.c_double), ptr(sp, C.c_uint8), ptr(so, C.c_int32)))

    def adjoint_image(self, W, npackets: int, seed: int, precision: int = MCRT_MIXED, first_packet: int = 0,
                      grad: bool = True):
        """Loss sum(W * image_I) over the peeled images (W: nband x nx x nx x
        npeel) and, with grad, dLoss/drho per cell (cell_shape), by path
        replay (mcrt_adjoint_image)."""
        m = self.model
        Wf = np.ascontiguousarray(np.asarray(W, dtype=np.float32).reshape(m.nband, m.nx, m.nx, m.npeel))
        loss = C.c_double()
        g = np.zeros(m.cell_shape) if grad else None
        self._check(self.lib.mcrt_adjoint_image(self.h, int(first_packet), int(npackets), C.c_uint64(int(seed)),
                                                int(precision), ptr(Wf, C.c_float), C.byref(loss),
                                                ptr(g, C.c_double) if grad else None))
        return float(loss.value), g

    def set_tdust(self, tdust):
        t = _f64(tdust)
        self._check(self.lib.mcrt_set_tdust(self.h, ptr(t, C.c_double)))

    # -- the packet loop -----------------------------------------------------
    def zero_tallies(self):
        self._check(self.lib.mcrt_zero_tallies(self.h))

    def run_packets(self, npackets: int, seed: int, precision: int = MCRT_FP64, first_packet: int = 0):
        self._check(self.lib.mcrt_run_packets(self.h, int(first_packet), int(npackets),
                                              C.c_uint64(int(seed)), int(precision)))

    def set_launch(self, lanes=0, queue_per_lane=0, events_per_round=0, peel_blocks_per_cu=0, regroup=0):
        """Launch shape of run_packets: resident packet lanes (multiple of
        256), peel-off queue records per lane, events per lane per round,
        peel-kernel blocks per CU, waiting lanes per wavefront that trigger
        an interaction / emission / tally pass.  0 = default."""
        self._check(self.lib.mcrt_set_launch(self.h, int(lanes), int(queue_per_lane),
                                             int(events_per_round), int(peel_blocks_per_cu), int(regroup)))

    def set_peel_gather(self, fetch_lanes=0, tally_lanes=0):
        """Peel-kernel gather thresholds: idle lanes of a wavefront that
        trigger taking new items, and lanes waiting for a tally pass.
        0 = default."""
        self._check(self.lib.mcrt_set_peel_gather(self.h, int(fetch_lanes), int(tally_lanes)))

    def synchronize(self):
        self._check(self.lib.mcrt_synchronize(self.h))

    def last_run_ms(self) -> float:
        v = C.c_float()
        self._check(self.lib.mcrt_last_run_ms(self.h, C.byref(v)))
        return float(v.value)

    def run_stats(self) -> dict:
        """Rounds, per-kernel HIP-event times and queue counts of the last run."""
        v = np.zeros(len(RUN_STATS))
        self._check(self.lib.mcrt_run_stats(self.h, ptr(v, C.c_double)))
        return {k: float(v[i]) for i, k in enumerate(RUN_STATS)}

    def tallies(self) -> Dict[str, np.ndarray]:
        """Post-reduction tallies named as in the reference's modules."""
        m = self.model
        cs = m.cell_shape
        out = dict(dtauabs=np.zeros(cs), nabs=np.zeros(cs, np.int32), dbig=np.zeros(cs, np.int32),
                   killed=np.zeros(cs, np.int32))
        sed = np.zeros((10, NO, m.nap, m.nph, m.nmu, m.nfreq), np.float32)
        psed = np.zeros((9, NO, m.nap, max(m.npeel, 1), m.nfreq), np.float32)
        img = np.zeros((8, max(m.nband, 1), m.nx, m.nx, max(m.npeel, 1)), np.float32)
        cnt = np.zeros(16)
        f = C.c_float
        self._check(self.lib.mcrt_get_tallies(
            self.h, ptr(out["dtauabs"], C.c_double), ptr(out["nabs"], C.c_int32),
            ptr(out["dbig"], C.c_int32), ptr(out["killed"], C.c_int32),
            ptr(sed, f), ptr(psed, f), ptr(img, f), ptr(cnt, C.c_double)))
        for i, n in enumerate(SED_NAMES):
            out[n] = sed[i]
        for i, n in enumerate(PSED_NAMES):
            out[n] = psed[i]
        for i, n in enumerate(IMAGE_NAMES):
            out[n] = img[i]
        out["counters"] = {k: cnt[i] for i, k in enumerate(COUNTERS)}
        return out

    def dusttemp(self, nphot: float, mode: int = MCRT_T_REFERENCE, update_tdust: bool = True):
        m = self.model
        t = np.zeros(m.cell_shape)
        self._check(self.lib.mcrt_dusttemp(self.h, float(nphot), int(mode), int(bool(update_tdust)),
                                           ptr(t, C.c_double)))
        return t

    def dusttemp_sigma(self) -> dict:
        """tfinal's convergence statistic of the last dusttemp call
        (mcrt_dusttemp_sigma): sigma, sigma_prev, ncells, converge."""
        sig, prev = C.c_double(), C.c_double()
        n, conv = C.c_int64(), C.c_int32()
        self._check(self.lib.mcrt_dusttemp_sigma(self.h, C.byref(sig), C.byref(prev), C.byref(n), C.byref(conv)))
        return dict(sigma=sig.value, sigma_prev=prev.value, ncells=int(n.value), converge=bool(conv.value))

    def disk(self, npackets: int, seed: int, precision: int = MCRT_FP64,
             tmode: int = MCRT_T_REFERENCE, first_packet: int = 0, iteration: Optional[int] = None
             ) -> Dict[str, np.ndarray]:
        """One Lucy iteration (disk, newdisktherm_5_fuv.f90:1): with the
        diffusion approximation on, setdiffus for this iteration (numbered
        from 1 by default, as gridset calls it); zero tallies, run every
        packet, read the reduced tallies, and invert the absorbed energy to
        a new dust temperature that the next iteration re-emits with;
        tal["converge"] is tfinal's verdict ("yes"/"no") and tal["sigma"]
        its statistic."""
        self._iter = self._iter + 1 if iteration is None else int(iteration)
        if self._diffusion:
            self.setdiffus(self._iter)
        self.zero_tallies()
        self.run_packets(npackets, seed, precision, first_packet=first_packet)
        tal = self.tallies()
        tal["tdave"] = self.dusttemp(npackets, tmode, True)
        st = self.dusttemp_sigma()
        tal["sigma"], tal["converge"] = st["sigma"], "yes" if st["converge"] else "no"
        return tal

    def lucy(self, npackets: int, seed: int, precision: int = MCRT_FP64, n_iter_min: int = 1,
             n_iter_max: int = 10, tmode: int = MCRT_T_REFERENCE):
        """The Lucy driver of newtts.f90:215-256: iterations of disk until
        tfinal says converge = 'yes', kept going below n_iter_min and
        stopped at n_iter_max.  Iteration `it` (1-based) runs packets
        it*npackets .. of the run seeded by `seed`.  Returns the list of
        per-iteration (sigma, converge) and the last iteration's tallies."""
        hist, it, converge, tal = [], 0, "no", None
        while converge == "no":
            it += 1
            tal = self.disk(npackets, seed, precision, tmode, first_packet=it * npackets, iteration=it)
            converge = tal["converge"]
            hist.append((tal["sigma"], converge))
            if it < n_iter_min:
                converge = "no"
            if it >= n_iter_max:
                converge = "yes"
        return hist, tal

    # -- deterministic entries ----------------------------------------------
    def trace_rays(self, pos, u, ijk, on_wall, tau_stop, max_steps=4096, precision=MCRT_FP64):
        pos = _f64(pos).reshape(-1, 3)
        n = pos.shape[0]
        u = _f64(u).reshape(-1, 3)
        ijk = np.ascontiguousarray(ijk, dtype=np.int32).reshape(-1, 3)
        ow = np.ascontiguousarray(on_wall, dtype=np.uint8).reshape(-1)
        ts = _f64(tau_stop).reshape(-1)
        faces = np.zeros((n, max_steps), np.int32)
        cells = np.zeros((n, max_steps, 3), np.int32)
        pos_out = np.zeros((n, 3))
        ijk_out = np.zeros((n, 3), np.int32)
        tau_out = np.zeros(n)
        nsteps = np.zeros(n, np.int32)
        status = np.zeros(n, np.int32)
        d, i = C.c_double, C.c_int32
        self._check(self.lib.mcrt_trace_rays(
            self.h, int(precision), n, ptr(pos, d), ptr(u, d), ptr(ijk, i), ptr(ow, C.c_uint8), ptr(ts, d),
            int(max_steps), ptr(faces, i), ptr(cells, i), ptr(pos_out, d), ptr(ijk_out, i),
            ptr(tau_out, d), ptr(nsteps, i), ptr(status, i)))
        return dict(faces=faces, cells=cells, pos=pos_out, ijk=ijk_out, tau=tau_out,
                    nsteps=nsteps, status=status)

    def mueller(self, par, stokes, direction=None, kind=0, precision=MCRT_FP64):
        par = _f64(par).reshape(-1, 9)
        n = par.shape[0]
        s = _f64(stokes).reshape(n, 4).copy()
        dirv = _f64(direction).reshape(n, 5).copy() if direction is not None else np.zeros((n, 5))
        d = C.c_double
        self._check(self.lib.mcrt_mueller(self.h, int(precision), n, int(kind), ptr(par, d),
                                          ptr(dirv, d), ptr(s, d)))
        return s, dirv

    def enable_trace(self, n: int):
        self._check(self.lib.mcrt_enable_trace(self.h, int(n)))

    def get_trace(self, n: int) -> np.ndarray:
        t = np.zeros((n, 8))
        self._check(self.lib.mcrt_get_trace(self.h, ptr(t, C.c_double), int(n)))
        return t



class BatchEngine(Engine):
    """A batch of independent axisymmetric models (model.Batch) in one
    context: the shared dust, star and SED binning from batch.shared, the
    per-model walls and media concatenated (mcrt_set_batch).  Packet p of
    model k runs with seed + k: the path it takes in Engine(model k).run_packets
    with that seed."""

    def __init__(self, batch: Batch, device: int = 0, stream: Optional[int] = None):
        super().__init__(batch.shared, device=device, stream=stream)
        self.batch = batch
        cat = batch.concat()
        self._cat = cat
        self.coff = batch.cell_offsets()
        d = C.c_double
        sh = batch.shared
        phi = [_f64(x) for x in (sh.phiarr, sh.aarr, sh.barr)]
        so = np.ascontiguousarray(sh.species_opac, dtype=np.int32)
        self._check(self.lib.mcrt_set_batch(
            self.h, batch.nmodels, ptr(cat["nrg"], C.c_int32), ptr(cat["ntg"], C.c_int32),
            *[ptr(cat[k], d) for k in ("rarr", "r2arr", "thetarr", "costarr", "tan2arr")],
            *[ptr(x, d) for x in phi], ptr(cat["vcell"], d), ptr(cat["rho"], d),
            ptr(cat["species"], C.c_uint8), ptr(so, C.c_int32)))
        self.set_batch_tdust(cat["tdust"])

    def set_batch_tdust(self, tdust):
        t = _f64(tdust)
        self._check(self.lib.mcrt_set_batch_tdust(self.h, ptr(t, C.c_double)))

    def set_batch_launch(self, block=0, waves_per_simd=0, regroup=0, combine=0, graph=1):
        self._check(self.lib.mcrt_set_batch_launch(self.h, int(block), int(waves_per_simd), int(regroup),
                                                   int(bool(combine)), int(bool(graph))))

    def run_batch(self, npackets, seed: int, precision: int = MCRT_MIXED):
        """One persistent launch over npackets[k] packets of every model k
        (an int gives every model that many)."""
        n = np.broadcast_to(np.asarray(npackets, dtype=np.int64), (self.batch.nmodels,))
        self._npk = np.ascontiguousarray(n)
        self._check(self.lib.mcrt_run_batch(self.h, self._npk.ctypes.data_as(C.POINTER(C.c_int64)),
                                            C.c_uint64(int(seed)), int(precision)))

    def batch_tallies(self):
        """Per-model tallies of the last run: list of dicts (dtauabs, nabs,
        dbig, killed in the model's cell_shape; the SED arrays as
        Engine.tallies), plus the per-model work and the summed counters."""
        b = self.batch
        nc = int(self.coff[-1])
        sh = b.shared
        nsed = 10 * NO * sh.nap * sh.nph * sh.nmu * sh.nfreq
        dt = np.zeros(nc)
        ints = [np.zeros(nc, np.int32) for _ in range(3)]
        sed = np.zeros(b.nmodels * nsed, np.float32)
        work = np.zeros(b.nmodels)
        cnt = np.zeros(16)
        self._check(self.lib.mcrt_get_batch_tallies(
            self.h, ptr(dt, C.c_double), *[ptr(x, C.c_int32) for x in ints], ptr(sed, C.c_float),
            ptr(work, C.c_double), ptr(cnt, C.c_double)))
        out = []
        for k, m in enumerate(b.models):
            a, e = int(self.coff[k]), int(self.coff[k + 1])
            t = dict(dtauabs=dt[a:e].reshape(m.cell_shape), nabs=ints[0][a:e].reshape(m.cell_shape),
                     dbig=ints[1][a:e].reshape(m.cell_shape), killed=ints[2][a:e].reshape(m.cell_shape))
            s = sed[k * nsed:(k + 1) * nsed].reshape(10, NO, sh.nap, sh.nph, sh.nmu, sh.nfreq)
            for i, name in enumerate(SED_NAMES):
                t[name] = s[i]
            out.append(t)
        return out, work, {k: cnt[i] for i, k in enumerate(COUNTERS)}

    def batch_dusttemp(self, mode: int = MCRT_T_REFERENCE, update_tdust: bool = True):
        """Lucy temperatures of every model from the last run; list of
        per-model arrays in cell_shape."""
        nc = int(self.coff[-1])
        t = np.zeros(nc)
        self._check(self.lib.mcrt_batch_dusttemp(self.h, int(mode), int(bool(update_tdust)), ptr(t, C.c_double)))
        return [t[int(self.coff[k]):int(self.coff[k + 1])].reshape(m.cell_shape)
                for k, m in enumerate(self.batch.models)]

    def autotune(self, packets_per_model: int, seed: int = 1, precision: int = MCRT_MIXED):
        """Time every launch shape on a sample and keep the fastest: returns
        (best (block, waves, regroup, combine), table rows of block, waves,
        regroup, combine, ms, packets/s)."""
        best = np.zeros(4, np.int32)
        table = np.zeros((64, 6))
        nrows = C.c_int32()
        self._check(self.lib.mcrt_batch_autotune(self.h, int(packets_per_model), C.c_uint64(int(seed)), int(precision),
                                                 ptr(best, C.c_int32), ptr(table, C.c_double), 64, C.byref(nrows)))
        return tuple(int(x) for x in best), table[:min(nrows.value, 64)]


def setemiss(nudust, kappad, adust, Tint, nuint, smooth=None, nlambda=None, device: int = 0) -> Dict[str, object]:
    """setemiss (dust_5.f90:507-699) on the GPU: the Planck-mean tables of
    dust tables of shape (nopac, nlam) (nudust descending, eV) at the
    temperatures Tint (K/11605) and the x = h nu/kT grid nuint.  Returns
    kapint (nopac, nnu, nT) normalized CDFs, kappap and kappar (nopac, nT),
    nfail (qsimp calls that did not converge), nevals (integrand
    evaluations) and the kernel time in ms."""
    lib = _lib.load()
    nud, kap, alb = (np.ascontiguousarray(np.atleast_2d(a), dtype=np.float64) for a in (nudust, kappad, adust))
    nopac, nlam = nud.shape
    if kap.shape != nud.shape or alb.shape != nud.shape:
        raise ValueError("nudust, kappad and adust must have the same shape")
    T, nu = _f64(Tint), _f64(nuint)
    nT, nnu = len(T), len(nu)
    out = dict(kapint=np.zeros((nopac, nnu, nT)), kappap=np.zeros((nopac, nT)), kappar=np.zeros((nopac, nT)))
    sm = None if smooth is None else np.ascontiguousarray(smooth, dtype=np.int32)
    nl = None if nlambda is None else np.ascontiguousarray(nlambda, dtype=np.int32)
    nfail, nevals, ms = C.c_int64(), C.c_int64(), C.c_float()
    h = C.c_void_p()
    rc = lib.mcrt_create(device, C.byref(h))
    if rc != 0:
        raise MCRTError(f"mcrt_create failed with code {rc} (no usable GPU?)")
    try:
        rc = lib.mcrt_setemiss(h, nopac, nlam, None if nl is None else ptr(nl, C.c_int32), ptr(nud, C.c_double),
                               ptr(kap, C.c_double), ptr(alb, C.c_double), nT, ptr(T, C.c_double), nnu,
                               ptr(nu, C.c_double), None if sm is None else ptr(sm, C.c_int32),
                               ptr(out["kapint"], C.c_double), ptr(out["kappap"], C.c_double),
                               ptr(out["kappar"], C.c_double), C.byref(nfail), C.byref(nevals), C.byref(ms))
        if rc != 0:
            msg = lib.mcrt_last_error(h)
            raise MCRTError(f"libmcrt error {rc}: {msg.decode() if msg else ''}")
    finally:
        lib.mcrt_destroy(h)
    out["nfail"], out["nevals"], out["ms"] = int(nfail.value), int(nevals.value), float(ms.value)
    return out


def build_models(walls, params, clumps=None, device: int = 0):
    """Cell arrays of many models in one GPU launch (mcrt_build_models):
    walls = [(rarr, thetarr, phiarr), ...] (R*, radians), params = array
    (nmodels, 12) in the order mcrt_build_models documents, clumps = per
    model a clumping factor of the cell shape, or None.  Returns
    ([(rho, species, vcell), ...] of shape (np-1, nt-1, nr-1), kernel ms)."""
    lib = _lib.load()
    n = len(walls)
    P = np.ascontiguousarray(params, dtype=np.float64).reshape(n, 12)
    dims = np.array([[len(w[0]), len(w[1]), len(w[2])] for w in walls], np.int32)
    nrg, ntg, npg = (np.ascontiguousarray(dims[:, i]) for i in range(3))
    R, T, F = (np.ascontiguousarray(np.concatenate([_f64(w[i]) for w in walls])) for i in range(3))
    shapes = [(int(d[2]) - 1, int(d[1]) - 1, int(d[0]) - 1) for d in dims]
    sizes = [a * b * c for a, b, c in shapes]
    ncell = int(sum(sizes))
    clump = None
    if clumps is not None:
        clump = np.ascontiguousarray(np.concatenate([_f64(np.broadcast_to(1.0 if c is None else c, sh)).ravel()
                                                     for c, sh in zip(clumps, shapes)]))
    rho, vcell = np.zeros(ncell), np.zeros(ncell)
    species = np.zeros(ncell, np.uint8)
    nflag, ms = C.c_int64(), C.c_float()
    h = C.c_void_p()
    rc = lib.mcrt_create(device, C.byref(h))
    if rc != 0:
        raise MCRTError(f"mcrt_create failed with code {rc} (no usable GPU?)")
    try:
        rc = lib.mcrt_build_models(h, n, ptr(nrg, C.c_int32), ptr(ntg, C.c_int32), ptr(npg, C.c_int32),
                                   ptr(R, C.c_double), ptr(T, C.c_double), ptr(F, C.c_double), ptr(P, C.c_double),
                                   None if clump is None else ptr(clump, C.c_double), ptr(rho, C.c_double),
                                   ptr(species, C.c_uint8), ptr(vcell, C.c_double), C.byref(nflag), C.byref(ms))
        if rc != 0:
            msg = lib.mcrt_last_error(h)
            raise MCRTError(f"libmcrt error {rc}: {msg.decode() if msg else ''}")
    finally:
        lib.mcrt_destroy(h)
    if nflag.value:
        raise MCRTError(f"mcrt_build_models: zerod found no single streamline root in {nflag.value} cells")
    out, o = [], 0
    for sh, sz in zip(shapes, sizes):
        out.append((rho[o:o + sz].reshape(sh), species[o:o + sz].reshape(sh), vcell[o:o + sz].reshape(sh)))
        o += sz
    return out, float(ms.value)


def diskflux(tal: Dict[str, np.ndarray], model: Model, npackets: int, ltot: Optional[float] = None):
    """flux_test.f90:19-80: normalize SEDs by np/nmu/nph/(ltot/lsol) and
    peeled SEDs and images by np/(ltot/lsol); turn the squared sums into
    relative errors.  Returns new arrays; the input is untouched."""
    ltot = model.lstar if ltot is None else ltot
    rnorm = np.float32(npackets / model.nmu / model.nph / (ltot / LSUN))
    image_norm = np.float32(npackets / (ltot / LSUN))
    out = {}
    si, nums = tal["si"].astype(np.float32), tal["nums"].astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        many = nums > 1.0
        for k, name in (("si2", "si"), ("sq2", "sq"), ("su2", "su"), ("sv2", "sv")):
            s2, s = tal[k].astype(np.float32), tal[name].astype(np.float32)
            rel = np.sqrt((s2 - s ** 2 / nums) / (nums - 1.0) * nums) / si
            if k == "si2":
                rel = rel + 1.0 / nums
            out[k] = np.where(many, rel, 0.0).astype(np.float32)
        out["aveinc"] = np.where(many, tal["aveinc"] / nums, tal["aveinc"]).astype(np.float32)
        for name in ("si", "sq", "su", "sv"):
            out[name] = (tal[name] / rnorm).astype(np.float32)
        numt = tal["numt"].astype(np.float32)
        mt = numt > 1.0
        for k, name in (("ti2", "ti"), ("tq2", "tq"), ("tu2", "tu"), ("tv2", "tv")):
            t2, t = tal[k].astype(np.float32), tal[name].astype(np.float32)
            out[k] = np.where(mt, np.sqrt((t2 * numt / t ** 2 - 1.0) / (numt - 1.0)), 0.0).astype(np.float32)
        for name in ("ti", "tq", "tu", "tv"):
            out[name] = (tal[name] / image_norm).astype(np.float32)
        for name in IMAGE_NAMES:
            out[name] = (tal[name] / image_norm).astype(np.float32)
    out["nums"], out["numt"] = tal["nums"], tal["numt"]
    return out


__all__ = ["Engine", "MCRTError", "diskflux", "setemiss", "build_models", "MCRT_FP64", "MCRT_MIXED", "MCRT_T_REFERENCE",
           "MCRT_T_PERCELL", "COUNTERS", "SED_NAMES", "PSED_NAMES", "IMAGE_NAMES"]
