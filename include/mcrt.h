/*
 * mcrt.h -- C ABI of the MI355X photon-packet engine (libmcrt.so).
 *
 * The reference (evc9pu/src3_mpi_cfitsio, Fortran 90 + MPI) has no plugin
 * API: its packet path is a set of Fortran subroutines sharing module-global
 * arrays.  Each entry point below replaces one of those routines or the
 * module state it reads, with plain pointers and sizes so that a Fortran
 * `interface ... bind(C)` block (src3_mpi_cfitsio_amd/fortran/mcrt_mod.f90)
 * or ctypes can call it.  Citations are relative to src3_mpi_cfitsio/.
 *
 * Conventions
 *  - every call returns 0 on success, a nonzero MCRT_E_* code otherwise
 *    (the reference `stop`s); mcrt_last_error() describes the failure;
 *  - host buffers are caller-owned and copied; device buffers are owned by
 *    the context and stay resident between calls (upload once per model);
 *  - walls keep the reference's 1-based meaning: array element k-1 holds
 *    wall k; cell (ir,it,ip) is element (ir-1)+(nrg-1)*((it-1)+(ntg-1)*(ip-1));
 *  - units are the reference's: lengths in stellar radii, photon energy in
 *    eV, temperature in K/11605, density in g/cm^3, volume in cm^3;
 *  - a context is used by one host thread at a time (the reference runs one
 *    thread per MPI rank); work is queued on the context's stream and every
 *    call that returns data to the host synchronizes that stream.  Inside
 *    mcrt_run_packets the peel-off kernels of each round also run on a
 *    second stream the context owns (default priority), ordered by events;
 *    the context's stream waits for them before the call returns, so work
 *    queued after it sees complete tallies.
 */
#ifndef MCRT_H
#define MCRT_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mcrt_ctx mcrt_ctx;

enum { MCRT_OK = 0, MCRT_E_ARG = 1, MCRT_E_STATE = 2, MCRT_E_HIP = 3, MCRT_E_NODEVICE = 4 };
enum { MCRT_FP64 = 0, MCRT_MIXED = 1 };           /* transport precision */
enum { MCRT_T_REFERENCE = 0, MCRT_T_PERCELL = 1 }; /* Lucy temperature mode */

/* Context on one GPU; stands in for one MPI rank's module state
 * (ttsre_mpi_mod.f90, grid_mod_5.f90, dust_mod_5.f90, tts_mod_5.f90). */
int mcrt_create(int device, mcrt_ctx **ctx);
int mcrt_destroy(mcrt_ctx *ctx);
const char *mcrt_last_error(const mcrt_ctx *ctx);
/* Queue work on an external HIP stream (hipStream_t as void*); NULL = own stream. */
int mcrt_set_stream(mcrt_ctx *ctx, void *stream);
int mcrt_synchronize(mcrt_ctx *ctx);

/* Grid walls and cell volumes: replaces grid_allocate + the wall section of
 * gridset (grid_mod_5.f90:69-122, gridset_5_model13new.f90:516,706-781,1152).
 * rarr,r2arr: nrg; thetarr,costarr,tan2arr: ntg (tan2arr=-1 flags the
 * equator, ntg odd); phiarr,aarr,barr: npg; vcell: cells. */
int mcrt_set_grid(mcrt_ctx *ctx, int32_t nrg, int32_t ntg, int32_t npg,
                  const double *rarr, const double *r2arr,
                  const double *thetarr, const double *costarr, const double *tan2arr,
                  const double *phiarr, const double *aarr, const double *barr,
                  const double *vcell);

/* Medium: densarr(nrg,ntg,npg,ndg+2) and findopac (grid_mod_5.f90:29-63)
 * compacted to one thermal species per cell, as select_dust_2.f90:22-34 and
 * tfinal_5_test1.f90:244-247 require.  species[c] in 0..10 (0 = empty),
 * species_opac[0..10] = findopac(.,.,.,id) opacity index (0-based). */
int mcrt_set_medium(mcrt_ctx *ctx, const double *rho, const uint8_t *species,
                    const int32_t *species_opac);

/* Dust tables (dust_mod_5.f90:7-21, filled by dustinit/setemiss,
 * dust_5.f90:345-699): per opacity index, nudust descending in eV with
 * kappa/kappav, albedo, g, pl; the Planck tables lTint,Tint (nT), nuint (nnu),
 * kapint[(iop*nnu + j)*nT + i] normalized CDFs and kappap[iop*nT + i].
 * Limit: nopac <= 4 opacity tables per model (a packet carries the
 * opacities of every table at its frequency, opacset's per-packet kapd);
 * the reference allows up to 187 (dust_mod_5.f90), one per species and
 * grain-size bin of findopac.  More are refused with MCRT_E_ARG. */
int mcrt_set_dust(mcrt_ctx *ctx, int32_t nopac, int32_t nlam,
                  const double *nudust, const double *kappad, const double *adust,
                  const double *gdust, const double *pldust, const double *kappav,
                  int32_t nT, int32_t nnu, const double *lTint, const double *Tint,
                  const double *nuint, const double *kapint, const double *kappap);

/* setemiss (dust_5.f90:507-699) on the GPU: the Planck-mean emissivity
 * tables that mcrt_set_dust takes, computed from the dust tables with the
 * reference's qsimp integration per (species, temperature, interval).
 * Inputs as in mcrt_set_dust (nudust/kappad/adust[iop*nlam + k]);
 * nlambda[iop] = table entries used per species (NULL: all nlam);
 * smooth[iop] != 0 smooths kappap as for species above ndg (needs nT >= 21;
 * NULL: none).  Outputs (host): kapint[(iop*nnu + j)*nT + i], kappap and
 * kappar [iop*nT + i] (kappar = the inverse Rosseland sum, as setemiss
 * leaves it).  nfail (optional) = qsimp calls that did not converge (the
 * reference's counta + countc); nevals (optional) = integrand evaluations;
 * ms (optional) = kernel time.  Replaces the
 * call `setemiss()` in dustinit (dust_5.f90:499) and its MPI split. */
int mcrt_setemiss(mcrt_ctx *ctx, int32_t nopac, int32_t nlam, const int32_t *nlambda,
                  const double *nudust, const double *kappad, const double *adust,
                  int32_t nT, const double *Tint, int32_t nnu, const double *nuint,
                  const int32_t *smooth, double *kapint, double *kappap, double *kappar,
                  int64_t *nfail, int64_t *nevals, float *ms);

/* setemiss without allocation (no counterpart in the reference: it is for
 * callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * kernels and results as mcrt_setemiss, on tables already in device memory:
 * nudust, kappad, adust, Tint, nuint and the outputs kapint, kappap, kappar
 * are device pointers (layouts as in mcrt_setemiss); nlambda, smooth (may be
 * NULL), nfail and nevals (may be NULL) are host memory.  Calls no hipMalloc
 * or hipFree: the interval integrals go through a fixed 16 MiB device
 * scratch array per GPU, species after species in chunks that fit.  Limits:
 * nopac <= 1024, nT <= 4096, nnu * nT <= 1048576 (2 nnu nT doubles of one
 * species); larger sizes are refused with MCRT_E_ARG.  Every call on a GPU
 * shares that GPU's scratch, so calls that overlap on one GPU (two host
 * threads, or two contexts on different streams) race; calls in sequence do
 * not.  Synchronizes the context stream. */
int mcrt_setemiss_noalloc(mcrt_ctx *ctx, int32_t nopac, int32_t nlam, const int32_t *nlambda,
                          const double *nudust, const double *kappad, const double *adust,
                          int32_t nT, const double *Tint, int32_t nnu, const double *nuint,
                          const int32_t *smooth, double *kapint, double *kappap, double *kappar,
                          int64_t *nfail, int64_t *nevals);

/* Cell arrays of nmodels synthetic models in one launch: gridset's cell
 * loop (gridset_5_model13new.f90:1105-1180) with cell volumes, the flared
 * disk of the synthetic workloads normalized to its mass, the analytic
 * Ulrich envelope of densenv (densenv.f90:30-45, 69-73, 111-113; zerod.f90)
 * at the cell centres, an optional clumping factor per cell, and the
 * species (1 disk-dominated, 3 envelope, 0 empty).  Walls of model k
 * (nrg[k], ntg[k], npg[k] of them, R* and radians) follow those of model
 * k-1 in rarr/thetarr/phiarr; cells likewise, each model in the cell order
 * of mcrt_set_grid.  params[k*12 + i]: rmincgs (R* in cm), autors (R* per
 * AU), disk Rin and Rout (AU), h0 (AU at 1 AU), flaring exponent, radial
 * exponent, envelope Rc (R*), rhoe0 (g/cm^3), envelope inner radius (R*),
 * disk mass (g), ambient density floor (g/cm^3).  clump may be NULL.
 * nflag = cells where zerod found no single streamline root (density NaN).
 * No counterpart in the reference as a whole: it replaces the grid set-up
 * of analytic models, one model per program run. */
int mcrt_build_models(mcrt_ctx *ctx, int32_t nmodels, const int32_t *nrg, const int32_t *ntg,
                      const int32_t *npg, const double *rarr, const double *thetarr,
                      const double *phiarr, const double *params, const double *clump,
                      double *rho, uint8_t *species, double *vcell, int64_t *nflag, float *ms);

/* External illumination (isrf_frac > 0 in the reference): the last npout
 * packets of every mcrt_run_packets call, npout = isrf_frac * npackets
 * rounded as newdisktherm_5_fuv.f90:158-163 rounds it, are photons of the
 * interstellar field (initpout.f90:18-57, newdisktherm_5_fuv.f90:416-455)
 * entering at 0.999999 rmax, with frequencies from rep_isrf.f90 on the
 * tables of isrf_mod (cdf_i, logcdf_i, logfreq_i of nisrf entries, as
 * isrf_cdf.f90 leaves them; cdf from 0 to 1).  isrf_frac = 0 turns it off.
 * Batched runs and the adjoint do not emit external photons. */
int mcrt_set_isrf(mcrt_ctx *ctx, int32_t nisrf, const double *cdf, const double *logcdf,
                  const double *logfreq, double isrf_frac);

/* Stellar hot spots (initpspot.f90 with ispot = 1; spotset.f90): nspot
 * (1, or 2 antipodal) caps covering the fraction fspot of the surface,
 * centred at latitude spotlat and longitude spotlon (degrees), emitting at
 * tshock (K/11605).  Every mcrt_run_packets call then starts with its
 * hot-spot photons, after any accretion photons (the iphot < npacc +
 * int(accfrac2 np) that are not accretion photons,
 * newdisktherm_5_fuv.f90:172-177 and :296-313: emitted inside a cap, half
 * of them as X-rays); with
 * fspot < 1 a stellar photon emitted inside a cap takes tshock, with
 * fspot = 1 every stellar photon does.  nspot = 0 turns spots off.
 * Batched runs and the adjoint do not use spots. */
int mcrt_set_spots(mcrt_ctx *ctx, int32_t nspot, double fspot, double spotlat, double spotlon,
                   double tshock, double accfrac2);

/* Disk-accretion photons (accfrac > 0 in the reference; initpacc_4.f90,
 * newdisktherm_5_fuv.f90:164-170 and :242-283): the first
 * int(accfrac * npackets) - 1 packets of every mcrt_run_packets call start
 * inside the disk and come before the hot-spot photons.  ldint holds the
 * cumulative accretion luminosity by radial cell and hdisk the scale height
 * (R*) of each radial cell (grid_mod's Ldint and hdisk as
 * diskprofile_model11.f90:158-227 leaves them: nrg - 1 values each, Ldint
 * from 0 in the first cell to 1 in the last).  A photon is placed in the
 * cell j with Ldint(j) <= xi < Ldint(j+1), at that cell's mid radius, a
 * random azimuth and a Gaussian height of scale hdisk(j), redrawn until
 * the cell holds dust; the dust absorbs it (nabs, dbig) and re-emits it at
 * the cell's temperature, peeled with origin 2 (disk).  Set the grid first.
 * accfrac = 0 turns it off.  Batched runs and the adjoint do not emit
 * accretion photons. */
int mcrt_set_accretion(mcrt_ctx *ctx, const double *ldint, const double *hdisk, double accfrac);

/* The diffusion approximation of very optically thick regions (DIFFUSION =
 * YES in the reference's parameter file; setdiffus_5.f90,
 * emit_common_2.f90:88-378, tfinal_5_test1.f90:232-233).  on = 0 turns it
 * off.  kappar[iop*nT + i]: the Rosseland means setemiss leaves (kappar of
 * mcrt_setemiss); hdisk[ir-1]: the disk scale height (R*) of radial cell
 * ir; rmaxd: the disk's outer radius (R*); nr0: the first radial cell
 * that may diffuse (1 when the dust destruction radius is the inner wall,
 * else 3).  Set grid, medium, dust and source first.  Theta diffusion
 * only (the reference's setdiffus sets no other direction). */
int mcrt_set_diffusion(mcrt_ctx *ctx, int32_t on, const double *kappar, const double *hdisk,
                       double rmaxd, int32_t nr0);
/* setdiffus for Lucy iteration iter (called by gridset at the start of each
 * disk call): the region of every radial column where the Rosseland
 * optical depth toward the pole reaches 1 within a twentieth of the disk
 * scale height, from the temperatures of the last mcrt_dusttemp (below
 * iter 2 floored at 50 K), mirrored about the midplane and the same in
 * every phi wedge; the dust temperature of those cells restarts at 0.1 K
 * (gridset_5_model13new.f90:1857-1868).  Then mcrt_run_packets moves every
 * photon absorbed there to the region's surface and re-emits it at the
 * surface temperature, which grows with each such photon (T^4 += 4 pi T*^4
 * / (ns dA), ns the run's stellar photons; the GPU orders the photons of a
 * surface cell by arrival), and mcrt_dusttemp keeps those temperatures for
 * the region.  *ncells (may be NULL): diffusion cells. */
int mcrt_setdiffus(mcrt_ctx *ctx, int32_t iter, int64_t *ncells);
/* The diffusion region and its photons: itmin[ir-1] (nrg-1) = the first
 * diffusion cell of column ir from the pole (0: none; cells itmin ..
 * ntg-itmin diffuse); count[cell] = photons re-emitted from each surface
 * cell since the tallies were zeroed.  Either pointer may be NULL. */
int mcrt_get_diffusion(mcrt_ctx *ctx, int32_t *itmin, uint32_t *count);

/* Star (tts_mod_5.f90: rstar [R_sun], tstar [K/11605]; rmin = rarr(1)). */
int mcrt_set_source(mcrt_ctx *ctx, double rstar, double tstar, double rmin);

/* Observer: peel-off directions and image/SED binning (tts_mod_5.f90:22-62,
 * setup.f90:36-41,115-130, filt_mod.f90).  thete/phie in radians (npeel);
 * images nx*nx over [-rmaxi, rmaxi]; nband filter curves of band_nw[b] lines
 * stored at band_freq/phi[b*nwmax + k] (Hz, ascending); SEDs nfreq x nmu x
 * nph x nap with aperture2 (R*^2) and u_mu (nmu bin centres of cos theta). */
int mcrt_set_observer(mcrt_ctx *ctx, int32_t ipeel, int32_t npeel,
                      const double *thete, const double *phie,
                      int32_t nx, double rmaxi,
                      int32_t nband, int32_t nwmax, const int32_t *band_nw,
                      const double *band_freq, const double *band_phi,
                      int32_t nfreq, int32_t nmu, int32_t nph, int32_t nap,
                      const double *aperture2, const double *u_mu,
                      double numin, double lnurat);

/* Dust temperature per cell used for re-emission: tdust(nrg,ntg,npg)
 * (grid_mod_5.f90; set from tdave in gridset_5_model13new.f90:1857-1868). */
int mcrt_set_tdust(mcrt_ctx *ctx, const double *tdust);

/* Zero every tally (initarr_2_new.f90:123-192 / the reset before disk). */
int mcrt_zero_tallies(mcrt_ctx *ctx);

/* The packet loop of disk (newdisktherm_5_fuv.f90:220-491, stellar photons)
 * with propagate_5_fuv.f90, tauint_3d_2_fuv.f90, stokes.f90 and
 * peeloff_3d_2_test.f90, for packets first_packet .. first_packet+npackets-1
 * of the run seeded by `seed` (counter-based RNG: results do not depend on
 * launch shape).  Tallies accumulate across calls until mcrt_zero_tallies. */
int mcrt_run_packets(mcrt_ctx *ctx, int64_t first_packet, int64_t npackets,
                     uint64_t seed, int32_t precision);

/* Launch configuration of mcrt_run_packets (no counterpart in the reference,
 * whose rank loop has none).  lanes: resident packet slots of the transport
 * kernel (a multiple of 256).  queue_per_lane: peel-off queue records per
 * lane (at least 8).  events_per_round: interaction events per lane between
 * two peel-kernel passes (at most queue_per_lane - 7, so the queue cannot
 * overflow).  peel_blocks_per_cu: peel-kernel blocks per CU.  regroup: lanes
 * of a wavefront (1..64) that must be waiting for an interaction, an
 * emission or a peel-off tally before that code runs while other lanes still
 * walk.  0 keeps the built-in default.  Results do not depend on these
 * values. */
int mcrt_set_launch(mcrt_ctx *ctx, int64_t lanes, int32_t queue_per_lane,
                    int32_t events_per_round, int32_t peel_blocks_per_cu, int32_t regroup);
/* Gather thresholds of the peel kernel (no counterpart in the reference):
 * idle lanes of a wavefront that trigger taking new (record, direction)
 * items with their first crossing, and lanes whose walks reached the
 * observer that trigger a tally pass.  0 keeps the defaults (4 and 8).
 * Results do not depend on these values. */
int mcrt_set_peel_gather(mcrt_ctx *ctx, int32_t fetch_lanes, int32_t tally_lanes);

/* Post-ALLREDUCE tallies (newdisktherm_5_fuv.f90:524-602): any pointer may be
 * NULL.  Layouts: dtauabs/nabs/dbig/killed per cell; sed
 * [10][8][nap][nph][nmu][nfreq] = si,sq,su,sv,si2,sq2,su2,sv2,nums,aveinc;
 * psed [9][8][nap][npeel][nfreq] = ti..tv,ti2..tv2,numt; image
 * [8][nband][nx][nx][npeel] = image_b_{i,q,u,v,d,s,e,o}; counters[16]. */
int mcrt_get_tallies(mcrt_ctx *ctx, double *dtauabs, int32_t *nabs, int32_t *dbig,
                     int32_t *killed, float *sed, float *psed, float *image,
                     double *counters);

/* Adjoint of the peeled images (no counterpart in the reference; the
 * gradient for model fitting).  For packets first_packet ..
 * first_packet+npackets-1 of the run `seed` (the same Philox streams as
 * mcrt_run_packets), *loss receives L = sum over the image pixels of
 * W * image_I (W: nband*nx*nx*npeel floats in the image layout of
 * mcrt_get_tallies, plane image_b_i) and, when grad is not NULL, grad[cell]
 * receives dL/drho for every cell (FP64, the density of mcrt_set_medium).
 * Each packet is run twice from its counters (path replay: a forward pass
 * for its loss, a replay pass that scatters the derivative of every flight's
 * sampling density and of every peel-off's exp(-tau)); no path is stored.
 * The gradient is taken at the current dust temperature.  The run's own
 * tallies are left untouched.  Only stellar photons are replayed: with
 * accretion, hot-spot or external photons set on the context the call
 * returns MCRT_E_STATE. */
int mcrt_adjoint_image(mcrt_ctx *ctx, int64_t first_packet, int64_t npackets,
                       uint64_t seed, int32_t precision, const float *W,
                       double *loss, double *grad);

/* ---- Batches of axisymmetric models in one persistent launch (no
 * counterpart in the reference, where every model of a grid is a separate
 * program run over the same packet loop).  The models share the context's
 * dust, source and SED binning (mcrt_set_dust / mcrt_set_source /
 * mcrt_set_observer with ipeel = 0: SEDs only); each has its own walls,
 * medium, dust temperatures and tallies.  Arrays are the per-model arrays of
 * mcrt_set_grid / mcrt_set_medium concatenated model after model: walls by
 * nrg[k] (rarr, r2arr) and ntg[k] (thetarr, costarr, tan2arr), cells by
 * (nrg[k]-1)*(ntg[k]-1).  Every model has one azimuthal cell: phiarr, aarr,
 * barr hold its two walls (0 and 2*pi).  Packet p of model k runs with seed
 * seed + k, so it follows the path it follows in a single-model run of
 * model k with that seed. */
int mcrt_set_batch(mcrt_ctx *ctx, int32_t nmodels, const int32_t *nrg, const int32_t *ntg,
                   const double *rarr, const double *r2arr, const double *thetarr,
                   const double *costarr, const double *tan2arr, const double *phiarr,
                   const double *aarr, const double *barr, const double *vcell,
                   const double *rho, const uint8_t *species, const int32_t *species_opac);
/* dust temperatures of every model, concatenated cells */
int mcrt_set_batch_tdust(mcrt_ctx *ctx, const double *tdust);
/* Launch shape of mcrt_run_batch: block 64, 128 or 256 threads; resident
 * waves per SIMD (1..8, clamped to what the kernel's registers allow);
 * regroup = waiting lanes per wavefront that trigger an interaction /
 * emission pass (1..64); combine != 0: write-combined dtauabs (one FP64
 * atomic per cell visit instead of per segment); graph != 0: the run is a
 * captured hipGraph (tally reset + launch), replayed while the shape and
 * buffers stay the same.  0 keeps the current value (block, waves,
 * regroup) — use mcrt_batch_autotune to choose them. */
int mcrt_set_batch_launch(mcrt_ctx *ctx, int32_t block, int32_t waves_per_simd, int32_t regroup,
                          int32_t combine, int32_t graph);
/* One persistent launch over npackets[k] packets of every model k (zeroes
 * the batch tallies first).  Models run most expensive first: by the
 * interactions per packet of the previous run, before any run by the
 * column density along the midplane. */
int mcrt_run_batch(mcrt_ctx *ctx, const int64_t *npackets, uint64_t seed, int32_t precision);
/* Batch tallies of the last run; any pointer may be NULL.  dtauabs, nabs,
 * dbig, killed: concatenated cells; sed: nmodels blocks of the
 * mcrt_get_tallies SED layout; work: interactions + packets per model;
 * counters[16] (the mcrt_get_tallies counters) summed over the batch. */
int mcrt_get_batch_tallies(mcrt_ctx *ctx, double *dtauabs, int32_t *nabs, int32_t *dbig,
                           int32_t *killed, float *sed, double *work, double *counters);
/* Lucy temperatures of every model from the last run (nphot = that
 * model's packets), as mcrt_dusttemp; tdave: concatenated cells, may be NULL. */
int mcrt_batch_dusttemp(mcrt_ctx *ctx, int32_t mode, int32_t update_tdust, double *tdave);
/* Autotuner: times the batch kernel over block {64,128,256} x waves per SIMD
 * {1,2,4} (where they fit) x regroup {16,32,48} x combine {0,1} on
 * packets_per_model packets of every model, measuring the bulk rate (each
 * trial stops when the queue is handed out), and keeps the fastest shape.
 * best[4] = block, waves, regroup, combine; table (may be NULL) receives up
 * to max_rows rows of 6 doubles (the four, ms, packets/s); *nrows = rows
 * tried.  Tallies of the trials are void. */
int mcrt_batch_autotune(mcrt_ctx *ctx, int64_t packets_per_model, uint64_t seed, int32_t precision,
                        int32_t *best, double *table, int32_t max_rows, int32_t *nrows);

/* Lucy temperature (tfinal_5_test1.f90 + dusttemp_5.f90) from the device
 * dtauabs tally, FP64.  nphot = number of stellar packets of the iteration.
 * mode MCRT_T_REFERENCE: upper hemisphere at ip=1 averaged with its mirror,
 * mirrored and copied to all phi; MCRT_T_PERCELL: every cell on its own.
 * update_tdust != 0 makes the result the re-emission temperature of the next
 * run.  tdave may be NULL. */
int mcrt_dusttemp(mcrt_ctx *ctx, double nphot, int32_t mode, int32_t update_tdust,
                  double *tdave);

/* Convergence test of the Lucy iteration (tfinal_5_test1.f90:334-368):
 * statistics of the last mcrt_dusttemp call.  Over the upper hemisphere at
 * ip = 1, every cell whose old temperature (the dust temperature the run
 * re-emitted with) and new one both exceed 0.1 K and differ gives
 * x = (T_new - T_old) / T_new; sigma = sqrt(<x^2> - <|x|>^2) over those
 * ncells (NaN when there are none, as 0/0 in the reference).  sigma_prev is
 * the sigma of the last mcrt_dusttemp call before it that updated the dust
 * temperature (the previous Lucy iteration; NaN before the second one), and
 * converge = 1 when |sigma_prev - sigma| / sigma < 0.25, the reference's
 * converge = 'yes'.  Any pointer may be NULL. */
int mcrt_dusttemp_sigma(mcrt_ctx *ctx, double *sigma, double *sigma_prev, int64_t *ncells,
                        int32_t *converge);

/* Deterministic entries for parity tests (no counterpart in the reference;
 * they expose tauint's face loop and the stokes/stokespeel Mueller update).
 * trace_rays: n rays from pos (3n) along u (3n) in cells ijk (3n) until
 * optical depth tau_stop (rho is taken as opacity per unit length), the grid
 * edge or the star.  faces (n*max_steps) and cells (3*n*max_steps) record each
 * crossing; status: 0 stopped, 1 escaped, 2 star, 3 max steps, -1 no wall. */
int mcrt_trace_rays(mcrt_ctx *ctx, int32_t precision, int32_t n, const double *pos,
                    const double *u, const int32_t *ijk, const uint8_t *on_wall,
                    const double *tau_stop, int32_t max_steps, int32_t *faces,
                    int32_t *cells, double *pos_out, int32_t *ijk_out, double *tau_out,
                    int32_t *nsteps, int32_t *status);
/* kind 0: stokes.f90:117-336 for an accepted (bmu, ri1); kind 1: stokespeel.f90.
 * par (9n): g, pl, bmu, ri1, cost, sint, phi, costp, sintp; dir (5n) in/out
 * cost,sint,phi,cosp,sinp (kind 0); stokes (4n) in/out I,Q,U,V. */
int mcrt_mueller(mcrt_ctx *ctx, int32_t precision, int32_t n, int32_t kind,
                 const double *par, double *dir, double *stokes);

/* Per-packet record of the next mcrt_run_packets calls (8 doubles per packet,
 * indexed by packet - first_packet): final nub, cost, phi, I, Q, U,
 * interactions, exit code (1 escaped, 2 killed, -1 error).  n = 0 disables. */
int mcrt_enable_trace(mcrt_ctx *ctx, int64_t n);
int mcrt_get_trace(mcrt_ctx *ctx, double *trace, int64_t n);

/* Duration in ms of the last mcrt_run_packets launch, from HIP events
 * recorded on the context stream (synchronizes). */
int mcrt_last_run_ms(mcrt_ctx *ctx, float *ms);

/* Measurement of the last mcrt_run_packets (no counterpart in the
 * reference): out[MCRT_NSTATS], indexed by MCRT_STAT_*.  Kernel times are
 * sums over rounds of HIP-event intervals bracketing each launch on the
 * context stream. */
enum {
  MCRT_STAT_ROUNDS = 0,        /* transport rounds */
  MCRT_STAT_TRANSPORT_MS = 1,  /* summed transport-kernel time */
  MCRT_STAT_PEEL_MS = 2,       /* summed peel-kernel time */
  MCRT_STAT_PEEL_LAUNCHES = 3, /* peel-kernel launches */
  MCRT_STAT_RECORDS = 4,       /* peel-off records queued (each walked in npeel directions) */
  MCRT_STAT_PARKED = 5,        /* lane-rounds ending with a parked packet */
  MCRT_STAT_LANES = 6,         /* resident lanes */
  MCRT_STAT_EVENTS0 = 7,       /* events per lane in the first round */
  MCRT_STAT_RECORD_BYTES = 8,  /* bytes of one peel-off record */
  MCRT_STAT_LANE_BYTES = 9,    /* bytes of one parked lane */
  /* errors since mcrt_zero_tallies (each also in counters[15]) by kind: */
  MCRT_STAT_WALK_ERRORS = 10,  /* a transport flight found no wall or passed the segment guard */
  MCRT_STAT_PEEL_ERRORS = 11,  /* a peel-off walk did */
  MCRT_STAT_OTHER_ERRORS = 12, /* an emission placed nowhere, a full peel-off queue */
  MCRT_STAT_TAIL_ROUNDS = 13,  /* rounds run with one wavefront per remaining packet */
  MCRT_NSTATS = 14
};
int mcrt_run_stats(mcrt_ctx *ctx, double *out);

#ifdef __cplusplus
}
#endif
#endif
