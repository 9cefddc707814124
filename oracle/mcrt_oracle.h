/*
 * mcrt_oracle.h -- CPU FP64 restatement of the reference photon-packet path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (libmcrt.so) never links or calls it.
 *
 * Every function restates one routine of the reference Fortran
 * (evc9pu/src3_mpi_cfitsio, files cited per function) in plain C99 with the
 * same floating-point operation order, so that a HIP FP64 kernel written
 * independently with the same order produces bit-identical face sequences.
 *
 * Index conventions follow the reference: walls are 1-based (rarr(1..nrg)),
 * cells 1..n-1, ir==0 means "inside the star", ir==nrg means "escaped".
 * Arrays are stored 0-based, so wall k lives at r2arr[k-1].
 * Cell (ir,it,ip) lives at (ir-1) + (nrg-1)*((it-1) + (ntg-1)*(ip-1)).
 */
#ifndef MCRT_ORACLE_H
#define MCRT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NSPECIES 10   /* ndg+2, grid_mod_5.f90:12 */
#define ORC_NO 8          /* SED components, tts_mod_5.f90:19 */

/* Model description: every pointer is caller-owned and read-only. */
typedef struct {
    /* grid walls (gridset_5_model13new.f90:516,706-781) */
    int32_t nrg, ntg, npg;
    const double *rarr, *r2arr;                 /* nrg */
    const double *thetarr, *costarr, *tan2arr;  /* ntg; tan2arr=-1 flags the equator */
    const double *phiarr, *aarr, *barr;         /* npg; aarr=sin(phi), barr=-cos(phi) */
    const double *vcell;                        /* cells, cm^3 */
    /* medium: one thermal species per cell (select_dust_2.f90:22-34) */
    const double *rho;                          /* cells, g/cm^3 */
    const uint8_t *species;                     /* cells: 0 = empty, else 1..10 */
    int32_t species_opac[ORC_NSPECIES + 1];     /* findopac(.,.,.,id) -> opacity index 0..nopac-1 */
    /* dust tables (dust_5.f90:384-455, dust_mod_5.f90) */
    int32_t nopac, nlam;
    const double *nudust, *kappad, *adust, *gdust, *pldust;   /* [iop*nlam + i], nudust descending (eV) */
    const double *kappav;                                      /* nopac, cm^2/g */
    int32_t nT, nnu;
    const double *lTint, *Tint, *nuint;                        /* nT, nT, nnu */
    const double *kapint;                                      /* [iop*nT*nnu + j*nT + i] = kapint(i,j,iop) */
    const double *kappap;                                      /* [iop*nT + i] */
    /* source: limb=0 blackbody star (initpspot.f90, initp_common.f90) */
    double rstar;      /* R_sun */
    double tstar;      /* K/11605 */
    double rmin;       /* stellar radius in R* (=rarr(1)) */
    /* observer (tts_mod_5.f90, setup.f90:36-41,115-130) */
    int32_t ipeel, npeel;
    const double *coste, *sinte, *cospe, *sinpe, *phie;        /* npeel */
    int32_t nx;        /* image pixels per side */
    double rmaxi;      /* image half-width, R* */
    int32_t nband, nwmax;
    const int32_t *band_nw;                     /* nband: lines in the filter table */
    const double *band_freq, *band_phi;         /* [b*nwmax + k], freq ascending (Hz) */
    int32_t nfreq, nmu, nph, nap;
    const double *aperture2;                    /* nap, R*^2 */
    const double *u_mu, *sini_mu;               /* nmu */
    double numin, lnurat;                       /* eV, ln(numax/numin) */
    int32_t iintmax;                            /* 10000, propagate_5_fuv.f90:32 */
    double taumax;                              /* 25, peeloff_3d_2_test.f90:52 */
    /* external illumination (initpout.f90, rep_isrf.f90, isrf_mod.f90):
     * packet i of a run (0-based) enters from outside when i >= isrf_from;
     * isrf_from < 0 turns it off */
    int32_t nisrf;
    const double *isrf_cdf, *isrf_logcdf, *isrf_logfreq;   /* nisrf */
    int64_t isrf_from;
    /* stellar hot spots (initpspot.f90 with ispot = 1, spotset.f90): packet
     * i of a run is a hot-spot photon when acc_to <= i < hot_to */
    int32_t ispot, nspot;
    double fspot, csprad, tshock;
    double scspot1, ssspot1, cspot1, sspot1, scspot2, ssspot2, cspot2, sspot2;
    int64_t hot_to;
    /* disk-accretion photons (initpacc_4.f90): packet i of a run is one when
     * i < acc_to; acc_ldint / acc_hdisk: Ldint and hdisk of nrg - 1 radial
     * cells (grid_mod) */
    int64_t acc_to;
    const double *acc_ldint, *acc_hdisk;
    /* diffusion approximation (setdiffus_5.f90, emit_common_2.f90:88-378):
     * diff_itmin[ir-1] = first diffusion cell of column ir from the pole (0
     * none; cells itmin..ntg-itmin diffuse), NULL: off; diff_delta = 4 pi
     * Tstar^4 / ns; diff_count: photons re-emitted per surface cell (shared
     * by all threads, written) */
    const int32_t *diff_itmin;
    double diff_delta;
    uint32_t *diff_count;
} orc_model;

/* Tallies: every pointer is caller-owned, zeroed by the caller, accumulated here.
 * Layouts are the reference's column-major arrays flattened:
 *   sed   [10][ORC_NO][nap][nph][nmu][nfreq]: si,sq,su,sv,si2,sq2,su2,sv2,nums,aveinc (propagate_5_fuv.f90:300-313)
 *   psed  [9][ORC_NO][nap][npeel][nfreq]:     ti,tq,tu,tv,ti2,tq2,tu2,tv2,numt (peeloff_3d_2_test.f90:327-338)
 *   image [8][nband][nx][nx][npeel]:          image_b_{i,q,u,v,d,s,e,o}(peelid,ix,iy,ift) (peeloff:265-277)
 *   cells: dtauabs (double), nabs, dbig, killed (int32) */
typedef struct {
    double *dtauabs;
    int32_t *nabs, *dbig, *killed;
    float *sed, *psed, *image;
    double counters[16];
} orc_tallies;

enum {  /* counters[] slots */
    ORC_C_TOT = 0, ORC_C_FLUX, ORC_C_SFLUX, ORC_C_AFLUX, ORC_C_ABFLUX, ORC_C_SCOUNT,
    ORC_C_NSCAT, ORC_C_HIT, ORC_C_HTOT, ORC_C_ESCAPED, ORC_C_KILLED, ORC_C_SEGMENTS,
    ORC_C_PEELSEG, ORC_C_PEELEVT, ORC_C_RESETS, ORC_C_ERRORS
};

/* ---- random numbers ---------------------------------------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* draw n of packet p for a run seeded with `seed`: one 32-bit word */
uint32_t orc_packet_word(uint64_t seed, uint64_t packet, uint32_t n);
/* L'Ecuyer & Cote combined MLCG (ecuyer_cote_mod.f90), 32 split generators */
void orc_lecuyer_set_all_seeds(int32_t s1, int32_t s2);
void orc_lecuyer_set_generator(int32_t g);
double orc_lecuyer_ran(void);

/* ---- deterministic pieces --------------------------------------------- */
typedef struct { double tn1, tp1, tp2; int32_t in1, ip1, ip2; } orc_walls;
void orc_reset_t(orc_walls *w);
void orc_insert_t(orc_walls *w, double t, int32_t i);
int  orc_find_next_wall(const orc_walls *w, int on_wall, double *t, int32_t *i);
void orc_radface(orc_walls *w, double ux, double uy, double uz, double x1, double y1, double z1, double R2_1, double R2_2);
void orc_thetaface(orc_walls *w, double ux, double uy, double uz, double x1, double y1, double z1, double t1, double t2);
void orc_phiface(orc_walls *w, double A1, double B1, double A2, double B2, double ux, double uy, double uz, double x1, double y1, double z1);
int32_t orc_locate(const double *xx, int32_t n, double x);
void orc_findangle3(double cost1, double sint1, double phi1, double cosp1, double sinp1,
                    double cost2, double sint2, double phi2, double cosp2, double sinp2,
                    double *costnew, double *sintnew, double *phinew, double *ri1);
void orc_dustmat(double g, double pl, double cost, double cost2, double p[4]);
/* stokes.f90 after the rejection loop accepted (bmu, ri1): updates direction and Stokes */
void orc_stokes_apply(double g, double pl, double bmu, double ri1,
                      double dir[5] /* cost,sint,phi,cosp,sinp in/out */, double s[4] /* in/out */);
/* stokespeel.f90 for fixed (bmu, ri1); cost/sint = observer, costp/sintp = incident */
void orc_stokespeel(double g, double pl, double s[4], double cost, double sint,
                    double ri1, double bmu, double costp, double sintp);
double orc_lucydustfreq(const orc_model *m, double T0, int32_t iop, double xi);
double orc_bbfreq_lecuyer(double teff);   /* bbfreq with the L'Ecuyer stream (random.f90:44-72) */
double orc_dusttemp(const orc_model *m, double na, double nphot, int32_t iop, double mcell);
/* tfinal with maxii=1 and no velocities: Lucy update, midplane mirror, phi replication */
void orc_tfinal(const orc_model *m, const double *dtauabs, double nphot, double *tdave);
/* setdiffus_5.f90:36-103: the diffusion flags of every cell (diffus,
 * cells, 1 = diffusion) and diffdir (-2 upper half, +2 lower half), and
 * itmin per radial column; tdave: the last Lucy temperatures; kappar
 * [iop*nT + i] the Rosseland means; hdisk (nrg-1) R*; rmaxd R*. */
void orc_setdiffus(const orc_model *m, const double *kappar, const double *hdisk, const double *tdave,
                   double rmaxd, int32_t nr0, int32_t iter, uint8_t *diffus, int8_t *diffdir, int32_t *itmin);
/* emit_common_2.f90:256-341 after a run: the temperatures of every diffusion
 * column from its surface cell's photon count and the accretion photons
 * (nabs) below it; writes tdust in the diffusion cells. */
void orc_diff_columns(const orc_model *m, const double *kappar, const uint32_t *count, const int32_t *nabs,
                      double delta, double *tdust);
/* tfinal with the dust temperatures of the run (a diffusion cell keeps its
 * own, tfinal_5_test1.f90:232-233; tdust may be NULL without diffusion) */
void orc_tfinal2(const orc_model *m, const double *dtauabs, double nphot, const double *tdust, double *tdave);
/* tfinal_5_test1.f90:334-368 convergence statistic (sums in out[3]) */
double orc_lucy_sigma(const orc_model *m, const double *told, const double *tnew, double out[3]);

/* Deterministic fixed-direction walk (tauint without interaction tallies):
 * walks from (x,y,z) in cell ijk along u until tau_stop is reached, the packet
 * escapes (ir==nrg) or hits the star (ir==0).  Records each crossed face code
 * (1..6) and cell after the crossing.  Returns the number of crossings, or -1
 * when no wall was found.  status: 0 stop inside, 1 escaped, 2 hit star, 3 max steps.
 * testgrid re-placements draw from Philox stream (seed 0, `stream`). */
int32_t orc_trace_ray(const orc_model *m, double pos[3], const double u[3], int32_t ijk[3], int on_wall,
                      double tau_stop, int32_t max_steps, int32_t *faces, int32_t *cells,
                      double *tau_out, int32_t *status, uint64_t stream);

/* Whole packets: newdisktherm_5_fuv.f90:220-491 (stellar photons) + propagate.
 * rng_mode 0 = counter-based per-packet stream (seed, first_packet+i);
 * rng_mode 1 = one L'Ecuyer stream (already seeded) consumed in packet order.
 * Tdust: per-cell dust temperature used for re-emission (K/11605).
 * nthreads > 1 runs packet ranges on threads with private tallies (rng_mode 0 only).
 * trace (optional, 8 doubles per packet): final nub, cost, phi, sip, sqp, sup,
 * number of interactions, exit code. */
int32_t orc_run_packets(const orc_model *m, const double *tdust, orc_tallies *t,
                        uint64_t first_packet, int64_t npackets, uint64_t seed,
                        int32_t rng_mode, int32_t nthreads, double *trace);

/* ---- emissivity tables (setemiss, dust_5.f90:507-699) ------------------- */
double orc_expm1_as(double x);                 /* mathsubs.f90:79-113 */
double orc_bnu(double nu, double T);           /* mathsubs.f90:4-22 */
double orc_dbnu(double nu, double T);          /* mathsubs.f90:29-47 */
/* one interval: qsimp (nr_2.f:209-231) of kappaB (kind 0) or chidB (kind 1);
 * returns 1 when qsimp did not converge */
int32_t orc_qsimp_seg(int32_t kind, double k0, double k1, double nu0, double dnu, double teff, double *out);
int64_t orc_setemiss(int32_t nopac, int32_t nlam, const int32_t *nlambda, const double *nudust,
                     const double *kappad, const double *adust, int32_t nT, const double *Tint,
                     int32_t nnu, const double *nuint, const int32_t *smooth,
                     double *kapint, double *kappap, double *kappar);

/* ---- model building (gridset cell loop, densenv, zerod) ------------------ */
int32_t orc_zerod(double xmu, double z, double *xmu0);               /* zerod.f90 */
double orc_densenv(double rad, double costin, double sina, double rc, double rhoe0, double rmine, double rmax,
                   double rhoamb, int32_t *flag);                     /* densenv.f90 core */
double orc_cell_volume(double r1, double r2, double t1, double t2, double p1, double p2, double rmincgs);

#ifdef __cplusplus
}
#endif
#endif
