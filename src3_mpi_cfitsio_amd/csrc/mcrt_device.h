// mcrt_device.h -- device-side photon-packet transport for CDNA4 (gfx950).
//
// Templated on the transport real type T: T=double is the FP64 path
// (compiled with -ffp-contract=off so it reproduces the reference arithmetic
// bit for bit), T=float is the mixed-precision path: FP32 opacities, optical
// depths, sampling, Stokes algebra and peel weights; FP64 packet geometry
// (position, direction, all face roots) and FP64 absorbed-energy tally.
//
// Reference routines restated here (paths relative to src3_mpi_cfitsio/):
//   closest_wall.f90, radface.f90, thetaface.f90, phiface.f90, testgrid.f90,
//   tauint_3d_2_fuv.f90, propagate_5_fuv.f90, peeloff_3d_2_test.f90,
//   stokes.f90, dustmat.f90, stokespeel.f90, findangle3.f90,
//   initpspot.f90, initp_common.f90, initpabs.f90, reemit.f90,
//   lucydustfreq_2_test.f90, random.f90 (bbfreq), dust_5.f90 (opacset),
//   newdisktherm_5_fuv.f90 (stellar packet set-up and direct peel).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>
#include <type_traits>

namespace mcrt {

constexpr int kMaxOpac = 4;     // opacity tables resident per packet
constexpr int kGuideBuckets = 128;   // buckets of the re-emission CDF guide table
constexpr int kNo = 8;          // SED components (tts_mod_5.f90:19)
constexpr double kPi = 3.1415926535897932384626433;
constexpr double kR2p = 2.0 * 3.1415926535897932384626433;
constexpr double kRsol = 6.9598e10;
constexpr double kTmaxK = 199999.0;

// Counters accumulated per launch (atomic, 64-bit).  The first C_NPUBLIC are
// mcrt_get_tallies' counters[16]; the error kinds after them (each error also
// counts in C_ERRORS) go to mcrt_run_stats.
enum Counter {
  C_TOT = 0, C_FLUX, C_SFLUX, C_AFLUX, C_ABFLUX, C_SCOUNT, C_NSCAT, C_HIT, C_HTOT,
  C_ESCAPED, C_KILLED, C_SEGMENTS, C_PEELSEG, C_PEELEVT, C_RESETS, C_ERRORS,
  C_ERR_WALK,      // a transport flight found no wall, or passed the segment guard
  C_ERR_PEEL,      // a peel-off walk did
  C_ERR_OTHER,     // an emission placed nowhere, a full peel-off queue
  C_NCOUNTERS
};
constexpr int C_NPUBLIC = C_ERRORS + 1;
static_assert(C_NPUBLIC == 16, "mcrt_get_tallies returns 16 counters");

struct DeviceModel {
  // grid walls, 1-based reference semantics stored 0-based (wall k at [k-1])
  int nrg, ntg, npg;
  const double *rarr, *r2arr, *thetarr, *costarr, *tan2arr, *phiarr, *aarr, *barr;
  const double *ccone, *scone;   // cos, sin of min(theta, pi - theta) per theta wall
  double phi_tol2;               // sin^2 of half testgrid's tolerance on the narrowest phi wedge
  const double* vcell;           // cm^3 per cell
  // medium: one thermal species per cell
  const double* rho64;
  const float* rho32;
  const uint8_t* species;
  const int* species_opac;        // [11] in device memory (a runtime-indexed array
                                 // inside the kernel argument would be copied to scratch)
  // dust
  int nopac, nlam;
  const double *nudust, *kappad, *adust, *gdust, *pldust;
  // log-uniform nudust of table id: log(nudust[0]) and 1/step (0: no guess)
  double nud_log0[kMaxOpac], nud_ilog[kMaxOpac];
  double kappav[kMaxOpac];
  int nT, nnu;
  const double *lTint, *Tint, *nuint, *kappap;
  // re-emission CDFs, kapint[(iop*nT + i)*nnu + j] (frequency contiguous), and
  // the same values rounded to FP32 for the mixed path
  const double* kapint;
  const float* kapint32;
  // guide table of the re-emission CDFs: kguide[(iop*nT + i)*(kguide_k+1) + k]
  // = the number of frequency entries of temperature row i with
  // kapint <= k/kguide_k (null: none)
  const uint16_t* kguide;
  int kguide_k;
  const double* tdust;
  // source
  double rstar, tstar, rmin;
  // observer
  int ipeel, npeel;
  const double* peel;            // [npeel][5]: coste, sinte, cospe, sinpe, phie
  int nx;
  double rmaxi;
  int nband, nwmax;
  const int* band_nw;
  const double *band_freq, *band_phi;
  int nfreq, nmu, nph, nap;
  const double *aperture2, *u_mu, *sini_mu;
  double numin, lnurat;
  int iintmax;
  double taumax;
  long long max_segments;        // per walk; guards against zero-length cycling
  // external illumination (initpout.f90, rep_isrf.f90): packet k of a run
  // (0-based) enters from outside when nisrf > 0 and k >= isrf_from
  int nisrf;
  const double *isrf_cdf, *isrf_logcdf, *isrf_logfreq;
  unsigned long long isrf_from;
  // stellar hot spots (initpspot.f90 with ispot = 1, spotset.f90): nspot
  // caps of cos radius csprad centred on (scspot, ssspot, cspot) unit
  // vectors; stellar photons emitted inside a cap (or all, fspot = 1) have
  // the shock temperature; packets acc_to <= k < hot_to of a run are
  // hot-spot photons
  int ispot, nspot;
  double fspot, csprad, tshock;
  double scspot1, ssspot1, cspot1, sspot1, scspot2, ssspot2, cspot2, sspot2;
  unsigned long long hot_to;
  // disk-accretion photons (initpacc_4.f90): packets k < acc_to of a run;
  // acc_ldint[nrg-1] is the cumulative accretion luminosity by radial cell
  // (0 first, 1 last), acc_hdisk[nrg-1] the disk scale height (R*)
  unsigned long long acc_to;
  const double *acc_ldint, *acc_hdisk;
  const struct AccStart* acc_start;   // [acc_to], written by k_acc_place
  // diffusion approximation (setdiffus_5.f90, emit_common_2.f90:88-378):
  // diff_itmin[ir-1] is the first diffusion cell of radial column ir counted
  // from the pole (0: none); cells diff_itmin..ntg-diff_itmin of the column
  // diffuse, the upper half toward the pole above, the lower half toward the
  // pole below.  Null: diffusion off.  diff_delta = 4 pi Tstar^4 / ns, the
  // surface temperature increment per photon times the surface area.
  const int32_t* diff_itmin;
  double diff_delta;
  // set per run by the host: the run's LDS plan stages the five dust tables
  // (the kernels for plain models then address them in LDS only)
  int dust_lds;
};

// where k_acc_place put accretion photon k: position, cell, species and the
// packet's next random draw (sp = -1: no cell with dust was found)
struct AccStart {
  double xp, yp, zp, rsq;
  int32_t ir, it, ip, sp;
  uint32_t n, pad;
};

constexpr int kAccTries = 1 << 16;   // position draws before an accretion photon is dropped

struct DeviceTallies {
  double* dtauabs;
  int *nabs, *dbig, *killed;
  float *sed, *psed, *image;
  unsigned long long* counters;  // C_NCOUNTERS
  double* trace;                 // optional: 8 doubles per packet
  unsigned* diff_count;          // photons re-emitted from each diffusion surface cell (diffusion on)
};

// With at most 1023 walls per axis (mcrt_set_grid) every factor is below
// 2^24 and the index below 2^30, so 24-bit multiplies give it exactly.
__device__ __forceinline__ size_t cell_index(const DeviceModel& m, int ir, int it, int ip) {
  const unsigned inner = (unsigned)(it - 1) + __umul24((unsigned)(m.ntg - 1), (unsigned)(ip - 1));
  return (size_t)((unsigned)(ir - 1) + __umul24((unsigned)(m.nrg - 1), inner));
}

// cell (ir, it) lies in the diffusion region of its column
__device__ __forceinline__ bool diffusing(const DeviceModel& m, int ir, int it) {
  if (!m.diff_itmin) return false;
  const int a = m.diff_itmin[ir - 1];
  return a > 0 && it >= a && it <= m.ntg - a;
}

// ---------------------------------------------------------------------------
// Counter-based RNG: Philox4x32-10, key = run seed, counter = (n/4, packet).
// ---------------------------------------------------------------------------
// coop: every lane of the wavefront carries the same stream (k_tail).  The
// lanes then hold 64 consecutive blocks, lane l block cb + l, computed in
// one pass, and a draw reads its word from the lane that holds it: one
// Philox evaluation per 256 draws on the packet's serial chain instead of
// one per 4.  The words are those of the serial stream.
struct Rng {
  uint32_t k0, k1, p0, p1, n;
  uint32_t b0, b1, b2, b3;      // current Philox block (scalars: no indexed private memory)
  bool coop = false;
  uint32_t cb = 0x80000000u;    // coop: block held by lane 0 (0x80000000: none yet)
  uint32_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;   // coop: this lane's block
};

__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                       uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // one 32x32 -> 64-bit product gives both halves (v_mad_u64_u32)
    const uint64_t m0 = (uint64_t)0xD2511F53u * c0, m1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(m0 >> 32), lo0 = (uint32_t)m0;
    uint32_t hi1 = (uint32_t)(m1 >> 32), lo1 = (uint32_t)m1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

__device__ __forceinline__ void rng_init(Rng& r, uint64_t seed, uint64_t packet) {
  r.k0 = (uint32_t)seed; r.k1 = (uint32_t)(seed >> 32);
  r.p0 = (uint32_t)packet; r.p1 = (uint32_t)(packet >> 32);
  r.n = 0;
  r.cb = 0x80000000u;
}

// Move a stream (keys kept) to draw n: the block holding draw n is
// regenerated when n is not at a block boundary; rng_word regenerates it
// otherwise.
__device__ __forceinline__ void rng_seek(Rng& r, uint32_t n) {
  r.n = n;
  if (!r.coop && (n & 3u)) {
    uint32_t o[4];
    philox(n >> 2, r.p0, r.p1, 0x4D435254u, r.k0, r.k1, o);
    r.b0 = o[0]; r.b1 = o[1]; r.b2 = o[2]; r.b3 = o[3];
  }
}

// Continue a stream at draw n
__device__ __forceinline__ void rng_resume(Rng& r, uint64_t seed, uint64_t packet, uint32_t n) {
  rng_init(r, seed, packet);
  rng_seek(r, n);
}

// coop: the lanes take blocks blk .. blk + 63 of the stream
__device__ __forceinline__ void rng_fill(Rng& r, uint32_t blk) {
  uint32_t o[4];
  philox(blk + (uint32_t)__lane_id(), r.p0, r.p1, 0x4D435254u, r.k0, r.k1, o);
  r.cb = blk;
  r.w0 = o[0]; r.w1 = o[1]; r.w2 = o[2]; r.w3 = o[3];
}

// coop: word p of the stream, p held by the lanes (cb*4 <= p < cb*4 + 256),
// read by every lane for its own p
__device__ __forceinline__ uint32_t rng_word_at(const Rng& r, uint32_t p) {
  const int s = (int)((p >> 2) - r.cb);
  const uint32_t x0 = (uint32_t)__shfl((int)r.w0, s, 64), x1 = (uint32_t)__shfl((int)r.w1, s, 64);
  const uint32_t x2 = (uint32_t)__shfl((int)r.w2, s, 64), x3 = (uint32_t)__shfl((int)r.w3, s, 64);
  const uint32_t c = p & 3u;
  return (c == 0u) ? x0 : (c == 1u) ? x1 : (c == 2u) ? x2 : x3;
}

__device__ __forceinline__ uint32_t rng_word(Rng& r) {
  const uint32_t j = r.n & 3u;
  if (r.coop) {
    const uint32_t blk = r.n >> 2;
    if (blk - r.cb >= 64u) rng_fill(r, blk);
    // (the words pass an empty asm, so the select stays one between values:
    // a select between fields is a load from a selected address, which
    // would put the whole state into private memory)
    uint32_t a0 = r.w0, a1 = r.w1, a2 = r.w2, a3 = r.w3;
    asm volatile("" : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3));
    const uint32_t v = (j == 0u) ? a0 : (j == 1u) ? a1 : (j == 2u) ? a2 : a3;
    r.n++;
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)(blk - r.cb));
  }
  if (j == 0u) {
    uint32_t o[4];
    philox(r.n >> 2, r.p0, r.p1, 0x4D435254u, r.k0, r.k1, o);
    r.b0 = o[0]; r.b1 = o[1]; r.b2 = o[2]; r.b3 = o[3];
  }
  uint32_t w = (j == 0u) ? r.b0 : (j == 1u) ? r.b1 : (j == 2u) ? r.b2 : r.b3;
  r.n++;
  return w;
}

// a draw in (0, 1) from one word of the stream
template <class T> __device__ __forceinline__ T u01(uint32_t w);
template <> __device__ __forceinline__ double u01<double>(uint32_t w) {
  return ((double)w + 0.5) * 2.3283064365386963e-10;
}
// 23 bits: (k + 1/2) 2^-23 is exact in FP32 for every k < 2^23, so the
// draw is strictly inside (0, 1).  (With 24 bits, k + 1/2 is not
// representable for k >= 2^23: it rounds to even and k = 2^24 - 1 gives 1.)
template <> __device__ __forceinline__ float u01<float>(uint32_t w) {
  return ((float)(w >> 9) + 0.5f) * 1.1920928955078125e-7f;
}
template <class T> __device__ __forceinline__ T uniform(Rng& r) { return u01<T>(rng_word(r)); }

// ---------------------------------------------------------------------------
// Real-type helpers
// ---------------------------------------------------------------------------
template <class T> struct RT;
template <> struct RT<double> {
  static constexpr double huge = DBL_MAX;
  __device__ static double sqrt_(double x) { return sqrt(x); }
  __device__ static double sin_(double x) { return sin(x); }
  __device__ static double cos_(double x) { return cos(x); }
  __device__ static double acos_(double x) { return acos(x); }
  __device__ static double exp_(double x) { return exp(x); }
  __device__ static double log_(double x) { return log(x); }
  __device__ static double atan2_(double y, double x) { return atan2(y, x); }
  __device__ static double pow_(double x, double y) { return pow(x, y); }
  __device__ static double fabs_(double x) { return fabs(x); }
  __device__ static double log_acc(double x) { return log(x); }
  __device__ static double tau_of(double u) { return -log(u); }
};
// Mixed path: the hardware transcendentals (v_sin/v_cos on revolutions,
// v_exp/v_log base 2; a few ulp over the ranges used here: angles in
// [-2pi, 2pi], exponents of optical depths up to 25) in place of the
// range-reduced library versions.  kernels_mixed.hip is also built with
// -fno-hip-fp32-correctly-rounded-divide-sqrt, so FP32 division and sqrt use
// v_rcp/v_sqrt.  The error this costs against the FP64 path is measured on
// the SEDs, images and temperatures.
template <> struct RT<float> {
  static constexpr float huge = FLT_MAX;
  __device__ static float sqrt_(float x) { return sqrtf(x); }
  __device__ static float sin_(float x) { return __sinf(x); }
  __device__ static float cos_(float x) { return __cosf(x); }
  __device__ static float acos_(float x) { return acosf(x); }
  __device__ static float exp_(float x) { return __expf(x); }
  __device__ static float log_(float x) { return __logf(x); }
  __device__ static float atan2_(float y, float x) { return atan2f(y, x); }
  __device__ static float pow_(float x, float y) { return __powf(x, y); }
  __device__ static float fabs_(float x) { return fabsf(x); }
  // the library log where a ratio near 1 matters: v_log_f32's absolute
  // error (~2^-22) can turn log(1 + 1e-7) into 0
  __device__ static float log_acc(float x) { return logf(x); }
  // optical depth of a flight, -log(u): the hardware log may return a tiny
  // positive value for u just below 1, so the result is clamped at 0
  __device__ static float tau_of(float u) { return fmaxf(-__logf(u), 0.0f); }
};

template <class T> __device__ __forceinline__ T cell_rho(const DeviceModel& m, size_t c);
template <> __device__ __forceinline__ double cell_rho<double>(const DeviceModel& m, size_t c) { return m.rho64[c]; }
template <> __device__ __forceinline__ float cell_rho<float>(const DeviceModel& m, size_t c) { return m.rho32[c]; }

// ---------------------------------------------------------------------------
// closest_wall.f90:13-66
// ---------------------------------------------------------------------------
struct Walls {
  double tn1, tp1, tp2;
  int in1, ip1, ip2;
  __device__ __forceinline__ void reset() {
    tn1 = -DBL_MAX; tp1 = DBL_MAX; tp2 = DBL_MAX;
    in1 = 0; ip1 = 0; ip2 = 0;
  }
  __device__ __forceinline__ void insert(double t, int i) {
    if (t < 0.0) {
      if (t > tn1) { tn1 = t; in1 = i; }
    } else {
      if (t < tp1) { tp2 = tp1; ip2 = ip1; tp1 = t; ip1 = i; }
      else if (t < tp2) { tp2 = t; ip2 = i; }
    }
  }
  // returns false when no wall exists ("ERROR - no valid solution")
  __device__ __forceinline__ bool next(bool on_wall, double& t, int& i) const {
    if (on_wall) {
      if (fabs(tn1) < fabs(tp1)) { t = tp1; i = ip1; }
      else { t = tp2; i = ip2; }
    } else { t = tp1; i = ip1; }
    return !(t == DBL_MAX);
  }
};

// radface.f90:44-62.  r spans 1..1e5 R*, so |x|^2 - R^2 on a shell cancels
// catastrophically below FP64.
__device__ __forceinline__ void radface(Walls& w, double ux, double uy, double uz, double x1, double y1, double z1,
                                        double R2_1, double R2_2) {
  double bb = 2.0 * (x1 * ux + y1 * uy + uz * z1);
  double cc = x1 * x1 + y1 * y1 + z1 * z1 - R2_1;
  double descr = bb * bb - 4.0 * cc;
  if (descr > 0.0) {
    descr = sqrt(descr);
    w.insert((-bb + descr) * 0.5, 1);
    w.insert((-bb - descr) * 0.5, 1);
  }
  cc = x1 * x1 + y1 * y1 + z1 * z1 - R2_2;
  descr = bb * bb - 4.0 * cc;
  if (descr > 0.0) {
    descr = sqrt(descr);
    w.insert((-bb + descr) * 0.5, 2);
    w.insert((-bb - descr) * 0.5, 2);
  }
}

// thetaface.f90:59-115 (equator flagged by tan2 = -1, poles by tan2 = 0).
// On a cone wall x^2+y^2 - tan^2 z^2 cancels like the radial quadratic.
__device__ __forceinline__ void thetaface(Walls& w, double ux, double uy, double uz, double x1, double y1, double z1,
                                          double t1, double t2) {
  double tmp1 = x1 * ux + y1 * uy, tmp2 = ux * ux + uy * uy, tmp3 = x1 * x1 + y1 * y1;
  double ta, tb, aa, bb, cc, descr;
  int i1, i2;
  if (t2 < t1) { ta = t2; tb = t1; i1 = 4; i2 = 3; }
  else { ta = t1; tb = t2; i1 = 3; i2 = 4; }
  if (ta < 0.0) {
    if (uz != 0.0) w.insert(-z1 / uz, i1);
  } else {
    bb = 2.0 * (tmp1 - ta * uz * z1);
    aa = tmp2 - ta * uz * uz;
    cc = tmp3 - z1 * z1 * ta;
    if (fabs(aa) > 1.0e-10) {
      descr = bb * bb - 4.0 * aa * cc;
      if (descr > 0.0 && ta != 0.0) {
        descr = sqrt(descr);
        w.insert((-bb + descr) * 0.5 / aa, i1);
        w.insert((-bb - descr) * 0.5 / aa, i1);
      }
    } else {
      if (fabs(bb) > 1.0e-10) w.insert(-cc / bb, i1);
    }
  }
  bb = 2.0 * (tmp1 - tb * uz * z1);
  aa = tmp2 - tb * uz * uz;
  cc = tmp3 - z1 * z1 * tb;
  if (fabs(aa) > 1.0e-10) {
    descr = bb * bb - 4.0 * aa * cc;
    if (descr > 0.0 && tb != 0.0) {
      descr = sqrt(descr);
      w.insert((-bb + descr) * 0.5 / aa, i2);
      w.insert((-bb - descr) * 0.5 / aa, i2);
    }
  } else {
    if (fabs(bb) > 1.0e-10) w.insert(-cc / bb, i2);
  }
}

// phiface.f90:48-56 with C = D = 0 (tauint_3d_2_fuv.f90:47)
__device__ __forceinline__ void phiface(Walls& w, double A1, double B1, double A2, double B2, double ux, double uy,
                                        double uz, double x1, double y1, double z1) {
  double denom1 = (A1 * ux + B1 * uy + 0.0 * uz);
  if (denom1 != 0.0) w.insert(-(A1 * x1 + B1 * y1 + 0.0 * z1 + 0.0) / denom1, 5);
  double denom2 = (A2 * ux + B2 * uy + 0.0 * uz);
  if (denom2 != 0.0) w.insert(-(A2 * x1 + B2 * y1 + 0.0 * z1 + 0.0) / denom2, 6);
}

// locate, nr_2.f:131-154
template <class T>
__device__ __forceinline__ int locate(const T* xx, int n, T x) {
  int jl = 0, ju = n + 1, jm;
  bool ascnd = xx[n - 1] > xx[0];
  while (ju - jl > 1) {
    jm = (ju + jl) / 2;
    if (ascnd == (x > xx[jm - 1])) jl = jm; else ju = jm;
  }
  if (jl == 0 && x == xx[0]) jl = 1;
  return jl;
}

template <class T> __device__ __forceinline__ T check1(T x) {  // check.f90
  if (x > (T)1) return (T)1;
  if (x < (T)-1) return (T)-1;
  return x;
}

// ---------------------------------------------------------------------------
// Packet state
// ---------------------------------------------------------------------------
template <class T> struct Opac { T kappa, albedo, g, pl; };

// Packet geometry (position, direction of flight, face distances) is FP64 in
// both precisions: at r ~ 1e5 R* an FP32 position carries ~1e-2 R* of
// rounding, enough for the on-wall tie rule of closest_wall.f90:52-59 to
// pick the wrong root near cell corners.  T is the type of everything else.
typedef double G;

template <class T> struct Pkt {
  G xp, yp, zp, rsq, rtot;
  G ux, uy, uz;
  T cost, sint, phi, cosp, sinp;
  T sip, sqp, sup, svp;
  T nub, wave;
  T cosb, sinb, lp;
  int ir, it, ip;
  bool on_wall;
  int iflag, i_orig, exitflag;
  T kapd[kMaxOpac];
  Opac<T> op[kMaxOpac];
};

// Runtime-indexed reads of the per-packet opacity arrays through an unrolled
// select chain, so the arrays stay in VGPRs instead of private memory.  Each
// element passes an empty asm first: without it the optimizer folds the
// chain of selects between loads into one load at a selected address, which
// forces the arrays into scratch (one private-memory load per use, on every
// walk segment).
__device__ __forceinline__ void in_vgpr(float& x) { asm volatile("" : "+v"(x)); }
__device__ __forceinline__ void in_vgpr(double& x) { asm volatile("" : "+v"(x)); }
template <class T>
__device__ __forceinline__ void in_vgpr(Opac<T>& o) { in_vgpr(o.kappa); in_vgpr(o.albedo); in_vgpr(o.g); in_vgpr(o.pl); }

template <class V>
__device__ __forceinline__ V pick(const V (&a)[kMaxOpac], int i) {
  V v = a[0];
  in_vgpr(v);
#pragma unroll
  for (int k = 1; k < kMaxOpac; ++k) {
    V x = a[k];
    in_vgpr(x);
    v = (i == k) ? x : v;
  }
  return v;
}
// The same, field by field: a select between two structures is a copy from
// a selected address, which would put both into private memory.
template <class T>
__device__ __forceinline__ Opac<T> pick(const Opac<T> (&a)[kMaxOpac], int i) {
  T kap[kMaxOpac], alb[kMaxOpac], g[kMaxOpac], pl[kMaxOpac];
#pragma unroll
  for (int k = 0; k < kMaxOpac; ++k) { kap[k] = a[k].kappa; alb[k] = a[k].albedo; g[k] = a[k].g; pl[k] = a[k].pl; }
  Opac<T> v;
  v.kappa = pick(kap, i);
  v.albedo = pick(alb, i);
  v.g = pick(g, i);
  v.pl = pick(pl, i);
  return v;
}

// Event counters of a workgroup, held in LDS (no registers per lane) and
// flushed to the device counters once per launch.
struct LocalCounts {
  unsigned long long* s;    // C_NCOUNTERS LDS slots of this workgroup
  // one: every lane of the wavefront carries the same packet (k_tail), which
  // counts once
  bool one = false;
  __device__ __forceinline__ void add(int i, unsigned long long v = 1ull) const {
    if (one && __lane_id() != 0) return;
    atomicAdd(&s[i], v);
  }
  // +1 from every active lane, as one atomic per wavefront: the per-segment
  // counters would otherwise serialize 64 same-address LDS atomics per step
  __device__ __forceinline__ void add_lanes(int i) const {
    if (one) { if (__lane_id() == 0) atomicAdd(&s[i], 1ull); return; }
    const unsigned long long mask = __ballot(1);
    if (__lane_id() == __ffsll((long long)mask) - 1) atomicAdd(&s[i], (unsigned long long)__popcll(mask));
  }
};

// Direction of flight from (cost, sint, cosp, sinp).  FP64 path: the
// reference's products.  Mixed path: FP32 angles give |u| = 1 only to ~1e-7
// while the face roots assume |u| = 1 (over a 1e4 R* step that error moves a
// crossing 1e-3 R* off its wall and trips testgrid), so u is renormalized in
// FP64.
template <class T> __device__ __forceinline__ void set_dir(Pkt<T>& p);
template <> __device__ __forceinline__ void set_dir<double>(Pkt<double>& p) {
  p.ux = p.sint * p.cosp; p.uy = p.sint * p.sinp; p.uz = p.cost;
}
template <> __device__ __forceinline__ void set_dir<float>(Pkt<float>& p) {
  const double ux = (double)p.sint * (double)p.cosp, uy = (double)p.sint * (double)p.sinp, uz = (double)p.cost;
  const double s = rsqrt(ux * ux + uy * uy + uz * uz);
  p.ux = ux * s; p.uy = uy * s; p.uz = uz * s;
}

// A lower bound on the distance from the packet to every wall of its cell,
// less a margin that covers the rounding of the face roots (the quadratics of
// radface/thetaface lose at most ~sqrt(eps)*r on grazing rays).  Spheres:
// r - r_in, r_out - r.  Cones: the quadratic of thetaface is the double cone
// of half-angle t = min(theta, pi - theta), at distance |rho cos t - |z| sin t|
// (the equator plane and the pole axis are its t = pi/2 and t = 0 cases).
// Phi walls: the full planes through the z axis, at |sin(phi) x - cos(phi) y|.
template <class T>
__device__ __forceinline__ double wall_clearance(const DeviceModel& m, const Pkt<T>& p) {
  const double r = p.rtot;
  double d = fmin(r - m.rarr[p.ir - 1], m.rarr[p.ir] - r);
  if (m.ntg > 2) {
    const double rho = sqrt(p.xp * p.xp + p.yp * p.yp), az = fabs(p.zp);
    d = fmin(d, fabs(rho * m.ccone[p.it - 1] - az * m.scone[p.it - 1]));
    d = fmin(d, fabs(rho * m.ccone[p.it] - az * m.scone[p.it]));
    if (m.npg > 2) {
      d = fmin(d, fabs(m.aarr[p.ip - 1] * p.xp + m.barr[p.ip - 1] * p.yp));
      d = fmin(d, fabs(m.aarr[p.ip] * p.xp + m.barr[p.ip] * p.yp));
    }
  }
  return d - 1.0e-6 * r;
}

template <class T>
__device__ __forceinline__ void set_kapd(Pkt<T>& p, const DeviceModel& m, int order) {
#pragma unroll
  for (int id = 0; id < kMaxOpac; ++id) {
    if (id >= m.nopac) continue;
    if (order == 0) p.kapd[id] = (T)(p.op[id].kappa * (T)(m.rstar * kRsol) * (T)m.kappav[id]);
    else if (order == 1) p.kapd[id] = (T)(p.op[id].kappa * (T)m.rstar * (T)kRsol * (T)m.kappav[id]);
    else p.kapd[id] = (T)(p.op[id].kappa * (T)kRsol * (T)m.rstar * (T)m.kappav[id]);
  }
}

// (1 - f) a + f b, as the reference interpolates its tables.  The FP32
// path fuses the second product explicitly, so that each search, serial
// or over a wavefront, forms the same value at the same index.
template <class T> __device__ __forceinline__ T lerp2(T fm1, T f, T a, T b) { return fm1 * a + f * b; }
template <> __device__ __forceinline__ float lerp2<float>(float fm1, float f, float a, float b) {
  return __builtin_fmaf(f, b, fm1 * a);
}

// dust_5.f90:9-109 (dust species, no gas mixing) for opacity table id:
// the table entries at index i (1-based), or interpolated inside the
// bracket nd(i0) >= nu > nd(i1 = i0 + 1) of the decreasing frequency table
template <class T>
__device__ __forceinline__ Opac<T> opac_entry(const DeviceModel& m, int id, int i) {
  const size_t o0 = (size_t)id * m.nlam + (size_t)(i - 1);
  Opac<T> o;
  o.g = (T)m.gdust[o0]; o.pl = (T)m.pldust[o0]; o.kappa = (T)m.kappad[o0]; o.albedo = (T)m.adust[o0];
  return o;
}
template <class T>
__device__ __forceinline__ Opac<T> opac_interp(const DeviceModel& m, T nu, int id, int i0, int i1) {
  const double* nd = m.nudust + (size_t)id * m.nlam;
  const double *kd = m.kappad + (size_t)id * m.nlam, *ad = m.adust + (size_t)id * m.nlam;
  const double *gd = m.gdust + (size_t)id * m.nlam, *pd = m.pldust + (size_t)id * m.nlam;
  T f = (nu - (T)nd[i1 - 1]) / ((T)nd[i0 - 1] - (T)nd[i1 - 1]);
  T fm1 = (T)1 - f;
  Opac<T> o;
  o.g = lerp2<T>(fm1, f, (T)gd[i1 - 1], (T)gd[i0 - 1]);
  o.pl = lerp2<T>(fm1, f, (T)pd[i1 - 1], (T)pd[i0 - 1]);
  o.kappa = lerp2<T>(fm1, f, (T)kd[i1 - 1], (T)kd[i0 - 1]);
  o.albedo = lerp2<T>(fm1, f, (T)ad[i1 - 1], (T)ad[i0 - 1]);
  return o;
}

template <class T>
__device__ __forceinline__ Opac<T> opac_one(const DeviceModel& m, T nu, int id, bool coop = false) {
  const double* nd = m.nudust + (size_t)id * m.nlam;
  const int nl = m.nlam;
  if (nu >= (T)nd[0]) return opac_entry<T>(m, id, 1);
  if (nu <= (T)nd[nl - 1]) return opac_entry<T>(m, id, nl);
  // The bisection of dust_5.f90:40-52 ends on the unique bracket
  // nd(i0) >= nu > nd(i0+1) of the decreasing table: i0 + 1 is the first
  // index with nd < nu.  On a log-uniform table that index is guessed from
  // log(nu) and stepped to the bracket, which gives the same i0 in one or
  // two probes.  With every lane of the wavefront carrying the packet
  // (coop), lane l tests indices 1 + l, 65 + l, ... and a ballot names it.
  int i0 = 1, i1 = nl, i;
  if (coop) {
    for (int base = 1; base <= nl; base += 64) {
      const int ix = base + __lane_id();
      const unsigned long long hit = __ballot(ix <= nl && nu > (T)nd[ix <= nl ? ix - 1 : 0]);
      if (hit) { i1 = base + __ffsll((long long)hit) - 1; break; }
    }
    i0 = i1 - 1;
  } else if (m.nud_ilog[id] > 0.0) {
    int g = 1 + (int)(((double)m.nud_log0[id] - (double)RT<T>::log_(nu)) * m.nud_ilog[id]);
    g = g < 1 ? 1 : (g > nl - 1 ? nl - 1 : g);
    while (g > 1 && !(nu <= (T)nd[g - 1])) g--;
    while (g < nl - 1 && nu <= (T)nd[g]) g++;
    i0 = g; i1 = g + 1;
  } else {
    while (i1 - i0 > 1) {
      i = (i1 + i0) / 2;
      if (nu <= (T)nd[i - 1]) i0 = i; else i1 = i;
    }
  }
  return opac_interp<T>(m, nu, id, i0, i1);
}

// every opacity table at frequency nu (constant indices: o stays in registers)
template <class T>
__device__ __forceinline__ void opacset(const DeviceModel& m, T nu, Opac<T> (&o)[kMaxOpac], bool coop = false) {
  static_assert(kMaxOpac == 4, "opacset unrolls four tables");
  o[0] = opac_one<T>(m, nu, 0, coop);
  if (m.nopac > 1) o[1] = opac_one<T>(m, nu, 1, coop);
  if (m.nopac > 2) o[2] = opac_one<T>(m, nu, 2, coop);
  if (m.nopac > 3) o[3] = opac_one<T>(m, nu, 3, coop);
}

// random.f90:44-92 bbfreq: the zeta(4) partial sums are computed once on the
// host with the reference's summation order and passed in `inczeta`.
template <class T>
__device__ __forceinline__ T bbfreq(Rng& r, T teff, const double* inczeta) {
  T z0 = (T)inczeta[144] * uniform<T>(r);
  T z1 = uniform<T>(r), z2 = uniform<T>(r), z3 = uniform<T>(r), z4 = uniform<T>(r);
  int i = 1;
  while (z0 > (T)inczeta[i - 1]) i++;
  return -RT<T>::log_(z1 * z2 * z3 * z4) * teff / (T)i;
}

// the re-emission CDF rows in the precision of the path
template <class T> struct KapRows { __device__ static const double* get(const DeviceModel& m) { return m.kapint; } };
template <> struct KapRows<float> { __device__ static const float* get(const DeviceModel& m) { return m.kapint32; } };
template <class T> __device__ __forceinline__ auto kap_rows(const DeviceModel& m) { return KapRows<T>::get(m); }

// lucydustfreq_2_test.f90:34-45: the temperature row j of the CDF tables
// and the interpolation fraction f toward row j + 1
template <class T>
__device__ __forceinline__ void lucy_row(const DeviceModel& m, T T0, int& j, T& f) {
  const int nT = m.nT;
  const double* lT = m.lTint;
  T lT0 = RT<T>::log_(T0);
  j = (int)((T)(nT - 1) * (lT0 - (T)lT[0]) / ((T)lT[nT - 1] - (T)lT[0])) + 1;
  if (j < 1) j = 1;
  if (j > nT - 1) j = nT - 1;
  f = (lT0 - (T)lT[j - 1]) / ((T)lT[j] - (T)lT[j - 1]);
  if (f < (T)0) f = (T)0;
  if (f > (T)1) f = (T)1;
}

// lucydustfreq_2_test.f90:140-160: the frequency inside the bracket
// P(i0) <= xi < P(i1); fT is the temperature interpolation fraction
template <class T>
__device__ __forceinline__ T lucy_finish(const double* nui, T T0, T xi, int i0, int i1, T P0, T P1, T fT) {
  T f = fT;                         // (kept where the reference leaves it)
  if (xi < P0) f = (T)0;
  else if (xi > P1) f = (T)1;
  else if (P0 <= (T)0) f = (T)0.5;
  else if (P1 <= (T)0) { }
  else if (xi <= (T)0) { }
  else {
    // lucydustfreq_2_test.f90:150.  P1/P0 can round to 1 when P0 and P1 are
    // adjacent floating-point values (the top of a CDF row): f = 0 then,
    // where the reference formula would divide 0 by 0.
    const T den = RT<T>::log_acc(P1 / P0);
    f = den > (T)0 ? RT<T>::log_acc(xi / P0) / den : (T)0;
  }
  return T0 * (T)nui[i0 - 1] * RT<T>::pow_((T)nui[i1 - 1] / (T)nui[i0 - 1], f);
}

// lucydustfreq_2_test.f90:19-160 with xi supplied
template <class T>
__device__ __forceinline__ T lucydustfreq_xi(const DeviceModel& m, T T0, int iop, T xi) {
  const int nT = m.nT, nnu = m.nnu;
  const double* nui = m.nuint;
  const auto* K = kap_rows<T>(m) + (size_t)iop * nT * nnu;
#define KAP(i, j) ((T)K[(size_t)((i) - 1) * nnu + ((j) - 1)])
#define CDFP(ix) lerp2<T>((T)1 - f, f, KAP(j, ix), KAP(j + 1, ix))
  int j;
  T f;
  lucy_row<T>(m, T0, j, f);
  int i0 = 1, i1 = nnu, i;
  T P0 = CDFP(i0), P1, P = (T)0;
  if (xi < P0) {
    if (P0 > (T)0) return T0 * xi * (T)nui[0] / P0;
    return T0 * (T)nui[0];
  }
  P1 = CDFP(i1);
  if (P1 > (T)1) P1 = (T)1;
  if (xi >= P1) {
    if (P1 < (T)1) {
      if (xi < (T)1) return T0 * (T)nui[nnu - 1] * ((T)1 - P1) / ((T)1 - xi);
      return T0 * (T)nui[nnu - 1];
    }
    return T0 * (T)nui[nnu - 1];
  }
  // The reference bisects (lucydustfreq_2_test.f90:96-126) with the
  // invariant P(i0) <= xi < P(i1).  Every row of the table is a running sum
  // over frequency divided by its total, so P(i) is non-decreasing in i (the
  // interpolation between two non-decreasing rows, rounded, stays
  // non-decreasing) and the bracket it ends with is the unique one with
  // i1 = i0 + 1.  The same bracket is found here by an 8-way search: each
  // pass probes up to 7 indices with independent loads, so a 513-entry row
  // takes 3 dependent rounds instead of 9.  P0 and P1 are the same values.
  //
  // Before that, a guide table (the index table of Chen & Asau, 1974)
  // narrows the start.  For xi in [k/K, (k+1)/K), every index whose entries
  // in both rows j and j+1 are <= k/K has P <= xi, and every index whose
  // entries in both rows exceed (k+1)/K has P > xi; the guide's counts of
  // entries <= k/K give those indices, widened by one for rounding.  The
  // range is checked against P at its two ends (P(lo) <= xi < P(hi)), read
  // in the same pass as 7 points inside it, and the full row is searched
  // when the check fails, so the bracket is the same either way.
  (void)i;
  if (m.kguide) {
    const int nb = m.kguide_k;
    int k = (int)(xi * (T)nb);
    k = k < 0 ? 0 : (k > nb - 1 ? nb - 1 : k);
    const uint16_t* ga = m.kguide + ((size_t)iop * nT + (size_t)(j - 1)) * (size_t)(nb + 1);
    const uint16_t* gb = ga + (nb + 1);
    const int a0 = ga[k], b0 = gb[k], a1 = ga[k + 1], b1 = gb[k + 1];
    int lo = (a0 < b0 ? a0 : b0) - 1, hi = (a1 > b1 ? a1 : b1) + 2;
    lo = lo < 1 ? 1 : lo;
    hi = hi > nnu ? nnu : hi;
    if (hi - lo >= 1 && (lo > 1 || hi < nnu)) {
      const int step = (hi - lo + 7) >> 3;
      T Pq[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        int ix = lo + q * step;
        ix = ix < hi ? ix : hi;
        T v = CDFP(ix);
        Pq[q] = (ix > 1 && v > (T)1) ? (T)1 : v;
      }
      if (!(Pq[0] > xi) && Pq[8] > xi) {
        int n0 = lo, n1 = hi;
        T p0 = Pq[0], p1 = Pq[8];
#pragma unroll
        for (int q = 7; q >= 1; --q) {
          const int ix = lo + q * step;
          if (ix < hi && Pq[q] > xi) { n1 = ix; p1 = Pq[q]; }
        }
#pragma unroll
        for (int q = 1; q < 8; ++q) {
          const int ix = lo + q * step;
          if (ix < hi && !(Pq[q] > xi)) { n0 = ix; p0 = Pq[q]; }
        }
        i0 = n0; i1 = n1; P0 = p0; P1 = p1;
      }
    }
  }
  while (i1 - i0 > 1) {
    const int step = (i1 - i0 + 7) >> 3;
    T Pq[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const int q = i0 + (k + 1) * step;
      const int qc = q < i1 ? q : i1;
      T v = CDFP(qc);
      Pq[k] = v > (T)1 ? (T)1 : v;
    }
    int n0 = i0, n1 = i1;
    T p0 = P0, p1 = P1;
#pragma unroll
    for (int k = 6; k >= 0; --k) {
      const int q = i0 + (k + 1) * step;
      if (q < i1) {
        if (Pq[k] > xi) { n1 = q; p1 = Pq[k]; }
      }
    }
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const int q = i0 + (k + 1) * step;
      if (q < i1 && !(Pq[k] > xi)) { n0 = q; p0 = Pq[k]; }
    }
    i0 = n0; i1 = n1; P0 = p0; P1 = p1;
  }
  (void)P;
  return lucy_finish<T>(nui, T0, xi, i0, i1, P0, P1, f);
#undef KAP
#undef CDFP
}

// The same draw with every lane of the wavefront carrying the packet
// (k_tail): lane l holds P at indices 2 + l, 66 + l, ... of the row and one
// ballot names the first index with P > xi.  With P non-decreasing (the
// search above relies on it) that index is the serial search's i1, so the
// bracket, P0, P1 and the result are the same.  The row (interpolated at
// the temperature, clamped as the serial search clamps it) is formed once
// per (temperature, table) in the wavefront's LDS row cache: a packet
// trapped in a thick cell re-emits there many times at one temperature,
// and the table rows would otherwise come from memory on every draw.
constexpr int kCdfCache = 512;      // cached P entries (indices 2..513)
template <class T>
struct CdfCache {
  T* row = nullptr;                 // this wavefront's LDS row, kCdfCache entries
  T T0 = (T)-1;                     // temperature and table of the cached row
  int iop = -1;
  T fT = (T)0, first = (T)0, last = (T)0;   // lucy_row's f, P(1) (unclamped), P(nnu)
};

template <class T>
__device__ __forceinline__ T lucydustfreq_coop(const DeviceModel& m, T T0, int iop, T xi, CdfCache<T>& cc) {
  const int nT = m.nT, nnu = m.nnu;
  const double* nui = m.nuint;
  const int lane = __lane_id();
  const bool cached = cc.row != nullptr && nnu - 1 <= kCdfCache;
  const auto* Ka = kap_rows<T>(m) + (size_t)iop * nT * nnu;
  const auto* Kb = Ka;
  T f, fm1 = (T)0, Pfirst, Plast;
  if (cached && cc.T0 == T0 && cc.iop == iop) {
    f = cc.fT; Pfirst = cc.first; Plast = cc.last;
  } else {
    int j;
    lucy_row<T>(m, T0, j, f);
    fm1 = (T)1 - f;
    Ka += (size_t)(j - 1) * nnu;
    Kb = Ka + nnu;
    Pfirst = lerp2<T>(fm1, f, (T)Ka[0], (T)Kb[0]);
    Plast = lerp2<T>(fm1, f, (T)Ka[nnu - 1], (T)Kb[nnu - 1]);
    if (Plast > (T)1) Plast = (T)1;
    if (cached) {
      for (int k = 2 + lane; k <= nnu; k += 64) {
        const T w = lerp2<T>(fm1, f, (T)Ka[k - 1], (T)Kb[k - 1]);
        cc.row[k - 2] = w > (T)1 ? (T)1 : w;
      }
      // the row is read below by other lanes of this wavefront
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      cc.T0 = T0; cc.iop = iop; cc.fT = f; cc.first = Pfirst; cc.last = Plast;
    }
  }
  if (xi < Pfirst) {
    if (Pfirst > (T)0) return T0 * xi * (T)nui[0] / Pfirst;
    return T0 * (T)nui[0];
  }
  if (xi >= Plast) {
    if (Plast < (T)1) {
      if (xi < (T)1) return T0 * (T)nui[nnu - 1] * ((T)1 - Plast) / ((T)1 - xi);
      return T0 * (T)nui[nnu - 1];
    }
    return T0 * (T)nui[nnu - 1];
  }
  T prev = Pfirst;                  // P at the index before this pass's first
  for (int base = 2;; base += 64) {
    const int ix = base + lane;
    T v = (T)1;
    if (ix <= nnu) {
      if (cached) v = cc.row[ix - 2];
      else {
        v = lerp2<T>(fm1, f, (T)Ka[ix - 1], (T)Kb[ix - 1]);
        if (v > (T)1) v = (T)1;
      }
    }
    const unsigned long long hit = __ballot(ix <= nnu && v > xi);
    if (hit) {                      // reached within the row: P(nnu) > xi
      const int t = __ffsll((long long)hit) - 1;
      const T P1 = __shfl(v, t, 64);
      const T Pt = __shfl(v, t > 0 ? t - 1 : 0, 64);
      return lucy_finish<T>(nui, T0, xi, base + t - 1, base + t, t > 0 ? Pt : prev, P1, f);
    }
    prev = __shfl(v, 63, 64);
  }
}

// dustmat.f90:42-67
template <class T>
__device__ __forceinline__ void dustmat(T g, T pl, T cost, T cost2, T p[4]) {
  T g2 = g * g, g2p1 = g2 + (T)1, onemg2 = (T)1 - g2, twog = (T)2 * g;
  T temp = g2p1 - twog * cost;
  T p1 = (T)1.5 * onemg2 / ((T)2 + g2) * ((T)1 + cost2) / (temp * RT<T>::sqrt_(temp));
  p[0] = p1;
  p[1] = -pl * p1 * ((T)1 - cost2) / ((T)1 + cost2);
  p[2] = p1 * (T)2 * cost / ((T)1 + cost2);
  p[3] = (T)0;
}

template <class T>
__device__ __forceinline__ T cs_peak(T g) {  // stokes.f90:54-70
  T g2 = g * g, g2p1 = g2 + (T)1, onemg2 = (T)1 - g2, twog = (T)2 * g, temp, peak;
  if (g > (T)0) { temp = g2p1 - twog; peak = (T)1.5 * (onemg2) / ((T)2 + g2) * (T)2 / temp / RT<T>::sqrt_(temp); }
  else { temp = g2p1 + twog; peak = (T)1.5 * (onemg2) / ((T)2 + g2) * (T)2 / temp / RT<T>::sqrt_(temp); }
  return peak * (T)1.01;
}

// stokes.f90:117-336 for an accepted (bmu, ri1).  dir = cost,sint,phi,cosp,sinp.
// Returns false (no change) for bmu == 1 as the reference does.
template <class T>
__device__ __forceinline__ void stokes_apply(T g, T pl, T bmu, T ri1, T dir[5], T s[4]) {
  const T one = (T)1, two = (T)2, zero = (T)0;
  const T PI = (T)kPi, R2P = (T)kR2p;
  T costp = dir[0], sintp = dir[1], phip = dir[2];
  T cost = dir[0], sint = dir[1], phi = dir[2];
  T sip = s[0], sqp = s[1], sup = s[2], svp = s[3];
  T p[4], cosb2 = bmu * bmu, sinbt;
  T a11, a12, a13, a21, a22, a23, a24, a31, a32, a33, a34, a42, a43, a44, a;
  T sini2 = zero, cosi2 = zero, bott, cosdph, sin2i2, cos2i2, sin2, cos2, sin2cos1, cos2sin1;
  dustmat(g, pl, bmu, cosb2, p);
  sinbt = RT<T>::sqrt_(one - cosb2);
  if (ri1 > PI) {
    T ri3 = R2P - ri1, cosi3 = RT<T>::cos_(ri3), sini3 = RT<T>::sin_(ri3);
    T sin2i3 = two * sini3 * cosi3, cos2i3 = two * (cosi3 * cosi3) - one;
    a11 = p[0]; a12 = p[1] * cos2i3; a13 = p[1] * sin2i3;
    a = (a11 * sip + a12 * sqp + a13 * sup) / sip;
    if (bmu == one) return;
    else if (bmu == -one) { sini3 = one; sini2 = one; cosi3 = zero; cosi2 = zero; cost = -costp; sint = sintp; }
    else {
      cost = costp * bmu + sintp * sinbt * cosi3;
      if (RT<T>::fabs_(cost) < one) {
        sint = RT<T>::fabs_(RT<T>::sqrt_(one - cost * cost));
        sini2 = sini3 * sintp / sint;
        bott = sint * sinbt;
        cosi2 = costp / bott - cost * bmu / bott;
      } else {
        sint = zero; sini2 = zero;
        if (cost >= one) cosi2 = -one;
        if (cost <= -one) cosi2 = one;
      }
    }
    cosdph = -cosi2 * cosi3 + sini2 * sini3 * bmu;
    if (RT<T>::fabs_(cosdph) > one) cosdph = (cosdph > one) ? one : -one;
    phi = phip + RT<T>::acos_(cosdph);
    if (phi > R2P) phi = phi - R2P;
    if (phi < zero) phi = phi + R2P;
    sin2i2 = two * sini2 * cosi2; cos2i2 = two * (cosi2 * cosi2) - one;
    sin2 = sin2i2 * sin2i3; cos2 = cos2i2 * cos2i3;
    sin2cos1 = sin2i2 * cos2i3; cos2sin1 = cos2i2 * sin2i3;
    a21 = p[1] * cos2i2; a22 = p[0] * cos2 - p[2] * sin2; a23 = p[0] * cos2sin1 + p[2] * sin2cos1; a24 = -p[3] * sin2i2;
    a31 = -p[1] * sin2i2; a32 = -p[0] * sin2cos1 - p[2] * cos2sin1; a33 = -p[0] * sin2 + p[2] * cos2; a34 = -p[3] * cos2i2;
    a42 = -p[3] * sin2i3; a43 = p[3] * cos2i3; a44 = p[2];
  } else {
    T cosi1 = RT<T>::cos_(ri1), sini1 = RT<T>::sin_(ri1);
    T sin2i1 = two * sini1 * cosi1, cos2i1 = two * (cosi1 * cosi1) - one;
    a11 = p[0]; a12 = p[1] * cos2i1; a13 = -p[1] * sin2i1;
    a = (a11 * sip + a12 * sqp + a13 * sup) / sip;
    if (bmu == one) return;
    else if (bmu == -one) { sini1 = one; sini2 = one; cosi1 = zero; cosi2 = zero; cost = -costp; sint = sintp; }
    else {
      cost = costp * bmu + sintp * sinbt * cosi1;
      if (RT<T>::fabs_(cost) < one) {
        sint = RT<T>::fabs_(RT<T>::sqrt_(one - cost * cost));
        sini2 = sini1 * sintp / sint;
        bott = sint * sinbt;
        cosi2 = costp / bott - cost * bmu / bott;
      } else {
        sint = zero; sini2 = zero;
        if (cost >= one) cosi2 = -one;
        if (cost <= -one) cosi2 = one;
      }
    }
    cosdph = -cosi1 * cosi2 + sini1 * sini2 * bmu;
    if (RT<T>::fabs_(cosdph) > one) cosdph = (cosdph > one) ? one : -one;
    phi = phip - RT<T>::acos_(cosdph);
    if (phi > R2P) phi = phi - R2P;
    if (phi < zero) phi = phi + R2P;
    sin2i2 = two * sini2 * cosi2; cos2i2 = two * (cosi2 * cosi2) - one;
    sin2 = sin2i2 * sin2i1; cos2 = cos2i2 * cos2i1;
    sin2cos1 = sin2i2 * cos2i1; cos2sin1 = cos2i2 * sin2i1;
    a21 = p[1] * cos2i2; a22 = p[0] * cos2 - p[2] * sin2; a23 = -p[0] * cos2sin1 - p[2] * sin2cos1; a24 = p[3] * sin2i2;
    a31 = p[1] * sin2i2; a32 = p[0] * sin2cos1 + p[2] * cos2sin1; a33 = -p[0] * sin2 + p[2] * cos2; a34 = -p[3] * cos2i2;
    a42 = p[3] * sin2i1; a43 = p[3] * cos2i1; a44 = p[2];
  }
  s[0] = (a11 * sip + a12 * sqp + a13 * sup) / a;
  s[1] = (a21 * sip + a22 * sqp + a23 * sup + a24 * svp) / a;
  s[2] = (a31 * sip + a32 * sqp + a33 * sup + a34 * svp) / a;
  s[3] = (a42 * sqp + a43 * sup + a44 * svp) / a;
  dir[0] = cost; dir[1] = sint; dir[2] = phi; dir[3] = RT<T>::cos_(phi); dir[4] = RT<T>::sin_(phi);
}

// stokespeel.f90:77-322: fixed (bmu, ri1), no division by the phase function.
template <class T>
__device__ __forceinline__ void stokespeel(T g, T pl, T s[4], T cost, T sint, T ri1, T bmu, T costp, T sintp) {
  const T one = (T)1, two = (T)2, zero = (T)0;
  const T PI = (T)kPi, R2P = (T)kR2p;
  T sip = s[0], sqp = s[1], sup = s[2], svp = s[3];
  T p[4], cosb2 = bmu * bmu, sinbt;
  T a11, a12, a13, a21, a22, a23, a24, a31, a32, a33, a34, a42, a43, a44;
  T sini2 = zero, cosi2 = zero, bott, sin2i2, cos2i2, sin2, cos2, sin2cos1, cos2sin1;
  dustmat(g, pl, bmu, cosb2, p);
  sinbt = RT<T>::sqrt_(one - cosb2);
  if (ri1 > PI) {
    T ri3 = R2P - ri1, cosi3 = RT<T>::cos_(ri3), sini3 = RT<T>::sin_(ri3);
    T sin2i3 = two * sini3 * cosi3, cos2i3 = two * (cosi3 * cosi3) - one;
    a11 = p[0]; a12 = p[1] * cos2i3; a13 = p[1] * sin2i3;
    if (bmu == one) return;
    else if (bmu == -one) { sini3 = one; sini2 = one; cosi3 = zero; cosi2 = zero; }
    else {
      if (RT<T>::fabs_(cost) < one) {
        sini2 = sini3 * sintp / sint;
        bott = sint * sinbt;
        cosi2 = costp / bott - cost * bmu / bott;
      } else {
        sini2 = zero;
        if (cost >= one) cosi2 = -one;
        if (cost <= -one) cosi2 = one;
      }
    }
    sin2i2 = two * sini2 * cosi2; cos2i2 = two * (cosi2 * cosi2) - one;
    sin2 = sin2i2 * sin2i3; cos2 = cos2i2 * cos2i3;
    sin2cos1 = sin2i2 * cos2i3; cos2sin1 = cos2i2 * sin2i3;
    a21 = p[1] * cos2i2; a22 = p[0] * cos2 - p[2] * sin2; a23 = p[0] * cos2sin1 + p[2] * sin2cos1; a24 = -p[3] * sin2i2;
    a31 = -p[1] * sin2i2; a32 = -p[0] * sin2cos1 - p[2] * cos2sin1; a33 = -p[0] * sin2 + p[2] * cos2; a34 = -p[3] * cos2i2;
    a42 = -p[3] * sin2i3; a43 = p[3] * cos2i3; a44 = p[2];
  } else {
    T cosi1 = RT<T>::cos_(ri1), sini1 = RT<T>::sin_(ri1);
    T sin2i1 = two * sini1 * cosi1, cos2i1 = two * (cosi1 * cosi1) - one;
    a11 = p[0]; a12 = p[1] * cos2i1; a13 = -p[1] * sin2i1;
    if (bmu == one) return;
    else if (bmu == -one) { sini1 = one; sini2 = one; cosi1 = zero; cosi2 = zero; }
    else {
      if (RT<T>::fabs_(cost) < one) {
        sini2 = sini1 * sintp / sint;
        bott = sint * sinbt;
        cosi2 = costp / bott - cost * bmu / bott;
      } else {
        sini2 = zero;
        if (cost >= one) cosi2 = -one;
        if (cost <= -one) cosi2 = one;
      }
    }
    sin2i2 = two * sini2 * cosi2; cos2i2 = two * (cosi2 * cosi2) - one;
    sin2 = sin2i2 * sin2i1; cos2 = cos2i2 * cos2i1;
    sin2cos1 = sin2i2 * cos2i1; cos2sin1 = cos2i2 * sin2i1;
    a21 = p[1] * cos2i2; a22 = p[0] * cos2 - p[2] * sin2; a23 = -p[0] * cos2sin1 - p[2] * sin2cos1; a24 = p[3] * sin2i2;
    a31 = p[1] * sin2i2; a32 = p[0] * sin2cos1 + p[2] * cos2sin1; a33 = -p[0] * sin2 + p[2] * cos2; a34 = -p[3] * cos2i2;
    a42 = p[3] * sin2i1; a43 = p[3] * cos2i1; a44 = p[2];
  }
  s[0] = (a11 * sip + a12 * sqp + a13 * sup);
  s[1] = (a21 * sip + a22 * sqp + a23 * sup + a24 * svp);
  s[2] = (a31 * sip + a32 * sqp + a33 * sup + a34 * svp);
  s[3] = (a42 * sqp + a43 * sup + a44 * svp);
}

// findangle3.f90:31-67 (ri1 = 0 where the reference leaves it unset)
template <class T>
__device__ __forceinline__ void findangle3(T cost1, T sint1, T phi1, T cosp1, T sinp1,
                                           T cost2, T sint2, T phi2, T cosp2, T sinp2,
                                           T& costnew, T& sintnew, T& phinew, T& ri1) {
  const T PI = (T)kPi, R2P = (T)kR2p;
  T cn = cost1 * cost2 + (sint1 * sint2 * (cosp1 * cosp2 + sinp1 * sinp2));
  cn = check1(cn);
  T sn = RT<T>::sqrt_((T)1 - cn * cn);
  costnew = cn; sintnew = sn; ri1 = (T)0;
  if (sint1 > (T)1.0e-8) {
    if (sn > (T)1.0e-8) {
      T cosx = (cost2 - (cost1 * cn)) / (sint1 * sn);
      cosx = check1(cosx);
      T x = RT<T>::acos_(cosx);
      if (phi2 > phi1) {
        T diff = phi2 - phi1;
        if (diff > PI) { ri1 = x; phinew = PI + x; }
        else { phinew = PI - x; ri1 = R2P - x; }
      } else {
        T diff = phi1 - phi2;
        if (diff > PI) { phinew = PI - x; ri1 = R2P - x; }
        else { phinew = PI + x; ri1 = x; }
      }
    } else {
      phinew = (T)0;
    }
  } else {
    phinew = phi2;
  }
}

// ---------------------------------------------------------------------------
// Rarely needed FP64 transcendentals of testgrid, kept out of line so their
// registers do not weigh on every walk step.
__device__ __attribute__((noinline)) double azimuth(double x, double y) {
  double phi = atan2(y, x);
  if (phi < 0.0) phi = phi + kR2p;
  return phi;
}
// testgrid.f90:61-63: radius of a random re-placement, uniform in volume
__device__ __attribute__((noinline)) double shell_radius(double u, double ra, double rb) {
  return pow(u * (pow(ra, 3.0) - pow(rb, 3.0)) + pow(rb, 3.0), (double)(1.0f / 3.0f));
}
// emit_common_2.f90:258-270, :343-361: the fourth roots of the diffusion
// surface's T^4 (only photons absorbed in a diffusion region reach them)
__device__ __attribute__((noinline)) double quarter_power(double x) { return pow(x, 0.25); }
// sin of a wall angle and the azimuth of a direction, for the same photons
__device__ __attribute__((noinline)) double sin_nl(double x) { return sin(x); }
__device__ __attribute__((noinline)) double atan2_nl(double y, double x) { return atan2(y, x); }
struct Point3 { double x, y, z; };
// testgrid.f90:80-86
__device__ __attribute__((noinline)) Point3 sphere_point(double rtot, double cosb, double phi) {
  double s = sqrt(1.0 - pow(cosb, 2.0));
  return Point3{rtot * s * cos(phi), rtot * s * sin(phi), rtot * cosb};
}

// testgrid.f90:40-124, on the FP64 geometry.  Returns true when state changed.
template <class T>
__device__ __forceinline__ bool testgrid(const DeviceModel& m, Pkt<T>& p, Rng& rng) {
  const double tol = 1.0e-4;
  const double PI = kPi, R2P = kR2p;
  const int nrg = m.nrg, ntg = m.ntg, npg = m.npg;
  const double* r2 = m.r2arr;
  const double* ca = m.costarr;
  const double* pa = m.phiarr;
  bool reset = false, changed = false;
  double cosb = check1(p.zp / p.rtot);
  double phi = 0.0;
  // The azimuth (atan2) only decides anything when the packet lies outside
  // its phi wedge by more than tol of the wedge: phi = atan2(..) is in
  // [0, 2pi] = [phiarr(1), phiarr(npg)], so the whole-grid phi test never
  // fires, and inside the widened wedge pfrac is within [-tol, 1+tol].  The
  // wall planes give s = sin(phi_wall) x - cos(phi_wall) y =
  // rho sin(phi_wall - phi); s1 <= rho sin(eps), s2 >= -rho sin(eps) with
  // eps = tol/2 of the narrowest wedge (phi_tol2 = sin(eps)^2) proves the
  // packet inside the wedge widened by eps (wedges narrower than pi, npg >= 4),
  // so atan2 is skipped, packets on a phi wall included.  Every decision is
  // the one atan2 would give.  On the z axis (rho = 0) atan2(0,0) = 0 does
  // decide, and is computed.
  bool have_phi = true;
  if (npg > 2) {
    have_phi = false;
    if (npg >= 4 && p.ip >= 1 && p.ip <= npg - 1) {
      const double s1 = m.aarr[p.ip - 1] * p.xp + m.barr[p.ip - 1] * p.yp;
      const double s2 = m.aarr[p.ip] * p.xp + m.barr[p.ip] * p.yp;
      const double rho2 = p.xp * p.xp + p.yp * p.yp;
      const double d2 = m.phi_tol2 * rho2;
      const bool in1 = s1 < 0.0 || s1 * s1 <= d2;
      const bool in2 = s2 > 0.0 || s2 * s2 <= d2;
      if (!(rho2 > 0.0 && in1 && in2)) have_phi = true;
    } else {
      have_phi = true;
    }
    if (have_phi) phi = azimuth(p.xp, p.yp);
  }
  if ((p.rsq - r2[0]) / (r2[1] - r2[0]) < -tol || (p.rsq - r2[nrg - 2]) / (r2[nrg - 1] - r2[nrg - 2]) > 1.0 + tol) {
    p.rtot = shell_radius((double)uniform<T>(rng), m.rarr[nrg - 1], m.rarr[0]);
    reset = true;
  }
  if ((cosb - ca[0]) / (ca[1] - ca[0]) < -tol || (cosb - ca[ntg - 2]) / (ca[ntg - 1] - ca[ntg - 2]) > 1.0 + tol) {
    cosb = ((double)uniform<T>(rng) * (ca[ntg - 1] - ca[0]) + ca[0]);
    reset = true;
  }
  if (npg > 2 && have_phi && ((phi - pa[0]) / (pa[1] - pa[0]) < -tol || (phi - pa[npg - 2]) / (pa[npg - 1] - pa[npg - 2]) > 1.0 + tol)) {
    phi = (double)uniform<T>(rng) * (pa[npg - 1] - pa[0]) + pa[0];
    reset = true;
  }
  if (reset && !have_phi) { phi = azimuth(p.xp, p.yp); have_phi = true; }
  if (reset) {
    const Point3 q = sphere_point(p.rtot, cosb, phi);
    p.xp = q.x;
    p.yp = q.y;
    p.zp = q.z;
    p.rsq = p.rtot * p.rtot;
    p.on_wall = false;
    changed = true;
  }
  if (p.ir < 1 || p.ir > nrg) { p.ir = locate<double>(r2, nrg, p.rsq); changed = true; }
  if (p.it < 1 || p.it > ntg) { p.it = locate<double>(ca, ntg, cosb); changed = true; }
  if (p.ip < 1 || p.ip > npg) {
    if (!have_phi) { phi = azimuth(p.xp, p.yp); have_phi = true; }
    p.ip = locate<double>(pa, npg, phi);
    changed = true;
  }
  double rfrac = (p.rsq - r2[p.ir - 1]) / (r2[p.ir] - r2[p.ir - 1]);
  double tfrac = (cosb - ca[p.it - 1]) / (ca[p.it] - ca[p.it - 1]);
  bool phierr = false;
  if (npg > 2 && have_phi) {
    double pfrac = (phi - pa[p.ip - 1]) / (pa[p.ip] - pa[p.ip - 1]);
    if (pfrac < -tol || pfrac > 1.0 + tol) {
      double d1 = phi - pa[p.ip - 1], d2 = phi - pa[p.ip];
      if (d1 > PI) d1 -= R2P;
      if (d1 < -PI) d1 += R2P;
      if (d2 > PI) d2 -= R2P;
      if (d2 < -PI) d2 += R2P;
      phierr = fabs(d1) > tol && fabs(d2) > tol;
    }
  }
  if (rfrac < -tol || rfrac > 1.0 + tol) { p.ir = locate<double>(r2, nrg, p.rsq); changed = true; }
  if (tfrac < -tol || tfrac > 1.0 + tol) { p.it = locate<double>(ca, ntg, cosb); changed = true; }
  if (phierr) { p.ip = locate<double>(pa, npg, phi); changed = true; }
  return changed;
}

// one face-selection step shared by the transport and peel-off walks
template <class T>
__device__ __forceinline__ bool next_face(const DeviceModel& m, const Pkt<T>& p, double ux, double uy, double uz,
                                          double& dt, int& iface) {
  Walls w;
  w.reset();
  radface(w, ux, uy, uz, p.xp, p.yp, p.zp, m.r2arr[p.ir - 1], m.r2arr[p.ir]);
  if (m.ntg > 2) thetaface(w, ux, uy, uz, p.xp, p.yp, p.zp, m.tan2arr[p.it - 1], m.tan2arr[p.it]);
  if (m.npg > 2 && m.ntg > 2)
    phiface(w, m.aarr[p.ip - 1], m.barr[p.ip - 1], m.aarr[p.ip], m.barr[p.ip], ux, uy, uz, p.xp, p.yp, p.zp);
  if (!w.next(p.on_wall, dt, iface)) return false;
  if (dt < 0.0) return false;
  return true;
}

// iface 1/2: ir -/+ 1, 3/4: it -/+ 1, 5/6: ip -/+ 1 (phi wraps).  Written as
// three additions rather than a switch: the optimizer turns a switch over
// three references into one store through a selected pointer, which keeps
// the caller's cell indices in private memory (a scratch load on every
// use along the packet's serial chain).
__device__ __forceinline__ void cross_face(const DeviceModel& m, int& ir, int& it, int& ip, int iface) {
  const int s = (iface & 1) ? -1 : 1;
  ir += (iface <= 2) ? s : 0;
  it += (iface == 3 || iface == 4) ? s : 0;
  ip += (iface >= 5) ? s : 0;
  if (ip == 0) ip = m.npg - 1;
  if (ip == m.npg) ip = 1;
}

template <class T>
__device__ __forceinline__ T cell_opac(const DeviceModel& m, const Pkt<T>& p, size_t c, int& iop, T& rho) {
  rho = cell_rho<T>(m, c);
  if (m.nopac == 1) {   // one opacity table: rho > 0 exactly when a species is present
    iop = rho > (T)0 ? 0 : -1;
    return rho > (T)0 ? (T)0 + p.kapd[0] * rho : (T)0;
  }
  int sp = m.species[c];
  iop = -1;
  if (sp > 0 && rho > (T)0) { iop = m.species_opac[sp]; return (T)0 + pick(p.kapd, iop) * rho; }
  return (T)0;
}

}  // namespace mcrt
