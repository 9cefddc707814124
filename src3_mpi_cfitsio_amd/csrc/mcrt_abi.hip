// mcrt_abi.hip -- the extern "C" boundary (include/mcrt.h): context,
// device-resident model state, launches, tally read-back.
#include "../../include/mcrt.h"
#include "mcrt_kernels.h"
#include "mcrt_launch.h"
#include "mcrt_emiss.h"
#include "mcrt_build.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

using namespace mcrt;

struct mcrt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t own_stream = nullptr;
  std::string err;
  std::map<std::string, void*> buf;       // device allocations by name
  std::map<std::string, size_t> bytes;
  DeviceModel dm{};
  DeviceTallies dt{};
  bool has_grid = false, has_medium = false, has_dust = false, has_source = false, has_observer = false, has_tdust = false;
  // external illumination: fraction of a run's packets, device tables
  double isrf_frac = 0.0;
  int nisrf = 0;
  const double *isrf_cdf = nullptr, *isrf_logcdf = nullptr, *isrf_logfreq = nullptr;
  // stellar hot spots (mcrt_set_spots): geometry in spot, hot-spot photon fraction
  DeviceModel spot{};
  double accfrac2 = 0.0;
  // disk-accretion photons (mcrt_set_accretion): fraction of a run's packets,
  // device tables of acc_n = nrg - 1 radial cells
  double accfrac = 0.0;
  int acc_n = 0;
  const double *acc_ldint = nullptr, *acc_hdisk = nullptr;
  size_t ncell = 0, nsed = 0, npsed = 0, nimage = 0;
  int64_t trace_n = 0;
  // diffusion approximation (mcrt_set_diffusion, mcrt_setdiffus)
  bool diff_on = false, diff_set = false;
  int diff_nr0 = 3;
  double diff_rmaxd = 0.0, diff_delta_run = 0.0;
  // Lucy convergence (mcrt_dusttemp_sigma): sums of the last mcrt_dusttemp,
  // sigma of the last call that updated tdust
  double lucy_sum[3] = {0.0, 0.0, 0.0};
  double lucy_sigma = NAN, lucy_sigma_iter = NAN, lucy_sigma_prev = NAN;
  bool lucy_valid = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool ev_valid = false;
  int cu_count = 256;
  // launch configuration (mcrt_set_launch); 0 = built-in default
  int64_t lanes = 0;
  int64_t queue_per_lane = 0;
  int events_per_round = 0;
  int peel_blocks_per_cu = 0;
  int regroup = 0;
  int peel_fetch = 0, peel_tally = 0;   // peel-kernel gather thresholds (0: defaults)
  RunCtl* host_ctl = nullptr;             // pinned read-back of the run control blocks (two)
  hipEvent_t evr[3] = {nullptr, nullptr, nullptr};   // per-round kernel bracketing
  // the peel kernels of round r run on peel_stream while round r + 1's
  // transport runs on stream (two peel-off queues, ordered by events)
  hipStream_t peel_stream = nullptr;
  hipEvent_t evp[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};   // peel start / end per queue
  double stats[MCRT_NSTATS] = {0};        // mcrt_run_stats of the last run
  // batch of axisymmetric models (mcrt_set_batch)
  bool has_batch = false, has_batch_tdust = false;
  int nmodels = 0;
  std::vector<int> b_nrg, b_ntg;
  std::vector<long long> b_coff;
  size_t b_ncell = 0, b_max_cells = 0, b_nsed = 0;
  BatchDesc bd{};
  DeviceTallies bt{};
  const int* b_species_opac = nullptr;
  BatchRun* b_run = nullptr;              // device
  BatchRun* b_run_h = nullptr;            // pinned staging
  std::vector<int> b_order;
  std::vector<unsigned long long> b_prefix, b_npk, b_work;
  std::vector<double> b_cost;             // per-packet cost of each model (ordering)
  int b_block = 256, b_waves = 2, b_regroup = 32, b_combine = 0, b_graph = 1;
  hipGraphExec_t b_exec = nullptr;        // captured run of the current shape
  unsigned long long b_exec_key[6] = {0, 0, 0, 0, 0, 0};
  unsigned long long b_gen = 1;           // bumped by every call that changes what a launch embeds
  const double *b_phiarr = nullptr, *b_aarr = nullptr, *b_barr = nullptr;
  unsigned long long* b_npk_dev = nullptr;
  bool b_ran = false;
};

namespace {

// LDS plan of the transport and peel kernels: the walls, Planck grids, peel
// directions and zeta sums always (at most kLdsAlways doubles), the dust
// tables when the total stays within kLdsMax doubles.
constexpr uint32_t kLdsMax = 62u * 1024u / 8u;
constexpr uint32_t kLdsAlways = 56u * 1024u / 8u;
LdsPlan plan_lds(const DeviceModel& m) {
  LdsPlan P{};
  const uint32_t nrg = (uint32_t)m.nrg, ntg = (uint32_t)m.ntg, npg = (uint32_t)m.npg;
  const uint32_t nd = (uint32_t)(m.nopac * m.nlam);
  const uint32_t sizes[LT_COUNT] = {nrg, nrg, ntg, ntg, ntg, ntg, npg, npg, npg,
                                    (uint32_t)m.nT, (uint32_t)m.nnu, 5u * (uint32_t)(m.npeel > 0 ? m.npeel : 1), 145u,
                                    nd, nd, nd, nd, nd};
  uint32_t used = 0;
  for (int k = 0; k < LT_COUNT; ++k) {
    if (k == LT_NUDUST && used + 5u * nd > kLdsMax) break;
    P.off[k] = used;
    P.n[k] = sizes[k];
    used += sizes[k];
  }
  P.bytes = used * 8u;
  return P;
}

uint32_t lds_always_doubles(const DeviceModel& m) {
  return 2u * (uint32_t)m.nrg + 4u * (uint32_t)m.ntg + 3u * (uint32_t)m.npg + (uint32_t)m.nT + (uint32_t)m.nnu +
         5u * (uint32_t)(m.npeel > 0 ? m.npeel : 1) + 145u;
}

int fail(mcrt_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hipfail(mcrt_ctx* c, hipError_t e, const char* what) {
  return fail(c, MCRT_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHECK(c, expr)                                  \
  do {                                                     \
    hipError_t _e = (expr);                                \
    if (_e != hipSuccess) return hipfail((c), _e, #expr);  \
  } while (0)

// (re)allocate a named device buffer and optionally upload host data
int dev_alloc(mcrt_ctx* c, const std::string& name, size_t nbytes, void** out) {
  auto it = c->buf.find(name);
  if (it != c->buf.end()) {
    if (c->bytes[name] == nbytes) { *out = it->second; return MCRT_OK; }
    HIPCHECK(c, hipFree(it->second));
    c->buf.erase(it);
  }
  void* p = nullptr;
  HIPCHECK(c, hipMalloc(&p, nbytes ? nbytes : 16));
  c->buf[name] = p;
  c->bytes[name] = nbytes;
  *out = p;
  return MCRT_OK;
}

template <class V>
int upload(mcrt_ctx* c, const std::string& name, const V* host, size_t n, const V** out) {
  void* p = nullptr;
  int rc = dev_alloc(c, name, n * sizeof(V), &p);
  if (rc) return rc;
  if (n) HIPCHECK(c, hipMemcpyAsync(p, host, n * sizeof(V), hipMemcpyHostToDevice, c->stream));
  *out = (const V*)p;
  return MCRT_OK;
}

template <class V>
int upload_cast(mcrt_ctx* c, const std::string& name, const double* host, size_t n, const V** out) {
  std::vector<V> tmp(n);
  for (size_t i = 0; i < n; ++i) tmp[i] = (V)host[i];
  int rc = upload<V>(c, name, tmp.data(), n, out);
  if (rc) return rc;
  HIPCHECK(c, hipStreamSynchronize(c->stream));   // tmp goes out of scope
  return MCRT_OK;
}

// random.f90:74-92: partial sums of 1/j^4, smallest term first
std::vector<double> zeta_partial_sums() {
  std::vector<double> z(145, 0.0);
  for (int j = 145; j >= 1; --j) {
    double temp = 1.0 / (double)(j * j * j * j);
    for (int i = 145; i >= j; --i) z[i - 1] += temp;
  }
  return z;
}

int ensure_tallies(mcrt_ctx* c) {
  if (!(c->has_grid && c->has_observer)) return fail(c, MCRT_E_STATE, "grid and observer must be set before tallies");
  DeviceModel& m = c->dm;
  c->ncell = (size_t)(m.nrg - 1) * (m.ntg - 1) * (m.npg - 1);
  c->nsed = (size_t)10 * m.nfreq * m.nmu * m.nph * m.nap * kNo;
  c->npsed = (size_t)9 * m.nfreq * m.npeel * m.nap * kNo;
  c->nimage = (size_t)8 * m.nband * m.nx * m.nx * m.npeel;
  void* p;
  int rc;
  if ((rc = dev_alloc(c, "t.dtauabs", c->ncell * 8, &p))) return rc;
  c->dt.dtauabs = (double*)p;
  if ((rc = dev_alloc(c, "t.nabs", c->ncell * 4, &p))) return rc;
  c->dt.nabs = (int*)p;
  if ((rc = dev_alloc(c, "t.dbig", c->ncell * 4, &p))) return rc;
  c->dt.dbig = (int*)p;
  if ((rc = dev_alloc(c, "t.killed", c->ncell * 4, &p))) return rc;
  c->dt.killed = (int*)p;
  if ((rc = dev_alloc(c, "t.sed", c->nsed * 4, &p))) return rc;
  c->dt.sed = (float*)p;
  if ((rc = dev_alloc(c, "t.psed", c->npsed * 4, &p))) return rc;
  c->dt.psed = (float*)p;
  if ((rc = dev_alloc(c, "t.image", c->nimage * 4, &p))) return rc;
  c->dt.image = (float*)p;
  if ((rc = dev_alloc(c, "t.counters", C_NCOUNTERS * 8, &p))) return rc;
  c->dt.counters = (unsigned long long*)p;
  if ((rc = dev_alloc(c, "t.diffcount", c->ncell * 4, &p))) return rc;
  c->dt.diff_count = (unsigned*)p;
  return MCRT_OK;
}

int zero_tallies(mcrt_ctx* c) {
  int rc = ensure_tallies(c);
  if (rc) return rc;
  HIPCHECK(c, launch_fill_f64(c->dt.dtauabs, c->ncell, 0.0, c->stream));
  HIPCHECK(c, launch_fill_i32(c->dt.nabs, c->ncell, 0, c->stream));
  HIPCHECK(c, launch_fill_i32(c->dt.dbig, c->ncell, 0, c->stream));
  HIPCHECK(c, launch_fill_i32(c->dt.killed, c->ncell, 0, c->stream));
  HIPCHECK(c, launch_fill_f32(c->dt.sed, c->nsed, 0.f, c->stream));
  HIPCHECK(c, launch_fill_f32(c->dt.psed, c->npsed, 0.f, c->stream));
  HIPCHECK(c, launch_fill_f32(c->dt.image, c->nimage, 0.f, c->stream));
  HIPCHECK(c, launch_fill_u64(c->dt.counters, C_NCOUNTERS, 0ull, c->stream));
  HIPCHECK(c, launch_fill_i32((int*)c->dt.diff_count, c->ncell, 0, c->stream));
  return MCRT_OK;
}

// ---- scratch of mcrt_setemiss_noalloc ------------------------------------
// A file-scope device array: every GPU the library runs on holds its own
// copy until the process ends.  On the host its name is only a stand-in, so
// host code uses the address hipGetSymbolAddress returns on each device
// (g_scratch[device], set up once per device by mcrt_create).  Layout: two
// 64-bit counters at 0, the per-species table lengths at kScrNlam and the
// smoothing flags at kScrSmooth (kScratchSpecies int32 each), then the
// interval integrals of one chunk of species at kScrSeg.
constexpr int kScratchSpecies = 1024;
constexpr size_t kScrNlam = 256, kScrSmooth = kScrNlam + 4u * kScratchSpecies;
constexpr size_t kScrSeg = 16384;
constexpr size_t kScratchSegBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kScrSeg + kScratchSegBytes];

constexpr int kMaxScratchDevices = 256;
unsigned char* g_scratch[kMaxScratchDevices] = {nullptr};
std::mutex g_scratch_mu;

// With `device` current: its copy of the array (once per device).
hipError_t scratch_init(int device) {
  if (device < 0 || device >= kMaxScratchDevices) return hipSuccess;   // the entry point then refuses
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return hipSuccess;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!ptr) return hipErrorInvalidSymbol;
  g_scratch[device] = (unsigned char*)ptr;
  return hipSuccess;
}

bool device_pointer(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess) { (void)hipGetLastError(); return false; }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

bool ready(mcrt_ctx* c) {
  return c->has_grid && c->has_medium && c->has_dust && c->has_source && c->has_observer && c->has_tdust;
}

}  // namespace

extern "C" {

int mcrt_create(int device, mcrt_ctx** out) {
  if (!out) return MCRT_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MCRT_E_NODEVICE;
  if (device < 0 || device >= n) return MCRT_E_ARG;
  mcrt_ctx* c = new mcrt_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) { delete c; return MCRT_E_HIP; }
  if (scratch_init(device) != hipSuccess) { delete c; return MCRT_E_HIP; }
  if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) { delete c; return MCRT_E_HIP; }
  c->stream = c->own_stream;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0) c->cu_count = prop.multiProcessorCount;
  (void)hipEventCreate(&c->ev0);
  (void)hipEventCreate(&c->ev1);
  for (auto& e : c->evr) (void)hipEventCreate(&e);
  for (auto& q : c->evp)
    for (auto& e : q) (void)hipEventCreate(&e);
  if (hipStreamCreateWithFlags(&c->peel_stream, hipStreamNonBlocking) != hipSuccess) {
    c->peel_stream = nullptr;
    mcrt_destroy(c);
    return MCRT_E_HIP;
  }
  if (hipHostMalloc((void**)&c->host_ctl, 2 * sizeof(RunCtl), hipHostMallocDefault) != hipSuccess) {
    c->host_ctl = nullptr;
    mcrt_destroy(c);
    return MCRT_E_HIP;
  }
  c->dm.max_segments = 1000000;
  c->dm.nisrf = 0;                          // external illumination off (mcrt_set_isrf)
  c->dm.isrf_from = ~0ull;
  c->dm.iintmax = 10000;
  c->dm.taumax = 25.0;
  std::vector<double> z = zeta_partial_sums();
  const double* dz;
  if (upload<double>(c, "inczeta", z.data(), z.size(), &dz) != MCRT_OK || hipStreamSynchronize(c->stream) != hipSuccess) {
    mcrt_destroy(c);
    return MCRT_E_HIP;
  }
  *out = c;
  return MCRT_OK;
}

int mcrt_destroy(mcrt_ctx* c) {
  if (!c) return MCRT_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->peel_stream) (void)hipStreamSynchronize(c->peel_stream);
  for (auto& kv : c->buf) (void)hipFree(kv.second);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  for (auto& e : c->evr) if (e) (void)hipEventDestroy(e);
  for (auto& q : c->evp)
    for (auto& e : q) if (e) (void)hipEventDestroy(e);
  if (c->host_ctl) (void)hipHostFree(c->host_ctl);
  if (c->b_exec) (void)hipGraphExecDestroy(c->b_exec);
  if (c->b_run_h) (void)hipHostFree(c->b_run_h);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  if (c->peel_stream) (void)hipStreamDestroy(c->peel_stream);
  delete c;
  return MCRT_OK;
}

const char* mcrt_last_error(const mcrt_ctx* c) { return c ? c->err.c_str() : "null context"; }

int mcrt_set_stream(mcrt_ctx* c, void* stream) {
  if (c) c->b_gen++;
  if (!c) return MCRT_E_ARG;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->stream = stream ? (hipStream_t)stream : c->own_stream;
  return MCRT_OK;
}

int mcrt_synchronize(mcrt_ctx* c) {
  if (!c) return MCRT_E_ARG;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return MCRT_OK;
}

int mcrt_set_grid(mcrt_ctx* c, int32_t nrg, int32_t ntg, int32_t npg, const double* rarr, const double* r2arr,
                  const double* thetarr, const double* costarr, const double* tan2arr, const double* phiarr,
                  const double* aarr, const double* barr, const double* vcell) {
  if (!c) return MCRT_E_ARG;
  if (nrg < 3 || ntg < 2 || npg < 2 || !rarr || !r2arr || !thetarr || !costarr || !tan2arr || !phiarr || !aarr || !barr || !vcell)
    return fail(c, MCRT_E_ARG, "mcrt_set_grid: need nrg>=3, ntg>=2, npg>=2 and all arrays");
  if (ntg > 2 && (ntg % 2) == 0) return fail(c, MCRT_E_ARG, "mcrt_set_grid: ntg must be odd (equator wall)");
  if (nrg > 1023 || ntg > 1023 || npg > 1023) return fail(c, MCRT_E_ARG, "mcrt_set_grid: at most 1023 walls per axis");
  double dphi_min = 2.0 * M_PI;
  if (npg > 2) {
    // testgrid takes phi = atan2 in [0, 2pi]: the walls must span exactly that
    if (std::fabs(phiarr[0]) > 1e-12 || std::fabs(phiarr[npg - 1] - 2.0 * M_PI) > 1e-9)
      return fail(c, MCRT_E_ARG, "mcrt_set_grid: phi walls must run from 0 to 2*pi");
    for (int k = 1; k < npg; ++k) {
      const double d = phiarr[k] - phiarr[k - 1];
      if (!(d > 0.0)) return fail(c, MCRT_E_ARG, "mcrt_set_grid: phi walls must increase");
      if (d < dphi_min) dphi_min = d;
    }
  }
  HIPCHECK(c, hipSetDevice(c->device));
  DeviceModel& m = c->dm;
  m.nrg = nrg; m.ntg = ntg; m.npg = npg;
  {
    const double e = std::sin(0.5 * 1.0e-4 * dphi_min);
    m.phi_tol2 = e * e;
  }
  size_t ncell = (size_t)(nrg - 1) * (ntg - 1) * (npg - 1);
  int rc = 0;
  rc |= upload(c, "rarr", rarr, nrg, &m.rarr);
  rc |= upload(c, "r2arr", r2arr, nrg, &m.r2arr);
  rc |= upload(c, "thetarr", thetarr, ntg, &m.thetarr);
  rc |= upload(c, "costarr", costarr, ntg, &m.costarr);
  rc |= upload(c, "tan2arr", tan2arr, ntg, &m.tan2arr);
  rc |= upload(c, "phiarr", phiarr, npg, &m.phiarr);
  rc |= upload(c, "aarr", aarr, npg, &m.aarr);
  rc |= upload(c, "barr", barr, npg, &m.barr);
  rc |= upload(c, "vcell", vcell, ncell, &m.vcell);
  {
    // double-cone half-angle t = min(theta, pi - theta) of each theta wall
    std::vector<double> cc(ntg), sc(ntg);
    for (int k = 0; k < ntg; ++k) {
      const double t = std::min(thetarr[k], M_PI - thetarr[k]);
      cc[k] = std::cos(t);
      sc[k] = std::sin(t);
    }
    rc |= upload(c, "ccone", cc.data(), ntg, &m.ccone);
    rc |= upload(c, "scone", sc.data(), ntg, &m.scone);
    if (rc) return rc;
    HIPCHECK(c, hipStreamSynchronize(c->stream));   // the host vectors are freed here
  }
  if (rc) return rc;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->has_grid = true;
  c->has_medium = c->has_tdust = false;   // per-cell arrays must match the new grid
  c->ncell = ncell;
  return MCRT_OK;
}

int mcrt_set_medium(mcrt_ctx* c, const double* rho, const uint8_t* species, const int32_t* species_opac) {
  if (!c) return MCRT_E_ARG;
  if (!c->has_grid) return fail(c, MCRT_E_STATE, "mcrt_set_medium: set the grid first");
  if (!rho || !species || !species_opac) return fail(c, MCRT_E_ARG, "mcrt_set_medium: null array");
  for (int i = 0; i <= 10; ++i)
    if (species_opac[i] < 0 || species_opac[i] >= kMaxOpac) return fail(c, MCRT_E_ARG, "mcrt_set_medium: species_opac out of range");
  for (size_t i = 0; i < c->ncell; ++i) {
    int s = species[i];
    if (s > 10) return fail(c, MCRT_E_ARG, "mcrt_set_medium: species code above 10");
    if (s >= 5 && s <= 8 && rho[i] > 0.0) return fail(c, MCRT_E_ARG, "mcrt_set_medium: small-grain species (5-8) are not supported");
    if (!(rho[i] >= 0.0)) return fail(c, MCRT_E_ARG, "mcrt_set_medium: negative or NaN density");
    if ((rho[i] > 0.0) != (s > 0)) return fail(c, MCRT_E_ARG, "mcrt_set_medium: a cell has density without species or species without density");
  }
  HIPCHECK(c, hipSetDevice(c->device));
  DeviceModel& m = c->dm;
  int rc = 0;
  rc |= upload(c, "rho64", rho, c->ncell, &m.rho64);
  rc |= upload_cast<float>(c, "rho32", rho, c->ncell, &m.rho32);
  rc |= upload(c, "species", species, c->ncell, &m.species);
  rc |= upload(c, "species_opac", (const int*)species_opac, 11, &m.species_opac);
  if (rc) return rc;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->has_medium = true;
  return MCRT_OK;
}

int mcrt_set_dust(mcrt_ctx* c, int32_t nopac, int32_t nlam, const double* nudust, const double* kappad, const double* adust,
                  const double* gdust, const double* pldust, const double* kappav, int32_t nT, int32_t nnu,
                  const double* lTint, const double* Tint, const double* nuint, const double* kapint, const double* kappap) {
  if (c) c->b_gen++;
  if (!c) return MCRT_E_ARG;
  if (nopac < 1 || nopac > kMaxOpac) return fail(c, MCRT_E_ARG, "mcrt_set_dust: nopac must be 1..4");
  if (nlam < 2 || nT < 2 || nnu < 2) return fail(c, MCRT_E_ARG, "mcrt_set_dust: nlam, nT, nnu must be >= 2");
  if (!nudust || !kappad || !adust || !gdust || !pldust || !kappav || !lTint || !Tint || !nuint || !kapint || !kappap)
    return fail(c, MCRT_E_ARG, "mcrt_set_dust: null array");
  HIPCHECK(c, hipSetDevice(c->device));
  DeviceModel& m = c->dm;
  m.nopac = nopac; m.nlam = nlam; m.nT = nT; m.nnu = nnu;
  size_t nl = (size_t)nopac * nlam;
  int rc = 0;
  rc |= upload(c, "nudust", nudust, nl, &m.nudust);
  rc |= upload(c, "kappad", kappad, nl, &m.kappad);
  rc |= upload(c, "adust", adust, nl, &m.adust);
  rc |= upload(c, "gdust", gdust, nl, &m.gdust);
  rc |= upload(c, "pldust", pldust, nl, &m.pldust);
  rc |= upload(c, "lTint", lTint, nT, &m.lTint);
  rc |= upload(c, "Tint", Tint, nT, &m.Tint);
  rc |= upload(c, "nuint", nuint, nnu, &m.nuint);
  {
    // the re-emission CDFs frequency-contiguous per temperature row
    // ([iop][iT][inu]; the reference's kapint(nT, nnu, nsp) has the
    // temperature fastest), so the entries one search probes share cache
    // lines; in FP64 and, for the mixed path, rounded to FP32
    const size_t n = (size_t)nopac * nT * nnu;
    std::vector<double> kt(n);
    std::vector<float> kf(n);
    for (int id = 0; id < nopac; ++id)
      for (int j = 0; j < nnu; ++j)
        for (int i = 0; i < nT; ++i) {
          const double v = kapint[((size_t)id * nnu + j) * nT + i];
          kt[((size_t)id * nT + i) * nnu + j] = v;
          kf[((size_t)id * nT + i) * nnu + j] = (float)v;
        }
    rc |= upload(c, "kapint", kt.data(), n, &m.kapint);
    rc |= upload(c, "kapint32", kf.data(), n, &m.kapint32);
    if (rc) return rc;
    HIPCHECK(c, hipStreamSynchronize(c->stream));   // kt, kf go out of scope
  }
  rc |= upload(c, "kappap", kappap, (size_t)nopac * nT, &m.kappap);
  if (rc) return rc;
  // guide table of the re-emission CDFs (lucydustfreq_xi): for every row,
  // the number of entries <= k/K, k = 0..K
  m.kguide = nullptr;
  m.kguide_k = 0;
  if (nnu <= 65535) {
    const int K = kGuideBuckets;
    std::vector<uint16_t> g((size_t)nopac * nT * (K + 1));
    for (int id = 0; id < nopac; ++id)
      for (int i = 0; i < nT; ++i) {
        const double* row = kapint + (size_t)id * nT * nnu + i;     // entry j at row[j * nT]
        uint16_t* out = g.data() + ((size_t)id * nT + i) * (K + 1);
        int n = 0;
        for (int k = 0; k <= K; ++k) {
          const double edge = (double)k / K;
          while (n < nnu && row[(size_t)n * nT] <= edge) ++n;
          out[k] = (uint16_t)n;
        }
      }
    rc = upload(c, "kguide", g.data(), g.size(), &m.kguide);
    if (rc) return rc;
    HIPCHECK(c, hipStreamSynchronize(c->stream));   // g goes out of scope
    m.kguide_k = K;
  }
  for (int i = 0; i < kMaxOpac; ++i) m.kappav[i] = (i < nopac) ? kappav[i] : 1.0;
  // opacset's index guess: only for a strictly decreasing, log-uniform nudust
  for (int id = 0; id < kMaxOpac; ++id) {
    m.nud_log0[id] = 0.0;
    m.nud_ilog[id] = 0.0;
    if (id >= nopac) continue;
    const double* nd = nudust + (size_t)id * nlam;
    bool ok = nd[0] > 0.0 && nd[nlam - 1] > 0.0;
    const double step = ok ? std::log(nd[0] / nd[nlam - 1]) / (nlam - 1) : 0.0;
    for (int k = 1; ok && k < nlam; ++k) {
      if (!(nd[k] < nd[k - 1])) ok = false;
      else if (std::fabs(std::log(nd[k - 1] / nd[k]) - step) > 1e-3 * step) ok = false;
    }
    if (ok && step > 0.0) {
      m.nud_log0[id] = std::log(nd[0]);
      m.nud_ilog[id] = 1.0 / step;
    }
  }
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->has_dust = true;
  return MCRT_OK;
}

int mcrt_setemiss(mcrt_ctx* c, int32_t nopac, int32_t nlam, const int32_t* nlambda, const double* nudust,
                  const double* kappad, const double* adust, int32_t nT, const double* Tint, int32_t nnu,
                  const double* nuint, const int32_t* smooth, double* kapint, double* kappap, double* kappar,
                  int64_t* nfail, int64_t* nevals, float* ms) {
  if (!c) return MCRT_E_ARG;
  if (nopac < 1 || nlam < 2 || nT < 2 || nnu < 2) return fail(c, MCRT_E_ARG, "mcrt_setemiss: nopac >= 1, nlam, nT, nnu >= 2");
  if (nT > 4096) return fail(c, MCRT_E_ARG, "mcrt_setemiss: nT must be <= 4096 (kappap smoothing in LDS)");
  if (!nudust || !kappad || !adust || !Tint || !nuint || !kapint || !kappap || !kappar)
    return fail(c, MCRT_E_ARG, "mcrt_setemiss: null array");
  std::vector<int> nl(nopac, nlam), sm(nopac, 0);
  for (int id = 0; id < nopac; ++id) {
    if (nlambda) nl[id] = nlambda[id];
    if (smooth) sm[id] = smooth[id] ? 1 : 0;
    if (nl[id] < 2 || nl[id] > nlam) return fail(c, MCRT_E_ARG, "mcrt_setemiss: nlambda must be 2..nlam");
    if (sm[id] && nT < 21) return fail(c, MCRT_E_ARG, "mcrt_setemiss: smoothed species need nT >= 21");
  }
  for (int i = 0; i < nT; ++i)
    if (!(Tint[i] > 0.0)) return fail(c, MCRT_E_ARG, "mcrt_setemiss: Tint must be > 0");
  HIPCHECK(c, hipSetDevice(c->device));
  const size_t nd = (size_t)nopac * nlam, nseg = (size_t)nopac * nnu * nT, nrow = (size_t)nopac * nT;
  const size_t sizes[] = {nopac * sizeof(int), nopac * sizeof(int), nd * 8, nd * 8, nd * 8, (size_t)nT * 8,
                          (size_t)nnu * 8, 2 * nseg * 8, nseg * 8, nrow * 8, nrow * 8, 16};
  const int nbuf = (int)(sizeof(sizes) / sizeof(sizes[0]));
  void* p[12] = {nullptr};
  hipError_t e = hipSuccess;
  for (int k = 0; k < nbuf && e == hipSuccess; ++k) e = hipMalloc(&p[k], sizes[k]);
  float t_ms = 0.0f;
  unsigned long long counts[2] = {0, 0};      // non-converged qsimp calls, integrand evaluations
  if (e == hipSuccess) {
    const void* src[] = {nl.data(), sm.data(), nudust, kappad, adust, Tint, nuint};
    for (int k = 0; k < 7 && e == hipSuccess; ++k) e = hipMemcpyAsync(p[k], src[k], sizes[k], hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(p[11], 0, 16, c->stream);
    EmissArgs a{nopac, nlam, nT, nnu, (const int*)p[0], (const double*)p[2], (const double*)p[3], (const double*)p[4],
                (const double*)p[5], (const double*)p[6], (const int*)p[1], (double*)p[7], (double*)p[8], (double*)p[9],
                (double*)p[10], (unsigned long long*)p[11], (unsigned long long*)p[11] + 1};
    if (e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
    if (e == hipSuccess) e = launch_setemiss(a, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(kapint, p[8], sizes[8], hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(kappap, p[9], sizes[9], hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(kappar, p[10], sizes[10], hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(counts, p[11], 16, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipEventElapsedTime(&t_ms, c->ev0, c->ev1);
  }
  for (int k = 0; k < nbuf; ++k)
    if (p[k]) (void)hipFree(p[k]);
  if (e != hipSuccess) return hipfail(c, e, "mcrt_setemiss");
  if (nfail) *nfail = (int64_t)counts[0];
  if (nevals) *nevals = (int64_t)counts[1];
  if (ms) *ms = t_ms;
  return MCRT_OK;
}

int mcrt_setemiss_noalloc(mcrt_ctx* c, int32_t nopac, int32_t nlam, const int32_t* nlambda, const double* nudust,
                          const double* kappad, const double* adust, int32_t nT, const double* Tint, int32_t nnu,
                          const double* nuint, const int32_t* smooth, double* kapint, double* kappap, double* kappar,
                          int64_t* nfail, int64_t* nevals) {
  if (!c) return MCRT_E_ARG;
  const char* me = "mcrt_setemiss_noalloc";
  if (nopac < 1 || nlam < 2 || nT < 2 || nnu < 2) return fail(c, MCRT_E_ARG, std::string(me) + ": nopac >= 1, nlam, nT, nnu >= 2");
  if (nopac > kScratchSpecies)
    return fail(c, MCRT_E_ARG, std::string(me) + ": nopac must be <= 1024 (the scratch holds 1024 species' flags)");
  if (nT > 4096) return fail(c, MCRT_E_ARG, std::string(me) + ": nT must be <= 4096 (kappap smoothing in LDS)");
  const size_t per_species = 2u * (size_t)nnu * (size_t)nT * sizeof(double);
  if (per_species > kScratchSegBytes)
    return fail(c, MCRT_E_ARG, std::string(me) + ": 2*nnu*nT doubles of one species exceed the 16 MiB scratch "
                               "(nnu*nT must be <= 1048576)");
  const void* dev[] = {nudust, kappad, adust, Tint, nuint, kapint, kappap, kappar};
  for (const void* q : dev)
    if (!device_pointer(q)) return fail(c, MCRT_E_ARG, std::string(me) + ": tables must be device pointers");
  HIPCHECK(c, hipSetDevice(c->device));
  unsigned char* s = (c->device < kMaxScratchDevices) ? g_scratch[c->device] : nullptr;
  if (!s) return fail(c, MCRT_E_STATE, std::string(me) + ": no scratch on this device");
  int32_t nl[kScratchSpecies], sm[kScratchSpecies];
  for (int id = 0; id < nopac; ++id) {
    nl[id] = nlambda ? nlambda[id] : nlam;
    sm[id] = (smooth && smooth[id]) ? 1 : 0;
    if (nl[id] < 2 || nl[id] > nlam) return fail(c, MCRT_E_ARG, std::string(me) + ": nlambda must be 2..nlam");
    if (sm[id] && nT < 21) return fail(c, MCRT_E_ARG, std::string(me) + ": smoothed species need nT >= 21");
  }
  double th[4096];
  HIPCHECK(c, hipMemcpyAsync(th, Tint, (size_t)nT * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < nT; ++i)
    if (!(th[i] > 0.0)) return fail(c, MCRT_E_ARG, std::string(me) + ": Tint must be > 0");
  int* d_nl = (int*)(s + kScrNlam);
  int* d_sm = (int*)(s + kScrSmooth);
  unsigned long long* cnt = (unsigned long long*)s;
  double* seg = (double*)(s + kScrSeg);
  HIPCHECK(c, hipMemcpyAsync(d_nl, nl, (size_t)nopac * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(d_sm, sm, (size_t)nopac * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, launch_fill_u64(cnt, 2, 0ull, c->stream));
  const int chunk = (int)std::min<size_t>((size_t)nopac, kScratchSegBytes / per_species);
  for (int id0 = 0; id0 < nopac; id0 += chunk) {
    const int n = std::min(chunk, nopac - id0);
    const size_t t0 = (size_t)id0 * nlam, k0 = (size_t)id0 * nnu * nT, r0 = (size_t)id0 * nT;
    EmissArgs a{n, nlam, nT, nnu, d_nl + id0, nudust + t0, kappad + t0, adust + t0, Tint, nuint, d_sm + id0, seg,
                kapint + k0, kappap + r0, kappar + r0, cnt, cnt + 1};
    HIPCHECK(c, launch_setemiss(a, c->stream));
  }
  unsigned long long counts[2] = {0, 0};
  HIPCHECK(c, hipMemcpyAsync(counts, cnt, sizeof(counts), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (nfail) *nfail = (int64_t)counts[0];
  if (nevals) *nevals = (int64_t)counts[1];
  return MCRT_OK;
}

int mcrt_build_models(mcrt_ctx* c, int32_t nmodels, const int32_t* nrg, const int32_t* ntg, const int32_t* npg,
                      const double* rarr, const double* thetarr, const double* phiarr, const double* params,
                      const double* clump, double* rho, uint8_t* species, double* vcell, int64_t* nflag, float* ms) {
  if (!c) return MCRT_E_ARG;
  if (nmodels < 1 || nmodels > 65535) return fail(c, MCRT_E_ARG, "mcrt_build_models: nmodels must be 1..65535");
  if (!nrg || !ntg || !npg || !rarr || !thetarr || !phiarr || !params || !rho || !species || !vcell)
    return fail(c, MCRT_E_ARG, "mcrt_build_models: null array");
  std::vector<int> nr(nmodels), nt(nmodels), np(nmodels);
  std::vector<long long> roff(nmodels), toff(nmodels), poff(nmodels), coff(nmodels);
  long long nrw = 0, ntw = 0, npw = 0, ncell = 0, maxc = 0;
  for (int k = 0; k < nmodels; ++k) {
    if (nrg[k] < 2 || ntg[k] < 2 || npg[k] < 2) return fail(c, MCRT_E_ARG, "mcrt_build_models: nrg, ntg, npg must be >= 2");
    nr[k] = nrg[k] - 1; nt[k] = ntg[k] - 1; np[k] = npg[k] - 1;
    roff[k] = nrw; toff[k] = ntw; poff[k] = npw; coff[k] = ncell;
    nrw += nrg[k]; ntw += ntg[k]; npw += npg[k];
    const long long n = (long long)nr[k] * nt[k] * np[k];
    ncell += n;
    maxc = std::max(maxc, n);
  }
  const unsigned nblk = (unsigned)((maxc + 255) / 256);
  if ((long long)nblk * 256 < maxc || nblk > 0x7fffffffu) return fail(c, MCRT_E_ARG, "mcrt_build_models: model too large");
  HIPCHECK(c, hipSetDevice(c->device));
  const size_t isz = (size_t)nmodels * sizeof(int), lsz = (size_t)nmodels * sizeof(long long);
  const size_t sizes[] = {isz, isz, isz, lsz, lsz, lsz, lsz, (size_t)nrw * 8, (size_t)ntw * 8, (size_t)npw * 8,
                          (size_t)nmodels * kBuildPar * 8, clump ? (size_t)ncell * 8 : 8, (size_t)ncell * 8,
                          (size_t)ncell * 8, (size_t)ncell * 8, (size_t)ncell, (size_t)nmodels * nblk * 8,
                          (size_t)nmodels * 8, 8};
  const int nbuf = (int)(sizeof(sizes) / sizeof(sizes[0]));
  void* p[19] = {nullptr};
  hipError_t e = hipSuccess;
  for (int k = 0; k < nbuf && e == hipSuccess; ++k) e = hipMalloc(&p[k], sizes[k]);
  float t_ms = 0.0f;
  unsigned long long flags = 0;
  if (e == hipSuccess) {
    const void* src[] = {nr.data(), nt.data(), np.data(), roff.data(), toff.data(), poff.data(), coff.data(),
                         rarr, thetarr, phiarr, params, clump};
    for (int k = 0; k < 12 && e == hipSuccess; ++k)
      if (src[k]) e = hipMemcpyAsync(p[k], src[k], sizes[k], hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(p[18], 0, 8, c->stream);
    BuildArgs a{nmodels, nblk, (const int*)p[0], (const int*)p[1], (const int*)p[2], (const long long*)p[3],
                (const long long*)p[4], (const long long*)p[5], (const long long*)p[6], (const double*)p[7],
                (const double*)p[8], (const double*)p[9], (const double*)p[10], clump ? (const double*)p[11] : nullptr,
                (double*)p[12], (double*)p[13], (double*)p[14], (uint8_t*)p[15], (double*)p[16], (double*)p[17],
                (unsigned long long*)p[18]};
    if (e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
    if (e == hipSuccess) e = launch_build(a, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(rho, p[12], sizes[12], hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(vcell, p[13], sizes[13], hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(species, p[15], sizes[15], hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&flags, p[18], 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipEventElapsedTime(&t_ms, c->ev0, c->ev1);
  }
  for (int k = 0; k < nbuf; ++k)
    if (p[k]) (void)hipFree(p[k]);
  if (e != hipSuccess) return hipfail(c, e, "mcrt_build_models");
  if (nflag) *nflag = (int64_t)flags;
  if (ms) *ms = t_ms;
  return MCRT_OK;
}

int mcrt_set_isrf(mcrt_ctx* c, int32_t nisrf, const double* cdf, const double* logcdf, const double* logfreq,
                  double isrf_frac) {
  if (!c) return MCRT_E_ARG;
  if (!(isrf_frac >= 0.0 && isrf_frac < 1.0)) return fail(c, MCRT_E_ARG, "mcrt_set_isrf: isrf_frac must be in [0, 1)");
  if (isrf_frac == 0.0) {
    c->isrf_frac = 0.0;
    c->nisrf = 0;
    return MCRT_OK;
  }
  if (nisrf < 2 || !cdf || !logcdf || !logfreq) return fail(c, MCRT_E_ARG, "mcrt_set_isrf: nisrf >= 2 tables needed");
  for (int i = 1; i < nisrf; ++i)
    if (!(cdf[i] >= cdf[i - 1])) return fail(c, MCRT_E_ARG, "mcrt_set_isrf: the CDF must be non-decreasing");
  if (cdf[0] != 0.0 || cdf[nisrf - 1] != 1.0) return fail(c, MCRT_E_ARG, "mcrt_set_isrf: the CDF must run from 0 to 1");
  HIPCHECK(c, hipSetDevice(c->device));
  int rc = 0;
  rc |= upload(c, "isrf.cdf", cdf, (size_t)nisrf, &c->isrf_cdf);
  rc |= upload(c, "isrf.logcdf", logcdf, (size_t)nisrf, &c->isrf_logcdf);
  rc |= upload(c, "isrf.logfreq", logfreq, (size_t)nisrf, &c->isrf_logfreq);
  if (rc) return rc;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->nisrf = nisrf;
  c->isrf_frac = isrf_frac;
  return MCRT_OK;
}

int mcrt_set_spots(mcrt_ctx* c, int32_t nspot, double fspot, double spotlat, double spotlon, double tshock,
                   double accfrac2) {
  if (!c) return MCRT_E_ARG;
  if (nspot == 0) {
    c->spot.ispot = 0;
    c->accfrac2 = 0.0;
    return MCRT_OK;
  }
  if (nspot != 1 && nspot != 2) return fail(c, MCRT_E_ARG, "mcrt_set_spots: nspot must be 0, 1 or 2");
  if (!(fspot > 0.0 && fspot <= 1.0)) return fail(c, MCRT_E_ARG, "mcrt_set_spots: fspot must be in (0, 1]");
  if (!(tshock > 0.0)) return fail(c, MCRT_E_ARG, "mcrt_set_spots: tshock must be > 0");
  if (!(accfrac2 >= 0.0 && accfrac2 < 1.0)) return fail(c, MCRT_E_ARG, "mcrt_set_spots: accfrac2 must be in [0, 1)");
  const double pi = 3.1415926535897932384626433, deg2rad = pi / 180.0, rad2deg = 180.0 / pi;
  // gridset_5_model13new.f90:159-165: the spot radius in degrees from the filling factor
  const double thspot = (nspot == 1 ? std::acos(1.0 - 2.0 * fspot) : std::acos(1.0 - fspot)) * rad2deg;
  // spotset.f90:9-27
  const double th0sp1 = (90.0 - spotlat) * deg2rad, phi0sp1 = spotlon * deg2rad;
  const double th0sp2 = pi - th0sp1, phi0sp2 = phi0sp1 + pi;
  DeviceModel& d = c->spot;
  d.ispot = 1;
  d.nspot = nspot;
  d.fspot = fspot;
  d.tshock = tshock;
  d.scspot1 = std::sin(th0sp1) * std::cos(phi0sp1);
  d.ssspot1 = std::sin(th0sp1) * std::sin(phi0sp1);
  d.cspot1 = std::cos(th0sp1);
  d.sspot1 = std::sin(th0sp1);
  d.scspot2 = std::sin(th0sp2) * std::cos(phi0sp2);
  d.ssspot2 = std::sin(th0sp2) * std::sin(phi0sp2);
  d.cspot2 = std::cos(th0sp2);
  d.sspot2 = std::sin(th0sp2);
  d.csprad = std::cos(thspot * deg2rad);
  c->accfrac2 = accfrac2;
  return MCRT_OK;
}

int mcrt_set_accretion(mcrt_ctx* c, const double* ldint, const double* hdisk, double accfrac) {
  if (!c) return MCRT_E_ARG;
  if (!(accfrac >= 0.0 && accfrac < 1.0)) return fail(c, MCRT_E_ARG, "mcrt_set_accretion: accfrac must be in [0, 1)");
  if (accfrac == 0.0) {
    c->accfrac = 0.0;
    return MCRT_OK;
  }
  if (!c->has_grid) return fail(c, MCRT_E_STATE, "mcrt_set_accretion: set the grid first");
  if (!ldint || !hdisk) return fail(c, MCRT_E_ARG, "mcrt_set_accretion: null table");
  const int n = c->dm.nrg - 1;
  if (n < 2) return fail(c, MCRT_E_ARG, "mcrt_set_accretion: at least two radial cells needed");
  // the cell j with Ldint(j) <= xi < Ldint(j+1) must exist for every xi in [0, 1)
  if (ldint[0] != 0.0 || ldint[n - 1] != 1.0)
    return fail(c, MCRT_E_ARG, "mcrt_set_accretion: Ldint must run from 0 (first cell) to 1 (last cell)");
  for (int i = 1; i < n; ++i)
    if (!(ldint[i] >= ldint[i - 1])) return fail(c, MCRT_E_ARG, "mcrt_set_accretion: Ldint must be non-decreasing");
  for (int i = 0; i < n; ++i)
    if (!(hdisk[i] >= 0.0) || !std::isfinite(hdisk[i]))
      return fail(c, MCRT_E_ARG, "mcrt_set_accretion: hdisk must be finite and >= 0");
  HIPCHECK(c, hipSetDevice(c->device));
  int rc = 0;
  rc |= upload(c, "acc.ldint", ldint, (size_t)n, &c->acc_ldint);
  rc |= upload(c, "acc.hdisk", hdisk, (size_t)n, &c->acc_hdisk);
  if (rc) return rc;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->acc_n = n;
  c->accfrac = accfrac;
  return MCRT_OK;
}

int mcrt_set_source(mcrt_ctx* c, double rstar, double tstar, double rmin) {
  if (c) c->b_gen++;
  if (!c) return MCRT_E_ARG;
  if (!(rstar > 0.0) || !(tstar > 0.0) || !(rmin > 0.0)) return fail(c, MCRT_E_ARG, "mcrt_set_source: rstar, tstar, rmin must be > 0");
  c->dm.rstar = rstar; c->dm.tstar = tstar; c->dm.rmin = rmin;
  c->has_source = true;
  return MCRT_OK;
}

int mcrt_set_observer(mcrt_ctx* c, int32_t ipeel, int32_t npeel, const double* thete, const double* phie, int32_t nx,
                      double rmaxi, int32_t nband, int32_t nwmax, const int32_t* band_nw, const double* band_freq,
                      const double* band_phi, int32_t nfreq, int32_t nmu, int32_t nph, int32_t nap,
                      const double* aperture2, const double* u_mu, double numin, double lnurat) {
  if (c) c->b_gen++;
  if (!c) return MCRT_E_ARG;
  if (npeel < 0 || (ipeel == 1 && npeel < 1) || nx < 1 || nband < 0 || nfreq < 1 || nmu < 1 || nph < 1 || nap < 1)
    return fail(c, MCRT_E_ARG, "mcrt_set_observer: bad sizes");
  if (npeel > 0 && (!thete || !phie)) return fail(c, MCRT_E_ARG, "mcrt_set_observer: null peel angles");
  if (nband > 0 && (!band_nw || !band_freq || !band_phi || nwmax < 3)) return fail(c, MCRT_E_ARG, "mcrt_set_observer: bad band tables");
  for (int b = 0; b < nband; ++b)
    if (band_nw[b] < 3 || band_nw[b] > nwmax) return fail(c, MCRT_E_ARG, "mcrt_set_observer: band_nw must be 3..nwmax");
  if (!aperture2 || !u_mu || !(numin > 0.0) || !(lnurat > 0.0)) return fail(c, MCRT_E_ARG, "mcrt_set_observer: bad SED binning");
  HIPCHECK(c, hipSetDevice(c->device));
  DeviceModel& m = c->dm;
  std::vector<double> pe((size_t)5 * (npeel > 0 ? npeel : 1), 0.0);
  for (int k = 0; k < npeel; ++k) {
    pe[5 * k + 0] = cos(thete[k]);
    pe[5 * k + 1] = sin(thete[k]);
    pe[5 * k + 2] = cos(phie[k]);
    pe[5 * k + 3] = sin(phie[k]);
    pe[5 * k + 4] = phie[k];
  }
  std::vector<double> sini(nmu);
  for (int i = 0; i < nmu; ++i) sini[i] = sqrt(1.0 - u_mu[i] * u_mu[i]);   // newdisktherm_5_fuv.f90:114-116
  m.ipeel = ipeel; m.npeel = npeel; m.nx = nx; m.rmaxi = rmaxi;
  m.nband = nband; m.nwmax = nwmax;
  m.nfreq = nfreq; m.nmu = nmu; m.nph = nph; m.nap = nap;
  m.numin = numin; m.lnurat = lnurat;
  int rc = 0;
  rc |= upload(c, "peel", pe.data(), pe.size(), &m.peel);
  if (nband > 0) {
    rc |= upload(c, "band_nw", (const int*)band_nw, nband, &m.band_nw);
    rc |= upload(c, "band_freq", band_freq, (size_t)nband * nwmax, &m.band_freq);
    rc |= upload(c, "band_phi", band_phi, (size_t)nband * nwmax, &m.band_phi);
  }
  rc |= upload(c, "aperture2", aperture2, nap, &m.aperture2);
  rc |= upload(c, "u_mu", u_mu, nmu, &m.u_mu);
  rc |= upload(c, "sini_mu", sini.data(), nmu, &m.sini_mu);
  if (rc) return rc;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->has_observer = true;
  return MCRT_OK;
}

int mcrt_set_tdust(mcrt_ctx* c, const double* tdust) {
  if (!c) return MCRT_E_ARG;
  if (!c->has_grid) return fail(c, MCRT_E_STATE, "mcrt_set_tdust: set the grid first");
  if (!tdust) return fail(c, MCRT_E_ARG, "mcrt_set_tdust: null array");
  HIPCHECK(c, hipSetDevice(c->device));
  int rc = upload(c, "tdust", tdust, c->ncell, &c->dm.tdust);
  if (rc) return rc;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->has_tdust = true;
  return MCRT_OK;
}

int mcrt_zero_tallies(mcrt_ctx* c) {
  if (!c) return MCRT_E_ARG;
  HIPCHECK(c, hipSetDevice(c->device));
  return zero_tallies(c);
}

int mcrt_run_packets(mcrt_ctx* c, int64_t first_packet, int64_t npackets, uint64_t seed, int32_t precision) {
  if (!c) return MCRT_E_ARG;
  if (!ready(c)) return fail(c, MCRT_E_STATE, "mcrt_run_packets: grid, medium, dust, source, observer and tdust must be set");
  if (npackets < 0 || first_packet < 0) return fail(c, MCRT_E_ARG, "mcrt_run_packets: negative packet range");
  if (precision != MCRT_FP64 && precision != MCRT_MIXED) return fail(c, MCRT_E_ARG, "mcrt_run_packets: unknown precision");
  if (c->trace_n > 0 && npackets > c->trace_n) return fail(c, MCRT_E_ARG, "mcrt_run_packets: more packets than the trace buffer holds");
  HIPCHECK(c, hipSetDevice(c->device));
  if (!c->dt.dtauabs) {
    int rc = zero_tallies(c);
    if (rc) return rc;
  }
  if (npackets == 0) return MCRT_OK;
  DeviceTallies t = c->dt;
  t.trace = (c->trace_n > 0) ? (double*)c->buf["trace"] : nullptr;
  const bool fp64 = (precision == MCRT_FP64);
  // newdisktherm_5_fuv.f90:158-187, :242, :296: of np packets the first
  // npacc - 1 = int(accfrac np) - 1 are disk-accretion photons, the next
  // npaccstar = int(accfrac2 np) hot-spot photons (npaccstar - 1 without
  // accretion photons), the last npout = nint(isrf_frac np) come from the
  // external field
  DeviceModel dm = c->dm;
  int64_t npacc = 0;
  if (c->accfrac > 0.0) {
    if (c->acc_n != c->dm.nrg - 1) return fail(c, MCRT_E_STATE, "mcrt_run_packets: the accretion tables predate the grid");
    npacc = (int64_t)(c->accfrac * (double)npackets);
    dm.acc_to = (unsigned long long)(npacc > 1 ? npacc - 1 : 0);
    dm.acc_ldint = c->acc_ldint;
    dm.acc_hdisk = c->acc_hdisk;
  }
  if (c->nisrf > 0 && c->isrf_frac > 0.0) {
    int64_t npout = (int64_t)(c->isrf_frac * (double)npackets);
    if (c->isrf_frac * (double)npackets - (double)npout > 0.5) npout += 1;
    dm.nisrf = c->nisrf;
    dm.isrf_cdf = c->isrf_cdf;
    dm.isrf_logcdf = c->isrf_logcdf;
    dm.isrf_logfreq = c->isrf_logfreq;
    dm.isrf_from = (unsigned long long)(npackets - npout);
  }
  if (c->diff_on) {
    if (!c->diff_set) return fail(c, MCRT_E_STATE, "mcrt_run_packets: diffusion is on: call mcrt_setdiffus first");
    // ns = np - npacc - npout, the stellar photons (newdisktherm_5_fuv.f90:158-190),
    // whose count scales the diffusion surface temperatures
    int64_t nout = 0;
    if (c->nisrf > 0 && c->isrf_frac > 0.0) {
      nout = (int64_t)(c->isrf_frac * (double)npackets);
      if (c->isrf_frac * (double)npackets - (double)nout > 0.5) nout += 1;
    }
    const int64_t ns = npackets - (int64_t)(c->accfrac * (double)npackets) - nout;
    if (ns < 1) return fail(c, MCRT_E_ARG, "mcrt_run_packets: no stellar photons to scale the diffusion surfaces");
    dm.diff_delta = 4.0 * kPi * std::pow(c->dm.tstar, 4.0) / (double)ns;
    c->diff_delta_run = dm.diff_delta;
  }
  if (c->spot.ispot) {
    const DeviceModel& sp = c->spot;
    dm.ispot = 1; dm.nspot = sp.nspot; dm.fspot = sp.fspot; dm.csprad = sp.csprad; dm.tshock = sp.tshock;
    dm.scspot1 = sp.scspot1; dm.ssspot1 = sp.ssspot1; dm.cspot1 = sp.cspot1; dm.sspot1 = sp.sspot1;
    dm.scspot2 = sp.scspot2; dm.ssspot2 = sp.ssspot2; dm.cspot2 = sp.cspot2; dm.sspot2 = sp.sspot2;
    // the photons iphot < npacc + npaccstar (1-based) that are not accretion
    // photons are hot-spot photons
    const int64_t npaccstar = (int64_t)(c->accfrac2 * (double)npackets);
    dm.hot_to = (unsigned long long)(npacc + npaccstar > 1 ? npacc + npaccstar - 1 : 0);
  }
  // Resident lanes: as many 256-thread blocks per CU as the transport
  // kernel's registers and LDS allow (never more than the packets need).
  // Each round every lane runs events while it can reserve slots in the
  // round's shared peel-off queue (one record per event); the peel kernel
  // then walks the round's records.  Packets alive at the end of a round
  // stay parked in their lane.
  if (lds_always_doubles(c->dm) > kLdsAlways)
    return fail(c, MCRT_E_ARG, "mcrt_run_packets: walls, Planck grids and observer directions exceed the 56 KB LDS budget "
                               "(2*nrg + 4*ntg + 3*npg + nT + nnu + 5*npeel must stay below 7023)");
  const LdsPlan lds = plan_lds(c->dm);
  dm.dust_lds = lds.n[LT_NUDUST] > 0 ? 1 : 0;
  int64_t lanes = c->lanes;
  if (lanes <= 0) {
    int bpc = fp64 ? transport_blocks_per_cu_fp64(dm, lds.bytes) : transport_blocks_per_cu_mixed(dm, lds.bytes);
    if (bpc < 1) bpc = 1;
    lanes = (int64_t)c->cu_count * bpc * 256;
  }
  const int64_t need = ((npackets + 255) / 256) * 256;
  if (lanes > need) lanes = need;
  lanes = ((lanes + 255) / 256) * 256;
  const bool queued = c->dm.ipeel == 1 && c->dm.npeel > 0;
  int64_t per = c->queue_per_lane > 0 ? c->queue_per_lane : 80;
  if (per < (int64_t)kSlotChunk) per = (int64_t)kSlotChunk;
  int64_t capacity = queued ? lanes * per : 0;
  if (capacity > ((int64_t)1 << 30)) capacity = ((int64_t)1 << 30) / lanes * lanes;
  int64_t events = c->events_per_round > 0 ? c->events_per_round : ((int64_t)1 << 20);
  const int words = fp64 ? lane_words_fp64() : lane_words_mixed();
  void* p;
  int rc;
  // ctl[0]: packets handed out, lanes alive, queue 0's count; ctl[1]: queue 1's count
  if ((rc = dev_alloc(c, "run.ctl", 2 * sizeof(RunCtl), &p))) return rc;
  RunCtl* ctl = (RunCtl*)p;
  if ((rc = dev_alloc(c, fp64 ? "lanes.words64" : "lanes.words32", (size_t)lanes * words * 4, &p))) return rc;
  uint32_t* lane_words = (uint32_t*)p;
  if ((rc = dev_alloc(c, "lanes.live", (size_t)lanes, &p))) return rc;
  uint8_t* lane_live = (uint8_t*)p;
  // Two peel-off queues: round r fills queue r % 2 while the peel kernels
  // of round r - 1 walk the other one.
  PeelQueue qb[2] = {{nullptr, &ctl[0].q_count, 0}, {nullptr, &ctl[1].q_count, 0}};
  if (queued) {
    const size_t rb = fp64 ? peel_record_bytes_fp64() : peel_record_bytes_mixed();
    static const char* qn[2][2] = {{"q.recs32", "q.recs32b"}, {"q.recs64", "q.recs64b"}};
    for (int b = 0; b < 2; ++b) {
      if ((rc = dev_alloc(c, qn[fp64 ? 1 : 0][b], (size_t)capacity * rb, &p))) return rc;
      qb[b].recs = p;
      qb[b].capacity = (unsigned)capacity;
    }
  }
  if (dm.acc_to > 0) {
    if ((rc = dev_alloc(c, "acc.start", (size_t)dm.acc_to * sizeof(AccStart), &p))) return rc;
    dm.acc_start = (const AccStart*)p;
  }
  const int pgrid = c->cu_count * (c->peel_blocks_per_cu > 0 ? c->peel_blocks_per_cu : 8);
  // The mixed path walks its peel-offs in three kernels (mcrt_peel.h) unless
  // MCRT_PEEL_SPLIT=0 asks for the one-kernel k_peel.  Ray states and hits
  // are sized for every (record, direction) item of a full round.
  bool split = queued && !fp64;
  if (const char* ps = std::getenv("MCRT_PEEL_SPLIT")) split = split && std::atoi(ps) != 0;
  PeelRays rays{};
  PeelHits hits{};
  unsigned* split_ctr = nullptr;
  int wgrid = 0;
  size_t chunk = 0;
  const int sgrid = pgrid * 4;            // blocks of k_peel_start
  // blocks of k_peel_hits: a part of the hit list per two CUs, one block per
  // direction and part (each block sums its direction's peeled SED in LDS,
  // so fewer, fuller blocks flush fewer bins)
  const int hgrid = (c->cu_count / 2 > 0 ? c->cu_count / 2 : 1) * (c->dm.npeel > 0 ? c->dm.npeel : 1);
  // (an item is named by a 32-bit index within its round)
  const size_t round_items = (size_t)capacity * (size_t)c->dm.npeel;
  if (round_items >= (size_t)0xFFFFFFFFu) split = false;
  if (split) {
    // a round's items in chunks of at most 128 Mi (about 21 GB of ray states
    // and hits): a full round of the 200x180x64 workload (105 M items) in
    // one chunk, so the walk kernel drains once per round (peel kernel time
    // per 10^8-packet step 21.5 -> 17.1 s against chunks of 64 Mi)
    size_t max_items = (size_t)128 << 20;
    if (const char* pc = std::getenv("MCRT_PEEL_CHUNK")) {     // smaller chunks, for tests
      const long long v = std::atoll(pc);
      if (v >= 256 && (size_t)v < max_items) max_items = (size_t)v;
    }
    chunk = round_items < max_items ? round_items : max_items;
    int bpc = peel_walk_blocks_per_cu_mixed(plan_lds(c->dm).bytes);
    if (bpc < 1) bpc = 1;
    wgrid = c->cu_count * bpc;
    // every wavefront of k_peel_start and k_peel_walk reserves slots 64 at a
    // time and may leave up to 63 of them unused
    const size_t n = chunk + (size_t)64 * 4 * ((size_t)sgrid + (size_t)wgrid);
    double** d[12] = {&rays.ox, &rays.oy, &rays.oz, &rays.b, &rays.h2, &rays.t, &rays.tr, &rays.tlo, &rays.tlo2,
                      &rays.thi, &rays.thi2, &rays.tp};
    static const char* dn[12] = {"ray.ox", "ray.oy", "ray.oz", "ray.b", "ray.h2", "ray.t", "ray.tr", "ray.tlo",
                                 "ray.tlo2", "ray.thi", "ray.thi2", "ray.tp"};
    for (int i = 0; i < 12; ++i) {
      if ((rc = dev_alloc(c, dn[i], n * 8, &p))) return rc;
      *d[i] = (double*)p;
    }
    static const char* kn[kMaxOpac] = {"ray.kapd0", "ray.kapd1", "ray.kapd2", "ray.kapd3"};
    for (int i = 0; i < kMaxOpac; ++i) {
      if ((rc = dev_alloc(c, kn[i], n * 4, &p))) return rc;
      rays.kapd[i] = (float*)p;
    }
    if ((rc = dev_alloc(c, "ray.tsum", n * 4, &p))) return rc;
    rays.tsum = (float*)p;
    if ((rc = dev_alloc(c, "ray.cell", n * 4, &p))) return rc;
    rays.cell = (uint32_t*)p;
    if ((rc = dev_alloc(c, "ray.item", n * 4, &p))) return rc;
    rays.item = (uint32_t*)p;
    if ((rc = dev_alloc(c, "ray.steps", n, &p))) return rc;
    rays.steps = (uint8_t*)p;
    if ((rc = dev_alloc(c, "hit.xe", n * 8, &p))) return rc;
    hits.xe = (double*)p;
    if ((rc = dev_alloc(c, "hit.ye", n * 8, &p))) return rc;
    hits.ye = (double*)p;
    if ((rc = dev_alloc(c, "hit.ze", n * 8, &p))) return rc;
    hits.ze = (double*)p;
    if ((rc = dev_alloc(c, "hit.tsum", n * 4, &p))) return rc;
    hits.tsum = (float*)p;
    if ((rc = dev_alloc(c, "hit.item", n * 4, &p))) return rc;
    hits.item = (uint32_t*)p;
    if ((rc = dev_alloc(c, "peel.ctr", 4 * sizeof(unsigned), &p))) return rc;
    split_ctr = (unsigned*)p;
    rays.count = split_ctr;
    rays.next = split_ctr + 1;
    hits.count = split_ctr + 2;
    rays.capacity = (unsigned)n;
    hits.capacity = (unsigned)n;
  }
  // The end of the run: once every packet is handed out and at most
  // tail_max remain, each runs in a wavefront of its own (k_tail).
  // (at most as many as k_tail keeps resident at once)
  int64_t tail_max = fp64 ? tail_capacity_fp64(dm, lds.bytes, c->cu_count) : tail_capacity_mixed(dm, lds.bytes, c->cu_count);
  if (const char* tl = std::getenv("MCRT_TAIL_LANES")) tail_max = std::atoll(tl);   // 0: never
  bool tail_mode = false;
  unsigned nlive = 0;
  uint32_t* tail_list = nullptr;
  unsigned* tail_count = nullptr;
  if (tail_max > 0) {
    if ((rc = dev_alloc(c, "tail.list", (size_t)lanes * 4, &p))) return rc;
    tail_list = (uint32_t*)p;
    if ((rc = dev_alloc(c, "tail.count", 16, &p))) return rc;
    tail_count = (unsigned*)p;
  }
  // k_peel_walk refills once this many lanes of a wavefront are idle: 8
  // measured fastest on the 200x180x64 workload (16: +2%, 4: +3%, 32: +7%
  // per step)
  unsigned walk_fetch = c->peel_fetch > 0 ? (unsigned)c->peel_fetch : 8u;
  if (const char* pw = std::getenv("MCRT_PEEL_WALK_FETCH")) {
    const int v = std::atoi(pw);
    if (v >= 1 && v <= 64) walk_fetch = (unsigned)v;
  }
  // Gather thresholds of the peel kernel.  Lanes whose walk reached the
  // observer wait for a tally pass and idle lanes for a fetch pass; while
  // they wait they do not walk, so high thresholds leave most of a
  // wavefront idle during the walk steps (that take ~84% of the kernel's
  // cycles).  4 idle / 8 done lanes measured fastest on the 200x180x64
  // clumpy workload (fetch 16 / tally 32 walked with 27 of 64 lanes).
  uint32_t peel_regroup = c->peel_fetch > 0 ? (uint32_t)c->peel_fetch : 4u;
  uint32_t peel_tally = c->peel_tally > 0 ? (uint32_t)c->peel_tally : 8u;
  if (const char* pr = std::getenv("MCRT_PEEL_REGROUP")) {
    const int v = std::atoi(pr);
    if (v >= 1 && v <= 64) peel_regroup = (uint32_t)v;
  }
  if (const char* pr = std::getenv("MCRT_PEEL_TALLY")) {
    const int v = std::atoi(pr);
    if (v >= 1 && v <= 64) peel_tally = (uint32_t)v;
  }
  double st[MCRT_NSTATS] = {0};
  const char* venv = std::getenv("MCRT_VERBOSE");
  const bool verbose = venv && venv[0] && venv[0] != '0';
  st[MCRT_STAT_LANES] = (double)lanes;
  st[MCRT_STAT_EVENTS0] = (double)events;
  st[MCRT_STAT_RECORD_BYTES] = (double)(fp64 ? peel_record_bytes_fp64() : peel_record_bytes_mixed());
  st[MCRT_STAT_LANE_BYTES] = 4.0 * words;
  // The peel kernels of a round run on the peel stream, concurrently with
  // the next round's transport (MCRT_OVERLAP=0: both on the run stream, one
  // after the other).  Round r writes queue r % 2; its transport first
  // waits for the peel kernels of round r - 2, which read that queue.
  hipStream_t ps = c->peel_stream;
  if (const char* ov = std::getenv("MCRT_OVERLAP")) if (std::atoi(ov) == 0) ps = c->stream;
  // A run that returns an error part-way drains the peel stream first, so
  // no peel kernel of it is left running into the context's next call; a
  // run that completes has made the context's stream wait for them instead.
  struct PeelDrain {
    hipStream_t s;
    bool armed;
    ~PeelDrain() { if (armed && s) (void)hipStreamSynchronize(s); }
  } drain{ps, true};
  struct Pending {
    bool on = false;
    int64_t round = 0;
    float ta = 0.f;
    unsigned records = 0, live = 0;
    unsigned long long handed = 0;
  } pend[2];
  // time of the finished peel kernels of queue b (the host has seen a later
  // event on the same stream complete, or synchronizes here)
  auto collect = [&](int b) -> int {
    if (!pend[b].on) return MCRT_OK;
    HIPCHECK(c, hipEventSynchronize(c->evp[b][1]));
    float tb = 0.f;
    HIPCHECK(c, hipEventElapsedTime(&tb, c->evp[b][0], c->evp[b][1]));
    st[MCRT_STAT_PEEL_MS] += tb;
    if (verbose)
      std::fprintf(stderr, "[mcrt] round %lld: transport %.3f ms, peel %.3f ms, records %u, parked %u, events %lld, handed out %llu/%lld\n",
                   (long long)pend[b].round, pend[b].ta, tb, pend[b].records, pend[b].live, (long long)events,
                   pend[b].handed, (long long)npackets);
    pend[b].on = false;
    return MCRT_OK;
  };
  HIPCHECK(c, hipEventRecord(c->ev0, c->stream));
  HIPCHECK(c, launch_fill_i32((int*)ctl, 2 * sizeof(RunCtl) / 4, 0, c->stream));
  if (dm.acc_to > 0) {
    AccStart* acc = (AccStart*)c->buf["acc.start"];
    HIPCHECK(c, fp64 ? launch_acc_place_fp64(dm, seed, (uint64_t)first_packet, acc, c->stream)
                     : launch_acc_place_mixed(dm, seed, (uint64_t)first_packet, acc, c->stream));
  }
  // Waiting lanes of a wavefront that trigger an interaction pass: on the
  // 200x180x64 workload with overlapped peel rounds, 16 and 24 measured
  // 27.5-28.2 s per 10^8-packet step, 32 28.4-28.7 s, 48 31.0 s.
  constexpr uint32_t kRegroupDefault = 24;
  int64_t round = 0;
  for (;; ++round) {
    if (round >= ((int64_t)1 << 24)) return fail(c, MCRT_E_HIP, "mcrt_run_packets: packets did not finish");
    const int b = (int)(round & 1);
    RunParams rp;
    rp.first_packet = (uint64_t)first_packet;
    rp.npackets = (uint64_t)npackets;
    rp.seed = seed;
    rp.inczeta = (const double*)c->buf["inczeta"];
    rp.q = qb[b];
    rp.lane_words = lane_words;
    rp.lane_live = lane_live;
    rp.lanes = (uint64_t)lanes;
    rp.events = (uint32_t)events;
    rp.round = (uint32_t)(round < 0xffffffffll ? round : 0xffffffffll);
    rp.ctl = ctl;
    rp.regroup = (uint32_t)(c->regroup > 0 ? c->regroup : kRegroupDefault);
    rp.peel_regroup = peel_regroup;
    rp.peel_tally = peel_tally;
    rp.lds = lds;
    // queue b is free once the peel kernels of round - 2 are done
    if (pend[b].on) {
      HIPCHECK(c, hipStreamWaitEvent(c->stream, c->evp[b][1], 0));
      if (int r2 = collect(b)) return r2;
    }
    // queue b's count and the live count restart every round (fill kernels, not memsets)
    HIPCHECK(c, launch_fill_i32((int*)&ctl[b].q_count, 1, 0, c->stream));
    HIPCHECK(c, launch_fill_i32((int*)&ctl[0].live, 1, 0, c->stream));
    HIPCHECK(c, hipEventRecord(c->evr[0], c->stream));
    hipError_t e;
    if (tail_mode) {
      HIPCHECK(c, launch_fill_i32((int*)tail_count, 1, 0, c->stream));
      HIPCHECK(c, launch_live_list(lane_live, (uint64_t)lanes, tail_list, tail_count, c->stream));
      e = fp64 ? launch_tail_fp64(dm, t, rp, tail_list, nlive, c->stream)
               : launch_tail_mixed(dm, t, rp, tail_list, nlive, c->stream);
      st[MCRT_STAT_TAIL_ROUNDS] += 1.0;
    } else {
      e = fp64 ? launch_transport_fp64(dm, t, rp, c->stream) : launch_transport_mixed(dm, t, rp, c->stream);
    }
    if (e != hipSuccess) return hipfail(c, e, "transport kernel launch");
    HIPCHECK(c, hipEventRecord(c->evr[1], c->stream));
    // the round's counts: packets handed out, lanes alive, records queued
    HIPCHECK(c, hipMemcpyAsync(c->host_ctl, ctl, 2 * sizeof(RunCtl), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    const RunCtl h = c->host_ctl[0];
    const unsigned qcount = c->host_ctl[b].q_count;
    const uint64_t nrec = qcount < (unsigned)capacity ? qcount : (unsigned)capacity;
    float ta = 0.f;
    HIPCHECK(c, hipEventElapsedTime(&ta, c->evr[0], c->evr[1]));
    st[MCRT_STAT_TRANSPORT_MS] += ta;
    st[MCRT_STAT_RECORDS] += (double)nrec;
    st[MCRT_STAT_PARKED] += (double)h.live;
    if (queued) {
      HIPCHECK(c, hipStreamWaitEvent(ps, c->evr[1], 0));
      HIPCHECK(c, hipEventRecord(c->evp[b][0], ps));
      if (split) {
        // the round's record count sizes the chunks
        const uint64_t total = nrec * (uint64_t)c->dm.npeel;
        for (uint64_t i0 = 0; i0 < total; i0 += chunk) {
          const uint64_t i1 = total - i0 < chunk ? total : i0 + chunk;
          HIPCHECK(c, launch_fill_i32((int*)split_ctr, 4, 0, ps));
          e = launch_peel_split_mixed(dm, t, rp, rays, hits, i0, i1, sgrid, wgrid, hgrid, walk_fetch, ps);
          if (e != hipSuccess) return hipfail(c, e, "peel kernels launch");
        }
      } else {
        e = fp64 ? launch_peel_fp64(dm, t, rp, pgrid, ps) : launch_peel_mixed(dm, t, rp, pgrid, ps);
        if (e != hipSuccess) return hipfail(c, e, "peel kernel launch");
      }
      HIPCHECK(c, hipEventRecord(c->evp[b][1], ps));
      st[MCRT_STAT_PEEL_LAUNCHES] += 1.0;
      pend[b].on = true;
      pend[b].round = round;
      pend[b].ta = ta;
      pend[b].records = qcount;
      pend[b].live = h.live;
      pend[b].handed = h.next_packet;
    } else if (verbose) {
      std::fprintf(stderr, "[mcrt] round %lld: transport %.3f ms, peel %.3f ms, records %u, parked %u, events %lld, handed out %llu/%lld\n",
                   (long long)round, ta, 0.0, qcount, h.live, (long long)events, h.next_packet, (long long)npackets);
    }
    if (h.next_packet < (unsigned long long)npackets) continue;
    if (h.live == 0) break;
    if ((int64_t)h.live <= tail_max) tail_mode = true;
    nlive = h.live;
  }
  // the run ends when the last peel kernels have: the run stream waits for them
  for (int k = 1; k <= 2; ++k) {
    const int b = (int)((round + k) & 1);     // the older pending queue first
    if (pend[b].on) {
      HIPCHECK(c, hipStreamWaitEvent(c->stream, c->evp[b][1], 0));
      if (int r2 = collect(b)) return r2;
    }
  }
  drain.armed = false;
  st[MCRT_STAT_ROUNDS] = (double)(round + 1);
  {
    unsigned long long cnt[C_NCOUNTERS];
    HIPCHECK(c, hipMemcpyAsync(cnt, c->dt.counters, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    st[MCRT_STAT_WALK_ERRORS] = (double)cnt[C_ERR_WALK];
    st[MCRT_STAT_PEEL_ERRORS] = (double)cnt[C_ERR_PEEL];
    st[MCRT_STAT_OTHER_ERRORS] = (double)cnt[C_ERR_OTHER];
  }
  std::memcpy(c->stats, st, sizeof(st));
  HIPCHECK(c, hipEventRecord(c->ev1, c->stream));
  c->ev_valid = true;
  return MCRT_OK;
}

int mcrt_set_launch(mcrt_ctx* c, int64_t lanes, int32_t queue_per_lane, int32_t events_per_round,
                    int32_t peel_blocks_per_cu, int32_t regroup) {
  if (!c) return MCRT_E_ARG;
  if (lanes < 0 || queue_per_lane < 0 || events_per_round < 0 || peel_blocks_per_cu < 0 || regroup < 0)
    return fail(c, MCRT_E_ARG, "mcrt_set_launch: negative value");
  if (regroup > 64) return fail(c, MCRT_E_ARG, "mcrt_set_launch: regroup is at most 64 (lanes of a wavefront)");
  c->regroup = regroup;
  if (lanes % 256 != 0) return fail(c, MCRT_E_ARG, "mcrt_set_launch: lanes must be a multiple of 256");
  c->lanes = lanes;
  c->queue_per_lane = queue_per_lane;
  c->events_per_round = events_per_round;
  c->peel_blocks_per_cu = peel_blocks_per_cu;
  return MCRT_OK;
}

int mcrt_set_peel_gather(mcrt_ctx* c, int32_t fetch_lanes, int32_t tally_lanes) {
  if (!c) return MCRT_E_ARG;
  if (fetch_lanes < 0 || fetch_lanes > 64 || tally_lanes < 0 || tally_lanes > 64)
    return fail(c, MCRT_E_ARG, "mcrt_set_peel_gather: thresholds are 0..64 lanes of a wavefront");
  c->peel_fetch = fetch_lanes;
  c->peel_tally = tally_lanes;
  return MCRT_OK;
}

int mcrt_adjoint_image(mcrt_ctx* c, int64_t first_packet, int64_t npackets, uint64_t seed, int32_t precision,
                       const float* W, double* loss, double* grad) {
  if (!c) return MCRT_E_ARG;
  if (!ready(c)) return fail(c, MCRT_E_STATE, "mcrt_adjoint_image: grid, medium, dust, source, observer and tdust must be set");
  if (npackets < 0 || first_packet < 0) return fail(c, MCRT_E_ARG, "mcrt_adjoint_image: negative packet range");
  if (precision != MCRT_FP64 && precision != MCRT_MIXED) return fail(c, MCRT_E_ARG, "mcrt_adjoint_image: unknown precision");
  if (!W || !loss) return fail(c, MCRT_E_ARG, "mcrt_adjoint_image: null W or loss");
  if (c->dm.ipeel != 1 || c->dm.npeel <= 0 || c->dm.nband <= 0)
    return fail(c, MCRT_E_STATE, "mcrt_adjoint_image: the observer has no peeled images");
  // the replay runs stellar photons only: a context whose mcrt_run_packets
  // would emit accretion, hot-spot or external photons describes another
  // forward model than the one the gradient would be of
  if (c->diff_on)
    return fail(c, MCRT_E_STATE, "mcrt_adjoint_image: the diffusion approximation's surface temperatures depend on the "
                                 "order photons arrive in, which a replay cannot reproduce; turn it off (mcrt_set_diffusion)");
  if (c->accfrac > 0.0 || c->spot.ispot || (c->nisrf > 0 && c->isrf_frac > 0.0))
    return fail(c, MCRT_E_STATE, "mcrt_adjoint_image: accretion, hot-spot and external photons are not part of the "
                                 "replayed model; turn them off (mcrt_set_accretion, mcrt_set_spots, mcrt_set_isrf)");
  if (lds_always_doubles(c->dm) > kLdsAlways) return fail(c, MCRT_E_ARG, "mcrt_adjoint_image: tables exceed the LDS budget");
  HIPCHECK(c, hipSetDevice(c->device));
  const bool fp64 = (precision == MCRT_FP64);
  const size_t nw = (size_t)c->dm.nband * c->dm.npeel * c->dm.nx * c->dm.nx;
  const float* dW = nullptr;
  int rc = upload(c, "adj.W", W, nw, &dW);
  if (rc) return rc;
  void* p;
  if ((rc = dev_alloc(c, "adj.loss", sizeof(double), &p))) return rc;
  double* dloss = (double*)p;
  double* dgrad = nullptr;
  if (grad) {
    if ((rc = dev_alloc(c, "adj.grad", c->ncell * sizeof(double), &p))) return rc;
    dgrad = (double*)p;
    HIPCHECK(c, launch_fill_f64(dgrad, c->ncell, 0.0, c->stream));
  }
  HIPCHECK(c, launch_fill_f64(dloss, 1, 0.0, c->stream));
  // the transport code writes the Lucy tallies and counters: scratch copies
  DeviceTallies t{};
  if ((rc = dev_alloc(c, "adj.dtauabs", c->ncell * sizeof(double), &p))) return rc;
  t.dtauabs = (double*)p;
  if ((rc = dev_alloc(c, "adj.int3", 3 * c->ncell * sizeof(int), &p))) return rc;
  t.nabs = (int*)p; t.dbig = t.nabs + c->ncell; t.killed = t.dbig + c->ncell;
  if ((rc = dev_alloc(c, "adj.counters", C_NCOUNTERS * sizeof(unsigned long long), &p))) return rc;
  t.counters = (unsigned long long*)p;
  t.sed = t.psed = t.image = nullptr;
  t.trace = nullptr;
  AdjParams ap;
  ap.first_packet = (uint64_t)first_packet;
  ap.npackets = (uint64_t)npackets;
  ap.seed = seed;
  ap.inczeta = (const double*)c->buf["inczeta"];
  ap.W = dW;
  ap.grad = dgrad;
  ap.loss = dloss;
  ap.lds = plan_lds(c->dm);
  if (npackets > 0) {
    const int64_t blocks = (npackets + 255) / 256;
    const int64_t cap = (int64_t)c->cu_count * 8;
    const int grid = (int)(blocks < cap ? blocks : cap);
    hipError_t e = fp64 ? launch_adjoint_fp64(c->dm, t, ap, grid, c->stream) : launch_adjoint_mixed(c->dm, t, ap, grid, c->stream);
    if (e != hipSuccess) return hipfail(c, e, "adjoint kernel launch");
  }
  HIPCHECK(c, hipMemcpyAsync(loss, dloss, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (grad) HIPCHECK(c, hipMemcpyAsync(grad, dgrad, c->ncell * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return MCRT_OK;
}

}  // extern "C"

namespace {

// LDS plan of the batch kernel: only the tables every model shares
LdsPlan plan_lds_shared(const DeviceModel& m) {
  LdsPlan P{};
  const uint32_t npg = (uint32_t)m.npg, nd = (uint32_t)(m.nopac * m.nlam);
  const uint32_t sizes[LT_COUNT] = {0u, 0u, 0u, 0u, 0u, 0u, npg, npg, npg,
                                    (uint32_t)m.nT, (uint32_t)m.nnu, 5u * (uint32_t)(m.npeel > 0 ? m.npeel : 1), 145u,
                                    nd, nd, nd, nd, nd};
  uint32_t used = 0;
  for (int k = 0; k < LT_COUNT; ++k) {
    if (k == LT_NUDUST && used + 5u * nd > kLdsMax) break;
    P.off[k] = used;
    P.n[k] = sizes[k];
    used += sizes[k];
  }
  P.bytes = used * 8u;
  return P;
}

BatchLaunch batch_launch(mcrt_ctx* c, int block, bool combine) {
  BatchLaunch L;
  L.m = c->dm;
  L.m.npg = 2;
  L.m.phiarr = c->b_phiarr;
  L.m.aarr = c->b_aarr;
  L.m.barr = c->b_barr;
  {
    const double e = std::sin(0.5 * 1.0e-4 * 2.0 * M_PI);
    L.m.phi_tol2 = e * e;
  }
  L.m.species_opac = c->b_species_opac;
  L.m.diff_itmin = nullptr;            // the diffusion approximation is a single-model option
  L.t = c->bt;
  L.bd = c->bd;
  L.run = c->b_run;
  L.inczeta = (const double*)c->buf["inczeta"];
  L.lds = plan_lds_shared(L.m);
  L.block = block;
  L.combine = combine;
  L.grid = 0;
  return L;
}

// grid of `waves` resident waves per SIMD on every CU, at most what the
// kernel's registers and LDS allow; returns the waves per SIMD launched
int batch_grid(mcrt_ctx* c, bool fp64, BatchLaunch& L, int waves) {
  int bpc = 0;
  hipError_t e = fp64 ? batch_occupancy_fp64(L, &bpc) : batch_occupancy_mixed(L, &bpc);
  if (e != hipSuccess || bpc <= 0) return -1;
  int want = waves * 4 * 64 / L.block;
  if (want < 1) want = 1;
  const int per_cu = want < bpc ? want : bpc;
  L.grid = (unsigned)(per_cu * c->cu_count);
  return per_cu * L.block / 256 > 0 ? per_cu * L.block / 256 : 1;
}

bool batch_ready(mcrt_ctx* c) {
  return c->has_batch && c->has_batch_tdust && c->has_dust && c->has_source && c->has_observer;
}

int batch_zero(mcrt_ctx* c, hipStream_t s) {
  HIPCHECK(c, launch_fill_f64(c->bt.dtauabs, c->b_ncell, 0.0, s));
  HIPCHECK(c, launch_fill_i32(c->bt.nabs, 3 * c->b_ncell, 0, s));
  HIPCHECK(c, launch_fill_f32(c->bt.sed, (size_t)c->nmodels * c->b_nsed, 0.f, s));
  HIPCHECK(c, launch_fill_u64(c->bt.counters, C_NCOUNTERS, 0ull, s));
  HIPCHECK(c, launch_fill_u64(c->bd.work, (size_t)c->nmodels, 0ull, s));
  return MCRT_OK;
}

// Queue order (most expensive model first) and packet prefix for npk, then
// the run block; trial runs stop once every packet is handed out.
int batch_stage(mcrt_ctx* c, const int64_t* npk, uint64_t seed, bool trial) {
  const int nm = c->nmodels;
  c->b_npk.assign(nm, 0ull);
  for (int k = 0; k < nm; ++k) c->b_npk[k] = (unsigned long long)npk[k];
  c->b_order.resize(nm);
  for (int k = 0; k < nm; ++k) c->b_order[k] = k;
  std::stable_sort(c->b_order.begin(), c->b_order.end(), [&](int a, int b) {
    return c->b_cost[a] * (double)c->b_npk[a] > c->b_cost[b] * (double)c->b_npk[b];
  });
  c->b_prefix.assign(nm + 1, 0ull);
  for (int j = 0; j < nm; ++j) c->b_prefix[j + 1] = c->b_prefix[j] + c->b_npk[c->b_order[j]];
  hipStream_t s = c->stream;
  HIPCHECK(c, hipMemcpyAsync((void*)c->bd.order, c->b_order.data(), nm * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHECK(c, hipMemcpyAsync((void*)c->bd.prefix, c->b_prefix.data(), (nm + 1) * sizeof(unsigned long long), hipMemcpyHostToDevice, s));
  HIPCHECK(c, hipMemcpyAsync(c->b_npk_dev, c->b_npk.data(), nm * sizeof(unsigned long long), hipMemcpyHostToDevice, s));
  BatchRun& r = *c->b_run_h;
  r.seed = seed;
  r.total = c->b_prefix[nm];
  r.head = 0ull;
  r.regroup = (uint32_t)c->b_regroup;
  r.trial = trial ? 1u : 0u;
  HIPCHECK(c, hipMemcpyAsync(c->b_run, c->b_run_h, sizeof(BatchRun), hipMemcpyHostToDevice, s));
  return MCRT_OK;
}

int batch_launch_direct(mcrt_ctx* c, bool fp64, const BatchLaunch& L, hipStream_t s) {
  int rc = batch_zero(c, s);
  if (rc) return rc;
  hipError_t e = fp64 ? launch_batch_fp64(L, s) : launch_batch_mixed(L, s);
  if (e != hipSuccess) return hipfail(c, e, "batch kernel launch");
  return MCRT_OK;
}

// The run as a graph: tally reset (fill kernels) and the batch kernel,
// captured once per launch shape and replayed.
int batch_launch_graph(mcrt_ctx* c, bool fp64, const BatchLaunch& L) {
  const unsigned long long key[6] = {c->b_gen, fp64 ? 1ull : 0ull, (unsigned long long)L.block,
                                     L.combine ? 1ull : 0ull, (unsigned long long)L.grid, (unsigned long long)(uintptr_t)c->stream};
  if (!c->b_exec || std::memcmp(key, c->b_exec_key, sizeof(key)) != 0) {
    if (c->b_exec) { HIPCHECK(c, hipGraphExecDestroy(c->b_exec)); c->b_exec = nullptr; }
    HIPCHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = batch_launch_direct(c, fp64, L, c->stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return hipfail(c, e, "batch graph capture");
    e = hipGraphInstantiate(&c->b_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { c->b_exec = nullptr; return hipfail(c, e, "batch graph instantiate"); }
    std::memcpy(c->b_exec_key, key, sizeof(key));
  }
  HIPCHECK(c, hipGraphLaunch(c->b_exec, c->stream));
  return MCRT_OK;
}

int batch_run(mcrt_ctx* c, const int64_t* npk, uint64_t seed, bool fp64, bool trial, int block, int waves, bool combine,
              bool graph, float* ms) {
  int rc = batch_stage(c, npk, seed, trial);
  if (rc) return rc;
  BatchLaunch L = batch_launch(c, block, combine);
  if (batch_grid(c, fp64, L, waves) < 0) return fail(c, MCRT_E_ARG, "batch launch shape does not fit the device");
  HIPCHECK(c, hipEventRecord(c->ev0, c->stream));
  rc = graph ? batch_launch_graph(c, fp64, L) : batch_launch_direct(c, fp64, L, c->stream);
  if (rc) return rc;
  HIPCHECK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->ev_valid = true;
  if (ms) HIPCHECK(c, hipEventElapsedTime(ms, c->ev0, c->ev1));
  return MCRT_OK;
}

}  // namespace

extern "C" {

int mcrt_set_batch(mcrt_ctx* c, int32_t nmodels, const int32_t* nrg, const int32_t* ntg, const double* rarr,
                   const double* r2arr, const double* thetarr, const double* costarr, const double* tan2arr,
                   const double* phiarr, const double* aarr, const double* barr, const double* vcell, const double* rho,
                   const uint8_t* species, const int32_t* species_opac) {
  if (!c) return MCRT_E_ARG;
  if (nmodels < 1 || !nrg || !ntg || !rarr || !r2arr || !thetarr || !costarr || !tan2arr || !phiarr || !aarr || !barr ||
      !vcell || !rho || !species || !species_opac)
    return fail(c, MCRT_E_ARG, "mcrt_set_batch: need nmodels >= 1 and all arrays");
  if (std::fabs(phiarr[0]) > 1e-12 || std::fabs(phiarr[1] - 2.0 * M_PI) > 1e-9)
    return fail(c, MCRT_E_ARG, "mcrt_set_batch: the azimuthal walls must be 0 and 2*pi");
  for (int i = 0; i <= 10; ++i)
    if (species_opac[i] < 0 || species_opac[i] >= kMaxOpac) return fail(c, MCRT_E_ARG, "mcrt_set_batch: species_opac out of range");
  std::vector<long long> roff(nmodels), toff(nmodels), coff(nmodels);
  long long nr = 0, nt = 0, nc = 0;
  size_t maxc = 0;
  for (int k = 0; k < nmodels; ++k) {
    if (nrg[k] < 3 || ntg[k] < 2 || nrg[k] > 1023 || ntg[k] > 1023 || (ntg[k] > 2 && ntg[k] % 2 == 0))
      return fail(c, MCRT_E_ARG, "mcrt_set_batch: model walls need 3 <= nrg <= 1023 and odd 3 <= ntg <= 1023 (or 2)");
    roff[k] = nr; toff[k] = nt; coff[k] = nc;
    nr += nrg[k];
    nt += ntg[k];
    const size_t cells = (size_t)(nrg[k] - 1) * (ntg[k] - 1);
    nc += (long long)cells;
    if (cells > maxc) maxc = cells;
  }
  if (nc >= (1ll << 31)) return fail(c, MCRT_E_ARG, "mcrt_set_batch: more than 2^31 cells");
  for (long long i = 0; i < nc; ++i) {
    const int sp = species[i];
    if (sp > 10) return fail(c, MCRT_E_ARG, "mcrt_set_batch: species code above 10");
    if (sp >= 5 && sp <= 8 && rho[i] > 0.0) return fail(c, MCRT_E_ARG, "mcrt_set_batch: small-grain species (5-8) are not supported");
    if (!(rho[i] >= 0.0)) return fail(c, MCRT_E_ARG, "mcrt_set_batch: negative or NaN density");
    if ((rho[i] > 0.0) != (sp > 0)) return fail(c, MCRT_E_ARG, "mcrt_set_batch: a cell has density without species or species without density");
  }
  HIPCHECK(c, hipSetDevice(c->device));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->has_batch = c->has_batch_tdust = false;
  c->b_gen++;
  BatchDesc& b = c->bd;
  b.nmodels = nmodels;
  int rc = 0;
  rc |= upload(c, "b.nrg", (const int*)nrg, nmodels, &b.nrg);
  rc |= upload(c, "b.ntg", (const int*)ntg, nmodels, &b.ntg);
  rc |= upload(c, "b.roff", roff.data(), nmodels, &b.roff);
  rc |= upload(c, "b.toff", toff.data(), nmodels, &b.toff);
  rc |= upload(c, "b.coff", coff.data(), nmodels, &b.coff);
  rc |= upload(c, "b.rarr", rarr, nr, &b.rarr);
  rc |= upload(c, "b.r2arr", r2arr, nr, &b.r2arr);
  rc |= upload(c, "b.thetarr", thetarr, nt, &b.thetarr);
  rc |= upload(c, "b.costarr", costarr, nt, &b.costarr);
  rc |= upload(c, "b.tan2arr", tan2arr, nt, &b.tan2arr);
  rc |= upload(c, "b.vcell", vcell, nc, &b.vcell);
  rc |= upload(c, "b.rho64", rho, nc, &b.rho64);
  rc |= upload_cast<float>(c, "b.rho32", rho, nc, &b.rho32);
  rc |= upload(c, "b.species", species, nc, &b.species);
  rc |= upload(c, "b.species_opac", (const int*)species_opac, 11, &c->b_species_opac);
  rc |= upload(c, "b.phiarr", phiarr, 2, &c->b_phiarr);
  rc |= upload(c, "b.aarr", aarr, 2, &c->b_aarr);
  rc |= upload(c, "b.barr", barr, 2, &c->b_barr);
  std::vector<double> cc(nt), sc(nt);
  for (long long i = 0; i < nt; ++i) {
    const double t = std::min(thetarr[i], M_PI - thetarr[i]);
    cc[i] = std::cos(t);
    sc[i] = std::sin(t);
  }
  rc |= upload(c, "b.ccone", cc.data(), nt, &b.ccone);
  rc |= upload(c, "b.scone", sc.data(), nt, &b.scone);
  if (rc) return rc;
  void* p;
  if ((rc = dev_alloc(c, "b.tdust", nc * 8, &p))) return rc;
  b.tdust = (const double*)p;
  if ((rc = dev_alloc(c, "b.order", nmodels * sizeof(int), &p))) return rc;
  b.order = (const int*)p;
  if ((rc = dev_alloc(c, "b.prefix", (nmodels + 1) * sizeof(unsigned long long), &p))) return rc;
  b.prefix = (const unsigned long long*)p;
  if ((rc = dev_alloc(c, "b.work", nmodels * sizeof(unsigned long long), &p))) return rc;
  b.work = (unsigned long long*)p;
  if ((rc = dev_alloc(c, "b.npk", nmodels * sizeof(unsigned long long), &p))) return rc;
  c->b_npk_dev = (unsigned long long*)p;
  if ((rc = dev_alloc(c, "b.run", sizeof(BatchRun), &p))) return rc;
  c->b_run = (BatchRun*)p;
  if (!c->b_run_h && hipHostMalloc((void**)&c->b_run_h, sizeof(BatchRun), hipHostMallocDefault) != hipSuccess) {
    c->b_run_h = nullptr;
    return fail(c, MCRT_E_HIP, "mcrt_set_batch: pinned host allocation");
  }
  // tallies
  c->nmodels = nmodels;
  c->b_ncell = (size_t)nc;
  c->b_max_cells = maxc;
  c->b_nsed = (size_t)10 * c->dm.nfreq * c->dm.nmu * c->dm.nph * c->dm.nap * kNo;
  b.sed_stride = c->b_nsed;
  if ((rc = dev_alloc(c, "b.t.dtauabs", nc * 8, &p))) return rc;
  c->bt.dtauabs = (double*)p;
  if ((rc = dev_alloc(c, "b.t.int3", 3 * nc * 4, &p))) return rc;
  c->bt.nabs = (int*)p;
  c->bt.dbig = c->bt.nabs + nc;
  c->bt.killed = c->bt.dbig + nc;
  if ((rc = dev_alloc(c, "b.t.sed", (size_t)nmodels * c->b_nsed * 4 + 4, &p))) return rc;
  c->bt.sed = (float*)p;
  if ((rc = dev_alloc(c, "b.t.counters", C_NCOUNTERS * 8, &p))) return rc;
  c->bt.counters = (unsigned long long*)p;
  c->bt.psed = nullptr;
  c->bt.image = nullptr;
  c->bt.trace = nullptr;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  // first queue order: column density along the row of cells above the
  // equator (optically thick models hold the longest-lived packets)
  c->b_nrg.assign(nrg, nrg + nmodels);
  c->b_ntg.assign(ntg, ntg + nmodels);
  c->b_coff = coff;
  c->b_cost.assign(nmodels, 0.0);
  for (int k = 0; k < nmodels; ++k) {
    const int ncr = nrg[k] - 1, it = (ntg[k] - 1) / 2 > 0 ? (ntg[k] - 1) / 2 : 1;
    double col = 0.0;
    for (int ir = 1; ir <= ncr; ++ir) {
      const size_t cell = (size_t)coff[k] + (size_t)(ir - 1) + (size_t)ncr * (size_t)(it - 1);
      col += rho[cell] * (rarr[roff[k] + ir] - rarr[roff[k] + ir - 1]);
    }
    c->b_cost[k] = col;
  }
  c->b_ran = false;
  c->has_batch = true;
  return MCRT_OK;
}

int mcrt_set_batch_tdust(mcrt_ctx* c, const double* tdust) {
  if (!c) return MCRT_E_ARG;
  if (!c->has_batch) return fail(c, MCRT_E_STATE, "mcrt_set_batch_tdust: set the batch first");
  if (!tdust) return fail(c, MCRT_E_ARG, "mcrt_set_batch_tdust: null array");
  HIPCHECK(c, hipSetDevice(c->device));
  HIPCHECK(c, hipMemcpyAsync((void*)c->bd.tdust, tdust, c->b_ncell * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->has_batch_tdust = true;
  return MCRT_OK;
}

int mcrt_set_batch_launch(mcrt_ctx* c, int32_t block, int32_t waves_per_simd, int32_t regroup, int32_t combine, int32_t graph) {
  if (!c) return MCRT_E_ARG;
  if (block != 0 && block != 64 && block != 128 && block != 256) return fail(c, MCRT_E_ARG, "mcrt_set_batch_launch: block must be 64, 128 or 256");
  if (waves_per_simd < 0 || waves_per_simd > 8) return fail(c, MCRT_E_ARG, "mcrt_set_batch_launch: waves per SIMD must be 1..8");
  if (regroup < 0 || regroup > 64) return fail(c, MCRT_E_ARG, "mcrt_set_batch_launch: regroup must be 1..64");
  if (block) c->b_block = block;
  if (waves_per_simd) c->b_waves = waves_per_simd;
  if (regroup) c->b_regroup = regroup;
  c->b_combine = combine ? 1 : 0;
  c->b_graph = graph ? 1 : 0;
  return MCRT_OK;
}

int mcrt_run_batch(mcrt_ctx* c, const int64_t* npackets, uint64_t seed, int32_t precision) {
  if (!c) return MCRT_E_ARG;
  if (!batch_ready(c)) return fail(c, MCRT_E_STATE, "mcrt_run_batch: batch, batch tdust, dust, source and observer must be set");
  if (c->dm.ipeel != 0) return fail(c, MCRT_E_STATE, "mcrt_run_batch: batches bin SEDs only; set the observer with ipeel = 0");
  if (!npackets) return fail(c, MCRT_E_ARG, "mcrt_run_batch: null packet counts");
  if (precision != MCRT_FP64 && precision != MCRT_MIXED) return fail(c, MCRT_E_ARG, "mcrt_run_batch: unknown precision");
  for (int k = 0; k < c->nmodels; ++k)
    if (npackets[k] < 0) return fail(c, MCRT_E_ARG, "mcrt_run_batch: negative packet count");
  if (c->b_nsed != (size_t)10 * c->dm.nfreq * c->dm.nmu * c->dm.nph * c->dm.nap * kNo)
    return fail(c, MCRT_E_STATE, "mcrt_run_batch: the SED binning changed after mcrt_set_batch; set the batch again");
  HIPCHECK(c, hipSetDevice(c->device));
  const bool fp64 = precision == MCRT_FP64;
  int rc = batch_run(c, npackets, seed, fp64, false, c->b_block, c->b_waves, c->b_combine != 0, c->b_graph != 0, nullptr);
  if (rc) return rc;
  // interactions per packet of each model: the next run's queue order
  c->b_work.assign(c->nmodels, 0ull);
  HIPCHECK(c, hipMemcpyAsync(c->b_work.data(), c->bd.work, c->nmodels * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < c->nmodels; ++k)
    if (c->b_npk[k] > 0) c->b_cost[k] = (double)c->b_work[k] / (double)c->b_npk[k];
  c->b_ran = true;
  return MCRT_OK;
}

int mcrt_get_batch_tallies(mcrt_ctx* c, double* dtauabs, int32_t* nabs, int32_t* dbig, int32_t* killed, float* sed,
                           double* work, double* counters) {
  if (!c) return MCRT_E_ARG;
  if (!c->has_batch) return fail(c, MCRT_E_STATE, "mcrt_get_batch_tallies: no batch");
  HIPCHECK(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const size_t nc = c->b_ncell;
  if (dtauabs) HIPCHECK(c, hipMemcpyAsync(dtauabs, c->bt.dtauabs, nc * 8, hipMemcpyDeviceToHost, s));
  if (nabs) HIPCHECK(c, hipMemcpyAsync(nabs, c->bt.nabs, nc * 4, hipMemcpyDeviceToHost, s));
  if (dbig) HIPCHECK(c, hipMemcpyAsync(dbig, c->bt.dbig, nc * 4, hipMemcpyDeviceToHost, s));
  if (killed) HIPCHECK(c, hipMemcpyAsync(killed, c->bt.killed, nc * 4, hipMemcpyDeviceToHost, s));
  if (sed) HIPCHECK(c, hipMemcpyAsync(sed, c->bt.sed, (size_t)c->nmodels * c->b_nsed * 4, hipMemcpyDeviceToHost, s));
  std::vector<unsigned long long> w(c->nmodels);
  unsigned long long cnt[C_NCOUNTERS];
  HIPCHECK(c, hipMemcpyAsync(w.data(), c->bd.work, w.size() * 8, hipMemcpyDeviceToHost, s));
  HIPCHECK(c, hipMemcpyAsync(cnt, c->bt.counters, sizeof(cnt), hipMemcpyDeviceToHost, s));
  HIPCHECK(c, hipStreamSynchronize(s));
  if (work)
    for (int k = 0; k < c->nmodels; ++k) work[k] = (double)w[k];
  if (counters)
    for (int i = 0; i < C_NPUBLIC; ++i) counters[i] = (double)cnt[i];
  return MCRT_OK;
}

int mcrt_batch_dusttemp(mcrt_ctx* c, int32_t mode, int32_t update_tdust, double* tdave) {
  if (!c) return MCRT_E_ARG;
  if (!(c->has_batch && c->has_dust && c->has_source)) return fail(c, MCRT_E_STATE, "mcrt_batch_dusttemp: batch not set");
  if (!c->b_ran) return fail(c, MCRT_E_STATE, "mcrt_batch_dusttemp: no batch run yet");
  if (mode != MCRT_T_REFERENCE && mode != MCRT_T_PERCELL) return fail(c, MCRT_E_ARG, "mcrt_batch_dusttemp: unknown mode");
  HIPCHECK(c, hipSetDevice(c->device));
  void* p;
  int rc = dev_alloc(c, "b.tdave", c->b_ncell * 8, &p);
  if (rc) return rc;
  double* d_tdave = (double*)p;
  BatchLaunch L = batch_launch(c, 256, false);
  HIPCHECK(c, launch_batch_tfinal(L, c->b_npk_dev, mode, d_tdave, c->b_max_cells, c->stream));
  if (update_tdust) {
    HIPCHECK(c, hipMemcpyAsync((void*)c->bd.tdust, d_tdave, c->b_ncell * 8, hipMemcpyDeviceToDevice, c->stream));
    c->has_batch_tdust = true;
  }
  if (tdave) HIPCHECK(c, hipMemcpyAsync(tdave, d_tdave, c->b_ncell * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return MCRT_OK;
}

int mcrt_batch_autotune(mcrt_ctx* c, int64_t packets_per_model, uint64_t seed, int32_t precision, int32_t* best,
                        double* table, int32_t max_rows, int32_t* nrows) {
  if (!c) return MCRT_E_ARG;
  if (!batch_ready(c)) return fail(c, MCRT_E_STATE, "mcrt_batch_autotune: batch, batch tdust, dust, source and observer must be set");
  if (c->dm.ipeel != 0) return fail(c, MCRT_E_STATE, "mcrt_batch_autotune: batches bin SEDs only; set the observer with ipeel = 0");
  if (packets_per_model < 1 || !best) return fail(c, MCRT_E_ARG, "mcrt_batch_autotune: need packets_per_model >= 1 and best[4]");
  if (precision != MCRT_FP64 && precision != MCRT_MIXED) return fail(c, MCRT_E_ARG, "mcrt_batch_autotune: unknown precision");
  HIPCHECK(c, hipSetDevice(c->device));
  const bool fp64 = precision == MCRT_FP64;
  std::vector<int64_t> npk(c->nmodels, packets_per_model);
  const double total = (double)packets_per_model * c->nmodels;
  const int blocks[3] = {64, 128, 256}, waves[3] = {1, 2, 4}, regroups[3] = {16, 32, 48};
  double best_rate = -1.0;
  int rows = 0;
  for (int ib = 0; ib < 3; ++ib)
    for (int iw = 0; iw < 3; ++iw)
      for (int ig = 0; ig < 3; ++ig)
        for (int comb = 0; comb < 2; ++comb) {
          BatchLaunch L = batch_launch(c, blocks[ib], comb != 0);
          const int got = batch_grid(c, fp64, L, waves[iw]);
          if (got < 0) return fail(c, MCRT_E_HIP, "mcrt_batch_autotune: occupancy query failed");
          if (got != waves[iw]) continue;        // does not fit: same as a smaller wave count
          const int keep = c->b_regroup;
          c->b_regroup = regroups[ig];
          float ms = 0.f, ms2 = 0.f;
          // the first launch of a shape loads its code object: time the second
          int rc = batch_run(c, npk.data(), seed, fp64, true, blocks[ib], waves[iw], comb != 0, false, &ms);
          if (!rc) rc = batch_run(c, npk.data(), seed, fp64, true, blocks[ib], waves[iw], comb != 0, false, &ms2);
          c->b_regroup = keep;
          if (rc) return rc;
          ms = ms2;
          const double rate = ms > 0.f ? total / (ms * 1e-3) : 0.0;
          if (table && rows < max_rows) {
            double* row = table + 6 * rows;
            row[0] = blocks[ib]; row[1] = waves[iw]; row[2] = regroups[ig]; row[3] = comb; row[4] = ms; row[5] = rate;
          }
          ++rows;
          if (rate > best_rate) {
            best_rate = rate;
            best[0] = blocks[ib]; best[1] = waves[iw]; best[2] = regroups[ig]; best[3] = comb;
          }
        }
  if (nrows) *nrows = rows;
  if (best_rate < 0.0) return fail(c, MCRT_E_HIP, "mcrt_batch_autotune: no launch shape fits");
  c->b_block = best[0];
  c->b_waves = best[1];
  c->b_regroup = best[2];
  c->b_combine = best[3];
  return MCRT_OK;
}

int mcrt_run_stats(mcrt_ctx* c, double* out) {
  if (!c || !out) return MCRT_E_ARG;
  if (!c->ev_valid) return fail(c, MCRT_E_STATE, "mcrt_run_stats: no run recorded");
  std::memcpy(out, c->stats, sizeof(c->stats));
  return MCRT_OK;
}

int mcrt_last_run_ms(mcrt_ctx* c, float* ms) {
  if (!c || !ms) return MCRT_E_ARG;
  if (!c->ev_valid) return fail(c, MCRT_E_STATE, "mcrt_last_run_ms: no run recorded");
  HIPCHECK(c, hipEventSynchronize(c->ev1));
  HIPCHECK(c, hipEventElapsedTime(ms, c->ev0, c->ev1));
  return MCRT_OK;
}

int mcrt_get_tallies(mcrt_ctx* c, double* dtauabs, int32_t* nabs, int32_t* dbig, int32_t* killed, float* sed, float* psed,
                     float* image, double* counters) {
  if (!c) return MCRT_E_ARG;
  HIPCHECK(c, hipSetDevice(c->device));
  if (!c->dt.dtauabs) {
    int rc = zero_tallies(c);
    if (rc) return rc;
  }
  hipStream_t s = c->stream;
  if (dtauabs) HIPCHECK(c, hipMemcpyAsync(dtauabs, c->dt.dtauabs, c->ncell * 8, hipMemcpyDeviceToHost, s));
  if (nabs) HIPCHECK(c, hipMemcpyAsync(nabs, c->dt.nabs, c->ncell * 4, hipMemcpyDeviceToHost, s));
  if (dbig) HIPCHECK(c, hipMemcpyAsync(dbig, c->dt.dbig, c->ncell * 4, hipMemcpyDeviceToHost, s));
  if (killed) HIPCHECK(c, hipMemcpyAsync(killed, c->dt.killed, c->ncell * 4, hipMemcpyDeviceToHost, s));
  if (sed) HIPCHECK(c, hipMemcpyAsync(sed, c->dt.sed, c->nsed * 4, hipMemcpyDeviceToHost, s));
  if (psed && c->npsed) HIPCHECK(c, hipMemcpyAsync(psed, c->dt.psed, c->npsed * 4, hipMemcpyDeviceToHost, s));
  if (image && c->nimage) HIPCHECK(c, hipMemcpyAsync(image, c->dt.image, c->nimage * 4, hipMemcpyDeviceToHost, s));
  unsigned long long cnt[C_NCOUNTERS];
  HIPCHECK(c, hipMemcpyAsync(cnt, c->dt.counters, sizeof(cnt), hipMemcpyDeviceToHost, s));
  HIPCHECK(c, hipStreamSynchronize(s));
  if (counters)
    for (int i = 0; i < C_NPUBLIC; ++i) counters[i] = (double)cnt[i];
  return MCRT_OK;
}

int mcrt_dusttemp(mcrt_ctx* c, double nphot, int32_t mode, int32_t update_tdust, double* tdave) {
  if (!c) return MCRT_E_ARG;
  if (!(c->has_grid && c->has_medium && c->has_dust && c->has_source)) return fail(c, MCRT_E_STATE, "mcrt_dusttemp: model not set");
  if (!(nphot > 0.0)) return fail(c, MCRT_E_ARG, "mcrt_dusttemp: nphot must be > 0");
  if (mode != MCRT_T_REFERENCE && mode != MCRT_T_PERCELL) return fail(c, MCRT_E_ARG, "mcrt_dusttemp: unknown mode");
  HIPCHECK(c, hipSetDevice(c->device));
  if (!c->dt.dtauabs) {
    int rc = zero_tallies(c);
    if (rc) return rc;
  }
  void* p;
  int rc = dev_alloc(c, "tdave", c->ncell * 8, &p);
  if (rc) return rc;
  double* d_tdave = (double*)p;
  const double* told = c->dm.tdust;
  if (c->diff_on && c->diff_set && c->dm.tdust) {
    // the temperatures the run's diffusion photons left in the diffusion
    // columns (they stand for the reference's in-run updates of tdust)
    DiffArgs a{(const double*)c->buf["diff.kappar"], (const double*)c->buf["diff.hdisk"], nullptr, c->diff_rmaxd,
               c->diff_nr0, 0, (int32_t*)c->buf["diff.itmin"]};
    HIPCHECK(c, launch_diff_columns(c->dm, a, c->dt.diff_count, c->dt.nabs, c->diff_delta_run, (double*)c->buf["tdust"],
                                    c->stream));
    told = (const double*)c->buf["diff.told"];
  }
  HIPCHECK(c, launch_tfinal(c->dm, c->dt.dtauabs, nphot, mode, d_tdave, c->stream));
  // tfinal's convergence statistic against the temperatures this run
  // re-emitted with (the previous iteration's tdave)
  if (told) {
    void* q;
    if ((rc = dev_alloc(c, "lucy.sums", 3 * 8, &q))) return rc;
    HIPCHECK(c, launch_lucy_sigma(c->dm, told, d_tdave, (double*)q, c->stream));
    HIPCHECK(c, hipMemcpyAsync(c->lucy_sum, q, 3 * 8, hipMemcpyDeviceToHost, c->stream));
  }
  if (update_tdust) {
    void* q;
    if ((rc = dev_alloc(c, "tdust", c->ncell * 8, &q))) return rc;
    HIPCHECK(c, hipMemcpyAsync(q, d_tdave, c->ncell * 8, hipMemcpyDeviceToDevice, c->stream));
    c->dm.tdust = (const double*)q;
    c->has_tdust = true;
  }
  if (tdave) HIPCHECK(c, hipMemcpyAsync(tdave, d_tdave, c->ncell * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  // sig = sqrt(<x^2> - <|x|>^2) (tfinal_5_test1.f90:337-345); no counted
  // cell gives 0/0 as in the reference
  const double n = c->lucy_sum[2];
  const double m1 = c->lucy_sum[0] / n, m2 = c->lucy_sum[1] / n;
  c->lucy_sigma = std::sqrt(m2 - m1 * m1);
  c->lucy_sigma_prev = c->lucy_sigma_iter;
  if (update_tdust) c->lucy_sigma_iter = c->lucy_sigma;
  c->lucy_valid = true;
  return MCRT_OK;
}

int mcrt_set_diffusion(mcrt_ctx* c, int32_t on, const double* kappar, const double* hdisk, double rmaxd, int32_t nr0) {
  if (!c) return MCRT_E_ARG;
  if (!on) {
    c->diff_on = c->diff_set = false;
    c->dm.diff_itmin = nullptr;
    return MCRT_OK;
  }
  if (!(c->has_grid && c->has_dust && c->has_medium && c->has_source))
    return fail(c, MCRT_E_STATE, "mcrt_set_diffusion: set the grid, medium, dust and source first");
  if (!kappar || !hdisk) return fail(c, MCRT_E_ARG, "mcrt_set_diffusion: null array");
  if (nr0 < 1 || !(rmaxd > 0.0)) return fail(c, MCRT_E_ARG, "mcrt_set_diffusion: nr0 >= 1 and rmaxd > 0");
  if ((c->dm.ntg - 1) % 2 != 0) return fail(c, MCRT_E_STATE, "mcrt_set_diffusion: theta diffusion needs an even number of theta cells");
  const size_t nk = (size_t)c->dm.nopac * c->dm.nT, nr = (size_t)c->dm.nrg - 1;
  for (size_t i = 0; i < nk; ++i)
    if (!(kappar[i] > 0.0)) return fail(c, MCRT_E_ARG, "mcrt_set_diffusion: kappar must be > 0");
  for (size_t i = 0; i < nr; ++i)
    if (!(hdisk[i] >= 0.0)) return fail(c, MCRT_E_ARG, "mcrt_set_diffusion: hdisk must be >= 0");
  HIPCHECK(c, hipSetDevice(c->device));
  const double* d;
  int rc = upload(c, "diff.kappar", kappar, nk, &d);
  if (!rc) rc = upload(c, "diff.hdisk", hdisk, nr, &d);
  void* p;
  if (!rc) rc = dev_alloc(c, "diff.itmin", nr * 4, &p);
  if (!rc) rc = dev_alloc(c, "diff.told", c->ncell * 8, &p);
  if (rc) return rc;
  HIPCHECK(c, launch_fill_i32((int*)c->buf["diff.itmin"], nr, 0, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->diff_on = true;
  c->diff_set = false;
  c->diff_nr0 = nr0;
  c->diff_rmaxd = rmaxd;
  c->dm.diff_itmin = nullptr;
  return MCRT_OK;
}

int mcrt_setdiffus(mcrt_ctx* c, int32_t iter, int64_t* ncells) {
  if (!c) return MCRT_E_ARG;
  if (!c->diff_on) return fail(c, MCRT_E_STATE, "mcrt_setdiffus: diffusion is off (mcrt_set_diffusion)");
  if (!c->has_tdust || !c->dm.tdust) return fail(c, MCRT_E_STATE, "mcrt_setdiffus: set the dust temperatures first");
  HIPCHECK(c, hipSetDevice(c->device));
  // the last Lucy update's temperatures, else the ones the run would re-emit with
  auto tv = c->buf.find("tdave");
  const double* tdave = (tv != c->buf.end() && c->lucy_valid) ? (const double*)tv->second : c->dm.tdust;
  DiffArgs a{(const double*)c->buf["diff.kappar"], (const double*)c->buf["diff.hdisk"], tdave, c->diff_rmaxd, c->diff_nr0,
             (int)iter, (int32_t*)c->buf["diff.itmin"]};
  HIPCHECK(c, launch_setdiffus(c->dm, a, c->stream));
  c->dm.diff_itmin = a.itmin;
  // the temperatures before the reset: tfinal's convergence test compares with them
  HIPCHECK(c, hipMemcpyAsync(c->buf["diff.told"], c->dm.tdust, c->ncell * 8, hipMemcpyDeviceToDevice, c->stream));
  HIPCHECK(c, launch_diff_reset(c->dm, (double*)c->buf["tdust"], c->stream));
  std::vector<int32_t> itmin((size_t)c->dm.nrg - 1);
  HIPCHECK(c, hipMemcpyAsync(itmin.data(), a.itmin, itmin.size() * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  int64_t n = 0;
  for (int32_t v : itmin)
    if (v > 0) n += (int64_t)(c->dm.ntg - 2 * v + 1) * (c->dm.npg - 1);
  if (ncells) *ncells = n;
  c->diff_set = true;
  return MCRT_OK;
}

int mcrt_get_diffusion(mcrt_ctx* c, int32_t* itmin, uint32_t* count) {
  if (!c) return MCRT_E_ARG;
  if (!c->diff_on || !c->diff_set) return fail(c, MCRT_E_STATE, "mcrt_get_diffusion: no diffusion region set");
  HIPCHECK(c, hipSetDevice(c->device));
  if (itmin) HIPCHECK(c, hipMemcpyAsync(itmin, c->buf["diff.itmin"], ((size_t)c->dm.nrg - 1) * 4, hipMemcpyDeviceToHost, c->stream));
  if (count && c->dt.diff_count)
    HIPCHECK(c, hipMemcpyAsync(count, c->dt.diff_count, c->ncell * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return MCRT_OK;
}

int mcrt_dusttemp_sigma(mcrt_ctx* c, double* sigma, double* sigma_prev, int64_t* ncells, int32_t* converge) {
  if (!c) return MCRT_E_ARG;
  if (!c->lucy_valid) return fail(c, MCRT_E_STATE, "mcrt_dusttemp_sigma: no mcrt_dusttemp call yet");
  if (sigma) *sigma = c->lucy_sigma;
  if (sigma_prev) *sigma_prev = c->lucy_sigma_prev;
  if (ncells) *ncells = (int64_t)c->lucy_sum[2];
  // tfinal_5_test1.f90:354: converged when sigma changed by less than a
  // quarter of itself since the previous Lucy iteration (NaN compares false)
  if (converge) *converge = (std::fabs(c->lucy_sigma_prev - c->lucy_sigma) / c->lucy_sigma < 0.25) ? 1 : 0;
  return MCRT_OK;
}

int mcrt_trace_rays(mcrt_ctx* c, int32_t precision, int32_t n, const double* pos, const double* u, const int32_t* ijk,
                    const uint8_t* on_wall, const double* tau_stop, int32_t max_steps, int32_t* faces, int32_t* cells,
                    double* pos_out, int32_t* ijk_out, double* tau_out, int32_t* nsteps, int32_t* status) {
  if (!c) return MCRT_E_ARG;
  if (!(c->has_grid && c->has_medium)) return fail(c, MCRT_E_STATE, "mcrt_trace_rays: grid and medium must be set");
  if (n < 0 || max_steps < 1) return fail(c, MCRT_E_ARG, "mcrt_trace_rays: bad sizes");
  if (n == 0) return MCRT_OK;
  if (!pos || !u || !ijk || !on_wall || !tau_stop || !faces || !cells || !pos_out || !ijk_out || !tau_out || !nsteps || !status)
    return fail(c, MCRT_E_ARG, "mcrt_trace_rays: null array");
  for (int r = 0; r < n; ++r) {
    const int32_t* q = ijk + 3 * r;
    if (q[0] < 1 || q[0] > c->dm.nrg - 1 || q[1] < 1 || q[1] > c->dm.ntg - 1 || q[2] < 1 || q[2] > c->dm.npg - 1)
      return fail(c, MCRT_E_ARG, "mcrt_trace_rays: start cell outside the grid");
  }
  HIPCHECK(c, hipSetDevice(c->device));
  TraceArgs a;
  a.n = n; a.max_steps = max_steps;
  int rc = 0;
  rc |= upload(c, "tr.pos", pos, (size_t)3 * n, &a.pos);
  rc |= upload(c, "tr.u", u, (size_t)3 * n, &a.u);
  rc |= upload(c, "tr.tau", tau_stop, (size_t)n, &a.tau_stop);
  rc |= upload(c, "tr.ijk", (const int*)ijk, (size_t)3 * n, &a.ijk);
  rc |= upload(c, "tr.onw", on_wall, (size_t)n, &a.on_wall);
  void* p;
  size_t steps = (size_t)n * max_steps;
  rc |= dev_alloc(c, "tr.faces", steps * 4, &p); a.faces = (int*)p;
  rc |= dev_alloc(c, "tr.cells", steps * 12, &p); a.cells = (int*)p;
  rc |= dev_alloc(c, "tr.ijko", (size_t)n * 12, &p); a.ijk_out = (int*)p;
  rc |= dev_alloc(c, "tr.nst", (size_t)n * 4, &p); a.nsteps = (int*)p;
  rc |= dev_alloc(c, "tr.st", (size_t)n * 4, &p); a.status = (int*)p;
  rc |= dev_alloc(c, "tr.poso", (size_t)n * 24, &p); a.pos_out = (double*)p;
  rc |= dev_alloc(c, "tr.tauo", (size_t)n * 8, &p); a.tau_out = (double*)p;
  if (rc) return rc;
  HIPCHECK(c, launch_fill_i32(a.faces, steps, 0, c->stream));
  HIPCHECK(c, launch_fill_i32(a.cells, steps * 3, 0, c->stream));
  hipError_t e = (precision == MCRT_FP64) ? launch_trace_fp64(c->dm, a, c->stream) : launch_trace_mixed(c->dm, a, c->stream);
  if (e != hipSuccess) return hipfail(c, e, "trace kernel launch");
  hipStream_t s = c->stream;
  HIPCHECK(c, hipMemcpyAsync(faces, a.faces, steps * 4, hipMemcpyDeviceToHost, s));
  HIPCHECK(c, hipMemcpyAsync(cells, a.cells, steps * 12, hipMemcpyDeviceToHost, s));
  HIPCHECK(c, hipMemcpyAsync(ijk_out, a.ijk_out, (size_t)n * 12, hipMemcpyDeviceToHost, s));
  HIPCHECK(c, hipMemcpyAsync(nsteps, a.nsteps, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  HIPCHECK(c, hipMemcpyAsync(status, a.status, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  HIPCHECK(c, hipMemcpyAsync(pos_out, a.pos_out, (size_t)n * 24, hipMemcpyDeviceToHost, s));
  HIPCHECK(c, hipMemcpyAsync(tau_out, a.tau_out, (size_t)n * 8, hipMemcpyDeviceToHost, s));
  HIPCHECK(c, hipStreamSynchronize(s));
  return MCRT_OK;
}

int mcrt_mueller(mcrt_ctx* c, int32_t precision, int32_t n, int32_t kind, const double* par, double* dir, double* stokes) {
  if (!c) return MCRT_E_ARG;
  if (n < 0 || (kind != 0 && kind != 1)) return fail(c, MCRT_E_ARG, "mcrt_mueller: bad n or kind");
  if (n == 0) return MCRT_OK;
  if (!par || !stokes || (kind == 0 && !dir)) return fail(c, MCRT_E_ARG, "mcrt_mueller: null array");
  HIPCHECK(c, hipSetDevice(c->device));
  const double* dpar;
  void *pd, *ps;
  int rc = upload(c, "mu.par", par, (size_t)9 * n, &dpar);
  rc |= dev_alloc(c, "mu.dir", (size_t)5 * n * 8, &pd);
  rc |= dev_alloc(c, "mu.s", (size_t)4 * n * 8, &ps);
  if (rc) return rc;
  if (kind == 0) HIPCHECK(c, hipMemcpyAsync(pd, dir, (size_t)5 * n * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(ps, stokes, (size_t)4 * n * 8, hipMemcpyHostToDevice, c->stream));
  hipError_t e = (precision == MCRT_FP64) ? launch_mueller_fp64(n, kind, dpar, (double*)pd, (double*)ps, c->stream)
                                          : launch_mueller_mixed(n, kind, dpar, (double*)pd, (double*)ps, c->stream);
  if (e != hipSuccess) return hipfail(c, e, "mueller kernel launch");
  if (kind == 0) HIPCHECK(c, hipMemcpyAsync(dir, pd, (size_t)5 * n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipMemcpyAsync(stokes, ps, (size_t)4 * n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return MCRT_OK;
}

int mcrt_enable_trace(mcrt_ctx* c, int64_t n) {
  if (!c || n < 0) return MCRT_E_ARG;
  HIPCHECK(c, hipSetDevice(c->device));
  if (n == 0) { c->trace_n = 0; return MCRT_OK; }
  void* p;
  int rc = dev_alloc(c, "trace", (size_t)n * 8 * 8, &p);
  if (rc) return rc;
  HIPCHECK(c, launch_fill_f64((double*)p, (size_t)n * 8, 0.0, c->stream));
  c->trace_n = n;
  return MCRT_OK;
}

int mcrt_get_trace(mcrt_ctx* c, double* trace, int64_t n) {
  if (!c || !trace) return MCRT_E_ARG;
  if (n > c->trace_n) return fail(c, MCRT_E_ARG, "mcrt_get_trace: n exceeds the trace buffer");
  HIPCHECK(c, hipMemcpyAsync(trace, c->buf["trace"], (size_t)n * 64, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return MCRT_OK;
}

}  // extern "C"
