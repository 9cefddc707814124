// mcrt_batch.h -- many independent axisymmetric models in one persistent
// launch (no counterpart in the reference, whose model grids are separate
// program runs).
//
// The models share the star, the dust, the SED binning and the launch; each
// has its own radial and polar walls, densities, species, dust temperatures
// and tallies, stored back to back (BatchDesc).  One device queue holds the
// packets of every model, model after model in `order` (most expensive
// first, so that the long-lived packets of optically thick models start
// early and finish beside the bulk of the others).  A resident lane takes
// its next packet from that queue and binds the packet's model to its own
// copy of the model view: the shared tables stay wavefront-uniform (LDS),
// the per-model walls and cells become per-lane pointers.  Packet p of
// model k draws from Philox key seed + k, counter p, so it follows exactly
// the path it follows in a single-model run of model k with seed + k.
#pragma once
#include "mcrt_transport.h"

namespace mcrt {

struct BatchDesc {
  int nmodels;
  const int* nrg;                     // per model
  const int* ntg;
  const long long* roff;              // first radial wall of model k
  const long long* toff;              // first polar wall of model k
  const long long* coff;              // first cell of model k
  const double *rarr, *r2arr, *thetarr, *costarr, *tan2arr, *ccone, *scone;
  const double* rho64;
  const float* rho32;
  const uint8_t* species;
  const double* tdust;
  const double* vcell;
  const int* order;                   // queue position j runs model order[j]
  const unsigned long long* prefix;   // packets queued before position j (nmodels + 1)
  unsigned long long* work;           // per model: interactions + packets of the last run
  unsigned long long sed_stride;      // SED floats per model
};

// Values a captured launch reads at run time (device memory, written before
// each launch), so one graph serves every seed and packet count.
struct BatchRun {
  uint64_t seed;
  unsigned long long total;           // prefix[nmodels]
  unsigned long long head;            // queue head (zeroed by the graph's first node)
  uint32_t regroup;
  uint32_t trial;                     // autotuner timing: stop once every packet is handed out
};

// Per-lane model view of model k
template <class T>
__device__ __forceinline__ void bind_model(DeviceModel& m, DeviceTallies& t, const DeviceTallies& t0, const BatchDesc& b, int k) {
  const long long ro = b.roff[k], to = b.toff[k], co = b.coff[k];
  m.nrg = b.nrg[k];
  m.ntg = b.ntg[k];
  m.rarr = b.rarr + ro;
  m.r2arr = b.r2arr + ro;
  m.thetarr = b.thetarr + to;
  m.costarr = b.costarr + to;
  m.tan2arr = b.tan2arr + to;
  m.ccone = b.ccone + to;
  m.scone = b.scone + to;
  m.rho64 = b.rho64 + co;
  m.rho32 = b.rho32 + co;
  m.species = b.species + co;
  m.tdust = b.tdust + co;
  t.dtauabs = t0.dtauabs + co;
  t.nabs = t0.nabs + co;
  t.dbig = t0.dbig + co;
  t.killed = t0.killed + co;
  t.sed = t0.sed + (size_t)k * b.sed_stride;
}

// queue position of flat packet index g: the last j with prefix[j] <= g
__device__ __forceinline__ int batch_slot(const BatchDesc& b, unsigned long long g) {
  int lo = 0, hi = b.nmodels;          // prefix[lo] <= g < prefix[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (b.prefix[mid] <= g) lo = mid; else hi = mid;
  }
  return lo;
}

// The shared tables in LDS: phi walls (every model has the same single
// azimuthal cell), Planck grids, zeta(4) sums and, when planned, the dust.
__device__ __forceinline__ void stage_shared_tables(DeviceModel& m, const double*& inczeta, const LdsPlan& P, double* lds) {
  const double* src[LT_COUNT] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, m.phiarr, m.aarr, m.barr,
                                 m.lTint, m.nuint, m.peel, inczeta,
                                 m.nudust, m.kappad, m.adust, m.gdust, m.pldust};
#pragma unroll
  for (int k = LT_PHIARR; k < LT_COUNT; ++k)
    for (uint32_t i = threadIdx.x; i < P.n[k]; i += blockDim.x) lds[P.off[k] + i] = src[k][i];
  __syncthreads();
  m.phiarr = lds + P.off[LT_PHIARR];
  m.aarr = lds + P.off[LT_AARR];
  m.barr = lds + P.off[LT_BARR];
  m.lTint = lds + P.off[LT_LTINT];
  m.nuint = lds + P.off[LT_NUINT];
  inczeta = lds + P.off[LT_INCZETA];
  if (P.n[LT_NUDUST]) {
    m.nudust = lds + P.off[LT_NUDUST];
    m.kappad = lds + P.off[LT_KAPPAD];
    m.adust = lds + P.off[LT_ADUST];
    m.gdust = lds + P.off[LT_GDUST];
    m.pldust = lds + P.off[LT_PLDUST];
  }
}

// The persistent batch kernel: lanes = grid * BS threads, each running
// packets from the queue until it is empty.  The lane loop is k_transport's
// (walk steps for every walking lane; interactions and emissions once
// `regroup` lanes of the wavefront wait for them, or nobody walks) without
// an event budget: there is no peel-off queue to drain, so nothing is
// parked and the launch ends when the last packet does.  kComb selects the
// write-combined cell tallies (tally_abs).
template <class T, int BS, bool kComb>
__global__ void __launch_bounds__(BS) k_batch(DeviceModel m0, DeviceTallies t0, BatchDesc bd, BatchRun* run,
                                               const double* inczeta0, LdsPlan lds) {
  __shared__ unsigned long long s_cnt[C_NCOUNTERS];
  extern __shared__ double lds_tables[];
  counts_begin(s_cnt);
  DeviceModel m = m0;
  const double* inczeta = inczeta0;
  stage_shared_tables(m, inczeta, lds, lds_tables);
  DeviceTallies t = t0;
  Ctx<T> C{m, t, {}, {s_cnt}, inczeta, PeelQueue{nullptr, nullptr, 0}};
  const uint64_t seed = run->seed;
  const unsigned long long total = run->total;
  const uint32_t regroup = run->regroup;
  Life<T> L;
  tally_open(L);
  const bool trial = run->trial != 0;
  int st = L_IDLE;
  int model = 0;
  bool dry = false;
  for (uint32_t step = 0;; ++step) {
    // a timing trial measures the bulk rate only: the packets still in
    // flight when the queue runs dry are dropped (their tallies are void)
    if (trial && (step & 31u) == 0u &&
        __hip_atomic_load(&run->head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= total) break;
    if (st == L_WALK) {
      const int r = walk_step<T, false, kComb>(C, L);
      if (r != 0) {
        const int code = (r == 1) ? flight_end<T>(C, L) : ((r == 2) ? 1 : -1);
        if (code == 0) {
          st = L_READY;
        } else {
          if constexpr (kComb) tally_flush<T>(C, L);
          atomicAdd(&bd.work[model], (unsigned long long)L.nint + 1ull);
          end_packet<T>(C, L, code);
          st = L_IDLE;
        }
      }
    }
    const unsigned long long walking = __ballot(st == L_WALK);
    const unsigned long long ready = __ballot(st == L_READY);
    const unsigned long long idle = __ballot(st == L_IDLE && !dry);
    if ((walking | ready | idle) == 0ull) break;
    if (ready && (walking == 0ull || (unsigned)__popcll(ready) >= regroup)) {
      if (st == L_READY) {
        interact<T, false, kComb>(C, L);
        st = L_WALK;
      }
    }
    if (idle && (walking == 0ull || (unsigned)__popcll(idle) >= regroup)) {
      if (st == L_IDLE && !dry) {
        const unsigned long long g = fetch_packets(&run->head);
        if (g >= total) {
          dry = true;
        } else {
          const int j = batch_slot(bd, g);
          model = bd.order[j];
          const unsigned long long pk = g - bd.prefix[j];
          bind_model<T>(m, t, t0, bd, model);
          rng_init(C.rng, seed + (uint64_t)model, pk);
          L.pid = pk;
          const int e = emit<T>(C, L);
          if (e == 0) end_packet<T>(C, L, -1, C_ERR_OTHER);
          else st = (e == 1) ? L_WALK : L_READY;
        }
      }
    }
  }
  counts_flush(t0, s_cnt);
}


}  // namespace mcrt
