// mcrt_transport.h -- whole-packet transport on the device: stellar emission,
// direct peel, propagate (tauint / scatter / absorb + re-emit), peel-off
// imaging and SED binning.  Restates newdisktherm_5_fuv.f90:294-491,
// propagate_5_fuv.f90, tauint_3d_2_fuv.f90, peeloff_3d_2_test.f90.
#pragma once
#include "mcrt_device.h"

namespace mcrt {

// A peel-off event queued by the transport kernel: everything peeloff
// (peeloff_3d_2_test.f90:65-89) copies in from the packet at the event.
// One record serves all npeel directions.  For the direct stellar peel
// (newdisktherm_5_fuv.f90:371-414) the incident-direction slots carry
// cosb, sinb, cos(lp), sin(lp) and the weight is formed per direction.
template <class T>
struct alignas(16) PeelRec {
  G xp, yp, zp, rsq, rtot;
  T sip, sqp, sup, svp;
  T cost, sint, phi, cosp, sinp;
  T nub, wave, g, pl;
  T kapd[kMaxOpac];
  uint32_t cell;     // ir | it << 10 | ip << 20 | on_wall << 30
  uint32_t flags;    // iflag | i_orig << 1 | direct << 4 | species << 8
};

struct PeelQueue {
  void* recs;              // PeelRec<T>[capacity]
  unsigned* count;         // slots reserved this round
  unsigned capacity;
};

// Small read-only tables staged in LDS by every workgroup of the transport
// and peel kernels: the walk reads walls at every face crossing and an
// interaction searches the dust and Planck tables, so their latency sits on
// every packet's serial chain.  The tables before LT_NUDUST are always staged
// (unconditional LDS pointers let the compiler use ds_read); the five dust
// tables are staged when they fit (n[LT_NUDUST] = 0: they stay in global
// memory).
enum LdsTab : int {
  LT_RARR = 0, LT_R2ARR, LT_COSTARR, LT_TAN2ARR, LT_CCONE, LT_SCONE, LT_PHIARR, LT_AARR, LT_BARR,
  LT_LTINT, LT_NUINT, LT_PEEL, LT_INCZETA,
  LT_NUDUST, LT_KAPPAD, LT_ADUST, LT_GDUST, LT_PLDUST, LT_COUNT
};
struct LdsPlan {
  uint32_t off[LT_COUNT];   // offset of table k in doubles
  uint32_t n[LT_COUNT];     // doubles of table k
  uint32_t bytes;           // dynamic LDS per workgroup
};

// Device control block of one run (16 bytes, zeroed by the host at the start).
struct RunCtl {
  unsigned long long next_packet;   // packets handed out so far (run-relative)
  unsigned q_count;                 // peel records reserved this round
  unsigned live;                    // lanes holding a packet at the end of this round
};

struct RunParams {
  uint64_t first_packet;
  uint64_t npackets;
  uint64_t seed;
  const double* inczeta;   // 145 partial sums of zeta(4), random.f90:74-92
  PeelQueue q;             // q.capacity == 0: every peel-off walk runs inline
  // resident packet lanes: a lane owns one packet at a time and keeps it
  // across rounds in lane_words (structure of arrays, word w of lane l at
  // lane_words[w * lanes + l]); lane_live[l] != 0 while it holds one
  uint32_t* lane_words;
  uint8_t* lane_live;
  uint64_t lanes;
  uint32_t events;         // interaction events per lane in this round
  uint32_t round;          // 0 = first round of the run (no saved lanes)
  uint32_t regroup;        // waiting lanes of a wavefront that trigger an interaction / emission / tally pass
  uint32_t peel_regroup;   // idle lanes of a peel wavefront that trigger taking new items (with their first crossing)
  uint32_t peel_tally;     // lanes of a peel wavefront whose walks reached the observer that trigger a tally pass
  RunCtl* ctl;
  LdsPlan lds;
};

// Copy the planned tables into this workgroup's LDS and point the kernel's
// model copy at them.  Every thread of the workgroup must call it.
// kDustLds: the plan is known to stage the dust tables, so their pointers
// are LDS addresses at compile time (ds_read, not flat loads that also wait
// on the wavefront's outstanding global stores).
template <bool kDustLds = false>
__device__ __forceinline__ void stage_tables(DeviceModel& m, const double*& inczeta, const LdsPlan& P, double* lds) {
  const double* src[LT_COUNT] = {m.rarr, m.r2arr, m.costarr, m.tan2arr, m.ccone, m.scone, m.phiarr, m.aarr, m.barr,
                                 m.lTint, m.nuint, m.peel, inczeta,
                                 m.nudust, m.kappad, m.adust, m.gdust, m.pldust};
#pragma unroll
  for (int k = 0; k < LT_COUNT; ++k)
    for (uint32_t i = threadIdx.x; i < P.n[k]; i += blockDim.x) lds[P.off[k] + i] = src[k][i];
  __syncthreads();
  m.rarr = lds + P.off[LT_RARR];
  m.r2arr = lds + P.off[LT_R2ARR];
  m.costarr = lds + P.off[LT_COSTARR];
  m.tan2arr = lds + P.off[LT_TAN2ARR];
  m.ccone = lds + P.off[LT_CCONE];
  m.scone = lds + P.off[LT_SCONE];
  m.phiarr = lds + P.off[LT_PHIARR];
  m.aarr = lds + P.off[LT_AARR];
  m.barr = lds + P.off[LT_BARR];
  m.lTint = lds + P.off[LT_LTINT];
  m.nuint = lds + P.off[LT_NUINT];
  m.peel = lds + P.off[LT_PEEL];
  inczeta = lds + P.off[LT_INCZETA];
  if (kDustLds || P.n[LT_NUDUST]) {
    m.nudust = lds + P.off[LT_NUDUST];
    m.kappad = lds + P.off[LT_KAPPAD];
    m.adust = lds + P.off[LT_ADUST];
    m.gdust = lds + P.off[LT_GDUST];
    m.pldust = lds + P.off[LT_PLDUST];
  }
}

// Path-replay adjoint of the peeled images (k_adjoint): the loss weights,
// the gradient, and the bookkeeping of one packet's two passes.
struct AdjState {
  const float* W;      // loss weights in the image layout [nband][nx][nx][npeel]
  double* grad;        // d loss / d rho per cell
  double R;            // replay: loss of the packet's events still ahead
  double total;        // forward: loss of the packet
  int replay;          // 0 forward pass, 1 replay pass
  uint64_t pkey;       // RNG key of the packet's peel walks
  uint32_t nevent;     // peel-off events of the packet so far
};

template <class T>
struct Ctx {
  const DeviceModel& m;
  const DeviceTallies& t;
  Rng rng;
  LocalCounts cnt;
  const double* inczeta;
  PeelQueue q;
  unsigned slot_next, slot_end;   // this lane's reserved, not yet written queue slots
  AdjState* adj;                  // null except in k_adjoint (the hooks then fold away)
  // k_tail: all 64 lanes of a wavefront carry the same packet; lane 0 alone
  // writes, and the rejection sampling of the scattering angle tries 64
  // candidates at once
  bool coop = false;
  CdfCache<T> cdf;                // k_tail: the re-emission row of the last draw (LDS)
};

// the lane that writes the packet's tallies and records
template <class T>
__device__ __forceinline__ bool writer(const Ctx<T>& C) { return !C.coop || __lane_id() == 0; }

template <class T> struct Life;
template <class T> struct Pkt;
// adjoint hooks (defined with k_adjoint)
template <class T>
__device__ __forceinline__ void adj_segment(Ctx<T>& C, size_t c, T kapd, G len, T rho, bool ends_here);
template <class T>
__device__ __forceinline__ void adj_peel_event(Ctx<T>& C, const Pkt<T>& p, int sp, int iop, bool direct, T coslp, T sinlp);

// Peel-off queue slots are reserved kSlotChunk at a time per lane, so an
// interaction usually writes its record without waiting on an atomic.  Slots
// a lane reserved but did not fill are marked kSlotHole at the end of the
// round; the peel kernel skips them.  The queue of a round is shared by all
// lanes: a lane reserves its next chunk before an event that needs a slot,
// and once the reservations pass the capacity the round is over for it.
constexpr unsigned kSlotChunk = 8;
constexpr uint32_t kSlotHole = 0xFFFFFFFFu;

template <class T> __device__ __forceinline__ void atomic_add_dtau(double* a, T v) { atomicAdd(a, (double)v); }

// One atomic per wavefront: each active lane reserves `per` consecutive slots
// (coop: the wavefront's one packet reserves them, every lane sees the base).
__device__ __forceinline__ unsigned reserve_slots(unsigned* counter, unsigned per, bool coop = false) {
  if (coop) {
    unsigned b = 0;
    if (__lane_id() == 0) b = atomicAdd(counter, per);
    return __shfl(b, 0, 64);
  }
  unsigned long long mask = __ballot(1);
  int lane = __lane_id();
  int leader = __ffsll((long long)mask) - 1;
  unsigned rank = (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
  unsigned base = 0;
  if (lane == leader) base = atomicAdd(counter, (unsigned)__popcll(mask) * per);
  base = __shfl(base, leader, 64);
  return base + rank * per;
}

// Slots per reservation: kSlotChunk per lane, 64 for a wavefront's one
// packet (coop), whose reservation is an atomic round trip on its serial
// chain.
template <class T>
__device__ __forceinline__ unsigned slot_chunk(const Ctx<T>& C) { return C.coop ? 64u : kSlotChunk; }

// the next queue slot of this lane (>= capacity: the queue is full)
template <class T>
__device__ __forceinline__ unsigned take_slot(Ctx<T>& C) {
  if (C.slot_next == C.slot_end) {
    C.slot_next = reserve_slots(C.q.count, slot_chunk(C), C.coop);
    C.slot_end = C.slot_next + slot_chunk(C);
  }
  return C.slot_next++;
}

// Make sure this lane holds a free slot for its next event's record; false:
// the round's queue is full (the lane stops for this round).  Every lane of
// the calling branch reserves together (one atomic per wavefront).
template <class T>
__device__ __forceinline__ bool hold_slot(Ctx<T>& C) {
  if (C.q.capacity == 0 || C.slot_next < C.slot_end) return true;
  const unsigned base = reserve_slots(C.q.count, slot_chunk(C), C.coop);
  if (base >= C.q.capacity) { C.slot_next = C.slot_end = 0u; return false; }
  C.slot_next = base;
  C.slot_end = (base + slot_chunk(C) < C.q.capacity) ? base + slot_chunk(C) : C.q.capacity;
  return true;
}

// initp_common.f90:1-31 (limb = 0)
template <class T>
__device__ __forceinline__ void initp_sample(Ctx<T>& C, Pkt<T>& p) {
  T xran = uniform<T>(C.rng);
  p.cost = RT<T>::sqrt_(xran);
  p.sint = RT<T>::sqrt_((T)1 - p.cost * p.cost);
  xran = uniform<T>(C.rng);
  p.phi = (T)kR2p * xran;
  p.cosp = RT<T>::cos_(p.phi);
  p.sinp = RT<T>::sin_(p.phi);
}

// initp_common.f90:31-92
template <class T>
__device__ __forceinline__ void initp_transform(Pkt<T>& p) {
  const T PI = (T)kPi, R2P = (T)kR2p, one = (T)1;
  T costi, sinti, cospn, phinew;
  if (RT<T>::fabs_(p.cosb) < (T)0.9999999) {
    costi = p.cost * p.cosb - p.sint * p.sinb * p.cosp;
    if (RT<T>::fabs_(costi) > one) { if (costi > one) costi = one; if (p.cost < -one) costi = -one; }
    sinti = RT<T>::sqrt_(one - costi * costi);
    if (sinti > (T)0.0000001) {
      cospn = (p.cost - costi * p.cosb) / (sinti * p.sinb);
      if (RT<T>::fabs_(cospn) > one) cospn = (cospn < (T)0) ? -one : one;
      if (p.phi < PI) { phinew = RT<T>::acos_(cospn); if (phinew > PI) phinew = PI; }
      else { phinew = R2P - RT<T>::acos_(cospn); if (phinew < PI) phinew = PI; }
    } else phinew = (T)0;
    p.phi = phinew; p.cost = costi; p.sint = sinti;
  } else {
    if (p.cosb < (T)0) p.cost = -p.cost;
  }
  p.phi = p.phi + p.lp;
  if (p.phi > R2P) p.phi = p.phi - R2P;
  if (p.phi < (T)0) p.phi = p.phi + R2P;
  p.sinp = RT<T>::sin_(p.phi);
  p.cosp = RT<T>::cos_(p.phi);
}

// initpabs.f90 (blackbody star: frequency from bbfreq)
template <class T>
__device__ __forceinline__ void initpabs(Ctx<T>& C, Pkt<T>& p) {
  p.sip = (T)1; p.sqp = (T)0; p.sup = (T)0; p.svp = (T)0;
  p.cosb = (T)(p.zp / p.rtot);
  p.sinb = RT<T>::sqrt_((T)1 - p.cosb * p.cosb);
  p.lp = RT<T>::atan2_((T)p.yp, (T)p.xp);
  initp_sample(C, p);
  initp_transform(p);
  set_dir<T>(p);
  p.nub = bbfreq<T>(C.rng, (T)C.m.tstar, C.inczeta);
}

// ---------------------------------------------------------------------------
// A packet's life, in resumable pieces.  propagate_5_fuv.f90:1-326 is a loop
// of (interaction, flight), and a flight (tauint_3d_2_fuv.f90) is a loop of
// face crossings.  Life<T> is the state between two crossings.  Run back to
// back, the pieces below are exactly propagate.  Split, they let a
// wavefront advance every walking lane by one crossing per step and run the
// costly interaction code only when enough of its lanes are waiting for it.
// ---------------------------------------------------------------------------
template <class T>
struct Life {
  Pkt<T> p;
  G xpold, ypold, zpold;
  int idust2, iscat, iint, nint;
  uint64_t pid;        // run-relative packet index (trace slot)
  T tau, tsum;         // current flight: optical depth drawn / integrated so far
  long long nseg;      // face crossings of the current flight
  int pir, pit, pip;   // cell the flight started in
  bool first;          // first flight after the emission
  // per-cell values already read for the current cell (a packet trapped in
  // an optically thick cell interacts there thousands of times): density,
  // species and dust temperature, each with the cell index it belongs to
  uint32_t c_rho, c_sp, c_td;
  T rho;
  int sp, spop;        // species code and its opacity table
  // write-combined tallies of the cell visit in progress (kComb): dtauabs,
  // and the absorption counters dbig and nabs
  uint32_t acc_c;
  double acc_v;
  uint32_t acc_dbig, acc_nabs;
  T td;
  // wall_clearance at the packet's position, from peel_hidden at an
  // interaction for the walk_step that starts the next flight there
  double wclear;
  bool wclear_ok;
};

template <class T>
__device__ __forceinline__ void cell_cache_reset(Life<T>& L) {
  L.c_rho = L.c_sp = L.c_td = 0xFFFFFFFFu;
  L.wclear_ok = false;
}

template <class T>
__device__ __forceinline__ int cell_species(const DeviceModel& m, Life<T>& L, size_t c) {
  if (L.c_sp != (uint32_t)c) { L.sp = m.species[c]; L.spop = m.species_opac[L.sp]; L.c_sp = (uint32_t)c; }
  return L.sp;
}

template <class T>
__device__ __forceinline__ T cell_tdust(const DeviceModel& m, Life<T>& L, size_t c) {
  if (L.c_td != (uint32_t)c) { L.td = (T)m.tdust[c]; L.c_td = (uint32_t)c; }
  return L.td;
}

// cell_opac through the cache
template <class T>
__device__ __forceinline__ T cell_opac_l(const DeviceModel& m, Life<T>& L, size_t c, int& iop, T& rho) {
  if (L.c_rho != (uint32_t)c) { L.rho = cell_rho<T>(m, c); L.c_rho = (uint32_t)c; }
  rho = L.rho;
  const Pkt<T>& p = L.p;
  if (m.nopac == 1) {
    iop = rho > (T)0 ? 0 : -1;
    return rho > (T)0 ? (T)0 + p.kapd[0] * rho : (T)0;
  }
  int sp = cell_species<T>(m, L, c);
  iop = -1;
  if (sp > 0 && rho > (T)0) { iop = m.species_opac[sp]; return (T)0 + pick(p.kapd, iop) * rho; }
  return (T)0;
}

// The Lucy tally of one segment (tauint_3d_2_fuv.f90:83-90,122-130) and the
// absorption counters of an interaction (propagate_5_fuv.f90:139-141).  With
// kComb the lane sums them over its visit to one cell and adds the sums when
// it leaves the cell, ends the packet or parks it (tally_flush): a packet
// trapped in a thick cell then issues a few atomics per visit instead of up
// to three per interaction.  This matters for the packets whose serial
// chains end a run: on this GPU a vector memory instruction's completion
// counter also counts stores and atomics, so every atomic issued before the
// next load of the chain adds its round trip to that load's wait.  Integer
// counts are unchanged; dtauabs changes only by the order of its FP64 sums.
template <class T>
__device__ __forceinline__ void tally_flush(Ctx<T>& C, Life<T>& L) {
  if (L.acc_c != 0xFFFFFFFFu && writer(C)) {
    if (L.acc_v != 0.0) atomicAdd(&C.t.dtauabs[L.acc_c], L.acc_v);
    if (L.acc_dbig) atomicAdd(&C.t.dbig[L.acc_c], (int)L.acc_dbig);
    if (L.acc_nabs) atomicAdd(&C.t.nabs[L.acc_c], (int)L.acc_nabs);
  }
  L.acc_c = 0xFFFFFFFFu;
  L.acc_v = 0.0;
  L.acc_dbig = 0u;
  L.acc_nabs = 0u;
}

template <class T>
__device__ __forceinline__ void tally_open(Life<T>& L) {
  L.acc_c = 0xFFFFFFFFu;
  L.acc_v = 0.0;
  L.acc_dbig = 0u;
  L.acc_nabs = 0u;
}

template <bool kComb, class T>
__device__ __forceinline__ void tally_abs(Ctx<T>& C, Life<T>& L, size_t c, double v) {
  if constexpr (kComb) {
    if (L.acc_c != (uint32_t)c) { tally_flush<T>(C, L); L.acc_c = (uint32_t)c; }
    L.acc_v += v;
  } else {
    if (writer(C)) atomicAdd(&C.t.dtauabs[c], v);
  }
}

// dbig (every absorption) and nabs (when counted) of cell c
template <bool kComb, class T>
__device__ __forceinline__ void tally_absorb(Ctx<T>& C, Life<T>& L, size_t c, bool nabs) {
  if constexpr (kComb) {
    if (L.acc_c != (uint32_t)c) { tally_flush<T>(C, L); L.acc_c = (uint32_t)c; }
    L.acc_dbig++;
    if (nabs) L.acc_nabs++;
  } else if (writer(C)) {
    atomicAdd(&C.t.dbig[c], 1);
    if (nabs) atomicAdd(&C.t.nabs[c], 1);
  }
}

template <class T>
__device__ __forceinline__ void flight_begin(Life<T>& L, T tau) {
  L.tau = tau; L.tsum = (T)0; L.nseg = 0;
  L.pir = L.p.ir; L.pit = L.p.it; L.pip = L.p.ip;
}

// One pass of tauint_3d_2_fuv.f90's face loop (:40-205).  0 = the flight
// goes on, 1 = the interaction point was reached inside the cell, 2 = the
// packet left the grid, -1 = walk error.
// The reference checks the position against the cell indices (testgrid)
// before every face crossing.  With FP64 positions the check finds nothing
// to repair in all but ~1e-7 of the steps, so the mixed path (T = float)
// runs it only when the face solvers find no wall, then solves again; the
// FP64 path keeps the reference's order exactly.
template <class T>
constexpr bool kLazyTestgrid = std::is_same<T, float>::value;

template <class T, bool kAdj = false, bool kComb = false>
__device__ __forceinline__ int walk_step(Ctx<T>& C, Life<T>& L) {
  const DeviceModel& m = C.m;
  Pkt<T>& p = L.p;
  G dt;
  int iface;
  bool cached = L.wclear_ok;            // set at this position and cell by peel_hidden
  L.wclear_ok = false;
  if constexpr (!kLazyTestgrid<T>) { if (testgrid<T>(m, p, C.rng)) { C.cnt.add(C_RESETS); cached = false; } }
  // In-cell fast path.  At the start of a flight (tsum = 0) off any wall, if
  // the flight length tau / opac is below half the wall clearance, every face
  // root is longer than twice the flight, so the reference's test
  // opac * dt > tau holds whatever root closest_wall would pick (rounding
  // included).  The interaction point is then computed exactly as below
  // without solving for the faces.  This is the common case of packets
  // trapped in optically thick cells.
  if (!p.on_wall && L.tsum == (T)0) {
    const size_t c = cell_index(m, p.ir, p.it, p.ip);
    int iop;
    T rho;
    const T opac = cell_opac_l<T>(m, L, c, iop, rho);
    if (opac > (T)0) {
      const T ds = (L.tau - L.tsum) / opac;
      if ((G)ds < 0.5 * (cached ? L.wclear : wall_clearance<T>(m, p))) {
        ++L.nseg;
        C.cnt.add_lanes(C_SEGMENTS);
        if (iop >= 0) tally_abs<kComb>(C, L, c, (double)ds * (double)pick(p.kapd, iop) * (double)rho * (1.0 - (double)pick(p.op, iop).albedo));
        if constexpr (kAdj) { if (iop >= 0) adj_segment<T>(C, c, pick(p.kapd, iop), (G)ds, rho, true); }
        p.xp = p.xp + (G)ds * p.ux;
        p.yp = p.yp + (G)ds * p.uy;
        p.zp = p.zp + (G)ds * p.uz;
        p.rsq = p.xp * p.xp + p.yp * p.yp + p.zp * p.zp;
        p.rtot = sqrt(p.rsq);
        p.on_wall = false;
        return 1;
      }
    }
  }
  if (!next_face<T>(m, p, p.ux, p.uy, p.uz, dt, iface)) {
    if constexpr (!kLazyTestgrid<T>) return -1;
    if (testgrid<T>(m, p, C.rng)) C.cnt.add(C_RESETS);
    if (!next_face<T>(m, p, p.ux, p.uy, p.uz, dt, iface)) return -1;
  }
  if (++L.nseg > m.max_segments) return -1;
  C.cnt.add_lanes(C_SEGMENTS);
  size_t c = cell_index(m, p.ir, p.it, p.ip);
  int iop;
  T rho;
  T opac = cell_opac_l<T>(m, L, c, iop, rho);
  T dtau = opac * (T)dt;
  if ((L.tsum + dtau) > L.tau) {
    T ds = (L.tau - L.tsum) / opac;
    if (iop >= 0) tally_abs<kComb>(C, L, c, (double)ds * (double)pick(p.kapd, iop) * (double)rho * (1.0 - (double)pick(p.op, iop).albedo));
    if constexpr (kAdj) { if (iop >= 0) adj_segment<T>(C, c, pick(p.kapd, iop), (G)ds, rho, true); }
    p.xp = p.xp + (G)ds * p.ux;
    p.yp = p.yp + (G)ds * p.uy;
    p.zp = p.zp + (G)ds * p.uz;
    p.rsq = p.xp * p.xp + p.yp * p.yp + p.zp * p.zp;
    p.rtot = sqrt(p.rsq);
    p.on_wall = false;
    return 1;
  }
  if (iop >= 0) tally_abs<kComb>(C, L, c, (double)dt * (double)pick(p.kapd, iop) * (double)rho * (1.0 - (double)pick(p.op, iop).albedo));
  if constexpr (kAdj) { if (iop >= 0) adj_segment<T>(C, c, pick(p.kapd, iop), dt, rho, false); }
  cross_face(m, p.ir, p.it, p.ip, iface);
  p.xp = p.xp + dt * p.ux;
  p.yp = p.yp + dt * p.uy;
  p.zp = p.zp + dt * p.uz;
  p.on_wall = true;
  p.rsq = p.xp * p.xp + p.yp * p.yp + p.zp * p.zp;
  p.rtot = sqrt(p.rsq);
  L.tsum = L.tsum + dtau;
  if (p.ir == m.nrg) { p.exitflag = 1; return 2; }
  if (p.ir == 0) {     // back into the star: re-emitted from its surface (:170-200)
    C.cnt.add(C_AFLUX);
    p.rtot = sqrt(p.rsq);
    initpabs(C, p);
    opacset<T>(m, p.nub, p.op, C.coop);
    p.wave = (T)1.2398 / p.nub;
    p.i_orig = 1;
    p.iflag = 0;
    set_kapd(p, m, 2);
    p.ir = 1;
    L.tsum = (T)0;
  }
  return 0;
}

__device__ __forceinline__ int i_orig_of(int sp) {  // propagate_5_fuv.f90:123-132
  if (sp == 1 || sp == 2 || sp == 5 || sp == 6 || sp == 9) return 2;
  if (sp == 3 || sp == 7) return 3;
  return 4;
}

template <class T>
__device__ __forceinline__ int freq_bin(const DeviceModel& m, T nub) {
  int inub = (int)((T)m.nfreq * RT<T>::log_(nub / (T)m.numin) / (T)m.lnurat) + 1;
  if (inub < 1) inub = 1;
  if (inub > m.nfreq) inub = m.nfreq;
  return inub;
}

// ---------------------------------------------------------------------------
// peeloff_3d_2_test.f90:7-347 in three pieces: the set-up of the walk toward
// observer direction k (peel_begin), one face crossing of that walk
// (peel_step), and the weighting and binning once it leaves the grid
// (peel_tally).  In order they are exactly peeloff.
// ---------------------------------------------------------------------------
template <class T>
struct PeelWalk {
  Pkt<T> p;          // position, cell, kapd of the walk
  G ux, uy, uz;      // observer direction
  T tsum;
  long long nseg;
};

// :65-89.  false: the walk starts on the stellar surface heading inwards
// (nothing reaches the observer).
template <class T>
__device__ __forceinline__ bool peel_begin(const DeviceModel& m, const Pkt<T>& q, int k, PeelWalk<T>& w) {
  const double* pe = m.peel + 5 * k;
  w.p = q;
  w.ux = pe[1] * pe[2]; w.uy = pe[1] * pe[3]; w.uz = pe[0];
  w.tsum = (T)0;
  w.nseg = 0;
  Pkt<T>& p = w.p;
  if (p.on_wall) {
    const double* r2 = m.r2arr;
    bool inner = fabs(p.rsq - r2[p.ir - 1]) < fabs(p.rsq - r2[p.ir]);
    bool outwards = w.ux * p.xp + w.uy * p.yp + w.uz * p.zp > 0.0;
    if (inner && !outwards) { if (p.ir == 1) return false; p.ir = p.ir - 1; }
    if (!inner && outwards) p.ir = p.ir + 1;
  }
  return true;
}

// :91-190.  0 = the walk goes on, 1 = it left the grid toward the observer,
// 2 = it ended (tau > taumax, the star, or a walk error).  rng: testgrid's
// re-placements.  grad != null (adjoint): adds coef * kapd * length to every
// crossed cell, the derivative of the walk's optical depth times coef.
template <class T>
__device__ __forceinline__ int peel_step(Ctx<T>& C, PeelWalk<T>& w, Rng& rng, double* grad = nullptr, double coef = 0.0) {
  const DeviceModel& m = C.m;
  Pkt<T>& p = w.p;
  G dt;
  int iface;
  if constexpr (!kLazyTestgrid<T>) { if (testgrid<T>(m, p, rng)) C.cnt.add(C_RESETS); }
  bool found = next_face<T>(m, p, w.ux, w.uy, w.uz, dt, iface);
  if constexpr (kLazyTestgrid<T>) {
    if (!found) {
      if (testgrid<T>(m, p, rng)) C.cnt.add(C_RESETS);
      found = next_face<T>(m, p, w.ux, w.uy, w.uz, dt, iface);
    }
  }
  if (!found || ++w.nseg > m.max_segments) { C.cnt.add(C_ERRORS); C.cnt.add(C_ERR_PEEL); return 2; }
  C.cnt.add_lanes(C_PEELSEG);
  size_t c = cell_index(m, p.ir, p.it, p.ip);
  int iop;
  T rho;
  T opac = cell_opac<T>(m, p, c, iop, rho);
  if (grad && iop >= 0) atomicAdd(&grad[c], coef * (double)pick(p.kapd, iop) * (double)dt);
  T dtau = opac * (T)dt;
  w.tsum = w.tsum + dtau;
  if (w.tsum > (T)m.taumax) return 2;
  cross_face(m, p.ir, p.it, p.ip, iface);
  p.xp = p.xp + dt * w.ux;
  p.yp = p.yp + dt * w.uy;
  p.zp = p.zp + dt * w.uz;
  p.on_wall = true;
  p.rsq = p.xp * p.xp + p.yp * p.yp + p.zp * p.zp;
  p.rtot = sqrt(p.rsq);
  if (p.ir == m.nrg) return 1;
  if (p.ir == 0) return 2;
  return 0;
}

// ---------------------------------------------------------------------------
// Peel-off walk by ray parameter (mixed path).  A peel walk is a straight ray
// x(t) = o + t u from the first face it crosses.  Every wall of the grid is a
// quadric in t: a shell |x|^2 = R^2 has roots -b +- sqrt(R^2 - h^2) (b = o.u,
// h = |o x u|), a cone x^2 + y^2 = tan^2(theta) z^2 a quadratic, a phi plane
// one root.  The walk keeps, for the current cell, the ray parameter at which
// it leaves through each of the three axes; a crossing changes one cell index,
// so only that axis is solved again.  The wall just crossed is recognized by
// its root (the one nearest the current parameter) instead of the reference's
// on-wall tie rule, and no per-step position is formed.  The FP64 path keeps
// the reference's per-step solvers from the current position (peel_step).
// The first crossing of every walk is the reference's (peel_ray_first), so
// walks that start on a wall resolve exactly as in peeloff.
// ---------------------------------------------------------------------------
template <class T>
constexpr bool kRayPeel = std::is_same<T, float>::value;

constexpr double kRayNone = 1.0e300;

template <class T>
struct PeelRay {
  G ox, oy, oz;        // origin: where the first crossing left the walk
  G b, h2;             // o.u and |o x u|^2
  G t;                 // current ray parameter
  G tr;                // exit through the radial walls
  G tlo, tlo2;         // next two crossings of the lower polar wall (thetarr[it-1]) beyond t
  G thi, thi2;         // the same for the upper polar wall (thetarr[it])
  G tp;                // exit through the azimuthal walls
  int dr, dp;          // cell-index step of the radial / azimuthal exit
  bool started;
};

// shell exit of cell ir at parameter t: the inner wall when the ray still
// descends and reaches it, else the outer wall (possibly the one just entered,
// at its second root)
template <class T>
__device__ __forceinline__ void ray_radial(const DeviceModel& m, const PeelRay<T>& Rc, int ir, G& tr, int& dr) {
  const G sc = -Rc.b;
  const G R2lo = m.r2arr[ir - 1], R2hi = m.r2arr[ir];
  if (Rc.t < sc && R2lo > Rc.h2) {
    const G q = sc - sqrt(R2lo - Rc.h2);
    if (q > Rc.t) { tr = q; dr = -1; return; }
  }
  const G d = R2hi - Rc.h2;
  tr = sc + sqrt(d > 0.0 ? d : 0.0);
  dr = 1;
}

// roots (ascending) of the cone tan2 (or the equator plane, tan2 < 0) along the ray
template <class T>
__device__ __forceinline__ int ray_cone_roots(const PeelRay<T>& R, G ux, G uy, G uz, G tan2, G& r1, G& r2) {
  if (tan2 < 0.0) {                       // equator: z = 0
    if (uz == 0.0) return 0;
    r1 = -R.oz / uz;
    return 1;
  }
  if (tan2 == 0.0) return 0;              // pole
  const G A = ux * ux + uy * uy - tan2 * uz * uz;
  const G B = R.ox * ux + R.oy * uy - tan2 * R.oz * uz;
  const G Cc = R.ox * R.ox + R.oy * R.oy - tan2 * R.oz * R.oz;
  if (fabs(A) > 1.0e-10) {
    const G disc = B * B - A * Cc;
    if (!(disc > 0.0)) return 0;
    const G q = -(B + copysign(sqrt(disc), B));
    const G a = q / A, c = (q != 0.0) ? Cc / q : a;
    r1 = fmin(a, c);
    r2 = fmax(a, c);
    return 2;
  }
  if (fabs(B) > 0.5e-10) { r1 = -Cc / (2.0 * B); return 1; }
  return 0;
}

// next two crossings beyond t of one polar wall with roots r1 <= r2; when the
// ray has just entered the cell through this wall, the root nearest t is that
// crossing and is left out
__device__ __forceinline__ void ray_wall_next(G t, bool entered, int n, G r1, G r2, G& nx, G& nx2) {
  nx = kRayNone;
  nx2 = kRayNone;
  if (n == 1) {
    if (!entered && r1 > t) nx = r1;
    return;
  }
  if (n == 2) {
    if (entered) {
      const G far = (fabs(r1 - t) < fabs(r2 - t)) ? r2 : r1;
      if (far > t) nx = far;
    } else if (r1 > t) {
      nx = r1;
      nx2 = r2;
    } else if (r2 > t) {
      nx = r2;
    }
  }
}

// azimuthal exit of cell ip (each plane is crossed once: the wall entered
// through is skipped); came: 0 none, 1 the lower wall, 2 the upper wall
template <class T>
__device__ __forceinline__ void ray_azimuth(const DeviceModel& m, const PeelRay<T>& R, G ux, G uy, int ip, int came,
                                            G& tp, int& dp) {
  tp = kRayNone;
  dp = 0;
  if (m.npg <= 2 || m.ntg <= 2) return;
  if (came != 1) {
    const G A = m.aarr[ip - 1], B = m.barr[ip - 1];
    const G den = A * ux + B * uy;
    if (den != 0.0) {
      const G r = -(A * R.ox + B * R.oy) / den;
      if (r > R.t && r < tp) { tp = r; dp = -1; }
    }
  }
  if (came != 2) {
    const G A = m.aarr[ip], B = m.barr[ip];
    const G den = A * ux + B * uy;
    if (den != 0.0) {
      const G r = -(A * R.ox + B * R.oy) / den;
      if (r > R.t && r < tp) { tp = r; dp = 1; }
    }
  }
}

// Set up the ray at the point where the first crossing (face iface) left the
// walk, in the cell it entered.
template <class T>
__device__ __forceinline__ void ray_start(const DeviceModel& m, PeelWalk<T>& w, PeelRay<T>& R, int iface) {
  const Pkt<T>& p = w.p;
  R.ox = p.xp; R.oy = p.yp; R.oz = p.zp;
  R.b = p.xp * w.ux + p.yp * w.uy + p.zp * w.uz;
  const G cx = p.yp * w.uz - p.zp * w.uy, cy = p.zp * w.ux - p.xp * w.uz, cz = p.xp * w.uy - p.yp * w.ux;
  R.h2 = cx * cx + cy * cy + cz * cz;
  R.t = 0.0;
  // a "-" face (1, 3, 5) enters the next cell through its upper wall, a "+" face through its lower wall
  ray_radial<T>(m, R, p.ir, R.tr, R.dr);
  R.tlo = R.tlo2 = R.thi = R.thi2 = kRayNone;
  if (m.ntg > 2) {
    G r1 = 0.0, r2 = 0.0;
    int n = ray_cone_roots<T>(R, w.ux, w.uy, w.uz, m.tan2arr[p.it - 1], r1, r2);
    ray_wall_next(R.t, iface == 4, n, r1, r2, R.tlo, R.tlo2);
    n = ray_cone_roots<T>(R, w.ux, w.uy, w.uz, m.tan2arr[p.it], r1, r2);
    ray_wall_next(R.t, iface == 3, n, r1, r2, R.thi, R.thi2);
  }
  ray_azimuth<T>(m, R, w.ux, w.uy, p.ip, (iface == 5) ? 2 : ((iface == 6) ? 1 : 0), R.tp, R.dp);
  R.started = true;
}

// the walk's first crossing: peel_step's reference solvers from the record
// position (on-wall tie rule included), keeping the face it crossed
template <class T>
__device__ __forceinline__ int peel_ray_first(Ctx<T>& C, PeelWalk<T>& w, PeelRay<T>& R, Rng& rng) {
  const DeviceModel& m = C.m;
  Pkt<T>& p = w.p;
  G dt;
  int iface;
  bool found = next_face<T>(m, p, w.ux, w.uy, w.uz, dt, iface);
  if (!found) {
    if (testgrid<T>(m, p, rng)) C.cnt.add(C_RESETS);
    found = next_face<T>(m, p, w.ux, w.uy, w.uz, dt, iface);
  }
  if (!found || ++w.nseg > m.max_segments) { C.cnt.add(C_ERRORS); C.cnt.add(C_ERR_PEEL); return 2; }
  C.cnt.add_lanes(C_PEELSEG);
  size_t c = cell_index(m, p.ir, p.it, p.ip);
  int iop;
  T rho;
  T opac = cell_opac<T>(m, p, c, iop, rho);
  w.tsum = w.tsum + opac * (T)dt;
  if (w.tsum > (T)m.taumax) return 2;
  cross_face(m, p.ir, p.it, p.ip, iface);
  p.xp = p.xp + dt * w.ux;
  p.yp = p.yp + dt * w.uy;
  p.zp = p.zp + dt * w.uz;
  if (p.ir == m.nrg) return 1;
  if (p.ir == 0) return 2;
  if constexpr (kRayPeel<T>) ray_start<T>(m, w, R, iface);
  return 0;
}

// One crossing of the ray walk: 0 goes on, 1 left the grid toward the
// observer (w.p holds the exit point), 2 ended (tau > taumax, the star, or
// the segment guard).  Whatever axis a lane crosses, it then solves one new
// wall: one square root and two divisions.  The step is written without
// branches on the axis: every lane loads the two table entries of its wall,
// forms the coefficients of all three wall kinds and keeps its own by
// selection, so a wavefront runs one stream of instructions with every
// walking lane active instead of one masked stream per axis.  The arithmetic
// of each lane is that of the wall it crosses.
// kCountSeg = false: the caller counts the walk's segments (w.nseg) itself.
template <class T, bool kCountSeg = true>
__device__ __forceinline__ int peel_ray_step(Ctx<T>& C, PeelWalk<T>& w, PeelRay<T>& R) {
  const DeviceModel& m = C.m;
  Pkt<T>& p = w.p;
  if (w.nseg >= m.max_segments) { C.cnt.add(C_ERRORS); C.cnt.add(C_ERR_PEEL); return 2; }
  ++w.nseg;                     // w.nseg - 1 ray steps walked (the first crossing is one more)
  if constexpr (kCountSeg) C.cnt.add_lanes(C_PEELSEG);
  // 0 radial, 1 lower polar wall, 2 upper polar wall, 3 azimuthal
  int axis = 0;
  G tn = R.tr;
  if (R.tlo < tn) { tn = R.tlo; axis = 1; }
  if (R.thi < tn) { tn = R.thi; axis = 2; }
  if (R.tp < tn) { tn = R.tp; axis = 3; }
  const G dt = tn > R.t ? tn - R.t : 0.0;
  {
    int iop;
    T rho;
    const T opac = cell_opac<T>(m, p, cell_index(m, p.ir, p.it, p.ip), iop, rho);
    w.tsum = w.tsum + opac * (T)dt;
  }
  if (w.tsum > (T)m.taumax) return 2;
  R.t = tn;
  const bool ax0 = axis == 0, ax1 = axis == 1, ax2 = axis == 2, ax3 = axis == 3, pol = ax1 || ax2;
  const G sc = -R.b;
  // the crossing
  const int ir = p.ir + (ax0 ? R.dr : 0);
  if (ax0 && (ir == m.nrg || ir == 0)) {
    p.ir = ir;
    p.xp = R.ox + tn * w.ux;
    p.yp = R.oy + tn * w.uy;
    p.zp = R.oz + tn * w.uz;
    return ir == 0 ? 2 : 1;
  }
  p.ir = ir;
  int ip = p.ip + (ax3 ? R.dp : 0);
  ip = (ip == 0) ? m.npg - 1 : ((ip == m.npg) ? 1 : ip);
  p.ip = ip;
  const int it = p.it + (ax1 ? -1 : (ax2 ? 1 : 0));
  p.it = it;
  // the wall to solve next and its two table entries: the shells r2arr(ir-1),
  // r2arr(ir); the phi plane (aarr, barr) ahead; the polar wall tan2arr
  const int jp = (R.dp < 0) ? ip - 1 : ip;
  const int jt = ax1 ? it - 1 : it;
  const double* e0 = ax0 ? m.r2arr + (ir - 1) : (ax3 ? m.aarr + jp : m.tan2arr + jt);
  const double* e1 = ax0 ? m.r2arr + ir : m.barr + jp;
  const G v0 = *e0, v1 = *e1;
  // shell
  const bool inward = ax0 && tn < sc && v0 > R.h2;
  // cone (tan2 > 0) or the equator plane (tan2 < 0); tan2 = 0 is a pole
  const G t2 = v0;
  const bool cone = pol && t2 > 0.0, eqp = pol && t2 < 0.0;
  const G A = w.ux * w.ux + w.uy * w.uy - t2 * w.uz * w.uz;
  const G B = R.ox * w.ux + R.oy * w.uy - t2 * R.oz * w.uz;
  const G Cq = R.ox * R.ox + R.oy * R.oy - t2 * R.oz * R.oz;
  const bool quad = cone && fabs(A) > 1.0e-10;
  const bool lin = cone && !quad && fabs(B) > 0.5e-10;
  // phi plane A x + B y = 0
  const G pn = -(v0 * R.ox + v1 * R.oy), pd = v0 * w.ux + v1 * w.uy;
  const G sqarg = ax0 ? (inward ? v0 : v1) - R.h2 : (quad ? B * B - A * Cq : 0.0);
  // the shared solve: 1 two quadratic roots, 2 one root n1/d1, 3 none
  G n1 = ax3 ? pn : (eqp ? -R.oz : (lin ? -Cq : 1.0));
  G d1 = ax3 ? pd : (eqp ? w.uz : (lin ? 2.0 * B : 1.0));
  int kind = ((ax3 || eqp) && d1 != 0.0) || lin ? 2 : 3;
  if (quad && sqarg > 0.0) kind = 1;
  const G sq = sqrt(sqarg > 0.0 ? sqarg : 0.0);
  const G q = -(B + copysign(sq, B));
  G n2 = 1.0, d2 = 1.0;
  if (kind == 1) {
    n1 = q; d1 = A;
    n2 = Cq; d2 = (q != 0.0) ? q : A;
  }
  const G x1 = n1 / d1, x2 = n2 / d2;
  // radial exit of the new cell
  G root = inward ? sc - sq : sc + sq;
  int dr = inward ? -1 : 1;
  if (inward && !(root > tn)) {        // rounding at a tangent: leave through the outer wall
    const G d = v1 - R.h2;
    root = sc + sqrt(d > 0.0 ? d : 0.0);
    dr = 1;
  }
  // next crossings of the polar wall just solved
  const G r1 = fmin(x1, x2), r2 = fmax(x1, x2);
  G nx = kRayNone, nx2 = kRayNone;
  if (kind == 1) {
    nx = (r1 > tn) ? r1 : ((r2 > tn) ? r2 : kRayNone);
    nx2 = (r1 > tn) ? r2 : kRayNone;
  } else if (kind == 2) {
    nx = (x1 > tn) ? x1 : kRayNone;
  }
  const G otlo2 = R.tlo2, othi2 = R.thi2;
  R.tr = ax0 ? root : R.tr;
  R.dr = ax0 ? dr : R.dr;
  R.tp = ax3 ? ((kind == 2 && x1 > tn) ? x1 : kRayNone) : R.tp;
  // crossing the lower wall (ax1) keeps its second root as the new upper
  // wall's next crossing; crossing the upper wall (ax2) likewise downward
  R.tlo = ax1 ? nx : (ax2 ? othi2 : R.tlo);
  R.tlo2 = ax1 ? nx2 : (ax2 ? kRayNone : R.tlo2);
  R.thi = ax2 ? nx : (ax1 ? otlo2 : R.thi);
  R.thi2 = ax2 ? nx2 : (ax1 ? kRayNone : R.thi2);
  return 0;
}

// :192-240.  The Stokes vector a walk brings to the observer (exponential
// weight, then stokespeel for scattered light) and where it lands.
// q: the packet at the event; s0: its (weighted) Stokes I; (xe, ye, ze):
// where the walk left the grid; tsum: its optical depth.
template <class T>
struct PeelHit {
  T s[4];
  G ximage, yimage;
  int ix, iy;
  T freq;
};

template <class T>
__device__ __forceinline__ void peel_weigh(const DeviceModel& m, const Pkt<T>& q, T s0, int k, T gsc, T plsc,
                                           G xe, G ye, G ze, T tsum, PeelHit<T>& h) {
  const double* pe = m.peel + 5 * k;
  const T coste = (T)pe[0], sinte = (T)pe[1], cospe = (T)pe[2], sinpe = (T)pe[3], phie = (T)pe[4];
  h.s[0] = s0; h.s[1] = q.sqp; h.s[2] = q.sup; h.s[3] = q.svp;
  T factor = RT<T>::exp_(-tsum);
  h.s[0] = h.s[0] * factor; h.s[1] = h.s[1] * factor; h.s[2] = h.s[2] * factor; h.s[3] = h.s[3] * factor;
  if (q.iflag == 1) {
    T bmu, sinbm, phinew, ri1;
    findangle3<T>(q.cost, q.sint, q.phi, q.cosp, q.sinp, coste, sinte, phie, cospe, sinpe, bmu, sinbm, phinew, ri1);
    stokespeel<T>(gsc, plsc, h.s, coste, sinte, ri1, bmu, q.cost, q.sint);
  }
  h.yimage = ze * pe[1] - ye * pe[0] * pe[3] - xe * pe[0] * pe[2];
  h.ximage = ye * pe[2] - xe * pe[3];
  G rmaxi = m.rmaxi;
  G fractxh = 2.0 * rmaxi / (G)m.nx;
  h.ix = (int)((h.ximage + rmaxi) / fractxh) + 1;
  h.iy = (int)((h.yimage + rmaxi) / fractxh) + 1;
  h.freq = (T)2.9979e14 / q.wave;
}

// peeloff_3d_2_test.f90:245-260: the filter weight of band b at the walk's
// frequency; false when the frequency is outside the filter
template <class T>
__device__ __forceinline__ bool band_weight(const DeviceModel& m, int b, T freq, T& wt) {
  const double* fw = m.band_freq + (size_t)b * m.nwmax;
  const double* fp = m.band_phi + (size_t)b * m.nwmax;
  int nf = m.band_nw[b] - 1;
  if (!(freq > (T)fw[0] && freq < (T)fw[nf - 1])) return false;
  int iw = 1;
  while (freq > (T)fw[iw - 1]) iw++;
  iw = iw - 1;
  wt = (T)fp[iw - 1];
  return true;
}

// :241-347: image and peeled-SED binning.  sed_k != null: the peeled SED
// of direction k goes to that slice ([9][8][nap][nfreq], the layout of
// t.psed without the direction; in LDS) instead of t.psed.
template <class T>
__device__ __forceinline__ void peel_tally(Ctx<T>& C, const Pkt<T>& q, T s0, int k, T gsc, T plsc,
                                           G xe, G ye, G ze, T tsum, float* sed_k = nullptr) {
  const DeviceModel& m = C.m;
  const DeviceTallies& t = C.t;
  C.cnt.add(C_PEELEVT);
  PeelHit<T> h;
  peel_weigh<T>(m, q, s0, k, gsc, plsc, xe, ye, ze, tsum, h);
  const T* s = h.s;
  const int ix = h.ix, iy = h.iy;
  if (m.nband > 0 && ix <= m.nx && iy <= m.nx && ix > 0 && iy > 0) {
    size_t npix = (size_t)m.npeel * m.nx * m.nx;
    size_t plane = (size_t)m.nband * npix;
    for (int b = 0; b < m.nband; ++b) {
      T wt;
      if (band_weight<T>(m, b, h.freq, wt)) {
        size_t o = (size_t)b * npix + (size_t)k + (size_t)m.npeel * ((size_t)(ix - 1) + (size_t)m.nx * (size_t)(iy - 1));
        atomicAdd(&t.image[0 * plane + o], (float)(s[0] * wt));
        atomicAdd(&t.image[1 * plane + o], (float)(s[1] * wt));
        atomicAdd(&t.image[2 * plane + o], (float)(s[2] * wt));
        atomicAdd(&t.image[3 * plane + o], (float)(s[3] * wt));
        int pl = (q.i_orig == 1) ? 5 : (q.i_orig == 2) ? 4 : (q.i_orig == 3) ? 6 : (q.i_orig == 4) ? 7 : -1;
        if (pl >= 0) atomicAdd(&t.image[(size_t)pl * plane + o], (float)(s[0] * wt));
      }
    }
  }
  G rho2 = h.ximage * h.ximage + h.yimage * h.yimage;
  int inub = freq_bin<T>(m, q.nub);
  int i_inter = (q.iflag == 1) ? 2 : ((q.i_orig == 1) ? 1 : 3);
  int indx[3] = {1, 1 + q.i_orig, 5 + i_inter};
  if (sed_k) {
    const int ns = m.nfreq * m.nap * kNo;
    for (int ia = 1; ia <= m.nap; ++ia) {
      if (rho2 < m.aperture2[ia - 1]) {
        for (int j = 0; j < 3; ++j) {
          float* o = sed_k + (inub - 1) + m.nfreq * ((ia - 1) + m.nap * (indx[j] - 1));
          atomicAdd(o + 0 * ns, (float)s[0]);
          atomicAdd(o + 1 * ns, (float)s[1]);
          atomicAdd(o + 2 * ns, (float)s[2]);
          atomicAdd(o + 3 * ns, (float)s[3]);
          atomicAdd(o + 4 * ns, (float)(s[0] * s[0]));
          atomicAdd(o + 5 * ns, (float)(s[1] * s[1]));
          atomicAdd(o + 6 * ns, (float)(s[2] * s[2]));
          atomicAdd(o + 7 * ns, (float)(s[3] * s[3]));
          atomicAdd(o + 8 * ns, 1.0f);
        }
      }
    }
    return;
  }
  size_t nb = (size_t)m.nfreq * m.npeel * m.nap * kNo;
  for (int ia = 1; ia <= m.nap; ++ia) {
    if (rho2 < m.aperture2[ia - 1]) {
      for (int j = 0; j < 3; ++j) {
        size_t o = (size_t)(inub - 1) + (size_t)m.nfreq * ((size_t)k + (size_t)m.npeel * ((size_t)(ia - 1) + (size_t)m.nap * (size_t)(indx[j] - 1)));
        atomicAdd(&t.psed[0 * nb + o], (float)s[0]);
        atomicAdd(&t.psed[1 * nb + o], (float)s[1]);
        atomicAdd(&t.psed[2 * nb + o], (float)s[2]);
        atomicAdd(&t.psed[3 * nb + o], (float)s[3]);
        atomicAdd(&t.psed[4 * nb + o], (float)(s[0] * s[0]));
        atomicAdd(&t.psed[5 * nb + o], (float)(s[1] * s[1]));
        atomicAdd(&t.psed[6 * nb + o], (float)(s[2] * s[2]));
        atomicAdd(&t.psed[7 * nb + o], (float)(s[3] * s[3]));
        atomicAdd(&t.psed[8 * nb + o], 1.0f);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Peel-off events: the transport kernel queues one record per event; the
// peel kernel walks every record in every observer direction.
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void make_rec(PeelRec<T>& r, const Pkt<T>& p, int sp, int iop, bool direct, T coslp, T sinlp) {
  r.xp = p.xp; r.yp = p.yp; r.zp = p.zp; r.rsq = p.rsq; r.rtot = p.rtot;
  r.sip = p.sip; r.sqp = p.sqp; r.sup = p.sup; r.svp = p.svp;
  if (direct) { r.cost = p.cosb; r.sint = p.sinb; r.phi = (T)0; r.cosp = coslp; r.sinp = sinlp; }
  else { r.cost = p.cost; r.sint = p.sint; r.phi = p.phi; r.cosp = p.cosp; r.sinp = p.sinp; }
  r.nub = p.nub; r.wave = p.wave;
  const Opac<T> o = pick(p.op, iop);   // iop = the opacity table of species sp
  r.g = o.g; r.pl = o.pl;
#pragma unroll
  for (int i = 0; i < kMaxOpac; ++i) r.kapd[i] = p.kapd[i];
  r.cell = (uint32_t)p.ir | ((uint32_t)p.it << 10) | ((uint32_t)p.ip << 20) | ((uint32_t)(p.on_wall ? 1 : 0) << 30);
  r.flags = (uint32_t)p.iflag | ((uint32_t)p.i_orig << 1) | ((uint32_t)(direct ? 1 : 0) << 4) | ((uint32_t)sp << 8);
}

template <class T>
__device__ __forceinline__ void write_rec(Ctx<T>& C, const Pkt<T>& p, unsigned slot, int sp, int iop, bool direct,
                                          T coslp, T sinlp) {
  PeelRec<T> r;
  make_rec<T>(r, p, sp, iop, direct, coslp, sinlp);
  ((PeelRec<T>*)C.q.recs)[slot] = r;
}

// scattering / re-emission peel (propagate_5_fuv.f90:83-90, :162-170)
template <class T, bool kAdj = false>
__device__ __forceinline__ void peel_event(Ctx<T>& C, const Pkt<T>& p, int sp, int iop) {
  if constexpr (kAdj) { adj_peel_event<T>(C, p, sp, iop, false, (T)0, (T)0); return; }
  if (C.q.capacity == 0) return;          // no observer directions
  const unsigned slot = take_slot(C);
  if (slot < C.q.capacity) { if (writer(C)) write_rec<T>(C, p, slot, sp, iop, false, (T)0, (T)0); }
  else { C.cnt.add(C_ERRORS); C.cnt.add(C_ERR_OTHER); }   // the host sizes the queue so this cannot happen
}

// direct stellar peel, weight 4 I cos(norm) (newdisktherm_5_fuv.f90:371-414)
template <class T, bool kAdj = false>
__device__ __forceinline__ void peel_direct(Ctx<T>& C, const Pkt<T>& p, T coslp, T sinlp) {
  if constexpr (kAdj) { adj_peel_event<T>(C, p, 0, 0, true, coslp, sinlp); return; }
  if (C.q.capacity == 0) return;
  const unsigned slot = take_slot(C);
  if (slot < C.q.capacity) { if (writer(C)) write_rec<T>(C, p, slot, 0, 0, true, coslp, sinlp); }
  else { C.cnt.add(C_ERRORS); C.cnt.add(C_ERR_OTHER); }
}

// One trial of stokes.f90:75-136: the candidate scattering angle bmu and
// rotation ri1 from the trial's first two draws, and its weight rprob for
// the incoming Stokes vector.
template <class T>
__device__ __forceinline__ void stokes_trial(T g, T pl, const Pkt<T>& p, T u1, T u2, T& bmu, T& ri1, T& rprob) {
  const T PI = (T)kPi, R2P = (T)kR2p, one = (T)1, two = (T)2;
  T pp[4];
  bmu = one - two * u1;
  T cosb2 = bmu * bmu;
  dustmat<T>(g, pl, bmu, cosb2, pp);
  ri1 = R2P * u2;
  if (ri1 > PI) {
    T ri3 = R2P - ri1, cosi3 = RT<T>::cos_(ri3), sini3 = RT<T>::sin_(ri3);
    T sin2i3 = two * sini3 * cosi3, cos2i3 = two * (cosi3 * cosi3) - one;
    rprob = (pp[0] * p.sip + pp[1] * cos2i3 * p.sqp + pp[1] * sin2i3 * p.sup) / p.sip;
  } else {
    T cosi1 = RT<T>::cos_(ri1), sini1 = RT<T>::sin_(ri1);
    T sin2i1 = two * sini1 * cosi1, cos2i1 = two * (cosi1 * cosi1) - one;
    rprob = (pp[0] * p.sip + pp[1] * cos2i1 * p.sqp + (-pp[1] * sin2i1) * p.sup) / p.sip;
  }
}

// The rejection loop of stokes.f90:75-136 with the 64 lanes of a wavefront
// that all carry the same packet (k_tail): lane j tries trial done + j on
// its own draws (3 per trial, found by seeking the packet's counter-based
// stream), the running peak of the serial loop is the prefix maximum of the
// trials' weights, and the first accepted trial is the serial loop's.  The
// stream then continues after that trial's draws, as after the serial loop.
template <class T>
__device__ __forceinline__ void stokes_coop(Ctx<T>& C, T g, T pl, const Pkt<T>& p, T peak, T& bmu, T& ri1) {
  const int lane = __lane_id();
  const uint32_t n0 = C.rng.n;
  for (uint32_t done = 0;; done += 64u) {
    // this pass's 192 draws from the lane-held blocks (refilled when they
    // do not hold them all), each lane reading its trial's three words
    const uint32_t pa = n0 + 3u * (done + (uint32_t)lane);
    T u1, u2, xran;
    if (C.rng.coop) {
      const uint32_t lo = (n0 + 3u * done) >> 2, hi = (n0 + 3u * (done + 64u) - 1u) >> 2;
      if (lo - C.rng.cb >= 64u || hi - C.rng.cb >= 64u) rng_fill(C.rng, lo);
      u1 = u01<T>(rng_word_at(C.rng, pa));
      u2 = u01<T>(rng_word_at(C.rng, pa + 1u));
      xran = u01<T>(rng_word_at(C.rng, pa + 2u));
    } else {
      Rng r = C.rng;
      rng_seek(r, pa);
      u1 = uniform<T>(r);
      u2 = uniform<T>(r);
      xran = uniform<T>(r);
    }
    T b, ri, rprob;
    stokes_trial<T>(g, pl, p, u1, u2, b, ri, rprob);
    T pm = rprob;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const T v = __shfl_up(pm, off, 64);
      if (lane >= off && v > pm) pm = v;
    }
    const T pk = (pm > peak) ? pm : peak;
    const unsigned long long acc = __ballot(!(pk * xran > rprob));
    if (acc) {
      const int t = __ffsll((long long)acc) - 1;
      bmu = __shfl(b, t, 64);
      ri1 = __shfl(ri, t, 64);
      C.cnt.add(C_HTOT, (unsigned long long)(done + (uint32_t)t + 1u));
      C.cnt.add(C_HIT);
      rng_seek(C.rng, n0 + 3u * (done + (uint32_t)t + 1u));
      return;
    }
    peak = __shfl(pk, 63, 64);
  }
}

// stokes.f90:46-137 rejection sampling, then the Mueller update
template <class T>
__device__ __forceinline__ void stokes(Ctx<T>& C, Pkt<T>& p, int iop) {
  const Opac<T> o = pick(p.op, iop);
  T g = o.g, pl = o.pl;
  T peak = cs_peak<T>(g), bmu, ri1, rprob;
  if (C.coop) {
    stokes_coop<T>(C, g, pl, p, peak, bmu, ri1);
  } else {
    for (;;) {
      C.cnt.add_lanes(C_HTOT);
      const T u1 = uniform<T>(C.rng);
      const T u2 = uniform<T>(C.rng);
      stokes_trial<T>(g, pl, p, u1, u2, bmu, ri1, rprob);
      if (rprob > peak) peak = rprob;
      T xran = uniform<T>(C.rng);
      if (peak * xran > rprob) continue;
      C.cnt.add_lanes(C_HIT);
      break;
    }
  }
  T dir[5] = {p.cost, p.sint, p.phi, p.cosp, p.sinp};
  T s[4] = {p.sip, p.sqp, p.sup, p.svp};
  stokes_apply<T>(g, pl, bmu, ri1, dir, s);
  p.cost = dir[0]; p.sint = dir[1]; p.phi = dir[2]; p.cosp = dir[3]; p.sinp = dir[4];
  p.sip = s[0]; p.sqp = s[1]; p.sup = s[2]; p.svp = s[3];
}

// After a flight that stopped inside the grid: propagate_5_fuv.f90:199-229
// (after the first flight, :52-60, where iint is still 0).  0 = the packet
// waits at the loop top for its next interaction, 2 = killed.
template <class T>
__device__ __forceinline__ int flight_end(Ctx<T>& C, Life<T>& L) {
  const DeviceModel& m = C.m;
  Pkt<T>& p = L.p;
  if (!L.first) {
    if (p.ir != L.pir || p.it != L.pit || p.ip != L.pip) L.iint = 0;
    if (L.iint > m.iintmax) {
      if (writer(C)) atomicAdd(&C.t.killed[cell_index(m, p.ir, p.it, p.ip)], 1);
      C.cnt.add(C_ABFLUX);
      return 2;
    }
  }
  (void)uniform<T>(C.rng);   // select_dust_2.f90:134 draw
  L.idust2 = cell_species<T>(m, L, cell_index(m, p.ir, p.it, p.ip));
  if (L.first) C.cnt.add(C_TOT);
  L.first = false;
  return 0;
}

// A peel-off walk from p crosses no wall before wall_clearance, so its
// optical depth exceeds taumax inside the first segment once opac times the
// clearance does (with a margin for rounding): the walk ends there with
// nothing for the observer (peeloff_3d_2_test.f90:175), and its record is
// not queued.  Deep inside optically thick cells, where the long-lived
// packets spend their interactions, that is every record.
template <class T>
__device__ __forceinline__ bool peel_hidden(const DeviceModel& m, Life<T>& L) {
  const Pkt<T>& p = L.p;
  if (p.on_wall) return false;
  int iop;
  T rho;
  const T opac = cell_opac_l<T>(m, L, cell_index(m, p.ir, p.it, p.ip), iop, rho);
  L.wclear = wall_clearance<T>(m, p);
  L.wclear_ok = true;
  return (double)opac * L.wclear > 1.001 * m.taumax;
}

// propagate_5_fuv.f90:62-198: one interaction (scatter, or absorb and
// re-emit), its peel-off record, and the start of the next flight.
// reemit.f90:17-34 with lucydustfreq_2_test.f90:34-45: isotropic and
// unpolarized, at the cell's dust temperature (a Planck frequency at or
// above Tmax)
template <class T>
__device__ __forceinline__ void reemit(Ctx<T>& C, Pkt<T>& p, T tave, int iop) {
  const DeviceModel& m = C.m;
  p.sip = (T)1; p.sqp = (T)0; p.sup = (T)0; p.svp = (T)0;
  p.cost = (T)1 - (T)2 * uniform<T>(C.rng);
  p.sint = RT<T>::sqrt_((T)1 - p.cost * p.cost);
  p.phi = (T)kR2p * uniform<T>(C.rng);
  p.cosp = RT<T>::cos_(p.phi);
  p.sinp = RT<T>::sin_(p.phi);
  if (tave >= (T)(kTmaxK / 11605.0)) p.nub = bbfreq<T>(C.rng, tave, C.inczeta);
  else {
    T xi = uniform<T>(C.rng);
    p.nub = C.coop ? lucydustfreq_coop<T>(m, tave, iop, xi, C.cdf) : lucydustfreq_xi<T>(m, tave, iop, xi);
  }
  set_dir<T>(p);
}

// A photon absorbed in a diffusion cell (emit_common_2.f90:226-378, theta
// diffusion).  It is moved at fixed radius and azimuth to just outside the
// wall between the diffusion region and the free medium above (upper
// half) or below it (lower half), emitted into the outer hemisphere with
// a cosine law (isotrp.f90), and its frequency drawn from a blackbody at
// the surface temperature.  That temperature grows with every photon the
// surface cell re-emits: the reference adds 4 pi T*^4 / (ns dA) to T^4
// per photon in the order the photons arrive; here the arrival order on
// the GPU is an atomic counter per surface cell, the k-th photon seeing
// T0^4 + (k+1) Delta / dA.  The accretion photons absorbed at the midplane
// add 0.75 mu T4eff (grey atmosphere).  The Stokes vector is left as the
// absorbed photon had it, as the reference leaves it.  The packet ends in
// the free cell next to the surface.
template <class T>
__device__ __forceinline__ void diffuse(Ctx<T>& C, Pkt<T>& p) {
  const DeviceModel& m = C.m;
  const int ntg = m.ntg;
  const int am = m.diff_itmin[p.ir - 1];
  const bool upper = p.it <= (ntg - 1) / 2;
  const int its = upper ? am : ntg - am;            // the surface diffusion cell
  const int w = upper ? its : its + 1;               // its wall toward the free medium (1-based)
  const double cw = m.costarr[w - 1], sw = sin_nl(m.thetarr[w - 1]);
  // move (emit_common_2.f90:230-237)
  const double cosb = p.zp / p.rtot;
  const double sinb = sqrt(1.0 - cosb * cosb);
  const double cosbnew = 1.00001 * cw;
  const double sinbnew = sqrt(1.0 - cosbnew * cosbnew);
  p.xp = p.xp * sinbnew / sinb;
  p.yp = p.yp * sinbnew / sinb;
  p.zp = p.rtot * cosbnew;
  // outward normal of the surface (:240-249) and a cosine-law direction about it (isotrp.f90)
  double sux, suy, suz;
  if (cw > 0.0) { sux = -cw * p.xp / sw / p.rtot; suy = -cw * p.yp / sw / p.rtot; suz = sw; }
  else { sux = cw * p.xp / sw / p.rtot; suy = cw * p.yp / sw / p.rtot; suz = -sw; }
  double x = 0.0, y = 0.0, z = 0.0, r2 = 2.0, mur = 0.0;
  while (r2 * r2 > fabs(mur)) {
    x = (double)uniform<T>(C.rng); x = x + x - 1.0;
    y = (double)uniform<T>(C.rng); y = y + y - 1.0;
    z = (double)uniform<T>(C.rng); z = z + z - 1.0;
    mur = x * sux + y * suy + z * suz;
    r2 = x * x + y * y + z * z;
  }
  double r = sqrt(r2);
  if (mur < 0.0) r = -r;
  const double ux = x / r, uy = y / r, uz = z / r;
  p.cost = (T)uz;
  const double sint = sqrt(1.0 - uz * uz);
  p.sint = (T)sint;
  p.cosp = (T)(ux / sint);
  p.sinp = (T)(uy / sint);
  p.phi = (T)atan2_nl(uy / sint, ux / sint);
  // surface temperature after this photon (:258-270) and the emission temperature (:343-361)
  const double dA = kPi * sw * (m.r2arr[p.ir] - m.r2arr[p.ir - 1]);
  const size_t cs = cell_index(m, p.ir, its, p.ip);
  unsigned k = 0;
  if (writer(C)) k = atomicAdd(&C.t.diff_count[cs], 1u);
  if (C.coop) k = __shfl(k, 0, 64);
  const double Tcap = kTmaxK / 11605.0, T0 = 0.1 / 11605.0;
  double Ts = quarter_power(T0 * T0 * T0 * T0 + (double)(k + 1u) * m.diff_delta / dA);
  if (Ts > Tcap) Ts = Tcap;
  const int mid = upper ? (ntg - 1) / 2 : (ntg + 1) / 2;
  double nacc = (double)*(volatile const int*)&C.t.nabs[cell_index(m, p.ir, mid, p.ip)];
  if (C.coop) nacc = __shfl(nacc, 0, 64);
  const double T4eff = m.diff_delta * nacc / dA;
  const double mu = ux * sux + uy * suy + uz * suz;
  double Te = quarter_power(Ts * Ts * Ts * Ts + 0.75 * mu * T4eff);
  if (Te > Tcap) Te = Tcap;
  p.nub = bbfreq<T>(C.rng, (T)Te, C.inczeta);
  p.it = upper ? its - 1 : its + 1;
  set_dir<T>(p);
}

template <class T, bool kAdj = false, bool kComb = false>
__device__ __forceinline__ void interact(Ctx<T>& C, Life<T>& L) {
  const DeviceModel& m = C.m;
  const DeviceTallies& t = C.t;
  Pkt<T>& p = L.p;
  size_t c = cell_index(m, p.ir, p.it, p.ip);
  // idust2 is the species of this cell (flight_end), so its opacity table is cached
  int iop = (L.c_sp == (uint32_t)c) ? L.spop : m.species_opac[L.idust2];
  // an accretion photon's absorption where it was emitted (emit; no albedo
  // draw, initpacc_4.f90:75-100, newdisktherm_5_fuv.f90:247-283)
  const bool start = (p.i_orig == 0);
  if (!start) L.nint++;
  if (!start && uniform<T>(C.rng) <= pick(p.op, iop).albedo) {
    p.iflag = 1; L.iscat++; L.iint++;
    if (m.ipeel == 1 && (kAdj || !peel_hidden<T>(m, L))) peel_event<T, kAdj>(C, p, L.idust2, iop);
    stokes<T>(C, p, iop);
  } else {
    p.i_orig = start ? 2 : i_orig_of(L.idust2);
    // propagate_5_fuv.f90:139: a diffusion cell counts only the accretion
    // photons placed in it (initpacc_4.f90); its absorptions re-emit from
    // the diffusion surface (emit_common_2.f90:88)
    const bool diff = diffusing(m, p.ir, p.it);
    tally_absorb<kComb>(C, L, c, start || !diff);
    // the surface temperature reads the midplane's nabs: this lane's own
    // counts are in it first
    if constexpr (kComb) { if (diff) tally_flush<T>(C, L); }
    if (diff) diffuse<T>(C, p);
    else reemit<T>(C, p, cell_tdust<T>(m, L, c), iop);
    p.iflag = 0;
    opacset<T>(m, p.nub, p.op, C.coop);
    p.wave = (T)1.2398 / p.nub;
    set_kapd(p, m, start ? 0 : 1);
    if (m.ipeel == 1 && (kAdj || !peel_hidden<T>(m, L))) peel_event<T, kAdj>(C, p, L.idust2, iop);
    if (!start) L.iint++;
  }
  T tau = RT<T>::tau_of(uniform<T>(C.rng));
  L.xpold = p.xp; L.ypold = p.yp; L.zpold = p.zp;
  set_dir<T>(p);
  if (start) L.first = true;
  flight_begin(L, tau);
}

// SED binning of an escaped packet, propagate_5_fuv.f90:231-320
template <class T>
__device__ __forceinline__ void prop_end(Ctx<T>& C, Life<T>& L) {
  const DeviceModel& m = C.m;
  const DeviceTallies& t = C.t;
  const T R2P = (T)kR2p;
  Pkt<T>& p = L.p;
  int itb = (int)((T)m.nmu * ((T)1 - p.cost) / (T)2) + 1;
  if (p.phi < (T)0) p.phi = p.phi + R2P;
  int ipb = (int)((T)m.nph * p.phi / R2P) + 1;
  if (itb < 1) itb = 1;
  if (itb > m.nmu) itb = m.nmu;
  if (ipb < 1) ipb = 1;
  if (ipb > m.nph) ipb = m.nph;
  G x = L.ypold * (G)p.cosp - L.xpold * (G)p.sinp;
  G y = L.zpold * m.sini_mu[itb - 1] - L.ypold * m.u_mu[itb - 1] * (G)p.sinp - L.xpold * m.u_mu[itb - 1] * (G)p.cosp;
  G rho2 = (x * x + y * y);
  if (rho2 < m.aperture2[m.nap - 1] * 1.0001) {
    C.cnt.add(C_FLUX);
    if (p.iflag == 1) { C.cnt.add(C_SFLUX); C.cnt.add(C_NSCAT, (unsigned long long)L.iscat); }
  }
  int inub = freq_bin<T>(m, p.nub);
  int i_inter = (p.iflag == 1) ? 2 : ((p.i_orig == 1) ? 1 : 3);
  int indx[3] = {1, 1 + p.i_orig, 5 + i_inter};
  size_t nb = (size_t)m.nfreq * m.nmu * m.nph * m.nap * kNo;
  for (int ia = 1; ia <= m.nap && writer(C); ++ia) {
    if (rho2 < m.aperture2[ia - 1]) {
      for (int j = 0; j < 3; ++j) {
        size_t o = (size_t)(inub - 1) + (size_t)m.nfreq * ((size_t)(itb - 1) + (size_t)m.nmu * ((size_t)(ipb - 1) + (size_t)m.nph * ((size_t)(ia - 1) + (size_t)m.nap * (size_t)(indx[j] - 1))));
        atomicAdd(&t.sed[0 * nb + o], (float)p.sip);
        atomicAdd(&t.sed[1 * nb + o], (float)p.sqp);
        atomicAdd(&t.sed[2 * nb + o], (float)p.sup);
        atomicAdd(&t.sed[3 * nb + o], (float)p.svp);
        atomicAdd(&t.sed[4 * nb + o], (float)(p.sip * p.sip));
        atomicAdd(&t.sed[5 * nb + o], (float)(p.sqp * p.sqp));
        atomicAdd(&t.sed[6 * nb + o], (float)(p.sup * p.sup));
        atomicAdd(&t.sed[7 * nb + o], (float)(p.svp * p.svp));
        atomicAdd(&t.sed[8 * nb + o], 1.0f);
        atomicAdd(&t.sed[9 * nb + o], (float)RT<T>::fabs_(p.cost));
      }
    }
  }
  if (p.iflag == 1) C.cnt.add(C_SCOUNT);
}

// rep_isrf.f90:13-29: ISRF frequency (eV) from the tabulated CDF.  FP64
// follows the reference's log10/pow; the mixed path interpolates the same
// table in FP32.
template <class T>
__device__ __forceinline__ T rep_isrf(Ctx<T>& C) {
  const DeviceModel& m = C.m;
  const T xran = uniform<T>(C.rng);
  int ifreq = locate<double>(m.isrf_cdf, m.nisrf, (double)xran);
  if (ifreq == m.nisrf) ifreq = ifreq - 1;
  if (ifreq == 0) ifreq = 1;
  const T lf0 = (T)m.isrf_logfreq[ifreq - 1], lf1 = (T)m.isrf_logfreq[ifreq];
  const T lc0 = (T)m.isrf_logcdf[ifreq - 1], lc1 = (T)m.isrf_logcdf[ifreq];
  if constexpr (std::is_same<T, double>::value) {
    const double logfreq = lf0 + (log10(xran) - lc0) / (lc1 - lc0) * (lf1 - lf0);
    return pow(10.0, logfreq);
  } else {
    const float ln10 = 2.302585093f;
    const float logfreq = lf0 + (RT<T>::log_acc(xran) / ln10 - lc0) / (lc1 - lc0) * (lf1 - lf0);
    return RT<T>::exp_(ln10 * logfreq);
  }
}

// initpacc_4.f90:25-100 with emit_common_2.f90:37-85 (a cell outside the
// diffusion region) and reemit.f90: a disk-accretion photon.  The radial
// cell follows the cumulative accretion luminosity (locate on Ldint: the
// cell j with Ldint(j) <= xi < Ldint(j+1)); the point lies at that cell's
// mid radius, a uniform azimuth (rantheta) and a Gaussian height of the
// cell's scale height (gasdev, its second deviate kept for the packet's next
// draw); a point in an empty cell is drawn again at the same radius (as is
// one beyond the outer wall, whose cell the reference would read past the
// grid).  The cell's dust absorbs the photon (dbig, nabs) and re-emits it
// at the cell's temperature, peeled with origin 2: emit loads the placement
// and interact does the rest (i_orig = 0 marks the photon).  After
// kAccTries draws without a cell holding dust the packet is dropped as an
// error (sp = -1).
template <class T>
__device__ __forceinline__ AccStart acc_place(const DeviceModel& m, Rng& rng) {
  AccStart a{};
  a.sp = -1;
  const double xi = (double)uniform<T>(rng);
  int j = locate<double>(m.acc_ldint, m.nrg - 1, xi);   // Ldint(1) = 0 and Ldint(nrg-1) = 1: 1 <= j <= nrg-2
  j = j < 1 ? 1 : (j > m.nrg - 1 ? m.nrg - 1 : j);
  const G rp = 0.5 * (m.rarr[j - 1] + m.rarr[j]);
  const G hd = m.acc_hdisk[j - 1];
  T gset = (T)0;
  bool iset = false;
  for (int tries = 0; tries < kAccTries; ++tries) {
    T x, y, r2;
    do {
      x = uniform<T>(rng); x = x + x - (T)1;
      y = uniform<T>(rng); y = y + y - (T)1;
      r2 = x * x + y * y;
    } while (r2 > (T)1);
    const T r = RT<T>::sqrt_(r2);
    const T sp = y / r, cp = x / r;
    T g;
    if (!iset) {
      T v1, v2, rsq;
      do {
        v1 = (T)2 * uniform<T>(rng) - (T)1;
        v2 = (T)2 * uniform<T>(rng) - (T)1;
        rsq = v1 * v1 + v2 * v2;
      } while (!(rsq < (T)1 && rsq > (T)0));
      const T fac = RT<T>::sqrt_((T)-2 * RT<T>::log_acc(rsq) / rsq);
      gset = v1 * fac;
      g = v2 * fac;
      iset = true;
    } else {
      g = gset;
      iset = false;
    }
    a.xp = rp * (G)cp;
    a.yp = rp * (G)sp;
    a.zp = hd * (G)g;
    a.rsq = a.xp * a.xp + a.yp * a.yp + a.zp * a.zp;
    const G rtot = sqrt(a.rsq);
    a.ir = locate<double>(m.r2arr, m.nrg, a.rsq);
    if (a.ir < 1 || a.ir > m.nrg - 1) continue;
    const int it = locate<double>(m.costarr, m.ntg, a.zp / rtot);
    G lp = atan2(a.yp, a.xp);
    if (lp < 0.0) lp = lp + kR2p;
    const int ip = locate<double>(m.phiarr, m.npg, lp);
    a.it = it < 1 ? 1 : (it > m.ntg - 1 ? m.ntg - 1 : it);
    a.ip = ip < 1 ? 1 : (ip > m.npg - 1 ? m.npg - 1 : ip);
    const size_t c = cell_index(m, a.ir, a.it, a.ip);
    if (cell_rho<T>(m, c) > (T)0) {
      (void)uniform<T>(rng);   // select_dust_2.f90:134 draw
      a.sp = m.species[c];
      break;
    }
  }
  a.n = rng.n;
  return a;
}

// k_acc_place: the placement of a run's accretion photons k < acc_to, each
// on its own packet stream, before the transport kernel starts them.  It is
// kept out of k_transport, whose occupancy its registers would halve.
template <class T>
__global__ void __launch_bounds__(256) k_acc_place(DeviceModel m, uint64_t seed, uint64_t first_packet, AccStart* out) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m.acc_to) return;
  Rng rng;
  rng_init(rng, seed, first_packet + k);
  out[k] = acc_place<T>(m, rng);
}

// newdisktherm_5_fuv.f90:294-413: one stellar photon (ips = 0, no spots),
// its direct peel, and the start of its first flight.  With external
// illumination the last packets of a run are photons of the external
// field instead (newdisktherm_5_fuv.f90:416-455, initpout.f90:18-57):
// entering at 0.999999 rmax, directed inward in the local frame, ISRF
// frequency, not peeled (the reference subtracts the sky).  Both share
// the transform, opacity, position and cell lookup.  Disk-accretion
// photons (newdisktherm_5_fuv.f90:242-283) come first in a run, then the
// hot-spot photons (:296-313).  An accretion photon, placed by k_acc_place,
// waits for interact (its absorption where it was emitted): returns 2; 1
// when the first flight has begun; 0 when the packet was dropped.
template <class T, bool kAdj = false>
__device__ __forceinline__ int emit(Ctx<T>& C, Life<T>& L) {
  const DeviceModel& m = C.m;
  Pkt<T>& p = L.p;
  const bool acc = L.pid < m.acc_to;
  const bool outside = !acc && m.nisrf > 0 && L.pid >= m.isrf_from;
  const bool hot = m.ispot != 0 && !outside && !acc && L.pid < m.hot_to;
  bool spot = hot;
  p.on_wall = false;
  p.sip = (T)1; p.sqp = (T)0; p.sup = (T)0; p.svp = (T)0;
  cell_cache_reset(L);
  L.iscat = 0; L.iint = 0; L.nint = 0;
  if (acc) {
    const AccStart a = m.acc_start[L.pid];
    if (a.sp < 0) {
      p.nub = (T)0; p.cost = (T)0; p.phi = (T)0;
      return 0;
    }
    p.xp = a.xp; p.yp = a.yp; p.zp = a.zp; p.rsq = a.rsq;
    p.rtot = sqrt(p.rsq);
    p.ir = a.ir; p.it = a.it; p.ip = a.ip;
    L.idust2 = a.sp;
    rng_seek(C.rng, a.n);
    p.nub = (T)m.nudust[0];   // a table frequency until the re-emission: unpark sets p.op from it
    p.i_orig = 0;
    p.iflag = 0;
    p.exitflag = 0;
    return 2;
  }
  if (!outside) {
    // initpspot.f90:27-125: a point on the star; a hot-spot photon redraws
    // until the point lies in a spot, and with fspot < 1 a stellar photon
    // emitted inside a spot takes the spot's temperature
    for (;;) {
      for (;;) {
        p.cosb = (T)1 - (T)2 * uniform<T>(C.rng);
        p.sinb = RT<T>::sqrt_((T)1 - p.cosb * p.cosb);
        if (p.cosb != (T)0 && RT<T>::fabs_(p.cosb) < (T)1) break;
      }
      p.lp = (T)kR2p * uniform<T>(C.rng);
      if (!hot) break;
      const T cphi = RT<T>::cos_(p.lp), sphi = RT<T>::sin_(p.lp);
      const T d1 = (T)m.scspot1 * p.sinb * cphi + (T)m.ssspot1 * p.sinb * sphi + (T)m.cspot1 * p.cosb;
      const T d2 = (T)m.scspot2 * p.sinb * cphi + (T)m.ssspot2 * p.sinb * sphi + (T)m.cspot2 * p.cosb;
      const bool out1 = d1 < (T)m.csprad;
      if (m.nspot == 1 ? !out1 : !(out1 && d2 < (T)m.csprad)) break;
    }
    if (m.ispot != 0 && m.fspot < 1.0) {
      const T pi = (T)(0.5 * kR2p);
      T dellon = (p.lp < pi) ? p.lp : (T)kR2p - p.lp;
      if ((T)m.cspot1 * p.cosb + (T)m.sspot1 * p.sinb * RT<T>::cos_(dellon) > (T)m.csprad) spot = true;
      if (m.nspot == 2) {
        dellon = (p.lp < pi) ? pi - p.lp : p.lp - pi;
        if ((T)m.cspot2 * p.cosb + (T)m.sspot2 * p.sinb * RT<T>::cos_(dellon) > (T)m.csprad) spot = true;
      }
    }
    initp_sample(C, p);
  } else {
    // (draws in the reference's order; the fields are set in the order the
    // stellar branch sets them, so that the optimizer does not merge the two
    // branches' last stores into one store through a selected address,
    // which would put those fields into private memory)
    const T cb = (T)1 - (T)2 * uniform<T>(C.rng);
    const T ct = -uniform<T>(C.rng);
    const T ph = (T)kR2p * uniform<T>(C.rng);
    const T lp = (T)kR2p * uniform<T>(C.rng);
    p.cosb = cb;
    p.sinb = RT<T>::sqrt_((T)1 - cb * cb);
    p.lp = lp;
    p.cost = ct;
    p.sint = RT<T>::sqrt_((T)1 - ct * ct);
    p.phi = ph;
    p.cosp = RT<T>::cos_(ph);
    p.sinp = RT<T>::sin_(ph);
  }
  initp_transform(p);
  const bool shock = m.ispot != 0 && (spot || m.fspot == 1.0);
  p.nub = outside ? rep_isrf<T>(C) : bbfreq<T>(C.rng, (T)(shock ? m.tshock : m.tstar), C.inczeta);
  // half the hot-spot photons are X-rays, uniform in h nu over 0.015..0.06 um
  if (hot && uniform<T>(C.rng) < (T)0.5)
    p.nub = (T)(1.2398 / 0.06) + uniform<T>(C.rng) * (T)(1.2398 / 0.015 - 1.2398 / 0.06);
  p.i_orig = outside ? 5 : 1;
  p.iflag = 0;
  p.exitflag = 0;
  opacset<T>(m, p.nub, p.op);
  p.wave = (T)1.2398 / p.nub;
  set_kapd(p, m, 0);
  T coslp = RT<T>::cos_(p.lp), sinlp = RT<T>::sin_(p.lp);
  const G rw = outside ? m.rarr[m.nrg - 1] : m.rmin;
  p.zp = (G)p.cosb * rw;
  G rp = (G)p.sinb * rw;
  if (outside) {                                  // cosb*rmax*0.999999, a default-real literal
    p.zp = p.zp * (G)0.999999f;
    rp = rp * (G)0.999999f;
  }
  p.xp = rp * (G)coslp;
  p.yp = rp * (G)sinlp;
  p.rsq = p.zp * p.zp + rp * rp;
  p.rtot = sqrt(p.rsq);
  set_dir<T>(p);
  p.ir = outside ? m.nrg - 1 : 1;
  p.on_wall = true;
  p.it = locate<double>(m.thetarr, m.ntg, (double)RT<T>::acos_(p.cosb));
  p.ip = locate<double>(m.phiarr, m.npg, (double)p.lp);
  if (m.ipeel == 1 && !outside) peel_direct<T, kAdj>(C, p, coslp, sinlp);
  // propagate_5_fuv.f90:20-40: the first flight
  cell_cache_reset(L);
  L.xpold = p.xp; L.ypold = p.yp; L.zpold = p.zp;
  p.iflag = 0; p.exitflag = 0;
  T tau = RT<T>::tau_of(uniform<T>(C.rng));
  L.first = true;
  flight_begin(L, tau);
  return 1;
}

// end of a packet: SED for escaped ones, outcome counters, trace row.
// code -1 is an error, counted as err_kind (a failed walk, or an emission
// that could not be placed).
template <class T>
__device__ __forceinline__ void end_packet(Ctx<T>& C, Life<T>& L, int code, int err_kind = C_ERR_WALK) {
  if (code == 1) { prop_end<T>(C, L); C.cnt.add(C_ESCAPED); }
  else if (code == 2) C.cnt.add(C_KILLED);
  else { C.cnt.add(C_ERRORS); C.cnt.add(err_kind); }
  if (C.t.trace && writer(C)) {
    const Pkt<T>& p = L.p;
    double* tr = C.t.trace + 8 * L.pid;
    tr[0] = (double)p.nub; tr[1] = (double)p.cost; tr[2] = (double)p.phi; tr[3] = (double)p.sip;
    tr[4] = (double)p.sqp; tr[5] = (double)p.sup; tr[6] = (double)L.nint; tr[7] = (double)code;
  }
}

// ---------------------------------------------------------------------------
// Parking a packet between rounds.  Only what the loop top needs is kept:
// p.op is opacset(p.nub) at every point of the loop, so it is recomputed;
// cosb/sinb/lp are rewritten by initpabs before any later use.
// ---------------------------------------------------------------------------
struct LaneIO {
  uint32_t* base;
  uint64_t stride;
  int w;
  __device__ __forceinline__ void put(uint32_t v) { base[(size_t)(w++) * stride] = v; }
  __device__ __forceinline__ uint32_t get() { return base[(size_t)(w++) * stride]; }
  __device__ __forceinline__ void put_v(double v) {
    uint64_t b = (uint64_t)__double_as_longlong(v);
    put((uint32_t)b); put((uint32_t)(b >> 32));
  }
  __device__ __forceinline__ void put_v(float v) { put(__float_as_uint(v)); }
  __device__ __forceinline__ void get_v(double& v) {
    uint32_t lo = get(), hi = get();
    v = __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
  }
  __device__ __forceinline__ void get_v(float& v) { v = __uint_as_float(get()); }
};

// words per lane: 11 FP64 geometry values, 15 T values, 8 integers
template <class T> constexpr int lane_words() { return 22 + 15 * (int)(sizeof(T) / 4) + 8; }

template <class T>
__device__ __forceinline__ void park(LaneIO& io, const Ctx<T>& C, const Life<T>& L) {
  const Pkt<T>& p = L.p;
  io.put_v(p.xp); io.put_v(p.yp); io.put_v(p.zp); io.put_v(p.rsq); io.put_v(p.rtot);
  io.put_v(p.ux); io.put_v(p.uy); io.put_v(p.uz);
  io.put_v(L.xpold); io.put_v(L.ypold); io.put_v(L.zpold);
  io.put_v(p.cost); io.put_v(p.sint); io.put_v(p.phi); io.put_v(p.cosp); io.put_v(p.sinp);
  io.put_v(p.sip); io.put_v(p.sqp); io.put_v(p.sup); io.put_v(p.svp);
  io.put_v(p.nub); io.put_v(p.wave);
#pragma unroll
  for (int i = 0; i < kMaxOpac; ++i) io.put_v(p.kapd[i]);
  io.put((uint32_t)p.ir | ((uint32_t)p.it << 10) | ((uint32_t)p.ip << 20) | ((uint32_t)(p.on_wall ? 1 : 0) << 30));
  io.put((uint32_t)p.iflag | ((uint32_t)p.i_orig << 1) | ((uint32_t)L.idust2 << 8));
  io.put((uint32_t)L.iscat); io.put((uint32_t)L.iint); io.put((uint32_t)L.nint);
  io.put((uint32_t)L.pid); io.put((uint32_t)(L.pid >> 32));
  io.put(C.rng.n);
}

template <class T>
__device__ __forceinline__ void unpark(LaneIO& io, Ctx<T>& C, Life<T>& L, uint64_t seed, uint64_t first_packet) {
  Pkt<T>& p = L.p;
  io.get_v(p.xp); io.get_v(p.yp); io.get_v(p.zp); io.get_v(p.rsq); io.get_v(p.rtot);
  io.get_v(p.ux); io.get_v(p.uy); io.get_v(p.uz);
  io.get_v(L.xpold); io.get_v(L.ypold); io.get_v(L.zpold);
  io.get_v(p.cost); io.get_v(p.sint); io.get_v(p.phi); io.get_v(p.cosp); io.get_v(p.sinp);
  io.get_v(p.sip); io.get_v(p.sqp); io.get_v(p.sup); io.get_v(p.svp);
  io.get_v(p.nub); io.get_v(p.wave);
#pragma unroll
  for (int i = 0; i < kMaxOpac; ++i) io.get_v(p.kapd[i]);
  uint32_t cell = io.get(), fl = io.get();
  p.ir = (int)(cell & 1023u); p.it = (int)((cell >> 10) & 1023u); p.ip = (int)((cell >> 20) & 1023u);
  p.on_wall = ((cell >> 30) & 1u) != 0;
  p.iflag = (int)(fl & 1u); p.i_orig = (int)((fl >> 1) & 127u); L.idust2 = (int)(fl >> 8);
  p.exitflag = 0;
  L.iscat = (int)io.get(); L.iint = (int)io.get(); L.nint = (int)io.get();
  uint32_t lo = io.get(), hi = io.get();
  L.pid = ((uint64_t)hi << 32) | lo;
  uint32_t n = io.get();
  rng_resume(C.rng, seed, first_packet + L.pid, n);
  opacset<T>(C.m, p.nub, p.op, C.coop);
  L.first = false;
  cell_cache_reset(L);
}

// Workgroup counters: zeroed at the start, flushed with one device atomic
// per counter at the end (every thread of the workgroup reaches both).
__device__ __forceinline__ void counts_begin(unsigned long long* s) {
  if (threadIdx.x < C_NCOUNTERS) s[threadIdx.x] = 0ull;
  __syncthreads();
}
__device__ __forceinline__ void counts_flush(const DeviceTallies& t, unsigned long long* s) {
  __syncthreads();
  if (threadIdx.x < C_NCOUNTERS && s[threadIdx.x]) atomicAdd(&t.counters[threadIdx.x], s[threadIdx.x]);
}

// The active lanes of a wavefront take consecutive packet indices with one
// atomic.
__device__ __forceinline__ unsigned long long fetch_packets(unsigned long long* counter) {
  unsigned long long mask = __ballot(1);
  int lane = __lane_id();
  int leader = __ffsll((long long)mask) - 1;
  unsigned long long rank = (unsigned long long)__popcll(mask & ((1ull << lane) - 1ull));
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(mask));
  base = __shfl(base, leader, 64);
  return base + rank;
}

// Transport kernel: one thread per resident lane, rp.lanes threads (a
// multiple of 256).  A lane owns one packet at a time: walking a flight,
// waiting at the loop top for its next interaction (READY), or idle.  Each
// step advances every walking lane of the wavefront by one face crossing.
// READY lanes run their interaction, and idle lanes take and emit a new
// packet, only when rp.regroup lanes of the wavefront are waiting for that
// (or when nobody walks): the long interaction code then runs with most
// lanes active instead of one lane at a time, and the walk (the bulk of
// the work) runs with all walking lanes in step.
// Every event (emission or interaction) queues one peel-off record, and the
// round's queue is shared: a lane runs events (at most rp.events) for as
// long as it can reserve queue slots, so a long-lived packet (hundreds of
// thousands of interactions in optically thick cells) advances at its own
// pace through the whole round instead of a fixed share of events.  A
// packet still waiting for an interaction when the queue is full (or its
// events are spent) is parked in HBM for the next round.
enum : int { L_IDLE = 0, L_WALK = 1, L_READY = 2 };

// The mixed path write-combines its cell tallies (tally_abs); the FP64 path
// adds every segment and absorption on its own, as the reference orders them.
template <class T>
constexpr bool kCombTally = std::is_same<T, float>::value;

// kPlain: the model has one opacity table and no accretion photons, ISRF,
// hot spots or diffusion region, and the run stages its dust tables in LDS
// (plain_model()).  Those fields of the
// kernel's model copy are then constants, so the compiler drops their code:
// k_transport<float> keeps ~15k instead of ~24k instructions and spills a
// third of the scalar registers (each spilled value costs a lane read where
// it is used).
__host__ __device__ inline bool plain_model(const DeviceModel& m) {
  return m.acc_to == 0 && m.nisrf == 0 && m.ispot == 0 && m.diff_itmin == nullptr && m.nopac == 1 && m.dust_lds;
}
template <bool kPlain>
__device__ __forceinline__ void specialize(DeviceModel& m) {
  if constexpr (kPlain) {
    m.acc_to = 0; m.nisrf = 0; m.ispot = 0; m.diff_itmin = nullptr; m.nopac = 1;
  }
}

template <class T, bool kPlain>
__global__ void __launch_bounds__(256) k_transport(DeviceModel m0, DeviceTallies t, RunParams rp) {
  constexpr bool kComb = kCombTally<T>;
  __shared__ unsigned long long s_cnt[C_NCOUNTERS];
  extern __shared__ double lds_tables[];
  counts_begin(s_cnt);
  DeviceModel m = m0;
  specialize<kPlain>(m);
  const double* inczeta = rp.inczeta;
  stage_tables<kPlain>(m, inczeta, rp.lds, lds_tables);
  Ctx<T> C{m, t, {}, {s_cnt}, inczeta, rp.q};
  const uint64_t lane = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  LaneIO io{rp.lane_words + lane, rp.lanes, 0};
  Life<T> L;
  tally_open(L);
  int st = L_IDLE;
  if (rp.round > 0 && rp.lane_live[lane]) {
    unpark<T>(io, C, L, rp.seed, rp.first_packet);
    st = L_READY;
  }
  uint32_t budget = rp.events;
  bool dry = false;                 // the packet counter ran out for this lane
  bool full = false;                // the round's peel-off queue is full
  for (;;) {
    if (st == L_WALK) {
      const int r = walk_step<T, false, kComb>(C, L);
      if (r != 0) {
        const int code = (r == 1) ? flight_end<T>(C, L) : ((r == 2) ? 1 : -1);
        if (code == 0) st = L_READY;
        else {
          if constexpr (kComb) tally_flush<T>(C, L);
          end_packet<T>(C, L, code);
          st = L_IDLE;
        }
      }
    }
    const unsigned long long walking = __ballot(st == L_WALK);
    const unsigned long long ready = __ballot(st == L_READY && budget > 0 && !full);
    const unsigned long long idle = __ballot(st == L_IDLE && budget > 0 && !dry && !full);
    if ((walking | ready | idle) == 0ull) break;
    if (ready && (walking == 0ull || (unsigned)__popcll(ready) >= rp.regroup)) {
      if (st == L_READY && budget > 0 && !full) {
        full = !hold_slot(C);
        if (!full) {
          interact<T, false, kComb>(C, L);
          budget--;
          st = L_WALK;
        }
      }
    }
    if (idle && (walking == 0ull || (unsigned)__popcll(idle) >= rp.regroup)) {
      if (st == L_IDLE && budget > 0 && !dry && !full) {
        full = !hold_slot(C);
      }
      if (st == L_IDLE && budget > 0 && !dry && !full) {
        const unsigned long long k = fetch_packets(&rp.ctl->next_packet);
        if (k >= rp.npackets) {
          dry = true;
        } else {
          rng_init(C.rng, rp.seed, rp.first_packet + k);
          L.pid = k;
          budget--;
          const int e = emit<T>(C, L);
          if (e == 0) end_packet<T>(C, L, -1, C_ERR_OTHER);
          else st = (e == 1) ? L_WALK : L_READY;
        }
      }
    }
  }
  const bool live = (st == L_READY);
  if constexpr (kComb) tally_flush<T>(C, L);
  if (live) { io.w = 0; park<T>(io, C, L); }
  rp.lane_live[lane] = live ? 1 : 0;
  for (unsigned sl = C.slot_next; sl < C.slot_end && sl < C.q.capacity; ++sl)
    ((PeelRec<T>*)C.q.recs)[sl].flags = kSlotHole;
  const unsigned long long mask = __ballot(live);
  if (__lane_id() == 0 && mask) atomicAdd(&rp.ctl->live, (unsigned)__popcll(mask));
  counts_flush(t, s_cnt);
}

// The lanes of the last round that hold a parked packet, in any order
// (count: zeroed by the host).  (A template only so that every translation
// unit that includes this header may carry it.)
template <int kUnused = 0>
__global__ void __launch_bounds__(256) k_live_list(const uint8_t* lane_live, uint64_t lanes, uint32_t* out,
                                                   unsigned* count) {
  const uint64_t l = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = l < lanes && lane_live[l] != 0;
  const unsigned long long mask = __ballot(live);
  if (!mask) return;
  const int leader = __ffsll((long long)mask) - 1;
  unsigned base = 0;
  if (__lane_id() == leader) base = atomicAdd(count, (unsigned)__popcll(mask));
  base = __shfl(base, leader, 64);
  if (live) out[base + (unsigned)__popcll(mask & ((1ull << __lane_id()) - 1ull))] = (uint32_t)l;
}

// The end of a run: every packet handed out and few still alive, each
// running a serial chain of interactions (a packet trapped in optically
// thick cells may have 10^5-10^6 of them), alone in its wavefront, so the
// chain's latency sets the run's length.  k_tail gives each of the
// `nlive` parked packets (lane slots live[]) a whole wavefront: all 64
// lanes carry the packet and compute identically, lane 0 alone writes its
// tallies, records and parked state, and the scattering angle's rejection
// sampling tries 64 candidates at once (stokes_coop).  Results are those of
// k_transport: same draws, same arithmetic, same order.
template <class T, bool kPlain>
__global__ void __launch_bounds__(256) k_tail(DeviceModel m0, DeviceTallies t, RunParams rp, const uint32_t* live,
                                              unsigned nlive) {
  constexpr bool kComb = kCombTally<T>;
  __shared__ unsigned long long s_cnt[C_NCOUNTERS];
  extern __shared__ double lds_tables[];
  counts_begin(s_cnt);
  DeviceModel m = m0;
  specialize<kPlain>(m);
  const double* inczeta = rp.inczeta;
  stage_tables<kPlain>(m, inczeta, rp.lds, lds_tables);
  Ctx<T> C{m, t, {}, {s_cnt}, inczeta, rp.q};
  C.coop = true;
  __shared__ T s_cdf[4][kCdfCache];
  C.cdf.row = s_cdf[threadIdx.x >> 6];
  C.cnt.one = true;
  const unsigned wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (wave < nlive) {
    const uint64_t slot = live[wave];
    LaneIO io{rp.lane_words + slot, rp.lanes, 0};
    Life<T> L;
    tally_open(L);
    unpark<T>(io, C, L, rp.seed, rp.first_packet);
    C.rng.coop = true;
    C.rng.cb = 0x80000000u;
    int st = L_READY;
    uint32_t budget = rp.events;
    for (;;) {
      if (st == L_WALK) {
        const int r = walk_step<T, false, kComb>(C, L);
        if (r != 0) {
          const int code = (r == 1) ? flight_end<T>(C, L) : ((r == 2) ? 1 : -1);
          if (code == 0) st = L_READY;
          else {
            if constexpr (kComb) tally_flush<T>(C, L);
            end_packet<T>(C, L, code);
            st = L_IDLE;
          }
        }
      }
      if (st == L_IDLE) break;
      if (st == L_READY) {
        if (budget == 0 || !hold_slot(C)) break;
        interact<T, false, kComb>(C, L);
        budget--;
        st = L_WALK;
      }
    }
    const bool alive = (st == L_READY);
    if constexpr (kComb) tally_flush<T>(C, L);
    if (__lane_id() == 0) {
      if (alive) { io.w = 0; park<T>(io, C, L); }
      rp.lane_live[slot] = alive ? 1 : 0;
      for (unsigned sl = C.slot_next; sl < C.slot_end && sl < C.q.capacity; ++sl)
        ((PeelRec<T>*)C.q.recs)[sl].flags = kSlotHole;
      if (alive) atomicAdd(&rp.ctl->live, 1u);
    }
  }
  counts_flush(t, s_cnt);
}

// the packet fields peeloff reads, from a queued record
template <class T>
__device__ __forceinline__ void rec_to_pkt(const PeelRec<T>& r, Pkt<T>& q) {
  q.xp = r.xp; q.yp = r.yp; q.zp = r.zp; q.rsq = r.rsq; q.rtot = r.rtot;
  q.sip = r.sip; q.sqp = r.sqp; q.sup = r.sup; q.svp = r.svp;
  q.cost = r.cost; q.sint = r.sint; q.phi = r.phi; q.cosp = r.cosp; q.sinp = r.sinp;
  q.nub = r.nub; q.wave = r.wave;
#pragma unroll
  for (int k = 0; k < kMaxOpac; ++k) q.kapd[k] = r.kapd[k];
  q.ir = (int)(r.cell & 1023u); q.it = (int)((r.cell >> 10) & 1023u); q.ip = (int)((r.cell >> 20) & 1023u);
  q.on_wall = ((r.cell >> 30) & 1u) != 0;
  q.iflag = (int)(r.flags & 1u);
  q.i_orig = (int)((r.flags >> 1) & 7u);
}

// Stokes I sent toward direction k.  For the direct stellar peel
// (newdisktherm_5_fuv.f90:371-414) it is 4 I cos(norm), and directions
// behind the emitting surface element get nothing (false).
template <class T>
__device__ __forceinline__ bool peel_weight(const DeviceModel& m, const PeelRec<T>& r, int k, T& s0) {
  s0 = r.sip;
  if ((r.flags >> 4) & 1u) {
    const double* pe = m.peel + 5 * k;
    T cosnorm = r.cost * (T)pe[0] + (r.sint * (T)pe[1] * ((T)pe[2] * r.cosp + (T)pe[3] * r.sinp));
    if (!(cosnorm > (T)0)) return false;
    s0 = (T)4 * r.sip * cosnorm;
  }
  return true;
}

// Peel kernel: each lane walks its share of the round's (record, direction)
// items, item = lane, lane + stride, ...  Every step advances each walking
// lane by one face crossing.  Two pieces of code run only for gathered
// lanes, so that they run with most of the wavefront active instead of
// stalling the walkers on nearly every step: taking the next item with the
// walk's first crossing (the reference solvers from the record position,
// several times the cost of a ray step), once rp.peel_regroup lanes are idle;
// and the tally of walks that reached the observer (exponential weight,
// findangle3, stokespeel, image and SED binning), once rp.regroup lanes wait
// for it.  Either also runs when no lane walks.
template <class T>
__global__ void __launch_bounds__(256) k_peel(DeviceModel m0, DeviceTallies t, RunParams rp) {
  __shared__ unsigned long long s_cnt[C_NCOUNTERS];
  extern __shared__ double lds_tables[];
  counts_begin(s_cnt);
  DeviceModel m = m0;
  const double* inczeta = rp.inczeta;
  stage_tables(m, inczeta, rp.lds, lds_tables);
  Ctx<T> C{m, t, {}, {s_cnt}, inczeta, PeelQueue{nullptr, nullptr, 0}};
  const unsigned nq = *rp.q.count;
  const uint64_t nrec = nq < rp.q.capacity ? nq : rp.q.capacity;
  const uint64_t npeel = (uint64_t)m.npeel;
  const uint64_t total = nrec * npeel;
  const PeelRec<T>* recs = (const PeelRec<T>*)rp.q.recs;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t rkey = rp.seed ^ 0x5045454C00000000ull ^ rp.first_packet;
  uint64_t item = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  enum : int { P_IDLE = 0, P_WALK = 1, P_DONE = 2 };
  int st = P_IDLE;
  PeelWalk<T> w;
  PeelRay<T> ray;
  ray.started = false;
  for (;;) {
    if (st == P_WALK) {
      int r;
      if constexpr (kRayPeel<T>) r = peel_ray_step<T>(C, w, ray);
      else r = peel_step<T>(C, w, C.rng);
      if (r == 1) st = P_DONE;
      else if (r == 2) { st = P_IDLE; item += stride; }
    }
    const unsigned long long walking = __ballot(st == P_WALK);
    const unsigned long long done = __ballot(st == P_DONE);
    unsigned long long more = __ballot(st == P_IDLE && item < total);
    if ((walking | done | more) == 0ull) break;
    if (done && (walking == 0ull || (unsigned)__popcll(done) >= rp.peel_tally)) {
      if (st == P_DONE) {
        const uint64_t ri = item / npeel;
        const int k = (int)(item - ri * npeel);
        const PeelRec<T>& r = recs[ri];
        T s0;
        (void)peel_weight<T>(m, r, k, s0);
        Pkt<T> q;
        rec_to_pkt<T>(r, q);
        peel_tally<T>(C, q, s0, k, r.g, r.pl, w.p.xp, w.p.yp, w.p.zp, w.tsum);
        st = P_IDLE;
        item += stride;
      }
      // the lanes just tallied take their next items in the same pass
      more = __ballot(st == P_IDLE && item < total);
    }
    if (more && (walking == 0ull || (unsigned)__popcll(more) >= rp.peel_regroup)) {
      bool found = false;
      while (st == P_IDLE && !found && item < total) {
        const uint64_t ri = item / npeel;
        const int k = (int)(item - ri * npeel);
        const PeelRec<T>& r = recs[ri];
        if (r.flags != kSlotHole) {
          T s0;
          if (peel_weight<T>(m, r, k, s0)) {
            Pkt<T> q;
            rec_to_pkt<T>(r, q);
            found = peel_begin<T>(m, q, k, w);
          }
        }
        if (!found) item += stride;
      }
      if (found) {
        // testgrid re-placements only: one stream per (run, round, walk)
        rng_init(C.rng, rkey, ((uint64_t)rp.round << 40) + item);
        int r;
        if constexpr (kRayPeel<T>) r = peel_ray_first<T>(C, w, ray, C.rng);
        else r = peel_step<T>(C, w, C.rng);
        if (r == 0) st = P_WALK;
        else if (r == 1) st = P_DONE;
        else item += stride;
      }
    }
  }
  counts_flush(t, s_cnt);
}

}  // namespace mcrt
