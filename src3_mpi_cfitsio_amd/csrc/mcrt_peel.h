// mcrt_peel.h -- the peel-off walks of the mixed path as three kernels, one
// per kind of work (peeloff_3d_2_test.f90:7-347 split at its loop):
//   k_peel_start  every (record, direction) item of the round: the weight
//                 toward the observer, the walk's first crossing by the
//                 reference's solvers from the record position and the ray
//                 set-up; walks still inside the grid are written out as ray
//                 states, walks that left it at once as hits;
//   k_peel_walk   the ray states, one face crossing per step (peel_ray_step)
//                 with every lane walking: a lane whose walk ends takes the
//                 next state from its wavefront's reserved batch;
//   k_peel_hits   the walks that reached the observer: exponential weight,
//                 findangle3, stokespeel, image and peeled-SED binning.
// In one kernel (k_peel) these run for the lanes that need them while the
// rest of the wavefront waits: the first crossing (several times a ray
// step) ran with ~9 of 64 lanes and a third of the kernel's cycles, and the
// walk itself with 42 of 64.  Each piece here runs with full wavefronts; the
// price is the ray state written and read once per walk (128 bytes).
// The arithmetic of every walk is k_peel's: same first crossing, same ray
// steps, same tally, so the images and SEDs differ only by the order of
// their FP32 atomic sums.
#pragma once
#include "mcrt_transport.h"

namespace mcrt {

// Ray states, structure of arrays over `capacity` entries.
struct PeelRays {
  double *ox, *oy, *oz, *b, *h2, *t, *tr, *tlo, *tlo2, *thi, *thi2, *tp;
  float* kapd[kMaxOpac];   // the walk's opacities per table (record's frequency)
  float* tsum;             // optical depth so far
  uint32_t* cell;          // ir | it << 10 | ip << 20
  uint32_t* item;          // record * npeel + direction; kSlotHole: an unused slot
  uint8_t* steps;          // (dr + 1) | (dp + 1) << 2
  unsigned* count;         // states written this round
  unsigned* next;          // states handed out to walking lanes
  unsigned capacity;
};

// Walks that reached the observer.
struct PeelHits {
  double *xe, *ye, *ze;    // where the walk left the grid
  float* tsum;
  uint32_t* item;          // kSlotHole: an unused slot
  unsigned* count;
  unsigned capacity;
};

// Slots for the lanes of a wavefront where `want` holds, from a pool the
// wavefront reserves 64 at a time with one atomic (base, left: the pool,
// wavefront-uniform).  Returns the lane's slot; lanes that do not want one
// get ~0u.  Slots reserved past `capacity` are never handed out.
__device__ __forceinline__ unsigned pool_take(unsigned* counter, bool want, unsigned& base, unsigned& left) {
  const unsigned long long mask = __ballot(want);
  const unsigned n = (unsigned)__popcll(mask);
  const int lane = __lane_id();
  const unsigned rank = (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
  unsigned slot = ~0u;
  if (n == 0) return slot;
  if (left >= n) {
    slot = base + rank;
    base += n;
    left -= n;
  } else {
    // the rest of the pool for the first `left` lanes, a new pool for the others
    unsigned nb = 0;
    if (lane == __ffsll((long long)mask) - 1) nb = atomicAdd(counter, 64u);
    nb = __shfl(nb, __ffsll((long long)mask) - 1, 64);
    if (rank < left) slot = base + rank;
    else slot = nb + (rank - left);
    const unsigned used = n - left;
    base = nb + used;
    left = 64u - used;
  }
  return want ? slot : ~0u;
}

template <class T>
__device__ __forceinline__ void write_ray(const PeelRays& R, unsigned s, const PeelWalk<T>& w, const PeelRay<T>& r,
                                          uint32_t item) {
  R.ox[s] = r.ox; R.oy[s] = r.oy; R.oz[s] = r.oz; R.b[s] = r.b; R.h2[s] = r.h2; R.t[s] = r.t;
  R.tr[s] = r.tr; R.tlo[s] = r.tlo; R.tlo2[s] = r.tlo2; R.thi[s] = r.thi; R.thi2[s] = r.thi2; R.tp[s] = r.tp;
#pragma unroll
  for (int i = 0; i < kMaxOpac; ++i) R.kapd[i][s] = (float)w.p.kapd[i];
  R.tsum[s] = (float)w.tsum;
  R.cell[s] = (uint32_t)w.p.ir | ((uint32_t)w.p.it << 10) | ((uint32_t)w.p.ip << 20);
  R.item[s] = item;
  R.steps[s] = (uint8_t)((r.dr + 1) | ((r.dp + 1) << 2));
}

template <class T>
__device__ __forceinline__ void read_ray(const PeelRays& R, unsigned s, const DeviceModel& m, PeelWalk<T>& w,
                                         PeelRay<T>& r, uint32_t& item) {
  r.ox = R.ox[s]; r.oy = R.oy[s]; r.oz = R.oz[s]; r.b = R.b[s]; r.h2 = R.h2[s]; r.t = R.t[s];
  r.tr = R.tr[s]; r.tlo = R.tlo[s]; r.tlo2 = R.tlo2[s]; r.thi = R.thi[s]; r.thi2 = R.thi2[s]; r.tp = R.tp[s];
#pragma unroll
  for (int i = 0; i < kMaxOpac; ++i) w.p.kapd[i] = (T)R.kapd[i][s];
  w.tsum = (T)R.tsum[s];
  const uint32_t cell = R.cell[s];
  w.p.ir = (int)(cell & 1023u); w.p.it = (int)((cell >> 10) & 1023u); w.p.ip = (int)((cell >> 20) & 1023u);
  item = R.item[s];
  const uint32_t st = R.steps[s];
  r.dr = (int)(st & 3u) - 1;
  r.dp = (int)((st >> 2) & 3u) - 1;
  r.started = true;
  const int k = (int)(item % (uint32_t)m.npeel);
  const double* pe = m.peel + 5 * k;
  w.ux = pe[1] * pe[2]; w.uy = pe[1] * pe[3]; w.uz = pe[0];
  w.nseg = 1;                 // the first crossing
}

template <class T>
__device__ __forceinline__ void write_hit(const PeelHits& H, unsigned s, const PeelWalk<T>& w, uint32_t item) {
  H.xe[s] = w.p.xp; H.ye[s] = w.p.yp; H.ze[s] = w.p.zp;
  H.tsum[s] = (float)w.tsum;
  H.item[s] = item;
}

// The round's (record, direction) items [item0, item1), grid-stride (the
// host reads the record count after the transport kernel and splits the
// items into chunks that fit the ray and hit buffers).  The start of each
// walk is k_peel's fetch: the record's weight toward direction k, the set-up
// of the walk, and its first crossing with testgrid's re-placement stream of
// the same (run, round, item).
template <class T>
__global__ void __launch_bounds__(256) k_peel_start(DeviceModel m0, DeviceTallies t, RunParams rp, PeelRays rays,
                                                    PeelHits hits, uint64_t item0, uint64_t item1) {
  __shared__ unsigned long long s_cnt[C_NCOUNTERS];
  extern __shared__ double lds_tables[];
  counts_begin(s_cnt);
  DeviceModel m = m0;
  const double* inczeta = rp.inczeta;
  stage_tables(m, inczeta, rp.lds, lds_tables);
  Ctx<T> C{m, t, {}, {s_cnt}, inczeta, PeelQueue{nullptr, nullptr, 0}};
  const uint64_t npeel = (uint64_t)m.npeel;
  const PeelRec<T>* recs = (const PeelRec<T>*)rp.q.recs;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t rkey = rp.seed ^ 0x5045454C00000000ull ^ rp.first_packet;
  unsigned rbase = 0, rleft = 0, hbase = 0, hleft = 0;
  for (uint64_t base = item0 + (uint64_t)blockIdx.x * blockDim.x; base < item1; base += stride) {
    const uint64_t item = base + threadIdx.x;
    int r = 2;
    PeelWalk<T> w;
    PeelRay<T> ray;
    ray.started = false;
    if (item < item1) {
      const uint64_t ri = item / npeel;
      const int k = (int)(item - ri * npeel);
      const PeelRec<T>& rec = recs[ri];
      if (rec.flags != kSlotHole) {
        T s0;
        if (peel_weight<T>(m, rec, k, s0)) {
          Pkt<T> q;
          rec_to_pkt<T>(rec, q);
          if (peel_begin<T>(m, q, k, w)) {
            rng_init(C.rng, rkey, ((uint64_t)rp.round << 40) + item);
            r = peel_ray_first<T>(C, w, ray, C.rng);
          }
        }
      }
    }
    // (the host sizes both buffers so that no slot can fall outside them)
    const unsigned rs = pool_take(rays.count, r == 0, rbase, rleft);
    if (r == 0) {
      if (rs < rays.capacity) write_ray<T>(rays, rs, w, ray, (uint32_t)item);
      else { C.cnt.add(C_ERRORS); C.cnt.add(C_ERR_OTHER); }
    }
    const unsigned hs = pool_take(hits.count, r == 1, hbase, hleft);
    if (r == 1) {
      if (hs < hits.capacity) write_hit<T>(hits, hs, w, (uint32_t)item);
      else { C.cnt.add(C_ERRORS); C.cnt.add(C_ERR_OTHER); }
    }
  }
  // the unused rest of this wavefront's pools
  const int lane = __lane_id();
  if ((unsigned)lane < rleft && rbase + lane < rays.capacity) rays.item[rbase + lane] = kSlotHole;
  if ((unsigned)lane < hleft && hbase + lane < hits.capacity) hits.item[hbase + lane] = kSlotHole;
  counts_flush(t, s_cnt);
}

// The ray states of the round, one face crossing per step for every walking
// lane.  A lane whose walk ended takes a new state once `fetch` lanes of its
// wavefront are idle (or none walks): the loads of a new state wait on the
// memory system, so fetching one lane at a time would stall the others.
template <class T>
__global__ void __launch_bounds__(256) k_peel_walk(DeviceModel m0, DeviceTallies t, RunParams rp, PeelRays rays,
                                                   PeelHits hits, unsigned fetch) {
  __shared__ unsigned long long s_cnt[C_NCOUNTERS];
  extern __shared__ double lds_tables[];
  counts_begin(s_cnt);
  DeviceModel m = m0;
  const double* inczeta = rp.inczeta;
  stage_tables(m, inczeta, rp.lds, lds_tables);
  Ctx<T> C{m, t, {}, {s_cnt}, inczeta, PeelQueue{nullptr, nullptr, 0}};
  const unsigned nr = *rays.count < rays.capacity ? *rays.count : rays.capacity;
  unsigned gbase = 0, gleft = 0, hbase = 0, hleft = 0;
  bool dry = false;
  bool walking = false;
  uint32_t item = kSlotHole;
  unsigned long long segs = 0;    // this lane's segments (the first crossings are counted at the start)
  PeelWalk<T> w;
  PeelRay<T> ray;
  for (;;) {
    int r = 0;
    if (walking) r = peel_ray_step<T, false>(C, w, ray);
    const unsigned hs = pool_take(hits.count, walking && r == 1, hbase, hleft);
    if (walking && r == 1) {
      if (hs < hits.capacity) write_hit<T>(hits, hs, w, item);
      else { C.cnt.add(C_ERRORS); C.cnt.add(C_ERR_OTHER); }
    }
    if (r != 0) {
      segs += (unsigned long long)(w.nseg - 1);
      walking = false;
    }
    const unsigned long long walk_mask = __ballot(walking);
    const unsigned long long idle = __ballot(!walking && !dry);
    if ((walk_mask | idle) == 0ull) break;
    if (idle && (walk_mask == 0ull || (unsigned)__popcll(idle) >= fetch)) {
      const unsigned s = pool_take(rays.next, !walking && !dry, gbase, gleft);
      if (!walking && !dry) {
        if (s >= nr) {
          dry = true;
        } else {
          read_ray<T>(rays, s, m, w, ray, item);
          walking = item != kSlotHole;
        }
      }
    }
  }
  const int lane = __lane_id();
  if ((unsigned)lane < hleft && hbase + lane < hits.capacity) hits.item[hbase + lane] = kSlotHole;
  for (int off = 32; off > 0; off >>= 1) segs += __shfl_xor(segs, off, 64);
  if (lane == 0 && segs) C.cnt.add(C_PEELSEG, segs);
  counts_flush(t, s_cnt);
}

// The walks that reached the observer: k_peel's tally.  Block b takes the
// hits of direction k = b % npeel in part b / npeel of the hit list (the
// other directions' hits of that part go to their own blocks).  With
// sed_lds, the peeled SED of direction k ([9][8][nap][nfreq] floats) is
// summed in LDS and added to t.psed once per block: every hit adds to 27
// SED bins, and the hits of a round crowd into the few (frequency,
// direction) bins where the light is, so global atomics on them serialize
// in L2 (the kernel ran at 0.7% VALU with 2% L2 hits); in LDS each CU sums
// its own copy.  The images, spread over nx*nx pixels, stay global atomics.
template <class T>
__global__ void __launch_bounds__(256) k_peel_hits(DeviceModel m0, DeviceTallies t, RunParams rp, PeelHits hits,
                                                   int sed_lds) {
  __shared__ unsigned long long s_cnt[C_NCOUNTERS];
  extern __shared__ double lds_tables[];
  counts_begin(s_cnt);
  DeviceModel m = m0;
  const double* inczeta = rp.inczeta;
  stage_tables(m, inczeta, rp.lds, lds_tables);
  Ctx<T> C{m, t, {}, {s_cnt}, inczeta, PeelQueue{nullptr, nullptr, 0}};
  const unsigned nh = *hits.count < hits.capacity ? *hits.count : hits.capacity;
  const uint64_t npeel = (uint64_t)m.npeel;
  const int k = (int)(blockIdx.x % (unsigned)npeel);
  const unsigned parts = gridDim.x / (unsigned)npeel;
  const unsigned part = blockIdx.x / (unsigned)npeel;
  const unsigned per = (nh + parts - 1) / parts;
  const unsigned h0 = part * per, h1 = (h0 + per < nh) ? h0 + per : nh;
  const int ns = m.nfreq * m.nap * kNo;            // SED bins of one direction and component
  float* sed = nullptr;
  if (sed_lds) {
    sed = (float*)((char*)lds_tables + rp.lds.bytes);
    for (int i = threadIdx.x; i < 9 * ns; i += blockDim.x) sed[i] = 0.0f;
    __syncthreads();
  }
  const PeelRec<T>* recs = (const PeelRec<T>*)rp.q.recs;
  for (unsigned h = h0 + threadIdx.x; h < h1; h += blockDim.x) {
    const uint32_t item = hits.item[h];
    if (item == kSlotHole) continue;
    const uint64_t ri = item / npeel;
    if ((int)(item - ri * npeel) != k) continue;
    const PeelRec<T>& rec = recs[ri];
    T s0;
    (void)peel_weight<T>(m, rec, k, s0);
    Pkt<T> q;
    rec_to_pkt<T>(rec, q);
    peel_tally<T>(C, q, s0, k, rec.g, rec.pl, hits.xe[h], hits.ye[h], hits.ze[h], (T)hits.tsum[h], sed);
  }
  if (sed_lds) {
    __syncthreads();
    // t.psed[c][indx][ia][k][inub]: slice entry (c, j) with j = inub-1 + nfreq*(ia-1 + nap*(indx-1))
    const size_t nb = (size_t)m.nfreq * npeel * m.nap * kNo;
    for (int i = threadIdx.x; i < 9 * ns; i += blockDim.x) {
      const float v = sed[i];
      if (v != 0.0f) {
        const int c = i / ns, j = i - c * ns;
        const int inub = j % m.nfreq, rest = j / m.nfreq;          // rest = ia-1 + nap*(indx-1)
        atomicAdd(&t.psed[(size_t)c * nb + (size_t)inub + (size_t)m.nfreq * ((size_t)k + npeel * (size_t)rest)], v);
      }
    }
  }
  counts_flush(t, s_cnt);
}

}  // namespace mcrt
