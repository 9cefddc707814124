// mcrt_adjoint.h -- path-replay adjoint of the peeled Stokes-I images with
// respect to the density of every cell (no counterpart in the reference).
//
// Loss L = sum over image pixels of W * image_I, W in the image layout
// [nband][nx][nx][npeel].  A packet's contribution is the sum over its
// peel-off walks e of C_e = W[pixel_e] * I_e * filter_e.  With
// free-flight sampling, the expected loss depends on rho through
//   - each walk's exp(-tau_e):        d C_e / d rho_c = -C_e kapd l_ec,
//   - the flight density sigma e^-tau of every flight f before e:
//       d log p_f / d rho_c = -kapd l_fc (+ 1/rho_c if f ends in c),
// so dL/drho_c = sum_e C_e [ -kapd l_ec + sum_{f before e} dlog p_f ].
// The second sum is rearranged per flight: flight f is weighted by
// R_f = sum of C_e over the events after f.  The packet is therefore run
// twice from its Philox counters: the forward pass sums C_total; the replay
// pass starts with R = C_total, lowers R by C_e at every event, adds
// R (-kapd l + [end] 1/rho) along every flight and -C_e kapd l along every
// walk that reaches the observer.  Nothing of the path is stored.  Albedo,
// re-emission and the dust temperature do not depend on rho (gradient at a
// fixed temperature field).
#pragma once
#include "mcrt_transport.h"

namespace mcrt {

template <class T>
__device__ __forceinline__ void adj_segment(Ctx<T>& C, size_t c, T kapd, G len, T rho, bool ends_here) {
  AdjState& A = *C.adj;
  if (!A.replay || A.R == 0.0) return;
  double g = -A.R * (double)kapd * (double)len;
  if (ends_here) g += A.R / (double)rho;
  atomicAdd(&A.grad[c], g);
}

// the loss of one walk that reached the observer
template <class T>
__device__ __forceinline__ double peel_loss(const DeviceModel& m, const PeelHit<T>& h, int k, const float* W) {
  if (m.nband <= 0 || h.ix > m.nx || h.iy > m.nx || h.ix <= 0 || h.iy <= 0) return 0.0;
  const size_t npix = (size_t)m.npeel * m.nx * m.nx;
  double l = 0.0;
  for (int b = 0; b < m.nband; ++b) {
    T wt;
    if (band_weight<T>(m, b, h.freq, wt)) {
      const size_t o = (size_t)b * npix + (size_t)k + (size_t)m.npeel * ((size_t)(h.ix - 1) + (size_t)m.nx * (size_t)(h.iy - 1));
      l += (double)W[o] * (double)(h.s[0] * wt);
    }
  }
  return l;
}

// One peel-off event: the record the transport kernel would queue, walked
// toward every observer direction exactly as k_peel walks it.
template <class T>
__device__ __forceinline__ void adj_peel_event(Ctx<T>& C, const Pkt<T>& p, int sp, int iop, bool direct, T coslp, T sinlp) {
  AdjState& A = *C.adj;
  const DeviceModel& m = C.m;
  PeelRec<T> r;
  make_rec<T>(r, p, sp, iop, direct, coslp, sinlp);
  Pkt<T> q;
  rec_to_pkt<T>(r, q);
  double lsum = 0.0;
  for (int k = 0; k < m.npeel; ++k) {
    T s0;
    if (!peel_weight<T>(m, r, k, s0)) continue;
    PeelWalk<T> w;
    if (!peel_begin<T>(m, q, k, w)) continue;
    Rng prng;   // testgrid re-placements of the walk: not the packet's stream
    rng_init(prng, A.pkey, ((uint64_t)A.nevent << 8) + (uint64_t)k);
    int st;
    while ((st = peel_step<T>(C, w, prng)) == 0) {}
    if (st != 1) continue;
    PeelHit<T> h;
    peel_weigh<T>(m, q, s0, k, r.g, r.pl, w.p.xp, w.p.yp, w.p.zp, w.tsum, h);
    const double le = peel_loss<T>(m, h, k, A.W);
    if (le == 0.0) continue;
    lsum += le;
    if (A.replay) {   // the same walk again, adding -le kapd l to every crossed cell
      PeelWalk<T> w2;
      (void)peel_begin<T>(m, q, k, w2);
      rng_init(prng, A.pkey, ((uint64_t)A.nevent << 8) + (uint64_t)k);
      while (peel_step<T>(C, w2, prng, A.grad, -le) == 0) {}
    }
  }
  if (A.replay) A.R -= lsum;
  else A.total += lsum;
  A.nevent++;
}

struct AdjParams {
  uint64_t first_packet;
  uint64_t npackets;
  uint64_t seed;
  const double* inczeta;
  const float* W;
  double* grad;        // null: loss only
  double* loss;        // accumulated with one atomic per wavefront
  LdsPlan lds;
};

// one packet's life, emission to escape / kill, with the adjoint hooks
template <class T>
__device__ __forceinline__ void adj_run_packet(Ctx<T>& C, Life<T>& L, uint64_t seed, uint64_t pkt) {
  rng_init(C.rng, seed, pkt);
  const int e = emit<T, true>(C, L);
  if (e == 0) return;                      // an accretion photon with no dust cell
  if (e == 2) interact<T, true>(C, L);
  for (;;) {
    const int r = walk_step<T, true>(C, L);
    if (r == 0) continue;
    if (r != 1) return;                     // escaped, or walk error
    if (flight_end<T>(C, L) != 0) return;   // killed
    interact<T, true>(C, L);
  }
}

// One thread per packet (grid-stride): forward pass, then, when the packet
// contributes to the loss and a gradient is wanted, the replay pass.
template <class T>
__global__ void __launch_bounds__(256) k_adjoint(DeviceModel m0, DeviceTallies t, AdjParams ap) {
  __shared__ unsigned long long s_cnt[C_NCOUNTERS];
  extern __shared__ double lds_tables[];
  counts_begin(s_cnt);
  DeviceModel m = m0;
  const double* inczeta = ap.inczeta;
  stage_tables(m, inczeta, ap.lds, lds_tables);
  AdjState A;
  A.W = ap.W;
  A.grad = ap.grad;
  Ctx<T> C{m, t, {}, {s_cnt}, inczeta, PeelQueue{nullptr, nullptr, 0}, 0u, 0u, &A};
  Life<T> L;
  double lacc = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ap.npackets; i += stride) {
    const uint64_t pkt = ap.first_packet + i;
    A.pkey = ap.seed ^ 0x41444A4F494E5400ull ^ (pkt * 0x9E3779B97F4A7C15ull);
    A.total = 0.0; A.R = 0.0; A.replay = 0; A.nevent = 0;
    L.pid = i;
    adj_run_packet<T>(C, L, ap.seed, pkt);
    lacc += A.total;
    if (ap.grad && A.total != 0.0) {
      A.replay = 1; A.R = A.total; A.nevent = 0;
      adj_run_packet<T>(C, L, ap.seed, pkt);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) lacc += __shfl_down(lacc, off, 64);
  if ((threadIdx.x & 63) == 0 && lacc != 0.0) atomicAdd(ap.loss, lacc);
  counts_flush(t, s_cnt);
}

}  // namespace mcrt
