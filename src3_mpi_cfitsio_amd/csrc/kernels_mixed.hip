// kernels_mixed.hip -- mixed-precision instantiation: FP32 opacities,
// optical depths, sampling and Stokes algebra; FP64 packet geometry (face
// roots, positions) and FP64 absorbed-energy atomics (dtauabs).
#include "mcrt_kernels.h"
#include "mcrt_launch.h"

namespace mcrt {

hipError_t launch_transport_mixed(const DeviceModel& m, const DeviceTallies& t, const RunParams& rp, hipStream_t s) {
  if (plain_model(m))
    hipLaunchKernelGGL((k_transport<float, true>), dim3((unsigned)(rp.lanes / 256)), dim3(256), rp.lds.bytes, s, m, t, rp);
  else
    hipLaunchKernelGGL((k_transport<float, false>), dim3((unsigned)(rp.lanes / 256)), dim3(256), rp.lds.bytes, s, m, t, rp);
  return hipGetLastError();
}

hipError_t launch_tail_mixed(const DeviceModel& m, const DeviceTallies& t, const RunParams& rp, const uint32_t* live,
                             unsigned nlive, hipStream_t s) {
  if (nlive == 0) return hipSuccess;
  const unsigned grid = (nlive + 3u) / 4u;      // four packets per 256-thread block
  if (plain_model(m))
    hipLaunchKernelGGL((k_tail<float, true>), dim3(grid), dim3(256), rp.lds.bytes, s, m, t, rp, live, nlive);
  else
    hipLaunchKernelGGL((k_tail<float, false>), dim3(grid), dim3(256), rp.lds.bytes, s, m, t, rp, live, nlive);
  return hipGetLastError();
}

int tail_capacity_mixed(const DeviceModel& m, unsigned lds_bytes, int cu_count) {
  int n = 0;
  const hipError_t e = plain_model(m)
      ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_tail<float, true>, 256, lds_bytes)
      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_tail<float, false>, 256, lds_bytes);
  if (e != hipSuccess) return 0;
  return n * 4 * cu_count;
}

hipError_t launch_acc_place_mixed(const DeviceModel& m, uint64_t seed, uint64_t first_packet, AccStart* out,
                                  hipStream_t s) {
  const unsigned grid = (unsigned)((m.acc_to + 255) / 256);
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(k_acc_place<float>, dim3(grid), dim3(256), 0, s, m, seed, first_packet, out);
  return hipGetLastError();
}

int lane_words_mixed() { return lane_words<float>(); }

int transport_blocks_per_cu_mixed(const DeviceModel& m, unsigned lds_bytes) {
  int n = 0;
  const hipError_t e = plain_model(m)
      ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_transport<float, true>, 256, lds_bytes)
      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_transport<float, false>, 256, lds_bytes);
  if (e != hipSuccess) return 0;
  return n;
}

hipError_t launch_adjoint_mixed(const DeviceModel& m, const DeviceTallies& t, const AdjParams& ap, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_adjoint<float>, dim3(grid), dim3(256), ap.lds.bytes, s, m, t, ap);
  return hipGetLastError();
}

hipError_t launch_peel_mixed(const DeviceModel& m, const DeviceTallies& t, const RunParams& rp, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_peel<float>, dim3(grid), dim3(256), rp.lds.bytes, s, m, t, rp);
  return hipGetLastError();
}

hipError_t launch_peel_split_mixed(const DeviceModel& m, const DeviceTallies& t, const RunParams& rp, const PeelRays& rays,
                                   const PeelHits& hits, uint64_t item0, uint64_t item1, int grid_start, int grid_walk,
                                   int grid_hits, unsigned fetch, hipStream_t s) {
  hipLaunchKernelGGL(k_peel_start<float>, dim3(grid_start), dim3(256), rp.lds.bytes, s, m, t, rp, rays, hits, item0, item1);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_peel_walk<float>, dim3(grid_walk), dim3(256), rp.lds.bytes, s, m, t, rp, rays, hits, fetch);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  // hits: npeel blocks per part of the hit list, the peeled SED of each
  // block's direction in LDS when that slice fits in 64 KB
  const unsigned slice = 9u * (unsigned)m.nfreq * (unsigned)m.nap * (unsigned)kNo * 4u;
  const int sed_lds = slice <= 64u * 1024u ? 1 : 0;
  const int parts = grid_hits / m.npeel > 0 ? grid_hits / m.npeel : 1;
  hipLaunchKernelGGL(k_peel_hits<float>, dim3((unsigned)(parts * m.npeel)), dim3(256),
                     rp.lds.bytes + (sed_lds ? slice : 0u), s, m, t, rp, hits, sed_lds);
  return hipGetLastError();
}

int peel_walk_blocks_per_cu_mixed(unsigned lds_bytes) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_peel_walk<float>, 256, lds_bytes) != hipSuccess) return 0;
  return n;
}

size_t peel_record_bytes_mixed() { return sizeof(PeelRec<float>); }

template <int BS, bool kComb>
static hipError_t batch_launch_mixed(const BatchLaunch& L, hipStream_t s) {
  hipLaunchKernelGGL((k_batch<float, BS, kComb>), dim3(L.grid), dim3(BS), L.lds.bytes, s, L.m, L.t, L.bd, L.run, L.inczeta, L.lds);
  return hipGetLastError();
}

template <int BS, bool kComb>
static hipError_t batch_occ_mixed(const BatchLaunch& L, int* blocks_per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_batch<float, BS, kComb>, BS, L.lds.bytes);
}

hipError_t launch_batch_mixed(const BatchLaunch& L, hipStream_t s) {
  switch (L.block * 2 + (L.combine ? 1 : 0)) {
    case 128: return batch_launch_mixed<64, false>(L, s);
    case 129: return batch_launch_mixed<64, true>(L, s);
    case 256: return batch_launch_mixed<128, false>(L, s);
    case 257: return batch_launch_mixed<128, true>(L, s);
    case 512: return batch_launch_mixed<256, false>(L, s);
    case 513: return batch_launch_mixed<256, true>(L, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t batch_occupancy_mixed(const BatchLaunch& L, int* blocks_per_cu) {
  switch (L.block * 2 + (L.combine ? 1 : 0)) {
    case 128: return batch_occ_mixed<64, false>(L, blocks_per_cu);
    case 129: return batch_occ_mixed<64, true>(L, blocks_per_cu);
    case 256: return batch_occ_mixed<128, false>(L, blocks_per_cu);
    case 257: return batch_occ_mixed<128, true>(L, blocks_per_cu);
    case 512: return batch_occ_mixed<256, false>(L, blocks_per_cu);
    case 513: return batch_occ_mixed<256, true>(L, blocks_per_cu);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_trace_mixed(const DeviceModel& m, const TraceArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_trace<float>, dim3((a.n + 255) / 256), dim3(256), 0, s, m, a);
  return hipGetLastError();
}

hipError_t launch_mueller_mixed(int n, int kind, const double* par, double* dir, double* st, hipStream_t s) {
  hipLaunchKernelGGL(k_mueller<float>, dim3((n + 255) / 256), dim3(256), 0, s, n, kind, par, dir, st);
  return hipGetLastError();
}

}  // namespace mcrt
