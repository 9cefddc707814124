// kernels_fp64.hip -- FP64 instantiation (built with -ffp-contract=off so
// every product and sum rounds exactly as in the reference's Fortran), plus
// the FP64-only Lucy temperature and fill kernels.
#define MCRT_FP64_ONLY_KERNELS 1
#include "mcrt_kernels.h"
#include "mcrt_launch.h"
#include "mcrt_emiss.h"
#include "mcrt_build.h"

namespace mcrt {

hipError_t launch_build(const BuildArgs& a, hipStream_t s) {
  const dim3 grid(a.nblk, (unsigned)a.nmodels);
  hipLaunchKernelGGL(k_build_cells, grid, dim3(256), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_build_mass, dim3((unsigned)a.nmodels), dim3(256), 0, s, a);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_build_finish, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_setemiss(const EmissArgs& a, hipStream_t s) {
  const size_t nseg = (size_t)a.nopac * a.nnu * a.nT;
  hipLaunchKernelGGL(k_emiss_seg, dim3((unsigned)((nseg + 255) / 256)), dim3(256), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_emiss_scan, dim3((unsigned)a.nopac), dim3(256), 2 * (size_t)a.nT * sizeof(double), s, a);
  return hipGetLastError();
}

hipError_t launch_transport_fp64(const DeviceModel& m, const DeviceTallies& t, const RunParams& rp, hipStream_t s) {
  if (plain_model(m))
    hipLaunchKernelGGL((k_transport<double, true>), dim3((unsigned)(rp.lanes / 256)), dim3(256), rp.lds.bytes, s, m, t, rp);
  else
    hipLaunchKernelGGL((k_transport<double, false>), dim3((unsigned)(rp.lanes / 256)), dim3(256), rp.lds.bytes, s, m, t, rp);
  return hipGetLastError();
}

hipError_t launch_tail_fp64(const DeviceModel& m, const DeviceTallies& t, const RunParams& rp, const uint32_t* live,
                             unsigned nlive, hipStream_t s) {
  if (nlive == 0) return hipSuccess;
  const unsigned grid = (nlive + 3u) / 4u;      // four packets per 256-thread block
  if (plain_model(m))
    hipLaunchKernelGGL((k_tail<double, true>), dim3(grid), dim3(256), rp.lds.bytes, s, m, t, rp, live, nlive);
  else
    hipLaunchKernelGGL((k_tail<double, false>), dim3(grid), dim3(256), rp.lds.bytes, s, m, t, rp, live, nlive);
  return hipGetLastError();
}

int tail_capacity_fp64(const DeviceModel& m, unsigned lds_bytes, int cu_count) {
  int n = 0;
  const hipError_t e = plain_model(m)
      ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_tail<double, true>, 256, lds_bytes)
      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_tail<double, false>, 256, lds_bytes);
  if (e != hipSuccess) return 0;
  return n * 4 * cu_count;
}

hipError_t launch_live_list(const uint8_t* lane_live, uint64_t lanes, uint32_t* out, unsigned* count, hipStream_t s) {
  hipLaunchKernelGGL(k_live_list<0>, dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0, s, lane_live, lanes, out, count);
  return hipGetLastError();
}

hipError_t launch_acc_place_fp64(const DeviceModel& m, uint64_t seed, uint64_t first_packet, AccStart* out,
                                  hipStream_t s) {
  const unsigned grid = (unsigned)((m.acc_to + 255) / 256);
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(k_acc_place<double>, dim3(grid), dim3(256), 0, s, m, seed, first_packet, out);
  return hipGetLastError();
}

int lane_words_fp64() { return lane_words<double>(); }

int transport_blocks_per_cu_fp64(const DeviceModel& m, unsigned lds_bytes) {
  int n = 0;
  const hipError_t e = plain_model(m)
      ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_transport<double, true>, 256, lds_bytes)
      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_transport<double, false>, 256, lds_bytes);
  if (e != hipSuccess) return 0;
  return n;
}

hipError_t launch_adjoint_fp64(const DeviceModel& m, const DeviceTallies& t, const AdjParams& ap, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_adjoint<double>, dim3(grid), dim3(256), ap.lds.bytes, s, m, t, ap);
  return hipGetLastError();
}

hipError_t launch_peel_fp64(const DeviceModel& m, const DeviceTallies& t, const RunParams& rp, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_peel<double>, dim3(grid), dim3(256), rp.lds.bytes, s, m, t, rp);
  return hipGetLastError();
}

size_t peel_record_bytes_fp64() { return sizeof(PeelRec<double>); }

template <int BS, bool kComb>
static hipError_t batch_launch_fp64(const BatchLaunch& L, hipStream_t s) {
  hipLaunchKernelGGL((k_batch<double, BS, kComb>), dim3(L.grid), dim3(BS), L.lds.bytes, s, L.m, L.t, L.bd, L.run, L.inczeta, L.lds);
  return hipGetLastError();
}

template <int BS, bool kComb>
static hipError_t batch_occ_fp64(const BatchLaunch& L, int* blocks_per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_batch<double, BS, kComb>, BS, L.lds.bytes);
}

hipError_t launch_batch_fp64(const BatchLaunch& L, hipStream_t s) {
  switch (L.block * 2 + (L.combine ? 1 : 0)) {
    case 128: return batch_launch_fp64<64, false>(L, s);
    case 129: return batch_launch_fp64<64, true>(L, s);
    case 256: return batch_launch_fp64<128, false>(L, s);
    case 257: return batch_launch_fp64<128, true>(L, s);
    case 512: return batch_launch_fp64<256, false>(L, s);
    case 513: return batch_launch_fp64<256, true>(L, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t batch_occupancy_fp64(const BatchLaunch& L, int* blocks_per_cu) {
  switch (L.block * 2 + (L.combine ? 1 : 0)) {
    case 128: return batch_occ_fp64<64, false>(L, blocks_per_cu);
    case 129: return batch_occ_fp64<64, true>(L, blocks_per_cu);
    case 256: return batch_occ_fp64<128, false>(L, blocks_per_cu);
    case 257: return batch_occ_fp64<128, true>(L, blocks_per_cu);
    case 512: return batch_occ_fp64<256, false>(L, blocks_per_cu);
    case 513: return batch_occ_fp64<256, true>(L, blocks_per_cu);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_trace_fp64(const DeviceModel& m, const TraceArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_trace<double>, dim3((a.n + 255) / 256), dim3(256), 0, s, m, a);
  return hipGetLastError();
}

hipError_t launch_mueller_fp64(int n, int kind, const double* par, double* dir, double* st, hipStream_t s) {
  hipLaunchKernelGGL(k_mueller<double>, dim3((n + 255) / 256), dim3(256), 0, s, n, kind, par, dir, st);
  return hipGetLastError();
}

hipError_t launch_tfinal(const DeviceModel& m, const double* dtauabs, double nphot, int mode, double* tdave, hipStream_t s) {
  size_t ncell = (size_t)(m.nrg - 1) * (m.ntg - 1) * (m.npg - 1);
  if (mode == 0) {
    int nh = (m.nrg - 1) * ((m.ntg - 1) / 2);
    hipLaunchKernelGGL(k_tfinal_half, dim3((nh + 255) / 256), dim3(256), 0, s, m, dtauabs, nphot, tdave);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_tfinal_fill, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, s, m, tdave);
  } else {
    hipLaunchKernelGGL(k_tcell, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, s, m, dtauabs, nphot, tdave);
  }
  return hipGetLastError();
}

hipError_t launch_lucy_sigma(const DeviceModel& m, const double* told, const double* tnew, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_lucy_sigma, dim3(1), dim3(256), 0, s, m, told, tnew, out);
  return hipGetLastError();
}

hipError_t launch_setdiffus(const DeviceModel& m, const DiffArgs& a, hipStream_t s) {
  const int nr = m.nrg - 1, nh = (m.ntg - 1) / 2;
  hipError_t e = launch_fill_i32(a.itmin, (size_t)nr, 0x7fffffff, s);
  if (e != hipSuccess) return e;
  if (nh >= 2) {
    const int n = nr * (nh - 1);
    hipLaunchKernelGGL(k_setdiffus, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  // columns where no cell qualifies: 0
  hipLaunchKernelGGL(k_itmin_none, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, s, a.itmin, nr);
  return hipGetLastError();
}

hipError_t launch_diff_reset(const DeviceModel& m, double* tdust, hipStream_t s) {
  const size_t n = (size_t)(m.nrg - 1) * (m.ntg - 1) * (m.npg - 1);
  hipLaunchKernelGGL(k_diff_reset, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m, tdust);
  return hipGetLastError();
}

hipError_t launch_diff_columns(const DeviceModel& m, const DiffArgs& a, const unsigned* count, const int* nabs,
                               double delta, double* tdust, hipStream_t s) {
  const int n = (m.nrg - 1) * (m.npg - 1) * 2;
  hipLaunchKernelGGL(k_diff_columns, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m, a, count, nabs, delta, tdust);
  return hipGetLastError();
}

hipError_t launch_batch_tfinal(const BatchLaunch& L, const unsigned long long* npk, int mode, double* tdave, size_t max_cells,
                               hipStream_t s) {
  const unsigned nm = (unsigned)L.bd.nmodels;
  const unsigned gx = (unsigned)((max_cells + 255) / 256);
  if (nm == 0 || gx == 0) return hipSuccess;
  if (mode == 0) {
    hipLaunchKernelGGL(k_batch_tfinal_half, dim3(gx, nm), dim3(256), 0, s, L.m, L.bd, npk, (const double*)L.t.dtauabs, tdave);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_batch_tfinal_fill, dim3(gx, nm), dim3(256), 0, s, L.m, L.bd, npk, tdave);
  } else {
    hipLaunchKernelGGL(k_batch_tcell, dim3(gx, nm), dim3(256), 0, s, L.m, L.bd, npk, (const double*)L.t.dtauabs, tdave);
  }
  return hipGetLastError();
}

template <class V>
static hipError_t fill(V* p, size_t n, V v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  size_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill<V>, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}
hipError_t launch_fill_f32(float* p, size_t n, float v, hipStream_t s) { return fill(p, n, v, s); }
hipError_t launch_fill_f64(double* p, size_t n, double v, hipStream_t s) { return fill(p, n, v, s); }
hipError_t launch_fill_i32(int* p, size_t n, int v, hipStream_t s) { return fill(p, n, v, s); }
hipError_t launch_fill_u64(unsigned long long* p, size_t n, unsigned long long v, hipStream_t s) { return fill(p, n, v, s); }

}  // namespace mcrt
