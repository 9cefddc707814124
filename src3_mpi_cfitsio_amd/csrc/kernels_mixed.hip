// kernels_mixed.hip -- mixed-precision instantiation: FP32 opacities,
// optical depths, sampling and Stokes algebra; FP64 packet geometry (face
// roots, positions) and FP64 absorbed-energy atomics (dtauabs).
#include "mcrt_kernels.h"
#include "mcrt_launch.h"

namespace mcrt {

hipError_t launch_transport_mixed(const DeviceModel& m, const DeviceTallies& t, const RunParams& rp, hipStream_t s) {
  hipLaunchKernelGGL(k_transport<float>, dim3((unsigned)(rp.lanes / 256)), dim3(256), rp.lds.bytes, s, m, t, rp);
  return hipGetLastError();
}

int lane_words_mixed() { return lane_words<float>(); }

hipError_t launch_adjoint_mixed(const DeviceModel& m, const DeviceTallies& t, const AdjParams& ap, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_adjoint<float>, dim3(grid), dim3(256), ap.lds.bytes, s, m, t, ap);
  return hipGetLastError();
}

hipError_t launch_peel_mixed(const DeviceModel& m, const DeviceTallies& t, const RunParams& rp, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_peel<float>, dim3(grid), dim3(256), rp.lds.bytes, s, m, t, rp);
  return hipGetLastError();
}

size_t peel_record_bytes_mixed() { return sizeof(PeelRec<float>); }

hipError_t launch_trace_mixed(const DeviceModel& m, const TraceArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_trace<float>, dim3((a.n + 255) / 256), dim3(256), 0, s, m, a);
  return hipGetLastError();
}

hipError_t launch_mueller_mixed(int n, int kind, const double* par, double* dir, double* st, hipStream_t s) {
  hipLaunchKernelGGL(k_mueller<float>, dim3((n + 255) / 256), dim3(256), 0, s, n, kind, par, dir, st);
  return hipGetLastError();
}

}  // namespace mcrt
