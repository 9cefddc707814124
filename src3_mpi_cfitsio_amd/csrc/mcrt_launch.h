// mcrt_launch.h -- host-side launchers exported by the per-precision kernel
// translation units (kernels_fp64.hip: -ffp-contract=off; kernels_mixed.hip).
#pragma once
#include "mcrt_batch.h"
#include "mcrt_peel.h"

namespace mcrt {

struct RunParams;
struct EmissArgs;
struct BuildArgs;
struct TraceArgs;
struct AdjParams;

// one launch of the batch kernel k_batch (block 64, 128 or 256)
struct BatchLaunch {
  DeviceModel m;
  DeviceTallies t;
  BatchDesc bd;
  BatchRun* run;
  const double* inczeta;
  LdsPlan lds;
  int block;
  bool combine;
  unsigned grid;
};

// rp.lanes must be a multiple of 256
hipError_t launch_transport_fp64(const DeviceModel& m, const DeviceTallies& t, const RunParams& rp, hipStream_t s);
hipError_t launch_transport_mixed(const DeviceModel& m, const DeviceTallies& t, const RunParams& rp, hipStream_t s);
// k_tail: one wavefront per parked packet of the lane slots live[0 .. nlive)
hipError_t launch_tail_fp64(const DeviceModel& m, const DeviceTallies& t, const RunParams& rp, const uint32_t* live,
                            unsigned nlive, hipStream_t s);
hipError_t launch_tail_mixed(const DeviceModel& m, const DeviceTallies& t, const RunParams& rp, const uint32_t* live,
                             unsigned nlive, hipStream_t s);
// packets k_tail keeps resident at once (one wavefront each)
int tail_capacity_fp64(const DeviceModel& m, unsigned lds_bytes, int cu_count);
int tail_capacity_mixed(const DeviceModel& m, unsigned lds_bytes, int cu_count);
// the lane slots holding a parked packet (count zeroed before)
hipError_t launch_live_list(const uint8_t* lane_live, uint64_t lanes, uint32_t* out, unsigned* count, hipStream_t s);
// k_acc_place over the run's accretion photons k < m.acc_to
hipError_t launch_acc_place_fp64(const DeviceModel& m, uint64_t seed, uint64_t first_packet, AccStart* out, hipStream_t s);
hipError_t launch_acc_place_mixed(const DeviceModel& m, uint64_t seed, uint64_t first_packet, AccStart* out, hipStream_t s);
int lane_words_fp64();
int lane_words_mixed();
// resident 256-thread blocks of k_transport per CU with lds_bytes of dynamic LDS (0: query failed)
int transport_blocks_per_cu_fp64(const DeviceModel& m, unsigned lds_bytes);
int transport_blocks_per_cu_mixed(const DeviceModel& m, unsigned lds_bytes);
hipError_t launch_adjoint_fp64(const DeviceModel& m, const DeviceTallies& t, const AdjParams& ap, int grid, hipStream_t s);
hipError_t launch_adjoint_mixed(const DeviceModel& m, const DeviceTallies& t, const AdjParams& ap, int grid, hipStream_t s);
hipError_t launch_peel_fp64(const DeviceModel& m, const DeviceTallies& t, const RunParams& rp, int grid, hipStream_t s);
hipError_t launch_peel_mixed(const DeviceModel& m, const DeviceTallies& t, const RunParams& rp, int grid, hipStream_t s);
// the mixed path's peel-off walks of the round's items [item0, item1) as
// k_peel_start, k_peel_walk (grid_walk resident blocks, new ray states once
// `fetch` lanes of a wavefront are idle) and k_peel_hits; item1 - item0 must
// not exceed the capacity of rays and hits, and their counters must be zero
hipError_t launch_peel_split_mixed(const DeviceModel& m, const DeviceTallies& t, const RunParams& rp, const PeelRays& rays,
                                   const PeelHits& hits, uint64_t item0, uint64_t item1, int grid_start, int grid_walk,
                                   int grid_hits, unsigned fetch, hipStream_t s);
// resident 256-thread blocks of k_peel_walk per CU
int peel_walk_blocks_per_cu_mixed(unsigned lds_bytes);
size_t peel_record_bytes_fp64();
size_t peel_record_bytes_mixed();
hipError_t launch_trace_fp64(const DeviceModel& m, const TraceArgs& a, hipStream_t s);
hipError_t launch_trace_mixed(const DeviceModel& m, const TraceArgs& a, hipStream_t s);
hipError_t launch_mueller_fp64(int n, int kind, const double* par, double* dir, double* st, hipStream_t s);
hipError_t launch_mueller_mixed(int n, int kind, const double* par, double* dir, double* st, hipStream_t s);
hipError_t launch_batch_fp64(const BatchLaunch& L, hipStream_t s);
hipError_t launch_batch_mixed(const BatchLaunch& L, hipStream_t s);
hipError_t batch_occupancy_fp64(const BatchLaunch& L, int* blocks_per_cu);
hipError_t batch_occupancy_mixed(const BatchLaunch& L, int* blocks_per_cu);
// Lucy temperatures of every model (npk: packets per model; max_cells: cells of the largest model)
hipError_t launch_batch_tfinal(const BatchLaunch& L, const unsigned long long* npk, int mode, double* tdave, size_t max_cells,
                               hipStream_t s);
// setemiss tables: k_emiss_seg then k_emiss_scan (FP64, contraction off)
hipError_t launch_setemiss(const EmissArgs& a, hipStream_t s);
// model cell arrays: k_build_cells, k_build_mass, k_build_finish (FP64, contraction off)
hipError_t launch_build(const BuildArgs& a, hipStream_t s);
hipError_t launch_lucy_sigma(const DeviceModel& m, const double* told, const double* tnew, double* out, hipStream_t s);
struct DiffArgs;
hipError_t launch_setdiffus(const DeviceModel& m, const DiffArgs& a, hipStream_t s);
hipError_t launch_diff_reset(const DeviceModel& m, double* tdust, hipStream_t s);
hipError_t launch_diff_columns(const DeviceModel& m, const DiffArgs& a, const unsigned* count, const int* nabs,
                               double delta, double* tdust, hipStream_t s);
hipError_t launch_tfinal(const DeviceModel& m, const double* dtauabs, double nphot, int mode, double* tdave, hipStream_t s);
hipError_t launch_fill_f32(float* p, size_t n, float v, hipStream_t s);
hipError_t launch_fill_f64(double* p, size_t n, double v, hipStream_t s);
hipError_t launch_fill_i32(int* p, size_t n, int v, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long* p, size_t n, unsigned long long v, hipStream_t s);

}  // namespace mcrt
