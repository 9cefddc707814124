// mcrt_kernels.h -- kernels around the packet transport: deterministic
// fixed-direction ray walks, deterministic Mueller updates, the Lucy
// temperature inversion, and fill kernels.  Instantiated per precision by
// kernels_fp64.hip and kernels_mixed.hip.
#pragma once
#include "mcrt_transport.h"
#include "mcrt_adjoint.h"
#include "mcrt_batch.h"

namespace mcrt {

struct TraceArgs {
  int n, max_steps;
  const double *pos, *u, *tau_stop;
  const int* ijk;
  const uint8_t* on_wall;
  int *faces, *cells, *ijk_out, *nsteps, *status;
  double *pos_out, *tau_out;
};

// Fixed-direction walk (tauint's face loop with unit kapd, no RNG tallies).
// status: 0 stopped inside, 1 escaped, 2 hit the star, 3 max steps, -1 no wall.
template <class T>
__global__ void __launch_bounds__(256) k_trace(DeviceModel m, TraceArgs a) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.n) return;
  Pkt<T> p;
  p.xp = a.pos[3 * r]; p.yp = a.pos[3 * r + 1]; p.zp = a.pos[3 * r + 2];
  p.rsq = p.xp * p.xp + p.yp * p.yp + p.zp * p.zp;
  p.rtot = sqrt(p.rsq);
  p.ux = a.u[3 * r]; p.uy = a.u[3 * r + 1]; p.uz = a.u[3 * r + 2];
  p.ir = a.ijk[3 * r]; p.it = a.ijk[3 * r + 1]; p.ip = a.ijk[3 * r + 2];
  p.on_wall = a.on_wall[r] != 0;
#pragma unroll
  for (int i = 0; i < kMaxOpac; ++i) p.kapd[i] = (T)1;
  Rng rng;
  rng_init(rng, 0, (uint64_t)r);      // testgrid re-placements: stream = ray index
  T tau_stop = (T)a.tau_stop[r], tsum = (T)0;
  G dt;
  int iface, n = 0, st = 3;
  while (n < a.max_steps) {
    testgrid<T>(m, p, rng);
    if (!next_face<T>(m, p, p.ux, p.uy, p.uz, dt, iface)) { st = -1; break; }
    size_t c = cell_index(m, p.ir, p.it, p.ip);
    int iop;
    T rho;
    T opac = cell_opac<T>(m, p, c, iop, rho);
    T dtau = opac * (T)dt;
    if ((tsum + dtau) > tau_stop) {
      T ds = (tau_stop - tsum) / opac;
      p.xp = p.xp + (G)ds * p.ux; p.yp = p.yp + (G)ds * p.uy; p.zp = p.zp + (G)ds * p.uz;
      tsum = tau_stop;
      st = 0;
      break;
    }
    cross_face(m, p.ir, p.it, p.ip, iface);
    p.xp = p.xp + dt * p.ux; p.yp = p.yp + dt * p.uy; p.zp = p.zp + dt * p.uz;
    p.on_wall = true;
    p.rsq = p.xp * p.xp + p.yp * p.yp + p.zp * p.zp;
    p.rtot = sqrt(p.rsq);
    tsum = tsum + dtau;
    size_t o = (size_t)r * a.max_steps + n;
    a.faces[o] = iface;
    a.cells[3 * o] = p.ir; a.cells[3 * o + 1] = p.it; a.cells[3 * o + 2] = p.ip;
    n++;
    if (p.ir == m.nrg) { st = 1; break; }
    if (p.ir == 0) { st = 2; break; }
  }
  a.pos_out[3 * r] = (double)p.xp; a.pos_out[3 * r + 1] = (double)p.yp; a.pos_out[3 * r + 2] = (double)p.zp;
  a.ijk_out[3 * r] = p.ir; a.ijk_out[3 * r + 1] = p.it; a.ijk_out[3 * r + 2] = p.ip;
  a.tau_out[r] = (double)tsum;
  a.nsteps[r] = n;
  a.status[r] = st;
}

// Deterministic Mueller updates.  kind 0: stokes.f90 update for accepted
// (bmu, ri1); kind 1: stokespeel.f90.  par per item (9 doubles):
// g, pl, bmu, ri1, cost, sint, phi, costp, sintp.  dir per item (5): cost,
// sint, phi, cosp, sinp (kind 0 in/out).  s per item (4) in/out.
template <class T>
__global__ void __launch_bounds__(256) k_mueller(int n, int kind, const double* par, double* dir, double* s) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const double* q = par + 9 * r;
  T st[4] = {(T)s[4 * r], (T)s[4 * r + 1], (T)s[4 * r + 2], (T)s[4 * r + 3]};
  if (kind == 0) {
    T d[5] = {(T)dir[5 * r], (T)dir[5 * r + 1], (T)dir[5 * r + 2], (T)dir[5 * r + 3], (T)dir[5 * r + 4]};
    stokes_apply<T>((T)q[0], (T)q[1], (T)q[2], (T)q[3], d, st);
    for (int i = 0; i < 5; ++i) dir[5 * r + i] = (double)d[i];
  } else {
    stokespeel<T>((T)q[0], (T)q[1], st, (T)q[4], (T)q[5], (T)q[3], (T)q[2], (T)q[7], (T)q[8]);
  }
  for (int i = 0; i < 4; ++i) s[4 * r + i] = (double)st[i];
}

#ifdef MCRT_FP64_ONLY_KERNELS
// ---------------------------------------------------------------------------
// Lucy temperature (FP64): dusttemp_5.f90:8-152 with zero velocities,
// tfinal_5_test1.f90:206-281 (mode 0) or every cell independently (mode 1).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double pow4d(double x) { double x2 = x * x; return x2 * x2; }

__device__ double dusttemp_dev(const DeviceModel& m, double na, double nphot, int iop, double mcell) {
  const double Tfloor = 0.1 / 11605.0, Tcap = kTmaxK / 11605.0;
  const int nT = m.nT;
  const double* kp = m.kappap + (size_t)iop * nT;
  const double* Ti = m.Tint;
  double Td = Tfloor;
  if (!(mcell > 0.0)) return Tfloor;
  double absv = pow4d(m.tstar) * (na * kPi * m.rstar * m.rstar) / (nphot * m.kappav[iop]);
  absv = absv * (kRsol * kRsol) / pow4d(mcell);
  if (na == 0.0) return Tfloor;
  double minkap = pow4d(Ti[0]) * kp[0], maxkap = minkap;
  int mini = 1, maxi = 1;
  for (int i = 1; i <= nT; ++i) {
    double kt = pow4d(Ti[i - 1]) * kp[i - 1];
    if (minkap > kt) { minkap = kt; mini = i; }
    if (maxkap < kt) { maxkap = kt; maxi = i; }
  }
  if (maxkap < absv) { Td = Ti[maxi - 1]; if (Td > Tcap) Td = Tcap; if (Td < Tfloor) Td = Tfloor; }
  if (minkap > absv) { Td = Ti[mini - 1]; if (Td > Tcap) Td = Tcap; if (Td < Tfloor) Td = Tfloor; }
  int lo = (mini < maxi) ? mini : maxi, hi = (mini < maxi) ? maxi : mini;
  bool up = mini < maxi;
  for (int i = lo; i <= hi - 1; ++i) {
    double kl = pow4d(Ti[i - 1]) * kp[i - 1], ku = pow4d(Ti[i]) * kp[i];
    bool hitv = up ? (kl < absv && ku >= absv) : (kl > absv && ku <= absv);
    if (hitv) {
      double Tl = Ti[i - 1], Tu = Ti[i];
      Td = Tl + (Tu - Tl) / (ku - kl) * (absv - kl);
      if (Td > Tcap) Td = Tcap;
      if (Td < Tfloor) Td = Tfloor;
      break;
    }
  }
  return Td;
}

// mode 0, step 1: upper hemisphere at ip = 1, averaged with the mirror cell;
// idx runs over (nrg-1) * ((ntg-1)/2) source cells
__device__ __forceinline__ void tfinal_half_at(const DeviceModel& m, const double* dtauabs, double nphot, double* tdave, int idx) {
  const int nr = m.nrg - 1;
  int ir = idx % nr + 1, it = idx / nr + 1;
  size_t c = cell_index(m, ir, it, 1), cm = cell_index(m, ir, m.ntg - it, 1);
  double t = 0.1 / 11605.0;
  int sp = m.species[c];
  if (sp > 0 && m.rho64[c] > 0.0) {
    int iop = m.species_opac[sp];
    double rho2 = (m.species[cm] == sp) ? m.rho64[cm] : 0.0;
    double mcell1 = pow(m.rho64[c] * m.vcell[c], 0.25);
    double mcell2 = pow(rho2 * m.vcell[c], 0.25);
    double t1 = dusttemp_dev(m, dtauabs[c], nphot, iop, mcell1);
    double t2 = dusttemp_dev(m, dtauabs[cm], nphot, iop, mcell2);
    t = (t1 + t2) / 2.0;
    if (t * 11605.0 >= kTmaxK) t = kTmaxK / 11605.0;
  }
  tdave[c] = t;
}

// mode 0, step 2: mirror to the lower hemisphere and replicate in phi
__device__ __forceinline__ void tfinal_fill_at(const DeviceModel& m, double* tdave, size_t idx) {
  const int nr = m.nrg - 1, nt = m.ntg - 1, nh = (m.ntg - 1) / 2;
  int ir = (int)(idx % nr) + 1;
  int it = (int)((idx / nr) % nt) + 1;
  int ip = (int)(idx / ((size_t)nr * nt)) + 1;
  if (ip == 1 && it <= nh) return;                 // source cells
  int its = (it <= nh) ? it : m.ntg - it;          // tfinal_5_test1.f90:267-271
  tdave[idx] = tdave[cell_index(m, ir, its, 1)];
}

// mode 1: every cell from its own absorbed energy (no mirror, no phi copy)
__device__ __forceinline__ void tcell_at(const DeviceModel& m, const double* dtauabs, double nphot, double* tdave, size_t c) {
  double t = 0.1 / 11605.0;
  int sp = m.species[c];
  if (sp > 0 && m.rho64[c] > 0.0) {
    int iop = m.species_opac[sp];
    t = dusttemp_dev(m, dtauabs[c], nphot, iop, pow(m.rho64[c] * m.vcell[c], 0.25));
    if (t * 11605.0 >= kTmaxK) t = kTmaxK / 11605.0;
  }
  tdave[c] = t;
}

__global__ void __launch_bounds__(256) k_tfinal_half(DeviceModel m, const double* dtauabs, double nphot, double* tdave) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (m.nrg - 1) * ((m.ntg - 1) / 2)) return;
  tfinal_half_at(m, dtauabs, nphot, tdave, idx);
}

__global__ void __launch_bounds__(256) k_tfinal_fill(DeviceModel m, double* tdave) {
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)(m.nrg - 1) * (m.ntg - 1) * (m.npg - 1)) return;
  tfinal_fill_at(m, tdave, idx);
}

__global__ void __launch_bounds__(256) k_tcell(DeviceModel m, const double* dtauabs, double nphot, double* tdave) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (size_t)(m.nrg - 1) * (m.ntg - 1) * (m.npg - 1)) return;
  tcell_at(m, dtauabs, nphot, tdave, c);
}

// Convergence statistic of tfinal (tfinal_5_test1.f90:334-368): over the
// upper hemisphere at ip = 1, the cells whose old and new temperatures both
// exceed the 0.1 K floor and differ contribute x = (tnew - told) / tnew;
// out[0] = sum |x|, out[1] = sum x^2, out[2] = their count.  One block:
// each thread sums a fixed strided subset in index order, then a fixed tree
// in LDS, so the sums are the same bits on every run.
__global__ void __launch_bounds__(256) k_lucy_sigma(DeviceModel m, const double* told, const double* tnew, double* out) {
  __shared__ double s0[256], s1[256], s2[256];
  const double floor_t = 0.1 / 11605.0;
  const int nr = m.nrg - 1, nh = (m.ntg - 1) / 2;
  const int n = nr * nh;
  double a = 0.0, b = 0.0, k = 0.0;
  for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
    const int ir = idx % nr + 1, it = idx / nr + 1;
    const size_t c = cell_index(m, ir, it, 1);
    const double to = told[c], tn = tnew[c];
    if (tn != to && tn > floor_t && to > floor_t) {
      const double x = (tn - to) / tn;
      a = fabs(x) + a;
      b = x * x + b;
      k = k + 1.0;
    }
  }
  s0[threadIdx.x] = a; s1[threadIdx.x] = b; s2[threadIdx.x] = k;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s0[threadIdx.x] += s0[threadIdx.x + w];
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] += s2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out[0] = s0[0]; out[1] = s1[0]; out[2] = s2[0]; }
}

// Lucy temperatures of every model of a batch: blockIdx.y = model, with the
// per-model packet count as nphot.  mode 0: tfinal's upper hemisphere and
// mirror; mode 1: every cell on its own.
__device__ __forceinline__ void bind_cells(DeviceModel& m, const BatchDesc& b, int k) {
  const long long ro = b.roff[k], to = b.toff[k], co = b.coff[k];
  m.nrg = b.nrg[k];
  m.ntg = b.ntg[k];
  m.rarr = b.rarr + ro;
  m.r2arr = b.r2arr + ro;
  m.thetarr = b.thetarr + to;
  m.costarr = b.costarr + to;
  m.tan2arr = b.tan2arr + to;
  m.rho64 = b.rho64 + co;
  m.rho32 = b.rho32 + co;
  m.species = b.species + co;
  m.vcell = b.vcell + co;
}

__global__ void __launch_bounds__(256) k_batch_tfinal_half(DeviceModel m0, BatchDesc bd, const unsigned long long* npk,
                                                           const double* dtauabs, double* tdave) {
  const int k = blockIdx.y;
  DeviceModel m = m0;
  bind_cells(m, bd, k);
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (m.nrg - 1) * ((m.ntg - 1) / 2) || npk[k] == 0) return;
  const long long co = bd.coff[k];
  tfinal_half_at(m, dtauabs + co, (double)npk[k], tdave + co, idx);
}

__global__ void __launch_bounds__(256) k_batch_tfinal_fill(DeviceModel m0, BatchDesc bd, const unsigned long long* npk,
                                                           double* tdave) {
  const int k = blockIdx.y;
  DeviceModel m = m0;
  bind_cells(m, bd, k);
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)(m.nrg - 1) * (m.ntg - 1) * (m.npg - 1) || npk[k] == 0) return;
  tfinal_fill_at(m, tdave + bd.coff[k], idx);
}

__global__ void __launch_bounds__(256) k_batch_tcell(DeviceModel m0, BatchDesc bd, const unsigned long long* npk,
                                                     const double* dtauabs, double* tdave) {
  const int k = blockIdx.y;
  DeviceModel m = m0;
  bind_cells(m, bd, k);
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (size_t)(m.nrg - 1) * (m.ntg - 1) * (m.npg - 1) || npk[k] == 0) return;
  const long long co = bd.coff[k];
  tcell_at(m, dtauabs + co, (double)npk[k], tdave + co, c);
}

template <class V>
__global__ void __launch_bounds__(256) k_fill(V* p, size_t n, V v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

#endif  // MCRT_FP64_ONLY_KERNELS

}  // namespace mcrt
