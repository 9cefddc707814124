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

// Arguments of the diffusion kernels.
struct DiffArgs {
  const double* kappar;      // [nopac * nT] Rosseland means (setemiss)
  const double* hdisk;       // [nrg - 1] disk scale height (R*) per radial cell
  const double* tdave;       // temperatures of the last Lucy update (K/11605)
  double rmaxd;              // disk outer radius (R*)
  int nr0;                   // first radial cell that may diffuse
  int iter;                  // Lucy iteration (below 2: temperatures floored at 50 K)
  int32_t* itmin;            // [nrg - 1] out: first diffusion cell from the pole, 0 none
};

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
  if (sp > 0 && m.rho64[c] > 0.0 && diffusing(m, ir, it)) {
    // a diffusion cell keeps the temperatures the run's diffusion photons
    // left (tfinal_5_test1.f90:232-233)
    t = (m.tdust[c] + m.tdust[cm]) / 2.0;
    if (t * 11605.0 >= kTmaxK) t = kTmaxK / 11605.0;
  } else if (sp > 0 && m.rho64[c] > 0.0) {
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
  const int nr = m.nrg - 1, nt = m.ntg - 1;
  if (sp > 0 && m.rho64[c] > 0.0 && diffusing(m, (int)(c % nr) + 1, (int)((c / nr) % nt) + 1)) {
    t = m.tdust[c];
  } else if (sp > 0 && m.rho64[c] > 0.0) {
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

// chiR (dust_5.f90:112-141): the Rosseland mean of table iop at T,
// linear in log T between the nT table temperatures (clamped to the table)
__device__ __forceinline__ double chiR_dev(const DeviceModel& m, const double* kappar, double T, int iop) {
  const int nT = m.nT;
  const double lT = log(T);
  int k = (int)((double)(nT - 1) * (lT - m.lTint[0]) / (m.lTint[nT - 1] - m.lTint[0])) + 1;
  if (k < 1) k = 1;
  if (k > nT - 1) k = nT - 1;
  double f = (lT - m.lTint[k - 1]) / (m.lTint[k] - m.lTint[k - 1]);
  const double* kr = kappar + (size_t)iop * nT;
  return (1.0 - f) * kr[k - 1] + f * kr[k];
}


// setdiffus_5.f90:36-103 for cell it of column ir (one thread each, it in
// 2..(ntg-1)/2): the Rosseland optical depth from the cell toward the pole
// and the mean free path at depth 1; the column diffuses from it to the
// midplane when the depth reaches 1 within less than a twentieth of the
// disk scale height.  The column's diffusion region starts at the smallest
// such it (atomicMin: the same result in any order).
__global__ void __launch_bounds__(256) k_setdiffus(DeviceModel m, DiffArgs a) {
  const int nr = m.nrg - 1, nh = (m.ntg - 1) / 2;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nr * (nh - 1)) return;
  const int ir = idx % nr + 1, it = idx / nr + 2;
  const double rad = m.rarr[ir - 1];
  if (ir < a.nr0 || !(rad < a.rmaxd)) return;
  const double kapd_cm = m.rstar * kRsol;           // kapd = kappav * rmincgs (gridset_5_model13new.f90:79)
  double tau = 0.0, tau1 = 0.0;
  int itt = it;
  while (tau < 1.0 && itt > 1) {
    const double dt = m.thetarr[itt] - m.thetarr[itt - 1];
    tau1 = tau;
    const size_t c = cell_index(m, ir, itt, 1);
    double tt = a.tdave[c];
    if (a.iter < 2) tt = fmax(tt, 50.0 / 11605.0);
    const int sp = m.species[c];
    if (sp > 0 && m.rho64[c] > 0.0) {
      const int iop = m.species_opac[sp];
      tau = tau + m.kappav[iop] * kapd_cm * chiR_dev(m, a.kappar, tt, iop) * m.rho64[c] * dt * rad;
    }
    itt = itt - 1;
  }
  double lmfp1;
  if (tau > 1.0) lmfp1 = rad * (m.thetarr[itt + 1] - m.thetarr[itt]) / (tau - tau1) * (1.0 - tau1);
  else lmfp1 = rad * (m.thetarr[itt + 1] - m.thetarr[itt]);
  const double lmfp = rad * (m.thetarr[it] - m.thetarr[itt + 1]) + lmfp1;
  if (tau >= 1.0 && lmfp < a.hdisk[ir - 1] / 20.0) atomicMin(&a.itmin[ir - 1], it);
}

__global__ void __launch_bounds__(256) k_itmin_none(int32_t* itmin, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && itmin[i] == 0x7fffffff) itmin[i] = 0;
}

// gridset_5_model13new.f90:1857-1868: a diffusion cell starts the run at the
// 0.1 K floor; its temperature is then built by the run's diffusion photons.
__global__ void __launch_bounds__(256) k_diff_reset(DeviceModel m, double* tdust) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nr = m.nrg - 1, nt = m.ntg - 1;
  if (c >= (size_t)nr * nt * (m.npg - 1)) return;
  if (diffusing(m, (int)(c % nr) + 1, (int)((c / nr) % nt) + 1)) tdust[c] = 0.1 / 11605.0;
}

// The temperatures the reference's diffusion photons leave in a column
// (emit_common_2.f90:256-341) after the run, one thread per (ir, ip,
// hemisphere): the surface cell's T^4 = T0^4 + N Delta / dA for its N
// photons, then the interior integrated toward the midplane with the
// accretion photons absorbed below (nabs; none without accretion, when the
// whole region takes the surface temperature).
__global__ void __launch_bounds__(256) k_diff_columns(DeviceModel m, const DiffArgs a, const unsigned* count,
                                                      const int* nabs, double delta, double* tdust) {
  const int nr = m.nrg - 1, np = m.npg - 1;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nr * np * 2) return;
  const int ir = idx % nr + 1, ip = (idx / nr) % np + 1;
  const bool upper = idx / (nr * np) == 0;
  const int am = m.diff_itmin[ir - 1];
  if (am <= 0) return;
  const int ntg = m.ntg, mid = upper ? (ntg - 1) / 2 : (ntg + 1) / 2;
  const int its = upper ? am : ntg - am, step = upper ? 1 : -1;
  const int w = upper ? its : its + 1;
  const double sth = sin(m.thetarr[w - 1]);
  const double dA = kPi * sth * (m.r2arr[ir] - m.r2arr[ir - 1]);
  const double Tcap = kTmaxK / 11605.0, T0 = 0.1 / 11605.0;
  const size_t cs = cell_index(m, ir, its, ip);
  const double T4 = pow4d(T0) + (double)count[cs] * delta / dA;
  double Ts = pow(T4, 0.25);
  if (Ts > Tcap) Ts = Tcap;
  if (!(m.species[cs] > 0 && m.rho64[cs] > 0.0)) Ts = T0;
  tdust[cs] = Ts;
  double nacc = 0.0;
  for (int i = its; upper ? i <= mid : i >= mid; i += step) nacc += (double)nabs[cell_index(m, ir, i, ip)];
  double T4d = T4;
  const double rc = 0.5 * (m.rarr[ir - 1] + m.rarr[ir]);
  for (int i = its + step; upper ? i <= mid : i >= mid; i += step) {
    const int ib = i - step;                                   // the cell above toward the surface
    const size_t cb = cell_index(m, ir, ib, ip);
    const int sp = m.species[cb];
    if (sp > 0 && m.rho64[cb] > 0.0) {
      const int iop = m.species_opac[sp];
      const double dth = upper ? m.thetarr[i - 1] - m.thetarr[i - 2] : m.thetarr[i + 1] - m.thetarr[i];
      T4d = T4d + 0.75 * delta / dA * nacc * m.kappav[iop] * kRsol * m.rstar * m.rho64[cb] *
                      chiR_dev(m, a.kappar, tdust[cb], iop) * rc * dth;
      double Ti = pow(T4d, 0.25);
      tdust[cell_index(m, ir, i, ip)] = Ti > Tcap ? Tcap : Ti;
      nacc = nacc - (double)nabs[cb];
    }
  }
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
