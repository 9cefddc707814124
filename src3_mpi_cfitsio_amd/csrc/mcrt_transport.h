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
  unsigned* count;         // slots reserved (may exceed capacity: overflow runs inline)
  unsigned capacity;
};

struct RunParams {
  uint64_t first_packet;
  uint64_t npackets;
  uint64_t seed;
  const double* inczeta;   // 145 partial sums of zeta(4), random.f90:74-92
  PeelQueue q;             // q.capacity == 0: every peel-off walk runs inline
  uint64_t trace_base;     // index of this launch's first packet in the trace buffer
};

template <class T>
struct Ctx {
  const DeviceModel& m;
  const DeviceTallies& t;
  Rng rng;
  LocalCounts cnt;
  const double* inczeta;
  PeelQueue q;
};

template <class T> __device__ __forceinline__ void atomic_add_dtau(double* a, T v) { atomicAdd(a, (double)v); }

// One atomic per wavefront: the active lanes reserve consecutive slots.
__device__ __forceinline__ unsigned reserve_slot(unsigned* counter) {
  unsigned long long mask = __ballot(1);
  int lane = __lane_id();
  int leader = __ffsll((long long)mask) - 1;
  unsigned rank = (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
  unsigned base = 0;
  if (lane == leader) base = atomicAdd(counter, (unsigned)__popcll(mask));
  base = __shfl(base, leader, 64);
  return base + rank;
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
  p.ux = p.sint * p.cosp; p.uy = p.sint * p.sinp; p.uz = p.cost;
  p.nub = bbfreq<T>(C.rng, (T)C.m.tstar, C.inczeta);
}

// tauint_3d_2_fuv.f90:5-210.  Returns false on a walk error.
template <class T>
__device__ __forceinline__ bool tauint(Ctx<T>& C, Pkt<T>& p, T tau) {
  const DeviceModel& m = C.m;
  T tsum = (T)0;
  G dt;
  int iface;
  long long nseg = 0;
  for (;;) {
    if (testgrid<T>(m, p, C.rng)) C.cnt.c[C_RESETS]++;
    if (!next_face<T>(m, p, p.ux, p.uy, p.uz, dt, iface)) return false;
    if (++nseg > m.max_segments) return false;
    C.cnt.c[C_SEGMENTS]++;
    size_t c = cell_index(m, p.ir, p.it, p.ip);
    int iop;
    T rho;
    T opac = cell_opac<T>(m, p, c, iop, rho);
    T dtau = opac * (T)dt;
    if ((tsum + dtau) > tau) {
      T ds = (tau - tsum) / opac;
      if (iop >= 0) atomicAdd(&C.t.dtauabs[c], (double)ds * (double)pick(p.kapd, iop) * (double)rho * (1.0 - (double)pick(p.op, iop).albedo));
      p.xp = p.xp + (G)ds * p.ux;
      p.yp = p.yp + (G)ds * p.uy;
      p.zp = p.zp + (G)ds * p.uz;
      p.rsq = p.xp * p.xp + p.yp * p.yp + p.zp * p.zp;
      p.rtot = sqrt(p.rsq);
      p.on_wall = false;
      return true;
    }
    if (iop >= 0) atomicAdd(&C.t.dtauabs[c], (double)dt * (double)pick(p.kapd, iop) * (double)rho * (1.0 - (double)pick(p.op, iop).albedo));
    cross_face(m, p.ir, p.it, p.ip, iface);
    p.xp = p.xp + dt * p.ux;
    p.yp = p.yp + dt * p.uy;
    p.zp = p.zp + dt * p.uz;
    p.on_wall = true;
    p.rsq = p.xp * p.xp + p.yp * p.yp + p.zp * p.zp;
    p.rtot = sqrt(p.rsq);
    tsum = tsum + dtau;
    if (p.ir == m.nrg) { p.exitflag = 1; return true; }
    if (p.ir == 0) {
      C.cnt.c[C_AFLUX]++;
      p.rtot = sqrt(p.rsq);
      initpabs(C, p);
      opacset<T>(m, p.nub, p.op);
      p.wave = (T)1.2398 / p.nub;
      p.i_orig = 1;
      p.iflag = 0;
      set_kapd(p, m, 2);
      p.ir = 1;
      tsum = (T)0;
    }
  }
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

// peeloff_3d_2_test.f90:7-347.  `q` is the caller's packet (copied).
template <class T>
__device__ __forceinline__ void peeloff(Ctx<T>& C, const Pkt<T>& q, T sipnew, int peelid, T gsc, T plsc) {
  const DeviceModel& m = C.m;
  const DeviceTallies& t = C.t;
  Pkt<T> p = q;
  T s[4] = {sipnew, q.sqp, q.sup, q.svp};
  T tsum = (T)0;
  G dt;
  int iface;
  const int k = peelid - 1;
  const double* pe = m.peel + 5 * k;
  const T coste = (T)pe[0], sinte = (T)pe[1], cospe = (T)pe[2], sinpe = (T)pe[3], phie = (T)pe[4];
  const G ux1 = pe[1] * pe[2], uy1 = pe[1] * pe[3], uz1 = pe[0];
  const double* r2 = m.r2arr;
  if (p.on_wall) {
    bool inner = fabs(p.rsq - r2[p.ir - 1]) < fabs(p.rsq - r2[p.ir]);
    bool outwards = ux1 * p.xp + uy1 * p.yp + uz1 * p.zp > 0.0;
    if (inner && !outwards) { if (p.ir == 1) return; p.ir = p.ir - 1; }
    if (!inner && outwards) p.ir = p.ir + 1;
  }
  long long nseg = 0;
  for (;;) {
    if (testgrid<T>(m, p, C.rng)) C.cnt.c[C_RESETS]++;
    if (!next_face<T>(m, p, ux1, uy1, uz1, dt, iface) || ++nseg > m.max_segments) { C.cnt.c[C_ERRORS]++; return; }
    C.cnt.c[C_PEELSEG]++;
    size_t c = cell_index(m, p.ir, p.it, p.ip);
    int iop;
    T rho;
    T opac = cell_opac<T>(m, p, c, iop, rho);
    T dtau = opac * (T)dt;
    tsum = tsum + dtau;
    if (tsum > (T)m.taumax) return;
    cross_face(m, p.ir, p.it, p.ip, iface);
    p.xp = p.xp + dt * ux1;
    p.yp = p.yp + dt * uy1;
    p.zp = p.zp + dt * uz1;
    p.on_wall = true;
    p.rsq = p.xp * p.xp + p.yp * p.yp + p.zp * p.zp;
    p.rtot = sqrt(p.rsq);
    if (p.ir == m.nrg) break;
    if (p.ir == 0) return;
  }
  C.cnt.c[C_PEELEVT]++;
  T factor = RT<T>::exp_(-tsum);
  s[0] = s[0] * factor; s[1] = s[1] * factor; s[2] = s[2] * factor; s[3] = s[3] * factor;
  if (q.iflag == 1) {
    T bmu, sinbm, phinew, ri1;
    findangle3<T>(q.cost, q.sint, q.phi, q.cosp, q.sinp, coste, sinte, phie, cospe, sinpe, bmu, sinbm, phinew, ri1);
    stokespeel<T>(gsc, plsc, s, coste, sinte, ri1, bmu, q.cost, q.sint);
  }
  G yimage = p.zp * pe[1] - p.yp * pe[0] * pe[3] - p.xp * pe[0] * pe[2];
  G ximage = p.yp * pe[2] - p.xp * pe[3];
  G rmaxi = m.rmaxi;
  G fractxh = 2.0 * rmaxi / (G)m.nx;
  int ix = (int)((ximage + rmaxi) / fractxh) + 1;
  int iy = (int)((yimage + rmaxi) / fractxh) + 1;
  T freq = (T)2.9979e14 / q.wave;
  if (m.nband > 0 && ix <= m.nx && iy <= m.nx && ix > 0 && iy > 0) {
    size_t npix = (size_t)m.npeel * m.nx * m.nx;
    size_t plane = (size_t)m.nband * npix;
    for (int b = 0; b < m.nband; ++b) {
      const double* fw = m.band_freq + (size_t)b * m.nwmax;
      const double* fp = m.band_phi + (size_t)b * m.nwmax;
      int nf = m.band_nw[b] - 1;
      if (freq > (T)fw[0] && freq < (T)fw[nf - 1]) {
        int iw = 1;
        while (freq > (T)fw[iw - 1]) iw++;
        iw = iw - 1;
        T wt = (T)fp[iw - 1];
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
  G rho2 = ximage * ximage + yimage * yimage;
  int inub = freq_bin<T>(m, q.nub);
  int i_inter = (q.iflag == 1) ? 2 : ((q.i_orig == 1) ? 1 : 3);
  int indx[3] = {1, 1 + q.i_orig, 5 + i_inter};
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
// Peel-off events: queued for the peel kernel, or walked inline when the
// queue is full or disabled (same peeloff arithmetic either way).
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void write_rec(Ctx<T>& C, const Pkt<T>& p, unsigned slot, int sp, bool direct,
                                          T coslp, T sinlp) {
  PeelRec<T>* r = (PeelRec<T>*)C.q.recs + slot;
  r->xp = p.xp; r->yp = p.yp; r->zp = p.zp; r->rsq = p.rsq; r->rtot = p.rtot;
  r->sip = p.sip; r->sqp = p.sqp; r->sup = p.sup; r->svp = p.svp;
  if (direct) { r->cost = p.cosb; r->sint = p.sinb; r->phi = (T)0; r->cosp = coslp; r->sinp = sinlp; }
  else { r->cost = p.cost; r->sint = p.sint; r->phi = p.phi; r->cosp = p.cosp; r->sinp = p.sinp; }
  r->nub = p.nub; r->wave = p.wave;
  const Opac<T> o = pick(p.op, C.m.species_opac[sp]);
  r->g = o.g; r->pl = o.pl;
#pragma unroll
  for (int i = 0; i < kMaxOpac; ++i) r->kapd[i] = p.kapd[i];
  r->cell = (uint32_t)p.ir | ((uint32_t)p.it << 10) | ((uint32_t)p.ip << 20) | ((uint32_t)(p.on_wall ? 1 : 0) << 30);
  r->flags = (uint32_t)p.iflag | ((uint32_t)p.i_orig << 1) | ((uint32_t)(direct ? 1 : 0) << 4) | ((uint32_t)sp << 8);
}

// scattering / re-emission peel (propagate_5_fuv.f90:83-90, :162-170)
template <class T>
__device__ __forceinline__ void peel_event(Ctx<T>& C, const Pkt<T>& p, int sp) {
  if (C.q.capacity) {
    unsigned slot = reserve_slot(C.q.count);
    if (slot < C.q.capacity) { write_rec<T>(C, p, slot, sp, false, (T)0, (T)0); return; }
  }
  const Opac<T> o = pick(p.op, C.m.species_opac[sp]);
  for (int pid = 1; pid <= C.m.npeel; ++pid) peeloff<T>(C, p, p.sip, pid, o.g, o.pl);
}

// direct stellar peel, weight 4 I cos(norm) (newdisktherm_5_fuv.f90:371-414)
template <class T>
__device__ __forceinline__ void peel_direct(Ctx<T>& C, const Pkt<T>& p, T coslp, T sinlp) {
  if (C.q.capacity) {
    unsigned slot = reserve_slot(C.q.count);
    if (slot < C.q.capacity) { write_rec<T>(C, p, slot, 0, true, coslp, sinlp); return; }
  }
  for (int pk = 1; pk <= C.m.npeel; ++pk) {
    const double* pe = C.m.peel + 5 * (pk - 1);
    T cosnorm = p.cosb * (T)pe[0] + (p.sinb * (T)pe[1] * ((T)pe[2] * coslp + (T)pe[3] * sinlp));
    T sipnew = (T)4 * p.sip * cosnorm;
    if (cosnorm > (T)0) peeloff<T>(C, p, sipnew, pk, (T)0, (T)0);
  }
}

// stokes.f90:46-137 rejection sampling, then the Mueller update
template <class T>
__device__ __forceinline__ void stokes(Ctx<T>& C, Pkt<T>& p, int iop) {
  const T PI = (T)kPi, R2P = (T)kR2p, one = (T)1, two = (T)2;
  const Opac<T> o = pick(p.op, iop);
  T g = o.g, pl = o.pl;
  T peak = cs_peak<T>(g), bmu, ri1, rprob, pp[4];
  for (;;) {
    C.cnt.c[C_HTOT]++;
    bmu = one - two * uniform<T>(C.rng);
    T cosb2 = bmu * bmu;
    dustmat<T>(g, pl, bmu, cosb2, pp);
    ri1 = R2P * uniform<T>(C.rng);
    if (ri1 > PI) {
      T ri3 = R2P - ri1, cosi3 = RT<T>::cos_(ri3), sini3 = RT<T>::sin_(ri3);
      T sin2i3 = two * sini3 * cosi3, cos2i3 = two * (cosi3 * cosi3) - one;
      rprob = (pp[0] * p.sip + pp[1] * cos2i3 * p.sqp + pp[1] * sin2i3 * p.sup) / p.sip;
    } else {
      T cosi1 = RT<T>::cos_(ri1), sini1 = RT<T>::sin_(ri1);
      T sin2i1 = two * sini1 * cosi1, cos2i1 = two * (cosi1 * cosi1) - one;
      rprob = (pp[0] * p.sip + pp[1] * cos2i1 * p.sqp + (-pp[1] * sin2i1) * p.sup) / p.sip;
    }
    if (rprob > peak) peak = rprob;
    T xran = uniform<T>(C.rng);
    if (peak * xran > rprob) continue;
    C.cnt.c[C_HIT]++;
    break;
  }
  T dir[5] = {p.cost, p.sint, p.phi, p.cosp, p.sinp};
  T s[4] = {p.sip, p.sqp, p.sup, p.svp};
  stokes_apply<T>(g, pl, bmu, ri1, dir, s);
  p.cost = dir[0]; p.sint = dir[1]; p.phi = dir[2]; p.cosp = dir[3]; p.sinp = dir[4];
  p.sip = s[0]; p.sqp = s[1]; p.sup = s[2]; p.svp = s[3];
}

// propagate_5_fuv.f90:1-326.  Returns 1 escaped, 2 killed, -1 error.
template <class T>
__device__ __forceinline__ int propagate(Ctx<T>& C, Pkt<T>& p, int& nint) {
  const DeviceModel& m = C.m;
  const DeviceTallies& t = C.t;
  const T R2P = (T)kR2p;
  G xpold = p.xp, ypold = p.yp, zpold = p.zp;
  int iscat = 0, iint = 0, idust2;
  nint = 0;
  p.iflag = 0; p.exitflag = 0;
  T tau = -RT<T>::log_(uniform<T>(C.rng));
  if (!tauint(C, p, tau)) return -1;
  if (!p.exitflag) {
    (void)uniform<T>(C.rng);   // select_dust_2.f90:134 draw
    idust2 = m.species[cell_index(m, p.ir, p.it, p.ip)];
    C.cnt.c[C_TOT]++;
    for (;;) {
      size_t c = cell_index(m, p.ir, p.it, p.ip);
      int iop = m.species_opac[idust2];
      nint++;
      if (uniform<T>(C.rng) <= pick(p.op, iop).albedo) {
        p.iflag = 1; iscat++; iint++;
        if (m.ipeel == 1) peel_event<T>(C, p, idust2);
        stokes<T>(C, p, iop);
      } else {
        atomicAdd(&t.dbig[c], 1);
        p.i_orig = i_orig_of(idust2);
        atomicAdd(&t.nabs[c], 1);
        T tave = (T)m.tdust[c];
        p.sip = (T)1; p.sqp = (T)0; p.sup = (T)0; p.svp = (T)0;
        p.cost = (T)1 - (T)2 * uniform<T>(C.rng);
        p.sint = RT<T>::sqrt_((T)1 - p.cost * p.cost);
        p.phi = R2P * uniform<T>(C.rng);
        p.cosp = RT<T>::cos_(p.phi);
        p.sinp = RT<T>::sin_(p.phi);
        if (tave >= (T)(kTmaxK / 11605.0)) p.nub = bbfreq<T>(C.rng, tave, C.inczeta);
        else { T xi = uniform<T>(C.rng); p.nub = lucydustfreq_xi<T>(m, tave, iop, xi); }
        p.uy = p.sint * p.sinp; p.ux = p.sint * p.cosp; p.uz = p.cost;
        p.iflag = 0;
        opacset<T>(m, p.nub, p.op);
        p.wave = (T)1.2398 / p.nub;
        set_kapd(p, m, 1);
        if (m.ipeel == 1) peel_event<T>(C, p, idust2);
        iint++;
      }
      tau = -RT<T>::log_(uniform<T>(C.rng));
      xpold = p.xp; ypold = p.yp; zpold = p.zp;
      p.ux = p.sint * p.cosp; p.uy = p.sint * p.sinp; p.uz = p.cost;
      int pir = p.ir, pit = p.it, pip = p.ip;
      if (!tauint(C, p, tau)) return -1;
      if (p.exitflag) break;
      if (p.ir != pir || p.it != pit || p.ip != pip) iint = 0;
      if (iint > m.iintmax) {
        atomicAdd(&t.killed[cell_index(m, p.ir, p.it, p.ip)], 1);
        C.cnt.c[C_ABFLUX]++;
        return 2;
      }
      (void)uniform<T>(C.rng);
      idust2 = m.species[cell_index(m, p.ir, p.it, p.ip)];
    }
  }
  // SED binning, propagate_5_fuv.f90:231-320
  int itb = (int)((T)m.nmu * ((T)1 - p.cost) / (T)2) + 1;
  if (p.phi < (T)0) p.phi = p.phi + R2P;
  int ipb = (int)((T)m.nph * p.phi / R2P) + 1;
  if (itb < 1) itb = 1;
  if (itb > m.nmu) itb = m.nmu;
  if (ipb < 1) ipb = 1;
  if (ipb > m.nph) ipb = m.nph;
  G x = ypold * (G)p.cosp - xpold * (G)p.sinp;
  G y = zpold * m.sini_mu[itb - 1] - ypold * m.u_mu[itb - 1] * (G)p.sinp - xpold * m.u_mu[itb - 1] * (G)p.cosp;
  G rho2 = (x * x + y * y);
  if (rho2 < m.aperture2[m.nap - 1] * 1.0001) {
    C.cnt.c[C_FLUX]++;
    if (p.iflag == 1) { C.cnt.c[C_SFLUX]++; C.cnt.c[C_NSCAT] += (unsigned long long)iscat; }
  }
  int inub = freq_bin<T>(m, p.nub);
  int i_inter = (p.iflag == 1) ? 2 : ((p.i_orig == 1) ? 1 : 3);
  int indx[3] = {1, 1 + p.i_orig, 5 + i_inter};
  size_t nb = (size_t)m.nfreq * m.nmu * m.nph * m.nap * kNo;
  for (int ia = 1; ia <= m.nap; ++ia) {
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
  if (p.iflag == 1) C.cnt.c[C_SCOUNT]++;
  return 1;
}

// newdisktherm_5_fuv.f90:294-413 one stellar photon (ips = 0, no spots)
template <class T>
__device__ __forceinline__ void one_packet(Ctx<T>& C, uint64_t pid) {
  const DeviceModel& m = C.m;
  Pkt<T> p;
  p.on_wall = false;
  p.sip = (T)1; p.sqp = (T)0; p.sup = (T)0; p.svp = (T)0;
  for (;;) {
    p.cosb = (T)1 - (T)2 * uniform<T>(C.rng);
    p.sinb = RT<T>::sqrt_((T)1 - p.cosb * p.cosb);
    if (p.cosb != (T)0 && RT<T>::fabs_(p.cosb) < (T)1) break;
  }
  p.lp = (T)kR2p * uniform<T>(C.rng);
  initp_sample(C, p);
  initp_transform(p);
  p.nub = bbfreq<T>(C.rng, (T)m.tstar, C.inczeta);
  p.i_orig = 1;
  p.iflag = 0;
  p.exitflag = 0;
  opacset<T>(m, p.nub, p.op);
  p.wave = (T)1.2398 / p.nub;
  set_kapd(p, m, 0);
  T coslp = RT<T>::cos_(p.lp), sinlp = RT<T>::sin_(p.lp);
  p.zp = (G)p.cosb * m.rmin;
  G rp = (G)p.sinb * m.rmin;
  p.xp = rp * (G)coslp;
  p.yp = rp * (G)sinlp;
  p.rsq = p.zp * p.zp + rp * rp;
  p.rtot = sqrt(p.rsq);
  p.ux = p.sint * p.cosp; p.uy = p.sint * p.sinp; p.uz = p.cost;
  p.ir = 1;
  p.on_wall = true;
  p.it = locate<double>(m.thetarr, m.ntg, (double)RT<T>::acos_(p.cosb));
  p.ip = locate<double>(m.phiarr, m.npg, (double)p.lp);
  if (m.ipeel == 1) peel_direct<T>(C, p, coslp, sinlp);
  int nint = 0;
  int code = propagate<T>(C, p, nint);
  if (code == 1) C.cnt.c[C_ESCAPED]++;
  else if (code == 2) C.cnt.c[C_KILLED]++;
  else C.cnt.c[C_ERRORS]++;
  if (C.t.trace) {
    double* tr = C.t.trace + 8 * pid;
    tr[0] = (double)p.nub; tr[1] = (double)p.cost; tr[2] = (double)p.phi; tr[3] = (double)p.sip;
    tr[4] = (double)p.sqp; tr[5] = (double)p.sup; tr[6] = (double)nint; tr[7] = (double)code;
  }
}

__device__ __forceinline__ void flush_counts(const DeviceTallies& t, LocalCounts& cnt) {
  // wavefront reduction then one atomic per counter per wavefront
#pragma unroll
  for (int i = 0; i < C_NCOUNTERS; ++i) {
    unsigned long long v = cnt.c[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(&t.counters[i], v);
  }
}

// Transport kernel: one packet per thread (grid-stride).  Peel-off events go
// to the queue; k_peel walks them.
template <class T>
__global__ void __launch_bounds__(256) k_packets(DeviceModel m, DeviceTallies t, RunParams rp) {
  Ctx<T> C{m, t, {}, {}, rp.inczeta, rp.q};
  C.cnt.zero();
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < rp.npackets; i += stride) {
    rng_init(C.rng, rp.seed, rp.first_packet + i);
    one_packet<T>(C, rp.trace_base + i);
  }
  flush_counts(t, C.cnt);
}

// Peel kernel: one (event, direction) pair per thread.  Lanes of a
// wavefront run the same face-walk loop from queued events, so the walk that
// dominates the path (npeel walks per interaction) runs without the
// transport kernel's emission / sampling divergence and register load.
template <class T>
__global__ void __launch_bounds__(256) k_peel(DeviceModel m, DeviceTallies t, RunParams rp) {
  Ctx<T> C{m, t, {}, {}, rp.inczeta, PeelQueue{nullptr, nullptr, 0}};
  C.cnt.zero();
  const unsigned nq = *rp.q.count;
  const uint64_t nrec = nq < rp.q.capacity ? nq : rp.q.capacity;
  const uint64_t total = nrec * (uint64_t)m.npeel;
  const PeelRec<T>* recs = (const PeelRec<T>*)rp.q.recs;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    const uint64_t ri = i / (uint64_t)m.npeel;
    const int pk = (int)(i - ri * (uint64_t)m.npeel) + 1;
    const PeelRec<T>& r = recs[ri];
    Pkt<T> q;
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
    const bool direct = ((r.flags >> 4) & 1u) != 0;
    const int sp = (int)((r.flags >> 8) & 255u);
    rng_init(C.rng, rp.seed ^ 0x5045454C00000000ull, rp.first_packet * 64u + i);   // testgrid resets only
    T sipnew = q.sip;
    if (direct) {
      const double* pe = m.peel + 5 * (pk - 1);
      T cosnorm = q.cost * (T)pe[0] + (q.sint * (T)pe[1] * ((T)pe[2] * q.cosp + (T)pe[3] * q.sinp));
      if (!(cosnorm > (T)0)) continue;
      sipnew = (T)4 * q.sip * cosnorm;
      q.cosb = q.cost; q.sinb = q.sint;
    }
    peeloff<T>(C, q, sipnew, pk, r.g, r.pl);
  }
  flush_counts(t, C.cnt);
}

}  // namespace mcrt
