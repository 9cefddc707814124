// mcrt_emiss.h -- Planck-mean emissivity tables of setemiss
// (dust_5.f90:507-699) on the GPU.
//
// For every dust species id, temperature Tint(i) and frequency interval j
// (x = h nu / kT from nuint(j-1) to nuint(j), nuint(0) = 0), setemiss
// integrates with Numerical Recipes' qsimp (nr_2.f:209-231, EPS 1e-5,
// JMAX 20) over f in [0,1]:
//   kappaB(f) = ((1-f) kabs0 + f kabs1) Bnu(nu0 + f dnu, T)   (dust_5.f90:782-794)
//   chidB(f)  = dBnu(nu0 + f dnu, T) / ((1-f) chi0 + f chi1)  (:814-826)
// with kabs = (1 - albedo) kappa and chi = kappa interpolated by opacset
// (:21-55) at the interval ends.  The cumulative sums over j give kapint
// (normalized to 1 at the last interval), kappap (its total) and kappar
// (the inverse of the chidB sum); species above ndg get kappap smoothed by
// two 21-point passes (:612-653).
//
// Work split: k_emiss_seg runs one interval per lane (nopac x nnu x nT
// lanes, temperature fastest, so a wavefront shares the interval in x and
// converges in step); k_emiss_scan runs one block per species and one
// thread per temperature for the running sums in the reference's order,
// the normalization and the smoothing (kappap in LDS).  Compiled with
// contraction off, the arithmetic is the reference's operation sequence;
// only exp() may differ in the last bit from the host's libm.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcrt {

struct EmissArgs {
  int nopac, nlam, nT, nnu;
  const int* nlambda;                        // per species: table entries used (<= nlam)
  const double *nudust, *kappad, *adust;     // [id*nlam + k], nudust descending
  const double *Tint, *nuint;
  const int* smooth;                         // per species: 1 = smooth kappap
  double* seg;                               // [2][nopac*nnu*nT]: kappaB, chidB interval integrals
  double *kapint, *kappap, *kappar;          // [(id*nnu + j)*nT + i], [id*nT + i], [id*nT + i]
  unsigned long long* nfail;                 // qsimp calls that did not converge
  unsigned long long* nevals;                // integrand evaluations (null: not counted)
};

#ifdef MCRT_FP64_ONLY_KERNELS
// mathsubs.f90:79-113
__device__ __forceinline__ double em_expm1(double x) {
  if (fabs(x) > 0.6931472) return exp(x) - 1.0;
  double s = 0.0;
  s = x * (0.0001413161 + s);
  s = x * (0.0013298820 + s);
  s = x * (0.0083013598 + s);
  s = x * (0.0416573475 + s);
  s = x * (0.1666653019 + s);
  s = x * (0.4999999206 + s);
  s = x * (0.9999999995 + s);
  return s;
}

// mathsubs.f90:4-22 (kKind 0) and :29-47 (kKind 1)
template <int kKind>
__device__ __forceinline__ double em_planck(double nu, double T) {
  const double x = nu / T;
  if (x == 0.0) return 0.0;
  if constexpr (kKind == 0) return 0.1539897 * x * x * x / em_expm1(x);
  else return -0.1539897 * x * x * x * x / (4.0 * em_expm1(x) * em_expm1(-x));
}

struct EmSeg {
  double k0, k1, nu0, dnu, teff;
};

template <int kKind>
__device__ __forceinline__ double em_f(const EmSeg& s, double f) {
  const double nu = s.nu0 + f * s.dnu;
  if constexpr (kKind == 0) return ((1.0 - f) * s.k0 + f * s.k1) * em_planck<0>(nu, s.teff);
  else return em_planck<1>(nu, s.teff) / ((1.0 - f) * s.k0 + f * s.k1);
}

// qsimp + trapzd over [0,1]; returns true when the loop did not converge
// (the reference's `converg`).  *levels = trapzd refinements done.
template <int kKind>
__device__ bool em_qsimp(const EmSeg& s, double* out, int* levels) {
  double ost = -1.0e30, os = -1.0e30, st = 0.0, r = 0.0;
  for (int n = 1; n <= 20; ++n) {
    *levels = n;
    if (n == 1) {
      st = 0.5 * (1.0 - 0.0) * (em_f<kKind>(s, 0.0) + em_f<kKind>(s, 1.0));
    } else {
      const int it = 1 << (n - 2);
      const double tnm = (double)it;
      const double del = (1.0 - 0.0) / tnm;
      double x = 0.0 + 0.5 * del, sum = 0.0;
      for (int j = 1; j <= it; ++j) {
        sum = sum + em_f<kKind>(s, x);
        x = x + del;
      }
      st = 0.5 * (st + (1.0 - 0.0) * sum / tnm);
    }
    r = (4.0 * st - ost) / 3.0;
    if (fabs(r - os) < 1.0e-5 * fabs(os)) {
      *out = r;
      return false;
    }
    os = r;
    ost = st;
  }
  *out = r;
  return true;
}

// opacset's kappa and albedo of one species at nu (dust_5.f90:21-55)
__device__ __forceinline__ void em_opac(const double* nud, const double* kap, const double* alb, int nl, double nu,
                                        double& kappa, double& albedo) {
  if (nu >= nud[0]) { kappa = kap[0]; albedo = alb[0]; return; }
  if (nu <= nud[nl - 1]) { kappa = kap[nl - 1]; albedo = alb[nl - 1]; return; }
  int i0 = 1, i1 = nl;
  while (i1 - i0 > 1) {
    const int i = (i1 + i0) >> 1;
    if (nu <= nud[i - 1]) i0 = i; else i1 = i;
  }
  const double f = (nu - nud[i1 - 1]) / (nud[i0 - 1] - nud[i1 - 1]);
  const double fm1 = 1.0 - f;
  kappa = fm1 * kap[i1 - 1] + f * kap[i0 - 1];
  albedo = fm1 * alb[i1 - 1] + f * alb[i0 - 1];
}

__global__ void __launch_bounds__(256) k_emiss_seg(EmissArgs a) {
  const size_t nseg = (size_t)a.nopac * a.nnu * a.nT;
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned fails = 0;
  unsigned long long evals = 0;
  if (g < nseg) {
    const int i = (int)(g % (size_t)a.nT);
    const size_t r = g / (size_t)a.nT;
    const int j = (int)(r % (size_t)a.nnu);
    const int id = (int)(r / (size_t)a.nnu);
    const int nl = a.nlambda[id];
    const double* nud = a.nudust + (size_t)id * a.nlam;
    const double* kap = a.kappad + (size_t)id * a.nlam;
    const double* alb = a.adust + (size_t)id * a.nlam;
    const double teff = a.Tint[i];
    const double nu0 = (j == 0) ? 0.0 : a.nuint[j - 1] * teff;
    const double nu1 = a.nuint[j] * teff;
    const double dnu = nu1 - nu0;
    double c0, al0, c1, al1;
    em_opac(nud, kap, alb, nl, nu0, c0, al0);
    em_opac(nud, kap, alb, nl, nu1, c1, al1);
    double sk = 0.0, sc = 0.0;
    if (nu0 / teff < 50.0) {
      int lk = 0, lc = 0;
      fails += em_qsimp<0>(EmSeg{(1.0 - al0) * c0, (1.0 - al1) * c1, nu0, dnu, teff}, &sk, &lk) ? 1u : 0u;
      fails += em_qsimp<1>(EmSeg{c0, c1, nu0, dnu, teff}, &sc, &lc) ? 1u : 0u;
      evals = (1ull << (lk - 1)) + (1ull << (lc - 1)) + 2ull;    // 2^(n-1) + 1 per qsimp of n levels
    }
    a.seg[g] = sk;
    a.seg[nseg + g] = sc;
  }
  if (fails) atomicAdd(a.nfail, (unsigned long long)fails);
  if (a.nevals) {
    for (int o = 32; o > 0; o >>= 1) evals += __shfl_xor(evals, o);
    if (__lane_id() == 0) atomicAdd(a.nevals, evals);
  }
}

// one block per species; dynamic LDS: 2 nT doubles.  Smoothed species need
// nT >= 21 (the windows stay inside 1..nT), which the launcher checks.
__global__ void __launch_bounds__(256) k_emiss_scan(EmissArgs a) {
  extern __shared__ double em_lds[];
  const int id = blockIdx.x;
  const int nT = a.nT, nnu = a.nnu;
  const size_t nseg = (size_t)a.nopac * nnu * nT;
  const size_t base = (size_t)id * nnu * nT;
  double* kp = em_lds;
  double* tmp = em_lds + nT;
  for (int i = threadIdx.x; i < nT; i += blockDim.x) {
    const double teff = a.Tint[i];
    double K = 0.0, kr = 0.0, nu1 = 0.0;
    for (int j = 0; j < nnu; ++j) {
      const double nu0 = nu1;
      nu1 = a.nuint[j] * teff;
      const double dx = (nu1 - nu0) / teff;
      const size_t g = base + (size_t)j * nT + i;
      const double sk = a.seg[g], sc = a.seg[nseg + g];
      if (j == 0) {
        K = dx * sk;
        kr = dx * sc;
      } else {
        K = dx * sk + K;
        kr = dx * sc + kr;
      }
      a.kapint[g] = K;
    }
    a.kappar[(size_t)id * nT + i] = 1.0 / kr;
    kp[i] = K;
    for (int j = 0; j < nnu - 1; ++j) {
      const size_t g = base + (size_t)j * nT + i;
      a.kapint[g] = a.kapint[g] / K;
    }
    a.kapint[base + (size_t)(nnu - 1) * nT + i] = 1.0;
  }
  __syncthreads();
  if (a.smooth[id]) {
    const int ns = 10;
    for (int pass = 0; pass < 2; ++pass) {
      for (int i0 = threadIdx.x; i0 < nT; i0 += blockDim.x) {
        const int i = i0 + 1;                       // the reference's 1-based index
        int lo, hi;
        if (i > ns && i <= nT - ns) { lo = i - ns; hi = i + ns; }
        else if (i <= ns) { lo = 1; hi = 2 * i + 1; }
        else { lo = 2 * i - nT; hi = nT; }
        double acc = 0.0;
        for (int ii = lo; ii <= hi; ++ii) acc = acc + kp[ii - 1];
        tmp[i0] = acc / (double)(hi - lo + 1);
      }
      __syncthreads();
      for (int i0 = threadIdx.x; i0 < nT; i0 += blockDim.x) kp[i0] = tmp[i0];
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < nT; i += blockDim.x) a.kappap[(size_t)id * nT + i] = kp[i];
}

#endif  // MCRT_FP64_ONLY_KERNELS

}  // namespace mcrt
