// mcrt_build.h -- the synthetic models' cell arrays on the GPU: gridset's
// cell loop (gridset_5_model13new.f90:1105-1180) with the analytic Ulrich
// envelope of densenv (densenv.f90:30-45, 69-73, 111-113; zerod.f90) and
// the flared disk of the synthetic workloads, for many models in one launch.
//
// Per cell (ir fastest, the engine's cell order): the volume
//   -(r2^3 - r1^3)(cos t2 - cos t1)(p2 - p1)/3 * rmincgs^3,
// the disk density
//   varpi^-alpha exp(-z^2 / 2h^2), h = h0 varpi^flare   (AU, inside Rin..Rout)
// normalized to the model's disk mass, the envelope density at the cell
// centre (radius in R*), an optional clumping factor on both, and the
// species (1 where the disk dominates, 3 envelope, 0 empty).
//
// k_build_cells computes volumes and both densities and one disk-mass
// partial sum per block (fixed-order tree); k_build_mass adds a model's
// partials in a fixed order; k_build_finish scales, clumps and assigns the
// species.  No atomics: the same inputs give the same bits on every run.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcrt {

// per-model parameters, [k * kBuildPar + BP_*]
enum {
  BP_RMINCGS = 0,   // R* in cm
  BP_AUTORS,        // R* per AU
  BP_RIN_AU,        // disk inner radius
  BP_ROUT_AU,       // disk outer radius
  BP_H0_AU,         // disk scale height at 1 AU
  BP_FLARE,
  BP_ALPHA,         // disk radial density exponent
  BP_RC,            // envelope centrifugal radius (R*)
  BP_RHOE0,         // envelope density scale (g/cm^3)
  BP_RMINE,         // envelope inner radius (R*)
  BP_MDISK_G,       // disk mass (g)
  BP_RHOAMB,        // ambient density floor (g/cm^3)
  kBuildPar
};

struct BuildArgs {
  int nmodels;
  unsigned nblk;                                   // blocks per model along x
  const int *nr, *nt, *np;                         // cells per axis
  const long long *roff, *toff, *poff, *coff;      // first wall / cell of each model
  const double *rarr, *thetarr, *phiarr;           // walls, concatenated
  const double* par;
  const double* clump;                             // per cell, or null
  double *rho, *vcell, *rho_d;                     // rho_d: scratch (unnormalized disk)
  uint8_t* species;
  double *partial, *scale;                         // [k * nblk + b], [k]
  unsigned long long* nflag;                       // cells where zerod found no single root
};

#ifdef MCRT_FP64_ONLY_KERNELS

// zerod.f90:3-69; returns 0 when *xmu0 was set
__device__ __forceinline__ int bd_zerod(double xmu, double z, double* xmu0) {
  const double pi = 3.1415926535897932384626433;
  const double a1 = 0.0, a2 = z - 1.0, a3 = -z * xmu;
  const double q = (a1 * a1 - 3.0 * a2) / 9.0;
  const double r = (2.0 * (a1 * a1 * a1) - 9.0 * a1 * a2 + 27.0 * a3) / 54.0;
  const double q3 = q * q * q;
  if (q3 - r * r >= 0.0) {
    const double qh = sqrt(q);
    const double theta = acos(r / sqrt(q3));
    const double x1 = -2.0 * qh * cos(theta / 3.0) - a1 / 3.0;
    const double x2 = -2.0 * qh * cos((theta + 2.0 * pi) / 3.0) - a1 / 3.0;
    const double x3 = -2.0 * qh * cos((theta + 4.0 * pi) / 3.0) - a1 / 3.0;
    if ((x1 > 0.0 && x2 > 0.0) || (x1 > 0.0 && x3 > 0.0) || (x2 > 0.0 && x3 > 0.0)) return 1;
    if (x3 > 0.0) { *xmu0 = x3; return 0; }
    if (x2 > 0.0) { *xmu0 = x2; return 0; }
    if (x1 > 0.0) { *xmu0 = x1; return 0; }
    return 2;
  }
  const double factor = sqrt(r * r - q3) + fabs(r);
  const double factor3 = pow(factor, 1.0 / 3.0);
  *xmu0 = -1.0 * copysign(1.0, r) * (factor3 + q / factor3) - a1 / 3.0;
  return 0;
}

__device__ __forceinline__ double bd_block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (unsigned s = blockDim.x >> 1; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + s];
    __syncthreads();
  }
  return sh[0];
}

__global__ void __launch_bounds__(256) k_build_cells(BuildArgs a) {
  __shared__ double sh[256];
  const int k = blockIdx.y;
  const int nr = a.nr[k], nt = a.nt[k], np = a.np[k];
  const long long ncell = (long long)nr * nt * np;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double dm = 0.0;
  unsigned bad = 0;
  if (c < ncell) {
    const double* P = a.par + (size_t)k * kBuildPar;
    const int ir = (int)(c % nr);
    const int it = (int)((c / nr) % nt);
    const int ip = (int)(c / ((long long)nr * nt));
    const double* R = a.rarr + a.roff[k];
    const double* T = a.thetarr + a.toff[k];
    const double* F = a.phiarr + a.poff[k];
    const double r1 = R[ir], r2 = R[ir + 1];
    const double dr3 = pow(r2, 3.0) - pow(r1, 3.0);
    const double dcost = cos(T[it + 1]) - cos(T[it]);
    const double dp = F[ip + 1] - F[ip];
    const double rm = P[BP_RMINCGS];
    const double vol = -(dr3 * dcost * dp) / 3.0 * pow(rm, 3.0);
    const double rad = 0.5 * (r2 + r1);
    const double tc = 0.5 * (T[it + 1] + T[it]);
    const double cost = cos(tc);
    // disk (synthetic flared disk, AU)
    const double r_au = rad / P[BP_AUTORS];
    const double varpi = r_au * sin(tc);
    const double z = r_au * cost;
    const double vp = fmax(varpi, 1e-30);
    const double h = P[BP_H0_AU] * pow(vp, P[BP_FLARE]);
    double rd = 0.0;
    if (r_au >= P[BP_RIN_AU] && varpi >= P[BP_RIN_AU] && r_au <= P[BP_ROUT_AU]) {
      const double zh = z / h;
      rd = pow(vp, -P[BP_ALPHA]) * exp(-0.5 * (zh * zh));
    }
    // envelope (densenv.f90:30-45, 69-73, 111-113)
    const double rmax = R[nr];
    const double rmine = P[BP_RMINE];
    double de = 0.0;
    const double cosa = fabs(cost);
    const double zp = rad * cosa;
    const double xmu = zp / rad;
    if (rad > rmine) {
      const double rx = rad / P[BP_RC];
      double xmu0 = 0.0;
      if (bd_zerod(xmu, rx, &xmu0) != 0) {
        bad = 1;
        xmu0 = __builtin_nan("");
      }
      const double factor = 1.0 / ((xmu / xmu0) + 2.0 * (xmu0 * xmu0) / rx);
      de = P[BP_RHOE0] * pow(rx, -1.5) * factor / sqrt(1.0 + xmu / xmu0);
      if (de < P[BP_RHOAMB]) de = P[BP_RHOAMB];
    }
    if (rad >= rmax || rad <= rmine) de = 0.0;
    const size_t g = (size_t)(a.coff[k] + c);
    a.vcell[g] = vol;
    a.rho_d[g] = rd;
    a.rho[g] = de;
    dm = rd * vol;
  }
  if (bad) atomicAdd(a.nflag, 1ull);
  const double s = bd_block_sum(dm, sh);
  if (threadIdx.x == 0) a.partial[(size_t)k * a.nblk + blockIdx.x] = s;
}

// one block per model: the disk mass from the partials in a fixed order
__global__ void __launch_bounds__(256) k_build_mass(BuildArgs a) {
  __shared__ double sh[256];
  const int k = blockIdx.x;
  double v = 0.0;
  for (unsigned b = threadIdx.x; b < a.nblk; b += blockDim.x) v = v + a.partial[(size_t)k * a.nblk + b];
  const double mass = bd_block_sum(v, sh);
  if (threadIdx.x == 0) a.scale[k] = (mass > 0.0) ? a.par[(size_t)k * kBuildPar + BP_MDISK_G] / mass : 1.0;
}

__global__ void __launch_bounds__(256) k_build_finish(BuildArgs a) {
  const int k = blockIdx.y;
  const long long ncell = (long long)a.nr[k] * a.nt[k] * a.np[k];
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const size_t g = (size_t)(a.coff[k] + c);
  double rd = a.rho_d[g] * a.scale[k];
  double re = a.rho[g];
  if (a.clump) {
    const double f = a.clump[g];
    rd = rd * f;
    re = re * f;
  }
  const double rho = rd + re;
  a.rho[g] = rho;
  a.species[g] = (rho > 0.0) ? (uint8_t)((rd >= re) ? 1 : 3) : (uint8_t)0;
}

#endif  // MCRT_FP64_ONLY_KERNELS

}  // namespace mcrt
