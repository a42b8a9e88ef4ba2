// Monin-Obukhov surface layer (src/surface/SurfaceFlux.py:53-443): the six
// heat-flux variants selected by (optSurfBC, optSurfFlux).  The work is a
// handful of reductions over one (nx, ny) plane, so one 1024-thread block
// does all of it in double with fixed-order block reductions.
#pragma once
#include "common.hpp"

template <class T> struct SurfArgs {
  const T* u; const T* v; const T* TH;   // 3-D fields, level 0 is read
  T* psim; T* psim0; T* psih; T* psih0; T* fim; T* fih;   // in/out
  T* M; T* ustar; T* qz; T* invOB;                          // out planes
  double* scal;   // [1] qz_sfc_avg, [2] sqrt(mean ustar^2), [3] mean M
  int ncol, nz, bc, hetero;
  double sfc_val, SHF, Ugal, T0, vonk, g, dz_half, z1, z0m_nd, zT_nd, logm, logh;
};

constexpr int SURF_BS = 1024;

__device__ double block_sum_1024(double v, double* sm) {
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int s = SURF_BS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  const double r = sm[0];
  __syncthreads();
  return r;
}

// psi_m, psi_h, fi_m, fi_h (MOSTstable :53-77 / MOSTunstable :83-113).
__device__ __forceinline__ void most(double zL, bool stable, double& pm, double& ph, double& fm,
                                     double& fh) {
  if (stable) {
    pm = 5.0 * zL; ph = 5.0 * zL; fm = 1.0 + pm; fh = 1.0 + ph;
  } else {
    const double x = sqrt(sqrt(1.0 - 15.0 * zL));
    pm = -2.0 * log(0.5 * (1.0 + x)) - log(0.5 * (1.0 + x * x)) + 2.0 * atan(x) - M_PI / 2;
    ph = -2.0 * log(0.5 * (1.0 + x * x));
    fm = 1.0 / x;
    fh = (1.0 / x) * (1.0 / x);
  }
}

template <class T>
__global__ void __launch_bounds__(SURF_BS) k_surface(SurfArgs<T> a) {
  __shared__ double sm[SURF_BS];
  const int n = a.ncol, nz = a.nz;
  const double invn = 1.0 / n;
  double sM = 0, sT = 0;
  for (int c = threadIdx.x; c < n; c += SURF_BS) {
    const size_t i = (size_t)c * nz;
    const double uu = (double)a.u[i] + a.Ugal, vv = a.v[i];
    sM += sqrt(uu * uu + vv * vv);
    sT += (double)a.TH[i];
  }
  const double Mavg = block_sum_1024(sM, sm) * invn;
  const double THavg = block_sum_1024(sT, sm) * invn;
  // heat flux average
  double qavg;
  if (a.bc == 0) qavg = a.SHF;
  else if (a.bc == 1) qavg = a.sfc_val;
  else {
    double sq = 0;
    for (int c = threadIdx.x; c < n; c += SURF_BS) {
      const size_t i = (size_t)c * nz;
      const double uu = (double)a.u[i] + a.Ugal, vv = a.v[i];
      const double M = a.hetero ? sqrt(uu * uu + vv * vv) : Mavg;
      const double dm = a.logm + (double)a.psim[c] - (double)a.psim0[c];
      const double us = fmax(a.vonk * M / dm, 1e-3);
      const double dh = a.logh + (double)a.psih[c] - (double)a.psih0[c];
      const double tha = a.hetero ? (double)a.TH[i] : THavg;
      sq += us * a.vonk * (a.sfc_val - tha) / dh;
    }
    qavg = block_sum_1024(sq, sm) * invn;
  }
  const bool stable = qavg <= 0.0;
  double sus = 0;
  for (int c = threadIdx.x; c < n; c += SURF_BS) {
    const size_t i = (size_t)c * nz;
    const double uu = (double)a.u[i] + a.Ugal, vv = a.v[i];
    const double M = a.hetero ? sqrt(uu * uu + vv * vv) : Mavg;
    const double dm = a.logm + (double)a.psim[c] - (double)a.psim0[c];
    const double us = fmax(a.vonk * M / dm, 1e-3);
    double q, thref;
    if (a.bc == 2) {
      const double dh = a.logh + (double)a.psih[c] - (double)a.psih0[c];
      const double tha = a.hetero ? (double)a.TH[i] : THavg;
      q = us * a.vonk * (a.sfc_val - tha) / dh;
      thref = THavg + a.T0;
    } else {
      q = (a.bc == 0) ? a.SHF : a.sfc_val;
      thref = a.hetero ? (double)a.TH[i] + a.T0 : THavg + a.T0;
    }
    const double iob = -(a.vonk * a.g * qavg) / ((us * us * us) * thref);
    double pm, ph, fm, fh, pm0, ph0, d1, d2;
    most(a.dz_half * iob, stable, pm, ph, fm, fh);
    most(a.z0m_nd * iob, stable, pm0, d1, d2, d2);
    most(a.zT_nd * iob, stable, d1, ph0, d2, d2);
    a.M[c] = (T)M;
    a.ustar[c] = (T)us;
    a.qz[c] = (T)q;
    a.invOB[c] = (T)iob;
    a.psim[c] = (T)pm; a.psim0[c] = (T)pm0; a.psih[c] = (T)ph; a.psih0[c] = (T)ph0;
    a.fim[c] = (T)fm; a.fih[c] = (T)fh;
    sus += us * us;
  }
  const double su2 = block_sum_1024(sus, sm) * invn;
  if (threadIdx.x == 0 && a.scal) {
    a.scal[1] = qavg;
    a.scal[2] = sqrt(su2);
    a.scal[3] = a.hetero ? -1.0 : Mavg;
  }
}
