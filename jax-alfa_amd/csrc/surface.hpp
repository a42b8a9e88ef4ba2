// Monin-Obukhov surface layer (src/surface/SurfaceFlux.py:53-443): the six
// heat-flux variants selected by (optSurfBC, optSurfFlux).  The work is a
// handful of reductions over one (nx, ny) plane, done in double.
#pragma once
#include "common.hpp"

template <class T> struct SurfArgs {
  const T* u; const T* v; const T* TH;   // 3-D fields, level 0 is read
  T* psim; T* psim0; T* psih; T* psih0; T* fim; T* fih;   // in/out
  T* M; T* ustar; T* qz; T* invOB;                          // out planes
  double* scal;   // [1] qz_sfc_avg, [2] sqrt(mean ustar^2), [3] mean M
  int ncol, nz, bc, hetero;
  double sfc_val, SHF, Ugal, T0, vonk, g, dz_half, z1, z0m_nd, zT_nd, logm, logh;
  const double* sfc_dev;   // when set, the surface value is read here at run time
  __device__ __forceinline__ double sval() const { return sfc_dev ? *sfc_dev : sfc_val; }
};


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

// The plane work is spread over SURF_NB blocks of 256 threads (one column
// per thread at 256^2).  The three plane means are two-stage reductions:
// every block writes one partial, and every block that needs a mean adds
// the partials itself in the same fixed order (so all blocks see the
// bitwise same mean, run to run).
constexpr int SURF_NB_MAX = 256;
constexpr int SURF_T = 256;

__device__ double block_reduce_256(double v, double* sm) {
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int s = SURF_T / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  const double r = sm[0];
  __syncthreads();
  return r;
}

// sum of part[k*stride] for k < n, same order in every block
__device__ double sum_parts(const double* part, int n, int stride, double* sm) {
  double v = 0.0;
  for (int k = threadIdx.x; k < n; k += SURF_T) v += part[(size_t)k * stride];
  return block_reduce_256(v, sm);
}

template <class T> __device__ __forceinline__ void surf_col(const SurfArgs<T>& a, int c, double& uu,
                                                            double& vv, double& th) {
  const size_t i = (size_t)c * a.nz;
  uu = (double)a.u[i] + a.Ugal;
  vv = (double)a.v[i];
  th = (double)a.TH[i];
}

template <class T>
__device__ __forceinline__ void k_surf_sums(const SurfArgs<T>& a, double* part) {
  __shared__ double sm[SURF_T];
  double sM = 0, sT = 0;
  for (int c = blockIdx.x * SURF_T + threadIdx.x; c < a.ncol; c += gridDim.x * SURF_T) {
    double uu, vv, th;
    surf_col(a, c, uu, vv, th);
    sM += sqrt(uu * uu + vv * vv);
    sT += th;
  }
  const double m = block_reduce_256(sM, sm);
  const double t = block_reduce_256(sT, sm);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = m; part[2 * blockIdx.x + 1] = t; }
}

// prescribed-temperature flux partials (optSurfBC = 2)
template <class T>
__device__ __forceinline__ void k_surf_q(const SurfArgs<T>& a, const double* part,
                                                   double* part2) {
  __shared__ double sm[SURF_T];
  const double invn = 1.0 / a.ncol;
  const double Mavg = sum_parts(part, gridDim.x, 2, sm) * invn;
  const double THavg = sum_parts(part + 1, gridDim.x, 2, sm) * invn;
  double sq = 0;
  for (int c = blockIdx.x * SURF_T + threadIdx.x; c < a.ncol; c += gridDim.x * SURF_T) {
    double uu, vv, th;
    surf_col(a, c, uu, vv, th);
    const double M = a.hetero ? sqrt(uu * uu + vv * vv) : Mavg;
    const double dm = a.logm + (double)a.psim[c] - (double)a.psim0[c];
    const double us = fmax(a.vonk * M / dm, 1e-3);
    const double dh = a.logh + (double)a.psih[c] - (double)a.psih0[c];
    const double tha = a.hetero ? th : THavg;
    sq += us * a.vonk * (a.sval() - tha) / dh;
  }
  const double q = block_reduce_256(sq, sm);
  if (threadIdx.x == 0) part2[blockIdx.x] = q;
}

template <class T>
__device__ __forceinline__ void k_surf_final(const SurfArgs<T>& a, const double* part,
                                                       const double* part2, double* part3) {
  __shared__ double sm[SURF_T];
  const double invn = 1.0 / a.ncol;
  const double Mavg = sum_parts(part, gridDim.x, 2, sm) * invn;
  const double THavg = sum_parts(part + 1, gridDim.x, 2, sm) * invn;
  double qavg;
  if (a.bc == 0) qavg = a.SHF;
  else if (a.bc == 1) qavg = a.sval();
  else qavg = sum_parts(part2, gridDim.x, 1, sm) * invn;
  const bool stable = qavg <= 0.0;
  double sus = 0;
  for (int c = blockIdx.x * SURF_T + threadIdx.x; c < a.ncol; c += gridDim.x * SURF_T) {
    double uu, vv, th;
    surf_col(a, c, uu, vv, th);
    const double M = a.hetero ? sqrt(uu * uu + vv * vv) : Mavg;
    const double dm = a.logm + (double)a.psim[c] - (double)a.psim0[c];
    const double us = fmax(a.vonk * M / dm, 1e-3);
    double q, thref;
    if (a.bc == 2) {
      const double dh = a.logh + (double)a.psih[c] - (double)a.psih0[c];
      const double tha = a.hetero ? th : THavg;
      q = us * a.vonk * (a.sval() - tha) / dh;
      thref = THavg + a.T0;
    } else {
      q = (a.bc == 0) ? a.SHF : a.sval();
      thref = a.hetero ? th + a.T0 : THavg + a.T0;
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
  const double su = block_reduce_256(sus, sm);
  if (threadIdx.x == 0) {
    part3[3 * blockIdx.x] = su;
    part3[3 * blockIdx.x + 1] = qavg;
    part3[3 * blockIdx.x + 2] = a.hetero ? -1.0 : Mavg;
  }
}

__device__ __forceinline__ void k_surf_scal(int nb, int ncol, const double* part3,
                                                      double* scal) {
  __shared__ double sm[SURF_T];
  const double su = sum_parts(part3, nb, 3, sm);
  if (threadIdx.x == 0 && scal) {
    scal[1] = part3[1];
    scal[2] = sqrt(su / ncol);
    scal[3] = part3[2];
  }
}

// Surface moisture flux plane (Main.py:253-266 with SurfaceFlux.py:451-504):
// mode 0 the constant MoistureFlux, 1 the series flux, 2 the flux
// diagnosed from the prescribed surface Q through MOST with the heat
// stability functions (homogeneous: against the plane mean of Q at level 0).
template <class T>
__device__ __forceinline__ void k_surf_moist(int ncol, int nz, int mode, int hetero, const T* __restrict__ Q,
                             const T* __restrict__ ustar, const T* __restrict__ psih,
                             const T* __restrict__ psih0, const double* __restrict__ qavg0,
                             const double* __restrict__ sval, double flux, double vonk,
                             double logq, T* qm) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncol; c += gridDim.x * blockDim.x) {
    double q;
    if (mode == 0) {
      q = flux;
    } else if (mode == 1) {
      q = *sval;
    } else {
      const double qa = hetero ? (double)Q[(size_t)c * nz] : *qavg0;
      const double denom = logq + (double)psih[c] - (double)psih0[c];
      q = (double)ustar[c] * vonk * (*sval - qa) / denom;
    }
    qm[c] = (T)q;
  }
}
