// Pointwise and z-stencil kernels of the time step.  Fields are (ncol, nz)
// with z fastest (ncol = nx*ny, or mx*my on the 3/2-padded grid); planes are
// (ncol,).  One thread per point, consecutive threads on consecutive z, so
// every load and store is a coalesced 128-byte segment and z+-1 neighbours
// come from the same cache lines.
#pragma once
#include "common.hpp"

#define PW_LOOP(N)                                                             \
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < (N);      \
       i += (size_t)gridDim.x * blockDim.x)

struct PwDims {
  size_t n;      // ncol * nz
  int nz;
};

// Derivz_M + Derivz_TH (src/operations/Derivatives.py:190-273).
template <class T>
__device__ __forceinline__ void k_derivz(const PwDims& d, const T* __restrict__ u, const T* __restrict__ v,
                         const T* __restrict__ w, const T* __restrict__ TH,
                         const T* __restrict__ ustar, const T* __restrict__ M,
                         const T* __restrict__ fim, const T* __restrict__ fih,
                         const T* __restrict__ qz, T* dudz, T* dvdz, T* dwdz, T* dTHdz,
                         T Ugal, T idz, T half_dz_vonk, T vonk_half_dz) {
  const int nz = d.nz;
  PW_LOOP(d.n) {
    const int z = (int)(i % nz);
    const size_t c = i / nz;
    if (dudz) {
      T a, b;
      if (z == 0) {
        const T den = M[c] * half_dz_vonk;  // M * vonk * 0.5 * dz
        a = fim[c] * ustar[c] * (u[i] + Ugal) / den;
        b = fim[c] * ustar[c] * v[i] / den;
      } else if (z < nz - 1) {
        a = (u[i] - u[i - 1]) * idz;
        b = (v[i] - v[i - 1]) * idz;
      } else {
        a = T(0); b = T(0);
      }
      dudz[i] = a;
      dvdz[i] = b;
      dwdz[i] = (z < nz - 1) ? (w[i + 1] - w[i]) * idz : T(0);
    }
    if (dTHdz) {
      dTHdz[i] = (z == 0) ? fih[c] * (-qz[c] / ustar[c]) / vonk_half_dz
                          : (TH[i] - TH[i - 1]) * idz;
    }
  }
}

// k_derivz with W = 16 / sizeof(T) consecutive levels per thread and
// 16-byte accesses (nz % W == 0, arrays 16-byte aligned): the same
// arithmetic, the neighbour levels across the vector edge read as scalars.
template <class T>
__device__ __forceinline__ void k_derivz_v(const PwDims& d, const T* __restrict__ u, const T* __restrict__ v,
                           const T* __restrict__ w, const T* __restrict__ TH,
                           const T* __restrict__ ustar, const T* __restrict__ M,
                           const T* __restrict__ fim, const T* __restrict__ fih,
                           const T* __restrict__ qz, T* dudz, T* dvdz, T* dwdz, T* dTHdz,
                           T Ugal, T idz, T half_dz_vonk, T vonk_half_dz) {
  constexpr int W = Vec16<T>::W;
  const int nz = d.nz;
  const size_t nv = d.n / W;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < nv;
       t += (size_t)gridDim.x * blockDim.x) {
    const size_t i0 = t * W;
    const int z0 = (int)(i0 % nz);
    const size_t c = i0 / nz;
    if (dudz) {
      const Vec16<T> uu = vld(u + i0), vv = vld(v + i0), ww = vld(w + i0);
      const T up = z0 > 0 ? u[i0 - 1] : T(0), vp = z0 > 0 ? v[i0 - 1] : T(0);
      const T wn = z0 + W < nz ? w[i0 + W] : T(0);
      Vec16<T> a, b, e;
#pragma unroll
      for (int j = 0; j < W; ++j) {
        const int z = z0 + j;
        if (z == 0) {
          const T den = M[c] * half_dz_vonk;
          a.v[j] = fim[c] * ustar[c] * (uu.v[j] + Ugal) / den;
          b.v[j] = fim[c] * ustar[c] * vv.v[j] / den;
        } else if (z < nz - 1) {
          a.v[j] = (uu.v[j] - (j > 0 ? uu.v[j - 1] : up)) * idz;
          b.v[j] = (vv.v[j] - (j > 0 ? vv.v[j - 1] : vp)) * idz;
        } else {
          a.v[j] = T(0); b.v[j] = T(0);
        }
        e.v[j] = (z < nz - 1) ? ((j + 1 < W ? ww.v[j + 1] : wn) - ww.v[j]) * idz : T(0);
      }
      vst(dudz + i0, a); vst(dvdz + i0, b); vst(dwdz + i0, e);
    }
    if (dTHdz) {
      const Vec16<T> th = vld(TH + i0);
      const T tp = z0 > 0 ? TH[i0 - 1] : T(0);
      Vec16<T> o;
#pragma unroll
      for (int j = 0; j < W; ++j) {
        const int z = z0 + j;
        o.v[j] = (z == 0) ? fih[c] * (-qz[c] / ustar[c]) / vonk_half_dz
                          : (th.v[j] - (j > 0 ? th.v[j - 1] : tp)) * idz;
      }
      vst(dTHdz + i0, o);
    }
  }
}

// Derivz_Generic_uvp (Derivatives.py:310) / Derivz_Generic_w (:340).
template <class T>
__device__ __forceinline__ void k_derivz_generic(const PwDims& d, const T* __restrict__ F, T* out, T idz, int wnode) {
  const int nz = d.nz;
  PW_LOOP(d.n) {
    const int z = (int)(i % nz);
    if (wnode) out[i] = (z < nz - 1) ? (F[i + 1] - F[i]) * idz : T(0);
    else out[i] = (z > 0) ? (F[i] - F[i - 1]) * idz : T(0);
  }
}

// StagGridAvg (Utilities.py:64): out is (ncol, nz-1).
template <class T>
__device__ __forceinline__ void k_stag_avg(const PwDims& d, const T* __restrict__ F, T* out) {
  const int nz = d.nz;
  const size_t n = d.n / nz * (nz - 1);
  PW_LOOP(n) {
    const size_t c = i / (nz - 1);
    const int z = (int)(i % (nz - 1));
    const T* f = F + c * nz + z;
    out[i] = T(0.5) * (f[0] + f[1]);
  }
}

// Advection_NoDealias (src/pde/NSE_AdvectionTerms.py:163-238); on the padded
// grid it is the product stage of Advection_Dealias (:65-155).
template <class T>
__device__ __forceinline__ void k_advection(const PwDims& d, const T* __restrict__ u, const T* __restrict__ v,
                            const T* __restrict__ w, const T* __restrict__ dudy,
                            const T* __restrict__ dudz, const T* __restrict__ dvdx,
                            const T* __restrict__ dvdz, const T* __restrict__ dwdx,
                            const T* __restrict__ dwdy, T* Cx, T* Cy, T* Cz) {
  const int nz = d.nz;
  PW_LOOP(d.n) {
    const int z = (int)(i % nz);
    const T a1x = v[i] * (dudy[i] - dvdx[i]);
    const T a1y = u[i] * (dvdx[i] - dudy[i]);
    T cx, cy, cz;
    if (z == 0) {
      cx = a1x + T(0.5) * (w[i + 1] * (dudz[i + 1] - dwdx[i + 1]));
      cy = a1y + T(0.5) * (w[i + 1] * (dvdz[i + 1] - dwdy[i + 1]));
      cz = T(0);
    } else if (z < nz - 1) {
      const T a2x0 = w[i] * (dudz[i] - dwdx[i]), a2x1 = w[i + 1] * (dudz[i + 1] - dwdx[i + 1]);
      const T a2y0 = w[i] * (dvdz[i] - dwdy[i]), a2y1 = w[i + 1] * (dvdz[i + 1] - dwdy[i + 1]);
      cx = a1x + T(0.5) * (a2x0 + a2x1);
      cy = a1y + T(0.5) * (a2y0 + a2y1);
      const T uG = T(0.5) * (u[i - 1] + u[i]);
      const T vG = T(0.5) * (v[i - 1] + v[i]);
      cz = uG * (dwdx[i] - dudz[i]) + vG * (dwdy[i] - dvdz[i]);
    } else {
      cx = a1x + w[i] * (dudz[i] - dwdx[i]);
      cy = a1y + w[i] * (dvdz[i] - dwdy[i]);
      cz = T(0);
    }
    Cx[i] = cx; Cy[i] = cy; Cz[i] = cz;
  }
}

// ScalarAdvection_NoDealias (src/pde/SCL_AdvectionTerms.py:114-150); product
// stage of the dealiased variant (:57-106) on the padded grid.
template <class T>
__device__ __forceinline__ void k_scalar_adv(const PwDims& d, const T* __restrict__ u, const T* __restrict__ v,
                             const T* __restrict__ w, const T* __restrict__ sx,
                             const T* __restrict__ sy, const T* __restrict__ sz, T* out) {
  const int nz = d.nz;
  PW_LOOP(d.n) {
    const int z = (int)(i % nz);
    const T cc1 = u[i] * sx[i] + v[i] * sy[i];
    T r;
    if (z < nz - 1) r = cc1 + T(0.5) * (w[i] * sz[i] + w[i + 1] * sz[i + 1]);
    else r = cc1 + w[i] * sz[i];
    out[i] = r;
  }
}

// Accessors of a z-derivative field at point i (level z, column c): read
// from memory (PtrAt), or formed on the fly with the arithmetic of k_derivz
// (DzUV: u or v on w nodes with the MOST wall value at level 0 and 0 at the
// top level; DzW: w on uvp nodes; DzTH: TH on w nodes with the surface flux
// at level 0).
template <class T> struct PtrAt {
  const T* __restrict__ p;
  __device__ __forceinline__ T operator()(size_t i, int, size_t) const { return p[i]; }
};
template <class T, bool ADD> struct DzUV {
  const T* __restrict__ f;
  const T* __restrict__ fim;
  const T* __restrict__ ustar;
  const T* __restrict__ M;
  T add, idz, half_dz_vonk;
  int nz;
  __device__ __forceinline__ T operator()(size_t i, int z, size_t c) const {
    if (z == 0) {
      const T den = M[c] * half_dz_vonk;
      if constexpr (ADD) return fim[c] * ustar[c] * (f[i] + add) / den;
      else return fim[c] * ustar[c] * f[i] / den;
    }
    return z < nz - 1 ? (f[i] - f[i - 1]) * idz : T(0);
  }
};
template <class T> struct DzW {
  const T* __restrict__ w;
  T idz;
  int nz;
  __device__ __forceinline__ T operator()(size_t i, int z, size_t) const {
    return z < nz - 1 ? (w[i + 1] - w[i]) * idz : T(0);
  }
};
template <class T> struct DzTH {
  const T* __restrict__ th;
  const T* __restrict__ fih;
  const T* __restrict__ qz;
  const T* __restrict__ ustar;
  T idz, vonk_half_dz;
  __device__ __forceinline__ T operator()(size_t i, int z, size_t c) const {
    return z == 0 ? fih[c] * (-qz[c] / ustar[c]) / vonk_half_dz : (th[i] - th[i - 1]) * idz;
  }
};

// Advection (rotational form) and scalar advection at point i (level z,
// column c): k_advection + k_scalar_adv of the unpadded grid.  uz, vz, sz:
// accessors of du/dz, dv/dz (w nodes) and d(scalar)/dz.
template <class T> struct AdvOut { T cx, cy, cz, r; };
template <class T, class UZ, class VZ, class SZ>
__device__ __forceinline__ AdvOut<T> adv_point(size_t i, int z, size_t c, int nz,
                                               const T* __restrict__ u, const T* __restrict__ v,
                                               const T* __restrict__ w,
                                               const T* __restrict__ dudy, const UZ& dudz,
                                               const T* __restrict__ dvdx, const VZ& dvdz,
                                               const T* __restrict__ dwdx,
                                               const T* __restrict__ dwdy,
                                               const T* __restrict__ sx, const T* __restrict__ sy,
                                               const SZ& sz) {
  const T ui = u[i], vi = v[i], wi = w[i];
  const T a1x = vi * (dudy[i] - dvdx[i]);
  const T a1y = ui * (dvdx[i] - dudy[i]);
  const T cc1 = ui * sx[i] + vi * sy[i];
  AdvOut<T> o;
  if (z == 0) {
    const T w1 = w[i + 1];
    o.cx = a1x + T(0.5) * (w1 * (dudz(i + 1, z + 1, c) - dwdx[i + 1]));
    o.cy = a1y + T(0.5) * (w1 * (dvdz(i + 1, z + 1, c) - dwdy[i + 1]));
    o.cz = T(0);
    o.r = cc1 + T(0.5) * (wi * sz(i, z, c) + w1 * sz(i + 1, z + 1, c));
  } else if (z < nz - 1) {
    const T w1 = w[i + 1];
    const T uz0 = dudz(i, z, c), vz0 = dvdz(i, z, c);
    const T a2x0 = wi * (uz0 - dwdx[i]), a2x1 = w1 * (dudz(i + 1, z + 1, c) - dwdx[i + 1]);
    const T a2y0 = wi * (vz0 - dwdy[i]), a2y1 = w1 * (dvdz(i + 1, z + 1, c) - dwdy[i + 1]);
    o.cx = a1x + T(0.5) * (a2x0 + a2x1);
    o.cy = a1y + T(0.5) * (a2y0 + a2y1);
    const T uG = T(0.5) * (u[i - 1] + ui);
    const T vG = T(0.5) * (v[i - 1] + vi);
    o.cz = uG * (dwdx[i] - uz0) + vG * (dwdy[i] - vz0);
    o.r = cc1 + T(0.5) * (wi * sz(i, z, c) + w1 * sz(i + 1, z + 1, c));
  } else {
    o.cx = a1x + wi * (dudz(i, z, c) - dwdx[i]);
    o.cy = a1y + wi * (dvdz(i, z, c) - dwdy[i]);
    o.cz = T(0);
    o.r = cc1 + wi * sz(i, z, c);
  }
  return o;
}

// BuoyancyOpt1/2 (src/pde/NSE_BuoyancyTerms.py:49-150), dry; bar = plane
// mean of TH + T_0 (double).
template <class T>
__device__ __forceinline__ void k_buoyancy(const PwDims& d, const T* __restrict__ TH, const double* __restrict__ bar,
                           T* out, double T0, double g, int opt) {
  const int nz = d.nz;
  PW_LOOP(d.n) {
    const int z = (int)(i % nz);
    if (z == 0) { out[i] = T(0); continue; }
    const double t1 = (double)TH[i] + T0, t0 = (double)TH[i - 1] + T0;
    double n1, n0;
    if (opt == 0) { n1 = (t1 - bar[z]) / T0; n0 = (t0 - bar[z - 1]) / T0; }
    else { n1 = (t1 - bar[z]) / bar[z]; n0 = (t0 - bar[z - 1]) / bar[z - 1]; }
    out[i] = (T)(0.5 * g * (n1 + n0));
  }
}

template <class T> __device__ __forceinline__ T smag(T a, T b, T c, T d, T e, T f) {
  return sqrt(T(2) * (a * a + b * b + c * c + T(2) * d * d + T(2) * e * e + T(2) * f * f));
}

// StrainsUVPnodes_NoDealias (src/subgridscale/StrainRates.py:149-211) at
// point i (level z, column c): S12, S13, S23 and |S| (S11 = dudx, S22 =
// dvdy, S33 = dwdz).  dudz, dvdz, dwdz: accessors (PtrAt or the on-the-fly
// derivatives).
template <class T> struct StrainOut { T s12, s13, s23, S; };
template <class T, class UZ, class VZ, class WZ>
__device__ __forceinline__ StrainOut<T> strains_point(size_t i, int z, size_t c, int nz,
                                                      const T* __restrict__ dudx,
                                                      const T* __restrict__ dvdx,
                                                      const T* __restrict__ dwdx,
                                                      const T* __restrict__ dudy,
                                                      const T* __restrict__ dvdy,
                                                      const T* __restrict__ dwdy,
                                                      const UZ& dudz, const VZ& dvdz,
                                                      const WZ& dwdz) {
  T uz, vz, wx, wy;
  if (z == 0) {
    uz = dudz(i, z, c); vz = dvdz(i, z, c);
    wx = T(0.5) * dwdx[i + 1]; wy = T(0.5) * dwdy[i + 1];
  } else if (z < nz - 1) {
    uz = T(0.5) * (dudz(i, z, c) + dudz(i + 1, z + 1, c));
    vz = T(0.5) * (dvdz(i, z, c) + dvdz(i + 1, z + 1, c));
    wx = T(0.5) * (dwdx[i] + dwdx[i + 1]); wy = T(0.5) * (dwdy[i] + dwdy[i + 1]);
  } else {
    uz = dudz(i, z, c); vz = dvdz(i, z, c); wx = dwdx[i]; wy = dwdy[i];
  }
  StrainOut<T> o;
  o.s12 = T(0.5) * (dudy[i] + dvdx[i]);
  o.s13 = T(0.5) * (uz + wx);
  o.s23 = T(0.5) * (vz + wy);
  o.S = smag<T>(dudx[i], dvdy[i], dwdz(i, z, c), o.s12, o.s13, o.s23);
  return o;
}

// StrainsUVPnodes_NoDealias: writes S12, S13, S23 and |S|.
template <class T>
__device__ __forceinline__ void k_strains_uvp(const PwDims& d, const T* __restrict__ dudx, const T* __restrict__ dvdx,
                              const T* __restrict__ dwdx, const T* __restrict__ dudy,
                              const T* __restrict__ dvdy, const T* __restrict__ dwdy,
                              const T* __restrict__ dudz, const T* __restrict__ dvdz,
                              const T* __restrict__ dwdz, T* S12, T* S13, T* S23, T* S) {
  const int nz = d.nz;
  PW_LOOP(d.n) {
    const int z = (int)(i % nz);
    const StrainOut<T> o = strains_point<T>(i, z, 0, nz, dudx, dvdx, dwdx, dudy, dvdy, dwdy,
                                            PtrAt<T>{dudz}, PtrAt<T>{dvdz}, PtrAt<T>{dwdz});
    S12[i] = o.s12; S13[i] = o.s13; S23[i] = o.s23; S[i] = o.S;
  }
}

// The vertical gradients (k_derivz), advection + scalar advection
// (adv_point) and the uvp-node strain rates (strains_point) in one pass on
// the unpadded grid.  du/dz, dv/dz, dw/dz and dTH/dz are formed from u, v,
// w, TH where they are needed (the values at level z + 1 by the same
// arithmetic as at z + 1 itself) and written once for the later stages;
// u, v, w and the six velocity derivatives that both parts need are read
// once.
template <class T> struct DzAdvStrainArgs {
  const T *u, *v, *w, *TH, *sx, *sy;
  const T *dudx, *dvdx, *dwdx, *dudy, *dvdy, *dwdy;
  const T *ustar, *M, *fim, *fih, *qz;   // surface planes (MOST wall values)
  T Ugal, idz, half_dz_vonk;
  T *dudz, *dvdz, *dwdz, *dTdz;
  T *Cx, *Cy, *Cz, *sadv, *S12, *S13, *S23, *S;
};
template <class T>
__device__ __forceinline__ void k_dz_adv_strains(const PwDims& d, const DzAdvStrainArgs<T>& a) {
  const int nz = d.nz;
  const DzUV<T, true> uz{a.u, a.fim, a.ustar, a.M, a.Ugal, a.idz, a.half_dz_vonk, nz};
  const DzUV<T, false> vz{a.v, a.fim, a.ustar, a.M, T(0), a.idz, a.half_dz_vonk, nz};
  const DzW<T> wz{a.w, a.idz, nz};
  const DzTH<T> tz{a.TH, a.fih, a.qz, a.ustar, a.idz, a.half_dz_vonk};
  PW_LOOP(d.n) {
    const int z = (int)(i % nz);
    const size_t c = z == 0 ? i / nz : 0;   // the column is read only at level 0
    const AdvOut<T> o = adv_point<T>(i, z, c, nz, a.u, a.v, a.w, a.dudy, uz, a.dvdx, vz, a.dwdx,
                                     a.dwdy, a.sx, a.sy, tz);
    const StrainOut<T> s = strains_point<T>(i, z, c, nz, a.dudx, a.dvdx, a.dwdx, a.dudy, a.dvdy,
                                            a.dwdy, uz, vz, wz);
    a.dudz[i] = uz(i, z, c); a.dvdz[i] = vz(i, z, c); a.dwdz[i] = wz(i, z, c);
    a.dTdz[i] = tz(i, z, c);
    a.Cx[i] = o.cx; a.Cy[i] = o.cy; a.Cz[i] = o.cz; a.sadv[i] = o.r;
    a.S12[i] = s.s12; a.S13[i] = s.s13; a.S23[i] = s.s23; a.S[i] = s.S;
  }
}

// k_dz_adv_strains with 16-byte accesses (nz % W == 0, 16-byte aligned
// arrays): each thread takes W consecutive levels of one column.  The fields
// read at neighbouring levels are held as windows of W + 2 values (index 0:
// level z0 - 1, index W + 1: level z0 + W, read only inside the column), and
// adv_point / strains_point run on the windows: the same arithmetic as the
// scalar kernel.
template <class T>
__device__ __forceinline__ void k_dz_adv_strains_v(const PwDims& d, const DzAdvStrainArgs<T>& a) {
  constexpr int W = Vec16<T>::W;
  const int nz = d.nz;
  const size_t nv = d.n / W;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < nv;
       t += (size_t)gridDim.x * blockDim.x) {
    const size_t i0 = t * W;
    const int z0 = (int)(i0 % nz);
    const size_t c = i0 / nz;
    const bool lo = z0 > 0, hi = z0 + W < nz;
    T u_[W + 2], v_[W + 2], w_[W + 2], th_[W + 2], wx_[W + 2], wy_[W + 2];
    auto window = [&](const T* __restrict__ f, T* o) {
      const Vec16<T> m = vld(f + i0);
#pragma unroll
      for (int j = 0; j < W; ++j) o[1 + j] = m.v[j];
      o[0] = lo ? f[i0 - 1] : T(0);
      o[W + 1] = hi ? f[i0 + W] : T(0);
    };
    window(a.u, u_); window(a.v, v_); window(a.w, w_); window(a.TH, th_);
    window(a.dwdx, wx_); window(a.dwdy, wy_);
    const Vec16<T> ux = vld(a.dudx + i0), vx = vld(a.dvdx + i0), uy = vld(a.dudy + i0),
                   vy = vld(a.dvdy + i0), sx = vld(a.sx + i0), sy = vld(a.sy + i0);
    // accessors on the windows: local index 1 + j is level z0 + j
    const DzUV<T, true> uz{u_ + 1, a.fim + c, a.ustar + c, a.M + c, a.Ugal, a.idz,
                           a.half_dz_vonk, nz};
    const DzUV<T, false> vz{v_ + 1, a.fim + c, a.ustar + c, a.M + c, T(0), a.idz,
                            a.half_dz_vonk, nz};
    const DzW<T> wz{w_ + 1, a.idz, nz};
    const DzTH<T> tz{th_ + 1, a.fih + c, a.qz + c, a.ustar + c, a.idz, a.half_dz_vonk};
    Vec16<T> oUz, oVz, oWz, oTz, oCx, oCy, oCz, oR, o12, o13, o23, oS;
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const int z = z0 + j;
      const AdvOut<T> o = adv_point<T>((size_t)j, z, 0, nz, u_ + 1, v_ + 1, w_ + 1, uy.v, uz,
                                       vx.v, vz, wx_ + 1, wy_ + 1, sx.v, sy.v, tz);
      const StrainOut<T> s = strains_point<T>((size_t)j, z, 0, nz, ux.v, vx.v, wx_ + 1, uy.v,
                                              vy.v, wy_ + 1, uz, vz, wz);
      oUz.v[j] = uz((size_t)j, z, 0); oVz.v[j] = vz((size_t)j, z, 0);
      oWz.v[j] = wz((size_t)j, z, 0); oTz.v[j] = tz((size_t)j, z, 0);
      oCx.v[j] = o.cx; oCy.v[j] = o.cy; oCz.v[j] = o.cz; oR.v[j] = o.r;
      o12.v[j] = s.s12; o13.v[j] = s.s13; o23.v[j] = s.s23; oS.v[j] = s.S;
    }
    vst(a.dudz + i0, oUz); vst(a.dvdz + i0, oVz); vst(a.dwdz + i0, oWz); vst(a.dTdz + i0, oTz);
    vst(a.Cx + i0, oCx); vst(a.Cy + i0, oCy); vst(a.Cz + i0, oCz); vst(a.sadv + i0, oR);
    vst(a.S12 + i0, o12); vst(a.S13 + i0, o13); vst(a.S23 + i0, o23); vst(a.S + i0, oS);
  }
}

// StrainsWnodes (src/subgridscale/StrainRates.py:219-368).  S13w, S23w and
// |S_w| always; the diagonal/S12 w-node components only when S11w != NULL
// (the dealiased branch transforms all six).
template <class T>
__device__ __forceinline__ void k_strains_w(const PwDims& d, const T* __restrict__ dudx, const T* __restrict__ dvdx,
                            const T* __restrict__ dwdx, const T* __restrict__ dudy,
                            const T* __restrict__ dvdy, const T* __restrict__ dwdy,
                            const T* __restrict__ dudz, const T* __restrict__ dvdz,
                            const T* __restrict__ dwdz, T* S13w, T* S23w, T* Sw,
                            T* S11w, T* S22w, T* S33w, T* S12w) {
  const int nz = d.nz;
  PW_LOOP(d.n) {
    const int z = (int)(i % nz);
    T ux, uy, vx, vy, wz, wx, wy;
    if (z == 0) {
      ux = dudx[i]; uy = dudy[i]; vx = dvdx[i]; vy = dvdy[i]; wz = dwdz[i];
      wx = T(0.5) * (dwdx[i] + dwdx[i + 1]); wy = T(0.5) * (dwdy[i] + dwdy[i + 1]);
    } else if (z < nz - 1) {
      ux = T(0.5) * (dudx[i - 1] + dudx[i]); uy = T(0.5) * (dudy[i - 1] + dudy[i]);
      vx = T(0.5) * (dvdx[i - 1] + dvdx[i]); vy = T(0.5) * (dvdy[i - 1] + dvdy[i]);
      wz = T(0.5) * (dwdz[i - 1] + dwdz[i]);
      wx = dwdx[i]; wy = dwdy[i];
    } else {
      ux = T(0); uy = T(0); vx = T(0); vy = T(0); wz = T(0);
      wx = dwdx[i]; wy = dwdy[i];
    }
    const T s12 = T(0.5) * (uy + vx);
    const T s13 = T(0.5) * (dudz[i] + wx);
    const T s23 = T(0.5) * (dvdz[i] + wy);
    S13w[i] = s13; S23w[i] = s23;
    Sw[i] = smag<T>(ux, vy, wz, s12, s13, s23);
    if (S11w) { S11w[i] = ux; S22w[i] = vy; S33w[i] = wz; S12w[i] = s12; }
  }
}

// |S| from six components (the padded magnitudes of StrainRates.py:127-133).
template <class T>
__device__ __forceinline__ void k_smag(const PwDims& d, const T* a, const T* b, const T* c, const T* e,
                       const T* f, const T* g, T* out) {
  PW_LOOP(d.n) out[i] = smag<T>(a[i], b[i], c[i], e[i], f[i], g[i]);
}

// StressesUVPnodes (src/subgridscale/SGSStresses_SM.py:98-182) before any
// Dealias2: tij = -2 L^2 Cs2 |S| Sij, top level 0.  With S == NULL the
// Wong-Lilly form (SGSStresses_WL.py:70-141): m2L2 = -2 L^(4/3), no |S|.
template <class T>
__device__ __forceinline__ void k_stress_uvp(const PwDims& d, const T* __restrict__ S11, const T* __restrict__ S22,
                             const T* __restrict__ S33, const T* __restrict__ S12,
                             const T* __restrict__ S, const T* __restrict__ Cs2, T m2L2,
                             T* txx, T* tyy, T* tzz, T* txy) {
  const int nz = d.nz;
  PW_LOOP(d.n) {
    const int z = (int)(i % nz);
    if (z == nz - 1) { txx[i] = tyy[i] = tzz[i] = txy[i] = T(0); continue; }
    const T pre = m2L2 * Cs2[i] * (S ? S[i] : T(1));
    txx[i] = pre * S11[i]; tyy[i] = pre * S22[i]; tzz[i] = pre * S33[i]; txy[i] = pre * S12[i];
  }
}

// StressesWnodes (SGSStresses_SM.py:190-297) interior; level 0 gets the wall
// stress of Wall() (:63-90) when `wall` is set, else 0 (padded branch).
// Sw == NULL: Wong-Lilly (SGSStresses_WL.py:146-239).
template <class T>
__device__ __forceinline__ void k_stress_w(const PwDims& d, const T* __restrict__ S13w, const T* __restrict__ S23w,
                           const T* __restrict__ Sw, const T* __restrict__ Cs2, T m2L2,
                           T* txz, T* tyz, int wall, const T* __restrict__ u,
                           const T* __restrict__ v, const T* __restrict__ M,
                           const T* __restrict__ psim, const T* __restrict__ psim0,
                           double logz1z0, double vonk, double Ugal) {
  const int nz = d.nz;
  PW_LOOP(d.n) {
    const int z = (int)(i % nz);
    if (z == nz - 1) { txz[i] = tyz[i] = T(0); continue; }
    if (z == 0) {
      if (wall) {
        const size_t c = i / nz;
        const T denom = (T)logz1z0 + psim[c] - psim0[c];
        const T us = (T)vonk * M[c] / denom;
        txz[i] = -(us * us) * (u[i] + (T)Ugal) / M[c];
        tyz[i] = -(us * us) * v[i] / M[c];
      } else {
        txz[i] = tyz[i] = T(0);
      }
      continue;
    }
    const T pre = m2L2 * (T(0.5) * (Cs2[i - 1] + Cs2[i])) * (Sw ? Sw[i] : T(1));
    txz[i] = pre * S13w[i];
    tyz[i] = pre * S23w[i];
  }
}

// Wall stress written into level 0 of existing txz/tyz (dealiased branch,
// SGSStresses_SM.py:240-241), and qz level 0 = surface flux
// (ScalarSGSFluxes_SM.py:175).
template <class T>
__device__ __forceinline__ void k_level0(int ncol, int nz, T* txz, T* tyz, const T* __restrict__ u,
                         const T* __restrict__ v, const T* __restrict__ M,
                         const T* __restrict__ psim, const T* __restrict__ psim0,
                         double logz1z0, double vonk, double Ugal, T* qz,
                         const T* __restrict__ qsfc) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncol; c += gridDim.x * blockDim.x) {
    const size_t i = (size_t)c * nz;
    if (txz) {
      const T denom = (T)logz1z0 + psim[c] - psim0[c];
      const T us = (T)vonk * M[c] / denom;
      txz[i] = -(us * us) * (u[i] + (T)Ugal) / M[c];
      tyz[i] = -(us * us) * v[i] / M[c];
    }
    if (qz) qz[i] = qsfc[c];
  }
}

// ScalarFluxes (src/subgridscale/ScalarSGSFluxes_SM.py:55-222) before any
// Dealias2; qz level 0 = qsfc when qsfc != NULL, else 0.  S == Sw == NULL:
// Wong-Lilly (ScalarSGSFluxes_WL.py:55-201) with m2L2 = -L^(4/3).
template <class T>
__device__ __forceinline__ void k_scalar_flux(const PwDims& d, const T* __restrict__ dx, const T* __restrict__ dy,
                              const T* __restrict__ dz, const T* __restrict__ S,
                              const T* __restrict__ Sw, const T* __restrict__ CsPr, T m2L2,
                              T* qx, T* qy, T* qz, const T* __restrict__ qsfc) {
  const int nz = d.nz;
  PW_LOOP(d.n) {
    const int z = (int)(i % nz);
    if (z == nz - 1) { qx[i] = qy[i] = qz[i] = T(0); continue; }
    const T pre = m2L2 * CsPr[i] * (S ? S[i] : T(1));
    qx[i] = pre * dx[i];
    qy[i] = pre * dy[i];
    if (z == 0) qz[i] = qsfc ? qsfc[i / nz] : T(0);
    else qz[i] = m2L2 * (T(0.5) * (CsPr[i - 1] + CsPr[i])) * (Sw ? Sw[i] : T(1)) * dz[i];
  }
}

// Unpadded SGS stresses and scalar fluxes in one pass (FGR = 2): the w-node
// strain rates of k_strains_w are formed in registers, then the stresses of
// k_stress_uvp / k_stress_w (wall row included) and the fluxes of
// k_scalar_flux, with the same arithmetic.  Sm == NULL: Wong-Lilly (no |S|).
template <class T> struct SgsFusedArgs {
  const T *dudx, *dvdx, *dwdx, *dudy, *dvdy, *dwdy, *dudz, *dvdz, *dwdz;
  const T *S12, *Sm, *Cs2, *CsPr, *dTx, *dTy, *dTz;
  T pre_s, pre_q;
  const T *u, *v, *M, *psim, *psim0, *qsfc;
  double logz1z0, vonk, Ugal;
  T *txx, *tyy, *tzz, *txy, *txz, *tyz, *qx, *qy, *qz;
};

template <class T>
__device__ __forceinline__ void k_sgs_fused(const PwDims& d, const SgsFusedArgs<T>& a) {
  const int nz = d.nz;
  PW_LOOP(d.n) {
    const int z = (int)(i % nz);
    if (z == nz - 1) {
      a.txx[i] = a.tyy[i] = a.tzz[i] = a.txy[i] = T(0);
      a.txz[i] = a.tyz[i] = T(0);
      a.qx[i] = a.qy[i] = a.qz[i] = T(0);
      continue;
    }
    const T sm = a.Sm ? a.Sm[i] : T(1);
    const T ps = a.pre_s * a.Cs2[i] * sm;
    a.txx[i] = ps * a.dudx[i]; a.tyy[i] = ps * a.dvdy[i]; a.tzz[i] = ps * a.dwdz[i];
    a.txy[i] = ps * a.S12[i];
    const T pq = a.pre_q * a.CsPr[i] * sm;
    a.qx[i] = pq * a.dTx[i];
    a.qy[i] = pq * a.dTy[i];
    if (z == 0) {
      const size_t c = i / nz;
      const T denom = (T)a.logz1z0 + a.psim[c] - a.psim0[c];
      const T us = (T)a.vonk * a.M[c] / denom;
      a.txz[i] = -(us * us) * (a.u[i] + (T)a.Ugal) / a.M[c];
      a.tyz[i] = -(us * us) * a.v[i] / a.M[c];
      a.qz[i] = a.qsfc ? a.qsfc[c] : T(0);
      continue;
    }
    // w-node strains (k_strains_w, interior levels)
    const T ux = T(0.5) * (a.dudx[i - 1] + a.dudx[i]), uy = T(0.5) * (a.dudy[i - 1] + a.dudy[i]);
    const T vx = T(0.5) * (a.dvdx[i - 1] + a.dvdx[i]), vy = T(0.5) * (a.dvdy[i - 1] + a.dvdy[i]);
    const T wz = T(0.5) * (a.dwdz[i - 1] + a.dwdz[i]);
    const T s12 = T(0.5) * (uy + vx);
    const T s13 = T(0.5) * (a.dudz[i] + a.dwdx[i]);
    const T s23 = T(0.5) * (a.dvdz[i] + a.dwdy[i]);
    const T sw = a.Sm ? smag<T>(ux, vy, wz, s12, s13, s23) : T(1);
    const T pw = a.pre_s * (T(0.5) * (a.Cs2[i - 1] + a.Cs2[i])) * sw;
    a.txz[i] = pw * s13;
    a.tyz[i] = pw * s23;
    a.qz[i] = a.pre_q * (T(0.5) * (a.CsPr[i - 1] + a.CsPr[i])) * sw * a.dTz[i];
  }
}

// k_sgs_fused with W = 16 / sizeof(T) consecutive levels per thread and
// 16-byte accesses (nz % W == 0, all arrays 16-byte aligned); the lower
// neighbours of the first level come from one scalar load per array.
template <class T>
__device__ __forceinline__ void k_sgs_fused_v(const PwDims& d, const SgsFusedArgs<T>& a) {
  constexpr int W = Vec16<T>::W;
  const int nz = d.nz;
  const size_t nv = d.n / W;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < nv;
       t += (size_t)gridDim.x * blockDim.x) {
    const size_t i0 = t * W;
    const int z0 = (int)(i0 % nz);
    const Vec16<T> ux = vld(a.dudx + i0), vy = vld(a.dvdy + i0), wz = vld(a.dwdz + i0);
    const Vec16<T> s12 = vld(a.S12 + i0), cs = vld(a.Cs2 + i0), cp = vld(a.CsPr + i0);
    const Vec16<T> tx = vld(a.dTx + i0), ty = vld(a.dTy + i0), tz = vld(a.dTz + i0);
    const Vec16<T> uy = vld(a.dudy + i0), vx = vld(a.dvdx + i0);
    const Vec16<T> uz = vld(a.dudz + i0), vz = vld(a.dvdz + i0);
    const Vec16<T> wx = vld(a.dwdx + i0), wy = vld(a.dwdy + i0);
    const Vec16<T> sm = a.Sm ? vld(a.Sm + i0) : Vec16<T>{};
    // level z0 - 1 (used only when z0 > 0)
    T pux = 0, puy = 0, pvx = 0, pvy = 0, pwz = 0, pcs = 0, pcp = 0;
    if (z0 > 0) {
      pux = a.dudx[i0 - 1]; puy = a.dudy[i0 - 1]; pvx = a.dvdx[i0 - 1]; pvy = a.dvdy[i0 - 1];
      pwz = a.dwdz[i0 - 1]; pcs = a.Cs2[i0 - 1]; pcp = a.CsPr[i0 - 1];
    }
    Vec16<T> txx, tyy, tzz, txy, txz, tyz, qx, qy, qz;
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const int z = z0 + j;
      if (z == nz - 1) {
        txx.v[j] = tyy.v[j] = tzz.v[j] = txy.v[j] = T(0);
        txz.v[j] = tyz.v[j] = T(0);
        qx.v[j] = qy.v[j] = qz.v[j] = T(0);
        continue;
      }
      const T smv = a.Sm ? sm.v[j] : T(1);
      const T ps = a.pre_s * cs.v[j] * smv;
      txx.v[j] = ps * ux.v[j]; tyy.v[j] = ps * vy.v[j]; tzz.v[j] = ps * wz.v[j];
      txy.v[j] = ps * s12.v[j];
      const T pq = a.pre_q * cp.v[j] * smv;
      qx.v[j] = pq * tx.v[j];
      qy.v[j] = pq * ty.v[j];
      if (z == 0) {
        const size_t c = (i0 + j) / nz;
        const T denom = (T)a.logz1z0 + a.psim[c] - a.psim0[c];
        const T us = (T)a.vonk * a.M[c] / denom;
        txz.v[j] = -(us * us) * (a.u[i0 + j] + (T)a.Ugal) / a.M[c];
        tyz.v[j] = -(us * us) * a.v[i0 + j] / a.M[c];
        qz.v[j] = a.qsfc ? a.qsfc[c] : T(0);
        continue;
      }
      const T lux = j > 0 ? ux.v[j - 1] : pux, luy = j > 0 ? uy.v[j - 1] : puy;
      const T lvx = j > 0 ? vx.v[j - 1] : pvx, lvy = j > 0 ? vy.v[j - 1] : pvy;
      const T lwz = j > 0 ? wz.v[j - 1] : pwz;
      const T lcs = j > 0 ? cs.v[j - 1] : pcs, lcp = j > 0 ? cp.v[j - 1] : pcp;
      const T aux = T(0.5) * (lux + ux.v[j]), auy = T(0.5) * (luy + uy.v[j]);
      const T avx = T(0.5) * (lvx + vx.v[j]), avy = T(0.5) * (lvy + vy.v[j]);
      const T awz = T(0.5) * (lwz + wz.v[j]);
      const T w12 = T(0.5) * (auy + avx);
      const T w13 = T(0.5) * (uz.v[j] + wx.v[j]);
      const T w23 = T(0.5) * (vz.v[j] + wy.v[j]);
      const T sw = a.Sm ? smag<T>(aux, avy, awz, w12, w13, w23) : T(1);
      const T pw = a.pre_s * (T(0.5) * (lcs + cs.v[j])) * sw;
      txz.v[j] = pw * w13;
      tyz.v[j] = pw * w23;
      qz.v[j] = a.pre_q * (T(0.5) * (lcp + cp.v[j])) * sw * tz.v[j];
    }
    vst(a.txx + i0, txx); vst(a.tyy + i0, tyy); vst(a.tzz + i0, tzz); vst(a.txy + i0, txy);
    vst(a.txz + i0, txz); vst(a.tyz + i0, tyz);
    vst(a.qx + i0, qx); vst(a.qy + i0, qy); vst(a.qz + i0, qz);
  }
}

// RHS_Momentum (src/pde/NSE_AllTerms.py:48-104) + RHS_Scalar
// (src/pde/SCL_AllTerms.py:48-82), optAdvection = 0.  Profiles are double.
template <class T>
__device__ __forceinline__ void k_rhs(const PwDims& d, const T* __restrict__ u, const T* __restrict__ v,
                      const T* __restrict__ w, const T* __restrict__ TH,
                      const T* __restrict__ Cx, const T* __restrict__ Cy,
                      const T* __restrict__ Cz, const T* __restrict__ buoy,
                      const T* __restrict__ dtx, const T* __restrict__ dty,
                      const T* __restrict__ dtz, const T* __restrict__ thadv,
                      const T* __restrict__ divq, const double* __restrict__ Ug,
                      const double* __restrict__ Vg, const double* __restrict__ rf,
                      const double* __restrict__ rs, const double* __restrict__ ubar,
                      const double* __restrict__ vbar, const double* __restrict__ thbar,
                      double f, double Ugal, int damping, T* Ru, T* Rv, T* Rw, T* RT,
                      const double* __restrict__ bbar = nullptr, double T0 = 0.0,
                      double g = 0.0, int bopt = 0, const double* __restrict__ fadv = nullptr,
                      const T* __restrict__ Hq = nullptr) {
  const int nz = d.nz;
  PW_LOOP(d.n) {
    const int z = (int)(i % nz);
    if (Ru) {
      T ru = -Cx[i] - dtx[i] - (T)f * ((T)Vg[z] - v[i]);
      T rv = -Cy[i] - dty[i] + (T)f * ((T)Ug[z] - (T)Ugal - u[i]);
      if (fadv) {   // large-scale advection forcing (optAdvection >= 1)
        ru += (T)fadv[z];
        rv += (T)fadv[nz + z];
      }
      T b;
      if (buoy) {
        b = buoy[i];
      } else if (z == 0) {
        b = T(0);
      } else {
        // k_buoyancy inline: bbar = plane mean of THv = (TH + T_0)(1 + 0.61 Q)
        double t1 = (double)TH[i] + T0, t0 = (double)TH[i - 1] + T0;
        if (Hq) {
          t1 = t1 * (1.0 + 0.61 * (double)Hq[i]);
          t0 = t0 * (1.0 + 0.61 * (double)Hq[i - 1]);
        }
        double n1, n0;
        if (bopt == 0) { n1 = (t1 - bbar[z]) / T0; n0 = (t0 - bbar[z - 1]) / T0; }
        else { n1 = (t1 - bbar[z]) / bbar[z]; n0 = (t0 - bbar[z - 1]) / bbar[z - 1]; }
        b = (T)(0.5 * g * (n1 + n0));
      }
      T rw = -Cz[i] - dtz[i] + b;
      if (damping) {
        const double s = rs[z];
        ru = (T)((double)ru - s * ((double)u[i] - ubar[z]));
        rv = (T)((double)rv - s * ((double)v[i] - vbar[z]));
        rw = (T)((double)rw - rf[z] * (double)w[i]);
      }
      Ru[i] = ru; Rv[i] = rv; Rw[i] = rw;
    }
    if (RT) {
      T rt = -thadv[i] - divq[i];
      if (fadv) rt += (T)fadv[2 * nz + z];
      if (damping) rt = (T)((double)rt - rs[z] * ((double)TH[i] - thbar[z]));
      RT[i] = rt;
    }
  }
}

// k_rhs (both outputs, buoyancy formed inline) with W consecutive levels per
// thread and 16-byte accesses (nz % W == 0, arrays 16-byte aligned).  Same
// arithmetic, element by element.
template <class T>
__device__ __forceinline__ void k_rhs_v(const PwDims& d, const T* __restrict__ u, const T* __restrict__ v,
                        const T* __restrict__ w, const T* __restrict__ TH,
                        const T* __restrict__ Cx, const T* __restrict__ Cy,
                        const T* __restrict__ Cz, const T* __restrict__ dtx,
                        const T* __restrict__ dty, const T* __restrict__ dtz,
                        const T* __restrict__ thadv, const T* __restrict__ divq,
                        const double* __restrict__ Ug, const double* __restrict__ Vg,
                        const double* __restrict__ rf, const double* __restrict__ rs,
                        const double* __restrict__ ubar, const double* __restrict__ vbar,
                        const double* __restrict__ thbar, double f, double Ugal, int damping,
                        T* Ru, T* Rv, T* Rw, T* RT, const double* __restrict__ bbar, double T0,
                        double g, int bopt, const double* __restrict__ fadv,
                        const T* __restrict__ Hq) {
  constexpr int W = Vec16<T>::W;
  const int nz = d.nz;
  const size_t nv = d.n / W;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < nv;
       t += (size_t)gridDim.x * blockDim.x) {
    const size_t i0 = t * W;
    const int z0 = (int)(i0 % nz);
    const Vec16<T> uu = vld(u + i0), vv = vld(v + i0), ww = vld(w + i0), th = vld(TH + i0);
    const Vec16<T> cx = vld(Cx + i0), cy = vld(Cy + i0), cz = vld(Cz + i0);
    const Vec16<T> tx = vld(dtx + i0), ty = vld(dty + i0), tz = vld(dtz + i0);
    const Vec16<T> ta = vld(thadv + i0), dq = vld(divq + i0);
    const T thp = z0 > 0 ? TH[i0 - 1] : T(0);
    Vec16<T> hq;
    T hqp = T(0);
    if (Hq) {
      hq = vld(Hq + i0);
      hqp = z0 > 0 ? Hq[i0 - 1] : T(0);
    }
    Vec16<T> ru, rv, rw, rt;
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const int z = z0 + j;
      T a = -cx.v[j] - tx.v[j] - (T)f * ((T)Vg[z] - vv.v[j]);
      T b = -cy.v[j] - ty.v[j] + (T)f * ((T)Ug[z] - (T)Ugal - uu.v[j]);
      if (fadv) {
        a += (T)fadv[z];
        b += (T)fadv[nz + z];
      }
      T bu;
      if (z == 0) {
        bu = T(0);
      } else {
        double t1 = (double)th.v[j] + T0;
        double t0 = (double)(j > 0 ? th.v[j - 1] : thp) + T0;
        if (Hq) {
          t1 = t1 * (1.0 + 0.61 * (double)hq.v[j]);
          t0 = t0 * (1.0 + 0.61 * (double)(j > 0 ? hq.v[j - 1] : hqp));
        }
        double n1, n0;
        if (bopt == 0) { n1 = (t1 - bbar[z]) / T0; n0 = (t0 - bbar[z - 1]) / T0; }
        else { n1 = (t1 - bbar[z]) / bbar[z]; n0 = (t0 - bbar[z - 1]) / bbar[z - 1]; }
        bu = (T)(0.5 * g * (n1 + n0));
      }
      T c = -cz.v[j] - tz.v[j] + bu;
      T e = -ta.v[j] - dq.v[j];
      if (fadv) e += (T)fadv[2 * nz + z];
      if (damping) {
        const double sr = rs[z];
        a = (T)((double)a - sr * ((double)uu.v[j] - ubar[z]));
        b = (T)((double)b - sr * ((double)vv.v[j] - vbar[z]));
        c = (T)((double)c - rf[z] * (double)ww.v[j]);
        e = (T)((double)e - sr * ((double)th.v[j] - thbar[z]));
      }
      ru.v[j] = a; rv.v[j] = b; rw.v[j] = c; rt.v[j] = e;
    }
    vst(Ru + i0, ru); vst(Rv + i0, rv); vst(Rw + i0, rw); vst(RT + i0, rt);
  }
}

// Per-iteration forcing profiles into the plan's profile table: Ug, Vg
// (slots 0, 1) and Uadv, Vadv, THadv, Qadv (slots 4-7); NULL keeps a slot.
__device__ __forceinline__ void k_set_forcing(int nz, const double* __restrict__ Ug, const double* __restrict__ Vg,
                              const double* __restrict__ Ua, const double* __restrict__ Va,
                              const double* __restrict__ Ta, const double* __restrict__ Qa,
                              double* prof) {
  for (int z = blockIdx.x * blockDim.x + threadIdx.x; z < nz; z += gridDim.x * blockDim.x) {
    if (Ug) prof[z] = Ug[z];
    if (Vg) prof[nz + z] = Vg[z];
    if (Ua) prof[4 * nz + z] = Ua[z];
    if (Va) prof[5 * nz + z] = Va[z];
    if (Ta) prof[6 * nz + z] = Ta[z];
    if (Qa) prof[7 * nz + z] = Qa[z];
  }
}

// AB2_uvw (NSE_TimeAdvancement.py:45-76) / AB2_TH (SCL_TimeAdvancement.py:
// 46-71) alone, in place; the thread of level nz-2 writes the top level.
template <class T>
__device__ __forceinline__ void k_ab2_plain(const PwDims& d, T* F, const T* __restrict__ R, const T* __restrict__ Rp,
                            T dt, int kind, T th_top) {
  // kind 0: u/v (top copies nz-2), 1: w (0 at bottom and top), 2: TH
  const int nz = d.nz;
  PW_LOOP(d.n) {
    const int z = (int)(i % nz);
    if (kind == 1) {
      F[i] = (z == 0 || z == nz - 1) ? T(0) : F[i] + dt * (T(1.5) * R[i] - T(0.5) * Rp[i]);
      continue;
    }
    if (z == nz - 1) continue;
    const T fn = F[i] + dt * (T(1.5) * R[i] - T(0.5) * Rp[i]);
    F[i] = fn;
    if (z == nz - 2) F[i + 1] = (kind == 2) ? fn + th_top : fn;
  }
}

// AB2_Q (SCL_TimeAdvancement.py:79-99) in place, with the first-step rule
// (RHS_Q_previous = RHS_Q, Main.py:506-512) and the rotation of RHS_Q
// into the previous-step buffer (Main.py:164-165); level nz-1 becomes
// level nz-2 plus q_top = q_inversion dz.
template <class T>
__device__ __forceinline__ void k_ab2_q(const PwDims& d, T* Q, const T* __restrict__ R, T* Rp, T dt, T q_top,
                        int first) {
  const int nz = d.nz;
  PW_LOOP(d.n) {
    const int z = (int)(i % nz);
    const T r = R[i];
    const T rp = first ? r : Rp[i];
    Rp[i] = r;
    if (z == nz - 1) continue;
    const T qn = Q[i] + dt * (T(1.5) * r - T(0.5) * rp);
    Q[i] = qn;
    if (z == nz - 2) Q[i + 1] = qn + q_top;
  }
}

// 3x3 periodic box mean of every level (Utilities.py:87-115).
template <class T>
__device__ __forceinline__ void k_imfilter(const PwDims& d, int nx, int ny, const T* __restrict__ F, T* out) {
  const int nz = d.nz;
  PW_LOOP(d.n) {
    const int z = (int)(i % nz);
    const size_t c = i / nz;
    const int x = (int)(c / ny), y = (int)(c % ny);
    const int xs[3] = {(x + nx - 1) % nx, x, (x + 1) % nx};
    const int ys[3] = {(y + ny - 1) % ny, y, (y + 1) % ny};
    const T inv9 = T(1.0 / 9.0);
    T a = T(0);
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q) a += F[((size_t)xs[p] * ny + ys[q]) * nz + z] * inv9;
    out[i] = a;
  }
}

// Wall stress planes (SGSStresses_SM.py:63-90).
template <class T>
__device__ __forceinline__ void k_wall_plane(int ncol, int nz, const T* __restrict__ u, const T* __restrict__ v,
                             const T* __restrict__ M, const T* __restrict__ psim,
                             const T* __restrict__ psim0, double logz1z0, double vonk,
                             double Ugal, T* txz, T* tyz) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncol; c += gridDim.x * blockDim.x) {
    const size_t i = (size_t)c * nz;
    const T denom = (T)logz1z0 + psim[c] - psim0[c];
    const T us = (T)vonk * M[c] / denom;
    txz[c] = -(us * us) * (u[i] + (T)Ugal) / M[c];
    tyz[c] = -(us * us) * v[i] / M[c];
  }
}

// Real part of a spectrum (fRz_real of NSE_PressureTerms.py:182).
template <class T>
__device__ __forceinline__ void k_real_part(size_t n, const c2<T>* __restrict__ a, T* out) {
  PW_LOOP(n) out[i] = a[i].x;
}

// RHS -= grad p (src/Main.py:498-500), then AB2_uvw and AB2_TH
// (NSE_TimeAdvancement.py:45-76, SCL_TimeAdvancement.py:46-71) in place.
// Rn* hold this step's RHS before the pressure correction; Rp* hold the
// previous step's RHS on entry and this step's corrected RHS on exit (the
// next step's RHS_previous).  On the first step the previous RHS is the
// current one (Main.py:506-512).  Each element is read and written by the
// same thread, except that the thread of level nz-2 also writes the top
// level (u, v copy; TH + inversion*dz), which no other thread touches.
// Per-block max |u|,|v|,|w| of the new fields go to cflpart[3*block + k].
template <class T>
__device__ __forceinline__ void k_ab2(const PwDims& d, T* u, T* v, T* w, T* TH, const T* __restrict__ Rnu,
                      const T* __restrict__ Rnv, const T* __restrict__ Rnw,
                      const T* __restrict__ RnT, T* Rpu, T* Rpv, T* Rpw, T* RpT,
                      const T* __restrict__ dpdx, const T* __restrict__ dpdy,
                      const T* __restrict__ dpdz, T dt, T th_top, int first,
                      double* cflpart) {
  const int nz = d.nz;
  T mu = T(0), mv = T(0), mw = T(0);   // maxima in the field type (exact)
  PW_LOOP(d.n) {
    const int z = (int)(i % nz);
    const T ru = Rnu[i] - dpdx[i];
    const T rv = Rnv[i] - dpdy[i];
    const T rw = Rnw[i] - dpdz[i];
    const T rt = RnT[i];
    const T rpu = first ? ru : Rpu[i];
    const T rpv = first ? rv : Rpv[i];
    const T rpw = first ? rw : Rpw[i];
    const T rpt = first ? rt : RpT[i];
    Rpu[i] = ru; Rpv[i] = rv; Rpw[i] = rw; RpT[i] = rt;
    if (z < nz - 1) {
      const T un = u[i] + dt * (T(1.5) * ru - T(0.5) * rpu);
      const T vn = v[i] + dt * (T(1.5) * rv - T(0.5) * rpv);
      const T tn = TH[i] + dt * (T(1.5) * rt - T(0.5) * rpt);
      u[i] = un; v[i] = vn; TH[i] = tn;
      if (z == nz - 2) { u[i + 1] = un; v[i + 1] = vn; TH[i + 1] = tn + th_top; }
      mu = fmax(mu, fabs(un));
      mv = fmax(mv, fabs(vn));
    }
    if (z == 0 || z == nz - 1) {
      w[i] = T(0);
    } else {
      const T wn = w[i] + dt * (T(1.5) * rw - T(0.5) * rpw);
      w[i] = wn;
      mw = fmax(mw, fabs(wn));
    }
  }
  __shared__ T sm[3][256];
  sm[0][threadIdx.x] = mu; sm[1][threadIdx.x] = mv; sm[2][threadIdx.x] = mw;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < 3; ++k) sm[k][threadIdx.x] = fmax(sm[k][threadIdx.x], sm[k][threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) cflpart[3 * blockIdx.x + k] = sm[k][0];
}

// k_ab2 with W consecutive levels per thread and 16-byte accesses (nz % W
// == 0, arrays 16-byte aligned); the top level copies level nz-2 from the
// same vector.  Same arithmetic and CFL maxima.
template <class T>
__device__ __forceinline__ void k_ab2_v(const PwDims& d, T* u, T* v, T* w, T* TH, const T* __restrict__ Rnu,
                        const T* __restrict__ Rnv, const T* __restrict__ Rnw,
                        const T* __restrict__ RnT, T* Rpu, T* Rpv, T* Rpw, T* RpT,
                        const T* __restrict__ dpdx, const T* __restrict__ dpdy,
                        const T* __restrict__ dpdz, T dt, T th_top, int first,
                        double* cflpart) {
  constexpr int W = Vec16<T>::W;
  const int nz = d.nz;
  const size_t nv = d.n / W;
  T mu = T(0), mv = T(0), mw = T(0);   // maxima in the field type (exact)
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < nv;
       t += (size_t)gridDim.x * blockDim.x) {
    const size_t i0 = t * W;
    const int z0 = (int)(i0 % nz);
    const Vec16<T> nu = vld(Rnu + i0), nv_ = vld(Rnv + i0), nw = vld(Rnw + i0), nt = vld(RnT + i0);
    const Vec16<T> px = vld(dpdx + i0), py = vld(dpdy + i0), pz = vld(dpdz + i0);
    Vec16<T> ru, rv, rw, rpu, rpv, rpw, rpt;
#pragma unroll
    for (int j = 0; j < W; ++j) {
      ru.v[j] = nu.v[j] - px.v[j];
      rv.v[j] = nv_.v[j] - py.v[j];
      rw.v[j] = nw.v[j] - pz.v[j];
    }
    if (first) {
      rpu = ru; rpv = rv; rpw = rw; rpt = nt;
    } else {
      rpu = vld(Rpu + i0); rpv = vld(Rpv + i0); rpw = vld(Rpw + i0); rpt = vld(RpT + i0);
    }
    vst(Rpu + i0, ru); vst(Rpv + i0, rv); vst(Rpw + i0, rw); vst(RpT + i0, nt);
    Vec16<T> uu = vld(u + i0), vv = vld(v + i0), tt = vld(TH + i0), ww = vld(w + i0);
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const int z = z0 + j;
      if (z < nz - 1) {
        uu.v[j] = uu.v[j] + dt * (T(1.5) * ru.v[j] - T(0.5) * rpu.v[j]);
        vv.v[j] = vv.v[j] + dt * (T(1.5) * rv.v[j] - T(0.5) * rpv.v[j]);
        tt.v[j] = tt.v[j] + dt * (T(1.5) * nt.v[j] - T(0.5) * rpt.v[j]);
        mu = fmax(mu, fabs(uu.v[j]));
        mv = fmax(mv, fabs(vv.v[j]));
      } else if (j > 0) {   // z == nz - 1: copies of level nz - 2
        uu.v[j] = uu.v[j - 1];
        vv.v[j] = vv.v[j - 1];
        tt.v[j] = tt.v[j - 1] + th_top;
      }
      if (z == 0 || z == nz - 1) {
        ww.v[j] = T(0);
      } else {
        ww.v[j] = ww.v[j] + dt * (T(1.5) * rw.v[j] - T(0.5) * rpw.v[j]);
        mw = fmax(mw, fabs(ww.v[j]));
      }
    }
    vst(u + i0, uu); vst(v + i0, vv); vst(TH + i0, tt); vst(w + i0, ww);
  }
  __shared__ T sm[3][256];
  sm[0][threadIdx.x] = mu; sm[1][threadIdx.x] = mv; sm[2][threadIdx.x] = mw;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < 3; ++k) sm[k][threadIdx.x] = fmax(sm[k][threadIdx.x], sm[k][threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) cflpart[3 * blockIdx.x + k] = sm[k][0];
}

template <class T>
__device__ __forceinline__ void k_copy(size_t n, const T* __restrict__ a, T* b) {
  PW_LOOP(n) b[i] = a[i];
}

template <class T>
__device__ __forceinline__ void k_fill(size_t n, T* a, T val) {
  PW_LOOP(n) a[i] = val;
}
