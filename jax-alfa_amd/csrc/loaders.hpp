// Fused R2C loaders: the product fields of the dynamic procedure are formed
// while the y-transform loads its tile, so they never exist in HBM.
#pragma once
#include "common.hpp"
#include <type_traits>

// The 21 momentum fields of DynamicSGS_LASDD_SM.py:184-233:
//   0..2  u_, v_, w_ (w_ = w averaged to uvp nodes, :187-190)
//   3..8  uu, vv, ww, uv, uw, vw
//   9..14 S11, S22, S33, S12, S13, S23
//   15..20 |S| S11 .. |S| S23
// In mixed precision the velocities are shifted by a per-level constant
// (sh*, an estimate of the plane mean) before the products are formed; the
// Leonard terms L_ij = (u_i u_j)^ - u_i^ u_j^ are exactly invariant under such
// a shift because the cut-off filters keep the (0,0) mode, and the shift
// removes the cancellation between the two terms in single precision.
template <class T> struct LoadLasdd {
  static constexpr const char* NAME = "LoadLasdd";
  static constexpr int BATCH_FAST = 1;
  static int nin(int) { return 10; }
  const T* u; const T* v; const T* w;
  const T* S[6]; const T* Sm;
  const double* shu; const double* shv;   // may be NULL
  int ny, nz;
  int b0;   // field b of the launch is virtual field b + b0
  __device__ __forceinline__ T wbar(size_t i, int z) const {
    if (z == 0) return T(0.5) * w[i + 1];
    if (z == nz - 1) return w[i - 1];
    return T(0.5) * (w[i] + w[i + 1]);
  }
  __device__ __forceinline__ T operator()(int b, int x, int y, int z) const {
    b += b0;
    const size_t i = ((size_t)x * ny + y) * nz + z;
    if (b >= 15) return Sm[i] * S[b - 15][i];
    if (b >= 9) return S[b - 9][i];
    const T uu = shu ? (T)((double)u[i] - shu[z]) : u[i];
    const T vv = shv ? (T)((double)v[i] - shv[z]) : v[i];
    const T ww = wbar(i, z);
    switch (b) {
      case 0: return uu;
      case 1: return vv;
      case 2: return ww;
      case 3: return uu * uu;
      case 4: return vv * vv;
      case 5: return ww * ww;
      case 6: return uu * vv;
      case 7: return uu * ww;
      default: return vv * ww;
    }
  }
  // 16-byte loads.  The field of a tile is uniform over the workgroup, so
  // the kernel picks the field's kind once per tile (with_kind) and runs a
  // load loop compiled for that kind (vload_k): no per-element branching on
  // the field.  Kinds: 0 u_, 1 v_, 2 w_, 3 uu, 4 vv, 5 ww, 6 uv, 7 uw, 8 vw,
  // 9 S_ij, 10 |S| S_ij.
  static constexpr int KINDS = 11;
  template <class F> __device__ __forceinline__ void with_kind(int b, F&& f) const {
    b += b0;
    switch (b < 9 ? b : (b < 15 ? 9 : 10)) {
      case 0: f(std::integral_constant<int, 0>{}); break;
      case 1: f(std::integral_constant<int, 1>{}); break;
      case 2: f(std::integral_constant<int, 2>{}); break;
      case 3: f(std::integral_constant<int, 3>{}); break;
      case 4: f(std::integral_constant<int, 4>{}); break;
      case 5: f(std::integral_constant<int, 5>{}); break;
      case 6: f(std::integral_constant<int, 6>{}); break;
      case 7: f(std::integral_constant<int, 7>{}); break;
      case 8: f(std::integral_constant<int, 8>{}); break;
      case 9: f(std::integral_constant<int, 9>{}); break;
      default: f(std::integral_constant<int, 10>{}); break;
    }
  }
  // a velocity with its level shift (W consecutive z from z, z % W == 0)
  __device__ __forceinline__ Vec16<T> vshift(const T* f, const double* sh, size_t i, int z) const {
    Vec16<T> r = vld(f + i);
    if (sh) {
#pragma unroll
      for (int j = 0; j < Vec16<T>::W; ++j) r.v[j] = (T)((double)r.v[j] - sh[z + j]);
    }
    return r;
  }
  // w averaged to uvp nodes: needs w[z .. z+W]
  __device__ __forceinline__ Vec16<T> vwbar(size_t i, int z) const {
    constexpr int W = Vec16<T>::W;
    const Vec16<T> w0 = vld(w + i);
    const T wn = (z + W < nz) ? w[i + W] : T(0);
    Vec16<T> r;
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const int zz = z + j;
      const T a = w0.v[j], c = (j + 1 < W) ? w0.v[j + 1] : wn;
      if (zz == 0) r.v[j] = T(0.5) * c;
      else if (zz == nz - 1) r.v[j] = (j > 0) ? w0.v[j - 1] : w[i - 1];
      else r.v[j] = T(0.5) * (a + c);
    }
    return r;
  }
  template <int K> __device__ __forceinline__ Vec16<T> vload_k(int b, int x, int y, int z) const {
    constexpr int W = Vec16<T>::W;
    const size_t i = ((size_t)x * ny + y) * nz + z;
    Vec16<T> r;
    if constexpr (K == 0) return vshift(u, shu, i, z);
    else if constexpr (K == 1) return vshift(v, shv, i, z);
    else if constexpr (K == 2) return vwbar(i, z);
    else if constexpr (K == 3) {
      const Vec16<T> a = vshift(u, shu, i, z);
#pragma unroll
      for (int j = 0; j < W; ++j) r.v[j] = a.v[j] * a.v[j];
      return r;
    } else if constexpr (K == 4) {
      const Vec16<T> a = vshift(v, shv, i, z);
#pragma unroll
      for (int j = 0; j < W; ++j) r.v[j] = a.v[j] * a.v[j];
      return r;
    } else if constexpr (K == 5) {
      const Vec16<T> a = vwbar(i, z);
#pragma unroll
      for (int j = 0; j < W; ++j) r.v[j] = a.v[j] * a.v[j];
      return r;
    } else if constexpr (K == 6) {
      const Vec16<T> a = vshift(u, shu, i, z), c = vshift(v, shv, i, z);
#pragma unroll
      for (int j = 0; j < W; ++j) r.v[j] = a.v[j] * c.v[j];
      return r;
    } else if constexpr (K == 7) {
      const Vec16<T> a = vshift(u, shu, i, z), c = vwbar(i, z);
#pragma unroll
      for (int j = 0; j < W; ++j) r.v[j] = a.v[j] * c.v[j];
      return r;
    } else if constexpr (K == 8) {
      const Vec16<T> a = vshift(v, shv, i, z), c = vwbar(i, z);
#pragma unroll
      for (int j = 0; j < W; ++j) r.v[j] = a.v[j] * c.v[j];
      return r;
    } else if constexpr (K == 9) {
      return vld(S[b + b0 - 9] + i);
    } else {
      const Vec16<T> a = vld(S[b + b0 - 15] + i), m = vld(Sm + i);
#pragma unroll
      for (int j = 0; j < W; ++j) r.v[j] = a.v[j] * m.v[j];
      return r;
    }
  }
  bool aligned(int) const {
    for (int n = 0; n < 6; ++n)
      if (!al16(S[n])) return false;
    return al16(u) && al16(v) && al16(w) && al16(Sm);
  }
};

// The 10 scalar fields of DynamicSGS_ScalarLASDD_SM.py:182-221:
//   0 TH_, 1..3 u_TH, v_TH, w_TH, 4..6 dTHdx, dTHdy, THz, 7..9 |S| x (4..6)
// THz is dTHdz averaged to uvp nodes (:182-186).  Same shift as above for
// u_, v_ and TH (the scalar Leonard flux is shift invariant too).
template <class T> struct LoadScalar {
  static constexpr const char* NAME = "LoadScalar";
  static constexpr int BATCH_FAST = 1;
  static int nin(int) { return 8; }
  const T* u; const T* v; const T* w; const T* TH;
  const T* dx; const T* dy; const T* dz; const T* Sm;
  const double* shu; const double* shv; const double* sht;
  int ny, nz;
  int b0;      // field b of the launch is virtual field b + b0
  int w_uvp;   // w is given on uvp nodes already (no averaging)
  __device__ __forceinline__ T thz(size_t i, int z) const {
    if (z == 0 || z == nz - 1) return dz[i];
    return T(0.5) * (dz[i] + dz[i + 1]);
  }
  __device__ __forceinline__ T operator()(int b, int x, int y, int z) const {
    b += b0;
    const size_t i = ((size_t)x * ny + y) * nz + z;
    switch (b) {
      case 4: return dx[i];
      case 5: return dy[i];
      case 6: return thz(i, z);
      case 7: return Sm[i] * dx[i];
      case 8: return Sm[i] * dy[i];
      case 9: return Sm[i] * thz(i, z);
      default: break;
    }
    const T th = sht ? (T)((double)TH[i] - sht[z]) : TH[i];
    if (b == 0) return th;
    if (b == 1) return (shu ? (T)((double)u[i] - shu[z]) : u[i]) * th;
    if (b == 2) return (shv ? (T)((double)v[i] - shv[z]) : v[i]) * th;
    T ww;
    if (w_uvp) ww = w[i];
    else if (z == 0) ww = T(0.5) * w[i + 1];
    else if (z == nz - 1) ww = w[i - 1];
    else ww = T(0.5) * (w[i] + w[i + 1]);
    return ww * th;
  }
  // v averaged from the staggered field f (z .. z+W): the THz / w_ rule
  __device__ __forceinline__ Vec16<T> stag(const T* f, size_t i, int z, bool wrule) const {
    constexpr int W = Vec16<T>::W;
    const Vec16<T> f0 = vld(f + i);
    const T fn = (z + W < nz) ? f[i + W] : T(0);
    Vec16<T> r;
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const int zz = z + j;
      const T a = f0.v[j], c = (j + 1 < W) ? f0.v[j + 1] : fn;
      if (wrule) {
        if (zz == 0) r.v[j] = T(0.5) * c;
        else if (zz == nz - 1) r.v[j] = (j > 0) ? f0.v[j - 1] : f[i - 1];
        else r.v[j] = T(0.5) * (a + c);
      } else {
        r.v[j] = (zz == 0 || zz == nz - 1) ? a : T(0.5) * (a + c);
      }
    }
    return r;
  }
  __device__ __forceinline__ Vec16<T> vshift(const T* f, const double* sh, size_t i, int z) const {
    Vec16<T> r = vld(f + i);
    if (sh) {
#pragma unroll
      for (int j = 0; j < Vec16<T>::W; ++j) r.v[j] = (T)((double)r.v[j] - sh[z + j]);
    }
    return r;
  }
  // 16-byte loads, one load loop per kind of field (as LoadLasdd): 0 TH_,
  // 1 u_TH, 2 v_TH, 3 w_TH, 4 dTHdx or dTHdy, 6 THz, 7 |S| dTHdx or |S|
  // dTHdy, 9 |S| THz
  static constexpr int KINDS = 8;
  template <class F> __device__ __forceinline__ void with_kind(int b, F&& f) const {
    b += b0;
    switch (b) {
      case 0: f(std::integral_constant<int, 0>{}); break;
      case 1: f(std::integral_constant<int, 1>{}); break;
      case 2: f(std::integral_constant<int, 2>{}); break;
      case 3: f(std::integral_constant<int, 3>{}); break;
      case 4: case 5: f(std::integral_constant<int, 4>{}); break;
      case 6: f(std::integral_constant<int, 6>{}); break;
      case 7: case 8: f(std::integral_constant<int, 7>{}); break;
      default: f(std::integral_constant<int, 9>{}); break;
    }
  }
  template <int K> __device__ __forceinline__ Vec16<T> vload_k(int b, int x, int y, int z) const {
    constexpr int W = Vec16<T>::W;
    b += b0;
    const size_t i = ((size_t)x * ny + y) * nz + z;
    if constexpr (K == 4) return vld((b == 4 ? dx : dy) + i);
    else if constexpr (K == 6) return stag(dz, i, z, false);
    else if constexpr (K == 7 || K == 9) {
      Vec16<T> r = K == 7 ? vld((b == 7 ? dx : dy) + i) : stag(dz, i, z, false);
      const Vec16<T> m = vld(Sm + i);
#pragma unroll
      for (int j = 0; j < W; ++j) r.v[j] *= m.v[j];
      return r;
    } else {
      const Vec16<T> th = vshift(TH, sht, i, z);
      if constexpr (K == 0) return th;
      else {
        Vec16<T> f;
        if constexpr (K == 3) f = w_uvp ? vld(w + i) : stag(w, i, z, true);
        else if constexpr (K == 1) f = vshift(u, shu, i, z);
        else f = vshift(v, shv, i, z);
        Vec16<T> r;
#pragma unroll
        for (int j = 0; j < W; ++j) r.v[j] = f.v[j] * th.v[j];
        return r;
      }
    }
  }
  bool aligned(int) const {
    return al16(u) && al16(v) && al16(w) && al16(TH) && al16(dx) && al16(dy) && al16(dz) &&
           al16(Sm);
  }
};

// Stress / flux divergence epilogue of the y-C2R: out = spectral part +
// the z-derivative of one field (Derivz_Generic_w for txz, tyz, qz;
// Derivz_Generic_uvp for tzz), NSE_SGSTerms.py:86-102, SCL_SGSTerms.py:82-87.
template <class T> struct StoreDiv {
  static constexpr const char* NAME = "StoreDiv";
  static constexpr int NOUT = 2;   // the output and the z-differenced field
  T* out[4]; const T* fz[4]; int wnode[4];
  int ny, nz; T idz;
  __device__ __forceinline__ void operator()(int b, int x, int y, int z, T v) const {
    const size_t i = ((size_t)x * ny + y) * nz + z;
    const T* f = fz[b];
    T dzv;
    if (wnode[b]) dzv = (z < nz - 1) ? (f[i + 1] - f[i]) * idz : T(0);
    else dzv = (z > 0) ? (f[i] - f[i - 1]) * idz : T(0);
    out[b][i] = v + dzv;
  }
  __device__ __forceinline__ void vstore(int b, int x, int y, int z, const Vec16<T>& v) const {
    constexpr int W = Vec16<T>::W;
    const size_t i = ((size_t)x * ny + y) * nz + z;
    const T* f = fz[b];
    const Vec16<T> f0 = vld(f + i);
    Vec16<T> r;
    if (wnode[b]) {
      const T fn = (z + W < nz) ? f[i + W] : T(0);
#pragma unroll
      for (int j = 0; j < W; ++j) {
        const T c = (j + 1 < W) ? f0.v[j + 1] : fn;
        r.v[j] = v.v[j] + ((z + j < nz - 1) ? (c - f0.v[j]) * idz : T(0));
      }
    } else {
      const T fp = (z > 0) ? f[i - 1] : T(0);
#pragma unroll
      for (int j = 0; j < W; ++j) {
        const T c = (j > 0) ? f0.v[j - 1] : fp;
        r.v[j] = v.v[j] + ((z + j > 0) ? (f0.v[j] - c) * idz : T(0));
      }
    }
    vst(out[b] + i, r);
  }
  // z and z + 1 (z even, nz even, aligned buffers)
  __device__ __forceinline__ void store2(int b, int x, int y, int z, T v0, T v1) const {
    const size_t i = ((size_t)x * ny + y) * nz + z;
    const T* f = fz[b];
    const c2<T> f01 = *reinterpret_cast<const c2<T>*>(f + i);
    T d0, d1;
    if (wnode[b]) {
      d0 = (z < nz - 1) ? (f01.y - f01.x) * idz : T(0);
      d1 = (z + 1 < nz - 1) ? (f[i + 2] - f01.y) * idz : T(0);
    } else {
      d0 = (z > 0) ? (f01.x - f[i - 1]) * idz : T(0);
      d1 = (f01.y - f01.x) * idz;
    }
    *reinterpret_cast<c2<T>*>(out[b] + i) = mkc<T>(v0 + d0, v1 + d1);
  }
  bool aligned(int nb) const {
    for (int b = 0; b < nb; ++b)
      if (!al16(out[b]) || !al16(fz[b])) return false;
    return true;
  }
};
