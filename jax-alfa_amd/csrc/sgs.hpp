// Plane reductions (PlanarMean, src/utilities/Utilities.py:41) and the
// dynamic-SGS contraction / box-filter / beta kernels
// (src/subgridscale/DynamicSGS_LASDD_SM.py, DynamicSGS_ScalarLASDD_SM.py).
//
// Reduction layout: a block of 256 threads is ZB z-lanes x (256/ZB) column
// lanes; block (bz, bc) owns levels [bz*ZB, bz*ZB+ZB) and a fixed range of
// columns.  Each thread accumulates in double over its columns (coalesced:
// the z-lanes of one column are contiguous), the column lanes are combined
// in LDS in a fixed order and each block writes one partial per (term,
// level).  k_plane_final adds the partials in block order, so every plane
// mean is bitwise reproducible run to run and independent of what else the
// device is doing.
#pragma once
#include "common.hpp"
#include "pointwise.hpp"

constexpr int RED_BS = 256;
// columns per thread of the plane reductions: short runs (more workgroups)
// for float fields, long runs for double (measured at 256^3)
inline int red_cpt(size_t esize) { return esize == 4 ? 16 : 64; }
constexpr int RED_MAXT = 10;

struct RedGeom {
  int ncol, nz, zb, lanes, cpb, nbz, nbc;
};

inline RedGeom red_geom(int ncol, int nz, size_t esize) {
  RedGeom g;
  g.ncol = ncol;
  g.nz = nz;
  g.zb = nz <= 32 ? 32 : 64;
  g.lanes = RED_BS / g.zb;
  g.nbz = (nz + g.zb - 1) / g.zb;
  g.cpb = g.lanes * red_cpt(esize);
  g.nbc = (ncol + g.cpb - 1) / g.cpb;
  return g;
}

// f(c, z, acc) adds NT term values of point (c, z) into acc[0..NT).
// Blocks are renumbered so that each XCD owns one contiguous range of
// column blocks: neighbouring columns (the box filter's x+-1 rows) share an
// L2.  The partials keep their (column block, term, level) slots, so the sums
// do not depend on the numbering.
__device__ __forceinline__ void red_block(int& bz, int& bc) {
  const unsigned tot = gridDim.x * gridDim.y;
  const unsigned lg = xcd_logical(blockIdx.x + blockIdx.y * gridDim.x, tot);
  bz = (int)(lg % gridDim.x);
  bc = (int)(lg / gridDim.x);
}

template <int NT, class F>
__device__ __forceinline__ void k_plane_partial(const RedGeom& g, const F& f, double* part) {
  __shared__ double sm[NT][RED_BS];
  int bz, bc;
  red_block(bz, bc);
  const int zl = threadIdx.x % g.zb, cl = threadIdx.x / g.zb;
  const int z = bz * g.zb + zl;
  double acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = 0.0;
  if (z < g.nz) {
    const int c0 = bc * g.cpb;
    const int c1 = imin(c0 + g.cpb, g.ncol);
    if constexpr (NT <= 5 && F::NFIELDS <= 5) {
      // few fields and terms: unrolled, so that the loads of several columns are in
      // flight at once (the additions keep their order)
#pragma unroll 4
      for (int c = c0 + cl; c < c1; c += g.lanes) f(c, z, acc);
    } else {
      for (int c = c0 + cl; c < c1; c += g.lanes) f(c, z, acc);
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) sm[t][threadIdx.x] = acc[t];
  __syncthreads();
  if (cl == 0 && z < g.nz) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      double s = 0.0;
      for (int l = 0; l < g.lanes; ++l) s += sm[t][l * g.zb + zl];
      part[((size_t)bc * NT + t) * g.nz + z] = s;
    }
  }
}

// out[t*nz + z] = sum_b part[b][t][z] * scale; one wavefront per output,
// lane-strided partial sums and a fixed shuffle tree (deterministic).
__device__ __forceinline__ void k_plane_final(int nbc, int nt, int nz, const double* __restrict__ part,
                              double* out, double scale) {
  const int i = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (i >= nt * nz) return;
  const int t = i / nz, z = i % nz;
  double s = 0.0;
  for (int b = lane; b < nbc; b += 64) s += part[((size_t)b * nt + t) * nz + z];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) out[i] = s * scale;
}

// ---------------------------------------------------------------------------
// plane-sum functors
// ---------------------------------------------------------------------------
template <class T> struct FSum1 {   // sum of F + add
  static constexpr const char* NAME = "FSum1";
  static constexpr int NFIELDS = 1;
  const T* F; int nz; double add;
  __device__ __forceinline__ void operator()(int c, int z, double* acc) const {
    acc[0] += (double)F[(size_t)c * nz + z] + add;
  }
};
template <class T> struct FSum3 {   // u, v, TH + T0 (RHS damping and buoyancy means)
  static constexpr const char* NAME = "FSum3";
  static constexpr int NFIELDS = 3;
  const T* a; const T* b; const T* c3; int nz; double add3;
  __device__ __forceinline__ void operator()(int c, int z, double* acc) const {
    const size_t i = (size_t)c * nz + z;
    acc[0] += (double)a[i];
    acc[1] += (double)b[i];
    acc[2] += (double)c3[i] + add3;
  }
};

template <class T> __device__ __forceinline__ double con6(const double* A, const double* B) {
  return A[0] * B[0] + A[1] * B[1] + A[2] * B[2] + 2.0 * (A[3] * B[3] + A[4] * B[4] + A[5] * B[5]);
}

// Filtered momentum fields of one test-filter level: u_,v_,w_ hats (3), the
// six u_i u_j products (uu,vv,ww,uv,uw,vw), S_ij (6) and |S| S_ij (6).
template <class T> struct MomLevel {
  const T* uh[3];
  const T* ph[6];
  const T* Sh[6];
  const T* SSh[6];
};

template <class T>
__device__ __forceinline__ void mom_load(const MomLevel<T>& m, size_t i, double* Lij,
                                         double* Sij, double* SSij, double& Sm) {
  const double u0 = m.uh[0][i], u1 = m.uh[1][i], u2 = m.uh[2][i];
  Lij[0] = (double)m.ph[0][i] - u0 * u0;
  Lij[1] = (double)m.ph[1][i] - u1 * u1;
  Lij[2] = (double)m.ph[2][i] - u2 * u2;
  Lij[3] = (double)m.ph[3][i] - u0 * u1;
  Lij[4] = (double)m.ph[4][i] - u0 * u2;
  Lij[5] = (double)m.ph[5][i] - u1 * u2;
#pragma unroll
  for (int n = 0; n < 6; ++n) { Sij[n] = m.Sh[n][i]; SSij[n] = m.SSh[n][i]; }
  Sm = sqrt(2.0 * (Sij[0] * Sij[0] + Sij[1] * Sij[1] + Sij[2] * Sij[2] + 2.0 * Sij[3] * Sij[3] +
                   2.0 * Sij[4] * Sij[4] + 2.0 * Sij[5] * Sij[5]));
}

// The ten plane means a1..e2 of DynamicSGS_LASDD_SM.py:266-312 (sums; the
// caller scales by 1/(nx*ny)).  Arithmetic in double.
template <class T> struct FLasddSums {
  static constexpr const char* NAME = "FLasddSums";
  static constexpr int NFIELDS = 42;
  MomLevel<T> l1, l2;
  int nz;
  double L, TFR;
  __device__ void operator()(int c, int z, double* acc) const {
    const size_t i = (size_t)c * nz + z;
    double L1[6], S1[6], SS1[6], Sh, L2[6], S2[6], SS2[6], Shd;
    mom_load<T>(l1, i, L1, S1, SS1, Sh);
    mom_load<T>(l2, i, L2, S2, SS2, Shd);
    const double L2v = L * L, T2 = TFR * TFR, T4 = T2 * T2, T8 = T4 * T4;
    const double k2 = 2.0 * L2v;
    acc[0] += k2 * con6<T>(L1, SS1);
    acc[1] += k2 * con6<T>(L2, SS2);
    acc[2] += k2 * T2 * Sh * con6<T>(L1, S1);
    acc[3] += k2 * T4 * Shd * con6<T>(L2, S2);
    acc[4] += (k2 * k2) * con6<T>(SS1, SS1);
    acc[5] += (k2 * k2) * con6<T>(SS2, SS2);
    acc[6] += (4.0 * L2v * L2v) * T4 * (Sh * Sh) * con6<T>(S1, S1);
    acc[7] += (4.0 * L2v * L2v) * T8 * (Shd * Shd) * con6<T>(S2, S2);
    acc[8] += (8.0 * L2v * L2v) * T2 * Sh * con6<T>(S1, SS1);
    acc[9] += (8.0 * L2v * L2v) * T4 * Shd * con6<T>(S2, SS2);
  }
};

// M_ij, LM, MM (DynamicSGS_LASDD_SM.py:327-355) and |S_hat| (:235) at point
// i; b = beta of the level.
template <class T>
__device__ __forceinline__ void lasdd_lm_point(const MomLevel<T>& l1, size_t i, double b,
                                               double L, double TFR, double& lm, double& mm,
                                               double& Sm) {
  double Lij[6], Sij[6], SSij[6];
  mom_load<T>(l1, i, Lij, Sij, SSij, Sm);
  const double T1 = 2.0 * L * L, T2 = 2.0 * (TFR * L) * (TFR * L);
  double M[6];
#pragma unroll
  for (int n = 0; n < 6; ++n) M[n] = T1 * SSij[n] - T2 * b * Sm * Sij[n];
  lm = (Lij[0] * M[0] + Lij[1] * M[1] + Lij[2] * M[2]) +
       2.0 * (Lij[3] * M[3] + Lij[4] * M[4] + Lij[5] * M[5]);
  mm = M[0] * M[0] + M[1] * M[1] + M[2] * M[2] + 2.0 * (M[3] * M[3] + M[4] * M[4] + M[5] * M[5]);
}

template <class T>
__device__ __forceinline__ void k_lasdd_lm(const PwDims& d, const MomLevel<T>& l1, const double* __restrict__ beta, double L,
                           double TFR, T* LM, T* MM, T* Shat) {
  const int nz = d.nz;
  PW_LOOP(d.n) {
    const int z = (int)(i % nz);
    double lm, mm, Sm;
    lasdd_lm_point<T>(l1, i, beta[z], L, TFR, lm, mm, Sm);
    LM[i] = (T)lm;
    MM[i] = (T)mm;
    if (Shat) Shat[i] = (T)Sm;
  }
}

// Vector form (nz % 2 == 0, 16-byte aligned arrays): two consecutive levels
// per thread, the 21 fields read as 2-element vectors into registers and the
// point function run on them (same arithmetic).  Two, not 16 bytes' worth:
// four floats of 21 fields would not fit the registers.
template <class T>
__device__ __forceinline__ void k_lasdd_lm_v(const PwDims& d, const MomLevel<T>& l1, const double* __restrict__ beta,
                             double L, double TFR, T* LM, T* MM, T* Shat) {
  constexpr int W = 2;
  const int nz = d.nz;
  const size_t nv = d.n / W;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < nv;
       t += (size_t)gridDim.x * blockDim.x) {
    const size_t i0 = t * W;
    const int z0 = (int)(i0 % nz);
    c2<T> f[21];
#pragma unroll
    for (int k = 0; k < 3; ++k) f[k] = *reinterpret_cast<const c2<T>*>(l1.uh[k] + i0);
#pragma unroll
    for (int n = 0; n < 6; ++n) {
      f[3 + n] = *reinterpret_cast<const c2<T>*>(l1.ph[n] + i0);
      f[9 + n] = *reinterpret_cast<const c2<T>*>(l1.Sh[n] + i0);
      f[15 + n] = *reinterpret_cast<const c2<T>*>(l1.SSh[n] + i0);
    }
    T v[21][W];
#pragma unroll
    for (int k = 0; k < 21; ++k) { v[k][0] = f[k].x; v[k][1] = f[k].y; }
    MomLevel<T> r;   // the level, on the registers
#pragma unroll
    for (int k = 0; k < 3; ++k) r.uh[k] = v[k];
#pragma unroll
    for (int n = 0; n < 6; ++n) {
      r.ph[n] = v[3 + n]; r.Sh[n] = v[9 + n]; r.SSh[n] = v[15 + n];
    }
    T oLM[W], oMM[W], oS[W];
#pragma unroll
    for (int j = 0; j < W; ++j) {
      double lm, mm, Sm;
      lasdd_lm_point<T>(r, (size_t)j, beta[z0 + j], L, TFR, lm, mm, Sm);
      oLM[j] = (T)lm; oMM[j] = (T)mm; oS[j] = (T)Sm;
    }
    *reinterpret_cast<c2<T>*>(LM + i0) = mkc<T>(oLM[0], oLM[1]);
    *reinterpret_cast<c2<T>*>(MM + i0) = mkc<T>(oMM[0], oMM[1]);
    if (Shat) *reinterpret_cast<c2<T>*>(Shat + i0) = mkc<T>(oS[0], oS[1]);
  }
}

// |S| of the level-2 filtered strains (DynamicSGS_LASDD_SM.py:240), kept for
// the scalar model.
template <class T>
__device__ __forceinline__ void k_smag_level(const PwDims& d, const MomLevel<T>& l, T* out) {
  PW_LOOP(d.n) {
    double s[6];
#pragma unroll
    for (int n = 0; n < 6; ++n) s[n] = l.Sh[n][i];
    out[i] = (T)sqrt(2.0 * (s[0] * s[0] + s[1] * s[1] + s[2] * s[2] + 2.0 * s[3] * s[3] +
                            2.0 * s[4] * s[4] + 2.0 * s[5] * s[5]));
  }
}

// Filtered scalar fields of one level: TH_hat, (u,v,w)TH_hat, grad hats (3),
// |S| grad hats (3); plus the velocity hats and |S_hat| of the same level.
template <class T> struct ScaLevel {
  const T* th;
  const T* uth[3];
  const T* dh[3];
  const T* sdh[3];
  const T* uh[3];
  const T* Sm;
};

template <class T>
__device__ __forceinline__ void sca_terms(const ScaLevel<T>& s, size_t i, double* t5) {
  const double th = s.th[i], Sm = s.Sm[i];
  double a = 0, b = 0, c = 0, d = 0, e = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double LT = (double)s.uth[k][i] - (double)s.uh[k][i] * th;
    const double sd = s.sdh[k][i], dh = s.dh[k][i];
    a += LT * sd;
    b += LT * Sm * dh;
    c += sd * sd;
    d += sd * Sm * dh;
    e += (Sm * dh) * (Sm * dh);
  }
  t5[0] = a; t5[1] = b; t5[2] = c; t5[3] = d; t5[4] = e;
}

// a2..e2, a4..e4 of DynamicSGS_ScalarLASDD_SM.py:241-302 (sums).
template <class T> struct FScalarSums {
  static constexpr const char* NAME = "FScalarSums";
  static constexpr int NFIELDS = 28;
  ScaLevel<T> l1, l2;
  int nz;
  double L, TFR;
  __device__ void operator()(int c, int z, double* acc) const {
    const size_t i = (size_t)c * nz + z;
    double p[5], q[5];
    sca_terms<T>(l1, i, p);
    sca_terms<T>(l2, i, q);
    const double L2 = L * L, L4 = L2 * L2, T2 = TFR * TFR, T4 = T2 * T2, T8 = T4 * T4;
    acc[0] += L2 * p[0];
    acc[1] += (-L2 * T2) * p[1];
    acc[2] += L4 * p[2];
    acc[3] += (-2.0 * L4 * T2) * p[3];
    acc[4] += (L4 * T4) * p[4];
    acc[5] += L2 * q[0];
    acc[6] += (-L2 * T4) * q[1];
    acc[7] += L4 * q[2];
    acc[8] += (-2.0 * L4 * T4) * q[3];
    acc[9] += (L4 * T8) * q[4];
  }
};

// T_up, T_dn (DynamicSGS_ScalarLASDD_SM.py:304-312) at point i.
template <class T>
__device__ __forceinline__ void scalar_updn_point(const ScaLevel<T>& l1, size_t i, double b,
                                                  double L, double TFR, T& up, T& dn) {
  double t[5];
  sca_terms<T>(l1, i, t);
  const double L2 = L * L, L4 = L2 * L2, T2 = TFR * TFR, T4 = T2 * T2;
  up = (T)(L2 * t[0] + (-L2 * T2) * t[1] * b);
  dn = (T)(L4 * t[2] + (-2.0 * L4 * T2) * t[3] * b + (L4 * T4) * t[4] * b * b);
}

template <class T>
__device__ __forceinline__ void k_scalar_updn(const PwDims& d, const ScaLevel<T>& l1, const double* __restrict__ beta,
                              double L, double TFR, T* Tup, T* Tdn) {
  const int nz = d.nz;
  PW_LOOP(d.n) {
    const int z = (int)(i % nz);
    scalar_updn_point<T>(l1, i, beta[z], L, TFR, Tup[i], Tdn[i]);
  }
}

// Vector form (nz % 2 == 0, 16-byte aligned arrays): two consecutive levels
// per thread, the 14 fields as 2-element vectors in registers and the point
// function run on them (same arithmetic).
template <class T>
__device__ __forceinline__ void k_scalar_updn_v(const PwDims& d, const ScaLevel<T>& l1, const double* __restrict__ beta,
                                double L, double TFR, T* Tup, T* Tdn) {
  constexpr int W = 2;
  const int nz = d.nz;
  const size_t nv = d.n / W;
  auto ld2 = [](const T* p) { return *reinterpret_cast<const c2<T>*>(p); };
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < nv;
       t += (size_t)gridDim.x * blockDim.x) {
    const size_t i0 = t * W;
    const int z0 = (int)(i0 % nz);
    c2<T> f[14];
    f[0] = ld2(l1.th + i0);
    f[1] = ld2(l1.Sm + i0);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      f[2 + k] = ld2(l1.uth[k] + i0);
      f[5 + k] = ld2(l1.dh[k] + i0);
      f[8 + k] = ld2(l1.sdh[k] + i0);
      f[11 + k] = ld2(l1.uh[k] + i0);
    }
    T v[14][W];
#pragma unroll
    for (int k = 0; k < 14; ++k) { v[k][0] = f[k].x; v[k][1] = f[k].y; }
    ScaLevel<T> r;   // the level, on the registers
    r.th = v[0];
    r.Sm = v[1];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      r.uth[k] = v[2 + k]; r.dh[k] = v[5 + k]; r.sdh[k] = v[8 + k]; r.uh[k] = v[11 + k];
    }
    T up[W], dn[W];
#pragma unroll
    for (int j = 0; j < W; ++j)
      scalar_updn_point<T>(r, (size_t)j, beta[z0 + j], L, TFR, up[j], dn[j]);
    *reinterpret_cast<c2<T>*>(Tup + i0) = mkc<T>(up[0], up[1]);
    *reinterpret_cast<c2<T>*>(Tdn + i0) = mkc<T>(dn[0], dn[1]);
  }
}

// ---------------------------------------------------------------------------
// Wong-Lilly base model (DynamicSGS_LASDD_WL.py:128-313,
// DynamicSGS_ScalarLASDD_WL.py:130-275).  Levels hold the filtered
// velocities, the six products and S_ij; the level-2 set may hold only S_ij
// (LAD-WL needs nothing else there).
// ---------------------------------------------------------------------------
template <class T> struct WlMom {
  const T* uh[3];
  const T* ph[6];
  const T* Sh[6];
};

template <class T>
__device__ __forceinline__ void wl_leonard(const WlMom<T>& m, size_t i, double* Lij) {
  const double u0 = m.uh[0][i], u1 = m.uh[1][i], u2 = m.uh[2][i];
  Lij[0] = (double)m.ph[0][i] - u0 * u0;
  Lij[1] = (double)m.ph[1][i] - u1 * u1;
  Lij[2] = (double)m.ph[2][i] - u2 * u2;
  Lij[3] = (double)m.ph[3][i] - u0 * u1;
  Lij[4] = (double)m.ph[4][i] - u0 * u2;
  Lij[5] = (double)m.ph[5][i] - u1 * u2;
}

// a1 = <Q:S^>, a3 = <S^:S^>, a6 = <L:S^>, a8 = <S^^:S^^> (LASDD_WL.py:249-261)
template <class T> struct FWlSums {
  static constexpr const char* NAME = "FWlSums";
  static constexpr int NFIELDS = 30;
  WlMom<T> l1, l2;
  int nz;
  __device__ void operator()(int c, int z, double* acc) const {
    const size_t i = (size_t)c * nz + z;
    double L1[6], Q[6], S1[6], S2[6];
    wl_leonard<T>(l1, i, L1);
    wl_leonard<T>(l2, i, Q);
#pragma unroll
    for (int n = 0; n < 6; ++n) { S1[n] = l1.Sh[n][i]; S2[n] = l2.Sh[n][i]; }
    acc[0] += con6<T>(Q, S1);
    acc[1] += con6<T>(S1, S1);
    acc[2] += con6<T>(L1, S1);
    acc[3] += con6<T>(S2, S2);
  }
};

// M_ij = 2 L^(4/3) (S^_ij - TFR^(4/3) beta S^^_ij), LM, MM (LASDD_WL.py:284-301)
template <class T>
__device__ __forceinline__ void k_wl_lm(const PwDims& d, const WlMom<T>& l1, const WlMom<T>& l2, const double* __restrict__ beta,
                        double L, double TFR, T* LM, T* MM) {
  const int nz = d.nz;
  const double T1 = 2.0 * pow(L, 4.0 / 3.0), T2 = 2.0 * pow(TFR * L, 4.0 / 3.0);
  PW_LOOP(d.n) {
    const int z = (int)(i % nz);
    double Lij[6], M[6];
    wl_leonard<T>(l1, i, Lij);
    const double b = beta[z];
#pragma unroll
    for (int n = 0; n < 6; ++n) M[n] = T1 * (double)l1.Sh[n][i] - T2 * b * (double)l2.Sh[n][i];
    LM[i] = (T)con6<T>(Lij, M);
    MM[i] = (T)con6<T>(M, M);
  }
}

// Scalar level: TH^, (u,v,w)TH^, the gradient hats, and the velocity hats.
template <class T> struct WlSca {
  const T* th;
  const T* uth[3];
  const T* dh[3];
  const T* uh[3];
};

template <class T>
__device__ __forceinline__ void wl_sca_leonard(const WlSca<T>& s, size_t i, double* LT) {
  const double th = s.th[i];
#pragma unroll
  for (int k = 0; k < 3; ++k) LT[k] = (double)s.uth[k][i] - (double)s.uh[k][i] * th;
}

// a1 = <Q.d^>, a3 = <d^.d^>, a6 = <L.d^>, a8 = <d^^.d^^> (ScalarLASDD_WL.py:218-228)
template <class T> struct FWlScalarSums {
  static constexpr const char* NAME = "FWlScalarSums";
  static constexpr int NFIELDS = 20;
  WlSca<T> l1, l2;
  int nz;
  __device__ void operator()(int c, int z, double* acc) const {
    const size_t i = (size_t)c * nz + z;
    double L1[3], Q[3];
    wl_sca_leonard<T>(l1, i, L1);
    wl_sca_leonard<T>(l2, i, Q);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double d1 = l1.dh[k][i], d2 = l2.dh[k][i];
      acc[0] += Q[k] * d1;
      acc[1] += d1 * d1;
      acc[2] += L1[k] * d1;
      acc[3] += d2 * d2;
    }
  }
};

// T_up, T_dn of ScalarLASDD_WL.py:252-268; l2 needs only the gradient hats
template <class T>
__device__ __forceinline__ void k_wl_scalar_updn(const PwDims& d, const WlSca<T>& l1, const WlSca<T>& l2,
                                 const double* __restrict__ beta, double L, double TFR, T* Tup,
                                 T* Tdn) {
  const int nz = d.nz;
  const double L43 = pow(L, 4.0 / 3.0), L83 = pow(L, 8.0 / 3.0);
  const double T43 = pow(TFR, 4.0 / 3.0), T83 = pow(TFR, 8.0 / 3.0);
  PW_LOOP(d.n) {
    const int z = (int)(i % nz);
    double LT[3];
    wl_sca_leonard<T>(l1, i, LT);
    double a6 = 0, b6 = 0, a3 = 0, c36 = 0, a8 = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double d1 = l1.dh[k][i], d2 = l2.dh[k][i];
      a6 += LT[k] * d1;
      b6 += LT[k] * d2;
      a3 += d1 * d1;
      c36 += d1 * d2;
      a8 += d2 * d2;
    }
    const double b = beta[z];
    Tup[i] = (T)(L43 * (a6 - T43 * b * b6));
    Tdn[i] = (T)(L83 * (a3 - 2.0 * T43 * b * c36 + T83 * b * b * a8));
  }
}

// Imfilter (Utilities.py:87-115) of both inputs at each level, the clipped
// ratio (DynamicSGS_LASDD_SM.py:103-130, ScalarLASDD_SM.py:104-131, and the
// WL twins), and the plane sums of sqrt(C) and C for the profiles
// (LASDD_SM.py:359-360).  Each thread walks a contiguous run of columns
// along y and keeps the 3-row column sums of y-1, y and y+1, so every step
// loads one new column of the 3x3 box (3 values per field) instead of nine.
// Partials in the layout of k_plane_partial (NT = 2: sqrt(C), C).
template <class T>
__device__ __forceinline__ void
k_box_clip(const RedGeom& g, const T* __restrict__ num, const T* __restrict__ den, T* __restrict__ out,
           int nx, int ny, double* part) {
  __shared__ double sm[2][RED_BS];
  int bz, bc;
  red_block(bz, bc);
  const int zl = threadIdx.x % g.zb, cl = threadIdx.x / g.zb;
  const int z = bz * g.zb + zl;
  const int nz = g.nz;
  const int chunk = g.cpb / g.lanes;
  double s1 = 0.0, s2 = 0.0;
  if (z < nz) {
    const int c0 = bc * g.cpb + cl * chunk;
    const int c1 = imin(c0 + chunk, g.ncol);
    const T inv9 = T(1.0 / 9.0);
    T a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0;   // column sums at y-1, y, y+1
    int xm = 0, x = -1, xp = 0;
    auto col = [&](int yy, T& sa, T& sb) {
      const size_t jm = ((size_t)xm * ny + yy) * nz + z;
      const size_t j0 = ((size_t)x * ny + yy) * nz + z;
      const size_t jp = ((size_t)xp * ny + yy) * nz + z;
      sa = num[jm] + num[j0] + num[jp];
      sb = den[jm] + den[j0] + den[jp];
    };
#pragma unroll 2
    for (int c = c0; c < c1; ++c) {
      const int cx = c / ny, y = c - cx * ny;
      const int yp = (y + 1 == ny) ? 0 : y + 1;
      if (cx != x || c == c0) {
        x = cx;
        xm = (x + nx - 1) % nx;
        xp = (x + 1) % nx;
        col((y + ny - 1) % ny, a0, b0);
        col(y, a1, b1);
      } else {
        a0 = a1; b0 = b1;
        a1 = a2; b1 = b2;
      }
      col(yp, a2, b2);
      const T a = (a0 + a1 + a2) * inv9, b = (b0 + b1 + b2) * inv9;
      T r = a / b;
      if ((fabs((double)b) < 1e-10) || (r < T(0)) || (r > T(1))) r = T(0);
      out[(size_t)c * nz + z] = r;
      // the root in the precision of the field (the sums are double)
      s1 += (double)sqrt(r);
      s2 += (double)r;
    }
  }
  sm[0][threadIdx.x] = s1;
  sm[1][threadIdx.x] = s2;
  __syncthreads();
  if (cl == 0 && z < nz) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      double s = 0.0;
      for (int l = 0; l < g.lanes; ++l) s += sm[t][l * g.zb + zl];
      part[((size_t)bc * 2 + t) * nz + z] = s;
    }
  }
}

// k_box_clip with 16-byte accesses (nz % W == 0, 16-byte aligned arrays):
// launched with RED_BS / W threads on the geometry of red_geom, each thread
// takes W consecutive levels of the columns that W threads of k_box_clip
// take, in the same order, so the partial sums are those of k_box_clip.
template <class T>
__device__ __forceinline__ void
k_box_clip_v(const RedGeom& g, const T* __restrict__ num, const T* __restrict__ den,
             T* __restrict__ out, int nx, int ny, double* part) {
  constexpr int W = Vec16<T>::W;
  __shared__ double sm[2][RED_BS];
  int bz, bc;
  red_block(bz, bc);
  const int tz = g.zb / W;
  const int zl = threadIdx.x % tz, cl = threadIdx.x / tz;
  const int z = bz * g.zb + zl * W;
  const int nz = g.nz;
  const int chunk = g.cpb / g.lanes;
  double s1[W], s2[W];
#pragma unroll
  for (int j = 0; j < W; ++j) { s1[j] = 0.0; s2[j] = 0.0; }
  if (z < nz) {
    const int c0 = bc * g.cpb + cl * chunk;
    const int c1 = imin(c0 + chunk, g.ncol);
    const T inv9 = T(1.0 / 9.0);
    Vec16<T> a0, a1, a2, b0, b1, b2;   // column sums at y-1, y, y+1
    // the x rows x-1, x, x+1 of the two fields at level z; the chunk walks
    // y with one division at its start and rebases the rows when it crosses
    // into the next x
    const T *nm = nullptr, *n0 = nullptr, *np = nullptr, *dm = nullptr, *d0 = nullptr, *dp = nullptr;
    auto col = [&](int yy, Vec16<T>& sa, Vec16<T>& sb) {
      const size_t o = (size_t)yy * nz;
      const Vec16<T> vm = vld(nm + o), v0 = vld(n0 + o), vp = vld(np + o);
      const Vec16<T> wm = vld(dm + o), w0 = vld(d0 + o), wp = vld(dp + o);
#pragma unroll
      for (int j = 0; j < W; ++j) {
        sa.v[j] = vm.v[j] + v0.v[j] + vp.v[j];
        sb.v[j] = wm.v[j] + w0.v[j] + wp.v[j];
      }
    };
    int x = c0 / ny, y = c0 - x * ny;
    bool fresh = true;
    for (int c = c0; c < c1; ++c) {
      const int yp = (y + 1 == ny) ? 0 : y + 1;
      if (fresh) {
        fresh = false;
        const int xm = (x == 0) ? nx - 1 : x - 1, xp = (x + 1 == nx) ? 0 : x + 1;
        const size_t pl = (size_t)ny * nz;
        nm = num + xm * pl + z; n0 = num + x * pl + z; np = num + xp * pl + z;
        dm = den + xm * pl + z; d0 = den + x * pl + z; dp = den + xp * pl + z;
        col((y == 0) ? ny - 1 : y - 1, a0, b0);
        col(y, a1, b1);
      } else {
        a0 = a1; b0 = b1;
        a1 = a2; b1 = b2;
      }
      col(yp, a2, b2);
      Vec16<T> o;
#pragma unroll
      for (int j = 0; j < W; ++j) {
        const T a = (a0.v[j] + a1.v[j] + a2.v[j]) * inv9, b = (b0.v[j] + b1.v[j] + b2.v[j]) * inv9;
        T r = a / b;
        if ((fabs((double)b) < 1e-10) || (r < T(0)) || (r > T(1))) r = T(0);
        o.v[j] = r;
        // the root in the precision of the field (the sums are double)
        s1[j] += (double)sqrt(r);
        s2[j] += (double)r;
      }
      vst(out + (size_t)c * nz + z, o);
      if (++y == ny) { y = 0; ++x; fresh = true; }
    }
  }
#pragma unroll
  for (int j = 0; j < W; ++j) {
    sm[0][threadIdx.x * W + j] = s1[j];
    sm[1][threadIdx.x * W + j] = s2[j];
  }
  __syncthreads();
  if (cl == 0 && z < nz) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
#pragma unroll
      for (int j = 0; j < W; ++j) {
        double s = 0.0;
        for (int l = 0; l < g.lanes; ++l) s += sm[t][(l * tz + zl) * W + j];
        part[((size_t)bc * 2 + t) * nz + z + j] = s;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// beta: ComputeBeta1/2 (LASDD_SM.py:53-95, ScalarLASDD_SM.py:54-96) with the
// Laguerre iteration of Roots (Utilities.py:123-241), in double complex.
// coef layout (6, n): ff, ee, dd, cc, bb, aa.  One thread per (level, start).
// ---------------------------------------------------------------------------
struct dcx { double x, y; };
__device__ __forceinline__ dcx dadd(dcx a, dcx b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ dcx dsub(dcx a, dcx b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ dcx dmul(dcx a, dcx b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ double dabs(dcx a) { return hypot(a.x, a.y); }
__device__ __forceinline__ dcx ddiv(dcx a, dcx b) {
  // Smith's algorithm
  if (fabs(b.x) >= fabs(b.y)) {
    const double r = b.y / b.x, den = b.x + b.y * r;
    return {(a.x + a.y * r) / den, (a.y - a.x * r) / den};
  } else {
    const double r = b.x / b.y, den = b.x * r + b.y;
    return {(a.x * r + a.y) / den, (a.y * r - a.x) / den};
  }
}
__device__ __forceinline__ dcx dsqrt(dcx z) {
  // principal branch; the sign of a zero imaginary part selects the side
  if (z.x == 0.0 && z.y == 0.0) return {0.0, z.y};
  const double r = hypot(z.x, z.y);
  const double s = sqrt(0.5 * (r + fabs(z.x)));
  if (z.x >= 0.0) return {s, z.y / (2.0 * s)};
  return {fabs(z.y) / (2.0 * s), copysign(s, z.y)};
}

__device__ dcx laguerre(const double* p, double guess, bool* ok) {
  const int n = 5;
  const double eps = 1e-16, tol = 1e-6;
  dcx x = {guess, 0.0};
  bool conv = false;
  for (int it = 0; it < 20 && !conv; ++it) {
    dcx f = {0, 0}, df = {0, 0}, d2f = {0, 0};
    for (int k = 0; k <= n; ++k) f = dadd(dmul(f, x), dcx{p[k], 0.0});
    for (int k = 0; k < n; ++k) df = dadd(dmul(df, x), dcx{p[k] * (n - k), 0.0});
    for (int k = 0; k < n - 1; ++k) d2f = dadd(dmul(d2f, x), dcx{p[k] * (n - k) * (n - 1 - k), 0.0});
    const dcx fs = dabs(f) < eps ? dcx{eps, 0.0} : f;
    const dcx G = ddiv(df, fs);
    const dcx H = dsub(dmul(G, G), ddiv(d2f, fs));
    const dcx disc = dmul(dcx{(double)(n - 1), 0.0}, dsub(dmul(dcx{(double)n, 0.0}, H), dmul(G, G)));
    const dcx sq = dsqrt(disc);
    const dcx d1 = dadd(G, sq), d2 = dsub(G, sq);
    dcx den = dabs(d1) > dabs(d2) ? d1 : d2;
    if (dabs(den) < eps) den = dcx{eps, 0.0};
    const dcx xn = dsub(x, ddiv(dcx{(double)n, 0.0}, den));
    conv = (dabs(dsub(xn, x)) < tol * (1.0 + dabs(x))) || (dabs(f) < tol);
    x = xn;
  }
  *ok = conv;
  return x;
}

__device__ __forceinline__ void k_beta(const double* __restrict__ coef, int n, double* beta) {
  const double guesses[8] = {0.5, 0.75, 1.0, 1.25, 1.5, 2.0, 3.0, 4.5};
  __shared__ double best[64];
  const int lv = threadIdx.x / 8, g = threadIdx.x % 8;
  const int k = blockIdx.x * 8 + lv;
  double val = -1.0;   // "no valid root"
  if (k < n) {
    double p[6];
    for (int j = 0; j < 6; ++j) p[j] = coef[(size_t)j * n + k];
    bool ok;
    const dcx r = laguerre(p, guesses[g], &ok);
    if (ok && fabs(r.y) < 1e-6 && r.x > 0.0 && r.x < 5.0) val = r.x;
  }
  best[threadIdx.x] = val;
  __syncthreads();
  if (g == 0 && k < n) {
    double m = -1.0;
    for (int j = 0; j < 8; ++j) m = fmax(m, best[lv * 8 + j]);
    beta[k] = (m > 0.0) ? m : 1.0;
  }
}

// Quintic coefficients (LASDD_SM.py:314-319 momentum, ScalarLASDD_SM.py:
// 296-302 scalar) from the plane means m[10][nz]; coef (6, nz).
// mode 0: SM momentum, 1: SM scalar, 2: Wong-Lilly (momentum or scalar,
// LASDD_WL.py:255-274 from m[4][nz] = a1, a3, a6, a8).
__device__ __forceinline__ void k_quintic(const double* __restrict__ m, int nz, int scalar, double TFR,
                          double* coef) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nz) return;
  double ff, ee, dd, cc, bb, aa;
  if (scalar == 2) {
    const double a1 = m[0 * nz + k], a3 = m[1 * nz + k], a6 = m[2 * nz + k], a8 = m[3 * nz + k];
    const double t43 = pow(TFR, 4.0 / 3.0), t83 = pow(TFR, 8.0 / 3.0), t163 = pow(TFR, 16.0 / 3.0);
    const double a2 = -t83 * a1, a4 = -2.0 * t43 * a3, a5 = t83 * a3, a7 = -t43 * a6;
    const double a9 = -2.0 * t83 * a8, a10 = t163 * a8;
    aa = a1 * a3 - a6 * a8;
    bb = a1 * a4 - a7 * a8;
    cc = a2 * a3 + a1 * a5 - a6 * a9;
    dd = a2 * a4 - a7 * a9;
    ee = a2 * a5 - a6 * a10;
    ff = -a7 * a10;
  } else if (!scalar) {
    const double a1 = m[0 * nz + k], a2 = m[1 * nz + k], b1 = m[2 * nz + k], b2 = m[3 * nz + k];
    const double c1 = m[4 * nz + k], c2 = m[5 * nz + k], d1 = m[6 * nz + k], d2 = m[7 * nz + k];
    const double e1 = m[8 * nz + k], e2 = m[9 * nz + k];
    aa = a1 * c2 - a2 * c1;
    bb = a2 * e1 - b1 * c2;
    cc = b2 * c1 - a1 * e2 - a2 * d1;
    dd = b1 * e2 - b2 * e1;
    ee = a1 * d2 + b2 * d1;
    ff = -b1 * d2;
  } else {
    const double a2 = m[0 * nz + k], b2 = m[1 * nz + k], c2 = m[2 * nz + k], d2 = m[3 * nz + k];
    const double e2 = m[4 * nz + k], a4 = m[5 * nz + k], b4 = m[6 * nz + k], c4 = m[7 * nz + k];
    const double d4 = m[8 * nz + k], e4 = m[9 * nz + k];
    aa = a2 * c4 - a4 * c2;
    bb = -a4 * d2 + b2 * c4;
    cc = -c2 * b4 + a2 * d4 - a4 * e2;
    dd = b2 * d4 - b4 * d2;
    ee = a2 * e4 - b4 * e2;
    ff = b2 * e4;
  }
  coef[0 * nz + k] = ff; coef[1 * nz + k] = ee; coef[2 * nz + k] = dd;
  coef[3 * nz + k] = cc; coef[4 * nz + k] = bb; coef[5 * nz + k] = aa;
}

// Profiles: avg1 = mean(sqrt C)^2, avg2 = mean(C) from the box-clip sums;
// with beta_out, the procedure's beta profile is copied out as well.
__device__ __forceinline__ void k_profiles(const double* __restrict__ s, int nz, double inv_n, double* avg1,
                           double* avg2, const double* __restrict__ beta, double* beta_out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nz) return;
  const double m1 = s[k] * inv_n;
  if (avg1) avg1[k] = m1 * m1;
  if (avg2) avg2[k] = s[nz + k] * inv_n;
  if (beta_out) beta_out[k] = beta[k];
}

// device copy of n doubles (a kernel rather than a memcpy node, so that
// captured steps hold kernels only)
__device__ __forceinline__ void k_copy_d(int n, const double* __restrict__ src, double* __restrict__ dst) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = src[k];
}

__device__ __forceinline__ void k_fill_d(int n, double* a, double v) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) a[k] = v;
}
