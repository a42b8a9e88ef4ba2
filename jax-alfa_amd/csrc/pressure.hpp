// Pressure Poisson solve (src/pde/NSE_PressureTerms.py:109-182 PressureRC,
// src/pde/NSE_PressureTerms_Thomas.py:99-292 ThomasPressureSolve).
//
// One 64-lane wavefront owns one (kx, ky) column.  It forms the right-hand
// side RC of the column on the fly from the transformed Rx, Ry, Rz and the
// transformed divtz boundary planes, and solves the pre-factored tridiagonal
// system (nz+1 unknowns) with both Thomas sweeps written as first-order
// linear recurrences y_k = a_k y_{k-1} + b_k: each lane composes the affine
// maps of a contiguous chunk of levels, a 6-step shuffle scan combines the
// lanes, and each lane then replays its chunk.  All solve arithmetic is
// double.  The column is contiguous in memory (z fastest), so the loads and
// stores of a wavefront are fully coalesced.  Outputs are p_hat and the
// spectrum of dp/dz (Derivz_Generic_uvp of p is linear in p_hat).
#pragma once
#include "common.hpp"

template <class T> struct ThomasArgs {
  const c2<T>* fRx; const c2<T>* fRy; const c2<T>* fRz;   // (nx, nyc, nz)
  const c2<T>* fbt;                                        // (nx, nyc, 2)
  // (nx/2+1, nyc, nz+1) pivots of ThomasPressureInit's elimination; they
  // depend on kx only through kx^2, so row (|kx|, ky) serves kx and -kx
  const double* b;
  c2<T>* phat; c2<T>* dpz;                                 // (nx, nyc, nz)
  T* RCr; T* RCi;             // optional out (ncol, nz+1): RC before masking
  const T* RCin_r; const T* RCin_i;   // optional in (ncol, nz+1): RC given
  const T* fRzr_in;                   // with RCin: Re fRz (ncol, nz) for the zero mode
  int nx, ny, nyc, nz;
  double dz, idz, c_in;       // c_in = 1/dz^2
};

constexpr int TH_MAXN = 520;   // nz + 1 <= 520
// levels of a lane whose right-hand-side loads are issued together
#ifndef ALFA_TH_TU
#define ALFA_TH_TU 3
#endif
// dynamic LDS of one column: 3 arrays of nz + 1 doubles
inline size_t thomas_lds(int nz) { return 3 * (size_t)(nz + 1) * sizeof(double); }

__device__ __forceinline__ double shfl_up_d(double v, int off) { return __shfl_up(v, off, 64); }
__device__ __forceinline__ double shfl_down_d(double v, int off) { return __shfl_down(v, off, 64); }

template <class T>
__device__ __forceinline__ void k_thomas(const ThomasArgs<T>& a) {
  extern __shared__ double th_smem[];
  const int col = blockIdx.x;
  const int lane = threadIdx.x;
  const int nz = a.nz, N = nz + 1;
  double* dr = th_smem;
  double* di = dr + N;
  double* sb = di + N;
  const int i = col / a.nyc, j = col % a.nyc;
  const double kr = (i < a.nx / 2) ? (double)i : (double)(i - a.nx);
  const double kc = (double)j;
  const bool special = (i == 0 && j == 0) || (2 * i == a.nx) || (2 * j == a.ny);
  const size_t cb = (size_t)col * nz;
  const int ia = (i <= a.nx / 2) ? i : a.nx - i;
  const double* brow = a.b + ((size_t)ia * a.nyc + j) * N;
  // right-hand side of the column (NSE_PressureTerms.py:150-180)
  if (a.RCin_r) {
    for (int k = lane; k < N; k += 64) {
      const double rr = (double)a.RCin_r[(size_t)col * N + k];
      const double ri = (double)a.RCin_i[(size_t)col * N + k];
      if (a.RCr) { a.RCr[(size_t)col * N + k] = (T)rr; a.RCi[(size_t)col * N + k] = (T)ri; }
      dr[k] = special ? 0.0 : rr;
      di[k] = special ? 0.0 : ri;
      sb[k] = brow[k];
    }
  } else {
    // the levels of a lane TU at a time, every load of the batch issued
    // first (indices clamped into the column, so the loads need no branch):
    // one memory round trip per batch instead of one per level
    constexpr int TU = ALFA_TH_TU;
    const c2<T> fb0 = a.fbt[(size_t)col * 2], fb1 = a.fbt[(size_t)col * 2 + 1];
    for (int kb = lane; kb < N; kb += 64 * TU) {
      c2<T> X[TU], Y[TU], Z1[TU], Z0[TU];
      double bv[TU];
#pragma unroll
      for (int t = 0; t < TU; ++t) {
        const int k = kb + 64 * t;
        const int kq = k < 1 ? 1 : (k > nz - 1 ? nz - 1 : k);   // an interior level
        bv[t] = brow[k < N ? k : N - 1];
        X[t] = a.fRx[cb + kq - 1]; Y[t] = a.fRy[cb + kq - 1];
        Z1[t] = a.fRz[cb + kq]; Z0[t] = a.fRz[cb + kq - 1];
      }
#pragma unroll
      for (int t = 0; t < TU; ++t) {
        const int k = kb + 64 * t;
        if (k < N) {
          double rr, ri;
          if (k == 0 || k == nz) {
            const c2<T> f = k == 0 ? fb0 : fb1;
            rr = -a.dz * (double)f.x;
            ri = -a.dz * (double)f.y;
          } else {
            double hr = -kr * (double)X[t].y;
            hr -= kc * (double)Y[t].y;
            double hi = kr * (double)X[t].x;
            hi += kc * (double)Y[t].x;
            rr = hr + ((double)Z1[t].x - (double)Z0[t].x) / a.dz;
            ri = hi + ((double)Z1[t].y - (double)Z0[t].y) / a.dz;
          }
          if (a.RCr) { a.RCr[(size_t)col * N + k] = (T)rr; a.RCi[(size_t)col * N + k] = (T)ri; }
          dr[k] = special ? 0.0 : rr;
          di[k] = special ? 0.0 : ri;
          sb[k] = bv[t];
        }
      }
    }
  }
  __syncthreads();
  if (special) {
    if (col == 0) {
      // zero mode by vertical integration (NSE_PressureTerms_Thomas.py:271-275)
      for (int k = lane; k < nz; k += 64)
        dr[k] = a.RCin_r ? (double)a.fRzr_in[cb + k] : (double)a.fRz[cb + k].x;
      __syncthreads();
      if (lane == 0) {
        const double zf = a.RCin_r ? (double)a.RCin_r[0] : -a.dz * (double)a.fbt[0].x;
        double S = 0.0, prev = zf;
        a.phat[cb] = mkc<T>((T)zf, T(0));
        a.dpz[cb] = mkc<T>(T(0), T(0));
        for (int k = 1; k < nz; ++k) {
          S += dr[k] * a.dz;
          const double p = zf + S;
          a.phat[cb + k] = mkc<T>((T)p, T(0));
          a.dpz[cb + k] = mkc<T>((T)((p - prev) * a.idz), T(0));
          prev = p;
        }
      }
    } else {
      for (int k = lane; k < nz; k += 64) {
        a.phat[cb + k] = mkc<T>(T(0), T(0));
        a.dpz[cb + k] = mkc<T>(T(0), T(0));
      }
    }
    return;
  }
  const int C = (N + 63) / 64;
  const int k0 = lane * C, k1 = imin(N, k0 + C);
  // elimination multipliers m_k = a_k / p_{k-1} from the pivots p (a_k =
  // 1/dz^2 inside, -1 at the top row), as ThomasPressureInit forms them:
  // the same division, so the same values.  Each lane forms those of the
  // levels it sweeps, into registers (C <= CMAX).
  constexpr int CMAX = (TH_MAXN + 63) / 64;
  double ml[CMAX];
#pragma unroll
  for (int t = 0; t < CMAX; ++t) {
    const int k = k0 + t;
    ml[t] = (k >= 1 && k < k1) ? (k == nz ? -1.0 : a.c_in) / sb[k - 1] : 0.0;
  }
  // forward sweep: y_k = d_k - m_k y_{k-1}   (_thomas_solve :125-131)
  double A = 1.0, Br = 0.0, Bi = 0.0;
#pragma unroll
  for (int t = 0; t < CMAX; ++t) {
    const int k = k0 + t;
    if (k >= k1) break;
    const double al = (k == 0) ? 0.0 : -ml[t];
    A *= al;
    Br = al * Br + dr[k];
    Bi = al * Bi + di[k];
  }
  for (int off = 1; off < 64; off <<= 1) {
    const double A2 = shfl_up_d(A, off), R2 = shfl_up_d(Br, off), I2 = shfl_up_d(Bi, off);
    if (lane >= off) {
      Br = A * R2 + Br;
      Bi = A * I2 + Bi;
      A = A * A2;
    }
  }
  double yr = shfl_up_d(Br, 1), yi = shfl_up_d(Bi, 1);
  if (lane == 0) { yr = 0.0; yi = 0.0; }
#pragma unroll
  for (int t = 0; t < CMAX; ++t) {
    const int k = k0 + t;
    if (k >= k1) break;
    const double al = (k == 0) ? 0.0 : -ml[t];
    yr = al * yr + dr[k];
    yi = al * yi + di[k];
    dr[k] = yr;
    di[k] = yi;
  }
  // back substitution: x_k = (y_k - c_k x_{k+1}) / b_k   (:135-142).  The
  // quotients -c_k/b_k and y_k/b_k are formed once, kept in the lane's own
  // levels of sb, dr and di, and read again by the replay below.
  A = 1.0; Br = 0.0; Bi = 0.0;
  for (int k = k1 - 1; k >= k0; --k) {
    const double ck = (k == 0) ? 1.0 : (k == nz ? 0.0 : a.c_in);
    const double pk = sb[k];
    const double al = (k == N - 1) ? 0.0 : -ck / pk;
    const double qr = dr[k] / pk, qi = di[k] / pk;
    A *= al;
    Br = al * Br + qr;
    Bi = al * Bi + qi;
    sb[k] = al;
    dr[k] = qr;
    di[k] = qi;
  }
  for (int off = 1; off < 64; off <<= 1) {
    const double A2 = shfl_down_d(A, off), R2 = shfl_down_d(Br, off), I2 = shfl_down_d(Bi, off);
    if (lane + off < 64) {
      Br = A * R2 + Br;
      Bi = A * I2 + Bi;
      A = A * A2;
    }
  }
  double xr = shfl_down_d(Br, 1), xi = shfl_down_d(Bi, 1);
  if (lane == 63) { xr = 0.0; xi = 0.0; }
  for (int k = k1 - 1; k >= k0; --k) {
    const double al = sb[k];
    xr = al * xr + dr[k];
    xi = al * xi + di[k];
    dr[k] = xr;
    di[k] = xi;
  }
  __syncthreads();
  // p_hat[k] = x[k+1]; dpz_hat[k] = (p_hat[k] - p_hat[k-1]) / dz, 0 at k = 0
  for (int k = lane; k < nz; k += 64) {
    a.phat[cb + k] = mkc<T>((T)dr[k + 1], (T)di[k + 1]);
    if (k == 0) a.dpz[cb] = mkc<T>(T(0), T(0));
    else a.dpz[cb + k] = mkc<T>((T)((dr[k + 1] - dr[k]) * a.idz), (T)((di[k + 1] - di[k]) * a.idz));
  }
}

// R = RHS - RHS_prev/3 + 2u/(3 dt) (NSE_PressureTerms.py:140-142) as an
// R2C loader: field b in {0,1,2} = x, y, z.
template <class T> struct LoadPressureR {
  static constexpr const char* NAME = "LoadPressureR";
  static constexpr int BATCH_FAST = 1;
  static int nin(int nb) { return 3 * nb; }
  const T* R[3]; const T* Rp[3]; const T* U[3];
  int ny, nz; T third, c2dt;
  __device__ __forceinline__ T operator()(int b, int x, int y, int z) const {
    const size_t i = ((size_t)x * ny + y) * nz + z;
    return R[b][i] - third * Rp[b][i] + c2dt * U[b][i];
  }
  __device__ __forceinline__ Vec16<T> vload(int b, int x, int y, int z) const {
    const size_t i = ((size_t)x * ny + y) * nz + z;
    const Vec16<T> r = vld(R[b] + i), rp = vld(Rp[b] + i), u = vld(U[b] + i);
    Vec16<T> o;
#pragma unroll
    for (int j = 0; j < Vec16<T>::W; ++j) o.v[j] = r.v[j] - third * rp.v[j] + c2dt * u.v[j];
    return o;
  }
  bool aligned(int nb) const {
    for (int b = 0; b < nb; ++b)
      if (!al16(R[b]) || !al16(Rp[b]) || !al16(U[b])) return false;
    return true;
  }
};

// Two horizontal planes of one field (levels 0 and nz-1) as a 2-level field.
template <class T> struct LoadPlanes2 {
  static constexpr const char* NAME = "LoadPlanes2";
  static constexpr int BATCH_FAST = 0;
  static int nin(int nb) { return nb; }   // 2 of nz levels, counted over the 2-level grid
  const T* F; int ny, nz;
  __device__ __forceinline__ T operator()(int, int x, int y, int z) const {
    return F[((size_t)x * ny + y) * nz + (z == 0 ? 0 : nz - 1)];
  }
  // the two levels are not adjacent in memory: never vectorised
  __device__ __forceinline__ Vec16<T> vload(int b, int x, int y, int z) const {
    Vec16<T> r;
#pragma unroll
    for (int j = 0; j < Vec16<T>::W; ++j) r.v[j] = (*this)(b, x, y, z + j);
    return r;
  }
  bool aligned(int) const { return false; }
};
