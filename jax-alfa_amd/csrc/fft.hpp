// Batched 2-D real FFT building blocks for z-fastest (nx, ny, nz) fields.
//
// A 2-D rfft2 over axes (0, 1) (src/operations/FFT.py:44) is split the way
// numpy/XLA define it: a real-to-complex transform along y, then a complex
// transform along x.  The inverse (irfft2, e.g. Filtering.py:87) is a complex
// inverse along x followed by a complex-to-real along y that reads only the
// real parts of the ky = 0 and ky = ny/2 coefficients.
//
// Every kernel works on a tile of z levels that are contiguous in memory, so
// each HBM access is a 128-byte segment; the transform itself runs in LDS as
// a mixed-radix (4, 2, 3, 5) Stockham FFT with LPB independent lines.
//   k_y_r2c : real rows  -> half spectra, two z levels packed per complex
//             line; only the first ncy_out ky columns are written (pruned).
//   k_x     : forward x FFT of one ky column tile, then up to 4 outputs,
//             each a masked / i*k-multiplied / Dealias-remapped copy that is
//             optionally inverse transformed.
//   k_y_c2r : sum of up to 3 pruned half spectra -> real rows, through a
//             store functor (fused epilogues).
#pragma once
#include "common.hpp"
#include <type_traits>

// ----------------------------------------------------------------------------
// small DFTs; sign = -1 forward, +1 inverse
// ----------------------------------------------------------------------------
template <class T, int R> struct Dft;

template <class T> struct Dft<T, 2> {
  static __device__ __forceinline__ void run(c2<T>* v, int) {
    c2<T> a = v[0], b = v[1];
    v[0] = cadd<T>(a, b);
    v[1] = csub<T>(a, b);
  }
};
template <class T> struct Dft<T, 4> {
  static __device__ __forceinline__ void run(c2<T>* v, int sign) {
    c2<T> t0 = cadd<T>(v[0], v[2]), t1 = csub<T>(v[0], v[2]);
    c2<T> t2 = cadd<T>(v[1], v[3]), t3 = csub<T>(v[1], v[3]);
    c2<T> it3 = cmul_i<T>(t3, sign);   // sign*i*t3
    v[0] = cadd<T>(t0, t2);
    v[2] = csub<T>(t0, t2);
    v[1] = cadd<T>(t1, it3);
    v[3] = csub<T>(t1, it3);
  }
};
template <class T> struct Dft<T, 3> {
  static __device__ __forceinline__ void run(c2<T>* v, int sign) {
    const T h = T(0.86602540378443864676);  // sqrt(3)/2
    c2<T> t1 = cadd<T>(v[1], v[2]), t2 = csub<T>(v[1], v[2]);
    c2<T> y0 = cadd<T>(v[0], t1);
    c2<T> m = mkc<T>(v[0].x - T(0.5) * t1.x, v[0].y - T(0.5) * t1.y);
    c2<T> s = cscale<T>(cmul_i<T>(t2, sign), h);
    v[0] = y0;
    v[1] = cadd<T>(m, s);
    v[2] = csub<T>(m, s);
  }
};
template <class T> struct Dft<T, 5> {
  static __device__ __forceinline__ void run(c2<T>* v, int sign) {
    const T c1 = T(0.30901699437494742410), c2v = T(-0.80901699437494742410);
    const T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);
    c2<T> b1 = cadd<T>(v[1], v[4]), d1 = csub<T>(v[1], v[4]);
    c2<T> b2 = cadd<T>(v[2], v[3]), d2 = csub<T>(v[2], v[3]);
    c2<T> a0 = v[0];
    c2<T> y0 = cadd<T>(a0, cadd<T>(b1, b2));
    c2<T> r1 = mkc<T>(a0.x + c1 * b1.x + c2v * b2.x, a0.y + c1 * b1.y + c2v * b2.y);
    c2<T> r2 = mkc<T>(a0.x + c2v * b1.x + c1 * b2.x, a0.y + c2v * b1.y + c1 * b2.y);
    c2<T> q1 = cmul_i<T>(mkc<T>(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y), sign);
    c2<T> q2 = cmul_i<T>(mkc<T>(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y), sign);
    v[0] = y0;
    v[1] = cadd<T>(r1, q1);
    v[4] = csub<T>(r1, q1);
    v[2] = cadd<T>(r2, q2);
    v[3] = csub<T>(r2, q2);
  }
};

// radix 8 = 2 x 4 and radix 16 = 4 x 4 in registers: y[k1 + Q k2] =
// sum_j2 w_R^(j2 k1) w_P^(j2 k2) sum_j1 x[P j1 + j2] w_Q^(j1 k1)  (R = P Q)
template <class T> __device__ __forceinline__ c2<T> twc(c2<T> a, T c, T s, int sign) {
  // a * (c + i sign s)
  const T ss = sign > 0 ? s : -s;
  return mkc<T>(a.x * c - a.y * ss, a.x * ss + a.y * c);
}
template <class T> struct Dft<T, 8> {
  static __device__ __forceinline__ void run(c2<T>* v, int sign) {
    const T h = T(0.70710678118654752440);
    c2<T> a[4] = {v[0], v[2], v[4], v[6]};     // j2 = 0: x[2 j1]
    c2<T> b[4] = {v[1], v[3], v[5], v[7]};     // j2 = 1: x[2 j1 + 1]
    Dft<T, 4>::run(a, sign);
    Dft<T, 4>::run(b, sign);
    // twiddle w8^(k1) on the j2 = 1 branch
    b[1] = twc<T>(b[1], h, h, sign);
    b[2] = cmul_i<T>(b[2], sign);
    b[3] = twc<T>(b[3], -h, h, sign);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      v[k1] = cadd<T>(a[k1], b[k1]);
      v[k1 + 4] = csub<T>(a[k1], b[k1]);
    }
  }
};
template <class T> struct Dft<T, 16> {
  static __device__ __forceinline__ void run(c2<T>* v, int sign) {
    const T c1 = T(0.92387953251128675613), s1 = T(0.38268343236508977173);
    const T h = T(0.70710678118654752440);
    c2<T> t[4][4];
#pragma unroll
    for (int j2 = 0; j2 < 4; ++j2) {
      t[j2][0] = v[j2]; t[j2][1] = v[4 + j2]; t[j2][2] = v[8 + j2]; t[j2][3] = v[12 + j2];
      Dft<T, 4>::run(t[j2], sign);
    }
    // twiddles w16^(j2 k1)
    t[1][1] = twc<T>(t[1][1], c1, s1, sign);
    t[1][2] = twc<T>(t[1][2], h, h, sign);
    t[1][3] = twc<T>(t[1][3], s1, c1, sign);
    t[2][1] = twc<T>(t[2][1], h, h, sign);
    t[2][2] = cmul_i<T>(t[2][2], sign);
    t[2][3] = twc<T>(t[2][3], -h, h, sign);
    t[3][1] = twc<T>(t[3][1], s1, c1, sign);
    t[3][2] = twc<T>(t[3][2], -h, h, sign);
    t[3][3] = twc<T>(t[3][3], -c1, -s1, sign);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      c2<T> u[4] = {t[0][k1], t[1][k1], t[2][k1], t[3][k1]};
      Dft<T, 4>::run(u, sign);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = u[k2];
    }
  }
};

// One Stockham pass of radix R over LPB interleaved lines held in LDS
// (element (j, line p) at buf[j*LPB + p]).  All inputs of the pass are read
// into registers before any output is written, so the pass runs in place.
template <class T, int R, int LPB, int MAXN>
__device__ __forceinline__ void fft_pass(c2<T>* buf, const c2<T>* tw, int n, int Ns, int sign) {
  constexpr int MAXIT = (MAXN / R * LPB + FFT_BS - 1) / FFT_BS;
  c2<T> v[MAXIT][R];
  const int nR = n / R;
  const int total = nR * LPB;
  const int tstep = n / (Ns * R);
#pragma unroll
  for (int q = 0; q < MAXIT; ++q) {
    const int it = threadIdx.x + q * FFT_BS;
    if (it < total) {
      const int j = it / LPB, p = it - j * LPB;
      const int jm = j % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        c2<T> a = buf[(j + r * nR) * LPB + p];
        if (r > 0 && jm > 0) {
          c2<T> w = tw[jm * r * tstep];
          if (sign > 0) w.y = -w.y;
          a = cmul<T>(a, w);
        }
        v[q][r] = a;
      }
      Dft<T, R>::run(v[q], sign);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXIT; ++q) {
    const int it = threadIdx.x + q * FFT_BS;
    if (it < total) {
      const int j = it / LPB, p = it - j * LPB;
      const int jm = j % Ns;
      const int base = (j / Ns) * Ns * R + jm;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(base + r * Ns) * LPB + p] = v[q][r];
    }
  }
  __syncthreads();
}

// Full transform of LPB lines of length f.n <= MAXN (tw: n-entry table
// exp(-2 pi i t/n)).  MAXN sizes the per-thread register arrays, so kernels
// are instantiated per length tier (<= 256 and <= 384).
template <class T, int LPB, int MAXN>
__device__ void fft_lines(c2<T>* buf, const c2<T>* tw, const Fft1D& f, int sign) {
  int Ns = 1;
  for (int s = 0; s < f.nrad; ++s) {
    const int R = f.rad[s];
#if ALFA_RADIX_MAX >= 16
    if (R == 16) fft_pass<T, 16, LPB, MAXN>(buf, tw, f.n, Ns, sign);
    else
#endif
#if ALFA_RADIX_MAX >= 8
    if (R == 8) fft_pass<T, 8, LPB, MAXN>(buf, tw, f.n, Ns, sign);
    else
#endif
    if (R == 4) fft_pass<T, 4, LPB, MAXN>(buf, tw, f.n, Ns, sign);
    else if (R == 2) fft_pass<T, 2, LPB, MAXN>(buf, tw, f.n, Ns, sign);
    else if (R == 3) fft_pass<T, 3, LPB, MAXN>(buf, tw, f.n, Ns, sign);
    else fft_pass<T, 5, LPB, MAXN>(buf, tw, f.n, Ns, sign);
    Ns *= R;
  }
}

// Compile-time length.  For the grid sizes the plans meet most (see
// ct_length in plan.hpp) the kernels are instantiated with the transform
// length, its radix plan and every pass's stride as constants: index
// arithmetic strength-reduces, loop bounds are exact and the register arrays
// have their exact size.  Same passes, same radix order (8, 4, 2, 3, 5) and
// same arithmetic as fft_lines.
__host__ __device__ constexpr int radix_of(int m) {
  return m % 8 == 0 ? 8 : m % 4 == 0 ? 4 : m % 2 == 0 ? 2 : m % 3 == 0 ? 3 : 5;
}

// Layouts of the radix-8 first pass's output (LAY).  The first pass writes
// output m of butterfly j to row 8j + m.  The LDS serves a 64-bit access 16
// lanes at a time (a 128-bit one 8 at a time); with rows of at most 64 bytes
// those lanes span several j, whose rows 8j + m lie a multiple of the bank
// count apart, so they all hit the same banks (measured: 0.3 of the LDS
// cycles of the y R2C and x kernels, 0.8 of the fused contraction kernels'
// are conflict cycles).  The pass that reads the first pass's output maps
// its rows the same way; later passes use plain rows.
//   LAY 0: plain, row r at r.
//   LAY 2: padded, row r at r + r/8 (a spare row after each group of 8):
//          consecutive j land in different banks, and the addresses keep
//          their constant offsets, so it costs no VGPRs; the buffer needs
//          n/8 more rows.  (A swizzle within each group of 8 rows, m ^ (j &
//          3), needs no extra LDS but a per-lane row map: about ten more
//          VGPRs, which cost most of these kernels occupancy.)
template <int LAY> __device__ __forceinline__ int lay_row(int row) {
  return LAY == 2 ? row + (row >> 3) : row;
}
// rows of a transform buffer of length n under layout lay
__host__ __device__ constexpr int lay_rows(int n, int lay) { return lay == 2 ? n + n / 8 : n; }
// a plan of length n whose radix-8 first pass is read by a middle pass
// (FftCTMid from NS = 1 runs both), or by any second pass (FftCT)
__host__ __device__ constexpr bool mid_second(int n) {
  return n > 8 && radix_of(n) == 8 && n / 8 != radix_of(n / 8);
}
__host__ __device__ constexpr bool ct_second(int n) { return n > 8 && radix_of(n) == 8; }
// the layout of the y R2C and x kernels at compile-time length n with lines
// of rowb bytes: padded where the first pass's writes conflict
__host__ __device__ constexpr int fft_lay(int n, int rowb) { return (mid_second(n) && rowb <= 64) ? 2 : 0; }

// LAY: the first pass's output layout (written so by the NS = 1 pass, read
// so by the NS = 8 pass)
template <class T, int R, int LPB, int N, int NS, int BS, int LAY = 0>
__device__ __forceinline__ void fft_pass_ct(c2<T>* buf, const c2<T>* tw, int sign) {
  static_assert(LAY == 0 || NS != 1 || R == 8, "the first-pass layouts are for a radix-8 first pass");
  constexpr int NR = N / R;
  constexpr int TOTAL = NR * LPB;
  constexpr int MAXIT = (TOTAL + BS - 1) / BS;
  constexpr int TSTEP = N / (NS * R);
  c2<T> v[MAXIT][R];
#pragma unroll
  for (int q = 0; q < MAXIT; ++q) {
    const int it = threadIdx.x + q * BS;
    if (TOTAL % BS == 0 || it < TOTAL) {
      const int j = it / LPB, p = it % LPB;
      const int jm = j % NS;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        int row = j + r * NR;
        if constexpr (LAY == 2 && NS == 8 && NR % 8 == 0) row = j + (j >> 3) + r * (NR + NR / 8);
        else if constexpr (LAY != 0 && NS == 8) row = lay_row<LAY>(row);
        c2<T> a = buf[row * LPB + p];
        if (NS > 1 && r > 0) {
          c2<T> w = tw[jm * r * TSTEP];
          if (sign > 0) w.y = -w.y;
          a = cmul<T>(a, w);
        }
        v[q][r] = a;
      }
      Dft<T, R>::run(v[q], sign);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXIT; ++q) {
    const int it = threadIdx.x + q * BS;
    if (TOTAL % BS == 0 || it < TOTAL) {
      const int j = it / LPB, p = it % LPB;
      const int jm = j % NS;
      const int base = (j / NS) * NS * R + jm;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        int row = base + r * NS;
        if constexpr (LAY == 2 && NS == 1) row = base + (base >> 3) + r;   // base = 8j
        else if constexpr (LAY != 0 && NS == 1) row = lay_row<LAY>(row);
        buf[row * LPB + p] = v[q][r];
      }
    }
  }
  __syncthreads();
}

template <class T, int LPB, int N, int NS, int M, int BS, int LAY = 0>
struct FftCT {
  static __device__ __forceinline__ void run(c2<T>* buf, const c2<T>* tw, int sign) {
    constexpr int R = radix_of(M);
    fft_pass_ct<T, R, LPB, N, NS, BS, LAY>(buf, tw, sign);
    FftCT<T, LPB, N, NS * R, M / R, BS, NS == 1 ? LAY : 0>::run(buf, tw, sign);
  }
};
template <class T, int LPB, int N, int NS, int BS, int LAY>
struct FftCT<T, LPB, N, NS, 1, BS, LAY> {
  static __device__ __forceinline__ void run(c2<T>*, const c2<T>*, int) {}
};

// the passes of a compile-time plan after the first and before the last
// (kernels that fuse the first pass into their loads and the last pass into
// their stores run these in between)
template <class T, int LPB, int N, int NS, int M, int BS, int LAY = 0>
struct FftCTMid {
  static __device__ __forceinline__ void run(c2<T>* buf, const c2<T>* tw, int sign) {
    constexpr int R = radix_of(M);
    if constexpr (M != R) {
      fft_pass_ct<T, R, LPB, N, NS, BS, LAY>(buf, tw, sign);
      FftCTMid<T, LPB, N, NS * R, M / R, BS, NS == 1 ? LAY : 0>::run(buf, tw, sign);
    }
  }
};

// radix of the last pass of the compile-time plan of length n
__host__ __device__ constexpr int last_radix(int m) {
  while (m != radix_of(m)) m /= radix_of(m);
  return m;
}

// ----------------------------------------------------------------------------
// Band-limited passes.  A test- or explicit-filtered spectrum is nonzero only
// at |k| < c, and the filters keep only those modes of a forward transform.
// With c <= N/4 (twice the stride N/8 of the radix-8 first pass of a
// compile-time plan), inputs r = 2..5 of every first inverse butterfly are
// zero: only r = 0, 1 (k = j, j + N/8) and r = 6, 7 (k = j - N/4, j - N/8)
// can be nonzero.  With c <= N/RL (the stride of the last forward pass,
// radix RL), the kept forward modes are outputs r = 0 and r = RL-1 of the
// last pass.  These helpers read and compute only those.
// ----------------------------------------------------------------------------

// Output ROUT of butterfly jj of the last pass of the compile-time forward
// plan of length N (radix RL, stride N/RL, twiddle step 1), read from LDS:
// the value fft_pass_ct writes to element jj + ROUT*N/RL, by the same
// arithmetic.
template <class T, int N, int L, int ROUT>
__device__ __forceinline__ c2<T> fwd_last_out(const c2<T>* buf, const c2<T>* tw, int jj, int p) {
  constexpr int RL = last_radix(N), NSL = N / RL;
  c2<T> v[RL];
#pragma unroll
  for (int r = 0; r < RL; ++r) {
    c2<T> a = buf[(jj + r * NSL) * L + p];
    if (r > 0) a = cmul<T>(a, tw[jj * r]);
    v[r] = a;
  }
  Dft<T, RL>::run(v, -1);
  return v[ROUT];
}

// Last pass of the compile-time forward plan of length N over L lines, with
// only the outputs k < ncy and N - k (0 < k < ncy) written (ncy <= N/RL):
// output 0 of butterflies j < ncy and output RL-1 of butterflies j > N/RL -
// ncy.  Ends with a barrier, like a full pass.
template <class T, int L, int N, int BS>
__device__ __forceinline__ void fwd_last_pass_band(c2<T>* buf, const c2<T>* tw, int ncy) {
  constexpr int RL = last_radix(N), NSL = N / RL, TOT = NSL * L, IT = (TOT + BS - 1) / BS;
  c2<T> lo[IT], hi[IT];
#pragma unroll
  for (int q = 0; q < IT; ++q) {
    const int it = threadIdx.x + q * BS;
    if (TOT % BS == 0 || it < TOT) {
      const int j = it / L, p = it % L;
      c2<T> v[RL];
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        c2<T> a = buf[(j + r * NSL) * L + p];
        if (r > 0) a = cmul<T>(a, tw[j * r]);
        v[r] = a;
      }
      Dft<T, RL>::run(v, -1);
      lo[q] = v[0];
      hi[q] = v[RL - 1];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < IT; ++q) {
    const int it = threadIdx.x + q * BS;
    if (TOT % BS == 0 || it < TOT) {
      const int j = it / L, p = it % L;
      if (j < ncy) buf[j * L + p] = lo[q];
      if (j > NSL - ncy) buf[(j + (RL - 1) * NSL) * L + p] = hi[q];
    }
  }
  __syncthreads();
}

// the inputs r of a radix-8 first-pass butterfly that a band |k| < c with
// c <= N/4 can make nonzero
__host__ __device__ constexpr bool band_r(int r) { return r < 2 || r > 5; }

// First pass (radix 8, stride 1) of the inverse transform of L lines whose
// modes are zero outside k < kmax and k > N - kmax (kmax <= N/4); only the
// nonzero elements of buf are read.  The butterfly is the full radix-8 one
// with zeros in place, so the result is that of the full pass.  Ends with a
// barrier, like a full pass.
template <class T, int L, int N, int BS, int LAY = 0>
__device__ __forceinline__ void band_first_pass_inv(c2<T>* buf, int kmax) {
  constexpr int NR1 = N / 8, TOT = NR1 * L, IT = (TOT + BS - 1) / BS;
  c2<T> v[IT][8];
#pragma unroll
  for (int q = 0; q < IT; ++q) {
    const int it = threadIdx.x + q * BS;
#pragma unroll
    for (int r = 0; r < 8; ++r) v[q][r] = mkc<T>(0, 0);
    if (TOT % BS == 0 || it < TOT) {
      const int j = it / L, p = it % L;
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        if (!band_r(r)) continue;
        const int k = j + r * NR1;
        if (r < 2 ? k < kmax : k > N - kmax) v[q][r] = buf[k * L + p];
      }
      Dft<T, 8>::run(v[q], +1);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < IT; ++q) {
    const int it = threadIdx.x + q * BS;
    if (TOT % BS == 0 || it < TOT) {
      const int j = it / L, p = it % L;
#pragma unroll
      for (int m = 0; m < 8; ++m)
        buf[(LAY == 2 ? 9 * j + m : lay_row<LAY>(8 * j + m)) * L + p] = v[q][m];
    }
  }
  __syncthreads();
}

// transform of LPB lines by BS threads: compile-time length NCT, or f.n when
// NCT == 0 (the generic passes run with FFT_BS threads)
template <class T, int LPB, int MAXN, int NCT, int BS = FFT_BS, int LAY = 0>
__device__ __forceinline__ void fft_any(c2<T>* buf, const c2<T>* tw, const Fft1D& f, int sign) {
  if constexpr (NCT > 0) FftCT<T, LPB, NCT, 1, NCT, BS, ct_second(NCT) ? LAY : 0>::run(buf, tw, sign);
  else fft_lines<T, LPB, MAXN>(buf, tw, f, sign);
}

// Twiddles of an n-point plan, loaded to registers at kernel start and
// written to LDS after the first tile's loads: the two loads are in flight
// together instead of one after the other.
// DEFER = false: written to LDS at once (where the held registers would
// cost occupancy).
template <class T, int BS, int MAXN, bool DEFER = true>
struct TwStage {
  static constexpr int K = DEFER ? (MAXN + BS - 1) / BS : 1;
  c2<T> r[K];
  int n;
  bool pending;
  __device__ __forceinline__ void load(const c2<T>* tw, int n_, c2<T>* stw) {
    n = n_;
    pending = DEFER;
    if (!DEFER) {
      for (int i = threadIdx.x; i < n; i += BS) stw[i] = tw[i];
      return;
    }
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const int i = threadIdx.x + q * BS;
      if (i < n) r[q] = tw[i];
    }
  }
  // call between a tile's LDS stores and the barrier before its transform
  __device__ __forceinline__ void store(c2<T>* stw) {
    if (!DEFER || !pending) return;
    pending = false;
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const int i = threadIdx.x + q * BS;
      if (i < n) stw[i] = r[q];
    }
  }
};

struct PtrBatch {
  void* p[MAXB];
};

// ----------------------------------------------------------------------------
// Tile geometry.  A launch covers ntz * rows * nb tiles (z tile, transform
// row, field), one tile per workgroup; the grid is rounded up to a multiple
// of the 8 XCDs.  The tiles are numbered so that each XCD (workgroup i runs
// on XCD i mod 8) owns one contiguous range of them: tiles that share HBM
// rows or input data run on the same L2 at about the same time.  With
// batch_fast the fields of one (row, z tile) are adjacent (fused loaders
// read the same inputs for all of them), otherwise the z tiles of one row
// are.  tile(i, ...) is the i-th tile of a workgroup (a grid smaller than
// the tile count walks its range gridDim.x/8 tiles at a time).
// ----------------------------------------------------------------------------
struct TileGrid {
  int ntz, rows, nb, batch_fast;
  __device__ __forceinline__ bool tile(int i, int& zt, int& row, int& b) const {
    const unsigned tot = (unsigned)ntz * rows * nb;
    const unsigned x = blockIdx.x % NUM_XCD, j = blockIdx.x / NUM_XCD;
    const unsigned per = tot / NUM_XCD, rem = tot % NUM_XCD;
    const unsigned lo = x * per + (x < rem ? x : rem), cnt = per + (x < rem ? 1u : 0u);
    const unsigned k = (unsigned)i * (gridDim.x / NUM_XCD) + j;
    if (k >= cnt) return false;
    unsigned l = lo + k;
    if (batch_fast) {
      b = l % nb; l /= nb;
      zt = l % ntz; row = l / ntz;
    } else {
      zt = l % ntz; l /= ntz;
      row = l % rows; b = l / rows;
    }
    return true;
  }
};

// ----------------------------------------------------------------------------
// y-direction R2C: row x of field b (read through the loader) -> the first
// ncy_out columns of its half spectrum, (rows, ncy_out, nz) complex.  With
// vec set (nz a multiple of the vector width, 16-byte aligned buffers) every
// thread loads 16 bytes of z at a time through ld.vload and stores the two
// complex values of one packed line together.  The compile-time-length
// kernels (NCT > 0) are built for vec only (the launcher sends other cases
// to the generic kernel), which keeps the element-wise path and its
// registers out of them.  BAND (compile-time lengths with ncy_out <= N/RL,
// chosen by the launcher): the last radix pass is evaluated only where the
// store reads it.
// ----------------------------------------------------------------------------
// number of field kinds of a loader with per-kind load loops (with_kind /
// vload_k), 0 for loaders with one vload for every field
template <class L, class = void> struct LoaderKinds { static constexpr int value = 0; };
template <class L> struct LoaderKinds<L, std::void_t<decltype(L::KINDS)>> {
  static constexpr int value = L::KINDS;
};

template <class T, class Loader, int MAXN, int NCT, int BS, bool BAND = false>
__device__ __forceinline__ void
k_y_r2c(const Loader& ld, const Fft1D& fy, const c2<T>* __restrict__ tw, int nz, int ncy_out, const PtrBatch& outb,
        const TileGrid& tg, int vec) {
  constexpr int L = fft_lines<T>(BS);
  constexpr int ZT = 2 * L;
  constexpr int W = Vec16<T>::W;
  // padded first-pass rows where its writes would conflict
  constexpr int LAY = NCT > 0 ? fft_lay(NCT, L * (int)sizeof(c2<T>)) : 0;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int n = NCT > 0 ? NCT : fy.n, nEl = n * ZT;
  c2<T>* buf = reinterpret_cast<c2<T>*>(smem_raw);
  c2<T>* stw = buf + lay_rows(n, LAY) * L;
  T* bufr = reinterpret_cast<T*>(buf);
  TwStage<T, BS, (NCT > 0 ? NCT : MAXN)> tws;
  tws.load(tw, n, stw);
  const bool vv = NCT > 0 || vec;
  // the transform of the staged tile (zt, x) of field b and the store of its
  // first ncy_out columns
  auto finish = [&](int zt, int x, int b) {
    tws.store(stw);
    __syncthreads();
    // band (compile-time lengths, ncy_out <= N/RL): the last pass is
    // evaluated only at the outputs k < ncy_out and n - k that the store reads
    if constexpr (BAND) {
      FftCTMid<T, L, NCT, 1, NCT, BS, LAY>::run(buf, stw, -1);
      fwd_last_pass_band<T, L, NCT, BS>(buf, stw, ncy_out);
    } else {
      fft_any<T, L, MAXN, NCT, BS, LAY>(buf, stw, fy, -1);
    }
    // line p holds z0+2p (real part) and z0+2p+1 (imaginary part)
    c2<T>* out = reinterpret_cast<c2<T>*>(outb.p[b]);
    const int z0 = zt * ZT;
    for (int i = threadIdx.x; i < ncy_out * L; i += BS) {
      const int k = i / L, p = i - k * L;
      const int z = z0 + 2 * p;
      if (z >= nz) continue;
      const c2<T> Zk = buf[k * L + p];
      const c2<T> Zn = buf[((n - k) % n) * L + p];
      const c2<T> r0 = mkc<T>(T(0.5) * (Zk.x + Zn.x), T(0.5) * (Zk.y - Zn.y));
      const c2<T> r1 = mkc<T>(T(0.5) * (Zk.y + Zn.y), T(-0.5) * (Zk.x - Zn.x));
      c2<T>* o = out + ((size_t)x * ncy_out + k) * nz + z;
      if (vv) {
        cst2<T>(o, r0, r1);
      } else {
        o[0] = r0;
        if (z + 1 < nz) o[1] = r1;
      }
    }
  };
  int zt, x, b;
  if constexpr (NCT > 0) {
    // Compile-time lengths: thread t stages the NV vectors e0 + q*BS*W of the
    // tile, all at the same z (BS*W is a multiple of the tile width ZT).
    // Loaders of stored fields issue all NV loads at once, before the barrier
    // that waits for the previous tile's output reads.  Loaders that form
    // products (more registers per vector) load one vector at a time after
    // the barrier: measured at 256^3, holding their NV vectors costs more in
    // occupancy than the overlap gains.
    constexpr int NEL = NCT * ZT, NV = (NEL + BS * W - 1) / (BS * W), YS = BS * W / ZT;
    static_assert((BS * W) % ZT == 0, "a thread's vectors share one z");
    const int e0 = threadIdx.x * W, y0 = e0 / ZT, zo = e0 - y0 * ZT;
    auto in_tile = [&](int q) { return NEL % (BS * W) == 0 || e0 + q * BS * W < NEL; };
    for (int it = 0; tg.tile(it, zt, x, b); ++it) {
      if constexpr (LoaderKinds<Loader>::value > 0) {
        __syncthreads();
        ld.with_kind(b, [&](auto kc) {
          constexpr int K = decltype(kc)::value;
#pragma unroll 1
          for (int e = e0; e < NEL; e += BS * W) {
            const int y = e / ZT, ze = zt * ZT + (e - y * ZT);
            vst(bufr + e, ze < nz ? ld.template vload_k<K>(b, x, y, ze) : vzero<T>());
          }
        });
      } else {
        const int z = zt * ZT + zo;
        Vec16<T> rq[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q)
          if (in_tile(q)) rq[q] = z < nz ? ld.vload(b, x, y0 + q * YS, z) : vzero<T>();
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NV; ++q)
          if (in_tile(q)) vst(bufr + e0 + q * BS * W, rq[q]);
      }
      finish(zt, x, b);
    }
  } else {
    for (int it = 0; tg.tile(it, zt, x, b); ++it) {
      __syncthreads();
      if (vv) {
        if constexpr (LoaderKinds<Loader>::value > 0) {
          ld.with_kind(b, [&](auto kc) {
            constexpr int K = decltype(kc)::value;
            for (int e = threadIdx.x * W; e < nEl; e += BS * W) {
              const int y = e / ZT, z = zt * ZT + (e - y * ZT);
              vst(bufr + e, z < nz ? ld.template vload_k<K>(b, x, y, z) : vzero<T>());
            }
          });
        } else {
          for (int e = threadIdx.x * W; e < nEl; e += BS * W) {
            const int y = e / ZT, z = zt * ZT + (e - y * ZT);
            vst(bufr + e, z < nz ? ld.vload(b, x, y, z) : vzero<T>());
          }
        }
      } else {
        for (int i = threadIdx.x; i < nEl; i += BS) {
          const int y = i / ZT, z = zt * ZT + (i - y * ZT);
          bufr[i] = (z < nz) ? ld(b, x, y, z) : T(0);
        }
      }
      finish(zt, x, b);
    }
  }
}

// ----------------------------------------------------------------------------
// x-direction multi-output kernel.
// ----------------------------------------------------------------------------
enum { XM_NONE = 0, XM_IKX = 1, XM_IKY = 2 };
enum { XMAP_ID = 0, XMAP_D1 = 1, XMAP_D2 = 2 };

struct XOp {
  int ncy;       // columns of dst; ky >= ncy is not written
  int inverse;   // 1: inverse x FFT before the store
  int cx;        // keep kx index in [0,cx) U [n-cx+1, n)
  int mult;      // XM_*
  int map;       // XMAP_*
  int nout;      // x length of dst (== nin unless map != XMAP_ID)
  double scale;
};

struct XArgs {
  const void* in[MAXB];
  void* dst[MAXB][4];
  XOp op[4];
  int nops;
  int ncy_in;
  int nz;
  int in_spectral;   // 1: input is already x-transformed
  int ny_full;       // regular-grid ny (ky Nyquist column = ny_full/2)
  int vec;           // 16-byte accesses (set by the launcher: nz even, aligned)
};

// value of op `op` at input mode k of a forward spectrum
template <class T>
__device__ __forceinline__ c2<T> x_apply(const XOp& op, c2<T> v, int k, int nin, int kyv) {
  const bool keep = (k < op.cx) || (k >= nin - op.cx + 1);
  if (!keep) return mkc<T>(0, 0);
  if (op.mult == XM_IKX) {
    int kx = (k < nin / 2) ? k : k - nin;
    if (2 * k == nin) kx = 0;
    v = cscale<T>(cmul_i<T>(v, 1), (T)kx);
  } else if (op.mult == XM_IKY) {
    v = cscale<T>(cmul_i<T>(v, 1), (T)kyv);
  }
  return cscale<T>(v, (T)op.scale);
}

// load tile (zt, ky) of field b into LDS, (k, line) at buf[k*L + line];
// with vec, float lines are moved in pairs (16 bytes).  NEL > 0 (float,
// compile-time nEl): every load of the tile is issued before the first LDS
// write, one HBM round trip per tile instead of one per loop iteration
// (measured faster for the single-op float kernel only: elsewhere the
// registers it holds cost more occupancy than it gains).
template <class T, int L, int BS, int NEL = 0>
__device__ __forceinline__ void x_load(c2<T>* buf, const XArgs& a, int nEl, int zt, int ky, int b) {
  const c2<T>* in = reinterpret_cast<const c2<T>*>(a.in[b]);
  const int nz = a.nz;
  if constexpr (NEL > 0 && sizeof(T) == 4) {
    constexpr int NV = (NEL + 2 * BS - 1) / (2 * BS);
    if (a.vec) {
      c2<T> v0[NV], v1[NV];
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        const int e = 2 * (threadIdx.x + q * BS);
        v0[q] = mkc<T>(0, 0);
        v1[q] = mkc<T>(0, 0);
        if (NEL % (2 * BS) == 0 || e < NEL) {
          const int k = e / L, z = zt * L + (e - k * L);
          if (z < nz) cld2<T>(in + ((size_t)k * a.ncy_in + ky) * nz + z, v0[q], v1[q]);
        }
      }
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        const int e = 2 * (threadIdx.x + q * BS);
        if (NEL % (2 * BS) == 0 || e < NEL) {
          buf[e] = v0[q];
          buf[e + 1] = v1[q];
        }
      }
      return;
    }
  }
  if (sizeof(T) == 4 && a.vec) {
    for (int e = 2 * threadIdx.x; e < nEl; e += 2 * BS) {
      const int k = e / L, z = zt * L + (e - k * L);
      c2<T> v0 = mkc<T>(0, 0), v1 = mkc<T>(0, 0);
      if (z < nz) cld2<T>(in + ((size_t)k * a.ncy_in + ky) * nz + z, v0, v1);
      buf[e] = v0;
      buf[e + 1] = v1;
    }
    return;
  }
  for (int i = threadIdx.x; i < nEl; i += BS) {
    const int k = i / L, z = zt * L + (i - k * L);
    buf[i] = (z < nz) ? in[((size_t)k * a.ncy_in + ky) * nz + z] : mkc<T>(0, 0);
  }
}

template <class T, int L, int BS>
__device__ __forceinline__ void x_store(const c2<T>* buf, void* dstp, int ncy, int nz, int nout,
                                        int zt, int ky, int vec) {
  c2<T>* dst = reinterpret_cast<c2<T>*>(dstp);
  const int z0 = zt * L;
  if (sizeof(T) == 4 && vec) {
    for (int e = 2 * threadIdx.x; e < nout * L; e += 2 * BS) {
      const int k = e / L, z = z0 + (e - k * L);
      if (z < nz) cst2<T>(dst + ((size_t)k * ncy + ky) * nz + z, buf[e], buf[e + 1]);
    }
    return;
  }
  for (int i = threadIdx.x; i < nout * L; i += BS) {
    const int k = i / L, z = z0 + (i - k * L);
    if (z < nz) dst[((size_t)k * ncy + ky) * nz + z] = buf[i];
  }
}

// Forward spectrum -> op -> first inverse pass, for a single op with mask
// c = op.cx <= min(N/4, N/RL) followed by an inverse transform: after the
// forward passes but the last, each item (j, line p) evaluates the forward
// outputs at the modes k = j + r N/8 (r = 0, 1, 6, 7) the mask keeps, applies
// the op and runs the radix-8 inverse butterfly j on them (zeros elsewhere,
// as the full path has after the mask).  Ends with a barrier.
template <class T, int L, int N, int BS, int LAY = 0>
__device__ __forceinline__ void x_band_mid(c2<T>* buf, const c2<T>* tw, const XOp& op, int kyv) {
  constexpr int RL = last_radix(N), NSL = N / RL, NR1 = N / 8;
  constexpr int TOT = NR1 * L, IT = (TOT + BS - 1) / BS;
  c2<T> v[IT][8];
#pragma unroll
  for (int q = 0; q < IT; ++q) {
    const int it = threadIdx.x + q * BS;
#pragma unroll
    for (int r = 0; r < 8; ++r) v[q][r] = mkc<T>(0, 0);
    if (TOT % BS == 0 || it < TOT) {
      const int j = it / L, p = it % L;
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        if (!band_r(r)) continue;
        const int k = j + r * NR1;
        if (r < 2) {   // k < c <= N/RL: output 0 of forward butterfly k
          if (k < op.cx)
            v[q][r] = x_apply<T>(op, fwd_last_out<T, N, L, 0>(buf, tw, k, p), k, N, kyv);
        } else {       // k > N - c >= N - N/RL: output RL-1
          if (k >= N - op.cx + 1)
            v[q][r] = x_apply<T>(op, fwd_last_out<T, N, L, RL - 1>(buf, tw, k - (RL - 1) * NSL, p),
                                 k, N, kyv);
        }
      }
      Dft<T, 8>::run(v[q], +1);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < IT; ++q) {
    const int it = threadIdx.x + q * BS;
    if (TOT % BS == 0 || it < TOT) {
      const int j = it / L, p = it % L;
#pragma unroll
      for (int m = 0; m < 8; ++m)
        buf[(LAY == 2 ? 9 * j + m : lay_row<LAY>(8 * j + m)) * L + p] = v[q][m];
    }
  }
  __syncthreads();
}

// One ky column tile (L z lines, full x length) per iteration of a
// workgroup.  MULTI = false: one op with the identity map, applied in place in
// LDS (the common case; small register footprint).  MULTI = true: the
// forward spectrum is kept in registers and every op (mask / derivative /
// dealias map, optional inverse) is built from it in turn.
template <class T, int MAXN, bool MULTI, int NCT, int BS, bool BAND = false>
__device__ __forceinline__ void
k_x(const XArgs& a, const Fft1D& fin, const Fft1D& fout, const c2<T>* __restrict__ tw_in,
    const c2<T>* __restrict__ tw_out, const TileGrid& tg) {
  constexpr int L = fft_lines<T>(BS);
  // spectrum elements per thread in the multi-op path: rounded up, since
  // MAXN * L need not be a multiple of BS (96 double lines at 256 threads)
  constexpr int MAXE = (MAXN * L + BS - 1) / BS;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int nin = NCT > 0 ? NCT : fin.n;   // NCT is set only when fin.n == fout.n
  const int nbuf = NCT > 0 ? NCT : imax(nin, fout.n);
  // padded first-pass rows where its writes would conflict
  constexpr int LAY = NCT > 0 ? fft_lay(NCT, L * (int)sizeof(c2<T>)) : 0;
  c2<T>* buf = reinterpret_cast<c2<T>*>(smem_raw);
  c2<T>* stw_in = buf + lay_rows(nbuf, LAY) * L;
  c2<T>* stw_out = stw_in + nin;   // (lengths without a compile-time plan only)
  const int nz = a.nz;
  const int nEl = nin * L;
  TwStage<T, BS, (NCT > 0 ? NCT : MAXN)> tws;
  tws.load(tw_in, nin, stw_in);
  if (MULTI && NCT == 0)
    for (int i = threadIdx.x; i < fout.n; i += BS) stw_out[i] = tw_out[i];
  int zt, ky, b;
  for (int it = 0; tg.tile(it, zt, ky, b); ++it) {
    if (!MULTI && ky >= a.op[0].ncy) continue;
    __syncthreads();
    x_load<T, L, BS, (!MULTI && ALFA_BATCH_LOADS) ? NCT * L : 0>(buf, a, nEl, zt, ky, b);
    tws.store(stw_in);
    __syncthreads();
    const int kyv = (2 * ky == a.ny_full) ? 0 : ky;
    if constexpr (BAND && !MULTI) {
      // (chosen by the launcher) x-space input, one op with the identity
      // map, mask c <= min(N/4, N/RL), inverse transform
      FftCTMid<T, L, NCT, 1, NCT, BS, LAY>::run(buf, stw_in, -1);
      x_band_mid<T, L, NCT, BS, LAY>(buf, stw_in, a.op[0], kyv);
      FftCT<T, L, NCT, 8, NCT / 8, BS, LAY>::run(buf, stw_in, +1);
      x_store<T, L, BS>(buf, a.dst[b][0], a.op[0].ncy, nz, nin, zt, ky, a.vec);
      continue;
    }
    if constexpr (BAND && MULTI) {
      // (chosen by the launcher) the same for several such ops: the forward
      // passes but the last run once, and their output, held in registers,
      // feeds each op's band middle (last forward pass at the kept modes,
      // op, first inverse pass) and inverse passes
      constexpr int ME = (NCT * L + BS - 1) / BS;
      FftCTMid<T, L, NCT, 1, NCT, BS, LAY>::run(buf, stw_in, -1);
      c2<T> mid[ME];
#pragma unroll
      for (int q = 0; q < ME; ++q) {
        const int i = threadIdx.x + q * BS;
        mid[q] = (NCT * L % BS == 0 || i < NCT * L) ? buf[i] : mkc<T>(0, 0);
      }
      bool used = false;   // buf no longer holds the forward passes' output
      for (int o = 0; o < a.nops; ++o) {
        const XOp op = a.op[o];
        if (ky >= op.ncy) continue;
        if (used) {
          __syncthreads();   // the previous op's store has read buf
#pragma unroll
          for (int q = 0; q < ME; ++q) {
            const int i = threadIdx.x + q * BS;
            if (NCT * L % BS == 0 || i < NCT * L) buf[i] = mid[q];
          }
          __syncthreads();
        }
        x_band_mid<T, L, NCT, BS, LAY>(buf, stw_in, op, kyv);
        FftCT<T, L, NCT, 8, NCT / 8, BS, LAY>::run(buf, stw_in, +1);
        x_store<T, L, BS>(buf, a.dst[b][o], op.ncy, nz, nin, zt, ky, a.vec);
        used = true;
      }
      continue;
    }
    if (!a.in_spectral) fft_any<T, L, MAXN, NCT, BS, LAY>(buf, stw_in, fin, -1);
    if (!MULTI) {
      const XOp op = a.op[0];
      for (int i = threadIdx.x; i < nEl; i += BS) buf[i] = x_apply<T>(op, buf[i], i / L, nin, kyv);
      __syncthreads();
      if (op.inverse) fft_any<T, L, MAXN, NCT, BS, LAY>(buf, stw_in, fin, +1);
      x_store<T, L, BS>(buf, a.dst[b][0], op.ncy, nz, nin, zt, ky, a.vec);
    } else {
      c2<T> spec[MAXE];
#pragma unroll
      for (int q = 0; q < MAXE; ++q) {
        const int i = threadIdx.x + q * BS;
        spec[q] = (i < nEl) ? buf[i] : mkc<T>(0, 0);
      }
      for (int o = 0; o < a.nops; ++o) {
        const XOp op = a.op[o];
        if (ky >= op.ncy) continue;
        const int nout = op.nout;
        __syncthreads();
        if (NCT == 0 && op.map != XMAP_ID) {
          for (int i = threadIdx.x; i < nout * L; i += BS) buf[i] = mkc<T>(0, 0);
          __syncthreads();
        }
#pragma unroll
        for (int q = 0; q < MAXE; ++q) {
          const int i = threadIdx.x + q * BS;
          if (i < nEl) {
            const int k = i / L, p = i - k * L;
            const c2<T> v = x_apply<T>(op, spec[q], k, nin, kyv);
            // compile-time plans have fin.n == fout.n: identity map only
            if (NCT > 0 || op.map == XMAP_ID) {
              buf[i] = v;
            } else if (op.map == XMAP_D1) {
              if (k <= nin / 2) buf[k * L + p] = v;
              if (k >= nin / 2) buf[(nout - nin + k) * L + p] = v;
            } else {
              const int h = nout / 2;
              if (k < h) buf[k * L + p] = v;
              else if (k >= nin - (h - 1)) buf[(k - nin + nout) * L + p] = v;
            }
          }
        }
        __syncthreads();
        if (op.inverse) {
          if (NCT > 0 || op.map == XMAP_ID) fft_any<T, L, MAXN, NCT, BS, LAY>(buf, stw_in, fin, +1);
          else fft_any<T, L, MAXN, NCT, BS>(buf, stw_out, fout, +1);
        }
        x_store<T, L, BS>(buf, a.dst[b][o], op.ncy, nz, nout, zt, ky, a.vec);
      }
    }
  }
}

// ----------------------------------------------------------------------------
// y-direction C2R.  Output b of row x is the sum of nsum pruned half
// spectra (rows, ncy_in, nz); values go to st(b, x, y, z, value*scale).
// ----------------------------------------------------------------------------
struct YC2RArgs {
  const void* in[MAXB][3];
  int dy[MAXB];     // bit s set: input s of output b is multiplied by i*ky*dyscale
  int nsum;
  int ncy_in;
  int nz;
  int vec;          // 16-byte accesses (set by the launcher: nz % W == 0, aligned)
  double scale;
  double dyscale;
};

// the sums of the nsum inputs of output b at (x, ky = k) for the z pair
// (z, z+1) of one packed line; the y wavenumber of the Nyquist column
// (2k == n) is taken as 0, as in the x pass.
template <class T>
__device__ __forceinline__ void yc2r_in2(const YC2RArgs& a, int b, int x, int k, int z, int n,
                                         c2<T>& X1, c2<T>& X2) {
  X1 = mkc<T>(0, 0);
  X2 = mkc<T>(0, 0);
  const size_t base = ((size_t)x * a.ncy_in + k) * a.nz + z;
  const T kyv = (T)(a.dyscale * ((2 * k == n) ? 0 : k));
  for (int s = 0; s < a.nsum; ++s) {
    const c2<T>* in = reinterpret_cast<const c2<T>*>(a.in[b][s]) + base;
    c2<T> v1, v2 = mkc<T>(0, 0);
    if (a.vec) {
      cld2<T>(in, v1, v2);
    } else {
      v1 = in[0];
      if (z + 1 < a.nz) v2 = in[1];
    }
    if ((a.dy[b] >> s) & 1) {
      v1 = cscale<T>(cmul_i<T>(v1, 1), kyv);
      v2 = cscale<T>(cmul_i<T>(v2, 1), kyv);
    }
    X1 = cadd<T>(X1, v1);
    X2 = cadd<T>(X2, v2);
  }
}

// BS = 512 (compile-time lengths only): twice the lines per tile, so each
// tile row is 128 bytes of z.
// Compile-time length, vector path: the last radix pass writes its outputs
// straight to the storer, two z levels per store, instead of through LDS.
// The arithmetic and its order are those of the LDS path below.
template <class T, class Storer, int N, int BS>
__device__ __forceinline__ void y_c2r_last_pass_store(const Storer& st, const YC2RArgs& a,
                                                      const c2<T>* buf, const c2<T>* stw, int z0,
                                                      int x, int b) {
  constexpr int L = fft_lines<T>(BS);
  constexpr int R = last_radix(N), NS = N / R, TOT = NS * L, IT = (TOT + BS - 1) / BS;
  const T sc = (T)a.scale;
  if constexpr (sizeof(T) == 4 && L % 2 == 0 && ALFA_C2R_PAIRS) {
    // float: two adjacent lines (four z levels) per item, so each output row
    // is one 16-byte store (and one 16-byte LDS read per input)
    constexpr int LP = L / 2, TP = NS * LP, ITP = (TP + BS - 1) / BS;
#pragma unroll
    for (int q = 0; q < ITP; ++q) {
      const int it = threadIdx.x + q * BS;
      if (TP % BS == 0 || it < TP) {
        const int j = it / LP, pp = it % LP, z = z0 + 4 * pp;
        c2<T> v0[R], v1[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          c2<T> e0, e1;
          cld2<T>(buf + (j + r * NS) * L + 2 * pp, e0, e1);
          if (r > 0) {
            c2<T> w = stw[j * r];
            w.y = -w.y;
            e0 = cmul<T>(e0, w);
            e1 = cmul<T>(e1, w);
          }
          v0[r] = e0;
          v1[r] = e1;
        }
        Dft<T, R>::run(v0, +1);
        Dft<T, R>::run(v1, +1);
        if (z < a.nz) {
#pragma unroll
          for (int r = 0; r < R; ++r) {
            Vec16<T> o;
            o.v[0] = v0[r].x * sc; o.v[1] = v0[r].y * sc;
            o.v[2] = v1[r].x * sc; o.v[3] = v1[r].y * sc;
            st.vstore(b, x, j + r * NS, z, o);
          }
        }
      }
    }
    return;
  }
#pragma unroll
  for (int q = 0; q < IT; ++q) {
    const int it = threadIdx.x + q * BS;
    if (TOT % BS == 0 || it < TOT) {
      const int j = it / L, p = it % L, z = z0 + 2 * p;
      c2<T> v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        c2<T> e = buf[(j + r * NS) * L + p];
        if (r > 0) {
          c2<T> w = stw[j * r];
          w.y = -w.y;
          e = cmul<T>(e, w);
        }
        v[r] = e;
      }
      Dft<T, R>::run(v, +1);
      if (z < a.nz) {
#pragma unroll
        for (int r = 0; r < R; ++r) st.store2(b, x, j + r * NS, z, v[r].x * sc, v[r].y * sc);
      }
    }
  }
}

// BAND (compile-time lengths with ncy_in <= N/4, chosen by the launcher):
// the first inverse pass reads only the nonzero modes.
template <class T, class Storer, int MAXN, int NCT, int BS, bool FUSED = false, bool BAND = false>
__device__ __forceinline__ void
k_y_c2r(const Storer& st, const YC2RArgs& a, const Fft1D& fy, const c2<T>* __restrict__ tw, const TileGrid& tg) {
  constexpr int L = fft_lines<T>(BS);
  constexpr int ZT = 2 * L;
  constexpr int W = Vec16<T>::W;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int n = NCT > 0 ? NCT : fy.n, half = n / 2;
  c2<T>* buf = reinterpret_cast<c2<T>*>(smem_raw);
  c2<T>* stw = buf + n * L;
  const int nz = a.nz;
  const int kmax = imin(a.ncy_in, half + 1);
  const int nIn = kmax * L;
  TwStage<T, BS, (NCT > 0 ? NCT : MAXN), sizeof(T) == 4> tws;
  tws.load(tw, n, stw);
  int zt, x, b;
  for (int it = 0; tg.tile(it, zt, x, b); ++it) {
    __syncthreads();
    // the modes kmax..n-kmax are zero (band: the first pass reads none of them)
    if constexpr (!BAND) {
      for (int i = threadIdx.x; i < n * L; i += BS) {
        const int k = i / L;
        if (k >= kmax && k <= n - kmax) buf[i] = mkc<T>(0, 0);
      }
    }
    for (int i = threadIdx.x; i < nIn; i += BS) {
      const int k = i / L, p = i - k * L;
      const int z = zt * ZT + 2 * p;
      c2<T> X1 = mkc<T>(0, 0), X2 = mkc<T>(0, 0);
      if (z < nz) yc2r_in2<T>(a, b, x, k, z, n, X1, X2);
      if (k == 0 || k == half) { X1.y = T(0); X2.y = T(0); }
      buf[k * L + p] = mkc<T>(X1.x - X2.y, X1.y + X2.x);
      if (k > 0 && k < half) buf[(n - k) * L + p] = mkc<T>(X1.x + X2.y, X2.x - X1.y);
    }
    tws.store(stw);
    __syncthreads();
    const int z0 = zt * ZT;
    if constexpr (FUSED) {  // NCT > 0 and a.vec
      if constexpr (BAND) {
        band_first_pass_inv<T, L, NCT, BS>(buf, kmax);
        FftCTMid<T, L, NCT, 8, NCT / 8, BS>::run(buf, stw, +1);
      } else {
        FftCTMid<T, L, NCT, 1, NCT, BS>::run(buf, stw, +1);
      }
      y_c2r_last_pass_store<T, Storer, NCT, BS>(st, a, buf, stw, z0, x, b);
      continue;
    }
    if constexpr (BAND) {
      band_first_pass_inv<T, L, NCT, BS>(buf, kmax);
      FftCT<T, L, NCT, 8, NCT / 8, BS>::run(buf, stw, +1);
    } else {
      fft_any<T, L, MAXN, NCT, BS>(buf, stw, fy, +1);
    }
    const T* bufr = reinterpret_cast<const T*>(buf);
    const T sc = (T)a.scale;
    if (a.vec) {
      for (int e = threadIdx.x * W; e < n * ZT; e += BS * W) {
        const int y = e / ZT, zl = e - y * ZT;
        const int z = z0 + zl;
        if (z >= nz) continue;
        Vec16<T> v = vld(bufr + e);
#pragma unroll
        for (int j = 0; j < W; ++j) v.v[j] *= sc;
        st.vstore(b, x, y, z, v);
      }
    } else {
      for (int i = threadIdx.x; i < n * ZT; i += BS) {
        const int y = i / ZT, zl = i - y * ZT;
        const int z = z0 + zl;
        if (z < nz) st(b, x, y, z, bufr[i] * sc);
      }
    }
  }
}

// ----------------------------------------------------------------------------
// plain load / store functors
// ----------------------------------------------------------------------------
template <class T> struct LoadPtr {
  static constexpr const char* NAME = "LoadPtr";
  static constexpr int BATCH_FAST = 0;
  static int nin(int nb) { return nb; }
  const T* p[MAXB];
  int ny, nz;
  __device__ __forceinline__ T operator()(int b, int x, int y, int z) const {
    return p[b][((size_t)x * ny + y) * nz + z];
  }
  __device__ __forceinline__ Vec16<T> vload(int b, int x, int y, int z) const {
    return vld(p[b] + ((size_t)x * ny + y) * nz + z);
  }
  bool aligned(int nb) const {
    for (int b = 0; b < nb; ++b)
      if (!al16(p[b])) return false;
    return true;
  }
};

template <class T> struct StorePtr {
  static constexpr const char* NAME = "StorePtr";
  static constexpr int NOUT = 1;
  T* p[MAXB];
  int ny, nz;
  __device__ __forceinline__ void operator()(int b, int x, int y, int z, T v) const {
    p[b][((size_t)x * ny + y) * nz + z] = v;
  }
  __device__ __forceinline__ void vstore(int b, int x, int y, int z, const Vec16<T>& v) const {
    vst(p[b] + ((size_t)x * ny + y) * nz + z, v);
  }
  // z and z + 1 (z even, nz even, aligned buffers)
  __device__ __forceinline__ void store2(int b, int x, int y, int z, T v0, T v1) const {
    *reinterpret_cast<c2<T>*>(p[b] + ((size_t)x * ny + y) * nz + z) = mkc<T>(v0, v1);
  }
  bool aligned(int nb) const {
    for (int b = 0; b < nb; ++b)
      if (!al16(p[b])) return false;
    return true;
  }
};
