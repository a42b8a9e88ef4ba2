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
    if (R == 4) fft_pass<T, 4, LPB, MAXN>(buf, tw, f.n, Ns, sign);
    else if (R == 2) fft_pass<T, 2, LPB, MAXN>(buf, tw, f.n, Ns, sign);
    else if (R == 3) fft_pass<T, 3, LPB, MAXN>(buf, tw, f.n, Ns, sign);
    else fft_pass<T, 5, LPB, MAXN>(buf, tw, f.n, Ns, sign);
    Ns *= R;
  }
}

struct PtrBatch {
  void* p[MAXB];
};

// ----------------------------------------------------------------------------
// y-direction R2C.  grid = (ceil(nz/ZT), rows, batch); row x of field b is
// read through ld(b, x, y, z); output (rows, ncy_out, nz) complex.
// ----------------------------------------------------------------------------
// Launch geometry of the tiled kernels: a 1-D grid of ntz * rows * nb
// workgroups, renumbered XCD-contiguously; with batch_fast the fields of one
// (row, z tile) are adjacent (they share input tiles through L2), otherwise
// the z tiles of one row are (they share HBM rows).
struct TileGrid {
  int ntz, rows, nb, batch_fast;
  __device__ __forceinline__ void decode(int& zt, int& row, int& b) const {
    const unsigned total = (unsigned)ntz * rows * nb;
    unsigned l = xcd_logical(blockIdx.x, total);
    if (batch_fast) {
      b = l % nb; l /= nb;
      zt = l % ntz; row = l / ntz;
    } else {
      zt = l % ntz; l /= ntz;
      row = l % rows; b = l / rows;
    }
  }
};

template <class T, class Loader, int MAXN>
__global__ void __launch_bounds__(FFT_BS)
k_y_r2c(Loader ld, Fft1D fy, const c2<T>* __restrict__ tw, int nz, int ncy_out, PtrBatch outb,
        TileGrid tg) {
  constexpr int L = Lines<T>::L;
  constexpr int ZT = 2 * L;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int n = fy.n;
  c2<T>* buf = reinterpret_cast<c2<T>*>(smem_raw);
  c2<T>* stw = buf + n * L;
  int zt, x, b;
  tg.decode(zt, x, b);
  const int z0 = zt * ZT;
  for (int i = threadIdx.x; i < n; i += FFT_BS) stw[i] = tw[i];
  T* bufr = reinterpret_cast<T*>(buf);
  for (int i = threadIdx.x; i < n * ZT; i += FFT_BS) {
    const int y = i / ZT, zl = i - y * ZT;
    const int z = z0 + zl;
    bufr[i] = (z < nz) ? ld(b, x, y, z) : T(0);
  }
  __syncthreads();
  fft_lines<T, L, MAXN>(buf, stw, fy, -1);
  c2<T>* out = reinterpret_cast<c2<T>*>(outb.p[b]);
  for (int i = threadIdx.x; i < ncy_out * ZT; i += FFT_BS) {
    const int k = i / ZT, zl = i - k * ZT;
    const int z = z0 + zl;
    if (z >= nz) continue;
    const int p = zl >> 1;
    const c2<T> Zk = buf[k * L + p];
    const c2<T> Zn = buf[((n - k) % n) * L + p];
    c2<T> r;
    if ((zl & 1) == 0) r = mkc<T>(T(0.5) * (Zk.x + Zn.x), T(0.5) * (Zk.y - Zn.y));
    else r = mkc<T>(T(0.5) * (Zk.y + Zn.y), T(-0.5) * (Zk.x - Zn.x));
    out[((size_t)x * ncy_out + k) * nz + z] = r;
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
};

// value of op `op` at input mode k (element i of line p) of a forward spectrum
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

template <class T, int MAXN>
__global__ void __launch_bounds__(FFT_BS)
k_x(XArgs a, Fft1D fin, Fft1D fout, const c2<T>* __restrict__ tw_in,
    const c2<T>* __restrict__ tw_out, TileGrid tg) {
  constexpr int L = Lines<T>::L;
  constexpr int MAXE = MAXN * L / FFT_BS;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int nin = fin.n;
  const int nbuf = imax(nin, fout.n);
  c2<T>* buf = reinterpret_cast<c2<T>*>(smem_raw);
  c2<T>* stw_in = buf + nbuf * L;
  c2<T>* stw_out = stw_in + nin;
  int zt, ky, b;
  tg.decode(zt, ky, b);
  const int z0 = zt * L;
  const int nz = a.nz;
  for (int i = threadIdx.x; i < nin; i += FFT_BS) stw_in[i] = tw_in[i];
  for (int i = threadIdx.x; i < fout.n; i += FFT_BS) stw_out[i] = tw_out[i];
  const c2<T>* in = reinterpret_cast<const c2<T>*>(a.in[b]);
  for (int i = threadIdx.x; i < nin * L; i += FFT_BS) {
    const int k = i / L, p = i - k * L;
    const int z = z0 + p;
    buf[i] = (z < nz) ? in[((size_t)k * a.ncy_in + ky) * nz + z] : mkc<T>(0, 0);
  }
  __syncthreads();
  if (!a.in_spectral) fft_lines<T, L, MAXN>(buf, stw_in, fin, -1);
  const int kyv = (2 * ky == a.ny_full) ? 0 : ky;
  const int nEl = nin * L;
  if (a.nops == 1 && a.op[0].map == XMAP_ID) {
    // one output: apply the op in place, no register copy of the spectrum
    const XOp op = a.op[0];
    if (ky >= op.ncy) return;
    for (int i = threadIdx.x; i < nEl; i += FFT_BS) buf[i] = x_apply<T>(op, buf[i], i / L, nin, kyv);
    __syncthreads();
    if (op.inverse) fft_lines<T, L, MAXN>(buf, stw_in, fin, +1);
    c2<T>* dst = reinterpret_cast<c2<T>*>(a.dst[b][0]);
    for (int i = threadIdx.x; i < nin * L; i += FFT_BS) {
      const int k = i / L, p = i - k * L;
      const int z = z0 + p;
      if (z < nz) dst[((size_t)k * op.ncy + ky) * nz + z] = buf[i];
    }
    return;
  }
  c2<T> spec[MAXE];
#pragma unroll
  for (int q = 0; q < MAXE; ++q) {
    const int i = threadIdx.x + q * FFT_BS;
    spec[q] = (i < nEl) ? buf[i] : mkc<T>(0, 0);
  }
  for (int o = 0; o < a.nops; ++o) {
    const XOp op = a.op[o];
    if (ky >= op.ncy) continue;
    const int nout = op.nout;
    __syncthreads();
    if (op.map != XMAP_ID) {
      for (int i = threadIdx.x; i < nout * L; i += FFT_BS) buf[i] = mkc<T>(0, 0);
      __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < MAXE; ++q) {
      const int i = threadIdx.x + q * FFT_BS;
      if (i < nEl) {
        const int k = i / L, p = i - k * L;
        const c2<T> v = x_apply<T>(op, spec[q], k, nin, kyv);
        if (op.map == XMAP_ID) {
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
      if (op.map == XMAP_ID) fft_lines<T, L, MAXN>(buf, stw_in, fin, +1);
      else fft_lines<T, L, MAXN>(buf, stw_out, fout, +1);
    }
    c2<T>* dst = reinterpret_cast<c2<T>*>(a.dst[b][o]);
    for (int i = threadIdx.x; i < nout * L; i += FFT_BS) {
      const int k = i / L, p = i - k * L;
      const int z = z0 + p;
      if (z < nz) dst[((size_t)k * op.ncy + ky) * nz + z] = buf[i];
    }
  }
}

// ----------------------------------------------------------------------------
// y-direction C2R.  grid = (ceil(nz/ZT), rows, batch).  Output b is the sum
// of nsum pruned half spectra (rows, ncy_in, nz); values go to
// st(b, x, y, z, value*scale).
// ----------------------------------------------------------------------------
struct YC2RArgs {
  const void* in[MAXB][3];
  int nsum;
  int ncy_in;
  int nz;
  double scale;
};

template <class T, class Storer, int MAXN>
__global__ void __launch_bounds__(FFT_BS)
k_y_c2r(Storer st, YC2RArgs a, Fft1D fy, const c2<T>* __restrict__ tw, TileGrid tg) {
  constexpr int L = Lines<T>::L;
  constexpr int ZT = 2 * L;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int n = fy.n, half = n / 2;
  c2<T>* buf = reinterpret_cast<c2<T>*>(smem_raw);
  c2<T>* stw = buf + n * L;
  int zt, x, b;
  tg.decode(zt, x, b);
  const int z0 = zt * ZT;
  const int nz = a.nz;
  for (int i = threadIdx.x; i < n; i += FFT_BS) stw[i] = tw[i];
  for (int i = threadIdx.x; i < n * L; i += FFT_BS) buf[i] = mkc<T>(0, 0);
  __syncthreads();
  const int kmax = imin(a.ncy_in, half + 1);
  for (int i = threadIdx.x; i < kmax * L; i += FFT_BS) {
    const int k = i / L, p = i - k * L;
    const int z = z0 + 2 * p;
    c2<T> X1 = mkc<T>(0, 0), X2 = mkc<T>(0, 0);
    for (int s = 0; s < a.nsum; ++s) {
      const c2<T>* in = reinterpret_cast<const c2<T>*>(a.in[b][s]);
      const size_t base = ((size_t)x * a.ncy_in + k) * nz;
      if (z < nz) X1 = cadd<T>(X1, in[base + z]);
      if (z + 1 < nz) X2 = cadd<T>(X2, in[base + z + 1]);
    }
    if (k == 0 || k == half) { X1.y = T(0); X2.y = T(0); }
    buf[k * L + p] = mkc<T>(X1.x - X2.y, X1.y + X2.x);
    if (k > 0 && k < half) buf[(n - k) * L + p] = mkc<T>(X1.x + X2.y, X2.x - X1.y);
  }
  __syncthreads();
  fft_lines<T, L, MAXN>(buf, stw, fy, +1);
  const T* bufr = reinterpret_cast<const T*>(buf);
  const T sc = (T)a.scale;
  for (int i = threadIdx.x; i < n * ZT; i += FFT_BS) {
    const int y = i / ZT, zl = i - y * ZT;
    const int z = z0 + zl;
    if (z < nz) st(b, x, y, z, bufr[i] * sc);
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
};

template <class T> struct StorePtr {
  static constexpr const char* NAME = "StorePtr";
  static constexpr int NOUT = 1;
  T* p[MAXB];
  int ny, nz;
  __device__ __forceinline__ void operator()(int b, int x, int y, int z, T v) const {
    p[b][((size_t)x * ny + y) * nz + z] = v;
  }
};
