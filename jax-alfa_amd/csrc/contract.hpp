// Test-filter inverse transforms fused with what the dynamic procedures do
// with the filtered fields: the pointwise contractions (LM/MM, T_up/T_dn) of
// level 1, and for LASDD (optSgs = 1) the plane sums of both levels that give
// beta (the sums kernels at the end of this file); LAD (optSgs = 3) has
// beta = 1 and one level.
//
// Unfused, the y C2R of the test filter writes every filtered field to HBM
// (21 momentum fields, DynamicSGS_LASDD_SM.py:184-233; 10 scalar fields,
// DynamicSGS_ScalarLASDD_SM.py:179-224) and a pointwise kernel reads them all
// back to form LM, MM, |S_hat| (:235, :327-355) or T_up, T_dn
// (DynamicSGS_ScalarLASDD_SM.py:241-312).  Here one workgroup inverts every
// field of one tile in turn and folds each into per-point state held in
// registers, so only the fields that later stages read reach HBM.
//
// Tile: BS = 256 threads (others at lengths 48 and 80, ctr_bs), L lines (2L z levels, two per complex line) of one
// x row, with L = BS R / N so that the last radix pass (radix R, stride N/R)
// is one butterfly per thread and field: thread (j, p) ends every field
// holding rows y = j + r N/R (r < R) of the level pair (z0 + 2p, z0 + 2p + 1),
// i.e. the same 2R points for every field.  Fields go through the passes in
// pairs, interleaved in one LDS buffer of 2L lines (line p of field h at
// 2p + h): the radix-8 passes then have a butterfly for every thread, each
// barrier serves two fields, and the pair's staging writes and last-pass
// reads are single 16-byte accesses on contiguous rows.  The
// passes are those of k_y_c2r (band first pass, compile-time radix plan) on
// the same values, so each field's float values are the ones the unfused
// kernel would store.
//
// Momentum (LAD): the six S_hat_ij are inverted first (their squares give
// |S_hat|), then u_hat_i (kept; stored with the outputs at the end of the
// tile: the scalar procedure reads them), then per two components ij:
// S_hat_ij again, (u_i u_j)_hat and (|S| S_ij)_hat, which complete L_ij and
// M_ij; LM accumulates in double and MM (a sum of squares) in float.
// Inverting S_hat_ij twice (27 transforms, 14 pair steps, instead of 21)
// keeps six fields per point out of the registers.
// Scalar: TH_hat, then per direction k: (u_k TH)_hat, (dTH/dx_k)_hat and
// (|S| dTH/dx_k)_hat (5 pair steps), with u_hat_k and |S_hat| read from HBM.  With
// D_k = sdh_k - TFR^2 beta |S_hat| dh_k, the reference's
//   T_up = L^2 a - L^2 TFR^2 beta b,  T_dn = L^4 c - 2 L^4 TFR^2 beta d + L^4 TFR^4 beta^2 e
// (a..e the sums of DynamicSGS_ScalarLASDD_SM.py:241-302 at level 1) are
// T_up = L^2 sum_k LT_k D_k and T_dn = L^4 sum_k D_k^2, which this kernel
// accumulates directly (T_up in double, the sum of squares in float).
#pragma once
#include <type_traits>
#include "fft.hpp"
#include "loaders.hpp"
#include "plan.hpp"

// Radix of the fused kernels' last pass at length N with BS threads: that of
// the compile-time plan, except 64 (8 x 8), which the fused kernels factor
// as 8 x 4 x 2, and 80 (8 x 2 x 5) at 320 threads, factored 8 x 5 x 2, so
// that a thread holds 4 points per field instead of 16 or 10.
__host__ __device__ constexpr int ctr_last_radix(int n, int bs) {
  return n == 64 || (n == 80 && bs == 320) ? 2 : last_radix(n);
}
// threads per workgroup of the fused kernels at length N: 256, except
//  - 320 (8 lines of the radix-2 last pass) for the scalar kernels at 80,
//    whose state at 10 points per thread needs 250+ VGPRs (1 wave per SIMD;
//    119-140 at 4).  Measured on the 64-member ensemble: the scalar
//    contraction and sums at 80 run 15-23% faster at 320; the momentum ones
//    34-47% slower (twice the tiles, half the threads idle in the first
//    pass);
//  - 128 for the momentum kernels at 80 and all of them at 48 (8 lines, 16
//    levels per tile instead of 16 lines, 32 levels): grids of 48 or 80
//    levels then fill their tiles exactly instead of computing 17-25% of
//    padding levels.
__host__ __device__ constexpr int ctr_bs(int n, bool scalar) {
  return n == 48 ? 128 : n == 80 ? (scalar ? 320 : 128) : 256;
}

// The passes between the first (radix 8, stride 1) and the last (radix RL):
// radix_of(M / RL) at each step, which for every length but 64 is the
// compile-time plan's own sequence (the same passes as k_y_c2r's).
template <class T, int LPB, int N, int NS, int M, int RL, int BS, int LAY = 0>
struct CtrMid {
  static __device__ __forceinline__ void run(c2<T>* buf, const c2<T>* tw, int sign) {
    if constexpr (M != RL) {
      constexpr int R = radix_of(M / RL);
      fft_pass_ct<T, R, LPB, N, NS, BS, LAY>(buf, tw, sign);
      CtrMid<T, LPB, N, NS * R, M / R, RL, BS, NS == 1 ? LAY : 0>::run(buf, tw, sign);
    }
  }
};

// geometry of the fused kernels at compile-time length N: L lines per field,
// LL = 2L lines per pair
template <int N, int BS> struct CtrGeom {
  static constexpr int R = N > 0 ? ctr_last_radix(N, BS) : 1;
  static constexpr int NS = N > 0 ? N / R : 1;
  static constexpr int L = BS / NS;
  static constexpr int LL = 2 * L;
  static constexpr bool OK = N > 0 && (BS % NS == 0) && R <= 5 && radix_of(N) == 8;
  // first-pass layout: padded at rows of 8 lines (64 bytes), where the first
  // pass's writes would conflict and a middle pass reads them; with 16 or
  // more lines a 16-lane group of its writes covers one row
  static constexpr int LAY = (N > 0 && N / 8 != R && LL <= 8) ? 2 : 0;
  static constexpr int ROWS = N > 0 ? lay_rows(N, LAY) : 1;   // rows of the pair buffer
};

// An empty asm that the compiler must assume changes v: values derived from
// a float held across the field loops (its double, products with it) are then
// formed where they are used instead of being hoisted out of the loop and
// held at twice the registers.
template <class T> __device__ __forceinline__ void ctr_keep(T& v) { asm volatile("" : "+v"(v)); }

// spectrum tile of one field, (k < kmax, line p < L) as the two levels
// (z, z + 1); a null field loads nothing
template <class T, int L, int BS, int IN_IT>
__device__ __forceinline__ void ctr_load(const c2<T>* __restrict__ in, int ncy_in, int nz, int kmax,
                                         int x, int z0, c2<T> (&X1)[IN_IT], c2<T> (&X2)[IN_IT]) {
#pragma unroll
  for (int q = 0; q < IN_IT; ++q) {
    const int e = threadIdx.x + q * BS;
    const int k = e / L, p = e % L, z = z0 + 2 * p;
    X1[q] = mkc<T>(0, 0);
    X2[q] = mkc<T>(0, 0);
    if (in != nullptr && k < kmax && z < nz)
      cld2<T>(in + ((size_t)x * ncy_in + k) * nz + z, X1[q], X2[q]);
  }
}

// the loaded pair into the LL-line buffer, packed as k_y_c2r packs a field
// (imaginary parts of ky = 0 and the Nyquist column dropped).  The two fields
// are interleaved, field h of line p at line 2p + h, so each (k, p) of the
// pair is one 16-byte LDS store: the lanes of a wave write contiguous rows.
template <class T, int N, int L, int LL, int BS, int IN_IT>
__device__ __forceinline__ void ctr_stage2(c2<T>* buf, const c2<T> (&A1)[IN_IT],
                                           const c2<T> (&A2)[IN_IT], const c2<T> (&B1)[IN_IT],
                                           const c2<T> (&B2)[IN_IT], int kmax) {
  constexpr int half = N / 2;
  auto pk = [](c2<T> a, c2<T> b, int k, c2<T>& lo, c2<T>& hi) {
    if (k == 0 || k == half) { a.y = T(0); b.y = T(0); }
    lo = mkc<T>(a.x - b.y, a.y + b.x);
    hi = mkc<T>(a.x + b.y, b.x - a.y);
  };
#pragma unroll
  for (int q = 0; q < IN_IT; ++q) {
    const int e = threadIdx.x + q * BS;
    const int k = e / L, p = e % L;
    if (k < kmax) {
      c2<T> al, ah, bl, bh;
      pk(A1[q], A2[q], k, al, ah);
      pk(B1[q], B2[q], k, bl, bh);
      cst2<T>(buf + k * LL + 2 * p, al, bl);
      if (k > 0 && k < half) cst2<T>(buf + (N - k) * LL + 2 * p, ah, bh);
    }
  }
}

// every pass but the last over the LL lines, then the last (radix R) from
// LDS into registers: o[h][r][0], o[h][r][1] = levels z, z + 1 of row
// j + r NS of field h of the pair (line 2p + h), times sc
template <class T, int N, int BS, bool BAND>
__device__ __forceinline__ void ctr_invert2(c2<T>* buf, const c2<T>* stw, int kmax, T sc,
                                            T (&o)[2][CtrGeom<N, BS>::R][2]) {
  using G = CtrGeom<N, BS>;
  constexpr int R = G::R, NS = G::NS, L = G::L, LL = G::LL, LAY = G::LAY;
  if constexpr (BAND) band_first_pass_inv<T, LL, N, BS, LAY>(buf, kmax);
  else fft_pass_ct<T, 8, LL, N, 1, BS, LAY>(buf, stw, +1);
  CtrMid<T, LL, N, 8, N / 8, R, BS, LAY>::run(buf, stw, +1);
  const int j = threadIdx.x / L, p = threadIdx.x % L;
  // line p of both fields (lines 2p, 2p + 1): one 16-byte read per r
  c2<T> v[2][R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    c2<T> e0, e1;
    cld2<T>(buf + (j + r * NS) * LL + 2 * p, e0, e1);
    if (r > 0) {
      c2<T> w = stw[j * r];
      w.y = -w.y;
      e0 = cmul<T>(e0, w);
      e1 = cmul<T>(e1, w);
    }
    v[0][r] = e0;
    v[1][r] = e1;
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    Dft<T, R>::run(v[h], +1);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      o[h][r][0] = v[h][r].x * sc;
      o[h][r][1] = v[h][r].y * sc;
    }
  }
}

// The pair-step driver of both kernels.  A kernel walks its tiles with
// first() / next() and, per tile, calls step(s, use) for its NP pair steps in
// order (from its own phase loops, so each phase keeps only its own state
// live).  Step s stages the loaded spectra of pair s (in[2s], in[2s+1], the
// second possibly null), issues the loads of pair s + 1 (or of the next
// tile's pair 0), inverts both fields and hands their values to use(o).
template <class T, int N, int BS, bool BAND, int NP>
struct CtrDriver {
  using G = CtrGeom<N, BS>;
  static constexpr int R = G::R, NS = G::NS, L = G::L, LL = G::LL, ZT = 2 * L;
  static constexpr int KM = BAND ? N / 4 : N / 2 + 1;
  static constexpr int IN_IT = (KM * L + BS - 1) / BS;
  const c2<T>* const* in;
  const TileGrid& tg;
  c2<T>* buf;
  c2<T>* stw;
  int ncy_in, nz, kmax;
  T sc;
  int zt = 0, x = 0, it = 0, zt2 = 0, x2 = 0;
  bool more = false;
  c2<T> A1[IN_IT], A2[IN_IT], B1[IN_IT], B2[IN_IT];

  __device__ __forceinline__ CtrDriver(const c2<T>* const* in_, int ncy_in_, int nz_, double scale,
                                       const c2<T>* __restrict__ tw, const TileGrid& tg_)
      : in(in_), tg(tg_), ncy_in(ncy_in_), nz(nz_) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    buf = reinterpret_cast<c2<T>*>(smem_raw);
    stw = buf + G::ROWS * LL;
    kmax = imin(ncy_in, N / 2 + 1);
    sc = (T)scale;
    for (int i = threadIdx.x; i < N; i += BS) stw[i] = tw[i];
  }
  __device__ __forceinline__ void load_pair(int s, int xr, int ztr) {
    ctr_load<T, L, BS, IN_IT>(in[2 * s], ncy_in, nz, kmax, xr, ztr * ZT, A1, A2);
    ctr_load<T, L, BS, IN_IT>(in[2 * s + 1], ncy_in, nz, kmax, xr, ztr * ZT, B1, B2);
  }
  // the workgroup's first tile (false: none)
  __device__ __forceinline__ bool first() {
    int b;
    if (!tg.tile(0, zt, x, b)) return false;
    load_pair(0, x, zt);
    return true;
  }
  // after a tile's last step: move to the next tile (false: none)
  __device__ __forceinline__ bool next() {
    if (!more) return false;
    zt = zt2;
    x = x2;
    ++it;
    return true;
  }
  template <class Use> __device__ __forceinline__ void step(int s, Use&& use) {
    __syncthreads();
    if constexpr (!BAND) {
      for (int i = threadIdx.x; i < N * LL; i += BS) {
        const int k = i / LL;
        if (k >= kmax && k <= N - kmax) buf[i] = mkc<T>(0, 0);
      }
    }
    ctr_stage2<T, N, L, LL, BS, IN_IT>(buf, A1, A2, B1, B2, kmax);
    if (s + 1 < NP) {
      load_pair(s + 1, x, zt);
    } else {
      int b;
      more = tg.tile(it + 1, zt2, x2, b);
      if (more) load_pair(0, x2, zt2);
    }
    __syncthreads();
    T o[2][R][2];
    ctr_invert2<T, N, BS, BAND>(buf, stw, kmax, sc, o);
    use(o);
  }
};

// ---------------------------------------------------------------------------
// momentum (LAD)
// ---------------------------------------------------------------------------
constexpr int LAD_NP = 14;
// field of slot h of pair step s, as an index of the 21 LoadLasdd fields
// (0..2 u_hat, 3..8 products uu vv ww uv uw vw, 9..14 S_ij, 15..20 |S| S_ij),
// -1 for the empty slot
__host__ __device__ constexpr int lad_field(int s, int h) {
  return s < 3 ? 9 + 2 * s + h
       : s == 3 ? h
       : s == 4 ? (h == 0 ? 2 : -1)
       : ((s - 5) % 3 == 0 ? 9 : (s - 5) % 3 == 1 ? 3 : 15) + 2 * ((s - 5) / 3) + h;
}

template <class T> struct LadArgs {
  const c2<T>* in[2 * LAD_NP];   // x-filtered level-1 spectra, pair by pair
  const double* beta;            // (nz)
  T* uh[3];                      // u_hat, v_hat, w_hat
  T* LM; T* MM; T* Shat;
  int ncy_in, ny, nz;
  double scale;                  // 1 / (nx ny)
  double T1, T2;                 // 2 L^2, 2 (TFR L)^2
};

template <class T, int N, int BS, bool BAND>
__device__ __forceinline__ void k_c2r_lad(const LadArgs<T>& a, const c2<T>* __restrict__ tw,
                                          const TileGrid& tg) {
  using D = CtrDriver<T, N, BS, BAND, LAD_NP>;
  constexpr int R = D::R, NS = D::NS, L = D::L;
  const int nz = a.nz;
  const int j = threadIdx.x / L, p = threadIdx.x % L;
  D d(a.in, a.ncy_in, nz, a.scale, tw, tg);
  if (!d.first()) return;
  do {
    const int z = d.zt * D::ZT + 2 * p;
    const bool zin = z < nz;
    const double bz[2] = {zin ? a.beta[z] : 0.0, zin ? a.beta[z + 1] : 0.0};
    // |S_hat| = sqrt(2 S_ij S_ij) from the six S_hat_ij (summed in double,
    // kept as the float the unfused path stores)
    double sm[R][2];
#pragma unroll
    for (int r = 0; r < R; ++r) { sm[r][0] = 0.0; sm[r][1] = 0.0; }
#pragma unroll 1
    for (int s = 0; s < 3; ++s) {
      // components 2s, 2s + 1; the off-diagonal ones (3, 4, 5) count twice
      const double w0 = s < 2 ? 1.0 : 2.0, w1 = s < 1 ? 1.0 : 2.0;
      d.step(s, [&](const T (&o)[2][R][2]) {
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const double s0 = o[0][r][q], s1 = o[1][r][q];
            sm[r][q] += w0 * (s0 * s0);
            sm[r][q] += w1 * (s1 * s1);
          }
      });
    }
    T smf[R][2];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int q = 0; q < 2; ++q) smf[r][q] = (T)sqrt(2.0 * sm[r][q]);
    // u_hat, v_hat, w_hat, held in LDS after the pair buffer and the
    // twiddles ([3][R][2][BS]: each thread's own values; registers held
    // across the whole tile would cost a wave per SIMD)
    T* uhs = reinterpret_cast<T*>(d.stw + N);
    auto UH = [&](int k, int r, int q) -> T& { return uhs[((k * R + r) * 2 + q) * BS + threadIdx.x]; };
    d.step(3, [&](const T (&o)[2][R][2]) {
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int q = 0; q < 2; ++q) { UH(0, r, q) = o[0][r][q]; UH(1, r, q) = o[1][r][q]; }
    });
    d.step(4, [&](const T (&o)[2][R][2]) {
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int q = 0; q < 2; ++q) UH(2, r, q) = o[0][r][q];
    });
    // per two components (uu vv | ww uv | uw vw): S_hat_ij, (u_i u_j)_hat and
    // (|S| S_ij)_hat complete L_ij and M_ij.  LM (a signed sum) accumulates
    // in double; MM, a sum of squares, in the field type, each term added in
    // double (at most 6 roundings of a sum of positive terms).
    double lm[R][2];
    T mm[R][2];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int q = 0; q < 2; ++q) { lm[r][q] = 0.0; mm[r][q] = T(0); }
#pragma unroll 1
    for (int m = 0; m < 3; ++m) {
      T sh[2][R][2], ph[2][R][2];
      d.step(5 + 3 * m, [&](const T (&o)[2][R][2]) {
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int r = 0; r < R; ++r) { sh[h][r][0] = o[h][r][0]; sh[h][r][1] = o[h][r][1]; }
      });
      d.step(6 + 3 * m, [&](const T (&o)[2][R][2]) {
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int r = 0; r < R; ++r) { ph[h][r][0] = o[h][r][0]; ph[h][r][1] = o[h][r][1]; }
      });
      d.step(7 + 3 * m, [&](const T (&o)[2][R][2]) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          // component n = 2m + h of uu vv ww uv uw vw
          const int n = 2 * m + h;
          const int ia = n < 3 ? n : (n < 5 ? 0 : 1);
          const int ib = n < 3 ? n : (n == 3 ? 1 : 2);
          const double wgt = n < 3 ? 1.0 : 2.0;
#pragma unroll
          for (int r = 0; r < R; ++r)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
              const T ta = UH(ia, r, q), tb = UH(ib, r, q);
              T smv = smf[r][q];
              ctr_keep(smv);
              const double Lij = (double)ph[h][r][q] - (double)ta * (double)tb;
              const double M = a.T1 * (double)o[h][r][q] -
                               a.T2 * bz[q] * (double)smv * (double)sh[h][r][q];
              lm[r][q] += wgt * (Lij * M);
              mm[r][q] = (T)((double)mm[r][q] + wgt * (M * M));
            }
        }
      });
    }
    if (zin) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const size_t i = ((size_t)d.x * a.ny + j + r * NS) * nz + z;
#pragma unroll
        for (int k = 0; k < 3; ++k)
          *reinterpret_cast<c2<T>*>(a.uh[k] + i) = mkc<T>(UH(k, r, 0), UH(k, r, 1));
        *reinterpret_cast<c2<T>*>(a.LM + i) = mkc<T>((T)lm[r][0], (T)lm[r][1]);
        *reinterpret_cast<c2<T>*>(a.MM + i) = mkc<T>(mm[r][0], mm[r][1]);
        *reinterpret_cast<c2<T>*>(a.Shat + i) = mkc<T>(smf[r][0], smf[r][1]);
      }
    }
  } while (d.next());
}

// ---------------------------------------------------------------------------
// scalar (LAD)
// ---------------------------------------------------------------------------
constexpr int SCA_NP = 5;
// field of slot h of pair step s, as an index of the 10 LoadScalar fields
// (0 TH_hat, 1..3 (u_k TH)_hat, 4..6 dTH/dx_k hats, 7..9 |S| dTH/dx_k hats):
// (TH, uTH_0), (dTH_0, SdTH_0), (uTH_1, dTH_1), (SdTH_1, uTH_2), (dTH_2, SdTH_2)
__host__ __device__ constexpr int sca_field(int s, int h) {
  return s == 0 ? (h == 0 ? 0 : 1)
       : s == 1 ? (h == 0 ? 4 : 7)
       : s == 2 ? (h == 0 ? 2 : 5)
       : s == 3 ? (h == 0 ? 8 : 3)
       : (h == 0 ? 6 : 9);
}

template <class T> struct ScaArgs {
  const c2<T>* in[2 * SCA_NP];   // x-filtered level-1 spectra, pair by pair
  const double* beta;            // (nz)
  const T* uh[3];                // the momentum procedure's u_hat, v_hat, w_hat
  const T* Sm;                   // and |S_hat|
  T* Tup; T* Tdn;
  int ncy_in, ny, nz;
  double scale;                  // 1 / (nx ny)
  double L2, L4, T2;             // L^2, L^4, TFR^2
};

template <class T, int N, int BS, bool BAND>
__device__ __forceinline__ void k_c2r_sca(const ScaArgs<T>& a, const c2<T>* __restrict__ tw,
                                          const TileGrid& tg) {
  using D = CtrDriver<T, N, BS, BAND, SCA_NP>;
  constexpr int R = D::R, NS = D::NS, L = D::L;
  const int nz = a.nz;
  const int j = threadIdx.x / L, p = threadIdx.x % L;
  D d(a.in, a.ncy_in, nz, a.scale, tw, tg);
  if (!d.first()) return;
  do {
    const int z = d.zt * D::ZT + 2 * p;
    const bool zin = z < nz;
    const double bz[2] = {zin ? a.beta[z] : 0.0, zin ? a.beta[z + 1] : 0.0};
    // the values of a momentum-procedure field at the thread's points
    auto point_load = [&](const T* f, T (&v)[R][2]) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        c2<T> t = mkc<T>(0, 0);
        if (zin) t = *reinterpret_cast<const c2<T>*>(f + ((size_t)d.x * a.ny + j + r * NS) * nz + z);
        v[r][0] = t.x;
        v[r][1] = t.y;
      }
    };
    // TH_hat and |S_hat| at the thread's points, held in LDS after the pair
    // buffer and the twiddles ([2][R][2][BS])
    T* sts = reinterpret_cast<T*>(d.stw + N);
    auto TH_ = [&](int r, int q) -> T& { return sts[((0 * R + r) * 2 + q) * BS + threadIdx.x]; };
    auto SM_ = [&](int r, int q) -> T& { return sts[((1 * R + r) * 2 + q) * BS + threadIdx.x]; };
    T um[R][2], uth[R][2], dh[R][2];
    {
      T smh[R][2];
      point_load(a.Sm, smh);
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int q = 0; q < 2; ++q) SM_(r, q) = smh[r][q];
    }
    point_load(a.uh[0], um);
    // T_up (a signed sum) accumulates in double, T_dn (a sum of squares) in
    // the field type, each term added in double
    double up[R][2];
    T dn[R][2];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int q = 0; q < 2; ++q) { up[r][q] = 0.0; dn[r][q] = T(0); }
    // direction k: T_up += LT D, T_dn += D^2 from (u_k TH)_hat = ut, the
    // momentum u_hat_k = u, dTH/dx_k hat = dd and |S| dTH/dx_k hat = sd
    auto fold = [&](const T (&ut)[R][2], const T (&u)[R][2], const T (&dd)[R][2],
                    const T (&sd)[R][2]) {
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          // each operand pinned here, so the conversions to double are
          // made point by point instead of all at once
          T thv = TH_(r, q), smv = SM_(r, q), utv = ut[r][q], uv = u[r][q], sdv = sd[r][q],
            ddv = dd[r][q];
          ctr_keep(thv);
          ctr_keep(smv);
          ctr_keep(utv);
          ctr_keep(uv);
          ctr_keep(sdv);
          ctr_keep(ddv);
          const double LT = (double)utv - (double)uv * (double)thv;
          const double Dk = (double)sdv - a.T2 * bz[q] * (double)smv * (double)ddv;
          up[r][q] += LT * Dk;
          dn[r][q] = (T)((double)dn[r][q] + Dk * Dk);
        }
    };
    d.step(0, [&](const T (&o)[2][R][2]) {     // TH_hat, (u TH)_hat
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int q = 0; q < 2; ++q) { TH_(r, q) = o[0][r][q]; uth[r][q] = o[1][r][q]; }
    });
    d.step(1, [&](const T (&o)[2][R][2]) {     // dTH/dx, |S| dTH/dx: direction x
      fold(uth, um, o[0], o[1]);
    });
    point_load(a.uh[1], um);
    d.step(2, [&](const T (&o)[2][R][2]) {     // (v TH)_hat, dTH/dy
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int q = 0; q < 2; ++q) { uth[r][q] = o[0][r][q]; dh[r][q] = o[1][r][q]; }
    });
    d.step(3, [&](const T (&o)[2][R][2]) {     // |S| dTH/dy: direction y; (w TH)_hat
      fold(uth, um, dh, o[0]);
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int q = 0; q < 2; ++q) uth[r][q] = o[1][r][q];
    });
    point_load(a.uh[2], um);
    d.step(4, [&](const T (&o)[2][R][2]) {     // dTH/dz, |S| dTH/dz: direction z
      fold(uth, um, o[0], o[1]);
    });
    if (zin) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const size_t i = ((size_t)d.x * a.ny + j + r * NS) * nz + z;
        *reinterpret_cast<c2<T>*>(a.Tup + i) =
            mkc<T>((T)(a.L2 * up[r][0]), (T)(a.L2 * up[r][1]));
        *reinterpret_cast<c2<T>*>(a.Tdn + i) =
            mkc<T>((T)(a.L4 * (double)dn[r][0]), (T)(a.L4 * (double)dn[r][1]));
      }
    }
  } while (d.next());
}

// ---------------------------------------------------------------------------
// plane sums of the LASDD procedures (optSgs = 1), one test-filter level per
// launch: the momentum a1..e2 of DynamicSGS_LASDD_SM.py:266-312 and the scalar
// a2..e4 of DynamicSGS_ScalarLASDD_SM.py:241-302, which the beta quintics
// need before any contraction.  The kernels invert the level's filtered
// spectra as the contraction kernels do and add each point's five terms of
// the level into per-thread sums (over the thread's R rows, per level of its
// z pair); the workgroup reduces those over the tile's rows (a fixed xor
// shuffle tree inside each wave, then the waves in order) and writes one
// partial per (x row, term, z).  k_plane_final adds the partials in x-row
// order, so the sums are reproducible run to run.
// ---------------------------------------------------------------------------
template <class T, int N, int BS>
__device__ __forceinline__ void ctr_tile_sums(c2<T>* buf, const double (&acc)[5][2],
                                              double* __restrict__ part, int x, int z0, int nz,
                                              int nt, int tbase, int tstride, const double (&c)[5]) {
  constexpr int L = CtrGeom<N, BS>::L, NW = BS / 64;
  double v[5][2];
#pragma unroll
  for (int t = 0; t < 5; ++t)
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      double s = acc[t][q];
      // lanes j*L + p of a wave hold rows j of level pair p: add over j
#pragma unroll
      for (int off = L; off < 64; off <<= 1) s += __shfl_xor(s, off, 64);
      v[t][q] = s;
    }
  // the transform buffer is free once every lane has left the last pass
  double* red = reinterpret_cast<double*>(buf);   // [NW][5][2][L]
  __syncthreads();
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane < L) {
#pragma unroll
    for (int t = 0; t < 5; ++t)
#pragma unroll
      for (int q = 0; q < 2; ++q) red[((wave * 5 + t) * 2 + q) * L + lane] = v[t][q];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 5 * 2 * L; i += BS) {
    const int t = i / (2 * L), q = (i / L) % 2, pl = i % L;
    double sum = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) sum += red[((w * 5 + t) * 2 + q) * L + pl];
    const int z = z0 + 2 * pl + q;
    if (z < nz) part[((size_t)x * nt + tbase + t * tstride) * nz + z] = c[t] * sum;
  }
}

template <class T> struct LasddSumArgs {
  const c2<T>* in[2 * LAD_NP];   // the level's x-filtered spectra, pair by pair
  T* uh[3];                      // level 2: u_hat, v_hat, w_hat out (else null)
  T* Shat;                       // level 2: |S_hat| out (else null)
  double* part;                  // (nx, 10, nz) partial sums
  int ncy_in, ny, nz, tbase, tstride;
  double scale;                  // 1 / (nx ny)
  // factors of the level's terms <L:|S|S>, |S^| <L:S^>, <|S|S:|S|S>,
  // |S^|^2 <S^:S^>, |S^| <S^:|S|S>
  double c[5];
};

template <class T, int N, int BS, bool BAND>
__device__ __forceinline__ void k_c2r_lasdd_sums(const LasddSumArgs<T>& a,
                                                 const c2<T>* __restrict__ tw, const TileGrid& tg) {
  using D = CtrDriver<T, N, BS, BAND, LAD_NP>;
  constexpr int R = D::R, NS = D::NS, L = D::L;
  const int nz = a.nz;
  const int j = threadIdx.x / L, p = threadIdx.x % L;
  D d(a.in, a.ncy_in, nz, a.scale, tw, tg);
  if (!d.first()) return;
  do {
    const int z = d.zt * D::ZT + 2 * p;
    const bool zin = z < nz;
    double acc[5][2];
#pragma unroll
    for (int t = 0; t < 5; ++t) { acc[t][0] = 0.0; acc[t][1] = 0.0; }
    // S_ij S_ij (summed in double) and |S_hat|
    double sm[R][2];
#pragma unroll
    for (int r = 0; r < R; ++r) { sm[r][0] = 0.0; sm[r][1] = 0.0; }
#pragma unroll 1
    for (int s = 0; s < 3; ++s) {
      const double w0 = s < 2 ? 1.0 : 2.0, w1 = s < 1 ? 1.0 : 2.0;
      d.step(s, [&](const T (&o)[2][R][2]) {
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const double s0 = o[0][r][q], s1 = o[1][r][q];
            sm[r][q] += w0 * (s0 * s0);
            sm[r][q] += w1 * (s1 * s1);
          }
      });
    }
    T smf[R][2];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        smf[r][q] = (T)sqrt(2.0 * sm[r][q]);
        // |S^|^2 <S^:S^> = 2 <S^:S^>^2
        acc[3][q] += 2.0 * (sm[r][q] * sm[r][q]);
      }
    // u_hat, v_hat, w_hat, held in LDS after the pair buffer and the twiddles
    T* uhs = reinterpret_cast<T*>(d.stw + N);
    auto UH = [&](int k, int r, int q) -> T& { return uhs[((k * R + r) * 2 + q) * BS + threadIdx.x]; };
    d.step(3, [&](const T (&o)[2][R][2]) {
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int q = 0; q < 2; ++q) { UH(0, r, q) = o[0][r][q]; UH(1, r, q) = o[1][r][q]; }
    });
    d.step(4, [&](const T (&o)[2][R][2]) {
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int q = 0; q < 2; ++q) UH(2, r, q) = o[0][r][q];
    });
#pragma unroll 1
    for (int m = 0; m < 3; ++m) {
      T sh[2][R][2], ph[2][R][2];
      d.step(5 + 3 * m, [&](const T (&o)[2][R][2]) {
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int r = 0; r < R; ++r) { sh[h][r][0] = o[h][r][0]; sh[h][r][1] = o[h][r][1]; }
      });
      d.step(6 + 3 * m, [&](const T (&o)[2][R][2]) {
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int r = 0; r < R; ++r) { ph[h][r][0] = o[h][r][0]; ph[h][r][1] = o[h][r][1]; }
      });
      d.step(7 + 3 * m, [&](const T (&o)[2][R][2]) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int n = 2 * m + h;
          const int ia = n < 3 ? n : (n < 5 ? 0 : 1);
          const int ib = n < 3 ? n : (n == 3 ? 1 : 2);
          const double wgt = n < 3 ? 1.0 : 2.0;
#pragma unroll
          for (int r = 0; r < R; ++r)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
              T shv = smf[r][q], ta = UH(ia, r, q), tb = UH(ib, r, q), sv = sh[h][r][q],
                pv = ph[h][r][q], ssv = o[h][r][q];
              ctr_keep(shv);
              ctr_keep(sv);
              ctr_keep(pv);
              const double Lij = (double)pv - (double)ta * (double)tb;
              const double S = sv, SS = ssv, Sh = shv;
              acc[0][q] += wgt * (Lij * SS);
              acc[1][q] += wgt * Sh * (Lij * S);
              acc[2][q] += wgt * (SS * SS);
              acc[4][q] += wgt * Sh * (S * SS);
            }
        }
      });
    }
    if (a.uh[0] != nullptr && zin) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const size_t i = ((size_t)d.x * a.ny + j + r * NS) * nz + z;
#pragma unroll
        for (int k = 0; k < 3; ++k)
          *reinterpret_cast<c2<T>*>(a.uh[k] + i) = mkc<T>(UH(k, r, 0), UH(k, r, 1));
        *reinterpret_cast<c2<T>*>(a.Shat + i) = mkc<T>(smf[r][0], smf[r][1]);
      }
    }
    ctr_tile_sums<T, N, BS>(d.buf, acc, a.part, d.x, d.zt * D::ZT, nz, 10, a.tbase, a.tstride, a.c);
  } while (d.next());
}

template <class T> struct ScaSumArgs {
  const c2<T>* in[2 * SCA_NP];   // the level's x-filtered scalar spectra, pair by pair
  const T* uh[3];                // the momentum u_hat, v_hat, w_hat of the level
  const T* Sm;                   // and its |S_hat|
  double* part;                  // (nx, 10, nz) partial sums
  int ncy_in, ny, nz, tbase, tstride;
  double scale;                  // 1 / (nx ny)
  // factors of the level's terms <LT.SdT>, <LT.|S^| dT>, <SdT.SdT>,
  // <SdT.|S^| dT>, <(|S^| dT)^2>
  double c[5];
};

template <class T, int N, int BS, bool BAND>
__device__ __forceinline__ void k_c2r_sca_sums(const ScaSumArgs<T>& a,
                                               const c2<T>* __restrict__ tw, const TileGrid& tg) {
  using D = CtrDriver<T, N, BS, BAND, SCA_NP>;
  constexpr int R = D::R, NS = D::NS, L = D::L;
  const int nz = a.nz;
  const int j = threadIdx.x / L, p = threadIdx.x % L;
  D d(a.in, a.ncy_in, nz, a.scale, tw, tg);
  if (!d.first()) return;
  do {
    const int z = d.zt * D::ZT + 2 * p;
    const bool zin = z < nz;
    auto point_load = [&](const T* f, T (&v)[R][2]) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        c2<T> t = mkc<T>(0, 0);
        if (zin) t = *reinterpret_cast<const c2<T>*>(f + ((size_t)d.x * a.ny + j + r * NS) * nz + z);
        v[r][0] = t.x;
        v[r][1] = t.y;
      }
    };
    // TH_hat and |S_hat| in LDS after the pair buffer and the twiddles
    T* sts = reinterpret_cast<T*>(d.stw + N);
    auto TH_ = [&](int r, int q) -> T& { return sts[((0 * R + r) * 2 + q) * BS + threadIdx.x]; };
    auto SM_ = [&](int r, int q) -> T& { return sts[((1 * R + r) * 2 + q) * BS + threadIdx.x]; };
    T um[R][2], uth[R][2], dh[R][2];
    {
      T smh[R][2];
      point_load(a.Sm, smh);
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int q = 0; q < 2; ++q) SM_(r, q) = smh[r][q];
    }
    point_load(a.uh[0], um);
    double acc[5][2];
#pragma unroll
    for (int t = 0; t < 5; ++t) { acc[t][0] = 0.0; acc[t][1] = 0.0; }
    // direction k: the five terms of sca_terms from (u_k TH)_hat = ut, the
    // momentum u_hat_k = u, dTH/dx_k hat = dd and |S| dTH/dx_k hat = sd
    auto fold = [&](const T (&ut)[R][2], const T (&u)[R][2], const T (&dd)[R][2],
                    const T (&sd)[R][2]) {
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          T thv = TH_(r, q), smv = SM_(r, q), utv = ut[r][q], uv = u[r][q], sdv = sd[r][q],
            ddv = dd[r][q];
          ctr_keep(thv);
          ctr_keep(smv);
          ctr_keep(utv);
          ctr_keep(uv);
          ctr_keep(sdv);
          ctr_keep(ddv);
          const double LT = (double)utv - (double)uv * (double)thv;
          const double S = sdv, Smd = (double)smv * (double)ddv;
          acc[0][q] += LT * S;
          acc[1][q] += LT * Smd;
          acc[2][q] += S * S;
          acc[3][q] += S * Smd;
          acc[4][q] += Smd * Smd;
        }
    };
    d.step(0, [&](const T (&o)[2][R][2]) {
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int q = 0; q < 2; ++q) { TH_(r, q) = o[0][r][q]; uth[r][q] = o[1][r][q]; }
    });
    d.step(1, [&](const T (&o)[2][R][2]) { fold(uth, um, o[0], o[1]); });
    point_load(a.uh[1], um);
    d.step(2, [&](const T (&o)[2][R][2]) {
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int q = 0; q < 2; ++q) { uth[r][q] = o[0][r][q]; dh[r][q] = o[1][r][q]; }
    });
    d.step(3, [&](const T (&o)[2][R][2]) {
      fold(uth, um, dh, o[0]);
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int q = 0; q < 2; ++q) uth[r][q] = o[1][r][q];
    });
    point_load(a.uh[2], um);
    d.step(4, [&](const T (&o)[2][R][2]) { fold(uth, um, o[0], o[1]); });
    ctr_tile_sums<T, N, BS>(d.buf, acc, a.part, d.x, d.zt * D::ZT, nz, 10, a.tbase, a.tstride, a.c);
  } while (d.next());
}

// ---------------------------------------------------------------------------
// launchers: false when the fused form does not apply (double fields, a
// length without a compile-time plan or with a last radix above 5, nz not a
// multiple of 4, unaligned buffers); the caller then runs the unfused C2R
// and pointwise kernel.
// ---------------------------------------------------------------------------
// whether the fused kernels serve a plan of element type T, transform length
// n (y) and nz levels; buffers must also be 16-byte aligned (ctr_aligned)
template <class T> inline bool ctr_applies(int nz, int n) {
  if constexpr (sizeof(T) != 4) {
    return false;
  } else {
    if (nz % 4 != 0) return false;
    bool ok = false;
    ct_length(n, [&](auto nct, auto) {
      constexpr int NCT = decltype(nct)::value;
      if constexpr (CtrGeom<NCT, ctr_bs(NCT, false)>::OK && CtrGeom<NCT, ctr_bs(NCT, true)>::OK)
        ok = true;
    });
    return ok;
  }
}
inline bool ctr_aligned(std::initializer_list<const void*> ps) {
  for (const void* q : ps)
    if (q != nullptr && !al16(q)) return false;
  return true;
}

// extra_lds: fields of per-thread point state the kernel holds in LDS;
// SCALAR: the scalar procedure's kernels (threads per workgroup ctr_bs)
template <class T, bool SCALAR, class Args, class K>
bool c2r_contract(alfa_plan* P, hipStream_t s, const Args& a, int nin, int nout_fields,
                  int ref_fields, int ref_pw_fields, int rows, const Fft1D& fyy, const void* tw,
                  const char* name, int extra_lds, K kern_sel) {
  if constexpr (sizeof(T) != 4) {
    return false;
  } else {
    if (a.nz % 4 != 0) return false;
    bool ok = false;
    ct_length(fyy.n, [&](auto nct, auto) {
      constexpr int NCT = decltype(nct)::value;
      if constexpr (CtrGeom<NCT, ctr_bs(NCT, SCALAR)>::OK) ok = true;
    });
    if (!ok) return false;
    if (!live(P)) return true;
    const double F = (double)rows * fyy.n * a.nz;
    ProfScope ps(P, s, name,
                 (double)nin * rows * a.ncy_in * a.nz * sizeof(c2<T>) + (double)nout_fields * F * sizeof(T),
                 (double)ref_fields * rows * a.nz *
                         ((double)(fyy.n / 2 + 1) * sizeof(c2<T>) + (double)fyy.n * sizeof(T)) +
                     (double)ref_pw_fields * F * sizeof(T));
    ct_length(fyy.n, [&](auto nct, auto) {
      constexpr int NCT = decltype(nct)::value;
      constexpr int BS = ctr_bs(NCT, SCALAR);
      if constexpr (CtrGeom<NCT, BS>::OK) {
        constexpr int L = CtrGeom<NCT, BS>::L, LL = CtrGeom<NCT, BS>::LL;
        TileGrid tg{(a.nz + 2 * L - 1) / (2 * L), rows, 1, 0};
        const unsigned grid = tile_grid((size_t)tg.ntz * rows);
        const size_t lds = (size_t)CtrGeom<NCT, BS>::ROWS * LL * sizeof(c2<T>) + (size_t)NCT * sizeof(c2<T>) +
                           (size_t)extra_lds * CtrGeom<NCT, BS>::R * 2 * BS * sizeof(T);
        const bool band = ALFA_BAND && a.ncy_in <= NCT / 4;
        if (band) kern_sel(std::integral_constant<int, NCT>{}, std::true_type{}, grid, lds, tg);
        else kern_sel(std::integral_constant<int, NCT>{}, std::false_type{}, grid, lds, tg);
      }
    });
    launch_ok(P);
    return true;
  }
}

// LM, MM, |S_hat| and u_hat from the 21 x-filtered level-1 spectra sp[]
// (LoadLasdd order); false if the fused form does not apply
template <class T>
bool c2r_lad(alfa_plan* P, hipStream_t s, c2<T>* const* sp, int ncy_in, const double* beta,
             T* const* uh, T* LM, T* MM, T* Shat, int rows, int ny, int nz, const Fft1D& fyy,
             const void* tw, double L, double TFR) {
  LadArgs<T> a{};
  for (int s = 0; s < LAD_NP; ++s)
    for (int h = 0; h < 2; ++h) a.in[2 * s + h] = lad_field(s, h) < 0 ? nullptr : sp[lad_field(s, h)];
  a.beta = beta;
  for (int k = 0; k < 3; ++k) a.uh[k] = uh[k];
  a.LM = LM; a.MM = MM; a.Shat = Shat;
  a.ncy_in = ncy_in; a.ny = ny; a.nz = nz;
  a.scale = 1.0 / ((double)rows * ny);
  a.T1 = 2.0 * L * L;
  a.T2 = 2.0 * (TFR * L) * (TFR * L);
  for (void* q : {(void*)uh[0], (void*)uh[1], (void*)uh[2], (void*)LM, (void*)MM, (void*)Shat})
    if (!al16(q)) return false;
  for (int b = 0; b < 21; ++b)
    if (!al16(sp[b])) return false;
  // reference data flow: 21 irfft2 outputs, then the contraction's 21 reads
  // and 3 writes
  return c2r_contract<T, false>(P, s, a, 27, 6, 21, 24, rows, fyy, tw, "k_y_c2r<LadContract>", 3,
                         [&](auto nct, auto band, unsigned grid, size_t lds, const TileGrid& tg) {
                           constexpr int NCT = decltype(nct)::value;
                           constexpr bool B = decltype(band)::value;
                           constexpr int BS = ctr_bs(NCT, false);
                           launch_k<BS>(P, s, dim3(grid), dim3(BS), lds,
                                        KBODY(k_c2r_lad<T, NCT, BS, B>), a, (const c2<T>*)tw, tg);
                         });
}

// T_up, T_dn from the 10 x-filtered level-1 scalar spectra sp[] (LoadScalar
// order), the momentum hats uh[] and |S_hat|; false if the fused form does
// not apply
template <class T>
bool c2r_sca(alfa_plan* P, hipStream_t s, c2<T>* const* sp, int ncy_in, const double* beta,
             const T* const* uh, const T* Sm, T* Tup, T* Tdn, int rows, int ny, int nz,
             const Fft1D& fyy, const void* tw, double L, double TFR) {
  ScaArgs<T> a{};
  for (int s = 0; s < SCA_NP; ++s)
    for (int h = 0; h < 2; ++h) a.in[2 * s + h] = sp[sca_field(s, h)];
  a.beta = beta;
  for (int k = 0; k < 3; ++k) a.uh[k] = uh[k];
  a.Sm = Sm; a.Tup = Tup; a.Tdn = Tdn;
  a.ncy_in = ncy_in; a.ny = ny; a.nz = nz;
  a.scale = 1.0 / ((double)rows * ny);
  a.L2 = L * L;
  a.L4 = a.L2 * a.L2;
  a.T2 = TFR * TFR;
  for (const void* q : {(const void*)uh[0], (const void*)uh[1], (const void*)uh[2],
                        (const void*)Sm, (const void*)Tup, (const void*)Tdn})
    if (!al16(q)) return false;
  for (int b = 0; b < 10; ++b)
    if (!al16(sp[b])) return false;
  // reference data flow: 10 irfft2 outputs, then the contraction's 14 reads
  // and 2 writes
  return c2r_contract<T, true>(P, s, a, 10, 6, 10, 16, rows, fyy, tw, "k_y_c2r<ScalarContract>", 2,
                         [&](auto nct, auto band, unsigned grid, size_t lds, const TileGrid& tg) {
                           constexpr int NCT = decltype(nct)::value;
                           constexpr bool B = decltype(band)::value;
                           constexpr int BS = ctr_bs(NCT, true);
                           launch_k<BS>(P, s, dim3(grid), dim3(BS), lds,
                                        KBODY(k_c2r_sca<T, NCT, BS, B>), a, (const c2<T>*)tw, tg);
                         });
}

// The level's five LASDD plane-sum terms into the partials part (nx, 10, nz)
// from its 21 x-filtered spectra sp[] (LoadLasdd order, pitch ncy_in);
// level 2 also writes u_hat (uh_out) and |S_hat| (Shat_out), which the scalar
// procedure reads.  The caller has checked ctr_applies and the alignment.
template <class T>
void c2r_lasdd_sums(alfa_plan* P, hipStream_t s, c2<T>* const* sp, int ncy_in, double* part,
                    int level, T* const* uh_out, T* Shat_out, int rows, int ny, int nz,
                    const Fft1D& fyy, const void* tw, double L, double TFR) {
  LasddSumArgs<T> a{};
  for (int st = 0; st < LAD_NP; ++st)
    for (int h = 0; h < 2; ++h)
      a.in[2 * st + h] = lad_field(st, h) < 0 ? nullptr : sp[lad_field(st, h)];
  for (int k = 0; k < 3; ++k) a.uh[k] = uh_out ? uh_out[k] : nullptr;
  a.Shat = Shat_out;
  a.part = part;
  a.ncy_in = ncy_in; a.ny = ny; a.nz = nz;
  a.tbase = level - 1; a.tstride = 2;
  a.scale = 1.0 / ((double)rows * ny);
  // FLasddSums (DynamicSGS_LASDD_SM.py:266-312): level 1 a1 c1 e1 g1 i1 (the
  // even terms), level 2 the odd ones
  const double L2v = L * L, T2 = TFR * TFR, T4 = T2 * T2, T8 = T4 * T4, k2 = 2.0 * L2v;
  const double c1[5] = {k2, k2 * T2, k2 * k2, 4.0 * L2v * L2v * T4, 8.0 * L2v * L2v * T2};
  const double c2v[5] = {k2, k2 * T4, k2 * k2, 4.0 * L2v * L2v * T8, 8.0 * L2v * L2v * T4};
  for (int t = 0; t < 5; ++t) a.c[t] = level == 1 ? c1[t] : c2v[t];
  // reference data flow: the level's 21 irfft2 outputs and their reads by the
  // plane sums
  c2r_contract<T, false>(P, s, a, 27, uh_out ? 4 : 0, 21, 21, rows, fyy, tw, "k_y_c2r<LasddSums>", 3,
                  [&](auto nct, auto band, unsigned grid, size_t lds, const TileGrid& tg) {
                    constexpr int NCT = decltype(nct)::value;
                    constexpr bool B = decltype(band)::value;
                    constexpr int BS = ctr_bs(NCT, false);
                    launch_k<BS>(P, s, dim3(grid), dim3(BS), lds,
                                 KBODY(k_c2r_lasdd_sums<T, NCT, BS, B>), a, (const c2<T>*)tw, tg);
                  });
}

// The level's five scalar LASDD plane-sum terms into part (nx, 10, nz) from
// its 10 x-filtered scalar spectra sp[] (LoadScalar order) and the level's
// momentum u_hat and |S_hat|.  The caller has checked ctr_applies and the
// alignment.
template <class T>
void c2r_sca_sums(alfa_plan* P, hipStream_t s, c2<T>* const* sp, int ncy_in, double* part,
                  int level, const T* const* uh, const T* Sm, int rows, int ny, int nz,
                  const Fft1D& fyy, const void* tw, double L, double TFR) {
  ScaSumArgs<T> a{};
  for (int st = 0; st < SCA_NP; ++st)
    for (int h = 0; h < 2; ++h) a.in[2 * st + h] = sp[sca_field(st, h)];
  for (int k = 0; k < 3; ++k) a.uh[k] = uh[k];
  a.Sm = Sm;
  a.part = part;
  a.ncy_in = ncy_in; a.ny = ny; a.nz = nz;
  a.tbase = 5 * (level - 1); a.tstride = 1;
  a.scale = 1.0 / ((double)rows * ny);
  // FScalarSums (DynamicSGS_ScalarLASDD_SM.py:241-302): a2..e2 at level 1,
  // a4..e4 at level 2
  const double L2 = L * L, L4 = L2 * L2, T2 = TFR * TFR, T4 = T2 * T2, T8 = T4 * T4;
  const double c1[5] = {L2, -L2 * T2, L4, -2.0 * L4 * T2, L4 * T4};
  const double c2v[5] = {L2, -L2 * T4, L4, -2.0 * L4 * T4, L4 * T8};
  for (int t = 0; t < 5; ++t) a.c[t] = level == 1 ? c1[t] : c2v[t];
  c2r_contract<T, true>(P, s, a, 10, 0, 10, 14, rows, fyy, tw, "k_y_c2r<ScalarSums>", 2,
                  [&](auto nct, auto band, unsigned grid, size_t lds, const TileGrid& tg) {
                    constexpr int NCT = decltype(nct)::value;
                    constexpr bool B = decltype(band)::value;
                    constexpr int BS = ctr_bs(NCT, true);
                    launch_k<BS>(P, s, dim3(grid), dim3(BS), lds,
                                 KBODY(k_c2r_sca_sums<T, NCT, BS, B>), a, (const c2<T>*)tw, tg);
                  });
}
