// Common device/host helpers for libalfa (MI355X / gfx950).
#pragma once
#include <initializer_list>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

template <class T> struct cvec;
template <> struct cvec<float> { using type = float2; };
template <> struct cvec<double> { using type = double2; };
template <class T> using c2 = typename cvec<T>::type;

template <class T> __host__ __device__ __forceinline__ c2<T> mkc(T x, T y) {
  c2<T> r; r.x = x; r.y = y; return r;
}
template <class T> __device__ __forceinline__ c2<T> cadd(c2<T> a, c2<T> b) { return mkc<T>(a.x + b.x, a.y + b.y); }
template <class T> __device__ __forceinline__ c2<T> csub(c2<T> a, c2<T> b) { return mkc<T>(a.x - b.x, a.y - b.y); }
template <class T> __device__ __forceinline__ c2<T> cmul(c2<T> a, c2<T> b) {
  return mkc<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <class T> __device__ __forceinline__ c2<T> cscale(c2<T> a, T s) { return mkc<T>(a.x * s, a.y * s); }
// multiply by i*sign (sign = +1 or -1)
template <class T> __device__ __forceinline__ c2<T> cmul_i(c2<T> a, int sign) {
  return sign > 0 ? mkc<T>(-a.y, a.x) : mkc<T>(a.y, -a.x);
}

// Block geometry shared by the FFT kernels: 256 threads; a block transforms
// L interleaved complex lines (L = 8 float or 4 double lines: 64 bytes of z
// per transform element, half of a 128-byte L2 line whose other half belongs
// to the neighbouring z tile, which runs on the same XCD next to it).
// Measured at 256^3 on MI355X, 64-byte rows beat 128-byte rows (higher
// occupancy) and 32-byte rows (more HBM transactions).
constexpr int FFT_BS = 256;
constexpr int FFT_MAXN = 384;      // longest transform: 1.5 * 256
constexpr int FFT_MAXRAD = 8;
// Build-time tuning of the FFT kernels (defaults are the measured best):
//   ALFA_LPB       complex lines per tile for float (double uses half)
//   ALFA_RADIX_MAX largest radix of the register butterflies (4, 8 or 16)
#ifndef ALFA_LPB
#define ALFA_LPB 8
#endif
#ifndef ALFA_RADIX_MAX
#define ALFA_RADIX_MAX 8
#endif
//   ALFA_C2R_BS, ALFA_R2C_BS, ALFA_X_BS: a new plan's threads per workgroup
//   of the compile-time-length y C2R, y R2C and single-op x kernels (256 or
//   512; the lines per tile scale with them, 512 threads give 128-byte tile
//   rows).  Both sizes are built; alfa_plan_autotune / alfa_plan_set_tiles
//   change a plan's choice.
#ifndef ALFA_C2R_BS
#define ALFA_C2R_BS 512
#endif
#ifndef ALFA_R2C_BS
#define ALFA_R2C_BS 256
#endif
#ifndef ALFA_X_BS
#define ALFA_X_BS 256
#endif
//   ALFA_X_BS_F64: the same for the double kernel (measured: 512 is 7% faster
//   for double, 256 is 8% faster for float)
#ifndef ALFA_X_BS_F64
#define ALFA_X_BS_F64 512
#endif
//   ALFA_BATCH_LOADS: the tile loads of the compile-time-length single-op
//   float x kernel are all issued before the first LDS write (1), or one loop
//   iteration at a time (0)
#ifndef ALFA_BATCH_LOADS
#define ALFA_BATCH_LOADS 1
#endif
//   ALFA_C2R_PAIRS: the float y C2R's last pass stores two lines (four z
//   levels, 16 bytes) per item (1) or one line (8 bytes) (0)
#ifndef ALFA_C2R_PAIRS
#define ALFA_C2R_PAIRS 1
#endif
//   ALFA_BAND: band-limited passes for filtered spectra (1) or full passes (0)
#ifndef ALFA_BAND
#define ALFA_BAND 1
#endif
template <class T> struct Lines { static constexpr int L = ALFA_LPB * 8 / (int)sizeof(c2<T>); };
// Lines per tile of an FFT kernel with BS threads: L at 256 threads, in
// proportion above (512 threads: 128-byte rows); below 256 (the 128-thread
// kernels of lengths up to FFT_SMALL_N) the rows stay 64 bytes and each
// thread takes more of the passes' items.
template <class T> constexpr int fft_lines(int bs) { return Lines<T>::L * (bs > FFT_BS ? bs / FFT_BS : 1); }
// Compile-time lengths up to this run their y R2C, x and y C2R kernels on
// 128 threads: at 256 most of a short transform's threads idle in its
// passes (length 96 at 8 lines: 96 radix-8 butterflies for 256 threads),
// while 128-thread workgroups fit twice as many tiles per CU.  Measured on
// one-grid ensembles (convective, mixed): 48^3 +8.5%, 64^3 +10%, 80^3 +4%,
// 96^3 +3%, but 128^3 -2%, which keeps the plan's tile choice.
constexpr int FFT_SMALL_N = 96;

struct Fft1D {
  int n;
  int nrad;
  int rad[FFT_MAXRAD];
};

// MI355X dispatches workgroup i to XCD (i mod 8), and each XCD has its own
// L2.  xcd_logical() renumbers the launch so that every XCD receives one
// contiguous range of logical workgroups: tiles that share HBM rows (the z
// tiles of one transform line set, or the fields of one tile) then run on
// the same L2 at about the same time.
constexpr int NUM_XCD = 8;
__device__ __forceinline__ unsigned xcd_logical(unsigned lin, unsigned total) {
  const unsigned xcd = lin % NUM_XCD, idx = lin / NUM_XCD;
  const unsigned per = total / NUM_XCD, rem = total % NUM_XCD;
  return xcd * per + (xcd < rem ? xcd : rem) + idx;
}

// 16-byte vector of W = 16 / sizeof(T) consecutive elements (4 float or 2
// double): the unit of every vectorised global load and store.
template <class T> struct Vec16 {
  static constexpr int W = 16 / (int)sizeof(T);
  T v[W];
};
template <class T> __device__ __forceinline__ Vec16<T> vld(const T* p) {
  const uint4 q = *reinterpret_cast<const uint4*>(p);
  Vec16<T> r;
  __builtin_memcpy(r.v, &q, 16);
  return r;
}
template <class T> __device__ __forceinline__ void vst(T* p, const Vec16<T>& r) {
  uint4 q;
  __builtin_memcpy(&q, r.v, 16);
  *reinterpret_cast<uint4*>(p) = q;
}
template <class T> __device__ __forceinline__ Vec16<T> vzero() {
  Vec16<T> r;
#pragma unroll
  for (int j = 0; j < Vec16<T>::W; ++j) r.v[j] = T(0);
  return r;
}
// two adjacent complex values (one 16-byte access for float, two for double)
template <class T> __device__ __forceinline__ void cld2(const c2<T>* p, c2<T>& a, c2<T>& b) {
  if constexpr (sizeof(T) == 4) {
    const float4 q = *reinterpret_cast<const float4*>(p);
    a = mkc<T>(q.x, q.y); b = mkc<T>(q.z, q.w);
  } else {
    a = p[0]; b = p[1];
  }
}
template <class T> __device__ __forceinline__ void cst2(c2<T>* p, c2<T> a, c2<T> b) {
  if constexpr (sizeof(T) == 4) {
    *reinterpret_cast<float4*>(p) = make_float4(a.x, a.y, b.x, b.y);
  } else {
    p[0] = a; p[1] = b;
  }
}
inline bool al16(const void* p) { return ((uintptr_t)p & 15u) == 0; }
// the 16-byte vector kernels apply: nz a multiple of the vector width and
// every (non-null) array 16-byte aligned
template <class T> inline bool vec_ok(int nz, std::initializer_list<const void*> ps) {
  if (nz % Vec16<T>::W != 0) return false;
  for (const void* p : ps)
    if (p != nullptr && !al16(p)) return false;
  return true;
}

// Maximum number of distinct fields one batched launch handles.
constexpr int MAXB = 24;

__host__ __device__ inline int imin(int a, int b) { return a < b ? a : b; }
__host__ __device__ inline int imax(int a, int b) { return a > b ? a : b; }

// ----------------------------------------------------------------------------
// Kernel entry points.  Every kernel body of the library is a __device__
// function; it runs as one of two __global__ entry points:
//   k_solo   : one run (one plan), the arguments passed by value;
//   k_member : several runs of the same shape in one launch, blockIdx.z
//              selecting the run, whose arguments the launch reads from a
//              device table of ArgPacks (one per run, in run order).
// Both call the same body with the same argument values, so a run's results
// are bitwise the same launched alone or with others.  The body sees the
// same blockIdx.x/y and gridDim.x/y either way.  BS is the workgroup-size
// bound the body was written for (1024, the compiler's default, where none
// was stated).
// ----------------------------------------------------------------------------
template <class... A> struct ArgPack;
template <> struct ArgPack<> {
  template <class F, class... B>
  __device__ __forceinline__ void call(F f, const B&... b) const {
    [[clang::always_inline]] f(b...);
  }
};
template <class H, class... R> struct ArgPack<H, R...> {
  H h;
  ArgPack<R...> r;
  template <class F, class... B>
  __device__ __forceinline__ void call(F f, const B&... b) const { r.call(f, b..., h); }
};
inline ArgPack<> make_pack() { return {}; }
template <class H, class... R> ArgPack<H, R...> make_pack(H h, R... r) {
  ArgPack<H, R...> p;
  p.h = h;
  p.r = make_pack(r...);
  return p;
}

// (the body is always inlined into the entry point: a call would pass the
// structures through memory and cost registers)
template <int BS, class F, class... A>
__global__ void __launch_bounds__(BS) k_solo(F f, A... a) {
  [[clang::always_inline]] f(a...);
}

// The table is read through the constant address space, as k_solo reads its
// kernel arguments: uniform reads of it become scalar loads into SGPRs
// instead of per-lane vector loads into VGPRs.
template <int BS, class F, class Pk>
__global__ void __launch_bounds__(BS) k_member(F f, const Pk* __restrict__ tab) {
  using CPk = const __attribute__((address_space(4))) Pk;
  const Pk& p = *(const Pk*)((CPk*)tab + blockIdx.z);
  [[clang::always_inline]] p.call(f);
}

// the body of kernel `kern` as a function object for k_solo / k_member.  The
// arguments are forwarded by reference and the bodies take their structures
// by const reference, so a structure indexed at run time (a field table, a
// radix plan) is read in place, from the kernel's argument segment (k_solo)
// or from the device table (k_member), never copied to private memory.
#define KBODY(...) \
  [] __device__(const auto&... a_) { [[clang::always_inline]] __VA_ARGS__(a_...); }
