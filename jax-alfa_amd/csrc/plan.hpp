// Plan object and launch helpers shared by the step driver and the
// per-operation entry points.
#pragma once
#include <map>
#include <string>
#include <vector>
#include "alfa.h"
#include "common.hpp"
#include "fft.hpp"
#include "pointwise.hpp"
#include "sgs.hpp"

// Device workspace.  Every buffer the step or an operation needs is named;
// plan creation runs the step once in "dry" mode, which launches nothing and
// only records (name, size); the recorded set is then allocated once, so no
// call after alfa_plan_create allocates device memory.
struct Workspace {
  bool dry = true;
  bool overflow = false;
  std::string overflow_name;
  std::map<std::string, std::pair<size_t, void*>> bufs;
  void* get(const std::string& name, size_t bytes) {
    auto it = bufs.find(name);
    if (it == bufs.end()) {
      if (!dry) { overflow = true; overflow_name = name; return nullptr; }
      bufs[name] = {bytes, (void*)(uintptr_t)0x1000};
      return (void*)(uintptr_t)0x1000;
    }
    if (bytes > it->second.first) {
      if (dry) it->second.first = bytes;
      else { overflow = true; overflow_name = name; return nullptr; }
    }
    return it->second.second;
  }
};

struct alfa_plan {
  alfa_config cfg;
  alfa_derived d;
  int prec;
  int nx, ny, nz, nyc, mx, my, myc;
  Fft1D fx, fy, fmx, fmy;
  void *tw_x = nullptr, *tw_y = nullptr, *tw_mx = nullptr, *tw_my = nullptr;
  double *th_m = nullptr, *th_b = nullptr;    // (nx*nyc, nz+1)
  double* prof = nullptr;                     // Ug, Vg, rf, rs: 4*nz
  double* scal = nullptr;                     // 8 scalars
  Workspace ws;
  std::vector<void*> allocs;
  hipError_t herr = hipSuccess;
};

int set_error(int code, const char* fmt, ...);

inline bool launch_ok(alfa_plan* P) {
  if (P->ws.dry) return true;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess && P->herr == hipSuccess) P->herr = e;
  return e == hipSuccess;
}

inline unsigned pw_blocks(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 32768) b = 32768;
  if (b == 0) b = 1;
  return (unsigned)b;
}

template <class T> inline size_t fft_lds(int n) { return (size_t)n * Lines<T>::L * sizeof(c2<T>) + (size_t)n * sizeof(c2<T>); }

// y R2C of nb fields, rows x [0, rows), transform length fyy.n, output pitch ncy.
template <class T, class Loader>
void y_r2c(alfa_plan* P, hipStream_t s, const Loader& ld, int rows, const Fft1D& fyy,
           const void* tw, int nz, int ncy, int nb, void* const* outs) {
  if (P->ws.dry) return;
  PtrBatch pb{};
  for (int b = 0; b < nb; ++b) pb.p[b] = outs[b];
  constexpr int ZT = 2 * Lines<T>::L;
  dim3 grid((nz + ZT - 1) / ZT, rows, nb);
  k_y_r2c<T, Loader><<<grid, FFT_BS, fft_lds<T>(fyy.n), s>>>(ld, fyy, (const c2<T>*)tw, nz, ncy, pb);
  launch_ok(P);
}

template <class T>
void x_pass(alfa_plan* P, hipStream_t s, const XArgs& a, int nb, const Fft1D& fin,
            const Fft1D& fout, const void* twin, const void* twout) {
  if (P->ws.dry) return;
  constexpr int L = Lines<T>::L;
  dim3 grid((a.nz + L - 1) / L, a.ncy_in, nb);
  const int nbuf = imax(fin.n, fout.n);
  const size_t lds = (size_t)nbuf * L * sizeof(c2<T>) + (size_t)(fin.n + fout.n) * sizeof(c2<T>);
  k_x<T><<<grid, FFT_BS, lds, s>>>(a, fin, fout, (const c2<T>*)twin, (const c2<T>*)twout);
  launch_ok(P);
}

template <class T, class Storer>
void y_c2r(alfa_plan* P, hipStream_t s, const Storer& st, const YC2RArgs& a, int nb, int rows,
           const Fft1D& fyy, const void* tw) {
  if (P->ws.dry) return;
  constexpr int ZT = 2 * Lines<T>::L;
  dim3 grid((a.nz + ZT - 1) / ZT, rows, nb);
  k_y_c2r<T, Storer><<<grid, FFT_BS, fft_lds<T>(fyy.n), s>>>(st, a, fyy, (const c2<T>*)tw);
  launch_ok(P);
}

// Plane sums of NT terms -> out[NT][nz] * scale (deterministic).
template <int NT, class F>
void plane_sums(alfa_plan* P, hipStream_t s, int ncol, int nz, const F& f, double* out,
                double scale, const char* tag) {
  RedGeom g = red_geom(ncol, nz);
  double* part = (double*)P->ws.get(std::string("red_") + tag, (size_t)g.nbc * NT * nz * sizeof(double));
  if (P->ws.dry) return;
  dim3 grid(g.nbz, g.nbc);
  k_plane_partial<NT, F><<<grid, RED_BS, 0, s>>>(g, f, part);
  launch_ok(P);
  k_plane_final<<<(NT * nz + 255) / 256, 256, 0, s>>>(g.nbc, NT, nz, part, out, scale);
  launch_ok(P);
}

#define PW_LAUNCH(P, s, n, kern, ...)                                     \
  do {                                                                    \
    if (!(P)->ws.dry) {                                                   \
      kern<<<pw_blocks(n), 256, 0, s>>>(__VA_ARGS__);                     \
      launch_ok(P);                                                       \
    }                                                                     \
  } while (0)
