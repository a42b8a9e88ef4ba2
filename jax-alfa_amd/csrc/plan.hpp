// Plan object and launch helpers shared by the step driver and the
// per-operation entry points.
#pragma once
#include <type_traits>
#include <cstring>
#include <functional>
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

// Optional per-launch timing with HIP events (alfa_profile_begin/end).  A
// record holds the launch's algorithmic bytes (distinct inputs + outputs).
struct ProfRec {
  std::string name;
  double bytes;       // distinct inputs + outputs of this launch
  double ref_bytes;   // the same work in the reference's data flow (see ProfScope)
  hipEvent_t a, b;
};
struct Prof {
  bool on = false;
  std::string filter;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<ProfRec> recs;
};

// One launch of a step recorded for a member group (alfa_group_step): the
// k_member instance it runs (its identity), the launch shape, this run's
// packed arguments, and a function that launches k_member over a table of
// `runs` packs.
struct LaunchRec {
  const void* id;
  dim3 grid, block;
  size_t lds;
  size_t align;
  std::vector<unsigned char> pack;
  std::function<void(const void* tab, int runs, hipStream_t s)> member;
};

struct alfa_plan {
  alfa_config cfg;
  alfa_derived d;
  int prec;
  int nx, ny, nz, nyc, mx, my, myc;
  Fft1D fx, fy, fmx, fmy;
  void *tw_x = nullptr, *tw_y = nullptr, *tw_mx = nullptr, *tw_my = nullptr;
  double* th_b = nullptr;   // Thomas pivots (nx/2+1, nyc, nz+1): row (|kx|, ky)
  double* prof = nullptr;                     // Ug, Vg, rf, rs, Uadv, Vadv, THadv, Qadv: 8*nz
  bool adv_on = false;                        // advection forcing set (alfa_set_forcing)
  const double* fadv() const { return adv_on ? prof + 4 * (size_t)nz : nullptr; }
  // for the scalar path of the RHS kernel on Q: its slot 2 is Qadv
  const double* fadvQ() const { return adv_on ? prof + 5 * (size_t)nz : nullptr; }
  double* scal = nullptr;                     // 8 scalars
  Workspace ws;
  Prof pf;
  std::vector<void*> allocs;
  hipError_t herr = hipSuccess;
  int ncu = 256;   // compute units of the plan's device
  // threads per workgroup of the compile-time-length FFT kernels (256 or
  // 512; the lines per tile scale with them): the build's defaults, or the
  // choices of alfa_plan_autotune / alfa_plan_set_tiles
  int tile_x = 256, tile_c2r = 512, tile_r2c = 256;
  // LAD models in float: the level-1 test-filter inverse transforms fused
  // with the LM/MM and T_up/T_dn contractions (1, the default) or the
  // separate C2R and pointwise kernels (0); alfa_plan_set_fusion
  int fuse_contract = 1;
  // set while a member group records this plan's step: launches are
  // appended here instead of being enqueued
  std::vector<LaunchRec>* rec = nullptr;
};

// Grid of the tiled FFT kernels: one workgroup per tile, rounded up to whole
// XCD rounds, at most
// TILE_GRID_MAX workgroups (each then walks several tiles, which spreads the
// per-workgroup set-up over them)
constexpr size_t TILE_GRID_MAX = 16384;
inline unsigned tile_grid(size_t tiles) {
  const size_t g = (tiles + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
  return (unsigned)(g < TILE_GRID_MAX ? g : TILE_GRID_MAX);
}

int set_error(int code, const char* fmt, ...);

// Launches are enqueued only outside the dry run and while every workspace
// buffer asked for so far exists: a buffer that the plan's dry run did not
// record makes Workspace::get return null, and no kernel is then launched
// with it (the entry point reports ALFA_ESTATE instead).
inline bool live(const alfa_plan* P) { return !P->ws.dry && !P->ws.overflow; }

inline bool launch_ok(alfa_plan* P) {
  if (P->ws.dry) return true;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess && P->herr == hipSuccess) P->herr = e;
  return e == hipSuccess;
}

// Launch body f (a KBODY) with arguments a: enqueued as k_solo, or, while
// the plan records for a member group, appended to its record list.
template <int BS, class F, class... A>
void launch_k(alfa_plan* P, hipStream_t s, dim3 grid, dim3 block, size_t lds, F f, A... a) {
  if (P->rec) {
    using Pk = ArgPack<A...>;
    const Pk pk = make_pack(a...);
    LaunchRec r;
    r.id = reinterpret_cast<const void*>(&k_member<BS, F, Pk>);
    r.grid = grid;
    r.block = block;
    r.lds = lds;
    r.align = alignof(Pk);
    r.pack.resize(sizeof(Pk));
    std::memcpy(r.pack.data(), &pk, sizeof(Pk));
    r.member = [f, grid, block, lds](const void* tab, int runs, hipStream_t st) {
      k_member<BS, F, Pk><<<dim3(grid.x, grid.y, runs), block, lds, st>>>(f, (const Pk*)tab);
    };
    P->rec->push_back(std::move(r));
    return;
  }
  k_solo<BS><<<grid, block, lds, s>>>(f, a...);
}

struct ProfScope {
  alfa_plan* P;
  hipStream_t s;
  long idx = -1;
  // ref_bytes < 0: the same as bytes.  The FFT launchers give the bytes of
  // the reference's data flow instead: each transformed field is one full
  // read of its input and one full write of its output, a half spectrum
  // (nx, ny/2+1, nz) being (ny+2)/ny of a field.
  ProfScope(alfa_plan* P_, hipStream_t s_, const std::string& name, double bytes,
            double ref_bytes = -1.0)
      : P(P_), s(s_) {
    Prof& f = P->pf;
    if (!f.on || P->ws.dry) return;
    if (!f.filter.empty() && f.filter != name) return;
    if (f.used + 2 > f.pool.size()) return;
    ProfRec r{name, bytes, ref_bytes < 0 ? bytes : ref_bytes, f.pool[f.used], f.pool[f.used + 1]};
    f.used += 2;
    (void)hipEventRecord(r.a, s);
    idx = (long)f.recs.size();
    f.recs.push_back(r);
  }
  ~ProfScope() {
    if (idx >= 0) (void)hipEventRecord(P->pf.recs[idx].b, s);
  }
};

inline size_t esize(const alfa_plan* P) { return P->prec == ALFA_FP64 ? 8 : 4; }

// full-field passes (reads + writes) of each pointwise kernel of the step
inline int pw_fields(const char* name) {
  static const struct { const char* n; int f; } tab[] = {
      {"k_derivz", 8}, {"k_derivz_generic", 2}, {"k_stag_avg", 2}, {"k_advection", 12}, {"k_dz_adv_strains", 24}, {"k_dz_adv_strains_v", 24}, {"k_sgs_fused", 25}, {"k_sgs_fused_v", 25}, {"k_ab2_q", 5}, {"k_derivz_v", 8}, {"k_rhs_v", 17},
      {"k_scalar_adv", 7}, {"k_buoyancy", 2}, {"k_strains_uvp", 13}, {"k_strains_w", 12},
      {"k_smag", 7}, {"k_stress_uvp", 10}, {"k_stress_w", 6}, {"k_scalar_flux", 9},
      {"k_rhs", 17}, {"k_ab2_plain", 4}, {"k_imfilter", 2}, {"k_lasdd_lm", 24}, {"k_lasdd_lm_v", 24},
      {"k_smag_level", 7}, {"k_scalar_updn", 16}, {"k_scalar_updn_v", 16}, {"k_real_part", 3}, {"k_fill", 1}};
  for (const auto& e : tab) {
    const size_t l = strlen(e.n);
    if (strncmp(name, e.n, l) == 0 && (name[l] == '<' || name[l] == 0)) return e.f;
  }
  return 0;
}

inline unsigned pw_blocks(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 32768) b = 32768;
  if (b == 0) b = 1;
  return (unsigned)b;
}

// Transform lengths with compile-time FFT kernels (the headline 256^3 grid
// and the common ensemble sizes); every other length runs the generic
// kernels of its register tier (<= 256, <= 384).  f(nct, maxn) is called
// with both as std::integral_constant.
template <class F> inline void ct_length(int n, F&& f) {
  using std::integral_constant;
  switch (n) {
    case 48: f(integral_constant<int, 48>{}, integral_constant<int, 48>{}); break;
    case 64: f(integral_constant<int, 64>{}, integral_constant<int, 64>{}); break;
    case 80: f(integral_constant<int, 80>{}, integral_constant<int, 80>{}); break;
    case 96: f(integral_constant<int, 96>{}, integral_constant<int, 96>{}); break;
    case 128: f(integral_constant<int, 128>{}, integral_constant<int, 128>{}); break;
    case 192: f(integral_constant<int, 192>{}, integral_constant<int, 192>{}); break;
    case 256: f(integral_constant<int, 256>{}, integral_constant<int, 256>{}); break;
    default:
      if (n <= 256) f(integral_constant<int, 0>{}, integral_constant<int, 256>{});
      else f(integral_constant<int, 0>{}, integral_constant<int, FFT_MAXN>{});
  }
}

// y R2C of nb fields, rows x [0, rows), transform length fyy.n, output pitch ncy.
template <class T, class Loader>
void y_r2c(alfa_plan* P, hipStream_t s, const Loader& ld, int rows, const Fft1D& fyy,
           const void* tw, int nz, int ncy, int nb, void* const* outs) {
  if (!live(P)) return;
  PtrBatch pb{};
  for (int b = 0; b < nb; ++b) pb.p[b] = outs[b];
  const double cells = (double)rows * fyy.n * nz;
  ProfScope ps(P, s, std::string("k_y_r2c<") + Loader::NAME + ">",
               Loader::nin(nb) * cells * sizeof(T) +
                   (double)nb * rows * ncy * nz * sizeof(c2<T>),
               (double)nb * rows * nz * ((double)fyy.n * sizeof(T) +
                                          (double)(fyy.n / 2 + 1) * sizeof(c2<T>)));
  // 16-byte loads need nz to be a multiple of the vector width and aligned
  // buffers; the paired complex stores need 16-byte aligned outputs
  bool vec = nz % Vec16<T>::W == 0 && ld.aligned(nb);
  for (int b = 0; b < nb; ++b) vec = vec && al16(outs[b]);
  // the compile-time-length kernels are built for 16-byte accesses only:
  // without them the generic kernel of the length's register tier runs
  const int nsel = vec ? fyy.n : (fyy.n <= 256 ? 1 : FFT_MAXN);
  ct_length(nsel, [&](auto nct, auto maxn) {
    constexpr int NCT = decltype(nct)::value;
    auto launch = [&](auto bs) {
      constexpr int BS = decltype(bs)::value;
      constexpr int L = fft_lines<T>(BS);
      TileGrid tg{(nz + 2 * L - 1) / (2 * L), rows, nb, Loader::BATCH_FAST};
      const unsigned grid = tile_grid((size_t)tg.ntz * rows * nb);
      // (the kernel's first-pass layout: padded rows where they conflict)
      constexpr int LAY = NCT > 0 ? fft_lay(NCT, L * (int)sizeof(c2<T>)) : 0;
      const size_t lds = (size_t)lay_rows(fyy.n, LAY) * L * sizeof(c2<T>) + (size_t)fyy.n * sizeof(c2<T>);
      // pruned output within the stride of the last radix pass: that pass is
      // evaluated only at the outputs the store reads
      if constexpr (NCT > 0) {
        if (ALFA_BAND && ncy <= NCT / last_radix(NCT)) {
          launch_k<BS>(P, s, dim3(grid), dim3(BS), lds, KBODY(k_y_r2c<T, Loader, decltype(maxn)::value, NCT, BS, true>), ld, fyy, (const c2<T>*)tw, nz, ncy, pb, tg, vec ? 1 : 0);
          return;
        }
      }
      launch_k<BS>(P, s, dim3(grid), dim3(BS), lds, KBODY(k_y_r2c<T, Loader, decltype(maxn)::value, NCT, BS>), ld, fyy, (const c2<T>*)tw, nz, ncy, pb, tg, vec ? 1 : 0);
    };
    if constexpr (NCT > 0 && NCT <= FFT_SMALL_N) {
      launch(std::integral_constant<int, 128>{});
    } else if constexpr (NCT > 0) {
      if (P->tile_r2c == 512) launch(std::integral_constant<int, 512>{});
      else launch(std::integral_constant<int, 256>{});
    } else {
      launch(std::integral_constant<int, FFT_BS>{});
    }
  });
  launch_ok(P);
}

// one x launch at BS threads (NCT, MAXN: the plan's length constants)
template <class T, bool MULTI, int BS, class NCTT, class MAXNT>
void launch_kx_bs(alfa_plan* P, hipStream_t s, const XArgs& a, int rows, int nb, const Fft1D& fin,
                  const Fft1D& fout, const void* twin, const void* twout, NCTT, MAXNT) {
  constexpr int NCT = NCTT::value;
  using maxn_t = MAXNT;
  constexpr int L = fft_lines<T>(BS);
  TileGrid tg{(a.nz + L - 1) / L, rows, nb, 0};
  const unsigned grid = tile_grid((size_t)tg.ntz * tg.rows * nb);
  const int nbuf = imax(fin.n, fout.n);
  // the kernel's first-pass layout (padded rows where they conflict); at a
  // compile-time length only the input twiddles are staged
  constexpr int LAY = NCT > 0 ? fft_lay(NCT, L * (int)sizeof(c2<T>)) : 0;
  const size_t lds = (size_t)lay_rows(nbuf, LAY) * L * sizeof(c2<T>) +
                     (size_t)(NCT > 0 ? fin.n : fin.n + fout.n) * sizeof(c2<T>);
  // masked ops between the forward and inverse transforms, each with its
  // kept modes within twice the stride of the radix-8 first pass and within
  // the stride of the last pass: band passes (float only: the double
  // kernel needs 14 more VGPRs for them and loses a wave per SIMD)
  if constexpr (NCT > 0 && sizeof(T) == 4) {
    bool band = ALFA_BAND && radix_of(NCT) == 8 && !a.in_spectral && a.nops >= 1 &&
                (MULTI || a.nops == 1);
    for (int k = 0; k < a.nops && band; ++k) {
      const XOp& o = a.op[k];
      band = o.inverse && o.map == XMAP_ID && o.nout == NCT && o.cx >= 1 && o.cx <= NCT / 4 &&
             o.cx <= NCT / last_radix(NCT);
    }
    if (band) {
      launch_k<BS>(P, s, dim3(grid), dim3(BS), lds, KBODY(k_x<T, maxn_t::value, MULTI, NCT, BS, true>), a, fin, fout, (const c2<T>*)twin, (const c2<T>*)twout, tg);
      return;
    }
  }
  launch_k<BS>(P, s, dim3(grid), dim3(BS), lds, KBODY(k_x<T, maxn_t::value, MULTI, NCT, BS>), a, fin, fout, (const c2<T>*)twin, (const c2<T>*)twout, tg);
}

template <class T, bool MULTI>
void launch_kx(alfa_plan* P, hipStream_t s, const XArgs& a, int rows, int nb, const Fft1D& fin,
               const Fft1D& fout, const void* twin, const void* twout) {
  // compile-time length only when input and output lengths agree
  const int n = fin.n == fout.n ? fin.n : -1;
  const int tier = imax(fin.n, fout.n);
  auto go = [&](auto nct, auto maxn) {
    constexpr int NCT = decltype(nct)::value;
    if constexpr (NCT > 0 && NCT <= FFT_SMALL_N) {
      launch_kx_bs<T, MULTI, 128>(P, s, a, rows, nb, fin, fout, twin, twout, nct, maxn);
    } else if constexpr (NCT > 0 && !MULTI) {
      if (P->tile_x == 512)
        launch_kx_bs<T, MULTI, 512>(P, s, a, rows, nb, fin, fout, twin, twout, nct, maxn);
      else
        launch_kx_bs<T, MULTI, 256>(P, s, a, rows, nb, fin, fout, twin, twout, nct, maxn);
    } else {
      launch_kx_bs<T, MULTI, FFT_BS>(P, s, a, rows, nb, fin, fout, twin, twout, nct, maxn);
    }
  };
  if (n > 0) ct_length(n, go);
  else ct_length(tier <= 256 ? 1 : 1000, go);   // generic kernel of the tier
}

// x pass over nb spectra; the tiles cover ky in [0, rows) (rows = ncy_in by
// default).
template <class T>
void x_pass(alfa_plan* P, hipStream_t s, const XArgs& a_in, int nb, const Fft1D& fin,
            const Fft1D& fout, const void* twin, const void* twout, int rows = -1) {
  if (!live(P)) return;
  XArgs a = a_in;
  // float lines move in 16-byte pairs: nz even and aligned buffers
  bool vec = a.nz % 2 == 0;
  for (int b = 0; b < nb; ++b) {
    vec = vec && al16(a.in[b]);
    for (int o = 0; o < a.nops; ++o) vec = vec && al16(a.dst[b][o]);
  }
  a.vec = vec ? 1 : 0;
  if (rows < 0) rows = a.ncy_in;
  const bool multi = a.nops > 1 || a.op[0].map != XMAP_ID;
  double bytes = (double)nb * fin.n * imin(rows, a.ncy_in) * a.nz * sizeof(c2<T>);
  for (int o = 0; o < a.nops; ++o)
    bytes += (double)nb * a.op[o].nout * imin(a.op[o].ncy, rows) * a.nz * sizeof(c2<T>);
  ProfScope ps(P, s, multi ? "k_x<multi>" : "k_x<single>", bytes);
  if (multi) launch_kx<T, true>(P, s, a, rows, nb, fin, fout, twin, twout);
  else launch_kx<T, false>(P, s, a, rows, nb, fin, fout, twin, twout);
  launch_ok(P);
}

template <class T, class Storer>
void y_c2r(alfa_plan* P, hipStream_t s, const Storer& st, const YC2RArgs& a_in, int nb, int rows,
           const Fft1D& fyy, const void* tw) {
  if (!live(P)) return;
  YC2RArgs a = a_in;
  bool vec = a.nz % Vec16<T>::W == 0 && st.aligned(nb);
  for (int b = 0; b < nb; ++b)
    for (int k = 0; k < a.nsum; ++k) vec = vec && al16(a.in[b][k]);
  a.vec = vec ? 1 : 0;
  ProfScope ps(P, s, std::string("k_y_c2r<") + Storer::NAME + ">",
               (double)nb * a.nsum * rows * a.ncy_in * a.nz * sizeof(c2<T>) +
                   (double)nb * Storer::NOUT * rows * fyy.n * a.nz * sizeof(T),
               (double)nb * a.nsum * rows * a.nz *
                   ((double)(fyy.n / 2 + 1) * sizeof(c2<T>) + (double)fyy.n * sizeof(T)));
  ct_length(fyy.n, [&](auto nct, auto maxn) {
    constexpr int NCT = decltype(nct)::value;
    auto launch = [&](auto bs) {
      // compile-time lengths: tile_c2r threads and lines in proportion
      constexpr int BS = decltype(bs)::value;
      constexpr int L = fft_lines<T>(BS);
      TileGrid tg{(a.nz + 2 * L - 1) / (2 * L), rows, nb, 0};
      const unsigned grid = tile_grid((size_t)tg.ntz * rows * nb);
      const size_t lds = (size_t)fyy.n * L * sizeof(c2<T>) + (size_t)fyy.n * sizeof(c2<T>);
      if constexpr (NCT > 0) {
        // input modes within twice the stride of the radix-8 first pass: band
        // first pass
        const bool band = ALFA_BAND && radix_of(NCT) == 8 && a.ncy_in <= NCT / 4;
        // last radix pass fused into the stores; not for the float
        // divergence storer, whose fused form needs 84 VGPRs against 62
        // and runs 11% slower at 256^3 (the double one gains 4% from it)
        constexpr bool fuse = !(sizeof(T) == 4 && std::is_same<Storer, StoreDiv<T>>::value);
        if (fuse && a.vec) {
          if (band)
            launch_k<BS>(P, s, dim3(grid), dim3(BS), lds, KBODY(k_y_c2r<T, Storer, decltype(maxn)::value, NCT, BS, true, true>), st, a, fyy, (const c2<T>*)tw, tg);
          else
            launch_k<BS>(P, s, dim3(grid), dim3(BS), lds, KBODY(k_y_c2r<T, Storer, decltype(maxn)::value, NCT, BS, true>), st, a, fyy, (const c2<T>*)tw, tg);
          return;
        }
        if (band) {
          launch_k<BS>(P, s, dim3(grid), dim3(BS), lds, KBODY(k_y_c2r<T, Storer, decltype(maxn)::value, NCT, BS, false, true>), st, a, fyy, (const c2<T>*)tw, tg);
          return;
        }
      }
      launch_k<BS>(P, s, dim3(grid), dim3(BS), lds, KBODY(k_y_c2r<T, Storer, decltype(maxn)::value, NCT, BS>), st, a, fyy, (const c2<T>*)tw, tg);
    };
    if constexpr (NCT > 0 && NCT <= FFT_SMALL_N) {
      launch(std::integral_constant<int, 128>{});
    } else if constexpr (NCT > 0) {
      if (P->tile_c2r == 256) launch(std::integral_constant<int, 256>{});
      else launch(std::integral_constant<int, 512>{});
    } else {
      launch(std::integral_constant<int, FFT_BS>{});
    }
  });
  launch_ok(P);
}

// Plane sums of NT terms -> out[NT][nz] * scale (deterministic).
template <int NT, class F>
void plane_sums(alfa_plan* P, hipStream_t s, int ncol, int nz, const F& f, double* out,
                double scale, const char* tag) {
  RedGeom g = red_geom(ncol, nz, esize(P));
  double* part = (double*)P->ws.get(std::string("red_") + tag, (size_t)g.nbc * NT * nz * sizeof(double));
  if (!live(P)) return;
  dim3 grid(g.nbz, g.nbc);
  {
    ProfScope ps(P, s, std::string("k_plane_partial<") + F::NAME + ">",
                 (double)F::NFIELDS * ncol * nz * esize(P));
    launch_k<RED_BS>(P, s, dim3(grid), dim3(RED_BS), 0, KBODY(k_plane_partial<NT, F>), g, f, part);
  }
  launch_ok(P);
  launch_k<1024>(P, s, dim3((NT * nz + 3) / 4), dim3(256), 0, KBODY(k_plane_final), g.nbc, NT, nz, part, out, scale);
  launch_ok(P);
}

// Imfilter of num and den per level, the clipped ratio into out and the
// plane sums of sqrt(out) and out (k_box_clip + k_plane_final).
template <class T>
void box_clip_sums(alfa_plan* P, hipStream_t s, int nx, int ny, int nz, const T* num,
                   const T* den, T* out, double* sums, const char* tag) {
  const int ncol = nx * ny;
  RedGeom g = red_geom(ncol, nz, esize(P));
  double* part = (double*)P->ws.get(std::string("red_") + tag, (size_t)g.nbc * 2 * nz * sizeof(double));
  if (!live(P)) return;
  dim3 grid(g.nbz, g.nbc);
  {
    ProfScope ps(P, s, "k_box_clip", 3.0 * ncol * nz * esize(P));
    // 16-byte form where the levels and buffers allow it (same partials)
    if (vec_ok<T>(nz, {num, den, out}))
      launch_k<RED_BS>(P, s, dim3(grid), dim3(RED_BS / Vec16<T>::W), 0, KBODY(k_box_clip_v<T>), g, num, den, out, nx, ny, part);
    else
      launch_k<RED_BS>(P, s, dim3(grid), dim3(RED_BS), 0, KBODY(k_box_clip<T>), g, num, den, out, nx, ny, part);
  }
  launch_ok(P);
  launch_k<1024>(P, s, dim3((2 * nz + 3) / 4), dim3(256), 0, KBODY(k_plane_final), g.nbc, 2, nz, part, sums, 1.0);
  launch_ok(P);
}

// n points handled w per thread
#define PW_LAUNCH_W(P, s, n, w, kern, ...)                                \
  do {                                                                    \
    if (live(P)) {                                                   \
      {                                                                   \
        ProfScope ps_(P, s, #kern, (double)pw_fields(#kern) * (n) * esize(P)); \
        launch_k<1024>(P, s, dim3(pw_blocks((n) / (w))), dim3(256), 0, KBODY(kern), __VA_ARGS__); \
      }                                                                   \
      launch_ok(P);                                                       \
    }                                                                     \
  } while (0)

#define PW_LAUNCH(P, s, n, kern, ...)                                     \
  do {                                                                    \
    if (live(P)) {                                                   \
      {                                                                   \
        ProfScope ps_(P, s, #kern, (double)pw_fields(#kern) * (n) * esize(P)); \
        launch_k<1024>(P, s, dim3(pw_blocks(n)), dim3(256), 0, KBODY(kern), __VA_ARGS__); \
      }                                                                   \
      launch_ok(P);                                                       \
    }                                                                     \
  } while (0)

// Single kernel launch with a profiling record (a body written for any
// workgroup size up to 1024).
#define KLAUNCH(P, s, name, bytes, kern, grid, block, ...)                \
  do {                                                                    \
    if (live(P)) {                                                   \
      {                                                                   \
        ProfScope ps_(P, s, name, bytes);                                 \
        launch_k<1024>(P, s, dim3(grid), dim3(block), 0, KBODY(kern), __VA_ARGS__); \
      }                                                                   \
      launch_ok(P);                                                       \
    }                                                                     \
  } while (0)

// KLAUNCH with dynamic LDS bytes and the body's workgroup-size bound BS
#define KLAUNCH_LDS(P, s, name, bytes, BS, kern, grid, block, lds, ...)   \
  do {                                                                    \
    if (live(P)) {                                                   \
      {                                                                   \
        ProfScope ps_(P, s, name, bytes);                                 \
        launch_k<BS>(P, s, dim3(grid), dim3(block), lds, KBODY(kern), __VA_ARGS__); \
      }                                                                   \
      launch_ok(P);                                                       \
    }                                                                     \
  } while (0)
