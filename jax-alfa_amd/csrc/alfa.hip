// libalfa: C ABI of the MI355X-native JAX-ALFA time step (include/alfa.h).
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>
#include "step.hpp"
#include "stats.hpp"

using namespace alfa;

static thread_local std::string g_err;

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

extern "C" const char* alfa_last_error(void) { return g_err.c_str(); }

static bool factor(int n, Fft1D& f) {
  f.n = n;
  f.nrad = 0;
  int m = n;
#if ALFA_RADIX_MAX >= 16
  const int order[6] = {16, 8, 4, 2, 3, 5};
#elif ALFA_RADIX_MAX >= 8
  const int order[6] = {8, 4, 2, 3, 5, 1};
#else
  const int order[6] = {4, 2, 3, 5, 1, 1};
#endif
  for (int r : order) {
    if (r == 1) continue;
    while (m % r == 0) {
      if (f.nrad >= FFT_MAXRAD) return false;
      f.rad[f.nrad++] = r;
      m /= r;
    }
  }
  return m == 1;
}

template <class T>
static void* make_twiddles(int n, std::vector<void*>& allocs, hipError_t& e) {
  std::vector<c2<T>> h(n);
  for (int t = 0; t < n; ++t) {
    const double a = 2.0 * M_PI * (double)t / (double)n;
    h[t] = mkc<T>((T)std::cos(a), (T)(-std::sin(a)));
  }
  void* p = nullptr;
  e = hipMalloc(&p, n * sizeof(c2<T>));
  if (e != hipSuccess) return nullptr;
  allocs.push_back(p);
  e = hipMemcpy(p, h.data(), n * sizeof(c2<T>), hipMemcpyHostToDevice);
  return p;
}

static int check_hip(alfa_plan* P, const char* where) {
  hipError_t e = P->herr;
  P->herr = hipSuccess;
  if (e != hipSuccess) return set_error(ALFA_EHIP, "%s: %s", where, hipGetErrorString(e));
  if (P->ws.overflow) {
    P->ws.overflow = false;   // reported once; the next call starts clean
    return set_error(ALFA_ESTATE, "%s: workspace buffer '%s' was not planned", where,
                     P->ws.overflow_name.c_str());
  }
  return ALFA_OK;
}

// ---------------------------------------------------------------------------
// per-operation implementations (templated on the field type)
// ---------------------------------------------------------------------------
namespace ops {

template <class T> int fft(alfa_plan* P, hipStream_t s, const void* F, void* out) {
  Ctx<T> c(P, s);
  LoadPtr<T> ld{};
  ld.p[0] = (const T*)F; ld.ny = c.ny; ld.nz = c.nz;
  c2<T>* o = (c2<T>*)out;
  rfft2_fields<T>(c, ld, c.nz, 1, &o);
  return 0;
}

template <class T> int fft_pad(alfa_plan* P, hipStream_t s, const void* F, void* out) {
  Ctx<T> c(P, s);
  const int myc = c.my / 2 + 1;
  LoadPtr<T> ld{};
  ld.p[0] = (const T*)F; ld.ny = c.my; ld.nz = c.nz;
  void* o[1] = {out};
  y_r2c<T>(P, s, ld, c.mx, P->fmy, P->tw_my, c.nz, myc, 1, o);
  XArgs xa{};
  xa.in[0] = out; xa.dst[0][0] = out;
  xa.op[0] = xop(myc, 0, BIGCUT, XM_NONE, XMAP_ID, c.mx);
  xa.nops = 1; xa.ncy_in = myc; xa.nz = c.nz; xa.ny_full = c.my;
  x_pass<T>(P, s, xa, 1, P->fmx, P->fmx, P->tw_mx, P->tw_mx);
  return 0;
}

// masked / derivative inverse of a full-pitch spectrum into a real field
template <class T>
void spectral_to_real(Ctx<T>& c, const void* Ffft, int ncy, int cx, int mult, T* out) {
  alfa_plan* P = c.P;
  c2<T>* tmp = c.S("op0");
  XArgs xa{};
  xa.in[0] = Ffft; xa.dst[0][0] = tmp;
  xa.op[0] = xop(ncy, 1, cx, mult, XMAP_ID, c.nx);
  xa.nops = 1; xa.ncy_in = c.nyc; xa.nz = c.nz; xa.in_spectral = 1; xa.ny_full = c.ny;
  x_pass<T>(P, c.s, xa, 1, P->fx, P->fx, P->tw_x, P->tw_x);
  YC2RArgs ya{};
  ya.in[0][0] = tmp; ya.nsum = 1; ya.ncy_in = ncy; ya.nz = c.nz;
  ya.scale = 1.0 / ((double)c.nx * c.ny);
  StorePtr<T> st{};
  st.p[0] = out; st.ny = c.ny; st.nz = c.nz;
  y_c2r<T>(P, c.s, st, ya, 1, c.nx, P->fy, P->tw_y);
}

template <class T>
int filtering_explicit(alfa_plan* P, hipStream_t s, const void* Ffft, void* F, void* Fnew) {
  Ctx<T> c(P, s);
  const auto& d = P->d;
  if (!live(P)) return 0;
  const size_t n = 2 * (size_t)c.nx * c.nyc * c.nz;
  launch_k<1024>(P, s, dim3(pw_blocks(n)), dim3(256), 0, KBODY(k_fill<T>), n, (T*)Fnew, T(0));
  launch_ok(P);
  // masked spectrum at full pitch: only the first cut0_y columns are
  // processed (grid), columns >= cut0_y keep the zeros written above
  XArgs xa{};
  xa.in[0] = Ffft; xa.dst[0][0] = Fnew;
  xa.op[0] = xop(c.nyc, 0, d.cut0_x, XM_NONE, XMAP_ID, c.nx);
  xa.nops = 1; xa.ncy_in = c.nyc; xa.nz = c.nz; xa.in_spectral = 1; xa.ny_full = c.ny;
  x_pass<T>(P, s, xa, 1, P->fx, P->fx, P->tw_x, P->tw_x, d.cut0_y);
  spectral_to_real<T>(c, Fnew, d.cut0_y, BIGCUT, XM_NONE, (T*)F);
  return 0;
}

template <class T> int filtering_level(alfa_plan* P, hipStream_t s, const void* Ffft, void* out,
                                       int level) {
  Ctx<T> c(P, s);
  const auto& d = P->d;
  if (level == 1) spectral_to_real<T>(c, Ffft, d.cut1_y, d.cut1_x, XM_NONE, (T*)out);
  else spectral_to_real<T>(c, Ffft, d.cut2_y, d.cut2_x, XM_NONE, (T*)out);
  return 0;
}

template <class T> int derivxy(alfa_plan* P, hipStream_t s, const void* Ffft, int axis, void* out) {
  Ctx<T> c(P, s);
  spectral_to_real<T>(c, Ffft, c.nyc, BIGCUT, axis == 0 ? XM_IKX : XM_IKY, (T*)out);
  return 0;
}

template <class T> int dealias1_op(alfa_plan* P, hipStream_t s, const void* Ffft, void* out) {
  Ctx<T> c(P, s);
  c2<T>* pp = c.PS("op0");
  XArgs xa{};
  xa.in[0] = Ffft; xa.dst[0][0] = pp;
  xa.op[0] = xop(c.nyc, 1, BIGCUT, XM_NONE, XMAP_D1, c.mx);
  xa.nops = 1; xa.ncy_in = c.nyc; xa.nz = c.nz; xa.in_spectral = 1; xa.ny_full = c.ny;
  x_pass<T>(P, s, xa, 1, P->fx, P->fmx, P->tw_x, P->tw_mx);
  YC2RArgs ya{};
  ya.in[0][0] = pp; ya.nsum = 1; ya.ncy_in = c.nyc; ya.nz = c.nz;
  ya.scale = 1.0 / ((double)c.mx * c.my);
  StorePtr<T> st{};
  st.p[0] = (T*)out; st.ny = c.my; st.nz = c.nz;
  y_c2r<T>(P, s, st, ya, 1, c.mx, P->fmy, P->tw_my);
  return 0;
}

template <class T> int dealias2_op(alfa_plan* P, hipStream_t s, const void* Fpfft, void* out) {
  Ctx<T> c(P, s);
  const int ncy = c.ny / 2;
  c2<T>* tmp = c.S("op0");
  XArgs xa{};
  xa.in[0] = Fpfft; xa.dst[0][0] = tmp;
  xa.op[0] = xop(ncy, 1, BIGCUT, XM_NONE, XMAP_D2, c.nx);
  xa.nops = 1; xa.ncy_in = c.my / 2 + 1; xa.nz = c.nz; xa.in_spectral = 1; xa.ny_full = c.ny;
  x_pass<T>(P, s, xa, 1, P->fmx, P->fx, P->tw_mx, P->tw_x);
  YC2RArgs ya{};
  ya.in[0][0] = tmp; ya.nsum = 1; ya.ncy_in = ncy; ya.nz = c.nz;
  ya.scale = 2.25 / ((double)c.nx * c.ny);
  StorePtr<T> st{};
  st.p[0] = (T*)out; st.ny = c.ny; st.nz = c.nz;
  y_c2r<T>(P, s, st, ya, 1, c.nx, P->fy, P->tw_y);
  return 0;
}

template <class T>
int advection(alfa_plan* P, hipStream_t s, const void* u, const void* v, const void* w,
              const void* dudy, const void* dudz, const void* dvdx, const void* dvdz,
              const void* dwdx, const void* dwdy, void* Cx, void* Cy, void* Cz) {
  Ctx<T> c(P, s);
  if (P->d.optDealias == 0) {
    PW_LAUNCH(P, s, c.F, k_advection<T>, c.dims(), (const T*)u, (const T*)v, (const T*)w,
              (const T*)dudy, (const T*)dudz, (const T*)dvdx, (const T*)dvdz, (const T*)dwdx,
              (const T*)dwdy, (T*)Cx, (T*)Cy, (T*)Cz);
    return 0;
  }
  const T* in9[9] = {(const T*)u, (const T*)v, (const T*)w, (const T*)dudy, (const T*)dudz,
                     (const T*)dvdx, (const T*)dvdz, (const T*)dwdx, (const T*)dwdy};
  T* p9[9];
  for (int b = 0; b < 9; ++b) p9[b] = c.Pad("adv" + std::to_string(b));
  dealias1<T>(c, 9, in9, p9);
  T *pcx = c.Pad("cx"), *pcy = c.Pad("cy"), *pcz = c.Pad("cz");
  PW_LAUNCH(P, s, (size_t)c.mx * c.my * c.nz, k_advection<T>, c.pdims(), p9[0], p9[1], p9[2],
            p9[3], p9[4], p9[5], p9[6], p9[7], p9[8], pcx, pcy, pcz);
  const T* pc3[3] = {pcx, pcy, pcz};
  T* o3[3] = {(T*)Cx, (T*)Cy, (T*)Cz};
  dealias2<T>(c, 3, pc3, o3);
  return 0;
}

template <class T>
int scalar_advection(alfa_plan* P, hipStream_t s, const void* u, const void* v, const void* w,
                     const void* sx, const void* sy, const void* sz, void* out) {
  Ctx<T> c(P, s);
  if (P->d.optDealias == 0) {
    PW_LAUNCH(P, s, c.F, k_scalar_adv<T>, c.dims(), (const T*)u, (const T*)v, (const T*)w,
              (const T*)sx, (const T*)sy, (const T*)sz, (T*)out);
    return 0;
  }
  const T* in6[6] = {(const T*)u, (const T*)v, (const T*)w, (const T*)sx, (const T*)sy, (const T*)sz};
  T* p6[6];
  for (int b = 0; b < 6; ++b) p6[b] = c.Pad("adv" + std::to_string(b));
  dealias1<T>(c, 6, in6, p6);
  T* pcs = c.Pad("cs");
  PW_LAUNCH(P, s, (size_t)c.mx * c.my * c.nz, k_scalar_adv<T>, c.pdims(), p6[0], p6[1], p6[2],
            p6[3], p6[4], p6[5], pcs);
  const T* pc1[1] = {pcs};
  T* o1[1] = {(T*)out};
  dealias2<T>(c, 1, pc1, o1);
  return 0;
}

template <class T> int buoyancy(alfa_plan* P, hipStream_t s, const void* TH, void* out) {
  Ctx<T> c(P, s);
  double* bar = c.D("op_bar", c.nz);
  FSum1<T> f{(const T*)TH, c.nz, P->cfg.T_0};
  plane_sums<1>(P, s, c.ncol, c.nz, f, bar, 1.0 / c.ncol, "op1");
  PW_LAUNCH(P, s, c.F, k_buoyancy<T>, c.dims(), (const T*)TH, bar, (T*)out, P->cfg.T_0,
            P->d.g_nondim, P->cfg.optBuoyancy);
  return 0;
}

template <class T> int planar_mean(alfa_plan* P, hipStream_t s, const void* F, double* out) {
  Ctx<T> c(P, s);
  FSum1<T> f{(const T*)F, c.nz, 0.0};
  plane_sums<1>(P, s, c.ncol, c.nz, f, out, 1.0 / c.ncol, "op1");
  return 0;
}

template <class T>
int rhs_momentum(alfa_plan* P, hipStream_t s, const void* u, const void* v, const void* w,
                 const void* Cx, const void* Cy, const void* Cz, const void* b,
                 const void* dtx, const void* dty, const void* dtz, void* Ru, void* Rv, void* Rw) {
  Ctx<T> c(P, s);
  double* m = c.D("op_uv", 2 * c.nz);
  FSum3<T> f{(const T*)u, (const T*)v, (const T*)u, c.nz, 0.0};
  double* m3 = c.D("op_m3", 3 * c.nz);
  plane_sums<3>(P, s, c.ncol, c.nz, f, m3, 1.0 / c.ncol, "op3");
  (void)m;
  const double* pr = P->prof;
  PW_LAUNCH(P, s, c.F, k_rhs<T>, c.dims(), (const T*)u, (const T*)v, (const T*)w, (const T*)nullptr,
            (const T*)Cx, (const T*)Cy, (const T*)Cz, (const T*)b, (const T*)dtx, (const T*)dty,
            (const T*)dtz, (const T*)nullptr, (const T*)nullptr, pr, pr + c.nz, pr + 2 * c.nz,
            pr + 3 * c.nz, m3, m3 + c.nz, m3 + 2 * c.nz, P->d.f_coriolis_nondim, P->cfg.Ugal,
            P->cfg.optDamping, (T*)Ru, (T*)Rv, (T*)Rw, (T*)nullptr, (const double*)nullptr, 0.0,
            0.0, 0, P->fadv());
  return 0;
}

template <class T>
int rhs_scalar(alfa_plan* P, hipStream_t s, const void* TH, const void* adv, const void* divq,
               void* RT) {
  Ctx<T> c(P, s);
  double* m = c.D("op_bar", c.nz);
  FSum1<T> f{(const T*)TH, c.nz, 0.0};
  plane_sums<1>(P, s, c.ncol, c.nz, f, m, 1.0 / c.ncol, "op1");
  const double* pr = P->prof;
  PW_LAUNCH(P, s, c.F, k_rhs<T>, c.dims(), (const T*)nullptr, (const T*)nullptr,
            (const T*)nullptr, (const T*)TH, (const T*)nullptr, (const T*)nullptr,
            (const T*)nullptr, (const T*)nullptr, (const T*)nullptr, (const T*)nullptr,
            (const T*)nullptr, (const T*)adv, (const T*)divq, pr, pr + c.nz, pr + 2 * c.nz,
            pr + 3 * c.nz, m, m, m, 0.0, 0.0, P->cfg.optDamping, (T*)nullptr, (T*)nullptr,
            (T*)nullptr, (T*)RT, (const double*)nullptr, 0.0, 0.0, 0, P->fadv());
  return 0;
}

template <class T>
int pressure_rc(alfa_plan* P, hipStream_t s, const void* const R[3], const void* const Rp[3],
                const void* const U[3], const void* divtz, void* RCr, void* RCi, void* fRzr) {
  Ctx<T> c(P, s);
  T* dp[3] = {c.W("op_dpx"), c.W("op_dpy"), c.W("op_dpz")};
  const T* r[3] = {(const T*)R[0], (const T*)R[1], (const T*)R[2]};
  const T* rp[3] = {(const T*)Rp[0], (const T*)Rp[1], (const T*)Rp[2]};
  const T* u[3] = {(const T*)U[0], (const T*)U[1], (const T*)U[2]};
  c2<T>* fR[3];
  stage_pressure<T>(c, r, rp, u, (const T*)divtz, dp[0], dp[1], dp[2], nullptr, (T*)RCr, (T*)RCi,
                    fR);
  PW_LAUNCH(P, s, (size_t)c.nx * c.nyc * c.nz, k_real_part<T>, (size_t)c.nx * c.nyc * c.nz,
            (const c2<T>*)fR[2], (T*)fRzr);
  return 0;
}

template <class T>
int thomas_solve(alfa_plan* P, hipStream_t s, const void* RCr, const void* RCi, const void* fRzr,
                 void* p, void* dpdx, void* dpdy, void* dpdz) {
  Ctx<T> c(P, s);
  const auto& d = P->d;
  c2<T>* phat = c.S("phat");
  c2<T>* dpz = c.S("dpzh");
  c2<T>* dpx = c.S("dpxh");
  ThomasArgs<T> ta{};
  ta.b = P->th_b; ta.phat = phat; ta.dpz = dpz;
  ta.RCin_r = (const T*)RCr; ta.RCin_i = (const T*)RCi; ta.fRzr_in = (const T*)fRzr;
  ta.nx = c.nx; ta.ny = c.ny; ta.nyc = c.nyc; ta.nz = c.nz;
  ta.dz = d.dz; ta.idz = 1.0 / d.dz; ta.c_in = 1.0 / (d.dz * d.dz);
  if (live(P)) {
    launch_k<64>(P, s, dim3(c.nx * c.nyc), dim3(64), thomas_lds(c.nz), KBODY(k_thomas<T>), ta);
    launch_ok(P);
  }
  XArgs xa{};
  xa.ncy_in = c.nyc; xa.nz = c.nz; xa.ny_full = c.ny; xa.in_spectral = 1;
  xa.in[0] = phat; xa.dst[0][0] = dpx; xa.dst[0][1] = phat;
  xa.op[0] = xop(c.nyc, 1, BIGCUT, XM_IKX, XMAP_ID, c.nx);
  xa.op[1] = xop(c.nyc, 1, BIGCUT, XM_NONE, XMAP_ID, c.nx);
  xa.nops = 2;
  x_pass<T>(P, s, xa, 1, P->fx, P->fx, P->tw_x, P->tw_x);
  xa.in[0] = dpz; xa.dst[0][0] = dpz; xa.nops = 1;
  xa.op[0] = xop(c.nyc, 1, BIGCUT, XM_NONE, XMAP_ID, c.nx);
  x_pass<T>(P, s, xa, 1, P->fx, P->fx, P->tw_x, P->tw_x);
  YC2RArgs ya{};
  ya.in[0][0] = dpx; ya.in[1][0] = phat; ya.dy[1] = 1; ya.in[2][0] = dpz; ya.in[3][0] = phat;
  ya.nsum = 1; ya.ncy_in = c.nyc; ya.nz = c.nz; ya.scale = 1.0 / ((double)c.nx * c.ny);
  ya.dyscale = 1.0;
  StorePtr<T> st{};
  st.p[0] = (T*)dpdx; st.p[1] = (T*)dpdy; st.p[2] = (T*)dpdz; st.p[3] = (T*)p;
  st.ny = c.ny; st.nz = c.nz;
  y_c2r<T>(P, s, st, ya, 4, c.nx, P->fy, P->tw_y);
  return 0;
}

template <class T>
int lasdd(alfa_plan* P, hipStream_t s, const void* u, const void* v, const void* w,
          const void* const str[7], void* Cs2, double* a1, double* a2, double* beta) {
  Ctx<T> c(P, s);
  const T* Sij[6];
  for (int n = 0; n < 6; ++n) Sij[n] = (const T*)str[n];
  stage_lasdd<T>(c, (const T*)u, (const T*)v, (const T*)w, Sij, (const T*)str[6], nullptr, nullptr,
                 (T*)Cs2, a1, a2, beta);
  return 0;
}

template <class T>
int lasdd_wl(alfa_plan* P, hipStream_t s, const void* u, const void* v, const void* w,
             const void* const str[6], void* Cwl, double* a1, double* a2, double* beta) {
  Ctx<T> c(P, s);
  const T* Sij[6];
  for (int n = 0; n < 6; ++n) Sij[n] = (const T*)str[n];
  stage_lasdd_wl<T>(c, (const T*)u, (const T*)v, (const T*)w, Sij, nullptr, nullptr, (T*)Cwl, a1,
                    a2, beta);
  return 0;
}

template <class T>
int scalar_lasdd(alfa_plan* P, hipStream_t s, const void* const vel[3], const void* const h1[3],
                 const void* const h2[3], const void* TH, const void* const dT[3], const void* S,
                 const void* Shat, const void* Shatd, void* CsPr, double* prof, double* beta) {
  Ctx<T> c(P, s);
  MomHats<T> mh{};
  for (int k = 0; k < 3; ++k) {
    mh.uh1[k] = (const T*)h1[k];
    mh.uh2[k] = h2 ? (const T*)h2[k] : nullptr;
  }
  mh.Shat = (const T*)Shat;
  mh.Shatd = (const T*)Shatd;
  if (wl_model(P))
    stage_scalar_wl<T>(c, (const T*)vel[0], (const T*)vel[1], (const T*)vel[2], 1, (const T*)TH,
                       (const T*)dT[0], (const T*)dT[1], (const T*)dT[2], mh, nullptr, nullptr,
                       nullptr, (T*)CsPr, prof, beta);
  else
    stage_scalar_lasdd<T>(c, (const T*)vel[0], (const T*)vel[1], (const T*)vel[2], 1,
                          (const T*)TH, (const T*)dT[0], (const T*)dT[1], (const T*)dT[2],
                          (const T*)S, mh, nullptr, nullptr, nullptr, (T*)CsPr, prof, beta);
  return 0;
}

template <class T>
int sgs_stresses(alfa_plan* P, hipStream_t s, const void* const g[9], const void* Cs2,
                 const void* u, const void* v, const void* M, const void* pm, const void* pm0,
                 void* const tau[6]) {
  Ctx<T> c(P, s);
  const T* gg[9];
  for (int n = 0; n < 9; ++n) gg[n] = (const T*)g[n];
  SgsStrains<T> ss{};
  stage_strains<T>(c, gg, ss);
  T* t[6];
  for (int n = 0; n < 6; ++n) t[n] = (T*)tau[n];
  stage_stresses<T>(c, ss, (const T*)Cs2, (const T*)u, (const T*)v, (const T*)M, (const T*)pm,
                    (const T*)pm0, t);
  return 0;
}

template <class T>
int scalar_fluxes(alfa_plan* P, hipStream_t s, const void* const g[9], const void* const dT[3],
                  const void* CsPr, const void* qsfc, void* const q[3]) {
  Ctx<T> c(P, s);
  const T* gg[9];
  for (int n = 0; n < 9; ++n) gg[n] = (const T*)g[n];
  SgsStrains<T> ss{};
  stage_strains<T>(c, gg, ss);
  const T* d3[3] = {(const T*)dT[0], (const T*)dT[1], (const T*)dT[2]};
  T* q3[3] = {(T*)q[0], (T*)q[1], (T*)q[2]};
  stage_fluxes<T>(c, ss, d3, (const T*)CsPr, (const T*)qsfc, q3);
  return 0;
}

// DivStressDynamicSGS / DivStressStaticSGS (NSE_SGSTerms.py:113,186): the
// dynamic coefficient (when `dynamic`), the stresses and their divergence.
template <class T>
int div_stress_sgs(alfa_plan* P, hipStream_t s, const void* const g[9], const void* u,
                   const void* v, const void* w, const void* M, const void* pm, const void* pm0,
                   int dynamic, void* Cs2, double* a1, double* a2, double* beta,
                   void* const divt[3]) {
  Ctx<T> c(P, s);
  const T* gg[9];
  for (int n = 0; n < 9; ++n) gg[n] = (const T*)g[n];
  SgsStrains<T> ss{};
  stage_strains<T>(c, gg, ss);
  if (dynamic && wl_model(P))
    stage_lasdd_wl<T>(c, (const T*)u, (const T*)v, (const T*)w, ss.Sij, nullptr, nullptr, (T*)Cs2,
                      a1, a2, beta);
  else if (dynamic)
    stage_lasdd<T>(c, (const T*)u, (const T*)v, (const T*)w, ss.Sij, ss.Sm, nullptr, nullptr,
                   (T*)Cs2, a1, a2, beta);
  T* tau[6] = {c.W("txx"), c.W("tyy"), c.W("tzz"), c.W("txy"), c.W("txz"), c.W("tyz")};
  stage_stresses<T>(c, ss, (const T*)Cs2, (const T*)u, (const T*)v, (const T*)M, (const T*)pm,
                    (const T*)pm0, tau);
  stage_div_stress<T>(c, tau[0], tau[1], tau[2], tau[3], tau[4], tau[5], (T*)divt[0],
                      (T*)divt[1], (T*)divt[2]);
  return 0;
}

// DivFluxDynamicSGS / DivFluxStaticSGS (SCL_SGSTerms.py:97,169).  The dynamic
// scalar procedure reuses the momentum procedure's filtered velocities and
// |S^| (DynamicSGS_Main.py:326-331), which stay in the plan's workspace from
// the preceding div_stress_sgs call on the same plan.
template <class T>
int div_flux_sgs(alfa_plan* P, hipStream_t s, const void* const g[9], const void* u,
                 const void* v, const void* w, const void* TH, const void* const dT[3],
                 const void* qsfc, int dynamic, void* CsPr, double* prof, double* beta,
                 void* qz, void* divq) {
  Ctx<T> c(P, s);
  const T* gg[9];
  for (int n = 0; n < 9; ++n) gg[n] = (const T*)g[n];
  SgsStrains<T> ss{};
  stage_strains<T>(c, gg, ss);
  const T* d3[3] = {(const T*)dT[0], (const T*)dT[1], (const T*)dT[2]};
  if (dynamic && wl_model(P))
    stage_scalar_wl<T>(c, (const T*)u, (const T*)v, (const T*)w, 0, (const T*)TH, d3[0], d3[1],
                       d3[2], step_hats(c), nullptr, nullptr, nullptr, (T*)CsPr, prof, beta);
  else if (dynamic)
    stage_scalar_lasdd<T>(c, (const T*)u, (const T*)v, (const T*)w, 0, (const T*)TH, d3[0],
                          d3[1], d3[2], ss.Sm, step_hats(c), nullptr, nullptr, nullptr,
                          (T*)CsPr, prof, beta);
  T* q[3] = {c.W("qx"), c.W("qy"), (T*)qz};
  stage_fluxes<T>(c, ss, d3, (const T*)CsPr, (const T*)qsfc, q);
  stage_div_flux<T>(c, q[0], q[1], q[2], (T*)divq);
  return 0;
}

template <class T>
int surface(alfa_plan* P, hipStream_t s, const void* u, const void* v, const void* TH,
            double sfc_val, void* pm, void* pm0, void* ph, void* ph0, void* fm, void* fh,
            void* M, void* ustar, void* qz, void* invOB, double* qavg) {
  Ctx<T> c(P, s);
  StateT<T> S{};
  S.u = (T*)u; S.v = (T*)v; S.TH = (T*)TH;
  S.psim = (T*)pm; S.psim0 = (T*)pm0; S.psih = (T*)ph; S.psih0 = (T*)ph0; S.fim = (T*)fm; S.fih = (T*)fh;
  Surf<T> sf{(T*)M, (T*)ustar, (T*)qz, (T*)invOB};
  stage_surface<T>(c, S, sf, sfc_val);
  if (live(P) && qavg)
    (void)hipMemcpyAsync(qavg, P->scal + 1, sizeof(double), hipMemcpyDeviceToDevice, s);
  return 0;
}

}  // namespace ops

// Run every operation once in dry mode so that its named buffers are planned.
template <class T> static void plan_dry_run_once(alfa_plan* P) {
  hipStream_t s = nullptr;
  void* f = (void*)(uintptr_t)0x1000;
  alfa_state st{};
  st.u = st.v = st.w = st.TH = f;
  st.RHS_u = st.RHS_v = st.RHS_w = st.RHS_TH = f;
  st.Cs2_3D = st.Cs2PrRatio_3D = f;
  st.psi_m = st.psi_m0 = st.psi_h = st.psi_h0 = st.fi_m = st.fi_h = f;
  if (P->cfg.optMoisture >= 1) {
    st.Q = st.RHS_Q = f;
    st.Q_sfc_val = (const double*)f;
  }
  step_t<T>(P, &st, P->cfg.istep, 0.0, s);
  if (P->cfg.dynamicSGS_call_time > 1) step_t<T>(P, &st, P->cfg.istep + 1, 0.0, s);
  ops::fft<T>(P, s, f, f);
  ops::fft_pad<T>(P, s, f, f);
  ops::filtering_explicit<T>(P, s, f, f, f);
  ops::filtering_level<T>(P, s, f, f, 1);
  ops::derivxy<T>(P, s, f, 0, f);
  ops::dealias1_op<T>(P, s, f, f);
  ops::dealias2_op<T>(P, s, f, f);
  ops::advection<T>(P, s, f, f, f, f, f, f, f, f, f, f, f, f);
  ops::scalar_advection<T>(P, s, f, f, f, f, f, f, f);
  ops::buoyancy<T>(P, s, f, f);
  ops::planar_mean<T>(P, s, f, (double*)f);
  ops::rhs_momentum<T>(P, s, f, f, f, f, f, f, f, f, f, f, f, f, f);
  ops::rhs_scalar<T>(P, s, f, f, f, f);
  const void* three[3] = {f, f, f};
  ops::pressure_rc<T>(P, s, three, three, three, f, f, f, f);
  ops::thomas_solve<T>(P, s, f, f, f, f, f, f, f);
  Ctx<T> c(P, s);
  stage_div_stress<T>(c, (T*)f, (T*)f, (T*)f, (T*)f, (T*)f, (T*)f, (T*)f, (T*)f, (T*)f);
  stage_div_flux<T>(c, (T*)f, (T*)f, (T*)f, (T*)f);
  const void* seven[7] = {f, f, f, f, f, f, f};
  const void* nine[9] = {f, f, f, f, f, f, f, f, f};
  void* six[6] = {f, f, f, f, f, f};
  void* three_o[3] = {f, f, f};
  const int o = P->cfg.optSgs;
  if (o == 1 || o == 3) ops::lasdd<T>(P, s, f, f, f, seven, f, nullptr, nullptr, nullptr);
  if (o == 2 || o == 4) ops::lasdd_wl<T>(P, s, f, f, f, seven, f, nullptr, nullptr, nullptr);
  if (o >= 1 && o <= 4)
    ops::scalar_lasdd<T>(P, s, three, three, three, f, three, f, f, f, f, nullptr, nullptr);
  ops::sgs_stresses<T>(P, s, nine, f, f, f, f, f, f, six);
  stats_accumulate<T>(P, s, &st, (double*)f);
  ops::div_stress_sgs<T>(P, s, nine, f, f, f, f, f, f, o >= 1 && o <= 4, f, nullptr, nullptr,
                         nullptr, three_o);
  ops::div_flux_sgs<T>(P, s, nine, f, f, f, f, three, f, o >= 1 && o <= 4, f, nullptr, nullptr, f,
                       f);
  ops::scalar_fluxes<T>(P, s, nine, three, f, f, three_o);
}

// Every buffer of the step and the operations, under both settings of
// alfa_plan_set_fusion (the fused and the separate LAD/LASDD kernels use
// different scratch), so the setting can change at any time.
template <class T> static void plan_dry_run(alfa_plan* P) {
  const int fuse = P->fuse_contract;
  for (int f : {1, 0}) {
    P->fuse_contract = f;
    plan_dry_run_once<T>(P);
  }
  P->fuse_contract = fuse;
}

template <class T> static int allocate_workspace(alfa_plan* P) {
  P->ws.dry = true;
  plan_dry_run<T>(P);
  size_t total = 0;
  for (auto& kv : P->ws.bufs) {
    const size_t bytes = (kv.second.first + 255) & ~(size_t)255;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess)
      return set_error(ALFA_ENOMEM, "hipMalloc(%zu) for '%s': %s", bytes, kv.first.c_str(),
                       hipGetErrorString(e));
    P->allocs.push_back(p);
    kv.second.second = p;
    total += bytes;
  }
  P->d.workspace_bytes = (int64_t)total;
  P->ws.dry = false;
  return ALFA_OK;
}

static void free_prof(alfa_plan* P) {
  for (hipEvent_t e : P->pf.pool) (void)hipEventDestroy(e);
  P->pf = Prof();
}

static void free_plan(alfa_plan* P) {
  if (!P) return;
  free_prof(P);
  for (void* p : P->allocs) (void)hipFree(p);
  delete P;
}

extern "C" int alfa_plan_create(const alfa_config* cfg, int precision, alfa_plan** out) {
  if (!cfg || !out) return set_error(ALFA_EINVAL, "alfa_plan_create: null argument");
  *out = nullptr;
  if (precision != ALFA_FP64 && precision != ALFA_MIXED)
    return set_error(ALFA_EINVAL, "unknown precision %d", precision);
  const int nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
  if (nx < 8 || ny < 8 || nz < 4 || (nx & 1) || (ny & 1) || (nz & 1))
    return set_error(ALFA_EINVAL, "grid %dx%dx%d: nx, ny >= 8 and nz >= 4, all even", nx, ny, nz);
  if (nz + 1 > TH_MAXN) return set_error(ALFA_EINVAL, "nz=%d exceeds %d", nz, TH_MAXN - 1);
  if (cfg->FGR != 1 && cfg->FGR != 2) return set_error(ALFA_EINVAL, "FGR must be 1 or 2");
  if (cfg->optSgs < 0 || cfg->optSgs > 4)
    return set_error(ALFA_EINVAL,
                     "optSgs=%d: supported are 0 (static SM), 1 (LASDD-SM), 2 (LASDD-WL), "
                     "3 (LAD-SM), 4 (LAD-WL)",
                     cfg->optSgs);
  if (cfg->optSurfBC < 0 || cfg->optSurfBC > 2) return set_error(ALFA_EINVAL, "optSurfBC must be 0..2");
  if (cfg->optMoisture < 0 || cfg->optMoisture > 1)
    return set_error(ALFA_EINVAL, "optMoisture must be 0 or 1");
  if (cfg->optMoistureSurfBC < 0 || cfg->optMoistureSurfBC > 2)
    return set_error(ALFA_EINVAL, "optMoistureSurfBC must be 0..2");
  alfa_plan* P = new alfa_plan();
  P->cfg = *cfg;
  if (P->cfg.istep == 0) P->cfg.istep = 1;
  P->prec = precision;
  P->tile_x = precision == ALFA_FP64 ? ALFA_X_BS_F64 : ALFA_X_BS;
  P->tile_c2r = ALFA_C2R_BS;
  P->tile_r2c = ALFA_R2C_BS;
  P->nx = nx; P->ny = ny; P->nz = nz; P->nyc = ny / 2 + 1;
  P->mx = (int)(1.5 * nx); P->my = (int)(1.5 * ny); P->myc = P->my / 2 + 1;
  alfa_derived& d = P->d;
  memset(&d, 0, sizeof d);
  d.z_scale = cfg->l_x / (2 * M_PI);
  d.g_nondim = 9.81 * d.z_scale;
  d.f_coriolis_nondim = cfg->f_coriolis * d.z_scale;
  d.inversion_nondim = cfg->inversion * d.z_scale;
  d.dx = 2 * M_PI / nx;
  d.dy = 2 * M_PI / ny;
  d.dz = (cfg->l_z / d.z_scale) / (nz - 1);
  d.L = cfg->FGR * std::pow(d.dx * d.dy * d.dz, 1.0 / 3.0);
  d.dt_nondim = cfg->dt / d.z_scale;
  d.TFR = cfg->FGR == 1 ? 2.0 : std::sqrt(2.0);
  d.optDealias = cfg->FGR == 1 ? 1 : 0;
  d.z_damping_nondim = cfg->z_damping / d.z_scale;
  d.RelaxTime_nondim = cfg->RelaxTime / d.z_scale;
  d.mx = P->mx; d.my = P->my; d.nx_rfft = nx / 2 + 1; d.ny_rfft = ny / 2 + 1;
  // cut-offs: Python round() (half to even) of Filtering.py:73-74,111-112,149-150
  auto pyround = [](double v) { return (int)std::nearbyint(v); };
  d.cut0_x = pyround(nx * (1.0 / (2.0 * cfg->FGR)));
  d.cut0_y = pyround(ny * (1.0 / (2.0 * cfg->FGR)));
  d.cut1_x = pyround(nx / (2 * cfg->FGR * d.TFR));
  d.cut1_y = pyround(ny / (2 * cfg->FGR * d.TFR));
  d.cut2_x = pyround(nx / (2 * cfg->FGR * d.TFR * d.TFR));
  d.cut2_y = pyround(ny / (2 * cfg->FGR * d.TFR * d.TFR));
  d.precision = precision;
  if (!factor(nx, P->fx) || !factor(ny, P->fy) || !factor(P->mx, P->fmx) || !factor(P->my, P->fmy)) {
    free_plan(P);
    return set_error(ALFA_ERADIX, "FFT lengths %d, %d, %d, %d must factor into 2, 3, 5", nx, ny,
                     P->mx, P->my);
  }
  const int maxn = d.optDealias ? std::max(P->mx, P->my) : std::max(nx, ny);
  if (maxn > FFT_MAXN) {
    free_plan(P);
    return set_error(ALFA_EINVAL, "transform length %d exceeds %d", maxn, FFT_MAXN);
  }
  hipError_t e = hipSuccess;
  {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
        prop.multiProcessorCount > 0)
      P->ncu = prop.multiProcessorCount;
  }
  auto tw = [&](int n) -> void* {
    return precision == ALFA_FP64 ? make_twiddles<double>(n, P->allocs, e)
                                  : make_twiddles<float>(n, P->allocs, e);
  };
  P->tw_x = tw(nx);
  if (e == hipSuccess) P->tw_y = tw(ny);
  if (e == hipSuccess) P->tw_mx = tw(P->mx);
  if (e == hipSuccess) P->tw_my = tw(P->my);
  // Thomas factors (ThomasPressureInit, NSE_PressureTerms_Thomas.py:151-214)
  // the pivots depend on kx only through kx^2: one row per (|kx|, ky)
  const int N = nz + 1, nyc = P->nyc;
  const size_t ncolc = (size_t)(nx / 2 + 1) * nyc;
  std::vector<double> hb(ncolc * N, 0.0);   // pivots; the kernel forms the multipliers
  {
    const double idz2 = 1.0 / (d.dz * d.dz);
    std::vector<double> a(N, idz2), c(N, idz2), b(N);
    a[0] = 0; a[nz] = -1; c[0] = 1; c[nz] = 0;
    for (int i = 0; i <= nx / 2; ++i)
      for (int j = 0; j < nyc; ++j) {
        const double kr = i, kc = j;
        const double bb = -(kr * kr + kc * kc + 2.0 / (d.dz * d.dz));
        for (int k = 0; k < N; ++k) b[k] = bb;
        if (i == 0 && j == 0)
          for (int k = 1; k < nz; ++k) b[k] = -(1.0 + 2.0 / (d.dz * d.dz));
        b[0] = -1.0; b[nz] = 1.0;
        double* pb = &hb[((size_t)i * nyc + j) * N];
        pb[0] = b[0];
        for (int k = 1; k < N; ++k) {
          const double m = a[k] / pb[k - 1];
          pb[k] = b[k] - m * c[k - 1];
        }
      }
  }
  auto upload = [&](const std::vector<double>& h) -> double* {
    void* p = nullptr;
    if (e != hipSuccess) return nullptr;
    e = hipMalloc(&p, h.size() * sizeof(double));
    if (e != hipSuccess) return nullptr;
    P->allocs.push_back(p);
    e = hipMemcpy(p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice);
    return (double*)p;
  };
  P->th_b = upload(hb);
  // profiles: Ug, Vg (Initialization.py:163-170), Rayleigh full / staggered (:395-460)
  {
    std::vector<double> pr(8 * (size_t)nz, 0.0);
    const double inv = 1.0 / d.RelaxTime_nondim;
    const double ztop = cfg->l_z / d.z_scale;
    const double thick = ztop - d.z_damping_nondim;
    for (int k = 0; k < nz; ++k) {
      pr[k] = cfg->Ug2;
      pr[nz + k] = cfg->Vg2;
      const double zf = k * d.dz, zs = (k + 0.5) * d.dz;
      pr[2 * nz + k] = (zf >= d.z_damping_nondim && zf <= ztop)
                           ? 0.5 * inv * (1.0 - std::cos(M_PI * (zf - d.z_damping_nondim) / thick))
                           : 0.0;
      pr[3 * nz + k] = (zs >= d.z_damping_nondim && zs <= ztop)
                           ? 0.5 * inv * (1.0 - std::cos(M_PI * (zs - d.z_damping_nondim) / thick))
                           : 0.0;
    }
    P->prof = upload(pr);
    std::vector<double> z8(8, 0.0);
    P->scal = upload(z8);
  }
  if (e != hipSuccess) {
    int r = set_error(ALFA_EHIP, "plan setup: %s", hipGetErrorString(e));
    free_plan(P);
    return r;
  }
  int r = precision == ALFA_FP64 ? allocate_workspace<double>(P) : allocate_workspace<float>(P);
  if (r != ALFA_OK) {
    free_plan(P);
    return r;
  }
  *out = P;
  return ALFA_OK;
}

extern "C" int alfa_plan_destroy(alfa_plan* P) {
  if (P) (void)hipDeviceSynchronize();
  free_plan(P);
  return ALFA_OK;
}

extern "C" int alfa_plan_derived(const alfa_plan* P, alfa_derived* out) {
  if (!P || !out) return set_error(ALFA_EINVAL, "null argument");
  *out = P->d;
  return ALFA_OK;
}

#define DISPATCH(P, where, ...)                                               \
  do {                                                                        \
    if (!(P)) return set_error(ALFA_EINVAL, "%s: null plan", where);         \
    if ((P)->prec == ALFA_FP64) { using T = double; __VA_ARGS__; }            \
    else { using T = float; __VA_ARGS__; }                                    \
    return check_hip(P, where);                                               \
  } while (0)

#define STREAM(s) ((hipStream_t)(s))

extern "C" int alfa_state_init(alfa_plan* P, const alfa_state* st, alfa_stream stream) {
  if (!P || !st) return set_error(ALFA_EINVAL, "alfa_state_init: null argument");
  hipStream_t s = STREAM(stream);
  const size_t F = (size_t)P->nx * P->ny * P->nz, pl = (size_t)P->nx * P->ny;
  DISPATCH(P, "alfa_state_init", {
    void* zero3[4] = {st->RHS_u, st->RHS_v, st->RHS_w, st->RHS_TH};
    for (void* z : zero3) launch_k<1024>(P, s, dim3(pw_blocks(F)), dim3(256), 0, KBODY(k_fill<T>), F, (T*)z, T(0));
    launch_k<1024>(P, s, dim3(pw_blocks(F)), dim3(256), 0, KBODY(k_fill<T>), F, (T*)st->Cs2_3D, (T)P->cfg.Cs2);
    launch_k<1024>(P, s, dim3(pw_blocks(F)), dim3(256), 0, KBODY(k_fill<T>), F, (T*)st->Cs2PrRatio_3D, (T)P->cfg.Cs2PrRatio);
    void* zero2[4] = {st->psi_m, st->psi_m0, st->psi_h, st->psi_h0};
    for (void* z : zero2) launch_k<1024>(P, s, dim3(pw_blocks(pl)), dim3(256), 0, KBODY(k_fill<T>), pl, (T*)z, T(0));
    launch_k<1024>(P, s, dim3(pw_blocks(pl)), dim3(256), 0, KBODY(k_fill<T>), pl, (T*)st->fi_m, T(1));
    launch_k<1024>(P, s, dim3(pw_blocks(pl)), dim3(256), 0, KBODY(k_fill<T>), pl, (T*)st->fi_h, T(1));
    if (P->cfg.optMoisture >= 1 && st->RHS_Q)   // RHS_Q_previous = 0 (Main.py:73)
      launch_k<1024>(P, s, dim3(pw_blocks(F)), dim3(256), 0, KBODY(k_fill<T>), F, (T*)st->RHS_Q, T(0));
    launch_ok(P);
  });
}

// the moisture fields and surface value a moisture plan needs
static int check_moist_state(const alfa_plan* P, const alfa_state* st, const char* where) {
  if (!P || P->cfg.optMoisture < 1) return ALFA_OK;
  if (!st->Q || !st->RHS_Q)
    return set_error(ALFA_EINVAL, "%s: optMoisture needs state Q and RHS_Q", where);
  if (P->cfg.optMoistureSurfBC >= 1 && !st->Q_sfc_val)
    return set_error(ALFA_EINVAL, "%s: optMoistureSurfBC=%d needs state Q_sfc_val", where,
                     P->cfg.optMoistureSurfBC);
  return ALFA_OK;
}

extern "C" int alfa_step(alfa_plan* P, const alfa_state* st, int iteration, double sfc_val,
                         alfa_stream stream) {
  if (!st) return set_error(ALFA_EINVAL, "alfa_step: null state");
  if (int rc = check_moist_state(P, st, "alfa_step")) return rc;
  DISPATCH(P, "alfa_step", step_t<T>(P, st, iteration, sfc_val, STREAM(stream)));
}
extern "C" int alfa_step_sfc_dev(alfa_plan* P, const alfa_state* st, int iteration,
                                 const double* sfc_val, alfa_stream stream) {
  if (!P || !st || !sfc_val) return set_error(ALFA_EINVAL, "alfa_step_sfc_dev: null argument");
  if (int rc = check_moist_state(P, st, "alfa_step_sfc_dev")) return rc;
  DISPATCH(P, "alfa_step_sfc_dev",
           step_t<T>(P, st, iteration, 0.0, STREAM(stream), sfc_val));
}

// Tile choices of the compile-time-length FFT kernels: threads per workgroup
// of the y R2C, x and y C2R kernels (256 or 512; lines per tile scale with
// them).  Each line is transformed by the same sequence of operations at
// either size, so the choice changes timing only, never the results.
static bool tile_ok(int t) { return t == 256 || t == 512; }

extern "C" int alfa_plan_set_tiles(alfa_plan* P, const int tiles[3]) {
  if (!P || !tiles) return set_error(ALFA_EINVAL, "alfa_plan_set_tiles: null argument");
  for (int i = 0; i < 3; ++i)
    if (tiles[i] != 0 && !tile_ok(tiles[i]))
      return set_error(ALFA_EINVAL, "alfa_plan_set_tiles: tile %d is %d (0, 256 or 512)", i,
                       tiles[i]);
  if (tiles[0]) P->tile_r2c = tiles[0];
  if (tiles[1]) P->tile_x = tiles[1];
  if (tiles[2]) P->tile_c2r = tiles[2];
  return ALFA_OK;
}

extern "C" int alfa_plan_set_fusion(alfa_plan* P, int on) {
  if (!P) return set_error(ALFA_EINVAL, "alfa_plan_set_fusion: null plan");
  P->fuse_contract = on ? 1 : 0;
  return ALFA_OK;
}

extern "C" int alfa_plan_get_fusion(const alfa_plan* P, int* on) {
  if (!P || !on) return set_error(ALFA_EINVAL, "alfa_plan_get_fusion: null argument");
  *on = P->fuse_contract;
  return ALFA_OK;
}

extern "C" int alfa_plan_get_tiles(const alfa_plan* P, int tiles[3]) {
  if (!P || !tiles) return set_error(ALFA_EINVAL, "alfa_plan_get_tiles: null argument");
  tiles[0] = P->tile_r2c;
  tiles[1] = P->tile_x;
  tiles[2] = P->tile_c2r;
  return ALFA_OK;
}

namespace {
// What the autotuner times: whole steps on a scratch copy of a state when
// the caller gives one, else one forward transform and three inverses (x
// derivative over all modes, the two test-filter levels with their band
// cut-offs) of a zero field.
template <class T>
struct TuneWork {
  alfa_plan* P;
  hipStream_t s;
  const alfa_state* st = nullptr;   // scratch copy, or null
  int iteration = 1, reps = 3;
  double sfc_val = 0.0;
  void *F = nullptr, *Ffft = nullptr, *out = nullptr;

  int once(int r) {
    if (st) return step_t<T>(P, st, iteration + r, sfc_val, s);
    ops::fft<T>(P, s, F, Ffft);
    ops::derivxy<T>(P, s, Ffft, 0, out);
    ops::filtering_level<T>(P, s, Ffft, out, 1);
    ops::filtering_level<T>(P, s, Ffft, out, 2);
    return 0;
  }
  // reps after one untimed call (the first launch of a tile choice), ms
  int time(hipEvent_t e0, hipEvent_t e1, float& ms) {
    if (int rc = once(0)) return rc;
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < reps; ++r)
      if (int rc = once(r)) return rc;
    (void)hipEventRecord(e1, s);
    hipError_t e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess)
      return set_error(ALFA_EHIP, "alfa_plan_autotune: %s", hipGetErrorString(e));
    return ALFA_OK;
  }
};

template <class T>
int autotune_tiles(alfa_plan* P, hipStream_t s, const alfa_state* st, int iteration,
                   double sfc_val) {
  bool ct_x = false, ct_y = false;
  ct_length(P->nx, [&](auto nct, auto) { ct_x = decltype(nct)::value > 0; });
  ct_length(P->ny, [&](auto nct, auto) { ct_y = decltype(nct)::value > 0; });
  if (!ct_x && !ct_y) return ALFA_OK;   // generic kernels only: nothing to choose
  const size_t nf = (size_t)P->nx * P->ny * P->nz * sizeof(T);
  const size_t npl = (size_t)P->nx * P->ny * sizeof(T);
  std::vector<void*> bufs;
  hipError_t e = hipSuccess;
  auto alloc = [&](size_t n) -> void* {
    void* p = nullptr;
    if (e == hipSuccess) e = hipMalloc(&p, n);
    if (e == hipSuccess) bufs.push_back(p);
    return e == hipSuccess ? p : nullptr;
  };
  TuneWork<T> w{P, s};
  alfa_state cp{};
  if (st) {
    // the step updates its state in place: time it on a copy
    cp = *st;
    cp.Cs2_1D_avg1 = cp.Cs2_1D_avg2 = cp.beta1_1D = cp.Cs2PrRatio_1D = cp.beta2_1D = nullptr;
    cp.scalars = nullptr;
    void** vol[] = {&cp.u, &cp.v, &cp.w, &cp.TH, &cp.RHS_u, &cp.RHS_v, &cp.RHS_w, &cp.RHS_TH,
                    &cp.Cs2_3D, &cp.Cs2PrRatio_3D, &cp.Q, &cp.RHS_Q};
    void** pl[] = {&cp.psi_m, &cp.psi_m0, &cp.psi_h, &cp.psi_h0, &cp.fi_m, &cp.fi_h};
    const int nvol = P->cfg.optMoisture >= 1 ? 12 : 10;
    for (int i = 0; i < nvol; ++i) {
      void* src = *vol[i];
      *vol[i] = alloc(nf);
      if (e == hipSuccess) e = hipMemcpyAsync(*vol[i], src, nf, hipMemcpyDeviceToDevice, s);
    }
    for (void** f : pl) {
      void* src = *f;
      *f = alloc(npl);
      if (e == hipSuccess) e = hipMemcpyAsync(*f, src, npl, hipMemcpyDeviceToDevice, s);
    }
    w.st = &cp;
    w.iteration = iteration;
    w.sfc_val = sfc_val;
    // cover a whole cycle of dynamic and static coefficient steps
    w.reps = std::min(8, std::max(3, P->cfg.dynamicSGS_call_time));
  } else {
    const size_t ns = (size_t)P->nx * P->nyc * P->nz * sizeof(c2<T>);
    w.F = alloc(nf);
    w.Ffft = alloc(ns);
    w.out = alloc(nf);
    if (e == hipSuccess) e = hipMemsetAsync(w.F, 0, nf, s);
    w.reps = 5;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  int rc = ALFA_OK;
  if (e == hipSuccess) {
    // one knob at a time (each moves one kernel), in pipeline order; each
    // choice timed twice, interleaved, and the sums compared
    int* knob[3] = {ct_y ? &P->tile_r2c : nullptr, ct_x ? &P->tile_x : nullptr,
                    ct_y ? &P->tile_c2r : nullptr};
    for (int k = 0; k < 3 && rc == ALFA_OK; ++k) {
      if (!knob[k]) continue;
      const int keep = *knob[k];
      float t[2] = {0.f, 0.f};
      for (int trial = 0; trial < 2 && rc == ALFA_OK; ++trial)
        for (int c = 0; c < 2 && rc == ALFA_OK; ++c) {
          *knob[k] = c ? 512 : 256;
          float ms = 0.f;
          rc = w.time(e0, e1, ms);
          t[c] += ms;
        }
      // the other size must win by 1%: whole steps at either size are
      // often within run-to-run noise, and the kept choice then does not
      // flip from run to run
      const float t_keep = t[keep == 512 ? 1 : 0], t_other = t[keep == 512 ? 0 : 1];
      *knob[k] = (rc == ALFA_OK && t_other < 0.99f * t_keep) ? (keep == 512 ? 256 : 512) : keep;
    }
    e = hipStreamSynchronize(s);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (void* p : bufs) (void)hipFree(p);
  if (rc != ALFA_OK) return rc;
  if (e != hipSuccess)
    return set_error(ALFA_EHIP, "alfa_plan_autotune: %s", hipGetErrorString(e));
  return ALFA_OK;
}
}  // namespace

extern "C" int alfa_plan_autotune(alfa_plan* P, const alfa_state* st, int iteration,
                                  double sfc_val, alfa_stream stream, int chosen[3]) {
  if (!P) return set_error(ALFA_EINVAL, "alfa_plan_autotune: null plan");
  if (st)
    if (int rc = check_moist_state(P, st, "alfa_plan_autotune")) return rc;
  hipStream_t s = STREAM(stream);
  const int rc = P->prec == ALFA_FP64 ? autotune_tiles<double>(P, s, st, iteration, sfc_val)
                                      : autotune_tiles<float>(P, s, st, iteration, sfc_val);
  if (rc) return rc;
  if (chosen) alfa_plan_get_tiles(P, chosen);
  return check_hip(P, "alfa_plan_autotune");
}

extern "C" int alfa_struct_sizes(size_t out[3]) {
  if (!out) return set_error(ALFA_EINVAL, "alfa_struct_sizes: null argument");
  out[0] = sizeof(alfa_config);
  out[1] = sizeof(alfa_state);
  out[2] = sizeof(alfa_derived);
  return ALFA_OK;
}

extern "C" int alfa_set_forcing(alfa_plan* P, const double* Ug, const double* Vg,
                                const double* Uadv, const double* Vadv, const double* THadv,
                                const double* Qadv, alfa_stream stream) {
  if (!P) return set_error(ALFA_EINVAL, "alfa_set_forcing: null plan");
  if (!Ug && !Vg && !Uadv && !Vadv && !THadv && !Qadv) return ALFA_OK;
  hipStream_t s = STREAM(stream);
  launch_k<1024>(P, s, dim3((P->nz + 255) / 256), dim3(256), 0, KBODY(k_set_forcing), P->nz, Ug, Vg, Uadv, Vadv, THadv, Qadv,
                                                     P->prof);
  launch_ok(P);
  if (Uadv || Vadv || THadv || Qadv) P->adv_on = true;
  return check_hip(P, "alfa_set_forcing");
}

static_assert(ALFA_STATS_NPROF == STATS_NPROF && ALFA_STATS_NSCAL == STATS_NSCAL,
              "statistics layout");
extern "C" int alfa_stats_accumulate(alfa_plan* P, const alfa_state* st, double* acc,
                                     alfa_stream stream) {
  if (!P || !st || !acc) return set_error(ALFA_EINVAL, "alfa_stats_accumulate: null argument");
  DISPATCH(P, "alfa_stats_accumulate", stats_accumulate<T>(P, STREAM(stream), st, acc));
}

extern "C" int alfa_batch_step(alfa_plan* const* plans, const alfa_state* states, int n,
                               const int* iterations, const double* sfc_vals,
                               alfa_stream stream) {
  if (n < 0 || (n > 0 && (!plans || !states || !iterations)))
    return set_error(ALFA_EINVAL, "alfa_batch_step: null member table");
  for (int m = 0; m < n; ++m) {
    alfa_plan* P = plans[m];
    if (!P) return set_error(ALFA_EINVAL, "alfa_batch_step: member %d has no plan", m);
    if (int rc = check_moist_state(P, &states[m], "alfa_batch_step")) return rc;
    const double sv = sfc_vals ? sfc_vals[m] : 0.0;
    int rc = ALFA_OK;
    if (P->prec == ALFA_FP64) step_t<double>(P, &states[m], iterations[m], sv, STREAM(stream));
    else step_t<float>(P, &states[m], iterations[m], sv, STREAM(stream));
    rc = check_hip(P, "alfa_batch_step");
    if (rc) return rc;
  }
  return ALFA_OK;
}

// ---------------------------------------------------------------------------
// Member groups: runs of one configuration shape stepped by shared launches.
// Each member's step is recorded once per step kind (first step; dynamic or
// static SGS step; forcing on or off; FFT tiles) and the records of all
// members must match launch for launch.  The packed arguments of every
// launch are laid out member by member in one device table, uploaded once
// per kind; a group step then enqueues each kernel once, as k_member over
// the group (blockIdx.z = member), with no host-device traffic, so it can be
// captured into a hipGraph after one step of each kind.
// ---------------------------------------------------------------------------
struct GroupProgram {
  std::vector<LaunchRec> recs;   // member 0's records (identity, shape, launcher)
  std::vector<size_t> off;       // table offset of each launch
  void* dev = nullptr;
};

struct alfa_group {
  std::vector<alfa_plan*> plans;
  std::vector<alfa_state> states;
  std::vector<const double*> sfc;
  int prec = 0;
  std::map<long long, GroupProgram> progs;
};

static bool step_dynamic(const alfa_plan* P, int it) {
  const auto& g = P->cfg;
  return (g.optSgs >= 1 && g.optSgs <= 4) &&
         (it == g.istep || (g.dynamicSGS_call_time > 0 && it % g.dynamicSGS_call_time == 0));
}

static long long step_kind(const alfa_plan* P, int it) {
  long long k = (it == P->cfg.istep ? 1 : 0) | (step_dynamic(P, it) ? 2 : 0) |
                (P->adv_on ? 4 : 0);
  k |= (long long)P->tile_r2c << 3;
  k |= (long long)P->tile_x << 13;
  k |= (long long)P->tile_c2r << 23;
  k |= (long long)P->fuse_contract << 33;
  return k;
}

static void free_group(alfa_group* G) {
  if (!G) return;
  for (auto& kv : G->progs)
    if (kv.second.dev) (void)hipFree(kv.second.dev);
  delete G;
}

extern "C" int alfa_group_create(alfa_plan* const* plans, const alfa_state* states,
                                 const double* const* sfc_vals, int n, alfa_group** out) {
  if (!out || n <= 0 || !plans || !states)
    return set_error(ALFA_EINVAL, "alfa_group_create: empty or null member table");
  *out = nullptr;
  alfa_group* G = new alfa_group();
  G->prec = plans[0] ? plans[0]->prec : 0;
  for (int m = 0; m < n; ++m) {
    alfa_plan* P = plans[m];
    if (!P) { free_group(G); return set_error(ALFA_EINVAL, "alfa_group_create: member %d has no plan", m); }
    if (P->prec != G->prec || P->nx != plans[0]->nx || P->ny != plans[0]->ny ||
        P->nz != plans[0]->nz) {
      free_group(G);
      return set_error(ALFA_EINVAL, "alfa_group_create: member %d differs from member 0 in "
                                    "grid or precision", m);
    }
    if (int rc = check_moist_state(P, &states[m], "alfa_group_create")) { free_group(G); return rc; }
    const double* sv = sfc_vals ? sfc_vals[m] : nullptr;
    if (!sv && P->cfg.optSurfBC != 0) {
      free_group(G);
      return set_error(ALFA_EINVAL, "alfa_group_create: member %d (optSurfBC %d) needs a "
                                    "device surface value", m, P->cfg.optSurfBC);
    }
    G->plans.push_back(P);
    G->states.push_back(states[m]);
    G->sfc.push_back(sv);
  }
  *out = G;
  return ALFA_OK;
}

template <class T>
static int group_record(alfa_group* G, int it, hipStream_t s, GroupProgram& pr) {
  const int M = (int)G->plans.size();
  std::vector<std::vector<LaunchRec>> recs(M);
  for (int m = 0; m < M; ++m) {
    alfa_plan* P = G->plans[m];
    P->rec = &recs[m];
    step_t<T>(P, &G->states[m], it, 0.0, s, G->sfc[m]);
    P->rec = nullptr;
    if (int rc = check_hip(P, "alfa_group_step (recording)")) return rc;
  }
  for (int m = 1; m < M; ++m) {
    bool same = recs[m].size() == recs[0].size();
    for (size_t r = 0; same && r < recs[0].size(); ++r) {
      const LaunchRec &a = recs[0][r], &b = recs[m][r];
      same = a.id == b.id && a.grid.x == b.grid.x && a.grid.y == b.grid.y &&
             a.grid.z == b.grid.z && a.block.x == b.block.x && a.block.y == b.block.y &&
             a.block.z == b.block.z && a.lds == b.lds && a.pack.size() == b.pack.size();
    }
    if (!same)
      return set_error(ALFA_EINVAL, "alfa_group_step: member %d's step launches differ from "
                                    "member 0's (the members' configurations must match in "
                                    "everything but their physical constants)", m);
  }
  // tables: launch r at a 256-byte aligned offset, one pack per member at
  // the pack's own size (its array stride in the kernel)
  size_t total = 0;
  pr.off.resize(recs[0].size());
  for (size_t r = 0; r < recs[0].size(); ++r) {
    total = (total + 255) / 256 * 256;
    pr.off[r] = total;
    total += (size_t)M * recs[0][r].pack.size();
  }
  std::vector<unsigned char> host(total ? total : 1, 0);
  for (size_t r = 0; r < recs[0].size(); ++r) {
    const size_t ps = recs[0][r].pack.size();
    for (int m = 0; m < M; ++m)
      std::memcpy(host.data() + pr.off[r] + (size_t)m * ps, recs[m][r].pack.data(), ps);
  }
  hipError_t e = hipMalloc(&pr.dev, host.size());
  if (e == hipSuccess) e = hipMemcpy(pr.dev, host.data(), host.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (pr.dev) (void)hipFree(pr.dev);
    pr.dev = nullptr;
    return set_error(ALFA_EHIP, "alfa_group_step: member tables: %s", hipGetErrorString(e));
  }
  pr.recs = std::move(recs[0]);
  return ALFA_OK;
}

extern "C" int alfa_group_step(alfa_group* G, int iteration, alfa_stream stream) {
  if (!G) return set_error(ALFA_EINVAL, "alfa_group_step: null group");
  hipStream_t s = STREAM(stream);
  const long long kind = step_kind(G->plans[0], iteration);
  for (size_t m = 1; m < G->plans.size(); ++m)
    if (step_kind(G->plans[m], iteration) != kind)
      return set_error(ALFA_EINVAL, "alfa_group_step: member %zu takes another kind of step "
                                    "at iteration %d (first/dynamic/forcing/tiles)", m, iteration);
  auto it = G->progs.find(kind);
  if (it == G->progs.end()) {
    GroupProgram pr;
    const int rc = G->prec == ALFA_FP64 ? group_record<double>(G, iteration, s, pr)
                                        : group_record<float>(G, iteration, s, pr);
    if (rc != ALFA_OK) return rc;
    it = G->progs.emplace(kind, std::move(pr)).first;
  }
  const GroupProgram& pr = it->second;
  const int M = (int)G->plans.size();
  for (size_t r = 0; r < pr.recs.size(); ++r)
    pr.recs[r].member((const unsigned char*)pr.dev + pr.off[r], M, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(ALFA_EHIP, "alfa_group_step: %s", hipGetErrorString(e));
  return ALFA_OK;
}

extern "C" int alfa_group_launches(const alfa_group* G, int iteration, int* n) {
  if (!G || !n) return set_error(ALFA_EINVAL, "alfa_group_launches: null argument");
  auto it = G->progs.find(step_kind(G->plans[0], iteration));
  *n = it == G->progs.end() ? -1 : (int)it->second.recs.size();
  return ALFA_OK;
}

extern "C" int alfa_group_destroy(alfa_group* G) {
  if (G) (void)hipDeviceSynchronize();
  free_group(G);
  return ALFA_OK;
}

extern "C" int alfa_fft(alfa_plan* P, const void* F, void* out, alfa_stream s) {
  DISPATCH(P, "alfa_fft", ops::fft<T>(P, STREAM(s), F, out));
}
extern "C" int alfa_fft_pad(alfa_plan* P, const void* F, void* out, alfa_stream s) {
  DISPATCH(P, "alfa_fft_pad", ops::fft_pad<T>(P, STREAM(s), F, out));
}
extern "C" int alfa_filtering_explicit(alfa_plan* P, const void* Ffft, void* F, void* Fnew,
                                       alfa_stream s) {
  DISPATCH(P, "alfa_filtering_explicit", ops::filtering_explicit<T>(P, STREAM(s), Ffft, F, Fnew));
}
extern "C" int alfa_filtering_level1(alfa_plan* P, const void* Ffft, void* out, alfa_stream s) {
  DISPATCH(P, "alfa_filtering_level1", ops::filtering_level<T>(P, STREAM(s), Ffft, out, 1));
}
extern "C" int alfa_filtering_level2(alfa_plan* P, const void* Ffft, void* out, alfa_stream s) {
  DISPATCH(P, "alfa_filtering_level2", ops::filtering_level<T>(P, STREAM(s), Ffft, out, 2));
}
extern "C" int alfa_dealias1(alfa_plan* P, const void* Ffft, void* out, alfa_stream s) {
  DISPATCH(P, "alfa_dealias1", ops::dealias1_op<T>(P, STREAM(s), Ffft, out));
}
extern "C" int alfa_dealias2(alfa_plan* P, const void* Fpfft, void* out, alfa_stream s) {
  DISPATCH(P, "alfa_dealias2", ops::dealias2_op<T>(P, STREAM(s), Fpfft, out));
}
extern "C" int alfa_derivxy(alfa_plan* P, const void* Ffft, int axis, void* out, alfa_stream s) {
  if (axis != 0 && axis != 1) return set_error(ALFA_EINVAL, "alfa_derivxy: axis must be 0 or 1");
  DISPATCH(P, "alfa_derivxy", ops::derivxy<T>(P, STREAM(s), Ffft, axis, out));
}
extern "C" int alfa_derivz_m(alfa_plan* P, const void* u, const void* v, const void* w,
                             const void* ustar, const void* M, const void* fim, void* dudz,
                             void* dvdz, void* dwdz, alfa_stream s) {
  DISPATCH(P, "alfa_derivz_m", {
    const size_t F = (size_t)P->nx * P->ny * P->nz;
    PwDims dd{F, P->nz};
    PW_LAUNCH(P, STREAM(s), F, k_derivz<T>, dd, (const T*)u, (const T*)v, (const T*)w,
              (const T*)nullptr, (const T*)ustar, (const T*)M, (const T*)fim, (const T*)nullptr,
              (const T*)nullptr, (T*)dudz, (T*)dvdz, (T*)dwdz, (T*)nullptr, (T)P->cfg.Ugal,
              (T)(1.0 / P->d.dz), (T)(0.4 * 0.5 * P->d.dz), (T)(0.4 * 0.5 * P->d.dz));
  });
}
extern "C" int alfa_derivz_th(alfa_plan* P, const void* TH, const void* ustar, const void* qz,
                              const void* fih, void* dTHdz, alfa_stream s) {
  DISPATCH(P, "alfa_derivz_th", {
    const size_t F = (size_t)P->nx * P->ny * P->nz;
    PwDims dd{F, P->nz};
    PW_LAUNCH(P, STREAM(s), F, k_derivz<T>, dd, (const T*)nullptr, (const T*)nullptr,
              (const T*)nullptr, (const T*)TH, (const T*)ustar, (const T*)nullptr,
              (const T*)nullptr, (const T*)fih, (const T*)qz, (T*)nullptr, (T*)nullptr,
              (T*)nullptr, (T*)dTHdz, (T)P->cfg.Ugal, (T)(1.0 / P->d.dz), (T)(0.4 * 0.5 * P->d.dz),
              (T)(0.4 * 0.5 * P->d.dz));
  });
}
extern "C" int alfa_derivz_generic_uvp(alfa_plan* P, const void* F, void* out, alfa_stream s) {
  DISPATCH(P, "alfa_derivz_generic_uvp", {
    const size_t n = (size_t)P->nx * P->ny * P->nz;
    PwDims dd{n, P->nz};
    PW_LAUNCH(P, STREAM(s), n, k_derivz_generic<T>, dd, (const T*)F, (T*)out, (T)(1.0 / P->d.dz), 0);
  });
}
extern "C" int alfa_derivz_generic_w(alfa_plan* P, const void* F, void* out, alfa_stream s) {
  DISPATCH(P, "alfa_derivz_generic_w", {
    const size_t n = (size_t)P->nx * P->ny * P->nz;
    PwDims dd{n, P->nz};
    PW_LAUNCH(P, STREAM(s), n, k_derivz_generic<T>, dd, (const T*)F, (T*)out, (T)(1.0 / P->d.dz), 1);
  });
}
extern "C" int alfa_planar_mean(alfa_plan* P, const void* F, double* mean, alfa_stream s) {
  DISPATCH(P, "alfa_planar_mean", ops::planar_mean<T>(P, STREAM(s), F, mean));
}
extern "C" int alfa_stag_grid_avg(alfa_plan* P, const void* F, void* out, alfa_stream s) {
  DISPATCH(P, "alfa_stag_grid_avg", {
    const size_t n = (size_t)P->nx * P->ny * P->nz;
    PwDims dd{n, P->nz};
    PW_LAUNCH(P, STREAM(s), n, k_stag_avg<T>, dd, (const T*)F, (T*)out);
  });
}
extern "C" int alfa_imfilter(alfa_plan* P, const void* F, void* out, alfa_stream s) {
  DISPATCH(P, "alfa_imfilter", {
    const size_t n = (size_t)P->nx * P->ny * P->nz;
    PwDims dd{n, P->nz};
    PW_LAUNCH(P, STREAM(s), n, k_imfilter<T>, dd, P->nx, P->ny, (const T*)F, (T*)out);
  });
}
extern "C" int alfa_compute_beta(alfa_plan* P, const double* coeffs, int n, double* beta,
                                 alfa_stream s) {
  if (!P) return set_error(ALFA_EINVAL, "alfa_compute_beta: null plan");
  if (n <= 0) return set_error(ALFA_EINVAL, "alfa_compute_beta: n must be positive");
  launch_k<1024>(P, STREAM(s), dim3((n + 7) / 8), dim3(64), 0, KBODY(k_beta), coeffs, n, beta);
  launch_ok(P);
  return check_hip(P, "alfa_compute_beta");
}
extern "C" int alfa_advection(alfa_plan* P, const void* u, const void* v, const void* w,
                              const void* dudy, const void* dudz, const void* dvdx,
                              const void* dvdz, const void* dwdx, const void* dwdy, void* Cx,
                              void* Cy, void* Cz, alfa_stream s) {
  DISPATCH(P, "alfa_advection",
           ops::advection<T>(P, STREAM(s), u, v, w, dudy, dudz, dvdx, dvdz, dwdx, dwdy, Cx, Cy, Cz));
}
extern "C" int alfa_scalar_advection(alfa_plan* P, const void* u, const void* v, const void* w,
                                     const void* sx, const void* sy, const void* sz, void* out,
                                     alfa_stream s) {
  DISPATCH(P, "alfa_scalar_advection",
           ops::scalar_advection<T>(P, STREAM(s), u, v, w, sx, sy, sz, out));
}
extern "C" int alfa_buoyancy(alfa_plan* P, const void* TH, void* out, alfa_stream s) {
  DISPATCH(P, "alfa_buoyancy", ops::buoyancy<T>(P, STREAM(s), TH, out));
}
extern "C" int alfa_div_stress(alfa_plan* P, const void* txx, const void* tyy, const void* tzz,
                               const void* txy, const void* txz, const void* tyz, void* dtx,
                               void* dty, void* dtz, alfa_stream s) {
  DISPATCH(P, "alfa_div_stress", {
    Ctx<T> c(P, STREAM(s));
    stage_div_stress<T>(c, (T*)txx, (T*)tyy, (T*)tzz, (T*)txy, (T*)txz, (T*)tyz, (T*)dtx,
                        (T*)dty, (T*)dtz);
  });
}
extern "C" int alfa_div_flux(alfa_plan* P, const void* qx, const void* qy, const void* qz,
                             void* divq, alfa_stream s) {
  DISPATCH(P, "alfa_div_flux", {
    Ctx<T> c(P, STREAM(s));
    stage_div_flux<T>(c, (T*)qx, (T*)qy, (T*)qz, (T*)divq);
  });
}
extern "C" int alfa_rhs_momentum(alfa_plan* P, const void* u, const void* v, const void* w,
                                 const void* Cx, const void* Cy, const void* Cz, const void* b,
                                 const void* dtx, const void* dty, const void* dtz, void* Ru,
                                 void* Rv, void* Rw, alfa_stream s) {
  DISPATCH(P, "alfa_rhs_momentum",
           ops::rhs_momentum<T>(P, STREAM(s), u, v, w, Cx, Cy, Cz, b, dtx, dty, dtz, Ru, Rv, Rw));
}
extern "C" int alfa_rhs_scalar(alfa_plan* P, const void* TH, const void* adv, const void* divq,
                               void* RT, alfa_stream s) {
  DISPATCH(P, "alfa_rhs_scalar", ops::rhs_scalar<T>(P, STREAM(s), TH, adv, divq, RT));
}
extern "C" int alfa_pressure_rc(alfa_plan* P, const void* u, const void* v, const void* w,
                                const void* Ru, const void* Rv, const void* Rw, const void* Rpu,
                                const void* Rpv, const void* Rpw, const void* divtz, void* RCr,
                                void* RCi, void* fRzr, alfa_stream s) {
  const void* R[3] = {Ru, Rv, Rw};
  const void* Rp[3] = {Rpu, Rpv, Rpw};
  const void* U[3] = {u, v, w};
  DISPATCH(P, "alfa_pressure_rc", ops::pressure_rc<T>(P, STREAM(s), R, Rp, U, divtz, RCr, RCi, fRzr));
}
extern "C" int alfa_thomas_pressure_solve(alfa_plan* P, const void* RCr, const void* RCi,
                                          const void* fRzr, void* p, void* dpdx, void* dpdy,
                                          void* dpdz, alfa_stream s) {
  DISPATCH(P, "alfa_thomas_pressure_solve",
           ops::thomas_solve<T>(P, STREAM(s), RCr, RCi, fRzr, p, dpdx, dpdy, dpdz));
}
extern "C" int alfa_ab2_uvw(alfa_plan* P, void* u, void* v, void* w, const void* Ru,
                            const void* Rup, const void* Rv, const void* Rvp, const void* Rw,
                            const void* Rwp, alfa_stream s) {
  DISPATCH(P, "alfa_ab2_uvw", {
    const size_t n = (size_t)P->nx * P->ny * P->nz;
    PwDims dd{n, P->nz};
    const T dt = (T)P->d.dt_nondim;
    PW_LAUNCH(P, STREAM(s), n, k_ab2_plain<T>, dd, (T*)u, (const T*)Ru, (const T*)Rup, dt, 0, T(0));
    PW_LAUNCH(P, STREAM(s), n, k_ab2_plain<T>, dd, (T*)v, (const T*)Rv, (const T*)Rvp, dt, 0, T(0));
    PW_LAUNCH(P, STREAM(s), n, k_ab2_plain<T>, dd, (T*)w, (const T*)Rw, (const T*)Rwp, dt, 1, T(0));
  });
}
extern "C" int alfa_ab2_th(alfa_plan* P, void* TH, const void* R, const void* Rp, alfa_stream s) {
  DISPATCH(P, "alfa_ab2_th", {
    const size_t n = (size_t)P->nx * P->ny * P->nz;
    PwDims dd{n, P->nz};
    PW_LAUNCH(P, STREAM(s), n, k_ab2_plain<T>, dd, (T*)TH, (const T*)R, (const T*)Rp,
              (T)P->d.dt_nondim, 2, (T)(P->d.inversion_nondim * P->d.dz));
  });
}
extern "C" int alfa_strains_uvp(alfa_plan* P, const void* const g[9], void* out[7], alfa_stream s) {
  DISPATCH(P, "alfa_strains_uvp", {
    const size_t n = (size_t)P->nx * P->ny * P->nz;
    PwDims dd{n, P->nz};
    (void)hipMemcpyAsync(out[0], g[0], n * sizeof(T), hipMemcpyDeviceToDevice, STREAM(s));
    (void)hipMemcpyAsync(out[1], g[4], n * sizeof(T), hipMemcpyDeviceToDevice, STREAM(s));
    (void)hipMemcpyAsync(out[2], g[8], n * sizeof(T), hipMemcpyDeviceToDevice, STREAM(s));
    PW_LAUNCH(P, STREAM(s), n, k_strains_uvp<T>, dd, (const T*)g[0], (const T*)g[1],
              (const T*)g[2], (const T*)g[3], (const T*)g[4], (const T*)g[5], (const T*)g[6],
              (const T*)g[7], (const T*)g[8], (T*)out[3], (T*)out[4], (T*)out[5], (T*)out[6]);
  });
}
extern "C" int alfa_strains_w(alfa_plan* P, const void* const g[9], void* out[3], alfa_stream s) {
  DISPATCH(P, "alfa_strains_w", {
    const size_t n = (size_t)P->nx * P->ny * P->nz;
    PwDims dd{n, P->nz};
    PW_LAUNCH(P, STREAM(s), n, k_strains_w<T>, dd, (const T*)g[0], (const T*)g[1], (const T*)g[2],
              (const T*)g[3], (const T*)g[4], (const T*)g[5], (const T*)g[6], (const T*)g[7],
              (const T*)g[8], (T*)out[0], (T*)out[1], (T*)out[2], (T*)nullptr, (T*)nullptr,
              (T*)nullptr, (T*)nullptr);
  });
}
extern "C" int alfa_lasdd_sm(alfa_plan* P, const void* u, const void* v, const void* w,
                             const void* const str[7], void* Cs2, double* a1, double* a2,
                             double* beta, alfa_stream s) {
  if (P && P->cfg.optSgs != 1 && P->cfg.optSgs != 3)
    return set_error(ALFA_EINVAL, "alfa_lasdd_sm: plan optSgs must be 1 or 3");
  DISPATCH(P, "alfa_lasdd_sm", ops::lasdd<T>(P, STREAM(s), u, v, w, str, Cs2, a1, a2, beta));
}
extern "C" int alfa_lasdd_wl(alfa_plan* P, const void* u, const void* v, const void* w,
                             const void* const str[6], void* Cwl, double* a1, double* a2,
                             double* beta, alfa_stream s) {
  if (P && P->cfg.optSgs != 2 && P->cfg.optSgs != 4)
    return set_error(ALFA_EINVAL, "alfa_lasdd_wl: plan optSgs must be 2 or 4");
  DISPATCH(P, "alfa_lasdd_wl", ops::lasdd_wl<T>(P, STREAM(s), u, v, w, str, Cwl, a1, a2, beta));
}
extern "C" int alfa_scalar_lasdd(alfa_plan* P, const void* const vel[3], const void* const hat1[3],
                                 const void* const hat2[3], const void* TH,
                                 const void* const dTH[3], const void* S, const void* S_hat,
                                 const void* S_hatd, void* Cs2PrRatio_3D, double* Cs2PrRatio_1D,
                                 double* beta2, alfa_stream s) {
  if (P) {
    const int o = P->cfg.optSgs;
    if (o < 1 || o > 4)
      return set_error(ALFA_EINVAL, "alfa_scalar_lasdd: plan optSgs must be 1..4");
    if ((o == 1 || o == 2) && !hat2)
      return set_error(ALFA_EINVAL, "alfa_scalar_lasdd: level-2 velocity hats are required for beta");
    if ((o == 1 || o == 3) && (!S || !S_hat))
      return set_error(ALFA_EINVAL, "alfa_scalar_lasdd: S and S_hat are required (SM)");
    if (o == 1 && !S_hatd)
      return set_error(ALFA_EINVAL, "alfa_scalar_lasdd: S_hatd is required (LASDD-SM)");
  }
  DISPATCH(P, "alfa_scalar_lasdd",
           ops::scalar_lasdd<T>(P, STREAM(s), vel, hat1, hat2, TH, dTH, S, S_hat, S_hatd,
                                Cs2PrRatio_3D, Cs2PrRatio_1D, beta2));
}
extern "C" int alfa_sgs_stresses(alfa_plan* P, const void* const grads[9], const void* Cs2_3D,
                                 const void* u, const void* v, const void* M_sfc_loc,
                                 const void* psi_m, const void* psi_m0, void* tau[6],
                                 alfa_stream s) {
  DISPATCH(P, "alfa_sgs_stresses",
           ops::sgs_stresses<T>(P, STREAM(s), grads, Cs2_3D, u, v, M_sfc_loc, psi_m, psi_m0, tau));
}
extern "C" int alfa_scalar_fluxes(alfa_plan* P, const void* const grads[9],
                                  const void* const dTH[3], const void* Cs2PrRatio_3D,
                                  const void* qz_sfc, void* q[3], alfa_stream s) {
  DISPATCH(P, "alfa_scalar_fluxes",
           ops::scalar_fluxes<T>(P, STREAM(s), grads, dTH, Cs2PrRatio_3D, qz_sfc, q));
}
extern "C" int alfa_div_stress_sgs(alfa_plan* P, const void* const grads[9], const void* u,
                                   const void* v, const void* w, const void* M_sfc_loc,
                                   const void* psi_m, const void* psi_m0, int dynamic,
                                   void* Cs2_3D, double* Cs2_1D_avg1, double* Cs2_1D_avg2,
                                   double* beta1_1D, void* divt[3], alfa_stream s) {
  if (P && dynamic && (P->cfg.optSgs < 1 || P->cfg.optSgs > 4))
    return set_error(ALFA_EINVAL, "alfa_div_stress_sgs: dynamic needs optSgs 1..4");
  DISPATCH(P, "alfa_div_stress_sgs",
           ops::div_stress_sgs<T>(P, STREAM(s), grads, u, v, w, M_sfc_loc, psi_m, psi_m0, dynamic,
                                  Cs2_3D, Cs2_1D_avg1, Cs2_1D_avg2, beta1_1D, divt));
}
extern "C" int alfa_div_flux_sgs(alfa_plan* P, const void* const grads[9], const void* u,
                                 const void* v, const void* w, const void* TH,
                                 const void* const dTH[3], const void* qz_sfc, int dynamic,
                                 void* Cs2PrRatio_3D, double* Cs2PrRatio_1D, double* beta2_1D,
                                 void* qz, void* divq, alfa_stream s) {
  if (P && dynamic && (P->cfg.optSgs < 1 || P->cfg.optSgs > 4))
    return set_error(ALFA_EINVAL, "alfa_div_flux_sgs: dynamic needs optSgs 1..4");
  DISPATCH(P, "alfa_div_flux_sgs",
           ops::div_flux_sgs<T>(P, STREAM(s), grads, u, v, w, TH, dTH, qz_sfc, dynamic,
                                Cs2PrRatio_3D, Cs2PrRatio_1D, beta2_1D, qz, divq));
}
extern "C" int alfa_wall(alfa_plan* P, const void* u, const void* v, const void* M,
                         const void* pm, const void* pm0, void* txz, void* tyz, alfa_stream s) {
  DISPATCH(P, "alfa_wall", {
    const int ncol = P->nx * P->ny;
    const double logm = std::log(0.5 * P->d.dz * P->d.z_scale / P->cfg.z0m);
    launch_k<1024>(P, STREAM(s), dim3(pw_blocks(ncol)), dim3(256), 0, KBODY(k_wall_plane<T>), ncol, P->nz, (const T*)u, (const T*)v,
                                                            (const T*)M, (const T*)pm,
                                                            (const T*)pm0, logm, 0.4, P->cfg.Ugal,
                                                            (T*)txz, (T*)tyz);
    launch_ok(P);
  });
}
extern "C" int alfa_surface_flux(alfa_plan* P, const void* u, const void* v, const void* TH,
                                 double sfc_val, void* pm, void* pm0, void* ph, void* ph0,
                                 void* fm, void* fh, void* M, void* ustar, void* qz, void* invOB,
                                 double* qavg, alfa_stream s) {
  DISPATCH(P, "alfa_surface_flux",
           ops::surface<T>(P, STREAM(s), u, v, TH, sfc_val, pm, pm0, ph, ph0, fm, fh, M, ustar, qz,
                           invOB, qavg));
}

extern "C" int alfa_profile_begin(alfa_plan* P, const char* filter, int max_launches) {
  if (!P) return set_error(ALFA_EINVAL, "alfa_profile_begin: null plan");
  if (max_launches <= 0) return set_error(ALFA_EINVAL, "alfa_profile_begin: max_launches <= 0");
  free_prof(P);
  P->pf.pool.resize(2 * (size_t)max_launches);
  for (auto& e : P->pf.pool) {
    hipError_t r = hipEventCreate(&e);
    if (r != hipSuccess) {
      P->pf.pool.clear();
      return set_error(ALFA_EHIP, "hipEventCreate: %s", hipGetErrorString(r));
    }
  }
  P->pf.filter = filter ? filter : "";
  P->pf.on = true;
  return ALFA_OK;
}

extern "C" int alfa_profile_end(alfa_plan* P, alfa_kernel_stat* out, int max_out, int* n_out) {
  if (!P) return set_error(ALFA_EINVAL, "alfa_profile_end: null plan");
  P->pf.on = false;
  std::vector<std::string> names;
  std::vector<alfa_kernel_stat> acc;
  if (!P->pf.recs.empty()) {
    hipError_t r = hipEventSynchronize(P->pf.recs.back().b);
    if (r != hipSuccess) return set_error(ALFA_EHIP, "hipEventSynchronize: %s", hipGetErrorString(r));
  }
  for (const ProfRec& rec : P->pf.recs) {
    float ms = 0.f;
    hipError_t r = hipEventElapsedTime(&ms, rec.a, rec.b);
    if (r != hipSuccess) return set_error(ALFA_EHIP, "hipEventElapsedTime: %s", hipGetErrorString(r));
    size_t k = 0;
    while (k < names.size() && names[k] != rec.name) ++k;
    if (k == names.size()) {
      names.push_back(rec.name);
      alfa_kernel_stat z{};
      snprintf(z.name, sizeof z.name, "%s", rec.name.c_str());
      acc.push_back(z);
    }
    acc[k].launches += 1;
    acc[k].total_ms += ms;
    acc[k].bytes += rec.bytes;
    acc[k].ref_bytes += rec.ref_bytes;
  }
  const int n = (int)acc.size() < max_out ? (int)acc.size() : max_out;
  for (int i = 0; i < n; ++i) out[i] = acc[i];
  if (n_out) *n_out = (int)acc.size();
  free_prof(P);
  return ALFA_OK;
}

// ---------------------------------------------------------------------------
// PlanarMean without a plan and without allocation.  The partial sums of
// the plane reduction go to `scratch`, a file-scope device array of a fixed
// size (every GPU that runs this library holds its own copy until the
// process ends).  Host code never uses the array's name as an address: on
// the first call on a device (with that device current) its device address
// is looked up with hipGetSymbolAddress and kept per device.  Every call on
// one GPU shares that GPU's array: calls that overlap on one GPU (two
// streams, two threads) race on it.  Every partial is written before it is
// read, so nothing has to be cleared between calls.
// ---------------------------------------------------------------------------
#include <mutex>

constexpr size_t NOALLOC_SCRATCH_DOUBLES = (size_t)1 << 19;   // 4 MiB
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

static std::mutex g_scratch_mu;
static std::vector<double*> g_scratch_ptr;   // per device ordinal

static int noalloc_scratch(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_error(ALFA_EHIP, "hipGetDevice: %s", hipGetErrorString(e));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    double* p = nullptr;
    e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(ALFA_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                       hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return ALFA_OK;
}

template <class T>
static void planar_mean_noalloc_t(int nx, int ny, int nz, const void* F, double* mean,
                                  double* part, hipStream_t s) {
  // the same geometry, kernels and summation order as alfa_planar_mean
  const int ncol = nx * ny;
  const RedGeom g = red_geom(ncol, nz, sizeof(T));
  FSum1<T> f{(const T*)F, nz, 0.0};
  k_solo<RED_BS><<<dim3(g.nbz, g.nbc), RED_BS, 0, s>>>(KBODY(k_plane_partial<1, FSum1<T>>), g, f,
                                                       part);
  k_solo<1024><<<(nz + 3) / 4, 256, 0, s>>>(KBODY(k_plane_final), g.nbc, 1, nz,
                                            (const double*)part, mean, 1.0 / ncol);
}

extern "C" int alfa_planar_mean_noalloc(int precision, int nx, int ny, int nz, const void* F,
                                        double* mean, alfa_stream stream) {
  if (precision != ALFA_FP64 && precision != ALFA_MIXED)
    return set_error(ALFA_EINVAL, "alfa_planar_mean_noalloc: precision %d", precision);
  if (nx <= 0 || ny <= 0 || nz <= 0 || (long long)nx * ny > (1LL << 30))
    return set_error(ALFA_EINVAL, "alfa_planar_mean_noalloc: bad shape (%d, %d, %d)", nx, ny, nz);
  if (!F || !mean) return set_error(ALFA_EINVAL, "alfa_planar_mean_noalloc: null argument");
  const RedGeom g = red_geom(nx * ny, nz, precision == ALFA_FP64 ? 8 : 4);
  const size_t need = (size_t)g.nbc * nz;
  if (need > NOALLOC_SCRATCH_DOUBLES)
    return set_error(ALFA_EINVAL,
                     "alfa_planar_mean_noalloc: (%d, %d, %d) needs %zu partial sums, more than "
                     "the %zu of its fixed scratch (use alfa_planar_mean with a plan)",
                     nx, ny, nz, need, NOALLOC_SCRATCH_DOUBLES);
  double* part = nullptr;
  int rc = noalloc_scratch(&part);
  if (rc != ALFA_OK) return rc;
  hipStream_t s = STREAM(stream);
  if (precision == ALFA_FP64) planar_mean_noalloc_t<double>(nx, ny, nz, F, mean, part, s);
  else planar_mean_noalloc_t<float>(nx, ny, nz, F, mean, part, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_error(ALFA_EHIP, "alfa_planar_mean_noalloc: %s", hipGetErrorString(e));
  return ALFA_OK;
}
