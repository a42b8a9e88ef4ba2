// The composed time step (src/Main.py:158-540) and its stages.  Each stage
// function serves both the step and the matching per-operation entry point.
#pragma once
#include <cmath>
#include "loaders.hpp"
#include "plan.hpp"
#include "pressure.hpp"
#include "surface.hpp"
#include "contract.hpp"

namespace alfa {

constexpr int BIGCUT = 1 << 28;

template <class T> struct Ctx {
  alfa_plan* P;
  hipStream_t s;
  int nx, ny, nz, nyc, mx, my, ncol;
  size_t F, FB, SPB, PADB, PSPB;
  explicit Ctx(alfa_plan* p, hipStream_t st) : P(p), s(st) {
    nx = p->nx; ny = p->ny; nz = p->nz; nyc = p->nyc; mx = p->mx; my = p->my;
    ncol = nx * ny;
    F = (size_t)ncol * nz;
    FB = F * sizeof(T);
    SPB = (size_t)nx * nyc * nz * sizeof(c2<T>);
    PADB = (size_t)mx * my * nz * sizeof(T);
    PSPB = (size_t)mx * (my / 2 + 1) * nz * sizeof(c2<T>);
  }
  T* W(const std::string& n) { return (T*)P->ws.get("f_" + n, FB); }
  c2<T>* S(const std::string& n) { return (c2<T>*)P->ws.get("s_" + n, SPB); }
  T* Pad(const std::string& n) { return (T*)P->ws.get("p_" + n, PADB); }
  c2<T>* PS(const std::string& n) { return (c2<T>*)P->ws.get("ps_" + n, PSPB); }
  double* D(const std::string& n, size_t cnt) { return (double*)P->ws.get("d_" + n, cnt * sizeof(double)); }
  PwDims dims() const { PwDims d; d.n = F; d.nz = nz; return d; }
  PwDims pdims() const { PwDims d; d.n = (size_t)mx * my * nz; d.nz = nz; return d; }
};

inline XOp xop(int ncy, int inverse, int cx, int mult, int map, int nout, double scale = 1.0) {
  XOp o;
  o.ncy = ncy; o.inverse = inverse; o.cx = cx; o.mult = mult; o.map = map; o.nout = nout;
  o.scale = scale;
  return o;
}

// forward rfft2 of nb real fields (through a loader) into full-pitch spectra
template <class T, class Loader>
void rfft2_fields(Ctx<T>& c, const Loader& ld, int nz, int nb, c2<T>* const* out) {
  alfa_plan* P = c.P;
  y_r2c<T>(P, c.s, ld, c.nx, P->fy, P->tw_y, nz, c.nyc, nb, (void* const*)out);
  XArgs xa{};
  for (int b = 0; b < nb; ++b) { xa.in[b] = out[b]; xa.dst[b][0] = out[b]; }
  xa.op[0] = xop(c.nyc, 0, BIGCUT, XM_NONE, XMAP_ID, c.nx);
  xa.nops = 1; xa.ncy_in = c.nyc; xa.nz = nz; xa.in_spectral = 0; xa.ny_full = c.ny;
  x_pass<T>(P, c.s, xa, nb, P->fx, P->fx, P->tw_x, P->tw_x);
}

// Dealias1(FFT(F)) for nb fields (Dealiasing.py:51-78 after FFT.py:44).
template <class T>
void dealias1(Ctx<T>& c, int nb, const T* const* in, T* const* out) {
  alfa_plan* P = c.P;
  c2<T>* sp[MAXB];
  c2<T>* pp[MAXB];
  for (int b = 0; b < nb; ++b) {
    sp[b] = c.S("d1_" + std::to_string(b));
    pp[b] = c.PS("d1_" + std::to_string(b));
  }
  LoadPtr<T> ld{};
  for (int b = 0; b < nb; ++b) ld.p[b] = in[b];
  ld.ny = c.ny; ld.nz = c.nz;
  y_r2c<T>(P, c.s, ld, c.nx, P->fy, P->tw_y, c.nz, c.nyc, nb, (void* const*)sp);
  XArgs xa{};
  for (int b = 0; b < nb; ++b) { xa.in[b] = sp[b]; xa.dst[b][0] = pp[b]; }
  xa.op[0] = xop(c.nyc, 1, BIGCUT, XM_NONE, XMAP_D1, c.mx);
  xa.nops = 1; xa.ncy_in = c.nyc; xa.nz = c.nz; xa.ny_full = c.ny;
  x_pass<T>(P, c.s, xa, nb, P->fx, P->fmx, P->tw_x, P->tw_mx);
  YC2RArgs ya{};
  for (int b = 0; b < nb; ++b) ya.in[b][0] = pp[b];
  ya.nsum = 1; ya.ncy_in = c.nyc; ya.nz = c.nz; ya.scale = 1.0 / ((double)c.mx * c.my);
  StorePtr<T> st{};
  for (int b = 0; b < nb; ++b) st.p[b] = out[b];
  st.ny = c.my; st.nz = c.nz;
  y_c2r<T>(P, c.s, st, ya, nb, c.mx, P->fmy, P->tw_my);
}

// Dealias2(FFT_pad(F_pad)) for nb padded fields (Dealiasing.py:86-114).
template <class T>
void dealias2(Ctx<T>& c, int nb, const T* const* in, T* const* out) {
  alfa_plan* P = c.P;
  const int ncy = c.ny / 2;
  c2<T>* sp[MAXB];
  c2<T>* op[MAXB];
  for (int b = 0; b < nb; ++b) {
    sp[b] = c.PS("d2_" + std::to_string(b));
    op[b] = c.S("d2_" + std::to_string(b));
  }
  LoadPtr<T> ld{};
  for (int b = 0; b < nb; ++b) ld.p[b] = in[b];
  ld.ny = c.my; ld.nz = c.nz;
  y_r2c<T>(P, c.s, ld, c.mx, P->fmy, P->tw_my, c.nz, ncy, nb, (void* const*)sp);
  XArgs xa{};
  for (int b = 0; b < nb; ++b) { xa.in[b] = sp[b]; xa.dst[b][0] = op[b]; }
  xa.op[0] = xop(ncy, 1, BIGCUT, XM_NONE, XMAP_D2, c.nx);
  xa.nops = 1; xa.ncy_in = ncy; xa.nz = c.nz; xa.ny_full = c.ny;
  x_pass<T>(P, c.s, xa, nb, P->fmx, P->fx, P->tw_mx, P->tw_x);
  YC2RArgs ya{};
  for (int b = 0; b < nb; ++b) ya.in[b][0] = op[b];
  ya.nsum = 1; ya.ncy_in = ncy; ya.nz = c.nz; ya.scale = 2.25 / ((double)c.nx * c.ny);
  StorePtr<T> st{};
  for (int b = 0; b < nb; ++b) st.p[b] = out[b];
  st.ny = c.ny; st.nz = c.nz;
  y_c2r<T>(P, c.s, st, ya, nb, c.nx, P->fy, P->tw_y);
}

// Test filters of nb virtual fields at one or two levels
// (Filtering_Level1/2 of FFT(field), Filtering.py:97-165), forward half: the
// y R2C and the x pass, leaving each level's filtered spectrum (inverse in x,
// pruned in y) in s1[b] / s2[b].
template <class T, class Loader>
void test_filter_fwd(Ctx<T>& c, const Loader& ld, int nb, bool level1, bool level2,
                     const std::string& tag, c2<T>** s1, c2<T>** s2) {
  alfa_plan* P = c.P;
  const auto& d = P->d;
  // the level-2 cut-off lies inside the level-1 one: the forward transform
  // is pruned to the widest level asked for
  const int ncy = level1 ? d.cut1_y : d.cut2_y;
  for (int b = 0; b < nb; ++b) {
    s1[b] = c.S(tag + (level1 ? "1_" : "2_") + std::to_string(b));
    if (level1 && level2) s2[b] = c.S(tag + "2_" + std::to_string(b));
    else s2[b] = s1[b];
  }
  y_r2c<T>(P, c.s, ld, c.nx, P->fy, P->tw_y, c.nz, ncy, nb, (void* const*)s1);
  XArgs xa{};
  int nops = 0;
  if (level1) xa.op[nops++] = xop(d.cut1_y, 1, d.cut1_x, XM_NONE, XMAP_ID, c.nx);
  if (level2) xa.op[nops++] = xop(d.cut2_y, 1, d.cut2_x, XM_NONE, XMAP_ID, c.nx);
  for (int b = 0; b < nb; ++b) {
    xa.in[b] = s1[b];
    xa.dst[b][0] = s1[b];
    if (level1 && level2) xa.dst[b][1] = s2[b];
  }
  xa.nops = nops; xa.ncy_in = ncy; xa.nz = c.nz; xa.ny_full = c.ny;
  x_pass<T>(P, c.s, xa, nb, P->fx, P->fx, P->tw_x, P->tw_x);
}

// inverse half: the y C2R of each level's spectra into out1[b] / out2[b]
template <class T>
void test_filter_inv(Ctx<T>& c, int nb, bool level1, bool level2, c2<T>* const* s1,
                     c2<T>* const* s2, T* const* out1, T* const* out2) {
  alfa_plan* P = c.P;
  const auto& d = P->d;
  const double sc = 1.0 / ((double)c.nx * c.ny);
  for (int lev = 1; lev <= 2; ++lev) {
    if ((lev == 1 && !level1) || (lev == 2 && !level2)) continue;
    YC2RArgs ya{};
    for (int b = 0; b < nb; ++b) ya.in[b][0] = lev == 1 ? s1[b] : s2[b];
    ya.nsum = 1; ya.ncy_in = lev == 1 ? d.cut1_y : d.cut2_y; ya.nz = c.nz; ya.scale = sc;
    StorePtr<T> st{};
    for (int b = 0; b < nb; ++b) st.p[b] = lev == 1 ? out1[b] : out2[b];
    st.ny = c.ny; st.nz = c.nz;
    y_c2r<T>(P, c.s, st, ya, nb, c.nx, P->fy, P->tw_y);
  }
}

template <class T, class Loader>
void test_filter(Ctx<T>& c, const Loader& ld, int nb, bool level1, bool level2,
                 const std::string& tag, T* const* out1, T* const* out2) {
  c2<T>* s1[MAXB];
  c2<T>* s2[MAXB];
  test_filter_fwd<T>(c, ld, nb, level1, level2, tag, s1, s2);
  test_filter_inv<T>(c, nb, level1, level2, s1, s2, out1, out2);
}

// Pointers of the carried state of one run.
template <class T> struct StateT {
  T *u, *v, *w, *TH, *Ru, *Rv, *Rw, *RT, *Cs2, *CsPr;
  T *psim, *psim0, *psih, *psih0, *fim, *fih;
};

// Per-step derived planes.
template <class T> struct Surf {
  T *M, *ustar, *qz, *invOB;
};

template <class T>
void stage_surface(Ctx<T>& c, const StateT<T>& S, const Surf<T>& sf, double sfc_val,
                   const double* sfc_dev = nullptr) {
  alfa_plan* P = c.P;
  const auto& g = P->cfg;
  const auto& d = P->d;
  SurfArgs<T> a;
  a.u = S.u; a.v = S.v; a.TH = S.TH;
  a.psim = S.psim; a.psim0 = S.psim0; a.psih = S.psih; a.psih0 = S.psih0; a.fim = S.fim; a.fih = S.fih;
  a.M = sf.M; a.ustar = sf.ustar; a.qz = sf.qz; a.invOB = sf.invOB;
  a.scal = P->scal;
  a.ncol = c.ncol; a.nz = c.nz; a.bc = g.optSurfBC; a.hetero = g.optSurfFlux != 0;
  a.sfc_val = sfc_val; a.sfc_dev = sfc_dev; a.SHF = g.SensibleHeatFlux; a.Ugal = g.Ugal; a.T0 = g.T_0;
  a.vonk = 0.4; a.g = d.g_nondim; a.dz_half = 0.5 * d.dz; a.z1 = 0.0;
  const double zref = g.zTemperature > 0.0 ? g.zTemperature : g.z0T;
  a.z0m_nd = g.z0m / d.z_scale; a.zT_nd = zref / d.z_scale;
  a.logm = std::log(0.5 * d.dz * d.z_scale / g.z0m);
  a.logh = std::log(0.5 * d.dz * d.z_scale / zref);
  const int nb = imin(SURF_NB_MAX, (c.ncol + SURF_T - 1) / SURF_T);
  double* part = c.D("surf_part", 2 * SURF_NB_MAX);
  double* part2 = c.D("surf_part2", SURF_NB_MAX);
  double* part3 = c.D("surf_part3", 3 * SURF_NB_MAX);
  const double pb = (double)c.ncol * sizeof(T);
  KLAUNCH(P, c.s, "k_surf_sums", 3 * pb, k_surf_sums<T>, nb, SURF_T, a, part);
  if (g.optSurfBC == 2) KLAUNCH(P, c.s, "k_surf_q", 7 * pb, k_surf_q<T>, nb, SURF_T, a, part, part2);
  KLAUNCH(P, c.s, "k_surf_final", 17 * pb, k_surf_final<T>, nb, SURF_T, a, part, part2, part3);
  KLAUNCH(P, c.s, "k_surf_scal", 0.0, k_surf_scal, 1, SURF_T, nb, c.ncol, part3, P->scal);
}

// DivStress (NSE_SGSTerms.py:55-105): 5 transforms, 6 spectral derivatives
// summed pairwise in the y-C2R, plus the z-derivatives in its epilogue.
template <class T>
void stage_div_stress(Ctx<T>& c, T* txx, T* tyy, T* tzz, T* txy, T* txz, T* tyz, T* dtx,
                      T* dty, T* dtz) {
  alfa_plan* P = c.P;
  c2<T>* sp[5] = {c.S("ds0"), c.S("ds1"), c.S("ds2"), c.S("ds3"), c.S("ds4")};
  c2<T>* sxy_x = c.S("ds5");
  // order: txx, txz | tyy, tyz | txy
  LoadPtr<T> ld{};
  ld.p[0] = txx; ld.p[1] = txz; ld.p[2] = tyy; ld.p[3] = tyz; ld.p[4] = txy;
  ld.ny = c.ny; ld.nz = c.nz;
  y_r2c<T>(P, c.s, ld, c.nx, P->fy, P->tw_y, c.nz, c.nyc, 5, (void* const*)sp);
  // x derivatives: one x pass over txx, txz, txy (txy into its own buffer,
  // its y spectrum is still needed).  The y derivatives need no x pass: the
  // x round trip of an unmasked spectrum is the identity times nx, so i*ky
  // is applied to the y spectra directly in the y C2R with that factor.
  XArgs xa{};
  xa.ncy_in = c.nyc; xa.nz = c.nz; xa.ny_full = c.ny; xa.nops = 1;
  xa.in[0] = sp[0]; xa.dst[0][0] = sp[0];
  xa.in[1] = sp[1]; xa.dst[1][0] = sp[1];
  xa.in[2] = sp[4]; xa.dst[2][0] = sxy_x;
  xa.op[0] = xop(c.nyc, 1, BIGCUT, XM_IKX, XMAP_ID, c.nx);
  x_pass<T>(P, c.s, xa, 3, P->fx, P->fx, P->tw_x, P->tw_x);
  YC2RArgs ya{};
  ya.in[0][0] = sp[0]; ya.in[0][1] = sp[4]; ya.dy[0] = 2;   // divtx = d/dx txx + d/dy txy
  ya.in[1][0] = sxy_x; ya.in[1][1] = sp[2]; ya.dy[1] = 2;   // divty = d/dx txy + d/dy tyy
  ya.in[2][0] = sp[1]; ya.in[2][1] = sp[3]; ya.dy[2] = 2;   // divtz = d/dx txz + d/dy tyz
  ya.nsum = 2; ya.ncy_in = c.nyc; ya.nz = c.nz; ya.scale = 1.0 / ((double)c.nx * c.ny);
  ya.dyscale = c.nx;
  StoreDiv<T> st{};
  st.out[0] = dtx; st.out[1] = dty; st.out[2] = dtz;
  st.fz[0] = txz; st.fz[1] = tyz; st.fz[2] = tzz;
  st.wnode[0] = 1; st.wnode[1] = 1; st.wnode[2] = 0;
  st.ny = c.ny; st.nz = c.nz; st.idz = (T)(1.0 / P->d.dz);
  y_c2r<T>(P, c.s, st, ya, 3, c.nx, P->fy, P->tw_y);
}

// DivFlux (SCL_SGSTerms.py:55-89).
template <class T>
void stage_div_flux(Ctx<T>& c, T* qx, T* qy, T* qz, T* divq) {
  alfa_plan* P = c.P;
  c2<T>* sp[2] = {c.S("dq0"), c.S("dq1")};
  LoadPtr<T> ld{};
  ld.p[0] = qx; ld.p[1] = qy; ld.ny = c.ny; ld.nz = c.nz;
  y_r2c<T>(P, c.s, ld, c.nx, P->fy, P->tw_y, c.nz, c.nyc, 2, (void* const*)sp);
  XArgs xa{};
  xa.ncy_in = c.nyc; xa.nz = c.nz; xa.ny_full = c.ny; xa.nops = 1;
  xa.in[0] = sp[0]; xa.dst[0][0] = sp[0];
  xa.op[0] = xop(c.nyc, 1, BIGCUT, XM_IKX, XMAP_ID, c.nx);
  x_pass<T>(P, c.s, xa, 1, P->fx, P->fx, P->tw_x, P->tw_x);
  YC2RArgs ya{};
  ya.in[0][0] = sp[0]; ya.in[0][1] = sp[1]; ya.dy[0] = 2;   // d/dy qy: y spectrum times nx
  ya.nsum = 2; ya.ncy_in = c.nyc; ya.nz = c.nz; ya.scale = 1.0 / ((double)c.nx * c.ny);
  ya.dyscale = c.nx;
  StoreDiv<T> st{};
  st.out[0] = divq; st.fz[0] = qz; st.wnode[0] = 1;
  st.ny = c.ny; st.nz = c.nz; st.idz = (T)(1.0 / P->d.dz);
  y_c2r<T>(P, c.s, st, ya, 1, c.nx, P->fy, P->tw_y);
}

// PressureRC + ThomasPressureSolve + the gradient (NSE_PressureTerms.py:109,
// NSE_PressureTerms_Thomas.py:218).  Writes dpdx, dpdy, dpdz; optionally the
// unmasked RC and p itself (for the per-operation entry points).
template <class T>
void stage_pressure_from_spectra(Ctx<T>& c, c2<T>* fR[3], c2<T>* fbt, T* dpdx, T* dpdy, T* dpdz,
                                 T* p, T* RCr, T* RCi) {
  alfa_plan* P = c.P;
  const auto& d = P->d;
  c2<T>* phat = c.S("phat");
  c2<T>* dpz = c.S("dpzh");
  c2<T>* dpx = c.S("dpxh");
  ThomasArgs<T> ta{};
  ta.fRx = fR[0]; ta.fRy = fR[1]; ta.fRz = fR[2]; ta.fbt = fbt;
  ta.b = P->th_b; ta.phat = phat; ta.dpz = dpz;
  ta.RCr = RCr; ta.RCi = RCi;
  ta.nx = c.nx; ta.ny = c.ny; ta.nyc = c.nyc; ta.nz = c.nz;
  ta.dz = d.dz; ta.idz = 1.0 / d.dz; ta.c_in = 1.0 / (d.dz * d.dz);
  {
    // 3 spectra in, 2 spectra out, the double pivot table (rows of |kx|)
    const double spb = (double)c.nx * c.nyc * c.nz * sizeof(c2<T>);
    KLAUNCH_LDS(P, c.s, "k_thomas", 5.0 * spb + 1.0 * (c.nx / 2 + 1) * c.nyc * (c.nz + 1) * 8.0, 64,
                k_thomas<T>, c.nx * c.nyc, 64, thomas_lds(c.nz), ta);
  }
  XArgs xa{};
  xa.ncy_in = c.nyc; xa.nz = c.nz; xa.ny_full = c.ny; xa.in_spectral = 1;
  // dp/dx from i*kx p^; p^ itself is x-inverted in place and gives dp/dy
  // (i*ky in the y C2R) and p
  xa.in[0] = phat; xa.dst[0][0] = dpx; xa.dst[0][1] = phat;
  xa.op[0] = xop(c.nyc, 1, BIGCUT, XM_IKX, XMAP_ID, c.nx);
  xa.op[1] = xop(c.nyc, 1, BIGCUT, XM_NONE, XMAP_ID, c.nx);
  xa.nops = 2;
  x_pass<T>(P, c.s, xa, 1, P->fx, P->fx, P->tw_x, P->tw_x);
  xa.in[0] = dpz; xa.dst[0][0] = dpz; xa.nops = 1;
  xa.op[0] = xop(c.nyc, 1, BIGCUT, XM_NONE, XMAP_ID, c.nx);
  x_pass<T>(P, c.s, xa, 1, P->fx, P->fx, P->tw_x, P->tw_x);
  YC2RArgs ya{};
  ya.in[0][0] = dpx; ya.in[1][0] = phat; ya.dy[1] = 1; ya.in[2][0] = dpz; ya.in[3][0] = phat;
  ya.nsum = 1; ya.ncy_in = c.nyc; ya.nz = c.nz; ya.scale = 1.0 / ((double)c.nx * c.ny);
  ya.dyscale = 1.0;
  StorePtr<T> st{};
  st.p[0] = dpdx; st.p[1] = dpdy; st.p[2] = dpdz; st.p[3] = p;
  st.ny = c.ny; st.nz = c.nz;
  y_c2r<T>(P, c.s, st, ya, p ? 4 : 3, c.nx, P->fy, P->tw_y);
}

template <class T>
void stage_pressure(Ctx<T>& c, const T* R[3], const T* Rp[3], const T* U[3], const T* divtz,
                    T* dpdx, T* dpdy, T* dpdz, T* p, T* RCr, T* RCi, c2<T>** fR_out) {
  alfa_plan* P = c.P;
  const auto& d = P->d;
  c2<T>* fR[3] = {c.S("fRx"), c.S("fRy"), c.S("fRz")};
  c2<T>* fbt = (c2<T>*)P->ws.get("s_fbt", (size_t)c.nx * c.nyc * 2 * sizeof(c2<T>));
  LoadPressureR<T> ld{};
  for (int k = 0; k < 3; ++k) { ld.R[k] = R[k]; ld.Rp[k] = Rp[k]; ld.U[k] = U[k]; }
  ld.ny = c.ny; ld.nz = c.nz;
  ld.third = (T)(1.0 / 3.0);
  ld.c2dt = (T)(2.0 / (3.0 * d.dt_nondim));
  rfft2_fields<T>(c, ld, c.nz, 3, fR);
  LoadPlanes2<T> lp{divtz, c.ny, c.nz};
  rfft2_fields<T>(c, lp, 2, 1, &fbt);
  if (fR_out) for (int k = 0; k < 3; ++k) fR_out[k] = fR[k];
  stage_pressure_from_spectra<T>(c, fR, fbt, dpdx, dpdy, dpdz, p, RCr, RCi);
}

// Dynamic procedure of the momentum model (DynamicSGS_LASDD_SM.py:138-367).
template <class T>
void stage_lasdd(Ctx<T>& c, const T* u, const T* v, const T* w, const T* const Sij[6],
                 const T* Sm, const double* shu, const double* shv, T* Cs2, double* avg1,
                 double* avg2, double* beta_out) {
  alfa_plan* P = c.P;
  const auto& d = P->d;
  const bool lasdd = P->cfg.optSgs == 1;
  LoadLasdd<T> ld{};
  ld.u = u; ld.v = v; ld.w = w;
  for (int n = 0; n < 6; ++n) ld.S[n] = Sij[n];
  ld.Sm = Sm; ld.shu = shu; ld.shv = shv; ld.ny = c.ny; ld.nz = c.nz;
  T* f1[21];
  T* f2[21];
  for (int b = 0; b < 21; ++b) {
    f1[b] = c.W("lf1_" + std::to_string(b));
    f2[b] = lasdd ? c.W("lf2_" + std::to_string(b)) : nullptr;
  }
  c2<T>* sp1[MAXB];
  c2<T>* sp2[MAXB];
  test_filter_fwd<T>(c, ld, 21, true, lasdd, "ls", sp1, sp2);
  T* LM = c.W("LM");
  T* MM = c.W("MM");
  T* Shat = c.W("Shat");
  double* beta = c.D("beta1", c.nz);
  if (!lasdd && live(P)) {
    launch_k<1024>(P, c.s, dim3((c.nz + 255) / 256), dim3(256), 0, KBODY(k_fill_d), c.nz, beta, 1.0);
    launch_ok(P);
  }
  // Fused form, where it applies: the level-1 inverse transforms and the
  // LM/MM contraction in one kernel (u_hat and |S_hat| are still written: the
  // scalar procedure reads them).  LASDD first takes the plane sums of both
  // levels the same way (level 2 also writing u_hat and |S_hat| of its
  // level), then beta, then the contraction with beta.
  bool fused = false;
  T* Shatd = lasdd ? c.W("Shatd") : nullptr;
  if (P->fuse_contract && ctr_applies<T>(c.nz, c.ny) &&
      ctr_aligned({f1[0], f1[1], f1[2], LM, MM, Shat, Shatd}) &&
      (!lasdd || ctr_aligned({f2[0], f2[1], f2[2]}))) {
    if (lasdd) {
      double* part = c.D("lpart", (size_t)c.nx * 10 * c.nz);
      double* sums = c.D("lsums", 10 * c.nz);
      double* coef = c.D("lcoef", 6 * c.nz);
      c2r_lasdd_sums<T>(P, c.s, sp1, d.cut1_y, part, 1, nullptr, nullptr, c.nx, c.ny, c.nz, P->fy,
                        P->tw_y, d.L, d.TFR);
      c2r_lasdd_sums<T>(P, c.s, sp2, d.cut2_y, part, 2, f2, Shatd, c.nx, c.ny, c.nz, P->fy,
                        P->tw_y, d.L, d.TFR);
      if (live(P)) {
        launch_k<1024>(P, c.s, dim3((10 * c.nz + 3) / 4), dim3(256), 0, KBODY(k_plane_final), c.nx, 10, c.nz, part, sums, 1.0 / c.ncol);
        launch_k<1024>(P, c.s, dim3((c.nz + 63) / 64), dim3(64), 0, KBODY(k_quintic), sums, c.nz, 0, d.TFR, coef);
        launch_k<1024>(P, c.s, dim3((c.nz + 7) / 8), dim3(64), 0, KBODY(k_beta), coef, c.nz, beta);
        launch_ok(P);
      }
    }
    fused = c2r_lad<T>(P, c.s, sp1, d.cut1_y, beta, f1, LM, MM, Shat, c.nx, c.ny, c.nz, P->fy,
                       P->tw_y, d.L, d.TFR);
  }
  MomLevel<T> l1{}, l2{};
  for (int k = 0; k < 3; ++k) { l1.uh[k] = f1[k]; l2.uh[k] = f2[k]; }
  for (int n = 0; n < 6; ++n) {
    l1.ph[n] = f1[3 + n]; l1.Sh[n] = f1[9 + n]; l1.SSh[n] = f1[15 + n];
    l2.ph[n] = f2[3 + n]; l2.Sh[n] = f2[9 + n]; l2.SSh[n] = f2[15 + n];
  }
  if (!fused) {
    test_filter_inv<T>(c, 21, true, lasdd, sp1, sp2, f1, f2);
    if (lasdd) {
      double* sums = c.D("lsums", 10 * c.nz);
      double* coef = c.D("lcoef", 6 * c.nz);
      FLasddSums<T> f{l1, l2, c.nz, d.L, d.TFR};
      plane_sums<10>(P, c.s, c.ncol, c.nz, f, sums, 1.0 / c.ncol, "lasdd");
      if (live(P)) {
        launch_k<1024>(P, c.s, dim3((c.nz + 63) / 64), dim3(64), 0, KBODY(k_quintic), sums, c.nz, 0, d.TFR, coef);
        launch_k<1024>(P, c.s, dim3((c.nz + 7) / 8), dim3(64), 0, KBODY(k_beta), coef, c.nz, beta);
        launch_ok(P);
      }
    }
    if (vec_ok<T>(c.nz, {l1.uh[0], l1.uh[1], l1.uh[2], l1.ph[0], l1.ph[1], l1.ph[2], l1.ph[3],
                         l1.ph[4], l1.ph[5], l1.Sh[0], l1.Sh[1], l1.Sh[2], l1.Sh[3], l1.Sh[4],
                         l1.Sh[5], l1.SSh[0], l1.SSh[1], l1.SSh[2], l1.SSh[3], l1.SSh[4],
                         l1.SSh[5], LM, MM, Shat}))
      PW_LAUNCH_W(P, c.s, c.F, 2, k_lasdd_lm_v<T>, c.dims(), l1, beta, d.L, d.TFR, LM, MM, Shat);
    else
      PW_LAUNCH(P, c.s, c.F, k_lasdd_lm<T>, c.dims(), l1, beta, d.L, d.TFR, LM, MM, Shat);
  }
  if (lasdd && !fused) PW_LAUNCH(P, c.s, c.F, k_smag_level<T>, c.dims(), l2, Shatd);
  double* s2 = c.D("bsums", 2 * c.nz);
  box_clip_sums<T>(P, c.s, c.nx, c.ny, c.nz, LM, MM, Cs2, s2, "box");
  if (live(P)) {
    launch_k<1024>(P, c.s, dim3((c.nz + 255) / 256), dim3(256), 0, KBODY(k_profiles), s2, c.nz, 1.0 / c.ncol, avg1, avg2, (const double*)beta, beta_out);
    launch_ok(P);
  }
}

// Dynamic procedure of the scalar model (DynamicSGS_ScalarLASDD_SM.py:139-316).
// The momentum stage's filtered velocities and |S^| (its outputs u_hat..
// S_hatd, DynamicSGS_Main.py:326-331) that the scalar models reuse.
template <class T> struct MomHats {
  const T* uh1[3];
  const T* uh2[3];    // NULL where not computed (beta = 1 models)
  const T* Shat;
  const T* Shatd;
};

template <class T> MomHats<T> step_hats(Ctx<T>& c) {
  MomHats<T> h{};
  const int o = c.P->cfg.optSgs;
  for (int k = 0; k < 3; ++k) {
    h.uh1[k] = c.W("lf1_" + std::to_string(k));
    h.uh2[k] = (o == 1 || o == 2) ? c.W("lf2_" + std::to_string(k)) : nullptr;
  }
  h.Shat = c.W("Shat");
  h.Shatd = o == 1 ? c.W("Shatd") : nullptr;
  return h;
}

// w_uvp: w is already averaged to uvp nodes (the reference's w_)
template <class T>
void stage_scalar_lasdd(Ctx<T>& c, const T* u, const T* v, const T* w, int w_uvp, const T* TH,
                        const T* dx, const T* dy, const T* dz, const T* Sm, const MomHats<T>& mh,
                        const double* shu, const double* shv, const double* sht, T* CsPr,
                        double* prof, double* beta_out) {
  alfa_plan* P = c.P;
  const auto& d = P->d;
  const bool lasdd = P->cfg.optSgs == 1;
  LoadScalar<T> ld{};
  ld.u = u; ld.v = v; ld.w = w; ld.TH = TH; ld.dx = dx; ld.dy = dy; ld.dz = dz; ld.Sm = Sm;
  ld.shu = shu; ld.shv = shv; ld.sht = sht; ld.ny = c.ny; ld.nz = c.nz; ld.w_uvp = w_uvp;
  T* f1[10];
  T* f2[10];
  for (int b = 0; b < 10; ++b) {
    f1[b] = c.W("sf1_" + std::to_string(b));
    f2[b] = lasdd ? c.W("sf2_" + std::to_string(b)) : nullptr;
  }
  c2<T>* sp1[MAXB];
  c2<T>* sp2[MAXB];
  test_filter_fwd<T>(c, ld, 10, true, lasdd, "ss", sp1, sp2);
  ScaLevel<T> l1{}, l2{};
  l1.th = f1[0]; l2.th = f2[0];
  for (int k = 0; k < 3; ++k) {
    l1.uth[k] = f1[1 + k]; l1.dh[k] = f1[4 + k]; l1.sdh[k] = f1[7 + k];
    l2.uth[k] = f2[1 + k]; l2.dh[k] = f2[4 + k]; l2.sdh[k] = f2[7 + k];
    l1.uh[k] = mh.uh1[k];
    l2.uh[k] = mh.uh2[k];
  }
  l1.Sm = mh.Shat;
  l2.Sm = mh.Shatd;
  double* beta = c.D("beta2", c.nz);
  T* Tup = c.W("Tup");
  T* Tdn = c.W("Tdn");
  if (!lasdd && live(P)) {
    launch_k<1024>(P, c.s, dim3((c.nz + 255) / 256), dim3(256), 0, KBODY(k_fill_d), c.nz, beta, 1.0);
    launch_ok(P);
  }
  // Fused form, where it applies: the level-1 inverse transforms and the
  // T_up/T_dn contraction in one kernel; LASDD first takes the plane sums of
  // both levels the same way, then beta
  bool fused = false;
  if (P->fuse_contract && ctr_applies<T>(c.nz, c.ny) &&
      ctr_aligned({l1.uh[0], l1.uh[1], l1.uh[2], l1.Sm, Tup, Tdn}) &&
      (!lasdd || ctr_aligned({l2.uh[0], l2.uh[1], l2.uh[2], l2.Sm}))) {
    if (lasdd) {
      double* part = c.D("spart", (size_t)c.nx * 10 * c.nz);
      double* sums = c.D("ssums", 10 * c.nz);
      double* coef = c.D("scoef", 6 * c.nz);
      c2r_sca_sums<T>(P, c.s, sp1, d.cut1_y, part, 1, l1.uh, l1.Sm, c.nx, c.ny, c.nz, P->fy,
                      P->tw_y, d.L, d.TFR);
      c2r_sca_sums<T>(P, c.s, sp2, d.cut2_y, part, 2, l2.uh, l2.Sm, c.nx, c.ny, c.nz, P->fy,
                      P->tw_y, d.L, d.TFR);
      if (live(P)) {
        launch_k<1024>(P, c.s, dim3((10 * c.nz + 3) / 4), dim3(256), 0, KBODY(k_plane_final), c.nx, 10, c.nz, part, sums, 1.0 / c.ncol);
        launch_k<1024>(P, c.s, dim3((c.nz + 63) / 64), dim3(64), 0, KBODY(k_quintic), sums, c.nz, 1, d.TFR, coef);
        launch_k<1024>(P, c.s, dim3((c.nz + 7) / 8), dim3(64), 0, KBODY(k_beta), coef, c.nz, beta);
        launch_ok(P);
      }
    }
    fused = c2r_sca<T>(P, c.s, sp1, d.cut1_y, beta, l1.uh, l1.Sm, Tup, Tdn, c.nx, c.ny, c.nz,
                       P->fy, P->tw_y, d.L, d.TFR);
  }
  if (!fused) {
    test_filter_inv<T>(c, 10, true, lasdd, sp1, sp2, f1, f2);
    if (lasdd) {
      double* sums = c.D("ssums", 10 * c.nz);
      double* coef = c.D("scoef", 6 * c.nz);
      FScalarSums<T> f{l1, l2, c.nz, d.L, d.TFR};
      plane_sums<10>(P, c.s, c.ncol, c.nz, f, sums, 1.0 / c.ncol, "slasdd");
      if (live(P)) {
        launch_k<1024>(P, c.s, dim3((c.nz + 63) / 64), dim3(64), 0, KBODY(k_quintic), sums, c.nz, 1, d.TFR, coef);
        launch_k<1024>(P, c.s, dim3((c.nz + 7) / 8), dim3(64), 0, KBODY(k_beta), coef, c.nz, beta);
        launch_ok(P);
      }
    }
    if (vec_ok<T>(c.nz, {l1.th, l1.Sm, l1.uth[0], l1.uth[1], l1.uth[2], l1.dh[0], l1.dh[1],
                         l1.dh[2], l1.sdh[0], l1.sdh[1], l1.sdh[2], l1.uh[0], l1.uh[1], l1.uh[2],
                         Tup, Tdn}))
      PW_LAUNCH_W(P, c.s, c.F, 2, k_scalar_updn_v<T>, c.dims(), l1, beta, d.L, d.TFR, Tup, Tdn);
    else
      PW_LAUNCH(P, c.s, c.F, k_scalar_updn<T>, c.dims(), l1, beta, d.L, d.TFR, Tup, Tdn);
  }
  double* s2 = c.D("ssum2", 2 * c.nz);
  box_clip_sums<T>(P, c.s, c.nx, c.ny, c.nz, Tup, Tdn, CsPr, s2, "sbox");
  if (live(P)) {
    launch_k<1024>(P, c.s, dim3((c.nz + 255) / 256), dim3(256), 0, KBODY(k_profiles), s2, c.nz, 1.0 / c.ncol, nullptr, prof, (const double*)beta, beta_out);
    launch_ok(P);
  }
}

// Dynamic procedure of the Wong-Lilly momentum model (DynamicSGS_LASDD_WL.py:
// 128-313): 15 filtered fields at level 1; at level 2 all 15 for LASDD-WL
// (optSgs 2, beta) or only S_ij for LAD-WL (optSgs 4).
template <class T>
void stage_lasdd_wl(Ctx<T>& c, const T* u, const T* v, const T* w, const T* const Sij[6],
                    const double* shu, const double* shv, T* Cwl, double* avg1, double* avg2,
                    double* beta_out) {
  alfa_plan* P = c.P;
  const auto& d = P->d;
  const bool lasdd = P->cfg.optSgs == 2;
  LoadLasdd<T> ld{};
  ld.u = u; ld.v = v; ld.w = w;
  for (int n = 0; n < 6; ++n) ld.S[n] = Sij[n];
  ld.Sm = nullptr; ld.shu = shu; ld.shv = shv; ld.ny = c.ny; ld.nz = c.nz;
  T* f1[15];
  T* f2[15];
  for (int b = 0; b < 15; ++b) {
    f1[b] = c.W("lf1_" + std::to_string(b));
    f2[b] = (lasdd || b >= 9) ? c.W("lf2_" + std::to_string(b)) : nullptr;
  }
  if (lasdd) {
    test_filter<T>(c, ld, 15, true, true, "ls", f1, f2);
  } else {
    test_filter<T>(c, ld, 15, true, false, "ls", f1, f2);
    ld.b0 = 9;
    test_filter<T>(c, ld, 6, false, true, "lw", f2 + 9, f2 + 9);
  }
  WlMom<T> l1{}, l2{};
  for (int k = 0; k < 3; ++k) { l1.uh[k] = f1[k]; l2.uh[k] = f2[k]; }
  for (int n = 0; n < 6; ++n) {
    l1.ph[n] = f1[3 + n]; l1.Sh[n] = f1[9 + n];
    l2.ph[n] = f2[3 + n]; l2.Sh[n] = f2[9 + n];
  }
  double* beta = c.D("beta1", c.nz);
  if (lasdd) {
    double* sums = c.D("lsums", 4 * c.nz);
    double* coef = c.D("lcoef", 6 * c.nz);
    FWlSums<T> f{l1, l2, c.nz};
    plane_sums<4>(P, c.s, c.ncol, c.nz, f, sums, 1.0 / c.ncol, "lasddwl");
    if (live(P)) {
      launch_k<1024>(P, c.s, dim3((c.nz + 63) / 64), dim3(64), 0, KBODY(k_quintic), sums, c.nz, 2, d.TFR, coef);
      launch_k<1024>(P, c.s, dim3((c.nz + 7) / 8), dim3(64), 0, KBODY(k_beta), coef, c.nz, beta);
      launch_ok(P);
    }
  } else if (live(P)) {
    launch_k<1024>(P, c.s, dim3((c.nz + 255) / 256), dim3(256), 0, KBODY(k_fill_d), c.nz, beta, 1.0);
    launch_ok(P);
  }
  T* LM = c.W("LM");
  T* MM = c.W("MM");
  PW_LAUNCH(P, c.s, c.F, k_wl_lm<T>, c.dims(), l1, l2, beta, d.L, d.TFR, LM, MM);
  double* s2 = c.D("bsums", 2 * c.nz);
  box_clip_sums<T>(P, c.s, c.nx, c.ny, c.nz, LM, MM, Cwl, s2, "box");
  if (live(P)) {
    launch_k<1024>(P, c.s, dim3((c.nz + 255) / 256), dim3(256), 0, KBODY(k_profiles), s2, c.nz, 1.0 / c.ncol, avg1, avg2, (const double*)beta, beta_out);
    launch_ok(P);
  }
}

// Dynamic procedure of the Wong-Lilly scalar model
// (DynamicSGS_ScalarLASDD_WL.py:130-275): TH, (u,v,w)TH and the gradients
// at level 1; at level 2 all 7 (optSgs 2) or the gradients only (optSgs 4).
template <class T>
void stage_scalar_wl(Ctx<T>& c, const T* u, const T* v, const T* w, int w_uvp, const T* TH,
                     const T* dx, const T* dy, const T* dz, const MomHats<T>& mh,
                     const double* shu, const double* shv, const double* sht, T* CwlPr,
                     double* prof, double* beta_out) {
  alfa_plan* P = c.P;
  const auto& d = P->d;
  const bool lasdd = P->cfg.optSgs == 2;
  LoadScalar<T> ld{};
  ld.u = u; ld.v = v; ld.w = w; ld.TH = TH; ld.dx = dx; ld.dy = dy; ld.dz = dz; ld.Sm = nullptr;
  ld.shu = shu; ld.shv = shv; ld.sht = sht; ld.ny = c.ny; ld.nz = c.nz; ld.w_uvp = w_uvp;
  T* f1[7];
  T* f2[7];
  for (int b = 0; b < 7; ++b) {
    f1[b] = c.W("sf1_" + std::to_string(b));
    f2[b] = (lasdd || b >= 4) ? c.W("sf2_" + std::to_string(b)) : nullptr;
  }
  if (lasdd) {
    test_filter<T>(c, ld, 7, true, true, "ss", f1, f2);
  } else {
    test_filter<T>(c, ld, 7, true, false, "ss", f1, f2);
    ld.b0 = 4;
    test_filter<T>(c, ld, 3, false, true, "sw", f2 + 4, f2 + 4);
  }
  WlSca<T> l1{}, l2{};
  l1.th = f1[0]; l2.th = f2[0];
  for (int k = 0; k < 3; ++k) {
    l1.uth[k] = f1[1 + k]; l1.dh[k] = f1[4 + k];
    l2.uth[k] = f2[1 + k]; l2.dh[k] = f2[4 + k];
    l1.uh[k] = mh.uh1[k];
    l2.uh[k] = mh.uh2[k];
  }
  double* beta = c.D("beta2", c.nz);
  if (lasdd) {
    double* sums = c.D("ssums", 4 * c.nz);
    double* coef = c.D("scoef", 6 * c.nz);
    FWlScalarSums<T> f{l1, l2, c.nz};
    plane_sums<4>(P, c.s, c.ncol, c.nz, f, sums, 1.0 / c.ncol, "slasddwl");
    if (live(P)) {
      launch_k<1024>(P, c.s, dim3((c.nz + 63) / 64), dim3(64), 0, KBODY(k_quintic), sums, c.nz, 2, d.TFR, coef);
      launch_k<1024>(P, c.s, dim3((c.nz + 7) / 8), dim3(64), 0, KBODY(k_beta), coef, c.nz, beta);
      launch_ok(P);
    }
  } else if (live(P)) {
    launch_k<1024>(P, c.s, dim3((c.nz + 255) / 256), dim3(256), 0, KBODY(k_fill_d), c.nz, beta, 1.0);
    launch_ok(P);
  }
  T* Tup = c.W("Tup");
  T* Tdn = c.W("Tdn");
  PW_LAUNCH(P, c.s, c.F, k_wl_scalar_updn<T>, c.dims(), l1, l2, beta, d.L, d.TFR, Tup, Tdn);
  double* s2 = c.D("ssum2", 2 * c.nz);
  box_clip_sums<T>(P, c.s, c.nx, c.ny, c.nz, Tup, Tdn, CwlPr, s2, "sbox");
  if (live(P)) {
    launch_k<1024>(P, c.s, dim3((c.nz + 255) / 256), dim3(256), 0, KBODY(k_profiles), s2, c.nz, 1.0 / c.ncol, nullptr, prof, (const double*)beta, beta_out);
    launch_ok(P);
  }
}

// Strain rates shared by the stress and flux stages (StrainRates.py:56-368):
// S_ij and |S| on uvp nodes, S13, S23, |S| on w nodes, and with
// dealiasing (FGR = 1) their padded copies.
template <class T> struct SgsStrains {
  const T* Sij[6];    // S11 (= du/dx), S22, S33, S12, S13, S23 on uvp nodes
  T* Sm;              // |S| on uvp nodes
  T *S13w, *S23w, *Sw;
  T* ps[6];           // padded uvp components (dealias)
  T* pw[6];           // padded w-node components (dealias)
  T *pS, *pSw;        // padded |S| (dealias)
};

// w_nodes = false: the unpadded w-node strains are left to k_sgs_fused.
// uvp_done: S12, S13, S23 and |S| on uvp nodes are already in their
// workspace buffers (k_dz_adv_strains).
template <class T>
void stage_strains(Ctx<T>& c, const T* const g[9], SgsStrains<T>& o, bool w_nodes = true,
                   bool uvp_done = false) {
  alfa_plan* P = c.P;
  hipStream_t s = c.s;
  T *S12 = c.W("S12"), *S13 = c.W("S13"), *S23 = c.W("S23");
  o.Sm = c.W("S");
  if (!uvp_done)
    PW_LAUNCH(P, s, c.F, k_strains_uvp<T>, c.dims(), g[0], g[1], g[2], g[3], g[4], g[5], g[6],
              g[7], g[8], S12, S13, S23, o.Sm);
  o.Sij[0] = g[0]; o.Sij[1] = g[4]; o.Sij[2] = g[8];
  o.Sij[3] = S12; o.Sij[4] = S13; o.Sij[5] = S23;
  o.S13w = c.W("S13w"); o.S23w = c.W("S23w"); o.Sw = c.W("Sw");
  if (P->d.optDealias != 1) {
    if (w_nodes)
      PW_LAUNCH(P, s, c.F, k_strains_w<T>, c.dims(), g[0], g[1], g[2], g[3], g[4], g[5], g[6],
                g[7], g[8], o.S13w, o.S23w, o.Sw, (T*)nullptr, (T*)nullptr, (T*)nullptr,
                (T*)nullptr);
    o.pS = o.pSw = nullptr;
    return;
  }
  const size_t PF = (size_t)c.mx * c.my * c.nz;
  for (int n = 0; n < 6; ++n) o.ps[n] = c.Pad("su" + std::to_string(n));
  dealias1<T>(c, 6, o.Sij, o.ps);
  o.pS = c.Pad("S");
  PW_LAUNCH(P, s, PF, k_smag<T>, c.pdims(), o.ps[0], o.ps[1], o.ps[2], o.ps[3], o.ps[4], o.ps[5],
            o.pS);
  T *S11w = c.W("S11w"), *S22w = c.W("S22w"), *S33w = c.W("S33w"), *S12w = c.W("S12w");
  PW_LAUNCH(P, s, c.F, k_strains_w<T>, c.dims(), g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7],
            g[8], o.S13w, o.S23w, o.Sw, S11w, S22w, S33w, S12w);
  const T* w6[6] = {S11w, S22w, S33w, S12w, o.S13w, o.S23w};
  for (int n = 0; n < 6; ++n) o.pw[n] = c.Pad("sw" + std::to_string(n));
  dealias1<T>(c, 6, w6, o.pw);
  o.pSw = c.Pad("Sw");
  PW_LAUNCH(P, s, PF, k_smag<T>, c.pdims(), o.pw[0], o.pw[1], o.pw[2], o.pw[3], o.pw[4], o.pw[5],
            o.pSw);
}

inline bool wl_model(const alfa_plan* P) { return P->cfg.optSgs == 2 || P->cfg.optSgs == 4; }

// SGS stresses txx, tyy, tzz, txy, txz, tyz with the wall stress at level 0
// (DynamicSGS_Main.py:136-282 / StaticSGS_Main.py:76-250; SGSStresses_SM.py,
// SGSStresses_WL.py).
template <class T>
void stage_stresses(Ctx<T>& c, const SgsStrains<T>& ss, const T* Cs2, const T* u, const T* v,
                    const T* M, const T* psim, const T* psim0, T* const tau[6]) {
  alfa_plan* P = c.P;
  hipStream_t s = c.s;
  const auto& d = P->d;
  const bool wl = wl_model(P);
  const double logm = std::log(0.5 * d.dz * d.z_scale / P->cfg.z0m);
  // prefactor: -2 L^2 (SM, times |S|) or -2 L^(4/3) (WL, no |S|)
  const T pre = wl ? (T)(-2.0 * std::pow(d.L, 4.0 / 3.0)) : (T)(-2.0 * d.L * d.L);
  const double Ugal = P->cfg.Ugal;
  if (d.optDealias != 1) {
    PW_LAUNCH(P, s, c.F, k_stress_uvp<T>, c.dims(), ss.Sij[0], ss.Sij[1], ss.Sij[2], ss.Sij[3],
              wl ? (const T*)nullptr : ss.Sm, Cs2, pre, tau[0], tau[1], tau[2], tau[3]);
    PW_LAUNCH(P, s, c.F, k_stress_w<T>, c.dims(), ss.S13w, ss.S23w,
              wl ? (const T*)nullptr : ss.Sw, Cs2, pre, tau[4], tau[5], 1, u, v, M, psim, psim0,
              logm, 0.4, Ugal);
    return;
  }
  const size_t PF = (size_t)c.mx * c.my * c.nz;
  T* pc = c.Pad("cs2");
  const T* cs[1] = {Cs2};
  dealias1<T>(c, 1, cs, &pc);
  T *ptxx = c.Pad("txx"), *ptyy = c.Pad("tyy"), *ptzz = c.Pad("tzz"), *ptxy = c.Pad("txy");
  PW_LAUNCH(P, s, PF, k_stress_uvp<T>, c.pdims(), ss.ps[0], ss.ps[1], ss.ps[2], ss.ps[3],
            wl ? (const T*)nullptr : ss.pS, pc, pre, ptxx, ptyy, ptzz, ptxy);
  const T* pt4[4] = {ptxx, ptyy, ptzz, ptxy};
  T* t4[4] = {tau[0], tau[1], tau[2], tau[3]};
  dealias2<T>(c, 4, pt4, t4);
  T *ptxz = c.Pad("txz"), *ptyz = c.Pad("tyz");
  PW_LAUNCH(P, s, PF, k_stress_w<T>, c.pdims(), ss.pw[4], ss.pw[5],
            wl ? (const T*)nullptr : ss.pSw, pc, pre, ptxz, ptyz, 0, (const T*)nullptr,
            (const T*)nullptr, (const T*)nullptr, (const T*)nullptr, (const T*)nullptr, logm, 0.4,
            Ugal);
  const T* pt2[2] = {ptxz, ptyz};
  T* t2[2] = {tau[4], tau[5]};
  dealias2<T>(c, 2, pt2, t2);
  PW_LAUNCH(P, s, (size_t)c.ncol, k_level0<T>, c.ncol, c.nz, tau[4], tau[5], u, v, M, psim, psim0,
            logm, 0.4, Ugal, (T*)nullptr, (const T*)nullptr);
}

// SGS scalar fluxes qx, qy, qz with qz = qsfc at level 0
// (DynamicSGS_Main.py:420-478 / StaticSGS_Main.py:258-360; ScalarSGSFluxes_SM.py,
// ScalarSGSFluxes_WL.py).
template <class T>
void stage_fluxes(Ctx<T>& c, const SgsStrains<T>& ss, const T* const dT[3], const T* CsPr,
                  const T* qsfc, T* const q[3]) {
  alfa_plan* P = c.P;
  hipStream_t s = c.s;
  const auto& d = P->d;
  const bool wl = wl_model(P);
  // prefactor: -2 L^2 (SM, times |S|) or -L^(4/3) (WL, no |S|)
  const T pre = wl ? (T)(-std::pow(d.L, 4.0 / 3.0)) : (T)(-2.0 * d.L * d.L);
  if (d.optDealias != 1) {
    PW_LAUNCH(P, s, c.F, k_scalar_flux<T>, c.dims(), dT[0], dT[1], dT[2],
              wl ? (const T*)nullptr : ss.Sm, wl ? (const T*)nullptr : ss.Sw, CsPr, pre, q[0], q[1],
              q[2], qsfc);
    return;
  }
  const size_t PF = (size_t)c.mx * c.my * c.nz;
  const T* in4[4] = {dT[0], dT[1], dT[2], CsPr};
  T* p4[4] = {c.Pad("q0"), c.Pad("q1"), c.Pad("q2"), c.Pad("q3")};
  dealias1<T>(c, 4, in4, p4);
  T *pqx = c.Pad("qx"), *pqy = c.Pad("qy"), *pqz = c.Pad("qz");
  PW_LAUNCH(P, s, PF, k_scalar_flux<T>, c.pdims(), p4[0], p4[1], p4[2],
            wl ? (const T*)nullptr : ss.pS, wl ? (const T*)nullptr : ss.pSw, p4[3], pre, pqx, pqy,
            pqz, (const T*)nullptr);
  const T* pq3[3] = {pqx, pqy, pqz};
  dealias2<T>(c, 3, pq3, q);
  PW_LAUNCH(P, s, (size_t)c.ncol, k_level0<T>, c.ncol, c.nz, (T*)nullptr, (T*)nullptr,
            (const T*)nullptr, (const T*)nullptr, (const T*)nullptr, (const T*)nullptr,
            (const T*)nullptr, 0.0, 0.4, 0.0, q[2], qsfc);
}

// u, v, TH, THv = (TH + T_0)(1 + 0.61 Q) and Q (moisture runs)
template <class T> struct FSum5 {
  static constexpr const char* NAME = "FSum5";
  static constexpr int NFIELDS = 4;
  const T* u; const T* v; const T* TH; const T* Q; int nz; double T0;
  __device__ __forceinline__ void operator()(int c, int z, double* acc) const {
    const size_t i = (size_t)c * nz + z;
    const double th = TH[i], q = Q[i];
    acc[0] += (double)u[i];
    acc[1] += (double)v[i];
    acc[2] += th;
    acc[3] += (th + T0) * (1.0 + 0.61 * q);
    acc[4] += q;
  }
};

// level 0 of a field, and a plane (plane sums with nz = 1)
template <class T> struct FLevel0 {
  static constexpr const char* NAME = "FLevel0";
  static constexpr int NFIELDS = 0;
  const T* F; int nz;
  __device__ __forceinline__ void operator()(int c, int, double* acc) const {
    acc[0] += (double)F[(size_t)c * nz];
  }
};

template <class T> struct FSum4 {
  static constexpr const char* NAME = "FSum4";
  static constexpr int NFIELDS = 3;
  const T* u; const T* v; const T* TH; int nz; double T0;
  __device__ __forceinline__ void operator()(int c, int z, double* acc) const {
    const size_t i = (size_t)c * nz + z;
    const double th = TH[i];
    acc[0] += (double)u[i];
    acc[1] += (double)v[i];
    acc[2] += th;
    acc[3] += th + T0;
  }
};

__device__ __forceinline__ void k_cfl_final(const double* __restrict__ part, int nb, double cx, double cy,
                            double cz, double* scal, double* out) {
  __shared__ double sm[3][256];
  double m[3] = {0, 0, 0};
  for (int b = threadIdx.x; b < nb; b += 256)
    for (int k = 0; k < 3; ++k) m[k] = fmax(m[k], part[3 * b + k]);
  for (int k = 0; k < 3; ++k) sm[k][threadIdx.x] = m[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < 3; ++k) sm[k][threadIdx.x] = fmax(sm[k][threadIdx.x], sm[k][threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double cfl = fmax(fmax(sm[0][0] * cx, sm[1][0] * cy), sm[2][0] * cz);
    scal[0] = cfl;
    if (out)
      for (int k = 0; k < 5; ++k) out[k] = scal[k];
  }
}

constexpr int AB2_BLOCKS = 2048;

// One time step.
template <class T>
int step_t(alfa_plan* P, const alfa_state* st, int it, double sfc_val, hipStream_t s,
           const double* sfc_dev = nullptr) {
  Ctx<T> c(P, s);
  const auto& g = P->cfg;
  const auto& d = P->d;
  StateT<T> S;
  S.u = (T*)st->u; S.v = (T*)st->v; S.w = (T*)st->w; S.TH = (T*)st->TH;
  S.Ru = (T*)st->RHS_u; S.Rv = (T*)st->RHS_v; S.Rw = (T*)st->RHS_w; S.RT = (T*)st->RHS_TH;
  S.Cs2 = (T*)st->Cs2_3D; S.CsPr = (T*)st->Cs2PrRatio_3D;
  S.psim = (T*)st->psi_m; S.psim0 = (T*)st->psi_m0; S.psih = (T*)st->psi_h;
  S.psih0 = (T*)st->psi_h0; S.fim = (T*)st->fi_m; S.fih = (T*)st->fi_h;
  const bool first = it == g.istep;
  const bool dealias = d.optDealias == 1;
  const bool mixed = P->prec == ALFA_MIXED;
  const size_t PB = (size_t)c.ncol * sizeof(T);
  Surf<T> sf;
  sf.M = (T*)P->ws.get("pl_M", PB); sf.ustar = (T*)P->ws.get("pl_ustar", PB);
  sf.qz = (T*)P->ws.get("pl_qz", PB); sf.invOB = (T*)P->ws.get("pl_invOB", PB);
  T *dudx = c.W("dudx"), *dvdx = c.W("dvdx"), *dwdx = c.W("dwdx");
  T *dudy = c.W("dudy"), *dvdy = c.W("dvdy"), *dwdy = c.W("dwdy");
  T *dudz = c.W("dudz"), *dvdz = c.W("dvdz"), *dwdz = c.W("dwdz");
  T *dTx = c.W("dTHdx"), *dTy = c.W("dTHdy"), *dTz = c.W("dTHdz");
  // moisture (optMoisture >= 1): Q and its RHS_Q_previous in the state
  const bool moist = g.optMoisture >= 1;
  T* SQ = moist ? (T*)st->Q : nullptr;
  T* SRQ = moist ? (T*)st->RHS_Q : nullptr;
  T *dQx = nullptr, *dQy = nullptr, *dQz = nullptr;
  if (moist) { dQx = c.W("dQdx"); dQy = c.W("dQdy"); dQz = c.W("dQdz"); }

  // 1. explicit filter of u, v, w, TH (and Q) and the spectral gradients
  //    (Main.py:193-199, Derivatives.py:47-153, 281-306)
  {
    const int nf = moist ? 5 : 4;
    c2<T>* ef[5];
    c2<T>* ex[5];
    for (int b = 0; b < nf; ++b) {
      ef[b] = c.S("ef" + std::to_string(b));
      ex[b] = c.S("ex" + std::to_string(b));
    }
    LoadPtr<T> ld{};
    ld.p[0] = S.u; ld.p[1] = S.v; ld.p[2] = S.w; ld.p[3] = S.TH; ld.p[4] = SQ;
    ld.ny = c.ny; ld.nz = c.nz;
    y_r2c<T>(P, s, ld, c.nx, P->fy, P->tw_y, c.nz, d.cut0_y, nf, (void* const*)ef);
    XArgs xa{};
    for (int b = 0; b < nf; ++b) {
      xa.in[b] = ef[b]; xa.dst[b][0] = ef[b]; xa.dst[b][1] = ex[b];
    }
    xa.op[0] = xop(d.cut0_y, 1, d.cut0_x, XM_NONE, XMAP_ID, c.nx);
    xa.op[1] = xop(d.cut0_y, 1, d.cut0_x, XM_IKX, XMAP_ID, c.nx);
    xa.nops = 2; xa.ncy_in = d.cut0_y; xa.nz = c.nz; xa.ny_full = c.ny;
    x_pass<T>(P, s, xa, nf, P->fx, P->fx, P->tw_x, P->tw_x);
    // the y derivatives are the filtered spectra times i*ky, applied in the
    // y C2R (i*ky commutes with the x transform)
    YC2RArgs ya{};
    T* outs[15] = {S.u, S.v, S.w, S.TH, SQ, dudx, dvdx, dwdx, dTx, dQx,
                   dudy, dvdy, dwdy, dTy, dQy};
    T* o[15];
    int no = 0;
    for (int part = 0; part < 3; ++part)
      for (int b = 0; b < nf; ++b) {
        o[no] = outs[part * 5 + b];
        ya.in[no][0] = part == 1 ? ex[b] : ef[b];
        ya.dy[no] = part == 2 ? 1 : 0;
        ++no;
      }
    ya.nsum = 1; ya.ncy_in = d.cut0_y; ya.nz = c.nz; ya.scale = 1.0 / ((double)c.nx * c.ny);
    ya.dyscale = 1.0;
    StorePtr<T> so{};
    for (int b = 0; b < no; ++b) so.p[b] = o[b];
    so.ny = c.ny; so.nz = c.nz;
    y_c2r<T>(P, s, so, ya, no, c.nx, P->fy, P->tw_y);
  }
  // 2. surface fluxes (Main.py:213-243) and the moisture flux (:253-266)
  stage_surface<T>(c, S, sf, sfc_val, sfc_dev);
  T* qm = nullptr;
  if (moist) {
    qm = (T*)P->ws.get("pl_qm", PB);
    double* qa0 = c.D("qm_avg0", 1);
    if (g.optMoistureSurfBC == 2 && g.optSurfFlux == 0) {
      FLevel0<T> f0{SQ, c.nz};
      plane_sums<1>(P, s, c.ncol, 1, f0, qa0, 1.0 / c.ncol, "qm0");
    }
    const double zref = g.zMoisture > 0.0 ? g.zMoisture : g.z0T;
    KLAUNCH(P, s, "k_surf_moist", 5.0 * PB, k_surf_moist<T>, pw_blocks(c.ncol), 256, c.ncol,
            c.nz, g.optMoistureSurfBC, g.optSurfFlux != 0 ? 1 : 0, (const T*)SQ,
            (const T*)sf.ustar, (const T*)S.psih, (const T*)S.psih0, (const double*)qa0,
            st->Q_sfc_val, g.MoistureFlux, 0.4, std::log(0.5 * d.dz * d.z_scale / zref), qm);
    FLevel0<T> fq{qm, 1};
    plane_sums<1>(P, s, c.ncol, 1, fq, P->scal + 4, 1.0 / c.ncol, "qmavg");
  }
  // 3. vertical gradients (Derivz_M, Derivz_TH); on the unpadded grid they
  // are formed in the advection pass of step 4 (k_dz_adv_strains)
  if (dealias) {
    if (vec_ok<T>(c.nz, {S.u, S.v, S.w, S.TH, dudz, dvdz, dwdz, dTz}))
      PW_LAUNCH_W(P, s, c.F, Vec16<T>::W, k_derivz_v<T>, c.dims(), S.u, S.v, S.w, S.TH,
                  sf.ustar, sf.M, S.fim, S.fih, sf.qz, dudz, dvdz, dwdz, dTz, (T)g.Ugal,
                  (T)(1.0 / d.dz), (T)(0.4 * 0.5 * d.dz), (T)(0.4 * 0.5 * d.dz));
    else
      PW_LAUNCH(P, s, c.F, k_derivz<T>, c.dims(), S.u, S.v, S.w, S.TH, sf.ustar, sf.M, S.fim,
                S.fih, sf.qz, dudz, dvdz, dwdz, dTz, (T)g.Ugal, (T)(1.0 / d.dz),
                (T)(0.4 * 0.5 * d.dz), (T)(0.4 * 0.5 * d.dz));
  }
  if (moist) {   // Derivz_TH of Q with the moisture flux (Derivatives.py:304)
    if (vec_ok<T>(c.nz, {SQ, dQz}))
      PW_LAUNCH_W(P, s, c.F, Vec16<T>::W, k_derivz_v<T>, c.dims(), S.u, S.v, S.w, SQ, sf.ustar,
                  sf.M, S.fim, S.fih, qm, (T*)nullptr, (T*)nullptr, (T*)nullptr, dQz,
                  (T)g.Ugal, (T)(1.0 / d.dz), (T)(0.4 * 0.5 * d.dz), (T)(0.4 * 0.5 * d.dz));
    else
      PW_LAUNCH(P, s, c.F, k_derivz<T>, c.dims(), S.u, S.v, S.w, SQ, sf.ustar, sf.M, S.fim,
                S.fih, qm, (T*)nullptr, (T*)nullptr, (T*)nullptr, dQz, (T)g.Ugal,
                (T)(1.0 / d.dz), (T)(0.4 * 0.5 * d.dz), (T)(0.4 * 0.5 * d.dz));
  }
  // plane means of u, v, TH, THv = (TH + T_0)(1 + 0.61 Q) (and Q)
  double* means = c.D("means", (moist ? 5 : 4) * (size_t)c.nz);
  if (moist) {
    FSum5<T> f{S.u, S.v, S.TH, SQ, c.nz, g.T_0};
    plane_sums<5>(P, s, c.ncol, c.nz, f, means, 1.0 / c.ncol, "means");
  } else {
    FSum4<T> f{S.u, S.v, S.TH, c.nz, g.T_0};
    plane_sums<4>(P, s, c.ncol, c.nz, f, means, 1.0 / c.ncol, "means");
  }
  // 4. advection (Main.py:301-311)
  T *Cx = c.W("Cx"), *Cy = c.W("Cy"), *Cz = c.W("Cz"), *THadv = c.W("THadv");
  T* QAdv = moist ? c.W("QAdv") : nullptr;
  if (!dealias) {
    // with the vertical gradients of step 3 and the uvp-node strain rates of
    // step 6 in the same pass
    DzAdvStrainArgs<T> aa{};
    aa.u = S.u; aa.v = S.v; aa.w = S.w; aa.TH = S.TH; aa.sx = dTx; aa.sy = dTy;
    aa.dudx = dudx; aa.dvdx = dvdx; aa.dwdx = dwdx; aa.dudy = dudy; aa.dvdy = dvdy;
    aa.dwdy = dwdy;
    aa.ustar = sf.ustar; aa.M = sf.M; aa.fim = S.fim; aa.fih = S.fih; aa.qz = sf.qz;
    aa.Ugal = (T)g.Ugal; aa.idz = (T)(1.0 / d.dz); aa.half_dz_vonk = (T)(0.4 * 0.5 * d.dz);
    aa.dudz = dudz; aa.dvdz = dvdz; aa.dwdz = dwdz; aa.dTdz = dTz;
    aa.Cx = Cx; aa.Cy = Cy; aa.Cz = Cz; aa.sadv = THadv;
    aa.S12 = c.W("S12"); aa.S13 = c.W("S13"); aa.S23 = c.W("S23"); aa.S = c.W("S");
    if (vec_ok<T>(c.nz, {S.u, S.v, S.w, S.TH, dTx, dTy, dudx, dvdx, dwdx, dudy, dvdy, dwdy,
                         dudz, dvdz, dwdz, dTz, Cx, Cy, Cz, THadv, aa.S12, aa.S13, aa.S23, aa.S}))
      PW_LAUNCH_W(P, s, c.F, Vec16<T>::W, k_dz_adv_strains_v<T>, c.dims(), aa);
    else
      PW_LAUNCH(P, s, c.F, k_dz_adv_strains<T>, c.dims(), aa);
    if (moist)   // ScalarAdvection of Q (Main.py:313-318)
      PW_LAUNCH(P, s, c.F, k_scalar_adv<T>, c.dims(), S.u, S.v, S.w, dQx, dQy, dQz, QAdv);
  } else {
    const T* in9[9] = {S.u, S.v, S.w, dudy, dudz, dvdx, dvdz, dwdx, dwdy};
    T* p9[9];
    for (int b = 0; b < 9; ++b) p9[b] = c.Pad("adv" + std::to_string(b));
    dealias1<T>(c, 9, in9, p9);
    T *pcx = c.Pad("cx"), *pcy = c.Pad("cy"), *pcz = c.Pad("cz");
    PW_LAUNCH(P, s, (size_t)c.mx * c.my * c.nz, k_advection<T>, c.pdims(), p9[0], p9[1], p9[2],
              p9[3], p9[4], p9[5], p9[6], p9[7], p9[8], pcx, pcy, pcz);
    const T* pc3[3] = {pcx, pcy, pcz};
    T* o3[3] = {Cx, Cy, Cz};
    dealias2<T>(c, 3, pc3, o3);
    const T* in3[3] = {dTx, dTy, dTz};
    T* q3[3] = {c.Pad("sadv0"), c.Pad("sadv1"), c.Pad("sadv2")};
    dealias1<T>(c, 3, in3, q3);
    T* pcs = c.Pad("cs");
    PW_LAUNCH(P, s, (size_t)c.mx * c.my * c.nz, k_scalar_adv<T>, c.pdims(), p9[0], p9[1], p9[2],
              q3[0], q3[1], q3[2], pcs);
    const T* pc1[1] = {pcs};
    T* o1[1] = {THadv};
    dealias2<T>(c, 1, pc1, o1);
    if (moist) {
      const T* inq[3] = {dQx, dQy, dQz};
      T* pq[3] = {c.Pad("qadv0"), c.Pad("qadv1"), c.Pad("qadv2")};
      dealias1<T>(c, 3, inq, pq);
      T* pcq = c.Pad("cq");
      PW_LAUNCH(P, s, (size_t)c.mx * c.my * c.nz, k_scalar_adv<T>, c.pdims(), p9[0], p9[1],
                p9[2], pq[0], pq[1], pq[2], pcq);
      const T* pq1[1] = {pcq};
      T* oq[1] = {QAdv};
      dealias2<T>(c, 1, pq1, oq);
    }
  }
  // 5. buoyancy (Main.py:322-327) is formed inside the momentum RHS kernel
  // 6. SGS (Main.py:333-443)
  const bool dynamic = (g.optSgs >= 1 && g.optSgs <= 4) &&
                       (first || (g.dynamicSGS_call_time > 0 && it % g.dynamicSGS_call_time == 0));
  const T* grads[9] = {dudx, dvdx, dwdx, dudy, dvdy, dwdy, dudz, dvdz, dwdz};
  SgsStrains<T> ss{};
  // the moisture fluxes need the w-node |S| of the unpadded grid as well
  stage_strains<T>(c, grads, ss, dealias || moist, !dealias);
  const double* shu = mixed ? means : nullptr;
  const double* shv = mixed ? means + c.nz : nullptr;
  const double* sht = mixed ? means + 2 * c.nz : nullptr;
  const bool wl = g.optSgs == 2 || g.optSgs == 4;
  if (dynamic && wl)
    stage_lasdd_wl<T>(c, S.u, S.v, S.w, ss.Sij, shu, shv, S.Cs2, st->Cs2_1D_avg1,
                      st->Cs2_1D_avg2, st->beta1_1D);
  else if (dynamic)
    stage_lasdd<T>(c, S.u, S.v, S.w, ss.Sij, ss.Sm, shu, shv, S.Cs2, st->Cs2_1D_avg1,
                   st->Cs2_1D_avg2, st->beta1_1D);
  T* tau[6] = {c.W("txx"), c.W("tyy"), c.W("tzz"), c.W("txy"), c.W("txz"), c.W("tyz")};
  if (dealias) stage_stresses<T>(c, ss, S.Cs2, S.u, S.v, sf.M, S.psim, S.psim0, tau);
  T *txx = tau[0], *tyy = tau[1], *tzz = tau[2], *txy = tau[3], *txz = tau[4], *tyz = tau[5];
  if (dynamic && wl)
    stage_scalar_wl<T>(c, S.u, S.v, S.w, 0, S.TH, dTx, dTy, dTz, step_hats(c), shu, shv, sht,
                       S.CsPr, st->Cs2PrRatio_1D, st->beta2_1D);
  else if (dynamic)
    stage_scalar_lasdd<T>(c, S.u, S.v, S.w, 0, S.TH, dTx, dTy, dTz, ss.Sm, step_hats(c), shu, shv,
                          sht, S.CsPr, st->Cs2PrRatio_1D, st->beta2_1D);
  T *qx = c.W("qx"), *qy = c.W("qy"), *qzf = c.W("qz");
  if (dealias) {
    const T* dT[3] = {dTx, dTy, dTz};
    T* q[3] = {qx, qy, qzf};
    stage_fluxes<T>(c, ss, dT, S.CsPr, sf.qz, q);
  } else {
    // stresses and fluxes in one pass (same arithmetic as stage_stresses +
    // stage_fluxes on the unpadded grid)
    SgsFusedArgs<T> fa{};
    fa.dudx = dudx; fa.dvdx = dvdx; fa.dwdx = dwdx; fa.dudy = dudy; fa.dvdy = dvdy;
    fa.dwdy = dwdy; fa.dudz = dudz; fa.dvdz = dvdz; fa.dwdz = dwdz;
    fa.S12 = ss.Sij[3]; fa.Sm = wl ? nullptr : ss.Sm; fa.Cs2 = S.Cs2; fa.CsPr = S.CsPr;
    fa.dTx = dTx; fa.dTy = dTy; fa.dTz = dTz;
    fa.pre_s = wl ? (T)(-2.0 * std::pow(d.L, 4.0 / 3.0)) : (T)(-2.0 * d.L * d.L);
    fa.pre_q = wl ? (T)(-std::pow(d.L, 4.0 / 3.0)) : (T)(-2.0 * d.L * d.L);
    fa.u = S.u; fa.v = S.v; fa.M = sf.M; fa.psim = S.psim; fa.psim0 = S.psim0; fa.qsfc = sf.qz;
    fa.logz1z0 = std::log(0.5 * d.dz * d.z_scale / g.z0m); fa.vonk = 0.4; fa.Ugal = g.Ugal;
    fa.txx = txx; fa.tyy = tyy; fa.tzz = tzz; fa.txy = txy; fa.txz = txz; fa.tyz = tyz;
    fa.qx = qx; fa.qy = qy; fa.qz = qzf;
    if (vec_ok<T>(c.nz, {dudx, dvdx, dwdx, dudy, dvdy, dwdy, dudz, dvdz, dwdz, fa.S12, fa.Sm,
                         fa.Cs2, fa.CsPr, dTx, dTy, dTz, txx, tyy, tzz, txy, txz, tyz, qx, qy,
                         qzf}))
      PW_LAUNCH_W(P, s, c.F, Vec16<T>::W, k_sgs_fused_v<T>, c.dims(), fa);
    else
      PW_LAUNCH(P, s, c.F, k_sgs_fused<T>, c.dims(), fa);
  }
  T *dtx = c.W("divtx"), *dty = c.W("divty"), *dtz = c.W("divtz"), *divq = c.W("divq");
  stage_div_stress<T>(c, txx, tyy, tzz, txy, txz, tyz, dtx, dty, dtz);
  stage_div_flux<T>(c, qx, qy, qzf, divq);
  // moisture SGS flux: the scalar model with Cs2PrRatio and qm at level 0
  // (Main.py:356-370, 441-452), and its divergence
  T* divqm = nullptr;
  if (moist) {
    T *qHx = c.W("qHx"), *qHy = c.W("qHy"), *qHz = c.W("qHz");
    const T* dQ[3] = {dQx, dQy, dQz};
    T* q[3] = {qHx, qHy, qHz};
    stage_fluxes<T>(c, ss, dQ, S.CsPr, qm, q);
    divqm = c.W("divqm");
    stage_div_flux<T>(c, qHx, qHy, qHz, divqm);
  }
  // 7. right-hand sides (Main.py:464-473)
  T *Rnu = c.W("Rnu"), *Rnv = c.W("Rnv"), *Rnw = c.W("Rnw"), *RnT = c.W("RnT");
  {
    const double* pr = P->prof;
    if (vec_ok<T>(c.nz, {S.u, S.v, S.w, S.TH, Cx, Cy, Cz, dtx, dty, dtz, THadv, divq, Rnu, Rnv,
                         Rnw, RnT}))
      PW_LAUNCH_W(P, s, c.F, Vec16<T>::W, k_rhs_v<T>, c.dims(), S.u, S.v, S.w, S.TH, Cx, Cy, Cz,
                  dtx, dty, dtz, THadv, divq, pr, pr + c.nz, pr + 2 * c.nz, pr + 3 * c.nz, means,
                  means + c.nz, means + 2 * c.nz, d.f_coriolis_nondim, g.Ugal, g.optDamping, Rnu,
                  Rnv, Rnw, RnT, means + 3 * c.nz, g.T_0, d.g_nondim, g.optBuoyancy, P->fadv(),
                  (const T*)SQ);
    else
      PW_LAUNCH(P, s, c.F, k_rhs<T>, c.dims(), S.u, S.v, S.w, S.TH, Cx, Cy, Cz, (const T*)nullptr,
                dtx, dty, dtz, THadv, divq, pr, pr + c.nz, pr + 2 * c.nz, pr + 3 * c.nz, means,
                means + c.nz, means + 2 * c.nz, d.f_coriolis_nondim, g.Ugal, g.optDamping, Rnu,
                Rnv, Rnw, RnT, means + 3 * c.nz, g.T_0, d.g_nondim, g.optBuoyancy,
                P->fadv(), (const T*)SQ);
  }
  // RHS_Moisture (SCL_AllTerms.py:86-110): the RT path of k_rhs on Q, its
  // advection, div qm, the Qadv forcing and the sponge about the Q mean
  T* RnQ = nullptr;
  if (moist) {
    RnQ = c.W("RnQ");
    const double* pr = P->prof;
    PW_LAUNCH(P, s, c.F, k_rhs<T>, c.dims(), (const T*)nullptr, (const T*)nullptr,
              (const T*)nullptr, (const T*)SQ, (const T*)nullptr, (const T*)nullptr,
              (const T*)nullptr, (const T*)nullptr, (const T*)nullptr, (const T*)nullptr,
              (const T*)nullptr, (const T*)QAdv, (const T*)divqm, pr, pr + c.nz, pr + 2 * c.nz,
              pr + 3 * c.nz, means, means + c.nz, means + 4 * c.nz, 0.0, 0.0, g.optDamping,
              (T*)nullptr, (T*)nullptr, (T*)nullptr, RnQ, (const double*)nullptr, 0.0, 0.0, 0,
              P->fadvQ());
  }
  // 8. pressure (Main.py:481-500)
  T *dpdx = c.W("dpdx"), *dpdy = c.W("dpdy"), *dpdz = c.W("dpdz");
  {
    const T* R[3] = {Rnu, Rnv, Rnw};
    const T* Rp[3] = {S.Ru, S.Rv, S.Rw};
    const T* U[3] = {S.u, S.v, S.w};
    stage_pressure<T>(c, R, Rp, U, dtz, dpdx, dpdy, dpdz, nullptr, nullptr, nullptr, nullptr);
  }
  // 9. AB2 + CFL (Main.py:506-540)
  double* cflpart = c.D("cflpart", 3 * AB2_BLOCKS);
  if (vec_ok<T>(c.nz, {S.u, S.v, S.w, S.TH, Rnu, Rnv, Rnw, RnT, S.Ru, S.Rv, S.Rw, S.RT, dpdx,
                       dpdy, dpdz}))
    KLAUNCH(P, s, "k_ab2", 23.0 * c.FB, k_ab2_v<T>, AB2_BLOCKS, 256, c.dims(), S.u, S.v, S.w, S.TH,
          Rnu, Rnv, Rnw, RnT, S.Ru, S.Rv, S.Rw, S.RT, dpdx, dpdy, dpdz, (T)d.dt_nondim,
          (T)(d.inversion_nondim * d.dz), first ? 1 : 0, cflpart);
  else
    KLAUNCH(P, s, "k_ab2", 23.0 * c.FB, k_ab2<T>, AB2_BLOCKS, 256, c.dims(), S.u, S.v, S.w, S.TH,
          Rnu, Rnv, Rnw, RnT, S.Ru, S.Rv, S.Rw, S.RT, dpdx, dpdy, dpdz, (T)d.dt_nondim,
          (T)(d.inversion_nondim * d.dz), first ? 1 : 0, cflpart);
  if (moist)   // AB2_Q (SCL_TimeAdvancement.py:79-99), RHS_Q rotation
    PW_LAUNCH(P, s, c.F, k_ab2_q<T>, c.dims(), SQ, (const T*)RnQ, SRQ, (T)d.dt_nondim,
              (T)(g.q_inversion * d.z_scale * d.dz), first ? 1 : 0);
  if (live(P)) {
    launch_k<1024>(P, s, dim3(1), dim3(256), 0, KBODY(k_cfl_final), cflpart, AB2_BLOCKS, d.dt_nondim / d.dx, d.dt_nondim / d.dy,
                                  d.dt_nondim / d.dz, P->scal, st->scalars);
    launch_ok(P);
  }
  return 0;
}

}  // namespace alfa
