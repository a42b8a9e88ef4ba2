// Planar statistics of one sample: ComputeStats with ResetFlag = 0
// (src/utilities/Statistics.py:47-330), u_scale = TH_scale = Q_scale = 1
// (src/config/DerivedVars.py:49-69); with moisture also the Q statistics.
//
// Two deterministic plane-sum passes in double over the state after the
// step and the step's own intermediates (still in the plan workspace):
//   pass 1, 14 level means: u, v, w, TH, dudz, dvdz, dTHdz, txy, txz, tyz,
//           qz and the w-node u, v, TH of the vertical fluxes (level 0 taken
//           as is, u with Ugal added; :226-240);
//   pass 2, 10 fluctuation moments about those means: u2, v2, w2, TH2, uv,
//           uTH, vTH (:188-213) and uw, vw, wTH on the w nodes (:245-265);
// a plane pass for the surface (mean M, mean ustar^2; :288-289), and one
// small kernel that adds the sample into the caller's accumulator.
#pragma once
#include "step.hpp"

namespace alfa {

// accumulator layout: STATS_NPROF profiles of nz and then STATS_NSCAL
// scalars (the order of the C header)
constexpr int STATS_NPROF = 31;   // 26 dry profiles, then Q, dQdz, Q2, wQ, qHz
constexpr int STATS_NSCAL = 4;    // M_sfc, ustar, qz_sfc, qm_sfc

template <class T> struct FStatsMeans {
  static constexpr const char* NAME = "FStatsMeans";
  static constexpr int NFIELDS = 11;
  const T *u, *v, *w, *TH, *dudz, *dvdz, *dTHdz, *txy, *txz, *tyz, *qz;
  int nz;
  double Ugal;
  __device__ __forceinline__ void operator()(int c, int z, double* acc) const {
    const size_t i = (size_t)c * nz + z;
    const double uu = u[i], vv = v[i], tt = TH[i];
    acc[0] += uu; acc[1] += vv; acc[2] += (double)w[i]; acc[3] += tt;
    acc[4] += (double)dudz[i]; acc[5] += (double)dvdz[i]; acc[6] += (double)dTHdz[i];
    acc[7] += (double)txy[i]; acc[8] += (double)txz[i]; acc[9] += (double)tyz[i];
    acc[10] += (double)qz[i];
    if (z == 0) {
      acc[11] += uu + Ugal; acc[12] += vv; acc[13] += tt;
    } else {
      // StagGridAvg in the working precision (Utilities.py:64-79)
      acc[11] += (double)(T(0.5) * (u[i - 1] + u[i]));
      acc[12] += (double)(T(0.5) * (v[i - 1] + v[i]));
      acc[13] += (double)(T(0.5) * (TH[i - 1] + TH[i]));
    }
  }
};

template <class T> struct FStatsMoments {
  static constexpr const char* NAME = "FStatsMoments";
  static constexpr int NFIELDS = 4;
  const T *u, *v, *w, *TH;
  const double* m;   // pass-1 means, (14, nz)
  int nz;
  double Ugal;
  __device__ __forceinline__ void operator()(int c, int z, double* acc) const {
    const size_t i = (size_t)c * nz + z;
    const double mu = m[z] + Ugal, mv = m[nz + z], mw = m[2 * nz + z], mt = m[3 * nz + z];
    const double uf = ((double)u[i] + Ugal) - mu, vf = (double)v[i] - mv;
    const double wf = (double)w[i] - mw, tf = (double)TH[i] - mt;
    acc[0] += uf * uf; acc[1] += vf * vf; acc[2] += wf * wf; acc[3] += tf * tf;
    acc[4] += uf * vf; acc[5] += uf * tf; acc[6] += vf * tf;
    double us, vs, ts;
    if (z == 0) {
      us = (double)u[i] + Ugal; vs = v[i]; ts = TH[i];
    } else {
      us = (double)(T(0.5) * (u[i - 1] + u[i]));
      vs = (double)(T(0.5) * (v[i - 1] + v[i]));
      ts = (double)(T(0.5) * (TH[i - 1] + TH[i]));
    }
    const double usf = (us + Ugal) - (m[11 * nz + z] + Ugal);
    const double vsf = vs - m[12 * nz + z], tsf = ts - m[13 * nz + z];
    acc[7] += usf * wf; acc[8] += vsf * wf; acc[9] += wf * tsf;
  }
};

// moisture (Statistics.py:175-176, 206-207, 235-264, 275): level means of Q,
// dQdz, qHz and the w-node Q; then Q2 and wQ about them
template <class T> struct FStatsMeansQ {
  static constexpr const char* NAME = "FStatsMeansQ";
  static constexpr int NFIELDS = 3;
  const T *Q, *dQdz, *qHz;
  int nz;
  __device__ __forceinline__ void operator()(int c, int z, double* acc) const {
    const size_t i = (size_t)c * nz + z;
    acc[0] += (double)Q[i];
    acc[1] += (double)dQdz[i];
    acc[2] += (double)qHz[i];
    acc[3] += z == 0 ? (double)Q[i] : (double)(T(0.5) * (Q[i - 1] + Q[i]));
  }
};

template <class T> struct FStatsMomentsQ {
  static constexpr const char* NAME = "FStatsMomentsQ";
  static constexpr int NFIELDS = 2;
  const T *Q, *w;
  const double* m;    // dry pass-1 means (w at row 2)
  const double* mq;   // moisture means (4, nz)
  int nz;
  __device__ __forceinline__ void operator()(int c, int z, double* acc) const {
    const size_t i = (size_t)c * nz + z;
    const double qf = (double)Q[i] - mq[z];
    const double wf = (double)w[i] - m[2 * nz + z];
    const double qs = z == 0 ? (double)Q[i] : (double)(T(0.5) * (Q[i - 1] + Q[i]));
    acc[0] += qf * qf;
    acc[1] += wf * (qs - mq[3 * nz + z]);
  }
};

// surface planes: M and ustar^2 (nz = 1 reduction over the columns)
template <class T> struct FStatsSurface {
  static constexpr const char* NAME = "FStatsSurface";
  static constexpr int NFIELDS = 0;
  const T *M, *ustar;
  __device__ __forceinline__ void operator()(int c, int, double* acc) const {
    const double us = ustar[c];
    acc[0] += (double)M[c];
    acc[1] += us * us;
  }
};

struct StatsProfiles {
  const double *cs1, *cs2, *cspr, *b1, *b2;   // may be NULL (0)
};

// acc += the sample (Statistics.py:158-305): means, gradients / z_scale,
// moments, SGS profiles, surface scalars and the coefficient profiles.
__device__ __forceinline__ void k_stats_add(int nz, const double* __restrict__ m, const double* __restrict__ mo,
                            const double* __restrict__ sfc, const double* __restrict__ scal,
                            const StatsProfiles& pr, double Ugal, double T0, double z_scale,
                            const double* __restrict__ mq, const double* __restrict__ moq,
                            double* acc) {
  for (int z = threadIdx.x; z < nz; z += blockDim.x) {
    double* a = acc;
    a[0 * nz + z] += m[z] + Ugal;                 // U
    a[1 * nz + z] += m[nz + z];                   // V
    a[2 * nz + z] += m[2 * nz + z];               // W
    a[3 * nz + z] += m[3 * nz + z] + T0;          // TH
    a[4 * nz + z] += m[4 * nz + z] / z_scale;     // dUdz
    a[5 * nz + z] += m[5 * nz + z] / z_scale;     // dVdz
    a[6 * nz + z] += m[6 * nz + z] / z_scale;     // dTHdz
    a[7 * nz + z] += mo[z];                       // u2
    a[8 * nz + z] += mo[nz + z];                  // v2
    a[9 * nz + z] += mo[2 * nz + z];              // w2
    a[10 * nz + z] += mo[3 * nz + z];             // TH2
    a[11 * nz + z] += mo[4 * nz + z];             // uv
    a[12 * nz + z] += mo[7 * nz + z];             // uw
    a[13 * nz + z] += mo[8 * nz + z];             // vw
    a[14 * nz + z] += mo[5 * nz + z];             // uTH
    a[15 * nz + z] += mo[6 * nz + z];             // vTH
    a[16 * nz + z] += mo[9 * nz + z];             // wTH
    a[17 * nz + z] += m[7 * nz + z];              // txy
    a[18 * nz + z] += m[8 * nz + z];              // txz
    a[19 * nz + z] += m[9 * nz + z];              // tyz
    a[20 * nz + z] += m[10 * nz + z];             // qz
    a[21 * nz + z] += pr.cs1 ? pr.cs1[z] : 0.0;   // Cs2_1
    a[22 * nz + z] += pr.cs2 ? pr.cs2[z] : 0.0;   // Cs2_2
    a[23 * nz + z] += pr.cspr ? pr.cspr[z] : 0.0; // Cs2PrRatio
    a[24 * nz + z] += pr.b1 ? pr.b1[z] : 0.0;     // Beta1
    a[25 * nz + z] += pr.b2 ? pr.b2[z] : 0.0;     // Beta2
    if (mq) {   // moisture (dry runs add nothing: the profiles stay 0)
      a[26 * nz + z] += mq[z];                    // Q
      a[27 * nz + z] += mq[nz + z] / z_scale;     // dQdz
      a[28 * nz + z] += moq[z];                   // Q2
      a[29 * nz + z] += moq[nz + z];              // wQ
      a[30 * nz + z] += mq[2 * nz + z];           // qHz
    }
  }
  if (threadIdx.x == 0) {
    double* s = acc + STATS_NPROF * nz;
    s[0] += sfc[0];             // M_sfc
    s[1] += sqrt(sfc[1]);       // ustar: sqrt of the mean of ustar^2
    s[2] += scal[1];            // qz_sfc (qz_sfc_avg of the step)
    if (mq) s[3] += scal[4];    // qm_sfc (qm_sfc_avg of the step)
  }
}

// One sample into acc (STATS_NPROF * nz + STATS_NSCAL doubles): must follow
// the alfa_step of the same plan on the same stream (it reads that step's
// gradients, stresses, flux and surface planes from the workspace).
template <class T>
void stats_accumulate(alfa_plan* P, hipStream_t s, const alfa_state* st, double* acc) {
  Ctx<T> c(P, s);
  const auto& g = P->cfg;
  const auto& d = P->d;
  const size_t PB = (size_t)c.ncol * sizeof(T);
  double* m = c.D("st_means", 14 * (size_t)c.nz);
  double* mo = c.D("st_moments", 10 * (size_t)c.nz);
  double* sfc = c.D("st_sfc", 2);
  FStatsMeans<T> f1{};
  f1.u = (const T*)st->u; f1.v = (const T*)st->v; f1.w = (const T*)st->w;
  f1.TH = (const T*)st->TH;
  f1.dudz = c.W("dudz"); f1.dvdz = c.W("dvdz"); f1.dTHdz = c.W("dTHdz");
  f1.txy = c.W("txy"); f1.txz = c.W("txz"); f1.tyz = c.W("tyz"); f1.qz = c.W("qz");
  f1.nz = c.nz; f1.Ugal = g.Ugal;
  plane_sums<14>(P, s, c.ncol, c.nz, f1, m, 1.0 / c.ncol, "stats1");
  FStatsMoments<T> f2{};
  f2.u = f1.u; f2.v = f1.v; f2.w = f1.w; f2.TH = f1.TH; f2.m = m; f2.nz = c.nz; f2.Ugal = g.Ugal;
  plane_sums<10>(P, s, c.ncol, c.nz, f2, mo, 1.0 / c.ncol, "stats2");
  FStatsSurface<T> f3{};
  f3.M = (const T*)P->ws.get("pl_M", PB);
  f3.ustar = (const T*)P->ws.get("pl_ustar", PB);
  plane_sums<2>(P, s, c.ncol, 1, f3, sfc, 1.0 / c.ncol, "stats3");
  double *mq = nullptr, *moq = nullptr;
  if (g.optMoisture >= 1) {
    mq = c.D("st_meansq", 4 * (size_t)c.nz);
    moq = c.D("st_momentsq", 2 * (size_t)c.nz);
    FStatsMeansQ<T> fq{(const T*)st->Q, c.W("dQdz"), c.W("qHz"), c.nz};
    plane_sums<4>(P, s, c.ncol, c.nz, fq, mq, 1.0 / c.ncol, "stats4");
    FStatsMomentsQ<T> fq2{(const T*)st->Q, f1.w, m, mq, c.nz};
    plane_sums<2>(P, s, c.ncol, c.nz, fq2, moq, 1.0 / c.ncol, "stats5");
  }
  if (!live(P)) return;
  StatsProfiles pr{st->Cs2_1D_avg1, st->Cs2_1D_avg2, st->Cs2PrRatio_1D, st->beta1_1D,
                   st->beta2_1D};
  launch_k<1024>(P, s, dim3(1), dim3(256), 0, KBODY(k_stats_add), c.nz, m, mo, sfc, P->scal, pr, g.Ugal, g.T_0, d.z_scale, mq,
                                moq, acc);
  launch_ok(P);
}

}  // namespace alfa
