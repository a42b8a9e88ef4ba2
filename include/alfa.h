/*
 * alfa.h — C ABI of the MI355X-native JAX-ALFA time-step library (libalfa.so).
 *
 * Every entry point is stream-ordered and asynchronous: it enqueues HIP
 * kernels on `stream` and returns.  No entry point allocates device memory
 * or synchronises the host after alfa_plan_create(), so a sequence of calls
 * can be captured into a hipGraph.  All array arguments are device pointers
 * owned by the caller; the library keeps none of them after return.
 *
 * Layout (identical to the reference's JAX arrays):
 *   real field   (nx, ny, nz)            row-major, z fastest
 *   spectrum     (nx, ny/2+1, nz)        interleaved complex, z fastest
 *   padded field (mx, my, nz)            mx = int(1.5 nx), my = int(1.5 ny)
 *   plane        (nx, ny)
 *   profile      (nz,)                   always double
 * Element type of fields/spectra/planes follows the plan precision:
 *   ALFA_FP64  -> double / complex double everywhere;
 *   ALFA_MIXED -> float / complex float for fields, spectra and planes;
 *                 plane-mean accumulators, the beta root solve, the pressure
 *                 tridiagonal factors and sweeps and the surface-layer
 *                 scalars are double.
 *
 * Return value: 0 on success, otherwise an ALFA_E* code; the message is in
 * alfa_last_error() (thread-local).  Kernels never abort.
 *
 * Citations are to the reference repository (Sukantabasu/jax-alfa).
 */
#ifndef ALFA_H
#define ALFA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct alfa_plan alfa_plan;
typedef void* alfa_stream; /* a hipStream_t; NULL = the null stream */

enum { ALFA_FP64 = 0, ALFA_MIXED = 1 };

enum {
  ALFA_OK = 0,
  ALFA_EINVAL = 1,     /* bad shape, option or argument            */
  ALFA_ERADIX = 2,     /* FFT length with a prime factor above 5    */
  ALFA_EHIP = 3,       /* HIP runtime error                         */
  ALFA_ENOMEM = 4,     /* device allocation failed at plan creation */
  ALFA_ESTATE = 5      /* call out of order (e.g. step on a bad plan) */
};

/* Run configuration; the field names and meanings are those of the
 * reference's Config.py (src/config/ConfigLoader.py:67-157,
 * src/config/DerivedVars.py:39-119).  Lengths in m, times in s. */
typedef struct alfa_config {
  int32_t nx, ny, nz;
  int32_t optSgs;               /* 0 static SM, 1 LASDD-SM, 2 LASDD-WL,
                                   3 LAD-SM, 4 LAD-WL (DynamicSGS_Main.py:25-27) */
  int32_t FGR;                  /* 1 (dealiased, TFR=2) or 2 (TFR=sqrt 2)   */
  int32_t dynamicSGS_call_time;
  int32_t optSurfFlux;          /* 0 homogeneous, 1 heterogeneous           */
  int32_t optSurfBC;            /* 0 const flux, 1 series flux, 2 series T  */
  int32_t optBuoyancy;          /* 0: /T_0, 1: /<theta_v>                   */
  int32_t optDamping;           /* 1: Rayleigh sponge                       */
  int32_t istep;                /* first iteration number                   */
  int32_t optMoisture;          /* 1: specific humidity Q is prognostic      */
  int32_t optMoistureSurfBC;    /* 0 const MoistureFlux, 1 series flux,
                                   2 series surface Q (ConfigLoader.py:102-124) */
  int32_t reserved_i[5];
  double l_x, l_y, l_z;
  double dt;
  double Ugal;
  double z0m, z0T, zTemperature;
  double SensibleHeatFlux;
  double Ug2, Vg2;
  double f_coriolis;
  double inversion;
  double T_0;
  double Cs2, Cs2PrRatio;       /* initial coefficient fields; for the WL
                                   models pass Cwl and CwlPrRatio here
                                   (DerivedVars.py:101-106) */
  double z_damping, RelaxTime;
  double MoistureFlux;          /* kg/kg m/s (optMoistureSurfBC 0)           */
  double zMoisture;             /* reference height of the Q flux, 0: z0T    */
  double q_inversion;           /* Q lapse above the top (kg/kg/m)           */
  double reserved_d[5];
} alfa_config;

/* Derived constants the plan computed (DerivedVars.py:49-119,
 * Preprocess.py:44-62, Filtering.py cut-offs). */
typedef struct alfa_derived {
  double z_scale, dx, dy, dz, L, TFR, dt_nondim, g_nondim, f_coriolis_nondim;
  double inversion_nondim, z_damping_nondim, RelaxTime_nondim;
  int32_t mx, my, nx_rfft, ny_rfft, optDealias;
  int32_t cut0_x, cut0_y, cut1_x, cut1_y, cut2_x, cut2_y;
  int32_t precision;
  int64_t workspace_bytes;
} alfa_derived;

/* Carried state of one run (src/Main.py:66-111).  Device pointers, caller
 * owned, element type per plan precision.  u is stored minus Ugal and TH
 * minus T_0 (src/initialization/Initialization.py:58-100).  The step
 * updates every field in place. */
typedef struct alfa_state {
  void *u, *v, *w, *TH;                          /* (nx,ny,nz)          */
  void *RHS_u, *RHS_v, *RHS_w, *RHS_TH;          /* previous-step RHS   */
  void *Cs2_3D, *Cs2PrRatio_3D;                  /* (nx,ny,nz)          */
  void *psi_m, *psi_m0, *psi_h, *psi_h0, *fi_m, *fi_h; /* (nx,ny) MOST   */
  /* optional diagnostics written by the step when non-NULL: */
  double *Cs2_1D_avg1, *Cs2_1D_avg2, *beta1_1D;  /* (nz,) */
  double *Cs2PrRatio_1D, *beta2_1D;              /* (nz,) */
  double *scalars;   /* [0] CFL, [1] qz_sfc_avg, [2] mean ustar, [3] M_sfc mean,
                        [4] qm_sfc_avg (moisture) */
  /* optMoisture >= 1 (Main.py:71-73, 164-165, 528-529): Q, RHS_Q_previous */
  void *Q, *RHS_Q;                               /* (nx,ny,nz)          */
  /* device pointer to MoistureSurfaceBC_series[iteration-1] (one double)
   * for optMoistureSurfBC 1 (flux) and 2 (surface Q); read by the step */
  const double* Q_sfc_val;
} alfa_state;

/* ---- plan --------------------------------------------------------------- */
/* sizeof(alfa_config), sizeof(alfa_state), sizeof(alfa_derived) as the
 * library was built, for bindings to check their struct layouts.  No
 * counterpart in the reference. */
int alfa_struct_sizes(size_t out[3]);
int alfa_plan_create(const alfa_config* cfg, int precision, alfa_plan** out);
int alfa_plan_destroy(alfa_plan* plan);
int alfa_plan_derived(const alfa_plan* plan, alfa_derived* out);
const char* alfa_last_error(void);
/* Tile choices of the compile-time-length FFT kernels (transform lengths
 * 128, 192, 256; lengths 48 to 96 always run 128-thread kernels): threads
 * per workgroup of the y R2C, x and y C2R kernels, in that order, each 256
 * or 512.  The choice changes timing only:
 * every line is transformed by the same operations, so results are bitwise
 * equal.  No counterpart in the reference (jnp.fft.rfft2/irfft2 leave such
 * choices to XLA's cuFFT plans); they exist for callers that reuse a plan
 * for many steps (ensembles, serving).
 *   alfa_plan_autotune: times each kernel's two choices, one kernel at a
 *   time, on `stream`, switches to the other choice where it is at least 1%
 *   faster and writes the kept choices to `chosen` (nullable).  With a state `st`, it times whole steps
 *   (iterations `iteration`, `iteration`+1, ..., surface value `sfc_val`)
 *   on a scratch copy of it: the caller's state is not modified.  With st
 *   NULL it times a forward transform and three inverses (an x derivative,
 *   the two test-filter levels) of a zero field.  Allocates and frees its
 *   scratch (a copy of the state, or three fields); synchronizes `stream`.
 *   alfa_plan_set_tiles: sets them (0 keeps a choice; else 256 or 512).
 *   alfa_plan_get_tiles: reads them. */
int alfa_plan_autotune(alfa_plan* plan, const alfa_state* st, int iteration, double sfc_val,
                       alfa_stream stream, int chosen[3]);
int alfa_plan_set_tiles(alfa_plan* plan, const int tiles[3]);
int alfa_plan_get_tiles(const alfa_plan* plan, int tiles[3]);
/* Kernel fusion of the dynamic procedures (no reference counterpart; the
 * reference forms every filtered field, DynamicSGS_LASDD_SM.py:184-233 and
 * DynamicSGS_ScalarLASDD_SM.py:179-224, then reduces and contracts them,
 * :266-355 and :241-312).  With fusion on (the default) a float plan with
 * optSgs = 3 (LAD) inverts the level-1 test-filtered fields and contracts
 * them in one kernel per procedure; with optSgs = 1 (LASDD) one kernel per
 * level and procedure first forms the plane sums that give beta the same
 * way.  The filtered fields other than u_hat never reach HBM.  Off, the
 * separate transform, reduction and pointwise kernels run.  The two agree to
 * float rounding (their double accumulations are ordered differently).
 * FP64 plans, the Wong-Lilly models and grids without a compile-time
 * transform length always run the separate kernels.
 *   alfa_plan_set_fusion: on = 1 (fused) or 0 (separate kernels).
 *   alfa_plan_get_fusion: reads the setting. */
int alfa_plan_set_fusion(alfa_plan* plan, int on);
int alfa_plan_get_fusion(const alfa_plan* plan, int* on);
/* Initial state values of Main.py:87-111 (RHS_prev = 0, Cs2 fields =
 * Cs2/Cs2PrRatio, psi = 0, fi = 1) written into caller buffers. */
int alfa_state_init(alfa_plan* plan, const alfa_state* st, alfa_stream stream);

/* ---- whole step (src/Main.py:158-540) ------------------------------------
 * iteration: the reference's loop counter (first step = cfg.istep).
 * sfc_val:   SurfaceBC_series[iteration-1] for optSurfBC 1/2 (Main.py:224,233);
 *            ignored for optSurfBC 0. */
int alfa_step(alfa_plan* plan, const alfa_state* st, int iteration, double sfc_val,
              alfa_stream stream);

/* One step of each member of an ensemble, in table order on one stream:
 * member m is (plans[m], states[m]) at iteration iterations[m] with surface
 * value sfc_vals[m] (sfc_vals may be NULL: 0 for all).  The same result as
 * n alfa_step calls; every member keeps its own plan (grid, precision,
 * workspace).  Replaces the per-run time loop of src/Main.py:156-540 for a
 * set of independent runs (the reference runs one per process). */
int alfa_batch_step(alfa_plan* const* plans, const alfa_state* states, int n,
                    const int* iterations, const double* sfc_vals, alfa_stream stream);

/* Member group: n runs of one configuration shape (same grid, precision,
 * options and FFT tiles; their physical constants, forcing profiles and
 * initial fields may differ) advanced by shared launches.  Each kernel of
 * the step is enqueued once for the whole group; the group's members are
 * the z index of the launch grid and every member reads its own arguments
 * (its state, workspace, constants) from a device table.  Each member's
 * fields are bitwise those of alfa_step on that member alone.
 * sfc_vals[m]: device pointer to member m's surface value of the step (as
 * alfa_step_sfc_dev; NULL allowed where optSurfBC is 0).  The group keeps
 * the plan and state pointers; they must outlive it.
 * alfa_group_step builds the member tables of a kind of step (first step,
 * dynamic or static SGS step, forcing and tiles as set) the first time it
 * meets it: that call allocates and copies (not capturable); later steps of
 * that kind enqueue kernels only and can be captured into a hipGraph.
 * alfa_group_launches: kernels per group step of that kind (-1 before its
 * first step).  Replaces, for ensembles, the per-run loop of
 * src/Main.py:156-540 (the reference runs one case per process). */
typedef struct alfa_group alfa_group;
int alfa_group_create(alfa_plan* const* plans, const alfa_state* states,
                      const double* const* sfc_vals, int n, alfa_group** out);
int alfa_group_step(alfa_group* group, int iteration, alfa_stream stream);
int alfa_group_launches(const alfa_group* group, int iteration, int* n);
int alfa_group_destroy(alfa_group* group);

/* ---- time-varying forcing (src/Main.py:167-188) --------------------------
 * Sets the geostrophic wind profiles Ug, Vg (optGeoWind >= 1:
 * GeoWind_U/V[iteration-1], Initialization.py:173-254) and the large-scale
 * advection tendencies Uadv, Vadv, THadv, Qadv (optAdvection >= 1:
 * AdvForcing_*[iteration-1], Initialization.py:253-323, added to
 * RHS_Momentum / RHS_Scalar / RHS_Moisture, NSE_AllTerms.py:86-88,
 * SCL_AllTerms.py:70-71, 103)
 * used by the following steps and RHS ops of this plan.  Each argument is a
 * device pointer to nz nondimensional doubles, or NULL to keep the current
 * profile (Ug2/Vg2 constants and zero advection at plan creation).  Once an
 * advection profile is given the plan adds the three tendencies to every
 * RHS.  A stream-ordered copy kernel: no allocation, capturable. */
int alfa_set_forcing(alfa_plan* plan, const double* Ug, const double* Vg, const double* Uadv,
                     const double* Vadv, const double* THadv, const double* Qadv,
                     alfa_stream stream);

/* ---- statistics (src/utilities/Statistics.py:47-330; Main.py:545-617) ----
 * Adds one sample of ComputeStats (ResetFlag 0, dry) into acc: the planar
 * statistics of the state after alfa_step and of that step's intermediates
 * (dudz, dvdz, dTHdz, txy, txz, tyz, qz, M_sfc_loc, ustar, qz_sfc_avg; with
 * moisture also Q, dQdz, qHz_q, qm_sfc_avg),
 * which the call reads from the plan workspace: call it after the
 * alfa_step of the same plan on the same stream, before the next step.
 * acc: ALFA_STATS_LEN(nz) doubles, zeroed by the caller to reset (the
 * ResetStats branch), laid out as the ALFA_STATS_NPROF profiles of nz in
 * the order U V W TH dUdz dVdz dTHdz u2 v2 w2 TH2 uv uw vw uTH vTH wTH txy
 * txz tyz qz Cs2_1 Cs2_2 Cs2PrRatio Beta1 Beta2 Q dQdz Q2 wQ qHz, then
 * M_sfc, ustar, qz_sfc, qm_sfc (the moisture entries stay 0 in dry runs).
 * The coefficient profiles are taken from the state's optional pointers
 * (0 where NULL).  Averages are acc / samples (Main.py:589-595). */
#define ALFA_STATS_NPROF 31
#define ALFA_STATS_NSCAL 4
#define ALFA_STATS_LEN(nz) (ALFA_STATS_NPROF * (nz) + ALFA_STATS_NSCAL)
int alfa_stats_accumulate(alfa_plan* plan, const alfa_state* st, double* acc,
                          alfa_stream stream);

/* As alfa_step, but the surface value is read from device memory (one
 * double) when the step's kernels run, so that a step captured once into a
 * hipGraph can be replayed with the next SurfaceBC_series value written
 * there before each replay.  No counterpart in the reference (whose jitted
 * step takes the value as an argument, Main.py:224,233). */
int alfa_step_sfc_dev(alfa_plan* plan, const alfa_state* st, int iteration,
                      const double* sfc_val, alfa_stream stream);

/* ---- operations (src/operations) ---------------------------------------- */
/* FFT (src/operations/FFT.py:44): F (nx,ny,nz) -> F_fft (nx,ny/2+1,nz). */
int alfa_fft(alfa_plan*, const void* F, void* F_fft, alfa_stream);
/* FFT_pad (FFT.py:68): (mx,my,nz) -> (mx,my/2+1,nz). */
int alfa_fft_pad(alfa_plan*, const void* F_pad, void* F_pad_fft, alfa_stream);
/* Filtering_Explicit (Filtering.py:50): returns F and the masked F_fft. */
int alfa_filtering_explicit(alfa_plan*, const void* F_fft, void* F, void* F_fft_new,
                            alfa_stream);
/* Filtering_Level1 (Filtering.py:97) / Filtering_Level2 (Filtering.py:135). */
int alfa_filtering_level1(alfa_plan*, const void* F_fft, void* F_hat, alfa_stream);
int alfa_filtering_level2(alfa_plan*, const void* F_fft, void* F_hatd, alfa_stream);
/* Dealias1 (Dealiasing.py:51): spectrum -> padded field (mx,my,nz). */
int alfa_dealias1(alfa_plan*, const void* F_fft, void* F_pad, alfa_stream);
/* Dealias2 (Dealiasing.py:86): padded spectrum (mx,my/2+1,nz) -> field. */
int alfa_dealias2(alfa_plan*, const void* F_pad_fft, void* F, alfa_stream);
/* Derivxy (Derivatives.py:161): axis 0 -> kx2 (d/dx), 1 -> ky2 (d/dy). */
int alfa_derivxy(alfa_plan*, const void* F_fft, int axis, void* dF, alfa_stream);
/* Derivz_M (Derivatives.py:190). */
int alfa_derivz_m(alfa_plan*, const void* u, const void* v, const void* w,
                  const void* ustar, const void* M_sfc_loc, const void* fi_m,
                  void* dudz, void* dvdz, void* dwdz, alfa_stream);
/* Derivz_TH (Derivatives.py:243). */
int alfa_derivz_th(alfa_plan*, const void* TH, const void* ustar, const void* qz_sfc,
                   const void* fi_h, void* dTHdz, alfa_stream);
/* Derivz_Generic_uvp (Derivatives.py:310), Derivz_Generic_w (:340). */
int alfa_derivz_generic_uvp(alfa_plan*, const void* F, void* dFdz, alfa_stream);
int alfa_derivz_generic_w(alfa_plan*, const void* F, void* dFdz, alfa_stream);

/* ---- utilities (src/utilities/Utilities.py) ------------------------------ */
/* PlanarMean (:41): (nx,ny,nz) -> (nz,) double. */
int alfa_planar_mean(alfa_plan*, const void* F, double* mean, alfa_stream);
/* alfa_planar_mean without a plan and without any allocation: for callers
 * that cannot allocate device memory (code under hipGraph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  precision ALFA_FP64 (double F) or ALFA_MIXED (float F); the
 * same kernels and summation order as alfa_planar_mean, so the means are
 * bitwise equal.  The partial sums go to a fixed device array of the
 * library (4 MiB: 2^19 doubles, one per column block and level); a shape
 * needing more returns ALFA_EINVAL.  Calls that overlap on one GPU share that
 * array and race. */
int alfa_planar_mean_noalloc(int precision, int nx, int ny, int nz, const void* F,
                             double* mean, alfa_stream);
/* StagGridAvg (:64): (nx,ny,nz) -> (nx,ny,nz-1). */
int alfa_stag_grid_avg(alfa_plan*, const void* F, void* out, alfa_stream);
/* Imfilter (:87) applied to every level of an (nx,ny,nz) field. */
int alfa_imfilter(alfa_plan*, const void* F, void* out, alfa_stream);
/* ComputeBeta1/2 (DynamicSGS_LASDD_SM.py:53, ScalarLASDD_SM.py:54) with
 * Roots (:123): coeffs is (6, n) double [ff,ee,dd,cc,bb,aa]; beta (n,). */
int alfa_compute_beta(alfa_plan*, const double* coeffs, int n, double* beta,
                      alfa_stream);

/* ---- pde (src/pde) -------------------------------------------------------- */
/* Advection (NSE_AdvectionTerms.py:246). */
int alfa_advection(alfa_plan*, const void* u, const void* v, const void* w,
                   const void* dudy, const void* dudz, const void* dvdx,
                   const void* dvdz, const void* dwdx, const void* dwdy,
                   void* Cx, void* Cy, void* Cz, alfa_stream);
/* ScalarAdvection (SCL_AdvectionTerms.py:158). */
int alfa_scalar_advection(alfa_plan*, const void* u, const void* v, const void* w,
                          const void* dsdx, const void* dsdy, const void* dsdz,
                          void* out, alfa_stream);
/* BuoyancyOpt1/2 (NSE_BuoyancyTerms.py:49,104) per cfg.optBuoyancy, dry. */
int alfa_buoyancy(alfa_plan*, const void* TH, void* buoyancy, alfa_stream);
/* DivStress (NSE_SGSTerms.py:55). */
int alfa_div_stress(alfa_plan*, const void* txx, const void* tyy, const void* tzz,
                    const void* txy, const void* txz, const void* tyz,
                    void* divtx, void* divty, void* divtz, alfa_stream);
/* DivFlux (SCL_SGSTerms.py:55). */
int alfa_div_flux(alfa_plan*, const void* qx, const void* qy, const void* qz,
                  void* divq, alfa_stream);
/* RHS_Momentum (NSE_AllTerms.py:48), optAdvection = 0. */
int alfa_rhs_momentum(alfa_plan*, const void* u, const void* v, const void* w,
                      const void* Cx, const void* Cy, const void* Cz,
                      const void* buoyancy, const void* divtx, const void* divty,
                      const void* divtz, void* RHS_u, void* RHS_v, void* RHS_w,
                      alfa_stream);
/* RHS_Scalar (SCL_AllTerms.py:48), optAdvection = 0. */
int alfa_rhs_scalar(alfa_plan*, const void* TH, const void* adv, const void* divq,
                    void* RHS_TH, alfa_stream);
/* PressureRC (NSE_PressureTerms.py:109): RC_real/RC_imag (nx,ny/2+1,nz+1)
 * and fRz_real (nx,ny/2+1,nz). */
int alfa_pressure_rc(alfa_plan*, const void* u, const void* v, const void* w,
                     const void* RHS_u, const void* RHS_v, const void* RHS_w,
                     const void* RHS_u_prev, const void* RHS_v_prev,
                     const void* RHS_w_prev, const void* divtz,
                     void* RC_real, void* RC_imag, void* fRz_real, alfa_stream);
/* ThomasPressureSolve (NSE_PressureTerms_Thomas.py:218); the plan holds
 * b_thomas/m_thomas of ThomasPressureInit (:151). */
int alfa_thomas_pressure_solve(alfa_plan*, const void* RC_real, const void* RC_imag,
                               const void* fRz_real, void* p, void* dpdx,
                               void* dpdy, void* dpdz, alfa_stream);
/* AB2_uvw (NSE_TimeAdvancement.py:45), in place on u, v, w. */
int alfa_ab2_uvw(alfa_plan*, void* u, void* v, void* w, const void* RHS_u,
                 const void* RHS_u_prev, const void* RHS_v, const void* RHS_v_prev,
                 const void* RHS_w, const void* RHS_w_prev, alfa_stream);
/* AB2_TH (SCL_TimeAdvancement.py:46), in place on TH. */
int alfa_ab2_th(alfa_plan*, void* TH, const void* RHS_TH, const void* RHS_TH_prev,
                alfa_stream);

/* ---- subgrid scale (src/subgridscale) ------------------------------------- */
/* StrainsUVPnodes_NoDealias (StrainRates.py:149): S11,S22,S33,S12,S13,S23,S. */
int alfa_strains_uvp(alfa_plan*, const void* const grads[9], void* out[7],
                     alfa_stream);
/* StrainsWnodes_NoDealias (StrainRates.py:307): S13, S23, S_w. */
int alfa_strains_w(alfa_plan*, const void* const grads[9], void* out[3], alfa_stream);
/* LASDD (DynamicSGS_LASDD_SM.py:138) for optSgs 1 (LASDD) or 3 (LAD):
 * strains = S11,S22,S33,S12,S13,S23,S; writes Cs2_3D and the profiles. */
int alfa_lasdd_sm(alfa_plan*, const void* u, const void* v, const void* w,
                  const void* const strains[7], void* Cs2_3D, double* Cs2_1D_avg1,
                  double* Cs2_1D_avg2, double* beta1_1D, alfa_stream);
/* LASDD (DynamicSGS_LASDD_WL.py:128) for optSgs 2 (LASDD-WL) or 4 (LAD-WL):
 * strains = S11,S22,S33,S12,S13,S23; writes C_WL (Cwl_3D) and the profiles. */
int alfa_lasdd_wl(alfa_plan*, const void* u, const void* v, const void* w,
                  const void* const strains[6], void* Cwl_3D, double* Cwl_1D_avg1,
                  double* Cwl_1D_avg2, double* beta1_1D, alfa_stream);
/* ScalarLASDD (DynamicSGS_ScalarLASDD_SM.py:139 for optSgs 1/3,
 * DynamicSGS_ScalarLASDD_WL.py:130 for optSgs 2/4), arguments as the
 * reference: vel = u_, v_, w_ (w_ on uvp nodes), hat1 = u_hat, v_hat, w_hat,
 * hat2 = u_hatd, v_hatd, w_hatd (may be NULL for optSgs 3, 4); S, S_hat,
 * S_hatd are |S|, |S^|, |S^^| (SM only; S_hatd only for optSgs 1). */
int alfa_scalar_lasdd(alfa_plan*, const void* const vel[3], const void* const hat1[3],
                      const void* const hat2[3], const void* TH, const void* const dTH[3],
                      const void* S, const void* S_hat, const void* S_hatd,
                      void* Cs2PrRatio_3D, double* Cs2PrRatio_1D, double* beta2_1D,
                      alfa_stream);
/* SGS stresses from the velocity gradients (dudx,dvdx,dwdx,dudy,dvdy,dwdy,
 * dudz,dvdz,dwdz): strain rates (StrainRates.py:56-368), then
 * Stresses{UVP,W}nodes_{Dealias,NoDealias} of SGSStresses_SM.py:98-297 or
 * SGSStresses_WL.py:70-239 per optSgs, with Wall (:63) at level 0, as in
 * DynamicSGS_Main.py:136-282 / StaticSGS_Main.py:76-250.
 * tau = txx, tyy, tzz, txy, txz, tyz. */
int alfa_sgs_stresses(alfa_plan*, const void* const grads[9], const void* Cs2_3D,
                      const void* u, const void* v, const void* M_sfc_loc,
                      const void* psi_m, const void* psi_m0, void* tau[6], alfa_stream);
/* SGS scalar fluxes q = qx, qy, qz (ScalarSGSFluxes_SM.py:55-222 or
 * ScalarSGSFluxes_WL.py:55-201 per optSgs; qz = qz_sfc at level 0), as in
 * DynamicSGS_Main.py:420-478 / StaticSGS_Main.py:258-360. */
int alfa_scalar_fluxes(alfa_plan*, const void* const grads[9], const void* const dTH[3],
                       const void* Cs2PrRatio_3D, const void* qz_sfc, void* q[3], alfa_stream);
/* The Main.py SGS call sites (Main.py:333-443): DivStressDynamicSGS
 * (NSE_SGSTerms.py:113) when `dynamic`, else DivStressStaticSGS (:186) with the
 * given Cs2_3D: strain rates, the dynamic coefficient (written to Cs2_3D and
 * the profiles), stresses with the wall stress, and their divergence
 * divt = divtx, divty, divtz. */
int alfa_div_stress_sgs(alfa_plan*, const void* const grads[9], const void* u, const void* v,
                        const void* w, const void* M_sfc_loc, const void* psi_m,
                        const void* psi_m0, int dynamic, void* Cs2_3D, double* Cs2_1D_avg1,
                        double* Cs2_1D_avg2, double* beta1_1D, void* divt[3], alfa_stream);
/* DivFluxDynamicSGS (SCL_SGSTerms.py:97) when `dynamic`, else
 * DivFluxStaticSGS (:169): scalar coefficient (dynamic; it reuses the
 * momentum procedure's filtered fields, so it follows alfa_div_stress_sgs on
 * the same plan, as the reference passes dynamicSGSmomentum), the fluxes
 * (qz out, with qz_sfc at level 0) and their divergence divq. */
int alfa_div_flux_sgs(alfa_plan*, const void* const grads[9], const void* u, const void* v,
                      const void* w, const void* TH, const void* const dTH[3],
                      const void* qz_sfc, int dynamic, void* Cs2PrRatio_3D,
                      double* Cs2PrRatio_1D, double* beta2_1D, void* qz, void* divq,
                      alfa_stream);
/* Wall (SGSStresses_SM.py:63): planes (nx,ny). */
int alfa_wall(alfa_plan*, const void* u, const void* v, const void* M_sfc_loc,
              const void* psi_m, const void* psi_m0, void* txz, void* tyz,
              alfa_stream);

/* ---- surface (src/surface/SurfaceFlux.py:166-443) -------------------------
 * The variant follows (cfg.optSurfBC, cfg.optSurfFlux).  MOST planes are
 * updated in place; outputs are planes (nx,ny) and a scalar qz_sfc_avg. */
int alfa_surface_flux(alfa_plan*, const void* u, const void* v, const void* TH,
                      double sfc_val, void* psi_m, void* psi_m0, void* psi_h,
                      void* psi_h0, void* fi_m, void* fi_h, void* M_sfc_loc,
                      void* ustar, void* qz_sfc, void* invOB, double* qz_sfc_avg,
                      alfa_stream);

/* ---- measurement ------------------------------------------------------------
 * No counterpart in the reference (which only prints wall time,
 * src/Main.py:562-564,637).  Between begin and end every launch whose kernel
 * name equals `name_filter` (all launches when it is NULL or "") is bracketed
 * by HIP events on its stream; end synchronises on the last event and
 * returns, per kernel name, the launch count, the summed event time and the
 * summed algorithmic bytes (distinct inputs + outputs of each launch).
 * begin creates the events (at most max_launches records), so it belongs
 * outside any captured region. */
typedef struct alfa_kernel_stat {
  char name[96];
  int64_t launches;
  double total_ms;
  double bytes;       /* summed distinct inputs + outputs of the launches */
  double ref_bytes;   /* the same launches in the reference's data flow: each
                         transformed field of an FFT kernel counts one full
                         read of its input and one full write of its output
                         (a half spectrum is (ny+2)/ny of a field); equal to
                         bytes for the other kernels */
} alfa_kernel_stat;
int alfa_profile_begin(alfa_plan* plan, const char* name_filter, int max_launches);
int alfa_profile_end(alfa_plan* plan, alfa_kernel_stat* out, int max_out, int* n_out);

#ifdef __cplusplus
}
#endif
#endif /* ALFA_H */
