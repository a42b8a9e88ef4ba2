"""Reference-named entry points of the per-step hot path, on torch device
tensors, each a thin call into libalfa (include/alfa.h).

Names, argument order and return values follow the reference's functions
(src/operations, src/utilities/Utilities.py, src/pde, src/subgridscale,
src/surface); arguments the reference needs only for its globals or
pre-allocated zero arrays (ZeRo3D, ZeRo3D_fft, kx2/ky2 tables, Thomas
factors, damping profiles) are accepted and ignored, because the plan holds
them.  Grid constants come from the current plan (``use_plan``), the way the
reference's modules read them from Config.py.  Results are new tensors
(the reference is functional), except where noted.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .core import Plan, current_stream

_cur: Plan | None = None


def use_plan(plan: Plan) -> Plan:
    global _cur
    _cur = plan
    return plan


def _P() -> Plan:
    if _cur is None:
        raise RuntimeError("no current plan: call alfa_amd.ops.use_plan(Plan(cfg))")
    return _cur


class WaveTable:
    """Stand-in for the reference's kx2 / ky2 tables (Preprocess.py:73-100):
    Derivxy only needs to know the axis."""

    def __init__(self, axis: int):
        self.axis = axis


def Wavenumber():
    """src/initialization/Preprocess.py:73."""
    return WaveTable(0), WaveTable(1)


def _wavenumbers(n, half):
    """Preprocess.py:89-94: integer wavenumbers, the Nyquist entry set to 0"""
    k = np.fft.rfftfreq(n, 1.0 / n) if half else np.fft.fftfreq(n, 1.0 / n)
    k[n // 2] = 0.0
    return k


def _axis(k):
    """Axis of a Derivxy wavenumber argument: a WaveTable, 0/1/"x"/"y", or
    the reference's own broadcast table kx2 / ky2 (Preprocess.py:96-100),
    shape (nx, ny//2+1[, nz]), numpy or torch.  A table is recognised by the
    index along which it varies, and its values must be the plan's
    wavenumbers (the kernels use the plan's own table)."""
    if isinstance(k, WaveTable):
        return k.axis
    if isinstance(k, (int, str)):
        if k in (0, "x"):
            return 0
        if k in (1, "y"):
            return 1
        raise TypeError("kxy2 must be a WaveTable from Wavenumber(), an axis or a kx2/ky2 table")
    shape = tuple(getattr(k, "shape", ()))
    if len(shape) not in (2, 3):
        raise TypeError("kxy2 must be a WaveTable from Wavenumber(), an axis or a kx2/ky2 table")
    P = _P()
    if shape[0] != P.nx or shape[1] != P.ny // 2 + 1 or (len(shape) == 3 and shape[2] != P.nz):
        raise ValueError(f"wavenumber table of shape {shape} does not match the plan's "
                         f"({P.nx}, {P.ny // 2 + 1}, {P.nz}) spectra")
    idx = (slice(None), 0, 0) if len(shape) == 3 else (slice(None), 0)
    idy = (0, slice(None), 0) if len(shape) == 3 else (0, slice(None))

    def host(a):
        a = a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)
        return a.astype(np.float64)

    along_x, along_y = host(k[idx]), host(k[idy])
    kx, ky = _wavenumbers(P.nx, False), _wavenumbers(P.ny, True)
    if np.array_equal(along_x, kx) and np.all(along_y == along_y[0]):
        return 0
    if np.array_equal(along_y, ky) and np.all(along_x == along_x[0]):
        return 1
    raise ValueError("wavenumber table is neither the plan's kx2 nor its ky2 "
                     "(Preprocess.py:89-100: integer wavenumbers, Nyquist set to 0)")


# ---------------------------------------------------------------- operations
def FFT(F):
    """src/operations/FFT.py:44."""
    P = _P()
    out = P.spectrum()
    P.call("alfa_fft", P.ptr(F), P.ptr(out), current_stream())
    return out


def FFT_pad(F_pad):
    """src/operations/FFT.py:68."""
    P = _P()
    out = P.zeros((P.mx, P.my // 2 + 1, P.nz), complex_=True)
    P.call("alfa_fft_pad", P.ptr(F_pad), P.ptr(out), current_stream())
    return out


def Filtering_Explicit(F_fft):
    """src/operations/Filtering.py:50 -> (F, F_fft_new)."""
    P = _P()
    F = P.field()
    Fn = P.spectrum()
    P.call("alfa_filtering_explicit", P.ptr(F_fft), P.ptr(F), P.ptr(Fn), current_stream())
    return F, Fn


def Filtering_Level1(F_fft):
    """src/operations/Filtering.py:97."""
    P = _P()
    out = P.field()
    P.call("alfa_filtering_level1", P.ptr(F_fft), P.ptr(out), current_stream())
    return out


def Filtering_Level2(F_fft):
    """src/operations/Filtering.py:135."""
    P = _P()
    out = P.field()
    P.call("alfa_filtering_level2", P.ptr(F_fft), P.ptr(out), current_stream())
    return out


def Dealias1(F_fft, ZeRo3D_pad_fft=None):
    """src/operations/Dealiasing.py:51."""
    P = _P()
    out = P.zeros((P.mx, P.my, P.nz))
    P.call("alfa_dealias1", P.ptr(F_fft), P.ptr(out), current_stream())
    return out


def Dealias2(F_pad_fft, ZeRo3D_fft=None):
    """src/operations/Dealiasing.py:86."""
    P = _P()
    out = P.field()
    P.call("alfa_dealias2", P.ptr(F_pad_fft), P.ptr(out), current_stream())
    return out


def Derivxy(F_fft, kxy2):
    """src/operations/Derivatives.py:161."""
    P = _P()
    out = P.field()
    P.call("alfa_derivxy", P.ptr(F_fft), _axis(kxy2), P.ptr(out), current_stream())
    return out


def Derivz_M(u, v, w, ustar, M_sfc_loc, fi2D_m, ZeRo3D=None):
    """src/operations/Derivatives.py:190."""
    P = _P()
    a, b, c = P.field(), P.field(), P.field()
    P.call("alfa_derivz_m", P.ptr(u), P.ptr(v), P.ptr(w), P.ptr(ustar), P.ptr(M_sfc_loc),
           P.ptr(fi2D_m), P.ptr(a), P.ptr(b), P.ptr(c), current_stream())
    return a, b, c


def Derivz_TH(TH, ustar, qz_sfc, fi2D_h, ZeRo3D=None):
    """src/operations/Derivatives.py:243."""
    P = _P()
    out = P.field()
    P.call("alfa_derivz_th", P.ptr(TH), P.ptr(ustar), P.ptr(qz_sfc), P.ptr(fi2D_h), P.ptr(out),
           current_stream())
    return out


def Derivz_Generic_uvp(F, ZeRo3D=None):
    """src/operations/Derivatives.py:310."""
    P = _P()
    out = P.field()
    P.call("alfa_derivz_generic_uvp", P.ptr(F), P.ptr(out), current_stream())
    return out


def Derivz_Generic_w(F, ZeRo3D=None):
    """src/operations/Derivatives.py:340."""
    P = _P()
    out = P.field()
    P.call("alfa_derivz_generic_w", P.ptr(F), P.ptr(out), current_stream())
    return out


def velocityGradients(u, v, w, u_fft, v_fft, w_fft, kx2, ky2, ustar, M_sfc_loc, MOSTfunctions,
                      ZeRo3D=None):
    """src/operations/Derivatives.py:47."""
    dx = [Derivxy(f, 0) for f in (u_fft, v_fft, w_fft)]
    dy = [Derivxy(f, 1) for f in (u_fft, v_fft, w_fft)]
    dz = Derivz_M(u, v, w, ustar, M_sfc_loc, MOSTfunctions[4])
    return (dx[0], dx[1], dx[2], dy[0], dy[1], dy[2], dz[0], dz[1], dz[2])


def potentialTemperatureGradients(TH, kx2, ky2, ustar, qz_sfc, MOSTfunctions, ZeRo3D=None):
    """src/operations/Derivatives.py:107."""
    f = FFT(TH)
    return Derivxy(f, 0), Derivxy(f, 1), Derivz_TH(TH, ustar, qz_sfc, MOSTfunctions[5])


# ----------------------------------------------------------------- utilities
def PlanarMean(F):
    """src/utilities/Utilities.py:41 (float64 result)."""
    P = _P()
    out = P.zeros((P.nz,), dtype=_torch().float64)
    P.call("alfa_planar_mean", P.ptr(F), P.ptr(out), current_stream())
    return out


def StagGridAvg(F):
    """src/utilities/Utilities.py:64."""
    P = _P()
    out = P.zeros((P.nx, P.ny, P.nz - 1))
    P.call("alfa_stag_grid_avg", P.ptr(F), P.ptr(out), current_stream())
    return out


def Imfilter3D(F):
    """Utilities.py:87 applied to every level of an (nx, ny, nz) field (the
    reference vmaps it over levels, DynamicSGS_LASDD_SM.py:357)."""
    P = _P()
    out = P.field()
    P.call("alfa_imfilter", P.ptr(F), P.ptr(out), current_stream())
    return out


def Imfilter(x):
    """src/utilities/Utilities.py:87 on one (nx, ny) plane."""
    P = _P()
    F = x.reshape(P.nx, P.ny, 1).expand(P.nx, P.ny, P.nz).contiguous()
    return Imfilter3D(F)[:, :, 0].contiguous()


def ComputeBeta(ff, ee, dd, cc, bb, aa):
    """ComputeBeta1 (DynamicSGS_LASDD_SM.py:53) / ComputeBeta2
    (DynamicSGS_ScalarLASDD_SM.py:54): Laguerre roots (Utilities.py:123)."""
    torch = _torch()
    P = _P()
    co = torch.stack([torch.as_tensor(a, dtype=torch.float64) for a in (ff, ee, dd, cc, bb, aa)])
    co = co.to(P.device).contiguous()
    n = co.shape[1]
    out = P.zeros((n,), dtype=torch.float64)
    P.call("alfa_compute_beta", P.ptr(co), int(n), P.ptr(out), current_stream())
    return out


ComputeBeta1 = ComputeBeta
ComputeBeta2 = ComputeBeta


# ----------------------------------------------------------------------- pde
def Advection(u, v, w, dudy, dudz, dvdx, dvdz, dwdx, dwdy, *zeros):
    """src/pde/NSE_AdvectionTerms.py:246."""
    P = _P()
    o = [P.field() for _ in range(3)]
    P.call("alfa_advection", *[P.ptr(t) for t in (u, v, w, dudy, dudz, dvdx, dvdz, dwdx, dwdy)],
           *[P.ptr(t) for t in o], current_stream())
    return tuple(o)


def ScalarAdvection(u, v, w, dsdx, dsdy, dsdz, *zeros):
    """src/pde/SCL_AdvectionTerms.py:158."""
    P = _P()
    out = P.field()
    P.call("alfa_scalar_advection", *[P.ptr(t) for t in (u, v, w, dsdx, dsdy, dsdz, out)],
           current_stream())
    return out


def _buoy(TH, want):
    P = _P()
    if int(P.cfg.get("optBuoyancy", 1)) != want:
        raise ValueError(f"plan has optBuoyancy={P.cfg.get('optBuoyancy', 1)}")
    out = P.field()
    P.call("alfa_buoyancy", P.ptr(TH), P.ptr(out), current_stream())
    return out


def BuoyancyOpt1(TH, H=None, ZeRo3D=None):
    """src/pde/NSE_BuoyancyTerms.py:49 (dry: H must be zero)."""
    return _buoy(TH, 0)


def BuoyancyOpt2(TH, H=None, ZeRo3D=None):
    """src/pde/NSE_BuoyancyTerms.py:104 (dry: H must be zero)."""
    return _buoy(TH, 1)


def DivStress(txx, tyy, tzz, txy, txz, tyz, ZeRo3D=None, kx2=None, ky2=None):
    """src/pde/NSE_SGSTerms.py:55."""
    P = _P()
    o = [P.field() for _ in range(3)]
    P.call("alfa_div_stress", *[P.ptr(t) for t in (txx, tyy, tzz, txy, txz, tyz)],
           *[P.ptr(t) for t in o], current_stream())
    return tuple(o)


def DivFlux(qx, qy, qz, ZeRo3D=None, kx2=None, ky2=None):
    """src/pde/SCL_SGSTerms.py:55."""
    P = _P()
    out = P.field()
    P.call("alfa_div_flux", P.ptr(qx), P.ptr(qy), P.ptr(qz), P.ptr(out), current_stream())
    return out


def RHS_Momentum(u, v, w, Ug, Vg, Cx, Cy, Cz, buoyancy, divtx, divty, divtz,
                 RayleighDampCoeff=None, RayleighDampCoeff_stag=None, Uadv=None, Vadv=None):
    """src/pde/NSE_AllTerms.py:48 (optAdvection = 0; Ug, Vg and the damping
    profiles are the plan's)."""
    P = _P()
    o = [P.field() for _ in range(3)]
    P.call("alfa_rhs_momentum",
           *[P.ptr(t) for t in (u, v, w, Cx, Cy, Cz, buoyancy, divtx, divty, divtz)],
           *[P.ptr(t) for t in o], current_stream())
    return tuple(o)


def RHS_Scalar(TH, THAdvectionSum, divq, RayleighDampCoeff_stag=None, THadv=None):
    """src/pde/SCL_AllTerms.py:48 (optAdvection = 0)."""
    P = _P()
    out = P.field()
    P.call("alfa_rhs_scalar", P.ptr(TH), P.ptr(THAdvectionSum), P.ptr(divq), P.ptr(out),
           current_stream())
    return out


def PressureRC(u, v, w, RHS_u, RHS_v, RHS_w, RHS_u_previous, RHS_v_previous, RHS_w_previous,
               divtz, kr2_pressure=None, kc2_pressure=None):
    """src/pde/NSE_PressureTerms.py:109 -> (RC_real, RC_imag, fRz_real)."""
    P = _P()
    RCr = P.zeros((P.nx, P.nyc, P.nz + 1))
    RCi = P.zeros((P.nx, P.nyc, P.nz + 1))
    fRz = P.zeros((P.nx, P.nyc, P.nz))
    P.call("alfa_pressure_rc",
           *[P.ptr(t) for t in (u, v, w, RHS_u, RHS_v, RHS_w, RHS_u_previous, RHS_v_previous,
                                RHS_w_previous, divtz, RCr, RCi, fRz)], current_stream())
    return RCr, RCi, fRz


def ThomasPressureSolve(RC_real, RC_imag, fRz_real, b_thomas=None, m_thomas=None,
                        c_pressure=None):
    """src/pde/NSE_PressureTerms_Thomas.py:218 -> (p, dpdx, dpdy, dpdz)."""
    P = _P()
    o = [P.field() for _ in range(4)]
    P.call("alfa_thomas_pressure_solve", P.ptr(RC_real), P.ptr(RC_imag), P.ptr(fRz_real),
           *[P.ptr(t) for t in o], current_stream())
    return tuple(o)


def AB2_uvw(u, v, w, RHS_u, RHS_u_previous, RHS_v, RHS_v_previous, RHS_w, RHS_w_previous):
    """src/pde/NSE_TimeAdvancement.py:45 (returns updated copies)."""
    P = _P()
    un, vn, wn = u.clone(), v.clone(), w.clone()
    P.call("alfa_ab2_uvw", *[P.ptr(t) for t in (un, vn, wn, RHS_u, RHS_u_previous, RHS_v,
                                                RHS_v_previous, RHS_w, RHS_w_previous)],
           current_stream())
    return un, vn, wn


def AB2_TH(TH, RHS_TH, RHS_TH_previous):
    """src/pde/SCL_TimeAdvancement.py:46 (returns an updated copy)."""
    P = _P()
    tn = TH.clone()
    P.call("alfa_ab2_th", P.ptr(tn), P.ptr(RHS_TH), P.ptr(RHS_TH_previous), current_stream())
    return tn


# ---------------------------------------------------------------- subgridscale
def _ptr_array(P, ts):
    arr = (C.c_void_p * len(ts))()
    for i, t in enumerate(ts):
        arr[i] = P.ptr(t).value
    return arr


def StrainsUVPnodes_NoDealias(dudx, dvdx, dwdx, dudy, dvdy, dwdy, dudz, dvdz, dwdz, ZeRo3D=None):
    """src/subgridscale/StrainRates.py:149 -> S11, S22, S33, S12, S13, S23, S."""
    P = _P()
    o = [P.field() for _ in range(7)]
    P.call("alfa_strains_uvp", _ptr_array(P, [dudx, dvdx, dwdx, dudy, dvdy, dwdy, dudz, dvdz, dwdz]),
           _ptr_array(P, o), current_stream())
    return tuple(o)


def StrainsWnodes_NoDealias(dudx, dvdx, dwdx, dudy, dvdy, dwdy, dudz, dvdz, dwdz, ZeRo3D=None):
    """src/subgridscale/StrainRates.py:307 -> S13, S23, S_w."""
    P = _P()
    o = [P.field() for _ in range(3)]
    P.call("alfa_strains_w", _ptr_array(P, [dudx, dvdx, dwdx, dudy, dvdy, dwdy, dudz, dvdz, dwdz]),
           _ptr_array(P, o), current_stream())
    return tuple(o)


def LASDD(u, v, w, S11, S22, S33, S12, S13, S23, S, ZeRo3D=None):
    """src/subgridscale/DynamicSGS_LASDD_SM.py:138 (optSgs 1 or 3 of the plan)
    -> (Cs2_3D, Cs2_1D_avg1, Cs2_1D_avg2, beta1_1D)."""
    torch = _torch()
    P = _P()
    cs = P.field()
    pr = [P.zeros((P.nz,), dtype=torch.float64) for _ in range(3)]
    P.call("alfa_lasdd_sm", P.ptr(u), P.ptr(v), P.ptr(w),
           _ptr_array(P, [S11, S22, S33, S12, S13, S23, S]), P.ptr(cs),
           *[P.ptr(t) for t in pr], current_stream())
    return cs, pr[0], pr[1], pr[2]


def LASDD_WL(u, v, w, S11, S22, S33, S12, S13, S23, ZeRo3D=None):
    """src/subgridscale/DynamicSGS_LASDD_WL.py:128 (optSgs 2 or 4 of the plan)
    -> (Cwl_3D, Cwl_1D_avg1, Cwl_1D_avg2, beta1_1D)."""
    torch = _torch()
    P = _P()
    cs = P.field()
    pr = [P.zeros((P.nz,), dtype=torch.float64) for _ in range(3)]
    P.call("alfa_lasdd_wl", P.ptr(u), P.ptr(v), P.ptr(w),
           _ptr_array(P, [S11, S22, S33, S12, S13, S23]), P.ptr(cs),
           *[P.ptr(t) for t in pr], current_stream())
    return cs, pr[0], pr[1], pr[2]


def ScalarLASDD(u_, v_, w_, u_hat, v_hat, w_hat, u_hatd, v_hatd, w_hatd, TH,
                dTHdx, dTHdy, dTHdz, S=None, S_hat=None, S_hatd=None, ZeRo3D=None):
    """DynamicSGS_ScalarLASDD_SM.py:139 (optSgs 1, 3) / DynamicSGS_ScalarLASDD_WL.py:130
    (optSgs 2, 4) -> (Cs2PrRatio_3D, Cs2PrRatio_1D, beta2_1D).  u_hatd.. may be
    None for the beta = 1 models; S, S_hat, S_hatd are used by SM only."""
    torch = _torch()
    P = _P()
    cs = P.field()
    pr = [P.zeros((P.nz,), dtype=torch.float64) for _ in range(2)]
    h2 = None if u_hatd is None else _ptr_array(P, [u_hatd, v_hatd, w_hatd])
    P.call("alfa_scalar_lasdd", _ptr_array(P, [u_, v_, w_]), _ptr_array(P, [u_hat, v_hat, w_hat]),
           h2, P.ptr(TH), _ptr_array(P, [dTHdx, dTHdy, dTHdz]), P.ptr(S), P.ptr(S_hat),
           P.ptr(S_hatd), P.ptr(cs), P.ptr(pr[0]), P.ptr(pr[1]), current_stream())
    return cs, pr[0], pr[1]


def SGSStresses(dudx, dvdx, dwdx, dudy, dvdy, dwdy, dudz, dvdz, dwdz, Cs2_3D,
                u, v, M_sfc_loc, psi2D_m, psi2D_m0):
    """Strain rates + Stresses{UVP,W}nodes (SM or WL per the plan's optSgs,
    dealiased when FGR = 1) with the wall stress, as DynamicSGS_Main.py:136-282
    and StaticSGS_Main.py:76-250 compute them -> (txx, tyy, tzz, txy, txz, tyz)."""
    P = _P()
    o = [P.field() for _ in range(6)]
    P.call("alfa_sgs_stresses",
           _ptr_array(P, [dudx, dvdx, dwdx, dudy, dvdy, dwdy, dudz, dvdz, dwdz]), P.ptr(Cs2_3D),
           *[P.ptr(t) for t in (u, v, M_sfc_loc, psi2D_m, psi2D_m0)], _ptr_array(P, o),
           current_stream())
    return tuple(o)


def ScalarFluxes(dudx, dvdx, dwdx, dudy, dvdy, dwdy, dudz, dvdz, dwdz, dTHdx, dTHdy, dTHdz,
                 Cs2PrRatio_3D, qz_sfc):
    """ScalarFluxes{UVP,W}nodes (SM or WL per the plan, dealiased when FGR = 1),
    DynamicSGS_Main.py:420-478 / StaticSGS_Main.py:258-360 -> (qx, qy, qz)."""
    P = _P()
    o = [P.field() for _ in range(3)]
    P.call("alfa_scalar_fluxes",
           _ptr_array(P, [dudx, dvdx, dwdx, dudy, dvdy, dwdy, dudz, dvdz, dwdz]),
           _ptr_array(P, [dTHdx, dTHdy, dTHdz]), P.ptr(Cs2PrRatio_3D), P.ptr(qz_sfc),
           _ptr_array(P, o), current_stream())
    return tuple(o)


class SGSMomentum:
    """What DivStress*SGS hands to DivFlux*SGS (the reference's
    dynamicSGSmomentum / staticSGSmomentum tuples, NSE_SGSTerms.py:178,249).
    Here the filtered fields stay in the plan's workspace; the handle keeps
    the inputs the flux call needs.  Main.py passes a slice of the tuple
    (Main.py:352, :438): slicing returns the handle itself."""

    def __init__(self, grads, u, v, w):
        self.grads, self.u, self.v, self.w = grads, u, v, w

    def __getitem__(self, k):
        return self


def _div_stress_sgs(grads, u, v, w, M_sfc_loc, MOSTfunctions, dynamic, Cs2_3D):
    torch = _torch()
    P = _P()
    cs = P.field() if Cs2_3D is None else Cs2_3D
    pr = [P.zeros((P.nz,), dtype=torch.float64) for _ in range(3)]
    o = [P.field() for _ in range(3)]
    P.call("alfa_div_stress_sgs", _ptr_array(P, grads), P.ptr(u), P.ptr(v), P.ptr(w),
           P.ptr(M_sfc_loc), P.ptr(MOSTfunctions[0]), P.ptr(MOSTfunctions[1]), int(dynamic),
           P.ptr(cs), *[P.ptr(t) for t in pr], _ptr_array(P, o), current_stream())
    return o, cs, pr


def DivStressDynamicSGS(dudx, dvdx, dwdx, dudy, dvdy, dwdy, dudz, dvdz, dwdz, u, v, w, M_sfc_loc,
                        MOSTfunctions, ZeRo3D=None, ZeRo3D_fft=None, ZeRo3D_pad_fft=None, kx2=None,
                        ky2=None):
    """src/pde/NSE_SGSTerms.py:113 -> (divtx, divty, divtz, Cs2_1D_avg1, Cs2_1D_avg2,
    beta1_1D, Cs2_3D, dynamicSGSmomentum)."""
    grads = [dudx, dvdx, dwdx, dudy, dvdy, dwdy, dudz, dvdz, dwdz]
    o, cs, pr = _div_stress_sgs(grads, u, v, w, M_sfc_loc, MOSTfunctions, 1, None)
    return o[0], o[1], o[2], pr[0], pr[1], pr[2], cs, SGSMomentum(grads, u, v, w)


def DivStressStaticSGS(dudx, dvdx, dwdx, dudy, dvdy, dwdy, dudz, dvdz, dwdz, Cs2_3D, u, v,
                       M_sfc_loc, MOSTfunctions, ZeRo3D=None, ZeRo3D_fft=None,
                       ZeRo3D_pad_fft=None, kx2=None, ky2=None):
    """src/pde/NSE_SGSTerms.py:186 -> (divtx, divty, divtz, staticSGSmomentum)."""
    grads = [dudx, dvdx, dwdx, dudy, dvdy, dwdy, dudz, dvdz, dwdz]
    o, _, _ = _div_stress_sgs(grads, u, v, u, M_sfc_loc, MOSTfunctions, 0, Cs2_3D)
    return o[0], o[1], o[2], SGSMomentum(grads, u, v, None)


def _div_flux_sgs(mom, TH, dTH, qz_sfc, dynamic, CsPr):
    torch = _torch()
    P = _P()
    cs = P.field() if CsPr is None else CsPr
    pr = [P.zeros((P.nz,), dtype=torch.float64) for _ in range(2)]
    qz, divq = P.field(), P.field()
    w = mom.w if mom.w is not None else mom.u
    P.call("alfa_div_flux_sgs", _ptr_array(P, mom.grads), P.ptr(mom.u), P.ptr(mom.v), P.ptr(w),
           P.ptr(TH if TH is not None else mom.u), _ptr_array(P, dTH), P.ptr(qz_sfc), int(dynamic),
           P.ptr(cs), *[P.ptr(t) for t in pr], P.ptr(qz), P.ptr(divq), current_stream())
    return qz, divq, cs, pr


def DivFluxDynamicSGS(dynamicSGSmomentum, TH, dTHdx, dTHdy, dTHdz, SHFX, ZeRo3D=None,
                      ZeRo3D_fft=None, ZeRo3D_pad_fft=None, kx2=None, ky2=None):
    """src/pde/SCL_SGSTerms.py:97 (after DivStressDynamicSGS on the same plan)
    -> (qz, divq, Cs2PrRatio_3D, Cs2PrRatio_1D, beta2_1D)."""
    qz, divq, cs, pr = _div_flux_sgs(dynamicSGSmomentum, TH, [dTHdx, dTHdy, dTHdz], SHFX, 1, None)
    return qz, divq, cs, pr[0], pr[1]


def DivFluxStaticSGS(staticSGSmomentum, Cs2PrRatio_3D, dTHdx, dTHdy, dTHdz, SHFX, ZeRo3D=None,
                     ZeRo3D_fft=None, ZeRo3D_pad_fft=None, kx2=None, ky2=None):
    """src/pde/SCL_SGSTerms.py:169 -> (qz, divq)."""
    qz, divq, _, _ = _div_flux_sgs(staticSGSmomentum, None, [dTHdx, dTHdy, dTHdz], SHFX, 0,
                                   Cs2PrRatio_3D)
    return qz, divq


def Wall(u, v, M_sfc_loc, psi2D_m, psi2D_m0):
    """src/subgridscale/SGSStresses_SM.py:63 -> (txz, tyz) planes."""
    P = _P()
    a, b = P.plane(), P.plane()
    P.call("alfa_wall", *[P.ptr(t) for t in (u, v, M_sfc_loc, psi2D_m, psi2D_m0, a, b)],
           current_stream())
    return a, b


# -------------------------------------------------------------------- surface
def _surface(u, v, TH, sfc_val, MOSTfunctions):
    torch = _torch()
    P = _P()
    most = [m.clone() for m in MOSTfunctions]
    M, us, qz, iob = P.plane(), P.plane(), P.plane(), P.plane()
    qavg = P.zeros((1,), dtype=torch.float64)
    P.call("alfa_surface_flux", P.ptr(u), P.ptr(v), P.ptr(TH), C.c_double(float(sfc_val)),
           *[P.ptr(t) for t in most], *[P.ptr(t) for t in (M, us, qz, iob, qavg)],
           current_stream())
    return M, us, qz, float(qavg.item()), iob, tuple(most)


def _check_surface(bc, flux):
    P = _P()
    if int(P.cfg.get("optSurfBC", 0)) != bc or int(P.cfg.get("optSurfFlux", 0)) != flux:
        raise ValueError("surface variant does not match the plan's optSurfBC/optSurfFlux")


def SurfaceFlux_HomogeneousConstantFlux(u, v, TH, MOSTfunctions):
    """src/surface/SurfaceFlux.py:166 -> (M, ustar, qz_sfc_avg, invOB, MOST)."""
    _check_surface(0, 0)
    M, us, qz, qa, iob, most = _surface(u, v, TH, 0.0, MOSTfunctions)
    return M, us, qa, iob, most


def SurfaceFlux_HeterogeneousConstantFlux(u, v, TH, MOSTfunctions):
    """src/surface/SurfaceFlux.py:209."""
    _check_surface(0, 1)
    M, us, qz, qa, iob, most = _surface(u, v, TH, 0.0, MOSTfunctions)
    return M, us, qa, iob, most


def SurfaceFlux_HomogeneousVaryingFlux(u, v, TH, qz_sfc_t, MOSTfunctions):
    """src/surface/SurfaceFlux.py:246."""
    _check_surface(1, 0)
    return _surface(u, v, TH, qz_sfc_t, MOSTfunctions)


def SurfaceFlux_HeterogeneousVaryingFlux(u, v, TH, qz_sfc_t, MOSTfunctions):
    """src/surface/SurfaceFlux.py:290."""
    _check_surface(1, 1)
    return _surface(u, v, TH, qz_sfc_t, MOSTfunctions)


def SurfaceFlux_HomogeneousPrescribedTemperature(u, v, TH, TH_sfc_t, MOSTfunctions):
    """src/surface/SurfaceFlux.py:335."""
    _check_surface(2, 0)
    return _surface(u, v, TH, TH_sfc_t, MOSTfunctions)


def SurfaceFlux_HeterogeneousPrescribedTemperature(u, v, TH, TH_sfc_t, MOSTfunctions):
    """src/surface/SurfaceFlux.py:393."""
    _check_surface(2, 1)
    return _surface(u, v, TH, TH_sfc_t, MOSTfunctions)


def _torch():
    import torch
    return torch
