"""FP64 NumPy restatement of the JAX-ALFA per-time-step hot path.

TEST INFRASTRUCTURE ONLY.  This module is the parity oracle: it is imported
by ``tests/``, by ``__graft_entry__.smoke()`` and by ``bench.py``'s CPU
baseline leg, and nowhere else.  The product path (``alfa_amd`` and the HIP
library) never imports, links or runs anything from ``oracle/``.

Every function restates the algorithm of one reference function; the
``file:line`` citations are relative to the reference repository
(Sukantabasu/jax-alfa).  Arithmetic is float64/complex128 throughout and the
FFTs are ``numpy.fft`` (pocketfft), which has the same normalisation and the
same real-axis conventions as ``jnp.fft.rfft2/irfft2`` over axes (0, 1).

Layout: real fields are ``(nx, ny, nz)`` (z fastest), spectra are
``(nx, ny//2+1, nz)`` complex, padded fields ``(mx, my, nz)``.

Pinning status of each part (what checks it besides this restatement) is in
``tests/test_oracle_pins.py``; parts no independent check covers are marked
"parity unpinned" in their docstring.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

G_GRAV = 9.81      # src/config/DerivedVars.py:39
VONK = 0.4         # src/config/DerivedVars.py:40


# ---------------------------------------------------------------------------
# Configuration and derived constants
# ---------------------------------------------------------------------------

DEFAULTS = dict(
    nx=32, ny=32, nz=32, l_x=2000.0, l_y=2000.0, l_z=1000.0,
    dt=0.5, Ugal=0.0, optSurfFlux=0, optSurfBC=0, z0m=0.1, z0T=0.1,
    zTemperature=0.0, SensibleHeatFlux=0.0, Ug2=10.0, Vg2=0.0,
    f_coriolis=1e-4, inversion=0.0, optBuoyancy=1, T_0=300.0,
    optSgs=3, dynamicSGS_call_time=1, FGR=1, Cs2=0.1 ** 2, Cs2PrRatio=0.1 ** 2,
    Cwl=0.1 ** 2, CwlPrRatio=0.1 ** 2,
    optDamping=1, z_damping=700.0, RelaxTime=300.0, optPressureSolver=1,
    optMoistureSurfBC=0, MoistureFlux=0.0, zMoisture=0.0, q_inversion=0.0,
    optAdvection=0, optGeoWind=0, optMoisture=0, istep=1,
)


class Grid:
    """Derived constants of src/config/DerivedVars.py:49-119 and the FFT
    constants of src/initialization/Preprocess.py:44-100."""

    def __init__(self, cfg: dict):
        c = dict(DEFAULTS)
        c.update(cfg)
        self.cfg = c
        for k, v in c.items():
            setattr(self, k, v)
        nx, ny, nz = c["nx"], c["ny"], c["nz"]
        pi = math.pi
        self.z_scale = c["l_x"] / (2 * pi)                       # DerivedVars.py:51
        self.g_nondim = G_GRAV * self.z_scale                    # :54
        self.f_coriolis_nondim = c["f_coriolis"] * self.z_scale  # :57
        self.T_0_nondim = c["T_0"]                               # :60
        self.inversion_nondim = c["inversion"] * self.z_scale    # :63
        self.qm_sfc = c["MoistureFlux"] * np.ones((nx, ny))      # :72 (Q_scale = 1)
        self.q_inversion_nondim = c["q_inversion"] * self.z_scale  # :75
        self.qz_sfc = c["SensibleHeatFlux"] * np.ones((nx, ny))  # :66
        self.dx = 2 * pi / nx                                    # :78
        self.dy = 2 * pi / ny                                    # :79
        self.dz = (c["l_z"] / self.z_scale) / (nz - 1)           # :80
        self.idz = 1.0 / self.dz                                 # :81
        self.L = c["FGR"] * (self.dx * self.dy * self.dz) ** (1 / 3)  # :84
        self.dt_nondim = c["dt"] / self.z_scale                  # :87
        if c["FGR"] == 1:                                        # :93-98
            self.TFR = 2
            self.optDealias = 1
        else:
            self.TFR = np.sqrt(2)
            self.optDealias = 0
        self.z_damping_nondim = c["z_damping"] / self.z_scale    # :110
        self.RelaxTime_nondim = c["RelaxTime"] / self.z_scale    # :113
        # Preprocess.py:44-62
        self.mx = int(1.5 * nx)
        self.my = int(1.5 * ny)
        self.nx_rfft = nx // 2 + 1
        self.ny_rfft = ny // 2 + 1
        # Preprocess.py:73-100 (kept 2-D; broadcast over z where used)
        kx = np.fft.fftfreq(nx, 1 / nx)
        ky = np.fft.rfftfreq(ny, 1 / ny)
        kx[nx // 2] = 0
        ky[ny // 2] = 0
        self.kx = kx
        self.ky = ky
        self.kx2 = np.broadcast_to(kx[:, None, None], (nx, self.ny_rfft, nz))
        self.ky2 = np.broadcast_to(ky[None, :, None], (nx, self.ny_rfft, nz))
        # Filter cut-offs (Filtering.py:73-74, :111-112, :149-150)
        self.cut0_x = round(nx * (1.0 / (2.0 * c["FGR"])))
        self.cut0_y = round(ny * (1.0 / (2.0 * c["FGR"])))
        self.cut1_x = round(nx / (2 * c["FGR"] * self.TFR))
        self.cut1_y = round(ny / (2 * c["FGR"] * self.TFR))
        self.cut2_x = round(nx / (2 * c["FGR"] * self.TFR * self.TFR))
        self.cut2_y = round(ny / (2 * c["FGR"] * self.TFR * self.TFR))
        self.RayleighDampCoeff, self.RayleighDampCoeff_stag = \
            Initialize_RayleighDampingLayer(self)
        self.Ug = c["Ug2"] * np.ones(nz)     # Initialization.py:163-170
        self.Vg = c["Vg2"] * np.ones(nz)


# ---------------------------------------------------------------------------
# src/operations
# ---------------------------------------------------------------------------

# FFT backend.  The parity tests use numpy.fft (pocketfft, one thread).  The
# CPU-baseline timing sets FFT_WORKERS to run the same transforms through
# scipy.fft (also pocketfft) on that many threads.
FFT_WORKERS = None


def set_fft_workers(n):
    """None: numpy.fft; n >= 1: scipy.fft with workers=n."""
    global FFT_WORKERS
    FFT_WORKERS = None if n is None else int(n)


def _rfft2(F):
    if FFT_WORKERS is None:
        return np.fft.rfft2(F, axes=(0, 1))
    import scipy.fft
    return scipy.fft.rfft2(F, axes=(0, 1), workers=FFT_WORKERS)


def FFT(F):
    """src/operations/FFT.py:44-60: rfft2 over axes (0, 1)."""
    return _rfft2(F)


def irfft2(Fh, s):
    if FFT_WORKERS is None:
        return np.fft.irfft2(Fh, s=s, axes=(0, 1))
    import scipy.fft
    return scipy.fft.irfft2(Fh, s=s, axes=(0, 1), workers=FFT_WORKERS)


def _cutoff_mask_apply(g, F_fft, cx, cy):
    """Keep rows [0,cx) U [nx-cx+1, nx) and columns [0, cy)."""
    nx = g.nx
    out = np.zeros_like(F_fft)
    out[:cx, :cy] = F_fft[:cx, :cy]
    out[nx - cx + 1:, :cy] = F_fft[nx - cx + 1:, :cy]
    return out


def Filtering_Explicit(g, F_fft):
    """src/operations/Filtering.py:50-89."""
    Fn = _cutoff_mask_apply(g, F_fft, g.cut0_x, g.cut0_y)
    return irfft2(Fn, (g.nx, g.ny)), Fn


def Filtering_Level1(g, F_fft):
    """src/operations/Filtering.py:97-127."""
    return irfft2(_cutoff_mask_apply(g, F_fft, g.cut1_x, g.cut1_y), (g.nx, g.ny))


def Filtering_Level2(g, F_fft):
    """src/operations/Filtering.py:135-165."""
    return irfft2(_cutoff_mask_apply(g, F_fft, g.cut2_x, g.cut2_y), (g.nx, g.ny))


def Dealias1(g, F_fft):
    """src/operations/Dealiasing.py:51-78 (3/2-rule zero pad; the kx-Nyquist
    row nx/2 lands in both padded rows nx/2 and mx-nx/2)."""
    nxr, nyr = g.nx_rfft, g.ny_rfft
    pad = np.zeros((g.mx, g.my // 2 + 1, F_fft.shape[2]), dtype=complex)
    pad[:nxr, :nyr] = F_fft[:nxr, :nyr]
    pad[g.mx - (nxr - 1):, :nyr] = F_fft[(nxr - 1):, :nyr]
    return irfft2(pad, (g.mx, g.my))


def FFT_pad(F_pad):
    """src/operations/FFT.py:68-84."""
    return _rfft2(F_pad)


def Dealias2(g, F_pad_fft):
    """src/operations/Dealiasing.py:86-114 (truncate, drop both Nyquists,
    scale by 9/4)."""
    nxr, nyr = g.nx_rfft, g.ny_rfft
    F = np.zeros((g.nx, nyr, F_pad_fft.shape[2]), dtype=complex)
    F[:nxr - 1, :nyr - 1] = F_pad_fft[:nxr - 1, :nyr - 1]
    F[nxr:, :nyr - 1] = F_pad_fft[g.mx - (nxr - 2):, :nyr - 1]
    return (9.0 / 4.0) * irfft2(F, (g.nx, g.ny))


def Derivxy(g, F_fft, k2):
    """src/operations/Derivatives.py:161-182: irfft2(i k F)."""
    return irfft2(1j * k2 * F_fft, (g.nx, g.ny))


def Derivz_M(g, u, v, w, ustar, M_sfc_loc, fi2D_m):
    """src/operations/Derivatives.py:190-235."""
    nz = g.nz
    dudz = np.zeros_like(u)
    dvdz = np.zeros_like(u)
    dwdz = np.zeros_like(u)
    dudz[:, :, 1:nz - 1] = np.diff(u[:, :, 0:nz - 1], axis=2) * g.idz
    dvdz[:, :, 1:nz - 1] = np.diff(v[:, :, 0:nz - 1], axis=2) * g.idz
    dudz[:, :, 0] = fi2D_m * ustar * (u[:, :, 0] + g.Ugal) / (M_sfc_loc * VONK * 0.5 * g.dz)
    dvdz[:, :, 0] = fi2D_m * ustar * v[:, :, 0] / (M_sfc_loc * VONK * 0.5 * g.dz)
    dwdz[:, :, 0:nz - 1] = np.diff(w, axis=2) * g.idz
    dwdz[:, :, nz - 1] = 0.0
    return dudz, dvdz, dwdz


def Derivz_TH(g, TH, ustar, qz_sfc, fi2D_h):
    """src/operations/Derivatives.py:243-273."""
    d = np.zeros_like(TH)
    d[:, :, 1:] = np.diff(TH, axis=2) * g.idz
    d[:, :, 0] = fi2D_h * (-qz_sfc / ustar) / (VONK * 0.5 * g.dz)
    return d


def Derivz_Generic_uvp(g, F):
    """src/operations/Derivatives.py:310-332."""
    d = np.zeros_like(F)
    d[:, :, 1:] = np.diff(F, axis=2) * g.idz
    d[:, :, 0] = 0
    return d


def Derivz_Generic_w(g, F):
    """src/operations/Derivatives.py:340-362."""
    d = np.zeros_like(F)
    d[:, :, :-1] = np.diff(F, axis=2) * g.idz
    d[:, :, -1] = 0
    return d


def velocityGradients(g, u, v, w, u_fft, v_fft, w_fft, ustar, M_sfc_loc, MOSTfunctions):
    """src/operations/Derivatives.py:47-103."""
    dudx, dvdx, dwdx = (Derivxy(g, u_fft, g.kx2), Derivxy(g, v_fft, g.kx2),
                        Derivxy(g, w_fft, g.kx2))
    dudy, dvdy, dwdy = (Derivxy(g, u_fft, g.ky2), Derivxy(g, v_fft, g.ky2),
                        Derivxy(g, w_fft, g.ky2))
    fi2D_m = MOSTfunctions[4]
    dudz, dvdz, dwdz = Derivz_M(g, u, v, w, ustar, M_sfc_loc, fi2D_m)
    return dudx, dvdx, dwdx, dudy, dvdy, dwdy, dudz, dvdz, dwdz


def potentialTemperatureGradients(g, TH, ustar, qz_sfc, MOSTfunctions):
    """src/operations/Derivatives.py:107-153 (re-transforms TH, :140)."""
    THf = _rfft2(TH)
    dTHdx = Derivxy(g, THf, g.kx2)
    dTHdy = Derivxy(g, THf, g.ky2)
    dTHdz = Derivz_TH(g, TH, ustar, qz_sfc, MOSTfunctions[5])
    return dTHdx, dTHdy, dTHdz


# ---------------------------------------------------------------------------
# src/utilities/Utilities.py
# ---------------------------------------------------------------------------

def PlanarMean(F):
    """Utilities.py:41-56."""
    return np.mean(F, axis=(0, 1))


def StagGridAvg(F):
    """Utilities.py:64-79."""
    return 0.5 * (F[:, :, :-1] + F[:, :, 1:])


def Imfilter(x):
    """Utilities.py:87-115: 3x3 periodic box mean of a 2-D plane."""
    acc = np.zeros_like(x)
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            acc = acc + np.roll(np.roll(x, di, axis=0), dj, axis=1)
    return acc / 9.0


def Roots(coeffs, init_guess=1.0 + 0j, tol=1e-6, max_iter=20):
    """Utilities.py:123-241: Laguerre iteration on a polynomial (descending
    coefficients); returns the root, or nan+0j when not converged."""
    p = [complex(c) for c in coeffs]
    n = len(p) - 1
    dp = [p[i] * (n - i) for i in range(n)]
    d2p = [dp[i] * (n - 1 - i) for i in range(n - 1)]

    def polyval(cs, x):
        acc = 0j
        for c in cs:
            acc = acc * x + c
        return acc

    eps = 1e-16 + 0j
    x = complex(init_guess)
    it = 0
    converged = False
    while it < max_iter and not converged:
        f = polyval(p, x)
        df = polyval(dp, x)
        d2f = polyval(d2p, x)
        fs = eps if abs(f) < abs(eps) else f
        G = df / fs
        H = G ** 2 - d2f / fs
        disc = (n - 1) * (n * H - G ** 2)
        sq = np.sqrt(np.complex128(disc))
        d1 = G + sq
        d2 = G - sq
        den = d1 if abs(d1) > abs(d2) else d2
        den = eps if abs(den) < abs(eps) else den
        x_new = x - n / den
        converged = (abs(x_new - x) < tol * (1 + abs(x))) or (abs(f) < tol)
        x = complex(x_new)
        it += 1
    return x if converged else complex(np.nan, 0.0)


BETA_GUESSES = (0.5, 0.75, 1.0, 1.25, 1.5, 2.0, 3.0, 4.5)


def ComputeBeta(ff, ee, dd, cc, bb, aa):
    """DynamicSGS_LASDD_SM.py:53-95 (ComputeBeta1) and
    DynamicSGS_ScalarLASDD_SM.py:54-96 (ComputeBeta2): the largest real root
    in (0, 5) over 8 Laguerre starts, default 1."""
    nz = len(ff)
    out = np.empty(nz)
    for k in range(nz):
        coeffs = [ff[k], ee[k], dd[k], cc[k], bb[k], aa[k]]
        best = np.nan
        for gs in BETA_GUESSES:
            r = Roots(coeffs, init_guess=gs)
            if np.isnan(r.real):
                continue
            if abs(r.imag) < 1e-6 and r.real > 0 and r.real < 5.0:
                best = r.real if np.isnan(best) else max(best, r.real)
        out[k] = 1.0 if np.isnan(best) else best
    return out


def _clip_ratio(num, den):
    r = num / den
    mask = (np.abs(den) < 1e-10) | (r < 0) | (r > 1)
    return np.where(mask, 0.0, r)


def Cs2_3D_from_LM_MM(LM, MM):
    """DynamicSGS_LASDD_SM.py:103-130 vmapped over levels (:357): box-filter
    LM and MM per level, ratio, clip to [0, 1] (0 where |MM|<1e-10)."""
    out = np.empty_like(LM)
    with np.errstate(divide="ignore", invalid="ignore"):
        for k in range(LM.shape[2]):
            a = Imfilter(LM[:, :, k])
            b = Imfilter(MM[:, :, k])
            out[:, :, k] = _clip_ratio(a, b)
    return out


# ---------------------------------------------------------------------------
# src/initialization
# ---------------------------------------------------------------------------

def Initialize_RayleighDampingLayer(g):
    """src/initialization/Initialization.py:395-460 (profiles only; the
    reference broadcasts them to 3-D)."""
    inv = 1.0 / g.RelaxTime_nondim
    ztop = g.l_z / g.z_scale
    thick = ztop - g.z_damping_nondim
    z = np.arange(g.nz) * g.dz
    m = (z >= g.z_damping_nondim) & (z <= ztop)
    full = np.where(m, 0.5 * inv * (1.0 - np.cos(np.pi * (z - g.z_damping_nondim) / thick)), 0.0)
    zs = (np.arange(g.nz) + 0.5) * g.dz
    ms = (zs >= g.z_damping_nondim) & (zs <= ztop)
    stag = np.where(ms, 0.5 * inv * (1.0 - np.cos(np.pi * (zs - g.z_damping_nondim) / thick)), 0.0)
    return full, stag


# ---------------------------------------------------------------------------
# src/pde
# ---------------------------------------------------------------------------

def Advection_NoDealias(g, u, v, w, dudy, dudz, dvdx, dvdz, dwdx, dwdy):
    """src/pde/NSE_AdvectionTerms.py:163-238."""
    nz = g.nz
    a1 = v * (dudy - dvdx)
    a2 = w * (dudz - dwdx)
    a2G = StagGridAvg(a2)
    Cx = np.zeros_like(u)
    Cx[:, :, 1:nz - 1] = a1[:, :, 1:nz - 1] + a2G[:, :, 1:nz - 1]
    Cx[:, :, 0] = a1[:, :, 0] + 0.5 * a2[:, :, 1]
    Cx[:, :, nz - 1] = a1[:, :, nz - 1] + a2[:, :, nz - 1]
    a1 = u * (dvdx - dudy)
    a2 = w * (dvdz - dwdy)
    a2G = StagGridAvg(a2)
    Cy = np.zeros_like(u)
    Cy[:, :, 1:nz - 1] = a1[:, :, 1:nz - 1] + a2G[:, :, 1:nz - 1]
    Cy[:, :, 0] = a1[:, :, 0] + 0.5 * a2[:, :, 1]
    Cy[:, :, nz - 1] = a1[:, :, nz - 1] + a2[:, :, nz - 1]
    uG = StagGridAvg(u)
    vG = StagGridAvg(v)
    Cz = np.zeros_like(u)
    Cz[:, :, 1:nz - 1] = (uG[:, :, 0:nz - 2] * (dwdx[:, :, 1:nz - 1] - dudz[:, :, 1:nz - 1])
                          + vG[:, :, 0:nz - 2] * (dwdy[:, :, 1:nz - 1] - dvdz[:, :, 1:nz - 1]))
    Cz[:, :, 0] = 0
    Cz[:, :, nz - 1] = 0
    return Cx, Cy, Cz


def Advection_Dealias(g, u, v, w, dudy, dudz, dvdx, dvdz, dwdx, dwdy):
    """src/pde/NSE_AdvectionTerms.py:65-155."""
    nz = g.nz
    D = lambda F: Dealias1(g, FFT(F))
    up, vp, wp = D(u), D(v), D(w)
    dudy_p, dudz_p, dvdx_p = D(dudy), D(dudz), D(dvdx)
    dvdz_p, dwdx_p, dwdy_p = D(dvdz), D(dwdx), D(dwdy)
    out = []
    for a1, a2 in ((vp * (dudy_p - dvdx_p), wp * (dudz_p - dwdx_p)),
                   (up * (dvdx_p - dudy_p), wp * (dvdz_p - dwdy_p))):
        a2G = StagGridAvg(a2)
        cc = np.zeros_like(up)
        cc[:, :, 1:nz - 1] = a1[:, :, 1:nz - 1] + a2G[:, :, 1:nz - 1]
        cc[:, :, 0] = a1[:, :, 0] + 0.5 * a2[:, :, 1]
        cc[:, :, nz - 1] = a1[:, :, nz - 1] + a2[:, :, nz - 1]
        out.append(Dealias2(g, FFT_pad(cc)))
    uG = StagGridAvg(up)
    vG = StagGridAvg(vp)
    cc = np.zeros_like(up)
    cc[:, :, 1:nz - 1] = (uG[:, :, 0:nz - 2] * (dwdx_p[:, :, 1:nz - 1] - dudz_p[:, :, 1:nz - 1])
                          + vG[:, :, 0:nz - 2] * (dwdy_p[:, :, 1:nz - 1] - dvdz_p[:, :, 1:nz - 1]))
    cc[:, :, 0] = 0
    cc[:, :, nz - 1] = 0
    out.append(Dealias2(g, FFT_pad(cc)))
    return tuple(out)


def Advection(g, *a):
    """src/pde/NSE_AdvectionTerms.py:246-275."""
    return Advection_Dealias(g, *a) if g.optDealias == 1 else Advection_NoDealias(g, *a)


def ScalarAdvection(g, u, v, w, dsdx, dsdy, dsdz):
    """src/pde/SCL_AdvectionTerms.py:57-201 (both branches)."""
    nz = g.nz
    if g.optDealias == 1:
        D = lambda F: Dealias1(g, FFT(F))
        u, v, w, dsdx, dsdy, dsdz = (D(u), D(v), D(w), D(dsdx), D(dsdy), D(dsdz))
    cc1 = u * dsdx + v * dsdy
    cc2 = w * dsdz
    cc = np.zeros_like(u)
    cc[:, :, 0:nz - 1] = cc1[:, :, 0:nz - 1] + StagGridAvg(cc2)
    cc[:, :, nz - 1] = cc1[:, :, nz - 1] + cc2[:, :, nz - 1]
    if g.optDealias == 1:
        return Dealias2(g, FFT_pad(cc))
    return cc


def Buoyancy(g, TH, H=None):
    """src/pde/NSE_BuoyancyTerms.py:49-96 (optBuoyancy=0) and :104-150
    (optBuoyancy=1): THv = (TH + T_0)(1 + 0.61 H), H = Q with moisture
    (Main.py:323), dry H = 0."""
    THv = TH + g.T_0_nondim
    if H is not None:
        THv = THv * (1 + 0.61 * H)
    bar = PlanarMean(THv).reshape(1, 1, -1)
    if g.optBuoyancy == 0:
        nrm = (THv - bar) / g.T_0_nondim
    else:
        nrm = (THv - bar) / bar
    b = np.zeros_like(TH)
    b[:, :, 1:] = 0.5 * g.g_nondim * (nrm[:, :, 1:] + nrm[:, :, :-1])
    b[:, :, 0] = 0.0
    return b


def DivStress(g, txx, tyy, tzz, txy, txz, tyz):
    """src/pde/NSE_SGSTerms.py:55-105."""
    dxtxx = Derivxy(g, FFT(txx), g.kx2)
    dxtxy = Derivxy(g, FFT(txy), g.kx2)
    dxtxz = Derivxy(g, FFT(txz), g.kx2)
    dytyy = Derivxy(g, FFT(tyy), g.ky2)
    dytxy = Derivxy(g, FFT(txy), g.ky2)
    dytyz = Derivxy(g, FFT(tyz), g.ky2)
    dztxz = Derivz_Generic_w(g, txz)
    dztyz = Derivz_Generic_w(g, tyz)
    dztzz = Derivz_Generic_uvp(g, tzz)
    return (dxtxx + dytxy + dztxz, dxtxy + dytyy + dztyz, dxtxz + dytyz + dztzz)


def DivFlux(g, qx, qy, qz):
    """src/pde/SCL_SGSTerms.py:55-89."""
    return (Derivxy(g, FFT(qx), g.kx2) + Derivxy(g, FFT(qy), g.ky2)
            + Derivz_Generic_w(g, qz))


def RHS_Momentum(g, u, v, w, Cx, Cy, Cz, buoyancy, divtx, divty, divtz, Ug=None, Vg=None,
                 Uadv=None, Vadv=None):
    """src/pde/NSE_AllTerms.py:48-104.  Ug, Vg: the iteration's geostrophic
    profiles (optGeoWind >= 1, Main.py:170-174; default the constant
    Ug2/Vg2); Uadv, Vadv: the large-scale advection tendencies
    (optAdvection >= 1, :86-88), None when off."""
    Ug = (g.Ug if Ug is None else np.asarray(Ug)).reshape(1, 1, -1)
    Vg = (g.Vg if Vg is None else np.asarray(Vg)).reshape(1, 1, -1)
    RHS_u = -Cx - divtx - g.f_coriolis_nondim * (Vg - v)
    RHS_v = -Cy - divty + g.f_coriolis_nondim * (Ug - g.Ugal - u)
    RHS_w = -Cz - divtz + buoyancy
    if Uadv is not None:
        RHS_u = RHS_u + np.asarray(Uadv).reshape(1, 1, -1)
        RHS_v = RHS_v + np.asarray(Vadv).reshape(1, 1, -1)
    if g.optDamping == 1:
        rs = g.RayleighDampCoeff_stag.reshape(1, 1, -1)
        rf = g.RayleighDampCoeff.reshape(1, 1, -1)
        RHS_u = RHS_u - rs * (u - PlanarMean(u).reshape(1, 1, -1))
        RHS_v = RHS_v - rs * (v - PlanarMean(v).reshape(1, 1, -1))
        RHS_w = RHS_w - rf * w
    return RHS_u, RHS_v, RHS_w


def RHS_Scalar(g, TH, adv, divq, THadv=None):
    """src/pde/SCL_AllTerms.py:48-82; THadv: the large-scale advection
    tendency (optAdvection >= 1, :70-71), None when off."""
    R = -adv - divq
    if THadv is not None:
        R = R + np.asarray(THadv).reshape(1, 1, -1)
    if g.optDamping == 1:
        rs = g.RayleighDampCoeff_stag.reshape(1, 1, -1)
        R = R - rs * (TH - PlanarMean(TH).reshape(1, 1, -1))
    return R


def RHS_Moisture(g, Q, adv, divqm, Qadv=None):
    """src/pde/SCL_AllTerms.py:86-110: -adv - div qm + Qadv, then the sponge
    on the fluctuation about the plane mean."""
    R = -adv - divqm
    R = R + (np.zeros(g.nz) if Qadv is None else np.asarray(Qadv)).reshape(1, 1, -1)
    if g.optDamping == 1:
        rs = g.RayleighDampCoeff_stag.reshape(1, 1, -1)
        R = R - rs * (Q - PlanarMean(Q).reshape(1, 1, -1))
    return R


def AB2_Q(g, Q, R, Rp):
    """src/pde/SCL_TimeAdvancement.py:79-99: AB2, top level = level nz-2 +
    q_inversion dz."""
    Qn = Q + g.dt_nondim * (1.5 * R - 0.5 * Rp)
    Qn[:, :, -1] = Qn[:, :, -2] + g.q_inversion_nondim * g.dz
    return Qn


def SurfaceMoistureFlux(g, Q, ustar, MOST, qsfc_val=None):
    """Main.py:253-266 with SurfaceFlux.py:451-504: the surface moisture
    flux plane for optMoistureSurfBC 0 (constant MoistureFlux), 1 (series
    flux) and 2 (prescribed surface Q, diagnosed through MOST with the heat
    stability functions; homogeneous uses the plane mean of Q at level 0)."""
    nx, ny = g.nx, g.ny
    if g.optMoistureSurfBC == 0:
        return g.qm_sfc.copy()
    if g.optMoistureSurfBC == 1:
        return qsfc_val * np.ones((nx, ny))
    psi_h, psi_h0 = MOST[2], MOST[3]
    zref = g.zMoisture if g.zMoisture > 0.0 else g.z0T
    denom = np.log(0.5 * g.dz * g.z_scale / zref) + psi_h - psi_h0
    if g.optSurfFlux == 0:
        return ustar * VONK * (qsfc_val - np.mean(Q[:, :, 0])) * np.ones((nx, ny)) / denom
    return ustar * VONK * (qsfc_val - Q[:, :, 0]) / denom


def PressureRC(g, u, v, w, RHS_u, RHS_v, RHS_w, Ru_p, Rv_p, Rw_p, divtz):
    """src/pde/NSE_PressureTerms.py:109-182."""
    nz = g.nz
    dtn = g.dt_nondim
    Rx = RHS_u - (1 / 3) * Ru_p + (2 / (3 * dtn)) * u
    Ry = RHS_v - (1 / 3) * Rv_p + (2 / (3 * dtn)) * v
    Rz = RHS_w - (1 / 3) * Rw_p + (2 / (3 * dtn)) * w
    fRx, fRy, fRz = FFT(Rx), FFT(Ry), FFT(Rz)
    nyr = g.ny_rfft
    RC_r = np.zeros((g.nx, nyr, nz + 1))
    RC_i = np.zeros((g.nx, nyr, nz + 1))
    fbot = _rfft2(divtz[:, :, 0])
    ftop = _rfft2(divtz[:, :, -1])
    RC_r[:, :, 0] = -g.dz * fbot.real
    RC_i[:, :, 0] = -g.dz * fbot.imag
    RC_r[:, :, nz] = -g.dz * ftop.real
    RC_i[:, :, nz] = -g.dz * ftop.imag
    kr = pressure_kr(g)[:, :, None]
    kc = pressure_kc(g)[:, :, None]
    hr = -kr * fRx[:, :, :-1].imag - kc * fRy[:, :, :-1].imag
    hi = kr * fRx[:, :, :-1].real + kc * fRy[:, :, :-1].real
    RC_r[:, :, 1:nz] = hr + np.diff(fRz.real, axis=2) / g.dz
    RC_i[:, :, 1:nz] = hi + np.diff(fRz.imag, axis=2) / g.dz
    return RC_r, RC_i, fRz.real


def pressure_kr(g):
    """NSE_PressureTerms_Thomas.py:170-176 (integer kx, Nyquist row = -nx/2)."""
    i = np.arange(g.nx)
    k = np.where(i < g.nx // 2, i, i - g.nx).astype(float)
    return np.broadcast_to(k[:, None], (g.nx, g.ny_rfft))


def pressure_kc(g):
    """NSE_PressureTerms_Thomas.py:178-180."""
    return np.broadcast_to(np.arange(g.ny_rfft, dtype=float)[None, :], (g.nx, g.ny_rfft))


def pressure_tridiag(g):
    """a, b, c of NSE_PressureTerms_Thomas.py:186-197."""
    nz = g.nz
    a = np.ones(nz + 1) / g.dz ** 2
    c = np.ones(nz + 1) / g.dz ** 2
    a[0] = 0
    a[nz] = -1
    c[0] = 1
    c[nz] = 0
    bb = -(pressure_kr(g) ** 2 + pressure_kc(g) ** 2 + 2.0 / g.dz ** 2)
    b = np.repeat(bb[:, :, None], nz + 1, axis=2)
    b[:, :, 0] = -1.0
    b[:, :, nz] = 1.0
    return a, b, c


def ThomasPressureInit(g):
    """NSE_PressureTerms_Thomas.py:151-214 + _thomas_factorize (:74-95)."""
    a, b, c = pressure_tridiag(g)
    bs = b.copy()
    bs[0, 0, 1:g.nz] = -(1.0 + 2.0 / g.dz ** 2)
    N = bs.shape[-1]
    bm = bs.copy()
    m = np.zeros((N - 1,) + bs.shape[:2])
    for k in range(1, N):
        mk = a[k] / bm[..., k - 1]
        bm[..., k] = bm[..., k] - mk * c[k - 1]
        m[k - 1] = mk
    return bm, m, c


def _thomas_solve(m_all, b_mod, c, d):
    """NSE_PressureTerms_Thomas.py:99-143."""
    d = d.copy()
    N = d.shape[-1]
    for k in range(1, N):
        d[..., k] = d[..., k] - m_all[k - 1] * d[..., k - 1]
    x = np.zeros_like(d)
    x[..., N - 1] = d[..., N - 1] / b_mod[..., N - 1]
    for k in range(N - 2, -1, -1):
        x[..., k] = (d[..., k] - c[k] * x[..., k + 1]) / b_mod[..., k]
    return x


def ThomasPressureSolve(g, RC_real, RC_imag, fRz_real, thomas):
    """NSE_PressureTerms_Thomas.py:218-292."""
    b_thomas, m_thomas, c = thomas
    nz = g.nz
    RC_r = RC_real.copy()
    RC_i = RC_imag.copy()
    for R in (RC_r, RC_i):
        R[0, 0, :] = 0.0
        R[g.nx_rfft - 1, :, :] = 0.0
        R[:, g.ny_rfft - 1, :] = 0.0
    xr = _thomas_solve(m_thomas, b_thomas, c, RC_r)
    xi = _thomas_solve(m_thomas, b_thomas, c, RC_i)
    fpr = xr[..., 1:].copy()
    fpi = xi[..., 1:].copy()
    zf = RC_real[0, 0, 0]
    zero_mode = np.concatenate([[zf], zf + np.cumsum(fRz_real[0, 0, 1:nz] * g.dz)])
    fpr[0, 0] = zero_mode
    fpi[0, 0] = 0.0
    fp = fpr + 1j * fpi
    p = irfft2(fp, (g.nx, g.ny))
    pf = FFT(p)
    dpdx = Derivxy(g, pf, g.kx2)
    dpdy = Derivxy(g, pf, g.ky2)
    dpdz = Derivz_Generic_uvp(g, p)
    return p, dpdx, dpdy, dpdz


def PressureSolveLU(g, RC_real, RC_imag, fRz_real):
    """NSE_PressureTerms.py:220-335: dense (nz+1)^2 solve per wavenumber,
    special modes skipped, zero mode by integration (legacy path; used as a
    regression oracle for the Thomas solve)."""
    a, b, c = pressure_tridiag(g)
    nz = g.nz
    fpr = np.zeros((g.nx, g.ny_rfft, nz))
    fpi = np.zeros((g.nx, g.ny_rfft, nz))
    idx = np.arange(nz + 1)
    for i in range(g.nx):
        for j in range(g.ny_rfft):
            if (i == 0 and j == 0) or i == g.nx_rfft - 1 or j == g.ny_rfft - 1:
                continue
            Lm = np.zeros((nz + 1, nz + 1))
            Lm[idx, idx] = b[i, j]
            Lm[idx[1:], idx[:-1]] = a[1:]
            Lm[idx[:-1], idx[1:]] = c[:nz]
            fpr[i, j] = np.linalg.solve(Lm, RC_real[i, j])[1:]
            fpi[i, j] = np.linalg.solve(Lm, RC_imag[i, j])[1:]
    zf = RC_real[0, 0, 0]
    fpr[0, 0] = np.concatenate([[zf], zf + np.cumsum(fRz_real[0, 0, 1:nz] * g.dz)])
    p = irfft2(fpr + 1j * fpi, (g.nx, g.ny))
    pf = FFT(p)
    return p, Derivxy(g, pf, g.kx2), Derivxy(g, pf, g.ky2), Derivz_Generic_uvp(g, p)


def AB2_uvw(g, u, v, w, Ru, Ru_p, Rv, Rv_p, Rw, Rw_p):
    """src/pde/NSE_TimeAdvancement.py:45-76."""
    dt = g.dt_nondim
    un = u + dt * (1.5 * Ru - 0.5 * Ru_p)
    vn = v + dt * (1.5 * Rv - 0.5 * Rv_p)
    wn = w + dt * (1.5 * Rw - 0.5 * Rw_p)
    un[:, :, -1] = un[:, :, -2]
    vn[:, :, -1] = vn[:, :, -2]
    wn[:, :, -1] = 0.0
    wn[:, :, 0] = 0.0
    return un, vn, wn


def AB2_TH(g, TH, R, R_p):
    """src/pde/SCL_TimeAdvancement.py:46-71."""
    Tn = TH + g.dt_nondim * (1.5 * R - 0.5 * R_p)
    Tn[:, :, -1] = Tn[:, :, -2] + g.inversion_nondim * g.dz
    return Tn


# ---------------------------------------------------------------------------
# src/subgridscale
# ---------------------------------------------------------------------------

def StrainsUVPnodes(g, dudx, dvdx, dwdx, dudy, dvdy, dwdy, dudz, dvdz, dwdz):
    """src/subgridscale/StrainRates.py:56-211 (the non-padded part, shared
    by both branches)."""
    nz = g.nz
    uz = np.zeros_like(dudx); vz = np.zeros_like(dudx)
    wx = np.zeros_like(dudx); wy = np.zeros_like(dudx)
    uz[:, :, 1:nz - 1] = StagGridAvg(dudz[:, :, 1:nz])
    vz[:, :, 1:nz - 1] = StagGridAvg(dvdz[:, :, 1:nz])
    wx[:, :, 1:nz - 1] = StagGridAvg(dwdx[:, :, 1:nz])
    wy[:, :, 1:nz - 1] = StagGridAvg(dwdy[:, :, 1:nz])
    uz[:, :, 0] = dudz[:, :, 0]
    vz[:, :, 0] = dvdz[:, :, 0]
    wx[:, :, 0] = 0.5 * dwdx[:, :, 1]
    wy[:, :, 0] = 0.5 * dwdy[:, :, 1]
    uz[:, :, nz - 1] = dudz[:, :, nz - 1]
    vz[:, :, nz - 1] = dvdz[:, :, nz - 1]
    wx[:, :, nz - 1] = dwdx[:, :, nz - 1]
    wy[:, :, nz - 1] = dwdy[:, :, nz - 1]
    S11, S22, S33 = dudx.copy(), dvdy.copy(), dwdz.copy()
    S12 = 0.5 * (dudy + dvdx)
    S13 = 0.5 * (uz + wx)
    S23 = 0.5 * (vz + wy)
    S = _Smag(S11, S22, S33, S12, S13, S23)
    return S11, S22, S33, S12, S13, S23, S


def _Smag(S11, S22, S33, S12, S13, S23):
    return np.sqrt(2 * (S11 ** 2 + S22 ** 2 + S33 ** 2 + 2 * S12 ** 2
                        + 2 * S13 ** 2 + 2 * S23 ** 2))


def StrainsWnodes(g, dudx, dvdx, dwdx, dudy, dvdy, dwdy, dudz, dvdz, dwdz):
    """src/subgridscale/StrainRates.py:219-368 (unpadded fields; S13, S23,
    S_w and also S11..S23 for the padded branch)."""
    nz = g.nz
    z = lambda: np.zeros_like(dudx)
    ux, uy, vx, vy, wz = z(), z(), z(), z(), z()
    ux[:, :, 1:nz - 1] = StagGridAvg(dudx[:, :, :nz - 1])
    uy[:, :, 1:nz - 1] = StagGridAvg(dudy[:, :, :nz - 1])
    vx[:, :, 1:nz - 1] = StagGridAvg(dvdx[:, :, :nz - 1])
    vy[:, :, 1:nz - 1] = StagGridAvg(dvdy[:, :, :nz - 1])
    uz, vz = dudz.copy(), dvdz.copy()
    wx, wy = dwdx.copy(), dwdy.copy()
    wz[:, :, 1:nz - 1] = StagGridAvg(dwdz[:, :, :nz - 1])
    ux[:, :, 0] = dudx[:, :, 0]
    uy[:, :, 0] = dudy[:, :, 0]
    vx[:, :, 0] = dvdx[:, :, 0]
    vy[:, :, 0] = dvdy[:, :, 0]
    wx[:, :, 0] = 0.5 * (dwdx[:, :, 0] + dwdx[:, :, 1])
    wy[:, :, 0] = 0.5 * (dwdy[:, :, 0] + dwdy[:, :, 1])
    wz[:, :, 0] = dwdz[:, :, 0]
    S11, S22, S33 = ux, vy, wz
    S12 = 0.5 * (uy + vx)
    S13 = 0.5 * (uz + wx)
    S23 = 0.5 * (vz + wy)
    return S11, S22, S33, S12, S13, S23, _Smag(S11, S22, S33, S12, S13, S23)


def Wall(g, u, v, M_sfc_loc, psi2D_m, psi2D_m0):
    """src/subgridscale/SGSStresses_SM.py:63-90."""
    denom = np.log(0.5 * g.dz * g.z_scale / g.z0m) + psi2D_m - psi2D_m0
    ustar = VONK * M_sfc_loc / denom
    txz = -(ustar ** 2) * (u[:, :, 0] + g.Ugal) / M_sfc_loc
    tyz = -(ustar ** 2) * v[:, :, 0] / M_sfc_loc
    return txz, tyz


def StressesUVP(g, S11, S22, S33, S12, S, Cs2):
    """SGSStresses_SM.py:149-182 (and :98-141 on padded inputs, without the
    final Dealias2)."""
    pre = -2 * (g.L ** 2) * Cs2 * S
    out = [pre * S11, pre * S22, pre * S33, pre * S12]
    for t in out:
        t[:, :, -1] = 0
    return out


def StressesW_interior(g, S13, S23, S_w, Cs2):
    """SGSStresses_SM.py:251-297 (and :190-243 on padded inputs) without the
    wall row."""
    nz = g.nz
    txz = np.zeros_like(S_w)
    tyz = np.zeros_like(S_w)
    pre = -2 * (g.L ** 2) * StagGridAvg(Cs2[:, :, :nz - 1]) * S_w[:, :, 1:nz - 1]
    txz[:, :, 1:nz - 1] = pre * S13[:, :, 1:nz - 1]
    tyz[:, :, 1:nz - 1] = pre * S23[:, :, 1:nz - 1]
    txz[:, :, -1] = 0
    tyz[:, :, -1] = 0
    return txz, tyz


def ScalarFluxesUVP(g, dTHdx, dTHdy, S, CsPr):
    """ScalarSGSFluxes_SM.py:98-126 (and :55-90 on padded inputs)."""
    pre = -2 * (g.L ** 2) * CsPr * S
    qx = pre * dTHdx
    qy = pre * dTHdy
    qx[:, :, -1] = 0
    qy[:, :, -1] = 0
    return qx, qy


def ScalarFluxesW_interior(g, dTHdz, S, CsPr):
    """ScalarSGSFluxes_SM.py:185-222 (and :134-177 on padded inputs) without
    the surface row."""
    nz = g.nz
    qz = np.zeros_like(S)
    qz[:, :, 1:nz - 1] = (-2 * (g.L ** 2) * StagGridAvg(CsPr[:, :, :nz - 1])
                          * S[:, :, 1:nz - 1] * dTHdz[:, :, 1:nz - 1])
    qz[:, :, -1] = 0
    return qz


def StressesUVP_WL(g, S11, S22, S33, S12, Cwl):
    """SGSStresses_WL.py:111-141 (and :70-103 on padded inputs, without the
    final Dealias2): tau_ij = -2 L^(4/3) C_WL S_ij."""
    pre = -2 * g.L ** (4 / 3) * Cwl
    out = [pre * S11, pre * S22, pre * S33, pre * S12]
    for t in out:
        t[:, :, -1] = 0
    return out


def StressesW_interior_WL(g, S13, S23, Cwl):
    """SGSStresses_WL.py:196-239 (and :146-191 padded) without the wall row."""
    nz = g.nz
    txz = np.zeros_like(S13)
    tyz = np.zeros_like(S13)
    pre = -2 * g.L ** (4 / 3) * StagGridAvg(Cwl[:, :, :nz - 1])
    txz[:, :, 1:nz - 1] = pre * S13[:, :, 1:nz - 1]
    tyz[:, :, 1:nz - 1] = pre * S23[:, :, 1:nz - 1]
    return txz, tyz


def ScalarFluxesUVP_WL(g, dTHdx, dTHdy, CwlPr):
    """ScalarSGSFluxes_WL.py:96-117 (and :55-91 padded): q = -L^(4/3) C grad."""
    pre = -g.L ** (4 / 3) * CwlPr
    qx = pre * dTHdx
    qy = pre * dTHdy
    qx[:, :, -1] = 0
    qy[:, :, -1] = 0
    return qx, qy


def ScalarFluxesW_interior_WL(g, dTHdz, CwlPr):
    """ScalarSGSFluxes_WL.py:169-201 (and :122-164 padded) without the
    surface row."""
    nz = g.nz
    qz = np.zeros_like(dTHdz)
    qz[:, :, 1:nz - 1] = (-g.L ** (4 / 3) * StagGridAvg(CwlPr[:, :, :nz - 1])
                          * dTHdz[:, :, 1:nz - 1])
    return qz


def LASDD_SM(g, u, v, w, S11, S22, S33, S12, S13, S23, S):
    """src/subgridscale/DynamicSGS_LASDD_SM.py:138-367 (optSgs 1 = LASDD,
    3 = LAD with beta = 1).

    The plane-mean sums a1..e2 and the quintic aa..ff (:314-319) are pinned
    against the two-level condition they encode, formed from M_ij(beta)
    directly (tests/test_oracle_beta_algebra.py); "parts" returns the
    filtered tensors for that check."""
    nz = g.nz
    L, TFR = g.L, g.TFR
    u_ = u.copy(); v_ = v.copy()
    w_ = np.zeros_like(w)
    w_[:, :, 1:nz - 1] = StagGridAvg(w[:, :, 1:nz])
    w_[:, :, 0] = 0.5 * w[:, :, 1]
    w_[:, :, nz - 1] = w[:, :, nz - 2]
    F1 = lambda F: Filtering_Level1(g, FFT(F))
    F2 = lambda F: Filtering_Level2(g, FFT(F))
    prods = [u_ * u_, v_ * v_, w_ * w_, u_ * v_, u_ * w_, v_ * w_]
    Sij = [S11, S22, S33, S12, S13, S23]
    wt = [1, 1, 1, 2, 2, 2]
    uh = [F1(x) for x in (u_, v_, w_)]
    ud = [F2(x) for x in (u_, v_, w_)]
    pp_h = [F1(x) for x in prods]
    pp_d = [F2(x) for x in prods]
    Sh = [F1(x) for x in Sij]
    Sd = [F2(x) for x in Sij]
    S_hat = _Smag(*Sh)
    S_hatd = _Smag(*Sd)
    SSh = [F1(S * x) for x in Sij]
    SSd = [F2(S * x) for x in Sij]
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    Lij = [pp_h[n] - uh[a] * uh[b] for n, (a, b) in enumerate(pairs)]
    Qij = [pp_d[n] - ud[a] * ud[b] for n, (a, b) in enumerate(pairs)]

    def contract(A, B):
        return (A[0] * B[0] + A[1] * B[1] + A[2] * B[2]
                + 2 * (A[3] * B[3] + A[4] * B[4] + A[5] * B[5]))

    if g.optSgs in (1, 2):
        a1 = PlanarMean(2 * L ** 2 * contract(Lij, SSh))
        a2 = PlanarMean(2 * L ** 2 * contract(Qij, SSd))
        b1 = PlanarMean(2 * L ** 2 * TFR ** 2 * S_hat * contract(Lij, Sh))
        b2 = PlanarMean(2 * L ** 2 * TFR ** 4 * S_hatd * contract(Qij, Sd))
        c1 = PlanarMean((2 * L ** 2) ** 2 * contract(SSh, SSh))
        c2 = PlanarMean((2 * L ** 2) ** 2 * contract(SSd, SSd))
        d1 = PlanarMean(4 * L ** 4 * TFR ** 4 * S_hat ** 2 * contract(Sh, Sh))
        d2 = PlanarMean(4 * L ** 4 * TFR ** 8 * S_hatd ** 2 * contract(Sd, Sd))
        e1 = PlanarMean(8 * L ** 4 * TFR ** 2 * S_hat * contract(Sh, SSh))
        e2 = PlanarMean(8 * L ** 4 * TFR ** 4 * S_hatd * contract(Sd, SSd))
        aa = a1 * c2 - a2 * c1
        bb = a2 * e1 - b1 * c2
        cc = b2 * c1 - a1 * e2 - a2 * d1
        dd = b1 * e2 - b2 * e1
        ee = a1 * d2 + b2 * d1
        ff = -b1 * d2
        beta1 = ComputeBeta(ff, ee, dd, cc, bb, aa)
        coeffs = np.stack([ff, ee, dd, cc, bb, aa])
    else:
        beta1 = np.ones(nz)
        coeffs = None
    b3 = beta1.reshape(1, 1, -1)
    T1 = 2 * L ** 2
    T2 = 2 * (TFR * L) ** 2
    Mij = [T1 * SSh[n] - T2 * b3 * S_hat * Sh[n] for n in range(6)]
    LM = contract(Lij, Mij)
    MM = contract(Mij, Mij)
    Cs2_3D = Cs2_3D_from_LM_MM(LM, MM)
    Cs2_1D_avg1 = PlanarMean(np.sqrt(Cs2_3D)) ** 2
    Cs2_1D_avg2 = PlanarMean(Cs2_3D)
    return dict(u_=u_, v_=v_, w_=w_, u_hat=uh, u_hatd=ud, S_hat=S_hat, S_hatd=S_hatd,
                Cs2_3D=Cs2_3D, Cs2_1D_avg1=Cs2_1D_avg1, Cs2_1D_avg2=Cs2_1D_avg2,
                beta1_1D=beta1, LM=LM, MM=MM, beta_coeffs=coeffs,
                parts=dict(Lij=Lij, Qij=Qij, Sh=Sh, Sd=Sd, SSh=SSh, SSd=SSd))


def ScalarLASDD_SM(g, mom, TH, dTHdx, dTHdy, dTHdz, S):
    """src/subgridscale/DynamicSGS_ScalarLASDD_SM.py:139-316.

    The scalar sums and quintic coefficients are pinned as in LASDD_SM
    (tests/test_oracle_beta_algebra.py)."""
    nz = g.nz
    L, TFR = g.L, g.TFR
    u_, v_, w_ = mom["u_"], mom["v_"], mom["w_"]
    uh, ud = mom["u_hat"], mom["u_hatd"]
    S_hat, S_hatd = mom["S_hat"], mom["S_hatd"]
    THz = np.zeros_like(TH)
    THz[:, :, 1:nz - 1] = StagGridAvg(dTHdz[:, :, 1:nz])
    THz[:, :, 0] = dTHdz[:, :, 0]
    THz[:, :, nz - 1] = dTHdz[:, :, nz - 1]
    F1 = lambda F: Filtering_Level1(g, FFT(F))
    F2 = lambda F: Filtering_Level2(g, FFT(F))
    TH_hat = F1(TH)
    uTH_hat = [F1(x * TH) for x in (u_, v_, w_)]
    grads = (dTHdx, dTHdy, THz)
    dh = [F1(x) for x in grads]
    SdTh = [F1(S * x) for x in grads]
    LT = [uTH_hat[i] - uh[i] * TH_hat for i in range(3)]
    a2t = LT[0] * SdTh[0] + LT[1] * SdTh[1] + LT[2] * SdTh[2]
    b2t = LT[0] * S_hat * dh[0] + LT[1] * S_hat * dh[1] + LT[2] * S_hat * dh[2]
    c2t = SdTh[0] ** 2 + SdTh[1] ** 2 + SdTh[2] ** 2
    d2t = (SdTh[0] * S_hat * dh[0] + SdTh[1] * S_hat * dh[1] + SdTh[2] * S_hat * dh[2])
    e2t = (S_hat * dh[0]) ** 2 + (S_hat * dh[1]) ** 2 + (S_hat * dh[2]) ** 2
    if g.optSgs in (1, 2):
        TH_hatd = F2(TH)
        uTH_hatd = [F2(x * TH) for x in (u_, v_, w_)]
        dd_ = [F2(x) for x in grads]
        SdTd = [F2(S * x) for x in grads]
        QT = [uTH_hatd[i] - ud[i] * TH_hatd for i in range(3)]
        a2 = PlanarMean(L ** 2 * a2t)
        b2 = PlanarMean(-L ** 2 * TFR ** 2 * b2t)
        c2 = PlanarMean(L ** 4 * c2t)
        d2 = PlanarMean(-2 * L ** 4 * TFR ** 2 * d2t)
        e2 = PlanarMean(L ** 4 * TFR ** 4 * e2t)
        a4 = PlanarMean(L ** 2 * (QT[0] * SdTd[0] + QT[1] * SdTd[1] + QT[2] * SdTd[2]))
        b4 = PlanarMean(-L ** 2 * TFR ** 4 * (QT[0] * S_hatd * dd_[0] + QT[1] * S_hatd * dd_[1]
                                              + QT[2] * S_hatd * dd_[2]))
        c4 = PlanarMean(L ** 4 * (SdTd[0] ** 2 + SdTd[1] ** 2 + SdTd[2] ** 2))
        d4 = PlanarMean(-2 * L ** 4 * TFR ** 4 * (SdTd[0] * S_hatd * dd_[0] + SdTd[1] * S_hatd * dd_[1]
                                                  + SdTd[2] * S_hatd * dd_[2]))
        e4 = PlanarMean(L ** 4 * TFR ** 8 * ((S_hatd * dd_[0]) ** 2 + (S_hatd * dd_[1]) ** 2
                                             + (S_hatd * dd_[2]) ** 2))
        aa = a2 * c4 - a4 * c2
        bb = -a4 * d2 + b2 * c4
        cc = -c2 * b4 + a2 * d4 - a4 * e2
        dd = b2 * d4 - b4 * d2
        ee = a2 * e4 - b4 * e2
        ff = b2 * e4
        beta2 = ComputeBeta(ff, ee, dd, cc, bb, aa)
        coeffs = np.stack([ff, ee, dd, cc, bb, aa])
    else:
        beta2 = np.ones(nz)
        coeffs = None
    b3 = beta2.reshape(1, 1, -1)
    T_up = L ** 2 * a2t + (-L ** 2 * TFR ** 2) * b2t * b3
    T_dn = (L ** 4 * c2t + (-2 * L ** 4 * TFR ** 2) * d2t * b3
            + (L ** 4 * TFR ** 4) * e2t * b3 ** 2)
    CsPr = Cs2_3D_from_LM_MM(T_up, T_dn)
    parts = dict(LT=LT, dh=dh, SdTh=SdTh)
    if g.optSgs in (1, 2):
        parts.update(QT=QT, dd=dd_, SdTd=SdTd)
    return dict(Cs2PrRatio_3D=CsPr, Cs2PrRatio_1D=PlanarMean(CsPr), beta2_1D=beta2,
                T_up=T_up, T_dn=T_dn, beta_coeffs=coeffs, parts=parts)


def _wl_quintic(g, a1, a3, a6, a8):
    """The WL polynomial of ABL07 (DynamicSGS_LASDD_WL.py:255-274,
    DynamicSGS_ScalarLASDD_WL.py:224-243): four independent plane means,
    the other six scaled by powers of TFR.  Returns (ff, ee, dd, cc, bb, aa).

    Pinned: the polynomial equals the cross-multiplied two-level condition
    of the scale-dependent procedure formed from the WL model tensors at
    several beta, and vanishes where either level's model term does
    (tests/test_oracle_surface_advection_pins.py)."""
    T = g.TFR
    a2 = -(T ** (8 / 3)) * a1
    a4 = -2 * T ** (4 / 3) * a3
    a5 = T ** (8 / 3) * a3
    a7 = -T ** (4 / 3) * a6
    a9 = -2 * T ** (8 / 3) * a8
    a10 = T ** (16 / 3) * a8
    aa = a1 * a3 - a6 * a8
    bb = a1 * a4 - a7 * a8
    cc = a2 * a3 + a1 * a5 - a6 * a9
    dd = a2 * a4 - a7 * a9
    ee = a2 * a5 - a6 * a10
    ff = -a7 * a10
    return ff, ee, dd, cc, bb, aa


def LASDD_WL(g, u, v, w, S11, S22, S33, S12, S13, S23):
    """src/subgridscale/DynamicSGS_LASDD_WL.py:128-313 (optSgs 2 = LASDD-WL,
    4 = LAD-WL with beta = 1): Wong-Lilly M_ij = 2 L^(4/3) (S^_ij -
    TFR^(4/3) beta S^^_ij)."""
    nz = g.nz
    L, TFR = g.L, g.TFR
    u_ = u.copy(); v_ = v.copy()
    w_ = np.zeros_like(w)
    w_[:, :, 1:nz - 1] = StagGridAvg(w[:, :, 1:nz])
    w_[:, :, 0] = 0.5 * w[:, :, 1]
    w_[:, :, nz - 1] = w[:, :, nz - 2]
    F1 = lambda F: Filtering_Level1(g, FFT(F))
    F2 = lambda F: Filtering_Level2(g, FFT(F))
    prods = [u_ * u_, v_ * v_, w_ * w_, u_ * v_, u_ * w_, v_ * w_]
    Sij = [S11, S22, S33, S12, S13, S23]
    uh = [F1(x) for x in (u_, v_, w_)]
    ud = [F2(x) for x in (u_, v_, w_)]
    pp_h = [F1(x) for x in prods]
    pp_d = [F2(x) for x in prods]
    Sh = [F1(x) for x in Sij]
    Sd = [F2(x) for x in Sij]
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    Lij = [pp_h[n] - uh[a] * uh[b] for n, (a, b) in enumerate(pairs)]
    Qij = [pp_d[n] - ud[a] * ud[b] for n, (a, b) in enumerate(pairs)]

    def contract(A, B):
        return (A[0] * B[0] + A[1] * B[1] + A[2] * B[2]
                + 2 * (A[3] * B[3] + A[4] * B[4] + A[5] * B[5]))

    if g.optSgs == 2:
        co = _wl_quintic(g, PlanarMean(contract(Qij, Sh)), PlanarMean(contract(Sh, Sh)),
                         PlanarMean(contract(Lij, Sh)), PlanarMean(contract(Sd, Sd)))
        beta1 = ComputeBeta(*co)
        coeffs = np.stack(co)
    else:
        beta1 = np.ones(nz)
        coeffs = None
    b3 = beta1.reshape(1, 1, -1)
    T1 = 2 * L ** (4 / 3)
    T2 = 2 * (TFR * L) ** (4 / 3)
    Mij = [T1 * Sh[n] - T2 * b3 * Sd[n] for n in range(6)]
    LM = contract(Lij, Mij)
    MM = contract(Mij, Mij)
    Cwl = Cs2_3D_from_LM_MM(LM, MM)
    return dict(u_=u_, v_=v_, w_=w_, u_hat=uh, u_hatd=ud, S_hat=_Smag(*Sh), S_hatd=_Smag(*Sd),
                Cs2_3D=Cwl, Cs2_1D_avg1=PlanarMean(np.sqrt(Cwl)) ** 2,
                Cs2_1D_avg2=PlanarMean(Cwl), beta1_1D=beta1, LM=LM, MM=MM, beta_coeffs=coeffs)


def ScalarLASDD_WL(g, mom, TH, dTHdx, dTHdy, dTHdz):
    """src/subgridscale/DynamicSGS_ScalarLASDD_WL.py:130-275."""
    nz = g.nz
    L, TFR = g.L, g.TFR
    u_, v_, w_ = mom["u_"], mom["v_"], mom["w_"]
    uh, ud = mom["u_hat"], mom["u_hatd"]
    THz = np.zeros_like(TH)
    THz[:, :, 1:nz - 1] = StagGridAvg(dTHdz[:, :, 1:nz])
    THz[:, :, 0] = dTHdz[:, :, 0]
    THz[:, :, nz - 1] = dTHdz[:, :, nz - 1]
    F1 = lambda F: Filtering_Level1(g, FFT(F))
    F2 = lambda F: Filtering_Level2(g, FFT(F))
    TH_hat, TH_hatd = F1(TH), F2(TH)
    uTH_hat = [F1(x * TH) for x in (u_, v_, w_)]
    uTH_hatd = [F2(x * TH) for x in (u_, v_, w_)]
    grads = (dTHdx, dTHdy, THz)
    dh = [F1(x) for x in grads]
    dd_ = [F2(x) for x in grads]
    LT = [uTH_hat[i] - uh[i] * TH_hat for i in range(3)]
    QT = [uTH_hatd[i] - ud[i] * TH_hatd for i in range(3)]
    dot = lambda A, B: A[0] * B[0] + A[1] * B[1] + A[2] * B[2]
    a3t = dot(dh, dh)
    a6t = dot(LT, dh)
    a8t = dot(dd_, dd_)
    if g.optSgs == 2:
        co = _wl_quintic(g, PlanarMean(dot(QT, dh)), PlanarMean(a3t), PlanarMean(a6t),
                         PlanarMean(a8t))
        beta2 = ComputeBeta(*co)
        coeffs = np.stack(co)
    else:
        beta2 = np.ones(nz)
        coeffs = None
    b3 = beta2.reshape(1, 1, -1)
    T_up = L ** (4 / 3) * (a6t - TFR ** (4 / 3) * b3 * dot(LT, dd_))
    T_dn = L ** (8 / 3) * (a3t - 2 * TFR ** (4 / 3) * b3 * dot(dh, dd_)
                           + TFR ** (8 / 3) * b3 ** 2 * a8t)
    CwlPr = Cs2_3D_from_LM_MM(T_up, T_dn)
    return dict(Cs2PrRatio_3D=CwlPr, Cs2PrRatio_1D=PlanarMean(CwlPr), beta2_1D=beta2,
                T_up=T_up, T_dn=T_dn, beta_coeffs=coeffs)


def SGS_momentum_stresses(g, grads, Cs2_3D, u, v, M_sfc_loc, psi2D_m, psi2D_m0):
    """Stresses on uvp and w nodes for either branch:
    DynamicSGS_Main.py:119-282 / StaticSGS_Main.py:76-250; the Wong-Lilly
    base model (optSgs 2, 4) drops |S| and uses 2 L^(4/3) C_WL."""
    wl = g.optSgs in (2, 4)
    uvp = StrainsUVPnodes(g, *grads)
    wn = StrainsWnodes(g, *grads)
    S11, S22, S33, S12, S13, S23, S_uvp = uvp
    if g.optDealias == 1:
        D = lambda F: Dealias1(g, FFT(F))
        P = [D(x) for x in (S11, S22, S33, S12, S13, S23)]
        S_uvp_pad = _Smag(*P)
        Cp = D(Cs2_3D)
        if wl:
            tt = StressesUVP_WL(g, P[0], P[1], P[2], P[3], Cp)
        else:
            tt = StressesUVP(g, P[0], P[1], P[2], P[3], S_uvp_pad, Cp)
        txx, tyy, tzz, txy = [Dealias2(g, FFT_pad(t)) for t in tt]
        Pw = [D(x) for x in wn[:6]]
        S_w_pad = _Smag(*Pw)
        if wl:
            txz_p, tyz_p = StressesW_interior_WL(g, Pw[4], Pw[5], Cp)
        else:
            txz_p, tyz_p = StressesW_interior(g, Pw[4], Pw[5], S_w_pad, Cp)
        txz = Dealias2(g, FFT_pad(txz_p))
        tyz = Dealias2(g, FFT_pad(tyz_p))
    else:
        S_uvp_pad = S_uvp
        S_w_pad = wn[6]
        if wl:
            txx, tyy, tzz, txy = StressesUVP_WL(g, S11, S22, S33, S12, Cs2_3D)
            txz, tyz = StressesW_interior_WL(g, wn[4], wn[5], Cs2_3D)
        else:
            txx, tyy, tzz, txy = StressesUVP(g, S11, S22, S33, S12, S_uvp, Cs2_3D)
            txz, tyz = StressesW_interior(g, wn[4], wn[5], wn[6], Cs2_3D)
    twx, twy = Wall(g, u, v, M_sfc_loc, psi2D_m, psi2D_m0)
    txz[:, :, 0] = twx
    tyz[:, :, 0] = twy
    return dict(txx=txx, tyy=tyy, tzz=tzz, txy=txy, txz=txz, tyz=tyz,
                S11=S11, S22=S22, S33=S33, S12=S12, S13=S13, S23=S23,
                S_uvp=S_uvp, S_uvp_pad=S_uvp_pad, S_w=wn[6], S_w_pad=S_w_pad)


def SGS_scalar_fluxes(g, st, CsPr, dTHdx, dTHdy, dTHdz, qz_sfc):
    """DynamicSGS_Main.py:420-478 / StaticSGS_Main.py:258-360 (SM, or WL
    for optSgs 2, 4)."""
    wl = g.optSgs in (2, 4)
    if g.optDealias == 1:
        D = lambda F: Dealias1(g, FFT(F))
        dxp, dyp, dzp, cp = D(dTHdx), D(dTHdy), D(dTHdz), D(CsPr)
        if wl:
            qxp, qyp = ScalarFluxesUVP_WL(g, dxp, dyp, cp)
            qzp = ScalarFluxesW_interior_WL(g, dzp, cp)
        else:
            qxp, qyp = ScalarFluxesUVP(g, dxp, dyp, st["S_uvp_pad"], cp)
            qzp = ScalarFluxesW_interior(g, dzp, st["S_w_pad"], cp)
        qx = Dealias2(g, FFT_pad(qxp))
        qy = Dealias2(g, FFT_pad(qyp))
        qz = Dealias2(g, FFT_pad(qzp))
    elif wl:
        qx, qy = ScalarFluxesUVP_WL(g, dTHdx, dTHdy, CsPr)
        qz = ScalarFluxesW_interior_WL(g, dTHdz, CsPr)
    else:
        qx, qy = ScalarFluxesUVP(g, dTHdx, dTHdy, st["S_uvp"], CsPr)
        qz = ScalarFluxesW_interior(g, dTHdz, st["S_w"], CsPr)
    qz[:, :, 0] = qz_sfc
    return qx, qy, qz


# ---------------------------------------------------------------------------
# src/surface/SurfaceFlux.py
# ---------------------------------------------------------------------------

def MOSTstable(zL):
    """SurfaceFlux.py:53-77."""
    psi_m = 5.0 * zL
    psi_h = 5.0 * zL
    return psi_m, psi_h, 1.0 + psi_m, 1.0 + psi_h


def MOSTunstable(zL):
    """SurfaceFlux.py:83-113."""
    x = (1 - 15 * zL) ** 0.25
    psi_m = (-2 * np.log(0.5 * (1 + x)) - np.log(0.5 * (1 + x ** 2))
             + 2 * np.arctan(x) - np.pi / 2)
    psi_h = -2 * np.log(0.5 * (1 + x ** 2))
    return psi_m, psi_h, 1.0 / x, (1.0 / x) ** 2


def _update_MOSTfunctions(g, ustar, qz_sfc_avg, TH_ref):
    """SurfaceFlux.py:119-158."""
    invOB = -(VONK * g.g_nondim * qz_sfc_avg) / ((ustar ** 3) * TH_ref)
    f = MOSTstable if qz_sfc_avg <= 0 else MOSTunstable
    z1 = (0.5 * g.dz) * invOB
    z0 = (g.z0m / g.z_scale) * invOB
    zref = g.zTemperature if g.zTemperature > 0.0 else g.z0T
    zT = (zref / g.z_scale) * invOB
    psi_m, psi_h, fi_m, fi_h = f(z1)
    psi_m0 = f(z0)[0]
    psi_h0 = f(zT)[1]
    return (psi_m, psi_m0, psi_h, psi_h0, fi_m, fi_h), invOB


def SurfaceFlux(g, u, v, TH, MOST, sfc_val=None):
    """SurfaceFlux.py:166-443: the six heat-flux variants selected by
    (optSurfBC, optSurfFlux); returns (M_sfc_loc, ustar, qz_sfc_2D,
    qz_sfc_avg, invOB, MOSTfunctions)."""
    nx, ny = g.nx, g.ny
    one = np.ones((nx, ny))
    psi_m, psi_m0, psi_h, psi_h0 = MOST[0], MOST[1], MOST[2], MOST[3]
    hetero = g.optSurfFlux != 0
    Mloc = np.sqrt((u[:, :, 0] + g.Ugal) ** 2 + v[:, :, 0] ** 2)
    denom_m = np.log(0.5 * g.dz * g.z_scale / g.z0m) + psi_m - psi_m0
    if g.optSurfBC == 0:
        qz_avg = np.mean(g.qz_sfc)
        qz2 = g.qz_sfc
        if hetero:
            M = Mloc
            TH_ref = TH[:, :, 0] + g.T_0_nondim
        else:
            M = np.mean(Mloc) * one
            TH_ref = (np.mean(TH[:, :, 0]) + g.T_0_nondim) * one
        ustar = np.maximum(VONK * M / denom_m, 1e-3)
    elif g.optSurfBC == 1:
        qz_avg = sfc_val
        qz2 = sfc_val * one
        if hetero:
            M = Mloc
            TH_ref = TH[:, :, 0] + g.T_0_nondim
        else:
            M = np.mean(Mloc) * one
            TH_ref = (np.mean(TH[:, :, 0]) + g.T_0_nondim) * one
        ustar = np.maximum(VONK * M / denom_m, 1e-3)
    else:
        zref = g.zTemperature if g.zTemperature > 0.0 else g.z0T
        denom_h = np.log(0.5 * g.dz * g.z_scale / zref) + psi_h - psi_h0
        if hetero:
            M = Mloc
            ustar = np.maximum(VONK * M / denom_m, 1e-3)
            THa = TH[:, :, 0]
            qz2 = ustar * VONK * (sfc_val - THa) / denom_h
            qz_avg = np.mean(qz2)
            TH_ref = (np.mean(THa) + g.T_0_nondim) * one
        else:
            M = np.mean(Mloc) * one
            ustar = np.maximum(VONK * M / denom_m, 1e-3)
            tha = np.mean(TH[:, :, 0])
            TH_ref = (tha + g.T_0_nondim) * one
            qz2 = ustar * VONK * (sfc_val - tha) * one / denom_h
            qz_avg = np.mean(qz2)
    MOSTn, invOB = _update_MOSTfunctions(g, ustar, qz_avg, TH_ref)
    return M, ustar, qz2, qz_avg, invOB, MOSTn


# ---------------------------------------------------------------------------
# Step composition (src/Main.py:156-540)
# ---------------------------------------------------------------------------

@dataclass
class State:
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    TH: np.ndarray
    RHS_u: Optional[np.ndarray] = None
    RHS_v: Optional[np.ndarray] = None
    RHS_w: Optional[np.ndarray] = None
    RHS_TH: Optional[np.ndarray] = None
    Cs2_3D: Optional[np.ndarray] = None
    Cs2PrRatio_3D: Optional[np.ndarray] = None
    MOST: tuple = ()
    Q: Optional[np.ndarray] = None        # optMoisture >= 1 (Main.py:71-73)
    RHS_Q: Optional[np.ndarray] = None
    iteration: int = 1
    CFLmax: float = 0.0
    diag: dict = field(default_factory=dict)


def initial_state(g: Grid, u, v, w, TH, Q=None) -> State:
    """Main.py:66-111 (state allocation; u is stored minus Ugal and TH minus
    T_0 by Initialization.py:58-100, the caller passes stored values)."""
    nx, ny, nz = g.nx, g.ny, g.nz
    z2 = np.zeros((nx, ny))
    if g.optSgs in (2, 4):
        c0, p0 = g.Cwl, g.CwlPrRatio
    else:
        c0, p0 = g.Cs2, g.Cs2PrRatio
    return State(u=u.astype(float).copy(), v=v.astype(float).copy(),
                 w=w.astype(float).copy(), TH=TH.astype(float).copy(),
                 RHS_u=np.zeros((nx, ny, nz)), RHS_v=np.zeros((nx, ny, nz)),
                 RHS_w=np.zeros((nx, ny, nz)), RHS_TH=np.zeros((nx, ny, nz)),
                 Cs2_3D=c0 * np.ones((nx, ny, nz)), Cs2PrRatio_3D=p0 * np.ones((nx, ny, nz)),
                 MOST=(z2.copy(), z2.copy(), z2.copy(), z2.copy(), 1.0 + z2, 1.0 + z2),
                 Q=None if Q is None else np.asarray(Q, dtype=float).copy(),
                 RHS_Q=None if Q is None else np.zeros((nx, ny, nz)),
                 iteration=g.istep)


def step(g: Grid, s: State, thomas=None, sfc_val=None, keep=False, forcing=None,
         qsfc_val=None) -> State:
    """One iteration of the time loop, src/Main.py:158-540 (SM or WL base
    models).  forcing: the iteration's profiles of Main.py:167-188, dict
    with Ug, Vg (optGeoWind >= 1) and/or Uadv, Vadv, THadv (optAdvection >=
    1; all three when any is given), nondimensional.  With optMoisture >= 1
    the state carries Q and RHS_Q; qsfc_val is the iteration's moisture
    surface value (MoistureSurfaceBC_series[iteration-1]) for
    optMoistureSurfBC 1 and 2, and forcing may hold Qadv."""
    fc = forcing or {}
    adv_on = any(k in fc for k in ("Uadv", "Vadv", "THadv"))
    zprof = np.zeros(g.nz)
    if thomas is None:
        thomas = ThomasPressureInit(g)
    it = s.iteration
    first = it == g.istep
    Ru_p, Rv_p, Rw_p, RT_p = s.RHS_u, s.RHS_v, s.RHS_w, s.RHS_TH
    # Main.py:193-197
    u, u_fft = Filtering_Explicit(g, FFT(s.u))
    v, v_fft = Filtering_Explicit(g, FFT(s.v))
    w, w_fft = Filtering_Explicit(g, FFT(s.w))
    TH, _ = Filtering_Explicit(g, FFT(s.TH))
    moist = g.optMoisture >= 1
    if moist:
        Q, _ = Filtering_Explicit(g, FFT(s.Q))               # Main.py:198-199
    # Main.py:213-243
    M, ustar, qz2, qz_avg, invOB, MOST = SurfaceFlux(g, u, v, TH, s.MOST, sfc_val)
    if moist:                                                  # Main.py:253-266
        qm2 = SurfaceMoistureFlux(g, Q, ustar, MOST, qsfc_val)
        qm_avg = float(np.mean(qm2))
    # Main.py:275-290
    grads = velocityGradients(g, u, v, w, u_fft, v_fft, w_fft, ustar, M, MOST)
    dudx, dvdx, dwdx, dudy, dvdy, dwdy, dudz, dvdz, dwdz = grads
    dTHdx, dTHdy, dTHdz = potentialTemperatureGradients(g, TH, ustar, qz2, MOST)
    if moist:   # moistureGradients (Derivatives.py:281-306): the TH rule with qm
        dQdx, dQdy, dQdz = potentialTemperatureGradients(g, Q, ustar, qm2, MOST)
    # Main.py:301-311
    Cx, Cy, Cz = Advection(g, u, v, w, dudy, dudz, dvdx, dvdz, dwdx, dwdy)
    THadv = ScalarAdvection(g, u, v, w, dTHdx, dTHdy, dTHdz)
    if moist:
        QAdv = ScalarAdvection(g, u, v, w, dQdx, dQdy, dQdz)   # Main.py:313-318
    # Main.py:322-327
    buoy = Buoyancy(g, TH, Q if moist else None)
    # Main.py:333-443 (SGS)
    dynamic = (1 <= g.optSgs <= 4) and (first or it % g.dynamicSGS_call_time == 0)
    diag = {}
    if dynamic:
        uvp = StrainsUVPnodes(g, *grads)
        if g.optSgs in (2, 4):
            mom = LASDD_WL(g, u, v, w, *uvp[:6])
        else:
            mom = LASDD_SM(g, u, v, w, *uvp)
        Cs2_3D = mom["Cs2_3D"]
        st = SGS_momentum_stresses(g, grads, Cs2_3D, u, v, M, MOST[0], MOST[1])
        if g.optSgs in (2, 4):
            sc = ScalarLASDD_WL(g, mom, TH, dTHdx, dTHdy, dTHdz)
        else:
            sc = ScalarLASDD_SM(g, mom, TH, dTHdx, dTHdy, dTHdz, uvp[6])
        CsPr = sc["Cs2PrRatio_3D"]
        diag.update(Cs2_1D_avg1=mom["Cs2_1D_avg1"], Cs2_1D_avg2=mom["Cs2_1D_avg2"],
                    beta1_1D=mom["beta1_1D"], Cs2PrRatio_1D=sc["Cs2PrRatio_1D"],
                    beta2_1D=sc["beta2_1D"])
        if keep:
            diag.update(LM=mom["LM"], MM=mom["MM"], T_up=sc["T_up"], T_dn=sc["T_dn"],
                        beta1_coeffs=mom["beta_coeffs"], beta2_coeffs=sc["beta_coeffs"])
    else:
        Cs2_3D = s.Cs2_3D
        CsPr = s.Cs2PrRatio_3D
        st = SGS_momentum_stresses(g, grads, Cs2_3D, u, v, M, MOST[0], MOST[1])
    divtx, divty, divtz = DivStress(g, st["txx"], st["tyy"], st["tzz"], st["txy"],
                                    st["txz"], st["tyz"])
    qx, qy, qz = SGS_scalar_fluxes(g, st, CsPr, dTHdx, dTHdy, dTHdz, qz2)
    divq = DivFlux(g, qx, qy, qz)
    if moist:   # Main.py:356-370: the scalar model with Cs2PrRatio, qm at k = 0
        qHx, qHy, qHz = SGS_scalar_fluxes(g, st, CsPr, dQdx, dQdy, dQdz, qm2)
        divqm = DivFlux(g, qHx, qHy, qHz)
    # Main.py:464-473
    RHS_u, RHS_v, RHS_w = RHS_Momentum(
        g, u, v, w, Cx, Cy, Cz, buoy, divtx, divty, divtz, Ug=fc.get("Ug"), Vg=fc.get("Vg"),
        Uadv=fc.get("Uadv", zprof) if adv_on else None,
        Vadv=fc.get("Vadv", zprof) if adv_on else None)
    RHS_TH = RHS_Scalar(g, TH, THadv, divq, THadv=fc.get("THadv", zprof) if adv_on else None)
    if moist:                                                    # Main.py:474-475
        RHS_Q = RHS_Moisture(g, Q, QAdv, divqm, fc.get("Qadv"))
    # Main.py:481-500
    RC_r, RC_i, fRz = PressureRC(g, u, v, w, RHS_u, RHS_v, RHS_w, Ru_p, Rv_p, Rw_p, divtz)
    if g.optPressureSolver == 1:
        p, dpdx, dpdy, dpdz = ThomasPressureSolve(g, RC_r, RC_i, fRz, thomas)
    else:
        p, dpdx, dpdy, dpdz = PressureSolveLU(g, RC_r, RC_i, fRz)
    RHS_u = RHS_u - dpdx
    RHS_v = RHS_v - dpdy
    RHS_w = RHS_w - dpdz
    # Main.py:506-512
    RQ_p = s.RHS_Q
    if first:
        Ru_p, Rv_p, Rw_p, RT_p = RHS_u, RHS_v, RHS_w, RHS_TH
        if moist:
            RQ_p = RHS_Q
    # Main.py:518-526
    un, vn, wn = AB2_uvw(g, u, v, w, RHS_u, Ru_p, RHS_v, Rv_p, RHS_w, Rw_p)
    THn = AB2_TH(g, TH, RHS_TH, RT_p)
    Qn = AB2_Q(g, Q, RHS_Q, RQ_p) if moist else None              # Main.py:528-529
    # Main.py:534-540
    CFL = max(np.max(np.abs(un)) * g.dt_nondim / g.dx,
              np.max(np.abs(vn)) * g.dt_nondim / g.dy,
              np.max(np.abs(wn)) * g.dt_nondim / g.dz)
    diag.update(M_sfc_loc=M, ustar=ustar, qz_sfc=qz2, qz_sfc_avg=qz_avg, invOB=invOB,
                CFL=CFL, txy=st["txy"], txz=st["txz"], tyz=st["tyz"], qz=qz,
                dudz=dudz, dvdz=dvdz, dTHdz=dTHdz)
    if moist:
        diag.update(qm_sfc=qm2, qm_sfc_avg=qm_avg, dQdz=dQdz, qHz=qHz)
        if keep:
            diag.update(Q_f=Q, dQ=(dQdx, dQdy, dQdz), QAdv=QAdv, qHx=qHx, qHy=qHy,
                        divqm=divqm, RHS_Q=RHS_Q)
    if keep:
        diag.update(u_f=u, v_f=v, w_f=w, TH_f=TH, grads=grads, dTH=(dTHdx, dTHdy, dTHdz),
                    Cx=Cx, Cy=Cy, Cz=Cz, THadv=THadv, buoyancy=buoy, stresses=st,
                    divt=(divtx, divty, divtz), qx=qx, qy=qy, divq=divq,
                    RC_r=RC_r, RC_i=RC_i, fRz=fRz, p=p, dp=(dpdx, dpdy, dpdz),
                    RHS=(RHS_u, RHS_v, RHS_w, RHS_TH))
    return State(u=un, v=vn, w=wn, TH=THn, RHS_u=RHS_u, RHS_v=RHS_v, RHS_w=RHS_w,
                 RHS_TH=RHS_TH, Cs2_3D=Cs2_3D, Cs2PrRatio_3D=CsPr, MOST=MOST,
                 Q=Qn, RHS_Q=RHS_Q if moist else None,
                 iteration=it + 1, CFLmax=max(s.CFLmax, CFL), diag=diag)


# ---------------------------------------------------------------------------
# Statistics (src/utilities/Statistics.py; Main.py:545-617)
# ---------------------------------------------------------------------------

# profile keys of the statistics dictionary in the order of Statistics.py:
# 319-330 (dry runs: the moisture keys Q, dQdz, Q2, wQ, qHz are zero
# profiles and qm_sfc is 0.0)
STATS_PROFILES = ("U", "V", "W", "TH", "dUdz", "dVdz", "dTHdz", "u2", "v2", "w2", "TH2",
                  "uv", "uw", "vw", "uTH", "vTH", "wTH", "txy", "txz", "tyz", "qz",
                  "Cs2_1", "Cs2_2", "Cs2PrRatio", "Beta1", "Beta2")
STATS_SCALARS = ("M_sfc", "ustar", "qz_sfc")
STATS_MOIST_PROFILES = ("Q", "dQdz", "Q2", "wQ", "qHz")


def InitializeStats(g: Grid) -> dict:
    """Statistics.py:333-382 (InitializeStats) and the ResetStats branch of
    ComputeStats (:134-149): zero accumulators plus the constants."""
    z = np.zeros(g.nz)
    d = {k: z.copy() for k in STATS_PROFILES + STATS_MOIST_PROFILES}
    d.update({k: 0.0 for k in STATS_SCALARS + ("qm_sfc",)})
    d["Ugal"] = g.Ugal
    d["ZeRo1D"] = z.copy()
    return d


def ComputeStats(g: Grid, acc: dict, u, v, w, TH, dudz, dvdz, dTHdz, M_sfc_loc, ustar,
                 qz_sfc_avg, txy, txz, tyz, qz, Cs2_1D_avg1, Cs2_1D_avg2, Cs2PrRatio_1D,
                 beta1_1D, beta2_1D, Q=None, dQdz=None, qHz=None, qm_sfc_avg=0.0) -> dict:
    """Statistics.py:47-330, ComputeStats with ResetFlag = 0, u_scale =
    TH_scale = Q_scale = 1 (DerivedVars.py:49-69).  u, v, w, TH (and Q) are
    the state after the step's AB2; the other inputs are the step's own
    intermediates (Main.py:550-557).  Q = None: dry (the moisture keys add
    zeros)."""
    Ugal = g.Ugal
    nz = g.nz
    out = dict(acc)
    # :158-183 mean profiles and mean gradients
    mU = np.mean(u, axis=(0, 1)) + Ugal
    mV = np.mean(v, axis=(0, 1))
    mW = np.mean(w, axis=(0, 1))
    mTH_anom = np.mean(TH, axis=(0, 1))
    mTH = mTH_anom + g.T_0_nondim
    out["U"] = acc["U"] + mU
    out["V"] = acc["V"] + mV
    out["W"] = acc["W"] + mW
    out["TH"] = acc["TH"] + mTH
    out["dUdz"] = acc["dUdz"] + np.mean(dudz, axis=(0, 1)) / g.z_scale
    out["dVdz"] = acc["dVdz"] + np.mean(dvdz, axis=(0, 1)) / g.z_scale
    out["dTHdz"] = acc["dTHdz"] + np.mean(dTHdz, axis=(0, 1)) / g.z_scale
    # :188-221 level variances and horizontal fluxes (fluctuations about the
    # level means)
    uf = u + Ugal - mU[None, None, :]
    vf = v - mV[None, None, :]
    wf = w - mW[None, None, :]
    tf = TH - mTH_anom[None, None, :]
    for key, a, b in (("u2", uf, uf), ("v2", vf, vf), ("w2", wf, wf), ("TH2", tf, tf),
                      ("uv", uf, vf), ("uTH", uf, tf), ("vTH", vf, tf)):
        out[key] = acc[key] + np.mean(a * b, axis=(0, 1))
    # :226-265 vertical fluxes on the staggered grid: u, v, TH averaged to
    # w nodes for k >= 1, level 0 taken as is (u with Ugal added)
    us = np.zeros_like(u)
    vs = np.zeros_like(v)
    ts = np.zeros_like(TH)
    us[:, :, 1:nz] = StagGridAvg(u)
    vs[:, :, 1:nz] = StagGridAvg(v)
    ts[:, :, 1:nz] = StagGridAvg(TH)
    us[:, :, 0] = u[:, :, 0] + Ugal
    vs[:, :, 0] = v[:, :, 0]
    ts[:, :, 0] = TH[:, :, 0]
    mus = np.mean(us, axis=(0, 1)) + Ugal
    mvs = np.mean(vs, axis=(0, 1))
    mws = np.mean(w, axis=(0, 1))
    mts = np.mean(ts, axis=(0, 1))
    usf = us + Ugal - mus[None, None, :]
    vsf = vs - mvs[None, None, :]
    wsf = w - mws[None, None, :]
    tsf = ts - mts[None, None, :]
    # moisture (:175-176, :206-207, :235-264, :275, :291)
    if Q is not None:
        mQ = np.mean(Q, axis=(0, 1))
        out["Q"] = acc["Q"] + mQ
        out["dQdz"] = acc["dQdz"] + np.mean(dQdz, axis=(0, 1)) / g.z_scale
        out["Q2"] = acc["Q2"] + np.mean((Q - mQ[None, None, :]) ** 2, axis=(0, 1))
        qs = np.zeros_like(Q)
        qs[:, :, 1:nz] = StagGridAvg(Q)
        qs[:, :, 0] = Q[:, :, 0]
        mqs = np.mean(qs, axis=(0, 1))
        out["wQ"] = acc["wQ"] + np.mean(wsf * (qs - mqs[None, None, :]), axis=(0, 1))
        out["qHz"] = acc["qHz"] + np.mean(qHz, axis=(0, 1))
        out["qm_sfc"] = acc["qm_sfc"] + float(qm_sfc_avg)
    out["uw"] = acc["uw"] + np.mean(usf * wsf, axis=(0, 1))
    out["vw"] = acc["vw"] + np.mean(vsf * wsf, axis=(0, 1))
    out["wTH"] = acc["wTH"] + np.mean(wsf * tsf, axis=(0, 1))
    # :270-283 SGS stress and flux profiles
    out["txy"] = acc["txy"] + np.mean(txy, axis=(0, 1))
    out["txz"] = acc["txz"] + np.mean(txz, axis=(0, 1))
    out["tyz"] = acc["tyz"] + np.mean(tyz, axis=(0, 1))
    out["qz"] = acc["qz"] + np.mean(qz, axis=(0, 1))
    # :288-296 surface variables
    out["M_sfc"] = acc["M_sfc"] + float(np.mean(M_sfc_loc))
    out["ustar"] = acc["ustar"] + float(np.sqrt(np.mean(ustar ** 2)))
    out["qz_sfc"] = acc["qz_sfc"] + float(qz_sfc_avg)
    # :301-305 SGS coefficient profiles
    z = np.zeros(nz)
    for key, prof in (("Cs2_1", Cs2_1D_avg1), ("Cs2_2", Cs2_1D_avg2),
                      ("Cs2PrRatio", Cs2PrRatio_1D), ("Beta1", beta1_1D), ("Beta2", beta2_1D)):
        out[key] = acc[key] + (z if prof is None else np.asarray(prof, dtype=np.float64))
    return out


def AverageStats(acc: dict, samples: int) -> dict:
    """Main.py:589-595: every accumulated key divided by the sample count,
    the constants Ugal and ZeRo1D kept as they are."""
    return {k: (v if k in ("Ugal", "ZeRo1D") else v / samples) for k, v in acc.items()}


def run(g: Grid, s: State, nsteps: int, sfc_series=None) -> State:
    """Advance ``nsteps`` iterations; ``sfc_series[i]`` is the surface BC
    value for iteration i+1 (Main.py:224, :233 index with iteration-1)."""
    th = ThomasPressureInit(g)
    for _ in range(nsteps):
        sv = None if sfc_series is None else sfc_series[s.iteration - 1]
        s = step(g, s, th, sv)
    return s
