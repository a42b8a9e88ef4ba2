"""The reference's on-disk run inputs, read and written in its formats.

* ``input/vel.npy``: an (N, 3) float64 column stack of u, v, w, each the
  Fortran-order flattening of an (nx, ny, nz) field (written by the
  examples' ``CreateInputs_*.py``, e.g. CreateInputs_NBL_A94.py:68-77; read
  by src/initialization/Initialization.py:58-79, which removes Ugal).
* ``input/TH.npy``: the flat Fortran-order potential temperature (K)
  (Initialization.py:86-101, which removes T_0).
* ``input/SurfaceBC.npz``: ``data_series`` and ``t_series`` of nsteps + 1
  values, ``dt_surf`` and ``optSurfBC`` (CreateSurfaceBC_GABLS1.py:121-125),
  validated against the run's dt, optSurfBC and length and made
  nondimensional by Initialization.py:107-160.
* the geostrophic-wind file (``Ug_series``/``Vg_series``, or compact
  ``t_profile``/``Ug_profile``/``Vg_profile`` interpolated in time) and the
  large-scale advection file (``Uadv_series``, ``Vadv_series``, optional
  ``THadv_series``), Initialization.py:173-323.

Fields come back in the stored form the step uses: u minus Ugal, TH minus
T_0 (u_scale = TH_scale = 1, src/config/DerivedVars.py:49-50).
"""
from __future__ import annotations

import math
import os

import numpy as np


def nsteps_of(cfg) -> int:
    """src/config/DerivedVars.py:90: ceil(SimTime / dt)."""
    return int(math.ceil(cfg["SimTime"] / cfg["dt"]))


def write_initial_fields(input_dir, u, v, w, TH):
    """u, v, w (m/s, u with Ugal included) and TH (K) as the examples write
    them: vel.npy (N, 3) and TH.npy, Fortran-order flattening."""
    os.makedirs(input_dir, exist_ok=True)
    vel = np.column_stack([np.asarray(a, dtype=np.float64).reshape(-1, order="F")
                           for a in (u, v, w)])
    np.save(os.path.join(input_dir, "vel.npy"), vel)
    np.save(os.path.join(input_dir, "TH.npy"),
            np.asarray(TH, dtype=np.float64).reshape(-1, order="F"))


def read_initial_fields(input_dir, cfg):
    """Initialize_uvw + Initialize_TH: -> (u - Ugal, v, w, TH - T_0), each
    (nx, ny, nz) float64."""
    nx, ny, nz = int(cfg["nx"]), int(cfg["ny"]), int(cfg["nz"])
    vel = np.load(os.path.join(input_dir, "vel.npy"), allow_pickle=False)
    if vel.shape != (nx * ny * nz, 3):
        raise ValueError(f"vel.npy has shape {vel.shape}, expected ({nx * ny * nz}, 3)")
    Ugal = float(cfg.get("Ugal", 0.0))
    u = np.reshape(vel[:, 0] - Ugal, (nx, ny, nz), order="F")
    v = np.reshape(vel[:, 1], (nx, ny, nz), order="F")
    w = np.reshape(vel[:, 2], (nx, ny, nz), order="F")
    TH = np.load(os.path.join(input_dir, "TH.npy"), allow_pickle=False)
    if TH.size != nx * ny * nz:
        raise ValueError(f"TH.npy has {TH.size} values, expected {nx * ny * nz}")
    TH = np.reshape(TH, (nx, ny, nz), order="F") - float(cfg["T_0"])
    return u, v, w, TH


def write_surface_bc(path, data_series, dt, optSurfBC):
    """SurfaceBC.npz as CreateSurfaceBC_GABLS1.py:121-125 writes it:
    data_series (K or K m/s), t_series = arange(len) * dt, dt_surf,
    optSurfBC (int32)."""
    data = np.asarray(data_series, dtype=np.float64)
    np.savez(path, data_series=data, t_series=np.arange(data.size) * float(dt),
             dt_surf=np.float64(dt), optSurfBC=np.int32(optSurfBC))


def read_surface_bc(path, cfg):
    """Initialize_SurfaceBC (Initialization.py:107-160): checks dt_surf,
    optSurfBC and the length nsteps + 1, then -> the nondimensional series
    (heat flux as is for optSurfBC 1, temperature minus T_0 for 2)."""
    with np.load(path, allow_pickle=False) as f:
        dt_surf = float(f["dt_surf"])
        opt = int(f["optSurfBC"])
        series = np.asarray(f["data_series"], dtype=np.float64)
    if abs(dt_surf - float(cfg["dt"])) > 1e-6:
        raise ValueError(f"SurfaceBC file dt_surf={dt_surf:.6f} s does not match "
                         f"dt={float(cfg['dt']):.6f} s")
    if opt != int(cfg["optSurfBC"]):
        raise ValueError(f"SurfaceBC file optSurfBC={opt} does not match "
                         f"optSurfBC={int(cfg['optSurfBC'])}")
    n = nsteps_of(cfg) + 1
    if series.size != n:
        raise ValueError(f"SurfaceBC series length {series.size} != nsteps+1={n}")
    if opt == 1:
        return series.copy()
    return series - float(cfg["T_0"])


def read_geowind(path, cfg):
    """Initialize_GeoWind_Varying (Initialization.py:173-250): -> (Ug, Vg),
    each (nsteps+1, nz) nondimensional (u_scale = 1).  Either full series or
    compact profiles linearly interpolated to t = i dt per level."""
    nz = int(cfg["nz"])
    n = nsteps_of(cfg) + 1
    dt = float(cfg["dt"])
    with np.load(path, allow_pickle=False) as f:
        if abs(float(f["dt_geo"]) - dt) > 1e-6:
            raise ValueError(f"GeoWind file dt_geo={float(f['dt_geo']):.6f} s does not match "
                             f"dt={dt:.6f} s")
        if int(f["optGeoWind"]) != int(cfg.get("optGeoWind", 0)):
            raise ValueError(f"GeoWind file optGeoWind={int(f['optGeoWind'])} does not match "
                             f"optGeoWind={int(cfg.get('optGeoWind', 0))}")
        if "Ug_series" in f and "Vg_series" in f:
            Ug = np.asarray(f["Ug_series"], dtype=np.float64)
            Vg = np.asarray(f["Vg_series"], dtype=np.float64)
            if Ug.shape[0] != n:
                raise ValueError(f"GeoWind series length {Ug.shape[0]} != nsteps+1={n}")
            if Ug.shape[1] != nz:
                raise ValueError(f"GeoWind nz={Ug.shape[1]} != nz={nz}")
        else:
            t = np.asarray(f["t_profile"], dtype=np.float64)
            Up = np.asarray(f["Ug_profile"], dtype=np.float64)
            Vp = np.asarray(f["Vg_profile"], dtype=np.float64)
            if Up.shape[1] != nz:
                raise ValueError(f"GeoWind nz={Up.shape[1]} != nz={nz}")
            tt = np.arange(n) * dt
            if tt[0] < t[0] - 1e-6 or tt[-1] > t[-1] + 1e-6:
                raise ValueError("GeoWind compact profiles do not cover the simulation interval")
            Ug = np.stack([np.interp(tt, t, Up[:, k]) for k in range(nz)], axis=1)
            Vg = np.stack([np.interp(tt, t, Vp[:, k]) for k in range(nz)], axis=1)
    return Ug, Vg


def read_adv_forcing(path, cfg):
    """Initialize_AdvForcing (Initialization.py:253-323), dry: -> (Uadv,
    Vadv, THadv), each (nsteps+1, nz), times z_scale = l_x / (2 pi)
    (u_scale = TH_scale = 1); a missing THadv series is zero."""
    nz = int(cfg["nz"])
    n = nsteps_of(cfg) + 1
    dt = float(cfg["dt"])
    zs = float(cfg["l_x"]) / (2 * math.pi)
    with np.load(path, allow_pickle=False) as f:
        if abs(float(f["dt_adv"]) - dt) > 1e-6:
            raise ValueError(f"AdvForcing file dt_adv={float(f['dt_adv']):.6f} s does not match "
                             f"dt={dt:.6f} s")
        if int(f["optAdvection"]) != int(cfg.get("optAdvection", 0)):
            raise ValueError(f"AdvForcing file optAdvection={int(f['optAdvection'])} does not "
                             f"match optAdvection={int(cfg.get('optAdvection', 0))}")
        U = np.asarray(f["Uadv_series"], dtype=np.float64)
        V = np.asarray(f["Vadv_series"], dtype=np.float64)
        if U.shape[0] != n:
            raise ValueError(f"AdvForcing series length {U.shape[0]} != nsteps+1={n}")
        if U.shape[1] != nz:
            raise ValueError(f"AdvForcing nz={U.shape[1]} != nz={nz}")
        TH = (np.asarray(f["THadv_series"], dtype=np.float64) if "THadv_series" in f
              else np.zeros((n, nz)))
    return U * zs, V * zs, TH * zs


def save_3d_fields(outdir, iteration, les):
    """Main.py:620-635: ALFA_3DFields_Iteration_<n>.npz with u + Ugal, v, w
    and TH + T_0, and Q when the run carries moisture (optMoisture >= 1,
    Main.py:630-631)."""
    os.makedirs(outdir, exist_ok=True)
    f = les.fields()
    cfg = les.plan.cfg
    path = os.path.join(outdir, f"ALFA_3DFields_Iteration_{iteration}.npz")
    out = dict(u=f["u"] + float(cfg.get("Ugal", 0.0)), v=f["v"], w=f["w"],
               TH=f["TH"] + float(cfg["T_0"]))
    if int(cfg.get("optMoisture", 0)) >= 1:
        out["Q"] = f["Q"]
    np.savez(path, **out)
    return path


# carried state of one run: the fields and planes of alfa_state, the
# coefficient profiles and scalars, the iteration counter and CFLmax
def save_checkpoint(path, les):
    """Exact-restart file of an LES: every carried array in its stored
    precision (bitwise), the iteration and CFLmax."""
    arrays = {f"t_{k}": v.cpu().numpy() for k, v in les.t.items()}
    arrays.update({f"p_{k}": v.cpu().numpy() for k, v in les.prof.items()})
    arrays["scalars"] = les.scalars.cpu().numpy()
    np.savez(path, iteration=np.int64(les.iteration), CFLmax=np.float64(les.CFLmax),
             precision=np.array(les.plan.precision), **arrays)


def load_checkpoint(path, les):
    """Restore save_checkpoint's file into an LES of the same configuration
    and precision; the next step continues the saved run bit for bit."""
    torch = __import__("torch")
    with np.load(path, allow_pickle=False) as f:
        if str(f["precision"]) != les.plan.precision:
            raise ValueError(f"checkpoint precision {f['precision']} != {les.plan.precision}")
        for k, v in les.t.items():
            a = f[f"t_{k}"]
            if a.shape != tuple(v.shape):
                raise ValueError(f"checkpoint field {k} has shape {a.shape}, expected "
                                 f"{tuple(v.shape)}")
            v.copy_(torch.from_numpy(a))
        for k, v in les.prof.items():
            v.copy_(torch.from_numpy(f[f"p_{k}"]))
        les.scalars.copy_(torch.from_numpy(f["scalars"]))
        les.iteration = int(f["iteration"])
        les.CFLmax = float(f["CFLmax"])


def write_moisture(input_dir, Q):
    """input/Q.npy: the flat Fortran-order specific humidity (kg/kg)."""
    os.makedirs(input_dir, exist_ok=True)
    np.save(os.path.join(input_dir, "Q.npy"), np.asarray(Q, dtype=np.float64).reshape(-1, order="F"))


def read_moisture(input_dir, cfg):
    """Initialize_Q (Initialization.py:330-343): Q.npy as (nx, ny, nz),
    absolute kg/kg (no base state removed)."""
    nx, ny, nz = int(cfg["nx"]), int(cfg["ny"]), int(cfg["nz"])
    Q = np.load(os.path.join(input_dir, "Q.npy"), allow_pickle=False)
    if Q.size != nx * ny * nz:
        raise ValueError(f"Q.npy has {Q.size} values, expected {nx * ny * nz}")
    return np.reshape(Q, (nx, ny, nz), order="F")


def read_moisture_surface_bc(path, cfg):
    """Initialize_MoistureSurfaceBC (Initialization.py:346-386): checks
    dt_moist, optMoistureSurfBC and the length nsteps + 1; the series is
    returned as stored (a flux for mode 1 with u_scale = Q_scale = 1, the
    surface Q in kg/kg for mode 2)."""
    with np.load(path, allow_pickle=False) as f:
        dt_m = float(f["dt_moist"])
        opt = int(f["optMoistureSurfBC"])
        series = np.asarray(f["data_series"], dtype=np.float64)
    if abs(dt_m - float(cfg["dt"])) > 1e-6:
        raise ValueError(f"MoistureSurfaceBC file dt_moist={dt_m:.6f} s does not match "
                         f"dt={float(cfg['dt']):.6f} s")
    if opt != int(cfg.get("optMoistureSurfBC", 0)):
        raise ValueError(f"MoistureSurfaceBC file optMoistureSurfBC={opt} does not match "
                         f"optMoistureSurfBC={int(cfg.get('optMoistureSurfBC', 0))}")
    n = nsteps_of(cfg) + 1
    if series.size != n:
        raise ValueError(f"MoistureSurfaceBC series length {series.size} != nsteps+1={n}")
    return series.copy()
