"""Seeded synthetic boundary-layer cases (initial fields and forcing series).

The generators follow the style of the reference's case scripts
(``examples/SBL_GABLS1/runs/*/CreateInputs_SBL_GABLS1.py``,
``examples/CBL_N91/runs/*/CreateInputs_CBL_N91.py``,
``examples/SBL_GABLS1/runs/*/CreateSurfaceBC_GABLS1.py``) but draw their
noise from ``numpy.random.default_rng(seed)``.  Fields are returned in the
stored form the time step works on (src/initialization/Initialization.py:
58-100): ``u`` minus ``Ugal`` and ``TH`` minus ``T_0``, shape (nx, ny, nz),
z fastest, float64.

Cases:
  * ``neutral``    — 32^3 neutral layer, 2000x2000x1000 m, LAD-SM, FGR=1.
  * ``convective`` — 128^3 convective layer, 5120x5120x2048 m, LASDD-SM, FGR=2.
  * ``stable``     — 256^3 GABLS1-like stable layer, 400^3 m, LAD-SM, FGR=2,
                     prescribed cooling surface temperature.
"""
from __future__ import annotations

import numpy as np

NEUTRAL = dict(
    nx=32, ny=32, nz=32, l_x=2000.0, l_y=2000.0, l_z=1000.0, dt=0.5,
    Ugal=10.0, optSurfFlux=0, optSurfBC=0, z0m=0.1, z0T=0.1, zTemperature=0.0,
    SensibleHeatFlux=0.0, Ug2=10.0, Vg2=0.0, f_coriolis=1e-4, inversion=0.0,
    optBuoyancy=1, T_0=300.0, optSgs=3, dynamicSGS_call_time=1, FGR=1,
    Cs2=0.01, Cs2PrRatio=0.01, optDamping=1, z_damping=700.0, RelaxTime=300.0,
    optPressureSolver=1, istep=1,
)

CONVECTIVE = dict(
    nx=128, ny=128, nz=128, l_x=5120.0, l_y=5120.0, l_z=2048.0, dt=0.2,
    Ugal=0.0, optSurfFlux=1, optSurfBC=0, z0m=0.1, z0T=0.1, zTemperature=0.0,
    SensibleHeatFlux=0.24, Ug2=1.0, Vg2=0.0, f_coriolis=1e-4, inversion=0.003,
    optBuoyancy=1, T_0=300.0, optSgs=1, dynamicSGS_call_time=1, FGR=2,
    Cs2=0.01, Cs2PrRatio=0.01, optDamping=1, z_damping=1600.0, RelaxTime=300.0,
    optPressureSolver=1, istep=1,
)

STABLE = dict(
    nx=256, ny=256, nz=256, l_x=400.0, l_y=400.0, l_z=400.0, dt=0.05,
    Ugal=5.0, optSurfFlux=0, optSurfBC=2, z0m=0.1, z0T=0.1, zTemperature=0.0,
    SensibleHeatFlux=0.0, Ug2=8.0, Vg2=0.0, f_coriolis=1.39e-4, inversion=0.01,
    optBuoyancy=1, T_0=265.0, optSgs=3, dynamicSGS_call_time=1, FGR=2,
    Cs2=0.01, Cs2PrRatio=0.01, optDamping=1, z_damping=300.0, RelaxTime=60.0,
    optPressureSolver=1, istep=1,
)

CASES = {"neutral": NEUTRAL, "convective": CONVECTIVE, "stable": STABLE}


def _z_uvp(cfg):
    dz = cfg["l_z"] / (cfg["nz"] - 1)
    return (np.arange(cfg["nz"]) + 0.5) * dz


def initial_fields(name: str, seed: int = 0, **over):
    """Return (cfg, u, v, w, TH) in stored form for case ``name``; keyword
    overrides replace config entries (e.g. a smaller grid)."""
    cfg = dict(CASES[name])
    cfg.update(over)
    nx, ny, nz = cfg["nx"], cfg["ny"], cfg["nz"]
    rng = np.random.default_rng(seed)
    z = _z_uvp(cfg)
    u = np.zeros((nx, ny, nz))
    v = np.zeros((nx, ny, nz))
    w = np.zeros((nx, ny, nz))
    TH = np.full((nx, ny, nz), cfg["T_0"])
    if name == "neutral":
        prof = cfg["Ug2"] * np.log(z / cfg["z0m"]) / np.log(cfg["l_z"] / cfg["z0m"])
        u[:] = prof[None, None, :]
        low = z < 100.0
        nlow = int(low.sum())
        u[:, :, low] += rng.uniform(-0.5, 0.5, (nx, ny, nlow))
        v[:, :, low] += rng.uniform(-0.5, 0.5, (nx, ny, nlow))
        w[:, :, low] += rng.uniform(-0.5, 0.5, (nx, ny, nlow))
    elif name == "convective":
        u[:] = cfg["Ug2"]
        TH += np.where(z > 1000.0, cfg["inversion"] * (z - 1000.0), 0.0)[None, None, :]
        low = z < 100.0
        TH[:, :, low] += rng.uniform(-0.1, 0.1, (nx, ny, int(low.sum())))
    elif name == "stable":
        u[:] = cfg["Ug2"]
        TH += np.where(z > 100.0, cfg["inversion"] * (z - 100.0), 0.0)[None, None, :]
        low = z < 50.0
        TH[:, :, low] += rng.uniform(-0.1, 0.1, (nx, ny, int(low.sum())))
    else:
        raise ValueError(f"unknown case {name!r}")
    w[:, :, 0] = 0.0
    w[:, :, nz - 1] = 0.0
    return cfg, u - cfg["Ugal"], v, w, TH - cfg["T_0"]


def surface_series(cfg, nsteps: int):
    """Non-dimensional surface BC series of length nsteps+1 (index i is time
    i*dt), as src/initialization/Initialization.py:107-160 returns it.
    optSurfBC=2: temperature anomaly T_s(t) - T_0 with T_s = T_0 - 0.25 K/h * t
    (CreateSurfaceBC_GABLS1.py linear cooling); optSurfBC=1: a heat-flux
    series that oscillates by +-30% about SensibleHeatFlux with a period of
    40 time steps (the per-step change, about 5%, exposes a series read at
    the wrong step); optSurfBC=0: None."""
    t = np.arange(nsteps + 1) * cfg["dt"]
    if cfg["optSurfBC"] == 2:
        return (cfg["T_0"] - 0.25 * t / 3600.0) - cfg["T_0"]
    if cfg["optSurfBC"] == 1:
        period = 40.0 * cfg["dt"]
        return cfg["SensibleHeatFlux"] * (1.0 + 0.3 * np.sin(2.0 * np.pi * t / period))
    return None
