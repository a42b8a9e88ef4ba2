"""Planar statistics of a run: ComputeStats / InitializeStats
(src/utilities/Statistics.py:47-382) and the sampling and output of the time
loop (src/Main.py:545-617).

The accumulator is one device vector of doubles (``alfa_stats_accumulate``
adds a sample of the state after a step and of that step's intermediates);
``averaged()`` divides by the sample count as Main.py:589-595 does and
returns the reference's dictionary keys, and ``save()`` writes
``ALFA_Statistics_Iteration_<n>.npz`` with them.  In dry runs the moisture
keys (Q, dQdz, Q2, wQ, qHz, qm_sfc) are zero, as in the reference with
optMoisture = 0.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .core import LES, _torch, current_stream

PROFILES = ("U", "V", "W", "TH", "dUdz", "dVdz", "dTHdz", "u2", "v2", "w2", "TH2", "uv", "uw",
            "vw", "uTH", "vTH", "wTH", "txy", "txz", "tyz", "qz", "Cs2_1", "Cs2_2",
            "Cs2PrRatio", "Beta1", "Beta2", "Q", "dQdz", "Q2", "wQ", "qHz")
SCALARS = ("M_sfc", "ustar", "qz_sfc", "qm_sfc")
# key order of the reference's dictionary (Statistics.py:319-330)
KEY_ORDER = ("U", "V", "W", "TH", "dUdz", "dVdz", "dTHdz", "u2", "v2", "w2", "TH2", "uv", "uw",
             "vw", "uTH", "vTH", "wTH", "txy", "txz", "tyz", "qz", "M_sfc", "ustar", "qz_sfc",
             "Cs2_1", "Cs2_2", "Cs2PrRatio", "Beta1", "Beta2", "Q", "dQdz", "Q2", "wQ", "qHz",
             "qm_sfc", "Ugal", "ZeRo1D")


class Statistics:
    """Statistics accumulator of one run (one LES)."""

    def __init__(self, les: LES):
        self.les = les
        P = les.plan
        self.nz = P.nz
        self.acc = P.zeros((len(PROFILES) * P.nz + len(SCALARS),), dtype=_torch().float64)
        self.samples = 0

    def sample(self):
        """Add the statistics of the state after the last step (call right
        after that step, on the same stream)."""
        les = self.les
        les.plan.call("alfa_stats_accumulate", C.byref(les.state),
                      C.c_void_p(self.acc.data_ptr()), current_stream())
        self.samples += 1

    def reset(self):
        """The ResetStats branch (Statistics.py:134-149)."""
        self.acc.zero_()
        self.samples = 0

    def sums(self) -> dict:
        a = self.acc.cpu().numpy()
        nz = self.nz
        d = {k: a[i * nz:(i + 1) * nz].copy() for i, k in enumerate(PROFILES)}
        base = len(PROFILES) * nz
        d.update({k: float(a[base + i]) for i, k in enumerate(SCALARS)})
        return d

    def averaged(self) -> dict:
        """Main.py:589-595: accumulated keys over the sample count; Ugal and
        ZeRo1D as they are; moisture keys zero."""
        if self.samples == 0:
            raise ValueError("no samples collected")
        s = self.sums()
        z = np.zeros(self.nz)
        out = {}
        for k in KEY_ORDER:
            if k in s:
                out[k] = s[k] / self.samples
            elif k == "Ugal":
                out[k] = float(self.les.plan.cfg.get("Ugal", 0.0))
            elif k == "ZeRo1D":
                out[k] = z.copy()
            elif k == "qm_sfc":
                out[k] = 0.0
            else:
                out[k] = z.copy()
        return out

    def save(self, outdir: str, iteration: int) -> str:
        """Main.py:597-600: ALFA_Statistics_Iteration_<iteration>.npz."""
        os.makedirs(outdir, exist_ok=True)
        path = os.path.join(outdir, f"ALFA_Statistics_Iteration_{iteration}.npz")
        np.savez(path, **self.averaged())
        return path


def run(les: LES, nsteps: int, SampleInterval: int, OutputInterval: int, outdir: str,
        stats: Statistics = None) -> list:
    """The statistics part of the time loop (Main.py:545-617): after each
    step of iteration i, sample when i % SampleInterval == 0; when
    i % OutputInterval == 0 and samples were taken, write the averages and
    reset.  Returns the paths written."""
    stats = stats or Statistics(les)
    paths = []
    for _ in range(nsteps):
        it = les.iteration
        les.step(1)
        if it % SampleInterval == 0:
            stats.sample()
        if it % OutputInterval == 0 and stats.samples > 0:
            paths.append(stats.save(outdir, it))
            stats.reset()
    return paths


def developed_state_comparison(n: int = 64, hours: float = 1.0, avg_min: float = 30.0,
                               samples: int = 60, case: str = "stable") -> dict:
    """Mixed precision against FP64 at a developed state.

    Three runs of case `case` at n^3 (dt scaled with the grid spacing from
    the case's base grid) for `hours` of simulated time: mixed, FP64, and
    FP64 with the initial theta perturbed by 2^-24 relative (float
    rounding).  Each averages the statistics of `ComputeStats` (on the
    device) over the last `avg_min` minutes.  After the flow has become
    turbulent the runs are decorrelated, so the time averages differ by
    their finite-sample variability; the perturbed FP64 run measures that
    variability, and a mixed run without a precision bias differs from FP64
    by about as much.  Returns, for both comparisons, the largest
    difference of the mean wind profile (relative to the largest mean wind
    speed), the theta profile (relative to the absolute theta), the
    resolved-TKE profile and the dynamic-coefficient profiles Cs2 and
    Cs2/Pr (relative to their largest value), on the levels whose resolved
    TKE is at least 1% of its largest.
    """
    from . import cases, core
    base = {"neutral": 32, "convective": 128, "stable": 256}[case]
    dt = cases.CASES[case]["dt"] * base / n
    nsteps = int(round(hours * 3600.0 / dt))
    navg = int(round(avg_min * 60.0 / dt))
    every = max(1, navg // samples)
    res = {}
    for run in ("mixed", "fp64", "fp64_perturbed"):
        cfg, u, v, w, TH = cases.initial_fields(case, nx=n, ny=n, nz=n, dt=dt)
        if run == "fp64_perturbed":
            rng = np.random.default_rng(7)
            TH = TH + 2.0 ** -24 * (TH + cfg["T_0"]) * rng.standard_normal(TH.shape)
        ser = cases.surface_series(cfg, nsteps + 2)
        les = core.LES(core.Plan(cfg, "mixed" if run == "mixed" else "fp64"), u, v, w, TH,
                       sfc_series=ser)
        st = Statistics(les)
        for k in range(nsteps):
            les.step(1)
            if k >= nsteps - navg and (nsteps - 1 - k) % every == 0:
                st.sample()
        res[run] = st.averaged()
        les.plan.close()
    ref = res["fp64"]
    tke = ref["u2"] + ref["v2"] + ref["w2"]
    turb = tke >= 0.01 * np.max(tke)
    speed = float(np.max(np.hypot(ref["U"], ref["V"])))

    def compare(a):
        out = {"mean_wind_profile": float(max(np.max(np.abs(a["U"] - ref["U"])),
                                              np.max(np.abs(a["V"] - ref["V"]))) / speed),
               "mean_theta_profile": float(np.max(np.abs(a["TH"] - ref["TH"]) /
                                                  np.abs(ref["TH"]))),
               "tke_profile": float(np.max(np.abs(a["u2"] + a["v2"] + a["w2"] - tke)) /
                                    np.max(tke))}
        for k in ("Cs2_2", "Cs2PrRatio"):
            d = np.abs(a[k] - ref[k])[turb]
            out[k + "_profile"] = float(np.max(d) / max(np.max(np.abs(ref[k][turb])), 1e-300))
        return {k: float("%.3g" % v) for k, v in out.items()}

    return {"case": case, "grid": n, "dt_s": dt, "steps": nsteps, "simulated_h": hours,
            "average_over_min": avg_min, "samples_per_run": len(range(0, navg, every)),
            "turbulent_levels": int(turb.sum()),
            "mixed_vs_fp64": compare(res["mixed"]),
            "fp64_perturbed_vs_fp64": compare(res["fp64_perturbed"])}
