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
