"""Statistics of a run (alfa_stats_accumulate, alfa_amd.stats) against the
oracle's ComputeStats (src/utilities/Statistics.py:47-330) sampled after
every step of the same run.

Each key is compared relative to the largest magnitude of its physical
group (mean wind for U, V, W; velocity moments together; ...), because some
profiles are zero up to rounding (W, the vertical fluxes at rest).  FP64:
1e-11 for the flow statistics; the dynamic coefficient profiles at the
first steps are ratios of rounding-level contractions (see the step tests)
and are held to 1e-6.  Mixed: the means within 1e-4 of their scale.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import alfa_oracle as O  # noqa: E402
from alfa_amd import cases, core, stats  # noqa: E402

CASES = {
    "convective32": ("convective", dict(nx=32, ny=32, nz=32)),       # FGR=2 LASDD, hetero
    "neutral32": ("neutral", dict()),                                # FGR=1 dealiased, LAD
    "stable_static": ("stable", dict(nx=32, ny=32, nz=32, dynamicSGS_call_time=3)),
    "stable64x96": ("stable", dict(nx=64, ny=96, nz=16)),            # compile-time lengths
}
GROUPS = [("U", "V", "W"), ("TH",), ("dUdz", "dVdz"), ("dTHdz",),
          ("u2", "v2", "w2", "uv", "uw", "vw"), ("TH2",), ("uTH", "vTH", "wTH"),
          ("txy", "txz", "tyz"), ("qz",), ("Cs2_1",), ("Cs2_2",), ("Cs2PrRatio",),
          ("Beta1",), ("Beta2",)]
COEF = {"Cs2_1", "Cs2_2", "Cs2PrRatio", "Beta1", "Beta2"}


def run_both(key, prec, nsteps):
    name, over = CASES[key]
    cfg, u, v, w, TH = cases.initial_fields(name, **over)
    ser = cases.surface_series(cfg, nsteps + 2)
    g = O.Grid(cfg)
    s = O.initial_state(g, u, v, w, TH)
    th = O.ThomasPressureInit(g)
    les = core.LES(core.Plan(cfg, prec), u, v, w, TH, sfc_series=ser)
    st = stats.Statistics(les)
    acc = O.InitializeStats(g)
    prof = dict(Cs2_1D_avg1=None, Cs2_1D_avg2=None, Cs2PrRatio_1D=None, beta1_1D=None,
                beta2_1D=None)
    for _ in range(nsteps):
        s = O.step(g, s, th, None if ser is None else ser[s.iteration - 1])
        d = s.diag
        for k in prof:   # the profiles persist through static steps (Main.py)
            if k in d:
                prof[k] = d[k]
        acc = O.ComputeStats(g, acc, s.u, s.v, s.w, s.TH, d["dudz"], d["dvdz"], d["dTHdz"],
                             d["M_sfc_loc"], d["ustar"], d["qz_sfc_avg"], d["txy"], d["txz"],
                             d["tyz"], d["qz"], prof["Cs2_1D_avg1"], prof["Cs2_1D_avg2"],
                             prof["Cs2PrRatio_1D"], prof["beta1_1D"], prof["beta2_1D"])
        les.step(1)
        st.sample()
    return O.AverageStats(acc, nsteps), st.averaged()


def errors(ref, got):
    out = {}
    for grp in GROUPS:
        scale = max(float(np.max(np.abs(ref[k]))) for k in grp)
        for k in grp:
            out[k] = float(np.max(np.abs(np.asarray(got[k]) - ref[k]))) / max(scale, 1e-300)
    for k in stats.SCALARS:
        out[k] = abs(got[k] - ref[k]) / max(abs(ref[k]), 1e-300)
    return out


@pytest.mark.parametrize("key", list(CASES))
def test_stats_fp64(key):
    ref, got = run_both(key, "fp64", 4)
    e = errors(ref, got)
    print(key, "fp64 stats errors:", {k: "%.1e" % v for k, v in e.items()})
    for k, v in e.items():
        tol = 1e-6 if k in COEF else 1e-11
        assert v < tol, (k, v)
    assert set(got) == set(stats.KEY_ORDER)


def test_stats_mixed_means():
    ref, got = run_both("convective32", "mixed", 4)
    e = errors(ref, got)
    print("convective32 mixed stats errors:", {k: "%.1e" % v for k, v in e.items()})
    for k in ("U", "V", "W", "TH", "txy", "txz", "tyz", "qz", "M_sfc", "ustar", "qz_sfc"):
        assert e[k] < 1e-4, (k, e[k])


def test_stats_output_file(tmp_path):
    """Main.py:545-617 loop: samples every 2nd iteration, output every 4th:
    the file holds the reference's keys, averaged over the samples."""
    name, over = CASES["convective32"]
    cfg, u, v, w, TH = cases.initial_fields(name, **over)
    les = core.LES(core.Plan(cfg, "fp64"), u, v, w, TH)
    st = stats.Statistics(les)
    paths = stats.run(les, 8, SampleInterval=2, OutputInterval=4, outdir=str(tmp_path), stats=st)
    assert [p.rsplit("/", 1)[1] for p in paths] == ["ALFA_Statistics_Iteration_4.npz",
                                                    "ALFA_Statistics_Iteration_8.npz"]
    with np.load(paths[1]) as f:
        assert list(f.keys()) == list(stats.KEY_ORDER)
        assert f["U"].shape == (les.plan.nz,) and f["M_sfc"].shape == ()
        assert float(f["Ugal"]) == cfg.get("Ugal", 0.0) and not np.any(f["ZeRo1D"])
        assert not np.any(f["Q"]) and float(f["qm_sfc"]) == 0.0
    assert st.samples == 0   # reset after the output
