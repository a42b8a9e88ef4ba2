"""Moisture (optMoisture = 1) through alfa_step against the oracle.

Q is filtered, differentiated (MOST wall rule with the moisture flux),
advected, enters the buoyancy through the virtual temperature, gets the SGS
flux of the scalar model with Cs2PrRatio and advances with AB2_Q (Main.py:
193-529).  The three surface moisture BCs are covered: a constant flux, a
flux series, and a prescribed surface Q (homogeneous and heterogeneous).
FP64: every prognostic field, Q included, within 1e-11 relative of the
oracle after 5 steps; mixed: the mean Q profile within 1e-4 relative; the
moisture statistics against the oracle's ComputeStats.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import alfa_oracle as O  # noqa: E402
from alfa_amd import cases, core, stats  # noqa: E402

CASES = {
    # FGR=2 LASDD, constant moisture flux
    "convective_flux": ("convective", dict(nx=32, ny=32, nz=32, optMoistureSurfBC=0,
                                           MoistureFlux=1e-4, q_inversion=-2e-6)),
    # FGR=2 LAD, prescribed surface Q (homogeneous), series
    "stable_q": ("stable", dict(nx=32, ny=32, nz=32, optMoistureSurfBC=2, q_inversion=0.0)),
    # FGR=1 dealiased LAD, flux series
    "neutral_series": ("neutral", dict(optMoistureSurfBC=1)),
    # heterogeneous surface, prescribed Q, static SGS steps in between
    "convective_het_q": ("convective", dict(nx=32, ny=32, nz=32, optSurfFlux=1,
                                            optMoistureSurfBC=2, zMoisture=2.0,
                                            dynamicSGS_call_time=2)),
}


def setup(key, nsteps):
    name, over = CASES[key]
    cfg, u, v, w, TH = cases.initial_fields(name, **over)
    cfg["optMoisture"] = 1
    rng = np.random.default_rng(17)
    g = O.Grid(cfg)
    zu = (np.arange(g.nz) + 0.5) * g.dz * g.z_scale
    Q = 0.008 - 2e-6 * zu[None, None, :] + np.where(zu < 100.0, 1e-4, 0.0) * \
        rng.standard_normal(u.shape)
    t = np.arange(nsteps + 2) * cfg["dt"]
    if cfg["optMoistureSurfBC"] == 1:
        qser = 5e-5 + 1e-5 * np.sin(t / 7.0)
    elif cfg["optMoistureSurfBC"] == 2:
        qser = 0.0095 - 1e-4 * t / 60.0
    else:
        qser = None
    return cfg, g, u, v, w, TH, Q, qser


def run_both(key, prec, nsteps, with_stats=False):
    cfg, g, u, v, w, TH, Q, qser = setup(key, nsteps)
    ser = cases.surface_series(cfg, nsteps + 2)
    s = O.initial_state(g, u, v, w, TH, Q)
    th = O.ThomasPressureInit(g)
    les = core.LES(core.Plan(cfg, prec), u, v, w, TH, sfc_series=ser, Q=Q, qsfc_series=qser)
    st = stats.Statistics(les) if with_stats else None
    acc = O.InitializeStats(g)
    prof = {}
    errs = []
    for _ in range(nsteps):
        r = s.iteration - 1
        s = O.step(g, s, th, None if ser is None else ser[r],
                   qsfc_val=None if qser is None else qser[r])
        les.step(1)
        f = les.fields()
        errs.append({k: rel_err(f[k], getattr(s, k)) for k in ("u", "v", "w", "TH", "Q")})
        if with_stats:
            d = s.diag
            for k in ("Cs2_1D_avg1", "Cs2_1D_avg2", "Cs2PrRatio_1D", "beta1_1D", "beta2_1D"):
                if k in d:
                    prof[k] = d[k]
            acc = O.ComputeStats(g, acc, s.u, s.v, s.w, s.TH, d["dudz"], d["dvdz"], d["dTHdz"],
                                 d["M_sfc_loc"], d["ustar"], d["qz_sfc_avg"], d["txy"],
                                 d["txz"], d["tyz"], d["qz"], prof.get("Cs2_1D_avg1"),
                                 prof.get("Cs2_1D_avg2"), prof.get("Cs2PrRatio_1D"),
                                 prof.get("beta1_1D"), prof.get("beta2_1D"), Q=s.Q,
                                 dQdz=d["dQdz"], qHz=d["qHz"], qm_sfc_avg=d["qm_sfc_avg"])
            st.sample()
    return s, les, errs, acc, st


@pytest.mark.parametrize("key", list(CASES))
def test_moisture_step_fp64(key, tmp_path):
    s, les, errs, _, _ = run_both(key, "fp64", 5)
    # the 3-D snapshot of a moist run carries Q (Main.py:630-631)
    from alfa_amd import io
    with np.load(io.save_3d_fields(str(tmp_path), 5, les)) as f:
        assert sorted(f.keys()) == ["Q", "TH", "u", "v", "w"]
        np.testing.assert_array_equal(f["Q"], les.host("Q"))
    worst = {k: max(e[k] for e in errs) for k in errs[0]}
    print(key, "fp64 max rel err:", {k: "%.2e" % v for k, v in worst.items()})
    assert max(worst.values()) < 1e-11
    # the surface moisture flux mean of the last step (state scalars[4])
    qa = float(les.scalars[4].cpu())
    assert abs(qa - s.diag["qm_sfc_avg"]) <= 1e-11 * max(abs(s.diag["qm_sfc_avg"]), 1e-30)


def test_moisture_mixed_profile():
    s, les, errs, _, _ = run_both("convective_flux", "mixed", 10)
    f = les.fields()
    e = np.max(np.abs(f["Q"].mean(axis=(0, 1)) - s.Q.mean(axis=(0, 1)))) / np.max(np.abs(s.Q))
    print("mixed mean Q profile rel err %.2e" % e)
    assert e < 1e-4


def test_moisture_statistics_fp64():
    _, _, _, acc, st = run_both("stable_q", "fp64", 4, with_stats=True)
    ref = O.AverageStats(acc, 4)
    got = st.averaged()
    for k in ("Q", "dQdz", "Q2", "wQ", "qHz", "qm_sfc", "U", "TH", "u2", "wTH"):
        scale = max(float(np.max(np.abs(ref[k]))), 1e-300)
        e = float(np.max(np.abs(np.asarray(got[k]) - ref[k]))) / scale
        print("moisture stats", k, "%.1e" % e)
        assert e < 1e-10, (k, e)
