"""The fused test-filter kernels against the separate y C2R and pointwise
kernels they replace, on the same plan configuration: for LAD (optSgs = 3)
the level-1 inverse transforms + the LM/MM and T_up/T_dn contractions in one
kernel each; for LASDD (optSgs = 1) also the plane sums of both levels (one
kernel per level and procedure) that give beta before the contractions.

Both paths transform each filtered field by the same passes (but for length
64, which the fused kernels factor as 8 x 4 x 2 instead of 8 x 8, and 80 in
the scalar kernels, 8 x 5 x 2 instead of 8 x 2 x 5), so their float field
values agree bitwise; the contractions differ only in how their double
arithmetic is ordered (and MM, T_dn are accumulated in float), so LM, MM,
T_up and T_dn agree to a few float roundings and the runs stay within float
rounding of each other.  The grids cover every transform length the fused
kernels serve (48, 64, 80, 96, 128, 192, 256), the band-limited first pass
(FGR = 2) and the full one (FGR = 1), and the grids where the fused form
does not apply (a length without a compile-time plan, nz not a multiple of
4), where both settings must give bitwise the same run.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import alfa_oracle as O  # noqa: E402
from alfa_amd import cases, core  # noqa: E402

FUSED_GRIDS = {
    "stable48": ("stable", dict(nx=48, ny=48, nz=48)),                 # band first pass, radix 3 last
    "neutral48": ("neutral", dict(nx=48, ny=48, nz=32)),               # FGR = 1: full first pass
    "stable64": ("stable", dict(nx=48, ny=64, nz=16)),                 # 8 x 4 x 2 in the fused kernels
    "stable80": ("stable", dict(nx=48, ny=80, nz=16)),                 # radix 5 (LAD), 2 (scalar, 320 threads)
    "stable96": ("stable", dict(nx=96, ny=96, nz=32)),
    "stable128": ("stable", dict(nx=64, ny=128, nz=16)),               # radix-2 last pass
    "stable192": ("stable", dict(nx=32, ny=192, nz=16)),
    "stable256": ("stable", dict(nx=64, ny=256, nz=24)),               # the benchmark's length
    "neutral256": ("neutral", dict(nx=32, ny=256, nz=12)),             # FGR = 1 at 256
}
LASDD_GRIDS = {
    "convective48": ("convective", dict(nx=48, ny=48, nz=48)),         # LASDD-SM, band first pass
    "convective64": ("convective", dict(nx=64, ny=64, nz=32)),
    "convective80": ("convective", dict(nx=48, ny=80, nz=16)),
    "convective96": ("convective", dict(nx=96, ny=96, nz=32)),
    "convective128": ("convective", dict(nx=64, ny=128, nz=16)),
    "convective192": ("convective", dict(nx=32, ny=192, nz=16)),
    "convective256": ("convective", dict(nx=64, ny=256, nz=24)),
    "neutral48_lasdd": ("neutral", dict(nx=48, ny=48, nz=32, optSgs=1)),   # FGR = 1
}
UNFUSED_GRIDS = {
    "stable40": ("stable", dict(nx=40, ny=40, nz=40)),                 # no compile-time plan
    "stable48_nz18": ("stable", dict(nx=48, ny=48, nz=18)),            # nz % 4 != 0
}


def pair(name, over, nsteps, stirred=False):
    """Two mixed runs of one case, fused and separate.  stirred: seeded
    fluctuations at every level (u, v 0.5 m/s, w 0.3 m/s, TH 0.2 K), so that
    the Leonard and model terms of both procedures are resolved signals
    rather than the rounding noise of the nearly uniform initial layers."""
    cfg, u, v, w, TH = cases.initial_fields(name, **over)
    if stirred:
        rng = np.random.default_rng(7)
        u = u + 0.5 * rng.standard_normal(u.shape)
        v = v + 0.5 * rng.standard_normal(v.shape)
        w = w.copy()
        w[:, :, 1:-1] += 0.3 * rng.standard_normal(w[:, :, 1:-1].shape)
        TH = TH + 0.2 * rng.standard_normal(TH.shape)
    ser = cases.surface_series(cfg, nsteps + 2)
    runs = []
    for fused in (True, False):
        P = core.Plan(cfg, "mixed")
        assert P.set_fusion(fused) == fused
        runs.append(core.LES(P, u, v, w, TH, sfc_series=ser))
    return runs


@pytest.mark.parametrize("key", list(FUSED_GRIDS))
def test_fused_contraction_matches_separate_kernels(key):
    """From the case's own initial layers: the prognostic fields of the two
    runs within float rounding of each other after 4 steps, Cs2 within 1e-4
    of its largest value.  (Cs2/Pr is not compared here: while the layers
    are laminar its T_up and T_dn are rounding noise, ~1e-17, and so is their
    clipped ratio, in either run.)"""
    name, over = FUSED_GRIDS[key]
    a, b = pair(name, over, 4)
    a.step(4)
    b.step(4)
    torch.cuda.synchronize()
    fa, fb = a.fields(), b.fields()
    e = {k: rel_err(fa[k], fb[k]) for k in ("u", "v", "w", "TH")}
    e["Cs2_3D"] = rel_err(a.host("Cs2_3D"), b.host("Cs2_3D"))
    print(key, "4 steps fused vs separate:", {k: "%.1e" % v for k, v in e.items()})
    for k in ("u", "v", "w", "TH"):
        assert e[k] < 1e-5, (k, e[k])
    assert e["Cs2_3D"] < 1e-4, e["Cs2_3D"]
    for r in (a, b):
        r.plan.close()


@pytest.mark.parametrize("key", list(FUSED_GRIDS))
def test_fused_coefficients_match_on_stirred_fields(key):
    """Both coefficient fields from stirred initial fields: equal to a few
    float roundings after the first step (the same filtered fields, the
    contractions ordered differently) and within 1e-4 of their largest
    values after 4 steps."""
    name, over = FUSED_GRIDS[key]
    a, b = pair(name, over, 4, stirred=True)
    a.step(1)
    b.step(1)
    torch.cuda.synchronize()
    e1 = {k: rel_err(a.host(k), b.host(k)) for k in ("Cs2_3D", "Cs2PrRatio_3D")}
    print(key, "stirred, step 1 fused vs separate:", {k: "%.1e" % v for k, v in e1.items()},
          "coefficient means %.2e %.2e" % (a.host("Cs2_3D").mean(), a.host("Cs2PrRatio_3D").mean()))
    for k, v in e1.items():
        assert v < 1e-5, (k, v)
    a.step(3)
    b.step(3)
    torch.cuda.synchronize()
    fa, fb = a.fields(), b.fields()
    e = {k: rel_err(fa[k], fb[k]) for k in ("u", "v", "w", "TH")}
    e.update({k: rel_err(a.host(k), b.host(k)) for k in ("Cs2_3D", "Cs2PrRatio_3D")})
    print(key, "stirred, 4 steps fused vs separate:", {k: "%.1e" % v for k, v in e.items()})
    for k in ("u", "v", "w", "TH"):
        assert e[k] < 1e-5, (k, e[k])
    for k in ("Cs2_3D", "Cs2PrRatio_3D"):
        assert e[k] < 1e-4, (k, e[k])
    for r in (a, b):
        r.plan.close()


@pytest.mark.parametrize("key", list(LASDD_GRIDS))
def test_fused_lasdd_matches_separate_kernels(key):
    """LASDD (beta from the plane sums of both test-filter levels): from
    stirred fields, both betas within 1e-4 of the separate kernels' after the
    first step (the sums differ only in the order of their double additions;
    beta is a root the reference's Laguerre iteration finds to 1e-6) and the
    coefficient fields within 1e-4 of their largest values; after 4 steps the
    prognostic fields within float rounding of each other."""
    name, over = LASDD_GRIDS[key]
    a, b = pair(name, over, 4, stirred=True)
    a.step(1)
    b.step(1)
    torch.cuda.synchronize()
    db = {k: float(np.max(np.abs(a.prof[k].cpu().numpy() - b.prof[k].cpu().numpy())))
          for k in ("beta1_1D", "beta2_1D")}
    e1 = {k: rel_err(a.host(k), b.host(k)) for k in ("Cs2_3D", "Cs2PrRatio_3D")}
    print(key, "LASDD step 1 fused vs separate: beta max diff",
          {k: "%.1e" % v for k, v in db.items()}, "coefficients", {k: "%.1e" % v for k, v in e1.items()},
          "beta ranges %.3f..%.3f" % (float(a.prof["beta1_1D"].min()), float(a.prof["beta1_1D"].max())))
    for k, v in db.items():
        assert v < 1e-4, (k, v)
    for k, v in e1.items():
        assert v < 1e-4, (k, v)
    a.step(3)
    b.step(3)
    torch.cuda.synchronize()
    fa, fb = a.fields(), b.fields()
    e = {k: rel_err(fa[k], fb[k]) for k in ("u", "v", "w", "TH")}
    print(key, "LASDD 4 steps fused vs separate:", {k: "%.1e" % v for k, v in e.items()})
    for k in ("u", "v", "w", "TH"):
        assert e[k] < 1e-5, (k, e[k])
    for r in (a, b):
        r.plan.close()


def test_fused_lasdd_mixed_profiles_vs_oracle():
    """Mixed precision with the fused LASDD kernels (convective 48^3): mean
    wind and potential temperature profiles within 1e-4 relative of the FP64
    oracle after 20 steps."""
    name, over = LASDD_GRIDS["convective48"]
    cfg, u, v, w, TH = cases.initial_fields(name, **over)
    nsteps = 20
    ser = cases.surface_series(cfg, nsteps + 2)
    g = O.Grid(cfg)
    s = O.initial_state(g, u, v, w, TH)
    th = O.ThomasPressureInit(g)
    les = core.LES(core.Plan(cfg, "mixed"), u, v, w, TH, sfc_series=ser)
    for _ in range(nsteps):
        s = O.step(g, s, th, None if ser is None else ser[s.iteration - 1])
    les.step(nsteps)
    f = les.fields()
    ug = {k: f[k].mean(axis=(0, 1)) for k in ("u", "v", "TH")}
    uo = {k: getattr(s, k).mean(axis=(0, 1)) for k in ("u", "v", "TH")}
    speed = np.max(np.hypot(uo["u"] + cfg["Ugal"], uo["v"]))
    for k in ("u", "v"):
        e = np.max(np.abs(ug[k] - uo[k])) / speed
        print("convective48 LASDD", k, "fused mixed profile err / max mean wind speed %.2e" % e)
        assert e < 1e-4
    e = np.max(np.abs(ug["TH"] - uo["TH"])) / np.max(np.abs(uo["TH"] + cfg["T_0"]))
    print("convective48 LASDD TH fused mixed profile rel err %.2e" % e)
    assert e < 1e-4
    db = np.max(np.abs(les.prof["beta1_1D"].cpu().numpy() - s.diag["beta1_1D"]))
    print("convective48 LASDD beta1 after 20 steps: max diff to the oracle %.2e" % db)


@pytest.mark.parametrize("key", list(UNFUSED_GRIDS))
def test_fusion_setting_is_inert_where_it_does_not_apply(key):
    name, over = UNFUSED_GRIDS[key]
    a, b = pair(name, over, 3)
    a.step(2)
    b.step(2)
    torch.cuda.synchronize()
    fa, fb = a.fields(), b.fields()
    for k in ("u", "v", "w", "TH"):
        assert np.array_equal(fa[k], fb[k]), k
    for k in ("Cs2_3D", "Cs2PrRatio_3D"):
        assert np.array_equal(a.host(k), b.host(k)), k


def test_fp64_plans_ignore_fusion():
    """FP64 plans run the separate kernels whatever the setting: bitwise the
    same run either way."""
    cfg, u, v, w, TH = cases.initial_fields("stable", nx=48, ny=48, nz=16)
    ser = cases.surface_series(cfg, 4)
    runs = []
    for fused in (True, False):
        P = core.Plan(cfg, "fp64")
        P.set_fusion(fused)
        runs.append(core.LES(P, u, v, w, TH, sfc_series=ser))
    for r in runs:
        r.step(2)
    torch.cuda.synchronize()
    fa, fb = runs[0].fields(), runs[1].fields()
    for k in ("u", "v", "w", "TH"):
        assert np.array_equal(fa[k], fb[k]), k


@pytest.mark.parametrize("key", ["stable48", "neutral48", "stable96"])
def test_fused_mixed_profiles_vs_oracle(key):
    """Mixed precision with the fused kernels: mean wind and potential
    temperature profiles within 1e-4 relative of the FP64 oracle after 20
    steps (the bar of the other mixed step tests)."""
    name, over = FUSED_GRIDS[key]
    cfg, u, v, w, TH = cases.initial_fields(name, **over)
    nsteps = 20
    ser = cases.surface_series(cfg, nsteps + 2)
    g = O.Grid(cfg)
    s = O.initial_state(g, u, v, w, TH)
    th = O.ThomasPressureInit(g)
    P = core.Plan(cfg, "mixed")
    assert P.fusion()
    les = core.LES(P, u, v, w, TH, sfc_series=ser)
    for _ in range(nsteps):
        s = O.step(g, s, th, None if ser is None else ser[s.iteration - 1])
    les.step(nsteps)
    f = les.fields()
    ug = {k: f[k].mean(axis=(0, 1)) for k in ("u", "v", "TH")}
    uo = {k: getattr(s, k).mean(axis=(0, 1)) for k in ("u", "v", "TH")}
    speed = np.max(np.hypot(uo["u"] + cfg["Ugal"], uo["v"]))
    for k in ("u", "v"):
        e = np.max(np.abs(ug[k] - uo[k])) / speed
        print(key, k, "fused mixed profile err / max mean wind speed %.2e" % e)
        assert e < 1e-4
    e = np.max(np.abs(ug["TH"] - uo["TH"])) / np.max(np.abs(uo["TH"] + cfg["T_0"]))
    print(key, "TH fused mixed profile rel err %.2e" % e)
    assert e < 1e-4
