"""Per-operation parity of the SGS entry points for every base model.

alfa_sgs_stresses, alfa_scalar_fluxes, alfa_lasdd_wl and alfa_scalar_lasdd
are called through the C ABI (alfa_amd.ops) on seeded smooth O(1) inputs and
compared with the FP64 oracle, for the SM and WL base models, with and
without dealiasing.  Bars: FP64 <= 1e-12 for the stresses and fluxes (no
division), <= 1e-9 for the dynamic coefficients (ratios and a root solve);
mixed: <= 2e-5 for stresses and fluxes, coefficient profiles <= 1e-2.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import alfa_oracle as O  # noqa: E402
from alfa_amd import cases, core, ops  # noqa: E402

GRIDS = {
    "lasdd_sm_fgr2": ("convective", dict(nx=32, ny=32, nz=24, optSgs=1)),
    "lasdd_wl_fgr2": ("convective", dict(nx=32, ny=32, nz=24, optSgs=2)),
    "lad_wl_fgr1": ("neutral", dict(optSgs=4)),
    "lad_sm_fgr1": ("neutral", dict(optSgs=3)),
}
_plans = {}


def get_plan(key, prec):
    if (key, prec) not in _plans:
        name, over = GRIDS[key]
        cfg, *_ = cases.initial_fields(name, **over)
        _plans[(key, prec)] = (core.Plan(cfg, prec), O.Grid(cfg))
    P, g = _plans[(key, prec)]
    ops.use_plan(P)
    return P, g


def smooth(g, seed, kmax):
    rng = np.random.default_rng(seed)
    F = rng.standard_normal((g.nx, g.ny, g.nz))
    Fh = np.fft.rfft2(F, axes=(0, 1))
    kx = np.abs(np.fft.fftfreq(g.nx, 1 / g.nx))[:, None, None]
    ky = np.fft.rfftfreq(g.ny, 1 / g.ny)[None, :, None]
    Fh *= np.exp(-((kx / kmax) ** 2 + (ky / kmax) ** 2))
    return np.fft.irfft2(Fh, s=(g.nx, g.ny), axes=(0, 1))


def host(t):
    return t.detach().to(torch.float64).cpu().numpy()


@pytest.fixture(params=["fp64", "mixed"])
def prec(request):
    return request.param


@pytest.fixture(params=list(GRIDS))
def grid(request):
    return request.param


def test_sgs_stresses_and_fluxes(grid, prec):
    P, g = get_plan(grid, prec)
    tol = 1e-12 if prec == "fp64" else 2e-5
    grads = [smooth(g, 10 + i, g.nx / 6) for i in range(9)]
    dTH = [smooth(g, 20 + i, g.nx / 6) for i in range(3)]
    rng = np.random.default_rng(7)
    Cs2 = 0.02 + 0.01 * smooth(g, 30, g.nx / 8)
    CsPr = 0.02 + 0.01 * smooth(g, 31, g.nx / 8)
    u, v = 3.0 + smooth(g, 32, g.nx / 8), smooth(g, 33, g.nx / 8)
    M = np.hypot(u[:, :, 0] + g.Ugal, v[:, :, 0])
    pm, pm0 = 0.05 * rng.standard_normal((2, g.nx, g.ny))
    qsfc = 0.1 + 0.01 * rng.standard_normal((g.nx, g.ny))
    st = O.SGS_momentum_stresses(g, grads, Cs2, u, v, M, pm, pm0)
    D = P.asdev
    got = ops.SGSStresses(*[D(a) for a in grads], D(Cs2), D(u), D(v), D(M), D(pm), D(pm0))
    for name, t in zip(("txx", "tyy", "tzz", "txy", "txz", "tyz"), got):
        e = rel_err(host(t), st[name])
        assert e < tol, (name, e)
    qref = O.SGS_scalar_fluxes(g, st, CsPr, *dTH, qsfc)
    got = ops.ScalarFluxes(*[D(a) for a in grads], *[D(a) for a in dTH], D(CsPr), D(qsfc))
    for name, t, r in zip(("qx", "qy", "qz"), got, qref):
        e = rel_err(host(t), r)
        assert e < tol, (name, e)


def _momentum(g, model_wl):
    u, v, w = (smooth(g, 90 + i, g.nx / 8) for i in range(3))
    Sij = [0.5 * smooth(g, 95 + i, g.nx / 6) for i in range(6)]
    S = O._Smag(*Sij)
    if model_wl:
        return u, v, w, Sij, S, O.LASDD_WL(g, u, v, w, *Sij)
    return u, v, w, Sij, S, O.LASDD_SM(g, u, v, w, *Sij, S)


def test_lasdd_wl(grid, prec):
    P, g = get_plan(grid, prec)
    if g.optSgs not in (2, 4):
        pytest.skip("Wong-Lilly plans only")
    u, v, w, Sij, S, ref = _momentum(g, True)
    cs, a1, a2, beta = ops.LASDD_WL(*[P.asdev(a) for a in (u, v, w, *Sij)])
    if prec == "fp64":
        for got, key in ((beta, "beta1_1D"), (cs, "Cs2_3D"), (a1, "Cs2_1D_avg1"),
                         (a2, "Cs2_1D_avg2")):
            e = rel_err(host(got), ref[key])
            assert e < 1e-9, (key, e)
    else:
        e2 = rel_err(host(a2), ref["Cs2_1D_avg2"])
        print(f"{grid} mixed: Cwl profile rel err {e2:.2e}")
        assert e2 < 1e-2


def test_scalar_lasdd(grid, prec):
    P, g = get_plan(grid, prec)
    wl = g.optSgs in (2, 4)
    u, v, w, Sij, S, mom = _momentum(g, wl)
    TH = smooth(g, 40, g.nx / 8)
    dTH = [smooth(g, 41 + i, g.nx / 6) for i in range(3)]
    if wl:
        ref = O.ScalarLASDD_WL(g, mom, TH, *dTH)
    else:
        ref = O.ScalarLASDD_SM(g, mom, TH, *dTH, S)
    D = P.asdev
    beta_models = g.optSgs in (1, 2)
    hat2 = [D(a) for a in mom["u_hatd"]] if beta_models else [None] * 3
    cs, prof, beta = ops.ScalarLASDD(
        D(mom["u_"]), D(mom["v_"]), D(mom["w_"]), *[D(a) for a in mom["u_hat"]], *hat2, D(TH),
        *[D(a) for a in dTH], S=None if wl else D(S), S_hat=None if wl else D(mom["S_hat"]),
        S_hatd=D(mom["S_hatd"]) if g.optSgs == 1 else None)
    if prec == "fp64":
        for got, key in ((beta, "beta2_1D"), (cs, "Cs2PrRatio_3D"), (prof, "Cs2PrRatio_1D")):
            e = rel_err(host(got), ref[key])
            assert e < 1e-9, (key, e)
    else:
        e = rel_err(host(prof), ref["Cs2PrRatio_1D"])
        print(f"{grid} mixed: Cs2/Pr profile rel err {e:.2e}")
        assert e < 1e-2


@pytest.mark.parametrize("key", ["lasdd_sm_fgr2", "lasdd_wl_fgr2", "lad_sm_fgr1"])
def test_sgs_call_sites_match_the_oracle_step(key, prec):
    """DivStressDynamicSGS -> DivFluxDynamicSGS and the static pair on the
    intermediate fields of an oracle step (Main.py:333-443)."""
    P, g = get_plan(key, prec)
    name, over = GRIDS[key]
    cfg, u, v, w, TH = cases.initial_fields(name, **over)
    th = O.ThomasPressureInit(g)
    ser = cases.surface_series(cfg, 4)
    s = O.step(g, O.initial_state(g, u, v, w, TH), th, None if ser is None else ser[0])
    s2 = O.step(g, s, th, None if ser is None else ser[1], keep=True)
    dg = s2.diag
    D = P.asdev
    grads = [D(a) for a in dg["grads"]]
    uf, vf, wf, Tf = D(dg["u_f"]), D(dg["v_f"]), D(dg["w_f"]), D(dg["TH_f"])
    most = [D(m) for m in s2.MOST]
    dtx, dty, dtz, a1, a2, b1, cs, mom = ops.DivStressDynamicSGS(*grads, uf, vf, wf, D(dg["M_sfc_loc"]),
                                                               most)
    qz, divq, cp, prof, b2 = ops.DivFluxDynamicSGS(mom[10:], Tf, *[D(a) for a in dg["dTH"]],
                                                    D(dg["qz_sfc"]))
    if prec == "fp64":
        for got, ref in ((dtx, dg["divt"][0]), (dty, dg["divt"][1]), (dtz, dg["divt"][2]),
                         (divq, dg["divq"]), (qz, dg["qz"])):
            e = rel_err(host(got), ref)
            print(key, prec, "dynamic call sites rel err %.2e" % e)
            assert e < 1e-9
    else:
        # Two steps after a smooth start the dynamic coefficients are ratios of
        # rounding-level contractions in single precision (the step-level
        # mixed tests hold the flow itself); here the call sites must equal
        # the same pipeline composed from the finer entry points.
        S = ops.StrainsUVPnodes_NoDealias(*grads)
        if g.optSgs in (2, 4):
            cs_ref = ops.LASDD_WL(uf, vf, wf, *S[:6])[0]
        else:
            cs_ref = ops.LASDD(uf, vf, wf, *S)[0]
        tau = ops.SGSStresses(*grads, cs_ref, uf, vf, D(dg["M_sfc_loc"]), most[0], most[1])
        for a, b in zip((dtx, dty, dtz), ops.DivStress(*tau)):
            assert torch.equal(a, b)
        assert torch.equal(cs, cs_ref)
    # the static pair with the coefficient fields the dynamic call produced
    sx, sy, sz, smom = ops.DivStressStaticSGS(*grads, cs, uf, vf, D(dg["M_sfc_loc"]), most)
    sq, sdq = ops.DivFluxStaticSGS(smom[6:], cp, *[D(a) for a in dg["dTH"]], D(dg["qz_sfc"]))
    for a, b in ((sx, dtx), (sy, dty), (sz, dtz), (sdq, divq), (sq, qz)):
        assert torch.equal(a, b)
