"""Tile choices of the compile-time-length FFT kernels (alfa_plan_set_tiles,
alfa_plan_autotune): 256 or 512 threads per workgroup for the y R2C, x and
y C2R kernels of lengths 128 to 256 (lengths up to 96 always run their
128-thread kernels).  Every transform line goes through the same operations
at either size, so whole steps must agree bitwise across the choices."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from alfa_amd import _lib, cases, core  # noqa: E402

# compile-time lengths 64 x 96 (FGR=2: band-limited passes at the filter
# cut-offs; 128-thread kernels whatever the choice), 64 x 256 (the y
# kernels at the benchmark's length) and 128 x 128 at FGR=1 (dealiasing pads
# to 192 x 192)
SETUPS = {
    "convective64x96": ("convective", dict(nx=64, ny=96, nz=16)),
    "stable64x256": ("stable", dict(nx=64, ny=256, nz=16)),
    "neutral128": ("neutral", dict(nx=128, ny=128, nz=16)),
}


def _steps(key, prec, tiles, nsteps=3, autotune=None):
    """nsteps steps at the given tiles, or at the autotuner's choices
    (autotune = "transforms": timed on transforms of a zero field; "steps":
    timed on steps from a copy of the state, which must leave the state as
    it was)."""
    name, over = SETUPS[key]
    cfg, u, v, w, TH = cases.initial_fields(name, **over)
    ser = cases.surface_series(cfg, nsteps + 2)
    P = core.Plan(cfg, prec)
    chosen = P.autotune() if autotune == "transforms" else None
    if autotune is None:
        chosen = P.set_tiles(*tiles)
    les = core.LES(P, u, v, w, TH, sfc_series=ser)
    les.step(1)
    if autotune == "steps":
        before = les.fields()
        chosen = P.autotune(les)
        torch.cuda.synchronize()
        after = les.fields()
        for k in before:
            np.testing.assert_array_equal(before[k], after[k], err_msg=f"state {k} changed")
        assert les.iteration == 2
    les.step(nsteps - 1)
    torch.cuda.synchronize()
    f = les.fields()
    P.close()
    return chosen, f


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("key", list(SETUPS))
def test_tile_choices_give_identical_steps(key, prec):
    t256, a = _steps(key, prec, (256, 256, 256))
    t512, b = _steps(key, prec, (512, 512, 512))
    assert t256 == dict(y_r2c=256, x=256, y_c2r=256)
    assert t512 == dict(y_r2c=512, x=512, y_c2r=512)
    for k in ("u", "v", "w", "TH"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"{key} {prec} {k}")


@pytest.mark.parametrize("mode", ["transforms", "steps"])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_autotune_keeps_valid_choices_and_results(prec, mode):
    chosen, a = _steps("convective64x96", prec, None, autotune=mode)
    print(prec, mode, "autotuned tiles", chosen)
    assert set(chosen) == {"y_r2c", "x", "y_c2r"}
    assert all(v in (256, 512) for v in chosen.values())
    _, b = _steps("convective64x96", prec, (256, 512, 256))
    for k in ("u", "v", "w", "TH"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"{prec} {k}")


def test_set_tiles_rejects_other_sizes():
    cfg, *_ = cases.initial_fields("neutral")
    P = core.Plan(cfg, "mixed")
    before = P.tiles()
    with pytest.raises(RuntimeError, match="256 or 512"):
        P.set_tiles(x=384)
    assert P.tiles() == before
    assert P.set_tiles(y_c2r=256)["y_c2r"] == 256
    lib = _lib.load()
    assert lib.alfa_plan_set_tiles(P.handle, None) != 0
    assert lib.alfa_plan_get_tiles(P.handle, (C.c_int * 3)()) == 0
    P.close()


def _moist_run(autotune, nsteps=3):
    cfg, u, v, w, TH = cases.initial_fields("stable", nx=64, ny=64, nz=16, optMoistureSurfBC=2,
                                            q_inversion=0.0)
    cfg["optMoisture"] = 1
    rng = np.random.default_rng(5)
    Q = 0.008 + 1e-4 * rng.standard_normal(u.shape)
    qser = 0.0095 - 1e-4 * np.arange(nsteps + 2) * cfg["dt"] / 60.0
    ser = cases.surface_series(cfg, nsteps + 2)
    P = core.Plan(cfg, "fp64")
    les = core.LES(P, u, v, w, TH, sfc_series=ser, Q=Q, qsfc_series=qser)
    les.step(1)
    chosen = None
    if autotune:
        before = les.fields()
        chosen = P.autotune(les)
        torch.cuda.synchronize()
        after = les.fields()
        assert set(before) == {"u", "v", "w", "TH", "Q"}
        for k in before:
            np.testing.assert_array_equal(before[k], after[k], err_msg=f"state {k} changed")
    les.step(nsteps - 1)
    torch.cuda.synchronize()
    f = les.fields()
    P.close()
    return chosen, f


def test_autotune_on_a_moist_state():
    """With optMoisture the autotuner's scratch state holds Q and RHS_Q too:
    the caller's state, Q included, is left as it was and the run continues
    bitwise as at the default tiles."""
    chosen, a = _moist_run(True)
    print("moist autotuned tiles", chosen)
    _, b = _moist_run(False)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
