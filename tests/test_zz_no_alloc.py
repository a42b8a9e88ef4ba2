"""alfa_planar_mean_noalloc: PlanarMean (Utilities.py:41) with no plan and no
allocation, its partial sums in the library's fixed device scratch array.

Checked through the C ABI against alfa_planar_mean on a plan (the existing
entry point, same kernels and order: bitwise equal) and against the oracle
(FP64 within 1e-13 relative), on two calls in a row on the same GPU, the
second with a smaller input, so the second call runs on the scratch the
first one left behind.  Shapes beyond the scratch are refused with an
error before any device work.
"""
import ctypes as C

import numpy as np
import pytest

from alfa_amd import _lib

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import alfa_oracle as O  # noqa: E402
from alfa_amd import cases, core  # noqa: E402

PREC = {"fp64": (0, torch.float64), "mixed": (1, torch.float32)}


def _noalloc(prec, F):
    lib = _lib.load()
    code, _ = PREC[prec]
    nx, ny, nz = F.shape
    out = torch.full((nz,), float("nan"), dtype=torch.float64, device=F.device)
    rc = lib.alfa_planar_mean_noalloc(code, nx, ny, nz, C.c_void_p(F.data_ptr()),
                                      C.c_void_p(out.data_ptr()), core.current_stream())
    assert rc == 0, lib.alfa_last_error().decode()
    return out


def _with_plan(prec, F, name, over):
    cfg, *_ = cases.initial_fields(name, **over)
    P = core.Plan(cfg, prec)
    out = torch.full((F.shape[2],), float("nan"), dtype=torch.float64, device=F.device)
    P.call("alfa_planar_mean", C.c_void_p(F.data_ptr()), C.c_void_p(out.data_ptr()),
           core.current_stream())
    torch.cuda.synchronize()
    P.close()
    return out


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_noalloc_planar_mean_matches_the_plan_entry_point(prec):
    _, dt = PREC[prec]
    rng = np.random.default_rng(5)
    shapes = [("stable", dict(nx=256, ny=256, nz=64)), ("neutral", dict(nx=32, ny=32, nz=20))]
    for name, over in shapes:          # large first, then smaller: scratch reused
        F = rng.standard_normal((over["nx"], over["ny"], over["nz"])) + 3.0
        Fd = torch.tensor(F, dtype=dt, device="cuda")
        got = _noalloc(prec, Fd)
        torch.cuda.synchronize()
        ref = _with_plan(prec, Fd, name, over)
        g = got.cpu().numpy()
        assert np.array_equal(g, ref.cpu().numpy()), (prec, over)
        ora = O.PlanarMean(Fd.double().cpu().numpy())
        tol = 1e-13 if prec == "fp64" else 1e-6
        assert np.max(np.abs(g - ora)) <= tol * np.max(np.abs(ora)), (prec, over)


def test_noalloc_planar_mean_refuses_shapes_beyond_its_scratch():
    lib = _lib.load()
    F = torch.zeros(8, device="cuda")
    out = torch.zeros(8, dtype=torch.float64, device="cuda")
    # 1024 x 1024 columns x 512 levels of float: 16384 column blocks x 512 levels
    rc = lib.alfa_planar_mean_noalloc(1, 1024, 1024, 512, C.c_void_p(F.data_ptr()),
                                      C.c_void_p(out.data_ptr()), core.current_stream())
    assert rc != 0 and "scratch" in lib.alfa_last_error().decode()
