"""Derivxy's wavenumber argument (CPU only: no kernel is called).

The reference passes its broadcast tables kx2 / ky2 (Preprocess.py:96-100;
Derivatives.py:161-182).  alfa_amd.ops accepts them, as numpy arrays or
torch tensors, and finds the axis from the index along which the table
varies; a table whose values are not the plan's wavenumbers is refused.
"""
import types

import numpy as np
import pytest

import alfa_oracle as O
from alfa_amd import cases, ops


@pytest.fixture
def plan_dims():
    cfg, *_ = cases.initial_fields("convective", nx=12, ny=10, nz=4)
    g = O.Grid(cfg)
    old = ops._cur
    ops._cur = types.SimpleNamespace(nx=g.nx, ny=g.ny, nz=g.nz)
    yield g
    ops._cur = old


def test_reference_tables_give_their_axis(plan_dims):
    g = plan_dims
    kx2, ky2 = np.ascontiguousarray(g.kx2), np.ascontiguousarray(g.ky2)
    assert ops._axis(kx2) == 0 and ops._axis(ky2) == 1
    assert ops._axis(kx2[:, :, 0]) == 0 and ops._axis(ky2[:, :, 0]) == 1
    torch = pytest.importorskip("torch")
    assert ops._axis(torch.from_numpy(kx2)) == 0 and ops._axis(torch.from_numpy(ky2)) == 1
    assert ops._axis(ops.WaveTable(1)) == 1 and ops._axis("x") == 0


def test_foreign_tables_are_refused(plan_dims):
    g = plan_dims
    with pytest.raises(ValueError):
        ops._axis(2.0 * g.kx2)                     # not the plan's wavenumbers
    kx = np.fft.fftfreq(g.nx, 1 / g.nx)            # Nyquist left in
    with pytest.raises(ValueError):
        ops._axis(np.broadcast_to(kx[:, None, None], g.kx2.shape))
    with pytest.raises(ValueError):
        ops._axis(np.zeros((g.nx + 2, g.ny // 2 + 1, g.nz)))
    with pytest.raises(TypeError):
        ops._axis(3)
