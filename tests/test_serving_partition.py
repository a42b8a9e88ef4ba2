"""Host logic of grouped serving (no GPU): which admitted requests may share
launches.  Members of one part must take the same kind of step at every
iteration from now on: same configuration shape, and, when the dynamic SGS
procedure runs every k-th step, the same iteration modulo k."""
from types import SimpleNamespace as NS

import pytest

from alfa_amd.ensemble import _sgs_phase, shape_key


def member(n=32, prec="mixed", it=2, optSgs=3, ct=1, **cfg):
    c = dict(FGR=2, optSgs=optSgs, dynamicSGS_call_time=ct, optSurfBC=2, optSurfFlux=0,
             optMoisture=0, optMoistureSurfBC=0, optBuoyancy=1, optDamping=1,
             optPressureSolver=1, istep=1)
    c.update(cfg)
    return NS(cfg=c, les=NS(iteration=it), spec=NS(n=n), plan=NS(precision=prec))


@pytest.mark.parametrize("it", [2, 3, 10])
def test_every_step_dynamic_has_one_phase(it):
    assert _sgs_phase(member(it=it, ct=1)) == 0
    assert _sgs_phase(member(it=it, ct=3, optSgs=0)) == 0   # static model: no procedure


def test_phase_follows_the_dynamic_period():
    a, b, c = member(it=5, ct=2), member(it=7, ct=2), member(it=6, ct=2)
    assert _sgs_phase(a) == _sgs_phase(b) != _sgs_phase(c)
    # members in one phase take the same kind of step at every later iteration
    for s in range(6):
        assert ((a.les.iteration + s) % 2 == 0) == ((b.les.iteration + s) % 2 == 0)


def test_shape_key_separates_grids_precisions_and_options():
    base = shape_key(member())
    assert shape_key(member(it=9)) == base               # iteration is not part of the shape
    assert shape_key(member(n=48)) != base
    assert shape_key(member(prec="fp64")) != base
    assert shape_key(member(FGR=1)) != base
    assert shape_key(member(optSurfBC=0)) != base
    assert shape_key(member(ct=2)) != base
