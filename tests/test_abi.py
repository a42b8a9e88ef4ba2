"""The C-ABI library loads and exports every entry point include/alfa.h
declares; host-side argument checks run without a GPU."""
import ctypes as C

import pytest

from alfa_amd import _lib


def test_library_exports_every_header_symbol():
    lib = _lib.load()
    syms = _lib.header_symbols()
    assert len(syms) >= 38
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_every_binding_is_declared():
    syms = set(_lib.header_symbols())
    assert set(_lib._SIGS) <= syms


def test_struct_layout_matches_header():
    # alfa_config: 13 int32 + 5 reserved, 21 doubles + 5 reserved; alfa_state:
    # 16 fields + 6 diagnostics + Q, RHS_Q, Q_sfc_val
    assert C.sizeof(_lib.AlfaConfig) == 18 * 4 + 26 * 8
    assert C.sizeof(_lib.AlfaState) == 25 * 8
    # and the library's own sizeof of each struct
    lib = _lib.load()
    sz = (C.c_size_t * 3)()
    assert lib.alfa_struct_sizes(C.cast(sz, C.c_void_p)) == 0
    assert list(sz) == [C.sizeof(_lib.AlfaConfig), C.sizeof(_lib.AlfaState),
                        C.sizeof(_lib.AlfaDerived)]


@pytest.mark.parametrize("over,frag", [
    (dict(nx=31), "even"),
    (dict(FGR=3), "FGR"),
    (dict(optSgs=5), "optSgs"),
    (dict(nx=14, ny=14), "factor"),
])
def test_plan_create_rejects_bad_configs(over, frag):
    """Validation happens before any device call, so it runs here."""
    from alfa_amd import cases
    cfg = dict(cases.NEUTRAL)
    cfg.update(over)
    lib = _lib.load()
    p = C.c_void_p()
    cs = _lib.config_struct(cfg)
    rc = lib.alfa_plan_create(C.byref(cs), 0, C.byref(p))
    assert rc != 0
    assert frag in lib.alfa_last_error().decode()
