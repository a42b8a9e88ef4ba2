"""Plan and run objects over libalfa (device memory and streams from torch).

``Plan`` owns one ``alfa_plan`` (twiddles, pressure factors, profiles and the
device workspace of one grid/configuration).  ``LES`` holds the carried state
of one run (src/Main.py:66-111) as torch device tensors and advances it with
``alfa_step``.  torch is used only for device allocations and streams; all
arithmetic is in libalfa's HIP kernels, and there is no CPU path: without a
GPU or without the built library every call raises.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib


def _torch():
    import torch  # imported first so that libalfa binds to torch's HIP runtime
    return torch


def current_stream():
    return C.c_void_p(_torch().cuda.current_stream().cuda_stream)


class Plan:
    """One alfa_plan for a configuration (reference Config.py names)."""

    def __init__(self, cfg: dict, precision: str = "fp64", device: str = "cuda:0"):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("alfa_amd needs a GPU (torch.cuda.is_available() is False)")
        self.lib = _lib.load()
        self.cfg = dict(cfg)
        self.precision = precision
        prec = {"fp64": _lib.ALFA_FP64, "mixed": _lib.ALFA_MIXED}[precision]
        self.device = torch.device(device)
        self.dtype = torch.float64 if precision == "fp64" else torch.float32
        self.cdtype = torch.complex128 if precision == "fp64" else torch.complex64
        with torch.cuda.device(self.device):
            cs = _lib.config_struct(self.cfg)
            p = C.c_void_p()
            _lib.check(self.lib.alfa_plan_create(C.byref(cs), prec, C.byref(p)), "alfa_plan_create")
        self._p = p
        d = _lib.AlfaDerived()
        _lib.check(self.lib.alfa_plan_derived(p, C.byref(d)), "alfa_plan_derived")
        self.derived = {k: getattr(d, k) for k, _ in d._fields_}
        self.nx, self.ny, self.nz = int(cfg["nx"]), int(cfg["ny"]), int(cfg["nz"])
        self.nyc = self.ny // 2 + 1
        self.mx, self.my = self.derived["mx"], self.derived["my"]

    @property
    def handle(self):
        return self._p

    # -- tile choices of the FFT kernels -----------------------------------
    TILE_KERNELS = ("y_r2c", "x", "y_c2r")

    def tiles(self):
        """Threads per workgroup of the compile-time-length y R2C, x and y
        C2R kernels (256 or 512 each)."""
        t = (C.c_int * 3)()
        _lib.check(self.lib.alfa_plan_get_tiles(self._p, t), "alfa_plan_get_tiles")
        return dict(zip(self.TILE_KERNELS, list(t)))

    def set_tiles(self, y_r2c=0, x=0, y_c2r=0):
        """Set tile choices (0 keeps one); results do not depend on them."""
        t = (C.c_int * 3)(int(y_r2c), int(x), int(y_c2r))
        _lib.check(self.lib.alfa_plan_set_tiles(self._p, t), "alfa_plan_set_tiles")
        return self.tiles()

    def fusion(self):
        """Whether the LAD dynamic procedure runs its fused transform +
        contraction kernels (float plans, optSgs = 3)."""
        v = C.c_int()
        _lib.check(self.lib.alfa_plan_get_fusion(self._p, C.byref(v)), "alfa_plan_get_fusion")
        return bool(v.value)

    def set_fusion(self, on):
        """Select the fused (True) or separate (False) LAD kernels; the
        results agree to float rounding."""
        _lib.check(self.lib.alfa_plan_set_fusion(self._p, 1 if on else 0), "alfa_plan_set_fusion")
        return self.fusion()

    def autotune(self, les=None, stream=None):
        """Time each FFT kernel at both tile choices on this device and keep
        the faster (alfa_plan_autotune); returns the kept choices.  With an
        LES of this plan, whole steps from its current state are timed on a
        scratch copy (the run itself does not advance); without one, a
        forward transform and three inverses of a zero field."""
        torch = _torch()
        with torch.cuda.device(self.device):
            s = stream if stream is not None else torch.cuda.current_stream(self.device)
            t = (C.c_int * 3)()
            st, it, sv = None, 1, 0.0
            if les is not None:
                if les.qsfc is not None:   # as LES.step sets it for this iteration
                    les.state.Q_sfc_val = les.qsfc[les.iteration - 1].data_ptr()
                st, it = C.byref(les.state), int(les.iteration)
                if les.sfc_series is not None:
                    sv = float(les.sfc_series[les.iteration - 1])
            _lib.check(self.lib.alfa_plan_autotune(self._p, st, it, C.c_double(sv),
                                                   C.c_void_p(s.cuda_stream), t),
                       "alfa_plan_autotune")
        return dict(zip(self.TILE_KERNELS, list(t)))

    def close(self):
        if getattr(self, "_p", None) is not None and self._p.value:
            self.lib.alfa_plan_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- allocation helpers -------------------------------------------------
    def empty(self, shape, complex_=False, dtype=None):
        torch = _torch()
        dt = dtype or (self.cdtype if complex_ else self.dtype)
        return torch.empty(shape, dtype=dt, device=self.device)

    def zeros(self, shape, complex_=False, dtype=None):
        torch = _torch()
        dt = dtype or (self.cdtype if complex_ else self.dtype)
        return torch.zeros(shape, dtype=dt, device=self.device)

    def field(self):
        return self.zeros((self.nx, self.ny, self.nz))

    def spectrum(self):
        return self.zeros((self.nx, self.nyc, self.nz), complex_=True)

    def plane(self):
        return self.zeros((self.nx, self.ny))

    def asdev(self, a, complex_=False):
        """numpy or torch array -> contiguous device tensor of the plan dtype."""
        torch = _torch()
        dt = self.cdtype if complex_ else self.dtype
        if isinstance(a, np.ndarray):
            a = torch.from_numpy(np.ascontiguousarray(a))
        return a.to(device=self.device, dtype=dt).contiguous()

    def ptr(self, t, complex_=None):
        """Device pointer of a tensor after checking device, dtype and layout."""
        torch = _torch()
        if t is None:
            return None
        if not isinstance(t, torch.Tensor) or t.device != self.device:
            raise TypeError("expected a tensor on the plan's device")
        if not t.is_contiguous():
            raise ValueError("tensor must be contiguous")
        ok = (self.dtype, self.cdtype, torch.float64)
        if t.dtype not in ok:
            raise TypeError(f"tensor dtype {t.dtype} does not match plan precision {self.precision}")
        return C.c_void_p(t.data_ptr())

    def profile_begin(self, name_filter: Optional[str] = None, max_launches: int = 8192):
        """Bracket matching launches with HIP events (see include/alfa.h)."""
        f = None if not name_filter else name_filter.encode()
        _lib.check(self.lib.alfa_profile_begin(self._p, f, int(max_launches)), "alfa_profile_begin")

    def profile_end(self) -> list:
        """-> [dict(name, launches, total_ms, bytes, ref_bytes)] per kernel
        name: bytes are the launches' distinct inputs + outputs, ref_bytes
        the same work in the reference's data flow (include/alfa.h)."""
        arr = (_lib.AlfaKernelStat * 256)()
        n = C.c_int(0)
        _lib.check(self.lib.alfa_profile_end(self._p, arr, 256, C.byref(n)), "alfa_profile_end")
        return [dict(name=arr[i].name.decode(), launches=int(arr[i].launches),
                     total_ms=float(arr[i].total_ms), bytes=float(arr[i].bytes),
                     ref_bytes=float(arr[i].ref_bytes))
                for i in range(min(n.value, 256))]

    def call(self, name, *args):
        f = getattr(self.lib, name)
        _lib.check(f(self._p, *args), name)


class LES:
    """Device-resident state of one run and its time loop (src/Main.py:156-540)."""

    def __init__(self, plan: Plan, u, v, w, TH, sfc_series=None, Q=None, qsfc_series=None):
        """Q, qsfc_series: the initial specific humidity and the
        MoistureSurfaceBC_series (nondimensional, nsteps+1 values) of a
        moisture plan (optMoisture >= 1; Initialization.py:330-386)."""
        self.plan = plan
        P = plan
        self.moist = int(P.cfg.get("optMoisture", 0)) >= 1
        self.t = {n: P.field() for n in _lib.STATE_FIELDS[:10]}
        for n in _lib.STATE_FIELDS[10:]:
            self.t[n] = P.plane()
        self.prof = {n: P.zeros((P.nz,), dtype=_torch().float64) for n in _lib.STATE_DIAG[:5]}
        self.scalars = P.zeros((8,), dtype=_torch().float64)
        for n, a in zip(("u", "v", "w", "TH"), (u, v, w, TH)):
            self.t[n].copy_(P.asdev(a))
        self.qsfc = None
        if self.moist:
            if Q is None:
                raise ValueError("a moisture plan (optMoisture >= 1) needs the initial Q")
            self.t["Q"] = P.asdev(Q).clone()
            self.t["RHS_Q"] = P.field()
            if int(P.cfg.get("optMoistureSurfBC", 0)) >= 1:
                if qsfc_series is None:
                    raise ValueError("optMoistureSurfBC >= 1 needs qsfc_series")
                self.qsfc = _torch().tensor(np.asarray(qsfc_series, dtype=np.float64),
                                            device=P.device)
        self.state = _lib.AlfaState()
        for n in _lib.STATE_FIELDS:
            setattr(self.state, n, self.t[n].data_ptr())
        if self.moist:
            self.state.Q = self.t["Q"].data_ptr()
            self.state.RHS_Q = self.t["RHS_Q"].data_ptr()
        for n in _lib.STATE_DIAG[:5]:
            setattr(self.state, n, self.prof[n].data_ptr())
        self.state.scalars = self.scalars.data_ptr()
        P.call("alfa_state_init", C.byref(self.state), current_stream())
        self.iteration = int(P.cfg.get("istep", 1))
        self.sfc_series = None if sfc_series is None else np.asarray(sfc_series, dtype=float)
        self.CFLmax = 0.0
        self.forcing = {}

    def set_forcing_series(self, Ug=None, Vg=None, Uadv=None, Vadv=None, THadv=None, Qadv=None):
        """Time-varying forcing of Main.py:167-188: (nsteps+1, nz)
        nondimensional series (GeoWind_U/V, AdvForcing_U/V/TH/Q); row
        iteration-1 is applied before each step.  Kept on the device."""
        torch = _torch()
        P = self.plan
        for k, a in (("Ug", Ug), ("Vg", Vg), ("Uadv", Uadv), ("Vadv", Vadv), ("THadv", THadv),
                     ("Qadv", Qadv)):
            if a is None:
                continue
            a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
            if a.ndim != 2 or a.shape[1] != P.nz:
                raise ValueError(f"{k} series must be (nsteps+1, nz={P.nz}), got {a.shape}")
            self.forcing[k] = torch.from_numpy(a).to(P.device)
        if any(k in self.forcing for k in ("Uadv", "Vadv", "THadv", "Qadv")):
            # optAdvection: the tendencies together (missing ones zero)
            n = next(iter(self.forcing.values())).shape[0]
            for k in ("Uadv", "Vadv", "THadv", "Qadv"):
                self.forcing.setdefault(k, torch.zeros((n, P.nz), dtype=torch.float64,
                                                       device=P.device))

    def _apply_forcing(self, st):
        if not self.forcing:
            return
        row = self.iteration - 1
        ptr = []
        for k in ("Ug", "Vg", "Uadv", "Vadv", "THadv", "Qadv"):
            t = self.forcing.get(k)
            if t is not None and row >= t.shape[0]:
                raise IndexError(f"{k} series has no row for iteration {self.iteration}")
            ptr.append(None if t is None else C.c_void_p(t[row].data_ptr()))
        self.plan.call("alfa_set_forcing", *ptr, st)

    def step(self, n: int = 1):
        P = self.plan
        st = current_stream()
        for _ in range(n):
            sv = 0.0
            if self.sfc_series is not None:
                sv = float(self.sfc_series[self.iteration - 1])
            self._apply_forcing(st)
            if self.qsfc is not None:   # MoistureSurfaceBC_series[iteration-1]
                self.state.Q_sfc_val = self.qsfc[self.iteration - 1].data_ptr()
            P.call("alfa_step", C.byref(self.state), int(self.iteration), C.c_double(sv), st)
            self.iteration += 1

    def fields(self):
        """Host copies (float64 numpy) of u, v, w, TH (and Q) in stored form."""
        keys = ("u", "v", "w", "TH", "Q") if self.moist else ("u", "v", "w", "TH")
        return {k: self.t[k].double().cpu().numpy() for k in keys}

    def host(self, name):
        return self.t[name].double().cpu().numpy()

    def diag(self):
        s = self.scalars.cpu().numpy()
        return dict(CFL=float(s[0]), qz_sfc_avg=float(s[1]), ustar_rms=float(s[2]))
