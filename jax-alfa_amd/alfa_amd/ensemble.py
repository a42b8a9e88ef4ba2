"""Many independent LES runs on one GPU: batched ensembles and a request loop.

The reference advances one run per process (src/Main.py).  Here every run
("member") owns one ``alfa_plan`` (its grid, constants and workspace) and its
state tensors; members of any grid size share the device.

* Each member's time step is captured once into a hipGraph
  (``torch.cuda.CUDAGraph`` around ``alfa_step_sfc_dev``) after its first,
  eager step.  A replay launches the whole step (about a hundred kernels)
  with one call, and the surface value of the step is written to device
  memory before the replay (``alfa_step_sfc_dev`` reads it there).
* Members are spread over a few streams (default priority), so the kernels
  of small members overlap on the 256 CUs instead of running one grid at a
  time.
* ``run_ensemble`` steps a fixed set of members together (a synchronous
  ensemble); ``Server`` admits requests as they arrive, within an HBM budget,
  steps every admitted member in rounds (continuous batching: a member that
  arrives joins the next round, a member that finishes leaves), and reports
  completion-latency percentiles and aggregate throughput.
* Members of one configuration shape share launches (``alfa_group``): one
  launch per kernel for all of them, each reading its own arguments from a
  device table (``run_ensemble_grouped``; the server regroups its admitted
  requests between rounds).

Results do not depend on the mix: every kernel is deterministic (fixed
reduction orders) and each member has its own workspace, so a member's
fields are bitwise the same alone and in any mix.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib, cases, core

# base grid of each case: the time step is scaled with the grid spacing
_BASE_N = {"neutral": 32, "convective": 128, "stable": 256}


@dataclass
class MemberSpec:
    """One run: case, grid, forcing, length and arrival time (serving)."""
    case: str
    n: int
    steps: int
    seed: int = 0
    over: dict = field(default_factory=dict)
    cooling_K_per_h: float = 0.25
    arrival: float = 0.0

    def build(self):
        """-> (cfg, u, v, w, TH, surface series) in stored form."""
        base = cases.CASES[self.case]
        over = dict(self.over)
        over.setdefault("dt", base["dt"] * _BASE_N[self.case] / self.n)
        cfg, u, v, w, TH = cases.initial_fields(self.case, seed=self.seed, nx=self.n, ny=self.n,
                                                nz=self.n, **over)
        ser = None
        if cfg["optSurfBC"] == 2:
            t = np.arange(self.steps + 2) * cfg["dt"]
            ser = -self.cooling_K_per_h * t / 3600.0
        elif cfg["optSurfBC"] == 1:
            ser = np.full(self.steps + 2, cfg["SensibleHeatFlux"])
        return cfg, u, v, w, TH, ser


def ensemble_specs(nmembers=64, sizes=(48, 64, 80, 96), steps=500, seed=0):
    """The ensemble workload: convective physics, per-member Ug in [2, 12] m/s and surface
    heat flux in [0.05, 0.3] K m/s (homogeneous flux), grids drawn from sizes."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(nmembers):
        n = int(rng.choice(sizes))
        out.append(MemberSpec("convective", n, steps, seed=seed * 1000 + m,
                              over=dict(Ug2=float(rng.uniform(2, 12)),
                                        SensibleHeatFlux=float(rng.uniform(0.05, 0.3)),
                                        optSurfFlux=0)))
    return out


def serving_requests(nreq=40, rate=2.0, sizes=(32, 64, 128, 192), steps=(200, 2000), seed=0):
    """The serving workload: a Poisson stream (rate per second) of requests mixing the
    neutral, convective and stable setups with seeded forcing."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for r in range(nreq):
        t += float(rng.exponential(1.0 / rate))
        case = str(rng.choice(["neutral", "convective", "stable"]))
        n = int(rng.choice(sizes))
        nsteps = int(rng.integers(steps[0], steps[1] + 1))
        if case == "neutral":
            over = dict(Ug2=float(rng.uniform(5, 15)))
        elif case == "convective":
            over = dict(SensibleHeatFlux=float(rng.uniform(0.05, 0.3)), Ug2=float(rng.uniform(1, 5)))
        else:
            over = dict(Ug2=float(rng.uniform(5, 10)))
        spec = MemberSpec(case, n, nsteps, seed=seed * 1000 + r, over=over,
                          cooling_K_per_h=float(rng.uniform(0.1, 0.5)), arrival=t)
        out.append(spec)
    return out


# FFT tile choices per (device, precision, grid, filter ratio, SGS model),
# autotuned on steps of the first member of that kind and reused for the rest
_TILES = {}


def _tuned_tiles(m):
    P, cfg = m.plan, m.cfg
    key = (str(P.device), P.precision, m.spec.n, cfg.get("FGR"), cfg.get("optSgs"))
    if key in _TILES:
        return P.set_tiles(**_TILES[key])
    _TILES[key] = P.autotune(m.les, stream=m.stream)
    return _TILES[key]


# The first streams a process creates were measured to run an ensemble about
# 25% slower than streams created after them (1319 against 1744 Mpoint-steps/s
# for 64 members at 48-96^3 on 4 streams): they share hardware queues with
# work the process issued before.  A first set of streams is therefore
# created, used once and kept aside, and members get streams made after it.
_SPARE_STREAMS = []


def member_streams(n):
    torch = core._torch()
    if not _SPARE_STREAMS:
        for _ in range(4):
            st = torch.cuda.Stream()
            with torch.cuda.stream(st):
                torch.ones(1, device="cuda").add_(1)
            _SPARE_STREAMS.append(st)
        torch.cuda.synchronize()
    return [torch.cuda.Stream() for _ in range(n)]


def state_bytes(n, precision):
    """Caller-owned state of one member: 10 fields + 6 planes + profiles."""
    es = 8 if precision == "fp64" else 4
    return 10 * n ** 3 * es + 6 * n * n * es + 6 * n * 8


class Member:
    """One run on the device: plan, state, stream and captured step graphs."""

    def __init__(self, spec: MemberSpec, precision="mixed", stream=None, graphs=True, tune=True):
        torch = core._torch()
        self.spec = spec
        cfg, u, v, w, TH, ser = spec.build()
        self.cfg = cfg
        self.stream = stream or torch.cuda.current_stream()
        with torch.cuda.stream(self.stream):
            self.plan = core.Plan(cfg, precision)
            self.les = core.LES(self.plan, u, v, w, TH, sfc_series=ser)
            self.sfc = torch.zeros(1, dtype=torch.float64, device=self.plan.device)
            if tune:
                self.tiles = _tuned_tiles(self)
        self.bytes = int(self.plan.derived["workspace_bytes"]) + state_bytes(spec.n, precision)
        self.done = 0
        self.use_graphs = graphs
        self.graphs = {}
        self.t_admit = None
        self.t_done = None

    @property
    def points(self):
        return self.spec.n ** 3

    def _dynamic(self, it):
        c = self.cfg
        return c["optSgs"] >= 1 and (it == c.get("istep", 1) or it % c["dynamicSGS_call_time"] == 0)

    def _launch(self, it):
        P = self.plan
        st = core.current_stream()
        _lib.check(P.lib.alfa_step_sfc_dev(P.handle, C.byref(self.les.state), int(it),
                                           C.c_void_p(self.sfc.data_ptr()), st),
                   "alfa_step_sfc_dev")

    def step(self, n=1):
        """Enqueue n steps on the member's stream (graph replays after the
        first step of each kind)."""
        torch = core._torch()
        les = self.les
        with torch.cuda.stream(self.stream):
            for _ in range(n):
                it = les.iteration
                sv = 0.0 if les.sfc_series is None else float(les.sfc_series[it - 1])
                self.sfc.fill_(sv)
                first = it == self.cfg.get("istep", 1)
                key = self._dynamic(it)
                if not self.use_graphs or first:
                    self._launch(it)
                elif key in self.graphs:
                    self.graphs[key].replay()
                else:
                    g = torch.cuda.CUDAGraph()
                    with torch.cuda.graph(g, stream=self.stream):
                        self._launch(it)
                    self.graphs[key] = g
                    g.replay()
                les.iteration += 1
                self.done += 1

    def close(self):
        self.graphs.clear()
        self.plan.close()


# configuration entries that decide a step's launch sequence: members that
# agree in these (and in grid and precision) can share launches
_SHAPE_KEYS = ("FGR", "optSgs", "dynamicSGS_call_time", "optSurfBC", "optSurfFlux",
               "optMoisture", "optMoistureSurfBC", "optBuoyancy", "optDamping",
               "optPressureSolver", "istep")


def shape_key(m):
    return (m.spec.n, m.plan.precision) + tuple(m.cfg.get(k) for k in _SHAPE_KEYS)


class MemberGroup:
    """Members of one configuration shape advanced by shared launches
    (alfa_group): every kernel of the step runs once for the whole group,
    each member reading its own arguments, so a member's fields are bitwise
    those of its run alone."""

    def __init__(self, members):
        self.members = members
        n = len(members)
        lib = members[0].plan.lib
        self.lib = lib
        plans = (C.c_void_p * n)(*[m.plan.handle.value for m in members])
        states = (_lib.AlfaState * n)(*[m.les.state for m in members])
        sfc = (C.c_void_p * n)(*[m.sfc.data_ptr() for m in members])
        self._h = C.c_void_p()
        _lib.check(lib.alfa_group_create(C.cast(plans, C.c_void_p), C.cast(states, C.c_void_p),
                                         C.cast(sfc, C.c_void_p), n, C.byref(self._h)),
                   "alfa_group_create")

    @property
    def iteration(self):
        return self.members[0].les.iteration

    def set_surface(self):
        """Each member's surface value of the coming step, on the device."""
        for m in self.members:
            if m.les.sfc_series is not None:
                m.sfc.fill_(float(m.les.sfc_series[m.les.iteration - 1]))

    def enqueue(self):
        """Enqueue one group step on the current stream (no host sync)."""
        _lib.check(self.lib.alfa_group_step(self._h, int(self.iteration), core.current_stream()),
                   "alfa_group_step")

    def advance(self):
        for m in self.members:
            m.les.iteration += 1
            m.done += 1

    def launches(self):
        n = C.c_int(-1)
        _lib.check(self.lib.alfa_group_launches(self._h, int(self.iteration), C.byref(n)),
                   "alfa_group_launches")
        return n.value

    def close(self):
        if self._h.value:
            self.lib.alfa_group_destroy(self._h)
            self._h = C.c_void_p()


def run_ensemble_grouped(specs: List[MemberSpec], steps=None, precision="mixed", graphs=True,
                         warmup=3, keep_fields=False, parallel=True):
    """An ensemble with shared launches: members are grouped by configuration
    shape (grid, precision, options) and each group's step is one launch
    per kernel over its members; after one eager step of each kind the
    whole ensemble step (every group) is captured into one hipGraph per
    kind and replayed.  parallel: each group on its own stream, forked from
    and joined back into the step's stream (parallel branches of the graph),
    so that the small groups' launches overlap the large ones.  Returns the
    run_ensemble figures."""
    torch = core._torch()
    stream = member_streams(1)[0]
    t0 = time.perf_counter()
    members = [Member(s, precision, stream, graphs=False) for s in specs]
    by_key = {}
    for m in members:
        by_key.setdefault(shape_key(m), []).append(m)
    with torch.cuda.stream(stream):
        # largest groups first: their branches start first
        groups = [MemberGroup(ms) for ms in sorted(by_key.values(),
                                                     key=lambda ms: -sum(m.points for m in ms))]
    side = member_streams(len(groups)) if parallel and len(groups) > 1 else None
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    nsteps = steps or specs[0].steps
    graphs_by_kind = {}

    def kind():
        # the step kind of every group (first step, dynamic or static SGS):
        # a captured graph replays each group's program of that kind
        out = []
        for g in groups:
            m = g.members[0]
            it = m.les.iteration
            out.append((it == m.cfg.get("istep", 1), m._dynamic(it)))
        return tuple(out)

    def enqueue_all():
        if side is None:
            for g in groups:
                g.enqueue()
            return
        fork = torch.cuda.Event()
        fork.record(stream)
        joins = []
        for g, st in zip(groups, side):
            st.wait_event(fork)
            with torch.cuda.stream(st):
                g.enqueue()
            e = torch.cuda.Event()
            e.record(st)
            joins.append(e)
        for e in joins:
            stream.wait_event(e)

    def one_step(capture):
        with torch.cuda.stream(stream):
            for g in groups:
                g.set_surface()
            k = kind()
            if not capture or not graphs:
                enqueue_all()
            elif k in graphs_by_kind:
                graphs_by_kind[k].replay()
            else:
                gr = torch.cuda.CUDAGraph()
                with torch.cuda.graph(gr, stream=stream):
                    enqueue_all()
                graphs_by_kind[k] = gr
                gr.replay()
            for g in groups:
                g.advance()

    # untimed: the first step of each kind builds its member tables (eager),
    # the next one is captured
    warmup = max(1, min(max(3, warmup), nsteps - 1))
    seen = set()
    for _ in range(warmup):
        k = kind()
        one_step(capture=k in seen)
        seen.add(k)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(nsteps - warmup):
        k = kind()
        one_step(capture=k in seen)
        seen.add(k)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    pts = sum(m.points for m in members) * (nsteps - warmup)
    out = dict(members=len(members), steps_timed=nsteps - warmup, wall_s=wall, setup_s=t_setup,
               mpoint_steps_per_s=pts / wall / 1e6, groups=len(groups),
               streams=len(side) if side else 1,
               launches_per_step=sum(g.launches() for g in groups), graphs=graphs,
               grids={str(n): sum(1 for s in specs if s.n == n) for n in sorted({s.n for s in specs})},
               hbm_bytes=sum(m.bytes for m in members))
    if keep_fields:
        out["fields"] = [dict({f: m.les.host(f) for f in ("u", "v", "w", "TH")},
                              CFL=m.les.diag()["CFL"]) for m in members]
    graphs_by_kind.clear()
    for g in groups:
        g.close()
    for m in members:
        m.close()
    return out


def run_ensemble(specs: List[MemberSpec], steps=None, nstreams=4, precision="mixed",
                 graphs=True, log=None, warmup=2, keep_fields=False):
    """Step all members together for `steps` steps (default: each spec's
    own length, all equal in the ensemble workload), each member with its own launches
    and step graph on one of nstreams streams.  Returns throughput figures,
    and with keep_fields each member's final u, v, w, TH and CFL (host
    copies) under "fields"."""
    torch = core._torch()
    streams = member_streams(nstreams)
    t0 = time.perf_counter()
    members = [Member(s, precision, streams[i % nstreams], graphs) for i, s in enumerate(specs)]
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    nsteps = steps or specs[0].steps
    # untimed: the eager first step, the graph captures (2 steps) and any
    # further warmup steps
    warmup = max(2, min(warmup, nsteps - 1))
    for m in members:
        m.step(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(nsteps - warmup):
        for m in members:
            m.step(1)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    pts = sum(m.points for m in members) * (nsteps - warmup)
    out = dict(members=len(members), steps_timed=nsteps - warmup, wall_s=wall, setup_s=t_setup,
               mpoint_steps_per_s=pts / wall / 1e6, streams=nstreams, graphs=graphs,
               grids={str(n): sum(1 for s in specs if s.n == n) for n in sorted({s.n for s in specs})},
               hbm_bytes=sum(m.bytes for m in members))
    if keep_fields:
        out["fields"] = [dict({f: m.les.host(f) for f in ("u", "v", "w", "TH")},
                              CFL=m.les.diag()["CFL"]) for m in members]
    for m in members:
        m.close()
    return out


class _ServedGroup:
    """Admitted requests of one configuration shape and SGS phase, advanced
    by shared launches (MemberGroup) for the rounds in which the set does not
    change: the first step of each step kind runs eagerly (it records the
    group's launch tables), the second is captured into a hipGraph, later
    ones replay it."""

    def __init__(self, members, stream, graphs):
        torch = core._torch()
        self.members = members
        self.stream = stream
        self.use_graphs = graphs
        self.graphs = {}
        self.seen = set()
        with torch.cuda.stream(stream):
            self.group = MemberGroup(members)

    def step(self, n):
        torch = core._torch()
        g = self.group
        with torch.cuda.stream(self.stream):
            for _ in range(n):
                g.set_surface()
                m = self.members[0]
                kind = m._dynamic(m.les.iteration)
                if not self.use_graphs or kind not in self.seen:
                    g.enqueue()
                    self.seen.add(kind)
                elif kind in self.graphs:
                    self.graphs[kind].replay()
                else:
                    gr = torch.cuda.CUDAGraph()
                    with torch.cuda.graph(gr, stream=self.stream):
                        g.enqueue()
                    self.graphs[kind] = gr
                    gr.replay()
                g.advance()

    def close(self):
        self.graphs.clear()
        self.group.close()


def _sgs_phase(m):
    """Members of one shape whose iterations agree modulo the dynamic-SGS
    period take the same kind of step at every later iteration."""
    c = m.cfg
    ct = int(c.get("dynamicSGS_call_time", 1) or 1)
    return m.les.iteration % ct if c["optSgs"] >= 1 and ct > 1 else 0


class Server:
    """Continuous-batching request loop with HBM-budget admission.

    grouped (default): between rounds the admitted requests are partitioned
    by configuration shape (grid, precision, options) and SGS phase, and each
    part advances by shared launches (one launch per kernel for all its
    members, alfa_group) on one of the streams; a part is kept, with its
    launch tables and graphs, for as long as its set of requests does not
    change.  A request's first step runs alone and eagerly.  Otherwise every
    admitted request advances by its own graph replays on one of the
    streams.  Either way a request's fields are bitwise those of its run
    alone."""

    def __init__(self, hbm_budget_bytes, precision="mixed", nstreams=4, round_steps=10,
                 graphs=True, grouped=True):
        torch = core._torch()
        self.budget = int(hbm_budget_bytes)
        self.precision = precision
        self.streams = member_streams(nstreams)
        self.round_steps = round_steps
        self.graphs = graphs
        self.grouped = grouped
        self.results = {}
        self._parts = {}
        self._formed = 0
        self._largest = 0

    def _round_grouped(self, active):
        """One round: first steps of new requests, then every part of the
        admitted set advances min(round_steps, its shortest remainder)."""
        new = [m for m in active.values() if m.done == 0]
        for m in new:
            m.step(1)
        if new:   # the parts below run on other streams
            core._torch().cuda.synchronize()
        parts = {}
        for i, m in active.items():
            if m.done < m.spec.steps:
                parts.setdefault(shape_key(m) + (_sgs_phase(m),), []).append(i)
        keys = {tuple(ids) for ids in parts.values()}
        for ids in [ids for ids in self._parts if ids not in keys]:
            self._parts.pop(ids).close()
        for k, ids in enumerate(sorted(keys, key=lambda t: -active[t[0]].points * len(t))):
            ms = [active[i] for i in ids]
            part = self._parts.get(ids)
            if part is None:
                part = self._parts[ids] = _ServedGroup(ms, self.streams[k % len(self.streams)],
                                                       self.graphs)
                self._formed += 1
                self._largest = max(self._largest, len(ms))
            part.step(min(self.round_steps, min(m.spec.steps - m.done for m in ms)))

    def estimate_bytes(self, spec):
        # workspace per grid point measured at plan creation (mixed ~560 B,
        # FP64 ~1120 B at FGR=2; FGR=1 carries padded buffers too)
        per = 560 if self.precision == "mixed" else 1120
        if cases.CASES[spec.case]["FGR"] == 1:
            per *= 2
        return per * spec.n ** 3 + state_bytes(spec.n, self.precision)

    def serve(self, requests: List[MemberSpec], keep=None, log=None):
        """Run the request stream in real time.  keep: request indices whose
        final fields are returned (for the alone-versus-mix check)."""
        torch = core._torch()
        keep = set(keep or [])
        pending = deque(sorted(range(len(requests)), key=lambda i: requests[i].arrival))
        active = {}
        used = 0
        lat = {}
        points = 0
        t0 = time.perf_counter()
        k = 0
        while pending or active:
            now = time.perf_counter() - t0
            # admission: arrived requests in order, while the budget allows
            while pending and requests[pending[0]].arrival <= now:
                i = pending[0]
                need = self.estimate_bytes(requests[i])
                if active and used + need > self.budget:
                    break
                pending.popleft()
                # no tile autotuning at admission: its timed steps hold the
                # request loop (and every queued request) for their duration
                m = Member(requests[i], self.precision, self.streams[k % len(self.streams)],
                           self.graphs and not self.grouped, tune=False)
                k += 1
                m.t_admit = time.perf_counter() - t0
                used += m.bytes
                active[i] = m
            if not active:
                time.sleep(max(0.0, min(0.05, requests[pending[0]].arrival - now)))
                continue
            # one round: every admitted member advances up to round_steps
            if self.grouped:
                self._round_grouped(active)
            else:
                for i, m in active.items():
                    m.step(min(self.round_steps, m.spec.steps - m.done))
            torch.cuda.synchronize()
            now = time.perf_counter() - t0
            for i in [i for i, m in active.items() if m.done >= m.spec.steps]:
                for ids in [ids for ids in self._parts if i in ids]:
                    self._parts.pop(ids).close()
                m = active.pop(i)
                lat[i] = now - requests[i].arrival
                points += m.points * m.spec.steps
                if i in keep:
                    self.results[i] = {f: m.les.host(f) for f in ("u", "v", "w", "TH")}
                used -= m.bytes
                m.close()
        wall = time.perf_counter() - t0
        for part in self._parts.values():
            part.close()
        self._parts.clear()
        v = np.array([lat[i] for i in sorted(lat)])
        return dict(requests=len(requests), wall_s=wall, p50_s=float(np.percentile(v, 50)),
                    p99_s=float(np.percentile(v, 99)), max_s=float(v.max()),
                    mpoint_steps_per_s=points / wall / 1e6, budget_bytes=self.budget,
                    streams=len(self.streams), round_steps=self.round_steps,
                    grouped=self.grouped, parts_formed=self._formed,
                    largest_part=self._largest)
