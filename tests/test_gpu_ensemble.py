"""Ensembles and the request loop on one GPU (alfa_amd.ensemble).

* a step captured into a hipGraph and replayed gives the same fields as the
  eager alfa_step (the surface value read from device memory included);
* each served request gives the same result alone and in a mix of
  concurrently running requests of other sizes and cases (bitwise: every
  kernel is deterministic and each run has its own workspace; the stated
  tolerance is 0);
* a small ensemble runs and reports its throughput.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from alfa_amd import core  # noqa: E402
from alfa_amd.ensemble import (MemberSpec, Member, Server, run_ensemble,  # noqa: E402
                               run_ensemble_grouped)


@pytest.mark.parametrize("case,prec", [("stable", "mixed"), ("convective", "fp64"),
                                       ("neutral", "mixed")])
def test_graph_replay_matches_eager(case, prec):
    spec = MemberSpec(case, 32, 8, seed=3)
    m = Member(spec, prec, torch.cuda.Stream(), graphs=True)
    m.step(8)
    torch.cuda.synchronize()
    cfg, u, v, w, TH, ser = spec.build()
    les = core.LES(core.Plan(cfg, prec), u, v, w, TH, sfc_series=ser)
    les.step(8)
    torch.cuda.synchronize()
    assert len(m.graphs) >= 1
    for k in ("u", "v", "w", "TH"):
        a, b = m.les.host(k), les.host(k)
        print(case, prec, k, "max abs diff graph vs eager", float(np.max(np.abs(a - b))))
        assert np.array_equal(a, b)
    m.close()


@pytest.mark.parametrize("grouped", [False, True])
def test_served_result_is_the_same_alone_and_in_the_mix(grouped):
    reqs = [MemberSpec("stable", 32, 40, seed=1, arrival=0.0, cooling_K_per_h=0.3),
            MemberSpec("convective", 48, 30, seed=2, arrival=0.02),
            MemberSpec("neutral", 32, 50, seed=3, arrival=0.04),
            MemberSpec("convective", 32, 20, seed=4, arrival=0.06,
                       over=dict(SensibleHeatFlux=0.1, optSurfFlux=0))]
    mix = Server(64 << 30, "mixed", nstreams=3, round_steps=7, grouped=grouped)
    rep = mix.serve(reqs, keep=range(len(reqs)))
    print("mix:", rep)
    for i, r in enumerate(reqs):
        alone = Server(64 << 30, "mixed", nstreams=1, round_steps=11, grouped=False)
        solo = MemberSpec(r.case, r.n, r.steps, seed=r.seed, over=r.over,
                          cooling_K_per_h=r.cooling_K_per_h, arrival=0.0)
        alone.serve([solo], keep=[0])
        for k in ("u", "v", "w", "TH"):
            d = float(np.max(np.abs(mix.results[i][k] - alone.results[0][k])))
            print("request", i, k, "alone vs mix max abs diff", d)
            assert d == 0.0


def test_grouped_serving_matches_runs_alone():
    """Requests of one shape share launches once admitted: two stable and
    two convective 32^3 requests arrive together and form parts of two (the
    convective ones with the dynamic procedure every second step), a third of
    each shape arrives later and joins or forms its own part by SGS phase,
    and parts are rebuilt as requests finish.  Every result is bitwise the
    request served alone with per-member graphs."""
    reqs = [MemberSpec("stable", 32, 40, seed=11, arrival=0.0, cooling_K_per_h=0.2),
            MemberSpec("stable", 32, 33, seed=12, arrival=0.0, cooling_K_per_h=0.5,
                       over=dict(Ug2=6.0)),
            MemberSpec("convective", 32, 36, seed=14, arrival=0.0,
                       over=dict(dynamicSGS_call_time=2)),
            MemberSpec("convective", 32, 29, seed=15, arrival=0.0,
                       over=dict(dynamicSGS_call_time=2, SensibleHeatFlux=0.1)),
            MemberSpec("stable", 32, 17, seed=13, arrival=0.05, cooling_K_per_h=0.3),
            MemberSpec("convective", 32, 19, seed=16, arrival=0.05,
                       over=dict(dynamicSGS_call_time=2))]
    mix = Server(64 << 30, "mixed", nstreams=3, round_steps=4)
    rep = mix.serve(reqs, keep=range(len(reqs)))
    print("grouped mix:", rep)
    assert rep["grouped"] and rep["largest_part"] >= 2
    for i, r in enumerate(reqs):
        alone = Server(64 << 30, "mixed", nstreams=1, round_steps=11, grouped=False)
        solo = MemberSpec(r.case, r.n, r.steps, seed=r.seed, over=r.over,
                          cooling_K_per_h=r.cooling_K_per_h, arrival=0.0)
        alone.serve([solo], keep=[0])
        for k in ("u", "v", "w", "TH"):
            assert np.array_equal(mix.results[i][k], alone.results[0][k]), (i, r.case, k)


def _solo(spec, steps, prec):
    cfg, u, v, w, TH, ser = spec.build()
    les = core.LES(core.Plan(cfg, prec), u, v, w, TH, sfc_series=ser)
    les.step(steps)
    torch.cuda.synchronize()
    out = {f: les.host(f) for f in ("u", "v", "w", "TH")}
    out["CFL"] = les.diag()["CFL"]
    les.plan.close()
    return out


def test_small_ensemble_runs():
    """Four members of three grid sizes stepped together equal each member
    stepped alone, bit for bit, and the throughput is reported."""
    specs = [MemberSpec("convective", n, 6, seed=i, over=dict(optSurfFlux=0))
             for i, n in enumerate((32, 48, 32, 64))]
    out = run_ensemble(specs, steps=6, nstreams=2, keep_fields=True)
    print({k: v for k, v in out.items() if k != "fields"})
    assert out["members"] == 4 and out["mpoint_steps_per_s"] > 0
    for spec, got in zip(specs, out["fields"]):
        ref = _solo(spec, 6, "mixed")
        for k in ("u", "v", "w", "TH", "CFL"):
            assert np.array_equal(got[k], ref[k]), (spec.n, k)


def test_ensemble_64_members_at_full_size():
    """The batched-ensemble configuration at its stated size: 64 convective
    members on grids drawn from 48/64/80/96^3 with seeded Ug in [2, 12] m/s
    and surface heat flux in [0.05, 0.3] K m/s, mixed precision, stepped
    together (graph replays after the first step).  After 4 steps every
    member's fields and CFL are finite and bitwise those of the member run
    alone."""
    from alfa_amd.ensemble import ensemble_specs
    specs = ensemble_specs(64, steps=4, seed=0)
    assert sorted({s.n for s in specs}) == [48, 64, 80, 96]
    out = run_ensemble(specs, steps=4, keep_fields=True)
    print({k: v for k, v in out.items() if k != "fields"})
    for i, (spec, got) in enumerate(zip(specs, out["fields"])):
        for k in ("u", "v", "w", "TH"):
            assert np.all(np.isfinite(got[k])), (i, k)
        assert np.isfinite(got["CFL"]) and got["CFL"] > 0
        ref = _solo(spec, 4, "mixed")
        for k in ("u", "v", "w", "TH", "CFL"):
            assert np.array_equal(got[k], ref[k]), (i, spec.n, k)


def test_checkpoint_restart_is_exact(tmp_path):
    """alfa_amd.io checkpoint: 3 steps, save, restore into a fresh run, 3
    more steps equals 6 uninterrupted steps bit for bit (dynamic and static
    steps alternate: dynamicSGS_call_time = 2)."""
    from alfa_amd import cases, io
    cfg, u, v, w, TH = cases.initial_fields("stable", nx=32, ny=32, nz=32,
                                            dynamicSGS_call_time=2)
    ser = cases.surface_series(cfg, 10)
    for prec in ("fp64", "mixed"):
        ref = core.LES(core.Plan(cfg, prec), u, v, w, TH, sfc_series=ser)
        ref.step(6)
        a = core.LES(core.Plan(cfg, prec), u, v, w, TH, sfc_series=ser)
        a.step(3)
        p = tmp_path / f"ck_{prec}.npz"
        io.save_checkpoint(p, a)
        b = core.LES(core.Plan(cfg, prec), u * 0, v * 0, w * 0, TH * 0, sfc_series=ser)
        io.load_checkpoint(p, b)
        assert b.iteration == a.iteration
        b.step(3)
        torch.cuda.synchronize()
        for k in ("u", "v", "w", "TH", "Cs2_3D", "Cs2PrRatio_3D"):
            assert np.array_equal(b.host(k), ref.host(k)), (prec, k)
        snap = io.save_3d_fields(str(tmp_path), 6, b)
        with np.load(snap) as f:
            assert sorted(f.keys()) == ["TH", "u", "v", "w"]
            np.testing.assert_array_equal(f["u"], ref.host("u") + cfg.get("Ugal", 0.0))


def test_batch_step_equals_member_steps():
    """alfa_batch_step over a member table (different grids, precisions and
    cases) gives bitwise the fields of stepping each member alone."""
    import ctypes as C
    from alfa_amd import _lib, cases
    specs = [("stable", dict(nx=32, ny=32, nz=32), "mixed"),
             ("convective", dict(nx=48, ny=48, nz=32), "fp64"),
             ("neutral", dict(), "mixed")]
    batch, alone = [], []
    for name, over, prec in specs:
        cfg, u, v, w, TH = cases.initial_fields(name, **over)
        ser = cases.surface_series(cfg, 8)
        batch.append(core.LES(core.Plan(cfg, prec), u, v, w, TH, sfc_series=ser))
        alone.append(core.LES(core.Plan(cfg, prec), u, v, w, TH, sfc_series=ser))
    lib = batch[0].plan.lib
    n = len(batch)
    plans = (C.c_void_p * n)(*[m.plan.handle.value for m in batch])
    states = (_lib.AlfaState * n)(*[m.state for m in batch])
    for _ in range(4):
        its = (C.c_int * n)(*[m.iteration for m in batch])
        sv = (C.c_double * n)(*[0.0 if m.sfc_series is None else float(m.sfc_series[m.iteration - 1])
                                for m in batch])
        rc = lib.alfa_batch_step(C.cast(plans, C.c_void_p), C.cast(states, C.c_void_p), n,
                                 C.cast(its, C.c_void_p), C.cast(sv, C.c_void_p),
                                 core.current_stream())
        assert rc == 0, lib.alfa_last_error().decode()
        for m in batch:
            m.iteration += 1
        for m in alone:
            m.step(1)
    torch.cuda.synchronize()
    for a, b in zip(batch, alone):
        for k in ("u", "v", "w", "TH"):
            assert np.array_equal(a.host(k), b.host(k)), k


def test_grouped_ensemble_64_members_at_full_size():
    """The same 64-member configuration advanced by shared launches
    (alfa_group: one launch per kernel for all members of a grid size, the
    whole ensemble step in one graph per step kind): every member bitwise
    equal to its run alone after 5 steps, and far fewer launches per step
    than 64 separate steps."""
    from alfa_amd.ensemble import ensemble_specs
    specs = ensemble_specs(64, steps=5, seed=0)
    out = run_ensemble_grouped(specs, steps=5, keep_fields=True)
    print({k: v for k, v in out.items() if k != "fields"})
    assert out["groups"] == 4
    for i, (spec, got) in enumerate(zip(specs, out["fields"])):
        for k in ("u", "v", "w", "TH"):
            assert np.all(np.isfinite(got[k])), (i, k)
        ref = _solo(spec, 5, "mixed")
        for k in ("u", "v", "w", "TH", "CFL"):
            assert np.array_equal(got[k], ref[k]), (i, spec.n, k)


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_grouped_members_of_mixed_shapes(prec):
    """Groups of several shapes at once: stable members with a cooling
    surface series (each its own rate), dealiased neutral members, and
    members whose dynamic and static SGS steps alternate (two step kinds,
    two graphs); each member bitwise its run alone after 7 steps."""
    specs = [MemberSpec("stable", 32, 9, seed=1, cooling_K_per_h=0.2),
             MemberSpec("stable", 32, 9, seed=2, cooling_K_per_h=0.6, over=dict(Ug2=6.0)),
             MemberSpec("neutral", 32, 9, seed=3, over=dict(Ug2=8.0)),
             MemberSpec("neutral", 32, 9, seed=4, over=dict(Ug2=12.0)),
             MemberSpec("convective", 48, 9, seed=5, over=dict(dynamicSGS_call_time=2)),
             MemberSpec("convective", 48, 9, seed=6, over=dict(dynamicSGS_call_time=2,
                                                               SensibleHeatFlux=0.1))]
    out = run_ensemble_grouped(specs, steps=7, precision=prec, keep_fields=True)
    print({k: v for k, v in out.items() if k != "fields"})
    assert out["groups"] == 3
    for i, (spec, got) in enumerate(zip(specs, out["fields"])):
        ref = _solo(spec, 7, prec)
        for k in ("u", "v", "w", "TH", "CFL"):
            assert np.array_equal(got[k], ref[k]), (i, spec.case, k)
