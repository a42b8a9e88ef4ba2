#!/usr/bin/env python3
"""Throughput of the composed LES time step on one MI355X.

Workload: the stable boundary layer at 256^3 (400^3 m, LAD-SM dynamic
Smagorinsky, FGR=2, Thomas pressure solve, prescribed cooling surface,
dt = 0.05 s), mixed FP32/FP64, fields resident in HBM.  The same workload in
plain FP64 is timed beside it.  One JSON line goes to stdout:
  value       = whole-job point-steps/s of the mixed run, in Mpoint-steps/s
  roofline    = the dominant kernel's algorithmic bytes (the reference data
                flow: one full read + one full write per transformed field)
                / its HIP-event time, measured on its launches inside the
                timed region; the kernel's own distinct bytes beside it
  cpu_baseline= the FP64 oracle (NumPy, scipy.fft on the host cores) on a
                bounded sample
  fft_tiles   = the FFT kernels' workgroup sizes, chosen after the warmup by
                the plan's autotuner (timed on steps from a copy of the
                state, outside the timed region), or given by --tiles
Diagnostics (per-kernel table) go to stderr, and to --kernels-json PATH if given.

    python bench.py [--gpus 1] [--steps 20] [--warmup 5]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "jax-alfa_amd"))

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per point-step of the reference data flow at FGR=2
# (one read + one write per batched FFT pass, distinct inputs + outputs per
# pointwise pass): 567 field passes x 4 B (FP32) or x 8 B (FP64)
B_ALG = {"mixed": 2270.0, "fp64": 4540.0}
# A100 80 GB, the reference's own run of this grid size (CBL_N91 256^3 LAD-SM SP),
# SimulationLog.ipynb:661 / examples/CBL_N91/runs/256x256x256_LASDD_SM_SP/Config.py:71-72
A100_REF = {"Mpoint_steps_per_s": 202.1, "ms_per_step": 83.0,
            "run": "CBL_N91 256^3 LAD-SM SP (A100 80 GB)"}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")

# --case: (description, domain and model, optSgs, FGR) of the timed step
CASE_TEXT = {
    "stable": ("GABLS1-like stable boundary layer", "400^3 m, LAD-SM, FGR=2, dt=0.05 s", 3, 2),
    "convective": ("convective boundary layer", "5120^2 x 2048 m, LASDD-SM, FGR=2, dt=0.2 s", 1, 2),
    "neutral": ("neutral boundary layer", "2000^2 x 1000 m, LAD-SM, FGR=1 (dealiased), dt=0.5 s",
                3, 1),
}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by what the
    counter reads for a plain device copy of known size), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    e = d.get("kernels", {}).get(kernel)
    return None if e is None else float(e["bytes_per_launch"])


def measured_step_traffic(precision):
    """HBM bytes per point-step of the whole step, from the committed counter
    summary: every library kernel's bytes per launch times its dispatches,
    over the steps of the profiled run and the points of its 256^3 grid; or
    None."""
    try:
        with open(TRAFFIC_FILE) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    sfx = "" if precision == "mixed" else "_fp64"
    k, steps = d.get("kernels" + sfx), d.get("steps_in_run" + sfx)
    if not k or not steps:
        return None
    return sum(e["bytes_per_launch"] * e["dispatches"] for e in k.values()) / steps / 256 ** 3


def step_hbm(rate_mpts, precision):
    """The step's measured HBM traffic at the measured rate, against peak."""
    b = measured_step_traffic(precision)
    if b is None:
        return None
    gbs = rate_mpts * 1e6 * b / 1e9
    return {"bytes_per_point_step": round(b, 1), "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
            "source": "rocprofv3 FETCH_SIZE + WRITE_SIZE of every library kernel of the 256^3 "
                      "step (profiles/hbm_traffic.json)"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def field_stats(les, cfg):
    """Host copies of what the error report compares: mean wind and
    temperature profiles, resolved variances, SGS coefficient fields."""
    import numpy as np
    f = les.fields()
    out = {}
    for k in ("u", "v", "w", "TH"):
        a = f[k]
        m = a.mean(axis=(0, 1))
        out["mean_" + k] = m
        out["var_" + k] = (a * a).mean(axis=(0, 1)) - m * m
    out["mean_u"] = out["mean_u"] + cfg["Ugal"]
    out["mean_TH"] = out["mean_TH"] + cfg["T_0"]
    out["Cs2"] = les.host("Cs2_3D")
    out["CsPr"] = les.host("Cs2PrRatio_3D")
    return out


def error_report(mx, fp):
    """Mixed-precision result against the FP64 path after the same steps."""
    import numpy as np
    speed = float(np.max(np.hypot(fp["mean_u"], fp["mean_v"])))
    r = {"mean_wind_profile": max(float(np.max(np.abs(mx["mean_" + k] - fp["mean_" + k])))
                                  for k in ("u", "v")) / speed,
         "mean_theta_profile": float(np.max(np.abs(mx["mean_TH"] - fp["mean_TH"])
                                            / np.abs(fp["mean_TH"])))}
    for k in ("u", "v", "w", "TH"):
        ref = fp["var_" + k]
        r["resolved_variance_" + k] = float(np.max(np.abs(mx["var_" + k] - ref)) /
                                            max(np.max(np.abs(ref)), 1e-300))
    # The dynamic coefficients are ratios of filtered contractions; where the
    # flow is still laminar both numerator and denominator are rounding noise
    # (and so is the coefficient, in either precision), while the stresses
    # there vanish with |S|.  They are therefore also compared on the levels
    # that carry resolved turbulence (TKE >= 1% of its largest level value).
    tke = fp["var_u"] + fp["var_v"] + fp["var_w"]
    turb = tke >= 0.01 * np.max(tke)
    for k, name in (("Cs2", "Cs2_3D"), ("CsPr", "Cs2PrRatio_3D")):
        ref = fp[k]
        r[name] = float(np.max(np.abs(mx[k] - ref)) / max(np.max(np.abs(ref)), 1e-300))
        pm, pr = mx[k].mean(axis=(0, 1)), ref.mean(axis=(0, 1))
        r[name + "_mean_profile"] = float(np.max(np.abs(pm - pr)) / max(np.max(np.abs(pr)), 1e-300))
        r[name + "_mean_profile_turbulent_levels"] = float(
            np.max(np.abs(pm - pr)[turb]) / max(np.max(np.abs(pr[turb])), 1e-300))
    r = {k: float("%.3e" % v) for k, v in r.items()}
    r["turbulent_levels"] = int(turb.sum())
    return r


def run_gpu(prec, n, steps, warmup, find_dominant=True, dominant=None, tiles=None, fusion=True,
            case="stable"):
    import torch
    from alfa_amd import cases, core
    cfg, u, v, w, TH = cases.initial_fields(case, nx=n, ny=n, nz=n)
    ser = cases.surface_series(cfg, warmup + steps + 8)
    P = core.Plan(cfg, prec)
    P.set_fusion(fusion)
    les = core.LES(P, u, v, w, TH, sfc_series=ser)
    del u, v, w, TH
    log(f"[bench] {prec} {n}^3 plan ready, workspace {P.derived['workspace_bytes'] / 2**30:.2f} GiB")
    les.step(warmup)
    torch.cuda.synchronize()
    if tiles:   # fixed FFT tile choices (y R2C, x, y C2R)
        tiles = P.set_tiles(*tiles)
        log(f"[bench] {prec} FFT tiles (given): {tiles}")
    else:       # timed on steps from a copy of the warmed-up state
        tiles = P.autotune(les)
        log(f"[bench] {prec} FFT tiles (autotuned on steps): {tiles}")
    table = None
    if find_dominant:
        P.profile_begin(None, 16384)
        les.step(2)
        table = P.profile_end()
        table.sort(key=lambda r: -r["total_ms"])
        dominant = table[0]["name"]
    P.profile_begin(dominant, 8192)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    les.step(steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    live = P.profile_end()
    d = les.diag()
    cfl = d["CFL"]
    stats = field_stats(les, cfg)
    P.close()
    del les
    torch.cuda.empty_cache()
    ms = (t1 - t0) * 1e3 / steps
    return dict(ms=ms, mpts=(n ** 3) / (ms * 1e-3) / 1e6, table=table, dominant=dominant,
                live=live, cfl=cfl, stats=stats, tiles=tiles,
                total_steps=warmup + (2 if find_dominant else 0) + steps)


def _oracle_rate(O, name, over, nsteps, budget_s):
    """Point-steps/s of the FP64 oracle stepping case `name` from its initial
    state: nsteps steps, or fewer if budget_s runs out (at least one)."""
    from alfa_amd import cases
    cfg, u, v, w, TH = cases.initial_fields(name, **over)
    g = O.Grid(cfg)
    s = O.initial_state(g, u, v, w, TH)
    ser = cases.surface_series(cfg, nsteps + 2)
    th = O.ThomasPressureInit(g)
    k = 0
    t0 = time.perf_counter()
    while k < nsteps:
        s = O.step(g, s, th, None if ser is None else ser[s.iteration - 1])
        k += 1
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    return g.nx * g.ny * g.nz * k / el / 1e6, k, el


def cpu_baseline():
    """The FP64 NumPy/SciPy oracle on the host cores: FFTs through scipy.fft
    with one worker per core of the process's CPU share (at most 16, the
    share of one GPU), the rest of the NumPy work on one thread.  Samples:
    the stable layer of the GPU workload (same physics: LAD-SM, FGR=2,
    prescribed cooling surface) at 128^3 for 10 steps, and the neutral 32^3
    configuration for its 100 steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import alfa_oracle as O
    workers = min(len(os.sched_getaffinity(0)), 16)
    O.set_fft_workers(workers)
    try:
        v128, k128, t128 = _oracle_rate(O, "stable", dict(nx=128, ny=128, nz=128), 10, 30.0)
        v32, k32, t32 = _oracle_rate(O, "neutral", {}, 100, 30.0)
    finally:
        O.set_fft_workers(None)
    return dict(value=round(v128, 4), unit="Mpoint-steps/s", cores=workers, kind="port",
                sample=f"FP64 NumPy/SciPy oracle, stable layer 128^3 (LAD-SM, FGR=2), {k128} "
                       f"steps in {t128:.1f} s from the initial state; scipy.fft workers="
                       f"{workers}, other NumPy work single-threaded",
                neutral32=dict(value=round(v32, 4), steps=k32, seconds=round(t32, 1),
                             workload="neutral 32^3, LAD-SM, FGR=1 (dealiased), dt=0.5 s"))


def copy_probe(mib):
    """3 device copies of a mib-MiB buffer: a plain copy kernel of known size
    beside the measured kernels, to scale the HBM traffic counters."""
    if mib <= 0:
        return
    import torch
    nel = mib * 2 ** 20 // 4
    src = torch.ones(nel, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    # the same source read as 64-byte rows (16 floats of every 32), the access
    # width of the FFT kernels' tiles, into a half-size destination
    half = torch.empty(nel // 2, dtype=torch.float32, device="cuda")
    rows = src.view(-1, 32)[:, :16]
    for _ in range(3):
        half.view(-1, 16).copy_(rows)
    torch.cuda.synchronize()
    log(f"[bench] copy probe: 3 copies of {mib} MiB, 3 copies of its 64-byte rows "
        f"({mib // 2} MiB read, {mib // 2} MiB written)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ensemble-members", type=int, default=64)
    ap.add_argument("--ensemble-steps", type=int, default=500,
                    help="steps per member (the ensemble configuration's 500)")
    ap.add_argument("--serve-requests", type=int, default=24)
    ap.add_argument("--serve-steps", type=int, nargs=2, default=(200, 2000), metavar=("MIN", "MAX"),
                    help="run-length range of served requests (the serving configuration's "
                         "200-2000)")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-serve-compare", action="store_true",
                    help="skip the serving run with per-request launches beside the grouped one")
    ap.add_argument("--no-developed", action="store_true",
                    help="skip the mixed-vs-FP64 comparison at a developed state (64^3, 1 h)")
    ap.add_argument("--tiles", default=None, metavar="R2C,X,C2R",
                    help="threads per workgroup of the y R2C, x and y C2R FFT kernels (256 or "
                         "512 each) instead of the plan's autotuner, e.g. under a profiler, "
                         "whose trace would also hold the autotuner's trial steps")
    ap.add_argument("--case", default="stable", choices=("stable", "convective", "neutral"),
                    help="the seeded case of the timed step (default: the stable layer of the "
                         "headline workload; convective runs the LASDD model)")
    ap.add_argument("--no-fusion", action="store_true",
                    help="run the LAD dynamic procedure's separate test-filter C2R and "
                         "contraction kernels instead of the fused ones (A/B comparisons)")
    ap.add_argument("--kernels-json", default=None, metavar="PATH",
                    help="write the per-kernel tables (HIP-event times, bytes) to PATH")
    ap.add_argument("--skip-mixed", action="store_true",
                    help="profiling: time only the FP64 step (the JSON line then carries it)")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="profiling: run only the ensemble and serving workloads")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="after the run, copy a MIB-MiB device buffer 3 times (a plain copy "
                         "kernel of known size, to scale HBM traffic counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the per-member step does not shard across GPUs in this "
              "build (ensembles shard by member). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    n = args.grid
    tiles = [int(t) for t in args.tiles.split(",")] if args.tiles else None
    if tiles and (len(tiles) != 3 or any(t not in (256, 512) for t in tiles)):
        ap.error("--tiles takes three of 256 or 512, e.g. 256,256,512")
    if args.ensemble_only:
        from alfa_amd import ensemble as E
        specs = E.ensemble_specs(args.ensemble_members, steps=args.ensemble_steps, seed=0)
        ens = E.run_ensemble_grouped(specs)
        log(f"[bench] ensemble: {ens}")
        print(json.dumps({"ensemble": ens}), flush=True)
        return
    if args.skip_mixed:
        fp64 = run_gpu("fp64", n, args.steps, args.warmup, find_dominant=True, tiles=tiles)
        log(f"[bench] fp64: {fp64['ms']:.3f} ms/step, {fp64['mpts']:.1f} Mpoint-steps/s")
        for r in (fp64["table"] or [])[:25]:
            bw = r["bytes"] / (r["total_ms"] * 1e-3) / 1e9 if r["total_ms"] > 0 else 0
            log(f"  {r['name']:40s} {r['launches']:5d} launches {r['total_ms'] / 2:9.3f} ms/step "
                f"{bw:8.0f} GB/s")
        copy_probe(args.copy_probe)
        print(json.dumps({"fp64_ms_per_step": fp64["ms"], "value": fp64["mpts"]}), flush=True)
        return
    mixed = run_gpu("mixed", n, args.steps, args.warmup, tiles=tiles, fusion=not args.no_fusion,
                    case=args.case)
    log(f"[bench] mixed: {mixed['ms']:.3f} ms/step, {mixed['mpts']:.1f} Mpoint-steps/s, "
        f"CFL {mixed['cfl']:.3f}")
    for r in (mixed["table"] or [])[:25]:
        bw = r["bytes"] / (r["total_ms"] * 1e-3) / 1e9 if r["total_ms"] > 0 else 0
        log(f"  {r['name']:40s} {r['launches']:5d} launches {r['total_ms'] / 2:9.3f} ms/step "
            f"{bw:8.0f} GB/s")
    fp64 = None
    errors = None
    if not args.no_fp64:
        # same step count as the mixed run, so the fields can be compared
        fp64 = run_gpu("fp64", n, args.steps, args.warmup, find_dominant=True, tiles=tiles,
                       case=args.case)
        log(f"[bench] fp64: {fp64['ms']:.3f} ms/step, {fp64['mpts']:.1f} Mpoint-steps/s")
        errors = error_report(mixed["stats"], fp64["stats"])
        errors["after_steps"] = mixed["total_steps"]
        log(f"[bench] mixed vs fp64 after {mixed['total_steps']} steps: {errors}")
    live = {r["name"]: r for r in mixed["live"]}
    dom = live.get(mixed["dominant"])
    roof = None
    if dom and dom["total_ms"] > 0:
        # achieved: the reference data flow's bytes of the launches (each
        # transformed field one full read + one full write, half spectra at
        # (ny+2)/ny of a field) over their HIP-event time; achieved_distinct:
        # the bytes the kernel itself must move (pruned spectra, fused loads)
        sec = dom["total_ms"] * 1e-3
        ach = dom["ref_bytes"] / sec / 1e9
        ach_d = dom["bytes"] / sec / 1e9
        roof = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 4), traffic=measured_traffic(mixed["dominant"]),
                    traffic_source="profiles/hbm_traffic.json", kernel=mixed["dominant"],
                    launches=dom["launches"],
                    avg_launch_us=round(dom["total_ms"] * 1e3 / dom["launches"], 2),
                    bytes_per_launch=dom["ref_bytes"] / dom["launches"],
                    achieved_distinct=round(ach_d, 1), frac_distinct=round(ach_d / HBM_PEAK_GBS, 4),
                    distinct_bytes_per_launch=dom["bytes"] / dom["launches"])
    ens = srv = None
    if not args.no_ensemble:
        from alfa_amd import ensemble as E
        specs = E.ensemble_specs(args.ensemble_members, steps=args.ensemble_steps, seed=0)
        # shared launches: one launch per kernel for all members of a grid
        # size, the whole ensemble step one graph replay
        ens = E.run_ensemble_grouped(specs)
        torch.cuda.empty_cache()
        # beside it, the members stepped one by one (a graph per member step
        # on 4 streams)
        solo = E.run_ensemble(E.ensemble_specs(args.ensemble_members, steps=args.ensemble_steps,
                                               seed=0))
        ens["per_member_launches"] = {k: solo[k] for k in ("mpoint_steps_per_s", "wall_s",
                                                           "streams")}
        torch.cuda.empty_cache()
        ens["workload"] = (f"{args.ensemble_members} convective members, grids from "
                           f"48/64/80/96^3, seeded Ug and surface flux, mixed, "
                           f"{args.ensemble_steps} steps; members of one grid size share every "
                           f"launch (alfa_group), the ensemble step is one hipGraph")
        log(f"[bench] ensemble: {ens}")
        reqs = E.serving_requests(args.serve_requests, rate=2.0, steps=tuple(args.serve_steps),
                                  seed=0)
        # admitted requests of one configuration shape share launches
        # (regrouped between rounds); beside it, every request on its own
        # graph replays
        server = E.Server(hbm_budget_bytes=96 << 30, precision="mixed")
        srv = server.serve(reqs)
        torch.cuda.empty_cache()
        if not args.no_serve_compare:
            solo = E.Server(hbm_budget_bytes=96 << 30, precision="mixed", grouped=False).serve(reqs)
            srv["per_request_launches"] = {k: solo[k] for k in ("mpoint_steps_per_s", "p50_s",
                                                                "p99_s", "wall_s")}
            torch.cuda.empty_cache()
        srv["workload"] = (f"Poisson stream (2/s) of {args.serve_requests} requests, grids from "
                           f"32/64/128/192^3, neutral/convective/stable with seeded forcing, "
                           f"{args.serve_steps[0]}-{args.serve_steps[1]} steps each, mixed, "
                           f"96 GiB HBM budget; admitted requests of one shape share launches")
        log(f"[bench] serving: {srv}")
    developed = None
    if not args.no_developed:
        # the coefficient error at a developed state, against the variability
        # of an FP64 run perturbed at float rounding
        from alfa_amd import stats as S
        developed = S.developed_state_comparison(n=64, hours=1.0, avg_min=30.0)
        log(f"[bench] developed state: {developed}")
    copy_probe(args.copy_probe)
    cpu = None if args.no_cpu else cpu_baseline()
    # the committed counter summary covers this workload only
    profiled_workload = n == 256 and args.case == "stable" and not args.no_fusion
    out = {
        "metric": f"Mpoint-steps/s at {n}^3 (FP32-mixed vs FP64)",
        "value": round(mixed["mpts"], 2),
        "unit": "Mpoint-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 fields, f64 reductions/pressure/beta (mixed)",
        "data": (f"synthetic: seeded {CASE_TEXT[args.case][0]} "
                 f"(cases.initial_fields('{args.case}'))"),
        "config": {"workload": f"{CASE_TEXT[args.case][0]} {n}^3, {CASE_TEXT[args.case][1]}, "
                               f"Thomas pressure, mixed precision",
                   "grid": [n, n, n], "optSgs": CASE_TEXT[args.case][2],
                   "FGR": CASE_TEXT[args.case][3], "precision": "mixed"},
        # the whole step's HBM traffic as measured (counter bytes), at the
        # measured rate, against peak
        "hbm_step": step_hbm(mixed["mpts"], "mixed") if profiled_workload else None,
        # the reference's data flow (567 field passes per point-step, each
        # FFT pass one read and one write) at the measured rate: an
        # equivalent bandwidth, not traffic (it counts the passes that fusion
        # and pruning remove, so it can exceed 1)
        "reference_dataflow_frac_step": round(mixed["mpts"] * 1e6 * B_ALG["mixed"] /
                                              (HBM_PEAK_GBS * 1e9), 4),
        "b_alg_bytes_per_point_step": B_ALG["mixed"],
        "fp64": None if fp64 is None else {
            "value": round(fp64["mpts"], 2), "ms_per_step": round(fp64["ms"], 4),
            "hbm_step": step_hbm(fp64["mpts"], "fp64") if profiled_workload else None,
            "reference_dataflow_frac_step": round(fp64["mpts"] * 1e6 * B_ALG["fp64"] /
                                                  (HBM_PEAK_GBS * 1e9), 4)},
        "mixed_over_fp64": None if fp64 is None else round(mixed["mpts"] / fp64["mpts"], 3),
        # threads per workgroup of the y R2C, x and y C2R kernels, as the
        # plan's autotuner chose them after the warmup
        "fft_tiles": {"mixed": mixed["tiles"], "fp64": None if fp64 is None else fp64["tiles"]},
        # the mixed step's LAD procedures: level-1 test-filter inverse
        # transforms fused with their contractions (FP64 runs them separately)
        "fused_lad_contraction": not args.no_fusion,
        # the FP64 GPU path is held to the FP64 oracle at <= 1e-11 by the parity
        # tests; at 256^3 it stands in for the oracle, which would take hours
        "error_mixed_vs_fp64": errors,
        "error_mixed_vs_fp64_developed": developed,
        "roofline": roof,
        "ensemble": ens,
        "serving": srv,
        "cpu_baseline": cpu,
        "reference_a100": A100_REF,
    }
    if args.kernels_json:
        d = os.path.dirname(os.path.abspath(args.kernels_json))
        os.makedirs(d, exist_ok=True)
        with open(args.kernels_json, "w") as f:
            json.dump(dict(mixed=mixed["table"], fp64=None if fp64 is None else fp64["table"],
                           mixed_live=mixed["live"]), f, indent=1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
