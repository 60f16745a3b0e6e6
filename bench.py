"""Benchmark of the split-step path on one MI355X.

Workload (one "step" = one round trip): the singly-resonant cavity solve,
64x64x256 complex64 fields, MgO:PPLN 5 mm, NZ = 100 (100 slices per trip),
Rs = 0.98, idler not resonant, GDD compensation, free-space feedback; the
fields start from a focused 2 ps pump and seeded noise and stay resident in
HBM for the whole timed region.  value = cell-steps/s (cells x slices x
round trips / wall time).

The roofline entry is for the dominant kernel, timed with HIP events on the
solve's stream (opo_profile_trips) over a second run of the same trips; its
algorithmic bytes are 16 B per cell per field touched (8 read + 8 written).
The CPU baseline is the FP64 oracle on the host cores over the first round
trips of the same workload.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "cell-steps/s & %HBM roofline; p50/p99 latency of batched OPO sweeps"
HBM_PEAK_GBS = 8000.0
VALU_PEAK_T = 256 * 64 * 2.4e9 / 1e12   # FP32 VALU lane operations, no packed math
# fields touched per launch of each stage kind
STAGE_FIELDS = {"nonlinear": 3, "time_op": 3, "xy_op": 3, "free_space": 2, "gdd": 2, "power": 3}
STAGE_RW = {"power": 1}  # the power reduction only reads


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-trips", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-trips", type=int, default=20)
    ap.add_argument("--no-plain", action="store_true", help="skip the rocFFT comparator")
    ap.add_argument("--sweep-trips", type=int, default=10, help="0 skips the batched sweep")
    ap.add_argument("--serve-requests", type=int, default=16, help="0 skips the serving run")
    ap.add_argument("--serve-rate", type=float, default=10.0)
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch variants")
    ap.add_argument("--tune-reps", type=int, default=5, help="timed stretches per tuning candidate")
    ap.add_argument("--horizon-trips", type=int, default=0,
                    help="also compare FP32 and FP64 device fields with the FP64 oracle over this "
                         "many round trips (checkpoints 1, 3, 10, 30, 100, 200, ...); off by default")
    ap.add_argument("--grid", type=int, nargs=2, default=(64, 256), metavar=("NXY", "NT"),
                    help="grid of the cavity workload (default 64 256); other grids are for "
                         "profiling and skip the CPU and comparator legs")
    ap.add_argument("--no-reference-grid", action="store_true",
                    help="skip the run on the reference's compiled-default 128x128x2048 grid")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 128 MiB device-to-device copy first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs one solve per GPU on one GPU; --gpus {args.gpus} is not "
              "supported (independent solves shard across GPUs by footprint, no collectives).",
              file=sys.stderr)
        sys.exit(2)

    import cuopo3d_amd as co
    from cuopo3d_amd import configs
    from cuopo3d_amd.capi import STAGES

    lib = co.load()
    if lib.opo_device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.calibrate_copy:
        # a plain device-to-device copy of known size (Api -> Ap of a
        # 128x128x1024 complex64 solve, 128 MiB): its FETCH_SIZE / WRITE_SIZE
        # counters scale those of the split-step kernels to bytes
        cg, cp, _, _ = configs.cavity_64(n=(128, 128, 1024), n_round_trips=1)
        cs = co.Solver(cg, cp, device=0)
        cs.set_fields(np.zeros((1024, 128, 128), np.complex64), None, None)
        cs.sync()
        cs.close()
    default_grid = tuple(args.grid) == (64, 256)
    if not default_grid:
        args.no_cpu = args.no_plain = True
        args.sweep_trips = args.serve_requests = 0
    n = (args.grid[0], args.grid[0], args.grid[1])
    grid, p, ph, fields = configs.cavity_64(n=n, n_round_trips=args.warmup + args.steps)
    cells = grid.nx * grid.ny * grid.nt
    S = p.n_slices
    s = co.Solver(grid, p, device=0)
    s.set_fields(*fields)
    tune_report = s.autotune(reps=args.tune_reps) if not args.no_tune else "not tuned"
    s.runOPO(args.warmup)
    s.sync()   # all of the solve's work is on its stream
    t0 = time.perf_counter()
    pw = s.runOPO(args.steps)
    s.sync()
    t1 = time.perf_counter()
    dt = t1 - t0
    value = cells * S * args.steps / dt

    # per-stage kernel time with HIP events on the solve's stream
    ms = (C.c_double * 6)()
    nl = (C.c_int * 6)()
    co.check(lib.opo_profile_trips(s.h, args.profile_trips, ms, nl), "opo_profile_trips")
    stages = {}
    for k, name in enumerate(STAGES):
        if nl[k] == 0:
            continue
        avg_ms = ms[k] / nl[k]
        alg = cells * STAGE_FIELDS[name] * 8 * (STAGE_RW.get(name, 2))
        stages[name] = {"launches": nl[k], "avg_us": avg_ms * 1e3, "total_ms": ms[k],
                        "alg_bytes": alg, "GBs": alg / (avg_ms * 1e-3) / 1e9}
    dom = max(stages, key=lambda k: stages[k]["total_ms"])
    d = stages[dom]
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(d["GBs"], 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(d["GBs"] / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(grid, dom)}
    gpu_kernel_ms = sum(v["total_ms"] for v in stages.values()) / args.profile_trips

    out = {
        "metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c64 (f32)",
        "data": "synthetic (focused Gaussian pump + seeded Philox noise)",
        "config": {"workload": f"singly-resonant cavity {grid.nx}x{grid.ny}x{grid.nt}, MgO:PPLN 5 mm, "
                               "100 slices/trip, window = round-trip time, Rs=0.98, idler not resonant "
                               "(multiplier 1), GDD, free space; step = one round trip",
                   "grid": [grid.nx, grid.ny, grid.nt], "slices": S},
        "roofline": roofline,
        "valu": valu_utilisation(grid, dom, d["avg_us"] * 1e-6),
        "stages": {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                   for k, v in stages.items()},
        "gpu_kernel_ms_per_trip": gpu_kernel_ms,
        # The same rate priced in the 5-pass accounting of a split-step step
        # (RK4 pair, forward xy, forward t with the propagator, inverse t,
        # inverse xy: 5 x 3 fields x 16 B = 240 B per cell-step): the HBM
        # rate an unfused schedule would need to match this throughput.
        "five_pass_equivalent": {"bytes_per_cell_step": 240, "GBs": round(value * 240 / 1e9, 1),
                                 "frac_of_peak": round(value * 240 / 1e9 / HBM_PEAK_GBS, 4)},
        "autotune": tune_report.strip().splitlines(),
        "signal_power_last_trip_W": float(pw[-1, 1]),
    }

    single_fields = {}
    if default_grid:
        out["single_pass"], single_fields = single_pass_leg(co, configs)
        out["idler_dumped"] = idler_dumped_leg(co, configs, args.steps)
        out["fp64_path"] = fp64_leg(co, configs, min(args.steps, 50))
        if not args.no_reference_grid:
            out["reference_default_grid"] = reference_grid_leg(co, configs, args.tune_reps)
    if not args.no_cpu:
        from oracle.pyoracle import Oracle, get_threads
        g2, p2, _, f2 = configs.cavity_64(n_round_trips=args.cpu_trips)
        o = Oracle(g2, p2)
        o.set_fields(*f2)
        c0 = time.perf_counter()
        pr = o.run_round_trips(args.cpu_trips)
        c1 = time.perf_counter()
        cpu_rate = cells * S * args.cpu_trips / (c1 - c0)
        # parity of the same trips on the GPU
        s2 = co.Solver(g2, p2, device=0)
        s2.set_fields(*f2)
        pg = s2.runOPO(args.cpu_trips)
        gf = s2.get_fields()
        rf = o.get_fields()
        errs = [float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(gf, rf)]
        perr = float(np.max(np.abs(pg[:, 1] - pr[:, 1]) / np.abs(pr[:, 1])))
        out["cpu_baseline"] = {"value": cpu_rate, "unit": "cell-steps/s", "cores": get_threads(),
                               "kind": "port",
                               "sample": f"FP64 oracle, first {args.cpu_trips} round trips of the "
                                         f"same 64x64x256 cavity workload ({c1 - c0:.1f} s)"}
        out["parity_vs_oracle"] = {"trips": args.cpu_trips, "field_rel_l2_p_s_i": errs,
                                   "signal_power_rel_max": perr}
        # the single-pass configs in full on the host, with the GPU's error; the
        # oracle starts from the FP64 set-up (the FP32 set-up is its rounding)
        from cuopo3d_amd.capi import OPO_FP64
        others = {}
        for (name, prec), gf in single_fields.items():
            if prec:
                continue
            make = configs.single_pass_32 if name.startswith("single") else configs.femto_64
            g3, p3, _, f3 = make(precision=OPO_FP64)
            o3 = Oracle(g3, p3)
            o3.set_fields(*f3)
            c2 = time.perf_counter()
            o3.ssfm()
            c3 = time.perf_counter()
            rf3 = o3.get_fields()
            entry = {"value": g3.nx * g3.ny * g3.nt * p3.n_slices / (c3 - c2), "unit": "cell-steps/s",
                     "seconds": c3 - c2, "cores": get_threads()}
            for (n2, q), fl in single_fields.items():
                if n2 == name:
                    entry["fp64_rel_l2_p_s_i" if q else "fp32_rel_l2_p_s_i"] = [
                        float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(fl, rf3)]
            others[name] = entry
        out["cpu_baseline"]["single_pass_configs"] = others
        # the smallest and largest grids of the sweep, 2 round trips each
        sweep_cpu = {}
        solves = [configs.sweep_solve(d, n_round_trips=2) for d in configs.sweep_draws(n_solves=64, seed=3)]
        cells_of = [g.nx * g.ny * g.nt for (g, _, _, _) in solves]
        for tag, k in (("smallest", int(np.argmin(cells_of))), ("largest", int(np.argmax(cells_of)))):
            g4, p4, _, f4 = solves[k]
            o4 = Oracle(g4, p4)
            o4.set_fields(*f4)
            c4 = time.perf_counter()
            pr4 = o4.run_round_trips(2)
            c5 = time.perf_counter()
            s4 = co.Solver(g4, p4, device=0)
            s4.set_fields(*f4)
            pg4 = s4.runOPO(2)
            e4 = [float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(s4.get_fields(), o4.get_fields())]
            s4.close()
            sweep_cpu[tag] = {"solve": k, "grid": [g4.nx, g4.ny, g4.nt], "trips": 2,
                              "value": cells_of[k] * p4.n_slices * 2 / (c5 - c4), "unit": "cell-steps/s",
                              "seconds": c5 - c4, "field_rel_l2_p_s_i": e4,
                              "signal_power_rel_max": float(np.max(np.abs(pg4[:, 1] - pr4[:, 1]) / np.abs(pr4[:, 1])))}
        out["cpu_baseline"]["sweep_grids"] = sweep_cpu
    if not args.no_plain:
        out["plain_rocfft"] = plain_leg()
        if "value" in out["plain_rocfft"]:
            out["plain_rocfft"]["fused_speedup"] = value / out["plain_rocfft"]["value"]
    if args.horizon_trips > 0:
        out["horizon"] = horizon_leg(co, configs, args.horizon_trips)
    if args.sweep_trips > 0:
        out["sweep"] = sweep_leg(co, configs, args.sweep_trips)
    if args.serve_requests > 0:
        out["serving"] = serving_leg(args.serve_requests, args.serve_rate)
    print(json.dumps(out))


def horizon_leg(co, configs, trips):
    """Error growth with the number of round trips: FP32 and FP64 device
    fields and per-trip powers against the FP64 oracle at checkpoints."""
    from oracle.pyoracle import Oracle
    from cuopo3d_amd.capi import OPO_FP64
    g, p32, _, f32 = configs.cavity_64(n_round_trips=trips)
    _, p64, _, _ = configs.cavity_64(n_round_trips=trips, precision=OPO_FP64)
    f64 = tuple(x.astype(np.complex128) for x in f32)
    o = Oracle(g, p64)
    o.set_fields(*f64)
    s32 = co.Solver(g, p32, device=0)
    s32.set_fields(*f32)
    s64 = co.Solver(g, p64, device=0)
    s64.set_fields(*f64)
    marks = [m for m in (1, 3, 10, 30, 100, 200, 500, 1000) if m < trips] + [trips]
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    res = {"checkpoints": marks, "fp32_field_rel_l2_p_s_i": [], "fp64_field_rel_l2_p_s_i": [],
           "fp32_signal_power_rel_max": [], "fp64_signal_power_rel_max": []}
    done, pw_o, pw_32, pw_64 = 0, [], [], []
    for m in marks:
        n = m - done
        pw_o.append(o.run_round_trips(n))
        pw_32.append(s32.runOPO(n))
        pw_64.append(s64.runOPO(n))
        done = m
        r = o.get_fields()
        res["fp32_field_rel_l2_p_s_i"].append([rel(a, b) for a, b in zip(s32.get_fields(), r)])
        res["fp64_field_rel_l2_p_s_i"].append([rel(a, b) for a, b in zip(s64.get_fields(), r)])
        po = np.concatenate(pw_o)[:, 1]
        res["fp32_signal_power_rel_max"].append(float(np.max(np.abs(np.concatenate(pw_32)[:, 1] - po) / po)))
        res["fp64_signal_power_rel_max"].append(float(np.max(np.abs(np.concatenate(pw_64)[:, 1] - po) / po)))
        print(f"horizon trip {m}: fp32 {res['fp32_field_rel_l2_p_s_i'][-1]}", file=sys.stderr, flush=True)
    s32.close()
    s64.close()
    return res


def single_pass_leg(co, configs):
    """The single-pass configs of the baseline plan on the GPU: the 32x32x64
    plain single pass (100 slices) and the femtosecond 64x64x1024 single
    pass (399 slices, also on the FP64 path).  Returns the JSON summary and
    the output fields (for the oracle comparison of the CPU leg)."""
    from cuopo3d_amd.capi import OPO_FP64
    summary, fields = {}, {}
    for name, make, reps, precs in (("single_pass_32x32x64", configs.single_pass_32, 20, (0,)),
                                    ("femto_64x64x1024", configs.femto_64, 3, (0, OPO_FP64))):
        for prec in precs:
            g, p, _, f = make(precision=prec)
            s = co.Solver(g, p, device=0)
            s.set_fields(*f)
            for line in s.autotune(reps=3).strip().splitlines():
                print(f"{name} autotune:", line.strip(), file=sys.stderr)
            ts = []
            for _ in range(reps + 1):   # the first pass warms up
                s.set_fields(*f)
                s.sync()
                t0 = time.perf_counter()
                s.runSinglePass()       # returns after the pass has finished on the GPU
                ts.append(time.perf_counter() - t0)
            dt = float(np.median(ts[1:]))
            fields[(name, prec)] = s.get_fields()
            s.close()
            leg = summary.setdefault(name, {"grid": [g.nx, g.ny, g.nt], "slices": p.n_slices})
            leg["fp64" if prec else "fp32"] = {"value": g.nx * g.ny * g.nt * p.n_slices / dt,
                                               "unit": "cell-steps/s", "ms_per_pass": dt * 1e3}
    return summary, fields


def reference_grid_leg(co, configs, tune_reps, trips=3):
    """The same cavity on the grid the reference compiles by default
    (NX = NY = 128, NT = 2048: DtypesConsts.cuh:30-54), tuned for it."""
    g, p, _, f = configs.cavity_64(n=(128, 128, 2048), n_round_trips=trips + 1)
    s = co.Solver(g, p, device=0)
    s.set_fields(*f)
    report = s.autotune(reps=max(1, min(tune_reps, 2)))
    s.runOPO(1)
    s.sync()
    t0 = time.perf_counter()
    s.runOPO(trips)
    s.sync()
    dt = time.perf_counter() - t0
    s.close()
    cells = g.nx * g.ny * g.nt
    return {"grid": [g.nx, g.ny, g.nt], "trips": trips, "value": cells * p.n_slices * trips / dt,
            "unit": "cell-steps/s", "ms_per_step": dt / trips * 1e3,
            "autotune": [ln[-12:] for ln in report.strip().splitlines()]}


def fp64_leg(co, configs, trips):
    """The same cavity on the all-FP64 device path (the long-horizon mode)."""
    from cuopo3d_amd.capi import OPO_FP64
    g, p, _, f = configs.cavity_64(n_round_trips=trips + 1, precision=OPO_FP64)
    s = co.Solver(g, p, device=0)
    s.set_fields(*f)
    s.autotune(reps=3)
    s.runOPO(1)
    s.sync()
    t0 = time.perf_counter()
    s.runOPO(trips)
    s.sync()
    dt = time.perf_counter() - t0
    s.close()
    cells = g.nx * g.ny * g.nt
    return {"trips": trips, "value": cells * p.n_slices * trips / dt, "unit": "cell-steps/s",
            "ms_per_step": dt / trips * 1e3}


def idler_dumped_leg(co, configs, trips):
    """The same cavity with the non-resonant idler removed at the mirror
    (multiplier 0) instead of the reference's multiplier 1."""
    g, p, _, f = configs.cavity_64(n_round_trips=trips + 1)
    p.fb_re[1], p.fb_im[1] = 0.0, 0.0
    s = co.Solver(g, p, device=0)
    s.set_fields(*f)
    s.runOPO(1)
    s.sync()
    t0 = time.perf_counter()
    pw = s.runOPO(trips)
    s.sync()
    dt = time.perf_counter() - t0
    s.close()
    return {"trips": trips, "ms_per_step": dt / trips * 1e3,
            "signal_power_last_trip_W": float(pw[-1, 1]), "idler_power_last_trip_W": float(pw[-1, 2])}


def _counter_summary(grid, key, stage):
    """Figure `key` of a stage's kernel from the committed counter summary
    of this grid, or None when no summary covers it."""
    import re
    path = os.path.join(ROOT, "profiles", f"counters_cavity_{grid.nx}x{grid.ny}x{grid.nt}.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(key + r":.*\b" + re.escape(stage) + r"=(\d+)", text)
    return int(m.group(1)) if m else None


def measured_traffic(grid, stage):
    """HBM bytes per launch of a stage's kernel from the committed counter
    summary (rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, factors from a
    128 MiB device copy), or None when no summary covers this grid."""
    return _counter_summary(grid, "traffic_bytes_per_launch", stage)


def valu_utilisation(grid, stage, avg_s):
    """The kernels are bound by FP32 vector arithmetic, not by HBM: VALU
    lane operations per cell (rocprofv3 SQ_INSTS_VALU x 64 / cells, from the
    committed counter summary) times the cells of a launch, over its mean
    duration, against 256 CUs x 64 lanes x 2.4 GHz (FP32 without packed
    math, one lane operation per lane per clock; an FMA counts once)."""
    ops = _counter_summary(grid, "valu_lane_ops_per_cell", stage)
    if ops is None or avg_s <= 0:
        return None
    cells = grid.nx * grid.ny * grid.nt
    achieved = ops * cells / avg_s
    return {"kernel": stage, "lane_ops_per_cell": ops, "achieved": round(achieved / 1e12, 2),
            "peak": VALU_PEAK_T, "unit": "T lane-ops/s", "frac": round(achieved / 1e12 / VALU_PEAK_T, 4)}


def plain_leg(trips=4):
    """rocFFT + separate elementwise kernels (torch) on the same cavity workload,
    in a child process (torch carries its own HIP runtime)."""
    import subprocess
    r = subprocess.run([sys.executable, "-m", "cuopo3d_amd.baseline_rocfft", "--trips", str(trips)],
                       cwd=ROOT, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        return {"error": r.stderr.strip().splitlines()[-1] if r.stderr.strip() else "failed"}
    return json.loads(r.stdout.strip().splitlines()[-1])


def sweep_leg(co, configs, trips, n_solves=64, plain_solves=8, plain_trips=1):
    """64 uneven cavity solves batched (one launch per stage per size class, trip as a
    graph) against the same 64 solves run one at a time through the single-solve
    API, and against the plain rocFFT path one solve at a time (first solves)."""
    draws = configs.sweep_draws(n_solves=n_solves, seed=3)
    solves = [configs.sweep_solve(d, n_round_trips=trips) for d in draws]
    for line in co.autotune_grids([(g.nx, g.ny, g.nt) for (g, _, _, _) in solves], reps=3):
        print("sweep autotune:", line.strip().replace("\n", " | "), file=sys.stderr)
    work = sum(g.nx * g.ny * g.nt * p.n_slices for (g, p, _, _) in solves)
    b = co.Batch(200 << 30)
    tickets = [b.submit(g, p, *f, trips) for (g, p, _, f) in solves]
    b.step()  # admission, graph capture, first trip (not timed)
    t0 = time.perf_counter()
    for _ in range(trips - 1):
        b.step()
    dt = time.perf_counter() - t0
    batched = work * (trips - 1) / dt
    for t in tickets:
        b.poll(t, want_fields=False)
    b.close()
    # the same 64 solves one at a time, each warmed up by one trip
    one_trips = 2
    tot = 0.0
    sub_rates = []
    for k, (g, p, _, f) in enumerate(solves):
        s = co.Solver(g, p, device=0)
        s.set_fields(*f)
        s.runOPO(1)
        t0 = time.perf_counter()
        s.runOPO(one_trips)
        s.sync()
        e = time.perf_counter() - t0
        tot += e
        if k < plain_solves:
            sub_rates.append((g.nx * g.ny * g.nt * p.n_slices * one_trips, e))
        s.close()
    one = work * one_trips / tot
    out = {"solves": n_solves, "trips_timed": trips - 1, "batched_cell_steps_per_s": batched,
           "one_at_a_time_cell_steps_per_s": one, "speedup_vs_fused_one_at_a_time": batched / one}
    # plain rocFFT path, one solve at a time, in a child process
    import subprocess
    r = subprocess.run([sys.executable, "-m", "cuopo3d_amd.baseline_rocfft", "--sweep-solves",
                        str(plain_solves), "--trips", str(plain_trips)],
                       cwd=ROOT, capture_output=True, text=True, timeout=900)
    if r.returncode == 0:
        pl = json.loads(r.stdout.strip().splitlines()[-1])
        fused_sub = sum(w for w, _ in sub_rates) / sum(e for _, e in sub_rates)
        out["plain_rocfft_one_at_a_time_cell_steps_per_s"] = pl["value"]
        out["plain_rocfft_sample"] = f"first {plain_solves} solves of the sweep, {plain_trips} trip each"
        # the plain rate on the sample, scaled to the whole set by the fused one-at-a-time
        # rates of the same sample and of the whole set (the sizes differ)
        out["plain_rocfft_whole_set_estimate"] = pl["value"] * one / fused_sub
        out["speedup_vs_plain_one_at_a_time"] = batched / out["plain_rocfft_whole_set_estimate"]
    else:
        out["plain_rocfft_error"] = r.stderr.strip().splitlines()[-1] if r.stderr.strip() else "failed"
    return out


def serving_leg(n, rate):
    import cuopo3d_amd as co
    from cuopo3d_amd import serving, configs
    for line in co.autotune_grids([(nxy, nxy, nt) for nxy, nt in configs.SERVING_GRIDS], reps=3):
        print("serving autotune:", line.strip().replace("\n", " | "), file=sys.stderr)
    reqs = serving.poisson_requests(n, rate_per_s=rate, seed=5)
    rep = serving.serve(reqs, budget_bytes=128 << 30, keep_fields=False)
    d = rep.as_dict()
    d["arrival_rate_per_s"] = rate
    d["errors"] = sum(1 for r in reqs if r.error)
    return d


if __name__ == "__main__":
    main()
