"""Benchmark of the split-step path on one MI355X.

Headline workload (one "step" = one round trip of every solve): the batched
parameter sweep, 64 singly-resonant cavity solves of the seed-3 draw (grids
32x32x128 ... 128x128x512, complex64, MgO:PPLN 5 mm, NZ = 100 so 100
slices per trip; pump power, waist, detuning and mirror reflectivity drawn
per solve), all resident in HBM in one batch, one launch per split-step
stage and size class, the round trip captured as a HIP graph.  It replaces
the reference's one-process-per-point sweep (src/cuOPO3D.sh:56-86).
value = cell-steps/s (cells x slices x trips of all 64 solves / wall time).

Kernel times come from HIP events recorded inside the replayed trip graph
around every launch, over the timed region (opo_batch_profile); the
roofline entry is the kernel with the most time, its algorithmic bytes
16 B per cell per field read and written over the cells of its launch.
The CPU baseline is the FP64 oracle on the host cores over the smallest and
largest grids of the same sweep, 2 round trips each.
"""
import argparse
import ctypes as C
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "cell-steps/s & %HBM roofline; p50/p99 latency of batched OPO sweeps"
HBM_PEAK_GBS = 8000.0
VALU_PEAK_T = 256 * 64 * 2.4e9 / 1e12   # FP32 VALU lane operations, no packed math
# fields touched per launch of each stage kind (single-solve profiling mode)
STAGE_FIELDS = {"nonlinear": 3, "time_op": 3, "xy_op": 3, "free_space": 2, "gdd": 2, "power": 3}
STAGE_RW = {"power": 1}  # the power reduction only reads
# kernel of each batched stage kind
BATCH_KERNEL = {"time_op": "kb_time", "gdd": "kb_time (trip end)", "xy_op": "kb_xy",
                "free_space": "kb_xy (free space)", "power": "kb_power"}
NT_LIST = (16, 32, 64, 128, 256, 512, 1024, 2048, 4096)
NX_LIST = (16, 32, 64, 96, 128, 256, 512)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed round trips of the batched sweep")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--headline-only", action="store_true",
                    help="only the batched sweep (for profiler runs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the rocFFT comparator")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch variants")
    ap.add_argument("--tune-reps", type=int, default=3, help="timed stretches per tuning candidate")
    ap.add_argument("--tune", choices=("batch", "grids"), default="batch",
                    help="tune the sweep's variants on the batch's own trip, or each grid alone")
    ap.add_argument("--variants-out", help="write the launch variants in use to this JSON file")
    ap.add_argument("--variants-in", help="set the launch variants from this JSON file (no tuning)")
    ap.add_argument("--variants", help="set launch variants inline instead of tuning, e.g. t512=7,xy128=7")
    ap.add_argument("--classes", choices=("both", "serial", "concurrent"), default="both",
                    help="size classes of a stage one after another, as parallel graph branches, or "
                         "measure both and time the faster (default)")
    ap.add_argument("--serve-overload", action="store_true",
                    help="also serve 16 requests at 10/s with 50-300 trips (about 50 s)")
    ap.add_argument("--horizon-trips", type=int, default=0,
                    help="also compare FP32 and FP64 device fields with the FP64 oracle over this "
                         "many round trips of the 64x64x256 cavity (checkpoints); off by default")
    ap.add_argument("--cavity-grid", type=int, nargs=2, metavar=("NXY", "NT"),
                    help="instead of the sweep, time one cavity solve on this grid (profiling)")
    ap.add_argument("--fp64", action="store_true", help="with --cavity-grid: the all-FP64 device path")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 128 MiB device-to-device copy first (counter calibration)")
    args = ap.parse_args()
    if args.variants:
        args.variants_in = parse_variants(args.variants)
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(independent solves shard across GPUs by footprint, no collectives).", file=sys.stderr)
        sys.exit(2)

    import cuopo3d_amd as co
    from cuopo3d_amd import configs

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
    if args.cavity_grid:
        print(json.dumps(cavity_profile_mode(co, configs, lib, args)))
        return

    t_start = time.perf_counter()

    def leg(name, fn, *a):
        t0 = time.perf_counter()
        log(f"bench: {name} ...")
        r = fn(*a)
        log(f"bench: {name} done in {time.perf_counter() - t0:.1f} s ({time.perf_counter() - t_start:.0f} s)")
        return r

    out = leg("batched sweep (headline)", sweep_headline, co, configs, lib, args)
    log(f"bench: headline {out['value']:.4g} cell-steps/s, roofline {out['roofline']}")
    if not args.no_cpu:
        out["cpu_baseline"] = leg("CPU baseline (FP64 oracle)", cpu_leg, co, configs)
    if not args.headline_only:
        out["cavity_64x64x256"] = leg("cavity 64x64x256", cavity_leg, co, configs, args)
        out["single_pass"] = leg("single passes", single_pass_leg, co, configs)
        out["small_grid_sweep"] = leg("small-grid sweep", small_sweep_leg, co, configs)
        out["one_at_a_time"] = leg("sweep one at a time", one_at_a_time_leg, co, configs, out["value"])
        out["reference_default_grid"] = leg("128x128x2048", reference_grid_leg, co, configs, args.tune_reps)
        if not args.no_plain:
            out["plain_rocfft"] = leg("plain rocFFT", plain_leg, out["one_at_a_time"], out["value"])
        out["serving"] = leg("serving", serving_leg, co, configs, args.serve_overload)
        if args.horizon_trips > 0:
            out["horizon"] = leg("horizon", horizon_leg, co, configs, args.horizon_trips)
    print(json.dumps(out))


# ------------------------------------------------------------ variants
def variants_get(lib, prec=0):
    from cuopo3d_amd.capi import OPO_TUNE_TIME_PASS, OPO_TUNE_XY_PASS
    return {"t": {str(n): lib.opo_tune_get(prec, OPO_TUNE_TIME_PASS, n) for n in NT_LIST},
            "xy": {str(n): lib.opo_tune_get(prec, OPO_TUNE_XY_PASS, n) for n in NX_LIST}}


def parse_variants(spec):
    """'t512=7,xy128=7' -> {"t": {"512": 7}, "xy": {"128": 7}}"""
    out = {"t": {}, "xy": {}}
    for item in spec.split(","):
        m = re.fullmatch(r"(t|xy)(\d+)=(\d+)", item.strip())
        if not m:
            raise SystemExit(f"bench.py: bad --variants item {item!r} (want t<NT>=<v> or xy<NX>=<v>)")
        out[m.group(1)][m.group(2)] = int(m.group(3))
    return out


def load_variants(src):
    return src if isinstance(src, dict) else json.load(open(src))


def variants_tag(src):
    return "inline --variants" if isinstance(src, dict) else "variants set from " + os.path.basename(src)


def variants_set(lib, v, prec=0):
    from cuopo3d_amd.capi import OPO_TUNE_TIME_PASS, OPO_TUNE_XY_PASS, OPO_OK
    for axis, key in ((OPO_TUNE_TIME_PASS, "t"), (OPO_TUNE_XY_PASS, "xy")):
        for n, k in v[key].items():
            if lib.opo_tune_set(prec, axis, int(n), int(k)) != OPO_OK:
                raise RuntimeError(f"variant {k} does not exist for {key} {n}")


# ---------------------------------------------------------- headline
def sweep_headline(co, configs, lib, args, n_solves=64, trips=100):
    """The batched 64-solve sweep: value, per-kernel event times inside the
    replayed graph, roofline of the dominant kernel."""
    draws = configs.sweep_draws(n_solves=n_solves, seed=3)
    solves = [configs.sweep_solve(d, n_round_trips=trips) for d in draws]
    grids = sorted({(g.nx, g.ny, g.nt) for (g, _, _, _) in solves})
    tune = []
    log(f"sweep: {n_solves} solves set up")
    work = sum(g.nx * g.ny * g.nt * p.n_slices for (g, p, _, _) in solves)   # cell-steps per trip
    b = co.Batch(200 << 30)
    tickets = [b.submit(g, p, *f, trips) for (g, p, _, f) in solves]
    assert args.warmup + args.steps + 9 <= trips, "the timed trips must fit the solves' 100 trips"
    tune_s = 0.0
    if args.variants_in:
        variants_set(lib, load_variants(args.variants_in))
        tune = [variants_tag(args.variants_in)]
    elif not args.no_tune:
        t0 = time.perf_counter()
        if args.tune == "batch":   # admits the 64 solves, times each class's variants in the trip
            tune = [ln.strip() for ln in b.autotune(reps=args.tune_reps).strip().splitlines()]
        else:
            for rep in co.autotune_grids(grids, reps=args.tune_reps):
                tune += [ln.strip() for ln in rep.strip().splitlines()]
        tune_s = time.perf_counter() - t0
        log(f"sweep autotune ({args.tune}): {tune_s:.1f} s")
    variants = variants_get(lib)
    if args.variants_out:
        with open(args.variants_out, "w") as fh:
            json.dump(variants, fh)
    t0 = time.perf_counter()
    b.step()   # admission (if not tuned), descriptors, graph capture, first trip
    admit_s = time.perf_counter() - t0
    for _ in range(max(0, args.warmup - 1)):
        b.step()
    # a few trips without events, size classes of a stage one after another
    # and then as parallel branches of the trip graph; the faster form runs
    # the timed trips
    def rate3():
        t0 = time.perf_counter()
        for _ in range(3):
            b.step()
        return work * 3 / (time.perf_counter() - t0)
    serial_rate = concurrent_rate = None
    if args.classes != "concurrent":
        serial_rate = rate3()
    if args.classes != "serial":
        b.concurrent_classes(True)
        b.step()   # re-capture with the branches
        concurrent_rate = rate3()
    concurrent = args.classes == "concurrent" or (args.classes == "both" and concurrent_rate > serial_rate)
    if not concurrent:
        b.concurrent_classes(False)
    plain_rate = max(r for r in (serial_rate, concurrent_rate) if r is not None)
    log(f"sweep: serial classes {serial_rate}, concurrent classes {concurrent_rate} cell-steps/s")
    b.profile(True)
    b.step()   # re-capture with the events (not timed)
    b.profile(True)   # sums back to zero
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.step()
    dt = time.perf_counter() - t0
    value = work * args.steps / dt
    try:
        prof = b.profile_read()
        timing = "HIP events recorded as event nodes of the trip graph around every launch, over the timed trips"
        per_trip_div = args.steps
    except RuntimeError as e:
        # the runtime cannot time event nodes of a graph: time each kernel class
        # as a replayed graph of its launches of one trip, between two events
        log(f"sweep: in-graph event timing unavailable ({e}); timing each kernel class as its own graph")
        b.profile(False)
        b.step()
        prof = b.time_stages(reps=3)
        per_trip_div = 3
        timing = ("HIP events around a replayed graph of each kernel class's launches of one trip "
                  "(3 replays), right after the timed trips")
    resident = b.stats()[0]
    b.close()
    assert resident == n_solves
    stages = []
    for e in prof:
        if e["launches"] == 0:
            continue
        avg_s = e["ms"] / e["launches"] * 1e-3
        stages.append({"kernel": BATCH_KERNEL[e["stage"]], "stage": e["stage"], "class": e["size"],
                       "launches": e["launches"], "avg_us": round(avg_s * 1e6, 2),
                       "total_ms": round(e["ms"], 3), "alg_bytes": int(e["alg_bytes"]),
                       "GBs": round(e["alg_bytes"] / avg_s / 1e9, 1)})
    stages.sort(key=lambda x: -x["total_ms"])
    dom = stages[0]
    key = f"{dom['stage']}/{dom['class']}"
    roofline = {"bound": "hbm", "kernel": f"{dom['kernel']} {'NT' if dom['stage'] in ('time_op', 'gdd') else 'NX'}"
                                          f"={dom['class']}",
                "achieved": dom["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(dom["GBs"] / HBM_PEAK_GBS, 4),
                "traffic": counter_summary("counters_sweep_64.md", "traffic_bytes_per_launch", key)}
    kernel_ms_trip = sum(s["total_ms"] for s in stages) / per_trip_div
    cls_cells = {}
    for (g, p, _, _) in solves:
        cls_cells.setdefault(f"{g.nx}x{g.ny}x{g.nt}", 0)
        cls_cells[f"{g.nx}x{g.ny}x{g.nt}"] += 1
    return {
        "metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c64 (f32)",
        "data": "synthetic (focused Gaussian pump + seeded Philox noise, parameters drawn from seed 3)",
        "config": {"workload": "batched parameter sweep: 64 singly-resonant cavity solves of the seed-3 "
                               "draw, grids 32x32x128 .. 128x128x512, MgO:PPLN 5 mm, 100 slices per trip, "
                               "P 0.5-5 W, w0 30-80 um, detuning -pi..pi, R 0.9-0.99; one launch per "
                               "stage and size class, trip replayed as a HIP graph; step = one round "
                               "trip of all 64 solves",
                   "solves": n_solves, "grids": cls_cells, "cell_steps_per_trip": work},
        "roofline": roofline,
        "valu": valu_from_counters("counters_sweep_64.md", key, dom["alg_bytes"] / 48.0,
                                   dom["avg_us"] * 1e-6),
        "kernels": stages,
        "kernel_timing": timing,
        "kernel_ms_per_step": round(kernel_ms_trip, 3),
        "rate_without_events": plain_rate,
        "classes": {"serial_rate": serial_rate, "concurrent_rate": concurrent_rate,
                    "timed_with": "concurrent" if concurrent else "serial"},
        "first_step_s": round(admit_s, 3),
        "variants": variants,
        "autotune": {"mode": variants_tag(args.variants_in) if args.variants_in else ("none" if args.no_tune else args.tune),
                     "seconds": round(tune_s, 2), "report": tune},
    }


def counter_summary(fname, key, entry):
    """Integer figure `key` of kernel entry `entry` (e.g. time_op/512) from a
    committed counter summary under profiles/, or None."""
    try:
        text = open(os.path.join(ROOT, "profiles", fname)).read()
    except OSError:
        return None
    m = re.search(re.escape(key) + r":.*?\b" + re.escape(entry) + r"=(\d+)", text)
    return int(m.group(1)) if m else None


def valu_from_counters(fname, entry, cells, avg_s):
    """The kernels are bound by FP32 vector arithmetic as much as by HBM:
    VALU lane operations per cell (rocprofv3 SQ_INSTS_VALU x 64 / cells,
    from the committed counter summary) times the cells of a launch, over
    its mean duration, against 256 CUs x 64 lanes x 2.4 GHz (FP32 without
    packed math; an FMA counts once)."""
    ops = counter_summary(fname, "valu_lane_ops_per_cell", entry)
    if ops is None or avg_s <= 0:
        return None
    achieved = ops * cells / avg_s
    return {"kernel": entry, "lane_ops_per_cell": ops, "achieved": round(achieved / 1e12, 2),
            "peak": VALU_PEAK_T, "unit": "T lane-ops/s", "frac": round(achieved / 1e12 / VALU_PEAK_T, 4)}


# ------------------------------------------------------------- CPU leg
def cpu_leg(co, configs):
    """FP64 oracle on the host cores: the smallest and largest grids of the
    same 64-solve draw, 2 round trips each, with the GPU's error on them."""
    from oracle.pyoracle import Oracle, get_threads
    solves = [configs.sweep_solve(d, n_round_trips=2) for d in configs.sweep_draws(n_solves=64, seed=3)]
    cells_of = [g.nx * g.ny * g.nt for (g, _, _, _) in solves]
    work, secs, parts = 0.0, 0.0, {}
    for tag, k in (("smallest", int(np.argmin(cells_of))), ("largest", int(np.argmax(cells_of)))):
        g, p, _, f = solves[k]
        o = Oracle(g, p)
        o.set_fields(*f)
        c0 = time.perf_counter()
        pr = o.run_round_trips(2)
        c1 = time.perf_counter()
        s = co.Solver(g, p, device=0)
        s.set_fields(*f)
        pg = s.runOPO(2)
        errs = [float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(s.get_fields(), o.get_fields())]
        s.close()
        w = cells_of[k] * p.n_slices * 2
        work += w
        secs += c1 - c0
        parts[tag] = {"solve": k, "grid": [g.nx, g.ny, g.nt], "trips": 2, "seconds": round(c1 - c0, 3),
                      "value": w / (c1 - c0), "gpu_field_rel_l2_p_s_i": errs,
                      "gpu_signal_power_rel_max": float(np.max(np.abs(pg[:, 1] - pr[:, 1]) / np.abs(pr[:, 1])))}
    return {"value": work / secs, "unit": "cell-steps/s", "cores": get_threads(), "kind": "port",
            "sample": f"FP64 oracle (own mixed-radix FFT; FFTW absent), smallest and largest grids of the "
                      f"64-solve sweep, 2 round trips each ({secs:.1f} s)", "grids": parts}


# --------------------------------------------------------------- extras
def cavity_leg(co, configs, args, trips=20):
    """The 64x64x256 singly-resonant cavity alone (FP32, tuned) and on the
    all-FP64 device path, the long-horizon mode (its 200 trips hold 1e-5
    field and 1e-6 power against the oracle, tests/test_gpu_horizon.py)."""
    from cuopo3d_amd.capi import OPO_FP64
    out = {}
    for prec, name in ((0, "fp32"), (OPO_FP64, "fp64")):
        g, p, _, f = configs.cavity_64(n_round_trips=trips + 1, precision=prec)
        s = co.Solver(g, p, device=0)
        s.set_fields(*f)
        if not args.no_tune:
            s.autotune(reps=args.tune_reps)
        s.runOPO(1)
        s.sync()
        t0 = time.perf_counter()
        s.runOPO(trips)
        s.sync()
        dt = time.perf_counter() - t0
        s.close()
        out[name] = {"value": g.nx * g.ny * g.nt * p.n_slices * trips / dt, "unit": "cell-steps/s",
                     "ms_per_trip": dt / trips * 1e3, "trips": trips}
    out["fp64_over_fp32"] = out["fp64"]["value"] / out["fp32"]["value"]
    return out


def single_pass_leg(co, configs):
    """The single-pass workloads: the 32x32x64 plain single pass (100
    slices) and the femtosecond 64x64x1024 pass (399 slices, also on the
    FP64 path).  Their parity against the oracle is in the GPU tests."""
    from cuopo3d_amd.capi import OPO_FP64
    summary = {}
    for name, make, reps, precs in (("single_pass_32x32x64", configs.single_pass_32, 20, (0,)),
                                    ("femto_64x64x1024", configs.femto_64, 3, (0, OPO_FP64))):
        for prec in precs:
            g, p, _, f = make(precision=prec)
            s = co.Solver(g, p, device=0)
            s.set_fields(*f)
            s.autotune(reps=3)
            ts = []
            for _ in range(reps + 1):   # the first pass warms up
                s.set_fields(*f)
                s.sync()
                t0 = time.perf_counter()
                s.runSinglePass()       # returns after the pass has finished on the GPU
                ts.append(time.perf_counter() - t0)
            dt = float(np.median(ts[1:]))
            s.close()
            leg = summary.setdefault(name, {"grid": [g.nx, g.ny, g.nt], "slices": p.n_slices})
            leg["fp64" if prec else "fp32"] = {"value": g.nx * g.ny * g.nt * p.n_slices / dt,
                                               "unit": "cell-steps/s", "ms_per_pass": dt * 1e3}
    return summary


def _batched_rate(co, solves, trips, concurrent=False, tune=False):
    work = sum(g.nx * g.ny * g.nt * p.n_slices for (g, p, _, _) in solves)
    b = co.Batch(200 << 30)
    if concurrent:
        b.concurrent_classes(True)
    for (g, p, _, f) in solves:
        b.submit(g, p, *f, trips + 2)
    if tune:
        b.autotune(reps=3)
    b.step()
    b.step()
    t0 = time.perf_counter()
    for _ in range(trips):
        b.step()
    dt = time.perf_counter() - t0
    b.close()
    return work * trips / dt


def _one_at_a_time_rate(co, solves, trips):
    work, tot = 0.0, 0.0
    for (g, p, _, f) in solves:
        s = co.Solver(g, p, device=0)
        s.set_fields(*f)
        s.runOPO(1)
        s.sync()
        t0 = time.perf_counter()
        s.runOPO(trips)
        s.sync()
        tot += time.perf_counter() - t0
        work += g.nx * g.ny * g.nt * p.n_slices * trips
        s.close()
    return work / tot


def small_sweep_leg(co, configs, n_solves=48, trips=5):
    """A sweep of small grids (32x32x128 .. 64x64x256), where one solve
    alone fills a fraction of the GPU: batched against one at a time."""
    draws = configs.sweep_draws(n_solves=n_solves, seed=21)
    sizes = [(32, 128), (32, 256), (64, 128), (64, 256)]
    solves = []
    for k, d in enumerate(draws):
        d = dict(d)
        d["nxy"], d["nt"] = sizes[k % len(sizes)]
        solves.append(configs.sweep_solve(d, n_round_trips=100))
    serial = _batched_rate(co, solves, trips, tune=True)
    concurrent = _batched_rate(co, solves, trips, concurrent=True)
    batched = max(serial, concurrent)
    co.autotune_grids(sorted({(g.nx, g.ny, g.nt) for (g, _, _, _) in solves}), reps=3)
    one = _one_at_a_time_rate(co, solves, trips)
    return {"solves": n_solves, "grids": ["32x32x128", "32x32x256", "64x64x128", "64x64x256"],
            "trips_timed": trips, "batched_cell_steps_per_s": batched,
            "batched_serial_classes": serial, "batched_concurrent_classes": concurrent,
            "one_at_a_time_cell_steps_per_s": one, "speedup": batched / one}


def one_at_a_time_leg(co, configs, batched, trips=3):
    """The 64 sweep solves one after another through the single-solve API
    (fused kernels, trip graphs), against the batched headline."""
    solves = [configs.sweep_solve(d, n_round_trips=100) for d in configs.sweep_draws(n_solves=64, seed=3)]
    co.autotune_grids(sorted({(g.nx, g.ny, g.nt) for (g, _, _, _) in solves}), reps=3)
    one = _one_at_a_time_rate(co, solves, trips)
    return {"trips_timed": trips, "cell_steps_per_s": one, "batched_speedup": batched / one}


def reference_grid_leg(co, configs, tune_reps, trips=3):
    """The cavity on the grid the reference compiles by default
    (NX = NY = 128, NT = 2048: DtypesConsts.cuh:30-54), tuned for it, on the
    FP32 path and on the FP64 path (the mode that holds 1e-5 over long runs)."""
    from cuopo3d_amd.capi import OPO_FP64
    out = {}
    for name, prec in (("fp32", 0), ("fp64", OPO_FP64)):
        g, p, _, f = configs.cavity_64(n=(128, 128, 2048), n_round_trips=trips + 1, precision=prec)
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
        out[name] = {"value": cells * p.n_slices * trips / dt, "unit": "cell-steps/s",
                     "ms_per_step": dt / trips * 1e3,
                     "autotune": [ln[-12:] for ln in report.strip().splitlines()]}
        log(f"128x128x2048 {name}: {dt / trips * 1e3:.1f} ms per trip")
    out.update({"grid": [128, 128, 2048], "trips": trips, "value": out["fp32"]["value"],
                "unit": "cell-steps/s", "ms_per_step": out["fp32"]["ms_per_step"]})
    return out


def plain_leg(one_at_a_time, batched, solves=8, trips=1):
    """rocFFT + separate elementwise kernels (torch) one solve at a time on
    the first sweep solves, in a child process (torch carries its own HIP
    runtime); scaled to the whole set by the fused one-at-a-time rates."""
    import subprocess
    r = subprocess.run([sys.executable, "-m", "cuopo3d_amd.baseline_rocfft", "--sweep-solves",
                        str(solves), "--trips", str(trips)],
                       cwd=ROOT, capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        return {"error": r.stderr.strip().splitlines()[-1] if r.stderr.strip() else "failed"}
    pl = json.loads(r.stdout.strip().splitlines()[-1])
    pl["sample"] = f"first {solves} solves of the sweep, {trips} trip each"
    pl["batched_speedup_on_sample_rate"] = batched / pl["value"]
    return pl


def serving_leg(co, configs, overload):
    """Continuous batching of Poisson arrivals over the serving mix
    (32^2x128, 64^2x256, 96^2x512, 128^2x1024), 128 GB admission budget.
    Under capacity: 1 arrival/s with 5-30 trips (about 40 % of the GPU),
    so latency is service time; overloaded (optional): 10/s with 50-300
    trips, where latency is queue length."""
    from cuopo3d_amd import serving
    co.autotune_grids([(nxy, nxy, nt) for nxy, nt in configs.SERVING_GRIDS], reps=3)
    out = {}
    cases = [("under_capacity", 1.0, (5, 30))]
    if overload:
        cases.append(("overloaded", 10.0, (50, 300)))
    for name, rate, trips in cases:
        reqs = serving.poisson_requests(16, rate_per_s=rate, seed=5, trips=trips)
        rep = serving.serve(reqs, budget_bytes=128 << 30, keep_fields=False)
        d = rep.as_dict()
        d["arrival_rate_per_s"] = rate
        d["trips"] = list(trips)
        d["last_arrival_s"] = max(r.arrival for r in reqs)
        d["errors"] = sum(1 for r in reqs if r.error)
        out[name] = d
    return out


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
        log(f"horizon trip {m}: fp32 {res['fp32_field_rel_l2_p_s_i'][-1]}")
    s32.close()
    s64.close()
    return res


# ------------------------------------------------ single-cavity profiling
def cavity_profile_mode(co, configs, lib, args):
    """One cavity solve on the given grid, trips replayed as a graph, and
    per-stage kernel times with HIP events around each launch of a second
    run without graph capture (opo_profile_trips)."""
    from cuopo3d_amd.capi import STAGES
    n = (args.cavity_grid[0], args.cavity_grid[0], args.cavity_grid[1])
    prec = 1 if args.fp64 else 0
    grid, p, ph, fields = configs.cavity_64(n=n, n_round_trips=args.warmup + args.steps + 1, precision=prec)
    cells = grid.nx * grid.ny * grid.nt
    s = co.Solver(grid, p, device=0)
    s.set_fields(*fields)
    if args.variants_in:
        variants_set(lib, load_variants(args.variants_in), prec)
    elif not args.no_tune:
        s.autotune(reps=args.tune_reps)
    if args.variants_out:
        json.dump(variants_get(lib, prec), open(args.variants_out, "w"))
    s.runOPO(args.warmup)
    s.sync()
    t0 = time.perf_counter()
    s.runOPO(args.steps)
    s.sync()
    dt = time.perf_counter() - t0
    ms = (C.c_double * 6)()
    nl = (C.c_int * 6)()
    co.check(lib.opo_profile_trips(s.h, 5, ms, nl), "opo_profile_trips")
    stages = {}
    for k, name in enumerate(STAGES):
        if nl[k]:
            avg = ms[k] / nl[k]
            alg = cells * STAGE_FIELDS[name] * (16 if prec else 8) * STAGE_RW.get(name, 2)
            stages[name] = {"launches": nl[k], "avg_us": avg * 1e3, "alg_bytes": alg,
                            "GBs": alg / (avg * 1e-3) / 1e9}
    return {"workload": f"cavity {grid.nx}x{grid.ny}x{grid.nt}" + (" fp64" if prec else ""),
            "value": cells * p.n_slices * args.steps / dt,
            "unit": "cell-steps/s", "ms_per_trip": dt / args.steps * 1e3, "stages_unbatched_events": stages,
            "variants": variants_get(lib, prec)}


if __name__ == "__main__":
    main()
