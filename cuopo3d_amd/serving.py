"""Continuous-batching request loop for OPO solves.

Requests (grid, parameters, initial fields, number of round trips) arrive
over time.  Each arrival is submitted to a resident opo_batch; every
``step`` advances all admitted solves by one round trip, and solves that
arrive while others run are admitted at the next round-trip boundary once
the HBM budget has room (admission by footprint, in arrival order).
Finished solves are collected as they complete.  Latency is measured from
arrival to the moment the result (per-trip powers and fields) is on the
host.

The reference has no serving mode: its only multi-point driver is a bash
loop starting one process per (power, waist) point (src/cuOPO3D.sh:56-86).
"""
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import Batch, configs
from .capi import OPO_FP32


@dataclass
class Request:
    arrival: float          # seconds after the start of the run
    grid: object
    params: object
    fields: tuple
    n_trips: int
    tag: object = None
    ticket: int = None
    done_at: float = None
    power: np.ndarray = None
    out_fields: list = None
    error: str = None

    @property
    def cell_steps(self):
        g = self.grid
        return g.nx * g.ny * g.nt * self.params.n_slices * self.n_trips


@dataclass
class ServeReport:
    n: int
    makespan_s: float
    cell_steps: float
    p50_s: float
    p99_s: float
    mean_s: float
    steps: int
    max_resident: int
    latencies: list = field(default_factory=list)
    rebuilds: int = 0            # re-captures of the trip graph (resident set changed)
    rebuilds_in_place: int = 0   # of which updated the instantiated graph in place
    rebuild_s: float = 0.0       # host seconds they took

    @property
    def rate(self):
        return self.cell_steps / self.makespan_s if self.makespan_s > 0 else 0.0

    def as_dict(self):
        return {"requests": self.n, "makespan_s": self.makespan_s,
                "cell_steps_per_s": self.rate, "p50_latency_s": self.p50_s,
                "p99_latency_s": self.p99_s, "mean_latency_s": self.mean_s,
                "batch_steps": self.steps, "max_resident": self.max_resident,
                "rebuilds": self.rebuilds, "rebuilds_in_place": self.rebuilds_in_place,
                "rebuild_ms_mean": 1e3 * self.rebuild_s / self.rebuilds if self.rebuilds else 0.0}


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return float("nan")
    k = (len(xs) - 1) * q / 100.0
    lo, hi = math.floor(k), math.ceil(k)
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def serve(requests, budget_bytes, precision=OPO_FP32, keep_fields=True, clock=time.perf_counter):
    """Run the requests through one continuously batched GPU queue."""
    batch = Batch(budget_bytes, precision)
    pending = sorted(requests, key=lambda r: r.arrival)
    outstanding = {}
    t0 = clock()
    steps = 0
    max_res = 0
    i = 0
    while i < len(pending) or outstanding:
        now = clock() - t0
        while i < len(pending) and pending[i].arrival <= now:
            r = pending[i]
            i += 1
            try:
                r.ticket = batch.submit(r.grid, r.params, *r.fields, r.n_trips)
            except RuntimeError as e:   # rejected (grid, budget, parameters): this request only
                r.error = str(e)
                r.done_at = clock() - t0
                continue
            outstanding[r.ticket] = r
        res, queued, _ = batch.stats()
        if res == 0 and queued == 0:
            if i < len(pending):   # idle until the next arrival
                time.sleep(max(0.0, min(0.01, pending[i].arrival - (clock() - t0))))
            continue
        try:
            batch.step()
        except RuntimeError as e:
            # the batch itself failed: every request not yet finished fails with it
            for r in list(outstanding.values()) + pending[i:]:
                r.error = f"batch step failed: {e}"
                r.done_at = clock() - t0
            outstanding.clear()
            i = len(pending)
            break
        steps += 1
        max_res = max(max_res, batch.stats()[0])
        for t in list(outstanding):
            r = outstanding[t]
            try:
                got = batch.poll(t, want_fields=keep_fields)
            except RuntimeError as e:
                r.error = str(e)
                r.done_at = clock() - t0
                batch.meta.pop(t, None)
                del outstanding[t]
                continue
            if got is not None:
                r.power, r.out_fields = got
                r.done_at = clock() - t0
                del outstanding[t]
    makespan = clock() - t0
    lat = [r.done_at - r.arrival for r in requests]
    rebuilds = batch.rebuilds()
    batch.close()
    return ServeReport(n=len(requests), makespan_s=makespan,
                       cell_steps=float(sum(r.cell_steps for r in requests if r.error is None)),
                       p50_s=percentile(lat, 50), p99_s=percentile(lat, 99),
                       mean_s=float(np.mean(lat)) if lat else float("nan"), steps=steps,
                       max_resident=max_res, latencies=lat, rebuilds=rebuilds[0],
                       rebuilds_in_place=rebuilds[1], rebuild_s=rebuilds[2])


def poisson_requests(n, rate_per_s, seed=5, grids=configs.SERVING_GRIDS, trips=(50, 300),
                     precision=OPO_FP32):
    """Requests with exponential inter-arrival gaps, grids uniform over
    ``grids`` ((nxy, nt) pairs), trips uniform in ``trips``, physics drawn as
    in the batched sweep."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    draws = configs.sweep_draws(n_solves=n, seed=seed + 1)
    for k in range(n):
        t += rng.exponential(1.0 / rate_per_s)
        nxy, nt = grids[rng.integers(len(grids))]
        n_trips = int(rng.integers(trips[0], trips[1] + 1))
        d = dict(draws[k])
        d["nxy"], d["nt"] = nxy, nt
        g, p, ph, f = configs.sweep_solve(d, n_round_trips=n_trips, precision=precision)
        out.append(Request(arrival=t, grid=g, params=p, fields=f, n_trips=n_trips, tag=k))
    return out


def main(argv=None):
    """Serving load over several Poisson arrival rates; one JSON line per rate.

    python -m cuopo3d_amd.serving --rates 2,5,10,20,50 --requests 16
    """
    import argparse
    import json
    from . import autotune_grids
    ap = argparse.ArgumentParser(description="continuous-batching serving load")
    ap.add_argument("--rates", default="2,5,10,20,50", help="mean arrivals per second")
    ap.add_argument("--requests", type=int, default=16)
    ap.add_argument("--budget-gb", type=float, default=128.0)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--no-tune", action="store_true")
    args = ap.parse_args(argv)
    if not args.no_tune:
        autotune_grids([(nxy, nxy, nt) for nxy, nt in configs.SERVING_GRIDS], reps=3)
    for rate in [float(x) for x in args.rates.split(",")]:
        reqs = poisson_requests(args.requests, rate_per_s=rate, seed=args.seed)
        rep = serve(reqs, budget_bytes=args.budget_gb * (1 << 30), keep_fields=False)
        d = rep.as_dict()
        d["arrival_rate_per_s"] = rate
        d["last_arrival_s"] = max(r.arrival for r in reqs)
        d["errors"] = sum(1 for r in reqs if r.error)
        print(json.dumps(d), flush=True)
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())
