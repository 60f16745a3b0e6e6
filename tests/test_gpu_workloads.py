"""The batched sweep and the serving mix at their own grid sizes.

Sweep: the 64 solves of the seed-3 draw (grids 32^2x128 ... 128^2x512,
including 96^2x512 and 128^2x512), batched in one resident set for 2 round
trips.  Every solve's fields and per-trip powers must equal, bitwise, the
same solve run alone through the single-solve API, and the smallest and
largest grids are checked against the FP64 oracle (reference trip
sequence, Solver3D.cuh:963-1023) to 1e-5 relative L2 and 1e-6 per-trip
signal power.  This replaces the reference's one-process-per-point sweep
(src/cuOPO3D.sh:56-86).

Serving: a Poisson request mix over {32^2x128, 64^2x256, 96^2x512,
128^2x1024} with 2-4 round trips of 2 slices each, admitted by HBM budget
while others run.  Each result equals the request run alone, bitwise, and
the 128^2x1024 and 32^2x128 requests match the oracle.
"""
import numpy as np
import pytest

import cuopo3d_amd as co
from cuopo3d_amd import configs, serving
from oracle.pyoracle import Oracle

from helpers import rel_l2, batch_run

pytestmark = pytest.mark.gpu

GB = 1 << 30


def _alone(g, p, f, trips):
    with co.Solver(g, p, device=0) as s:
        s.set_fields(*f)
        pw = s.runOPO(trips)
        return s.get_fields(), pw


def _oracle_check(g, p, f, trips, fields, power, label):
    o = Oracle(g, p)
    o.set_fields(*f)
    pr = o.run_round_trips(trips)
    ref = o.get_fields()
    errs = [rel_l2(a, b) for a, b in zip(fields, ref)]
    perr = float(np.max(np.abs(power[:, 1] - pr[:, 1]) / np.abs(pr[:, 1])))
    print(label, (g.nx, g.ny, g.nt), "rel L2 vs oracle", errs, "signal power", perr, flush=True)
    assert max(errs) < 1e-5, (label, errs)
    assert perr < 1e-6, (label, perr)


def test_sweep_64_solves_batched(require_gpu):
    trips = 2
    draws = configs.sweep_draws(64, seed=3)
    solves = [configs.sweep_solve(d, n_round_trips=trips) for d in draws]
    grids = {(g.nx, g.nt) for (g, _, _, _) in solves}
    assert (96, 512) in grids and (128, 512) in grids and (32, 128) in grids
    print("sweep: 64 solves over", sorted(grids), flush=True)
    results, b = batch_run(solves, trips, budget=128 * GB)
    b.close()
    for k, ((g, p, ph, f), (pb, fb)) in enumerate(zip(solves, results)):
        fa, pa = _alone(g, p, f, trips)
        for a, c in zip(fa, fb):
            assert np.array_equal(a, c), (k, g.nx, g.nt)
        assert np.array_equal(pa, pb), (k, g.nx, g.nt)
    print("sweep: 64 batched solves == alone, bitwise", flush=True)
    cells = [g.nx * g.ny * g.nt for (g, _, _, _) in solves]
    for k, label in ((int(np.argmin(cells)), "smallest"), (int(np.argmax(cells)), "largest")):
        g, p, ph, f = solves[k]
        pb, fb = results[k]
        _oracle_check(g, p, f, trips, fb, pb, "sweep " + label)


def test_serving_mix_with_large_grids(require_gpu):
    reqs = serving.poisson_requests(8, rate_per_s=50.0, seed=2, grids=configs.SERVING_GRIDS,
                                    trips=(2, 4))
    for r in reqs:
        r.params.n_slices = 2
    grids = [(r.grid.nx, r.grid.nt) for r in reqs]
    assert (96, 512) in grids and (128, 1024) in grids
    fp = max(co.Batch(GB).footprint(r.grid, r.n_trips) for r in reqs)
    rep = serving.serve(reqs, budget_bytes=int(3.5 * fp))   # at most three resident at once
    print(rep.as_dict(), flush=True)
    assert all(r.error is None for r in reqs), [r.error for r in reqs]
    assert rep.max_resident > 1   # requests overlapped on the device
    for r in reqs:
        fa, pa = _alone(r.grid, r.params, r.fields, r.n_trips)
        for a, c in zip(fa, r.out_fields):
            assert np.array_equal(a, c), (r.grid.nx, r.grid.nt)
        assert np.array_equal(pa, r.power)
    print("serving: every request == alone, bitwise", flush=True)
    for want in ((128, 1024), (32, 128)):
        r = next(r for r in reqs if (r.grid.nx, r.grid.nt) == want)
        _oracle_check(r.grid, r.params, r.fields, r.n_trips, r.out_fields, r.power, "serving")
