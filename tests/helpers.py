"""Shared helpers of the test suite."""
import numpy as np

from cuopo3d_amd import configs
from cuopo3d_amd.capi import Grid, OPO_FP32, OPO_FP64


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / (nb if nb > 0 else 1.0))


def random_fields(grid, seed, amp=(1.0, 1.0, 1.0), precision=OPO_FP32):
    rng = np.random.default_rng(seed)
    shape = (grid.nt, grid.ny, grid.nx)
    dt = np.complex128 if precision == OPO_FP64 else np.complex64
    out = []
    for a in amp:
        z = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * a
        out.append(z.astype(dt))
    return out


def cavity_params(n, precision=OPO_FP32, **kw):
    """A cavity solve on grid n = (nx, ny, nt) with the 64^3-class physics."""
    grid = Grid(*n)
    ph = configs.physical(precision=precision, **kw)
    p = configs.make_params(grid, ph)
    return grid, p, ph


def batch_mix(n=4, trips=3, seed=3):
    """n cavity solves of uneven grids drawn like the sweep's parameters."""
    out = []
    draws = configs.sweep_draws(n_solves=16, seed=seed)
    sizes = [(32, 128), (64, 64), (96, 128), (32, 256)]
    for k in range(n):
        d = dict(draws[k])
        d["nxy"], d["nt"] = sizes[k % len(sizes)]
        out.append(configs.sweep_solve(d, n_round_trips=trips))
    return out


def batch_run(solves, trips, budget=8 << 30, concurrent=False):
    """Submit the solves to one batch, step until all finish; results in order."""
    import cuopo3d_amd as co
    b = co.Batch(budget)
    if concurrent:
        b.concurrent_classes(True)
    tickets = [b.submit(g, p, *f, trips) for (g, p, ph, f) in solves]
    results = {}
    for _ in range(10 * trips * len(solves) + 10):
        b.step()
        for t in tickets:
            if t not in results:
                r = b.poll(t)
                if r is not None:
                    results[t] = r
        if len(results) == len(tickets):
            break
    return [results[t] for t in tickets], b
