"""The continuous-batching request loop on the GPU: requests that arrive
while others run are admitted as budget frees up, every request completes,
and each result equals the same solve run alone."""
import numpy as np
import pytest

import cuopo3d_amd as co
from cuopo3d_amd import serving

pytestmark = pytest.mark.gpu

GB = 1 << 30


def test_serving_results_equal_alone(require_gpu):
    reqs = serving.poisson_requests(6, rate_per_s=200.0, seed=11,
                                    grids=((32, 128), (64, 64), (32, 256)), trips=(2, 5))
    fp = max(co.Batch(GB).footprint(r.grid) for r in reqs)
    rep = serving.serve(reqs, budget_bytes=int(3.5 * fp))
    print(rep.as_dict(), flush=True)
    assert all(r.error is None for r in reqs)
    assert rep.max_resident <= 3
    assert rep.p99_s >= rep.p50_s > 0
    for r in reqs:
        s = co.Solver(r.grid, r.params, device=0)
        s.set_fields(*r.fields)
        s.runOPO(r.n_trips)
        for a, b in zip(s.get_fields(), r.out_fields):
            assert np.array_equal(a, b)
        assert r.power.shape == (r.n_trips, 3) and np.all(np.isfinite(r.power))
