"""Launch variants and the autotuner (opo_autotune, opo_tune_get/set).

A variant changes only how lines are grouped into workgroups (tile width,
block size, persistent walk), never the arithmetic of a line, and every
multiply-add is an explicit fused operation (no compiler contraction), so
every variant must give bitwise the same fields; the tuner must leave the
solve's fields as it found them.
"""
import re

import numpy as np
import pytest

import cuopo3d_amd as co
from cuopo3d_amd import configs
from cuopo3d_amd.capi import OPO_FP32, OPO_OK, OPO_TUNE_TIME_PASS, OPO_TUNE_XY_PASS

from helpers import rel_l2, batch_mix, batch_run

pytestmark = pytest.mark.gpu

GRIDS = [(64, 64, 256), (32, 32, 1024), (128, 128, 64), (16, 16, 512), (16, 16, 2048), (16, 16, 4096),
         (256, 256, 32)]


@pytest.fixture
def restore_tuning():
    lib = co.load()
    saved = {}
    for nt in (16, 32, 64, 128, 256, 512, 1024, 2048, 4096):
        saved[(OPO_TUNE_TIME_PASS, nt)] = lib.opo_tune_get(OPO_FP32, OPO_TUNE_TIME_PASS, nt)
    for nx in (16, 32, 64, 96, 128, 256):
        saved[(OPO_TUNE_XY_PASS, nx)] = lib.opo_tune_get(OPO_FP32, OPO_TUNE_XY_PASS, nx)
    yield lib
    for (axis, n), v in saved.items():
        assert lib.opo_tune_set(OPO_FP32, axis, n, v) == OPO_OK


def _valid(lib, axis, n):
    """Variants that exist for size n."""
    keep = lib.opo_tune_get(OPO_FP32, axis, n)
    out = []
    for v in range(64):
        if lib.opo_tune_set(OPO_FP32, axis, n, v) == OPO_OK:
            out.append(v)
    lib.opo_tune_set(OPO_FP32, axis, n, keep)
    return out


def _run(grid, p, fields, trips=2):
    with co.Solver(grid, p, device=0) as s:
        s.set_fields(*fields)
        pw = s.runOPO(trips)
        return s.get_fields(), pw


@pytest.mark.parametrize("n", GRIDS, ids=lambda n: "x".join(map(str, n)))
def test_variants_give_the_same_fields(n, restore_tuning, require_gpu):
    lib = restore_tuning
    nx, ny, nt = n
    grid, p, _, fields = configs.cavity_64(n=n, n_round_trips=2)
    lib.opo_tune_set(OPO_FP32, OPO_TUNE_TIME_PASS, nt, 0)
    lib.opo_tune_set(OPO_FP32, OPO_TUNE_XY_PASS, nx, 0)
    ref, pref = _run(grid, p, fields)
    for axis, size in ((OPO_TUNE_TIME_PASS, nt), (OPO_TUNE_XY_PASS, nx)):
        for v in _valid(lib, axis, size):
            assert lib.opo_tune_set(OPO_FP32, axis, size, v) == OPO_OK
            try:
                got, pw = _run(grid, p, fields)
            except RuntimeError as e:  # a tile that does not divide NX is refused, not run
                assert "does not divide" in str(e), e
                continue
            errs = [rel_l2(a, b) for a, b in zip(got, ref)]
            same = all(np.array_equal(a, b) for a, b in zip(got, ref))
            print(f"{n} axis {axis} v{v}: bitwise {same}, rel L2 {max(errs):.1e}", flush=True)
            assert same, (axis, v, errs)
            assert np.array_equal(pw, pref), (axis, v)
        lib.opo_tune_set(OPO_FP32, axis, size, 0)


def test_autotune_keeps_fields_and_picks_a_variant(restore_tuning, require_gpu):
    lib = restore_tuning
    grid, p, _, fields = configs.cavity_64(n=(64, 64, 256), n_round_trips=3)
    ref, pref = _run(grid, p, fields, trips=3)
    with co.Solver(grid, p, device=0) as s:
        s.set_fields(*fields)
        s.runOPO(1)
        before = s.get_fields()
        report = s.autotune(reps=3)
        print(report, flush=True)
        assert "t pass NT=256" in report and "xy pass NX=64" in report, report
        after = s.get_fields()
        for a, b in zip(before, after):
            assert np.array_equal(a, b)
        # the run continues with the tuned variants and matches the untuned run
        s.runOPO(2)
        got = s.get_fields()
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    v = lib.opo_tune_get(OPO_FP32, OPO_TUNE_TIME_PASS, 256)
    assert v in _valid(lib, OPO_TUNE_TIME_PASS, 256)



def _persistent(lib, axis, n):
    """Valid variants of size n past the one-unit-per-workgroup ones (time:
    V >= 9, persistent; xy: V >= 3, persistent for NX >= 64, plane groups
    for NX <= 32)."""
    first = 9 if axis == OPO_TUNE_TIME_PASS else 3
    return [v for v in _valid(lib, axis, n) if v >= first]


def test_batch_with_persistent_variants(restore_tuning, require_gpu):
    """The batched passes on persistent launches give the fields of the
    one-item-per-workgroup launches, bitwise, for every size class."""
    lib = restore_tuning
    solves = batch_mix(4, trips=2)
    nts = sorted({g.nt for (g, _, _, _) in solves})
    nxs = sorted({g.nx for (g, _, _, _) in solves})
    for nt in nts:
        lib.opo_tune_set(OPO_FP32, OPO_TUNE_TIME_PASS, nt, 0)
    for nx in nxs:
        lib.opo_tune_set(OPO_FP32, OPO_TUNE_XY_PASS, nx, 0)
    ref, _ = batch_run(solves, 2)
    # two rounds of persistent choices: first and last valid of each class
    for pick in (0, -1):
        for nt in nts:
            vs = _persistent(lib, OPO_TUNE_TIME_PASS, nt)
            assert vs, nt
            assert lib.opo_tune_set(OPO_FP32, OPO_TUNE_TIME_PASS, nt, vs[pick]) == OPO_OK
        for nx in nxs:
            vs = _persistent(lib, OPO_TUNE_XY_PASS, nx)
            if vs:
                assert lib.opo_tune_set(OPO_FP32, OPO_TUNE_XY_PASS, nx, vs[pick]) == OPO_OK
        got, _ = batch_run(solves, 2)
        for k, ((pr, fr), (pg, fg)) in enumerate(zip(ref, got)):
            for a, b in zip(fr, fg):
                assert np.array_equal(a, b), (pick, k)
            assert np.array_equal(pr, pg), (pick, k)
        print("persistent batch == classic batch, pick", pick, flush=True)


def test_trip_graph_follows_tuning_changes(restore_tuning, require_gpu):
    """A solve's captured trip graph is re-captured when the process-wide
    variants change between trips (here to and from a plane-group xy
    variant, which changes the launch grid); the fields stay bitwise those
    of a run on one variant."""
    lib = restore_tuning
    grid, p, _, fields = configs.cavity_64(n=(32, 32, 128), n_round_trips=3)
    lib.opo_tune_set(OPO_FP32, OPO_TUNE_XY_PASS, 32, 0)
    ref, pref = _run(grid, p, fields, trips=3)
    with co.Solver(grid, p, device=0) as s:
        s.set_fields(*fields)
        pw = [s.runOPO(1)]
        for v in (4, 0):   # 512 threads on 8 planes, then one plane again
            assert lib.opo_tune_set(OPO_FP32, OPO_TUNE_XY_PASS, 32, v) == OPO_OK
            pw.append(s.runOPO(1))
        got = s.get_fields()
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    assert np.array_equal(np.concatenate(pw), pref)


def test_batch_follows_tuning_changes_between_trips(restore_tuning, require_gpu):
    """A resident batch whose xy variant changes between trips (one plane
    per workgroup to groups of planes, which changes its work list) rebuilds
    and gives the fields of a batch that kept one variant, bitwise."""
    lib = restore_tuning
    solves = [s for s in batch_mix(4, trips=3) if s[0].nx == 32]
    assert solves
    lib.opo_tune_set(OPO_FP32, OPO_TUNE_XY_PASS, 32, 0)
    ref, _ = batch_run(solves, 3)
    b = co.Batch(8 << 30)
    tickets = [b.submit(g, p, *f, 3) for (g, p, ph, f) in solves]
    b.step()
    assert lib.opo_tune_set(OPO_FP32, OPO_TUNE_XY_PASS, 32, 4) == OPO_OK
    results = {}
    for _ in range(20):
        b.step()
        for t in tickets:
            if t not in results:
                r = b.poll(t)
                if r is not None:
                    results[t] = r
        if len(results) == len(tickets):
            break
    b.close()
    for t, (pr, fr) in zip(tickets, ref):
        pg, fg = results[t]
        for a, c in zip(fg, fr):
            assert np.array_equal(a, c)
        assert np.array_equal(pg, pr)


@pytest.mark.parametrize("n", [(32, 32, 64), (64, 64, 256), (16, 16, 2048), (128, 128, 64), (256, 256, 64)],
                         ids=lambda n: "x".join(map(str, n)))
def test_fp64_variants_give_the_same_fields(n, require_gpu):
    """The FP64 path's launch variants (one tile or plane per workgroup,
    groups of small planes, one-pixel 2048-sample tiles on 256, 512 or 384
    threads, the band shapes of the two-pass 128 x 128 and 256 x 256 xy
    operators) give bitwise the fields of variant 0."""
    from cuopo3d_amd.capi import OPO_FP64
    lib = co.load()
    nx, ny, nt = n
    keep = {a: lib.opo_tune_get(OPO_FP64, a, sz) for a, sz in ((OPO_TUNE_TIME_PASS, nt), (OPO_TUNE_XY_PASS, nx))}
    grid, p, _, fields = configs.cavity_64(precision=OPO_FP64, n=n, n_round_trips=2)
    try:
        for axis, size in ((OPO_TUNE_TIME_PASS, nt), (OPO_TUNE_XY_PASS, nx)):
            lib.opo_tune_set(OPO_FP64, OPO_TUNE_TIME_PASS, nt, 0)
            lib.opo_tune_set(OPO_FP64, OPO_TUNE_XY_PASS, nx, 0)
            ref, pref = _run(grid, p, fields)
            vs = [v for v in range(64) if lib.opo_tune_set(OPO_FP64, axis, size, v) == OPO_OK]
            assert len(vs) > 1, (axis, size)
            for v in vs:
                assert lib.opo_tune_set(OPO_FP64, axis, size, v) == OPO_OK
                try:
                    got, pw = _run(grid, p, fields)
                except RuntimeError as e:   # a tile that does not divide NX is refused, not run
                    assert "does not divide" in str(e), e
                    continue
                for a, b in zip(got, ref):
                    assert np.array_equal(a, b), (axis, v)
                assert np.array_equal(pw, pref), (axis, v)
            print(f"fp64 {n} axis {axis}: variants {vs} bitwise equal", flush=True)
    finally:
        for a, v in keep.items():
            lib.opo_tune_set(OPO_FP64, a, nt if a == OPO_TUNE_TIME_PASS else nx, v)


def test_wide_tuned_tile_falls_back_on_narrow_grids(restore_tuning, require_gpu):
    """The t-pass choice is keyed by NT only: a 32-pixel tile tuned on a wide
    grid does not divide a 16-pixel row.  Single solves and a resident batch
    of 16 x 16 x 128 solves then run the default tile (variant 0) and give
    its fields bitwise; a 64-pixel grid in the same batch class keeps
    working too."""
    lib = restore_tuning
    draws = configs.sweep_draws(n_solves=3, seed=13)
    solves = []
    for d, nxy in zip(draws, (16, 16, 64)):
        d = dict(d)
        d["nxy"], d["nt"] = nxy, 128
        solves.append(configs.sweep_solve(d, n_round_trips=2))
    assert lib.opo_tune_set(OPO_FP32, OPO_TUNE_TIME_PASS, 128, 0) == OPO_OK
    ref = [_run(g, p, f, trips=2) for (g, p, ph, f) in solves]
    ref_batch, _ = batch_run(solves, 2)
    # variants 6-8 of a 128-sample line: tiles of twice the default 16 pixels
    wide = [v for v in (7, 6, 8) if v in _valid(lib, OPO_TUNE_TIME_PASS, 128)]
    assert wide, "no 32-pixel variant for NT=128"
    assert lib.opo_tune_set(OPO_FP32, OPO_TUNE_TIME_PASS, 128, wide[0]) == OPO_OK
    for (g, p, ph, f), (rf, rp) in zip(solves, ref):
        got, pw = _run(g, p, f, trips=2)
        for a, b in zip(got, rf):
            assert np.array_equal(a, b), (g.nx, "single solve")
        assert np.array_equal(pw, rp)
    got_batch, _ = batch_run(solves, 2)
    for (pg, fg), (pr, fr), (g, _, _, _) in zip(got_batch, ref_batch, solves):
        for a, c in zip(fg, fr):
            assert np.array_equal(a, c), (g.nx, "batch")
        assert np.array_equal(pg, pr)
    for (pg, fg), (rf, rp) in zip(got_batch, ref):
        for a, c in zip(fg, rf):
            assert np.array_equal(a, c)


def test_batch_autotune_keeps_fields(restore_tuning, require_gpu):
    """opo_batch_autotune times each size class's variants on the batch's own
    trip, before the first trip (it admits the queue itself) and between
    trips: the fields and per-trip powers afterwards are bitwise those of a
    batch that was never tuned (the resident fields are saved and put back),
    and each NT and NX class of the mix has a report line whose pick is the
    variant now in the process-wide table."""
    lib = restore_tuning
    solves = batch_mix(4, trips=3)
    ref, _ = batch_run(solves, 3)
    b = co.Batch(8 << 30)
    tickets = [b.submit(g, p, *f, 3) for (g, p, ph, f) in solves]
    reports = [b.autotune(reps=1)]
    b.step()
    reports.append(b.autotune(reps=1))
    results = {}
    for _ in range(10):
        b.step()
        for t in tickets:
            if t not in results:
                r = b.poll(t)
                if r is not None:
                    results[t] = r
        if len(results) == len(tickets):
            break
    b.close()
    for k, ((pr, fr), t) in enumerate(zip(ref, tickets)):
        pg, fg = results[t]
        for x, y in zip(fr, fg):
            assert np.array_equal(x, y), k
        assert np.array_equal(pr, pg), k
    print(reports[-1], flush=True)
    for axis, name, sizes in ((OPO_TUNE_TIME_PASS, "t", {g.nt for (g, _, _, _) in solves}),
                              (OPO_TUNE_XY_PASS, "xy", {g.nx for (g, _, _, _) in solves})):
        for n in sizes:
            m = re.search(rf"batch {name} pass N[TX]={n} .*-> v(\d+)", reports[-1])
            assert m, (name, n)
            assert lib.opo_tune_get(OPO_FP32, axis, n) == int(m.group(1)), (name, n)
