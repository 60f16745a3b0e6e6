"""Batched solves (opo_batch_*) on the GPU.

Each solve's arithmetic in a batch is the same as alone, so a solve's fields
must be bitwise equal whether it runs alone, in a mixed batch, or through
the single-solve API.  Admission waits for HBM budget, an over-budget solve
is rejected, and a solve that produces non-finite power is evicted without
touching its neighbours.
"""
import numpy as np
import pytest

import cuopo3d_amd as co
from cuopo3d_amd import configs
from cuopo3d_amd.capi import OPO_FP32
from oracle.pyoracle import Oracle

from helpers import rel_l2, batch_mix, batch_run

pytestmark = pytest.mark.gpu

GB = 1 << 30


def mix(n=4, trips=3, seed=3):
    return batch_mix(n, trips, seed)


def run_and_collect(solves, trips, budget=8 * GB):
    return batch_run(solves, trips, budget)


def test_batch_alone_equals_mixed_bitwise(require_gpu):
    solves = mix(4, trips=3)
    mixed, _ = run_and_collect(solves, 3)
    for k, s in enumerate(solves):
        alone, _ = run_and_collect([s], 3)
        (pa, fa), (pm, fm) = alone[0], mixed[k]
        for a, m in zip(fa, fm):
            assert np.array_equal(a, m), k
        assert np.array_equal(pa, pm), k
    print("alone == mixed bitwise for", [(s[0].nx, s[0].nt) for s in solves], flush=True)


def test_batch_equals_single_solve_and_oracle(require_gpu):
    solves = mix(3, trips=2)
    mixed, _ = run_and_collect(solves, 2)
    for (g, p, ph, f), (pb, fb) in zip(solves, mixed):
        s = co.Solver(g, p, device=0)
        s.set_fields(*f)
        ps = s.runOPO(2)
        fs = s.get_fields()
        for a, b in zip(fs, fb):
            assert np.array_equal(a, b)
        assert np.array_equal(ps, pb)   # the same reduction order alone and batched
        o = Oracle(g, p)
        o.set_fields(*f)
        pr = o.run_round_trips(2)
        fr = o.get_fields()
        errs = [rel_l2(a, b) for a, b in zip(fb, fr)]
        print((g.nx, g.nt), "batch vs oracle rel L2", errs, flush=True)
        assert max(errs) < 1e-5
        assert np.max(np.abs(pb - pr) / np.abs(pr)) < 1e-6


def test_batch_long_and_grouped_lines_equal_single_solves(require_gpu):
    """2048- and 4096-sample solves (three-level t operators) and 16- and
    32-pixel solves (xy groups of planes when tuned so) batched together give the
    single-solve fields bitwise, with the small-plane xy variant set to a
    group of planes and to one plane."""
    from cuopo3d_amd.capi import OPO_FP32, OPO_TUNE_XY_PASS, OPO_OK
    lib = co.load()
    draws = configs.sweep_draws(n_solves=4, seed=5)
    solves = []
    for d, (nxy, nt) in zip(draws, [(16, 2048), (32, 128), (16, 256), (16, 4096)]):
        d = dict(d)
        d["nxy"], d["nt"] = nxy, nt
        solves.append(configs.sweep_solve(d, n_round_trips=2))
    keep = {nx: lib.opo_tune_get(OPO_FP32, OPO_TUNE_XY_PASS, nx) for nx in (16, 32)}
    try:
        for v in (4, 0):   # 512 threads on 256/NX planes, then one plane of 256
            for nx in (16, 32):
                assert lib.opo_tune_set(OPO_FP32, OPO_TUNE_XY_PASS, nx, v) == OPO_OK
            mixed, _ = run_and_collect(solves, 2)
            for (g, p, ph, f), (pb, fb) in zip(solves, mixed):
                s = co.Solver(g, p, device=0)
                s.set_fields(*f)
                ps = s.runOPO(2)
                for a, b in zip(s.get_fields(), fb):
                    assert np.array_equal(a, b), (v, g.nx, g.nt)
                assert np.array_equal(ps, pb)   # the same reduction order alone and batched
                s.close()
            print("xy variant", v, ": batch == single for", [(s[0].nx, s[0].nt) for s in solves], flush=True)
    finally:
        for nx, v in keep.items():
            lib.opo_tune_set(OPO_FP32, OPO_TUNE_XY_PASS, nx, v)


def test_batch_admission_waits_for_budget(require_gpu):
    draws = configs.sweep_draws(n_solves=3, seed=9)
    for d in draws:
        d["nxy"], d["nt"] = 64, 128      # equal footprints
    solves = [configs.sweep_solve(d, n_round_trips=2) for d in draws]
    fp = max(co.Batch(GB).footprint(g) for (g, _, _, _) in solves)
    b = co.Batch(int(fp * 1.5))   # room for one solve at a time
    tickets = [b.submit(g, p, *f, 2) for (g, p, ph, f) in solves]
    b.step()
    r, q, used = b.stats()
    assert r == 1 and q == 2 and 0 < used <= fp * 1.5
    results = {}
    for _ in range(40):
        b.step()
        for t in tickets:
            if t not in results:
                x = b.poll(t)
                if x is not None:
                    results[t] = x
        if len(results) == 3:
            break
    assert len(results) == 3
    ref, _ = run_and_collect(solves, 2)
    for t, (pr, fr) in zip(tickets, ref):
        for a, c in zip(results[t][1], fr):
            assert np.array_equal(a, c)


def test_batch_rejects_over_budget(require_gpu):
    (g, p, ph, f), = mix(1)
    b = co.Batch(1 << 20)
    with pytest.raises(RuntimeError, match="budget"):
        b.submit(g, p, *f, 1)


def test_batch_evicts_nan_solve_only(require_gpu):
    solves = mix(3, trips=3)
    g, p, ph, (api, as_, ai) = solves[1]
    bad = as_.copy()
    bad[0, 0, 0] = np.nan
    poisoned = list(solves)
    poisoned[1] = (g, p, ph, (api, bad, ai))
    b = co.Batch(8 * GB)
    tickets = [b.submit(gg, pp, *ff, 3) for (gg, pp, _, ff) in poisoned]
    out = {}
    for _ in range(20):
        b.step()
        for t in tickets:
            if t in b.meta and t not in out:
                try:
                    r = b.poll(t)
                except RuntimeError as e:
                    out[t] = e
                    b.meta.pop(t, None)
                    continue
                if r is not None:
                    out[t] = r
        if len(out) == 3:
            break
    assert isinstance(out[tickets[1]], RuntimeError) and "non-finite" in str(out[tickets[1]])
    clean, _ = run_and_collect([solves[0], solves[2]], 3)
    for t, ref in zip((tickets[0], tickets[2]), clean):
        for a, c in zip(out[t][1], ref[1]):
            assert np.array_equal(a, c)


def test_sweep_cli_matches_single_solves(tmp_path, require_gpu):
    """cuopo3d_amd.sweep over 2 powers x 2 waists: every point's per-trip
    power equals the same point run alone through the single-solve API."""
    from cuopo3d_amd import sweep
    from cuopo3d_amd.capi import Grid
    rc = sweep.main(["--powers", "500:600:100", "--waists", "31,40", "--grid", "32", "64",
                     "--trips", "3", "--out", str(tmp_path), "--npy"])
    assert rc == 0
    grid = Grid(32, 32, 64)
    ph = sweep.reference_physical(3)
    for w in (31, 40):
        for pw in (500, 600):
            folder = tmp_path / f"waist_{w}" / f"sPPLT_N_{pw}"
            rows = np.loadtxt(folder / "power_per_trip.dat")
            assert rows.shape == (3, 4)
            assert (folder / f"OutputSignalPower_PumpPower{pw}_W_slice32_r.dat").exists()
            assert (folder / "As_temporal_x_16_y_16_i.dat").exists()
            p, f = sweep.point(grid, ph, pw, w)
            s = co.Solver(grid, p, device=0)
            s.set_fields(*f)
            ref = s.runOPO(3)
            # the file holds 17 significant digits of the batch's doubles
            assert np.allclose(rows[:, 1:], ref, rtol=1e-15, atol=0), (w, pw)
            # --npy: the whole fields, bitwise those of the single solve
            for name, fld in zip(("pump", "signal", "idler"), s.get_fields()):
                saved = np.load(folder / f"{name}.npy")
                assert saved.shape == (64, 32, 32) and np.array_equal(saved, fld), (w, pw, name)
            s.close()


def test_batch_with_eom_equals_single_solves(require_gpu):
    """Solves with the modulator in a batch give the single-solve fields bitwise."""
    solves = mix(3, trips=2)
    for (g, p, ph, f) in solves:
        p.eom_on, p.eom_beta, p.eom_fpm = 1, 0.6, 2.0 / p.t_window
    mixed, _ = run_and_collect(solves, 2)
    for (g, p, ph, f), (pb, fb) in zip(solves, mixed):
        s = co.Solver(g, p, device=0)
        s.set_fields(*f)
        ps = s.runOPO(2)
        for a, b in zip(s.get_fields(), fb):
            assert np.array_equal(a, b)
        assert np.array_equal(ps, pb)   # the same reduction order alone and batched


def test_batch_rejects_1x1_grid_with_diffraction():
    """A 1 x 1 grid runs in a batch only as the plane-wave model."""
    g, p, ph, f = configs.cavity_64(n=(1, 1, 256), n_round_trips=1)
    b = co.Batch(1 << 30)
    with pytest.raises(RuntimeError, match="plane-wave"):
        b.submit(g, p, *f, 1)
    b.close()


def test_cpp_driver_runs(tmp_path, require_gpu):
    """apps/opo3d_main (the cuOPO3D.cu counterpart) through the C++ wrapper:
    writes the per-trip powers and the reference-format signal files."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "apps", "opo3d_main")
    if not os.path.exists(exe):
        pytest.skip("apps/opo3d_main is not built (make -C apps)")
    r = subprocess.run([exe, "500", "31", "32", "64", "3", str(tmp_path), "--npy"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "trips 3" in r.stdout
    rows = np.loadtxt(tmp_path / "power_per_trip.dat")
    assert rows.shape == (3, 4) and np.all(np.isfinite(rows))
    assert (tmp_path / "OutputSignalPower_PumpPower500_W_slice32_r.dat").exists()
    assert (tmp_path / "As_temporal_x_16_y_16_i.dat").exists()
    # the .npy signal holds the same values as the text slice (20 digits)
    sig = np.load(tmp_path / "signal_PumpPower500_W.npy")
    assert sig.shape == (64, 32, 32) and sig.dtype == np.complex64
    re = np.loadtxt(tmp_path / "OutputSignalPower_PumpPower500_W_slice32_r.dat")
    assert np.array_equal(re.astype(np.float32), sig[32].real)


def test_fp64_batch_equals_single_solves(require_gpu):
    """An FP64 batch (all-FP64 device path) gives the FP64 single-solve
    fields bitwise and the oracle's to 1e-11, 128 x 128 planes (row bands
    then column bands) included."""
    from cuopo3d_amd.capi import OPO_FP64
    solves = []
    for n in ((32, 32, 64), (64, 64, 32), (128, 128, 32)):
        g, p, ph, f = configs.cavity_64(n=n, n_round_trips=2, precision=OPO_FP64)
        solves.append((g, p, ph, f))
    b = co.Batch(4 * GB, precision=OPO_FP64)
    tickets = [b.submit(g, p, *f, 2) for (g, p, ph, f) in solves]
    res = {}
    for _ in range(20):
        b.step()
        for t in tickets:
            if t not in res:
                r = b.poll(t)
                if r is not None:
                    res[t] = r
        if len(res) == len(tickets):
            break
    b.close()
    for (g, p, ph, f), t in zip(solves, tickets):
        pb, fb = res[t]
        s = co.Solver(g, p, device=0)
        s.set_fields(*f)
        ps = s.runOPO(2)
        for a, c in zip(s.get_fields(), fb):
            assert np.array_equal(a, c)
        o = Oracle(g, p)
        o.set_fields(*f)
        o.run_round_trips(2)
        errs = [rel_l2(a, c) for a, c in zip(fb, o.get_fields())]
        print((g.nx, g.nt), "fp64 batch vs oracle", errs, flush=True)
        assert max(errs) < 1e-11


@pytest.mark.parametrize("prec,nxy", [("fp32", 256), ("fp64", 256), ("fp32", 512), ("fp64", 512)])
def test_batch_large_planes_equal_single_solves(prec, nxy, require_gpu):
    """A 256 x 256 or 512 x 512 solve (row-band then column-band xy
    passes) batched with a 32 x 32 one gives the single-solve fields and
    per-trip power bitwise, in either precision."""
    from cuopo3d_amd.capi import OPO_FP64
    pr = OPO_FP64 if prec == "fp64" else OPO_FP32
    solves = [configs.cavity_64(n=n, n_round_trips=2, precision=pr) for n in ((nxy, nxy, 16), (32, 32, 64))]
    b = co.Batch(4 * GB, precision=pr)
    tickets = [b.submit(g, p, *f, 2) for (g, p, ph, f) in solves]
    res = {}
    for _ in range(20):
        b.step()
        for t in tickets:
            if t not in res:
                r = b.poll(t)
                if r is not None:
                    res[t] = r
        if len(res) == len(tickets):
            break
    b.close()
    assert len(res) == len(tickets)
    for (g, p, ph, f), t in zip(solves, tickets):
        pb, fb = res[t]
        s = co.Solver(g, p, device=0)
        s.set_fields(*f)
        ps = s.runOPO(2)
        for a, c in zip(s.get_fields(), fb):
            assert np.array_equal(a, c), (prec, g.nx)
        assert np.array_equal(ps, pb), (prec, g.nx)
        s.close()
    print(prec, nxy, "batch == single solve", flush=True)


def test_batch_concurrent_classes_bitwise(require_gpu):
    """Size classes of a stage run as parallel branches of the trip graph
    (opo_batch_concurrent_classes): the fields and per-trip powers of every
    solve are bitwise those of the serial trip, with a solve finishing early
    (a change of the resident set) along the way."""
    solves = mix(4, trips=3)
    ref, _ = run_and_collect(solves, 3)
    got, _ = batch_run(solves, 3, concurrent=True)
    for k, ((pr, fr), (pg, fg)) in enumerate(zip(ref, got)):
        for a, b in zip(fr, fg):
            assert np.array_equal(a, b), k
        assert np.array_equal(pr, pg), k
    # uneven trip counts: solves leave the batch at different steps
    short = [s if k % 2 else (s[0], s[1], s[2], s[3]) for k, s in enumerate(solves)]
    bt = co.Batch(8 * GB)
    bt.concurrent_classes(True)
    tickets = [bt.submit(g, p, *f, 2 + (k % 2)) for k, (g, p, ph, f) in enumerate(short)]
    results = {}
    for _ in range(20):
        bt.step()
        for t in tickets:
            if t not in results:
                r = bt.poll(t)
                if r is not None:
                    results[t] = r
        if len(results) == len(tickets):
            break
    bt.close()
    for k, t in enumerate(tickets):
        g, p, ph, f = short[k]
        with co.Solver(g, p, device=0) as s:
            s.set_fields(*f)
            pw = s.runOPO(2 + (k % 2))
            fa = s.get_fields()
        pg, fg = results[t]
        for a, b in zip(fa, fg):
            assert np.array_equal(a, b), k
        assert np.array_equal(pw, pg), k


def test_batch_mixed_slice_counts_bitwise(require_gpu):
    """Solves of different crystal lengths and slice counts (the float slice
    rule of Solver3D.cuh:949) in one batch: a size class's launch at slice k
    covers only its solves with more than k slices, so each solve's fields
    and per-trip powers are bitwise those of the solve alone through the
    single-solve API, also after the shortest solve has left the batch."""
    from cuopo3d_amd.capi import Grid
    draws = configs.sweep_draws(n_solves=4, seed=11)
    specs = [(32, 128, 5000.0, 100, 3), (32, 128, 3000.0, 31, 2), (64, 64, 2000.0, 41, 3), (32, 256, 5000.0, 60, 3)]
    solves = []
    for d, (nxy, nt, lcr, nz, trips) in zip(draws, specs):
        grid = Grid(nxy, nxy, nt)
        ph = configs.physical(lcr=lcr, nz=nz, rs=d["R"], phase_s=d["phase"], n_round_trips=trips)
        p = configs.make_params(grid, ph)
        f = configs.make_fields(grid, p, ph, power=d["power"], waist=d["waist"], seed=d["seed"])
        solves.append((grid, p, ph, f, trips))
    slices = [s[1].n_slices for s in solves]
    print("slices per solve", slices, flush=True)
    assert len(set(slices)) == 4
    b = co.Batch(8 * GB)
    tickets = [b.submit(g, p, *f, trips) for (g, p, ph, f, trips) in solves]
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
    for k, (g, p, ph, f, trips) in enumerate(solves):
        with co.Solver(g, p, device=0) as s:
            s.set_fields(*f)
            ps = s.runOPO(trips)
            fs = s.get_fields()
        pb, fb = results[tickets[k]]
        for a, c in zip(fs, fb):
            assert np.array_equal(a, c), k
        assert np.array_equal(ps, pb), k


def test_batch_plane_wave_solves_bitwise(require_gpu):
    """Plane-wave solves (NX = NY = 1, no transverse step, with and without
    the modulator and GDD; Solver1D.cuh:436-445, 62-72) batched beside 3-D
    solves of the same NT: their t pass is a size class of its own (one
    pixel per workgroup) and they skip the xy pass and free space, so each
    solve's fields and per-trip powers are bitwise those of the solve alone
    through the single-solve API."""
    solves = []
    for nt, gdd, beta in ((256, 1, 0.5), (256, 0, 0.3), (1024, 1, 0.0)):
        grid, p, ph, f = configs.cavity_64(n=(1, 1, nt), n_round_trips=3)
        p.no_diffraction = 1
        p.gdd_on = gdd
        if beta:
            p.eom_on, p.eom_beta, p.eom_fpm = 1, beta, 2.0 / p.t_window
        solves.append((grid, p, ph, f))
    # the plane-wave model on a 3-D grid: the usual t pass, no xy pass
    g, p, ph, f = configs.cavity_64(n=(32, 32, 64), n_round_trips=3)
    p.no_diffraction = 1
    solves.append((g, p, ph, f))
    solves += mix(4, trips=3)   # 3-D solves, one of them 32 x 32 x 256
    got, _ = run_and_collect(solves, 3)
    for k, (g, p, ph, f) in enumerate(solves):
        with co.Solver(g, p, device=0) as s:
            s.set_fields(*f)
            ps = s.runOPO(3)
            fs = s.get_fields()
        pb, fb = got[k]
        for a, c in zip(fs, fb):
            assert np.array_equal(a, c), k
        assert np.array_equal(ps, pb), k
    print("plane-wave and 3-D solves batched == alone, bitwise", flush=True)
