"""MI355X split-step Fourier path for (3+1)D chi(2) OPOs.

Python view of the C ABI in ``include/opo3d.h`` (library ``libopo3d.so``
next to this file).  ``Solver`` mirrors the member functions of the
reference's ``Solver<Crystal>`` (src/headers/Solver3D.cuh:429-491) so a
script written against the reference reads the same here; ``configs``
builds the benchmark workloads.
"""
import ctypes as C

import numpy as np

from . import capi
from .capi import (OPO_FP32, OPO_FP64, OPO_OK, OPO_ERR_NAN, OPO_ERR_UNSUPPORTED,  # noqa: F401
                   Grid, Params, Physical, Crystal, check, load)

__all__ = ["Solver", "Batch", "autotune_grids", "Grid", "Params", "Physical", "load", "check",
           "OPO_FP32", "OPO_FP64"]


def _cdtype(precision):
    return np.complex128 if precision == OPO_FP64 else np.complex64


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Solver:
    """One OPO solve resident on a GPU.

    Methods follow Solver<Crystal> (Solver3D.cuh): ``solverRK4``,
    ``dispersion``, ``diffractionInCrystal``, ``diffractionInFreeSpace``,
    ``GDDCompensation``, ``applyReflectivity``, ``SSFM``, ``runSinglePass``,
    ``runOPO``, ``averagePowerCutXY``.
    """

    def __init__(self, grid, params, device=-1):
        self.lib = load()
        self.grid = grid
        self.params = params
        self.precision = params.precision
        self.shape = (grid.nt, grid.ny, grid.nx)
        h = C.c_void_p()
        check(self.lib.opo_create(C.byref(grid), C.byref(params), device, C.byref(h)),
              "opo_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.opo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _as(self, a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=_cdtype(self.precision)).reshape(self.shape)
        return a

    def set_fields(self, api=None, as_=None, ai=None):
        bufs = [self._as(x) for x in (api, as_, ai)]
        check(self.lib.opo_set_fields(self.h, *[_ptr(b) for b in bufs]), "opo_set_fields")

    def get_fields(self):
        out = [np.empty(self.shape, dtype=_cdtype(self.precision)) for _ in range(3)]
        check(self.lib.opo_get_fields(self.h, *[_ptr(b) for b in out]), "opo_get_fields")
        return out

    def device_ptr(self, which):
        return self.lib.opo_field_device_ptr(self.h, which)

    # reference member functions ------------------------------------------
    def solverRK4(self):
        check(self.lib.opo_step_nonlinear(self.h), "opo_step_nonlinear")

    def dispersion(self):
        check(self.lib.opo_step_dispersion(self.h), "opo_step_dispersion")

    def diffractionInCrystal(self):
        check(self.lib.opo_step_diffraction(self.h), "opo_step_diffraction")

    def diffractionInFreeSpace(self):
        check(self.lib.opo_step_free_space(self.h), "opo_step_free_space")

    def GDDCompensation(self):
        check(self.lib.opo_step_gdd(self.h), "opo_step_gdd")

    def applyReflectivity(self):
        check(self.lib.opo_step_mirror(self.h), "opo_step_mirror")

    def SSFM(self):
        check(self.lib.opo_ssfm(self.h), "opo_ssfm")

    def runSinglePass(self):
        check(self.lib.opo_single_pass(self.h), "opo_single_pass")

    def runOPO(self, n_trips=None, raise_on_nan=True):
        """Run round trips; returns the per-trip (p, s, i) powers, shape (n, 3)."""
        n = self.params.n_round_trips if n_trips is None else n_trips
        pw = np.zeros((n, 3), dtype=np.float64)
        rc = self.lib.opo_run_round_trips(self.h, n, pw.ctypes.data_as(C.POINTER(C.c_double)))
        if rc == OPO_ERR_NAN and not raise_on_nan:
            return pw[: self.trips_done]
        check(rc, "opo_run_round_trips")
        return pw

    @property
    def trips_done(self):
        return self.lib.opo_trips_done(self.h)

    def averagePowerCutXY(self):
        out = (C.c_double * 3)()
        check(self.lib.opo_power(self.h, out), "opo_power")
        return np.array(list(out))

    def averagePowerTime(self, field, x, y, prefactor):
        """prefactor * sum_t |A(x, y, t)|^2 of field 0/1/2 (opo_power_time)."""
        out = C.c_double()
        check(self.lib.opo_power_time(self.h, field, x, y, prefactor, C.byref(out)), "opo_power_time")
        return out.value

    def checkCourantStability(self):
        """dz < 0.2 kp dx dy (opo_courant_ok); True when it holds."""
        return bool(self.lib.opo_courant_ok(C.byref(self.params)))

    def sync(self):
        check(self.lib.opo_sync(self.h), "opo_sync")

    def autotune(self, reps=10):
        """Pick the fastest launch variants for this grid (opo_autotune)."""
        buf = C.create_string_buffer(4096)
        check(self.lib.opo_autotune(self.h, reps, buf, len(buf)), "opo_autotune")
        return buf.value.decode()


def autotune_grids(grids, precision=OPO_FP32, reps=10, device=0):
    """Tune the launch variants for a set of (NX, NY, NT) grids, e.g. the
    size classes of a sweep or a serving mix, before they are run.

    The xy pass is tuned per NX on the grid of that NX with the most time
    samples (the most planes: a batch holds many planes of each NX, and
    small planes favour other launch shapes in bulk than a few alone); the
    t pass per NT on the widest grid of that NT (it carries most of the
    work), last, so that later calls do not overwrite it.  Returns the
    tuner's reports."""
    from . import configs
    grids = sorted({(int(g[0]), int(g[1]), int(g[2])) for g in grids})
    widest = {}
    for nx, ny, nt in grids:
        widest[nt] = max(widest.get(nt, (0, 0, 0)), (nx, ny, nt))
    by_nx = {}
    for g in grids:   # sorted: the last grid of each NX has the longest lines
        by_nx[g[0]] = g
    order = [g for g in by_nx.values() if g not in widest.values()] + list(widest.values())
    reports = []
    for nx, ny, nt in order:
        grid, p, _, fields = configs.cavity_64(n=(nx, ny, nt), n_round_trips=1, precision=precision)
        with Solver(grid, p, device=device) as s:
            s.set_fields(*fields)
            reports.append(s.autotune(reps))
    return reports


class Batch:
    """Resident batch of solves with HBM-budget admission (opo_batch_*)."""

    def __init__(self, budget_bytes, precision=OPO_FP32, device=0):
        self.lib = load()
        self.precision = precision
        h = C.c_void_p()
        check(self.lib.opo_batch_create(device, C.c_size_t(int(budget_bytes)), precision,
                                        C.byref(h)), "opo_batch_create")
        self.h = h
        self.meta = {}

    def close(self):
        if getattr(self, "h", None):
            self.lib.opo_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, grid, params, api, as_, ai, n_trips):
        dt = _cdtype(self.precision)
        bufs = [np.ascontiguousarray(x, dtype=dt) for x in (api, as_, ai)]
        t = C.c_uint64()
        check(self.lib.opo_batch_submit(self.h, C.byref(grid), C.byref(params),
                                        *[_ptr(b) for b in bufs], n_trips, C.byref(t)),
              "opo_batch_submit")
        self.meta[t.value] = ((grid.nt, grid.ny, grid.nx), n_trips)
        return t.value

    def step(self):
        n = C.c_int()
        check(self.lib.opo_batch_step(self.h, C.byref(n)), "opo_batch_step")
        return n.value

    def poll(self, ticket, want_fields=True):
        shape, n_trips = self.meta[ticket]
        done = C.c_int()
        pw = np.zeros((n_trips, 3))
        dt = _cdtype(self.precision)
        out = [np.empty(shape, dtype=dt) for _ in range(3)] if want_fields else [None] * 3
        check(self.lib.opo_batch_poll(self.h, ticket, C.byref(done),
                                      pw.ctypes.data_as(C.POINTER(C.c_double)),
                                      *[_ptr(b) for b in out]), "opo_batch_poll")
        if not done.value:
            return None
        del self.meta[ticket]
        return pw, out

    def stats(self):
        r, q, u = C.c_int(), C.c_int(), C.c_size_t()
        check(self.lib.opo_batch_stats(self.h, C.byref(r), C.byref(q), C.byref(u)),
              "opo_batch_stats")
        return r.value, q.value, u.value

    def concurrent_classes(self, on=True):
        """Run the size classes of each stage as parallel branches of the trip graph."""
        check(self.lib.opo_batch_concurrent_classes(self.h, 1 if on else 0), "opo_batch_concurrent_classes")

    def autotune(self, reps=5):
        """Pick the fastest launch variant of each size class on this batch's
        own trip (opo_batch_autotune); the solves' fields are kept."""
        buf = C.create_string_buffer(1 << 16)
        check(self.lib.opo_batch_autotune(self.h, reps, buf, len(buf)), "opo_batch_autotune")
        return buf.value.decode()

    def rebuilds(self):
        """(re-captures of the trip graph, of which updated in place, host seconds)."""
        n, u, sec = C.c_int(), C.c_int(), C.c_double()
        check(self.lib.opo_batch_rebuilds(self.h, C.byref(n), C.byref(u), C.byref(sec)),
              "opo_batch_rebuilds")
        return n.value, u.value, sec.value

    def profile(self, on=True):
        """Time every launch inside the replayed trip graph from the next step on."""
        check(self.lib.opo_batch_profile(self.h, 1 if on else 0), "opo_batch_profile")

    def profile_read(self):
        """Per (stage, size class): launches, summed ms, algorithmic bytes per launch."""
        from .capi import StageTime, STAGES
        n = C.c_int()
        check(self.lib.opo_batch_profile_read(self.h, None, 0, C.byref(n)), "opo_batch_profile_read")
        buf = (StageTime * max(n.value, 1))()
        check(self.lib.opo_batch_profile_read(self.h, buf, n.value, C.byref(n)), "opo_batch_profile_read")
        return [{"stage": STAGES[e.stage], "size": e.size, "launches": e.launches, "ms": e.ms,
                 "alg_bytes": e.alg_bytes} for e in buf[:n.value]]

    def time_stages(self, reps=3):
        """Each (stage, size class) timed alone as a replayed graph of its
        launches of one trip; advances the resident fields."""
        from .capi import StageTime, STAGES
        buf = (StageTime * 64)()
        n = C.c_int()
        check(self.lib.opo_batch_time_stages(self.h, int(reps), buf, 64, C.byref(n)), "opo_batch_time_stages")
        return [{"stage": STAGES[e.stage], "size": e.size, "launches": e.launches, "ms": e.ms,
                 "alg_bytes": e.alg_bytes} for e in buf[:min(n.value, 64)]]

    def footprint(self, grid, n_trips=1):
        return self.lib.opo_batch_footprint(C.byref(grid), self.precision, int(n_trips))
