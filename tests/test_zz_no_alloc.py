"""The no-allocation round-trip entry point (opo_run_round_trips_noalloc)
against opo_run_round_trips: two calls in a row through the C ABI, the
second on a smaller grid, each bitwise equal (fields and per-trip power) to
the allocating entry point on the same inputs.  The device buffers the
no-allocation call works on are the fields of a second solve, reached with
opo_field_device_ptr."""
import ctypes as C

import numpy as np
import pytest

import cuopo3d_amd as co
from cuopo3d_amd import configs
from cuopo3d_amd.capi import OPO_OK, OPO_FP64

pytestmark = pytest.mark.gpu


def _reference(grid, p, fields, trips):
    with co.Solver(grid, p, device=0) as s:
        s.set_fields(*fields)
        pw = s.runOPO(trips)
        return s.get_fields(), pw


def _noalloc(lib, grid, p, fields, trips):
    """Run the no-allocation entry point on the device fields of a holder solve."""
    with co.Solver(grid, p, device=0) as holder:
        holder.set_fields(*fields)
        h = holder.h
        ptr = [lib.opo_field_device_ptr(h, k) for k in (3, 0, 1, 2)]   # Api, Ap, As, Ai
        assert all(ptr)
        pw = np.zeros((trips, 3))
        rc = lib.opo_run_round_trips_noalloc(C.byref(grid), C.byref(p), 0, *ptr, trips,
                                             pw.ctypes.data_as(C.POINTER(C.c_double)))
        assert rc == OPO_OK, lib.opo_last_error()
        return holder.get_fields(), pw


@pytest.mark.parametrize("prec", [0, OPO_FP64], ids=["fp32", "fp64"])
def test_noalloc_equals_run_round_trips(prec, require_gpu):
    lib = co.load()
    # a 64x64x256 cavity for 3 trips, then a smaller 32x32x128 one for 2
    for n, trips in (((64, 64, 256), 3), ((32, 32, 128), 2)):
        grid, p, _, fields = configs.cavity_64(n=n, n_round_trips=trips, precision=prec)
        print("no-alloc vs run_round_trips", n, trips, "fp64" if prec else "fp32", flush=True)
        ref, pref = _reference(grid, p, fields, trips)
        got, pw = _noalloc(lib, grid, p, fields, trips)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), n
        assert np.array_equal(pw, pref), n
        # the free-space step is skipped on the run's last trip, as in the reference
        assert trips == p.n_round_trips
