"""The C-ABI library on a host without a GPU: it loads, exports every
function the headers declare, and rejects bad input with error codes."""
import ctypes as C
import os
import re

import pytest

from cuopo3d_amd import capi, configs
from cuopo3d_amd.capi import Grid, OPO_ERR_ARG, OPO_ERR_NOGPU, OPO_ERR_UNSUPPORTED

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(opo_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(capi.LIB_PATH)
    names = declared()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_symbol():
    lib = capi.load()   # binds argtypes for every name it knows
    for n in declared():
        assert getattr(lib, n).argtypes is not None, n


def test_create_rejects_bad_input():
    lib = capi.load()
    g, p, _, _ = configs.single_pass_32()
    h = C.c_void_p()
    assert lib.opo_create(C.byref(Grid(0, 32, 64)), C.byref(p), 0, C.byref(h)) == OPO_ERR_ARG
    assert lib.opo_create(C.byref(Grid(100, 100, 64)), C.byref(p), 0, C.byref(h)) == OPO_ERR_UNSUPPORTED
    assert b"not supported" in lib.opo_last_error()
    # NX != NY: both lengths must be supported line lengths
    assert lib.opo_create(C.byref(Grid(32, 48, 64)), C.byref(p), 0, C.byref(h)) == OPO_ERR_UNSUPPORTED
    assert lib.opo_create(C.byref(Grid(1, 32, 64)), C.byref(p), 0, C.byref(h)) == OPO_ERR_UNSUPPORTED
    if lib.opo_device_count() == 0:   # a supported rectangular grid gets past validation
        assert lib.opo_create(C.byref(Grid(32, 64, 64)), C.byref(p), 0, C.byref(h)) == OPO_ERR_NOGPU
    bad = p.copy()
    bad.precision = 7
    assert lib.opo_create(C.byref(g), C.byref(bad), 0, C.byref(h)) == OPO_ERR_ARG
    bad = p.copy()
    bad.dz = 0.0
    assert lib.opo_create(C.byref(g), C.byref(bad), 0, C.byref(h)) == OPO_ERR_ARG
    bad = p.copy()   # the modulator needs the time window of its time grid
    bad.eom_on, bad.eom_beta, bad.eom_fpm = 1, 0.5, 0.1
    bad.t_window = 0.0
    bad.omega = capi.C.cast(capi.C.pointer((capi.C.c_double * 64)()), capi.C.POINTER(capi.C.c_double))
    assert lib.opo_create(C.byref(g), C.byref(bad), 0, C.byref(h)) == OPO_ERR_ARG
    assert b"t_window" in lib.opo_last_error()
    if lib.opo_device_count() == 0:
        assert lib.opo_create(C.byref(g), C.byref(p), 0, C.byref(h)) == OPO_ERR_NOGPU
    assert not h.value


def test_null_handles_are_errors():
    lib = capi.load()
    assert lib.opo_ssfm(None) == OPO_ERR_ARG
    assert lib.opo_run_round_trips(None, 1, None) == OPO_ERR_ARG
    assert lib.opo_batch_step(None, None) == OPO_ERR_ARG
    assert lib.opo_field_bytes(None) == 0
    assert lib.opo_power_time(None, 0, 0, 0, 1.0, None) == OPO_ERR_ARG
    lib.opo_destroy(None)


def test_batch_footprint():
    lib = capi.load()
    g = Grid(128, 128, 1024)
    fp = lib.opo_batch_footprint(C.byref(g), 0, 300)
    assert 4 * 8 * 128 * 128 * 1024 <= fp < 4 * 8 * 128 * 128 * 1024 * 1.01
    # tables (6 NX + 6 NY + 6 NT entries, hi/lo pairs), partials, counters and
    # the 3-double power series of every trip each take at least one 64 KiB unit
    small = Grid(16, 16, 16)
    f1 = lib.opo_batch_footprint(C.byref(small), 0, 1)
    assert f1 == 4 * 65536 + 4 * 65536
    f10k = lib.opo_batch_footprint(C.byref(small), 0, 10000)
    assert f10k - f1 == 3 * 65536   # 240,000 B of series: four units instead of one


def test_tune_set_rejects_missing_variants():
    """The variant tables are host state: checkable without a GPU."""
    from cuopo3d_amd.capi import OPO_FP32, OPO_TUNE_TIME_PASS, OPO_TUNE_XY_PASS
    lib = capi.load()
    keep = lib.opo_tune_get(OPO_FP32, OPO_TUNE_TIME_PASS, 256)
    assert lib.opo_tune_set(OPO_FP32, OPO_TUNE_TIME_PASS, 256, 99) != capi.OPO_OK
    assert lib.opo_tune_set(OPO_FP32, OPO_TUNE_TIME_PASS, 16, 1) != capi.OPO_OK  # NT < 64: one variant
    assert lib.opo_tune_set(OPO_FP32, OPO_TUNE_XY_PASS, 48, 0) != capi.OPO_OK    # no such NX
    assert lib.opo_tune_set(OPO_FP32, OPO_TUNE_XY_PASS, 256, 3) != capi.OPO_OK   # 256: three band shapes
    assert lib.opo_tune_get(OPO_FP32, OPO_TUNE_XY_PASS, 256) in (0, 1, 2)
    assert lib.opo_tune_set(OPO_FP32, OPO_TUNE_XY_PASS, 512, 1) != capi.OPO_OK   # 512: one band shape
    assert lib.opo_tune_get(OPO_FP32, OPO_TUNE_XY_PASS, 512) == 0
    assert lib.opo_tune_set(OPO_FP32, OPO_TUNE_TIME_PASS, 256, keep) == capi.OPO_OK
    assert lib.opo_tune_get(OPO_FP32, OPO_TUNE_TIME_PASS, 256) == keep


def test_noalloc_rejects_oversized_calls_before_touching_the_gpu():
    """The no-allocation entry point checks its fixed scratch (tables,
    partials, counters, 3 doubles of power per trip) before any HIP call."""
    lib = capi.load()
    assert capi.OPO_NOALLOC_SCRATCH_BYTES == 1 << 20
    g, p, _, _ = configs.cavity_64(n=(128, 128, 2048), n_round_trips=1)
    dummy = C.c_void_p(16)   # never dereferenced: the size check comes first
    rc = lib.opo_run_round_trips_noalloc(C.byref(g), C.byref(p), 0, dummy, dummy, dummy, dummy,
                                         40000, None)
    assert rc == OPO_ERR_UNSUPPORTED
    assert b"scratch" in lib.opo_last_error()
    assert lib.opo_run_round_trips_noalloc(C.byref(g), C.byref(p), 0, None, dummy, dummy, dummy,
                                           1, None) == OPO_ERR_ARG
    assert lib.opo_run_round_trips_noalloc(C.byref(g), C.byref(p), 0, dummy, dummy, dummy, dummy,
                                           0, None) == OPO_ERR_ARG
    if lib.opo_device_count() == 0:
        assert lib.opo_run_round_trips_noalloc(C.byref(g), C.byref(p), 0, dummy, dummy, dummy,
                                               dummy, 10, None) == OPO_ERR_NOGPU
