"""ctypes view of the C ABI declared in include/opo3d.h.

The structures mirror ``opo_grid`` and ``opo_params`` field for field.  The
library is loaded from this package directory (``libopo3d.so``); a missing
library is an error, never a silent fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libopo3d.so")

OPO_OK = 0
OPO_ERR_ARG = -1
OPO_ERR_HIP = -2
OPO_ERR_NOMEM = -3
OPO_ERR_UNSUPPORTED = -4
OPO_ERR_NAN = -5
OPO_ERR_BUDGET = -6
OPO_ERR_STATE = -7
OPO_ERR_NOGPU = -8

OPO_FP32 = 0
OPO_FP64 = 1
OPO_TUNE_TIME_PASS = 0
OPO_TUNE_XY_PASS = 1
OPO_NOALLOC_SCRATCH_BYTES = 1 << 20   # include/opo3d.h

STAGES = ("nonlinear", "time_op", "xy_op", "free_space", "gdd", "power")


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nt", C.c_int)]


class Params(C.Structure):
    _fields_ = [
        ("lx", C.c_double), ("ly", C.c_double),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("n_slices", C.c_int),
        ("lcr", C.c_double),
        ("kappa", C.c_double * 3),
        ("kdiff", C.c_double * 3),
        ("delta", C.c_double * 3),
        ("beta2", C.c_double * 3),
        ("alpha", C.c_double * 3),
        ("t_window", C.c_double),
        ("omega", C.POINTER(C.c_double)),
        ("dk", C.c_double),
        ("dk_phase_scale", C.c_double),
        ("rk4_h_in_stages", C.c_int),
        ("fs_on", C.c_int),
        ("fs_length", C.c_double),
        ("skip_fs_on_last_trip", C.c_int),
        ("gdd_on", C.c_int),
        ("gdd", C.c_double),
        ("fb_re", C.c_double * 2),
        ("fb_im", C.c_double * 2),
        ("n_round_trips", C.c_int),
        ("power_prefactor", C.c_double * 3),
        ("precision", C.c_int),
        ("eom_on", C.c_int),
        ("eom_beta", C.c_double),
        ("eom_fpm", C.c_double),
        ("no_diffraction", C.c_int),
    ]

    def copy(self):
        other = Params()
        C.pointer(other)[0] = self
        return other


class Physical(C.Structure):
    """Mirror of ``opo_physical`` (include/opo3d_setup.h)."""
    _fields_ = [
        ("crystal", C.c_int),
        ("temperature", C.c_double),
        ("grating_period", C.c_double),
        ("lp", C.c_double), ("ls", C.c_double), ("li", C.c_double),
        ("lx", C.c_double), ("ly", C.c_double), ("lcr", C.c_double),
        ("nz", C.c_int),
        ("dq", C.c_double),
        ("lcav", C.c_double),
        ("rs", C.c_double), ("ri", C.c_double),
        ("phase_s", C.c_double), ("phase_i", C.c_double),
        ("resonant_s", C.c_int), ("resonant_i", C.c_int),
        ("gdd_gamma", C.c_double), ("gdd_on", C.c_int),
        ("fs_on", C.c_int),
        ("n_round_trips", C.c_int),
        ("t_window", C.c_double),
        ("precision", C.c_int),
        ("pol", C.c_char * 4),
    ]


class Crystal(C.Structure):
    """Mirror of ``opo_crystal_props`` (include/opo3d_setup.h)."""
    _fields_ = [
        ("n", C.c_double * 3), ("vg", C.c_double * 3),
        ("b2", C.c_double * 3), ("b3", C.c_double * 3),
        ("alpha", C.c_double * 3),
        ("d33", C.c_double), ("dq", C.c_double),
        ("kp", C.c_double), ("ks", C.c_double), ("ki", C.c_double),
        ("dk", C.c_double),
    ]


class StageTime(C.Structure):
    """opo_stage_time (include/opo3d.h)."""
    _fields_ = [("stage", C.c_int), ("size", C.c_int), ("launches", C.c_int),
                ("ms", C.c_double), ("alg_bytes", C.c_double)]


_lib = None


def load():
    """Load libopo3d.so from the package directory (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(LIB_PATH)
    vp, ip, sz = C.c_void_p, C.c_int, C.c_size_t
    dp = C.POINTER(C.c_double)
    sig = {
        "opo_last_error": (C.c_char_p, []),
        "opo_version": (C.c_char_p, []),
        "opo_device_count": (ip, []),
        "opo_create": (ip, [C.POINTER(Grid), C.POINTER(Params), ip, C.POINTER(vp)]),
        "opo_destroy": (None, [vp]),
        "opo_set_fields": (ip, [vp, vp, vp, vp]),
        "opo_get_fields": (ip, [vp, vp, vp, vp]),
        "opo_field_device_ptr": (vp, [vp, ip]),
        "opo_field_bytes": (sz, [vp]),
        "opo_step_nonlinear": (ip, [vp]),
        "opo_step_dispersion": (ip, [vp]),
        "opo_step_diffraction": (ip, [vp]),
        "opo_step_free_space": (ip, [vp]),
        "opo_step_gdd": (ip, [vp]),
        "opo_step_mirror": (ip, [vp]),
        "opo_ssfm": (ip, [vp]),
        "opo_single_pass": (ip, [vp]),
        "opo_run_round_trips": (ip, [vp, ip, dp]),
        "opo_run_round_trips_noalloc": (ip, [C.POINTER(Grid), C.POINTER(Params), ip, vp, vp, vp, vp,
                                             ip, dp]),
        "opo_trips_done": (ip, [vp]),
        "opo_power": (ip, [vp, dp]),
        "opo_power_time": (ip, [vp, C.c_int, C.c_int, C.c_int, C.c_double, dp]),
        "opo_courant_ok": (ip, [C.POINTER(Params)]),
        "opo_sync": (ip, [vp]),
        "opo_profile_trips": (ip, [vp, ip, dp, C.POINTER(ip)]),
        "opo_autotune": (ip, [vp, ip, C.c_char_p, sz]),
        "opo_tune_get": (ip, [ip, ip, ip]),
        "opo_tune_set": (ip, [ip, ip, ip, ip]),
        "opo_batch_create": (ip, [ip, sz, ip, C.POINTER(vp)]),
        "opo_batch_destroy": (None, [vp]),
        "opo_batch_submit": (ip, [vp, C.POINTER(Grid), C.POINTER(Params), vp, vp, vp, ip,
                                  C.POINTER(C.c_uint64)]),
        "opo_batch_step": (ip, [vp, C.POINTER(ip)]),
        "opo_batch_poll": (ip, [vp, C.c_uint64, C.POINTER(ip), dp, vp, vp, vp]),
        "opo_batch_footprint": (sz, [C.POINTER(Grid), ip, ip]),
        "opo_batch_stats": (ip, [vp, C.POINTER(ip), C.POINTER(ip), C.POINTER(sz)]),
        "opo_batch_profile": (ip, [vp, ip]),
        "opo_batch_concurrent_classes": (ip, [vp, ip]),
        "opo_batch_autotune": (ip, [vp, ip, C.c_char_p, sz]),
        "opo_batch_rebuilds": (ip, [vp, C.POINTER(ip), C.POINTER(ip), dp]),
        "opo_batch_profile_read": (ip, [vp, C.POINTER(StageTime), ip, C.POINTER(ip)]),
        "opo_batch_time_stages": (ip, [vp, ip, C.POINTER(StageTime), ip, C.POINTER(ip)]),
        # host set-up (include/opo3d_setup.h)
        "opo_slice_count": (ip, [C.c_double, ip]),
        "opo_crystal_properties": (ip, [ip, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.POINTER(Crystal)]),
        "opo_crystal_properties_pol": (ip, [ip, C.c_double, C.c_double, C.c_double, C.c_double,
                                            C.c_double, C.c_char_p, C.POINTER(Crystal)]),
        "opo_setup_params": (ip, [C.POINTER(Grid), C.POINTER(Physical), C.POINTER(Params)]),
        "opo_setup_pump": (ip, [C.POINTER(Grid), C.POINTER(Params), C.POINTER(Physical),
                                ip, C.c_double, C.c_double, C.c_double, C.c_double, ip, vp]),
        "opo_setup_noise": (ip, [C.POINTER(Grid), C.c_uint64, C.c_double, ip, vp]),
        "opo_time_grid": (ip, [ip, C.c_double, dp]),
        "opo_omega_grid": (ip, [ip, C.c_double, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)  # AttributeError if the symbol is missing
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    """Names of the functions this binding requires from the library."""
    import re
    names = []
    hdr_dir = os.path.join(os.path.dirname(HERE), "include")
    for h in sorted(os.listdir(hdr_dir)):
        if h.endswith(".h"):
            txt = open(os.path.join(hdr_dir, h)).read()
            names += re.findall(r"\b(opo_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def check(rc, what=""):
    if rc != OPO_OK:
        msg = load().opo_last_error()
        raise RuntimeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
    return rc
