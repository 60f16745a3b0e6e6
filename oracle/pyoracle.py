"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker of the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    vp, ip, dp = C.c_void_p, C.c_int, C.POINTER(C.c_double)
    sigs = {
        "oracle_create": (vp, [vp, vp, ip]),
        "oracle_destroy": (None, [vp]),
        "oracle_set_threads": (None, [ip]),
        "oracle_get_threads": (ip, []),
        "oracle_set_fields": (None, [vp, dp, dp, dp]),
        "oracle_get_fields": (None, [vp, dp, dp, dp]),
        "oracle_step_nonlinear": (None, [vp]),
        "oracle_step_dispersion": (None, [vp]),
        "oracle_step_diffraction": (None, [vp]),
        "oracle_step_free_space": (None, [vp]),
        "oracle_step_gdd": (None, [vp]),
        "oracle_step_mirror": (None, [vp]),
        "oracle_step_eom": (None, [vp]),
        "oracle_ssfm": (None, [vp]),
        "oracle_run_round_trips": (ip, [vp, ip, dp]),
        "oracle_power": (None, [vp, dp]),
        "oracle_table": (None, [vp, ip, dp]),
        "oracle_omega": (None, [vp, dp]),
        "oracle_fft": (None, [ip, ip, dp, dp, ip]),
        "oracle_dft_slow": (None, [ip, ip, dp, dp]),
        "oracle_slice_count": (ip, [C.c_double, ip]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def set_threads(n):
    load().oracle_set_threads(int(n))


def get_threads():
    return load().oracle_get_threads()


def fft(x, sign, precision=1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    load().oracle_fft(len(x), sign, _dp(x.view(np.float64)), _dp(out.view(np.float64)), precision)
    return out


def dft_slow(x, sign):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    load().oracle_dft_slow(len(x), sign, _dp(x.view(np.float64)), _dp(out.view(np.float64)))
    return out


def slice_count(lcr, nz):
    return load().oracle_slice_count(lcr, nz)


class Oracle:
    """FP64 (precision=1) or FP32 (precision=0) CPU restatement of one solve."""

    def __init__(self, grid, params, precision=1):
        self.lib = load()
        self.shape = (grid.nt, grid.ny, grid.nx)
        self.grid = grid
        self.h = self.lib.oracle_create(C.byref(grid), C.byref(params), precision)
        if not self.h:
            raise RuntimeError("oracle_create failed")

    def close(self):
        if self.h:
            self.lib.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_fields(self, api=None, as_=None, ai=None):
        bufs = [None if a is None else np.ascontiguousarray(a, dtype=np.complex128).reshape(self.shape)
                for a in (api, as_, ai)]
        self.lib.oracle_set_fields(self.h, *[None if b is None else _dp(b.view(np.float64))
                                             for b in bufs])

    def get_fields(self):
        out = [np.empty(self.shape, dtype=np.complex128) for _ in range(3)]
        self.lib.oracle_get_fields(self.h, *[_dp(b.view(np.float64)) for b in out])
        return out

    def nonlinear(self):
        self.lib.oracle_step_nonlinear(self.h)

    def dispersion(self):
        self.lib.oracle_step_dispersion(self.h)

    def diffraction(self):
        self.lib.oracle_step_diffraction(self.h)

    def free_space(self):
        self.lib.oracle_step_free_space(self.h)

    def gdd(self):
        self.lib.oracle_step_gdd(self.h)

    def mirror(self):
        self.lib.oracle_step_mirror(self.h)

    def eom(self):
        self.lib.oracle_step_eom(self.h)

    def ssfm(self):
        self.lib.oracle_ssfm(self.h)

    def run_round_trips(self, n):
        pw = np.zeros((n, 3))
        done = self.lib.oracle_run_round_trips(self.h, n, _dp(pw))
        return pw[:done]

    def power(self):
        out = np.zeros(3)
        self.lib.oracle_power(self.h, _dp(out))
        return out

    def table(self, which):
        n = self.grid.nt if which in (0, 1, 2, 8) else self.grid.nx * self.grid.ny
        out = np.empty(n, dtype=np.complex128)
        self.lib.oracle_table(self.h, which, _dp(out.view(np.float64)))
        return out

    def omega(self):
        out = np.empty(self.grid.nt)
        self.lib.oracle_omega(self.h, _dp(out))
        return out
