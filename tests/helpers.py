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
