"""Workload builders: physical description -> (Grid, Params, Api, As, Ai).

All set-up arithmetic runs in the library's FP64 host code
(include/opo3d_setup.h); nothing here touches a GPU.

Workloads (MgO:PPLN at 27 C, Lambda = 6.97 um, dQ = 14 pm/V, pump 532 nm,
signal and idler 1064 nm, 400 um transverse window):
  single_pass_32   32x32x64, 20 ps window, 5 mm crystal, NZ = 100, focused
                   Gaussian pump (w0 = 50 um, 2 ps FWHM, 1 W), single pass.
  cavity_64        64x64x256, singly resonant ring cavity (Lcav = 4 Lcr,
                   window = round-trip time), Rs = 0.98, idler not resonant,
                   GDD compensation 100 %, 200 round trips.
  sweep            64 cavity solves, grids 32^2x128 .. 128^2x512, parameters
                   drawn from a seed, 100 round trips each.
  femto_64         64x64x1024, 4 ps window, 150 fs sech^2 pump, 2 mm crystal,
                   NZ = 400, single pass.
  serving mix      grids {32^2x128, 64^2x256, 96^2x512, 128^2x1024}, 50-300
                   round trips.
"""
import ctypes as C
import math

import numpy as np

from . import capi
from .capi import Grid, Params, Physical, OPO_FP32, OPO_FP64, check, load

CRYSTAL_MGOPPLN = 0
CRYSTAL_PPLN = 1
CRYSTAL_MGOSPPLT = 2
CRYSTAL_ZGP = 3

PUMP_WAVEPLANE_CW = 0
PUMP_WAVEPLANE_PULSED = 1
PUMP_FOCUSED_CW = 2
PUMP_FOCUSED_PULSED = 3
PUMP_FOCUSED_SECH2 = 4

EPS0 = 8.8541878128e-12 * 1e12 / 1e6
CLIGHT = 299792458.0 * 1e6 / 1e12


def physical(lcr=5000.0, nz=100, lx=400.0, t_window=0.0, rs=0.98, ri=0.98, phase_s=0.0,
             phase_i=0.0, resonant_s=1, resonant_i=0, gdd_on=1, gdd_gamma=1.0, fs_on=1,
             n_round_trips=200, precision=OPO_FP32, dq=14e-6, lcav=None):
    ph = Physical()
    ph.crystal = CRYSTAL_MGOPPLN
    ph.temperature = 27.0
    ph.grating_period = 6.97
    ph.lp = 0.532
    ph.ls = 2 * 0.532
    ph.li = ph.ls * ph.lp / (ph.ls - ph.lp)
    ph.lx = lx
    ph.ly = lx
    ph.lcr = lcr
    ph.nz = nz
    ph.dq = dq
    ph.lcav = 4.0 * lcr if lcav is None else lcav
    ph.rs, ph.ri = rs, ri
    ph.phase_s, ph.phase_i = phase_s, phase_i
    ph.resonant_s, ph.resonant_i = resonant_s, resonant_i
    ph.gdd_gamma, ph.gdd_on = gdd_gamma, gdd_on
    ph.fs_on = fs_on
    ph.n_round_trips = n_round_trips
    ph.t_window = t_window
    ph.precision = precision
    return ph


def make_params(grid, ph):
    lib = load()
    p = Params()
    check(lib.opo_setup_params(C.byref(grid), C.byref(ph), C.byref(p)), "opo_setup_params")
    return p


def peak_field(ph, power, waist):
    """Ap0 = sqrt(4 P/(eps0 c pi np w0^2)) (EFields3D.cuh:252)."""
    cr = crystal(ph)
    return math.sqrt(4 * power / (EPS0 * CLIGHT * math.pi * cr.n[0] * waist * waist))


def crystal(ph):
    lib = load()
    cr = capi.Crystal()
    check(lib.opo_crystal_properties(ph.crystal, ph.temperature, ph.grating_period, ph.lp, ph.ls,
                                     ph.li, C.byref(cr)), "opo_crystal_properties")
    return cr


def make_fields(grid, params, ph, mode=PUMP_FOCUSED_PULSED, power=1.0, fwhm=2.0, waist=50.0,
                focal=None, noise_rel=1e-10, seed=1, precision=None):
    """Api from the pump mode; As noise (sigma = noise_rel * Ap0); Ai = As (cuOPO3D.cu:51)."""
    lib = load()
    prec = params.precision if precision is None else precision
    dt = np.complex128 if prec == OPO_FP64 else np.complex64
    shape = (grid.nt, grid.ny, grid.nx)
    api = np.empty(shape, dtype=dt)
    focal = ph.lcr / 2 if focal is None else focal
    check(lib.opo_setup_pump(C.byref(grid), C.byref(params), C.byref(ph), mode, power, fwhm,
                             waist, focal, prec, api.ctypes.data_as(C.c_void_p)),
          "opo_setup_pump")
    sigma = noise_rel * peak_field(ph, power, waist)
    as_ = np.empty(shape, dtype=dt)
    check(lib.opo_setup_noise(C.byref(grid), C.c_uint64(seed), sigma, prec,
                              as_.ctypes.data_as(C.c_void_p)), "opo_setup_noise")
    return api, as_, as_.copy()


def single_pass_32(precision=OPO_FP32, n=(32, 32, 64)):
    """Plain single-pass degenerate OPO, 32x32x64, 5 mm, NZ = 100."""
    grid = Grid(n[0], n[1], n[2])
    ph = physical(t_window=20.0, n_round_trips=1, precision=precision)
    p = make_params(grid, ph)
    api, as_, ai = make_fields(grid, p, ph)
    return grid, p, ph, (api, as_, ai)


def cavity_64(precision=OPO_FP32, n=(64, 64, 256), n_round_trips=200):
    """Singly-resonant ring cavity, 64x64x256, Rs = 0.98, idler not resonant."""
    grid = Grid(n[0], n[1], n[2])
    ph = physical(n_round_trips=n_round_trips, precision=precision)
    p = make_params(grid, ph)
    api, as_, ai = make_fields(grid, p, ph)
    return grid, p, ph, (api, as_, ai)


def femto_64(precision=OPO_FP32, n=(64, 64, 1024)):
    """Femtosecond single pass: 4 ps window, 150 fs sech^2 pump, 2 mm, NZ = 400."""
    grid = Grid(n[0], n[1], n[2])
    ph = physical(lcr=2000.0, nz=400, t_window=4.0, n_round_trips=1, precision=precision)
    p = make_params(grid, ph)
    api, as_, ai = make_fields(grid, p, ph, mode=PUMP_FOCUSED_SECH2, fwhm=0.150)
    return grid, p, ph, (api, as_, ai)


SWEEP_NXY = (32, 64, 96, 128)
SWEEP_NT = (128, 256, 512)


def sweep_draws(n_solves=64, seed=3):
    """Per-solve draws of the batched sweep: grid and (P, w0, phi, R)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_solves):
        nxy = int(rng.choice(SWEEP_NXY))
        nt = int(rng.choice(SWEEP_NT))
        P = float(rng.uniform(0.5, 5.0))
        w0 = float(rng.uniform(30.0, 80.0))
        phi = float(rng.uniform(-math.pi, math.pi))
        R = float(rng.uniform(0.9, 0.99))
        out.append(dict(nxy=nxy, nt=nt, power=P, waist=w0, phase=phi, R=R, seed=1000 + i))
    return out


def sweep_solve(d, n_round_trips=100, precision=OPO_FP32):
    grid = Grid(d["nxy"], d["nxy"], d["nt"])
    ph = physical(rs=d["R"], phase_s=d["phase"], n_round_trips=n_round_trips, precision=precision)
    p = make_params(grid, ph)
    fields = make_fields(grid, p, ph, power=d["power"], waist=d["waist"], seed=d["seed"])
    return grid, p, ph, fields


SERVING_GRIDS = ((32, 128), (64, 256), (96, 512), (128, 1024))
