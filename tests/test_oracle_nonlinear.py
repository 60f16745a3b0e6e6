"""Pins the oracle's nonlinear half step (solverRK4, Solver3D.cuh:872-943).

Checks: 50-digit mpmath evaluations of both stage rules (golden vectors);
zero signal and idler leave every field unchanged; Manley-Rowe invariants
of the coupled-wave equations (a sign error in any of the three right-hand
sides breaks them); the undepleted-pump gain closed form, with and without
phase mismatch, and fourth-order convergence of the textbook scheme; the
reference's step-free stage points coincide with the textbook ones at h = 1;
the degenerate OPO keeps As and Ai bitwise equal.
"""
import os

import numpy as np
import pytest

from cuopo3d_amd import configs
from oracle.pyoracle import Oracle

from test_oracle_operators import simple_params

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "rk4_golden.npz"))


def cells(a, shape):
    return [np.ascontiguousarray(a[:, j]).reshape(shape) for j in range(3)]


@pytest.mark.parametrize("rule", ["reference", "textbook"])
def test_rk4_matches_mpmath_golden(rule):
    A = GOLD["A"]
    n = A.shape[0]
    g, p, _ = simple_params(nt=n, dz=2 * float(GOLD["h"]), kappa=tuple(GOLD["kappa"]))
    p.rk4_h_in_stages = 1 if rule == "textbook" else 0
    o = Oracle(g, p)
    o.set_fields(*cells(A, (n, 1, 1)))
    o.nonlinear()
    out = np.stack([f.ravel() for f in o.get_fields()], axis=1)
    ref = GOLD[rule]
    assert np.abs(out - ref).max() / np.abs(ref).max() < 1e-14
    # the step changes the fields well above the tolerance
    assert np.abs(ref - A).max() / np.abs(ref).max() > 1e-4


def test_zero_signal_and_idler_are_fixed_points():
    rng = np.random.default_rng(1)
    ap = rng.standard_normal((8, 1, 1)) + 1j * rng.standard_normal((8, 1, 1))
    z = np.zeros_like(ap)
    g, p, _ = simple_params(nt=8, dz=10.0, kappa=(1e-2, 1e-2, 1e-2))
    o = Oracle(g, p)
    o.set_fields(ap, z, z)
    o.nonlinear()
    f = o.get_fields()
    assert np.array_equal(f[0], ap) and not f[1].any() and not f[2].any()


def test_manley_rowe_textbook_rk4():
    rng = np.random.default_rng(2)
    n = 16
    A = [rng.standard_normal((n, 1, 1)) + 1j * rng.standard_normal((n, 1, 1)) for _ in range(3)]
    kp, ks, ki = 3e-2, 2e-2, 1.5e-2
    g, p, _ = simple_params(nt=n, dz=0.4, n_slices=50, kappa=(kp, ks, ki))
    p.rk4_h_in_stages = 1
    o = Oracle(g, p)
    o.set_fields(*A)
    o.ssfm()   # linear operators are identities on a 1x1xN grid with zero omega
    B = o.get_fields()
    m1 = lambda F: np.abs(F[0]) ** 2 / kp + np.abs(F[1]) ** 2 / ks
    m2 = lambda F: np.abs(F[0]) ** 2 / kp + np.abs(F[2]) ** 2 / ki
    change = np.abs(np.abs(B[1]) ** 2 - np.abs(A[1]) ** 2).max()
    assert change > 1e-2   # the fields do exchange energy
    assert (np.abs(m1(B) - m1(A)) / m1(A)).max() < 1e-9
    assert (np.abs(m2(B) - m2(A)) / m2(A)).max() < 1e-9


def undepleted(as0, ai0, ap, ks, ki, dk, z):
    g = np.sqrt(ks * ki * abs(ap) ** 2 - (dk / 2) ** 2 + 0j)
    ch, sh = np.cosh(g * z), np.sinh(g * z) / g
    a = ch * as0 + sh * (-0.5j * dk * as0 + 1j * ks * ap * np.conj(ai0))
    b = ch * ai0 + sh * (-0.5j * dk * ai0 + 1j * ki * ap * np.conj(as0))
    return a * np.exp(0.5j * dk * z), b * np.exp(0.5j * dk * z)


def gain_case(h, S):
    ap = np.full((1, 1, 1), 1.0 + 0.5j)
    as0 = np.full((1, 1, 1), 1e-7 * (0.3 - 0.8j))
    ai0 = np.full((1, 1, 1), 1e-7 * (-0.5 + 0.1j))
    ks, ki = 0.02, 0.015
    g, p, _ = simple_params(nt=1, dz=2 * h, n_slices=S, kappa=(0.01, ks, ki))
    p.rk4_h_in_stages = 1
    o = Oracle(g, p)
    o.set_fields(ap, as0, ai0)
    o.ssfm()
    f = o.get_fields()
    ra, rb = undepleted(as0[0, 0, 0], ai0[0, 0, 0], ap[0, 0, 0], ks, ki, 0.0, 2 * h * S)
    return abs(f[1].ravel()[0] - ra) / abs(ra), abs(f[2].ravel()[0] - rb) / abs(rb), abs(ra / as0.ravel()[0])


def test_undepleted_gain_closed_form_and_order():
    e1s, e1i, gain = gain_case(1.0, 40)     # z = 80, g z ~ 1.4
    assert gain > 2.0
    assert e1s < 1e-7 and e1i < 1e-7
    e2s, _, _ = gain_case(0.5, 80)
    assert 12 < e1s / e2s < 20   # fourth-order global convergence


def test_phase_mismatch_single_step_closed_form():
    # one RK4 step from z = 0 with the stage phases e^{+i dk z} (s, i), e^{-i dk z} (p)
    h, dk = 0.5, 0.4
    ap, as0, ai0 = 1.0 + 0.2j, 1e-7 * (0.2 + 0.9j), 1e-7 * (0.7 - 0.3j)
    ks, ki = 0.3, 0.25
    g, p, _ = simple_params(nt=1, dz=2 * h, kappa=(0.01, ks, ki))
    p.rk4_h_in_stages = 1
    p.dk, p.dk_phase_scale = dk, 1.0
    o = Oracle(g, p)
    o.set_fields(*[np.full((1, 1, 1), v) for v in (ap, as0, ai0)])
    o.nonlinear()
    f = o.get_fields()
    ra, rb = undepleted(as0, ai0, ap, ks, ki, dk, h)
    assert abs(f[1].ravel()[0] - ra) / abs(ra - as0) < 1e-3
    assert abs(f[2].ravel()[0] - rb) / abs(rb - ai0) < 1e-3
    # the opposite sign of the phase is far off
    rm, _ = undepleted(as0, ai0, ap, ks, ki, -dk, h)
    assert abs(f[1].ravel()[0] - rm) / abs(rm - as0) > 1e-2


def test_reference_stage_rule_equals_textbook_at_unit_step():
    rng = np.random.default_rng(4)
    A = [0.3 * (rng.standard_normal((6, 1, 1)) + 1j * rng.standard_normal((6, 1, 1))) for _ in range(3)]
    out = []
    for rule in (0, 1):
        g, p, _ = simple_params(nt=6, dz=2.0, kappa=(0.05, 0.04, 0.03))
        p.rk4_h_in_stages = rule
        o = Oracle(g, p)
        o.set_fields(*A)
        o.nonlinear()
        out.append(o.get_fields())
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_degenerate_opo_keeps_signal_equal_to_idler():
    grid, p, ph, fields = configs.single_pass_32()
    p.n_slices = 10
    o = Oracle(grid, p)
    o.set_fields(*fields)
    o.ssfm()
    f = o.get_fields()
    assert np.array_equal(f[1], f[2])
    assert not np.array_equal(f[1], fields[1])
