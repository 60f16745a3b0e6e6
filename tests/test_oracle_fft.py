"""Pins the oracle's FFT (oracle/oracle.cpp, Stockham mixed radix).

Against: numpy/pocketfft golden vectors (tests/golden/fft_golden.npz), an
O(n^2) long-double DFT, and the DFT's defining properties (Parseval,
linearity, shift theorem, impulse).  Sign convention: -1 is cuFFT's
FORWARD (e^{-i}), +1 its INVERSE (e^{+i}), unnormalised
(Solver3D.cuh:606-626, 757-777).
"""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz"))
SIZES = (1, 2, 3, 8, 12, 16, 32, 64, 96, 128, 256, 512, 1024)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("n", SIZES)
def test_fft_matches_pocketfft_golden(n):
    x = GOLD[f"x{n}"]
    assert rel(po.fft(x, -1), GOLD[f"fwd{n}"]) < 1e-14
    assert rel(po.fft(x, +1), GOLD[f"inv{n}"]) < 1e-14


@pytest.mark.parametrize("n", (5, 6, 7, 10, 24, 48, 96, 100, 128))
def test_fft_matches_long_double_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for s in (-1, 1):
        assert rel(po.fft(x, s), po.dft_slow(x, s)) < 1e-14


def test_fft_properties():
    rng = np.random.default_rng(7)
    n = 96
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    X = po.fft(x, -1)
    # Parseval
    assert abs(np.vdot(X, X).real / n - np.vdot(x, x).real) < 1e-12 * np.vdot(x, x).real
    # linearity
    a, b = 0.3 - 1.1j, 2.0 + 0.5j
    assert rel(po.fft(a * x + b * y, -1), a * X + b * po.fft(y, -1)) < 1e-14
    # shift theorem: x[j - m] -> X[k] e^{-2 pi i k m / n}
    m = 5
    k = np.arange(n)
    assert rel(po.fft(np.roll(x, m), -1), X * np.exp(-2j * np.pi * k * m / n)) < 1e-13
    # impulse at j0 -> e^{-2 pi i k j0 / n}; inverse sign conjugates
    e = np.zeros(n, complex)
    e[3] = 1
    ph = 2 * np.pi * ((k * 3) % n) / n  # reduced exactly so numpy's exp is accurate
    assert rel(po.fft(e, -1), np.exp(-1j * ph)) < 2e-15
    assert rel(po.fft(e, +1), np.exp(+1j * ph)) < 2e-15
    # forward then inverse returns n x
    assert rel(po.fft(X, +1), n * x) < 1e-14


def test_fp32_instance_accuracy():
    # the FP32 build of the same FFT stays at single-precision accuracy
    x = GOLD["x1024"]
    assert rel(po.fft(x, -1, precision=0), GOLD["fwd1024"]) < 2e-6
