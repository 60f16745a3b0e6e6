"""Writes the golden vectors under tests/golden/ (run from the repo root).

fft_golden.npz   seeded complex inputs and their DFTs from numpy's pocketfft
                 (an FFT implementation independent of the oracle's), for the
                 transform lengths the path uses.
rk4_golden.npz   one RK4 half step of seeded (Ap, As, Ai) cell triples,
                 evaluated in 50-digit mpmath arithmetic: the reference's
                 stage points A + 0.5 k1, A + 0.5 k2, A + k3
                 (Solver3D.cuh:915, 921, 927) and the textbook ones
                 A + h/2 k1, A + h/2 k2, A + h k3, with the right-hand side
                 i kp As Ai, i ks Ap Ai*, i ki Ap As* (Solver3D.cuh:375-377)
                 and the update A + (k1 + 2 k2 + 2 k3 + k4) h / 6.
"""
import os

import mpmath
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def fft_golden():
    rng = np.random.default_rng(20260101)
    out = {}
    for n in (1, 2, 3, 8, 12, 16, 32, 64, 96, 128, 256, 512, 1024):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"x{n}"] = x
        out[f"fwd{n}"] = np.fft.fft(x)
        out[f"inv{n}"] = np.fft.ifft(x) * n
    # a 3D array: unnormalised xy transform per t-plane and t transform per pixel
    x3 = rng.standard_normal((16, 8, 8)) + 1j * rng.standard_normal((16, 8, 8))
    out["x3d"] = x3
    out["fft2_xy"] = np.fft.fft2(x3, axes=(1, 2))
    out["ifft_t"] = np.fft.ifft(x3, axis=0) * 16
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)


def rk4_golden():
    mpmath.mp.dps = 50
    rng = np.random.default_rng(20260102)
    n = 48
    A = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    A[:, 0] *= 0.5
    A[:, 1:] *= 0.2
    kap = np.array([7.430813e-3, 3.847380e-3, 3.1e-3])
    h = 0.75

    def rhs(a):
        I = mpmath.mpc(0, 1)
        kp, ks, ki = (mpmath.mpf(float(k)) for k in kap)
        return (I * kp * a[1] * a[2], I * ks * a[0] * mpmath.conj(a[2]),
                I * ki * a[0] * mpmath.conj(a[1]))

    def step(a, c1, c3, hh):
        k1 = rhs(a)
        k2 = rhs([a[j] + c1 * k1[j] for j in range(3)])
        k3 = rhs([a[j] + c1 * k2[j] for j in range(3)])
        k4 = rhs([a[j] + c3 * k3[j] for j in range(3)])
        return [a[j] + (k1[j] + 2 * k2[j] + 2 * k3[j] + k4[j]) * hh / 6 for j in range(3)]

    ref = np.zeros_like(A)
    txt = np.zeros_like(A)
    hh = mpmath.mpf(h)
    for r in range(n):
        a = [mpmath.mpc(float(A[r, j].real), float(A[r, j].imag)) for j in range(3)]
        b = step(a, mpmath.mpf("0.5"), mpmath.mpf(1), hh)
        c = step(a, hh / 2, hh, hh)
        ref[r] = [complex(v) for v in b]
        txt[r] = [complex(v) for v in c]
    np.savez_compressed(os.path.join(HERE, "rk4_golden.npz"), A=A, kappa=kap, h=h,
                        reference=ref, textbook=txt)


if __name__ == "__main__":
    fft_golden()
    rk4_golden()
    print("wrote", os.listdir(HERE))
