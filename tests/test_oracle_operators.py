"""Pins the oracle's linear operators and round-trip sequence.

D (Solver3D.cuh:153-169, 571-639), Q (:282-301, 725-788), free space
(:305-323, 817-869), GDD (:261-277, 643-694), mirror (Cavity3D.cuh:12-28),
power (:529-545) and the trip order (:979-1011) are checked against closed
forms of the continuous problem (Gaussian pulse under group delay and GVD,
Gaussian beam under paraxial diffraction), exact invariants (unitarity,
back-propagation), and hand-evaluated values of the reference's omega rule.
"""
import numpy as np
import pytest

from cuopo3d_amd.capi import Grid, Params
from oracle.pyoracle import Oracle


def simple_params(nt=1, nx=1, ny=1, dx=1.0, dz=1.0, n_slices=1, kdiff=(10.0, 10.0, 10.0),
                  beta2=(0.0, 0.0, 0.0), delta=(0.0, 0.0, 0.0), alpha=(0.0, 0.0, 0.0),
                  kappa=(0.0, 0.0, 0.0), t_window=1.0, omega=None):
    p = Params()
    p.lx = p.ly = dx * max(nx - 1, 1)
    p.dx = p.dy = dx
    p.dz = dz
    p.n_slices = n_slices
    p.lcr = dz * n_slices
    for f in range(3):
        p.kappa[f], p.kdiff[f], p.delta[f] = kappa[f], kdiff[f], delta[f]
        p.beta2[f], p.alpha[f] = beta2[f], alpha[f]
        p.power_prefactor[f] = 1.0
    p.t_window = t_window
    p.fb_re[0] = p.fb_re[1] = 1.0
    p.n_round_trips = 1
    p.precision = 1
    keep = None
    if omega is not None:
        keep = np.ascontiguousarray(omega, dtype=np.float64)
        import ctypes as C
        p.omega = keep.ctypes.data_as(C.POINTER(C.c_double))
    return Grid(nx, ny, nt), p, keep


def dft_omega(nt, dt):
    """Angular frequency of bin k of the +i transform on the lattice t_j = j dt."""
    k = np.arange(nt)
    kk = np.where(k < nt // 2, k, k - nt)
    return 2 * np.pi * kk / (nt * dt)


def test_omega_rule_hand_values():
    # EFields3D.cuh:323-325: F = linspace(-NT/(2W), NT/(2W), NT), w = 2 pi fftshift(F)
    g, p, _ = simple_params(nt=4, t_window=1.0)
    o = Oracle(g, p)
    assert np.allclose(o.omega(), 2 * np.pi * np.array([2 / 3, 2.0, -2.0, -2 / 3]), rtol=1e-15)
    g, p, _ = simple_params(nt=8, t_window=2.0)
    F = -2.0 + np.arange(8) * 4.0 / 7.0
    assert np.allclose(Oracle(g, p).omega(), 2 * np.pi * np.r_[F[4:], F[:4]], rtol=1e-15)
    g, p, _ = simple_params(nt=1, t_window=1.0)
    assert Oracle(g, p).omega()[0] == 0.0   # size-1 linspace guard


def gaussian_pulse_after(t, T0, b, tau):
    # A(t) = exp(-t^2/(2 T0^2)) under e^{i w tau} e^{i b w^2 / 2} in the +i convention
    q = T0 ** 2 - 1j * b
    return T0 / np.sqrt(q) * np.exp(-((t - tau) ** 2) / (2 * q))


@pytest.mark.parametrize("field", [0, 1, 2])
def test_dispersion_gaussian_pulse_closed_form(field):
    nt, dt, T0 = 512, 0.05, 0.6
    t = (np.arange(nt) - nt // 2) * dt
    dz = 3.0
    beta2 = [0.0, 0.0, 0.0]
    delta = [0.0, 0.0, 0.0]
    alpha = [0.0, 0.0, 0.0]
    beta2[field] = 0.02 * (1 + field)
    delta[field] = 0.15 * (1 - field)
    alpha[field] = 0.01
    g, p, keep = simple_params(nt=nt, dz=dz, beta2=beta2, delta=delta, alpha=alpha,
                               omega=dft_omega(nt, dt))
    a0 = np.exp(-t ** 2 / (2 * T0 ** 2)).astype(complex).reshape(nt, 1, 1)
    o = Oracle(g, p)
    o.set_fields(a0, a0, a0)
    o.dispersion()
    out = o.get_fields()[field].ravel()
    ref = gaussian_pulse_after(t, T0, beta2[field] * dz, delta[field] * dz) * np.exp(-0.5 * alpha[field] * dz)
    assert np.abs(out - ref).max() < 1e-12


def test_gdd_gaussian_pulse_closed_form():
    nt, dt, T0, gdd = 512, 0.05, 0.5, 0.07
    t = (np.arange(nt) - nt // 2) * dt
    g, p, keep = simple_params(nt=nt, omega=dft_omega(nt, dt))
    p.gdd, p.gdd_on = gdd, 1
    a0 = np.exp(-t ** 2 / (2 * T0 ** 2)).astype(complex).reshape(nt, 1, 1)
    o = Oracle(g, p)
    o.set_fields(a0, a0, a0)
    o.gdd()
    f = o.get_fields()
    assert np.abs(f[1].ravel() - gaussian_pulse_after(t, T0, gdd, 0.0)).max() < 1e-12
    assert np.abs(f[2].ravel() - gaussian_pulse_after(t, T0, gdd, 0.0)).max() < 1e-12
    assert np.array_equal(f[0], a0)   # the pump is not compensated


def gaussian_beam(x, y, w0, k, z):
    # paraxial dA/dz = (i / 2k) lap A, A(z=0) = exp(-r^2/w0^2)
    zr = k * w0 ** 2 / 2
    q = 1 + 1j * z / zr
    return np.exp(-(x ** 2 + y ** 2) / (w0 ** 2 * q)) / q


def beam_grid(n, dx):
    c = (np.arange(n) - n // 2) * dx   # centred on index NX/2 as the reference (uX = NX/2)
    return np.meshgrid(c, c, indexing="xy")


@pytest.mark.parametrize("n", [64, 96, 128])
def test_diffraction_gaussian_beam_closed_form(n):
    dx, w0, dz = 4.0, 14.0, 300.0
    k = (25.0, 20.0, 15.0)
    alpha = (2e-4, 0.0, 1e-4)
    g, p, _ = simple_params(nx=n, ny=n, dx=dx, dz=dz, kdiff=k, alpha=alpha)
    X, Y = beam_grid(n, dx)
    a0 = np.exp(-(X ** 2 + Y ** 2) / w0 ** 2).astype(complex).reshape(1, n, n)
    o = Oracle(g, p)
    o.set_fields(a0, a0, a0)
    o.diffraction()
    out = o.get_fields()
    for f in range(3):
        ref = gaussian_beam(X, Y, w0, k[f], dz) * np.exp(-0.5 * alpha[f] * dz)
        assert np.abs(out[f][0] - ref).max() < 1e-11, f


def test_free_space_is_back_propagation():
    # R12: exp(+i L 2 pi^2 f^2 / k): a beam propagated by -L; no absorption; s and i only
    n, dx, w0, L = 64, 4.0, 14.0, 250.0
    g, p, _ = simple_params(nx=n, ny=n, dx=dx, kdiff=(30.0, 20.0, 15.0), alpha=(1e-3,) * 3)
    p.fs_length, p.fs_on = L, 1
    X, Y = beam_grid(n, dx)
    a0 = np.exp(-(X ** 2 + Y ** 2) / w0 ** 2).astype(complex).reshape(1, n, n)
    o = Oracle(g, p)
    o.set_fields(a0, a0, a0)
    o.free_space()
    out = o.get_fields()
    assert np.array_equal(out[0], a0)
    assert np.abs(out[1][0] - gaussian_beam(X, Y, w0, 20.0, -L)).max() < 1e-11
    assert np.abs(out[2][0] - gaussian_beam(X, Y, w0, 15.0, -L)).max() < 1e-11


def test_unitarity_without_absorption():
    rng = np.random.default_rng(3)
    g, p, _ = simple_params(nx=32, ny=32, nt=64, dx=3.0, dz=40.0, beta2=(1e-3, 2e-3, 3e-3),
                            delta=(0.01, 0.0, -0.01), t_window=20.0)
    a = rng.standard_normal((64, 32, 32)) + 1j * rng.standard_normal((64, 32, 32))
    o = Oracle(g, p)
    o.set_fields(a, a, a)
    o.dispersion()
    o.diffraction()
    for f in o.get_fields():
        assert abs(np.linalg.norm(f) / np.linalg.norm(a) - 1) < 1e-13


def test_linear_round_trip_invariant():
    # kappa = 0, alpha = 0: S slices of Q(dz) then the free space Q_fs(L) equal
    # Q(S dz - L); with D and GDD off the time axis is untouched.
    n, dx, w0, dz, S = 64, 4.0, 14.0, 50.0, 7
    L = 200.0
    g, p, _ = simple_params(nx=n, ny=n, nt=4, dx=dx, dz=dz, n_slices=S, kdiff=(25.0, 20.0, 20.0),
                            t_window=10.0)
    p.fs_on, p.fs_length = 1, L
    p.skip_fs_on_last_trip = 0
    p.n_round_trips = 2
    X, Y = beam_grid(n, dx)
    beam = np.exp(-(X ** 2 + Y ** 2) / w0 ** 2)
    a0 = np.broadcast_to(beam, (4, n, n)).astype(complex)
    o = Oracle(g, p)
    o.set_fields(a0, a0, a0)
    pw = o.run_round_trips(1)
    out = o.get_fields()
    ref = gaussian_beam(X, Y, w0, 20.0, S * dz - L)
    for tt in range(4):
        assert np.abs(out[1][tt] - ref).max() < 1e-11
    # the pump is re-injected and propagated S dz without the free-space step
    assert np.abs(out[0][0] - gaussian_beam(X, Y, w0, 25.0, S * dz)).max() < 1e-11
    # per-trip power: pref * sum |A|^2 (pref = 1 here), unitary evolution keeps it
    assert abs(pw[0, 1] - 4 * np.sum(np.abs(beam) ** 2)) < 1e-9 * pw[0, 1]


def test_free_space_skipped_on_last_trip():
    # R17: no free space on trip NRT-1
    n = 32
    g, p, _ = simple_params(nx=n, ny=n, dx=4.0, dz=50.0, n_slices=3, kdiff=(25.0, 20.0, 20.0))
    p.fs_on, p.fs_length, p.skip_fs_on_last_trip, p.n_round_trips = 1, 150.0, 1, 1
    X, Y = beam_grid(n, 4.0)
    a0 = np.exp(-(X ** 2 + Y ** 2) / 12.0 ** 2).astype(complex).reshape(1, n, n)
    o = Oracle(g, p)
    o.set_fields(a0, a0, a0)
    o.run_round_trips(1)
    assert np.abs(o.get_fields()[1][0] - gaussian_beam(X, Y, 12.0, 20.0, 150.0)).max() < 1e-11


def test_mirror_and_power():
    g, p, _ = simple_params(nx=4, ny=4, nt=8)
    p.fb_re[0], p.fb_im[0] = np.sqrt(0.9) * np.cos(0.4), np.sqrt(0.9) * np.sin(0.4)
    p.fb_re[1], p.fb_im[1] = 1.0, 0.0
    p.power_prefactor[0], p.power_prefactor[1], p.power_prefactor[2] = 2.0, 3.0, 5.0
    c = 0.25 - 0.5j
    a = np.full((8, 4, 4), c)
    o = Oracle(g, p)
    o.set_fields(a, a, a)
    o.mirror()
    f = o.get_fields()
    assert np.allclose(f[1], np.sqrt(0.9) * np.exp(0.4j) * c, rtol=0, atol=1e-16)
    assert np.array_equal(f[2], a) and np.array_equal(f[0], a)
    pw = o.power()
    n = 128 * abs(c) ** 2
    assert np.allclose(pw, [2 * n, 3 * 0.9 * n, 5 * n], rtol=1e-15)


def test_nan_stops_round_trips():
    g, p, _ = simple_params(nx=4, ny=4, nt=4, t_window=1.0)
    p.n_round_trips = 10
    a = np.ones((4, 4, 4), complex)
    s = a.copy()
    s[0, 0, 0] = np.nan
    o = Oracle(g, p)
    o.set_fields(a, s, a)
    pw = o.run_round_trips(5)
    assert pw.shape[0] == 1 and np.isnan(pw[0, 1])


def test_eom_closed_form():
    """KernelEOM (Solver1D.cuh:62-72): s and i times exp(i beta sin(2 pi fpm t))
    on t = linspace(-W/2, W/2, NT); the pump untouched."""
    nt, nx, W, beta, fpm = 64, 4, 20.0, 0.7, 0.3
    g, p, _ = simple_params(nt=nt, nx=nx, ny=nx, t_window=W)
    p.eom_on, p.eom_beta, p.eom_fpm = 1, beta, fpm
    rng = np.random.default_rng(21)
    f0 = [rng.standard_normal((nt, nx, nx)) + 1j * rng.standard_normal((nt, nx, nx)) for _ in range(3)]
    o = Oracle(g, p)
    o.set_fields(*f0)
    o.eom()
    f = o.get_fields()
    t = np.linspace(-W / 2, W / 2, nt).reshape(nt, 1, 1)
    m = np.exp(1j * beta * np.sin(2 * np.pi * fpm * t))
    assert np.array_equal(f[0], f0[0])
    assert np.abs(f[1] - m * f0[1]).max() < 1e-13
    assert np.abs(f[2] - m * f0[2]).max() < 1e-13


def test_trip_applies_eom_before_gdd():
    """A trip with no slices is Api -> Ap, then EOM, then GDD (Solver1D.cuh:467-468):
    the modulated pulse must come out dispersed, which EOM after GDD would not give."""
    nt, dt, T0, gdd, beta, fpm = 256, 0.05, 0.5, 0.05, 1.3, 0.8
    W = dt * (nt - 1)
    t = (np.arange(nt) - (nt - 1) / 2) * dt          # linspace(-W/2, W/2, NT)
    g, p, _ = simple_params(nt=nt, omega=dft_omega(nt, dt), t_window=W, n_slices=0)
    p.gdd, p.gdd_on = gdd, 1
    p.eom_on, p.eom_beta, p.eom_fpm = 1, beta, fpm
    a0 = np.exp(-t ** 2 / (2 * T0 ** 2)).astype(complex)
    o = Oracle(g, p)
    o.set_fields(a0.reshape(nt, 1, 1), a0.reshape(nt, 1, 1), a0.reshape(nt, 1, 1))
    o.run_round_trips(1)
    got = o.get_fields()[1].ravel()
    # closed-form order: modulate in time, then exp(i w^2 GDD / 2) on the spectrum
    m = a0 * np.exp(1j * beta * np.sin(2 * np.pi * fpm * t))
    w = dft_omega(nt, dt)
    want = np.fft.fft(np.fft.ifft(m) * np.exp(0.5j * w * w * gdd))
    assert np.abs(got - want).max() < 1e-12
    wrong = np.fft.fft(np.fft.ifft(a0) * np.exp(0.5j * w * w * gdd)) * np.exp(1j * beta * np.sin(2 * np.pi * fpm * t))
    assert np.abs(got - wrong).max() > 1e-3


def test_plane_wave_mode_equals_one_pixel_without_absorption():
    """no_diffraction (the plane-wave model, Solver1D.cuh:436-445) skips Q; on a
    1x1 grid without absorption Q is the identity, so both give the same fields."""
    nt = 64
    g, p, _ = simple_params(nt=nt, dz=5.0, n_slices=4, beta2=(1e-2, 2e-2, 3e-2),
                            delta=(0.1, 0.0, -0.1), kappa=(1e-3, 2e-3, 2e-3), t_window=10.0)
    rng = np.random.default_rng(4)
    f0 = [(rng.standard_normal(nt) + 1j * rng.standard_normal(nt)).reshape(nt, 1, 1) for _ in range(3)]
    outs = []
    for nd in (0, 1):
        p.no_diffraction = nd
        o = Oracle(g, p)
        o.set_fields(*f0)
        o.ssfm()
        outs.append(o.get_fields())
    for a, b in zip(*outs):
        assert np.abs(a - b).max() < 1e-13
    # with absorption the transverse step carries exp(-alpha dz/2) (R8), so they differ
    p.alpha[1] = 1e-3
    res = []
    for nd in (0, 1):
        p.no_diffraction = nd
        o = Oracle(g, p)
        o.set_fields(*f0)
        o.ssfm()
        res.append(o.get_fields()[1])
    assert np.abs(res[0] - res[1]).max() > 1e-6
