"""Host set-up of a solve (include/opo3d_setup.h), all on the CPU.

Crystal group velocity, GVD and TOD come from analytic wavelength
derivatives of the Sellmeier index (MgOPPLN.cuh:148-259); they are checked
against finite differences of the index itself.  Pump fields against their
closed forms (EFields3D.cuh:72-137, 247-253), the Philox generator against
the published known-answer vector, the slice-count rule against hand
evaluation of the float loop, and opo_setup_params against the reference
expressions it carries.
"""
import ctypes as C
import math

import numpy as np
import pytest

from cuopo3d_amd import configs
from cuopo3d_amd.capi import Crystal, Grid, load, check

LIB = load()
CLIGHT = 299.792458


def props(crystal, lam, T=27.0):
    cr = Crystal()
    check(LIB.opo_crystal_properties(crystal, T, 6.97, lam, 2 * 0.532, 2 * 0.532, C.byref(cr)), "props")
    return cr


@pytest.mark.parametrize("crystal", [configs.CRYSTAL_MGOPPLN, configs.CRYSTAL_PPLN,
                                     configs.CRYSTAL_MGOSPPLT])
@pytest.mark.parametrize("lam", [0.532, 1.064, 1.55])
def test_dispersion_from_finite_differences(crystal, lam):
    n = lambda L: props(crystal, L).n[0]
    h1, h2 = 1e-4, 5e-4
    d1 = (n(lam + h1) - n(lam - h1)) / (2 * h1)
    d2 = (n(lam + h2) - 2 * n(lam) + n(lam - h2)) / h2 ** 2
    cr = props(crystal, lam)
    assert abs(cr.vg[0] - CLIGHT / (n(lam) - lam * d1)) / cr.vg[0] < 1e-7
    b2 = lam ** 3 * d2 / (2 * math.pi * CLIGHT ** 2)
    assert abs(cr.b2[0] - b2) / abs(b2) < 1e-5
    # TOD follows the reference's d3ndl3 expression (MgOPPLN.cuh:206-234), which
    # is not the third derivative of n; it is carried for completeness only and
    # never enters the split-step path.
    assert math.isfinite(cr.b3[0])


def test_crystal_index_ranges():
    # extraordinary indices of LiNbO3 / LiTaO3 near room temperature
    assert 2.22 < props(configs.CRYSTAL_MGOPPLN, 0.532).n[0] < 2.23
    assert 2.14 < props(configs.CRYSTAL_MGOPPLN, 1.064).n[0] < 2.16
    assert 2.12 < props(configs.CRYSTAL_MGOSPPLT, 1.064).n[0] < 2.15
    # normal dispersion in the visible: beta2 > 0, vg(532) < vg(1064)
    cr = props(configs.CRYSTAL_MGOPPLN, 0.532)
    assert cr.b2[0] > cr.b2[1] > 0 and cr.vg[0] < cr.vg[1]


def test_setup_params_carry_reference_expressions():
    grid, p, ph, _ = configs.cavity_64()
    cr = configs.crystal(ph)
    ns, ni, np_ = cr.n[1], cr.n[2], cr.n[0]
    dq = 14e-6
    assert p.kappa[0] == pytest.approx(2 * math.pi * dq / (np_ * 0.532), rel=1e-15)
    assert p.kappa[1] == pytest.approx(2 * math.pi * dq / (ns * 1.064), rel=1e-15)
    assert p.kdiff[0] == pytest.approx(2 * math.pi * ns / 0.532, rel=1e-15)   # R6: ns for the pump
    assert p.kdiff[2] == pytest.approx(2 * math.pi * ni / 1.064, rel=1e-15)
    assert p.delta[0] == pytest.approx(1 / cr.vg[0] - 1 / cr.vg[2], rel=1e-15)  # R7
    assert p.delta[1] == 0.0
    assert p.t_window == pytest.approx((20000.0 + 5000.0 * (ns - 1)) / CLIGHT, rel=1e-15)
    assert p.t_window == pytest.approx(85.873, abs=1e-3)
    assert p.gdd == pytest.approx(-cr.b2[1] * 5000.0, rel=1e-15)
    assert p.fb_re[0] == pytest.approx(math.sqrt(0.98), rel=1e-15) and p.fb_im[0] == 0.0
    assert (p.fb_re[1], p.fb_im[1]) == (1.0, 0.0)   # idler not resonant: untouched (R13)
    assert p.dx == pytest.approx(400.0 / 63, rel=1e-15) and p.dz == pytest.approx(5000 / 99, rel=1e-15)
    eps0 = 8.8541878128e-6
    assert p.power_prefactor[1] == pytest.approx(0.5 * eps0 * CLIGHT * ns * p.dx * p.dy / 400.0 ** 2,
                                                 rel=1e-15)
    assert p.n_slices == 100 and p.dk_phase_scale == 0.0 and p.rk4_h_in_stages == 0


@pytest.mark.parametrize("lcr,nz,s", [(5000, 100, 100), (2000, 400, 399), (29000, 100, 99),
                                      (5000, 50, 49), (1000, 100, 99)])
def test_slice_count_float_rule(lcr, nz, s):
    assert LIB.opo_slice_count(lcr, nz) == s
    # the float loop by hand
    L = np.float32(lcr)
    dz = np.float32(L / (np.float32(nz) - np.float32(1)))
    z, k = np.float32(0), 0
    while z <= L:
        k += 1
        z = np.float32(z + dz)
    assert k == s


def test_time_and_omega_grids():
    t = np.zeros(5)
    check(LIB.opo_time_grid(5, 2.0, t.ctypes.data_as(C.POINTER(C.c_double))), "t")
    assert np.allclose(t, [-1, -0.5, 0, 0.5, 1])
    w = np.zeros(4)
    check(LIB.opo_omega_grid(4, 1.0, w.ctypes.data_as(C.POINTER(C.c_double))), "w")
    assert np.allclose(w, 2 * np.pi * np.array([2 / 3, 2, -2, -2 / 3]), rtol=1e-15)


def test_pump_peak_and_envelopes():
    grid, p, ph, (api, _, _) = configs.single_pass_32(precision=1)
    ap0 = configs.peak_field(ph, 1.0, 50.0)
    assert ap0 == pytest.approx(0.2936435, rel=1e-6)
    cr = configs.crystal(ph)
    zr = math.pi * cr.n[0] * 50.0 ** 2 / 0.532
    mx = 1 - 1j * (2500.0 / zr)
    # focused-pulsed at the beam centre (NX/2, NY/2): (Ap0/MX) exp(-(t/tau)^2)
    t = np.zeros(64)
    LIB.opo_time_grid(64, p.t_window, t.ctypes.data_as(C.POINTER(C.c_double)))
    tau = 2.0 * math.sqrt(2) / (2 * math.sqrt(2 * math.log(2)))
    assert tau == pytest.approx(1.201122, abs=1e-6)
    centre = api[:, 16, 16]
    assert np.allclose(centre, ap0 / mx * np.exp(-(t / tau) ** 2), rtol=1e-12, atol=1e-300)
    # |A| has full width at half maximum equal to the FWHM parameter (2 ps)
    tt = np.linspace(-5, 5, 200001)
    env = np.exp(-(tt / tau) ** 2)
    above = tt[env >= 0.5]
    assert above[-1] - above[0] == pytest.approx(2.0, abs=1e-4)


def test_sech2_pump_fwhm():
    grid, p, ph, (api, _, _) = configs.femto_64(precision=1)
    t = np.zeros(1024)
    LIB.opo_time_grid(1024, p.t_window, t.ctypes.data_as(C.POINTER(C.c_double)))
    inten = np.abs(api[:, 32, 32]) ** 2
    inten /= inten.max()
    # linear interpolation of the half-maximum crossings
    i = np.where(inten >= 0.5)[0]
    def cross(a, b):
        return t[a] + (0.5 - inten[a]) * (t[b] - t[a]) / (inten[b] - inten[a])
    fw = cross(i[-1], i[-1] + 1) - cross(i[0], i[0] - 1)
    assert fw == pytest.approx(0.150, rel=2e-3)


def philox_ref(ctr, key):
    M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
    c = list(ctr)
    k0, k1 = key
    for _ in range(10):
        p0, p1 = M0 * c[0], M1 * c[2]
        c = [((p1 >> 32) ^ c[1] ^ k0) & 0xFFFFFFFF, p1 & 0xFFFFFFFF,
             ((p0 >> 32) ^ c[3] ^ k1) & 0xFFFFFFFF, p0 & 0xFFFFFFFF]
        k0, k1 = (k0 + W0) & 0xFFFFFFFF, (k1 + W1) & 0xFFFFFFFF
    return c


def test_philox_known_answer_and_noise_cell():
    # Random123 known-answer vectors for philox4x32-10
    assert philox_ref([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert philox_ref([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e,
                                                              0xa20bc7c6, 0x6d5451fd]
    assert philox_ref([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                      [0xa4093822, 0x299f31d0]) == [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]
    # cell i of opo_setup_noise: counter (i, 0, 0, 0), key = seed; two 53-bit
    # uniforms, Box-Muller
    seed, sigma = 0x1234_5678_9abc_def0, 0.5
    g = Grid(4, 2, 3)
    out = np.empty((3, 2, 4), np.complex128)
    check(LIB.opo_setup_noise(C.byref(g), C.c_uint64(seed), sigma, 1, out.ctypes.data_as(C.c_void_p)), "n")
    for i in (0, 5, 23):
        r = philox_ref([i, 0, 0, 0], [seed & 0xFFFFFFFF, seed >> 32])
        a = ((r[0] << 21) ^ (r[1] >> 11)) & ((1 << 53) - 1)
        b = ((r[2] << 21) ^ (r[3] >> 11)) & ((1 << 53) - 1)
        u1, u2 = (a + 0.5) / 2.0 ** 53, (b + 0.5) / 2.0 ** 53
        rad = math.sqrt(-2 * math.log(u1))
        ref = sigma * rad * complex(math.cos(2 * math.pi * u2), math.sin(2 * math.pi * u2))
        assert abs(out.ravel()[i] - ref) < 1e-15


def test_noise_statistics_and_idler_copy():
    grid, p, ph, (api, as_, ai) = configs.cavity_64(precision=1)
    sigma = 1e-10 * configs.peak_field(ph, 1.0, 50.0)
    for part in (as_.real, as_.imag):
        assert abs(part.std() / sigma - 1) < 3e-3 and abs(part.mean()) < 3e-3 * sigma
    assert np.array_equal(as_, ai)   # cuOPO3D.cu:51 (R14)


def test_courant_gate():
    # Solver3D.cuh:495-504: dz < 0.2 kp dx dy; at a 400 um window it fails for 128^2
    for n, ok in ((32, 1), (64, 1), (96, 1), (128, 0)):
        g = Grid(n, n, 128)
        p = configs.make_params(g, configs.physical())
        assert LIB.opo_courant_ok(C.byref(p)) == ok, n


def zgp(pol, lp=2.09, ls=4.18, li=4.18):
    cr = Crystal()
    check(LIB.opo_crystal_properties_pol(configs.CRYSTAL_ZGP, 27.0, 0.0, lp, ls, li, pol, C.byref(cr)),
          "zgp")
    return cr


@pytest.mark.parametrize("pol", [b"ooo", b"eee"])
@pytest.mark.parametrize("lam", [2.09, 3.5, 8.0])
def test_zgp_dispersion_from_finite_differences(pol, lam):
    """ZGP (ZGP.cuh:101-256): group velocity and GVD against finite
    differences of its own Sellmeier index, for both polarisations."""
    n = lambda L: zgp(pol, lp=L).n[0]
    h1, h2 = 1e-4, 5e-4
    d1 = (n(lam + h1) - n(lam - h1)) / (2 * h1)
    d2 = (n(lam + h2) - 2 * n(lam) + n(lam - h2)) / h2 ** 2
    cr = zgp(pol, lp=lam)
    assert abs(cr.vg[0] - CLIGHT / (n(lam) - lam * d1)) / cr.vg[0] < 1e-7
    b2 = lam ** 3 * d2 / (2 * math.pi * CLIGHT ** 2)
    assert abs(cr.b2[0] - b2) / abs(b2) < 1e-4
    assert cr.b3[0] == 0.0


def test_zgp_index_closed_form_and_reference_quirks():
    # the Zelmon expression itself at 2.09 um, 'e' and 'o'
    for pol, (A, B, Cc, D, E) in ((b"eoo", (8.0929, 1.8649, 0.41468, 0.84052, 452.05)),
                                  (b"oee", (8.0409, 1.68625, 0.40824, 1.2880, 611.05))):
        L2 = 2.09 ** 2
        n_ref = math.sqrt(A + B * L2 / (L2 - Cc) + D * L2 / (L2 - E))
        cr = zgp(pol)
        assert abs(cr.n[0] - n_ref) < 1e-14
        assert 3.1 < cr.n[0] < 3.2                       # ZnGeP2 near 2 um
    cr = zgp(b"eoo")
    assert cr.n[0] > zgp(b"ooo").n[0]                    # positive uniaxial
    # ZGP.cuh:71-74: kp uses ns, ki uses ls, no grating term (QPM = false)
    assert abs(cr.kp - 2 * math.pi * cr.n[1] / 2.09) < 1e-12
    assert abs(cr.ki - 2 * math.pi * cr.n[2] / 4.18) < 1e-12
    assert abs(cr.dk - (cr.kp - cr.ks - cr.ki)) < 1e-12
    assert abs(cr.dq - 2 * 75e-6 / math.pi) < 1e-18
    assert list(cr.alpha) == [1.57e-4, 0.17e-6, 0.17e-6]


def test_zgp_requires_polarisations():
    cr = Crystal()
    assert LIB.opo_crystal_properties(configs.CRYSTAL_ZGP, 27.0, 0.0, 2.09, 4.18, 4.18, C.byref(cr)) != 0
    assert LIB.opo_crystal_properties_pol(configs.CRYSTAL_ZGP, 27.0, 0.0, 2.09, 4.18, 4.18, b"exo",
                                          C.byref(cr)) != 0
    # a full solve set-up with ZGP: kappa from dQ = 2 d33/pi and n per polarisation
    ph = configs.physical(lcr=5000.0, dq=0.0)
    ph.crystal = configs.CRYSTAL_ZGP
    ph.lp, ph.ls, ph.li = 2.09, 4.18, 4.18
    ph.pol = b"eoo"
    p = configs.make_params(Grid(32, 32, 64), ph)
    cr = zgp(b"eoo")
    for f, lam in enumerate((2.09, 4.18, 4.18)):
        assert abs(p.kappa[f] - 2 * math.pi * cr.dq / (cr.n[f] * lam)) < 1e-15
    assert p.kdiff[0] == cr.kp and p.dk == cr.dk
