// opo_setup.cpp -- FP64 host set-up of an OPO solve (crystal, cavity, grids,
// pump, noise).  Host-only C++; no device code.
#include "../../include/opo3d_setup.h"

#include <cmath>
#include <cstring>
#include <vector>

#include "opo_internal.hpp"

namespace {

const double kPi = 3.14159265358979323846264338327950288;
const double kC = 299792458.0 * 1e6 / 1e12;         // um/ps   DtypesConsts.cuh:56
const double kEps0 = 8.8541878128e-12 * 1e12 / 1e6;  // W ps/(V^2 um)  :57

// Sellmeier model with analytic wavelength derivatives.  Both families of
// the reference share the structure n^2 = a + sum_j B_j/(L^2 - C_j^2) + D L^2
// (MgOPPLN.cuh:96-116 with D = -a6; MgOsPPLT.cuh n() with D = D), whose
// derivative terms are "resonances" R_p = sum_j B_j/(L^2 - C_j^2)^p.
struct Sellmeier {
    double a, D;
    double B[3], Cc[3];
    int nres;
    double R(double L, int p) const {
        double s = 0;
        for (int j = 0; j < nres; ++j) s += B[j] / std::pow(L * L - Cc[j] * Cc[j], p);
        return s;
    }
    double n(double L) const {
        double s = a + D * L * L;
        for (int j = 0; j < nres; ++j) s += B[j] / (L * L - Cc[j] * Cc[j]);
        return std::sqrt(s);
    }
    // dn/dL = -L (R2 - D) / n   (MgOPPLN.cuh:167 with a6 = -D; MgOsPPLT.cuh dndl)
    double dn(double L) const { return -L * (R(L, 2) - D) / n(L); }
    // MgOPPLN.cuh:195-198
    double d2n(double L) const {
        double nn = n(L);
        return (L * dn(L) / (nn * nn) - 1.0 / nn) * (R(L, 2) - D) + 4.0 * L * L / nn * R(L, 3);
    }
    // MgOPPLN.cuh:225-233
    double d3n(double L) const {
        double nn = n(L), d1 = dn(L), d2 = d2n(L);
        double A1 = (2 * d1 + L * d2) / (nn * nn);
        double A2 = -2 * L * d1 * d1 / (nn * nn * nn);
        double AA = (A1 + A2) * (R(L, 2) - D);
        double B1 = 12 * L / nn;
        double B2 = -4 * L * L / nn * d2 * (1 - 1 / nn);
        double BB = (B1 + B2) * R(L, 3);
        double CC = -24 * L * L * L / nn * R(L, 4);
        return AA + BB + CC;
    }
    double gv(double L) const { return kC / (n(L) - L * dn(L)); }                 // :239-243
    double gvd(double L) const { return L * L * L * d2n(L) / (2 * kPi * kC * kC); }  // :248-251
    double tod(double L) const {                                                  // :256-259
        return -std::pow(L, 4) / (4 * kPi * kPi * kC * kC * kC) * (3 * d2n(L) + L * d3n(L));
    }
};

// Gayer-type model (MgOPPLN.cuh:96-116, PPLN.cuh same form).
Sellmeier gayer(double T, const double a[6], const double b[4]) {
    double f = (T - 24.5) * (T + 570.82);
    Sellmeier s{};
    s.a = a[0] + b[0] * f;
    s.B[0] = a[1] + b[1] * f;
    s.Cc[0] = a[2] + b[2] * f;
    s.B[1] = a[3] + b[3] * f;
    s.Cc[1] = a[4];
    s.nres = 2;
    s.D = -a[5];
    return s;
}

// Bruner model (MgOsPPLT.cuh n(), resonances()).
Sellmeier bruner(double T) {
    double TK2 = (T + 273.15) * (T + 273.15);
    Sellmeier s{};
    s.a = 4.502483;
    s.B[0] = 0.007294 + 3.483933e-8 * TK2;
    s.Cc[0] = 0.185087 + 1.607839e-8 * TK2;
    s.B[1] = 0.073423;
    s.Cc[1] = 0.199595;
    s.B[2] = 0.001;
    s.Cc[2] = 7.99724;
    s.nres = 3;
    s.D = -0.02357;
    return s;
}

// Zelmon model of ZnGeP2 (ZGP.cuh:101-125): n^2 = A + B L^2/(L^2 - C)
// + D L^2/(L^2 - E) per polarisation.  With g = B C L/(L^2-C)^2
// + D E L/(L^2-E)^2: dn/dL = -g/n and d2n/dL2 = (-g' - (dn/dL)^2)/n, the
// closed forms the reference evaluates (ZGP.cuh:183-244).
struct Zelmon {
    double A, B, Cc, D, E;
    double n(double L) const {
        const double L2 = L * L;
        return std::sqrt(A + B * L2 / (L2 - Cc) + D * L2 / (L2 - E));
    }
    double dn(double L) const {
        const double L2 = L * L, u = L2 - Cc, v = L2 - E;
        return -(B * Cc * L / (u * u) + D * E * L / (v * v)) / n(L);
    }
    double d2n(double L) const {
        const double L2 = L * L, u = L2 - Cc, v = L2 - E, d1 = dn(L);
        const double mg1 = B * Cc * (3 * L2 + Cc) / (u * u * u) + D * E * (3 * L2 + E) / (v * v * v);
        return (mg1 - d1 * d1) / n(L);
    }
    double gv(double L) const { return kC / (n(L) - L * dn(L)); }
    double gvd(double L) const { return L * L * L * d2n(L) / (2 * kPi * kC * kC); }
};

Zelmon zelmon(char pol) {
    if (pol == 'o') return Zelmon{8.0409, 1.68625, 0.40824, 1.2880, 611.05};
    return Zelmon{8.0929, 1.8649, 0.41468, 0.84052, 452.05};
}

// ------------------------------------------------------ Philox4x32-10
struct Philox {
    uint32_t k0, k1;
    void round(uint32_t c[4], uint32_t a0, uint32_t a1) const {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ a0, n1 = lo1, n2 = hi0 ^ c[3] ^ a1, n3 = lo0;
        c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
    }
    void gen(uint64_t ctr, uint32_t out[4]) const {
        uint32_t c[4] = {(uint32_t)ctr, (uint32_t)(ctr >> 32), 0u, 0u};
        uint32_t a0 = k0, a1 = k1;
        for (int r = 0; r < 10; ++r) {
            round(c, a0, a1);
            a0 += 0x9E3779B9u;
            a1 += 0xBB67AE85u;
        }
        std::memcpy(out, c, sizeof(c));
    }
};

template <typename T> void store_field(void *out, size_t i, double re, double im) {
    T *o = static_cast<T *>(out);
    o[2 * i] = (T)re;
    o[2 * i + 1] = (T)im;
}

}  // namespace

extern "C" {

int opo_slice_count(double lcr, int nz) {
    if (nz < 2 || !(lcr > 0)) return 0;
    volatile float L = (float)lcr;
    volatile float dz = (float)(L / ((float)nz - 1.0f));
    int count = 0;
    for (volatile float z = 0.0f; z <= L; z = z + dz) ++count;
    return count;
}

int opo_crystal_properties(int crystal, double T, double Lambda, double lp, double ls, double li,
                           opo_crystal_props *out) {
    return opo_crystal_properties_pol(crystal, T, Lambda, lp, ls, li, nullptr, out);
}

int opo_crystal_properties_pol(int crystal, double T, double Lambda, double lp, double ls, double li,
                               const char *pol, opo_crystal_props *out) {
    if (!out) return opo::fail(OPO_ERR_ARG, "opo_crystal_properties: null output");
    if (crystal == OPO_CRYSTAL_ZGP) {  // ZGP.cuh:49-77
        if (!pol) return opo::fail(OPO_ERR_ARG, "ZGP needs the polarisations of p, s, i");
        for (int f = 0; f < 3; ++f)
            if (pol[f] != 'o' && pol[f] != 'e')
                return opo::fail(OPO_ERR_ARG, "ZGP polarisation %d is '%c', not 'o' or 'e'", f, pol[f]);
        const double L[3] = {lp, ls, li};
        const double al[3] = {1.57e-4, 0.17e-6, 0.17e-6};
        for (int f = 0; f < 3; ++f) {
            const Zelmon z = zelmon(pol[f]);
            out->n[f] = z.n(L[f]);
            out->vg[f] = z.gv(L[f]);
            out->b2[f] = z.gvd(L[f]);
            out->b3[f] = 0.0;  // the reference's ZGP has no TOD
            out->alpha[f] = al[f];
        }
        out->d33 = 75.0e-6;
        out->dq = 2.0 * out->d33 / kPi;
        out->kp = 2.0 * kPi * out->n[1] / lp;  // ZGP.cuh:71 (ns, as R6)
        out->ks = 2.0 * kPi * out->n[1] / ls;
        out->ki = 2.0 * kPi * out->n[2] / ls;  // ZGP.cuh:73 (ls)
        out->dk = out->kp - out->ks - out->ki;  // no QPM term (QPM = false, :58)
        return OPO_OK;
    }
    Sellmeier s;
    double d33, al[3];
    if (crystal == OPO_CRYSTAL_MGOPPLN) {  // MgOPPLN.cuh:59-60, 99-113
        const double a[6] = {5.756, 0.0983, 0.2020, 189.32, 12.52, 1.32e-2};
        const double b[4] = {2.860e-6, 4.700e-8, 6.113e-8, 1.516e-4};
        s = gayer(T, a, b);
        d33 = 25.20e-6;
        al[0] = 0.025e-4; al[1] = 0.002e-4; al[2] = 0.002e-4;
    } else if (crystal == OPO_CRYSTAL_PPLN) {  // PPLN.cuh:58-59 and its n()
        const double a[6] = {5.35583, 0.100473, 0.20692, 100.0, 11.34927, 1.5334e-2};
        const double b[4] = {4.629e-7, 3.862e-8, -0.89e-8, 2.657e-5};
        s = gayer(T, a, b);
        d33 = 25.20e-6;
        al[0] = 0.025e-4; al[1] = 0.002e-4; al[2] = 0.002e-4;
    } else if (crystal == OPO_CRYSTAL_MGOSPPLT) {  // MgOsPPLT.cuh:58-59 and its n()
        s = bruner(T);
        d33 = 11.0e-6;
        al[0] = 1.57e-4; al[1] = 0.17e-6; al[2] = 0.17e-6;
    } else {
        return opo::fail(OPO_ERR_ARG, "opo_crystal_properties: unknown crystal %d", crystal);
    }
    const double L[3] = {lp, ls, li};
    for (int f = 0; f < 3; ++f) {
        out->n[f] = s.n(L[f]);
        out->vg[f] = s.gv(L[f]);
        out->b2[f] = s.gvd(L[f]);
        out->b3[f] = s.tod(L[f]);
        out->alpha[f] = al[f];
    }
    out->d33 = d33;
    out->dq = 2.0 * d33 / kPi;
    out->kp = 2.0 * kPi * out->n[1] / lp;  // MgOPPLN.cuh:74 (ns, R6)
    out->ks = 2.0 * kPi * out->n[1] / ls;
    out->ki = 2.0 * kPi * out->n[2] / ls;
    out->dk = out->kp - out->ks - out->ki - 2.0 * kPi / Lambda;
    return OPO_OK;
}

int opo_setup_params(const opo_grid *g, const opo_physical *ph, opo_params *p) {
    if (!g || !ph || !p) return opo::fail(OPO_ERR_ARG, "opo_setup_params: null argument");
    if (g->nx < 1 || g->ny < 1 || g->nt < 1 || ph->nz < 2)
        return opo::fail(OPO_ERR_ARG, "opo_setup_params: bad grid or NZ");
    opo_crystal_props cr;
    int rc = opo_crystal_properties_pol(ph->crystal, ph->temperature, ph->grating_period, ph->lp,
                                        ph->ls, ph->li, ph->pol, &cr);
    if (rc) return rc;
    std::memset(p, 0, sizeof(*p));
    p->lx = ph->lx;
    p->ly = ph->ly;
    p->dx = g->nx > 1 ? ph->lx / (g->nx - 1.0) : ph->lx;  // MgOPPLN.cuh:62-63
    p->dy = g->ny > 1 ? ph->ly / (g->ny - 1.0) : ph->ly;
    p->dz = ph->lcr / (ph->nz - 1.0);                     // MgOPPLN.cuh:65
    p->n_slices = opo_slice_count(ph->lcr, ph->nz);
    p->lcr = ph->lcr;
    const double dq = ph->dq > 0 ? ph->dq : cr.dq;
    const double lam[3] = {ph->lp, ph->ls, ph->li};
    for (int f = 0; f < 3; ++f) {
        p->kappa[f] = 2.0 * kPi * dq / (cr.n[f] * lam[f]);  // EFields3D.cuh:219-221
        p->beta2[f] = cr.b2[f];
        p->alpha[f] = cr.alpha[f];
    }
    p->kdiff[0] = cr.kp;  // EFields3D.cuh:222-224 (R6)
    p->kdiff[1] = cr.ks;
    p->kdiff[2] = cr.ki;
    p->delta[0] = 1.0 / cr.vg[0] - 1.0 / cr.vg[2];  // Solver3D.cuh:163-165 (R7)
    p->delta[1] = 0.0;
    p->delta[2] = 1.0 / cr.vg[2] - 1.0 / cr.vg[1];
    p->t_window = ph->t_window > 0 ? ph->t_window
                                   : (ph->lcav + ph->lcr * (cr.n[1] - 1.0)) / kC;  // Cavity3D.cuh:44
    p->omega = nullptr;
    p->dk = cr.dk;
    p->dk_phase_scale = 0.0;  // Solver3D.cuh:375-377 (R4)
    p->rk4_h_in_stages = 0;   // Solver3D.cuh:915-927 (R3)
    p->fs_on = ph->fs_on;
    p->fs_length = ph->lcr;   // Solver3D.cuh:989 (R12)
    p->skip_fs_on_last_trip = 1;
    p->gdd_on = ph->gdd_on;
    p->gdd = -ph->gdd_gamma * cr.b2[1] * ph->lcr;  // Solver3D.cuh:676
    // Cavity3D.cuh:12-28 with Solver3D.cuh:996-997: resonant fields get
    // sqrt(R) e^{i phi}; non-resonant fields are left untouched (R13).
    p->fb_re[0] = ph->resonant_s ? std::sqrt(ph->rs) * std::cos(ph->phase_s) : 1.0;
    p->fb_im[0] = ph->resonant_s ? std::sqrt(ph->rs) * std::sin(ph->phase_s) : 0.0;
    p->fb_re[1] = ph->resonant_i ? std::sqrt(ph->ri) * std::cos(ph->phase_i) : 1.0;
    p->fb_im[1] = ph->resonant_i ? std::sqrt(ph->ri) * std::sin(ph->phase_i) : 0.0;
    p->n_round_trips = ph->n_round_trips;
    const double pref = 0.5 * kEps0 * kC * cr.n[1] * p->dx * p->dy / (ph->lx * ph->ly);  // :542
    for (int f = 0; f < 3; ++f) p->power_prefactor[f] = pref;
    p->precision = ph->precision;
    return OPO_OK;
}

int opo_time_grid(int nt, double W, double *t) {
    if (nt < 1 || !t) return opo::fail(OPO_ERR_ARG, "opo_time_grid: bad argument");
    if (nt == 1) { t[0] = 0.0; return OPO_OK; }  // guard for linspace(size 1)
    for (int i = 0; i < nt; ++i) t[i] = -0.5 * W + i * W / (nt - 1);
    return OPO_OK;
}

int opo_omega_grid(int nt, double W, double *w) {
    if (nt < 1 || !w) return opo::fail(OPO_ERR_ARG, "opo_omega_grid: bad argument");
    opo::omega_rule(nt, W, w);
    return OPO_OK;
}

int opo_setup_pump(const opo_grid *g, const opo_params *p, const opo_physical *ph, int mode,
                   double power, double fwhm, double waist, double focal, int precision,
                   void *out) {
    if (!g || !p || !ph || !out) return opo::fail(OPO_ERR_ARG, "opo_setup_pump: null argument");
    opo_crystal_props cr;
    int rc = opo_crystal_properties_pol(ph->crystal, ph->temperature, ph->grating_period, ph->lp,
                                        ph->ls, ph->li, ph->pol, &cr);
    if (rc) return rc;
    const int nx = g->nx, ny = g->ny, nt = g->nt;
    const double np = cr.n[0];
    // EFields3D.cuh:247-253
    const double tau = fwhm * std::sqrt(2.0) / (2.0 * std::sqrt(2.0 * std::log(2.0)));
    const double w02 = waist * waist;
    const double zR = kPi * np * w02 / ph->lp;
    const double eta = focal / zR;
    const double Ap0 = std::sqrt(4.0 * power / (kEps0 * kC * kPi * np * w02));
    const double T0 = fwhm / (2.0 * std::acosh(std::sqrt(2.0)));
    std::vector<double> t(nt);
    opo_time_grid(nt, p->t_window, t.data());
    const double uX = 0.5 * nx, uY = 0.5 * ny;
    // MX = 1 - i eta; a/MX and CpxExp(real/(w0^2 MX)) as in EFields3D.cuh:107-137
    const double mr = 1.0, mi = -eta, md = mr * mr + mi * mi;
    const double inv_re = mr / md, inv_im = -mi / md;  // 1/MX
    const bool focused = mode == OPO_PUMP_FOCUSED_CW || mode == OPO_PUMP_FOCUSED_PULSED ||
                         mode == OPO_PUMP_FOCUSED_SECH2;
    if (mode < 0 || mode > OPO_PUMP_FOCUSED_SECH2)
        return opo::fail(OPO_ERR_ARG, "opo_setup_pump: unknown mode %d", mode);
    for (int it = 0; it < nt; ++it) {
        double env = 1.0;
        if (mode == OPO_PUMP_WAVEPLANE_PULSED || mode == OPO_PUMP_FOCUSED_PULSED)
            env = std::exp(-(t[it] / tau) * (t[it] / tau));
        else if (mode == OPO_PUMP_FOCUSED_SECH2)
            env = 1.0 / std::cosh(t[it] / T0);
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const double X = (ix - uX) * p->dx, Y = (iy - uY) * p->dy;
                const double r2 = -(X * X) - (Y * Y);
                double re, im;
                if (!focused) {
                    re = Ap0 * env * std::exp(r2 / w02);
                    im = 0.0;
                } else {
                    // CpxExp(r2/(w0^2 MX)) = exp(a) (cos b, sin b), a + ib = (r2/w0^2)/MX
                    const double a = r2 / w02 * inv_re, b = r2 / w02 * inv_im;
                    const double ea = std::exp(a);
                    const double er = ea * std::cos(b), ei = ea * std::sin(b);
                    const double cr_ = Ap0 * inv_re, ci_ = Ap0 * inv_im;  // Ap0/MX
                    re = env * (cr_ * er - ci_ * ei);
                    im = env * (cr_ * ei + ci_ * er);
                }
                const size_t i = ((size_t)it * ny + iy) * nx + ix;
                if (precision == OPO_FP64) store_field<double>(out, i, re, im);
                else store_field<float>(out, i, re, im);
            }
    }
    return OPO_OK;
}

int opo_setup_noise(const opo_grid *g, uint64_t seed, double sigma, int precision, void *out) {
    if (!g || !out) return opo::fail(OPO_ERR_ARG, "opo_setup_noise: null argument");
    Philox ph{(uint32_t)seed, (uint32_t)(seed >> 32)};
    const size_t n = (size_t)g->nx * g->ny * g->nt;
    const double two53 = 9007199254740992.0;
    for (size_t i = 0; i < n; ++i) {
        uint32_t r[4];
        ph.gen((uint64_t)i, r);
        // two 53-bit uniforms in (0, 1)
        uint64_t a = ((uint64_t)r[0] << 21) ^ (r[1] >> 11);
        uint64_t b = ((uint64_t)r[2] << 21) ^ (r[3] >> 11);
        a &= (1ull << 53) - 1;
        b &= (1ull << 53) - 1;
        double u1 = ((double)a + 0.5) / two53, u2 = ((double)b + 0.5) / two53;
        double rad = std::sqrt(-2.0 * std::log(u1)), th = 2.0 * kPi * u2;
        double re = sigma * rad * std::cos(th), im = sigma * rad * std::sin(th);
        if (precision == OPO_FP64) store_field<double>(out, i, re, im);
        else store_field<float>(out, i, re, im);
    }
    return OPO_OK;
}

}  // extern "C"
