// oracle.cpp -- CPU restatement of the reference split-step path.
//
// TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
// CPU-baseline leg of bench.py as the checker.  The product library never
// links or calls it.
//
// Reference: alfredos84/cuOPO3D, src/headers/Solver3D.cuh, Cavity3D.cuh,
// EFields3D.cuh, Operators.cuh, Common.cuh.  Each function cites the lines it
// restates.  Defects of the reference that make its output degenerate are
// fixed here (normalisations, all-pixel dispersion, built propagators), and
// its non-degenerate choices are kept as parameter values (opo_params).
//
// FFT: self-sorting mixed-radix Stockham (Van Loan, "Computational
// Frameworks for the Fast Fourier Transform", SIAM 1992, sec. 1.7), written
// here from the recurrence Y_{Lp}[k + L r][j] = sum_q w_p^{rq} w_{Lp}^{kq}
// Y_L[k][j + m q], m = n/(L p).  Twiddles come from long-double sin/cos of
// exact fractions of 2*pi.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <functional>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>
#include <omp.h>

namespace {

template <typename R> struct cx { R re, im; };
template <typename R> inline cx<R> operator+(cx<R> a, cx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R> inline cx<R> operator-(cx<R> a, cx<R> b) { return {a.re - b.re, a.im - b.im}; }
// Operators.cuh:109-116: (a.x b.x - a.y b.y, a.x b.y + a.y b.x)
template <typename R> inline cx<R> operator*(cx<R> a, cx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R> inline cx<R> operator*(R s, cx<R> a) { return {s * a.re, s * a.im}; }
template <typename R> inline cx<R> operator*(cx<R> a, R s) { return {a.re * s, a.im * s}; }
template <typename R> inline cx<R> operator/(cx<R> a, R s) { return {a.re / s, a.im / s}; }
template <typename R> inline cx<R> conj(cx<R> a) { return {a.re, -a.im}; }

const long double kPiL = 3.141592653589793238462643383279502884L;
const double kTwoPi = 6.283185307179586476925286766559;

// exp(i * 2*pi * num/den) from long double, exact-fraction argument.
inline void unit_root(long long num, long long den, long double &c, long double &s) {
    long long r = num % den;
    if (r < 0) r += den;
    long double a = 2.0L * kPiL * (long double)r / (long double)den;
    c = cosl(a);
    s = sinl(a);
}

// exp(i*phi) with phi built in FP64 and reduced mod 2*pi before cos/sin.
inline void cis_reduced(double phi, double &c, double &s) {
    double r = std::fmod(phi, kTwoPi);
    c = std::cos(r);
    s = std::sin(r);
}

// ------------------------------------------------------------------ FFT
template <typename R> struct Plan {
    int n = 1, sign = -1;
    std::vector<int> fac;
    std::vector<std::vector<cx<R>>> tw;    // per stage: tw[k*p + q] = w_{Lp}^{kq}
    std::vector<std::vector<cx<R>>> root;  // per stage: w_p^{j}, j < p (generic radix)
};

std::vector<int> factorize(int n) {
    std::vector<int> f;
    while (n % 4 == 0) { f.push_back(4); n /= 4; }
    while (n % 2 == 0) { f.push_back(2); n /= 2; }
    while (n % 3 == 0) { f.push_back(3); n /= 3; }
    for (int p = 5; n > 1; p += 2)
        while (n % p == 0) { f.push_back(p); n /= p; }
    return f;
}

template <typename R> std::shared_ptr<Plan<R>> make_plan(int n, int sign) {
    auto P = std::make_shared<Plan<R>>();
    P->n = n;
    P->sign = sign;
    P->fac = factorize(n);
    int L = 1;
    for (int p : P->fac) {
        int Lp = L * p;
        std::vector<cx<R>> t((size_t)L * p);
        for (int k = 0; k < L; ++k)
            for (int q = 0; q < p; ++q) {
                long double c, s;
                unit_root((long long)sign * k * q, Lp, c, s);
                t[(size_t)k * p + q] = {(R)c, (R)s};
            }
        std::vector<cx<R>> rt(p);
        for (int j = 0; j < p; ++j) {
            long double c, s;
            unit_root((long long)sign * j, p, c, s);
            rt[j] = {(R)c, (R)s};
        }
        P->tw.push_back(std::move(t));
        P->root.push_back(std::move(rt));
        L = Lp;
    }
    return P;
}

template <typename R> std::shared_ptr<Plan<R>> get_plan(int n, int sign) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, std::shared_ptr<Plan<R>>> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_pair(n, sign);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    auto P = make_plan<R>(n, sign);
    cache[key] = P;
    return P;
}

// In-place transform of x (length n) using scratch y (length n).
template <typename R> void fft_exec(const Plan<R> &P, cx<R> *x, cx<R> *y) {
    const int n = P.n;
    if (n == 1) return;
    const R s = (R)P.sign;
    cx<R> *src = x, *dst = y;
    int L = 1;
    for (size_t st = 0; st < P.fac.size(); ++st) {
        const int p = P.fac[st];
        const int Lp = L * p, m = n / Lp;
        const cx<R> *tw = P.tw[st].data();
        for (int k = 0; k < L; ++k) {
            const cx<R> *twk = tw + (size_t)k * p;
            const cx<R> *in = src + (size_t)k * m * p;
            if (p == 4) {
                cx<R> w1 = twk[1], w2 = twk[2], w3 = twk[3];
                for (int j = 0; j < m; ++j) {
                    cx<R> a0 = in[j], a1 = in[j + m] * w1, a2 = in[j + 2 * m] * w2,
                          a3 = in[j + 3 * m] * w3;
                    cx<R> t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = a1 - a3;
                    // (s*i) * t3
                    cx<R> it3 = {-s * t3.im, s * t3.re};
                    dst[(size_t)(k) * m + j] = t0 + t2;
                    dst[(size_t)(k + L) * m + j] = t1 + it3;
                    dst[(size_t)(k + 2 * L) * m + j] = t0 - t2;
                    dst[(size_t)(k + 3 * L) * m + j] = t1 - it3;
                }
            } else if (p == 2) {
                cx<R> w1 = twk[1];
                for (int j = 0; j < m; ++j) {
                    cx<R> a0 = in[j], a1 = in[j + m] * w1;
                    dst[(size_t)k * m + j] = a0 + a1;
                    dst[(size_t)(k + L) * m + j] = a0 - a1;
                }
            } else if (p == 3) {
                cx<R> w1 = twk[1], w2 = twk[2];
                const R h = (R)0.5;
                const R r3 = (R)(s * (R)0.86602540378443864676372317075294L);
                for (int j = 0; j < m; ++j) {
                    cx<R> a0 = in[j], a1 = in[j + m] * w1, a2 = in[j + 2 * m] * w2;
                    cx<R> t1 = a1 + a2, t2 = a1 - a2;
                    cx<R> u = a0 - h * t1;
                    cx<R> v = {-r3 * t2.im, r3 * t2.re};  // i*s*sqrt(3)/2 * t2
                    dst[(size_t)k * m + j] = a0 + t1;
                    dst[(size_t)(k + L) * m + j] = u + v;
                    dst[(size_t)(k + 2 * L) * m + j] = u - v;
                }
            } else {
                const cx<R> *root = P.root[st].data();
                std::vector<cx<R>> a(p);
                for (int j = 0; j < m; ++j) {
                    for (int q = 0; q < p; ++q) a[q] = in[j + (size_t)q * m] * twk[q];
                    for (int r = 0; r < p; ++r) {
                        cx<R> acc = {0, 0};
                        for (int q = 0; q < p; ++q) acc = acc + a[q] * root[(r * q) % p];
                        dst[(size_t)(k + L * r) * m + j] = acc;
                    }
                }
            }
        }
        std::swap(src, dst);
        L = Lp;
    }
    if (src != x) std::memcpy(x, src, sizeof(cx<R>) * (size_t)n);
}

// ------------------------------------------------------------ the solve
int g_threads = 0;
int nthreads() { return g_threads > 0 ? g_threads : omp_get_max_threads(); }

struct Base {
    virtual ~Base() {}
    virtual void set_fields(const double *, const double *, const double *) = 0;
    virtual void get_fields(double *, double *, double *) = 0;
    virtual void nonlinear(double z) = 0;
    virtual void dispersion() = 0;
    virtual void diffraction() = 0;
    virtual void free_space() = 0;
    virtual void gdd() = 0;
    virtual void eom() = 0;
    virtual void mirror() = 0;
    virtual void power(double out[3]) = 0;
    virtual void table(int which, double *out) = 0;
    virtual void omega(double *out) = 0;
    opo_grid g{};
    opo_params p{};
    int trips_done = 0;
};

template <typename R> struct Solve final : Base {
    size_t P = 0, N = 0;  // pixels per plane, cells
    std::vector<cx<R>> Api, A[3];
    std::vector<double> w;                       // omega grid (FP64)
    std::vector<cx<R>> eiLz[3], eiQz[3], eiQfs[2], egdd;
    std::vector<double> zs;                      // slice positions

    Solve(const opo_grid &g_, const opo_params &p_) {
        g = g_;
        p = p_;
        p.omega = nullptr;
        P = (size_t)g.nx * g.ny;
        N = P * g.nt;
        Api.assign(N, {0, 0});
        for (auto &a : A) a.assign(N, {0, 0});
        build_omega(p_.omega);
        build_tables();
        zs.resize(p.n_slices);
        for (int k = 0; k < p.n_slices; ++k) zs[k] = k * p.dz;
    }

    // EFields3D.cuh:318-328 with Common.cuh:120-128 and EFields3D.cuh:58-68:
    // F = linspace(-NT/(2W), NT/(2W), NT); w = 2*pi*fftshift(F).  For NT = 1
    // the reference leaves w[0] at its zero initial value.
    void build_omega(const double *explicit_w) {
        const int nt = g.nt;
        w.assign(nt, 0.0);
        if (explicit_w) {
            for (int i = 0; i < nt; ++i) w[i] = explicit_w[i];
            return;
        }
        if (nt < 2) return;
        const double W = p.t_window;
        std::vector<double> F(nt);
        for (int i = 0; i < nt; ++i)
            F[i] = -0.5 * nt / W + i * ((double)nt / W) / (nt - 1);
        const int c = nt / 2;
        for (int i = 0; i < c; ++i) {
            w[i + c] = F[i];
            w[i] = F[i + c];
        }
        for (int i = 0; i < nt; ++i) w[i] *= kTwoPi;
    }

    // Signed spatial frequency index of FFT-order bin j after the reference's
    // centred table and fftShift2D (Solver3D.cuh:288-296, EFields3D.cuh:141-171,
    // 331-351): bin j holds the centred entry (j + N/2) mod N, i.e. index
    // ((j + N/2) mod N) - N/2.
    static double signed_freq(int j, int n) {
        if (n < 2) return 0.0;
        int c = n / 2;
        return (double)(((j + c) % n) - c);
    }

    void build_tables() {
        const int nx = g.nx, ny = g.ny, nt = g.nt;
        // Solver3D.cuh:153-169: eiLz_X(w) = exp(i dz w (delta_X + w b2_X / 2)) exp(-alpha_X dz/2)
        for (int f = 0; f < 3; ++f) {
            eiLz[f].resize(nt);
            double att = std::exp(-0.5 * p.alpha[f] * p.dz);
            for (int k = 0; k < nt; ++k) {
                double phi = p.dz * w[k] * (p.delta[f] + 0.5 * w[k] * p.beta2[f]);
                double c, s;
                cis_reduced(phi, c, s);
                eiLz[f][k] = {(R)(c * att), (R)(s * att)};
            }
        }
        // Solver3D.cuh:282-301: eiQz_X = exp(-i dz 2 pi^2/k_X (dfX^2 fx^2 + dfY^2 fy^2)) exp(-alpha_X dz/2)
        const double dfX = nx > 1 ? 1.0 / (p.dx * nx) : 0.0;
        const double dfY = ny > 1 ? 1.0 / (p.dy * ny) : 0.0;
        const double pi2 = (double)(kPiL * kPiL);
        for (int f = 0; f < 3; ++f) {
            eiQz[f].resize(P);
            double att = std::exp(-0.5 * p.alpha[f] * p.dz);
            for (int iy = 0; iy < ny; ++iy)
                for (int ix = 0; ix < nx; ++ix) {
                    double fx = signed_freq(ix, nx), fy = signed_freq(iy, ny);
                    double q = dfX * dfX * fx * fx + dfY * dfY * fy * fy;
                    double phi = -p.dz * (2.0 * pi2 / p.kdiff[f] * q);
                    double c, s;
                    cis_reduced(phi, c, s);
                    eiQz[f][(size_t)iy * nx + ix] = {(R)(c * att), (R)(s * att)};
                }
        }
        // Solver3D.cuh:305-323: free space, +i, length L, no absorption, s and i.
        for (int f = 0; f < 2; ++f) {
            eiQfs[f].resize(P);
            for (int iy = 0; iy < ny; ++iy)
                for (int ix = 0; ix < nx; ++ix) {
                    double fx = signed_freq(ix, nx), fy = signed_freq(iy, ny);
                    double q = dfX * dfX * fx * fx + dfY * dfY * fy * fy;
                    double phi = p.fs_length * (2.0 * pi2 / p.kdiff[1 + f] * q);
                    double c, s;
                    cis_reduced(phi, c, s);
                    eiQfs[f][(size_t)iy * nx + ix] = {(R)c, (R)s};
                }
        }
        // Solver3D.cuh:261-277: exp(i * w^2 * GDD / 2)
        egdd.resize(nt);
        for (int k = 0; k < nt; ++k) {
            double c, s;
            cis_reduced(0.5 * w[k] * w[k] * p.gdd, c, s);
            egdd[k] = {(R)c, (R)s};
        }
    }

    void set_fields(const double *api, const double *as, const double *ai) override {
        auto load = [&](std::vector<cx<R>> &v, const double *src) {
            if (!src) return;
            for (size_t i = 0; i < N; ++i) v[i] = {(R)src[2 * i], (R)src[2 * i + 1]};
        };
        load(Api, api);
        if (api) A[0] = Api;
        load(A[1], as);
        load(A[2], ai);
    }
    void get_fields(double *ap, double *as, double *ai) override {
        double *dst[3] = {ap, as, ai};
        for (int f = 0; f < 3; ++f) {
            if (!dst[f]) continue;
            for (size_t i = 0; i < N; ++i) {
                dst[f][2 * i] = (double)A[f][i].re;
                dst[f][2 * i + 1] = (double)A[f][i].im;
            }
        }
    }

    // dAdz (Solver3D.cuh:365-382): dAp = i kp As Ai e^{-i s dk z},
    // dAs = i ks Ap Ai* e^{+i s dk z}, dAi = i ki Ap As* e^{+i s dk z}.
    struct Rhs {
        cx<R> ikp, iks, iki, ep, es;
    };
    Rhs make_rhs(double z) const {
        double c, s;
        cis_reduced(p.dk_phase_scale * p.dk * z, c, s);
        Rhs r;
        r.ikp = {(R)0, (R)p.kappa[0]};
        r.iks = {(R)0, (R)p.kappa[1]};
        r.iki = {(R)0, (R)p.kappa[2]};
        r.es = {(R)c, (R)s};
        r.ep = {(R)c, (R)-s};
        return r;
    }
    static inline void rhs(const Rhs &r, cx<R> ap, cx<R> as, cx<R> ai, cx<R> &dp, cx<R> &ds,
                           cx<R> &di) {
        dp = r.ikp * as * ai * r.ep;
        ds = r.iks * ap * conj(ai) * r.es;
        di = r.iki * ap * conj(as) * r.es;
    }

    // solverRK4 (Solver3D.cuh:872-943) with h = dz/2 (:938), stage points
    // A + c1 k1, A + c2 k2, A + c3 k3 (:915, :921, :927) and the update
    // A += (k1 + 2 k2 + 2 k3 + k4) h / 6 (:418-420).
    void nonlinear(double z) override {
        const double h = 0.5 * p.dz;
        const R c1 = (R)(p.rk4_h_in_stages ? 0.5 * h : 0.5);
        const R c3 = (R)(p.rk4_h_in_stages ? h : 1.0);
        const Rhs r1 = make_rhs(z), r2 = make_rhs(z + p.dz / 4), r4 = make_rhs(z + p.dz / 2);
        const R hr = (R)h;
        cx<R> *ap = A[0].data(), *as = A[1].data(), *ai = A[2].data();
#pragma omp parallel for schedule(static) num_threads(nthreads())
        for (long long i = 0; i < (long long)N; ++i) {
            cx<R> p0 = ap[i], s0 = as[i], i0 = ai[i];
            cx<R> k1p, k1s, k1i, k2p, k2s, k2i, k3p, k3s, k3i, k4p, k4s, k4i;
            rhs(r1, p0, s0, i0, k1p, k1s, k1i);
            rhs(r2, p0 + k1p * c1, s0 + k1s * c1, i0 + k1i * c1, k2p, k2s, k2i);
            rhs(r2, p0 + k2p * c1, s0 + k2s * c1, i0 + k2i * c1, k3p, k3s, k3i);
            rhs(r4, p0 + k3p * c3, s0 + k3s * c3, i0 + k3i * c3, k4p, k4s, k4i);
            const R two = (R)2, six = (R)6;
            ap[i] = p0 + ((k1p + two * k2p + two * k3p + k4p) * hr) / six;
            as[i] = s0 + ((k1s + two * k2s + two * k3s + k4s) * hr) / six;
            ai[i] = i0 + ((k1i + two * k2i + two * k3i + k4i) * hr) / six;
        }
    }

    // Along-t operator: IFFT_t (+i), /NT, x table, FFT_t (-i), on every pixel
    // (Solver3D.cuh:571-639, with the y = 0-only launch and the 1/SIZE scale
    // fixed to all pixels and 1/NT).
    void t_operator(std::vector<cx<R>> &a, const std::vector<cx<R>> &tab) {
        const int nt = g.nt;
        if (nt == 1) {
            for (size_t i = 0; i < N; ++i) a[i] = a[i] * tab[0];
            return;
        }
        auto pi = get_plan<R>(nt, +1), pf = get_plan<R>(nt, -1);
        const size_t B = 8;
        const long long nblk = (long long)((P + B - 1) / B);
        const R rnt = (R)nt;
#pragma omp parallel num_threads(nthreads())
        {
            std::vector<cx<R>> buf((size_t)nt * B), scr(nt);
#pragma omp for schedule(static)
            for (long long b = 0; b < nblk; ++b) {
                size_t p0 = (size_t)b * B, nb = std::min(B, P - p0);
                for (int t = 0; t < nt; ++t)
                    for (size_t j = 0; j < nb; ++j) buf[j * nt + t] = a[(size_t)t * P + p0 + j];
                for (size_t j = 0; j < nb; ++j) {
                    cx<R> *x = &buf[j * nt];
                    fft_exec(*pi, x, scr.data());
                    for (int k = 0; k < nt; ++k) x[k] = (x[k] / rnt) * tab[k];
                    fft_exec(*pf, x, scr.data());
                }
                for (int t = 0; t < nt; ++t)
                    for (size_t j = 0; j < nb; ++j) a[(size_t)t * P + p0 + j] = buf[j * nt + t];
            }
        }
    }

    // xy operator: FFT_xy (-i), table x field, IFFT_xy (+i), /(NX*NY)
    // (Solver3D.cuh:725-788 and 817-869, with the 1/SIZE scale of :780 fixed
    // to 1/(NX*NY) as in :102-119).
    void xy_operator(std::vector<cx<R>> &a, const std::vector<cx<R>> &tab) {
        const int nx = g.nx, ny = g.ny, nt = g.nt;
        auto fx = get_plan<R>(nx, -1), ix = get_plan<R>(nx, +1);
        auto fy = get_plan<R>(ny, -1), iy = get_plan<R>(ny, +1);
        const R rsz = (R)((double)nx * ny);
#pragma omp parallel num_threads(nthreads())
        {
            std::vector<cx<R>> col(ny), scr(std::max(nx, ny));
#pragma omp for schedule(static)
            for (int t = 0; t < nt; ++t) {
                cx<R> *pl = &a[(size_t)t * P];
                for (int y = 0; y < ny; ++y) fft_exec(*fx, pl + (size_t)y * nx, scr.data());
                for (int x = 0; x < nx; ++x) {
                    for (int y = 0; y < ny; ++y) col[y] = pl[(size_t)y * nx + x];
                    fft_exec(*fy, col.data(), scr.data());
                    for (int y = 0; y < ny; ++y) pl[(size_t)y * nx + x] = col[y];
                }
                for (size_t i = 0; i < P; ++i) pl[i] = tab[i] * pl[i];
                for (int x = 0; x < nx; ++x) {
                    for (int y = 0; y < ny; ++y) col[y] = pl[(size_t)y * nx + x];
                    fft_exec(*iy, col.data(), scr.data());
                    for (int y = 0; y < ny; ++y) pl[(size_t)y * nx + x] = col[y];
                }
                for (int y = 0; y < ny; ++y) fft_exec(*ix, pl + (size_t)y * nx, scr.data());
                for (size_t i = 0; i < P; ++i) pl[i] = pl[i] / rsz;
            }
        }
    }

    void dispersion() override {
        for (int f = 0; f < 3; ++f) t_operator(A[f], eiLz[f]);
    }
    void diffraction() override {
        for (int f = 0; f < 3; ++f) xy_operator(A[f], eiQz[f]);
    }
    void free_space() override {
        for (int f = 0; f < 2; ++f) xy_operator(A[1 + f], eiQfs[f]);
    }
    // GDDCompensation (Solver3D.cuh:643-694) on As and Ai, all pixels, 1/NT.
    void gdd() override {
        t_operator(A[1], egdd);
        t_operator(A[2], egdd);
    }
    // KernelEOM (Solver1D.cuh:62-72) on As and Ai: A(t) *= exp(i beta
    // sin(2 pi fpm t)), t = linspace(-W/2, W/2, NT) (Common.cuh:108-118).
    void eom() override {
        const int nt = g.nt;
        for (int k = 0; k < nt; ++k) {
            const double t = nt > 1 ? -0.5 * p.t_window + k * p.t_window / (nt - 1) : 0.0;
            const double ph = p.eom_beta * std::sin(2.0 * M_PI * p.eom_fpm * t);
            const cx<R> m = {(R)std::cos(ph), (R)std::sin(ph)};
            for (int f = 1; f < 3; ++f) {
                cx<R> *a = A[f].data() + (size_t)k * P;
                for (size_t i = 0; i < P; ++i) a[i] = m * a[i];
            }
        }
    }
    // roundTripChange (Cavity3D.cuh:12-28): A <- (sqrt(R) e^{i phi}) A for
    // the s and i multipliers (1 + 0i for a non-resonant field, R13).
    void mirror() override {
        for (int f = 0; f < 2; ++f) {
            cx<R> m = {(R)p.fb_re[f], (R)p.fb_im[f]};
            auto &a = A[1 + f];
#pragma omp parallel for schedule(static) num_threads(nthreads())
            for (long long i = 0; i < (long long)N; ++i) a[i] = m * a[i];
        }
    }
    // averagePowerCutXY (Solver3D.cuh:529-545): pref * sum |A|^2, summed in
    // long double (the reference's FP32 thrust::reduce is replaced).
    void power(double out[3]) override {
        for (int f = 0; f < 3; ++f) {
            long double acc = 0.0L;
            const auto &a = A[f];
            for (size_t i = 0; i < N; ++i) {
                long double re = a[i].re, im = a[i].im;
                acc += re * re + im * im;
            }
            out[f] = (double)(acc * (long double)p.power_prefactor[f]);
        }
    }
    void table(int which, double *out) override {
        const std::vector<cx<R>> *t = nullptr;
        if (which >= 0 && which < 3) t = &eiLz[which];
        else if (which >= 3 && which < 6) t = &eiQz[which - 3];
        else if (which == 6 || which == 7) t = &eiQfs[which - 6];
        else if (which == 8) t = &egdd;
        if (!t) return;
        for (size_t i = 0; i < t->size(); ++i) {
            out[2 * i] = (double)(*t)[i].re;
            out[2 * i + 1] = (double)(*t)[i].im;
        }
    }
    void omega(double *out) override {
        for (int i = 0; i < g.nt; ++i) out[i] = w[i];
    }

    void pump_reinject() { A[0] = Api; }  // Solver3D.cuh:981
};

}  // namespace

struct oracle_solve {
    std::unique_ptr<Base> s;
    std::function<void()> reinject;
};

extern "C" {

oracle_solve *oracle_create(const opo_grid *g, const opo_params *p, int precision) {
    if (!g || !p || g->nx < 1 || g->ny < 1 || g->nt < 1) return nullptr;
    auto *o = new oracle_solve;
    if (precision) {
        auto *s = new Solve<double>(*g, *p);
        o->s.reset(s);
        o->reinject = [s]() { s->pump_reinject(); };
    } else {
        auto *s = new Solve<float>(*g, *p);
        o->s.reset(s);
        o->reinject = [s]() { s->pump_reinject(); };
    }
    return o;
}
void oracle_destroy(oracle_solve *o) { delete o; }
void oracle_set_threads(int n) { g_threads = n; }
int oracle_get_threads(void) { return nthreads(); }

void oracle_set_fields(oracle_solve *o, const double *api, const double *as, const double *ai) {
    o->s->set_fields(api, as, ai);
}
void oracle_get_fields(oracle_solve *o, double *ap, double *as, double *ai) {
    o->s->get_fields(ap, as, ai);
}
void oracle_step_nonlinear(oracle_solve *o) { o->s->nonlinear(0.0); }
void oracle_step_dispersion(oracle_solve *o) { o->s->dispersion(); }
void oracle_step_diffraction(oracle_solve *o) { o->s->diffraction(); }
void oracle_step_free_space(oracle_solve *o) { o->s->free_space(); }
void oracle_step_gdd(oracle_solve *o) { o->s->gdd(); }
void oracle_step_mirror(oracle_solve *o) { o->s->mirror(); }
void oracle_step_eom(oracle_solve *o) { o->s->eom(); }

// SSFM (Solver3D.cuh:946-960): per slice N(dz/2) D(dz) Q(dz) N(dz/2), with
// the slice position z_k = k*dz passed to both half steps.
void oracle_ssfm(oracle_solve *o) {
    Base &s = *o->s;
    for (int k = 0; k < s.p.n_slices; ++k) {
        double z = k * s.p.dz;
        s.nonlinear(z);
        s.dispersion();
        if (!s.p.no_diffraction) s.diffraction();  // plane-wave model: Solver1D.cuh:436-445
        s.nonlinear(z);
    }
}

// runOPO body (Solver3D.cuh:979-1011): Ap <- Api; SSFM; free space unless
// this is trip NRT-1 (R17); GDD on s and i; mirror; per-trip power.
int oracle_run_round_trips(oracle_solve *o, int n, double *power) {
    Base &s = *o->s;
    for (int i = 0; i < n; ++i) {
        int rt = s.trips_done;
        o->reinject();
        oracle_ssfm(o);
        bool last = s.p.skip_fs_on_last_trip && rt == s.p.n_round_trips - 1;
        if (s.p.fs_on && !last && !s.p.no_diffraction) s.free_space();
        if (s.p.eom_on) s.eom();  // Solver1D.cuh:467, before GDD
        if (s.p.gdd_on) s.gdd();
        s.mirror();
        double pw[3];
        s.power(pw);
        if (power) for (int f = 0; f < 3; ++f) power[3 * i + f] = pw[f];
        s.trips_done++;
        if (!std::isfinite(pw[1])) return i + 1;
    }
    return n;
}
void oracle_power(oracle_solve *o, double out[3]) { o->s->power(out); }
void oracle_table(oracle_solve *o, int which, double *out) { o->s->table(which, out); }
void oracle_omega(oracle_solve *o, double *out) { o->s->omega(out); }

void oracle_fft(int n, int sign, const double *in, double *out, int precision) {
    if (precision) {
        auto P = get_plan<double>(n, sign);
        std::vector<cx<double>> x(n), y(n);
        for (int i = 0; i < n; ++i) x[i] = {in[2 * i], in[2 * i + 1]};
        fft_exec(*P, x.data(), y.data());
        for (int i = 0; i < n; ++i) { out[2 * i] = x[i].re; out[2 * i + 1] = x[i].im; }
    } else {
        auto P = get_plan<float>(n, sign);
        std::vector<cx<float>> x(n), y(n);
        for (int i = 0; i < n; ++i) x[i] = {(float)in[2 * i], (float)in[2 * i + 1]};
        fft_exec(*P, x.data(), y.data());
        for (int i = 0; i < n; ++i) { out[2 * i] = x[i].re; out[2 * i + 1] = x[i].im; }
    }
}

void oracle_dft_slow(int n, int sign, const double *in, double *out) {
    for (int k = 0; k < n; ++k) {
        long double ar = 0, ai = 0;
        for (int j = 0; j < n; ++j) {
            long double c, s;
            unit_root((long long)sign * j * k, n, c, s);
            long double xr = in[2 * j], xi = in[2 * j + 1];
            ar += xr * c - xi * s;
            ai += xr * s + xi * c;
        }
        out[2 * k] = (double)ar;
        out[2 * k + 1] = (double)ai;
    }
}

// MgOPPLN.cuh:65 (dz = float(Lcr/(NZ-1.0f))) and Solver3D.cuh:949 (the
// float loop), evaluated in IEEE single precision.
int oracle_slice_count(double lcr, int nz) {
    volatile float L = (float)lcr;
    volatile float dz = (float)(L / ((float)nz - 1.0f));
    int count = 0;
    for (volatile float z = 0.0f; z <= L; z = z + dz) ++count;
    return count;
}

}  // extern "C"
