// opo_solve.hpp -- one OPO solve on one device, behind the C ABI (include/opo3d.h).
//
// One opo_solve owns the device state of one OPO solve: the injected pump
// Api, the fields Ap, As, Ai (layout [t][y][x], interleaved complex), the
// propagator tables (built once in FP64, phases reduced mod 2*pi, then
// rounded to the solve's precision) and the power/NaN bookkeeping.  Work is
// queued on the solve's own stream; a round trip is captured once as a HIP
// graph and replayed.
//
// Slice schedule (reference SSFM, Solver3D.cuh:946-960):
//   N(h) D Q N(h) | N(h) D Q N(h) | ... is run as
//   [N]  [D] [Q]  [N N] [D] [Q] ... [N N] [D] [Q]  [N]
// where consecutive half steps of adjacent slices share one pointwise pass.
// Within a round trip the first pass reads the pump from Api (re-injection,
// Solver3D.cuh:981) and the last one applies the mirror factors
// (Cavity3D.cuh:12-28), which commute with the free-space and GDD operators
// that follow.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/opo3d.h"
#include "../../include/opo3d_setup.h"
#include "opo_device.hpp"
#include "opo_internal.hpp"
#include "opo_kernels.hpp"
#include "opo_power.hpp"

namespace opo {

// Text of the last error recorded by fail() on the calling thread.
const std::string &last_error_text();

inline constexpr double kTwoPi = 6.283185307179586476925286766559;

#define CK(call)                                                                           \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(OPO_ERR_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_),       \
                        __FILE__, __LINE__);                                               \
    } while (0)

// exp(-2 pi i j/n) from long double on exact fractions.
inline void unit_root_neg(long j, long n, double &c, double &s) {
    const long double pi = 3.141592653589793238462643383279502884L;
    long double a = -2.0L * pi * (long double)(j % n) / (long double)n;
    c = (double)cosl(a);
    s = (double)sinl(a);
}

// Signed frequency index of FFT-order bin j: the reference builds centred
// tables and fftShift2D's them (Solver3D.cuh:288-296, EFields3D.cuh:141-171).
inline double signed_freq(int j, int n) {
    if (n < 2) return 0.0;
    int c = n / 2;
    return (double)(((j + c) % n) - c);
}

inline bool grid_supported(const opo_grid &g, bool fp64) {
    if (!tline_supported(g.nt, fp64)) return false;
    if (!xyplane_supported(g.nx, g.ny, fp64)) return false;
    return true;
}

// -------------------------------------------------------------- solve
struct SolveImpl {
    opo_grid g{};
    opo_params p{};
    std::vector<double> omega;
    int device = 0;
    int prec = OPO_FP32;
    size_t N = 0, P = 0;
    int trips_done = 0;
    hipStream_t stream = nullptr;
    bool use_graphs = true;
    virtual ~SolveImpl() {}
    virtual int init() = 0;
    virtual int set_fields(const void *, const void *, const void *) = 0;
    virtual int get_fields(void *, void *, void *) = 0;
    virtual void *dev_ptr(int which) = 0;
    virtual int step(int which) = 0;
    virtual int ssfm() = 0;
    virtual int enqueue_trips(int n) = 0;
    virtual int finish_trips(int n, double *power) = 0;
    virtual int power(double out[3]) = 0;
    virtual int power_time(int f, int x, int y, double pref, double *out) = 0;
    virtual int profile_trips(int n, double *ms, int *launches) = 0;
    // Write this solve's SolveDesc<T> (T = the solve's precision) to out.
    virtual void fill_desc(void *out, double *series, int series_cap, int nchunk) const = 0;
    virtual int autotune(int reps, std::string &report) = 0;
    size_t field_bytes() const { return N * (prec == OPO_FP64 ? 16 : 8); }
    int select() {
        CK(hipSetDevice(device));
        return OPO_OK;
    }
};

enum { STEP_NONLINEAR, STEP_DISPERSION, STEP_DIFFRACTION, STEP_FREE_SPACE, STEP_GDD, STEP_MIRROR };

template <typename T> struct SolveT final : SolveImpl {
    using C = cplx<T>;
    C *Api = nullptr, *A[3] = {nullptr, nullptr, nullptr};
    C *tables = nullptr;
    C *tw_x = nullptr, *tw_y = nullptr, *tw_t = nullptr, *tabT[3] = {}, *tabG = nullptr, *qx[3] = {}, *qy[3] = {},
      *qfsx[2] = {}, *qfsy[2] = {}, *tabE = nullptr;
    double *dbuf = nullptr;   // partials | pw3
    double *partial = nullptr, *pw3 = nullptr;
    double *series = nullptr;
    int series_cap = 0;
    int *flags = nullptr;     // [0] status, [1] trip counter
    int nblk = 64;
    // Captured work: [0] a round trip with free space, [1] one without,
    // [2] a single pass (opo_single_pass).
    hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};
    int graph_variants[2] = {-1, -1};   // (t, xy) launch variants the graphs were captured with

    bool owns = true;   // false: every device buffer belongs to someone else

    ~SolveT() override {
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto &gx : graph)
            if (gx) (void)hipGraphExecDestroy(gx);
        if (!owns) return;
        (void)hipFree(Api);
        for (auto *a : A) (void)hipFree(a);
        (void)hipFree(tables);
        (void)hipFree(dbuf);
        (void)hipFree(series);
        (void)hipFree(flags);
        for (auto e : ev_pool) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int init() override {
        P = (size_t)g.nx * g.ny;
        N = P * g.nt;
        int rc = select();
        if (rc) return rc;
        CK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        const size_t fb = N * sizeof(C);
        if (hipMalloc(&Api, fb) != hipSuccess) return fail(OPO_ERR_NOMEM, "hipMalloc(%zu) failed", fb);
        for (auto &a : A)
            if (hipMalloc(&a, fb) != hipSuccess) return fail(OPO_ERR_NOMEM, "hipMalloc(%zu) failed", fb);
        CK(hipMemsetAsync(Api, 0, fb, stream));
        for (auto *a : A) CK(hipMemsetAsync(a, 0, fb, stream));
        nblk = power_blocks(N);
        CK(hipMalloc(&dbuf, sizeof(double) * (3 * (size_t)nblk + 4)));
        partial = dbuf;
        pw3 = dbuf + 3 * (size_t)nblk;
        CK(hipMalloc(&flags, 4 * sizeof(int)));
        CK(hipMemsetAsync(flags, 0, 4 * sizeof(int), stream));
        rc = build_tables();
        if (rc) return rc;
        CK(hipStreamSynchronize(stream));
        const char *ng = std::getenv("OPO_NO_GRAPH");
        use_graphs = !(ng && ng[0] == '1');
        return OPO_OK;
    }

    // Entries (cplx<T>) of the table block: twiddles for NX, NY and NT, the
    // three D tables, GDD, the x and y factors of Q for three fields, the
    // free-space x and y factors for s and i, and the modulator table, each
    // entry a (hi, lo) pair.
    static size_t table_entries(const opo_grid &g) {
        return 2 * ((size_t)g.nx + g.ny + g.nt + 3 * (size_t)g.nt + g.nt + 3 * (size_t)g.nx + 3 * (size_t)g.ny +
                    2 * (size_t)g.nx + 2 * (size_t)g.ny + g.nt);
    }
    struct TableOffsets {
        size_t twx, twy, twt, D, G, qx, qy, fx, fy, E;
    };
    static TableOffsets table_offsets(const opo_grid &g) {
        const int nx = g.nx, ny = g.ny, nt = g.nt;
        size_t off = 0;
        auto take = [&](size_t n) { size_t o = off; off += 2 * n; return o; };
        TableOffsets o;
        o.twx = take(nx);
        o.twy = take(ny);
        o.twt = take(nt);
        o.D = take(3 * (size_t)nt);
        o.G = take(nt);
        o.qx = take(3 * (size_t)nx);
        o.qy = take(3 * (size_t)ny);
        o.fx = take(2 * (size_t)nx);
        o.fy = take(2 * (size_t)ny);
        o.E = take(nt);
        return o;
    }

    // Fill the host image h (table_entries(g) elements) of the table block
    // from the parameters and the NT angular frequencies w.  With p.eom_on,
    // tgrid (NT doubles of host scratch) receives the time grid.
    static void fill_tables(const opo_grid &g, const opo_params &p, const double *w, C *h, double *tgrid) {
        const int nx = g.nx, ny = g.ny, nt = g.nt;
        const TableOffsets o = table_offsets(g);
        const size_t o_twx = o.twx, o_twt = o.twt, o_T = o.D, o_G = o.G, o_qx = o.qx, o_qy = o.qy,
                     o_fx = o.fx, o_fy = o.fy, o_E = o.E;
        // put(i, re, im): the pair of one entry at h[i] (hi) and h[i + 1] (lo);
        // entry j of a table starting at o lives at i = o + 2 j.  hi = value
        // rounded to T, lo = rounding residual (zero for T = double).
        auto put = [&](size_t i, double re, double im) {
            const T hr = (T)re, hi = (T)im;
            h[i] = C{hr, hi};
            h[i + 1] = C{(T)(re - (double)hr), (T)(im - (double)hi)};
        };
        for (size_t i = 0; i < table_entries(g); ++i) h[i] = C{(T)0, (T)0};
        for (int j = 0; j < nx; ++j) {
            double c, s;
            unit_root_neg(j, nx, c, s);
            put(o_twx + 2 * (size_t)j, c, s);
        }
        for (int j = 0; j < ny; ++j) {
            double c, s;
            unit_root_neg(j, ny, c, s);
            put(o.twy + 2 * (size_t)j, c, s);
        }
        for (int j = 0; j < nt; ++j) {
            double c, s;
            unit_root_neg(j, nt, c, s);
            put(o_twt + 2 * (size_t)j, c, s);
        }
        // D: eiLz_X / NT (Solver3D.cuh:153-169; the 1/NT of the time pair folded in).
        for (int f = 0; f < 3; ++f) {
            const double att = std::exp(-0.5 * p.alpha[f] * p.dz) / nt;
            for (int k = 0; k < nt; ++k) {
                double c, s;
                cis(p.dz * w[k] * (p.delta[f] + 0.5 * w[k] * p.beta2[f]), c, s);
                put(o_T + 2 * ((size_t)f * nt + k), c * att, s * att);
            }
        }
        // GDD: exp(i w^2 GDD / 2) / NT (Solver3D.cuh:261-277, 676).
        for (int k = 0; k < nt; ++k) {
            double c, s;
            cis(0.5 * w[k] * w[k] * p.gdd, c, s);
            put(o_G + 2 * (size_t)k, c / nt, s / nt);
        }
        // Q: exp(-i dz 2 pi^2/k (dfX^2 fx^2 + dfY^2 fy^2)) exp(-alpha dz/2) /(NX NY)
        // (Solver3D.cuh:282-301), split exactly into an x factor (carrying the
        // absorption and 1/NX) and a y factor (carrying 1/NY).
        const double dfX = nx > 1 ? 1.0 / (p.dx * nx) : 0.0;
        const double dfY = ny > 1 ? 1.0 / (p.dy * ny) : 0.0;
        const double pi2 = 9.869604401089358618834490999876151;  // pi^2
        for (int f = 0; f < 3; ++f) {
            const double att = std::exp(-0.5 * p.alpha[f] * p.dz);
            for (int j = 0; j < nx; ++j) {
                double fx = signed_freq(j, nx), c, s;
                cis(-p.dz * (2.0 * pi2 / p.kdiff[f] * (dfX * dfX * fx * fx)), c, s);
                put(o_qx + 2 * ((size_t)f * nx + j), c * att / nx, s * att / nx);
            }
            for (int j = 0; j < ny; ++j) {
                double fy = signed_freq(j, ny), c, s;
                cis(-p.dz * (2.0 * pi2 / p.kdiff[f] * (dfY * dfY * fy * fy)), c, s);
                put(o_qy + 2 * ((size_t)f * ny + j), c / ny, s / ny);
            }
        }
        // Free space on s and i: exp(+i L 2 pi^2/k (dfX^2 fx^2 + dfY^2 fy^2)), no
        // absorption (Solver3D.cuh:305-323), normalised by 1/(NX NY) (:102-119).
        for (int f = 0; f < 2; ++f) {
            for (int j = 0; j < nx; ++j) {
                double fx = signed_freq(j, nx), c, s;
                cis(p.fs_length * (2.0 * pi2 / p.kdiff[1 + f] * (dfX * dfX * fx * fx)), c, s);
                put(o_fx + 2 * ((size_t)f * nx + j), c / nx, s / nx);
            }
            for (int j = 0; j < ny; ++j) {
                double fy = signed_freq(j, ny), c, s;
                cis(p.fs_length * (2.0 * pi2 / p.kdiff[1 + f] * (dfY * dfY * fy * fy)), c, s);
                put(o_fy + 2 * ((size_t)f * ny + j), c / ny, s / ny);
            }
        }
        // Modulator on s and i: exp(i beta sin(2 pi fpm t_j)) on the time grid
        // t_j = -W/2 + j W/(NT-1) (KernelEOM, Solver1D.cuh:62-72).
        if (p.eom_on) {
            opo_time_grid(nt, p.t_window, tgrid);
            for (int k = 0; k < nt; ++k) {
                double c, s2;
                cis(p.eom_beta * std::sin(kTwoPi * std::fmod(p.eom_fpm * tgrid[k], 1.0)), c, s2);
                put(o_E + 2 * (size_t)k, c, s2);
            }
        }
    }

    // Point the table pointers into a device table block.
    void bind_tables(C *block) {
        const TableOffsets o = table_offsets(g);
        tables = block;
        tw_x = tables + o.twx;
        tw_y = tables + o.twy;
        tw_t = tables + o.twt;
        for (int f = 0; f < 3; ++f) {
            tabT[f] = tables + o.D + 2 * (size_t)f * g.nt;
            qx[f] = tables + o.qx + 2 * (size_t)f * g.nx;
            qy[f] = tables + o.qy + 2 * (size_t)f * g.ny;
        }
        tabG = tables + o.G;
        tabE = p.eom_on ? tables + o.E : nullptr;
        for (int f = 0; f < 2; ++f) {
            qfsx[f] = tables + o.fx + 2 * (size_t)f * g.nx;
            qfsy[f] = tables + o.fy + 2 * (size_t)f * g.ny;
        }
    }

    // Run on buffers owned by the caller (the no-allocation entry point):
    // fields api, a0..a2, a table block already filled, dbuf_ of
    // 3 power_blocks(N) + 4 doubles, flags_ of 4 ints and a power series of
    // series_cap_ trips.  Launches go to the null stream, without graphs.
    void init_borrowed(C *api, C *a0, C *a1, C *a2, C *block, double *dbuf_, int *flags_, double *series_,
                       int series_cap_) {
        owns = false;
        P = (size_t)g.nx * g.ny;
        N = P * g.nt;
        Api = api;
        A[0] = a0;
        A[1] = a1;
        A[2] = a2;
        bind_tables(block);
        nblk = power_blocks(N);
        dbuf = dbuf_;
        partial = dbuf;
        pw3 = dbuf + 3 * (size_t)nblk;
        flags = flags_;
        series = series_;
        series_cap = series_cap_;
        use_graphs = false;
    }

    int build_tables() {
        const size_t total = table_entries(g);
        std::vector<C> h(total);
        std::vector<double> tg(g.nt);
        fill_tables(g, p, omega.data(), h.data(), tg.data());
        C *block = nullptr;
        if (hipMalloc(&block, total * sizeof(C)) != hipSuccess)
            return fail(OPO_ERR_NOMEM, "hipMalloc(tables) failed");
        bind_tables(block);
        CK(hipMemcpy(tables, h.data(), total * sizeof(C), hipMemcpyHostToDevice));
        return OPO_OK;
    }

    // RK4 arguments for nsteps consecutive half steps at slice positions z0, z1.
    Rk4Args<T> rk_args(int nsteps, double z0, double z1, bool mirror) const {
        Rk4Args<T> a{};
        const double h = 0.5 * p.dz;
        const double c1 = p.rk4_h_in_stages ? 0.5 * h : 0.5, c3 = p.rk4_h_in_stages ? h : 1.0;
        for (int f = 0; f < 3; ++f) {
            a.kc1[f] = (T)(p.kappa[f] * c1);
            a.kc3[f] = (T)(p.kappa[f] * c3);
            a.kh6[f] = (T)(p.kappa[f] * h / 6.0);
        }
        a.nsteps = nsteps;
        a.has_phase = (p.dk_phase_scale != 0.0 && p.dk != 0.0) ? 1 : 0;
        for (int st = 0; st < 2; ++st) {
            const double z = st == 0 ? z0 : z1;
            // stage positions z, z + dz/4, z + dz/2 (Solver3D.cuh:911-929)
            const double zs[3] = {z, z + 0.25 * p.dz, z + 0.5 * p.dz};
            for (int k = 0; k < 3; ++k) {
                double c = 1.0, s = 0.0;
                if (a.has_phase) cis(p.dk_phase_scale * p.dk * zs[k], c, s);
                a.es_re[st][k] = (T)c;
                a.es_im[st][k] = (T)s;
            }
        }
        a.out_scale = mirror ? 1 : 0;
        a.ms_re = (T)p.fb_re[0];
        a.ms_im = (T)p.fb_im[0];
        a.mi_re = (T)p.fb_re[1];
        a.mi_im = (T)p.fb_im[1];
        a.ms_lo_re = (T)(p.fb_re[0] - (double)a.ms_re);
        a.ms_lo_im = (T)(p.fb_im[0] - (double)a.ms_im);
        a.mi_lo_re = (T)(p.fb_re[1] - (double)a.mi_re);
        a.mi_lo_im = (T)(p.fb_im[1] - (double)a.mi_im);
        a.eom = mirror ? tabE : nullptr;
        return a;
    }

    Geometry geo() const { return Geometry{g.nx, g.ny, g.nt}; }

    // Stage timing (opo_profile_trips): when profiling, every launch is
    // bracketed by a pair of events recorded on the solve's stream.
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> ev_stage;  // stage of the pair (2k, 2k+1)
    size_t ev_used = 0;
    int prof_mark(int stage) {
        if (!profiling) return OPO_OK;
        if (ev_used + 1 >= ev_pool.size()) {
            for (int k = 0; k < 64; ++k) {
                hipEvent_t e;
                CK(hipEventCreate(&e));
                ev_pool.push_back(e);
            }
        }
        if (ev_used % 2 == 0) ev_stage.push_back(stage);
        CK(hipEventRecord(ev_pool[ev_used++], stream));
        return OPO_OK;
    }
#define PROF(stage, launch_expr)                 \
    do {                                         \
        int prc_ = prof_mark(stage);             \
        if (prc_) return prc_;                   \
        CK(launch_expr);                         \
        if ((prc_ = prof_mark(stage))) return prc_; \
    } while (0)

    int q_rk4(const C *psrc, const Rk4Args<T> &a, const int *st) {
        PROF(OPO_STAGE_NONLINEAR, (launch_rk4<T>(stream, psrc, A[0], A[1], A[2], N, P, a, st)));
        return OPO_OK;
    }
    int q_dispersion(const int *st) {
        FieldOps<T> ops{};
        for (int f = 0; f < 3; ++f) { ops.f[f] = A[f]; ops.tab[f] = tabT[f]; }
        PROF(OPO_STAGE_TIME_OP, (launch_tline<T>(stream, geo(), ops, 3, tw_t, st)));
        return OPO_OK;
    }
    int q_gdd(const int *st) {
        FieldOps<T> ops{};
        for (int f = 0; f < 2; ++f) { ops.f[f] = A[1 + f]; ops.tab[f] = tabG; }
        PROF(OPO_STAGE_GDD, (launch_tline<T>(stream, geo(), ops, 2, tw_t, st)));
        return OPO_OK;
    }
    int q_diffraction(const int *st) {
        FieldOps<T> ops{};
        for (int f = 0; f < 3; ++f) { ops.f[f] = A[f]; ops.tab[f] = qx[f]; ops.taby[f] = qy[f]; }
        PROF(OPO_STAGE_XY_OP, (launch_xyplane<T>(stream, geo(), ops, 3, tw_x, tw_y, st)));
        return OPO_OK;
    }
    int q_free_space(const int *st) {
        FieldOps<T> ops{};
        for (int f = 0; f < 2; ++f) { ops.f[f] = A[1 + f]; ops.tab[f] = qfsx[f]; ops.taby[f] = qfsy[f]; }
        PROF(OPO_STAGE_FREE_SPACE, (launch_xyplane<T>(stream, geo(), ops, 2, tw_x, tw_y, st)));
        return OPO_OK;
    }

    // Fused pass: pointwise half step(s) then the along-t operator on fields
    // first..2 with tables tab0 + (f - first) * tab_stride.
    int q_nl_time(const C *psrc, const Rk4Args<T> &rk, const C *tab0, int tab_stride, int first,
                  int stage, const int *st) {
        TLine3Args<T> a{};
        a.psrc = psrc;
        for (int f = 0; f < 3; ++f) a.f[f] = A[f];
        a.tab0 = tab0;
        a.tab_stride = tab_stride;
        a.first_field = first;
        a.rk = rk;
        PROF(stage, (launch_tline3<T>(stream, geo(), a, tw_t, st)));
        return OPO_OK;
    }

    // SSFM (Solver3D.cuh:946-960) on the fused schedule: per slice one pass
    // [N(h) of slice k-1, N(h) of slice k, D] and one pass [Q].  trip: read
    // the pump from Api, and end with [N(h), mirror, GDD of s and i] (the
    // mirror and GDD commute with the free-space step that follows).
    int q_ssfm(bool trip, const int *st) {
        const int S = p.n_slices;
        int rc;
        if (S <= 0) {
            if (trip) CK(hipMemcpyAsync(A[0], Api, N * sizeof(C), hipMemcpyDeviceToDevice, stream));
            return OPO_OK;
        }
        for (int k = 0; k < S; ++k) {
            const double z = k * p.dz;
            const C *psrc = (k == 0 && trip) ? Api : A[0];
            Rk4Args<T> rk = k == 0 ? rk_args(1, z, 0.0, false) : rk_args(2, (k - 1) * p.dz, z, false);
            if ((rc = q_nl_time(psrc, rk, tabT[0], (int)(tabT[1] - tabT[0]), 0, OPO_STAGE_TIME_OP, st)))
                return rc;
            if (!p.no_diffraction && (rc = q_diffraction(st))) return rc;
        }
        const Rk4Args<T> last = rk_args(1, (S - 1) * p.dz, 0.0, trip);
        if (trip && p.gdd_on) return q_nl_time(A[0], last, tabG, 0, 1, OPO_STAGE_GDD, st);
        return q_rk4(A[0], last, st);
    }

    // One round trip (Solver3D.cuh:979-1011) with the per-trip power.
    int q_trip(bool fs) {
        int rc = q_ssfm(true, flags);  // ends with the mirror and GDD
        if (rc) return rc;
        if (fs && !p.no_diffraction && (rc = q_free_space(flags))) return rc;
        int prc = prof_mark(OPO_STAGE_POWER);
        if (prc) return prc;
        CK(launch_power_partial<T>(stream, A[0], A[1], A[2], N, nblk, partial, flags));
        CK(launch_power_final(stream, partial, nblk, p.power_prefactor[0], p.power_prefactor[1],
                              p.power_prefactor[2], series, flags + 1, flags, 1, series_cap));
        return prof_mark(OPO_STAGE_POWER);
    }

    // n trips without graphs, events around every launch; sums per stage.
    int profile_trips(int n, double *ms, int *launches) override {
        int rc = select();
        if (rc) return rc;
        for (int k = 0; k < OPO_N_STAGES; ++k) { ms[k] = 0.0; launches[k] = 0; }
        if ((rc = ensure_series(n))) return rc;
        CK(hipMemsetAsync(flags, 0, 2 * sizeof(int), stream));
        profiling = true;
        for (int i = 0; i < n && !rc; ++i) {
            ev_used = 0;
            ev_stage.clear();
            const int rt = trips_done;
            const bool fs = p.fs_on && !(p.skip_fs_on_last_trip && rt == p.n_round_trips - 1);
            rc = q_trip(fs);
            if (rc) break;
            hipError_t e = hipStreamSynchronize(stream);
            if (e != hipSuccess) { rc = fail(OPO_ERR_HIP, "profile sync: %s", hipGetErrorString(e)); break; }
            for (size_t k = 0; k + 1 < ev_used; k += 2) {
                float t = 0.f;
                e = hipEventElapsedTime(&t, ev_pool[k], ev_pool[k + 1]);
                if (e != hipSuccess) { rc = fail(OPO_ERR_HIP, "event time: %s", hipGetErrorString(e)); break; }
                const int st = ev_stage[k / 2];
                ms[st] += t;
                launches[st] += 1;
            }
            trips_done++;
        }
        profiling = false;
        return rc;
    }

    int set_fields(const void *api, const void *as, const void *ai) override {
        int rc = select();
        if (rc) return rc;
        const size_t fb = N * sizeof(C);
        CK(hipStreamSynchronize(stream));
        if (api) {
            CK(hipMemcpy(Api, api, fb, hipMemcpyHostToDevice));
            CK(hipMemcpy(A[0], Api, fb, hipMemcpyDeviceToDevice));
        }
        if (as) CK(hipMemcpy(A[1], as, fb, hipMemcpyHostToDevice));
        if (ai) CK(hipMemcpy(A[2], ai, fb, hipMemcpyHostToDevice));
        return OPO_OK;
    }
    int get_fields(void *ap, void *as, void *ai) override {
        int rc = select();
        if (rc) return rc;
        const size_t fb = N * sizeof(C);
        CK(hipStreamSynchronize(stream));
        void *dst[3] = {ap, as, ai};
        for (int f = 0; f < 3; ++f)
            if (dst[f]) CK(hipMemcpy(dst[f], A[f], fb, hipMemcpyDeviceToHost));
        return OPO_OK;
    }
    void *dev_ptr(int which) override {
        if (which >= 0 && which < 3) return A[which];
        if (which == 3) return Api;
        return nullptr;
    }

    int step(int which) override {
        int rc = select();
        if (rc) return rc;
        switch (which) {
            case STEP_NONLINEAR: rc = q_rk4(A[0], rk_args(1, 0.0, 0.0, false), nullptr); break;
            case STEP_DISPERSION: rc = q_dispersion(nullptr); break;
            case STEP_DIFFRACTION: rc = q_diffraction(nullptr); break;
            case STEP_FREE_SPACE: rc = q_free_space(nullptr); break;
            case STEP_GDD: rc = p.gdd_on ? q_gdd(nullptr) : OPO_OK; break;
            case STEP_MIRROR: {
                // mirror alone: a zero-coupling RK4 pass is the identity, then the factors.
                Rk4Args<T> a = rk_args(1, 0.0, 0.0, true);
                for (int f = 0; f < 3; ++f) a.kc1[f] = a.kc3[f] = a.kh6[f] = (T)0;
                rc = q_rk4(A[0], a, nullptr);
                break;
            }
            default: return fail(OPO_ERR_ARG, "unknown step %d", which);
        }
        if (rc) return rc;
        CK(hipStreamSynchronize(stream));
        return OPO_OK;
    }

    int ssfm() override {
        int rc = select();
        if (rc) return rc;
        if (p.n_slices > 0 && (rc = launch_graph(2))) return rc;   // no slices: nothing to capture
        CK(hipStreamSynchronize(stream));
        return OPO_OK;
    }

    int ensure_series(int n) {
        if (n <= series_cap) return OPO_OK;
        if (!owns) return fail(OPO_ERR_ARG, "%d trips exceed the %d-trip power series provided", n, series_cap);
        for (auto &gx : graph)
            if (gx) { (void)hipGraphExecDestroy(gx); gx = nullptr; }
        CK(hipStreamSynchronize(stream));
        (void)hipFree(series);
        series = nullptr;
        int cap = n < 64 ? 64 : n;
        if (hipMalloc(&series, sizeof(double) * 3 * (size_t)cap) != hipSuccess)
            return fail(OPO_ERR_NOMEM, "hipMalloc(series) failed");
        series_cap = cap;
        return OPO_OK;
    }

    // Enqueue graph v (see graph[]), capturing it on first use.
    int launch_graph(int v) {
        auto enqueue = [&]() { return v == 2 ? q_ssfm(false, nullptr) : q_trip(v == 0); };
        if (!use_graphs) return enqueue();
        // the launch variants are process-wide: re-capture when they changed
        const bool fp64 = sizeof(T) == 8;
        const int cur[2] = {time_variant_for(fp64, g.nt, g.nx), tune_get(fp64, 1, g.nx)};
        if (cur[0] != graph_variants[0] || cur[1] != graph_variants[1]) {
            for (auto &gx : graph)
                if (gx) { (void)hipGraphExecDestroy(gx); gx = nullptr; }
            graph_variants[0] = cur[0];
            graph_variants[1] = cur[1];
        }
        if (!graph[v]) {
            hipGraph_t gr = nullptr;
            CK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
            int rc = enqueue();
            hipError_t e = hipStreamEndCapture(stream, &gr);
            if (rc) {
                if (gr) (void)hipGraphDestroy(gr);
                return rc;
            }
            if (e != hipSuccess) return fail(OPO_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
            e = hipGraphInstantiate(&graph[v], gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (e != hipSuccess) return fail(OPO_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
        }
        CK(hipGraphLaunch(graph[v], stream));
        return OPO_OK;
    }

    int pending = 0;
    int enqueue_trips(int n) override {
        int rc = select();
        if (rc) return rc;
        if ((rc = ensure_series(n))) return rc;
        CK(hipMemsetAsync(flags, 0, 2 * sizeof(int), stream));
        for (int i = 0; i < n; ++i) {
            const int rt = trips_done + i;
            const bool fs = p.fs_on && !(p.skip_fs_on_last_trip && rt == p.n_round_trips - 1);
            if ((rc = launch_graph(fs ? 0 : 1))) return rc;
        }
        pending = n;
        return OPO_OK;
    }
    int finish_trips(int n, double *power) override {
        int rc = select();
        if (rc) return rc;
        int hflags[2] = {0, 0};
        CK(hipMemcpyAsync(hflags, flags, sizeof(hflags), hipMemcpyDeviceToHost, stream));
        std::vector<double> hs(3 * (size_t)n);
        if (power && n > 0)
            CK(hipMemcpyAsync(hs.data(), series, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost,
                              stream));
        CK(hipStreamSynchronize(stream));
        pending = 0;
        const int done = hflags[0] ? hflags[0] : n;
        if (power) std::memcpy(power, hs.data(), sizeof(double) * 3 * (size_t)done);
        trips_done += done;
        if (hflags[0])
            return fail(OPO_ERR_NAN, "non-finite signal power at round trip %d", trips_done - 1);
        return OPO_OK;
    }

    // Time every variant of the two passes on this grid, keep the fastest.
    int autotune(int reps, std::string &report) override {
        int rc = select();
        if (rc) return rc;
        if (reps < 1) reps = 1;
        const bool fp64 = sizeof(T) == 8;
        const int orig[2] = {tune_get(fp64, 0, g.nt), tune_get(fp64, 1, g.nx)};
        const size_t fb = N * sizeof(C);
        C *save[3] = {nullptr, nullptr, nullptr};
        hipEvent_t ev[2] = {nullptr, nullptr};
        bool saved = false;
        rc = autotune_run(reps, report, save, ev, saved);
        // Whatever happened, put the fields back and release the scratch;
        // on failure the previous choices stand.
        if (saved)
            for (int f = 0; f < 3; ++f)
                (void)hipMemcpyAsync(A[f], save[f], fb, hipMemcpyDeviceToDevice, stream);
        const hipError_t es = hipStreamSynchronize(stream);
        for (auto *b : save) (void)hipFree(b);
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
        if (rc == OPO_OK && es != hipSuccess) rc = fail(OPO_ERR_HIP, "autotune: %s", hipGetErrorString(es));
        if (rc != OPO_OK) {
            tune_set(fp64, 0, g.nt, orig[0]);
            tune_set(fp64, 1, g.nx, orig[1]);
        }
        for (auto &gx : graph)   // re-capture with the chosen variants
            if (gx) { (void)hipGraphExecDestroy(gx); gx = nullptr; }
        return rc;
    }

    // Fastest of reps timed runs of the stretch (robust against other work
    // sharing the device during one of them).  The stretch is captured as a
    // HIP graph and replayed, as a round trip is, unless graphs are off.
    template <typename F> int time_stretch(F &stretch, int reps, hipEvent_t ev[2], float &ms) {
        hipGraphExec_t ge = nullptr;
        if (use_graphs) {
            hipGraph_t gr = nullptr;
            CK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
            const int rc = stretch();
            const hipError_t e = hipStreamEndCapture(stream, &gr);
            if (rc) {
                if (gr) (void)hipGraphDestroy(gr);
                return rc;
            }
            if (e != hipSuccess) return fail(OPO_ERR_HIP, "autotune capture: %s", hipGetErrorString(e));
            const hipError_t ei = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (ei != hipSuccess) return fail(OPO_ERR_HIP, "autotune instantiate: %s", hipGetErrorString(ei));
        }
        auto once = [&]() -> int {
            if (ge) return hipGraphLaunch(ge, stream) == hipSuccess ? OPO_OK : fail(OPO_ERR_HIP, "autotune launch");
            return stretch();
        };
        int rc = once();   // warm-up
        for (int r = 0; r < reps && !rc; ++r) {
            hipError_t e = hipEventRecord(ev[0], stream);
            if (e == hipSuccess) rc = once();
            if (!rc && e == hipSuccess) e = hipEventRecord(ev[1], stream);
            if (!rc && e == hipSuccess) e = hipEventSynchronize(ev[1]);
            float t = 0.f;
            if (!rc && e == hipSuccess) e = hipEventElapsedTime(&t, ev[0], ev[1]);
            if (!rc && e != hipSuccess) rc = fail(OPO_ERR_HIP, "autotune timing: %s", hipGetErrorString(e));
            if (!rc) ms = t < ms ? t : ms;
        }
        if (ge) (void)hipGraphExecDestroy(ge);
        return rc;
    }

    // The measurement of autotune(); scratch is owned by the caller.
    int autotune_run(int reps, std::string &report, C *save[3], hipEvent_t ev[2], bool &saved) {
        const bool fp64 = sizeof(T) == 8;
        const size_t fb = N * sizeof(C);
        for (int f = 0; f < 3; ++f)
            if (hipMalloc(&save[f], fb) != hipSuccess)
                return fail(OPO_ERR_NOMEM, "autotune: no room to save the fields");
        for (int f = 0; f < 3; ++f) CK(hipMemcpyAsync(save[f], A[f], fb, hipMemcpyDeviceToDevice, stream));
        saved = true;
        CK(hipEventCreate(&ev[0]));
        CK(hipEventCreate(&ev[1]));
        // Candidates are timed in context: a stretch of the slice schedule
        // ([t pass][xy pass] per slice), so each pass runs after the other
        // as it does in a round trip.  Coordinate descent: the t pass with
        // the xy pass fixed, then the xy pass, then the t pass again.
        const Rk4Args<T> rk = rk_args(2, 0.0, p.dz, false);
        const int L = p.n_slices < 1 ? 1 : (p.n_slices < 16 ? p.n_slices : 16);
        auto stretch = [&]() {
            int r = OPO_OK;
            for (int k = 0; k < L && !r; ++k) {
                r = q_nl_time(A[0], rk, tabT[0], (int)(tabT[1] - tabT[0]), 0, OPO_STAGE_TIME_OP, nullptr);
                if (!r) r = q_diffraction(nullptr);
            }
            return r;
        };
        char buf[160];
        const int order[3] = {0, 1, 0};
        for (int round = 0; round < 3; ++round) {
            const int axis = order[round];
            if (axis == 1 && g.nx != g.ny) continue;   // NX != NY planes have one launch shape
            const int n = axis == 0 ? g.nt : g.nx;
            int best = tune_get(fp64, axis, n);
            double best_t = 1e30;
            report += axis == 0 ? "t pass NT=" : "xy pass NX=";
            report += std::to_string(n) + " (us per slice):";
            for (int v = 0; v < tune_variant_count(axis); ++v) {
                if (!tune_variant_valid(fp64, axis, n, v)) continue;
                if (axis == 0 && (g.nx == 1 || g.nx % tune_tile_px(fp64, n, v) != 0)) continue;
                tune_set(fp64, axis, n, v);
                float ms = 1e30f;
                const int rc = time_stretch(stretch, reps, ev, ms);
                if (rc) return rc;
                const double us = 1e3 * ms / (double)L;
                const int pc = tune_per_cu(fp64, axis, n, v);
                char pers[32] = "";
                if (pc) std::snprintf(pers, sizeof(pers), ", persistent %d/CU", pc);
                else if (axis == 1 && xy_group(n, v) > 1)
                    std::snprintf(pers, sizeof(pers), ", %d planes", xy_group(n, v));
                std::snprintf(buf, sizeof(buf), " v%d(tile %d, block %d%s) %.2f;", v,
                              axis == 0 ? tune_tile_px(fp64, n, v) : n, tune_block(fp64, axis, n, v), pers, us);
                report += buf;
                if (us < best_t) {
                    best_t = us;
                    best = v;
                }
            }
            tune_set(fp64, axis, n, best);
            report += " -> v" + std::to_string(best) + "\n";
        }
        return OPO_OK;
    }

    void fill_desc(void *out, double *series_out, int cap, int nchunk) const override {
        SolveDesc<T> d{};
        d.api = Api;
        for (int f = 0; f < 3; ++f) d.f[f] = A[f];
        d.twx = tw_x;
        d.twt = tw_t;
        d.tabT = tabT[0];
        d.tabG = tabG;
        d.qx = qx[0];
        d.qy = qy[0];
        d.qfsx = qfsx[0];
        d.qfsy = qfsy[0];
        d.nx = g.nx;
        d.ny = g.ny;
        d.nt = g.nt;
        d.gdd_on = p.gdd_on;
        d.eom = tabE;
        const Rk4Args<T> r = rk_args(1, 0.0, 0.0, true);
        for (int f = 0; f < 3; ++f) {
            d.kc1[f] = r.kc1[f];
            d.kc3[f] = r.kc3[f];
            d.kh6[f] = r.kh6[f];
        }
        d.ms_re = r.ms_re;
        d.ms_im = r.ms_im;
        d.ms_lo_re = r.ms_lo_re;
        d.ms_lo_im = r.ms_lo_im;
        d.mi_re = r.mi_re;
        d.mi_im = r.mi_im;
        d.mi_lo_re = r.mi_lo_re;
        d.mi_lo_im = r.mi_lo_im;
        for (int f = 0; f < 3; ++f) d.pref[f] = p.power_prefactor[f];
        d.series = series_out;
        d.series_cap = cap;
        d.nchunk = nchunk;
        std::memcpy(out, &d, sizeof(d));
    }

    int power_time(int f, int x, int y, double pref, double *out) override {
        int rc = select();
        if (rc) return rc;
        const size_t plane = (size_t)g.nx * g.ny;
        CK(launch_power_pixel<T>(stream, A[f] + (size_t)y * g.nx + x, plane, g.nt, pref, pw3 + 3));
        CK(hipMemcpyAsync(out, pw3 + 3, sizeof(double), hipMemcpyDeviceToHost, stream));
        CK(hipStreamSynchronize(stream));
        return OPO_OK;
    }

    int power(double out[3]) override {
        int rc = select();
        if (rc) return rc;
        CK(launch_power_partial<T>(stream, A[0], A[1], A[2], N, nblk, partial, nullptr));
        CK(launch_power_final(stream, partial, nblk, p.power_prefactor[0], p.power_prefactor[1],
                              p.power_prefactor[2], pw3, nullptr, nullptr, 0, 1));
        CK(hipMemcpyAsync(out, pw3, 3 * sizeof(double), hipMemcpyDeviceToHost, stream));
        CK(hipStreamSynchronize(stream));
        return OPO_OK;
    }
};

// Check a grid and its parameters; create a solve on a device (its tables
// built, its fields zero).
int validate(const opo_grid *g, const opo_params *p);
int make_solve(const opo_grid *g, const opo_params *p, int device, std::unique_ptr<SolveImpl> &out);

}  // namespace opo
