// opo_capi.cpp -- C ABI of the split-step path (include/opo3d.h).
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

static thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
void clear_error() { g_err.clear(); }

static const double kTwoPi = 6.283185307179586476925286766559;

void cis(double phi, double &c, double &s) {
    double r = std::fmod(phi, kTwoPi);
    c = std::cos(r);
    s = std::sin(r);
}

void omega_rule(int nt, double W, double *w) {
    for (int i = 0; i < nt; ++i) w[i] = 0.0;
    if (nt < 2) return;  // the reference leaves w[0] = 0 for NT = 1
    std::vector<double> F(nt);
    for (int i = 0; i < nt; ++i) F[i] = -0.5 * nt / W + i * ((double)nt / W) / (nt - 1);
    const int c = nt / 2;
    for (int i = 0; i < c; ++i) {
        w[i + c] = F[i];
        w[i] = F[i + c];
    }
    for (int i = 0; i < nt; ++i) w[i] *= kTwoPi;
}

#define CK(call)                                                                           \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(OPO_ERR_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_),       \
                        __FILE__, __LINE__);                                               \
    } while (0)

// exp(-2 pi i j/n) from long double on exact fractions.
static void unit_root_neg(long j, long n, double &c, double &s) {
    const long double pi = 3.141592653589793238462643383279502884L;
    long double a = -2.0L * pi * (long double)(j % n) / (long double)n;
    c = (double)cosl(a);
    s = (double)sinl(a);
}

// Signed frequency index of FFT-order bin j: the reference builds centred
// tables and fftShift2D's them (Solver3D.cuh:288-296, EFields3D.cuh:141-171).
static double signed_freq(int j, int n) {
    if (n < 2) return 0.0;
    int c = n / 2;
    return (double)(((j + c) % n) - c);
}

static bool grid_supported(const opo_grid &g, bool fp64) {
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

static int validate(const opo_grid *g, const opo_params *p) {
    if (!g || !p) return fail(OPO_ERR_ARG, "null grid or params");
    if (g->nx < 1 || g->ny < 1 || g->nt < 1) return fail(OPO_ERR_ARG, "grid dimensions must be >= 1");
    if (p->precision != OPO_FP32 && p->precision != OPO_FP64)
        return fail(OPO_ERR_ARG, "precision must be OPO_FP32 or OPO_FP64");
    if (!grid_supported(*g, p->precision == OPO_FP64))
        return fail(OPO_ERR_UNSUPPORTED, "grid %dx%dx%d (%s) is not supported", g->nx, g->ny, g->nt,
                    p->precision == OPO_FP64 ? "fp64" : "fp32");
    if (!(p->dz > 0) || p->n_slices < 0) return fail(OPO_ERR_ARG, "dz must be > 0 and n_slices >= 0");
    if (!p->omega && g->nt > 1 && !(p->t_window > 0))
        return fail(OPO_ERR_ARG, "t_window must be > 0 when omega is not given");
    for (int f = 0; f < 3; ++f)
        if (!(p->kdiff[f] > 0)) return fail(OPO_ERR_ARG, "kdiff must be > 0");
    if ((g->nx > 1 && !(p->dx > 0)) || (g->ny > 1 && !(p->dy > 0)))
        return fail(OPO_ERR_ARG, "dx, dy must be > 0");
    if (p->eom_on && g->nt > 1 && !(p->t_window > 0))
        return fail(OPO_ERR_ARG, "the modulator needs t_window > 0 (its time grid)");
    return OPO_OK;
}

static int make_solve(const opo_grid *g, const opo_params *p, int device,
                      std::unique_ptr<SolveImpl> &out) {
    int rc = validate(g, p);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(OPO_ERR_NOGPU, "no HIP device");
    if (device < 0) {
        if (hipGetDevice(&device) != hipSuccess) device = 0;
    }
    if (device >= ndev) return fail(OPO_ERR_ARG, "device %d out of range (%d devices)", device, ndev);
    std::unique_ptr<SolveImpl> s;
    if (p->precision == OPO_FP64) s.reset(new SolveT<double>());
    else s.reset(new SolveT<float>());
    s->g = *g;
    s->p = *p;
    s->p.omega = nullptr;
    s->prec = p->precision;
    s->device = device;
    s->omega.assign(g->nt, 0.0);
    if (p->omega) for (int i = 0; i < g->nt; ++i) s->omega[i] = p->omega[i];
    else omega_rule(g->nt, p->t_window, s->omega.data());
    rc = s->init();
    if (rc) return rc;
    out = std::move(s);
    return OPO_OK;
}

// ------------------------------------------------ no-allocation scratch
// One fixed block of device memory per device for the entry point that
// allocates nothing (opo_run_round_trips_noalloc): the table block, the
// power partials and counters, and the per-trip power series of the call.
// Each device holds its own copy of the array at its own address, found
// once per device with hipGetSymbolAddress; host code never uses the
// array's name as an address.
static __device__ __align__(256) unsigned char scratch[OPO_NOALLOC_SCRATCH_BYTES];

static std::mutex g_noalloc_mu;              // one no-allocation call at a time per process
static void *g_scratch_ptr[64];              // device address of scratch per device
alignas(64) static unsigned char g_stage[OPO_NOALLOC_SCRATCH_BYTES];  // host image of the tables

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Bytes of scratch a call needs: tables, partials + pw3, flags, series.
static size_t noalloc_bytes(const opo_grid &g, int prec, int n) {
    const size_t cb = prec == OPO_FP64 ? 16 : 8;
    const size_t N = (size_t)g.nx * g.ny * g.nt;
    const int nblk = power_blocks(N);
    return align256(SolveT<float>::table_entries(g) * cb) + align256(sizeof(double) * (3 * (size_t)nblk + 4)) +
           align256(4 * sizeof(int)) + align256(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
}

template <typename T>
static int noalloc_run(const opo_grid &g, const opo_params &p, int device, unsigned char *base, const void *api,
                       void *ap, void *as, void *ai, int n, double *power) {
    using C = cplx<T>;
    SolveT<T> s;
    s.g = g;
    s.p = p;
    s.p.omega = nullptr;
    s.prec = p.precision;
    s.device = device;
    // host image: omega (NT doubles) and the time grid first, then the tables
    double *w = reinterpret_cast<double *>(g_stage);
    double *tg = w + g.nt;
    C *h = reinterpret_cast<C *>(g_stage + align256(2 * sizeof(double) * (size_t)g.nt));
    const size_t ne = SolveT<T>::table_entries(g);
    if ((unsigned char *)(h + ne) > g_stage + sizeof(g_stage))
        return fail(OPO_ERR_UNSUPPORTED, "grid %dx%dx%d: tables exceed the %zu-byte staging area", g.nx, g.ny,
                    g.nt, sizeof(g_stage));
    if (p.omega) for (int i = 0; i < g.nt; ++i) w[i] = p.omega[i];
    else omega_rule(g.nt, p.t_window, w);
    SolveT<T>::fill_tables(g, s.p, w, h, tg);
    // device layout inside scratch: tables | partials + pw3 | flags | series
    const size_t N = (size_t)g.nx * g.ny * g.nt;
    const int nblk = power_blocks(N);
    unsigned char *q = base;
    C *block = reinterpret_cast<C *>(q);
    q += align256(ne * sizeof(C));
    double *dbuf = reinterpret_cast<double *>(q);
    q += align256(sizeof(double) * (3 * (size_t)nblk + 4));
    int *flags = reinterpret_cast<int *>(q);
    q += align256(4 * sizeof(int));
    double *series = reinterpret_cast<double *>(q);
    s.init_borrowed((C *)api, (C *)ap, (C *)as, (C *)ai, block, dbuf, flags, series, n);
    CK(hipMemcpy(block, h, ne * sizeof(C), hipMemcpyHostToDevice));
    int rc = s.enqueue_trips(n);   // clears the status word and trip counter first
    if (rc) {
        (void)hipStreamSynchronize(nullptr);
        return rc;
    }
    int hflags[2] = {0, 0};
    CK(hipMemcpyAsync(hflags, flags, sizeof(hflags), hipMemcpyDeviceToHost, nullptr));
    CK(hipStreamSynchronize(nullptr));
    const int done = hflags[0] ? hflags[0] : n;
    if (power) CK(hipMemcpy(power, series, sizeof(double) * 3 * (size_t)done, hipMemcpyDeviceToHost));
    if (hflags[0]) return fail(OPO_ERR_NAN, "non-finite signal power at round trip %d", done - 1);
    return OPO_OK;
}

}  // namespace opo

using namespace opo;

struct opo_solve {
    std::unique_ptr<SolveImpl> s;
};

// ----------------------------------------------------------------- batch
// A batch keeps many solves resident and advances all of them one round trip
// per opo_batch_step.  Each pass is one launch per size class over a work
// list of (solve, tile) items; the trip is captured once as a HIP graph and
// re-captured only when the set of resident solves changes.
struct BatchEntry {
    uint64_t ticket = 0;
    opo_grid g{};
    opo_params p{};
    std::vector<unsigned char> api, as, ai;  // host copies while queued
    std::unique_ptr<SolveImpl> s;
    double *d_series = nullptr;
    int n_trips = 0, done_trips = 0;
    size_t footprint = 0;
    bool finished = false;
    int error = OPO_OK;
    std::string error_msg;
};

struct opo_batch {
    int device = 0;
    size_t budget = 0, used = 0;
    int prec = OPO_FP32;
    uint64_t next_ticket = 1;
    std::deque<BatchEntry *> queue;
    std::vector<BatchEntry *> resident;
    std::map<uint64_t, std::unique_ptr<BatchEntry>> all;
    std::mutex mu;
    hipStream_t stream = nullptr;
    // Concurrent size classes (opo_batch_concurrent_classes): within each
    // stage of the trip graph the launches of the different size classes
    // are captured as parallel branches on side streams (the classes of a
    // stage touch disjoint solves), joined before the next stage.
    bool concurrent = false;
    static constexpr int kSide = 4;
    hipStream_t side[kSide] = {};
    std::vector<hipEvent_t> sync_ev;   // fork/join events, distinct per use in one capture
    size_t sync_used = 0;
    int next_sync_event(hipEvent_t &e) {
        if (sync_used >= sync_ev.size()) {
            hipEvent_t x;
            CK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
            sync_ev.push_back(x);
        }
        e = sync_ev[sync_used++];
        return OPO_OK;
    }
    // device-side descriptors, rebuilt when the resident set changes
    void *d_desc = nullptr;
    int *d_status = nullptr;
    TripState *d_ts = nullptr;
    WorkItem *d_work = nullptr;
    double *d_partial = nullptr;
    struct Range {
        int cls, off, n;
        int v;   // launch variant the class's work list was built for
        // Work items of slice k: the class's solves are listed by slice
        // count, most first, so slice k runs the prefix nk[k] of the list
        // (the solves with more than k slices).
        std::vector<int> nk;
    };
    std::vector<Range> trange, xyrange, fsrange;
    Range prange{0, 0, 0, 0, {}};
    int n_slices = -1, max_chunk = 1;
    hipGraphExec_t graph = nullptr;
    bool dirty = true;
    std::vector<int> built_variants;  // tuned launch variants the graph uses

    // Variant of the fused pass over NT class nt: the tuned one when its
    // tile divides the NX of every resident solve of the class, else
    // variant 0 (its tile divides every supported NX).
    int class_time_variant(int cls) const {
        if (cls < 0) return kBatchPlaneVariant;
        const bool fp64 = prec == OPO_FP64;
        const int v = tune_get(fp64, 0, cls);
        for (const BatchEntry *e : resident)
            if (tclass(e) == cls && time_variant_for(fp64, cls, e->g.nx) != v) return 0;
        return v;
    }
    // Size class of a solve's t pass: NT, or -NT for plane-wave solves
    // (NX = NY = 1: one pixel per workgroup, kBatchPlaneVariant).
    static int tclass(const BatchEntry *e) { return e->g.nx == 1 ? -e->g.nt : e->g.nt; }
    // Whether a solve takes the xy pass of class nx (plane-wave solves skip it).
    static bool in_xy(const BatchEntry *e, int nx) { return e->g.nx == nx && !e->p.no_diffraction; }
    std::vector<int> current_variants() const {
        std::vector<int> v;
        const bool fp64 = prec == OPO_FP64;
        for (const Range &r : trange) v.push_back(class_time_variant(r.cls));
        for (const Range &r : xyrange) v.push_back(tune_get(fp64, 1, r.cls));
        return v;
    }
    std::vector<TripState> h_ts;
    std::vector<int> h_status;

    // Kernel timing inside the replayed trip graph (opo_batch_profile):
    // events recorded around every launch at capture, read after each step.
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> ev_entry;    // entry index of the pair (2k, 2k + 1)
    size_t ev_used = 0;
    std::vector<opo_stage_time> prof;
    std::vector<double> prof_alg;   // algorithmic bytes summed over the timed launches of an entry
    std::vector<double> ev_alg;     // algorithmic bytes of the launch of pair k
    int prof_entry(int stage, int size, double alg) {
        for (size_t k = 0; k < prof.size(); ++k)
            if (prof[k].stage == stage && prof[k].size == size) {
                prof[k].alg_bytes = alg;
                return (int)k;
            }
        prof.push_back(opo_stage_time{stage, size, 0, 0.0, alg});
        prof_alg.push_back(0.0);
        return (int)prof.size() - 1;
    }
    void prof_reset() {
        for (opo_stage_time &x : prof) x.launches = 0, x.ms = 0.0;
        prof_alg.assign(prof.size(), 0.0);
    }
    int mark(int stage, int size, double alg, hipStream_t st) {
        if (!profiling) return OPO_OK;
        if (ev_used + 1 >= ev_pool.size()) {
            for (int k = 0; k < 64; ++k) {
                hipEvent_t e;
                CK(hipEventCreate(&e));
                ev_pool.push_back(e);
            }
        }
        if (ev_used % 2 == 0) {
            ev_entry.push_back(prof_entry(stage, size, alg));
            ev_alg.push_back(alg);
        }
        // recorded as an event node of the captured trip graph
        CK(hipEventRecordWithFlags(ev_pool[ev_used++], st, hipEventRecordExternal));
        return OPO_OK;
    }
    // After a step: add the elapsed time of every bracketed launch.
    // A runtime that cannot time event nodes of a graph leaves the sums at
    // zero and the reason in prof_error; the step itself is unaffected.
    std::string prof_error;
    int collect() {
        if (!profiling || !prof_error.empty()) return OPO_OK;
        for (size_t k = 0; k + 1 < ev_used; k += 2) {
            float t = 0.f;
            const hipError_t e = hipEventElapsedTime(&t, ev_pool[k], ev_pool[k + 1]);
            if (e != hipSuccess) {
                prof_error = std::string("timing the events of the trip graph: ") + hipGetErrorString(e);
                (void)hipGetLastError();
                prof_reset();
                return OPO_OK;
            }
            prof_alg[ev_entry[k / 2]] += ev_alg[k / 2];
            opo_stage_time &x = prof[ev_entry[k / 2]];
            x.ms += t;
            x.launches += 1;
        }
        return OPO_OK;
    }
    // Algorithmic bytes of one launch over the resident solves selected.
    template <typename Pred> double class_bytes(int fields, int rw, Pred pick) const {
        const double cb = prec == OPO_FP64 ? 16.0 : 8.0;
        double cells = 0.0;
        for (const BatchEntry *e : resident)
            if (pick(e)) cells += (double)e->g.nx * e->g.ny * e->g.nt;
        return cells * fields * cb * rw;
    }

    ~opo_batch() {
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        if (graph) (void)hipGraphExecDestroy(graph);
        for (auto e : ev_pool) (void)hipEventDestroy(e);
        for (auto e : sync_ev) (void)hipEventDestroy(e);
        for (auto &x : side)
            if (x) (void)hipStreamDestroy(x);
        for (auto &kv : all)
            if (kv.second->d_series) (void)hipFree(kv.second->d_series);
        all.clear();
        (void)hipFree(d_desc);
        (void)hipFree(d_status);
        (void)hipFree(d_ts);
        (void)hipFree(d_work);
        (void)hipFree(d_partial);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Power partial blocks of a solve: as many as the solve uses alone, so
    // its per-trip powers are bitwise the same in the batch.
    static int chunks_for(const opo_grid &g) { return power_blocks((size_t)g.nx * g.ny * g.nt); }

    // Device buffers of the resident set, grown (never shrunk) as it needs:
    // an admission or a finished solve reuses them without freeing.
    size_t cap_desc = 0, cap_status = 0, cap_ts = 0, cap_work = 0, cap_partial = 0;
    template <typename U> int ensure_buf(U *&ptr, size_t &cap, size_t bytes) {
        if (bytes <= cap) return OPO_OK;
        (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
        const size_t want = std::max(bytes, (size_t)4096);
        if (hipMalloc((void **)&ptr, want) != hipSuccess)
            return fail(OPO_ERR_NOMEM, "batch descriptor allocation of %zu bytes failed", want);
        cap = want;
        return OPO_OK;
    }
    // Rebuild bookkeeping: count, graphs updated in place, host seconds.
    int n_rebuilds = 0, n_updated = 0;
    double rebuild_s = 0.0;

    int rebuild() {
        const auto t0 = std::chrono::steady_clock::now();
        const int rc = rebuild_impl();
        rebuild_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        n_rebuilds += 1;
        return rc;
    }

    int rebuild_impl() {
        const int n = (int)resident.size();
        const size_t dsz = prec == OPO_FP64 ? sizeof(SolveDesc<double>) : sizeof(SolveDesc<float>);
        std::vector<unsigned char> hdesc(dsz * n);
        max_chunk = 1;
        for (int i = 0; i < n; ++i) max_chunk = std::max(max_chunk, chunks_for(resident[i]->g));
        for (int i = 0; i < n; ++i) {
            BatchEntry *e = resident[i];
            e->s->fill_desc(hdesc.data() + dsz * i, e->d_series, e->n_trips, chunks_for(e->g));
        }
        // work lists grouped by size class
        std::vector<WorkItem> work;
        trange.clear();
        xyrange.clear();
        fsrange.clear();
        const bool fp64 = prec == OPO_FP64;
        std::map<int, std::vector<int>> by_nt, by_nx;
        n_slices = 0;
        for (int i = 0; i < n; ++i) {
            by_nt[tclass(resident[i])].push_back(i);
            if (!resident[i]->p.no_diffraction) by_nx[resident[i]->g.nx].push_back(i);
            n_slices = std::max(n_slices, resident[i]->p.n_slices);
        }
        // within a class, solves with more slices first (stable)
        auto by_slices = [&](std::vector<int> &v) {
            std::stable_sort(v.begin(), v.end(), [&](int a, int b) {
                return resident[a]->p.n_slices > resident[b]->p.n_slices;
            });
        };
        // per-slice prefix lengths of a class's work list
        auto prefixes = [&](Range &r, const std::vector<std::pair<int, int>> &items) {
            r.nk.assign(n_slices, 0);
            for (const auto &si : items)   // (slices, work items) per solve
                for (int k = 0; k < si.first; ++k) r.nk[k] += si.second;
        };
        // Each class's variant is read once here: the work list and the
        // launches captured below both follow it, whatever the process-wide
        // choice becomes meanwhile (a change makes the next step rebuild).
        for (auto &kv : by_nt) {
            by_slices(kv.second);
            Range r{kv.first, (int)work.size(), 0, class_time_variant(kv.first), {}};
            const int tx = kv.first < 0 ? 1 : tune_tile_px(fp64, kv.first, r.v);
            std::vector<std::pair<int, int>> items;
            for (int i : kv.second) {
                const opo_grid &g = resident[i]->g;
                if (tx == 0 || g.nx % tx != 0)
                    return fail(OPO_ERR_UNSUPPORTED, "tuned tile of %d pixels does not divide NX=%d (NT=%d)",
                                tx, g.nx, kv.first);
                const int tiles = (g.nx / tx) * g.ny;
                for (int t = 0; t < tiles; ++t) work.push_back(WorkItem{i, t});
                items.push_back({resident[i]->p.n_slices, tiles});
            }
            r.n = (int)work.size() - r.off;
            prefixes(r, items);
            trange.push_back(r);
        }
        for (auto &kv : by_nx) {
            by_slices(kv.second);
            // one item per group of consecutive planes of a field (xy_group)
            const int xv = tune_get(fp64, 1, kv.first);
            const int pg = xy_group(kv.first, xv);
            Range r{kv.first, (int)work.size(), 0, xv, {}};
            std::vector<std::pair<int, int>> items;
            for (int i : kv.second) {
                if (resident[i]->g.nt % pg != 0)
                    return fail(OPO_ERR_UNSUPPORTED, "NT=%d is not a multiple of the %d-plane xy group (NX=%d)",
                                resident[i]->g.nt, pg, kv.first);
                for (int it = 0; it < 3 * resident[i]->g.nt / pg; ++it) work.push_back(WorkItem{i, it});
                items.push_back({resident[i]->p.n_slices, 3 * resident[i]->g.nt / pg});
            }
            r.n = (int)work.size() - r.off;
            prefixes(r, items);
            xyrange.push_back(r);
            Range q{kv.first, (int)work.size(), 0, xv, {}};
            for (int i : kv.second)
                if (resident[i]->p.fs_on)   // in the NX class: not a plane-wave solve
                    for (int it = 0; it < 2 * resident[i]->g.nt / pg; ++it) work.push_back(WorkItem{i, it});
            q.n = (int)work.size() - q.off;
            if (q.n) fsrange.push_back(q);
        }
        prange = Range{0, (int)work.size(), 0, 0, {}};
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < chunks_for(resident[i]->g); ++c) work.push_back(WorkItem{i, c});
        prange.n = (int)work.size() - prange.off;
        // Grow the buffers where needed.  A buffer that moves changes the
        // captured launch arguments, which the update below carries over.
        int rc0;
        if ((rc0 = ensure_buf(d_desc, cap_desc, std::max<size_t>(hdesc.size(), 1))) ||
            (rc0 = ensure_buf(d_status, cap_status, sizeof(int) * std::max(n, 1))) ||
            (rc0 = ensure_buf(d_ts, cap_ts, sizeof(TripState) * std::max(n, 1))) ||
            (rc0 = ensure_buf(d_work, cap_work, sizeof(WorkItem) * std::max<size_t>(work.size(), 1))) ||
            (rc0 = ensure_buf(d_partial, cap_partial, sizeof(double) * 3 * (size_t)std::max(n, 1) * max_chunk)))
            return rc0;
        CK(hipMemcpyAsync(d_desc, hdesc.data(), hdesc.size(), hipMemcpyHostToDevice, stream));
        CK(hipMemcpyAsync(d_work, work.data(), sizeof(WorkItem) * work.size(), hipMemcpyHostToDevice, stream));
        CK(hipMemsetAsync(d_status, 0, sizeof(int) * std::max(n, 1), stream));
        h_ts.assign(n, TripState{0, 0});
        h_status.assign(n, 0);
        // capture one round trip of every resident solve
        ev_used = 0;
        ev_entry.clear();
        ev_alg.clear();
        sync_used = 0;
        hipGraph_t gr = nullptr;
        CK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
        int rc = fp64 ? enqueue_trip<double>() : enqueue_trip<float>();
        hipError_t e = hipStreamEndCapture(stream, &gr);
        if (rc) {
            if (gr) (void)hipGraphDestroy(gr);
            return rc;
        }
        if (e != hipSuccess) return fail(OPO_ERR_HIP, "batch capture: %s", hipGetErrorString(e));
        // Update the instantiated graph in place when the new trip has the
        // same shape (same launches in the same order: the size classes did
        // not change, only their work lists); instantiate it otherwise.
        if (graph) {
            hipGraphNode_t err_node = nullptr;
            hipGraphExecUpdateResult res = hipGraphExecUpdateError;
            const hipError_t eu = hipGraphExecUpdate(graph, gr, &err_node, &res);
            if (eu == hipSuccess && res == hipGraphExecUpdateSuccess) {
                n_updated += 1;
            } else {
                (void)hipGetLastError();   // a refused update is not an error of the batch
                (void)hipGraphExecDestroy(graph);
                graph = nullptr;
            }
        }
        if (!graph) {
            e = hipGraphInstantiate(&graph, gr, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(gr);
                return fail(OPO_ERR_HIP, "batch graph: %s", hipGetErrorString(e));
            }
        }
        (void)hipGraphDestroy(gr);
        built_variants.clear();
        for (const Range &r : trange) built_variants.push_back(r.v);
        for (const Range &r : xyrange) built_variants.push_back(r.v);
        dirty = false;
        return OPO_OK;
    }

    // Launch filter of enqueue_trip: with only_stage >= 0, only the launches
    // of that (stage, size class) are enqueued (the stage timing graphs).
    int only_stage = -1, only_size = -1, n_enqueued = 0;
    bool wanted(int stage, int size) const {
        return only_stage < 0 || (stage == only_stage && size == only_size);
    }

    // Enqueue one round trip of every resident solve: per slice the fused
    // t pass of each NT class then the xy pass of each NX class; then the
    // trip-end pass, free space and the power sums.  Each launch goes through
    // launch(): the filter above, and the profiling events around it.
    template <typename T> int enqueue_trip() {
        const SolveDesc<T> *sd = static_cast<const SolveDesc<T> *>(d_desc);
        n_enqueued = 0;
        // one launch of a stage on stream st: the filter, and the profiling
        // events around it
        auto launch = [&](int stage, int size, double alg, hipStream_t st, auto &&fn) -> int {
            if (!wanted(stage, size)) return OPO_OK;
            int rc = mark(stage, size, alg, st);
            if (rc) return rc;
            CK(fn(st));
            n_enqueued += 1;
            return mark(stage, size, alg, st);
        };
        // the launches of one stage, class c on stream c (side streams for
        // c >= 1 when the classes run concurrently), joined at the end
        auto stage_of = [&](int nclass, auto &&one) -> int {
            const bool par = concurrent && only_stage < 0 && nclass > 1;
            const int nside = par ? std::min(nclass - 1, (int)kSide) : 0;
            hipEvent_t fork = nullptr;
            int rc;
            if (par) {
                if ((rc = next_sync_event(fork))) return rc;
                CK(hipEventRecord(fork, stream));
                for (int k = 0; k < nside; ++k) CK(hipStreamWaitEvent(side[k], fork, 0));
            }
            for (int c = 0; c < nclass; ++c) {
                const hipStream_t st = (par && c > 0) ? side[(c - 1) % kSide] : stream;
                if ((rc = one(c, st))) return rc;
            }
            for (int k = 0; k < nside; ++k) {
                hipEvent_t join = nullptr;
                if ((rc = next_sync_event(join))) return rc;
                CK(hipEventRecord(join, side[k]));
                CK(hipStreamWaitEvent(stream, join, 0));
            }
            return OPO_OK;
        };
        std::vector<double> tb, fb;
        for (const Range &r : trange) tb.push_back(class_bytes(3, 2, [&](const BatchEntry *e) { return tclass(e) == r.cls; }));
        // bytes of slice k's launch: the class's solves with more than k slices
        auto slice_bytes = [&](bool t_pass, int cls, int k) {
            return class_bytes(3, 2, [&](const BatchEntry *e) {
                return (t_pass ? tclass(e) == cls : in_xy(e, cls)) && e->p.n_slices > k;
            });
        };
        for (const Range &r : fsrange)
            fb.push_back(class_bytes(2, 2, [&](const BatchEntry *e) { return in_xy(e, r.cls) && e->p.fs_on; }));
        const double pb = class_bytes(3, 1, [](const BatchEntry *) { return true; });
        int rc;
        for (int k = 0; k < n_slices; ++k) {
            const int mode = k == 0 ? BATCH_FIRST : BATCH_MID;
            if ((rc = stage_of((int)trange.size(), [&](int c, hipStream_t st) {
                     const Range &r = trange[c];
                     const int nw = r.nk[k];
                     if (nw == 0) return (int)OPO_OK;   // every solve of the class has fewer slices
                     return launch(OPO_STAGE_TIME_OP, r.cls, slice_bytes(true, r.cls, k), st, [&](hipStream_t s2) {
                         return launch_batch_time<T>(s2, std::abs(r.cls), nw, sd, d_status, d_work + r.off, mode, r.v);
                     });
                 })))
                return rc;
            if ((rc = stage_of((int)xyrange.size(), [&](int c, hipStream_t st) {
                     const Range &r = xyrange[c];
                     const int nw = r.nk[k];
                     if (nw == 0) return (int)OPO_OK;
                     return launch(OPO_STAGE_XY_OP, r.cls, slice_bytes(false, r.cls, k), st, [&](hipStream_t s2) {
                         return launch_batch_xy<T>(s2, r.cls, nw, sd, d_ts, d_status, d_work + r.off, 0, r.v);
                     });
                 })))
                return rc;
        }
        if ((rc = stage_of((int)trange.size(), [&](int c, hipStream_t st) {
                 const Range &r = trange[c];
                 return launch(OPO_STAGE_GDD, r.cls, tb[c], st, [&](hipStream_t s2) {
                     return launch_batch_time<T>(s2, std::abs(r.cls), r.n, sd, d_status, d_work + r.off, BATCH_LAST,
                                                 r.v);
                 });
             })))
            return rc;
        if ((rc = stage_of((int)fsrange.size(), [&](int c, hipStream_t st) {
                 const Range &r = fsrange[c];
                 return launch(OPO_STAGE_FREE_SPACE, r.cls, fb[c], st, [&](hipStream_t s2) {
                     return launch_batch_xy<T>(s2, r.cls, r.n, sd, d_ts, d_status, d_work + r.off, 1, r.v);
                 });
             })))
            return rc;
        return launch(OPO_STAGE_POWER, 0, pb, stream, [&](hipStream_t s2) {
            return launch_batch_power<T>(s2, prange.n, (int)resident.size(), sd, d_ts, d_status,
                                         d_work + prange.off, d_partial, max_chunk);
        });
    }

    // Time each (stage, size class) of the trip by itself: a graph holding
    // that class's launches of one round trip, in trip order and with the
    // trip's arguments, replayed reps times between two HIP events on the
    // batch stream.  The solves' fields go through these launches too.
    // One (stage, size class) of the trip by itself: its launches of one
    // round trip captured as a graph and replayed reps times (after one
    // untimed replay) between two events on the batch stream.
    int time_key(int stage, int size, int reps, float &ms, int &launches) {
        ms = 0.f;
        launches = 0;
        const bool keep_prof = profiling;
        profiling = false;
        only_stage = stage;
        only_size = size;
        hipGraph_t gr = nullptr;
        hipGraphExec_t ge = nullptr;
        hipEvent_t ev[2] = {nullptr, nullptr};
        int rc = OPO_OK;
        if (hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed) != hipSuccess) {
            rc = fail(OPO_ERR_HIP, "stage timing: capture failed");
        } else {
            rc = prec == OPO_FP64 ? enqueue_trip<double>() : enqueue_trip<float>();
            launches = n_enqueued;
            const hipError_t e = hipStreamEndCapture(stream, &gr);
            if (!rc && e != hipSuccess) rc = fail(OPO_ERR_HIP, "stage timing capture: %s", hipGetErrorString(e));
        }
        only_stage = only_size = -1;
        profiling = keep_prof;
        if (!rc && hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0) != hipSuccess)
            rc = fail(OPO_ERR_HIP, "stage timing: instantiate failed");
        if (gr) (void)hipGraphDestroy(gr);
        if (!rc) {
            hipError_t e2 = hipEventCreate(&ev[0]);
            if (e2 == hipSuccess) e2 = hipEventCreate(&ev[1]);
            if (e2 == hipSuccess) e2 = hipGraphLaunch(ge, stream);   // warm-up
            if (e2 == hipSuccess) e2 = hipEventRecord(ev[0], stream);
            for (int r = 0; r < reps && e2 == hipSuccess; ++r) e2 = hipGraphLaunch(ge, stream);
            if (e2 == hipSuccess) e2 = hipEventRecord(ev[1], stream);
            if (e2 == hipSuccess) e2 = hipEventSynchronize(ev[1]);
            if (e2 == hipSuccess) e2 = hipEventElapsedTime(&ms, ev[0], ev[1]);
            if (e2 != hipSuccess) rc = fail(OPO_ERR_HIP, "stage timing: %s", hipGetErrorString(e2));
        }
        if (ge) (void)hipGraphExecDestroy(ge);
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
        return rc;
    }

    // Time each (stage, size class) of the trip by itself (time_key).
    // The solves' fields go through these launches too.
    int time_stages(int reps, std::vector<opo_stage_time> &out) {
        out.clear();
        // the (stage, size) entries of the trip, in trip order
        std::vector<std::pair<int, int>> keys;
        for (const Range &r : trange) keys.push_back({OPO_STAGE_TIME_OP, r.cls});
        for (const Range &r : xyrange) keys.push_back({OPO_STAGE_XY_OP, r.cls});
        for (const Range &r : trange) keys.push_back({OPO_STAGE_GDD, r.cls});
        for (const Range &r : fsrange) keys.push_back({OPO_STAGE_FREE_SPACE, r.cls});
        keys.push_back({OPO_STAGE_POWER, 0});
        for (const auto &key : keys) {
            float ms = 0.f;
            int launches = 0;
            const int rc = time_key(key.first, key.second, reps, ms, launches);
            if (rc) return rc;
            // per-slice stages: the mean over the trip's launches (solves
            // with fewer slices drop out of the later launches)
            auto per_slice = [&](bool t_pass) {
                double sum = 0.0;
                int nl = 0;
                for (int k = 0; k < n_slices; ++k) {
                    const double b = class_bytes(3, 2, [&](const BatchEntry *e) {
                        return (t_pass ? tclass(e) == key.second : in_xy(e, key.second)) && e->p.n_slices > k;
                    });
                    if (b > 0) sum += b, nl += 1;
                }
                return nl ? sum / nl : 0.0;
            };
            const double alg = key.first == OPO_STAGE_POWER
                                   ? class_bytes(3, 1, [](const BatchEntry *) { return true; })
                               : key.first == OPO_STAGE_XY_OP ? per_slice(false)
                               : key.first == OPO_STAGE_FREE_SPACE
                                   ? class_bytes(2, 2, [&](const BatchEntry *e) { return in_xy(e, key.second) && e->p.fs_on; })
                               : key.first == OPO_STAGE_TIME_OP ? per_slice(true)
                                   : class_bytes(3, 2, [&](const BatchEntry *e) { return tclass(e) == key.second; });
            out.push_back(opo_stage_time{key.first, key.second, launches * reps, (double)ms, alg});
        }
        return OPO_OK;
    }

    // Tune the launch variants in this batch's own trip: for each NT class
    // the t-pass variants whose tile divides the NX of every resident solve
    // of the class, for each NX class the xy-pass variants, each timed as
    // that class's launches of one trip over all its solves (time_key), the
    // fastest kept in the process-wide tables.  The resident fields are
    // saved first and put back after, so the solves continue from where
    // they were; this needs free device memory of the size of those fields.
    int autotune(int reps, std::string &report) {
        const bool fp64 = prec == OPO_FP64;
        int rc;
        if (dirty && (rc = rebuild())) return rc;
        size_t total = 0;
        for (const BatchEntry *e : resident) total += 3 * e->s->field_bytes();
        unsigned char *save = nullptr;
        if (hipMalloc(&save, total) != hipSuccess)
            return fail(OPO_ERR_NOMEM, "batch autotune: no room to save the resident fields (%zu bytes)", total);
        size_t off = 0;
        for (const BatchEntry *e : resident)
            for (int f = 0; f < 3; ++f) {
                const size_t fb = e->s->field_bytes();
                if (hipMemcpyAsync(save + off, e->s->dev_ptr(f), fb, hipMemcpyDeviceToDevice, stream) != hipSuccess) {
                    (void)hipFree(save);
                    return fail(OPO_ERR_HIP, "batch autotune: saving the fields failed");
                }
                off += fb;
            }
        std::vector<int> nts, nxs;
        for (const Range &r : trange)
            if (r.cls > 0) nts.push_back(r.cls);   // the plane-wave classes have one launch shape
        for (const Range &r : xyrange) nxs.push_back(r.cls);
        rc = OPO_OK;
        char buf[160];
        for (int axis = 0; axis < 2 && !rc; ++axis) {
            for (int n : axis == 0 ? nts : nxs) {
                const int stage = axis == 0 ? OPO_STAGE_TIME_OP : OPO_STAGE_XY_OP;
                int best = tune_get(fp64, axis, n);
                double best_us = 1e30;
                report += axis == 0 ? "batch t pass NT=" : "batch xy pass NX=";
                report += std::to_string(n) + " (us per launch):";
                for (int v = 0; v < tune_variant_count(axis) && !rc; ++v) {
                    if (!tune_variant_valid(fp64, axis, n, v)) continue;
                    bool fits = true;
                    if (axis == 0)
                        for (const BatchEntry *e : resident)
                            if (tclass(e) == n && e->g.nx % tune_tile_px(fp64, n, v) != 0) fits = false;
                    if (!fits) continue;
                    tune_set(fp64, axis, n, v);
                    if ((rc = rebuild())) break;
                    float ms = 0.f;
                    int launches = 0;
                    if ((rc = time_key(stage, n, reps, ms, launches))) break;
                    const double us = launches ? 1e3 * ms / ((double)launches * reps) : 1e30;
                    std::snprintf(buf, sizeof(buf), " v%d %.2f;", v, us);
                    report += buf;
                    if (us < best_us) {
                        best_us = us;
                        best = v;
                    }
                }
                tune_set(fp64, axis, n, best);
                report += " -> v" + std::to_string(best) + "\n";
                if (rc) break;
            }
        }
        off = 0;
        for (const BatchEntry *e : resident)
            for (int f = 0; f < 3; ++f) {
                const size_t fb = e->s->field_bytes();
                if (hipMemcpyAsync(e->s->dev_ptr(f), save + off, fb, hipMemcpyDeviceToDevice, stream) != hipSuccess &&
                    !rc)
                    rc = fail(OPO_ERR_HIP, "batch autotune: restoring the fields failed");
                off += fb;
            }
        const hipError_t es = hipStreamSynchronize(stream);
        (void)hipFree(save);
        if (!rc && es != hipSuccess) rc = fail(OPO_ERR_HIP, "batch autotune: %s", hipGetErrorString(es));
        dirty = true;
        return rc;
    }

    // Admit queued solves in arrival order while the head of the queue fits
    // the budget.
    void admit() {
        while (!queue.empty() && used + queue.front()->footprint <= budget) {
            BatchEntry *e = queue.front();
            std::unique_ptr<SolveImpl> s;
            int rc = make_solve(&e->g, &e->p, device, s);
            if (rc == OPO_ERR_NOMEM && !resident.empty()) break;  // wait for memory to free up
            queue.pop_front();
            if (!rc) rc = s->set_fields(e->api.data(), e->as.data(), e->ai.data());
            if (!rc && hipMalloc(&e->d_series, sizeof(double) * 3 * (size_t)e->n_trips) != hipSuccess)
                rc = fail(OPO_ERR_NOMEM, "series allocation failed");
            if (rc) {
                e->finished = true;
                e->error = rc;
                e->error_msg = g_err;
                continue;
            }
            e->api = {};
            e->as = {};
            e->ai = {};
            e->s = std::move(s);
            used += e->footprint;
            resident.push_back(e);
            dirty = true;
        }
    }
};

extern "C" {

const char *opo_last_error(void) { return g_err.c_str(); }
const char *opo_version(void) { return "opo3d-mi355x 0.1 (gfx950)"; }
int opo_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int opo_create(const opo_grid *grid, const opo_params *params, int device, opo_solve **out) {
    if (!out) return fail(OPO_ERR_ARG, "opo_create: null output handle");
    *out = nullptr;
    std::unique_ptr<SolveImpl> s;
    int rc = make_solve(grid, params, device, s);
    if (rc) return rc;
    *out = new opo_solve{std::move(s)};
    return OPO_OK;
}

void opo_destroy(opo_solve *s) { delete s; }

#define SOLVE_OR_FAIL(s) \
    if (!(s) || !(s)->s) return fail(OPO_ERR_ARG, "null solve handle")

int opo_set_fields(opo_solve *s, const void *api, const void *as, const void *ai) {
    SOLVE_OR_FAIL(s);
    return s->s->set_fields(api, as, ai);
}
int opo_get_fields(opo_solve *s, void *ap, void *as, void *ai) {
    SOLVE_OR_FAIL(s);
    return s->s->get_fields(ap, as, ai);
}
void *opo_field_device_ptr(opo_solve *s, int which) {
    if (!s || !s->s) return nullptr;
    return s->s->dev_ptr(which);
}
size_t opo_field_bytes(const opo_solve *s) { return (s && s->s) ? s->s->field_bytes() : 0; }

int opo_step_nonlinear(opo_solve *s) { SOLVE_OR_FAIL(s); return s->s->step(STEP_NONLINEAR); }
int opo_step_dispersion(opo_solve *s) { SOLVE_OR_FAIL(s); return s->s->step(STEP_DISPERSION); }
int opo_step_diffraction(opo_solve *s) { SOLVE_OR_FAIL(s); return s->s->step(STEP_DIFFRACTION); }
int opo_step_free_space(opo_solve *s) { SOLVE_OR_FAIL(s); return s->s->step(STEP_FREE_SPACE); }
int opo_step_gdd(opo_solve *s) { SOLVE_OR_FAIL(s); return s->s->step(STEP_GDD); }
int opo_step_mirror(opo_solve *s) { SOLVE_OR_FAIL(s); return s->s->step(STEP_MIRROR); }
int opo_ssfm(opo_solve *s) { SOLVE_OR_FAIL(s); return s->s->ssfm(); }
int opo_single_pass(opo_solve *s) { SOLVE_OR_FAIL(s); return s->s->ssfm(); }

int opo_run_round_trips(opo_solve *s, int n, double *power) {
    SOLVE_OR_FAIL(s);
    if (n < 0) return fail(OPO_ERR_ARG, "n must be >= 0");
    // A round trip of the reference runs its slice loop at least once
    // (Solver3D.cuh:949, z = 0 <= Lcr) before the feedback.
    if (s->s->p.n_slices < 1) return fail(OPO_ERR_ARG, "round trips need n_slices >= 1");
    if (n == 0) return OPO_OK;
    int rc = s->s->enqueue_trips(n);
    if (rc) return rc;
    return s->s->finish_trips(n, power);
}
int opo_run_round_trips_noalloc(const opo_grid *grid, const opo_params *params, int device, const void *api,
                                void *ap, void *as, void *ai, int n, double *power) {
    int rc = validate(grid, params);
    if (rc) return rc;
    if (!api || !ap || !as || !ai) return fail(OPO_ERR_ARG, "null field pointer");
    if (n < 1) return fail(OPO_ERR_ARG, "n must be >= 1");
    if (params->n_slices < 1) return fail(OPO_ERR_ARG, "round trips need n_slices >= 1");
    const size_t need = noalloc_bytes(*grid, params->precision, n);
    if (need > OPO_NOALLOC_SCRATCH_BYTES)
        return fail(OPO_ERR_UNSUPPORTED,
                    "grid %dx%dx%d with %d trips needs %zu bytes of scratch; the no-allocation entry point has %d",
                    grid->nx, grid->ny, grid->nt, n, need, (int)OPO_NOALLOC_SCRATCH_BYTES);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(OPO_ERR_NOGPU, "no HIP device");
    if (device < 0 && hipGetDevice(&device) != hipSuccess) device = 0;
    if (device >= ndev || device >= 64) return fail(OPO_ERR_ARG, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    CK(hipSetDevice(device));
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        CK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(OPO_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_ptr[device] = ptr;
    }
    unsigned char *base = static_cast<unsigned char *>(g_scratch_ptr[device]);
    if (params->precision == OPO_FP64) return noalloc_run<double>(*grid, *params, device, base, api, ap, as, ai, n, power);
    return noalloc_run<float>(*grid, *params, device, base, api, ap, as, ai, n, power);
}

int opo_trips_done(const opo_solve *s) { return (s && s->s) ? s->s->trips_done : 0; }
int opo_power(opo_solve *s, double out[3]) {
    SOLVE_OR_FAIL(s);
    if (!out) return fail(OPO_ERR_ARG, "null output");
    return s->s->power(out);
}
int opo_power_time(opo_solve *s, int field, int x, int y, double prefactor, double *out) {
    SOLVE_OR_FAIL(s);
    if (!out) return fail(OPO_ERR_ARG, "null output");
    if (field < 0 || field > 2) return fail(OPO_ERR_ARG, "field must be 0 (pump), 1 (signal) or 2 (idler)");
    if (x < 0 || x >= s->s->g.nx || y < 0 || y >= s->s->g.ny)
        return fail(OPO_ERR_ARG, "pixel (%d, %d) outside the %d x %d grid", x, y, s->s->g.nx, s->s->g.ny);
    return s->s->power_time(field, x, y, prefactor, out);
}
int opo_courant_ok(const opo_params *p) {
    if (!p) return 0;
    // Solver3D.cuh:495-504, 973: dz < 0.2 kp dx dy
    return p->dz < 0.2 * p->kdiff[0] * p->dx * p->dy ? 1 : 0;
}
int opo_profile_trips(opo_solve *s, int n, double ms[OPO_N_STAGES], int launches[OPO_N_STAGES]) {
    SOLVE_OR_FAIL(s);
    if (n < 1 || !ms || !launches) return fail(OPO_ERR_ARG, "opo_profile_trips: bad argument");
    if (s->s->p.n_slices < 1) return fail(OPO_ERR_ARG, "round trips need n_slices >= 1");
    return s->s->profile_trips(n, ms, launches);
}
int opo_autotune(opo_solve *s, int reps, char *report, size_t report_len) {
    SOLVE_OR_FAIL(s);
    std::string rep;
    int rc = s->s->autotune(reps, rep);
    if (report && report_len) {
        std::snprintf(report, report_len, "%s", rep.c_str());
    }
    return rc;
}
int opo_tune_get(int precision, int axis, int n) {
    return tune_get(precision == OPO_FP64, axis, n);
}
int opo_tune_set(int precision, int axis, int n, int variant) {
    if (!tune_set(precision == OPO_FP64, axis, n, variant))
        return fail(OPO_ERR_ARG, "no variant %d for axis %d, size %d", variant, axis, n);
    return OPO_OK;
}
int opo_sync(opo_solve *s) {
    SOLVE_OR_FAIL(s);
    int rc = s->s->select();
    if (rc) return rc;
    CK(hipStreamSynchronize(s->s->stream));
    return OPO_OK;
}

// ------------------------------------------------------------- batching
size_t opo_batch_footprint(const opo_grid *g, int precision, int n_trips) {
    if (!g || n_trips < 0) return 0;
    const size_t n = (size_t)g->nx * g->ny * g->nt;
    const size_t cb = precision == OPO_FP64 ? 16 : 8;
    // Every device allocation of a resident solve, each rounded up to the
    // allocation granularity assumed here: Api, Ap, As, Ai; the table block
    // (twiddles, D, GDD, Q, free-space and modulator tables as hi/lo pairs);
    // the power partials (one per 4096 cells, 1 to 1024, three fields) and
    // counters; the per-trip power series of the batch.
    const size_t kGran = 64 * 1024;
    auto up = [&](size_t b) { return (b + kGran - 1) / kGran * kGran; };
    const size_t tabs = 2 * (size_t)(6 * g->nx + 5 * g->ny + 6 * g->nt) * cb;
    size_t nblk = (n + 4095) / 4096;
    nblk = nblk < 1 ? 1 : (nblk > 1024 ? 1024 : nblk);
    const size_t dbuf = sizeof(double) * (3 * nblk + 4);
    const size_t series = sizeof(double) * 3 * (size_t)(n_trips > 0 ? n_trips : 1);
    return 4 * up(n * cb) + up(tabs) + up(dbuf) + up(4 * sizeof(int)) + up(series);
}

int opo_batch_create(int device, size_t budget, int precision, opo_batch **out) {
    if (!out) return fail(OPO_ERR_ARG, "null output");
    *out = nullptr;
    if (precision != OPO_FP32 && precision != OPO_FP64) return fail(OPO_ERR_ARG, "bad precision");
    int ndev = opo_device_count();
    if (ndev < 1) return fail(OPO_ERR_NOGPU, "no HIP device");
    if (device < 0) device = 0;
    if (device >= ndev) return fail(OPO_ERR_ARG, "device out of range");
    std::unique_ptr<opo_batch> b(new opo_batch);
    b->device = device;
    b->budget = budget;
    b->prec = precision;
    CK(hipSetDevice(device));
    CK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    *out = b.release();
    return OPO_OK;
}

void opo_batch_destroy(opo_batch *b) { delete b; }

int opo_batch_submit(opo_batch *b, const opo_grid *g, const opo_params *p, const void *api,
                     const void *as, const void *ai, int n_trips, uint64_t *ticket) {
    if (!b || !g || !p || !api || !as || !ai || n_trips < 1)
        return fail(OPO_ERR_ARG, "opo_batch_submit: bad argument");
    int rc = validate(g, p);
    if (rc) return rc;
    if (p->n_slices < 1) return fail(OPO_ERR_ARG, "round trips need n_slices >= 1");
    if (p->precision != b->prec) return fail(OPO_ERR_ARG, "precision differs from the batch's");
    if (!batch_supported(g->nx, g->ny, g->nt, p->precision == OPO_FP64))
        return fail(OPO_ERR_UNSUPPORTED, "grid %dx%dx%d is not supported in a batch", g->nx, g->ny, g->nt);
    if (p->dk_phase_scale != 0.0 && p->dk != 0.0)
        return fail(OPO_ERR_UNSUPPORTED, "a batch runs solves without the dk phase term");
    if (g->nx == 1 && !p->no_diffraction)
        return fail(OPO_ERR_UNSUPPORTED, "a 1x1 grid runs in a batch as the plane-wave model (no_diffraction)");
    const size_t fp = opo_batch_footprint(g, p->precision, n_trips);
    if (fp > b->budget) return fail(OPO_ERR_BUDGET, "solve needs %zu bytes, budget is %zu", fp, b->budget);
    std::lock_guard<std::mutex> lk(b->mu);
    auto e = std::make_unique<BatchEntry>();
    e->ticket = b->next_ticket++;
    e->g = *g;
    e->p = *p;
    e->p.omega = nullptr;
    e->n_trips = n_trips;
    e->footprint = fp;
    const size_t fb = (size_t)g->nx * g->ny * g->nt * (p->precision == OPO_FP64 ? 16 : 8);
    e->api.assign((const unsigned char *)api, (const unsigned char *)api + fb);
    e->as.assign((const unsigned char *)as, (const unsigned char *)as + fb);
    e->ai.assign((const unsigned char *)ai, (const unsigned char *)ai + fb);
    if (ticket) *ticket = e->ticket;
    b->queue.push_back(e.get());
    b->all[e->ticket] = std::move(e);
    return OPO_OK;
}

int opo_batch_step(opo_batch *b, int *n_active) {
    if (!b) return fail(OPO_ERR_ARG, "null batch");
    std::lock_guard<std::mutex> lk(b->mu);
    CK(hipSetDevice(b->device));
    if (n_active) *n_active = 0;
    b->admit();
    const int n = (int)b->resident.size();
    if (n == 0) return OPO_OK;
    int rc;
    if (!b->dirty && b->current_variants() != b->built_variants) b->dirty = true;  // re-tuned
    if (b->dirty && (rc = b->rebuild())) return rc;
    for (int i = 0; i < n; ++i) {
        const BatchEntry *e = b->resident[i];
        const int rt = e->done_trips;
        b->h_ts[i].trip = rt;
        b->h_ts[i].fs = e->p.fs_on && !(e->p.skip_fs_on_last_trip && rt == e->p.n_round_trips - 1);
    }
    CK(hipMemcpyAsync(b->d_ts, b->h_ts.data(), sizeof(TripState) * n, hipMemcpyHostToDevice, b->stream));
    CK(hipGraphLaunch(b->graph, b->stream));
    CK(hipMemcpyAsync(b->h_status.data(), b->d_status, sizeof(int) * n, hipMemcpyDeviceToHost, b->stream));
    CK(hipStreamSynchronize(b->stream));
    if ((rc = b->collect())) return rc;
    std::vector<BatchEntry *> keep;
    for (int i = 0; i < n; ++i) {
        BatchEntry *e = b->resident[i];
        if (b->h_status[i]) {
            e->done_trips += 1;
            e->error = OPO_ERR_NAN;
            e->error_msg = "non-finite signal power at round trip " + std::to_string(e->done_trips - 1);
        } else {
            e->done_trips += 1;
        }
        if (e->error || e->done_trips >= e->n_trips) e->finished = true;
        else keep.push_back(e);
    }
    if (keep.size() != b->resident.size()) b->dirty = true;
    b->resident.swap(keep);
    if (n_active) *n_active = n;
    return OPO_OK;
}

int opo_batch_poll(opo_batch *b, uint64_t ticket, int *done, double *power, void *ap, void *as,
                   void *ai) {
    if (!b || !done) return fail(OPO_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(b->mu);
    auto it = b->all.find(ticket);
    if (it == b->all.end()) return fail(OPO_ERR_ARG, "unknown ticket %llu", (unsigned long long)ticket);
    BatchEntry *e = it->second.get();
    *done = e->finished ? 1 : 0;
    if (!e->finished) return OPO_OK;
    int rc = e->error;
    std::string msg = e->error_msg;
    if (e->s) {
        CK(hipSetDevice(b->device));
        if (power && e->done_trips > 0)
            CK(hipMemcpy(power, e->d_series, sizeof(double) * 3 * (size_t)e->done_trips,
                         hipMemcpyDeviceToHost));
        int r2 = e->s->get_fields(ap, as, ai);
        if (!rc) rc = r2;
        b->used -= e->footprint;
        e->s.reset();
        (void)hipFree(e->d_series);
        e->d_series = nullptr;
    }
    b->all.erase(it);
    if (rc) return fail(rc, "%s", msg.c_str());
    return OPO_OK;
}

int opo_batch_rebuilds(opo_batch *b, int *rebuilds, int *updated, double *seconds) {
    if (!b) return fail(OPO_ERR_ARG, "null batch");
    std::lock_guard<std::mutex> lk(b->mu);
    if (rebuilds) *rebuilds = b->n_rebuilds;
    if (updated) *updated = b->n_updated;
    if (seconds) *seconds = b->rebuild_s;
    return OPO_OK;
}

int opo_batch_concurrent_classes(opo_batch *b, int on) {
    if (!b) return fail(OPO_ERR_ARG, "null batch");
    std::lock_guard<std::mutex> lk(b->mu);
    const bool want = on != 0;
    if (want && !b->side[0]) {
        CK(hipSetDevice(b->device));
        for (auto &x : b->side) CK(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
    }
    if (want != b->concurrent) {
        b->concurrent = want;
        b->dirty = true;   // re-capture with or without the branches
    }
    return OPO_OK;
}

int opo_batch_autotune(opo_batch *b, int reps, char *report, size_t report_len) {
    if (!b || reps < 1) return fail(OPO_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(b->mu);
    CK(hipSetDevice(b->device));
    b->admit();
    if (b->resident.empty()) return fail(OPO_ERR_STATE, "no resident solves to tune on");
    std::string rep;
    const int rc = b->autotune(reps, rep);
    if (report && report_len) {
        std::strncpy(report, rep.c_str(), report_len - 1);
        report[report_len - 1] = 0;
    }
    return rc;
}

int opo_batch_time_stages(opo_batch *b, int reps, opo_stage_time *out, int cap, int *n) {
    if (!b || !n || reps < 1 || (cap > 0 && !out)) return fail(OPO_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(b->mu);
    CK(hipSetDevice(b->device));
    if (b->resident.empty()) return fail(OPO_ERR_STATE, "no resident solves to time");
    int rc;
    if (b->dirty && (rc = b->rebuild())) return rc;
    std::vector<opo_stage_time> v;
    if ((rc = b->time_stages(reps, v))) return rc;
    *n = (int)v.size();
    for (int k = 0; k < cap && k < *n; ++k) out[k] = v[k];
    return OPO_OK;
}

int opo_batch_profile(opo_batch *b, int on) {
    if (!b) return fail(OPO_ERR_ARG, "null batch");
    std::lock_guard<std::mutex> lk(b->mu);
    const bool want = on != 0;
    if (want != b->profiling) {   // re-capture with or without the events
        b->profiling = want;
        b->dirty = true;
        b->prof.clear();
        b->prof_alg.clear();
        b->prof_error.clear();
    }
    b->prof_reset();   // start the sums again
    return OPO_OK;
}

int opo_batch_profile_read(opo_batch *b, opo_stage_time *out, int cap, int *n) {
    if (!b || !n || (cap > 0 && !out)) return fail(OPO_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(b->mu);
    if (!b->prof_error.empty()) return fail(OPO_ERR_HIP, "%s", b->prof_error.c_str());
    *n = (int)b->prof.size();
    for (int k = 0; k < cap && k < *n; ++k) {
        out[k] = b->prof[k];
        if (out[k].launches) out[k].alg_bytes = b->prof_alg[k] / out[k].launches;
    }
    return OPO_OK;
}

int opo_batch_stats(opo_batch *b, int *resident, int *queued, size_t *bytes_used) {
    if (!b) return fail(OPO_ERR_ARG, "null batch");
    std::lock_guard<std::mutex> lk(b->mu);
    if (resident) *resident = (int)b->resident.size();
    if (queued) *queued = (int)b->queue.size();
    if (bytes_used) *bytes_used = b->used;
    return OPO_OK;
}

}  // extern "C"
