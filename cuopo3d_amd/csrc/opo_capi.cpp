// opo_capi.cpp -- C ABI of the split-step path (include/opo3d.h): errors,
// the single-solve entry points, the launch-variant tables and the entry
// point that allocates nothing.  The solve itself is opo_solve.hpp; batches
// are opo_batch.cpp.
#include "opo_solve.hpp"

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
const std::string &last_error_text() { return g_err; }

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

int validate(const opo_grid *g, const opo_params *p) {
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

int make_solve(const opo_grid *g, const opo_params *p, int device,
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

}  // extern "C"
