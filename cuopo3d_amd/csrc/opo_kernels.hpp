// opo_kernels.hpp -- host-side launch interface of the device kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace opo {

template <typename T> struct cplx;

// Arguments of the pointwise nonlinear step (reference solverRK4,
// Solver3D.cuh:872-943; right-hand side dAdz, :365-382).
template <typename T> struct Rk4Args {
    // Per field (p, s, i), with kappa the coupling and h = dz/2: kappa*c1 and
    // kappa*c3 (stage points A + c1 k1, A + c1 k2, A + c3 k3) and kappa*h/6
    // (the update weight), formed in FP64.
    T kc1[3], kc3[3], kh6[3];
    int nsteps;        // consecutive half steps fused (1 or 2)
    int has_phase;     // dk phase present (dk_phase_scale != 0)
    T es_re[2][3], es_im[2][3];  // e^{+i s dk z} at the 3 stage positions of each step
    int out_scale;     // multiply s and i by the mirror factors after the step
    T ms_re, ms_im, mi_re, mi_im;          // mirror factors (rounded)
    T ms_lo_re, ms_lo_im, mi_lo_re, mi_lo_im;  // their rounding residuals
    // With out_scale: the modulator factor of time sample t as a (hi, lo)
    // pair at eom[2t], eom[2t+1], applied to s and i before the mirror;
    // nullptr when there is no modulator.
    const cplx<T> *eom;
};

// A set of up to 3 fields, each with its own diagonal table.
template <typename T> struct FieldOps {
    cplx<T> *f[3];
    const cplx<T> *tab[3];    // along-axis tables (t) or x tables (xy)
    const cplx<T> *taby[3];   // y tables (xy kernel only)
};

struct Geometry {
    int nx, ny, nt;
};

// Fused pass: nonlinear half step(s) of all three fields followed by the
// along-t operator on fields first_field..2, per tile of pixels (field f
// uses the table tab0 + (f - first_field)*tab_stride).  Fields below
// first_field are stored right after the pointwise step.  rk.nsteps == 0
// skips the pointwise step.
template <typename T> struct TLine3Args {
    const cplx<T> *psrc;   // pump read from here (Api on a trip's first slice)
    cplx<T> *f[3];         // Ap, As, Ai
    const cplx<T> *tab0;
    int tab_stride;
    int first_field;       // 0: all three fields; 1: signal and idler
    Rk4Args<T> rk;
};

template <typename T>
hipError_t launch_tline3(hipStream_t st, const Geometry &g, const TLine3Args<T> &a,
                         const cplx<T> *tw_t, const int *status);

// Launchers; return hipError_t of the launch.
template <typename T>
hipError_t launch_rk4(hipStream_t st, const cplx<T> *psrc, cplx<T> *ap, cplx<T> *as, cplx<T> *ai,
                      size_t n, size_t plane, const Rk4Args<T> &a, const int *status);

// Along-t operator F_t^{-}( tab . F_t^{+}(A) ) for nf fields (blockIdx.y = field).
template <typename T>
hipError_t launch_tline(hipStream_t st, const Geometry &g, const FieldOps<T> &ops, int nf,
                        const cplx<T> *tw_t, const int *status);

// xy operator Q_y Q_x (each F^{+}( tab . F^{-}(A) )) on every t-plane of nf
// fields; tw_x, tw_y the NX- and NY-point twiddles (the same table when
// NX = NY).  Square planes run as one pass through LDS, NX != NY planes as
// a pass over bands of rows then one over bands of columns.
template <typename T>
hipError_t launch_xyplane(hipStream_t st, const Geometry &g, const FieldOps<T> &ops, int nf,
                          const cplx<T> *tw_x, const cplx<T> *tw_y, const int *status);

// Power partial sums: partial[f*nblk + b], double, Neumaier-compensated per
// thread and reduced in a fixed tree order per block.
template <typename T>
hipError_t launch_power_partial(hipStream_t st, const cplx<T> *f0, const cplx<T> *f1,
                                const cplx<T> *f2, size_t n, int nblk, double *partial,
                                const int *status);
// Final reduction of the partials in fixed order; writes pref[f]*sum to
// series[3*(*trip) + f] (trip == nullptr: series[f]), increments *trip, and
// sets *status when the signal power is not finite and check_nan is set.
hipError_t launch_power_final(hipStream_t st, const double *partial, int nblk, double pref0,
                              double pref1, double pref2, double *series, int *trip, int *status,
                              int check_nan, int series_cap);

// pref * sum_t |f[t*plane]|^2 into out[0] (one workgroup; f points at the
// pixel's t = 0 sample).
template <typename T>
hipError_t launch_power_pixel(hipStream_t st, const cplx<T> *f, size_t plane, int nt, double pref,
                              double *out);

// Launch-variant tables (the autotuner's choices), per precision and size.
// axis 0: fused t pass, keyed by NT; axis 1: xy pass, keyed by NX.
int tune_variant_count(int axis);
bool tune_variant_valid(bool fp64, int axis, int n, int v);
int tune_get(bool fp64, int axis, int n);
bool tune_set(bool fp64, int axis, int n, int v);
int tune_tile_px(bool fp64, int nt, int v);   // pixels per tile of a t-pass variant
// The t-pass variant a launch over pixel rows of nx pixels uses: the tuned
// choice for NT when its tile divides nx, else variant 0, whose tile (16,
// 8, 4, 2 or 1 pixels) divides every supported NX.  The choice is
// process-wide and keyed by NT alone, so a tile tuned on a wide grid may
// not fit a narrow one.
int time_variant_for(bool fp64, int nt, int nx);
int tune_block(bool fp64, int axis, int n, int v);
int tune_per_cu(bool fp64, int axis, int n, int v);
// Grid of a persistent launch: per_cu workgroups per CU of the current
// device, at most one per unit.
unsigned persistent_grid(int units, int per_cu);  // persistent workgroups per CU, 0: one unit each

// Consecutive t-planes of one field that one workgroup of xy variant v
// transforms.  Lines of at most 32 points are single register DFTs, one
// task per line, so one 32 x 32 plane gives a 256-thread workgroup 32
// tasks per phase; variants 3-5 of NX <= 32 take 256 / NX planes (256
// tasks), the others one plane.  Every supported NT (16 and up) is a
// multiple of the group.
__host__ __device__ constexpr int xy_group(int nx, int v) {
    return (nx > 1 && nx <= 32 && v >= 3 && v < 6) ? 256 / nx : 1;
}

bool tline_supported(int nt, bool fp64);
bool xyplane_supported(int nx, int ny, bool fp64);

}  // namespace opo

namespace opo {

// ------------------------------------------------------------ batching
// Everything a batched kernel needs about one solve (device memory).
template <typename T> struct SolveDesc {
    const cplx<T> *api;
    cplx<T> *f[3];
    const cplx<T> *twx, *twt;      // twiddles (hi/lo pairs)
    const cplx<T> *tabT, *tabG;    // D tables (3 x 2nt, stride 2nt), GDD table
    const cplx<T> *qx, *qy;        // Q factor tables (3 fields, stride 2nx / 2ny)
    const cplx<T> *qfsx, *qfsy;    // free-space factor tables (s, i)
    int nx, ny, nt, gdd_on;
    const cplx<T> *eom;            // modulator table (hi/lo pairs), nullptr when off
    T kc1[3], kc3[3], kh6[3];      // nonlinear step weights (see Rk4Args)
    T ms_re, ms_im, ms_lo_re, ms_lo_im, mi_re, mi_im, mi_lo_re, mi_lo_im;  // mirror
    double pref[3];                // power prefactors
    double *series;                // per-trip power, 3 per trip
    int series_cap;
    int nchunk;                    // power partial sums per field
};

// Host-written per-trip state of a solve.
struct TripState {
    int trip;   // index of the trip being run (power goes to series[3*trip])
    int fs;     // free-space step on this trip
};

// One workgroup's work: a solve and an item within it (tile, plane, chunk).
struct WorkItem {
    int solve;
    int item;
};

enum { BATCH_FIRST = 0, BATCH_MID = 1, BATCH_LAST = 2 };

// Fused pass for every work item of one NT class.  mode: BATCH_FIRST (pump
// from Api, one half step), BATCH_MID (two half steps), BATCH_LAST (one half
// step, mirror, GDD of s and i when the solve has it on).  variant: the
// launch variant the work list was built for (its tile width); the
// process-wide choice may have changed since.  kBatchPlaneVariant: the
// plane-wave solves of the class (NX = NY = 1), one pixel per workgroup.
constexpr int kBatchPlaneVariant = -1;
template <typename T>
hipError_t launch_batch_time(hipStream_t st, int nt, int n_work, const SolveDesc<T> *sd,
                             const int *status, const WorkItem *work, int mode, int variant);
// xy operator for every work item of one NX class: item = f*groups + group
// of xy_group(NX, variant) consecutive planes, over the three fields
// (free_space = 0) or over s and i with the free-space tables on solves
// whose TripState.fs is set (free_space = 1).
template <typename T>
hipError_t launch_batch_xy(hipStream_t st, int nx, int n_work, const SolveDesc<T> *sd,
                           const TripState *ts, const int *status, const WorkItem *work,
                           int free_space, int variant);
// Per-solve power: partial sums (item = chunk) then one workgroup per solve
// writing series[3*trip + f]; a non-finite signal power sets status.
template <typename T>
hipError_t launch_batch_power(hipStream_t st, int n_work, int n_solves, const SolveDesc<T> *sd,
                              const TripState *ts, int *status, const WorkItem *work,
                              double *partial, int max_chunk);
bool batch_supported(int nx, int ny, int nt, bool fp64);

}  // namespace opo
