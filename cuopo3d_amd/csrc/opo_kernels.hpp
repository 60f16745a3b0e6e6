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
    T kp, ks, ki;      // kappa of pump, signal, idler (1/V)
    T c1, c3;          // stage coefficients: A + c1 k1, A + c1 k2, A + c3 k3
    T h;               // half-step dz/2
    int nsteps;        // consecutive half steps fused (1 or 2)
    int has_phase;     // dk phase present (dk_phase_scale != 0)
    T es_re[2][3], es_im[2][3];  // e^{+i s dk z} at the 3 stage positions of each step
    int out_scale;     // multiply s and i by the mirror factors after the step
    T ms_re, ms_im, mi_re, mi_im;          // mirror factors (rounded)
    T ms_lo_re, ms_lo_im, mi_lo_re, mi_lo_im;  // their rounding residuals
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
                      size_t n, const Rk4Args<T> &a, const int *status);

// Along-t operator F_t^{-}( tab . F_t^{+}(A) ) for nf fields (blockIdx.y = field).
template <typename T>
hipError_t launch_tline(hipStream_t st, const Geometry &g, const FieldOps<T> &ops, int nf,
                        const cplx<T> *tw_t, const int *status);

// xy operator Q_y Q_x (each F^{+}( tab . F^{-}(A) )) on every t-plane of nf fields.
template <typename T>
hipError_t launch_xyplane(hipStream_t st, const Geometry &g, const FieldOps<T> &ops, int nf,
                          const cplx<T> *tw_x, const int *status);

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

bool tline_supported(int nt, bool fp64);
bool xyplane_supported(int nx, int ny, bool fp64);

}  // namespace opo
