// opo_kernels.hip -- CDNA4 kernels of the split-step Fourier path.
//
// Per z-slice the reference applies N(dz/2) D(dz) Q(dz) N(dz/2)
// (Solver3D.cuh:949-957).  D is diagonal in omega (a line operator along t),
// Q is diagonal in (kx, ky) and its table factorises exactly into an x
// factor and a y factor, so Q = Q_y Q_x is two line operators on each
// t-plane.  The kernels here:
//   k_rk4      pointwise RK4 half step(s) of all three fields, optionally
//              reading the pump from the injection buffer and applying the
//              mirror factors to s and i on the way out;
//   k_tline    along-t operator on tiles of TX consecutive x pixels x NT,
//              staged in LDS (coalesced TX*8-byte rows);
//   k_xyplane  Q_x then Q_y on a whole t-plane staged in LDS (rows padded
//              by one element so column and row walks are bank-spread);
//   k_power_*  compensated power sums with a fixed reduction order.
#include "opo_device.hpp"
#include "opo_kernels.hpp"
#include "opo_rk4.hpp"
#include "opo_passes.hpp"

namespace opo {

template <typename T, bool PHASE>
__global__ void __launch_bounds__(256) k_rk4(const cplx<T> *__restrict__ psrc, cplx<T> *ap,
                                             cplx<T> *as, cplx<T> *ai, size_t n, Rk4Args<T> a,
                                             const int *status) {
    if (status && *status) return;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        Tri<T> A = {psrc[i], as[i], ai[i]};
        rk4_half_step<T, PHASE>(a, 0, A);
        if (a.nsteps > 1) rk4_half_step<T, PHASE>(a, 1, A);
        if (a.out_scale) {
            // roundTripChange (Cavity3D.cuh:12-28): (sqrt(R) e^{i phi}) * A
            A.s = cmul_hl(A.s, cplx<T>{a.ms_re, a.ms_im}, cplx<T>{a.ms_lo_re, a.ms_lo_im});
            A.i = cmul_hl(A.i, cplx<T>{a.mi_re, a.mi_im}, cplx<T>{a.mi_lo_re, a.mi_lo_im});
        }
        ap[i] = A.p;
        as[i] = A.s;
        ai[i] = A.i;
    }
}

template <typename T>
hipError_t launch_rk4(hipStream_t st, const cplx<T> *psrc, cplx<T> *ap, cplx<T> *as, cplx<T> *ai,
                      size_t n, const Rk4Args<T> &a, const int *status) {
    const int bs = 256;
    size_t want = (n + bs - 1) / bs;
    int grid = (int)(want < 8192 ? want : 8192);
    if (grid < 1) grid = 1;
    if (a.has_phase)
        hipLaunchKernelGGL((k_rk4<T, true>), dim3(grid), dim3(bs), 0, st, psrc, ap, as, ai, n, a, status);
    else
        hipLaunchKernelGGL((k_rk4<T, false>), dim3(grid), dim3(bs), 0, st, psrc, ap, as, ai, n, a, status);
    return hipGetLastError();
}

// ---------------------------------------------------------- t operator
// Tile width (pixels along x) per time length and precision: keep the tile
// at <= 64 KiB so two or more workgroups share a CU.
template <typename T, int NT> constexpr int tline_tx() {
    constexpr int bytes = NT * (int)sizeof(cplx<T>);
    return bytes <= 4096 ? 16 : (65536 / bytes >= 16 ? 16 : 65536 / bytes);
}

template <typename T, int NT, int TX, int BS>
__global__ void __launch_bounds__(BS) k_tline(FieldOps<T> ops, int nx, int ny, const cplx<T> *__restrict__ tw,
                                              const int *status) {
    if (status && *status) return;
    __shared__ cplx<T> lds[NT * TX];
    const int tid = threadIdx.x;
    const int tiles_x = nx / TX;
    const int y = blockIdx.x / tiles_x, x0 = (blockIdx.x % tiles_x) * TX;
    const int f = blockIdx.y;
    cplx<T> *A = ops.f[f];
    const size_t plane = (size_t)nx * ny;
    const size_t base = (size_t)y * nx + x0;
    for (int i = tid; i < NT * TX; i += BS) {
        const int t = i / TX, xl = i % TX;
        lds[i] = A[(size_t)t * plane + base + xl];
    }
    __syncthreads();
    line_op<T, NT, +1, 1, TX>(lds, TX, tw, ops.tab[f], tid, BS);
    for (int i = tid; i < NT * TX; i += BS) {
        const int t = i / TX, xl = i % TX;
        A[(size_t)t * plane + base + xl] = lds[i];
    }
}

template <typename T, int NT>
static hipError_t tline_dispatch(hipStream_t st, const Geometry &g, const FieldOps<T> &ops, int nf,
                                 const cplx<T> *tw, const int *status) {
    constexpr int TXmax = tline_tx<T, NT>();
    if (g.nx % TXmax == 0) {
        dim3 grid((unsigned)((g.nx / TXmax) * g.ny), (unsigned)nf);
        hipLaunchKernelGGL((k_tline<T, NT, TXmax, 256>), grid, dim3(256), 0, st, ops, g.nx, g.ny, tw, status);
    } else if (g.nx == 1) {
        dim3 grid((unsigned)g.ny, (unsigned)nf);
        hipLaunchKernelGGL((k_tline<T, NT, 1, 256>), grid, dim3(256), 0, st, ops, g.nx, g.ny, tw, status);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

bool tline_supported(int nt, bool) {
    switch (nt) {
        case 1: case 16: case 32: case 64: case 128: case 256: case 512: case 1024: return true;
        default: return false;
    }
}

template <typename T>
hipError_t launch_tline(hipStream_t st, const Geometry &g, const FieldOps<T> &ops, int nf,
                        const cplx<T> *tw, const int *status) {
    switch (g.nt) {
        case 1: return tline_dispatch<T, 1>(st, g, ops, nf, tw, status);
        case 16: return tline_dispatch<T, 16>(st, g, ops, nf, tw, status);
        case 32: return tline_dispatch<T, 32>(st, g, ops, nf, tw, status);
        case 64: return tline_dispatch<T, 64>(st, g, ops, nf, tw, status);
        case 128: return tline_dispatch<T, 128>(st, g, ops, nf, tw, status);
        case 256: return tline_dispatch<T, 256>(st, g, ops, nf, tw, status);
        case 512: return tline_dispatch<T, 512>(st, g, ops, nf, tw, status);
        case 1024: return tline_dispatch<T, 1024>(st, g, ops, nf, tw, status);
        default: return hipErrorInvalidValue;
    }
}

template <typename T, int NT, int TX, int BS, bool PHASE>
__global__ void __launch_bounds__(BS) k_tline3(TLine3Args<T> a, int nx, int ny,
                                               const cplx<T> *__restrict__ tw, const int *status) {
    if (status && *status) return;
    __shared__ cplx<T> lds[3 * NT * TX];
    nl_time_tile<T, NT, TX, BS, PHASE>(a, nx, ny, xcd_local_tile(blockIdx.x, gridDim.x), tw, lds);
}

template <typename T, int NT>
static hipError_t tline3_dispatch(hipStream_t st, const Geometry &g, const TLine3Args<T> &a,
                                  const cplx<T> *tw, const int *status) {
    constexpr int TX = tline3_tx<T, NT>();
    const bool ph = a.rk.has_phase != 0;
    if (g.nx % TX == 0) {
        dim3 grid((unsigned)((g.nx / TX) * g.ny));
        if (ph) hipLaunchKernelGGL((k_tline3<T, NT, TX, 256, true>), grid, dim3(256), 0, st, a, g.nx, g.ny, tw, status);
        else hipLaunchKernelGGL((k_tline3<T, NT, TX, 256, false>), grid, dim3(256), 0, st, a, g.nx, g.ny, tw, status);
    } else if (g.nx == 1) {
        dim3 grid((unsigned)g.ny);
        if (ph) hipLaunchKernelGGL((k_tline3<T, NT, 1, 256, true>), grid, dim3(256), 0, st, a, g.nx, g.ny, tw, status);
        else hipLaunchKernelGGL((k_tline3<T, NT, 1, 256, false>), grid, dim3(256), 0, st, a, g.nx, g.ny, tw, status);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_tline3(hipStream_t st, const Geometry &g, const TLine3Args<T> &a,
                         const cplx<T> *tw, const int *status) {
    switch (g.nt) {
        case 1: return tline3_dispatch<T, 1>(st, g, a, tw, status);
        case 16: return tline3_dispatch<T, 16>(st, g, a, tw, status);
        case 32: return tline3_dispatch<T, 32>(st, g, a, tw, status);
        case 64: return tline3_dispatch<T, 64>(st, g, a, tw, status);
        case 128: return tline3_dispatch<T, 128>(st, g, a, tw, status);
        case 256: return tline3_dispatch<T, 256>(st, g, a, tw, status);
        case 512: return tline3_dispatch<T, 512>(st, g, a, tw, status);
        case 1024: return tline3_dispatch<T, 1024>(st, g, a, tw, status);
        default: return hipErrorInvalidValue;
    }
}

// --------------------------------------------------------- xy operator
template <typename T, int NX, int BS>
__global__ void __launch_bounds__(BS) k_xyplane(FieldOps<T> ops, int nt, const cplx<T> *__restrict__ tw,
                                                const int *status) {
    if (status && *status) return;
    __shared__ cplx<T> lds[NX * (NX + 1)];
    const int t = blockIdx.x, f = blockIdx.y;
    xy_plane<T, NX, BS>(ops.f[f] + (size_t)t * NX * NX, ops.tab[f], ops.taby[f], tw, lds);
}

template <typename T, int NX>
static hipError_t xy_dispatch(hipStream_t st, const Geometry &g, const FieldOps<T> &ops, int nf,
                              const cplx<T> *tw, const int *status) {
    dim3 grid((unsigned)g.nt, (unsigned)nf);
    constexpr int BS = NX >= 96 ? 512 : 256;
    hipLaunchKernelGGL((k_xyplane<T, NX, BS>), grid, dim3(BS), 0, st, ops, g.nt, tw, status);
    return hipGetLastError();
}

bool xyplane_supported(int nx, int ny, bool fp64) {
    if (nx != ny) return false;
    switch (nx) {
        case 1: case 16: case 32: case 64: case 96: return true;
        case 128: return !fp64;
        default: return false;
    }
}

template <typename T>
hipError_t launch_xyplane(hipStream_t st, const Geometry &g, const FieldOps<T> &ops, int nf,
                          const cplx<T> *tw, const int *status) {
    if (g.nx != g.ny) return hipErrorInvalidValue;
    switch (g.nx) {
        case 1: return xy_dispatch<T, 1>(st, g, ops, nf, tw, status);
        case 16: return xy_dispatch<T, 16>(st, g, ops, nf, tw, status);
        case 32: return xy_dispatch<T, 32>(st, g, ops, nf, tw, status);
        case 64: return xy_dispatch<T, 64>(st, g, ops, nf, tw, status);
        case 96: return xy_dispatch<T, 96>(st, g, ops, nf, tw, status);
        case 128:
            if constexpr (sizeof(T) == 4) return xy_dispatch<T, 128>(st, g, ops, nf, tw, status);
            else return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
}

// --------------------------------------------------------------- power
__device__ __forceinline__ void neumaier_add(double &sum, double &comp, double v) {
    double t = sum + v;
    if (fabs(sum) >= fabs(v)) comp += (sum - t) + v;
    else comp += (v - t) + sum;
    sum = t;
}

template <typename T>
__global__ void __launch_bounds__(256) k_power_partial(const cplx<T> *f0, const cplx<T> *f1,
                                                       const cplx<T> *f2, size_t n, double *partial,
                                                       const int *status) {
    if (status && *status) return;
    __shared__ double red[3][256];
    const cplx<T> *F[3] = {f0, f1, f2};
    const int tid = threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (int f = 0; f < 3; ++f) {
        double s = 0.0, c = 0.0;
        if (F[f]) {
            for (size_t i = (size_t)blockIdx.x * blockDim.x + tid; i < n; i += stride) {
                cplx<T> a = F[f][i];
                double re = (double)a.x, im = (double)a.y;
                neumaier_add(s, c, re * re + im * im);
            }
        }
        red[f][tid] = s + c;
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w)
            for (int f = 0; f < 3; ++f) red[f][tid] += red[f][tid + w];
        __syncthreads();
    }
    if (tid == 0)
        for (int f = 0; f < 3; ++f) partial[(size_t)f * gridDim.x + blockIdx.x] = red[f][0];
}

__global__ void __launch_bounds__(256) k_power_final(const double *partial, int nblk, double pref0,
                                                     double pref1, double pref2, double *series,
                                                     int *trip, int *status, int check_nan,
                                                     int series_cap) {
    if (status && *status) return;
    __shared__ double red[3][256];
    const int tid = threadIdx.x;
    for (int f = 0; f < 3; ++f) {
        double s = 0.0, c = 0.0;
        for (int b = tid; b < nblk; b += 256) neumaier_add(s, c, partial[(size_t)f * nblk + b]);
        red[f][tid] = s + c;
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w)
            for (int f = 0; f < 3; ++f) red[f][tid] += red[f][tid + w];
        __syncthreads();
    }
    if (tid == 0) {
        double pw[3] = {pref0 * red[0][0], pref1 * red[1][0], pref2 * red[2][0]};
        int k = 0;
        if (trip) {
            k = *trip;
            *trip = k + 1;
        }
        if (k < series_cap)
            for (int f = 0; f < 3; ++f) series[3 * (size_t)k + f] = pw[f];
        if (check_nan && status && !isfinite(pw[1])) *status = k + 1;
    }
}

template <typename T>
hipError_t launch_power_partial(hipStream_t st, const cplx<T> *f0, const cplx<T> *f1,
                                const cplx<T> *f2, size_t n, int nblk, double *partial,
                                const int *status) {
    hipLaunchKernelGGL((k_power_partial<T>), dim3(nblk), dim3(256), 0, st, f0, f1, f2, n, partial, status);
    return hipGetLastError();
}

hipError_t launch_power_final(hipStream_t st, const double *partial, int nblk, double pref0,
                              double pref1, double pref2, double *series, int *trip, int *status,
                              int check_nan, int series_cap) {
    hipLaunchKernelGGL(k_power_final, dim3(1), dim3(256), 0, st, partial, nblk, pref0, pref1, pref2,
                       series, trip, status, check_nan, series_cap);
    return hipGetLastError();
}

// ------------------------------------------------------- instantiation
#define OPO_INST(T)                                                                               \
    template hipError_t launch_rk4<T>(hipStream_t, const cplx<T> *, cplx<T> *, cplx<T> *,          \
                                      cplx<T> *, size_t, const Rk4Args<T> &, const int *);        \
    template hipError_t launch_tline<T>(hipStream_t, const Geometry &, const FieldOps<T> &, int,   \
                                        const cplx<T> *, const int *);                            \
    template hipError_t launch_tline3<T>(hipStream_t, const Geometry &, const TLine3Args<T> &,      \
                                         const cplx<T> *, const int *);                           \
    template hipError_t launch_xyplane<T>(hipStream_t, const Geometry &, const FieldOps<T> &, int, \
                                          const cplx<T> *, const int *);                          \
    template hipError_t launch_power_partial<T>(hipStream_t, const cplx<T> *, const cplx<T> *,     \
                                                const cplx<T> *, size_t, int, double *, const int *);
OPO_INST(float)
OPO_INST(double)

}  // namespace opo
