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

// ------------------------------------------------ fused nonlinear + t op
// Blocks are handed to the 8 XCDs round-robin (block b runs on XCD b % 8).
// Remap so that neighbouring tiles, which share cache lines at the tile
// edges, are processed on the same XCD and hit its L2.
__device__ __forceinline__ int xcd_local_tile(int b, int nb) {
    if (nb % 8 != 0) return b;
    return (b % 8) * (nb / 8) + b / 8;
}

// LDS budget of the fused pass: three tiles of NT x TX <= 48 KiB (two or
// more workgroups per CU), at least 4 pixels wide.
template <typename T, int NT> constexpr int tline3_tx() {
    constexpr int per_px = 3 * NT * (int)sizeof(cplx<T>);
    constexpr int tx = (per_px * 4 > 49152) ? 98304 / per_px : 49152 / per_px;
    return tx >= 16 ? 16 : (tx >= 8 ? 8 : (tx >= 4 ? 4 : (tx >= 2 ? 2 : 1)));
}

template <typename T> struct alignas(4 * sizeof(T)) cpair {
    cplx<T> a, b;
};

template <typename T, int NT, int TX, int BS, bool PHASE>
__global__ void __launch_bounds__(BS) k_tline3(TLine3Args<T> a, int nx, int ny,
                                               const cplx<T> *__restrict__ tw, const int *status) {
    if (status && *status) return;
    constexpr int TILE = NT * TX;
    __shared__ cplx<T> lds[3 * TILE];
    const int tid = threadIdx.x;
    const int tiles_x = nx / TX;
    const int tile = xcd_local_tile(blockIdx.x, gridDim.x);
    const int y = tile / tiles_x, x0 = (tile % tiles_x) * TX;
    const size_t plane = (size_t)nx * ny;
    const size_t base = (size_t)y * nx + x0;
    const Rk4Args<T> &rk = a.rk;
    // pointwise step in registers, two pixels per thread per iteration
    if constexpr (TX % 2 == 0) {
        for (int i = 2 * tid; i < TILE; i += 2 * BS) {
            const int t = i / TX, xl = i % TX;
            const size_t gi = (size_t)t * plane + base + xl;
            cpair<T> p = *reinterpret_cast<const cpair<T> *>(a.psrc + gi);
            cpair<T> s = *reinterpret_cast<const cpair<T> *>(a.f[1] + gi);
            cpair<T> q = *reinterpret_cast<const cpair<T> *>(a.f[2] + gi);
            Tri<T> A0 = {p.a, s.a, q.a}, A1 = {p.b, s.b, q.b};
            if (rk.nsteps > 0) {
                rk4_half_step<T, PHASE>(rk, 0, A0);
                rk4_half_step<T, PHASE>(rk, 0, A1);
                if (rk.nsteps > 1) {
                    rk4_half_step<T, PHASE>(rk, 1, A0);
                    rk4_half_step<T, PHASE>(rk, 1, A1);
                }
            }
            if (rk.out_scale) {
                const cplx<T> ms = {rk.ms_re, rk.ms_im}, msl = {rk.ms_lo_re, rk.ms_lo_im};
                const cplx<T> mi = {rk.mi_re, rk.mi_im}, mil = {rk.mi_lo_re, rk.mi_lo_im};
                A0.s = cmul_hl(A0.s, ms, msl);
                A1.s = cmul_hl(A1.s, ms, msl);
                A0.i = cmul_hl(A0.i, mi, mil);
                A1.i = cmul_hl(A1.i, mi, mil);
            }
            lds[i] = A0.p;
            lds[i + 1] = A1.p;
            lds[TILE + i] = A0.s;
            lds[TILE + i + 1] = A1.s;
            lds[2 * TILE + i] = A0.i;
            lds[2 * TILE + i + 1] = A1.i;
        }
    } else {
        for (int i = tid; i < TILE; i += BS) {
            const int t = i / TX, xl = i % TX;
            const size_t gi = (size_t)t * plane + base + xl;
            Tri<T> A = {a.psrc[gi], a.f[1][gi], a.f[2][gi]};
            if (rk.nsteps > 0) rk4_half_step<T, PHASE>(rk, 0, A);
            if (rk.nsteps > 1) rk4_half_step<T, PHASE>(rk, 1, A);
            if (rk.out_scale) {
                A.s = cmul_hl(A.s, cplx<T>{rk.ms_re, rk.ms_im}, cplx<T>{rk.ms_lo_re, rk.ms_lo_im});
                A.i = cmul_hl(A.i, cplx<T>{rk.mi_re, rk.mi_im}, cplx<T>{rk.mi_lo_re, rk.mi_lo_im});
            }
            lds[i] = A.p;
            lds[TILE + i] = A.s;
            lds[2 * TILE + i] = A.i;
        }
    }
    __syncthreads();
    const LineSet<T> ls{lds + a.first_field * TILE, TILE, a.tab0, a.tab_stride, 3 - a.first_field};
    line_op_set<T, NT, +1, 1, TX>(ls, TX, tw, tid, BS);
    if constexpr (TX % 2 == 0) {
        for (int i = 2 * tid; i < TILE; i += 2 * BS) {
            const int t = i / TX, xl = i % TX;
            const size_t gi = (size_t)t * plane + base + xl;
            *reinterpret_cast<cpair<T> *>(a.f[0] + gi) = cpair<T>{lds[i], lds[i + 1]};
            *reinterpret_cast<cpair<T> *>(a.f[1] + gi) = cpair<T>{lds[TILE + i], lds[TILE + i + 1]};
            *reinterpret_cast<cpair<T> *>(a.f[2] + gi) =
                cpair<T>{lds[2 * TILE + i], lds[2 * TILE + i + 1]};
        }
    } else {
        for (int i = tid; i < TILE; i += BS) {
            const int t = i / TX, xl = i % TX;
            const size_t gi = (size_t)t * plane + base + xl;
            for (int f = 0; f < 3; ++f) a.f[f][gi] = lds[f * TILE + i];
        }
    }
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
    constexpr int NXP = NX + 1;  // padded row pitch
    __shared__ cplx<T> lds[NX * NXP];
    const int tid = threadIdx.x;
    const int t = blockIdx.x, f = blockIdx.y;
    cplx<T> *A = ops.f[f] + (size_t)t * NX * NX;
    for (int i = tid; i < NX * NX; i += BS) lds[(i / NX) * NXP + (i % NX)] = A[i];
    __syncthreads();
    // Q_x: lines are rows (pitch NXP), elements contiguous.
    line_op<T, NX, -1, NXP, 1>(lds, NX, tw, ops.tab[f], tid, BS);
    // Q_y: lines are columns, elements one row apart.
    line_op<T, NX, -1, 1, NXP>(lds, NX, tw, ops.taby[f], tid, BS);
    for (int i = tid; i < NX * NX; i += BS) A[i] = lds[(i / NX) * NXP + (i % NX)];
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
