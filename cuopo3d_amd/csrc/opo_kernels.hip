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
#include <atomic>

#include "opo_device.hpp"
#include "opo_kernels.hpp"
#include "opo_rk4.hpp"
#include "opo_passes.hpp"
#include "opo_power.hpp"

namespace opo {

template <typename T, bool PHASE>
__global__ void __launch_bounds__(256) k_rk4(const cplx<T> *__restrict__ psrc, cplx<T> *ap,
                                             cplx<T> *as, cplx<T> *ai, size_t n, size_t plane, Rk4Args<T> a,
                                             const int *status) {
    if (status && *status) return;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        Tri<T> A = {psrc[i], as[i], ai[i]};
        rk4_half_step<T, PHASE>(a, 0, A);
        if (a.nsteps > 1) rk4_half_step<T, PHASE>(a, 1, A);
        if (a.out_scale) {
            if (a.eom) {  // KernelEOM (Solver1D.cuh:62-72) at time sample i / plane
                const size_t t = i / plane;
                A.s = cmul_hl(A.s, a.eom[2 * t], a.eom[2 * t + 1]);
                A.i = cmul_hl(A.i, a.eom[2 * t], a.eom[2 * t + 1]);
            }
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
                      size_t n, size_t plane, const Rk4Args<T> &a, const int *status) {
    const int bs = 256;
    size_t want = (n + bs - 1) / bs;
    int grid = (int)(want < 8192 ? want : 8192);
    if (grid < 1) grid = 1;
    if (a.has_phase)
        hipLaunchKernelGGL((k_rk4<T, true>), dim3(grid), dim3(bs), 0, st, psrc, ap, as, ai, n, plane, a, status);
    else
        hipLaunchKernelGGL((k_rk4<T, false>), dim3(grid), dim3(bs), 0, st, psrc, ap, as, ai, n, plane, a, status);
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
    chunked_copy<T, NT * TX, BS>(tid, [&](int i) { return A[(size_t)(i / TX) * plane + base + i % TX]; },
                                 [&](int i, cplx<T> v) { lds[i] = v; });
    __syncthreads();
    line_op<T, NT, +1, 1, TX>(lds, TX, tw, ops.tab[f], tid, BS);
    chunked_copy<T, NT * TX, BS>(tid, [&](int i) { return lds[i]; },
                                 [&](int i, cplx<T> v) { A[(size_t)(i / TX) * plane + base + i % TX] = v; });
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

bool tline_supported(int nt, bool fp64) {
    switch (nt) {
        case 1: case 16: case 32: case 64: case 128: case 256: case 512: case 1024: case 2048: return true;
        case 4096: return !fp64;   // three fields of a 4096-sample line fill 102 KiB of LDS in FP32
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
        case 2048: return tline_dispatch<T, 2048>(st, g, ops, nf, tw, status);
        case 4096:
            if constexpr (sizeof(T) == 4) return tline_dispatch<T, 4096>(st, g, ops, nf, tw, status);
            else return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
}

template <typename T, int NT, int TX, int BS, bool PHASE>
__global__ void __launch_bounds__(BS) k_tline3(TLine3Args<T> a, int nx, int ny,
                                               const cplx<T> *__restrict__ tw, const int *status) {
    if (status && *status) return;
    __shared__ cplx<T> lds[3 * TimeTile<T, NT, TX>::FT];
    __shared__ cplx<T> twl[time_stage_twiddles<T, NT, TX>() ? 2 * NT : 1];
    nl_time_tile<T, NT, TX, BS, PHASE>(a, nx, ny, xcd_local_tile(blockIdx.x, gridDim.x), tw, lds, twl);
}

template <typename T, int NT, int TX, int BS>
__global__ void __launch_bounds__(BS) k_tline3p(TLine3Args<T> a, int nx, int ny,
                                                const cplx<T> *__restrict__ tw, const int *status) {
    if (status && *status) return;
    __shared__ cplx<T> lds[3 * TimeTile<T, NT, TX>::FT];
    __shared__ cplx<T> stab[8 * NT];
    nl_time_run<T, NT, TX, BS>(TimeSolo<T, NT, TX>{a, nx, ny, tw}, lds, stab);
}

// Compute units of the current device (persistent grids are sized by it).
static int cu_count() {
    static std::atomic<int> cached[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int n = cached[dev].load(std::memory_order_relaxed);
    if (n <= 0) {
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;
        cached[dev].store(n, std::memory_order_relaxed);
    }
    return n;
}

unsigned persistent_grid(int units, int per_cu) {
    const long long g = (long long)per_cu * cu_count();
    return (unsigned)(units < g ? units : g);
}

// ----------------------------------------------------- variant tables
constexpr int kSplitVariants = 3;   // xy variants of the two-pass form (xy_split_launch)

// Chosen variant per (precision, size): process-wide, read at every launch
// and written by the tuner, so the entries are atomics (a launch sees either
// the old or the new choice, both of which give the same results).
struct TuneTables {
    std::atomic<int> time[2][13];   // [fp64][log2 NT]
    std::atomic<int> xy[2][8];      // [fp64][NX index]
    TuneTables() {
        const int nxs[8] = {1, 16, 32, 64, 96, 128, 256, 512};
        for (int f = 0; f < 2; ++f) {
            for (int k = 0; k < 13; ++k) time[f][k].store(time_default_variant(f != 0, 1 << k));
            for (int k = 0; k < 8; ++k) xy[f][k].store(nxs[k] == 512 ? 0 : xy_default_variant(nxs[k]));
        }
    }
};
static TuneTables &tune_tables() {
    static TuneTables tt;   // initialised once, thread-safe
    return tt;
}

static int log2i(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return (1 << l) == n ? l : -1;
}
static int nx_index(int nx) {
    switch (nx) {
        case 1: return 0; case 16: return 1; case 32: return 2; case 64: return 3;
        case 96: return 4; case 128: return 5; case 256: return 6; case 512: return 7; default: return -1;
    }
}

template <typename T, int NT>
static void time_variant_info(int v, bool &valid, int &tx, int &bs, int &per_cu) {
    valid = false;
    static_for<0, kTimeVariants>([&](auto vv) {
        constexpr int V = decltype(vv)::value;
        using Vr = TimeVariant<T, NT, V>;
        if (V == v) {
            valid = Vr::valid;
            tx = Vr::tx;
            bs = Vr::bs;
            per_cu = Vr::per_cu;
        }
    });
}

template <typename T>
static bool time_info(int nt, int v, int &tx, int &bs, int &per_cu) {
    bool valid = false;
    switch (nt) {
#define TI(N) case N: time_variant_info<T, N>(v, valid, tx, bs, per_cu); break;
        TI(1) TI(16) TI(32) TI(64) TI(128) TI(256) TI(512) TI(1024) TI(2048)
        case 4096:
            if constexpr (sizeof(T) == 4) time_variant_info<T, 4096>(v, valid, tx, bs, per_cu);
            break;
#undef TI
        default: return false;
    }
    return valid;
}

template <typename T>
static bool time_info(int nt, int v, int &tx, int &bs) {
    int per_cu = 0;
    return time_info<T>(nt, v, tx, bs, per_cu);
}



int tune_variant_count(int axis) { return axis == 0 ? kTimeVariants : kXyVariants; }

bool tune_variant_valid(bool fp64, int axis, int n, int v) {
    if (axis == 0) {
        if (v < 0 || v >= kTimeVariants) return false;
        int tx = 0, bs = 0;
        return fp64 ? time_info<double>(n, v, tx, bs) : time_info<float>(n, v, tx, bs);
    }
    if (v < 0 || v >= kXyVariants || nx_index(n) < 0) return false;
    if (n < 16 || n == 512) return v == 0;   // 512: the two-pass form, one band shape
    if (n == 256 || (fp64 && n == 128)) return v < kSplitVariants;   // the two-pass form
    if (n <= 32) return v < 6;
    if (xy_variant_persistent(n, v)) return !fp64 && n * n / xy_variant_bs(n, v) <= 16;
    return true;
}

int tune_get(bool fp64, int axis, int n) {
    TuneTables &tt = tune_tables();
    if (axis == 0) {
        const int l = log2i(n);
        return (l >= 0 && l < 13) ? tt.time[fp64][l].load(std::memory_order_relaxed) : 0;
    }
    const int k = nx_index(n);
    return k >= 0 ? tt.xy[fp64][k].load(std::memory_order_relaxed) : 0;
}

bool tune_set(bool fp64, int axis, int n, int v) {
    if (!tune_variant_valid(fp64, axis, n, v)) return false;
    TuneTables &tt = tune_tables();
    if (axis == 0) tt.time[fp64][log2i(n)].store(v, std::memory_order_relaxed);
    else tt.xy[fp64][nx_index(n)].store(v, std::memory_order_relaxed);
    return true;
}

int tune_tile_px(bool fp64, int nt, int v) {
    int tx = 0, bs = 0;
    const bool ok = fp64 ? time_info<double>(nt, v, tx, bs) : time_info<float>(nt, v, tx, bs);
    return ok ? tx : 0;
}

int time_variant_for(bool fp64, int nt, int nx) {
    const int v = tune_get(fp64, 0, nt);
    const int tx = tune_tile_px(fp64, nt, v);
    return (tx > 0 && nx % tx == 0) ? v : 0;
}

int tune_per_cu(bool fp64, int axis, int n, int v) {
    if (axis == 1) return xy_variant_per_cu(n, v);
    int tx = 0, bs = 0, per_cu = 0;
    const bool ok = fp64 ? time_info<double>(n, v, tx, bs, per_cu) : time_info<float>(n, v, tx, bs, per_cu);
    return ok ? per_cu : 0;
}

int tune_block(bool fp64, int axis, int n, int v) {
    if (axis == 1) return (n == 256 || (fp64 && n == 128)) ? (v == 1 ? 512 : 256) : xy_variant_bs(n, v);
    int tx = 0, bs = 0;
    const bool ok = fp64 ? time_info<double>(n, v, tx, bs) : time_info<float>(n, v, tx, bs);
    return ok ? bs : 0;
}

template <typename T, int NT, int V>
static hipError_t tline3_launch_v(hipStream_t st, const Geometry &g, const TLine3Args<T> &a,
                                  const cplx<T> *tw, const int *status) {
    using Vr = TimeVariant<T, NT, V>;
    if constexpr (!Vr::valid) {
        return hipErrorInvalidValue;
    } else {
        if (g.nx % Vr::tx != 0) return hipErrorInvalidValue;
        const int units = (g.nx / Vr::tx) * g.ny;
        if constexpr (Vr::persistent) {
            if (a.rk.has_phase) return hipErrorInvalidValue;
            hipLaunchKernelGGL((k_tline3p<T, NT, Vr::tx, Vr::bs>), dim3(persistent_grid(units, Vr::per_cu)),
                               dim3(Vr::bs), 0, st, a, g.nx, g.ny, tw, status);
            return hipGetLastError();
        }
        dim3 grid((unsigned)units);
        if (a.rk.has_phase) {
            if constexpr (V == 0)
                hipLaunchKernelGGL((k_tline3<T, NT, Vr::tx, Vr::bs, true>), grid, dim3(Vr::bs), 0, st, a,
                                   g.nx, g.ny, tw, status);
            else
                return hipErrorInvalidValue;
        } else {
            hipLaunchKernelGGL((k_tline3<T, NT, Vr::tx, Vr::bs, false>), grid, dim3(Vr::bs), 0, st, a, g.nx,
                               g.ny, tw, status);
        }
        return hipGetLastError();
    }
}

template <typename T, int NT>
static hipError_t tline3_dispatch(hipStream_t st, const Geometry &g, const TLine3Args<T> &a,
                                  const cplx<T> *tw, const int *status) {
    if (g.nx == 1) {
        dim3 grid((unsigned)g.ny);
        if (a.rk.has_phase)
            hipLaunchKernelGGL((k_tline3<T, NT, 1, 256, true>), grid, dim3(256), 0, st, a, g.nx, g.ny, tw, status);
        else
            hipLaunchKernelGGL((k_tline3<T, NT, 1, 256, false>), grid, dim3(256), 0, st, a, g.nx, g.ny, tw, status);
        return hipGetLastError();
    }
    const int v = a.rk.has_phase ? 0 : time_variant_for(sizeof(T) == 8, NT, g.nx);
    hipError_t e = hipErrorInvalidValue;
    static_for<0, kTimeVariants>([&](auto vv) {
        constexpr int V = decltype(vv)::value;
        if (V == v) e = tline3_launch_v<T, NT, V>(st, g, a, tw, status);
    });
    return e;
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
        case 2048: return tline3_dispatch<T, 2048>(st, g, a, tw, status);
        case 4096:
            if constexpr (sizeof(T) == 4) return tline3_dispatch<T, 4096>(st, g, a, tw, status);
            else return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
}

// --------------------------------------------------------- xy operator
template <typename T, int NX, int BS, int P>
__global__ void __launch_bounds__(BS) k_xyplane(FieldOps<T> ops, int nt, const cplx<T> *__restrict__ tw,
                                                const int *status) {
    if (status && *status) return;
    __shared__ cplx<T> lds[P * NX * (NX + 1)];
    const int t = blockIdx.x * P, f = blockIdx.y;
    xy_plane<T, NX, BS, P>(ops.f[f] + (size_t)t * NX * NX, ops.tab[f], ops.taby[f], tw, lds);
}

template <typename T, int NX, int BS>
__global__ void __launch_bounds__(BS) k_xyplanep(FieldOps<T> ops, int nt, int nf,
                                                 const cplx<T> *__restrict__ tw, const int *status) {
    if (status && *status) return;
    __shared__ cplx<T> lds[NX * (NX + 1)];
    __shared__ cplx<T> stab[xy_uniform<NX, BS>() ? 1 : 14 * NX];
    xy_run<T, NX, BS>(XySolo<T, NX>{ops, nt, nf, tw}, lds, stab);
}

// FP64 planes of 128 x 128 (256 KiB) do not fit LDS, so their xy operator
// runs as two passes through HBM: Q_x on bands of RB consecutive rows, then
// Q_y on bands of CB consecutive columns (rows of the band are CB * 16 B
// contiguous in HBM).  Each line goes through exactly the operations of the
// one-plane kernel.
template <typename T, int NX, int RB, int BS>
__global__ void __launch_bounds__(BS) k_xrows(FieldOps<T> ops, int ny, const cplx<T> *__restrict__ tw,
                                              const int *status) {
    if (status && *status) return;
    constexpr int NXP = NX + 1;
    __shared__ cplx<T> lds[RB * NXP];
    const int bands = ny / RB;
    const int t = blockIdx.x / bands, r0 = (blockIdx.x % bands) * RB, f = blockIdx.y;
    cplx<T> *A = ops.f[f] + ((size_t)t * ny + r0) * NX;
    chunked_copy<T, RB * NX, BS>(threadIdx.x, [&](int i) { return A[i]; },
                                 [&](int i, cplx<T> v) { lds[(i / NX) * NXP + (i % NX)] = v; });
    __syncthreads();
    band_line_op<T, NX, -1, NXP, 1>(lds, RB, tw, ops.tab[f], threadIdx.x, BS);
    chunked_copy<T, RB * NX, BS>(threadIdx.x, [&](int i) { return lds[(i / NX) * NXP + (i % NX)]; },
                                 [&](int i, cplx<T> v) { A[i] = v; });
}

template <typename T, int NY, int CB, int BS>
__global__ void __launch_bounds__(BS) k_ycols(FieldOps<T> ops, int nx, const cplx<T> *__restrict__ tw,
                                              const int *status) {
    if (status && *status) return;
    constexpr int CBP = CB + 1;
    __shared__ cplx<T> lds[NY * CBP];
    const int bands = nx / CB;
    const int t = blockIdx.x / bands, c0 = (blockIdx.x % bands) * CB, f = blockIdx.y;
    cplx<T> *A = ops.f[f] + (size_t)t * NY * nx + c0;
    chunked_copy<T, NY * CB, BS>(threadIdx.x, [&](int i) { return A[(size_t)(i / CB) * nx + (i % CB)]; },
                                 [&](int i, cplx<T> v) { lds[(i / CB) * CBP + (i % CB)] = v; });
    __syncthreads();
    // lines are the CB columns (line c starts at c, its elements CBP apart)
    band_line_op<T, NY, -1, 1, CBP>(lds, CB, tw, ops.taby[f], threadIdx.x, BS);
    chunked_copy<T, NY * CB, BS>(threadIdx.x, [&](int i) { return lds[(i / CB) * CBP + (i % CB)]; },
                                 [&](int i, cplx<T> v) { A[(size_t)(i / CB) * nx + (i % CB)] = v; });
}

// Bands of 32 rows or columns: 256 line tasks per phase of a 128-point
// line operator (16 x 8 split) for 256 threads, 66 KiB of LDS (two
// workgroups per CU).
constexpr int kSplitBand = 32;

template <typename T, int NX, int B, int BS>
static hipError_t xy_split_launch_b(hipStream_t st, const Geometry &g, const FieldOps<T> &ops, int nf,
                                    const cplx<T> *tw, const int *status) {
    if (g.ny != NX || NX % B != 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_xrows<T, NX, B, BS>), dim3((unsigned)(g.nt * (NX / B)), (unsigned)nf), dim3(BS), 0, st,
                       ops, g.ny, tw, status);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_ycols<T, NX, B, BS>), dim3((unsigned)(g.nt * (NX / B)), (unsigned)nf), dim3(BS), 0, st,
                       ops, g.nx, tw, status);
    return hipGetLastError();
}
// Variants of the two-pass form (FP64 128 x 128 and all 256 x 256 planes): v = 0 bands of 32
// with 256 threads, 1 bands of 32 with 512, 2 bands of 16 with 256 (four
// workgroups per CU); kSplitVariants of them.
template <typename T, int NX>
static hipError_t xy_split_launch(hipStream_t st, const Geometry &g, const FieldOps<T> &ops, int nf,
                                  const cplx<T> *tw, const int *status, int v = 0) {
    if (v == 1) return xy_split_launch_b<T, NX, kSplitBand, 512>(st, g, ops, nf, tw, status);
    if (v == 2) return xy_split_launch_b<T, NX, 16, 256>(st, g, ops, nf, tw, status);
    return xy_split_launch_b<T, NX, kSplitBand, 256>(st, g, ops, nf, tw, status);
}

template <typename T, int NX, int BS, int P>
static hipError_t xy_launch_groups(hipStream_t st, const Geometry &g, const FieldOps<T> &ops, int nf,
                                   const cplx<T> *tw, const int *status) {
    if (g.nt % P != 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_xyplane<T, NX, BS, P>), dim3((unsigned)(g.nt / P), (unsigned)nf), dim3(BS), 0, st, ops,
                       g.nt, tw, status);
    return hipGetLastError();
}

template <typename T, int NX, int P>
static hipError_t xy_launch_bs(int bs, hipStream_t st, const Geometry &g, const FieldOps<T> &ops, int nf,
                               const cplx<T> *tw, const int *status) {
    if (bs == 256) return xy_launch_groups<T, NX, 256, P>(st, g, ops, nf, tw, status);
    if (bs == 512) return xy_launch_groups<T, NX, 512, P>(st, g, ops, nf, tw, status);
    return xy_launch_groups<T, NX, 1024, P>(st, g, ops, nf, tw, status);
}

template <typename T, int NX>
static hipError_t xy_dispatch(hipStream_t st, const Geometry &g, const FieldOps<T> &ops, int nf,
                              const cplx<T> *tw, const int *status) {
    if constexpr (NX < 16) {
        return xy_launch_groups<T, NX, 256, 1>(st, g, ops, nf, tw, status);
    } else {
        const int v = tune_get(sizeof(T) == 8, 1, NX);
        const int bs = xy_variant_bs(NX, v);
        if constexpr (NX >= 64) {
            if constexpr (sizeof(T) == 4) {   // persistent planes: FP32 only
                if (xy_variant_persistent(NX, v)) {
                    const dim3 pg(persistent_grid(g.nt * nf, xy_variant_per_cu(NX, v)));
                    if (bs == 256) hipLaunchKernelGGL((k_xyplanep<T, NX, 256>), pg, dim3(256), 0, st, ops, g.nt, nf, tw, status);
                    else if (bs == 512) hipLaunchKernelGGL((k_xyplanep<T, NX, 512>), pg, dim3(512), 0, st, ops, g.nt, nf, tw, status);
                    else hipLaunchKernelGGL((k_xyplanep<T, NX, 1024>), pg, dim3(1024), 0, st, ops, g.nt, nf, tw, status);
                    return hipGetLastError();
                }
            }
            return xy_launch_bs<T, NX, 1>(bs, st, g, ops, nf, tw, status);
        } else {
            if (xy_group(NX, v) > 1) return xy_launch_bs<T, NX, xy_group(NX, 3)>(bs, st, g, ops, nf, tw, status);
            return xy_launch_bs<T, NX, 1>(bs, st, g, ops, nf, tw, status);
        }
    }
}

static bool line_len_supported(int n) {
    switch (n) {
        case 16: case 32: case 64: case 96: case 128: case 256: case 512: return true;
        default: return false;
    }
}

bool xyplane_supported(int nx, int ny, bool fp64) {
    if (nx != ny) return line_len_supported(nx) && line_len_supported(ny);
    return nx == 1 || line_len_supported(nx);
}

// NX != NY: Q_x on bands of rows, then Q_y on bands of columns, each band
// in LDS (k_xrows, k_ycols); bands of 32 rows or columns where the other
// dimension allows, else 16 (every supported length is a multiple of 16).
// 256- and 512-point lines take the fastest square-plane band shapes:
// FP32 bands of 32 on 512 threads (256: one 16-point DFT per thread and
// phase; 512: 131 KiB of LDS, as fast as bands of 16 on 256 threads), FP64
// bands of 16 (256: 64 KiB of LDS, two workgroups per CU; 512: 131 KiB).
template <typename T, int N> constexpr int max_band() {
    return (sizeof(T) == 8 && N >= 256) ? 16 : 32;
}
template <typename T, int N> constexpr int band32_threads() { return (sizeof(T) == 4 && N >= 256) ? 512 : 256; }

template <typename T, int NX>
static hipError_t rows_launch(hipStream_t st, const Geometry &g, const FieldOps<T> &ops, int nf,
                              const cplx<T> *tw, const int *status) {
    if constexpr (max_band<T, NX>() == 32) {
        constexpr int B32 = band32_threads<T, NX>();
        if (g.ny % 32 == 0) {
            hipLaunchKernelGGL((k_xrows<T, NX, 32, B32>), dim3((unsigned)(g.nt * (g.ny / 32)), (unsigned)nf),
                               dim3(B32), 0, st, ops, g.ny, tw, status);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((k_xrows<T, NX, 16, 256>), dim3((unsigned)(g.nt * (g.ny / 16)), (unsigned)nf), dim3(256),
                       0, st, ops, g.ny, tw, status);
    return hipGetLastError();
}
template <typename T, int NY>
static hipError_t cols_launch(hipStream_t st, const Geometry &g, const FieldOps<T> &ops, int nf,
                              const cplx<T> *tw, const int *status) {
    if constexpr (max_band<T, NY>() == 32) {
        constexpr int B32 = band32_threads<T, NY>();
        if (g.nx % 32 == 0) {
            hipLaunchKernelGGL((k_ycols<T, NY, 32, B32>), dim3((unsigned)(g.nt * (g.nx / 32)), (unsigned)nf),
                               dim3(B32), 0, st, ops, g.nx, tw, status);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((k_ycols<T, NY, 16, 256>), dim3((unsigned)(g.nt * (g.nx / 16)), (unsigned)nf), dim3(256),
                       0, st, ops, g.nx, tw, status);
    return hipGetLastError();
}
template <typename T>
static hipError_t xy_rect_launch(hipStream_t st, const Geometry &g, const FieldOps<T> &ops, int nf,
                                 const cplx<T> *tw_x, const cplx<T> *tw_y, const int *status) {
    if (!line_len_supported(g.nx) || !line_len_supported(g.ny) || g.nx % 16 || g.ny % 16)
        return hipErrorInvalidValue;
    hipError_t e = hipErrorInvalidValue;
    switch (g.nx) {
        case 16: e = rows_launch<T, 16>(st, g, ops, nf, tw_x, status); break;
        case 32: e = rows_launch<T, 32>(st, g, ops, nf, tw_x, status); break;
        case 64: e = rows_launch<T, 64>(st, g, ops, nf, tw_x, status); break;
        case 96: e = rows_launch<T, 96>(st, g, ops, nf, tw_x, status); break;
        case 128: e = rows_launch<T, 128>(st, g, ops, nf, tw_x, status); break;
        case 256: e = rows_launch<T, 256>(st, g, ops, nf, tw_x, status); break;
        case 512: e = rows_launch<T, 512>(st, g, ops, nf, tw_x, status); break;
    }
    if (e != hipSuccess) return e;
    switch (g.ny) {
        case 16: return cols_launch<T, 16>(st, g, ops, nf, tw_y, status);
        case 32: return cols_launch<T, 32>(st, g, ops, nf, tw_y, status);
        case 64: return cols_launch<T, 64>(st, g, ops, nf, tw_y, status);
        case 96: return cols_launch<T, 96>(st, g, ops, nf, tw_y, status);
        case 128: return cols_launch<T, 128>(st, g, ops, nf, tw_y, status);
        case 256: return cols_launch<T, 256>(st, g, ops, nf, tw_y, status);
        case 512: return cols_launch<T, 512>(st, g, ops, nf, tw_y, status);
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t launch_xyplane(hipStream_t st, const Geometry &g, const FieldOps<T> &ops, int nf,
                          const cplx<T> *tw, const cplx<T> *tw_y, const int *status) {
    if (g.nx != g.ny) return xy_rect_launch<T>(st, g, ops, nf, tw, tw_y, status);
    switch (g.nx) {
        case 1: return xy_dispatch<T, 1>(st, g, ops, nf, tw, status);
        case 16: return xy_dispatch<T, 16>(st, g, ops, nf, tw, status);
        case 32: return xy_dispatch<T, 32>(st, g, ops, nf, tw, status);
        case 64: return xy_dispatch<T, 64>(st, g, ops, nf, tw, status);
        case 96: return xy_dispatch<T, 96>(st, g, ops, nf, tw, status);
        case 128:
            if constexpr (sizeof(T) == 4) return xy_dispatch<T, 128>(st, g, ops, nf, tw, status);
            else return xy_split_launch<T, 128>(st, g, ops, nf, tw, status, tune_get(true, 1, 128));
        case 256: return xy_split_launch<T, 256>(st, g, ops, nf, tw, status, tune_get(sizeof(T) == 8, 1, 256));
        case 512: return xy_rect_launch<T>(st, g, ops, nf, tw, tw_y, status);   // bands of 16, one shape
        default: return hipErrorInvalidValue;
    }
}

// --------------------------------------------------------------- power
template <typename T>
__global__ void __launch_bounds__(256) k_power_partial(const cplx<T> *f0, const cplx<T> *f1,
                                                       const cplx<T> *f2, size_t n, double *partial,
                                                       const int *status) {
    if (status && *status) return;
    __shared__ double red[3][256];
    power_partial_block<T>(f0, f1, f2, n, blockIdx.x, gridDim.x, red, partial + blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(256) k_power_final(const double *partial, int nblk, double pref0,
                                                     double pref1, double pref2, double *series,
                                                     int *trip, int *status, int check_nan,
                                                     int series_cap) {
    if (status && *status) return;
    __shared__ double red[3][256];
    power_final_sum(partial, nblk, nblk, red);
    if (threadIdx.x == 0) {
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

// Power of one pixel's time series: pref * sum_t |A(x, y, t)|^2, one
// workgroup, Neumaier-compensated per thread and reduced in a fixed tree
// order (averagePowerTime, Solver3D.cuh:509-523).
template <typename T>
__global__ void __launch_bounds__(256) k_power_pixel(const cplx<T> *f, size_t plane, int nt,
                                                     double pref, double *out) {
    __shared__ double red[256];
    const int tid = threadIdx.x;
    double s = 0.0, c = 0.0;
    for (int t = tid; t < nt; t += 256) {
        const cplx<T> a = f[(size_t)t * plane];
        const double re = (double)a.x, im = (double)a.y;
        neumaier_add(s, c, fma(re, re, im * im));
    }
    red[tid] = s + c;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
    if (tid == 0) out[0] = pref * red[0];
}

template <typename T>
hipError_t launch_power_pixel(hipStream_t st, const cplx<T> *f, size_t plane, int nt, double pref,
                              double *out) {
    hipLaunchKernelGGL((k_power_pixel<T>), dim3(1), dim3(256), 0, st, f, plane, nt, pref, out);
    return hipGetLastError();
}

// ------------------------------------------------------- instantiation
#define OPO_INST(T)                                                                               \
    template hipError_t launch_rk4<T>(hipStream_t, const cplx<T> *, cplx<T> *, cplx<T> *,          \
                                      cplx<T> *, size_t, size_t, const Rk4Args<T> &, const int *); \
    template hipError_t launch_tline<T>(hipStream_t, const Geometry &, const FieldOps<T> &, int,   \
                                        const cplx<T> *, const int *);                            \
    template hipError_t launch_tline3<T>(hipStream_t, const Geometry &, const TLine3Args<T> &,      \
                                         const cplx<T> *, const int *);                           \
    template hipError_t launch_xyplane<T>(hipStream_t, const Geometry &, const FieldOps<T> &, int, \
                                          const cplx<T> *, const cplx<T> *, const int *);         \
    template hipError_t launch_power_partial<T>(hipStream_t, const cplx<T> *, const cplx<T> *,     \
                                                const cplx<T> *, size_t, int, double *, const int *); \
    template hipError_t launch_power_pixel<T>(hipStream_t, const cplx<T> *, size_t, int, double,   \
                                              double *);
OPO_INST(float)
OPO_INST(double)

}  // namespace opo
