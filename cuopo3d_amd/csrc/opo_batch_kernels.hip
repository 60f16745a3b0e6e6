// opo_batch_kernels.hip -- split-step passes over many solves at once.
//
// A batch holds solves of different grids resident in HBM.  Each pass is
// one launch per size class (the NT of the along-t pass, the NX of the xy
// pass) whose workgroups walk a work list of (solve, tile) items, so small
// solves share the GPU instead of each under-filling it.  Every solve's
// arithmetic (tiles, line operators, reduction order) is the same as when it
// runs alone, so its results do not depend on its neighbours.
#include "opo_device.hpp"
#include "opo_kernels.hpp"
#include "opo_passes.hpp"
#include "opo_rk4.hpp"
#include "opo_power.hpp"

namespace opo {

template <typename T>
__device__ __forceinline__ Rk4Args<T> batch_rk(const SolveDesc<T> &d, int mode) {
    Rk4Args<T> r{};
    for (int f = 0; f < 3; ++f) {
        r.kc1[f] = d.kc1[f];
        r.kc3[f] = d.kc3[f];
        r.kh6[f] = d.kh6[f];
    }
    r.nsteps = mode == BATCH_MID ? 2 : 1;
    r.has_phase = 0;
    r.out_scale = mode == BATCH_LAST ? 1 : 0;
    r.ms_re = d.ms_re;
    r.ms_im = d.ms_im;
    r.ms_lo_re = d.ms_lo_re;
    r.ms_lo_im = d.ms_lo_im;
    r.mi_re = d.mi_re;
    r.mi_im = d.mi_im;
    r.mi_lo_re = d.mi_lo_re;
    r.mi_lo_im = d.mi_lo_im;
    r.eom = mode == BATCH_LAST ? d.eom : nullptr;
    return r;
}

// Arguments of the fused pass of one solve in a batch (mode: BATCH_FIRST,
// BATCH_MID or BATCH_LAST).
template <typename T, int NT>
__device__ __forceinline__ TLine3Args<T> batch_time_args(const SolveDesc<T> &d, int mode) {
    TLine3Args<T> a;
    a.psrc = mode == BATCH_FIRST ? d.api : d.f[0];
    a.f[0] = d.f[0];
    a.f[1] = d.f[1];
    a.f[2] = d.f[2];
    if (mode == BATCH_LAST) {
        a.tab0 = d.tabG;
        a.tab_stride = 0;
        a.first_field = d.gdd_on ? 1 : 3;
    } else {
        a.tab0 = d.tabT;
        a.tab_stride = 2 * NT;
        a.first_field = 0;
    }
    a.rk = batch_rk(d, mode);
    return a;
}

template <typename T, int NT, int TX, int BS>
__global__ void __launch_bounds__(BS) kb_time(const SolveDesc<T> *__restrict__ sd,
                                              const int *__restrict__ status,
                                              const WorkItem *__restrict__ work, int mode) {
    __shared__ cplx<T> lds[3 * TimeTile<T, NT, TX>::FT];
    __shared__ cplx<T> twl[time_stage_twiddles<T, NT, TX>() ? 2 * NT : 1];
    const WorkItem w = work[xcd_local_tile(blockIdx.x, gridDim.x)];
    if (status[w.solve]) return;
    const SolveDesc<T> &d = sd[w.solve];
    nl_time_tile<T, NT, TX, BS, false>(batch_time_args<T, NT>(d, mode), d.nx, d.ny, w.item, d.twt, lds, twl);
}

// Work units of the fused pass over a work list: tile w.item of solve w.solve.
template <typename T, int NT, int TX> struct TimeBatch {
    const SolveDesc<T> *sd;
    const int *status;
    const WorkItem *work;
    int n, mode;
    __device__ int units() const { return n; }
    __device__ bool skip(int u) const { return status[work[u].solve] != 0; }
    __device__ int key(int u) const { return work[u].solve; }
    __device__ int nx(int u) const { return sd[work[u].solve].nx; }
    __device__ int ny(int u) const { return sd[work[u].solve].ny; }
    __device__ int tile(int u) const { return work[u].item; }
    __device__ TLine3Args<T> args(int u) const { return batch_time_args<T, NT>(sd[work[u].solve], mode); }
    __device__ void stage(int u, cplx<T> *stab, int tid, int nthr) const {
        stage_time_tables<T, NT>(sd[work[u].solve].twt, args(u), stab, tid, nthr);
    }
};

template <typename T, int NT, int TX, int BS>
__global__ void __launch_bounds__(BS) kb_timep(const SolveDesc<T> *__restrict__ sd,
                                               const int *__restrict__ status,
                                               const WorkItem *__restrict__ work, int n_work, int mode) {
    __shared__ cplx<T> lds[3 * TimeTile<T, NT, TX>::FT];
    __shared__ cplx<T> stab[8 * NT];
    nl_time_run<T, NT, TX, BS>(TimeBatch<T, NT, TX>{sd, status, work, n_work, mode}, lds, stab);
}

// Work units of the xy pass over a work list: item = k nt + t, plane t of
// field k (or of s and i with the free-space tables).
template <typename T, int NX> struct XyBatch {
    const SolveDesc<T> *sd;
    const TripState *ts;
    const int *status;
    const WorkItem *work;
    int n, free_space;
    __device__ int units() const { return n; }
    __device__ bool skip(int u) const {
        const int s = work[u].solve;
        return status[s] != 0 || (free_space && !ts[s].fs);
    }
    __device__ int key(int u) const { return work[u].solve; }
    __device__ int slot(int u) const { return work[u].item / sd[work[u].solve].nt; }
    __device__ cplx<T> *plane(int u) const {
        const SolveDesc<T> &d = sd[work[u].solve];
        const int k = work[u].item / d.nt, t = work[u].item % d.nt;
        const int f = free_space ? 1 + k : k;
        return (f == 0 ? d.f[0] : (f == 1 ? d.f[1] : d.f[2])) + (size_t)t * NX * NX;
    }
    __device__ void stage(int u, cplx<T> *stab, int tid, int nthr) const {
        const SolveDesc<T> &d = sd[work[u].solve];
        const cplx<T> *bx = free_space ? d.qfsx : d.qx, *by = free_space ? d.qfsy : d.qy;
        const cplx<T> *tx[3] = {bx, bx + 2 * NX, bx + 4 * NX};
        const cplx<T> *ty[3] = {by, by + 2 * NX, by + 4 * NX};
        stage_xy_tables<T, NX>(d.twx, tx, ty, free_space ? 2 : 3, stab, tid, nthr);
    }
    __device__ void tables(int u, const cplx<T> *&tw, const cplx<T> *&tx, const cplx<T> *&ty) const {
        const SolveDesc<T> &d = sd[work[u].solve];
        const int k = work[u].item / d.nt;
        tw = d.twx;
        tx = (free_space ? d.qfsx : d.qx) + 2 * NX * k;
        ty = (free_space ? d.qfsy : d.qy) + 2 * NX * k;
    }
};

template <typename T, int NX, int BS>
__global__ void __launch_bounds__(BS) kb_xyp(const SolveDesc<T> *__restrict__ sd,
                                             const TripState *__restrict__ ts,
                                             const int *__restrict__ status,
                                             const WorkItem *__restrict__ work, int n_work, int free_space) {
    __shared__ cplx<T> lds[NX * (NX + 1)];
    __shared__ cplx<T> stab[xy_uniform<NX, BS>() ? 1 : 14 * NX];
    xy_run<T, NX, BS>(XyBatch<T, NX>{sd, ts, status, work, n_work, free_space}, lds, stab);
}

template <typename T, int NX, int BS, int P>
__global__ void __launch_bounds__(BS) kb_xy(const SolveDesc<T> *__restrict__ sd,
                                            const TripState *__restrict__ ts,
                                            const int *__restrict__ status,
                                            const WorkItem *__restrict__ work, int free_space) {
    __shared__ cplx<T> lds[P * NX * (NX + 1)];
    const WorkItem w = work[xcd_local_tile(blockIdx.x, gridDim.x)];
    if (status[w.solve]) return;
    if (free_space && !ts[w.solve].fs) return;
    const SolveDesc<T> &d = sd[w.solve];
    const int groups = d.nt / P;   // item = k * groups + group of P planes
    const int k = w.item / groups, t = (w.item % groups) * P;
    const int f = free_space ? 1 + k : k;
    const cplx<T> *tx = free_space ? d.qfsx + 2 * NX * k : d.qx + 2 * NX * k;
    const cplx<T> *ty = free_space ? d.qfsy + 2 * NX * k : d.qy + 2 * NX * k;
    cplx<T> *A = (f == 0 ? d.f[0] : (f == 1 ? d.f[1] : d.f[2])) + (size_t)t * NX * NX;
    xy_plane<T, NX, BS, P>(A, tx, ty, d.twx, lds);
}

// FP64 128 x 128 and all 256 x 256 and 512 x 512 planes (beyond the LDS): Q_x on bands of RB rows,
// then (a second launch over the same work list) Q_y on bands of RB
// columns, each workgroup walking the bands of its plane; every line gets
// the operations of the single-solve row and column passes, so the fields
// are bitwise those of the solve alone.
template <typename T, int N, int RB, int BS, bool ROWS>
__global__ void __launch_bounds__(BS) kb_xyband(const SolveDesc<T> *__restrict__ sd,
                                                const TripState *__restrict__ ts,
                                                const int *__restrict__ status,
                                                const WorkItem *__restrict__ work, int free_space) {
    constexpr int BP = RB + 1;
    __shared__ cplx<T> lds[ROWS ? RB * (N + 1) : N * BP];
    const WorkItem w = work[xcd_local_tile(blockIdx.x, gridDim.x)];
    if (status[w.solve]) return;
    if (free_space && !ts[w.solve].fs) return;
    const SolveDesc<T> &d = sd[w.solve];
    const int k = w.item / d.nt, t = w.item % d.nt;
    const int f = free_space ? 1 + k : k;
    const cplx<T> *tab = ROWS ? (free_space ? d.qfsx : d.qx) + 2 * N * k : (free_space ? d.qfsy : d.qy) + 2 * N * k;
    cplx<T> *A = (f == 0 ? d.f[0] : (f == 1 ? d.f[1] : d.f[2])) + (size_t)t * N * N;
    for (int b0 = 0; b0 < N; b0 += RB) {
        if constexpr (ROWS) {   // rows b0 .. b0 + RB - 1, pitch N + 1
            cplx<T> *R = A + (size_t)b0 * N;
            chunked_copy<T, RB * N, BS>(threadIdx.x, [&](int i) { return gld_c(R + i); },
                                        [&](int i, cplx<T> v) { lds[(i / N) * (N + 1) + i % N] = v; });
            __syncthreads();
            band_line_op<T, N, -1, N + 1, 1>(lds, RB, d.twx, tab, threadIdx.x, BS);
            chunked_copy<T, RB * N, BS>(threadIdx.x, [&](int i) { return lds[(i / N) * (N + 1) + i % N]; },
                                        [&](int i, cplx<T> v) { gst_c(R + i, v); });
        } else {                // columns b0 .. b0 + RB - 1, each line down a column
            cplx<T> *Cb = A + b0;
            chunked_copy<T, N * RB, BS>(threadIdx.x, [&](int i) { return gld_c(Cb + (size_t)(i / RB) * N + i % RB); },
                                        [&](int i, cplx<T> v) { lds[(i / RB) * BP + i % RB] = v; });
            __syncthreads();
            band_line_op<T, N, -1, 1, BP>(lds, RB, d.twx, tab, threadIdx.x, BS);
            chunked_copy<T, N * RB, BS>(threadIdx.x, [&](int i) { return lds[(i / RB) * BP + i % RB]; },
                                        [&](int i, cplx<T> v) { gst_c(Cb + (size_t)(i / RB) * N + i % RB, v); });
        }
        __syncthreads();
    }
}

template <typename T, int N, int RB, int BS>
static hipError_t bx_band_launch(hipStream_t st, int n_work, const SolveDesc<T> *sd, const TripState *ts,
                                 const int *status, const WorkItem *work, int fs) {
    hipLaunchKernelGGL((kb_xyband<T, N, RB, BS, true>), dim3((unsigned)n_work), dim3(BS), 0, st, sd, ts, status,
                       work, fs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((kb_xyband<T, N, RB, BS, false>), dim3((unsigned)n_work), dim3(BS), 0, st, sd, ts, status,
                       work, fs);
    return hipGetLastError();
}
// The band shapes of the single-solve two-pass variants (FP64 128 x 128 and
// all 256 x 256 planes): 1 bands of 32 on 512 threads, 2 bands of 16 on
// 256, else bands of 32 on 256.
template <typename T, int N>
static hipError_t bx_band_variant(hipStream_t st, int n_work, const SolveDesc<T> *sd, const TripState *ts,
                                  const int *status, const WorkItem *work, int fs, int variant) {
    if (variant == 1) return bx_band_launch<T, N, 32, 512>(st, n_work, sd, ts, status, work, fs);
    if (variant == 2) return bx_band_launch<T, N, 16, 256>(st, n_work, sd, ts, status, work, fs);
    return bx_band_launch<T, N, 32, 256>(st, n_work, sd, ts, status, work, fs);
}

// Power of every resident solve, in the order a solve uses alone
// (opo_power.hpp): work item (solve, b) is partial block b of the solve's
// power_blocks(cells) blocks, then one 256-thread workgroup per solve.
template <typename T>
__global__ void __launch_bounds__(256) kb_power_partial(const SolveDesc<T> *__restrict__ sd,
                                                        const int *__restrict__ status,
                                                        const WorkItem *__restrict__ work,
                                                        double *partial, int max_chunk) {
    __shared__ double red[3][256];
    const WorkItem w = work[blockIdx.x];
    if (status[w.solve]) return;
    const SolveDesc<T> &d = sd[w.solve];
    const size_t n = (size_t)d.nx * d.ny * d.nt;
    power_partial_block<T>(d.f[0], d.f[1], d.f[2], n, w.item, d.nchunk, red,
                           partial + (size_t)w.solve * 3 * max_chunk + w.item, max_chunk);
}

template <typename T>
__global__ void __launch_bounds__(256) kb_power_final(const SolveDesc<T> *__restrict__ sd,
                                                      const TripState *__restrict__ ts, int *status,
                                                      const double *partial, int max_chunk) {
    __shared__ double red[3][256];
    const int s = blockIdx.x;
    if (status[s]) return;
    const SolveDesc<T> &d = sd[s];
    power_final_sum(partial + (size_t)s * 3 * max_chunk, max_chunk, d.nchunk, red);
    if (threadIdx.x == 0) {
        double pw[3];
        for (int f = 0; f < 3; ++f) pw[f] = d.pref[f] * red[f][0];
        const int k = ts[s].trip;
        if (k < d.series_cap)
            for (int f = 0; f < 3; ++f) d.series[3 * (size_t)k + f] = pw[f];
        if (!isfinite(pw[1])) status[s] = k + 1;
    }
}

bool batch_supported(int nx, int ny, int nt, bool fp64) {
    if (nx != ny || !tline_supported(nt, fp64) || nt < 16) return false;
    switch (nx) {
        case 1: case 16: case 32: case 64: case 96: case 128: case 256: case 512: return true;
        default: return false;
    }
}

template <typename T, int NT>
static hipError_t bt_dispatch(hipStream_t st, int n_work, const SolveDesc<T> *sd, const int *status,
                              const WorkItem *work, int mode, int v) {
    if (v == kBatchPlaneVariant) {   // plane-wave solves: one pixel per workgroup
        hipLaunchKernelGGL((kb_time<T, NT, 1, 256>), dim3((unsigned)n_work), dim3(256), 0, st, sd, status, work,
                           mode);
        return hipGetLastError();
    }
    hipError_t e = hipErrorInvalidValue;
    static_for<0, kTimeVariants>([&](auto vv) {
        constexpr int V = decltype(vv)::value;
        using Vr = TimeVariant<T, NT, V>;
        if constexpr (Vr::valid) {
            if (V == v) {
                if constexpr (Vr::persistent)
                    hipLaunchKernelGGL((kb_timep<T, NT, Vr::tx, Vr::bs>), dim3(persistent_grid(n_work, Vr::per_cu)),
                                       dim3(Vr::bs), 0, st, sd, status, work, n_work, mode);
                else
                    hipLaunchKernelGGL((kb_time<T, NT, Vr::tx, Vr::bs>), dim3((unsigned)n_work), dim3(Vr::bs), 0,
                                       st, sd, status, work, mode);
                e = hipGetLastError();
            }
        }
    });
    return e;
}

template <typename T>
hipError_t launch_batch_time(hipStream_t st, int nt, int n_work, const SolveDesc<T> *sd,
                             const int *status, const WorkItem *work, int mode, int variant) {
    if (n_work <= 0) return hipSuccess;
    switch (nt) {
        case 16: return bt_dispatch<T, 16>(st, n_work, sd, status, work, mode, variant);
        case 32: return bt_dispatch<T, 32>(st, n_work, sd, status, work, mode, variant);
        case 64: return bt_dispatch<T, 64>(st, n_work, sd, status, work, mode, variant);
        case 128: return bt_dispatch<T, 128>(st, n_work, sd, status, work, mode, variant);
        case 256: return bt_dispatch<T, 256>(st, n_work, sd, status, work, mode, variant);
        case 512: return bt_dispatch<T, 512>(st, n_work, sd, status, work, mode, variant);
        case 1024: return bt_dispatch<T, 1024>(st, n_work, sd, status, work, mode, variant);
        case 2048: return bt_dispatch<T, 2048>(st, n_work, sd, status, work, mode, variant);
        case 4096:
            if constexpr (sizeof(T) == 4) return bt_dispatch<T, 4096>(st, n_work, sd, status, work, mode, variant);
            else return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
}

template <typename T, int NX, int P>
static hipError_t bx_launch_bs(int bs, hipStream_t st, int n_work, const SolveDesc<T> *sd, const TripState *ts,
                               const int *status, const WorkItem *work, int fs) {
    const dim3 g((unsigned)n_work);
    if (bs == 256) hipLaunchKernelGGL((kb_xy<T, NX, 256, P>), g, dim3(256), 0, st, sd, ts, status, work, fs);
    else if (bs == 512) hipLaunchKernelGGL((kb_xy<T, NX, 512, P>), g, dim3(512), 0, st, sd, ts, status, work, fs);
    else hipLaunchKernelGGL((kb_xy<T, NX, 1024, P>), g, dim3(1024), 0, st, sd, ts, status, work, fs);
    return hipGetLastError();
}

template <typename T, int NX>
static hipError_t bx_dispatch(hipStream_t st, int n_work, const SolveDesc<T> *sd, const TripState *ts,
                              const int *status, const WorkItem *work, int fs, int v) {
    if (!tune_variant_valid(sizeof(T) == 8, 1, NX, v)) return hipErrorInvalidValue;
    const int bs = xy_variant_bs(NX, v);
    if constexpr (NX >= 64) {
        if constexpr (sizeof(T) == 4) {   // persistent planes: FP32 only
            if (xy_variant_persistent(NX, v)) {
                const dim3 pg(persistent_grid(n_work, xy_variant_per_cu(NX, v)));
                if (bs == 256) hipLaunchKernelGGL((kb_xyp<T, NX, 256>), pg, dim3(256), 0, st, sd, ts, status, work, n_work, fs);
                else if (bs == 512) hipLaunchKernelGGL((kb_xyp<T, NX, 512>), pg, dim3(512), 0, st, sd, ts, status, work, n_work, fs);
                else hipLaunchKernelGGL((kb_xyp<T, NX, 1024>), pg, dim3(1024), 0, st, sd, ts, status, work, n_work, fs);
                return hipGetLastError();
            }
        }
        return bx_launch_bs<T, NX, 1>(bs, st, n_work, sd, ts, status, work, fs);
    } else {
        if (xy_group(NX, v) > 1) return bx_launch_bs<T, NX, xy_group(NX, 3)>(bs, st, n_work, sd, ts, status, work, fs);
        return bx_launch_bs<T, NX, 1>(bs, st, n_work, sd, ts, status, work, fs);
    }
}

template <typename T>
hipError_t launch_batch_xy(hipStream_t st, int nx, int n_work, const SolveDesc<T> *sd,
                           const TripState *ts, const int *status, const WorkItem *work,
                           int free_space, int variant) {
    if (n_work <= 0) return hipSuccess;
    switch (nx) {
        case 16: return bx_dispatch<T, 16>(st, n_work, sd, ts, status, work, free_space, variant);
        case 32: return bx_dispatch<T, 32>(st, n_work, sd, ts, status, work, free_space, variant);
        case 64: return bx_dispatch<T, 64>(st, n_work, sd, ts, status, work, free_space, variant);
        case 96: return bx_dispatch<T, 96>(st, n_work, sd, ts, status, work, free_space, variant);
        case 128:
            if constexpr (sizeof(T) == 4) return bx_dispatch<T, 128>(st, n_work, sd, ts, status, work, free_space, variant);
            else return bx_band_variant<T, 128>(st, n_work, sd, ts, status, work, free_space, variant);
        case 256: return bx_band_variant<T, 256>(st, n_work, sd, ts, status, work, free_space, variant);
        case 512: return bx_band_launch<T, 512, 16, 256>(st, n_work, sd, ts, status, work, free_space);
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t launch_batch_power(hipStream_t st, int n_work, int n_solves, const SolveDesc<T> *sd,
                              const TripState *ts, int *status, const WorkItem *work,
                              double *partial, int max_chunk) {
    if (n_work <= 0) return hipSuccess;
    hipLaunchKernelGGL((kb_power_partial<T>), dim3((unsigned)n_work), dim3(256), 0, st, sd, status, work,
                       partial, max_chunk);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((kb_power_final<T>), dim3((unsigned)n_solves), dim3(256), 0, st, sd, ts, status,
                       partial, max_chunk);
    return hipGetLastError();
}

#define OPO_BATCH_INST(T)                                                                            \
    template hipError_t launch_batch_time<T>(hipStream_t, int, int, const SolveDesc<T> *, const int *, \
                                             const WorkItem *, int, int);                            \
    template hipError_t launch_batch_xy<T>(hipStream_t, int, int, const SolveDesc<T> *,              \
                                           const TripState *, const int *, const WorkItem *, int,    \
                                           int);                                                     \
    template hipError_t launch_batch_power<T>(hipStream_t, int, int, const SolveDesc<T> *,           \
                                              const TripState *, int *, const WorkItem *, double *,  \
                                              int);
OPO_BATCH_INST(float)
OPO_BATCH_INST(double)

}  // namespace opo
