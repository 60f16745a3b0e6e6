// opo_passes.hpp -- the two passes of a z-slice as device functions, shared
// by the single-solve kernels and the batched (many-solve) kernels.
//
//   nl_time_tile  one tile of the fused pass: pointwise RK4 half step(s) of
//                 the three fields, then the along-t line operator
//                 (dispersion/walk-off/absorption, or GDD at a trip's end);
//   xy_plane      Q_x then Q_y on one t-plane of one field.
#pragma once
#include "opo_device.hpp"
#include "opo_kernels.hpp"
#include "opo_rk4.hpp"

namespace opo {

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

// Launch variants of the fused pass for one NT (chosen by the autotuner):
// V = 2 m + b with tile width m = 0: TXd, 1: TXd/2, 2: 2 TXd pixels and
// b = 0: 256, 1: 512 threads.  Only variant 0 exists for FP64 and for
// NT < 64; a variant whose three tiles exceed 96 KiB of LDS does not exist.
constexpr int kTimeVariants = 6;
template <typename T, int NT, int V> struct TimeVariant {
    static constexpr int txd = tline3_tx<T, NT>();
    static constexpr int m = V / 2, b = V % 2;
    static constexpr int tx = m == 0 ? txd : (m == 1 ? txd / 2 : txd * 2);
    static constexpr int bs = b == 0 ? 256 : 512;
    static constexpr bool valid =
        V == 0 || (sizeof(T) == 4 && NT >= 64 && tx >= 2 && tx <= 32 &&
                   3 * NT * tx * (int)sizeof(cplx<T>) <= 98304);
};

// Block-size variants of the xy pass: 256, 512 or 1024 threads.
constexpr int kXyVariants = 3;
constexpr int xy_variant_bs(int v) { return v == 0 ? 256 : (v == 1 ? 512 : 1024); }
constexpr int xy_default_variant(int nx) { return nx >= 96 ? 1 : 0; }

template <typename T> struct alignas(4 * sizeof(T)) cpair {
    cplx<T> a, b;
};

// One tile of the fused pass: TX consecutive pixels of row y, all NT time
// samples, all three fields.  lds holds 3 * NT * TX elements.
template <typename T, int NT, int TX, int BS, bool PHASE>
__device__ __forceinline__ void nl_time_tile(const TLine3Args<T> &a, int nx, int ny, int tile,
                                             const cplx<T> *__restrict__ tw, cplx<T> *lds) {
    constexpr int TILE = NT * TX;
    const int tid = threadIdx.x;
    const int tiles_x = nx / TX;
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


// Q_x then Q_y on the NX x NX plane A (x fastest).  lds holds NX*(NX+1)
// elements: rows padded by one element so row and column walks spread over
// the LDS banks.
template <typename T, int NX, int BS>
__device__ __forceinline__ void xy_plane(cplx<T> *A, const cplx<T> *__restrict__ tabx,
                                         const cplx<T> *__restrict__ taby,
                                         const cplx<T> *__restrict__ tw, cplx<T> *lds) {
    constexpr int NXP = NX + 1;
    const int tid = threadIdx.x;
    for (int i = tid; i < NX * NX; i += BS) lds[(i / NX) * NXP + (i % NX)] = A[i];
    __syncthreads();
    // Q_x: lines are rows (pitch NXP), elements contiguous.
    line_op<T, NX, -1, NXP, 1>(lds, NX, tw, tabx, tid, BS);
    // Q_y: lines are columns, elements one row apart.
    line_op<T, NX, -1, 1, NXP>(lds, NX, tw, taby, tid, BS);
    for (int i = tid; i < NX * NX; i += BS) A[i] = lds[(i / NX) * NXP + (i % NX)];
}

}  // namespace opo
