// opo_power.hpp -- compensated power sums shared by the single-solve and
// the batched kernels (averagePowerCutXY, Solver3D.cuh:529-545, with the
// FP32 thrust::reduce replaced by an FP64 Neumaier sum in a fixed order).
//
// The order is a function of the solve alone: nblk = power_blocks(cells)
// partial blocks of 256 threads, thread t of block b summing cells
// b*256 + t + k*nblk*256, a fixed tree inside the block, then one
// workgroup of 256 threads over the nblk partials with the same tree.  A
// solve therefore gets bitwise the same powers alone and in a batch.
#pragma once
#include "opo_device.hpp"

namespace opo {

// Partial sums of one field: as many blocks as one per 4096 cells, 1 to 1024.
__host__ __device__ inline int power_blocks(size_t cells) {
    const size_t want = (cells + 256 * 16 - 1) / (256 * 16);
    return (int)(want < 1 ? 1 : (want > 1024 ? 1024 : want));
}

__device__ __forceinline__ void neumaier_add(double &sum, double &comp, double v) {
    double t = sum + v;
    if (fabs(sum) >= fabs(v)) comp += (sum - t) + v;
    else comp += (v - t) + sum;
    sum = t;
}

// Block b of nblk: the three fields' partial sums of |A|^2 land in
// out[f * ostride] (fields that are null give 0).  red: 3 x 256 doubles of
// LDS.  256 threads.
template <typename T>
__device__ __forceinline__ void power_partial_block(const cplx<T> *f0, const cplx<T> *f1, const cplx<T> *f2,
                                                    size_t n, int b, int nblk, double (*red)[256],
                                                    double *out, size_t ostride) {
    const int tid = threadIdx.x;
    const size_t stride = (size_t)nblk * 256;
    for (int f = 0; f < 3; ++f) {
        const cplx<T> *F = f == 0 ? f0 : (f == 1 ? f1 : f2);
        double s = 0.0, c = 0.0;
        if (F) {
            for (size_t i = (size_t)b * 256 + tid; i < n; i += stride) {
                const cplx<T> a = F[i];
                const double re = (double)a.x, im = (double)a.y;
                neumaier_add(s, c, fma(re, re, im * im));
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
        for (int f = 0; f < 3; ++f) out[(size_t)f * ostride] = red[f][0];
}

// Sum of the nblk partials of each field (field f's at partial[f * pstride
// + b]) into red[f][0].  256 threads.
__device__ __forceinline__ void power_final_sum(const double *partial, size_t pstride, int nblk,
                                                double (*red)[256]) {
    const int tid = threadIdx.x;
    for (int f = 0; f < 3; ++f) {
        double s = 0.0, c = 0.0;
        for (int b = tid; b < nblk; b += 256) neumaier_add(s, c, partial[(size_t)f * pstride + b]);
        red[f][tid] = s + c;
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w)
            for (int f = 0; f < 3; ++f) red[f][tid] += red[f][tid + w];
        __syncthreads();
    }
}

}  // namespace opo
