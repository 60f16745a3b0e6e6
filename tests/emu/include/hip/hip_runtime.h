// Host stand-in for <hip/hip_runtime.h>: lets the CPU tests compile the
// device templates of cuopo3d_amd/csrc with g++ and run them as one thread
// (barriers are no-ops; a single thread runs every task of a phase in turn).
#pragma once
#include <cmath>
#include <cstddef>
#define __device__
#define __host__
#define __forceinline__ inline
#define __global__
#define address_space(n)
inline void __syncthreads() {}
inline int __builtin_amdgcn_readfirstlane(int x) { return x; }
typedef int hipError_t;
typedef void *hipStream_t;
// One workgroup of one thread: index 0 everywhere.
struct EmuDim3 {
    unsigned x = 0, y = 0, z = 0;
};
inline EmuDim3 threadIdx, blockIdx, blockDim, gridDim;
