// opo_rk4.hpp -- pointwise nonlinear step of the three coupled fields.
//
// Restates solverRK4 (Solver3D.cuh:872-943) for one cell: the right-hand
// side dAdz (:365-382), the stage points of kernelLinealCombination
// (:386-402) and the update of rk4 (:406-425), evaluated in registers so
// that consecutive half steps of adjacent slices cost one pass over HBM.
#pragma once
#include "opo_device.hpp"
#include "opo_kernels.hpp"

namespace opo {

template <typename T> struct Tri {
    cplx<T> p, s, i;
};

// dAdz (Solver3D.cuh:365-382): i kp As Ai e^{-i s dk z}, i ks Ap Ai* e^{+i s dk z},
// i ki Ap As* e^{+i s dk z}; products evaluated in the reference's order.
template <typename T, bool PHASE>
__device__ __forceinline__ Tri<T> rhs(const Rk4Args<T> &a, int st, int stage, const Tri<T> &A) {
    Tri<T> d;
    cplx<T> ip = {-a.kp * A.s.y, a.kp * A.s.x};  // (i kp) As
    cplx<T> is = {-a.ks * A.p.y, a.ks * A.p.x};  // (i ks) Ap
    cplx<T> ii = {-a.ki * A.p.y, a.ki * A.p.x};  // (i ki) Ap
    d.p = cmul(ip, A.i);
    d.s = cmulc(is, A.i);
    d.i = cmulc(ii, A.s);
    if constexpr (PHASE) {
        cplx<T> e = {a.es_re[st][stage], a.es_im[st][stage]};
        d.p = cmulc(d.p, e);
        d.s = cmul(d.s, e);
        d.i = cmul(d.i, e);
    }
    return d;
}

template <typename T> __device__ __forceinline__ Tri<T> axpy(const Tri<T> &A, const Tri<T> &k, T c) {
    // kernelLinealCombination (Solver3D.cuh:386-402): A + k*s
    return {cadd(A.p, cscale(k.p, c)), cadd(A.s, cscale(k.s, c)), cadd(A.i, cscale(k.i, c))};
}

template <typename T> __device__ __forceinline__ cplx<T> rk4_comb(cplx<T> a, cplx<T> k1, cplx<T> k2,
                                                                  cplx<T> k3, cplx<T> k4, T h) {
    // rk4 (Solver3D.cuh:418-420): A += (k1 + 2 k2 + 2 k3 + k4) * dz / 6
    const T two = (T)2, six = (T)6;
    cplx<T> s = cadd(cadd(cadd(k1, cscale(k2, two)), cscale(k3, two)), k4);
    return {a.x + (s.x * h) / six, a.y + (s.y * h) / six};
}

template <typename T, bool PHASE>
__device__ __forceinline__ void rk4_half_step(const Rk4Args<T> &a, int st, Tri<T> &A) {
    Tri<T> k1 = rhs<T, PHASE>(a, st, 0, A);
    Tri<T> k2 = rhs<T, PHASE>(a, st, 1, axpy(A, k1, a.c1));
    Tri<T> k3 = rhs<T, PHASE>(a, st, 1, axpy(A, k2, a.c1));
    Tri<T> k4 = rhs<T, PHASE>(a, st, 2, axpy(A, k3, a.c3));
    A.p = rk4_comb(A.p, k1.p, k2.p, k3.p, k4.p, a.h);
    A.s = rk4_comb(A.s, k1.s, k2.s, k3.s, k4.s, a.h);
    A.i = rk4_comb(A.i, k1.i, k2.i, k3.i, k4.i, a.h);
}

}  // namespace opo
