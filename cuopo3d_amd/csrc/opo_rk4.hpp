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

// dAdz (Solver3D.cuh:365-382): dAp = i kp As Ai e^{-i s dk z},
// dAs = i ks Ap Ai* e^{+i s dk z}, dAi = i ki Ap As* e^{+i s dk z}.
// The slopes are carried without their factor i*kappa (q = k / (i kappa));
// kappa is folded into the real stage and update weights, so a stage point
// A + c k is A + i (c kappa) q and costs two FMAs per field.
template <typename T, bool PHASE>
__device__ __forceinline__ Tri<T> slopes(const Rk4Args<T> &a, int st, int stage, const Tri<T> &A) {
    Tri<T> q;
    q.p = cmul(A.s, A.i);
    q.s = cmulc(A.p, A.i);
    q.i = cmulc(A.p, A.s);
    if constexpr (PHASE) {
        cplx<T> e = {a.es_re[st][stage], a.es_im[st][stage]};
        q.p = cmulc(q.p, e);
        q.s = cmul(q.s, e);
        q.i = cmul(q.i, e);
    }
    return q;
}

// a + i w q for a real weight w
template <typename T> __device__ __forceinline__ cplx<T> add_iw(cplx<T> a, cplx<T> q, T w) {
    return {fma_t(-w, q.y, a.x), fma_t(w, q.x, a.y)};
}

// kernelLinealCombination (Solver3D.cuh:386-402): the stage point A + c k
template <typename T>
__device__ __forceinline__ Tri<T> stage_point(const Tri<T> &A, const Tri<T> &q, const T *w) {
    return {add_iw(A.p, q.p, w[0]), add_iw(A.s, q.s, w[1]), add_iw(A.i, q.i, w[2])};
}

// rk4 (Solver3D.cuh:418-420): A += (k1 + 2 k2 + 2 k3 + k4) h / 6, here
// A + i (kappa h/6) (q1 + 2 q2 + 2 q3 + q4) with kappa h/6 formed in FP64 on
// the host (no divisions in the kernel).
template <typename T>
__device__ __forceinline__ cplx<T> rk4_comb(cplx<T> a, cplx<T> q1, cplx<T> q2, cplx<T> q3, cplx<T> q4,
                                            T w) {
    const T two = (T)2;
    cplx<T> s = {fma_t(two, q2.x, q1.x), fma_t(two, q2.y, q1.y)};
    s = {fma_t(two, q3.x, s.x), fma_t(two, q3.y, s.y)};
    return add_iw(a, cadd(s, q4), w);
}

template <typename T, bool PHASE>
__device__ __forceinline__ void rk4_half_step(const Rk4Args<T> &a, int st, Tri<T> &A) {
    Tri<T> q1 = slopes<T, PHASE>(a, st, 0, A);
    Tri<T> q2 = slopes<T, PHASE>(a, st, 1, stage_point(A, q1, a.kc1));
    Tri<T> q3 = slopes<T, PHASE>(a, st, 1, stage_point(A, q2, a.kc1));
    Tri<T> q4 = slopes<T, PHASE>(a, st, 2, stage_point(A, q3, a.kc3));
    A.p = rk4_comb(A.p, q1.p, q2.p, q3.p, q4.p, a.kh6[0]);
    A.s = rk4_comb(A.s, q1.s, q2.s, q3.s, q4.s, a.kh6[1]);
    A.i = rk4_comb(A.i, q1.i, q2.i, q3.i, q4.i, a.kh6[2]);
}

}  // namespace opo
