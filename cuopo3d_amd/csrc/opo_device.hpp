// opo_device.hpp -- device building blocks of the split-step path on CDNA4.
//
//  * cplx<T>: interleaved (re, im) complex in T = float or double, the
//    reference's complex_t layout (DtypesConsts.cuh:20).
//  * Compile-time twiddles: cos/sin(2*pi*j/n) evaluated by the compiler
//    from an exact octant reduction and a Taylor series, so every register
//    DFT uses correctly rounded constants (no recurrences).
//  * dft<N,S>: natural-order DFT of N values held in registers, exponent
//    sign S, built from radix-2/3/4 codelets by the two-level split
//    X[k1 + A k2] = sum_b w_B^{b k2} w_N^{b k1} sum_q w_A^{q k1} x[B q + b].
//  * line_op: y = F_{-S}( tab . F_S(x) ) along lines held in LDS, with the
//    two-level split N = N1*N2 across threads.  Phase 1 (N1-point DFTs and
//    inter-level twiddles), phase 2 (N2-point DFT, table multiply, inverse
//    N2-point DFT, inverse twiddles) and phase 1' (inverse N1-point DFTs)
//    each read and write the same LDS slots, so the whole operator costs
//    two LDS exchanges and three barriers, and the spectrum never needs to
//    be put back in natural order.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include <utility>

namespace opo {

template <typename T> struct alignas(2 * sizeof(T)) cplx {
    T x, y;
};

// Global-memory accesses through pointers whose address space the compiler
// cannot see (field and table pointers read from a solve descriptor in
// memory).  Left generic, they compile to flat instructions, which count
// against the LDS wait counter too: every LDS wait of a line operator would
// then also wait for the prefetch of the next tile or plane.  The pointer
// must address global memory; values move component-wise (see ld_c).
#define OPO_GLOBAL __attribute__((address_space(1)))
template <typename T> __device__ __forceinline__ cplx<T> gld_c(const cplx<T> *p) {
    const OPO_GLOBAL cplx<T> *q = (const OPO_GLOBAL cplx<T> *)p;
    return cplx<T>{q->x, q->y};
}
template <typename T> __device__ __forceinline__ void gst_c(cplx<T> *p, cplx<T> v) {
    OPO_GLOBAL cplx<T> *q = (OPO_GLOBAL cplx<T> *)p;
    q->x = v.x;
    q->y = v.y;
}


// Every multiply-add is written as an explicit fused operation and the
// library is compiled with floating-point contraction off, so each result
// is one fixed sequence of IEEE operations: tile shapes, block sizes and
// launch forms cannot change a single bit of it.
template <typename T> __device__ __forceinline__ T fma_t(T a, T b, T c) {
    if constexpr (sizeof(T) == 4) return fmaf(a, b, c);
    else return fma(a, b, c);
}

template <typename T> __device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) {
    return {a.x + b.x, a.y + b.y};
}
template <typename T> __device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) {
    return {a.x - b.x, a.y - b.y};
}
// (a.x b.x - a.y b.y, a.x b.y + a.y b.x): the product of Operators.cuh:109-116.
template <typename T> __device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return {fma_t(a.x, b.x, -(a.y * b.y)), fma_t(a.x, b.y, a.y * b.x)};
}
// a * conj(b)
template <typename T> __device__ __forceinline__ cplx<T> cmulc(cplx<T> a, cplx<T> b) {
    return {fma_t(a.x, b.x, a.y * b.y), fma_t(a.y, b.x, -(a.x * b.y))};
}
template <typename T> __device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s) {
    return {a.x * s, a.y * s};
}
template <typename T> __device__ __forceinline__ cplx<T> cconj(cplx<T> a) { return {a.x, -a.y}; }
// (S*i) * a
template <int S, typename T> __device__ __forceinline__ cplx<T> mul_si(cplx<T> a) {
    return S > 0 ? cplx<T>{-a.y, a.x} : cplx<T>{a.y, -a.x};
}

// Twiddles and diagonal tables are applied as a * (hi + lo), where hi is the
// constant rounded to T and lo the rounding residual (FP32 only).  A plain
// FP32 twiddle has |w|^2 != 1 by up to ~6e-8, so a forward/inverse pair
// scales the field by a fixed factor; that factor is the same on every
// z-slice and accumulates linearly over the S slices.  The small product
// a*lo is folded into the FMA chain ahead of the exact products a*hi, so
// it survives their rounding (added after rounding, a sub-half-ulp term is
// simply lost).  For T = double lo is zero and skipped.
template <typename T>
__device__ __forceinline__ cplx<T> cmul_hl(cplx<T> a, cplx<T> hi, cplx<T> lo) {
    if constexpr (sizeof(T) == 8) {
        return cmul(a, hi);
    } else {
        const T cr = fmaf(a.x, lo.x, -(a.y * lo.y)), ci = fmaf(a.x, lo.y, a.y * lo.x);
        return {fmaf(a.x, hi.x, fmaf(-a.y, hi.y, cr)), fmaf(a.x, hi.y, fmaf(a.y, hi.x, ci))};
    }
}
// a * conj(hi + lo)
template <typename T>
__device__ __forceinline__ cplx<T> cmulc_hl(cplx<T> a, cplx<T> hi, cplx<T> lo) {
    if constexpr (sizeof(T) == 8) {
        return cmulc(a, hi);
    } else {
        const T cr = fmaf(a.x, lo.x, a.y * lo.y), ci = fmaf(a.y, lo.x, -(a.x * lo.y));
        return {fmaf(a.x, hi.x, fmaf(a.y, hi.y, cr)), fmaf(a.y, hi.x, fmaf(-a.x, hi.y, ci))};
    }
}

// ------------------------------------------------- compile-time twiddles
namespace ct {
constexpr double kPi4 = 0.78539816339744830961566084581987572;  // pi/4

constexpr double taylor_sin(double x) {
    double x2 = x * x, term = x, sum = x;
    for (int k = 1; k < 14; ++k) {
        term *= -x2 / ((2.0 * k) * (2.0 * k + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double taylor_cos(double x) {
    double x2 = x * x, term = 1.0, sum = 1.0;
    for (int k = 1; k < 14; ++k) {
        term *= -x2 / ((2.0 * k - 1.0) * (2.0 * k));
        sum += term;
    }
    return sum;
}
// cos and sin of 2*pi*j/n with the octant found in exact integer arithmetic.
constexpr void cis2pi(long long j, long long n, double &c, double &s) {
    j %= n;
    if (j < 0) j += n;
    long long m = 8 * j;
    long long o = m / n;
    long long r = m - o * n;                   // fraction r/n of the octant
    double x = kPi4 * ((double)r / (double)n);
    double xp = kPi4 * ((double)(n - r) / (double)n);
    double C = taylor_cos(x), S = taylor_sin(x), Cp = taylor_cos(xp), Sp = taylor_sin(xp);
    switch (o) {
        case 0: c = C; s = S; break;
        case 1: c = Sp; s = Cp; break;
        case 2: c = -S; s = C; break;
        case 3: c = -Cp; s = Sp; break;
        case 4: c = -C; s = -S; break;
        case 5: c = -Sp; s = -Cp; break;
        case 6: c = S; s = -C; break;
        default: c = Cp; s = -Sp; break;
    }
    if (r == 0) {  // exact multiples of pi/4 where a component is exactly 0 or +-1
        if (o == 0) { c = 1.0; s = 0.0; }
        if (o == 2) { c = 0.0; s = 1.0; }
        if (o == 4) { c = -1.0; s = 0.0; }
        if (o == 6) { c = 0.0; s = -1.0; }
    }
}
constexpr double cos2pi(long long j, long long n) {
    double c = 0, s = 0;
    cis2pi(j, n, c, s);
    return c;
}
constexpr double sin2pi(long long j, long long n) {
    double c = 0, s = 0;
    cis2pi(j, n, c, s);
    return s;
}
}  // namespace ct

template <int V> using ic = std::integral_constant<int, V>;

template <int B, int E, typename F> __host__ __device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(ic<B>{});
        static_for<B + 1, E>(f);
    }
}

// a * exp(i * 2*pi * J / N), J and N compile-time, with the trivial cases
// (+-1, +-i) done by swaps so no rounding is introduced.
template <int N, int J, typename T> __device__ __forceinline__ cplx<T> twiddle_ct(cplx<T> a) {
    constexpr int j = ((J % N) + N) % N;
    if constexpr (j == 0) {
        return a;
    } else if constexpr (4 * j == N) {
        return mul_si<+1>(a);
    } else if constexpr (2 * j == N) {
        return cplx<T>{-a.x, -a.y};
    } else if constexpr (4 * j == 3 * N) {
        return mul_si<-1>(a);
    } else {
        constexpr double c = ct::cos2pi(j, N), s = ct::sin2pi(j, N);
        constexpr T ch = (T)c, sh = (T)s;
        constexpr T cl = (T)(c - (double)ch), sl = (T)(s - (double)sh);
        return cmul_hl(a, cplx<T>{ch, sh}, cplx<T>{cl, sl});
    }
}

// ------------------------------------------------------ register DFTs
template <int N> constexpr int first_factor() {
    if constexpr (N % 4 == 0 && N > 4) return 4;
    else if constexpr (N % 2 == 0 && N > 2) return 2;
    else if constexpr (N % 3 == 0 && N > 3) return 3;
    else return N;
}

template <int N, int S, typename T> __device__ __forceinline__ void dft(cplx<T> *v);

template <int S, typename T> __device__ __forceinline__ void dft2(cplx<T> *v) {
    cplx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}
template <int S, typename T> __device__ __forceinline__ void dft3(cplx<T> *v) {
    constexpr double r3 = 0.86602540378443864676372317075294;  // sqrt(3)/2
    constexpr T r3h = (T)r3, r3l = (T)(r3 - (double)r3h);
    cplx<T> a0 = v[0], t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    cplx<T> u = {fma_t((T)-0.5, t1.x, a0.x), fma_t((T)-0.5, t1.y, a0.y)};
    cplx<T> w = mul_si<S>(cmul_hl(t2, cplx<T>{r3h, (T)0}, cplx<T>{r3l, (T)0}));
    v[0] = cadd(a0, t1);
    v[1] = cadd(u, w);
    v[2] = csub(u, w);
}
template <int S, typename T> __device__ __forceinline__ void dft4(cplx<T> *v) {
    cplx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    cplx<T> t2 = cadd(v[1], v[3]), t3 = mul_si<S>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[1] = cadd(t1, t3);
    v[2] = csub(t0, t2);
    v[3] = csub(t1, t3);
}

template <int N, int S, typename T> __device__ __forceinline__ void dft(cplx<T> *v) {
    if constexpr (N == 1) {
        return;
    } else if constexpr (N == 2) {
        dft2<S>(v);
    } else if constexpr (N == 3) {
        dft3<S>(v);
    } else if constexpr (N == 4) {
        dft4<S>(v);
    } else {
        constexpr int A = first_factor<N>();
        constexpr int B = N / A;
        static_assert(A > 1 && B > 1, "unsupported register DFT length");
        cplx<T> y[N];
        static_for<0, B>([&](auto b_) {
            constexpr int b = decltype(b_)::value;
            cplx<T> a[A];
            static_for<0, A>([&](auto q_) {
                constexpr int q = decltype(q_)::value;
                a[q] = v[B * q + b];
            });
            dft<A, S>(a);
            static_for<0, A>([&](auto k_) {
                constexpr int k1 = decltype(k_)::value;
                y[k1 * B + b] = twiddle_ct<N, S * b * k1>(a[k1]);
            });
        });
        static_for<0, A>([&](auto k_) {
            constexpr int k1 = decltype(k_)::value;
            cplx<T> c[B];
            static_for<0, B>([&](auto b_) {
                constexpr int b = decltype(b_)::value;
                c[b] = y[k1 * B + b];
            });
            dft<B, S>(c);
            static_for<0, B>([&](auto k2_) {
                constexpr int k2 = decltype(k2_)::value;
                v[k1 + A * k2] = c[k2];
            });
        });
    }
}

// ----------------------------------------------------- line operator
// Two-level split of each line length N into N1 * N2 (N1, N2 <= 32); the
// 2048-sample lines use the three-level split (Split3) instead.
template <int N> struct Split {
    static constexpr int N1 = (N <= 32) ? N : (N == 64) ? 8 : (N == 96) ? 12 : (N == 128) ? 16
                            : (N == 256) ? 16 : (N == 512) ? 32 : (N == 1024) ? 32 : 0;
    static constexpr int N2 = N / N1;
    static_assert(N1 * N2 == N, "unsupported line length");
};

// A batch of lines in LDS: element n of line l is at lds[l*LS + n*ES].
// tw: W_N^j = exp(-2 pi i j / N), j < N, as (hi, lo) pairs tw[2j], tw[2j+1]
// (FP64-built).  tab: the diagonal operator in the bin order of the first
// transform, already holding any normalisation, also as (hi, lo) pairs.
// S is the exponent sign of the first
// transform: +1 for the time axis (inverse first, Solver3D.cuh:606-626),
// -1 for x and y (forward first, Solver3D.cuh:757-777).
// Threads map line-fastest, so consecutive lanes touch consecutive lines.
// Up to three fields share each phase (and its barrier): field f's lines
// start at b0 + f*bstride in LDS and its table at t0 + f*tstride; tasks
// enumerate (line, sub-index, field) line-fastest.  Strides rather than
// pointer arrays keep the per-field addresses in registers (an indexed
// array, or selects between pointers, end up in scratch memory).
template <typename T> struct LineSet {
    cplx<T> *b0;
    int bstride;
    const cplx<T> *t0;
    int tstride;
    int nf;
    __device__ __forceinline__ cplx<T> *base(int f) const { return b0 + f * bstride; }
    __device__ __forceinline__ const cplx<T> *tab(int f) const { return t0 + f * tstride; }
};

// PAD: extra elements after every N2 consecutive elements of a line
// (element n sits at n*ES + (n / N2)*PAD), so that the N1 rows walked in
// phase 2 start in different LDS banks when N2*ES is a multiple of the
// bank count.
// UNI: the caller guarantees nlines and nthr are multiples of the wave
// size, so the 64 tasks of a wave share their (sub-index, field) and every
// twiddle and table entry a wave reads is one address: the index is made
// wave-uniform and the constants come through the scalar cache instead of
// 64 vector lanes each.
template <typename T> using const_space_ptr = const __attribute__((address_space(4))) T *;

// Element i of a read-only table.  With UNI the table is read through the
// constant address space: the kernel never writes it, and a wave-uniform
// index then becomes a scalar load even where the kernel stores to global
// memory between two reads (which otherwise forces vector loads).
// GT: the table is in global memory (see gld_c).
template <typename T, bool UNI, bool GT = false>
__device__ __forceinline__ cplx<T> table_at(const cplx<T> *__restrict__ p, int i) {
    if constexpr (UNI) {
        const const_space_ptr<T> q = (const_space_ptr<T>)p;
        return cplx<T>{q[2 * i], q[2 * i + 1]};
    } else if constexpr (GT) {
        return gld_c(p + i);
    } else {
        return p[i];
    }
}

// Register part of phase 1 (two-level split) for sub-index n2: N1-point
// DFT of v[n1] = x[N2 n1 + n2], then the inter-level twiddles W_N^(n2 k1).
template <typename T, int N, int S, bool UNI = false>
__device__ __forceinline__ void phase1_regs(cplx<T> *v, int n2, const cplx<T> *__restrict__ tw) {
    constexpr int N1 = Split<N>::N1;
    dft<N1, S>(v);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) {
        const int j = n2 * k1;
        v[k1] = S < 0 ? cmul_hl(v[k1], table_at<T, UNI>(tw, 2 * j), table_at<T, UNI>(tw, 2 * j + 1))
                      : cmulc_hl(v[k1], table_at<T, UNI>(tw, 2 * j), table_at<T, UNI>(tw, 2 * j + 1));
    }
}

// GT: the line tables are in global memory (twiddles may be staged in LDS).
template <typename T, int N, int S, int LS, int ES, int PAD = 0, bool UNI = false, bool GT = false>
__device__ __forceinline__ void line_op_set(const LineSet<T> &ls, int nlines,
                                            const cplx<T> *__restrict__ tw, int tid, int nthr) {
    constexpr int N1 = Split<N>::N1, N2 = Split<N>::N2;
    constexpr int RS = N2 * ES + PAD;   // stride between the N1 rows of N2 elements
    static_assert(N2 > 1 || PAD == 0, "padding applies to two-level lines");
    auto uni = [](int r) { return UNI ? __builtin_amdgcn_readfirstlane(r) : r; };
    if constexpr (N2 == 1) {
        for (int task = tid; task < nlines * ls.nf; task += nthr) {
            const int l = task % nlines, f = task / nlines;
            const cplx<T> *tab = ls.tab(f);
            cplx<T> v[N];
            cplx<T> *base = ls.base(f) + l * LS;
#pragma unroll
            for (int n = 0; n < N; ++n) v[n] = base[n * ES];
            dft<N, S>(v);
#pragma unroll
            for (int k = 0; k < N; ++k) v[k] = cmul_hl(v[k], table_at<T, UNI, GT>(tab, 2 * k), table_at<T, UNI, GT>(tab, 2 * k + 1));
            dft<N, -S>(v);
#pragma unroll
            for (int n = 0; n < N; ++n) base[n * ES] = v[n];
        }
        __syncthreads();
    } else {
        // phase 1: task (l, n2, f)
        {
            for (int task = tid; task < nlines * N2 * ls.nf; task += nthr) {
                const int l = task % nlines, r = uni(task / nlines), n2 = r % N2, f = r / N2;
                cplx<T> *base = ls.base(f) + l * LS + n2 * ES;
                cplx<T> v[N1];
#pragma unroll
                for (int n1 = 0; n1 < N1; ++n1) v[n1] = base[n1 * RS];
                phase1_regs<T, N, S, UNI>(v, n2, tw);
#pragma unroll
                for (int k1 = 0; k1 < N1; ++k1) base[k1 * RS] = v[k1];
            }
            __syncthreads();
        }
        // phase 2: task (l, k1, f)
        for (int task = tid; task < nlines * N1 * ls.nf; task += nthr) {
            const int l = task % nlines, r = uni(task / nlines), k1 = r % N1, f = r / N1;
            const cplx<T> *tab = ls.tab(f);
            cplx<T> *base = ls.base(f) + l * LS + k1 * RS;
            cplx<T> v[N2];
#pragma unroll
            for (int n2 = 0; n2 < N2; ++n2) v[n2] = base[n2 * ES];
            dft<N2, S>(v);
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) {
                const int k = k1 + N1 * k2;
                v[k2] = cmul_hl(v[k2], table_at<T, UNI, GT>(tab, 2 * k), table_at<T, UNI, GT>(tab, 2 * k + 1));
            }
            dft<N2, -S>(v);
#pragma unroll
            for (int n2 = 1; n2 < N2; ++n2) {
                const int j = n2 * k1;
                v[n2] = S < 0 ? cmulc_hl(v[n2], table_at<T, UNI>(tw, 2 * j), table_at<T, UNI>(tw, 2 * j + 1))
                              : cmul_hl(v[n2], table_at<T, UNI>(tw, 2 * j), table_at<T, UNI>(tw, 2 * j + 1));
            }
#pragma unroll
            for (int n2 = 0; n2 < N2; ++n2) base[n2 * ES] = v[n2];
        }
        __syncthreads();
        // phase 1': task (l, n2, f)
        {
            for (int task = tid; task < nlines * N2 * ls.nf; task += nthr) {
                const int l = task % nlines, r = uni(task / nlines), n2 = r % N2, f = r / N2;
                cplx<T> *base = ls.base(f) + l * LS + n2 * ES;
                cplx<T> v[N1];
#pragma unroll
                for (int k1 = 0; k1 < N1; ++k1) v[k1] = base[k1 * RS];
                dft<N1, -S>(v);
#pragma unroll
                for (int n1 = 0; n1 < N1; ++n1) base[n1 * RS] = v[n1];
            }
            __syncthreads();
        }
    }
}

// Three-level split of long lines, N = N1 * N2 * N3 (M = N2 * N3): every
// phase works on register DFTs of at most 16 points, so a tile of a few
// pixels still gives each thread a task in every phase, at the price of two
// more LDS exchanges and one more pair of twiddle multiplies.
template <int N> struct Split3 {
    static constexpr int N1 = (N == 256) ? 8 : (N == 512) ? 8 : (N == 1024) ? 16 : (N == 2048) ? 16 : (N == 4096) ? 16 : 0;
    static constexpr int N2 = (N == 256) ? 8 : (N == 512) ? 8 : (N == 1024) ? 8 : (N == 2048) ? 16 : (N == 4096) ? 16 : 0;
    static constexpr int N3 = (N == 256) ? 4 : (N == 512) ? 8 : (N == 1024) ? 8 : (N == 2048) ? 8 : (N == 4096) ? 16 : 0;
    static constexpr int M = N2 * N3;
    static_assert(N1 * N2 * N3 == N, "unsupported three-level line length");
};

// Line operator y = F_{-S}(tab . F_S(x)) on three levels.  Element n of
// line l sits at l*LS + n*ES + (n / N3)*PAD (PAD separates the groups of N3
// elements that phase C walks).  With n = M n1 + m, m = N3 n2 + n3 and bin
// k = k1 + N1 (k2 + N2 k3):
//   A  per (l, m):      N1-point DFT over n1, times W_N^(m k1)
//   B  per (l, k1, n3): N2-point DFT over n2, times W_M^(n3 k2)
//   C  per (l, k1, k2): N3-point DFT over n3, table, inverse DFT, times
//                       W_M^(-n3 k2)
//   B' per (l, k1, n3): inverse N2-point DFT, times W_N^(-m k1)
//   A' per (l, m):      inverse N1-point DFT
// W_M^j is the N-point twiddle W_N^(N1 j).  The same LineSet conventions
// (fields, tables, task order line-fastest) as line_op_set.
// Register part of phase A (three-level split) for sub-index m: N1-point
// DFT of v[n1] = x[M n1 + m], then the twiddles W_N^(m k1).
template <typename T, int N, int S>
__device__ __forceinline__ void phaseA_regs(cplx<T> *v, int m, const cplx<T> *__restrict__ tw) {
    constexpr int N1 = Split3<N>::N1;
    dft<N1, S>(v);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) {
        const int j = m * k1;
        v[k1] = S < 0 ? cmul_hl(v[k1], tw[2 * j], tw[2 * j + 1]) : cmulc_hl(v[k1], tw[2 * j], tw[2 * j + 1]);
    }
}

// PG: elements between two pads of PAD (the innermost group N3 by default).
template <typename T, int N, int S, int LS, int ES, int PAD = 0, int PG = 0, bool GT = false>
__device__ __forceinline__ void line_op_set3(const LineSet<T> &ls, int nlines,
                                             const cplx<T> *__restrict__ tw, int tid, int nthr) {
    using Q = Split3<N>;
    constexpr int N1 = Q::N1, N2 = Q::N2, N3 = Q::N3, M = Q::M;
    constexpr int G = PG > 0 ? PG : N3;   // elements between two pads
    static_assert(N % G == 0, "the pad group divides the line");
    auto pos = [](int n) { return n * ES + (n / G) * PAD; };
    auto fwd = [&](cplx<T> a, int j) {
        return S < 0 ? cmul_hl(a, tw[2 * j], tw[2 * j + 1]) : cmulc_hl(a, tw[2 * j], tw[2 * j + 1]);
    };
    auto inv = [&](cplx<T> a, int j) {
        return S < 0 ? cmulc_hl(a, tw[2 * j], tw[2 * j + 1]) : cmul_hl(a, tw[2 * j], tw[2 * j + 1]);
    };
    // A: task (l, m, f)
    {
        for (int task = tid; task < nlines * M * ls.nf; task += nthr) {
            const int l = task % nlines, r = task / nlines, m = r % M, f = r / M;
            cplx<T> *base = ls.base(f) + l * LS;
            cplx<T> v[N1];
#pragma unroll
            for (int n1 = 0; n1 < N1; ++n1) v[n1] = base[pos(M * n1 + m)];
            phaseA_regs<T, N, S>(v, m, tw);
#pragma unroll
            for (int k1 = 0; k1 < N1; ++k1) base[pos(M * k1 + m)] = v[k1];
        }
        __syncthreads();
    }
    // B: task (l, n3, k1, f)
    for (int task = tid; task < nlines * N3 * N1 * ls.nf; task += nthr) {
        const int l = task % nlines, r = task / nlines, n3 = r % N3, r2 = r / N3, k1 = r2 % N1, f = r2 / N1;
        cplx<T> *base = ls.base(f) + l * LS;
        const int o = M * k1 + n3;
        cplx<T> v[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) v[n2] = base[pos(o + N3 * n2)];
        dft<N2, S>(v);
#pragma unroll
        for (int k2 = 1; k2 < N2; ++k2) v[k2] = fwd(v[k2], N1 * n3 * k2);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) base[pos(o + N3 * k2)] = v[k2];
    }
    __syncthreads();
    // C: task (l, k2, k1, f)
    for (int task = tid; task < nlines * N2 * N1 * ls.nf; task += nthr) {
        const int l = task % nlines, r = task / nlines, k2 = r % N2, r2 = r / N2, k1 = r2 % N1, f = r2 / N1;
        const cplx<T> *tab = ls.tab(f);
        cplx<T> *base = ls.base(f) + l * LS;
        const int o = M * k1 + N3 * k2;
        cplx<T> v[N3];
#pragma unroll
        for (int n3 = 0; n3 < N3; ++n3) v[n3] = base[pos(o + n3)];
        dft<N3, S>(v);
#pragma unroll
        for (int k3 = 0; k3 < N3; ++k3) {
            const int k = k1 + N1 * (k2 + N2 * k3);
            v[k3] = cmul_hl(v[k3], table_at<T, false, GT>(tab, 2 * k), table_at<T, false, GT>(tab, 2 * k + 1));
        }
        dft<N3, -S>(v);
#pragma unroll
        for (int n3 = 1; n3 < N3; ++n3) v[n3] = inv(v[n3], N1 * n3 * k2);
#pragma unroll
        for (int n3 = 0; n3 < N3; ++n3) base[pos(o + n3)] = v[n3];
    }
    __syncthreads();
    // B': task (l, n3, k1, f)
    for (int task = tid; task < nlines * N3 * N1 * ls.nf; task += nthr) {
        const int l = task % nlines, r = task / nlines, n3 = r % N3, r2 = r / N3, k1 = r2 % N1, f = r2 / N1;
        cplx<T> *base = ls.base(f) + l * LS;
        const int o = M * k1 + n3;
        cplx<T> v[N2];
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) v[k2] = base[pos(o + N3 * k2)];
        dft<N2, -S>(v);
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) v[n2] = inv(v[n2], (N3 * n2 + n3) * k1);
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) base[pos(o + N3 * n2)] = v[n2];
    }
    __syncthreads();
    // A': task (l, m, f)
    {
        for (int task = tid; task < nlines * M * ls.nf; task += nthr) {
            const int l = task % nlines, r = task / nlines, m = r % M, f = r / M;
            cplx<T> *base = ls.base(f) + l * LS;
            cplx<T> v[N1];
#pragma unroll
            for (int k1 = 0; k1 < N1; ++k1) v[k1] = base[pos(M * k1 + m)];
            dft<N1, -S>(v);
#pragma unroll
            for (int n1 = 0; n1 < N1; ++n1) base[pos(M * n1 + m)] = v[n1];
        }
        __syncthreads();
    }
}

// Which lines take the three-level operator (measured per slice, t pass
// and xy pass together, on MI355X):
//  * 2048 samples: as N1 * N2 = 64 * 32 the first phase would have 64-point
//    register DFTs and a quarter of the tasks of the other lengths (1.8x
//    slower per tile);
//  * 1024 samples: the 32-point register DFTs of the two-level split hold
//    about 250 VGPRs per lane in FP32 (two waves per SIMD); with 16-point
//    DFTs the cavity runs 10-11 % faster at 32 x 32 and 128 x 128, FP32
//    and FP64;
//  * 512 and 256 samples in FP64 (64 x 64 cavity 2-4 % faster, 32 x 32
//    256-sample lines 1.4 %); FP64 128-sample lines stay two-level (three
//    levels 8 * 4 * 4: 3 % faster at 64 x 64, 4 % slower at 96 x 96), and
//    so do FP32 512-sample lines (three levels within +-4 % by grid).
template <typename T, int N> constexpr bool three_level() {
    return N >= 1024 || (sizeof(T) == 8 && (N == 512 || N == 256));
}
// Length of the innermost group of a line (the groups that padding
// separates in LDS).
template <typename T, int N> constexpr int inner_len() {
    if constexpr (three_level<T, N>()) return Split3<N>::N3;
    else return Split<N>::N2;
}

// PG: pad group of a three-level line (line_op_set3); two-level lines pad
// every N2 elements.
template <typename T, int N, int S, int LS, int ES, int PAD = 0, bool UNI = false, int PG = 0, bool GT = false>
__device__ __forceinline__ void line_op_any(const LineSet<T> &ls, int nlines, const cplx<T> *__restrict__ tw,
                                            int tid, int nthr) {
    if constexpr (three_level<T, N>()) line_op_set3<T, N, S, LS, ES, PAD, PG, GT>(ls, nlines, tw, tid, nthr);
    else line_op_set<T, N, S, LS, ES, PAD, UNI, GT>(ls, nlines, tw, tid, nthr);
}

// Copy of TOT elements by BS threads, element j of thread tid being
// tid + j BS: ld(e) reads element e, st(e, v) writes it.  The reads of a
// chunk of up to 32 VGPRs of elements are all issued before its first
// write, so a thread waits on one memory latency per chunk rather than per
// element (a plain strided loop waits after every load).  Threads with
// fewer than 8 elements keep the plain loop: their few waits cost less than
// the registers (FP64 32 x 32 planes measured 20 % slower chunked).
template <typename T, int TOT, int BS, typename Ld, typename St>
__device__ __forceinline__ void chunked_copy(int tid, Ld ld, St st) {
    constexpr int IT = (TOT + BS - 1) / BS;
    constexpr int CH = 32 / (int)(sizeof(cplx<T>) / 4);   // elements per chunk
    constexpr bool EVEN = TOT % BS == 0;
    if constexpr (IT < 8) {
        for (int e = tid; e < TOT; e += BS) st(e, ld(e));
        return;
    }
#pragma unroll 1
    for (int c = 0; c < IT; c += CH) {   // kept rolled: one chunk of registers live at a time
        cplx<T> r[CH];
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            const int e = tid + (c + k) * BS;
            if (c + k < IT && (EVEN || e < TOT)) r[k] = ld(e);
        }
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            const int e = tid + (c + k) * BS;
            if (c + k < IT && (EVEN || e < TOT)) st(e, r[k]);
        }
    }
}

// Lines of the row-band and column-band kernels (planes beyond the LDS):
// 512-point lines take the three-level 8 x 8 x 8 operator in both
// precisions (FP32 too: its two-level 32 x 16 split holds ~215 VGPRs).
template <typename T, int N> constexpr bool band_three_level() { return three_level<T, N>() || N == 512; }
template <typename T, int N, int S, int LS, int ES>
__device__ __forceinline__ void band_line_op(cplx<T> *lds, int nlines, const cplx<T> *__restrict__ tw,
                                             const cplx<T> *__restrict__ tab, int tid, int nthr) {
    LineSet<T> ls{lds, 0, tab, 0, 1};
    if constexpr (band_three_level<T, N>()) line_op_set3<T, N, S, LS, ES, 0, 0>(ls, nlines, tw, tid, nthr);
    else line_op_set<T, N, S, LS, ES, 0, false>(ls, nlines, tw, tid, nthr);
}

// One field.
template <typename T, int N, int S, int LS, int ES, bool UNI = false>
__device__ __forceinline__ void line_op(cplx<T> *lds, int nlines, const cplx<T> *__restrict__ tw,
                                        const cplx<T> *__restrict__ tab, int tid, int nthr) {
    LineSet<T> ls{lds, 0, tab, 0, 1};
    line_op_any<T, N, S, LS, ES, 0, UNI>(ls, nlines, tw, tid, nthr);
}

}  // namespace opo
