// CPU check of the device line operators (cuopo3d_amd/csrc/opo_device.hpp):
// y = F_{-S}(tab . F_S(x)) on every line of a tile, against an FP64 DFT
// evaluated term by term.  Prints one line per case:
//   <label> <relative L2 error>
#include <hip/hip_runtime.h>
#include <complex>
#include <cstdio>
#include <random>
#include <vector>
#include "../../cuopo3d_amd/csrc/opo_device.hpp"
using namespace opo;
typedef std::complex<double> cd;

struct Case {
    std::vector<cd> x, tab, ref;
    std::vector<cplx<float>> tw, tabf;
};

// nl lines of nf fields of length N, random data and tables (|tab| ~ 1/N).
template <int N, int S> Case make(int nl, int nf, unsigned seed) {
    Case c;
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> g;
    c.x.resize((size_t)nf * nl * N);
    c.tab.resize((size_t)nf * N);
    for (auto &v : c.x) v = cd(g(rng), g(rng));
    for (auto &v : c.tab) v = std::polar(1.0 / N, 3.0 * g(rng)) * (1.0 + 0.1 * g(rng));
    c.tw.resize(2 * N);
    c.tabf.resize(2 * (size_t)nf * N);
    for (int j = 0; j < N; ++j) {
        const double a = -2 * M_PI * (double)j / N, cr = cos(a), ci = sin(a);
        const float hr = (float)cr, hi = (float)ci;
        c.tw[2 * j] = {hr, hi};
        c.tw[2 * j + 1] = {(float)(cr - hr), (float)(ci - hi)};
    }
    for (size_t k = 0; k < c.tab.size(); ++k) {
        const float hr = (float)c.tab[k].real(), hi = (float)c.tab[k].imag();
        c.tabf[2 * k] = {hr, hi};
        c.tabf[2 * k + 1] = {(float)(c.tab[k].real() - hr), (float)(c.tab[k].imag() - hi)};
    }
    c.ref.resize(c.x.size());
    std::vector<cd> X(N);
    for (int f = 0; f < nf; ++f)
        for (int l = 0; l < nl; ++l) {
            const cd *x = &c.x[((size_t)f * nl + l) * N];
            for (int k = 0; k < N; ++k) {
                cd acc = 0;
                for (int n = 0; n < N; ++n) acc += x[n] * std::polar(1.0, S * 2 * M_PI * (double)((long)n * k % N) / N);
                X[k] = acc * c.tab[(size_t)f * N + k];
            }
            for (int n = 0; n < N; ++n) {
                cd acc = 0;
                for (int k = 0; k < N; ++k) acc += X[k] * std::polar(1.0, -S * 2 * M_PI * (double)((long)n * k % N) / N);
                c.ref[((size_t)f * nl + l) * N + n] = acc;
            }
        }
    return c;
}

double rel(const std::vector<cd> &a, const std::vector<cd> &b) {
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        num += std::norm(a[i] - b[i]);
        den += std::norm(b[i]);
    }
    return std::sqrt(num / den);
}

template <int N, int LEVELS> constexpr int inner_group() {
    if constexpr (LEVELS == 3) return Split3<N>::N3;
    else return Split<N>::N2;
}

// Time-tile layout: element n of line (pixel) l at n*TX + l + (n / G)*PAD,
// G the innermost group of the operator or the pad group PG.
template <int N, int TX, int S, int LEVELS, int PG = 0> void time_tile(const char *label) {
    constexpr int NF = 3, PAD = TX < 16 ? TX : 0;
    constexpr int G = PG > 0 ? PG : inner_group<N, LEVELS>();
    constexpr int FT = N * TX + (N / G) * PAD;
    Case c = make<N, S>(TX, NF, 1000 * N + TX);
    std::vector<cplx<float>> lds((size_t)NF * FT);
    auto at = [](int n, int l) { return n * TX + l + (n / G) * PAD; };
    for (int f = 0; f < NF; ++f)
        for (int l = 0; l < TX; ++l)
            for (int n = 0; n < N; ++n) {
                const cd v = c.x[((size_t)f * TX + l) * N + n];
                lds[(size_t)f * FT + at(n, l)] = {(float)v.real(), (float)v.imag()};
            }
    const LineSet<float> ls{lds.data(), FT, c.tabf.data(), 2 * N, NF};
    if constexpr (LEVELS == 3) line_op_set3<float, N, S, 1, TX, PAD, PG>(ls, TX, c.tw.data(), 0, 1);
    else line_op_set<float, N, S, 1, TX, PAD>(ls, TX, c.tw.data(), 0, 1);
    std::vector<cd> out(c.x.size());
    for (int f = 0; f < NF; ++f)
        for (int l = 0; l < TX; ++l)
            for (int n = 0; n < N; ++n) {
                const cplx<float> v = lds[(size_t)f * FT + at(n, l)];
                out[((size_t)f * TX + l) * N + n] = cd(v.x, v.y);
            }
    std::printf("%s %.6e\n", label, rel(out, c.ref));
}

// xy-plane layout: NX rows of pitch NX + 1; lines are rows (Q_x) or columns
// (Q_y); UNI reads the constants wave-uniformly as the 64- and 128-point
// planes do.
template <int NX, bool ROWS, bool UNI> void xy_plane(const char *label) {
    constexpr int NXP = NX + 1;
    Case c = make<NX, -1>(NX, 1, 7 * NX + ROWS);
    std::vector<cplx<float>> lds((size_t)NX * NXP);
    auto at = [](int l, int n) { return ROWS ? l * NXP + n : n * NXP + l; };
    for (int l = 0; l < NX; ++l)
        for (int n = 0; n < NX; ++n) {
            const cd v = c.x[(size_t)l * NX + n];
            lds[at(l, n)] = {(float)v.real(), (float)v.imag()};
        }
    if constexpr (ROWS) line_op<float, NX, -1, NXP, 1, UNI>(lds.data(), NX, c.tw.data(), c.tabf.data(), 0, 1);
    else line_op<float, NX, -1, 1, NXP, UNI>(lds.data(), NX, c.tw.data(), c.tabf.data(), 0, 1);
    std::vector<cd> out(c.x.size());
    for (int l = 0; l < NX; ++l)
        for (int n = 0; n < NX; ++n) {
            const cplx<float> v = lds[at(l, n)];
            out[(size_t)l * NX + n] = cd(v.x, v.y);
        }
    std::printf("%s %.6e\n", label, rel(out, c.ref));
}

// Band layout of the row-band and column-band kernels: RB lines of N
// points, rows of pitch N + 1 (ROWS) or columns of a band of RB + 1 pitch.
template <int N, int RB, bool ROWS> void band(const char *label) {
    constexpr int NP = N + 1, BP = RB + 1;
    Case c = make<N, -1>(RB, 1, 11 * N + RB + ROWS);
    std::vector<cplx<float>> lds((size_t)(ROWS ? RB * NP : N * BP));
    auto at = [](int l, int n) { return ROWS ? l * NP + n : n * BP + l; };
    for (int l = 0; l < RB; ++l)
        for (int n = 0; n < N; ++n) {
            const cd v = c.x[(size_t)l * N + n];
            lds[at(l, n)] = {(float)v.real(), (float)v.imag()};
        }
    if constexpr (ROWS) band_line_op<float, N, -1, NP, 1>(lds.data(), RB, c.tw.data(), c.tabf.data(), 0, 1);
    else band_line_op<float, N, -1, 1, BP>(lds.data(), RB, c.tw.data(), c.tabf.data(), 0, 1);
    std::vector<cd> out(c.x.size());
    for (int l = 0; l < RB; ++l)
        for (int n = 0; n < N; ++n) {
            const cplx<float> v = lds[at(l, n)];
            out[(size_t)l * N + n] = cd(v.x, v.y);
        }
    std::printf("%s %.6e\n", label, rel(out, c.ref));
}

// chunked_copy run for every thread of a BS-thread workgroup in turn:
// prints the number of elements of [0, TOT) not copied exactly once plus
// the number of accesses outside it (0 when the index map is a bijection).
template <typename T, int TOT, int BS> void copy_cover(const char *label) {
    std::vector<int> rd(TOT, 0), wr(TOT, 0);
    long bad = 0;
    for (int tid = 0; tid < BS; ++tid)
        chunked_copy<T, TOT, BS>(
            tid,
            [&](int e) {
                if (e < 0 || e >= TOT) { ++bad; return cplx<T>{}; }
                ++rd[e];
                return cplx<T>{(T)e, (T)-e};
            },
            [&](int e, cplx<T> v) {
                if (e < 0 || e >= TOT || v.x != (T)e || v.y != (T)-e) { ++bad; return; }
                ++wr[e];
            });
    for (int e = 0; e < TOT; ++e) bad += (rd[e] != 1) + (wr[e] != 1);
    std::printf("%s %ld\n", label, bad);
}

int main() {
    time_tile<256, 4, +1, 2>("t256_two_level");
    time_tile<512, 4, +1, 2>("t512_two_level");
    time_tile<1024, 2, +1, 2>("t1024_two_level");
    time_tile<256, 4, -1, 3>("t256_three_level");
    time_tile<512, 4, +1, 3>("t512_three_level");
    time_tile<1024, 2, -1, 3>("t1024_three_level");
    time_tile<2048, 2, +1, 3>("t2048_three_level");
    time_tile<2048, 1, -1, 3>("t2048_three_level_one_pixel");
    time_tile<4096, 1, +1, 3>("t4096_three_level_one_pixel");
    time_tile<2048, 2, -1, 3, 16>("t2048_three_level_pad16");
    xy_plane<32, true, false>("x32_rows");
    xy_plane<64, true, true>("x64_rows_uniform");
    xy_plane<64, false, true>("x64_columns_uniform");
    xy_plane<96, false, false>("x96_columns");
    xy_plane<128, true, true>("x128_rows_uniform");
    band<256, 32, true>("b256_rows");
    band<256, 16, false>("b256_columns");
    band<512, 16, true>("b512_rows_three_level");
    band<512, 16, false>("b512_columns_three_level");
    copy_cover<float, 512 * 32, 512>("copy_f32_16384_512");    // 32 per thread: two chunks
    copy_cover<double, 512 * 16, 256>("copy_f64_8192_256");    // 32 per thread: four chunks
    copy_cover<float, 96 * 96, 256>("copy_f32_9216_256");      // 36 per thread: last chunk partial
    copy_cover<double, 96 * 96, 1024>("copy_f64_9216_1024");   // 9 per thread
    copy_cover<float, 3 * 1000, 256>("copy_f32_3000_256");     // ragged: 12 per thread, tail masked
    copy_cover<double, 32 * 32, 256>("copy_f64_1024_256");     // 4 per thread: the plain loop
    return 0;
}
