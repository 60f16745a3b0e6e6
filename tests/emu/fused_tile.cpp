// CPU check of the fused form of the t pass (opo_passes.hpp): every tile of
// a small grid through nl_time_tile and through nl_time_tile_fused, one
// thread (BS = 1), from the same random fields, tables and step weights.
// The two forms do the same operations on every element, so the outputs
// must be bitwise equal.  Prints one line per case:
//   <label> <number of differing output values>
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>
#include "../../cuopo3d_amd/csrc/opo_passes.hpp"
using namespace opo;

template <typename T> std::vector<cplx<T>> rnd(size_t n, double scale, std::mt19937_64 &rng) {
    std::normal_distribution<double> g;
    std::vector<cplx<T>> v(n);
    for (auto &x : v) x = {(T)(scale * g(rng)), (T)(scale * g(rng))};
    return v;
}

// hi/lo pairs of n entries exp(i a_j) * mag
template <typename T> std::vector<cplx<T>> table(int n, double mag, std::mt19937_64 &rng, bool twiddles) {
    std::uniform_real_distribution<double> u(-3.14159, 3.14159);
    std::vector<cplx<T>> t(2 * (size_t)n);
    for (int j = 0; j < n; ++j) {
        const double a = twiddles ? -2 * 3.14159265358979323846 * j / n : u(rng);
        const double re = mag * std::cos(a), im = mag * std::sin(a);
        const T hr = (T)re, hi = (T)im;
        t[2 * j] = {hr, hi};
        t[2 * j + 1] = {(T)(re - (double)hr), (T)(im - (double)hi)};
    }
    return t;
}

template <typename T, int NT, int TX>
void run_case(const char *label, int first_field, int nsteps, bool mirror) {
    std::mt19937_64 rng(NT * 131 + TX * 7 + first_field * 3 + nsteps);
    const int nx = 2 * TX, ny = 2, tiles = (nx / TX) * ny;
    const size_t n = (size_t)nx * ny * NT;
    auto api = rnd<T>(n, 0.3, rng), f0 = rnd<T>(n, 0.3, rng), f1 = rnd<T>(n, 1e-3, rng), f2 = rnd<T>(n, 1e-3, rng);
    auto tw = table<T>(NT, 1.0, rng, true);
    auto tabs = table<T>(3 * NT, 1.0 / NT, rng, false);   // three tables, 2 NT apart
    auto eom = table<T>(NT, 1.0, rng, false);
    Rk4Args<T> rk{};
    for (int f = 0; f < 3; ++f) {
        rk.kc1[f] = (T)(0.5 * (1e-4 + 1e-5 * f));
        rk.kc3[f] = (T)(1.0 * (1e-4 + 1e-5 * f));
        rk.kh6[f] = (T)(25.0 / 6.0 * (1e-4 + 1e-5 * f));
    }
    rk.nsteps = nsteps;
    rk.has_phase = 0;
    rk.out_scale = mirror ? 1 : 0;
    rk.ms_re = (T)0.98994949;
    rk.ms_im = (T)0.01;
    rk.mi_re = (T)1.0;
    rk.eom = mirror ? eom.data() : nullptr;
    std::vector<cplx<T>> out[2][3];
    for (int form = 0; form < 2; ++form) {
        std::vector<cplx<T>> g[3] = {f0, f1, f2};
        TLine3Args<T> a{};
        a.psrc = api.data();
        for (int f = 0; f < 3; ++f) a.f[f] = g[f].data();
        a.tab0 = tabs.data();
        a.tab_stride = 2 * NT;
        a.first_field = first_field;
        a.rk = rk;
        std::vector<cplx<T>> lds(3 * (size_t)TimeTile<T, NT, TX>::FT), twl(2 * (size_t)NT);
        for (int t = 0; t < tiles; ++t) {
            if (form == 0) nl_time_tile<T, NT, TX, 1, false>(a, nx, ny, t, tw.data(), lds.data(), twl.data());
            else nl_time_tile_fused<T, NT, TX, 1, false>(a, nx, ny, t, tw.data(), lds.data(), twl.data());
        }
        for (int f = 0; f < 3; ++f) out[form][f] = g[f];
    }
    long diff = 0;
    for (int f = 0; f < 3; ++f)
        for (size_t i = 0; i < n; ++i)
            diff += std::memcmp(&out[0][f][i], &out[1][f][i], sizeof(cplx<T>)) != 0;
    long changed = 0;   // the pass must have done something
    for (size_t i = 0; i < n; ++i) changed += std::memcmp(&out[0][1][i], &f1[i], sizeof(cplx<T>)) != 0;
    std::printf("%s %ld %ld\n", label, diff, changed);
}

int main() {
    run_case<float, 128, 8>("f32_nt128_tx8_mid", 0, 2, false);
    run_case<float, 256, 16>("f32_nt256_tx16_mid", 0, 2, false);
    run_case<float, 256, 4>("f32_nt256_tx4_first", 0, 1, false);
    run_case<float, 1024, 4>("f32_nt1024_tx4_mid", 0, 2, false);
    run_case<float, 2048, 2>("f32_nt2048_tx2_mid", 0, 2, false);
    run_case<float, 256, 8>("f32_nt256_tx8_tripend_gdd", 1, 1, true);
    run_case<float, 1024, 2>("f32_nt1024_tx2_tripend_gdd", 1, 1, true);
    run_case<float, 128, 16>("f32_nt128_tx16_tripend_nogdd", 3, 1, true);
    run_case<double, 256, 4>("f64_nt256_tx4_mid", 0, 2, false);
    run_case<double, 512, 2>("f64_nt512_tx2_mid", 0, 2, false);
    run_case<double, 1024, 2>("f64_nt1024_tx2_tripend_gdd", 1, 1, true);
    return 0;
}
