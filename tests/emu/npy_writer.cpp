// Writes a small complex64 and a complex128 field with opo3d::saveNpy
// (include/opo3d.hpp): value (t, y, x) = (t + 0.5 y) + i (x - 0.25 t).
#include <complex>
#include <string>
#include <vector>
#include "../../include/opo3d.hpp"

template <typename C> void fill_and_save(const opo_grid &g, const std::string &path) {
    std::vector<C> v((size_t)g.nt * g.ny * g.nx);
    for (int t = 0; t < g.nt; ++t)
        for (int y = 0; y < g.ny; ++y)
            for (int x = 0; x < g.nx; ++x)
                v[((size_t)t * g.ny + y) * g.nx + x] = C(t + 0.5 * y, x - 0.25 * t);
    opo3d::saveNpy(v, g, path);
}

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    const opo_grid g{5, 3, 7};
    fill_and_save<std::complex<float>>(g, argv[1]);
    fill_and_save<std::complex<double>>(g, argv[2]);
    return 0;
}
