// opo3d.hpp -- C++ view of the C ABI, shaped like the reference's
// Solver<Crystal> (src/headers/Solver3D.cuh:429-491) so a driver written
// against the reference reads the same.  Header-only; link libopo3d.so.
//
// Also provides the reference's text output formats (src/headers/Files.cuh):
// an x-y slice at one time sample (saveMatrixComplex_TimeSlice, :73-94) and
// the time trace of one pixel (SaveElectricField3D_time, :180-197), written
// from a single device-to-host copy of the field.
#pragma once

#include <complex>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <stdexcept>
#include <string>
#include <vector>

#include "opo3d.h"
#include "opo3d_setup.h"

namespace opo3d {

inline void check(int rc, const char *what) {
    if (rc != OPO_OK)
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) +
                                 "): " + opo_last_error());
}

using field32 = std::vector<std::complex<float>>;

class Solver {
  public:
    Solver(const opo_grid &g, const opo_params &p, int device = -1) : g_(g), p_(p) {
        check(opo_create(&g_, &p_, device, &h_), "opo_create");
    }
    ~Solver() { opo_destroy(h_); }
    Solver(const Solver &) = delete;
    Solver &operator=(const Solver &) = delete;

    size_t cells() const { return (size_t)g_.nx * g_.ny * g_.nt; }
    const opo_grid &grid() const { return g_; }
    const opo_params &params() const { return p_; }
    opo_solve *handle() { return h_; }

    // EFields::Api/As/Ai assignment (cuOPO3D.cu:44-51); pass nullptr to keep a field.
    void setFields(const void *api, const void *as, const void *ai) {
        check(opo_set_fields(h_, api, as, ai), "opo_set_fields");
    }
    void getFields(void *ap, void *as, void *ai) { check(opo_get_fields(h_, ap, as, ai), "opo_get_fields"); }

    // Reference member functions.
    void solverRK4() { check(opo_step_nonlinear(h_), "solverRK4"); }
    void dispersion() { check(opo_step_dispersion(h_), "dispersion"); }
    void diffractionInCrystal() { check(opo_step_diffraction(h_), "diffractionInCrystal"); }
    void diffractionInFreeSpace() { check(opo_step_free_space(h_), "diffractionInFreeSpace"); }
    void GDDCompensation() { check(opo_step_gdd(h_), "GDDCompensation"); }
    void applyReflectivity() { check(opo_step_mirror(h_), "applyReflectivity"); }
    void SSFM() { check(opo_ssfm(h_), "SSFM"); }
    void runSinglePass() { check(opo_single_pass(h_), "runSinglePass"); }
    bool checkCourantStability() const { return opo_courant_ok(&p_) == 1; }
    // runOPO(): n_round_trips trips; returns the per-trip (p, s, i) powers.
    std::vector<double> runOPO(int n = -1) {
        if (n < 0) n = p_.n_round_trips;
        std::vector<double> pw(3 * (size_t)n);
        check(opo_run_round_trips(h_, n, pw.data()), "runOPO");
        return pw;
    }
    std::vector<double> averagePowerCutXY() {
        std::vector<double> out(3);
        check(opo_power(h_, out.data()), "averagePowerCutXY");
        return out;
    }

  private:
    opo_grid g_;
    opo_params p_;
    opo_solve *h_ = nullptr;
};

// Text writers with the reference's layout: setprecision(20), "_r.dat" and
// "_i.dat" files, tab-separated rows over x, one row per y.
template <typename C>
void saveMatrixComplex_TimeSlice(const std::vector<C> &V, const opo_grid &g, int slice,
                                 const std::string &name) {
    const char *ext[2] = {"_r.dat", "_i.dat"};
    for (int part = 0; part < 2; ++part) {
        std::ofstream f(name + ext[part]);
        f << std::setprecision(20);
        for (int iy = 0; iy < g.ny; ++iy) {
            for (int ix = 0; ix < g.nx; ++ix) {
                const C &v = V[((size_t)slice * g.ny + iy) * g.nx + ix];
                f << (part ? v.imag() : v.real()) << "\t";
            }
            f << "\n";
        }
    }
}

// One value per line over t at pixel (x, y).
template <typename C>
void SaveElectricField3D_time(const std::vector<C> &V, const opo_grid &g, int x, int y,
                              const std::string &name) {
    const char *ext[2] = {"_r.dat", "_i.dat"};
    for (int part = 0; part < 2; ++part) {
        std::ofstream f(name + ext[part]);
        f << std::setprecision(20);
        for (int it = 0; it < g.nt; ++it) {
            const C &v = V[((size_t)it * g.ny + y) * g.nx + x];
            f << (part ? v.imag() : v.real()) << "\n";
        }
    }
}

// Binary alternative to the text writers: the whole field as a NumPy .npy
// file (format version 1.0), shape (NT, NY, NX), little-endian complex64
// ("<c8") or complex128 ("<c16") in the [t][y][x] layout of the solver.
template <typename C>
void saveNpy(const std::vector<C> &V, const opo_grid &g, const std::string &path) {
    static_assert(sizeof(C) == 8 || sizeof(C) == 16, "complex<float> or complex<double>");
    std::string header = std::string("{'descr': '") + (sizeof(C) == 8 ? "<c8" : "<c16") +
                         "', 'fortran_order': False, 'shape': (" + std::to_string(g.nt) + ", " +
                         std::to_string(g.ny) + ", " + std::to_string(g.nx) + "), }";
    // magic (6) + version (2) + header length (2) + header, padded with
    // spaces to a multiple of 64 bytes and ended by a newline
    const size_t total = 10 + header.size() + 1;
    header.append((64 - total % 64) % 64, ' ');
    header.push_back('\n');
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("saveNpy: cannot open " + path);
    const unsigned char pre[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    f.write(reinterpret_cast<const char *>(pre), 8);
    const unsigned short hl = (unsigned short)header.size();
    const unsigned char len[2] = {(unsigned char)(hl & 0xff), (unsigned char)(hl >> 8)};
    f.write(reinterpret_cast<const char *>(len), 2);
    f.write(header.data(), (std::streamsize)header.size());
    f.write(reinterpret_cast<const char *>(V.data()), (std::streamsize)(V.size() * sizeof(C)));
    if (!f) throw std::runtime_error("saveNpy: write failed for " + path);
}

}  // namespace opo3d
