// opo3d_main.cpp -- command-line driver, the counterpart of the reference's
// src/cuOPO3D.cu: builds the crystal, cavity and fields on the host, runs
// the OPO on the GPU through the C ABI, and writes the signal field in the
// reference's text format plus the per-trip powers.
//
// usage: opo3d_main POWER_mW WAIST_um [NX NT TRIPS OUTDIR [--all] [--npy]]
//   Crystal and cavity as in cuOPO3D.cu:19-46 (MgO:sPPLT, 29 mm, T = 27 C,
//   Lambda = 6.97 um, Lcav = 4 Lcr, Rs = Ri = 0.98, 100 % GDD compensation,
//   focused-cw pump at 532 nm focused at Lcr/2, signal and idler resonant).
//   The grid is a runtime choice here (the reference fixes 128x128x2048 at
//   compile time).  --all writes every time slice and every pixel's time
//   trace as the reference does (cuOPO3D.cu:75-85); by default only the
//   slice NT/2 and the trace of the centre pixel are written.  --npy also
//   writes the three whole fields as NumPy .npy files.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "../include/opo3d.hpp"

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s POWER_mW WAIST_um [NX NT TRIPS OUTDIR [--all] [--npy]]\n", argv[0]);
        return 2;
    }
    const double power = std::atof(argv[1]) * 1e-3;  // cuOPO3D.cu:36 (mW -> W)
    const double waist = std::atof(argv[2]);
    const int nx = argc > 3 ? std::atoi(argv[3]) : 64;
    const int nt = argc > 4 ? std::atoi(argv[4]) : 256;
    const int trips = argc > 5 ? std::atoi(argv[5]) : 200;
    const std::string out = argc > 6 ? std::string(argv[6]) + "/" : std::string("./");
    bool all = false, npy = false;   // --all: every slice and pixel as .dat; --npy: whole fields as .npy
    for (int k = 7; k < argc; ++k) {
        all = all || std::strcmp(argv[k], "--all") == 0;
        npy = npy || std::strcmp(argv[k], "--npy") == 0;
    }
    try {
        if (opo_device_count() < 1) throw std::runtime_error("no HIP device");
        opo_grid g{nx, nx, nt};
        opo_physical ph{};
        ph.crystal = OPO_CRYSTAL_MGOSPPLT;
        ph.temperature = 27.0;
        ph.grating_period = 6.97;
        ph.lp = 0.532;
        ph.ls = 2 * ph.lp;
        ph.li = ph.ls * ph.lp / (ph.ls - ph.lp);
        ph.lx = ph.ly = 1.0e3;
        ph.lcr = 29.0e3;
        ph.nz = 100;
        ph.lcav = 4 * ph.lcr;
        ph.rs = ph.ri = 0.98;
        ph.resonant_s = ph.resonant_i = 1;
        ph.gdd_gamma = 1.0;
        ph.gdd_on = 1;
        ph.fs_on = 1;
        ph.n_round_trips = trips;
        ph.precision = OPO_FP32;
        opo_params p{};
        opo3d::check(opo_setup_params(&g, &ph, &p), "opo_setup_params");
        const size_t n = (size_t)nx * nx * nt;
        opo3d::field32 api(n), as(n);
        opo3d::check(opo_setup_pump(&g, &p, &ph, OPO_PUMP_FOCUSED_CW, power, 0.20, waist, ph.lcr / 2,
                                    OPO_FP32, api.data()), "opo_setup_pump");
        const double eps0 = 8.8541878128e-6, c = 299.792458;
        opo_crystal_props cr{};
        opo3d::check(opo_crystal_properties(ph.crystal, ph.temperature, ph.grating_period, ph.lp,
                                            ph.ls, ph.li, &cr), "opo_crystal_properties");
        const double ap0 = std::sqrt(4 * power / (eps0 * c * 3.141592653589793 * cr.n[0] * waist * waist));
        opo3d::check(opo_setup_noise(&g, 1, 1e-10 * ap0, OPO_FP32, as.data()), "opo_setup_noise");
        opo3d::Solver s(g, p);
        if (!s.checkCourantStability())
            std::printf("        ---> Courant convergence criteria: failed (dz = %g)\n", p.dz);
        s.setFields(api.data(), as.data(), as.data());  // Ai = As (cuOPO3D.cu:51)
        std::vector<double> pw;
        int rc = OPO_OK;
        try {
            pw = s.runOPO(trips);
        } catch (const std::exception &e) {
            std::fprintf(stderr, "%s\n", e.what());
            rc = OPO_ERR_NAN;
        }
        opo3d::field32 ap(n), sig(n), idl(n);
        s.getFields(ap.data(), sig.data(), idl.data());
        std::ofstream fp(out + "power_per_trip.dat");
        fp << std::setprecision(17);
        for (size_t k = 0; k < pw.size() / 3; ++k)
            fp << k << "\t" << pw[3 * k] << "\t" << pw[3 * k + 1] << "\t" << pw[3 * k + 2] << "\n";
        const std::string tag = "_PumpPower" + std::string(argv[1]) + "_W";
        if (all) {
            for (int sl = 0; sl < nt; ++sl)
                opo3d::saveMatrixComplex_TimeSlice(sig, g, sl, out + "OutputSignalPower" + tag + "_slice" + std::to_string(sl));
            for (int x = 0; x < nx; ++x)
                for (int y = 0; y < nx; ++y)
                    opo3d::SaveElectricField3D_time(sig, g, x, y, out + "As_temporal_x_" + std::to_string(x) + "_y_" + std::to_string(y));
        } else {
            opo3d::saveMatrixComplex_TimeSlice(sig, g, nt / 2, out + "OutputSignalPower" + tag + "_slice" + std::to_string(nt / 2));
            opo3d::SaveElectricField3D_time(sig, g, nx / 2, nx / 2, out + "As_temporal_x_" + std::to_string(nx / 2) + "_y_" + std::to_string(nx / 2));
        }
        if (npy) {
            opo3d::saveNpy(ap, g, out + "pump" + tag + ".npy");
            opo3d::saveNpy(sig, g, out + "signal" + tag + ".npy");
            opo3d::saveNpy(idl, g, out + "idler" + tag + ".npy");
        }
        std::printf("trips %zu, final signal power %.6e W\n", pw.size() / 3, pw.empty() ? 0.0 : pw[pw.size() - 2]);
        return rc == OPO_OK ? 0 : 1;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
