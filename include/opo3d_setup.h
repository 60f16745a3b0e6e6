/*
 * opo3d_setup.h -- host-side set-up of an OPO solve (FP64).
 *
 * The reference computes these scalars and initial fields once, before the
 * split-step path, in its crystal classes (src/headers/CrystalsBase/),
 * EFields<Crystal> (src/headers/EFields3D.cuh:71-137, 199-328) and
 * Cavity<Crystal> (src/headers/Cavity3D.cuh:42-50).  These functions produce
 * the same quantities in FP64 so the C++ driver and the tests can fill an
 * opo_params and the initial fields without any CUDA code.
 */
#ifndef OPO3D_SETUP_H
#define OPO3D_SETUP_H

#include "opo3d.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Crystal models (src/headers/CrystalsBase). */
#define OPO_CRYSTAL_MGOPPLN  0  /* MgO:PPLN, Gayer et al.   MgOPPLN.cuh  */
#define OPO_CRYSTAL_PPLN     1  /* congruent PPLN           PPLN.cuh     */
#define OPO_CRYSTAL_MGOSPPLT 2  /* MgO:sPPLT, Bruner et al. MgOsPPLT.cuh */
#define OPO_CRYSTAL_ZGP      3  /* ZnGeP2, Zelmon et al.    ZGP.cuh: birefringent
                                   phase matching (no grating), polarisations
                                   of p, s, i chosen by opo_physical.pol    */

/* Pump field modes (EFields3D.cuh:259-293) plus a sech^2 pulse. */
#define OPO_PUMP_WAVEPLANE_CW     0  /* setWavePlaneCW     EFields3D.cuh:72-85   */
#define OPO_PUMP_WAVEPLANE_PULSED 1  /* setWavePlanePulsed EFields3D.cuh:89-103  */
#define OPO_PUMP_FOCUSED_CW       2  /* setFocusedCW       EFields3D.cuh:107-120 */
#define OPO_PUMP_FOCUSED_PULSED   3  /* setFocusedPulsed   EFields3D.cuh:124-137 */
#define OPO_PUMP_FOCUSED_SECH2    4  /* focused beam x sech(t/T0), T0 = FWHM/(2 acosh sqrt2) */

/* Physical description of one solve. */
typedef struct {
    int    crystal;         /* OPO_CRYSTAL_*                                 */
    double temperature;     /* deg C                                         */
    double grating_period;  /* Lambda (um)                                   */
    double lp, ls, li;      /* wavelengths (um)                              */
    double lx, ly, lcr;     /* window and crystal length (um)                */
    int    nz;              /* NZ: dz = Lcr/(NZ-1)                           */
    double dq;              /* > 0 overrides dQ = 2*d33/pi (um/V)            */
    double lcav;            /* cavity length (um); trt = (Lcav + Lcr(ns-1))/C */
    double rs, ri;          /* mirror reflectivities                         */
    double phase_s, phase_i;/* round-trip phases (rad)                       */
    int    resonant_s, resonant_i; /* EFields::resFields (cuOPO3D.cu:354)    */
    double gdd_gamma;       /* Cavity::gamma (fraction of GDD compensated)   */
    int    gdd_on;          /* Cavity::gdd                                   */
    int    fs_on;           /* free-space step each trip                     */
    int    n_round_trips;   /* NRT                                           */
    double t_window;        /* > 0 overrides the window trt (ps)             */
    int    precision;       /* OPO_FP32 | OPO_FP64                           */
    char   pol[4];          /* OPO_CRYSTAL_ZGP: polarisations of p, s, i as
                               'o'/'e' (ZGP.cuh constructor tuple, e.g.
                               "eoo"); ignored by the periodically poled
                               crystals                                      */
} opo_physical;

/* Crystal properties at (T, Lambda) for the three wavelengths. */
typedef struct {
    double n[3], vg[3], b2[3], b3[3], alpha[3]; /* p, s, i                  */
    double d33, dq;         /* um/V                                          */
    double kp, ks, ki;      /* 2 pi ns/lp, 2 pi ns/ls, 2 pi ni/ls (MgOPPLN.cuh:74-76) */
    double dk;              /* kp - ks - ki - 2 pi/Lambda (MgOPPLN.cuh:77)   */
} opo_crystal_props;

/* Slice count S of `for (float z = 0; z <= Lcr; z += dz)` with
 * dz = float(Lcr/(NZ - 1.0f)) (Solver3D.cuh:949, MgOPPLN.cuh:65). */
int opo_slice_count(double lcr, int nz);

/* n, group velocity, GVD, TOD and absorption (MgOPPLN.cuh:96-259,
 * PPLN.cuh, MgOsPPLT.cuh) in FP64. */
int opo_crystal_properties(int crystal, double temperature, double grating_period, double lp,
                           double ls, double li, opo_crystal_props *out);
/* The same with polarisations (three 'o'/'e' characters), required for
 * OPO_CRYSTAL_ZGP (ZGP.cuh:49-77): n, group velocity and GVD per
 * polarisation, TOD not modelled (0), kp = 2 pi ns/lp, ks = 2 pi ns/ls,
 * ki = 2 pi ni/ls and dk = kp - ks - ki with no grating term (:71-74). */
int opo_crystal_properties_pol(int crystal, double temperature, double grating_period, double lp,
                               double ls, double li, const char *pol, opo_crystal_props *out);

/* Fill every field of an opo_params from a physical description,
 * following EFields3D.cuh:214-226, Cavity3D.cuh:44-47, Solver3D.cuh:163-165,
 * 294-296, 317-318, 542, 676, 987-997. */
int opo_setup_params(const opo_grid *grid, const opo_physical *phys, opo_params *out);

/* Initial pump field Api (EFields::setPumpField, EFields3D.cuh:244-296) on
 * the grid: power in W (used as peak power, as the reference does), FWHM in
 * ps, waist in um, focal point in um from the crystal input.  out is
 * complex64 (precision OPO_FP32) or complex128 (OPO_FP64). */
int opo_setup_pump(const opo_grid *grid, const opo_params *params, const opo_physical *phys,
                   int mode, double power, double fwhm, double waist, double focal_point,
                   int precision, void *out);

/* Seeded complex Gaussian noise, Re and Im iid N(0, sigma^2), from
 * Philox4x32-10 (Salmon et al., SC'11) keyed by seed and counted by the
 * cell index, with Box-Muller.  Replaces the clock-seeded generator of
 * EFields3D.cuh:299-315. */
int opo_setup_noise(const opo_grid *grid, uint64_t seed, double sigma, int precision, void *out);

/* t = linspace(-W/2, W/2, NT) (EFields3D.cuh:322); NT = 1 gives t = 0. */
int opo_time_grid(int nt, double t_window, double *t);
/* omega = 2 pi fftshift(linspace(-NT/(2W), NT/(2W), NT)) (EFields3D.cuh:323-325). */
int opo_omega_grid(int nt, double t_window, double *omega);

#ifdef __cplusplus
}
#endif
#endif /* OPO3D_SETUP_H */
