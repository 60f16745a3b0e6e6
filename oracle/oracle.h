/*
 * oracle.h -- CPU restatement of the reference split-step path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it, as the checker.
 *
 * It follows the reference's operator order (alfredos84/cuOPO3D,
 * src/headers/Solver3D.cuh) with the documented defects fixed (1/(NX*NY)
 * after the xy transform pair, dispersion on every pixel, 1/NT in the time
 * pair, propagators built before a single pass).  Arithmetic is FP64 by
 * default; an FP32 instance gives the single-precision floor.
 *
 * Field buffers at this interface are always complex128 (re, im) in the
 * reference layout IDX(x,y,t) = t*NX*NY + y*NX + x.
 */
#ifndef OPO3D_ORACLE_H
#define OPO3D_ORACLE_H

#include "../include/opo3d.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_solve oracle_solve;

/* precision: 1 = FP64 arithmetic, 0 = FP32 arithmetic. */
oracle_solve *oracle_create(const opo_grid *g, const opo_params *p,
                            int precision);
void oracle_destroy(oracle_solve *o);
void oracle_set_threads(int n);
int  oracle_get_threads(void);

void oracle_set_fields(oracle_solve *o, const double *api, const double *as,
                       const double *ai);
void oracle_get_fields(oracle_solve *o, double *ap, double *as, double *ai);

void oracle_step_nonlinear(oracle_solve *o);   /* Solver3D.cuh:872-943 */
void oracle_step_dispersion(oracle_solve *o);  /* Solver3D.cuh:571-639 */
void oracle_step_diffraction(oracle_solve *o); /* Solver3D.cuh:725-788 */
void oracle_step_free_space(oracle_solve *o);  /* Solver3D.cuh:817-869 */
void oracle_step_gdd(oracle_solve *o);         /* Solver3D.cuh:643-694 */
void oracle_step_mirror(oracle_solve *o);      /* Cavity3D.cuh:67-80   */
void oracle_step_eom(oracle_solve *o);         /* Solver1D.cuh:62-72   */
void oracle_ssfm(oracle_solve *o);             /* Solver3D.cuh:946-960 */
/* n round trips (Solver3D.cuh:979-1011); power: n*3 doubles or NULL.
 * Returns the number of trips completed (stops after a non-finite signal
 * power). */
int  oracle_run_round_trips(oracle_solve *o, int n, double *power);
void oracle_power(oracle_solve *o, double out[3]);

/* Propagator tables as the oracle builds them (complex128):
 * which 0..2: eiLz_p/s/i (NT), 3..5: eiQz_p/s/i (NX*NY, FFT order),
 * 6..7: free-space s/i (NX*NY), 8: GDD (NT). */
void oracle_table(oracle_solve *o, int which, double *out);
/* omega grid (NT doubles). */
void oracle_omega(oracle_solve *o, double *out);

/* Plain transforms for pinning: unnormalised DFT of n complex128 values with
 * exponent sign (-1 forward, +1 inverse), in FP64 (precision 1) or FP32. */
void oracle_fft(int n, int sign, const double *in, double *out,
                int precision);
/* Reference O(n^2) DFT in long double. */
void oracle_dft_slow(int n, int sign, const double *in, double *out);

/* Slice count of the float loop `for (float z = 0; z <= Lcr; z += dz)` with
 * dz = float(Lcr/(NZ-1.0f)) (Solver3D.cuh:949, MgOPPLN.cuh:65). */
int oracle_slice_count(double lcr, int nz);

#ifdef __cplusplus
}
#endif
#endif
