/*
 * opo3d.h -- C ABI of the MI355X split-step Fourier path for (3+1)D chi(2)
 * optical parametric oscillators.
 *
 * The reference (alfredos84/cuOPO3D) exposes this path as the C++ class
 * template Solver<Crystal> (src/headers/Solver3D.cuh:429-491) operating on
 * Thrust vectors owned by EFields<Crystal> (src/headers/EFields3D.cuh:175-239)
 * and Cavity<Crystal> (src/headers/Cavity3D.cuh:31-62).  Every entry point
 * below names the reference member it replaces.
 *
 * Conventions
 *  - Plain C: opaque handles, int status (0 = OPO_OK, negative = error), no
 *    exceptions or exit() across the ABI.  opo_last_error() returns the text
 *    of the most recent failure on the calling thread.
 *  - Units follow the reference (src/headers/DtypesConsts.cuh:55-57): um, ps, V.
 *  - Field layout follows the reference IDX(x,y,t) = t*NX*NY + y*NX + x
 *    (src/headers/Common.cuh:167-170): x fastest, interleaved (re, im).
 *    Host field buffers are complex64 for OPO_FP32 solves and complex128 for
 *    OPO_FP64 solves.
 *  - Field index order everywhere: 0 = pump, 1 = signal, 2 = idler.
 *  - Every parameter is an explicit double; the reference's arithmetic quirks
 *    are parameter values (see opo_params), never alternative code paths.
 */
#ifndef OPO3D_H
#define OPO3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define OPO_OK               0
#define OPO_ERR_ARG         -1  /* invalid argument / null handle            */
#define OPO_ERR_HIP         -2  /* HIP runtime error                          */
#define OPO_ERR_NOMEM       -3  /* device allocation failed                   */
#define OPO_ERR_UNSUPPORTED -4  /* grid size the kernels are not built for    */
#define OPO_ERR_NAN         -5  /* non-finite field detected (run stopped)    */
#define OPO_ERR_BUDGET      -6  /* solve does not fit the batch HBM budget    */
#define OPO_ERR_STATE       -7  /* call not valid in the current state        */
#define OPO_ERR_NOGPU       -8  /* no HIP device available                    */

/* ------------------------------------------------------------- precision */
#define OPO_FP32 0   /* FP32 bulk; FP64-built phases reduced mod 2pi        */
#define OPO_FP64 1   /* all-FP64 device path                                 */

/* Supported axis lengths of the device kernels.  nx, ny in {16, 32, 64,
 * 96, 128, 256, 512}, independently (the reference declares NX and NY
 * separately, DtypesConsts.cuh:28-31; NX != NY planes and planes of 256
 * and 512 points take two passes through HBM for the xy operator instead
 * of one), or nx = ny = 1; nt in {1, 16, 32, 64,
 * 128, 256, 512, 1024, 2048}, and 4096 on the FP32 path (three fields of
 * one pixel's 4096-sample line fill 102 KiB of LDS; FP64 would need twice
 * that).  A length of 1 turns the axis off (NX = NY =
 * 1 is the reference's 1D plane-wave build, src/headers/Solver1D.cuh;
 * NT = 1 its CW build, DtypesConsts.cuh:38).  Batches take NX = NY. */

typedef struct {
    int nx, ny, nt;
} opo_grid;

/*
 * opo_params -- every scalar the path consumes, in FP64.
 * Built by opo_setup_params() from physical inputs, or filled by the caller.
 * Citations give the reference expression each field carries.
 */
typedef struct {
    /* geometry -------------------------------------------------------- */
    double lx, ly;        /* transverse window LX, LY (um)                  */
    double dx, dy;        /* LX/(NX-1), LY/(NY-1)      MgOPPLN.cuh:62-63    */
    double dz;            /* Lcr/(NZ-1)                MgOPPLN.cuh:65       */
    int    n_slices;      /* S: iterations of the float loop Solver3D.cuh:949 */
    double lcr;           /* crystal length (um)                            */

    /* per field: 0 pump, 1 signal, 2 idler ----------------------------- */
    double kappa[3];      /* 2*pi*dQ/(n*lambda)        EFields3D.cuh:219-221 */
    double kdiff[3];      /* diffraction wavevectors   EFields3D.cuh:222-224 */
    double delta[3];      /* group-delay offsets (ps/um): 1/vp-1/vi, 0,
                             1/vi-1/vs                 Solver3D.cuh:163-165 */
    double beta2[3];      /* GVD (ps^2/um)             Solver3D.cuh:163-165 */
    double alpha[3];      /* linear absorption (1/um), applied as
                             exp(-alpha*dz/2) in D and again in Q
                             (Solver3D.cuh:163-165, 294-296)              */

    /* time / frequency grid -------------------------------------------- */
    double t_window;      /* time window W (ps): t = linspace(-W/2, W/2, NT),
                             omega by the rule of EFields3D.cuh:318-328     */
    const double *omega;  /* optional: NT angular frequencies (rad/ps) in the
                             order of the +i transform's output bins; NULL
                             selects the reference rule above              */

    /* nonlinear step ---------------------------------------------------- */
    double dk;            /* phase mismatch (1/um)     MgOPPLN.cuh:77       */
    double dk_phase_scale;/* multiplies dk*z in the RHS phase; 0 reproduces
                             the reference (Solver3D.cuh:375-377)           */
    int    rk4_h_in_stages;/* 0: stage points A+0.5k1, A+0.5k2, A+k3 without
                             the step (Solver3D.cuh:915,921,927); 1: textbook
                             RK4 stage points A+h/2 k1, A+h/2 k2, A+h k3    */

    /* cavity feedback (once per round trip) ---------------------------- */
    int    fs_on;         /* free-space diffraction of s and i       */
    double fs_length;     /* its length (um); the reference uses Lcr
                             (Solver3D.cuh:989)                             */
    int    skip_fs_on_last_trip; /* Solver3D.cuh:987                        */
    int    gdd_on;        /* GDD compensation of s and i  Solver3D.cuh:993  */
    double gdd;           /* GDD (ps^2): -gamma*b2s*Lcr  Solver3D.cuh:676   */
    double fb_re[2];      /* s, i mirror multipliers sqrt(R)*e^{i phi}; 1+0i
                             leaves a non-resonant field untouched
                             (Cavity3D.cuh:12-28, Solver3D.cuh:996-997)     */
    double fb_im[2];
    int    n_round_trips; /* NRT; the free-space step is skipped on trip
                             NRT-1 when skip_fs_on_last_trip is set         */

    /* diagnostics ------------------------------------------------------ */
    double power_prefactor[3]; /* P = pref * sum|A|^2; the reference uses
                             0.5*EPS0*C*ns*dx*dy/(LX*LY)  Solver3D.cuh:542  */

    int    precision;     /* OPO_FP32 or OPO_FP64                           */

    /* intracavity elements of the plane-wave model (Solver1D.cuh) ------- */
    int    eom_on;        /* electro-optic modulator on s and i once per
                             round trip, after the crystal and before GDD:
                             A(t) *= exp(i eom_beta sin(2 pi eom_fpm t))
                             (KernelEOM, Solver1D.cuh:62-72, 467)           */
    double eom_beta;      /* modulation depth (rad)                         */
    double eom_fpm;       /* modulation frequency (THz = 1/ps)              */
    int    no_diffraction;/* 1: the plane-wave model, no transverse step in
                             the crystal or in free space (Solver1D.cuh
                             SSFM :436-445); use with NX = NY = 1           */
} opo_params;

typedef struct opo_solve opo_solve;
typedef struct opo_batch opo_batch;

/* ---------------------------------------------------------- diagnostics */
const char *opo_last_error(void);
const char *opo_version(void);
/* Number of HIP devices (0 when none); never fails. */
int opo_device_count(void);

/* ------------------------------------------------------- single solves */

/* Replaces the constructor Solver(Crystal*, Cavity*, EFields*) plus the
 * per-run propagator set-up (Solver3D.cuh:456-468, 548-567, 699-721):
 * allocates the device state of one solve and builds every propagator
 * table once (FP64 phases reduced mod 2*pi).  device < 0 selects the
 * current device. */
int opo_create(const opo_grid *grid, const opo_params *params, int device,
               opo_solve **out);
void opo_destroy(opo_solve *s);

/* Host -> device copy of the initial pump Api and the signal/idler fields
 * (reference: EFields::Api/As/Ai assignment, cuOPO3D.cu:44-51).  Also sets
 * Ap = Api.  Any pointer may be NULL to leave that field unchanged. */
int opo_set_fields(opo_solve *s, const void *api, const void *as,
                   const void *ai);
/* Device -> host copy of Ap, As, Ai (cuOPO3D.cu:75-85 read A->As). */
int opo_get_fields(opo_solve *s, void *ap, void *as, void *ai);
/* Device pointers of the resident fields (which: 0 Ap, 1 As, 2 Ai, 3 Api),
 * for callers that keep their data in HBM. */
void *opo_field_device_ptr(opo_solve *s, int which);
/* Bytes of one field on the device (nx*ny*nt*8 or *16). */
size_t opo_field_bytes(const opo_solve *s);

/* Individual operators, each one reference member function: */
int opo_step_nonlinear(opo_solve *s);   /* solverRK4(z)   Solver3D.cuh:872-943 */
int opo_step_dispersion(opo_solve *s);  /* dispersion()   Solver3D.cuh:571-639 */
int opo_step_diffraction(opo_solve *s); /* diffractionInCrystal() :725-788     */
int opo_step_free_space(opo_solve *s);  /* diffractionInFreeSpace() :817-869   */
int opo_step_gdd(opo_solve *s);         /* GDDCompensation(As), (Ai) :643-694  */
int opo_step_mirror(opo_solve *s);      /* applyReflectivity Cavity3D.cuh:67-80 */

/* SSFM(): the S-slice loop N(dz/2) D(dz) Q(dz) N(dz/2)  Solver3D.cuh:946-960,
 * on the fused kernels. */
int opo_ssfm(opo_solve *s);
/* runSinglePass() (Solver3D.cuh:1031-1043); the propagators are built at
 * create time, so the zero-table defect of the reference does not occur. */
int opo_single_pass(opo_solve *s);
/* n round trips of runOPO() (Solver3D.cuh:963-1023): pump re-injection,
 * SSFM, free space (skipped on trip NRT-1), GDD, mirror.  power (nullable)
 * receives n*3 doubles, the per-trip power of p, s, i after the mirror
 * (averagePowerCutXY, Solver3D.cuh:529-545) from a compensated FP64 sum.
 * Returns OPO_ERR_NAN (and stops) when a trip's signal power is not finite;
 * opo_trips_done() then tells how many trips completed. */
int opo_run_round_trips(opo_solve *s, int n, double *power);
int opo_trips_done(const opo_solve *s);

/* opo_run_round_trips without allocating device memory.  No counterpart in
 * the reference: it is for callers that cannot call hipMalloc (code
 * between HIP graph captures, tight serving loops, harnesses that forbid
 * allocation).  Runs round trips 0..n-1 of the solve described by grid and
 * params (the free-space step is skipped on trip params->n_round_trips - 1
 * as in opo_run_round_trips) on fields the caller keeps in device memory of
 * `device` (< 0: the current device): api, the injected pump (read), and
 * ap, as, ai (read and written), [t][y][x] interleaved complex of the
 * solve's precision.  Same kernels, same launch variants and the same
 * fields as opo_create + opo_set_fields + opo_run_round_trips, bitwise.
 * power (nullable, host) receives n*3 doubles as there.  The tables, power
 * partials, counters and power series live in one fixed per-device array
 * of OPO_NOALLOC_SCRATCH_BYTES, which the device holds until the process
 * ends; a grid and trip count whose scratch exceeds it is rejected with
 * OPO_ERR_UNSUPPORTED before the GPU is touched (FP32 128x128x2048 leaves
 * room for about 33,000 trips, FP64 for about 24,000).  Calls in one
 * process are serialised; calls that overlap on one GPU from different
 * processes each have their own array.  Synchronous: launches go one by
 * one (no graph) to the device's null stream, and the call returns when
 * they have finished. */
#define OPO_NOALLOC_SCRATCH_BYTES (1 << 20)
int opo_run_round_trips_noalloc(const opo_grid *grid, const opo_params *params,
                                int device, const void *api, void *ap,
                                void *as, void *ai, int n, double *power);
/* averagePowerCutXY(A) for p, s, i (Solver3D.cuh:529-545), compensated. */
int opo_power(opo_solve *s, double out[3]);
/* averagePowerTime(A, x, y) (Solver3D.cuh:509-523): *out = prefactor *
 * sum_t |A(x, y, t)|^2 of field 0 (p), 1 (s) or 2 (i), compensated FP64
 * sum over the NT samples of the pixel.  The reference's prefactor is
 * 0.5*EPS0*C*ns*dt/trt (Solver3D.cuh:521); its reduction runs over NX*NY*NT
 * entries of an NT-entry buffer, past its end, so only the NT samples it
 * wrote are summed here. */
int opo_power_time(opo_solve *s, int field, int x, int y, double prefactor, double *out);
/* checkCourantStability(dz, 0.2*kp*dx*dy) (Solver3D.cuh:495-504), with kp
 * = kdiff[0].  Returns 1 when passed, 0 when failed. */
int opo_courant_ok(const opo_params *p);
/* Wait for all work queued on the solve's stream. */
int opo_sync(opo_solve *s);

/* Per-stage kernel timing.  The reference only prints wall time around the
 * whole run (TimingCode, Common.cuh:83-90; Solver3D.cuh:977, 1013-1014).
 * This runs n round trips without graph capture, brackets every launch with
 * HIP events on the solve's stream, and returns per stage kind the summed
 * kernel time (ms) and the launch count.  The fields advance as in
 * opo_run_round_trips. */
#define OPO_STAGE_NONLINEAR  0  /* stand-alone pointwise RK4 half step     */
#define OPO_STAGE_TIME_OP    1  /* RK4 half steps + along-t operator D,
                                   all fields, one pass                      */
#define OPO_STAGE_XY_OP      2  /* xy operator Q on all fields               */
#define OPO_STAGE_FREE_SPACE 3  /* free-space xy operator on s, i            */
#define OPO_STAGE_GDD        4  /* last RK4 half step + mirror + GDD on s, i */
#define OPO_STAGE_POWER      5  /* per-trip power reduction                  */
#define OPO_N_STAGES         6
int opo_profile_trips(opo_solve *s, int n, double ms[OPO_N_STAGES],
                      int launches[OPO_N_STAGES]);

/* Launch-variant autotuner (no counterpart in the reference, whose launch
 * shapes are compile-time constants, DtypesConsts.cuh:45-48).  For the
 * solve's grid and precision, times every compiled variant of the fused
 * t pass (tile width x block size, one tile per workgroup or persistent
 * workgroups) and of the xy pass (block size, persistent workgroups or
 * groups of small planes) on stretches of up to 16 slices of the real
 * schedule, the fastest of `reps` stretches counting, by coordinate
 * descent (t, xy, t), and keeps the fastest for all later
 * launches of that size in this process.  The solve's fields are saved and
 * restored.  Results are bitwise independent of the variant (tiles only
 * group lines; every multiply-add is explicit).  report (nullable)
 * receives one line per descent step. */
#define OPO_TUNE_TIME_PASS 0
#define OPO_TUNE_XY_PASS   1
int opo_autotune(opo_solve *s, int reps, char *report, size_t report_len);
/* Current variant for (precision, axis, size n = NT or NX); set returns
 * OPO_ERR_ARG for a variant that does not exist. */
int opo_tune_get(int precision, int axis, int n);
int opo_tune_set(int precision, int axis, int n, int variant);

/* ------------------------------------------------- batching and serving */

/* A batch holds many solves of different grids resident in HBM and advances
 * all of them one round trip per opo_batch_step (one launch per split-step
 * stage, the trip captured as a HIP graph).  Admission is by HBM footprint:
 * a submission that would exceed hbm_budget_bytes waits in a queue until
 * finished solves free their memory.  Replaces the one-process-per-point
 * sweep of src/cuOPO3D.sh:56-86. */
int opo_batch_create(int device, size_t hbm_budget_bytes, int precision,
                     opo_batch **out);
void opo_batch_destroy(opo_batch *b);
/* Queue a solve for n_trips round trips; fields are host buffers (copied).
 * Solves of one batch may differ in grid, slice count (crystal length, NZ)
 * and cavity parameters; they share the batch's precision.  Grids:
 * NX = NY in {16, 32, 64, 96, 128, 256, 512}, or 1 x 1 with no_diffraction
 * (the plane-wave model of Solver1D.cuh, modulator included); NT a power
 * of two from 16 to 2048 (4096 in FP32).  Solves with the dk phase term
 * (dk_phase_scale != 0) are refused: the reference multiplies it by 0
 * (Solver3D.cuh:375-377). */
int opo_batch_submit(opo_batch *b, const opo_grid *grid,
                     const opo_params *params, const void *api,
                     const void *as, const void *ai, int n_trips,
                     uint64_t *ticket);
/* Admit queued solves that fit, then run one round trip of every resident
 * solve.  *n_active (nullable) receives the number of solves stepped. */
int opo_batch_step(opo_batch *b, int *n_active);
/* done = 1 when the solve has finished; then power (nullable, n_trips*3
 * doubles) and the fields (nullable host buffers) are copied out and its
 * device memory is released. */
int opo_batch_poll(opo_batch *b, uint64_t ticket, int *done, double *power,
                   void *ap, void *as, void *ai);
/* Device bytes a solve of this grid/precision run for n_trips round trips
 * occupies in a batch: its four fields, tables, power partials, counters
 * and power series, each allocation rounded up to 64 KiB. */
size_t opo_batch_footprint(const opo_grid *grid, int precision, int n_trips);
/* Counts: resident solves, queued solves, bytes in use. */
int opo_batch_stats(opo_batch *b, int *resident, int *queued,
                    size_t *bytes_used);

/* Run the size classes of each stage of the trip concurrently: the
 * launches of a stage (the fused t pass of each NT class, the xy pass of
 * each NX class, ...) are captured as parallel branches of the trip graph
 * on side streams of default priority, joined before the next stage.  The
 * classes of one stage touch disjoint solves, so results are bitwise the
 * same either way.  Off by default; takes effect at the next step. */
int opo_batch_concurrent_classes(opo_batch *b, int on);

/* Tune the launch variants on this batch's own trip (the batched
 * counterpart of opo_autotune; the reference has no launch variants).
 * Admits what the budget allows from the queue, then for each NT class
 * times every t-pass variant whose tile divides the NX of each of the
 * class's resident solves, and for each NX class every xy-pass variant:
 * the class's launches of one round trip over all its solves, replayed
 * reps times.  The fastest of each class is kept in the process-wide
 * variant tables (opo_tune_get).  The resident fields are saved before and
 * put back after, so the solves' results are unchanged (all variants are
 * bitwise equal anyway); this needs free device memory of the size of the
 * resident fields (OPO_ERR_NOMEM otherwise).  A per-class report is
 * written to report (report_len bytes, may be NULL). */
int opo_batch_autotune(opo_batch *b, int reps, char *report, size_t report_len);

/* Re-captures of the trip graph since the batch was created (every change
 * of the resident set causes one at the next step), how many of them
 * updated the instantiated graph in place instead of instantiating a new
 * one, and the host seconds they took in total.  No counterpart in the
 * reference (one process per sweep point, src/cuOPO3D.sh:56-86). */
int opo_batch_rebuilds(opo_batch *b, int *rebuilds, int *updated, double *seconds);

/* Kernel timing of batched round trips (no counterpart in the reference,
 * which only prints wall time, Common.cuh:83-90).  With profiling on, the
 * trip graph is captured with a HIP event recorded before and after every
 * launch, so each launch is timed inside the replayed graph, and every
 * opo_batch_step adds the elapsed times to one entry per (stage, size
 * class).  Turning it on or off re-captures the graph at the next step
 * and drops the entries; every call sets the sums back to zero. */
typedef struct {
    int stage;          /* OPO_STAGE_TIME_OP (first and middle slices),
                           OPO_STAGE_GDD (trip-end pass), OPO_STAGE_XY_OP,
                           OPO_STAGE_FREE_SPACE or OPO_STAGE_POWER          */
    int size;           /* NT of a t-pass class (-NT: its plane-wave solves,
                           NX = NY = 1), NX of an xy class, 0 power        */
    int launches;       /* launches timed                                   */
    double ms;          /* their summed duration                            */
    double alg_bytes;   /* algorithmic bytes of one launch: 16 B (FP64 32 B)
                           per cell per field read and written (the power
                           sum reads only), over the class's solves         */
} opo_stage_time;
int opo_batch_profile(opo_batch *b, int on);
/* Copies up to cap entries to out; *n receives the number of entries.
 * OPO_ERR_HIP when the runtime could not time the graph's event nodes. */
int opo_batch_profile_read(opo_batch *b, opo_stage_time *out, int cap, int *n);
/* Time each (stage, size class) of the current trip on its own: a graph of
 * that class's launches of one round trip (trip order, trip arguments),
 * replayed reps times between two HIP events on the batch stream.  One
 * entry per class as above (launches = launches per trip x reps).  The
 * resident solves' fields go through the timed launches as through extra
 * slices, so time a batch whose results are not needed afterwards. */
int opo_batch_time_stages(opo_batch *b, int reps, opo_stage_time *out, int cap, int *n);

#ifdef __cplusplus
}
#endif
#endif /* OPO3D_H */
