/*
 * creep_hip.h -- C-ABI of libcreep_hip.so, the MI355X implementation of the
 * LS-EVPFFT per-increment fixed-point loop.
 *
 * The reference has no C interface on this path: its "operator API" is the set
 * of Fortran module procedures `subroutine X(micro, props)` in
 * lsevpfft/src/various_functions.f90 called from the increment loop of
 * lsevpfft/src/LS-EVPFFT.f90:196-589, with state in the derived types of
 * lsevpfft/src/types.f90:5-172 and iteration globals in global.f90:19-27.
 * Each entry point below names the reference procedures it replaces.  Plain
 * pointers and sizes only; a Fortran host binds it with ISO_C_BINDING
 * (see creep_modeling_amd/host/creep_hip_iface.f90).
 *
 * Conventions
 *   - voxel index v = i + n1*(j + n2*k), i.e. micro%voxel(i,j,k) with i fastest;
 *   - 3x3 tensors row-major (a[3*r+c] = A(r+1,c+1)); 4th-order tensors
 *     c[27*i+9*j+3*k+l]; 6x6 matrices in the orthonormal b-basis of chg_basis
 *     (tensor_functions.f90:21-198), row-major;
 *   - return codes: 0 OK, <0 HIP/runtime/argument error, >0 numerical failure
 *     (1 singular LU as the reference's `stop 1`, 2 non-finite errors);
 *     cg_last_error() gives the message;
 *   - one host thread per context; every call is synchronous with respect to
 *     the values it returns; contexts are independent.
 */
#ifndef CREEP_HIP_H
#define CREEP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cg_ctx cg_ctx;

/* One process block of fft.in after load_input's post-processing
 * (IO_functions.f90:202-345; types.f90:120-152).  dsim and scauchy are
 * updated by cg_iterate (state_6x6_evpsc writes them, various_functions.f90:1601). */
typedef struct {
    int iudot[9];
    double udot[9];
    int idsim[6], iscau[6];
    double scauchy[9], dsim[9];
    double tdot, tdotref, tdotmin, devmmx, devmapp;
    int ictrl, nsteps;
    double error, erroral;
    int itmax, iupdate, iuphard, iwfields, iwstep;
    double xc0;
    int igamma, ithermo;
} cg_process;

/* Iteration state carried by the caller between cg_iterate calls
 * (the driver locals `iter` and `mix` of LS-EVPFFT.f90:15, 218-233). */
typedef struct {
    int iter;      /* 0 at the start of an increment or after a rollback */
    double mix;    /* relaxation factor applied at the next pass           */
    int imicro;    /* increment number within the process (1 = first)    */
} cg_iter_state;

/* Results of one fixed-point pass (LS-EVPFFT.f90:238-358). */
typedef struct {
    double errs, erre, errsbc;        /* normalised errors (:341-343)      */
    double velgradavg[9];             /* micro%velgradavg                  */
    double sgavg[9];                  /* micro%sgavg (solid, unnormalised) */
    double sgavgsol[9];
    double dvelgradavg[9];
    double svmavg, dL_iter;
    int64_t nr_steps, evals, outer;   /* constitutive work counters        */
} cg_iter_out;

/* Macroscopic quantities after an accepted increment (LS-EVPFFT.f90:460-523,
 * results.csv and diagnostics.csv columns). */
typedef struct {
    double eavg[9], sgsolavg[9], epavg[9], eelavg[9], edotpavg[9], velgradavg[9], sgavg[9];
    double evmavg, evmpavg, dvmavg, dvmpavg, svmavg, lvmavg, edotvmmx;
    double gammaavg, gavg, geffavg;
    /* update_diagnostics (types.f90:345-758) */
    double detF_min, detF_min_sol, detF_min_gas;
    double volinc_max_sol, volinc_max_gas, stretchinc_max_sol, stretchinc_max_gas;
    double rotang_max_sol, rotang_max_gas;
    double dL_step_rms_sol, dL_step_rms_gas, dL_step_rms, dL_step_max_sol, dL_step_max_gas, dL_step_max;
    double dvelgradavg_mag, cond_c066mod_max_sol, cond_c066mod_max_gas, cond_cgas_max;
    int imin_sol[3], imin_gas[3];
} cg_incr_out;

/* Replaces init_micro (types.f90:256-343) and allocation of the FFTW buffers
 * (fourier_functions.f90:69-196).  precision: 0 = FP64 form, 1 = mixed form
 * (FFT + Green operator in FP32).  modal/do_mix: the --modal/--mix flags
 * (LS-EVPFFT.f90:114-126). */
int cg_create(int n1, int n2, int n3, int nph, double complgas, int modal, int do_mix, int precision,
              cg_ctx **out);
int cg_destroy(cg_ctx *ctx);
const char *cg_last_error(cg_ctx *ctx);

/* Phase tables after data_crystal / data_crystal_elast (IO_functions.f90:367-719):
 * ns systems; dnca/dbca (ns,3) unit vectors, tau (ns,3) = tau0f,tau0b,tau1,
 * thet (ns,2), nrs, gamd0, hpfac (ns), hard (ns,ns), C (81) crystal stiffness. */
int cg_set_phase(cg_ctx *ctx, int iph, int igas, int nsyst, int nmodes, const int *nsm, const double *dnca,
                 const double *dbca, const double *tau, const double *thet, const double *nrs,
                 const double *gamd0, const double *hpfac, const double *hard, const double *C);
/* Microstructure (data_grain, IO_functions.f90:759-809): euler (V,3) Bunge degrees,
 * grain and phase ids (1-based). */
int cg_set_voxels(cg_ctx *ctx, const double *euler, const int *grain, const int *phase);
/* Initial displacement field (IDISPINI and FILEDISPL of fft.in; init_defgrad,
 * IO_functions.f90:147-154, 815-936), before cg_init: idispini 1, disp[3 V] cell values, or
 * idispini 2, disp[3 (n1+1)(n2+1)(n3+1 or 1 for n3 = 1)] node values, x fastest, as the file
 * lists them.  cg_init then starts from F = grad(u0) + 4 I (the reference's delt), keeps F0 = F for
 * the elastic deformation gradient F (Fp F0)^-1, weights the voxels by J and moves the nodal
 * coordinates by the nodal displacement. */
int cg_set_displacement(cg_ctx *ctx, int idispini, const double *disp);
/* Rotations, elastic stiffness, Hall-Petch strengths (IO_functions.f90:100-199,
 * 347-363), initial_guess (various_functions.f90:8-53) and the initial C0
 * (LS-EVPFFT.f90:156-173) for the first process p1. */
int cg_init(cg_ctx *ctx, const cg_process *p1);

/* take_snapshot (various_functions.f90:2839; every field only when the increment can roll
 * back, ictrl = -1 and tdot > tdotmin, else L alone), update_schmid (:472) if
 * do_schmid, form_G (:609; Gamma is evaluated on the fly in the FFT stage and
 * Goperr0 by a reduction) and calc_c066mod_Goperr066mod (:812). */
int cg_begin_increment(cg_ctx *ctx, const cg_process *p, int do_schmid);
/* One pass of the fixed-point loop (LS-EVPFFT.f90:222-358: <sigma>, PK1,
 * convolution_with_Goper, correction_to_current, dL_iter + mixing, solve_res,
 * calc_velgradavg_corr, error normalisation).  Updates st->iter, st->mix and
 * p->dsim / p->scauchy. */
int cg_iterate(cg_ctx *ctx, cg_process *p, cg_iter_state *st, cg_iter_out *out);
/* Several RVEs in one constitutive launch.  The reference runs a sweep or a calibration as
 * independent solver processes (utils/calibrate.py:1045-1133, utils/sweep_microstructures.py:24-29);
 * here the RVEs' contexts are grouped and cg_batch_iterate runs one fixed-point pass of every
 * active RVE: each RVE's spectral stage on its own stream, then one launch of the correction +
 * constitutive kernel over a per-RVE work table (one per kernel instance present), then each
 * RVE's reductions.  Per RVE the result is exactly that of cg_iterate (same kernels, same
 * blocks, same reduction order); rc[i] is cg_iterate's return code for RVE i, st/out/p as
 * there.  RVEs whose increment or run is over are left out through active[i] = 0 (they retire
 * from the work table).  All contexts must live on the current device.  Returns <0 on a
 * runtime error (message in cg_last_error(NULL)). */
typedef struct cg_batch cg_batch;
int cg_batch_create(cg_ctx *const *ctxs, int n, cg_batch **out);
int cg_batch_destroy(cg_batch *b);
int cg_batch_iterate(cg_batch *b, cg_process *const *p, cg_iter_state *st, cg_iter_out *out, const int *active,
                     int *rc);
/* cg_batch_iterate in two halves: cg_batch_launch enqueues the pass and returns at once (p, st
 * and out must stay valid until the wait); cg_batch_wait waits for it and does the host part,
 * filling rc.  Two batches launched in turn keep the GPU busy while the host works on the
 * other one's results. */
int cg_batch_launch(cg_batch *b, cg_process *const *p, cg_iter_state *st, cg_iter_out *out, const int *active);
int cg_batch_wait(cg_batch *b, int *rc);
/* rollback (various_functions.f90:2868-2895), restoring p->dsim / p->scauchy; -1 when the
 * increment took no snapshot (it began with ictrl != -1 or tdot <= tdotmin). */
int cg_rollback(cg_ctx *ctx, cg_process *p);
/* kinematic_increments + kinematic_substep (various_functions.f90:2745-2837):
 * out[4] = max volinc, max rotang (deg), max stretchinc, min detF_next; returns nsub in *nsub. */
int cg_kinematic_substep(cg_ctx *ctx, double tdot, double *out4, int *nsub);
/* The kinematic check of cg_kinematic_substep without a context and without allocating: no
 * counterpart in the reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  Inputs and outputs are
 * device pointers: velgrad = 9 columns of stride ld (row-major 3x3 components, column c at
 * velgrad + c*ld), detF = nvox values; out5 (device) receives max volinc, max rotang (deg),
 * max stretchinc, min detF_next and nsub (as a double), exactly as cg_kinematic_substep
 * returns them for the same fields.  Work is only enqueued on `stream` (a hipStream_t, NULL
 * for the default stream): nothing is synchronised, allocated or freed.  The block partials
 * live in a fixed device array of 2 MiB per GPU, so nvox <= 2^24 (256^3); larger grids return
 * -6.  Calls that overlap on one GPU share that array and race.  Errors: -1 arguments, -2 HIP,
 * -4 no device, -6 too large; cg_last_error(NULL) gives the message (per host thread). */
int cg_kinematic_substep_noalloc(long nvox, const double *velgrad, long ld, const double *detF, double tdot,
                                 double *out5, void *stream);
/* Accepted-increment updates (LS-EVPFFT.f90:431-523 and types.f90:345-758):
 * accumulation, update_plastic_defgrad, update_grid_velgrad_node, update_defgrad,
 * update_elstic_defgrad, update_el_stiff, update_tensor_config (iupdate=1),
 * update_hardening (iuphard=1), macro averages, update_diagnostics. */
int cg_end_increment(cg_ctx *ctx, const cg_process *p, int scan_diagnostics, cg_incr_out *out);
/* calc_c0 (various_functions.f90:2537-2649) with the process' current tdot/tdotref/xc0. */
int cg_calc_c0(cg_ctx *ctx, const cg_process *p);

/* Field access in voxel-major order (v*ncomp + comp), symmetric tensors as full 3x3.
 * Names: sg sgt velgrad velgradold dvelgradref defgrad defgradinv detF wgtc
 * defgrade defgradp ag sg66 c066mod cgm cgas sch crss trialtau gamdot gacumgr
 * edotp ept disgradsym epvm defgradinc defgradeinc xnode.
 * cg_field_ncomp returns the component count (-1 unknown). */
int cg_field_ncomp(cg_ctx *ctx, const char *name);
int cg_get_field(cg_ctx *ctx, const char *name, double *out, size_t n);
int cg_set_field(cg_ctx *ctx, const char *name, const double *in, size_t n);
/* The eigenstrain rate of ithermo = 1 (IO_functions.f90:316-341: one 3x3 tensor per grain,
 * assigned to the grain's voxels; here given per voxel, (V,9) voxel-major, row-major 3x3).  It
 * enters solve_res (various_functions.f90:1226-1227, 1265) in the first increment
 * (cg_iter_state.imicro == 1) of every process whose ithermo is 1. */
int cg_set_thermo(cg_ctx *ctx, const double *eth);
/* Macroscopic operators: c0[81], s0[81], de_ds[81], c066[36], goperr0[81] (any may be NULL). */
int cg_get_macro(cg_ctx *ctx, double *c0, double *s0, double *de_ds, double *c066, double *goperr0);

/* Per-stage device time (ms, HIP events) accumulated since the last reset:
 * t[0] FFT+Green stage, t[1] correction+constitutive kernel, t[2] reductions and
 * copies, t[3] per-increment updates; launches[0..1] kernel launch counts of the
 * first two stages; passes = fixed-point passes, voxel_iters = passes * V. */
typedef struct {
    double t_ms[4];
    int64_t launches[4];
    int64_t passes, voxel_iters, nr_steps, evals, outer;
} cg_stats;
int cg_get_stats(cg_ctx *ctx, cg_stats *st, int reset);
/* HIP-event timing (default 0, off): 1 = every stage (the FFT + Gamma stage, the constitutive
 * solve, the sums, the per-increment work); 2 = the constitutive solve and the per-increment work
 * only.  The solve's events are attached to its launch; the other stages' events are stream
 * markers, which leave a few microseconds between kernels, so mode 2 keeps them out of the
 * passes. */
int cg_set_timing(cg_ctx *ctx, int on);

/* FFT shape tuning.  Before a context's first pass the block shapes and block placement of its
 * register FFT passes are timed on the device and the fastest kept (every candidate computes
 * the same values in the same order).  Replaces the FFTW planning of init_fftw /
 * init_fftw_wisdom_path (fourier_functions.f90:35-66, 68-194): with a tuning file (the path of
 * --fftw_save), a line for the same device, grid, precision and Green-operator kind is read
 * instead of timing, and a new result is appended.  Call before the first cg_iterate. */
int cg_set_tuning_file(cg_ctx *ctx, const char *path);
/* The shapes in use: shape8 = {x forward, y forward, z, y inverse, x inverse block placement
 * (1 XCD-contiguous, 0 launch order), y forward half blocks, y inverse half blocks, z one kx
 * column}; us5 = time per launch of each chosen pass (microseconds, 0 if not timed).  Returns 1
 * if the shapes were tuned or read, 0 if they are the defaults, <0 on error.  No reference
 * counterpart (inspection). */
int cg_get_fft_tuning(cg_ctx *ctx, int *shape8, double *us5);
/* Use the given shapes (entries 0 or 1, as cg_get_fft_tuning) and skip the tuning.  No reference
 * counterpart. */
int cg_set_fft_tuning(cg_ctx *ctx, const int *shape8);

/* Replay each pass' fixed device work as HIP graphs (1) or launch it kernel by kernel (0, the
 * default): the FFT + Gamma stage with its reduction, and the reductions and host copy after
 * the solve; captured at their first use and again when C0, igamma or the FFT shapes change.
 * The results are the same either way.  No reference counterpart. */
int cg_set_graphs(cg_ctx *ctx, int on);

/* Overlap each pass' host part with the next pass' transforms (1, the default) or not (0).
 * With overlap on, cg_iterate enqueues the x, y, z + Gamma and y passes of the next pass after
 * the pass' sums and waits only for the sums; the next cg_iterate then runs the x inverse alone.
 * Any entry point that writes state, a new increment, or another igamma makes the next pass run
 * all five.  The results are bitwise the same either way.  Not with graphs (cg_set_graphs).  No
 * reference counterpart. */
int cg_set_overlap(cg_ctx *ctx, int on);

#ifdef __cplusplus
}
#endif
#endif
