/*
 * evpfft_oracle.h -- CPU restatement of the LS-EVPFFT increment loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation (libcreep_hip.so) and the CPU baseline timed beside it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it; the product never links, imports or calls it.
 *
 * It follows lsevpfft/src/LS-EVPFFT.f90, various_functions.f90,
 * fourier_functions.f90, tensor_functions.f90, types.f90 and
 * IO_functions.f90 operation for operation (FP64 throughout), with its own
 * FFT (mixed radix 2/3 + direct DFT for other factors) and its own small dense
 * solvers in place of FFTW and LAPACK.  Each function cites the reference
 * lines it restates.
 *
 * Reductions are deterministic: each z-plane is summed serially, planes are
 * then summed in order, independent of the OpenMP thread count.
 */
#ifndef EVPFFT_ORACLE_H
#define EVPFFT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

/* One macro-step output row: the results.csv and diagnostics.csv columns
 * (IO_functions.f90:1264-1271, types.f90:787-798) plus work counters. */
typedef struct {
    int iter, imacroloop, ipr, imicro;
    double sim_time, wall_time, tdot, tdot_ref;
    double eavg[6], sgsolavg[6], epavg[6], eelavg[6], velgradavg[6], edotpavg[6];
    double evm, evmp, dvm, dvmp, svm, gamma, g, geff;
    /* diagnostics.csv */
    double devm_inc_max, diag_tdotref, erre, errs, errsbc, s33_target, s33, s33_sol;
    double detF_min_sol, detF_min_gas, detF_min_sol_rel, detF_min_gas_rel;
    double volinc_max_sol, volinc_max_gas, volinc_max_sol_rel, volinc_max_gas_rel;
    double stretchinc_max_sol, stretchinc_max_gas, stretchinc_max_sol_rel, stretchinc_max_gas_rel;
    double rotang_max_sol, rotang_max_gas;
    double dL_step_rms_sol, dL_step_rms_gas, dL_step_rms, dL_step_max_sol, dL_step_max_gas, dL_step_max;
    double dvelgradavg_mag, cond_c066mod_max_sol, cond_c066mod_max_gas, cond_cgas_max;
    int imin_sol[3], imin_gas[3];
    /* counters over all fixed-point passes of this macro step (incl. rolled back) */
    long long passes;      /* fixed-point iterations executed                  */
    long long nr_steps;    /* Newton steps summed over voxels                  */
    long long evals;       /* constitutive evaluations summed over voxels      */
    long long outer;       /* hardening outer iterations summed over voxels    */
    long long rollbacks;
} orc_row;

/* Process block of fft.in (IO_functions.f90:202-345), already post-processed
 * (idsim, dsim, ictrl==0 conversion). */
typedef struct {
    int iudot[9];        /* row-major (i,j)                                  */
    double udot[9];
    int idsim[6], iscau[6];
    double scauchy[9], dsim[9];
    double tdot, tdotref, tdotmin, devmmx, devmapp;
    int ictrl, nsteps;
    double error, erroral;
    int itmax, iupdate, iuphard, iwfields, iwstep;
    double xc0;
    int igamma, ithermo;
} orc_process;

orc_ctx *orc_create(int n1, int n2, int n3, int nph, double complgas, int modal, int do_mix,
                    int nthreads);
void orc_destroy(orc_ctx *c);
const char *orc_error(orc_ctx *c);

/* Phase tables (IO_functions.f90:367-637): ns systems; arrays row-major:
 * dnca/dbca (ns,3), tau (ns,3) = tau0f,tau0b,tau1, thet (ns,2), hard (ns,ns),
 * C (3,3,3,3) crystal stiffness. */
int orc_set_phase(orc_ctx *c, int iph, int igas, int nsyst, int nmodes, const int *nsm,
                  const double *dnca, const double *dbca, const double *tau, const double *thet,
                  const double *nrs, const double *gamd0, const double *hpfac, const double *hard,
                  const double *C);
/* Voxels in x-fastest order: euler (V,3) degrees, grain/phase 1-based. */
int orc_set_voxels(orc_ctx *c, const double *euler, const int *grain, const int *phase);
int orc_add_process(orc_ctx *c, const orc_process *p);
/* initial displacement field (idispini 1: 3 values per cell, 2: per node), before orc_init */
int orc_set_displacement(orc_ctx *c, int idispini, const double *disp);
/* gradient_tensor1 (various_functions.f90:56-112): u [v][3] -> grad [v][9], grad(l,m) = d_m u_l */
void orc_gradient_tensor1(const double *u, double *grad, int ideriv, int n1, int n2, int n3);
/* Initial state: load_input tail (Hall-Petch, wphsol), initial_guess and the
 * initial C0 (LS-EVPFFT.f90:149-173). */
int orc_init(orc_ctx *c);
/* One accepted macro step (LS-EVPFFT.f90:198-589).  Returns 1 when a step was
 * made, 0 when all processes are done, <0 on error. */
int orc_step(orc_ctx *c, orc_row *row);

/* Field access (V-major: value index v*ncomp + comp; 3x3 tensors row-major). */
int orc_get_field(orc_ctx *c, const char *name, double *out);
int orc_field_ncomp(orc_ctx *c, const char *name);
int orc_set_field(orc_ctx *c, const char *name, const double *in);

/* Building blocks exposed for closed-form tests. */
void orc_chg_basis(double *ce2, double *c2, double *ce4, double *c4, int iopt, int kdim);
int orc_fft3d(double *re_im, int n1, int n2, int n3, int ncomp, int sign);
void orc_spatial_freq(int x, int y, int z, int n1, int n2, int n3, double *xi);
int orc_green_at(orc_ctx *c, int kx, int ky, int kz, double *g81);
void orc_plastic_strain_rate(const double *sg6, const double *sch, const double *trialtau,
                             const double *gamd0, const double *nrs, int ns, double *edotp6,
                             double *dedotp, double *gamdot);
void orc_expm_adaptive(const double *a, double *e);
void orc_polar(const double *F, double *R, double *U);
double orc_cond1_est(const double *a66);
int orc_solve_voxel(orc_ctx *c, long v, double *sg_out6);
void orc_state_6x6(const int *idsim, double *dbar, const int *iscau, double *sbar,
                   const double *mevptg, const double *devp0, double *scauchy, double *dsim);
void orc_green_op(const double *c0, const double *xi, double *g81);
int orc_form_G(orc_ctx *c, int ipr0);
int orc_convolve(orc_ctx *c, const double *in, double *out);
int orc_harden(orc_ctx *c, long v, double *tau_out, double *gamma1);
int orc_set_c066(orc_ctx *c, const double *c066);
int orc_set_ipr(orc_ctx *c, int ipr0);
int orc_calc_c0(orc_ctx *c);
int orc_kinematic(orc_ctx *c, double tdot, double *out4);
int orc_stage(orc_ctx *c, const char *name, double tdot);
int orc_set_macro(orc_ctx *c, const char *name, const double *v);
int orc_get_lavg(orc_ctx *c, double *out27);

#ifdef __cplusplus
}
#endif
#endif
