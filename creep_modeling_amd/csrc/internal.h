// internal.h -- device state of one RVE context and the kernel launchers
// shared between the translation units of libcreep_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cg {

constexpr int MAXPH = 4;
constexpr int MAXMODES = 8;
constexpr int MAXSYS = 48;  // slip systems per phase

// Per-phase constants, resident in device memory (read-only in kernels).
struct PhaseDev {
    int igas, nsyst, nmodes;
    int nsm[MAXMODES];
    double c66cr[36];     // crystal stiffness in the b-basis (crystal axes)
    double mode_tau1[MAXMODES], mode_thet0[MAXMODES], mode_thet1[MAXMODES];
    // offsets (in doubles) into PhaseTables::data
    int off_dnca, off_dbca, off_nrs, off_gamd0, off_hard, off_tau, off_hpfac;
    int off_sys;          // per system 8 doubles: b (3), n (3), nrs, gamd0 (one scalar load per system)
};

struct PhaseTables {
    const PhaseDev* ph;   // device array [MAXPH]
    const double* data;   // device: concatenated per-system tables
};

// Field arrays (SoA: component-major, a[c*V + v]); symmetric tensors as
// b-basis 6-vectors, general 3x3 tensors as 9 row-major components.
struct Fields {
    int n1, n2, n3, h1;   // h1 = n1/2 + 1 (half spectrum along x)
    long V;
    long ld;              // column stride of every per-voxel field (a[c*ld + v]), V padded so that the
                          // columns of a field start at staggered HBM channel offsets
    int ns;               // slip systems per voxel (max over phases)
    uint8_t* phase;       // 0-based phase id
    uint8_t* gas;
    double *w;            // wgtc
    double *sg, *sgt;     // 6
    double *L, *Lold, *dLref;  // 9
    double *F, *Fi, *J, *Fe, *Fp, *ag;  // 9,9,1,9,9,9
    double *sg66, *c66m, *cgm, *cgas;   // 36 each (cgm only if modal, cgas only if gas voxels)
    double *sop;          // 42 (without modified AL): L D L^T of cm^-1 and sym(s66)/tdot of the symmetric solve
    double *sch, *crss, *ttau, *gdot, *gacc;  // 5ns, 2ns, 2ns, ns, 1
    double *taun;         // 1: |ttau| (columns in order), written with ttau by every harden and by k_taun
    double *schA;         // 18: A = Fe ag and A^-1 of update_schmid (sch_s = sym(A b_s x A^-T n_s))
    double *edotp, *ept, *dgs, *epvm;   // 6,6,6,1
    double *Finc, *Feinc;               // 9,9
    double *eth;          // 6 (b-basis): eigenstrain rate of the voxel's grain (ithermo = 1), or null
    double *Fini;         // 9: initial deformation gradient of an initial displacement field, or null (F0 = I)
    double *s_L, *s_Lold, *s_sg, *s_edotp, *s_gdot, *s_ttau;  // snapshots
    double2* spec;        // FP64 half spectrum, [point][9] (point = (k*n2+j)*h1+kx)
    float2* specf;        // mixed form: FP32 half spectrum
    double* xnode;        // 3*(n1+1)*(n2+1)*n3node
    double* part;         // block partial sums
    double* red;          // final reductions (device)
};

// Constants of one fixed-point pass.
struct PassConst {
    double tdot, complgas, error, erroral, mix, lscale;
    int modal, do_mix, kin_ctrl, lold_same, iuphard;
    int thermo;           // the eigenstrain rate f.eth enters the solve (first increment, ithermo = 1)
};

// Arguments of the fused correction + constitutive kernel for one pass of one RVE.
struct SolveArgs {
    PassConst pc;
    double dvelgradavg[9];
    double finv_inv[9];
    const double* fluct;  // 9 reduced sums of w dLref Fi
};
// One RVE's entry of the batched constitutive launch: its blocks are blk0 .. blk0 + nblk - 1
// of the launch.
struct SolveWork {
    Fields f;
    PhaseTables pt;
    SolveArgs a;
    long blk0, nblk;
};

// reduction slots written by the fused correction + constitutive kernel
// (slots 0..22 come from the constitutive phase, 23..53 from the correction phase)
enum {
    R_ERRS = 0, R_ERRE = 1,
    R_SGSOL = 2,     // 9: sum_solid w sigma
    R_SGALL = 11,    // 9: sum_all w sigma (next pass' <sigma>)
    R_NR = 20, R_EVAL = 21, R_OUTER = 22,
    R_NPHASEB = 23,
    R_LALL = 23,     // 9: sum w L (before mixing)
    R_LGAS = 32,     // 9
    R_LSOL = 41,     // 9
    R_DLSOL = 50, R_WSOL = 51, R_DLGAS = 52, R_WGAS = 53,
    R_NSOLVE = 54
};
// Status slots of Fields::red, above every reduction's output range (at most 81 values): a
// kernel that meets an exactly singular system (the reference's lu_inverse / lu_eqsystem
// `stop 1`, tensor_functions.f90:323-334, 353-356) stores the voxel index and a 1.  A pass
// copies red[0, R_NCOPY) to the host in one transfer, so the flag travels with its sums.
enum { R_FLAG_VOX = 126, R_FLAG_SING = 127, R_NCOPY = 128 };

// ---- solver_kernels.hip ----
void launch_reduce_final(const double* part, int nblocks, int nvals, double* out, hipStream_t s);
void launch_reduce_final_cm(double* part, int nblocks, int nvals, double* out, hipStream_t s);
void launch_sigma_sum(const Fields& f, hipStream_t s, int* nblocks_out);
void launch_correct_solve_full(const Fields& f, const PhaseTables& pt, const PassConst& pc,
                               const double* dvelgradavg, const double* finv_inv, const double* fluct_dev,
                               hipStream_t s, int* nb, int int_pow, int sym_tau, int small_ns,
                               hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// a pass' R_NSOLVE sums: red, the host's copy of red[0, ncopy) (host-mapped memory), <sigma> for
// the next pass and, with dscale, the next pass' mixed-form stress scale {1/s, s/V}, s = max(von
// Mises of the solid stress sum, 1), written by the block that finishes last (counter: zero
// between calls); done (or null): recorded at the end of the last kernel
void launch_pass_sums(double* part, int nblocks, int nvals, int ncopy, double* red, double* host, double* sgsum,
                      double* dscale, unsigned* counter, long V, hipStream_t s, hipEvent_t done);
// the correction + constitutive kernel over several RVEs: block b belongs to RVE blk_rve[b]
void launch_correct_solve_batch(const SolveWork* work, const int* blk_rve, int nblocks, int modal, int int_pow,
                                int sym_tau, int small_ns, hipStream_t s);
int solve_blocks(long V);  // blocks of one RVE in the constitutive launch
// write_sch: also store the Schmid tensors (only the modified-AL solve reads them; the others
// rebuild them from A and A^-1)
void launch_update_schmid(const Fields& f, const PhaseTables& pt, int write_sch, hipStream_t s);
void launch_sch_from_A(const Fields& f, const PhaseTables& pt, hipStream_t s);
void launch_c066mod(const Fields& f, const double* c066, const double* goperr0, int modal, hipStream_t s);
void launch_gas_compl(const Fields& f, double tdot, double complgas, int modal, hipStream_t s);
void launch_goperr0(const Fields& f, const double* c0, const double* s0, int igamma, hipStream_t s, int* nb);
// max_out: the largest von Mises strain rate of a solid voxel (device)
void launch_accumulate(const Fields& f, double tdot, double* max_out, hipStream_t s, int* nb);
// out[0..13): disgradsym (6), ept (6) and w of voxel v
void launch_pick_last(const Fields& f, long v, double* out, hipStream_t s);
void launch_plastic_defgrad(const Fields& f, const PhaseTables& pt, double tdot, hipStream_t s);
void launch_update_defgrad(const Fields& f, double tdot, hipStream_t s, int* nb);
void launch_div_w(const Fields& f, double detfavg, hipStream_t s);
void launch_elastic_stiff(const Fields& f, const PhaseTables& pt, hipStream_t s);
void launch_tensor_config_a(const Fields& f, hipStream_t s, int* nb);
void launch_tensor_config_b(const Fields& f, const double* dud, hipStream_t s);
void launch_hardening(const Fields& f, const PhaseTables& pt, double tdot, int small_ns, int sym_tau, hipStream_t s);
void launch_solve_ops(const Fields& f, double tdot, hipStream_t s);
void launch_taun(const Fields& f, const PhaseTables& pt, int sym, hipStream_t s);
void launch_averages(const Fields& f, const PhaseTables& pt, hipStream_t s, int* nb);
void launch_diagnostics(const Fields& f, double lfloor, int scan_cond, hipStream_t s, int* nb);
void launch_calc_c0(const Fields& f, const PhaseTables& pt, double tdot, double tdotref, hipStream_t s, int* nb,
                    int int_pow);
void launch_kinematic(const Fields& f, double tdot, hipStream_t s, int* nb);
// the same kernels on raw device pointers; part holds kinematic_scratch_doubles(V) partials and
// red receives 5 values (max volinc, rotang, stretchinc, min detF_next, nsub)
void launch_kinematic_raw(const double* L, long ld, const double* J, long V, double tdot, double* part, double* red,
                          hipStream_t s);
long kinematic_scratch_doubles(long V);
// take_snapshot's copies (restore: back); full = 0: L only (the diagnostics compare L with it)
void launch_snapshot(const Fields& f, int restore, hipStream_t s, int full = 1);
void launch_init_voxels(const Fields& f, const PhaseTables& pt, const double* euler, const int* grain,
                        const double* hp_voxel, const double* udot, double tdot, int keep_defgrad, hipStream_t s,
                        int* nb);
void launch_init_defgrad(const Fields& f, const double* disp, int idispini, double delt, hipStream_t s, int* nb);
void launch_xnode_disp(const Fields& f, const double* disp, int idispini, int n3node, hipStream_t s);
void launch_copy(double* dst, const double* src, size_t n, hipStream_t s);
void launch_fill(double* dst, double val, size_t n, hipStream_t s);
void launch_init_xnode(const Fields& f, int n3node, hipStream_t s);
void launch_xnode(const Fields& f, const double* dv, const double* lavg, double tdot, int ideriv, int n3node,
                  hipStream_t s);

// ---- fft_kernels.hip ----
struct FftCtx;
// the reference medium's stiffness and compliance (4th-order, 81 each) and the Green operator kind
struct GreenDev {
    double c0[81];
    double s0[81];
    int igamma;
};
// One RVE of a batched pass (cg_batch_*): what its FFT + Gamma stage and its sums read and write.
struct FftWork {
    Fields f;
    const void* tw[3];   // twiddles of the x, y, z transforms (the group's precision)
    const double* sgsum; // <sigma> (9)
    double* dscale;      // the mixed form's stress scale {1/s, s/V} (written by the pass sums)
    unsigned* counter;   // the pass sums' last-block counter (zero between passes)
    double* fluct;       // 9 sums of w dLref F^-1
    double* host;        // the pass' sums, host-mapped
    GreenDev g;
};
// The FFT + Gamma stage of nrve RVEs with the same grid (f0's dimensions) and precision, w on the
// device: one launch per pass over all of them (default block shapes), then their sums of
// w dLref F^-1 (two launches).  Each RVE computes what fft_green computes for it alone.
int fft_green_batch(const FftWork* w, int nrve, const Fields& f0, int mixed, int general, hipStream_t s);
// (solver_kernels.hip) the two-level sums of w dLref F^-1 of such a group into each RVE's fluct;
// nblocks: the x inverse's blocks per RVE
void launch_reduce_fluct_batch(const FftWork* w, int nrve, int nblocks, hipStream_t s);
// a pass' sums of a group of RVEs with nblocks solve blocks each (launch_pass_sums for each RVE)
void launch_pass_sums_batch(const FftWork* w, int nrve, int nblocks, int nvals, int ncopy, hipStream_t s,
                            hipEvent_t done);
// Block shapes and placement of the register FFT passes, picked by the autotuner (stages:
// 0 x forward, 1 y forward, 2 z + Gamma, 3 y inverse, 4 x inverse); none changes the results.
struct FftTune {
    int xcd[5] = {1, 1, 1, 1, 1};   // blocks placed so each XCD gets a contiguous share of the tiles
    int yhalf[2] = {0, 0};          // y passes (forward, inverse) with half-size blocks
    int kt1 = 0;                    // z pass with one kx column a block
    int tuned = 0;                  // 1: picked by timing (or read from a tuning file)
    float us[5] = {0, 0, 0, 0, 0};  // time per launch of the chosen shapes (microseconds)
};
FftCtx* fft_create(const Fields& f, char* err, int errlen);
void fft_destroy(FftCtx* c);
// PK1 -> Gamma convolution -> dLref; writes the 9 block partial sums of
// w dLref F^-1 to f.part, column-major (part[i * nblocks + b]), and returns the number of blocks.
// scale: the mixed form's stress scale (1 in FP64), used by the LDS kernels; the register
// kernels read {1/scale, scale/V} from the device array dscale (the pass sums write it), so a
// captured stage stays valid when the scale changes.
// ev0, ev1 (or null): events recorded at the start of the first and the end of the last register
// pass (the stage's timers, no marker packets between the kernels)
int fft_green(const Fields& f, FftCtx* c, const double* sgsum_dev, const double* c0, const double* s0, int igamma,
              int mixed, double scale, const double* dscale, hipStream_t s, hipEvent_t ev0 = nullptr,
              hipEvent_t ev1 = nullptr);
// Passes st0 .. st1-1 of the register form (0 x forward, 1 y forward, 2 z + Gamma, 3 y inverse,
// 4 x inverse); returns the last pass' block count.  Only for grids with fft_tunable().
int fft_green_passes(const Fields& f, FftCtx* c, const double* sgsum_dev, const double* c0, const double* s0,
                     int igamma, int mixed, const double* dscale, int st0, int st1, hipStream_t s,
                     hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
// Times the candidate shapes of each register pass on stream s (the passes overwrite the
// spectrum, ΔL^ref and the block partials, which the next fft_green recomputes) and keeps the
// fastest; 0 on success or when the lengths use the LDS kernels.
int fft_tune(const Fields& f, FftCtx* c, const double* sgsum_dev, const double* c0, const double* s0, int igamma,
             int mixed, const double* dscale, hipStream_t s);
FftTune fft_get_tune(const FftCtx* c);
// the register passes' twiddle tables (x, y, z) of a context, for its FftWork entry
void fft_twiddles(const FftCtx* c, int mixed, const void** tw);
void fft_set_tune(FftCtx* c, const FftTune& t);
bool fft_tunable(const Fields& f, const FftCtx* c);
// spectral integration of L F for the nodal update: dv (3 x V) and 9 block partial
// sums of L F / V into f.part; returns the number of blocks, *cm = 1 when the partials are
// column-major (part[i * nblocks + b]: the register passes), 0 when row-major.
int fft_integrate(const Fields& f, FftCtx* c, int ideriv, double* dv, hipStream_t s, int* cm);

}  // namespace cg
