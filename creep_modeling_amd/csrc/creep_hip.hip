// creep_hip.hip -- C-ABI of libcreep_hip.so (include/creep_hip.h): context
// management, host-side macroscopic updates and the per-increment / per-pass
// orchestration of the HIP kernels on one stream.
//
// Host-side restatements here: calc_velgradavg_corr + state_6x6_evpsc
// (various_functions.f90:1526-1705), the dL_iter / mixing rule and error
// normalisation (LS-EVPFFT.f90:313-343), the macro averages
// (LS-EVPFFT.f90:460-523) and the macro tangents of calc_c0 (:2633-2647).
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/creep_hip.h"
#include "internal.h"
#include "tensor.h"

using namespace cg;

namespace {

struct HostPhase {
    int igas = 0, nsyst = 0, nmodes = 0;
    int nsm[MAXMODES] = {0};
    std::vector<double> dnca, dbca, tau, thet, nrs, gamd0, hpfac, hard;
    double C[81] = {0};
    bool set = false;
};

// b-basis expansion of a 6x6 into a 4th-order tensor: C4 = sum_ab c_ab b_a (x) b_b
void to4(const double* c66, double* c4) {
    double B[6][9];
    for (int a = 0; a < 6; a++) basis(a, B[a]);
    for (int p = 0; p < 9; p++)
        for (int q = 0; q < 9; q++) {
            double s = 0.0;
            for (int a = 0; a < 6; a++) {
                if (B[a][p] == 0.0) continue;
                for (int b = 0; b < 6; b++) s += c66[6 * a + b] * B[a][p] * B[b][q];
            }
            c4[9 * p + q] = s;
        }
}
// b-basis components of a 4th-order tensor: c_ab = b_a : C4 : b_b
void from4(const double* c4, double* c66) {
    double B[6][9];
    for (int a = 0; a < 6; a++) basis(a, B[a]);
    for (int a = 0; a < 6; a++)
        for (int b = 0; b < 6; b++) {
            double s = 0.0;
            for (int p = 0; p < 9; p++) {
                if (B[a][p] == 0.0) continue;
                for (int q = 0; q < 9; q++) s += B[a][p] * c4[9 * p + q] * B[b][q];
            }
            c66[6 * a + b] = s;
        }
}
const int IJV[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};

// state_6x6_evpsc (various_functions.f90:1638-1705)
bool state6x6(const int* idsim, const int* iscau, const double* mev66, const double* devp0, double* scauchy,
              double* dsim) {
    double profac[6], dv[6], sv[6], m4[81], mv[36], d33[9], devv[6], a6[6], aa[6], A[36];
    for (int i = 0; i < 6; i++) profac[i] = (i >= 3) ? 2.0 : 1.0;
    for (int i = 0; i < 6; i++) {
        dv[i] = dsim[3 * IJV[i][0] + IJV[i][1]];
        sv[i] = scauchy[3 * IJV[i][0] + IJV[i][1]];
    }
    to4(mev66, m4);
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) mv[6 * i + j] = m4[27 * IJV[i][0] + 9 * IJV[i][1] + 3 * IJV[j][0] + IJV[j][1]];
    from6(devp0, d33);
    for (int i = 0; i < 6; i++) devv[i] = d33[3 * IJV[i][0] + IJV[i][1]];
    for (int i = 0; i < 6; i++) a6[i] = dv[i] - devv[i];
    for (int i = 0; i < 6; i++) {
        double r = -1.0 * idsim[i] * a6[i];
        for (int j = 0; j < 6; j++) {
            r += mv[6 * i + j] * iscau[j] * sv[j] * profac[j];
            A[6 * i + j] = iscau[j] * (i == j ? 1.0 : 0.0) - idsim[j] * mv[6 * i + j] * profac[j];
        }
        aa[i] = r;
    }
    bool ok = solve6(A, aa);
    for (int i = 0; i < 6; i++) {
        a6[i] = idsim[i] * a6[i] + iscau[i] * aa[i];
        sv[i] = iscau[i] * sv[i] + idsim[i] * aa[i];
    }
    for (int i = 0; i < 6; i++) dv[i] = a6[i] + devv[i];
    for (int i = 0; i < 6; i++) {
        dsim[3 * IJV[i][0] + IJV[i][1]] = dv[i];
        dsim[3 * IJV[i][1] + IJV[i][0]] = dv[i];
        scauchy[3 * IJV[i][0] + IJV[i][1]] = sv[i];
        scauchy[3 * IJV[i][1] + IJV[i][0]] = sv[i];
    }
    return ok;
}

void sym3(const double* a, double* s) {
    double t[9];
    t[0] = a[0]; t[4] = a[4]; t[8] = a[8];
    t[1] = t[3] = 0.5 * (a[1] + a[3]);
    t[2] = t[6] = 0.5 * (a[2] + a[6]);
    t[5] = t[7] = 0.5 * (a[5] + a[7]);
    for (int i = 0; i < 9; i++) s[i] = t[i];
}
void skw3(const double* a, double* s) {
    double t[9];
    t[0] = t[4] = t[8] = 0.0;
    t[1] = 0.5 * (a[1] - a[3]); t[3] = -t[1];
    t[2] = 0.5 * (a[2] - a[6]); t[6] = -t[2];
    t[5] = 0.5 * (a[5] - a[7]); t[7] = -t[5];
    for (int i = 0; i < 9; i++) s[i] = t[i];
}

}  // namespace

struct cg_ctx {
    int n1, n2, n3, nph, modal, do_mix, mixed;
    long V;
    double complgas;
    char err[512];
    HostPhase ph[MAXPH];
    int ns = 0;
    std::vector<double> euler;
    std::vector<int> grain, phase;
    bool voxels_set = false, initialised = false;
    Fields f;
    PhaseDev* d_ph = nullptr;
    double* d_tab = nullptr;
    FftCtx* fft = nullptr;
    hipStream_t s = nullptr;
    double* h_red = nullptr;   // pinned
    // a pass' sums as the reduction kernel writes them (host-mapped, fine-grained) and its
    // device address
    double* h_pass = nullptr;
    double* h_pass_dev = nullptr;


    double* d_sgsum = nullptr;
    double* d_fluct = nullptr;
    size_t part_cap = 0;
    int sop_valid = 0;     // f.sop holds the operators for sop_tdot and the current c066mod, sg66
    int taun_valid = 0;    // f.taun = |ttau| (false after ttau was written other than by a harden)
    double sop_tdot = 0.0;
    int snap_valid = 0;   // the increment's snapshot was taken (it can roll back)
    int sgsum_valid = 0, lold_same = 0, int_pow = 0, sym_tau = 0, small_ns = 0, ngas = 0, n3node = 1;
    long last_solid = -1;
    double nsys_solid = 0.0;
    // macroscopic state (micro_type, types.f90:55-94)
    double c066[36], c0[81], s0[81], de_ds[81], goperr0[81];
    double wphsol, wphgas, wphcsol, wphcgas;
    double defgradavg[9], defgradinvavgc[9], defgradinvavgcs[9], defgradinvavgcg[9];
    double dvelgradavg[9], velgradavg[9], velgradavgsol[9], velgradavggas[9];
    double sgavg[9], sgavgsol[9], sgsolavg[9];
    double svmavg, dvmavg, lvmavg, evmavg, edotvmmx;
    double epavg[9], eavg[9], eelavg[9], edotpavg[9], evmpavg, dvmpavg;
    double dsim_snap[9], scauchy_snap[9];
    // timing
    int timing = 0;
    hipEvent_t ev[12];
    cg_stats st;
    // FFT shape tuning: done at the first pass (or given by cg_set_fft_tuning), cached in the
    // tuning file when one is set
    int fft_tuned = 0;
    std::string tune_path;
    // initial displacement field (cg_set_displacement): 0 none, 1 cell values, 2 node values
    int idispini = 0;
    std::vector<double> dispini;
    // the mixed form's stress scale as the register FFT kernels read it: {1/scale, scale/V}
    // the mixed form's stress scale {1/s, s/V} as the register FFT kernels read it (the pass sums
    // write it: s = max(svm, 1) of the pass' solid stress sums; 1 before the first pass) and the
    // pass sums' last-block counter
    double* d_scale = nullptr;
    unsigned* d_cnt = nullptr;
    // HIP graphs of a pass' fixed device work: pre = FFT + Gamma stage and the sums of
    // w dLref F^-1; post = the solve's sums, <sigma> for the next pass and their copy to the host.
    // pre is captured again when C0 / S0, igamma or the FFT shapes change.  Off by default: the
    // stream never drains between a pass' kernels, and a graph launch delays the first kernel of
    // the stage (1% slower at 128^3, no gain on the 32^3-96^3 sweep).
    int use_graphs = 0;
    // Overlap of the host's part of a pass with the next pass' transforms: after cg_iterate has
    // enqueued a pass' sums (written straight to host memory), it enqueues the x, y, z + Gamma
    // and y passes of the next pass (they read only sigma, <sigma>, F^-1, J, C0 and, in the
    // mixed form, the scale that the device derives from the same sums) and waits for the sums
    // alone.  The next cg_iterate runs only the x inverse then, if nothing changed the state in
    // between (any other entry point that writes state drops the early passes, and the next
    // pass runs all five).  The x inverse, the only pass that writes a readable field
    // (dvelgradref), always runs in its own pass.
    int overlap = 1;
    int early = 0;                  // the next pass' first four passes are enqueued
    int early_igamma = -1, early_tune = -1;
    long early_green = -1;
    int early_slot = 0;             // timing events of the early passes: ev[8 + 2 slot], ev[9 + 2 slot]
    int early_timed[2] = {0, 0};
    hipEvent_t ev_sums = nullptr;   // the pass' sums have reached the host
    hipGraphExec_t g_pre = nullptr, g_post = nullptr;
    long green_version = 0, g_pre_version = -1;
    int g_pre_igamma = -1, tune_version = 0, g_pre_tune = -1;
};

static void drop_early(cg_ctx* c);

static int fail(cg_ctx* c, int code, const char* fmt, ...) {
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->err, sizeof c->err, fmt, ap);
        va_end(ap);
    }
    return code;
}
#define HIPCHK(c, x)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (x);                                                                  \
        if (e_ != hipSuccess) return fail((c), -2, "%s: %s", #x, hipGetErrorString(e_));      \
    } while (0)

static int dalloc(cg_ctx* c, double** p, size_t n) {
    if (n == 0) { *p = nullptr; return 0; }
    hipError_t e = hipMalloc(p, n * sizeof(double));
    if (e != hipSuccess) return fail(c, -3, "hipMalloc(%zu doubles): %s", n, hipGetErrorString(e));
    return 0;
}

static int ensure_part(cg_ctx* c, size_t n) {
    if (n <= c->part_cap) return 0;
    if (c->f.part) hipFree(c->f.part);
    c->f.part = nullptr;
    if (dalloc(c, &c->f.part, n)) return -3;
    c->part_cap = n;
    return 0;
}

// The singular-system flag the kernels raise (R_FLAG_SING) as the reference's `stop 1`:
// returns 1 with the voxel's 1-based (i, j, k) and clears the flag, or 0.
static int singular_from(cg_ctx* c, const double* h, const char* where) {
    if (h[R_FLAG_SING] == 0.0) return 0;
    const long v = (long)h[R_FLAG_VOX];
    launch_fill(c->f.red + R_FLAG_VOX, 0.0, 2, c->s);
    return fail(c, 1, "Singular matrix in lu_inverse / lu_eqsystem (%s) at voxel (%ld,%ld,%ld)", where,
                v % c->n1 + 1, (v / c->n1) % c->n2 + 1, v / ((long)c->n1 * c->n2) + 1);
}
static int check_singular(cg_ctx* c, const char* where) {
    HIPCHK(c, hipMemcpyAsync(c->h_red + R_FLAG_VOX, c->f.red + R_FLAG_VOX, 2 * sizeof(double),
                             hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return singular_from(c, c->h_red, where);
}

// reduce nvals partial columns into host_out[0..nvals) (synchronous); h_red then holds red[0,
// R_NCOPY), the singular-system flags included
// done: red already holds the sums (a column-major reduction ran)
static int reduce_to_host(cg_ctx* c, int nblocks, int nvals, double* host_out, int done = 0) {
    if (!done) launch_reduce_final(c->f.part, nblocks, nvals, c->f.red, c->s);
    HIPCHK(c, hipMemcpyAsync(c->h_red, c->f.red, R_NCOPY * sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    memcpy(host_out, c->h_red, nvals * sizeof(double));
    return 0;
}

// ---------------------------------------------------------------------------
// Device scratch of the no-allocation entry point (cg_kinematic_substep_noalloc): the
// block partials of its reductions, 4 per 256-voxel block, for grids of at most
// NOALLOC_MAX_VOX voxels (256^3): 2 MiB on every GPU the library runs on, held until the
// process ends.  Host code reaches it only through the per-device address that
// hipGetSymbolAddress returns (the array's name is a host-side stand-in).  Every call on a
// GPU shares that GPU's array: calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr long NOALLOC_MAX_VOX = 1L << 24;
constexpr long NOALLOC_PART = 4 * (NOALLOC_MAX_VOX / 256);
static __device__ __align__(256) double scratch[NOALLOC_PART];
constexpr int NOALLOC_MAX_DEV = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[NOALLOC_MAX_DEV];
static thread_local char g_noctx_err[256] = "";

static int noctx_fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_noctx_err, sizeof g_noctx_err, fmt, ap);
    va_end(ap);
    return code;
}

// The current device's copy of `scratch`, resolved once per device with that device current.
static int scratch_on_current_device(double** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEV)
        return noctx_fail(-4, "no usable current device (%d)", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return noctx_fail(-2, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

extern "C" {

int cg_kinematic_substep_noalloc(long nvox, const double* velgrad, long ld, const double* detF, double tdot,
                                 double* out5, void* stream) {
    if (nvox < 1 || ld < nvox || !velgrad || !detF || !out5)
        return noctx_fail(-1, "cg_kinematic_substep_noalloc: bad arguments (nvox=%ld, ld=%ld)", nvox, ld);
    if (nvox > NOALLOC_MAX_VOX)
        return noctx_fail(-6, "cg_kinematic_substep_noalloc: %ld voxels exceed the scratch capacity of %ld voxels",
                          nvox, NOALLOC_MAX_VOX);
    double* sp = nullptr;
    if (int e = scratch_on_current_device(&sp)) return e;
    launch_kinematic_raw(velgrad, ld, detF, nvox, tdot, sp, out5, (hipStream_t)stream);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return noctx_fail(-2, "cg_kinematic_substep_noalloc: %s", hipGetErrorString(e));
    return 0;
}

int cg_create(int n1, int n2, int n3, int nph, double complgas, int modal, int do_mix, int precision,
              cg_ctx** out) {
    if (!out) return -1;
    *out = nullptr;
    if (n1 < 2 || n2 < 1 || n3 < 1 || nph < 1 || nph > MAXPH) return -1;
    // field columns are addressed through buffer resources over 8 columns: 64 ld < 2^32
    if ((long)n1 * n2 * n3 >= (1L << 26) - 8192) return -1;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return -4;
    cg_ctx* c = new cg_ctx();
    memset(c->err, 0, sizeof c->err);
    c->n1 = n1; c->n2 = n2; c->n3 = n3; c->nph = nph;
    c->V = (long)n1 * n2 * n3;
    c->complgas = complgas; c->modal = modal; c->do_mix = do_mix; c->mixed = precision ? 1 : 0;
    memset(&c->f, 0, sizeof c->f);
    memset(&c->st, 0, sizeof c->st);
    if (hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess) { delete c; return -2; }
    for (int i = 0; i < 12; i++) hipEventCreate(&c->ev[i]);
    hipEventCreateWithFlags(&c->ev_sums, hipEventDisableTiming);
    *out = c;
    return 0;
}

// Column stride of the per-voxel fields: V rounded up to 4096 doubles plus 256 (2 KiB), so that
// the columns of a field (and of different fields) do not all start at the same offset modulo
// the HBM channel interleave; at power-of-two V an unpadded stride serialises a kernel's column
// streams on a few channels.
static long field_stride(long V) { return (V + 4095) / 4096 * 4096 + 256; }

static void free_fields(cg_ctx* c) {
    Fields& f = c->f;
    void* ptrs[] = {f.phase, f.gas, f.w, f.sg, f.sgt, f.L, f.Lold, f.dLref, f.F, f.Fi, f.J, f.Fe, f.Fp, f.ag,
                    f.sg66, f.c66m, f.cgm, f.cgas, f.sop, f.sch, f.schA, f.crss, f.ttau, f.taun, f.gdot, f.gacc, f.edotp, f.ept, f.dgs,
                    f.epvm, f.Finc, f.Feinc, f.eth, f.Fini, f.s_L, f.s_Lold, f.s_sg, f.s_edotp, f.s_gdot, f.s_ttau, f.spec,
                    f.specf, f.xnode, f.part, f.red};
    for (void* p : ptrs)
        if (p) hipFree(p);
    memset(&c->f, 0, sizeof c->f);
}

int cg_destroy(cg_ctx* c) {
    if (!c) return 0;
    if (c->s) hipStreamSynchronize(c->s);
    free_fields(c);
    if (c->d_ph) hipFree(c->d_ph);
    if (c->d_tab) hipFree(c->d_tab);
    if (c->d_sgsum) hipFree(c->d_sgsum);
    if (c->d_fluct) hipFree(c->d_fluct);
    if (c->h_red) hipHostFree(c->h_red);
    if (c->h_pass) hipHostFree(c->h_pass);
    if (c->d_scale) hipFree(c->d_scale);
    if (c->d_cnt) hipFree(c->d_cnt);
    if (c->g_pre) hipGraphExecDestroy(c->g_pre);
    if (c->g_post) hipGraphExecDestroy(c->g_post);
    if (c->fft) fft_destroy(c->fft);
    for (int i = 0; i < 12; i++) hipEventDestroy(c->ev[i]);
    if (c->ev_sums) hipEventDestroy(c->ev_sums);
    if (c->s) hipStreamDestroy(c->s);
    delete c;
    return 0;
}

const char* cg_last_error(cg_ctx* c) { return c ? c->err : g_noctx_err; }

int cg_set_phase(cg_ctx* c, int iph, int igas, int nsyst, int nmodes, const int* nsm, const double* dnca,
                 const double* dbca, const double* tau, const double* thet, const double* nrs,
                 const double* gamd0, const double* hpfac, const double* hard, const double* C) {
    drop_early(c);
    if (!c) return -1;
    if (iph < 1 || iph > c->nph) return fail(c, -1, "phase index %d out of range", iph);
    HostPhase& P = c->ph[iph - 1];
    P.igas = igas;
    P.set = true;
    if (igas) return 0;
    if (nsyst < 1 || nsyst > MAXSYS || nmodes < 1 || nmodes > MAXMODES) return fail(c, -1, "bad phase sizes");
    int tot = 0;
    for (int m = 0; m < nmodes; m++) { P.nsm[m] = nsm[m]; tot += nsm[m]; }
    if (tot != nsyst) return fail(c, -1, "sum(nsm) != nsyst");
    P.nsyst = nsyst;
    P.nmodes = nmodes;
    P.dnca.assign(dnca, dnca + 3 * nsyst);
    P.dbca.assign(dbca, dbca + 3 * nsyst);
    P.tau.assign(tau, tau + 3 * nsyst);
    P.thet.assign(thet, thet + 2 * nsyst);
    P.nrs.assign(nrs, nrs + nsyst);
    P.gamd0.assign(gamd0, gamd0 + nsyst);
    P.hpfac.assign(hpfac, hpfac + nsyst);
    P.hard.assign(hard, hard + (size_t)nsyst * nsyst);
    memcpy(P.C, C, sizeof P.C);
    c->ns = std::max(c->ns, nsyst);
    return 0;
}

int cg_set_voxels(cg_ctx* c, const double* euler, const int* grain, const int* phase) {
    drop_early(c);
    if (!c) return -1;
    c->euler.assign(euler, euler + 3 * c->V);
    c->grain.assign(grain, grain + c->V);
    c->phase.assign(phase, phase + c->V);
    for (long v = 0; v < c->V; v++)
        if (c->phase[v] < 1 || c->phase[v] > c->nph) return fail(c, -1, "phase id out of range at voxel %ld", v);
    c->voxels_set = true;
    return 0;
}

int cg_init(cg_ctx* c, const cg_process* p1) {
    drop_early(c);
    if (!c || !p1) return -1;
    if (!c->voxels_set) return fail(c, -1, "cg_set_voxels not called");
    for (int i = 0; i < c->nph; i++)
        if (!c->ph[i].set) return fail(c, -1, "phase %d not set", i + 1);
    const long V = c->V;
    Fields& f = c->f;
    f.n1 = c->n1; f.n2 = c->n2; f.n3 = c->n3; f.h1 = c->n1 / 2 + 1; f.V = V;
    f.ld = field_stride(V);
    const long LD = f.ld;
    f.ns = std::max(c->ns, 1);
    const size_t ns = f.ns;
    // phase tables
    PhaseDev hp[MAXPH];
    memset(hp, 0, sizeof hp);
    std::vector<double> tab;
    int all_int = 1, all_sym = 1, small_ns = 1;
    for (int i = 0; i < c->nph; i++) {
        HostPhase& P = c->ph[i];
        PhaseDev& D = hp[i];
        D.igas = P.igas;
        if (P.igas) continue;
        D.nsyst = P.nsyst;
        D.nmodes = P.nmodes;
        int last = 0;
        for (int m = 0; m < P.nmodes; m++) {
            D.nsm[m] = P.nsm[m];
            last += P.nsm[m];
            // harden reads the mode parameters from the mode's last system (various_functions.f90:1466-1470)
            D.mode_tau1[m] = P.tau[3 * (last - 1) + 2];
            D.mode_thet0[m] = P.thet[2 * (last - 1)];
            D.mode_thet1[m] = P.thet[2 * (last - 1) + 1];
        }
        from4(P.C, D.c66cr);
        auto put = [&](const std::vector<double>& v) { int off = (int)tab.size(); tab.insert(tab.end(), v.begin(), v.end()); return off; };
        D.off_dnca = put(P.dnca);
        D.off_dbca = put(P.dbca);
        D.off_nrs = put(P.nrs);
        D.off_gamd0 = put(P.gamd0);
        D.off_hard = put(P.hard);
        D.off_tau = put(P.tau);
        D.off_hpfac = put(P.hpfac);
        {
            std::vector<double> rec(8 * (size_t)P.nsyst);
            for (int is = 0; is < P.nsyst; is++) {
                for (int q = 0; q < 3; q++) {
                    rec[8 * is + q] = P.dbca[3 * is + q];
                    rec[8 * is + 3 + q] = P.dnca[3 * is + q];
                }
                rec[8 * is + 6] = P.nrs[is];
                rec[8 * is + 7] = P.gamd0[is];
            }
            if (tab.size() % 2) tab.push_back(0.0);  // records on 16-byte boundaries
            D.off_sys = put(rec);
        }
        for (double n : P.nrs)
            if (!(n == floor(n) && n >= 1.0 && n <= 64.0)) all_int = 0;
        for (int is = 0; is < P.nsyst; is++)
            if (P.tau[3 * is] != P.tau[3 * is + 1]) all_sym = 0;
        if (P.nsyst > 12 || P.nmodes != 1) small_ns = 0;  // the one-mode, at most 12 systems solve
    }
    c->int_pow = all_int;
    c->sym_tau = all_sym;
    c->small_ns = small_ns;
    tab.insert(tab.end(), 16, 0.0);  // a row of the interaction matrix may be read 12 wide
    HIPCHK(c, hipMalloc(&c->d_ph, sizeof hp));
    HIPCHK(c, hipMemcpy(c->d_ph, hp, sizeof hp, hipMemcpyHostToDevice));
    if (dalloc(c, &c->d_tab, std::max<size_t>(tab.size(), 1))) return -3;
    if (!tab.empty()) HIPCHK(c, hipMemcpy(c->d_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));

    // per-voxel ids
    std::vector<uint8_t> hph(V), hgas(V);
    c->ngas = 0;
    c->last_solid = -1;
    c->nsys_solid = 0.0;
    for (long v = 0; v < V; v++) {
        hph[v] = (uint8_t)(c->phase[v] - 1);
        hgas[v] = (uint8_t)(c->ph[c->phase[v] - 1].igas ? 1 : 0);
        if (hgas[v]) c->ngas++;
        else { c->last_solid = v; c->nsys_solid += c->ph[c->phase[v] - 1].nsyst; }
    }
    HIPCHK(c, hipMalloc(&f.phase, V));
    HIPCHK(c, hipMalloc(&f.gas, V));
    HIPCHK(c, hipMemcpy(f.phase, hph.data(), V, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(f.gas, hgas.data(), V, hipMemcpyHostToDevice));
    int rc = 0;
    rc |= dalloc(c, &f.w, LD); rc |= dalloc(c, &f.sg, 6 * LD); rc |= dalloc(c, &f.sgt, 6 * LD);
    rc |= dalloc(c, &f.L, 9 * LD); rc |= dalloc(c, &f.Lold, 9 * LD); rc |= dalloc(c, &f.dLref, 9 * LD);
    rc |= dalloc(c, &f.F, 9 * LD); rc |= dalloc(c, &f.Fi, 9 * LD); rc |= dalloc(c, &f.J, LD);
    rc |= dalloc(c, &f.Fe, 9 * LD); rc |= dalloc(c, &f.Fp, 9 * LD); rc |= dalloc(c, &f.ag, 9 * LD);
    rc |= dalloc(c, &f.sg66, 36 * LD); rc |= dalloc(c, &f.c66m, 36 * LD);
    if (c->modal) rc |= dalloc(c, &f.cgm, 36 * LD);
    else rc |= dalloc(c, &f.sop, 42 * LD);
    if (c->ngas) rc |= dalloc(c, &f.cgas, 36 * LD);
    rc |= dalloc(c, &f.sch, 5 * ns * LD); rc |= dalloc(c, &f.schA, 18 * LD); rc |= dalloc(c, &f.crss, 2 * ns * LD); rc |= dalloc(c, &f.ttau, 2 * ns * LD);
    rc |= dalloc(c, &f.taun, LD);
    rc |= dalloc(c, &f.gdot, ns * LD); rc |= dalloc(c, &f.gacc, LD);
    rc |= dalloc(c, &f.edotp, 6 * LD); rc |= dalloc(c, &f.ept, 6 * LD); rc |= dalloc(c, &f.dgs, 6 * LD);
    rc |= dalloc(c, &f.epvm, LD); rc |= dalloc(c, &f.Finc, 9 * LD); rc |= dalloc(c, &f.Feinc, 9 * LD);
    rc |= dalloc(c, &f.s_L, 9 * LD); rc |= dalloc(c, &f.s_Lold, 9 * LD); rc |= dalloc(c, &f.s_sg, 6 * LD);
    rc |= dalloc(c, &f.s_edotp, 6 * LD); rc |= dalloc(c, &f.s_gdot, ns * LD); rc |= dalloc(c, &f.s_ttau, 2 * ns * LD);
    const size_t nspec = (size_t)f.h1 * f.n2 * f.n3 * 9;
    // the FP64 spectrum is also used by the nodal integration in the mixed form
    if (c->mixed) HIPCHK(c, hipMalloc(&f.specf, nspec * sizeof(float2)));
    HIPCHK(c, hipMalloc(&f.spec, nspec * sizeof(double2)));
    c->n3node = f.n3 == 1 ? 1 : f.n3 + 1;
    rc |= dalloc(c, &f.xnode, 3 * (size_t)(f.n1 + 1) * (f.n2 + 1) * c->n3node);
    rc |= dalloc(c, &f.red, 256);
    if (rc) return rc;
    launch_fill(f.red, 0.0, 256, c->s);  // the status slots start clear
    if (ensure_part(c, (size_t)((V + 63) / 64 + 64) * 81)) return -3;  // the solve may run 64-thread blocks
    HIPCHK(c, hipHostMalloc(&c->h_red, 256 * sizeof(double)));
    HIPCHK(c, hipHostMalloc(&c->h_pass, R_NCOPY * sizeof(double), hipHostMallocMapped | hipHostMallocCoherent));
    HIPCHK(c, hipHostGetDevicePointer((void**)&c->h_pass_dev, c->h_pass, 0));
    if (dalloc(c, &c->d_sgsum, 16) || dalloc(c, &c->d_fluct, 16) || dalloc(c, &c->d_scale, 2)) return -3;
    {
        const double sc0[2] = {1.0, 1.0 / (double)V};
        HIPCHK(c, hipMemcpy(c->d_scale, sc0, sizeof sc0, hipMemcpyHostToDevice));
    }
    HIPCHK(c, hipMalloc(&c->d_cnt, sizeof(unsigned)));
    HIPCHK(c, hipMemset(c->d_cnt, 0, sizeof(unsigned)));
    c->fft = fft_create(f, c->err, sizeof c->err);
    if (!c->fft) return -5;

    // voxel initialisation
    int gmax = 0;
    for (long v = 0; v < V; v++) gmax = std::max(gmax, c->grain[v]);
    std::vector<long> nvox((size_t)(gmax + 1) * c->nph, 0);
    for (long v = 0; v < V; v++)
        if (!hgas[v]) {
            if (c->grain[v] < 1) return fail(c, -1, "grain id out of bounds for Hall-Petch init");
            nvox[(size_t)c->grain[v] * c->nph + c->phase[v] - 1]++;
        }
    // Hall-Petch shift 252.8/sqrt(max((64 N)^(1/3), 12)) with N the voxel count of the voxel's
    // (grain, phase) pair (IO_functions.f90:107-169: a grain id spread over several phases has
    // one size per phase), tabulated per voxel
    std::vector<double> hpg((size_t)(gmax + 1) * c->nph, 0.0), hpv(V, 0.0);
    const double hpk = 632.0 / 2.5, vvol = 64.0, dmin = 3.0 * pow(vvol, 1.0 / 3.0);
    for (size_t q = 0; q < hpg.size(); q++)
        if (nvox[q] > 0) hpg[q] = hpk / sqrt(std::max(pow((double)nvox[q] * vvol, 1.0 / 3.0), dmin));
    for (long v = 0; v < V; v++)
        if (!hgas[v]) hpv[v] = hpg[(size_t)c->grain[v] * c->nph + c->phase[v] - 1];
    double *d_eul = nullptr, *d_hp = nullptr;
    int* d_gr = nullptr;
    if (dalloc(c, &d_eul, 3 * V) || dalloc(c, &d_hp, V)) return -3;
    HIPCHK(c, hipMalloc(&d_gr, V * sizeof(int)));
    HIPCHK(c, hipMemcpy(d_eul, c->euler.data(), 3 * V * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(d_hp, hpv.data(), V * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(d_gr, c->grain.data(), V * sizeof(int), hipMemcpyHostToDevice));
    PhaseTables pt{c->d_ph, c->d_tab};
    int nb;
    launch_init_xnode(f, c->n3node, c->s);
    // init_defgrad (IO_functions.f90:815-936), between data_grain and the weighted initial C0:
    // F, F0, J, F^-1 and w_c from the initial displacement field; delt is (4, 4, 4), the value
    // load_input gives it whatever the file says (IO_functions.f90:62-66)
    double dsum[30];
    if (c->idispini > 0) {
        double* d_disp = nullptr;
        if (!f.Fini && dalloc(c, &f.Fini, 9 * f.ld)) return -3;
        if (dalloc(c, &d_disp, c->dispini.size())) return -3;
        HIPCHK(c, hipMemcpy(d_disp, c->dispini.data(), c->dispini.size() * sizeof(double), hipMemcpyHostToDevice));
        launch_init_defgrad(f, d_disp, c->idispini, 4.0, c->s, &nb);
        launch_xnode_disp(f, d_disp, c->idispini, c->n3node, c->s);
        const int rr = reduce_to_host(c, nb, 30, dsum);
        hipFree(d_disp);
        if (rr) return -2;
        if (int e = check_singular(c, "init_defgrad: lu_inverse of F")) return e;
        launch_div_w(f, dsum[29], c->s);  // w_c = wgt J / <J>
    }
    launch_init_voxels(f, pt, d_eul, d_gr, d_hp, p1->udot, p1->tdot, c->idispini > 0, c->s, &nb);
    c->taun_valid = 0;
    double sums[36];
    if (reduce_to_host(c, nb, 36, sums)) return -2;
    hipFree(d_eul); hipFree(d_hp); hipFree(d_gr);
    if (int e = check_singular(c, "data_grain: sg66 = cg66^-1")) return e;
    // wphsol / wphgas (IO_functions.f90:347-363) and micro-level defaults (types.f90:338-341)
    c->wphsol = 0.0; c->wphgas = 0.0;
    const double wgt = 1.0 / (double)V;
    for (long v = 0; v < V; v++) { if (hgas[v]) c->wphgas += wgt; else c->wphsol += wgt; }
    c->wphcsol = c->wphsol; c->wphcgas = c->wphgas;
    const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    memcpy(c->defgradavg, I3, sizeof I3); memcpy(c->defgradinvavgc, I3, sizeof I3);
    memcpy(c->defgradinvavgcs, I3, sizeof I3); memcpy(c->defgradinvavgcg, I3, sizeof I3);
    if (c->idispini > 0) {  // init_defgrad's averages (IO_functions.f90:887-935)
        const double dfa = dsum[29];
        c->wphcsol = dsum[27] / dfa;
        c->wphcgas = dsum[28] / dfa;
        memset(c->defgradavg, 0, sizeof c->defgradavg);
        c->defgradavg[0] = c->defgradavg[4] = c->defgradavg[8] = 4.0;
        for (int t = 0; t < 9; t++) {
            c->defgradinvavgc[t] = dsum[t] / dfa;
            c->defgradinvavgcs[t] = dsum[9 + t] / dfa / c->wphcsol;
            c->defgradinvavgcg[t] = c->wphcgas > 0.0 ? dsum[18 + t] / dfa / c->wphcgas : I3[t];
        }
    }
    memset(c->dvelgradavg, 0, sizeof c->dvelgradavg);
    memset(c->velgradavg, 0, sizeof c->velgradavg);
    memset(c->sgavg, 0, sizeof c->sgavg);
    memset(c->sgsolavg, 0, sizeof c->sgsolavg);
    memset(c->goperr0, 0, sizeof c->goperr0);
    memset(c->epavg, 0, sizeof c->epavg); memset(c->eavg, 0, sizeof c->eavg);
    memset(c->eelavg, 0, sizeof c->eelavg); memset(c->edotpavg, 0, sizeof c->edotpavg);
    c->svmavg = 0.0; c->evmavg = 0.0; c->edotvmmx = 0.0; c->evmpavg = 0.0; c->dvmpavg = 0.0;
    // initial_guess tail: dvmavg = vm(dsim), lvmavg = dvmavg (various_functions.f90:50-51)
    c->dvmavg = vm_strain(p1->dsim);
    c->lvmavg = c->dvmavg;
    // initial C0 (LS-EVPFFT.f90:156-173)
    double s066[36];
    for (int i = 0; i < 36; i++) c->c066[i] = sums[i] / c->wphcsol * p1->tdot;
    memcpy(s066, c->c066, sizeof s066);
    bool ok = inv6(s066);
    to4(c->c066, c->c0);
    to4(s066, c->s0);
    c->green_version++;
    memcpy(c->de_ds, c->s0, sizeof c->s0);
    c->sgsum_valid = 0;
    c->lold_same = 0;
    c->initialised = true;
    HIPCHK(c, hipGetLastError());
    if (!ok) return fail(c, 1, "Singular matrix in lu_inverse (initial C0)");
    return 0;
}

// Timing (cg_set_timing): 1 = every stage (marker events around the FFT + Gamma stage and the
// sums, and around the per-increment work), 2 = the constitutive solve and the per-increment work
// only.  The solve's events are attached to its launch in both modes; slot k uses ev[2k], ev[2k+1].
static bool timed(const cg_ctx* c, int k) { return c->timing == 1 || (c->timing == 2 && k == 3); }
static void tstart(cg_ctx* c, int k) {
    if (timed(c, k)) hipEventRecord(c->ev[2 * k], c->s);
}
static void tstop(cg_ctx* c, int k) {
    if (timed(c, k)) hipEventRecord(c->ev[2 * k + 1], c->s);
}
static void tcollect(cg_ctx* c, int k) {
    if (!c->timing || (k != 1 && !timed(c, k))) return;
    float ms = 0.0f;
    hipEventSynchronize(c->ev[2 * k + 1]);
    hipEventElapsedTime(&ms, c->ev[2 * k], c->ev[2 * k + 1]);
    c->st.t_ms[k] += ms;
}
// the early passes' time (slot 0 or 1) counts as FFT + Gamma stage time
static void tcollect_early(cg_ctx* c, int slot) {
    if (!c->early_timed[slot]) return;
    c->early_timed[slot] = 0;
    float ms = 0.0f;
    hipEventSynchronize(c->ev[9 + 2 * slot]);
    hipEventElapsedTime(&ms, c->ev[8 + 2 * slot], c->ev[9 + 2 * slot]);
    c->st.t_ms[0] += ms;
}
// Drops the early passes of the next pass (the state they read is about to change, or the
// host runs something else next).  Their writes (the spectrum) are overwritten by the next
// pass' own transforms, in stream order; their time is not counted as a pass' stage time.
static void drop_early(cg_ctx* c) {
    if (!c) return;
    c->early = 0;
    c->early_timed[0] = c->early_timed[1] = 0;
}

int cg_begin_increment(cg_ctx* c, const cg_process* p, int do_schmid) {
    drop_early(c);
    if (!c || !c->initialised || !p) return -1;
    tstart(c, 3);
    memcpy(c->dsim_snap, p->dsim, sizeof c->dsim_snap);
    memcpy(c->scauchy_snap, p->scauchy, sizeof c->scauchy_snap);
    // take_snapshot copies 66 columns per voxel (1.1 GB at 128^3).  Only an increment under
    // Delta-t control with tdot > tdotmin can roll back (LS-EVPFFT.f90:361-368: the cut-back
    // needs tdot > tdotmin, and tdot only shrinks within an increment), so only those copy
    // every column; the others copy L alone, which the diagnostics compare with the increment's
    // end.  cg_rollback refuses an increment without the full snapshot.
    c->snap_valid = (p->ictrl == -1 && p->tdot > p->tdotmin) ? 1 : 0;
    launch_snapshot(c->f, 0, c->s, c->snap_valid);
    PhaseTables pt{c->d_ph, c->d_tab};
    if (do_schmid) launch_update_schmid(c->f, pt, c->modal, c->s);
    if (c->modal) {  // form_G: Goperr0 = (1/V) sum Gamma(xi)
        int nb;
        launch_goperr0(c->f, c->c0, c->s0, p->igamma, c->s, &nb);
        double g[81];
        if (reduce_to_host(c, nb, 81, g)) return -2;
        for (int i = 0; i < 81; i++) c->goperr0[i] = g[i] / (double)c->V;
    }
    launch_c066mod(c->f, c->c066, c->goperr0, c->modal, c->s);
    c->sop_valid = 0;
    tstop(c, 3);
    HIPCHK(c, hipGetLastError());
    tcollect(c, 3);
    return 0;
}

}  // extern "C"

// The FFT shape tuning of a context, done before its first pass.  The tuning file (the
// reference's FFTW wisdom file, --fftw_save, fourier_functions.f90:35-66, 106-127) holds one
// line per (device, grid, precision, Green-operator kind):
//   creep_hip_fft_tune <arch> <CUs> <n1> <n2> <n3> <precision> <igamma>0> : <xcd x5> <yhalf x2> <kt1>
// A matching line is used as it is; otherwise the shapes are timed and the line appended.
static std::string tune_key(const cg_ctx* c, int igamma) {
    hipDeviceProp_t pr;
    int dev = 0;
    hipGetDevice(&dev);
    char arch[256] = "unknown";
    int cus = 0;
    if (hipGetDeviceProperties(&pr, dev) == hipSuccess) {
        snprintf(arch, sizeof arch, "%s", pr.gcnArchName);
        for (char* q = arch; *q; q++)
            if (*q == ' ' || *q == ':') *q = '_';
        cus = pr.multiProcessorCount;
    }
    char k[512];
    snprintf(k, sizeof k, "creep_hip_fft_tune %s %d %d %d %d %d %d", arch, cus, c->n1, c->n2, c->n3, c->mixed,
             igamma > 0 ? 1 : 0);
    return k;
}
static bool tune_read(const std::string& path, const std::string& key, FftTune& t) {
    FILE* fp = fopen(path.c_str(), "r");
    if (!fp) return false;
    char line[1024];
    bool found = false;
    while (!found && fgets(line, sizeof line, fp)) {
        const size_t kl = key.size();
        if (strncmp(line, key.c_str(), kl) != 0 || strncmp(line + kl, " : ", 3) != 0) continue;
        FftTune u;
        int n = sscanf(line + kl + 3, "%d %d %d %d %d %d %d %d", &u.xcd[0], &u.xcd[1], &u.xcd[2], &u.xcd[3],
                       &u.xcd[4], &u.yhalf[0], &u.yhalf[1], &u.kt1);
        if (n == 8) {
            u.tuned = 1;
            t = u;
            found = true;
        }
    }
    fclose(fp);
    return found;
}
static void tune_write(const std::string& path, const std::string& key, const FftTune& t) {
    FILE* fp = fopen(path.c_str(), "a");
    if (!fp) return;  // an unwritable cache only costs the tuning of the next run
    fprintf(fp, "%s : %d %d %d %d %d %d %d %d\n", key.c_str(), t.xcd[0], t.xcd[1], t.xcd[2], t.xcd[3], t.xcd[4],
            t.yhalf[0], t.yhalf[1], t.kt1);
    fclose(fp);
}
// the mixed form's stress scale for the LDS FFT kernels (the pass sums write the same max(svm, 1)
// into d_scale for the register kernels)
static double pass_scale(const cg_ctx* c) { return c->mixed ? std::max(c->svmavg, 1.0) : 1.0; }
// Captures what body() enqueues on the context's stream into an executable graph.
template <class Body>
static int capture(cg_ctx* c, hipGraphExec_t* ex, Body body) {
    if (*ex) {
        hipGraphExecDestroy(*ex);
        *ex = nullptr;
    }
    HIPCHK(c, hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
    body();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->s, &g);
    if (e != hipSuccess || !g) return fail(c, -2, "graph capture failed: %s", hipGetErrorString(e));
    const hipError_t e2 = hipGraphInstantiate(ex, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e2 != hipSuccess) {
        *ex = nullptr;
        return fail(c, -2, "graph instantiation failed: %s", hipGetErrorString(e2));
    }
    return 0;
}

static int ensure_fft_tuned(cg_ctx* c, const cg_process* p) {
    if (c->fft_tuned) return 0;
    c->fft_tuned = 1;
    if (!fft_tunable(c->f, c->fft)) return 0;
    const std::string key = tune_key(c, p->igamma);
    FftTune t;
    if (!c->tune_path.empty() && tune_read(c->tune_path, key, t)) {
        fft_set_tune(c->fft, t);
        return 0;
    }
    if (!c->sgsum_valid) {  // the x forward pass reads <sigma>
        int nb;
        launch_sigma_sum(c->f, c->s, &nb);
        launch_reduce_final(c->f.part, nb, 9, c->d_sgsum, c->s);
    }
    if (fft_tune(c->f, c->fft, c->d_sgsum, c->c0, c->s0, p->igamma, c->mixed, c->d_scale, c->s) != 0)
        return fail(c, -2, "FFT shape tuning failed");
    c->tune_version++;
    if (!c->tune_path.empty()) tune_write(c->tune_path, key, fft_get_tune(c->fft));
    return 0;
}

// One fixed-point pass is split in three so that several RVEs can share the constitutive
// launch (cg_batch_iterate): pass_begin enqueues the work before it on the context's stream
// and fills the solve's arguments; pass_reduce enqueues the reductions of the solve's sums
// and their copy to the host; pass_finish (after the stream has finished) is the host part.
// run_stage = false: the FFT + Gamma stage and its sums are left to the caller (the batched
// passes of cg_batch_launch run them for several RVEs at once).
static int pass_begin(cg_ctx* c, cg_process* p, cg_iter_state* it, SolveArgs& sa, bool run_stage = true) {
    Fields& f = c->f;
    // the early passes enqueued by the previous cg_iterate hold for this pass if the Green
    // operator, its kind and the FFT shapes are those they ran with, and this is not the first
    // pass of an increment
    const bool early = run_stage && c->early && it->iter > 0 && c->early_igamma == p->igamma &&
                       c->early_green == c->green_version && c->early_tune == c->tune_version;
    if (c->early && !early) drop_early(c);
    c->early = 0;
    if (it->iter == 0) {  // LS-EVPFFT.f90:222-231
        if (c->ngas) launch_gas_compl(f, p->tdot, c->complgas, c->modal, c->s);
        memset(c->dvelgradavg, 0, sizeof c->dvelgradavg);
        c->edotvmmx = 0.0;
        it->mix = 1.0;
    }
    it->iter++;
    int nb;
    // stage timers: events attached to the stage's first and last kernels where the register
    // passes run outside a graph, else stream markers around the stage
    const bool attach = c->timing == 1 && fft_tunable(f, c->fft) && !c->use_graphs;
    if (!attach) tstart(c, 0);
    if (!c->sgsum_valid) {
        launch_sigma_sum(f, c->s, &nb);
        launch_reduce_final(f.part, nb, 9, c->d_sgsum, c->s);
    }
    // stress norm used to scale the FP32 spectral stage of the mixed form (written by the
    // device after each pass once device_scale is set: the same max(svm, 1) of the pass' sums)
    const double scale = pass_scale(c);
    auto stage = [&]() {
        int n;
        hipEvent_t e0 = attach ? c->ev[0] : nullptr, e1 = attach ? c->ev[1] : nullptr;
        if (early)
            n = fft_green_passes(f, c->fft, c->d_sgsum, c->c0, c->s0, p->igamma, c->mixed, c->d_scale, 4, 5, c->s, e0,
                                 e1);
        else
            n = fft_green(f, c->fft, c->d_sgsum, c->c0, c->s0, p->igamma, c->mixed, scale, c->d_scale, c->s, e0, e1);
        launch_reduce_final_cm(f.part, n, 9, c->d_fluct, c->s);
    };
    // the register kernels take the scale from d_scale, so their stage can be replayed as a graph
    if (!run_stage) {
    } else if (c->use_graphs && fft_tunable(f, c->fft) && !early) {
        if (!c->g_pre || c->g_pre_version != c->green_version || c->g_pre_igamma != p->igamma ||
            c->g_pre_tune != c->tune_version) {
            if (int e = capture(c, &c->g_pre, stage)) return e;
            c->g_pre_version = c->green_version;
            c->g_pre_igamma = p->igamma;
            c->g_pre_tune = c->tune_version;
        }
        HIPCHK(c, hipGraphLaunch(c->g_pre, c->s));
    } else {
        stage();
    }
    if (!attach) tstop(c, 0);
    // correction + mixing + constitutive solve (one kernel)
    PassConst pc;
    pc.tdot = p->tdot; pc.complgas = c->complgas; pc.error = p->error; pc.erroral = p->erroral;
    pc.mix = it->mix; pc.lscale = std::max(c->lvmavg, 2.2250738585072014e-308);
    pc.modal = c->modal; pc.do_mix = c->do_mix; pc.kin_ctrl = (p->iscau[2] == 0);
    pc.lold_same = c->lold_same; pc.iuphard = p->iuphard;
    // ithermo = 1: the eigenstrain rate acts in the first increment of the process only
    // (various_functions.f90:1226: imicro == 1 .and. ithermo == 1)
    pc.thermo = (f.eth != nullptr && p->ithermo == 1 && it->imicro == 1) ? 1 : 0;
    PhaseTables pt{c->d_ph, c->d_tab};
    if (!c->taun_valid) {
        launch_taun(f, pt, c->sym_tau, c->s);
        c->taun_valid = 1;
    }
    if (!c->modal && (!c->sop_valid || c->sop_tdot != p->tdot)) {
        launch_solve_ops(f, p->tdot, c->s);
        c->sop_valid = 1;
        c->sop_tdot = p->tdot;
    }
    sa.pc = pc;
    for (int i = 0; i < 9; i++) sa.dvelgradavg[i] = c->dvelgradavg[i];
    inv3(c->defgradinvavgc, sa.finv_inv);
    sa.fluct = c->d_fluct;
    return 0;
}

static int pass_reduce(cg_ctx* c) {
    Fields& f = c->f;
    // the solve's sums, written also to the host's copy, the next pass' <sigma> (= sum_all w sigma)
    // and the solid sums its mixed-form scale derives from; with the overlap, ev_sums is recorded
    // at the end of the last kernel
    auto sums = [&](hipEvent_t done) {
        launch_pass_sums(f.part, solve_blocks(c->V), R_NSOLVE, R_NCOPY, f.red, c->h_pass_dev, c->d_sgsum,
                         c->mixed ? c->d_scale : nullptr, c->d_cnt, c->V, c->s, done);
    };
    if (!c->use_graphs) {
        tstart(c, 2);
        sums(c->overlap && fft_tunable(f, c->fft) ? c->ev_sums : nullptr);
        tstop(c, 2);
        return 0;
    }
    if (!c->g_post)
        if (int e = capture(c, &c->g_post, [&]() { sums(nullptr); })) return e;
    HIPCHK(c, hipGraphLaunch(c->g_post, c->s));
    return 0;
}

static int pass_finish(cg_ctx* c, cg_process* p, cg_iter_state* it, cg_iter_out* out) {
    const long V = c->V;
    if (int e = singular_from(c, c->h_pass, "update_schmid / calc_gas_compl / solve_res")) return e;
    const bool kin_ctrl = (p->iscau[2] == 0);
    c->sgsum_valid = 1;
    c->lold_same = 1;
    const double* r = c->h_pass;
    c->st.passes++;
    c->st.voxel_iters += V;
    c->st.nr_steps += (int64_t)r[R_NR];
    c->st.evals += (int64_t)r[R_EVAL];
    c->st.outer += (int64_t)r[R_OUTER];
    // correction_to_current averages (various_functions.f90:1132-1137)
    for (int i = 0; i < 9; i++) {
        c->velgradavg[i] = r[R_LALL + i];
        c->velgradavggas[i] = c->wphcgas > 0.0 ? r[R_LGAS + i] / c->wphcgas : r[R_LGAS + i];
        c->velgradavgsol[i] = r[R_LSOL + i] / c->wphcsol;
    }
    // dL_iter -> mix for the next pass (LS-EVPFFT.f90:313-327)
    double dl = 0.0;
    if (!kin_ctrl) {
        const double ds = r[R_WSOL] > 0.0 ? sqrt(r[R_DLSOL] / r[R_WSOL]) : 0.0;
        const double dg = r[R_WGAS] > 0.0 ? sqrt(r[R_DLGAS] / r[R_WGAS]) : 0.0;
        dl = std::max(ds, dg);
        it->mix = std::max(1.0e-8, std::min(1.0, 5.0e-2 / std::max(dl, 5.0e-2)));
    }
    // solve_res error sums (various_functions.f90:1357-1360)
    if (r[R_ERRE] < 0.0 || r[R_ERRS] < 0.0) return fail(c, 2, "ERRE/ERRS negative -> convergence failed!");
    double erre = sqrt(r[R_ERRE]), errs = sqrt(r[R_ERRS]);
    // calc_velgradavg_corr (various_functions.f90:1526-1635)
    for (int i = 0; i < 9; i++) {
        c->sgavg[i] = r[R_SGSOL + i];
        c->sgavgsol[i] = r[R_SGSOL + i] / c->wphcsol;
    }
    c->svmavg = vm_stress(c->sgavg);
    double a33[9];
    for (int i = 0; i < 6; i++) {
        const int ii = IJV[i][0], jj = IJV[i][1];
        a33[3 * ii + jj] = p->iscau[i] * (p->scauchy[3 * ii + jj] - c->sgavg[3 * ii + jj]);
        a33[3 * jj + ii] = p->iscau[i] * (p->scauchy[3 * jj + ii] - c->sgavg[3 * jj + ii]);
    }
    double errsbc = 0.0;
    for (int i = 0; i < 9; i++) errsbc += a33[i] * a33[i];
    errsbc = sqrt(errsbc);
    double vs[9], va[9], dbar[6], sbar[6], mev[36], devp0[6];
    sym3(c->velgradavg, vs);
    skw3(c->velgradavg, va);
    to6(vs, dbar);
    to6(c->sgavg, sbar);
    from4(c->de_ds, mev);
    for (int a = 0; a < 6; a++) {
        double s = 0.0;
        for (int b = 0; b < 6; b++) s += mev[6 * a + b] * sbar[b];
        devp0[a] = dbar[a] - s;
    }
    bool ok = state6x6(p->idsim, p->iscau, mev, devp0, p->scauchy, p->dsim);
    double wasim[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int ii = 0; ii < 3; ii++)
        for (int jj = 0; jj < 3; jj++) {
            if (ii == jj) continue;
            const int a = p->iudot[3 * ii + jj], b = p->iudot[3 * jj + ii];
            if (a == 1 && b == 0) {
                wasim[3 * ii + jj] = p->udot[3 * ii + jj] - p->dsim[3 * ii + jj];
                wasim[3 * jj + ii] = -wasim[3 * ii + jj];
            } else if (a == 1 && b == 1) {
                wasim[3 * ii + jj] = 0.5 * (p->udot[3 * ii + jj] - p->udot[3 * jj + ii]);
            }
        }
    for (int i = 0; i < 9; i++) c->dvelgradavg[i] = (p->dsim[i] - vs[i]) + (wasim[i] - va[i]);
    // error normalisation (LS-EVPFFT.f90:341-343)
    erre = erre / c->lvmavg;
    errs = errs / c->svmavg;
    errsbc = errsbc / c->svmavg;
    // A diverged pass leaves NaN or Inf in the error sums.  The reference's `< 0` test
    // (various_functions.f90:1357) lets NaN through and the loop then runs to itmax on
    // garbage; here it ends the run with the non-finite code.
    const bool finite = std::isfinite(erre) && std::isfinite(errs) && std::isfinite(errsbc);
    if (out) {
        out->errs = errs; out->erre = erre; out->errsbc = errsbc;
        memcpy(out->velgradavg, c->velgradavg, sizeof out->velgradavg);
        memcpy(out->sgavg, c->sgavg, sizeof out->sgavg);
        memcpy(out->sgavgsol, c->sgavgsol, sizeof out->sgavgsol);
        memcpy(out->dvelgradavg, c->dvelgradavg, sizeof out->dvelgradavg);
        out->svmavg = c->svmavg;
        out->dL_iter = dl;
        out->nr_steps = (int64_t)r[R_NR];
        out->evals = (int64_t)r[R_EVAL];
        out->outer = (int64_t)r[R_OUTER];
    }
    if (!ok) return fail(c, 1, "Failed system solve in lu_eqsystem (state_6x6_evpsc)");
    if (!finite)
        return fail(c, 2, "non-finite fixed-point errors (errs=%g erre=%g errsbc=%g) at pass %d", errs, erre, errsbc,
                    it->iter);
    return 0;
}

extern "C" {

int cg_iterate(cg_ctx* c, cg_process* p, cg_iter_state* it, cg_iter_out* out) {
    if (!c || !c->initialised || !p || !it) return -1;
    if (int e = ensure_fft_tuned(c, p)) return e;
    SolveArgs sa;
    const int used_slot = c->early ? c->early_slot : -1;
    if (int e = pass_begin(c, p, it, sa)) return e;
    PhaseTables pt{c->d_ph, c->d_tab};
    int nb;
    // the solve's own start and end events (no marker packets between the kernels)
    launch_correct_solve_full(c->f, pt, sa.pc, sa.dvelgradavg, sa.finv_inv, sa.fluct, c->s, &nb, c->int_pow,
                              c->sym_tau, c->small_ns, c->timing ? c->ev[2] : nullptr, c->timing ? c->ev[3] : nullptr);
    if (int e = pass_reduce(c)) return e;
    if (c->overlap && !c->use_graphs && fft_tunable(c->f, c->fft)) {
        // the next pass' transforms run while the host does its part of this one (ev_sums is
        // recorded at the end of the sums' kernel)
        const int slot = used_slot == 0 ? 1 : 0;
        const bool t1 = c->timing == 1;
        c->early_timed[slot] = t1 ? 1 : 0;
        fft_green_passes(c->f, c->fft, c->d_sgsum, c->c0, c->s0, p->igamma, c->mixed, c->d_scale, 0, 4, c->s,
                         t1 ? c->ev[8 + 2 * slot] : nullptr, t1 ? c->ev[9 + 2 * slot] : nullptr);
        c->early = 1;
        c->early_slot = slot;
        c->early_igamma = p->igamma;
        c->early_green = c->green_version;
        c->early_tune = c->tune_version;
        HIPCHK(c, hipEventSynchronize(c->ev_sums));
    } else {
        HIPCHK(c, hipStreamSynchronize(c->s));
    }
    HIPCHK(c, hipGetLastError());
    tcollect(c, 0); tcollect(c, 1); tcollect(c, 2);
    if (used_slot >= 0) tcollect_early(c, used_slot);
    const int rc = pass_finish(c, p, it, out);
    if (rc) drop_early(c);
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched passes: several RVEs share the constitutive launch
// ---------------------------------------------------------------------------
struct cg_batch {
    std::vector<cg_ctx*> ctx;
    hipStream_t s = nullptr;          // the shared constitutive launches
    hipEvent_t ev_solve = nullptr;
    std::vector<hipEvent_t> ev_pre;   // per RVE: its pre-solve work is enqueued
    SolveWork* h_work = nullptr;      // pinned host copy of the work table
    SolveWork* d_work = nullptr;
    int* d_blk_rve = nullptr;         // RVE of every block of the launches, groups one after another
    size_t blk_cap = 0;
    std::vector<int> table_active;    // active set d_blk_rve was built for
    std::vector<int> key;             // kernel instance of every RVE: modal, integer exponents, equal strengths, <= 12 systems
    std::vector<int> grp_key, grp_off, grp_nblk;  // launches of the current table
    // batched FFT + Gamma stages and sums: the entries of RVEs with the same grid, precision and
    // Green-operator kind one after another, one group per launch set
    FftWork* h_fw = nullptr;          // pinned
    FftWork* d_fw = nullptr;
    // the pass in flight (cg_batch_launch -> cg_batch_wait)
    std::vector<cg_process*> pend_p;
    cg_iter_state* pend_st = nullptr;
    cg_iter_out* pend_out = nullptr;
    std::vector<int> pend_active;
};

static int solve_key(const cg_ctx* c) {
    return (c->modal ? 8 : 0) + (c->int_pow ? 4 : 0) + (c->sym_tau ? 2 : 0) + (c->small_ns ? 1 : 0);
}

// d_blk_rve and the RVEs' first blocks for the active set: one launch per kernel instance
static int batch_table(cg_batch* b, const int* active) {
    const int n = (int)b->ctx.size();
    std::vector<int> act(active, active + n);
    for (int& a : act) a = a ? 1 : 0;
    if (act == b->table_active) return 0;
    std::vector<int> blk;
    b->grp_key.clear(); b->grp_off.clear(); b->grp_nblk.clear();
    for (int k = 0; k < 16; k++) {
        const size_t off = blk.size();
        for (int i = 0; i < n; i++) {
            if (!act[i] || b->key[i] != k) continue;
            const long nb = solve_blocks(b->ctx[i]->V);
            b->h_work[i].blk0 = (long)(blk.size() - off);
            b->h_work[i].nblk = nb;
            blk.insert(blk.end(), (size_t)nb, i);
        }
        if (blk.size() > off) {
            b->grp_key.push_back(k);
            b->grp_off.push_back((int)off);
            b->grp_nblk.push_back((int)(blk.size() - off));
        }
    }
    if (blk.size() > b->blk_cap) {
        if (b->d_blk_rve) hipFree(b->d_blk_rve);
        b->d_blk_rve = nullptr;
        if (hipMalloc(&b->d_blk_rve, blk.size() * sizeof(int)) != hipSuccess)
            return noctx_fail(-3, "cg_batch_iterate: hipMalloc of the block table failed");
        b->blk_cap = blk.size();
    }
    if (!blk.empty() &&
        hipMemcpyAsync(b->d_blk_rve, blk.data(), blk.size() * sizeof(int), hipMemcpyHostToDevice, b->s) != hipSuccess)
        return noctx_fail(-2, "cg_batch_iterate: block table upload failed");
    if (hipStreamSynchronize(b->s) != hipSuccess) return noctx_fail(-2, "cg_batch_iterate: synchronisation failed");
    b->table_active = act;
    return 0;
}

extern "C" {

int cg_batch_create(cg_ctx* const* ctxs, int n, cg_batch** out) {
    if (!out || !ctxs || n < 1) return noctx_fail(-1, "cg_batch_create: bad arguments");
    *out = nullptr;
    for (int i = 0; i < n; i++)
        if (!ctxs[i] || !ctxs[i]->initialised) return noctx_fail(-1, "cg_batch_create: context %d not initialised", i);
    cg_batch* b = new cg_batch();
    b->ctx.assign(ctxs, ctxs + n);
    b->ev_pre.assign(n, nullptr);
    for (int i = 0; i < n; i++) b->key.push_back(solve_key(ctxs[i]));
    bool ok = hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&b->ev_solve, hipEventDisableTiming) == hipSuccess &&
              hipHostMalloc(&b->h_work, n * sizeof(SolveWork)) == hipSuccess &&
              hipMalloc(&b->d_work, n * sizeof(SolveWork)) == hipSuccess &&
              hipHostMalloc(&b->h_fw, n * sizeof(FftWork)) == hipSuccess &&
              hipMalloc(&b->d_fw, n * sizeof(FftWork)) == hipSuccess;
    for (int i = 0; ok && i < n; i++) ok = hipEventCreateWithFlags(&b->ev_pre[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        cg_batch_destroy(b);
        return noctx_fail(-2, "cg_batch_create: stream, event or table allocation failed");
    }
    memset(b->h_work, 0, n * sizeof(SolveWork));
    *out = b;
    return 0;
}

int cg_batch_destroy(cg_batch* b) {
    if (!b) return 0;
    if (b->s) hipStreamSynchronize(b->s);
    for (hipEvent_t e : b->ev_pre)
        if (e) hipEventDestroy(e);
    if (b->ev_solve) hipEventDestroy(b->ev_solve);
    if (b->h_work) hipHostFree(b->h_work);
    if (b->d_work) hipFree(b->d_work);
    if (b->h_fw) hipHostFree(b->h_fw);
    if (b->d_fw) hipFree(b->d_fw);
    if (b->d_blk_rve) hipFree(b->d_blk_rve);
    if (b->s) hipStreamDestroy(b->s);
    delete b;
    return 0;
}

int cg_batch_launch(cg_batch* b, cg_process* const* p, cg_iter_state* st, cg_iter_out* out, const int* active) {
    if (!b || !p || !st || !active) return noctx_fail(-1, "cg_batch_launch: bad arguments");
    const int n = (int)b->ctx.size();
    int nact = 0;
    for (int i = 0; i < n; i++)
        if (active[i]) {
            if (!p[i]) return noctx_fail(-1, "cg_batch_launch: no process for RVE %d", i);
            nact++;
        }
    for (int i = 0; i < n; i++)
        if (active[i])
            if (int e = ensure_fft_tuned(b->ctx[i], p[i]))
                return noctx_fail(e, "cg_batch_launch: RVE %d: %s", i, b->ctx[i]->err);
    b->pend_p.assign(p, p + n);
    b->pend_st = st;
    b->pend_out = out;
    b->pend_active.assign(active, active + n);
    if (!nact) return 0;
    if (int e = batch_table(b, active)) return e;
    // The RVEs whose lengths have register FFT kernels (and no graphs) run their FFT + Gamma
    // stages and their sums batched: one launch per pass for all RVEs of a grid, precision and
    // Green-operator kind.  Others run theirs on their own streams.
    std::vector<int> bat(n, 0);
    std::vector<std::pair<std::array<int, 5>, int>> order;
    for (int i = 0; i < n; i++) {
        if (!active[i]) continue;
        cg_ctx* c = b->ctx[i];
        bat[i] = fft_tunable(c->f, c->fft) && !c->use_graphs;
        if (bat[i]) order.push_back({{c->f.n1, c->f.n2, c->f.n3, c->mixed, p[i]->igamma ? 1 : 0}, i});
    }
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<std::array<int, 5>, int>& a, const std::pair<std::array<int, 5>, int>& q) {
                         return a.first < q.first;
                     });
    // work before the solve, on every RVE's own stream
    for (int i = 0; i < n; i++) {
        if (!active[i]) continue;
        cg_ctx* c = b->ctx[i];
        SolveArgs sa;
        if (int e = pass_begin(c, p[i], &st[i], sa, !bat[i]))
            return noctx_fail(e, "cg_batch_launch: RVE %d: %s", i, c->err);
        b->h_work[i].f = c->f;
        b->h_work[i].pt = PhaseTables{c->d_ph, c->d_tab};
        b->h_work[i].a = sa;
        hipEventRecord(b->ev_pre[i], c->s);
        hipStreamWaitEvent(b->s, b->ev_pre[i], 0);
    }
    std::vector<size_t> gstart;  // first entry of each group in order / h_fw
    for (size_t k = 0; k < order.size(); k++) {
        if (k == 0 || order[k].first != order[k - 1].first) gstart.push_back(k);
        cg_ctx* c = b->ctx[order[k].second];
        FftWork& w = b->h_fw[k];
        w.f = c->f;
        fft_twiddles(c->fft, c->mixed, w.tw);
        w.sgsum = c->d_sgsum;
        w.dscale = c->mixed ? c->d_scale : nullptr;
        w.counter = c->d_cnt;
        w.fluct = c->d_fluct;
        w.host = c->h_pass_dev;
        memcpy(w.g.c0, c->c0, sizeof w.g.c0);
        memcpy(w.g.s0, c->s0, sizeof w.g.s0);
        w.g.igamma = p[order[k].second]->igamma;
    }
    gstart.push_back(order.size());
    if (!order.empty() &&
        hipMemcpyAsync(b->d_fw, b->h_fw, order.size() * sizeof(FftWork), hipMemcpyHostToDevice, b->s) != hipSuccess)
        return noctx_fail(-2, "cg_batch_iterate: FFT work table upload failed");
    for (size_t g = 0; g + 1 < gstart.size(); g++) {
        const int k0 = (int)gstart[g], cnt = (int)(gstart[g + 1] - gstart[g]);
        const cg_ctx* c = b->ctx[order[k0].second];
        if (fft_green_batch(b->d_fw + k0, cnt, c->f, c->mixed, order[k0].first[4], b->s) < 0)
            return noctx_fail(-2, "cg_batch_iterate: batched FFT stage failed");
    }
    // the shared constitutive launches (one per kernel instance present)
    if (hipMemcpyAsync(b->d_work, b->h_work, n * sizeof(SolveWork), hipMemcpyHostToDevice, b->s) != hipSuccess)
        return noctx_fail(-2, "cg_batch_iterate: work table upload failed");
    for (size_t g = 0; g < b->grp_key.size(); g++) {
        const int k = b->grp_key[g];
        launch_correct_solve_batch(b->d_work, b->d_blk_rve + b->grp_off[g], b->grp_nblk[g], (k >> 3) & 1, (k >> 2) & 1,
                                   (k >> 1) & 1, k & 1, b->s);
    }
    // the batched RVEs' sums (straight to their host copies)
    for (size_t g = 0; g + 1 < gstart.size(); g++) {
        const int k0 = (int)gstart[g], cnt = (int)(gstart[g + 1] - gstart[g]);
        launch_pass_sums_batch(b->d_fw + k0, cnt, solve_blocks(b->ctx[order[k0].second]->V), R_NSOLVE, R_NCOPY, b->s,
                               nullptr);
    }
    hipEventRecord(b->ev_solve, b->s);
    // the other RVEs' sums on their own streams; every RVE's later work waits for the round
    for (int i = 0; i < n; i++) {
        if (!active[i]) continue;
        hipStreamWaitEvent(b->ctx[i]->s, b->ev_solve, 0);
        if (!bat[i])
            if (int e = pass_reduce(b->ctx[i]))
                return noctx_fail(e, "cg_batch_launch: RVE %d: %s", i, b->ctx[i]->err);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return noctx_fail(-2, "cg_batch_launch: %s", hipGetErrorString(e));
    return 0;
}

int cg_batch_wait(cg_batch* b, int* rc) {
    if (!b || !rc) return noctx_fail(-1, "cg_batch_wait: bad arguments");
    const int n = (int)b->ctx.size();
    for (int i = 0; i < n; i++) rc[i] = 0;
    if ((int)b->pend_active.size() != n) return 0;  // nothing launched
    for (int i = 0; i < n; i++)
        if (b->pend_active[i] && hipStreamSynchronize(b->ctx[i]->s) != hipSuccess)
            return noctx_fail(-2, "cg_batch_wait: synchronisation of RVE %d failed", i);
    if (hipStreamSynchronize(b->s) != hipSuccess) return noctx_fail(-2, "cg_batch_wait: synchronisation failed");
    for (int i = 0; i < n; i++)
        if (b->pend_active[i])
            rc[i] = pass_finish(b->ctx[i], b->pend_p[i], &b->pend_st[i], b->pend_out ? &b->pend_out[i] : nullptr);
    b->pend_active.clear();
    return 0;
}

int cg_batch_iterate(cg_batch* b, cg_process* const* p, cg_iter_state* st, cg_iter_out* out, const int* active,
                     int* rc) {
    if (int e = cg_batch_launch(b, p, st, out, active)) return e;
    return cg_batch_wait(b, rc);
}

int cg_rollback(cg_ctx* c, cg_process* p) {
    drop_early(c);
    if (!c || !c->initialised || !p) return -1;
    if (!c->snap_valid)
        return fail(c, -1, "cg_rollback: this increment took no snapshot (only ictrl = -1 with tdot > tdotmin can roll back)");
    memcpy(p->dsim, c->dsim_snap, sizeof c->dsim_snap);
    memcpy(p->scauchy, c->scauchy_snap, sizeof c->scauchy_snap);
    launch_snapshot(c->f, 1, c->s);
    c->sgsum_valid = 0;
    c->taun_valid = 0;
    c->lold_same = 0;
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}

int cg_kinematic_substep(cg_ctx* c, double tdot, double* out4, int* nsub) {
    if (!c || !c->initialised) return -1;
    int nb;
    if (ensure_part(c, (size_t)((c->V + 63) / 64 + 64) * 81)) return -3;
    launch_kinematic(c->f, tdot, c->s, &nb);
    double r[5];
    HIPCHK(c, hipMemcpyAsync(c->h_red, c->f.red, 5 * sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    memcpy(r, c->h_red, sizeof r);
    if (out4) memcpy(out4, r, 4 * sizeof(double));
    // r[4] = max(1, ceiling(worst ratio)), formed on the device; a non-finite ratio gives 1
    if (nsub) *nsub = (r[4] >= 1.0 && r[4] < 2.0e9) ? (int)r[4] : 1;
    return 0;
}

int cg_end_increment(cg_ctx* c, const cg_process* p, int scan_diagnostics, cg_incr_out* out) {
    drop_early(c);
    if (!c || !c->initialised || !p) return -1;
    Fields& f = c->f;
    const long V = c->V;
    const double tdot = p->tdot;
    PhaseTables pt{c->d_ph, c->d_tab};
    tstart(c, 3);
    if (c->lold_same) {  // velgradold == velgrad after the last pass
        launch_copy(f.Lold, f.L, 9 * f.ld, c->s);
        c->lold_same = 0;
    }
    int nb;
    double r[128];
    // accumulation (LS-EVPFFT.f90:431-446); the largest strain rate stays on the device until the
    // averages' copy below
    constexpr int R_EDOTMX = 100, R_LAST = 101;  // free slots of red (13 values from R_LAST)
    launch_accumulate(f, tdot, f.red + R_EDOTMX, c->s, &nb);
    if (p->iupdate == 1) {
        launch_plastic_defgrad(f, pt, tdot, c->s);
        // update_grid_velgrad_node: dv in the (now free) dLref buffer
        int cm = 0;
        nb = fft_integrate(f, c->fft, p->igamma, f.dLref, c->s, &cm);
        if (cm) launch_reduce_final_cm(f.part, nb, 9, f.red, c->s);
        if (reduce_to_host(c, cm ? 0 : nb, 9, r, cm)) return -2;
        launch_xnode(f, f.dLref, r, tdot, p->igamma, c->n3node, c->s);
        launch_update_defgrad(f, tdot, c->s, &nb);
        if (reduce_to_host(c, nb, 39, r)) return -2;
        const double detfavg = r[38];
        memcpy(c->defgradavg, r, 9 * sizeof(double));
        c->wphcsol = r[36] / detfavg;
        c->wphcgas = r[37] / detfavg;
        const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        for (int i = 0; i < 9; i++) {
            c->defgradinvavgc[i] = r[9 + i] / detfavg;
            c->defgradinvavgcs[i] = r[18 + i] / detfavg / c->wphcsol;
            c->defgradinvavgcg[i] = c->wphcgas > 0.0 ? r[27 + i] / detfavg / c->wphcgas : I3[i];
        }
        launch_div_w(f, detfavg, c->s);
        launch_elastic_stiff(f, pt, c->s);
        c->sop_valid = 0;
        launch_tensor_config_a(f, c->s, &nb);
        if (reduce_to_host(c, nb, 9, r)) return -2;
        double inv[9], dm[9], dud[9];
        inv3(c->defgradinvavgc, inv);
        for (int i = 0; i < 9; i++) dm[i] = c->velgradavg[i] - r[i];
        mm3(dm, inv, dud);
        launch_tensor_config_b(f, dud, c->s);
    }
    if (p->iuphard == 1) launch_hardening(f, pt, tdot, c->small_ns, c->sym_tau, c->s);
    // macro averages (LS-EVPFFT.f90:460-523)
    double vs[9];
    sym3(c->velgradavg, vs);
    c->lvmavg = vm_strain(c->velgradavg);
    c->dvmavg = vm_strain(vs);
    c->evmavg += c->dvmavg * tdot;
    launch_averages(f, pt, c->s, &nb);
    // eelavg = eavg + (disgradsym - ept) w of the last solid voxel in loop order (:482): its 13
    // values travel with the averages and the largest strain rate, in one copy
    launch_reduce_final(f.part, nb, 50, f.red, c->s);
    if (c->last_solid >= 0) launch_pick_last(f, c->last_solid, f.red + R_LAST, c->s);
    HIPCHK(c, hipMemcpyAsync(c->h_red, f.red, (R_LAST + 13) * sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    memcpy(r, c->h_red, 50 * sizeof(double));
    c->edotvmmx = std::max(c->edotvmmx, c->h_red[R_EDOTMX]);
    double last[13] = {0};
    if (c->last_solid >= 0) memcpy(last, c->h_red + R_LAST, sizeof last);
    double dl6[6], dl3[9];
    for (int k = 0; k < 6; k++) dl6[k] = last[k] - last[6 + k];
    from6(dl6, dl3);
    for (int i = 0; i < 9; i++) {
        c->epavg[i] = r[i] / c->wphcsol;
        c->eavg[i] = r[9 + i] / c->wphcsol;
        c->eelavg[i] = c->last_solid >= 0 ? (r[9 + i] + dl3[i] * last[12]) / c->wphcsol : 0.0;
        c->edotpavg[i] = r[27 + i] / c->wphcsol;
        c->sgsolavg[i] = r[36 + i] / c->wphcsol;
    }
    const double gammaavg = r[46] > 0.0 ? r[45] / r[46] : 0.0;
    const double gavg = c->nsys_solid > 0.0 ? r[47] / c->nsys_solid : 0.0;
    const double geffavg = r[48] > 0.0 ? r[49] / r[48] : gavg;
    c->evmpavg = vm_strain(c->epavg);
    c->dvmpavg = vm_strain(c->edotpavg);
    if (out) {
        memset(out, 0, sizeof *out);
        memcpy(out->eavg, c->eavg, 72); memcpy(out->sgsolavg, c->sgsolavg, 72); memcpy(out->epavg, c->epavg, 72);
        memcpy(out->eelavg, c->eelavg, 72); memcpy(out->edotpavg, c->edotpavg, 72);
        memcpy(out->velgradavg, c->velgradavg, 72); memcpy(out->sgavg, c->sgavg, 72);
        out->evmavg = c->evmavg; out->evmpavg = c->evmpavg; out->dvmavg = c->dvmavg; out->dvmpavg = c->dvmpavg;
        out->svmavg = c->svmavg; out->lvmavg = c->lvmavg; out->edotvmmx = c->edotvmmx;
        out->gammaavg = gammaavg; out->gavg = gavg; out->geffavg = geffavg;
    }
    // update_diagnostics (types.f90:345-758)
    if (out) {
        const double Lref = std::max(c->lvmavg, 2.2250738585072014e-308);
        const double lfloor = std::max(1.0e-12 * Lref, 2.2250738585072014e-308);
        if (ensure_part(c, (size_t)((V + 255) / 256) * 24 + 64)) return -3;
        launch_diagnostics(f, lfloor, scan_diagnostics, c->s, &nb);
        // with the singular-system flags of the updates above (the diagnostics raise none)
        HIPCHK(c, hipMemcpyAsync(c->h_red, f.red, R_NCOPY * sizeof(double), hipMemcpyDeviceToHost, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
        const double* d = c->h_red;
        const double HUGEV = 1.7976931348623157e308;
        out->dL_step_rms = d[1] > 0.0 ? sqrt(d[0] / d[1]) : 0.0;
        out->dL_step_rms_sol = d[3] > 0.0 ? sqrt(d[2] / d[3]) : 0.0;
        out->dL_step_rms_gas = d[5] > 0.0 ? sqrt(d[4] / d[5]) : 0.0;
        out->dL_step_max = d[6]; out->dL_step_max_sol = d[7]; out->dL_step_max_gas = d[8];
        out->volinc_max_sol = d[9]; out->volinc_max_gas = d[10];
        out->stretchinc_max_sol = d[11]; out->stretchinc_max_gas = d[12];
        out->rotang_max_sol = d[13]; out->rotang_max_gas = d[14];
        out->cond_c066mod_max_sol = d[15]; out->cond_c066mod_max_gas = d[16]; out->cond_cgas_max = d[17];
        out->detF_min = d[18] >= HUGEV * 0.5 ? 0.0 : d[18];
        out->detF_min_sol = (d[20] < 0 || d[19] >= HUGEV * 0.5) ? 0.0 : d[19];
        out->detF_min_gas = (d[22] < 0 || d[21] >= HUGEV * 0.5) ? 0.0 : d[21];
        for (int k = 0; k < 2; k++) {
            const long idx = (long)d[20 + 2 * k];
            int* im = k == 0 ? out->imin_sol : out->imin_gas;
            if (idx >= 0) {
                im[0] = (int)(idx % c->n1) + 1;
                im[1] = (int)((idx / c->n1) % c->n2) + 1;
                im[2] = (int)(idx / ((long)c->n1 * c->n2)) + 1;
            } else {
                im[0] = im[1] = im[2] = 0;
            }
        }
        double n2 = 0.0;
        for (int i = 0; i < 9; i++) n2 += c->dvelgradavg[i] * c->dvelgradavg[i];
        out->dvelgradavg_mag = sqrt(n2);
    }
    c->sgsum_valid = 0;  // weights changed
    tstop(c, 3);
    HIPCHK(c, hipStreamSynchronize(c->s));
    HIPCHK(c, hipGetLastError());
    tcollect(c, 3);
    const char* where = "update_defgrad / update_elstic_defgrad / update_el_stiff / update_tensor_config";
    return out ? singular_from(c, c->h_red, where) : check_singular(c, where);
}

int cg_calc_c0(cg_ctx* c, const cg_process* p) {
    drop_early(c);
    if (!c || !c->initialised || !p) return -1;
    PhaseTables pt{c->d_ph, c->d_tab};
    int nb;
    tstart(c, 3);
    launch_calc_c0(c->f, pt, p->tdot, p->tdotref, c->s, &nb, c->int_pow);
    double r[72];
    if (reduce_to_host(c, nb, 72, r)) return -2;
    tstop(c, 3);
    tcollect(c, 3);
    if (int e = singular_from(c, c->h_red, "calc_c0")) return e;
    double dsde[36], dP[36], a66[36];
    for (int i = 0; i < 36; i++) {
        dsde[i] = r[i] / c->wphcsol;
        dP[i] = r[36 + i] / c->wphsol;
    }
    bool ok = inv6(dsde);
    to4(dsde, c->de_ds);
    for (int i = 0; i < 36; i++) {
        c->c066[i] = dP[i] * p->xc0;
        a66[i] = dP[i] * p->xc0;
    }
    ok = inv6(a66) && ok;
    to4(c->c066, c->c0);
    to4(a66, c->s0);
    c->green_version++;
    if (!ok) return fail(c, 1, "Singular matrix in lu_inverse (calc_c0)");
    return 0;
}

// ---------------------------------------------------------------------------
// fields
// ---------------------------------------------------------------------------
namespace {
struct FieldDesc {
    double* p;
    int stored;   // components stored on device
    int ncomp;    // components exposed
    int sym;      // 1: stored as b-basis 6-vector, exposed as 3x3
};
bool field_desc(cg_ctx* c, const char* name, FieldDesc& d) {
    Fields& f = c->f;
    const int ns = f.ns;
    struct E { const char* n; double* p; int st; int nc; int sym; } T[] = {
        {"sg", f.sg, 6, 9, 1}, {"sgt", f.sgt, 6, 9, 1}, {"velgrad", f.L, 9, 9, 0}, {"velgradold", f.Lold, 9, 9, 0},
        {"dvelgradref", f.dLref, 9, 9, 0}, {"defgrad", f.F, 9, 9, 0}, {"defgradinv", f.Fi, 9, 9, 0},
        {"detF", f.J, 1, 1, 0}, {"wgtc", f.w, 1, 1, 0}, {"defgrade", f.Fe, 9, 9, 0}, {"defgradp", f.Fp, 9, 9, 0},
        {"ag", f.ag, 9, 9, 0}, {"sg66", f.sg66, 36, 36, 0}, {"c066mod", f.c66m, 36, 36, 0}, {"cgm", f.cgm, 36, 36, 0},
        {"cgas", f.cgas, 36, 36, 0}, {"sch", f.sch, 5 * ns, 5 * ns, 0}, {"crss", f.crss, 2 * ns, 2 * ns, 0},
        {"trialtau", f.ttau, 2 * ns, 2 * ns, 0}, {"gamdot", f.gdot, ns, ns, 0}, {"gacumgr", f.gacc, 1, 1, 0},
        {"edotp", f.edotp, 6, 9, 1}, {"ept", f.ept, 6, 9, 1}, {"disgradsym", f.dgs, 6, 9, 1}, {"epvm", f.epvm, 1, 1, 0},
        {"defgradinc", f.Finc, 9, 9, 0}, {"defgradeinc", f.Feinc, 9, 9, 0}, {"defgradini", f.Fini, 9, 9, 0},
    };
    for (auto& e : T)
        if (!strcmp(e.n, name)) {
            d.p = e.p; d.stored = e.st; d.ncomp = e.nc; d.sym = e.sym;
            return true;
        }
    return false;
}
}  // namespace

int cg_field_ncomp(cg_ctx* c, const char* name) {
    if (!c || !name) return -1;
    if (!strcmp(name, "xnode")) return 3;
    FieldDesc d;
    if (!field_desc(c, name, d)) return -1;
    return d.ncomp;
}

int cg_get_field(cg_ctx* c, const char* name, double* out, size_t n) {
    if (!c || !c->initialised || !name || !out) return -1;
    if (!strcmp(name, "xnode")) {  // node coordinates (3, n1+1, n2+1, n3+1), x fastest
        const size_t nn = 3 * (size_t)(c->n1 + 1) * (c->n2 + 1) * c->n3node;
        if (n < nn) return fail(c, -1, "output buffer too small");
        HIPCHK(c, hipStreamSynchronize(c->s));
        HIPCHK(c, hipMemcpy(out, c->f.xnode, nn * sizeof(double), hipMemcpyDeviceToHost));
        return 0;
    }
    FieldDesc d;
    if (!field_desc(c, name, d) || !d.p) return fail(c, -1, "unknown or absent field %s", name);
    const long V = c->V;
    if (n < (size_t)d.ncomp * V) return fail(c, -1, "output buffer too small");
    if (d.p == c->f.sch && !c->modal) {  // not kept by this context's solve: expand A, A^-1
        PhaseTables pt{c->d_ph, c->d_tab};
        launch_sch_from_A(c->f, pt, c->s);
    }
    if (d.p == c->f.Lold && c->lold_same) {  // velgradold == velgrad since the last pass
        launch_copy(c->f.Lold, c->f.L, 9 * c->f.ld, c->s);
        c->lold_same = 0;
    }
    std::vector<double> h((size_t)d.stored * V);
    HIPCHK(c, hipStreamSynchronize(c->s));
    for (int k = 0; k < d.stored; k++)
        HIPCHK(c, hipMemcpy(h.data() + (size_t)k * V, d.p + (size_t)k * c->f.ld, V * sizeof(double), hipMemcpyDeviceToHost));
    // equal forward and backward strengths: the hardens inside the solve write only the
    // forward column of trialtau (the increment's last harden writes both)
    if (c->sym_tau && d.p == c->f.ttau)
        for (int k = 0; k + 1 < d.stored; k += 2)
            std::copy(h.begin() + (size_t)k * V, h.begin() + (size_t)(k + 1) * V, h.begin() + (size_t)(k + 1) * V);
    for (long v = 0; v < V; v++) {
        if (d.sym) {
            double s6[6];
            for (int k = 0; k < 6; k++) s6[k] = h[(size_t)k * V + v];
            from6(s6, out + 9 * v);
        } else {
            for (int k = 0; k < d.stored; k++) out[(size_t)d.ncomp * v + k] = h[(size_t)k * V + v];
        }
    }
    return 0;
}

int cg_set_field(cg_ctx* c, const char* name, const double* in, size_t n) {
    drop_early(c);
    if (!c || !c->initialised || !name || !in) return -1;
    c->sop_valid = 0;
    c->taun_valid = 0;
    FieldDesc d;
    if (!field_desc(c, name, d) || !d.p) return fail(c, -1, "unknown or absent field %s", name);
    const long V = c->V;
    if (n < (size_t)d.ncomp * V) return fail(c, -1, "input buffer too small");
    if (c->lold_same) {  // keep velgradold (== velgrad since the last pass) before any field changes
        launch_copy(c->f.Lold, c->f.L, 9 * c->f.ld, c->s);
        c->lold_same = 0;
    }
    // The symmetric solve of a context without the modified-AL term rebuilds the Schmid tensors
    // from A = Fe ag (schA, written by update_schmid) and never reads sch: a host write to sch
    // would be silently ignored there, so it is refused.  Writes to defgrade and ag reach the
    // Schmid tensors at the next update_schmid (cg_begin_increment with do_schmid), as in the
    // reference, where sch changes only in update_schmid.
    if (d.p == c->f.sch && !c->modal)
        return fail(c, -1, "sch is rebuilt from defgrade and ag by update_schmid in this context; write those instead");
    if (c->sym_tau && (d.p == c->f.crss || d.p == c->f.ttau)) {
        // the kernels of a context whose forward and backward strengths are equal read one
        // column of each pair
        for (long v = 0; v < V; v++)
            for (int k = 0; k + 1 < d.ncomp; k += 2)
                if (in[(size_t)d.ncomp * v + k] != in[(size_t)d.ncomp * v + k + 1])
                    return fail(c, -1, "%s: forward and backward values differ in a context built with equal ones", name);
    }
    std::vector<double> h((size_t)d.stored * V);
    for (long v = 0; v < V; v++) {
        if (d.sym) {
            double s6[6];
            to6(in + 9 * v, s6);
            for (int k = 0; k < 6; k++) h[(size_t)k * V + v] = s6[k];
        } else {
            for (int k = 0; k < d.stored; k++) h[(size_t)k * V + v] = in[(size_t)d.ncomp * v + k];
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->s));
    for (int k = 0; k < d.stored; k++)
        HIPCHK(c, hipMemcpy(d.p + (size_t)k * c->f.ld, h.data() + (size_t)k * V, V * sizeof(double), hipMemcpyHostToDevice));
    c->sgsum_valid = 0;
    return 0;
}

int cg_set_thermo(cg_ctx* c, const double* eth) {
    drop_early(c);
    if (!c || !c->initialised || !eth) return -1;
    Fields& f = c->f;
    const long V = c->V;
    if (!f.eth && dalloc(c, &f.eth, 6 * f.ld)) return -3;
    std::vector<double> h((size_t)6 * V, 0.0);
    for (long v = 0; v < V; v++) {
        double s6[6];
        to6(eth + 9 * v, s6);  // chg_basis(edotth, eth, ..., 2, 6): the symmetric part in the b-basis
        for (int k = 0; k < 6; k++) h[(size_t)k * V + v] = c->ph[c->phase[v] - 1].igas ? 0.0 : s6[k];
    }
    HIPCHK(c, hipStreamSynchronize(c->s));
    for (int k = 0; k < 6; k++)
        HIPCHK(c, hipMemcpy(f.eth + (size_t)k * f.ld, h.data() + (size_t)k * V, V * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

int cg_get_macro(cg_ctx* c, double* c0, double* s0, double* de_ds, double* c066, double* goperr0) {
    if (!c) return -1;
    if (c0) memcpy(c0, c->c0, sizeof c->c0);
    if (s0) memcpy(s0, c->s0, sizeof c->s0);
    if (de_ds) memcpy(de_ds, c->de_ds, sizeof c->de_ds);
    if (c066) memcpy(c066, c->c066, sizeof c->c066);
    if (goperr0) memcpy(goperr0, c->goperr0, sizeof c->goperr0);
    return 0;
}

int cg_get_stats(cg_ctx* c, cg_stats* st, int reset) {
    if (!c) return -1;
    if (st) *st = c->st;
    if (reset) memset(&c->st, 0, sizeof c->st);
    return 0;
}

int cg_set_timing(cg_ctx* c, int on) {
    drop_early(c);
    if (!c) return -1;
    c->timing = on == 2 ? 2 : on ? 1 : 0;
    return 0;
}

int cg_set_tuning_file(cg_ctx* c, const char* path) {
    drop_early(c);
    if (!c) return -1;
    c->tune_path = path ? path : "";
    return 0;
}

int cg_get_fft_tuning(cg_ctx* c, int* shape8, double* us5) {
    if (!c || !c->fft) return -1;
    const FftTune t = fft_get_tune(c->fft);
    if (shape8) {
        for (int k = 0; k < 5; k++) shape8[k] = t.xcd[k];
        shape8[5] = t.yhalf[0];
        shape8[6] = t.yhalf[1];
        shape8[7] = t.kt1;
    }
    if (us5)
        for (int k = 0; k < 5; k++) us5[k] = t.us[k];
    return t.tuned;
}

int cg_set_fft_tuning(cg_ctx* c, const int* shape8) {
    drop_early(c);
    if (!c || !c->fft || !shape8) return -1;
    FftTune t;
    for (int k = 0; k < 8; k++)
        if (shape8[k] != 0 && shape8[k] != 1) return fail(c, -1, "cg_set_fft_tuning: entries are 0 or 1");
    for (int k = 0; k < 5; k++) t.xcd[k] = shape8[k];
    t.yhalf[0] = shape8[5];
    t.yhalf[1] = shape8[6];
    t.kt1 = shape8[7];
    fft_set_tune(c->fft, t);
    c->fft_tuned = 1;  // given, not timed
    c->tune_version++;
    return 0;
}

int cg_set_displacement(cg_ctx* c, int idispini, const double* disp) {
    drop_early(c);
    if (!c || !disp) return -1;
    if (c->initialised) return fail(c, -1, "cg_set_displacement: call before cg_init");
    if (idispini != 1 && idispini != 2) return fail(c, -1, "cg_set_displacement: idispini must be 1 or 2");
    const size_t n = idispini == 1 ? (size_t)c->V : (size_t)(c->n1 + 1) * (c->n2 + 1) * (c->n3 == 1 ? 1 : c->n3 + 1);
    c->dispini.assign(disp, disp + 3 * n);
    c->idispini = idispini;
    return 0;
}

int cg_set_graphs(cg_ctx* c, int on) {
    drop_early(c);
    if (!c) return -1;
    c->use_graphs = on ? 1 : 0;
    return 0;
}

int cg_set_overlap(cg_ctx* c, int on) {
    drop_early(c);
    if (!c) return -1;
    c->overlap = on ? 1 : 0;
    return 0;
}

}  // extern "C"
