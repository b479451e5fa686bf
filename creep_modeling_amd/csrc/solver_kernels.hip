// solver_kernels.hip -- per-voxel kernels of the fixed-point loop and of the
// accepted-increment updates (one thread per voxel, FP64).
//
// Reference operators: various_functions.f90 (solve_res :1142-1364,
// plastic_strain_rate :1367-1419, harden :1439-1491, correction_to_current
// :1062-1139, update_* :1748-2492, calc_c0 :2537-2649, kinematic_increments
// :2766-2837), LS-EVPFFT.f90 (mixing :277-327, accumulation :431-446,
// averages :460-523) and types.f90 (update_diagnostics :345-758).
#include <hip/hip_ext.h>

#include "internal.h"
#include "reduce.h"
#include "tensor.h"
#include "vox.h"

namespace cg {

constexpr int BS = 256;
// threads per block of the correction + constitutive kernel
#ifndef CG_SOLVE_BS
#define CG_SOLVE_BS 64
#endif
constexpr int SBS = CG_SOLVE_BS;
// Block b of a launch runs on XCD b % 8.  With CG_XCD_MAP the solve kernel gives each XCD a
// contiguous eighth of the voxels, walked in order.
#ifndef CG_XCD_MAP
#define CG_XCD_MAP 0
#endif
__device__ __forceinline__ long solve_block(long b, long nb) {
    if (!CG_XCD_MAP) return b;
    const long q = nb / 8;
    return b < 8 * q ? (b % 8) * q + b / 8 : b;
}

// Phase tables seen through the constant address space: with a wave-uniform phase id
// and uniform indices (readfirstlane) the compiler reads them with scalar loads, which
// neither occupy vector registers nor join the in-order vector-memory counter.
typedef __attribute__((address_space(4))) const double cdbl;
typedef __attribute__((address_space(4))) const PhaseDev cphase;
template <bool UNI>
__device__ __forceinline__ int uidx(int i) {
    if constexpr (UNI) return __builtin_amdgcn_readfirstlane(i);
    else return i;
}

static inline int nblk(long V) { return (int)((V + BS - 1) / BS); }

// An exactly singular per-voxel system: the reference stops with status 1 there
// (lu_inverse / lu_eqsystem, tensor_functions.f90:323-334, 353-356).  Every thread that
// meets one stores the same flag (plain stores, no ordering needed); the host reads it
// with the pass' or the update's sums and returns the numerical-failure code.
__device__ __forceinline__ void flag_singular(const Fields& f, long v) {
    f.red[R_FLAG_VOX] = (double)v;
    f.red[R_FLAG_SING] = 1.0;
}

// |x|^(n-1) for an integer n (exponentiation by squaring) or through xpow (exp(y ln x)).
// byte v of a per-voxel byte array of V entries through a buffer resource that ends at V
// (an index past V reads 0)
__device__ __forceinline__ int ld_byte(const uint8_t* a, long V, long v) {
    return __builtin_amdgcn_raw_buffer_load_b8(Vox::rs(a, (uint32_t)V), (int)v, 0, 0);
}
template <bool INTPOW>
__device__ __forceinline__ double powm1(double x, double n) {
    if (INTPOW) {
        int e = (int)n - 1;
        double r = 1.0, b = x;
        while (e > 0) {
            if (e & 1) r *= b;
            b *= b;
            e >>= 1;
        }
        return r;
    } else {
        return xpow(x, n - 1.0);
    }
}

// ---------------------------------------------------------------------------
// final reduction of block partials: out[i] = sum_b part[b*nvals + i]
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BS) void k_reduce_final(const double* __restrict__ part, int nblocks, int nvals,
                                                     double* __restrict__ out) {
    const int i = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += BS) s += part[(size_t)b * nvals + i];
    __shared__ double sm[BS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (lane == 0) sm[wid] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = sm[0];
        for (int w = 1; w < BS / 64; w++) t += sm[w];
        out[i] = t;
    }
}
// the same over a column-major partial array part[i * nblocks + b] (block i sums column i)
__device__ __forceinline__ void reduce_final_cm_blk(const double* __restrict__ part, int nblocks,
                                                    double* __restrict__ out, int i) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += BS) s += part[(size_t)i * nblocks + b];
    __shared__ double sm[BS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (lane == 0) sm[wid] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = sm[0];
        for (int w = 1; w < BS / 64; w++) t += sm[w];
        out[i] = t;
    }
}
__global__ __launch_bounds__(BS) void k_reduce_final_cm(const double* __restrict__ part, int nblocks,
                                                        double* __restrict__ out) {
    reduce_final_cm_blk(part, nblocks, out, blockIdx.x);
}
// First level of the column-major reduction: block (c, i) sums the partials
// [c*CH, (c+1)*CH) of value i into tmp[i * nch + c] (fixed order).
constexpr int RCH = 1024;
__device__ __forceinline__ void reduce_cols_blk(const double* __restrict__ part, int nblocks, double* __restrict__ tmp,
                                                int c, int i, int nch) {
    const long b0 = (long)c * RCH;
    const long n = nblocks - b0 < RCH ? nblocks - b0 : RCH;
    const double* col = part + (size_t)i * nblocks + b0;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < RCH / BS; k++) {
        const long j = threadIdx.x + (long)k * BS;
        if (j < n) s += col[j];
    }
    __shared__ double sm[BS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (lane == 0) sm[wid] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = sm[0];
        for (int w = 1; w < BS / 64; w++) t += sm[w];
        tmp[(size_t)i * nch + c] = t;
    }
}
__global__ __launch_bounds__(BS) void k_reduce_cols(const double* __restrict__ part, int nblocks,
                                                    double* __restrict__ tmp) {
    reduce_cols_blk(part, nblocks, tmp, blockIdx.x, blockIdx.y, gridDim.x);
}
// The same two levels for a group of RVEs (FftWork entries, RVE = blockIdx.z, resp. blockIdx.y):
// nvals columns of nblocks partials in each RVE's f.part, chunk sums after them.
__global__ __launch_bounds__(BS) void k_reduce_cols_bat(const FftWork* __restrict__ w, int nblocks, int nvals) {
    const FftWork& e = w[blockIdx.z];
    reduce_cols_blk(e.f.part, nblocks, e.f.part + (size_t)nvals * nblocks, blockIdx.x, blockIdx.y, gridDim.x);
}
// second level into each RVE's fluct (9 sums of w dLref F^-1)
__global__ __launch_bounds__(BS) void k_reduce_fluct_bat(const FftWork* __restrict__ w, int nblocks, int nch) {
    const FftWork& e = w[blockIdx.y];
    if (nch <= 1) reduce_final_cm_blk(e.f.part, nblocks, e.fluct, blockIdx.x);
    else reduce_final_cm_blk(e.f.part + (size_t)9 * nblocks, nch, e.fluct, blockIdx.x);
}
void launch_reduce_fluct_batch(const FftWork* w, int nrve, int nblocks, hipStream_t s) {
    const int nch = (nblocks + RCH - 1) / RCH;
    if (nch > 1) hipLaunchKernelGGL(k_reduce_cols_bat, dim3(nch, 9, nrve), dim3(BS), 0, s, w, nblocks, 9);
    hipLaunchKernelGGL(k_reduce_fluct_bat, dim3(9, nrve), dim3(BS), 0, s, w, nblocks, nch);
}
// part holds nvals columns of nblocks partials; the space after them (nvals * ceil(nblocks /
// 1024) doubles) is scratch for the two-level reduction of long columns
void launch_reduce_final_cm(double* part, int nblocks, int nvals, double* out, hipStream_t s) {
    const int nch = (nblocks + RCH - 1) / RCH;
    if (nch <= 1) {
        hipLaunchKernelGGL(k_reduce_final_cm, dim3(nvals), dim3(BS), 0, s, part, nblocks, out);
        return;
    }
    double* tmp = part + (size_t)nvals * nblocks;
    hipLaunchKernelGGL(k_reduce_cols, dim3(nch, nvals), dim3(BS), 0, s, part, nblocks, tmp);
    hipLaunchKernelGGL(k_reduce_final_cm, dim3(nvals), dim3(BS), 0, s, tmp, nch, out);
}
// The second level of a pass' sums (after k_reduce_cols, or directly over the partials when one
// chunk covers them): block i sums column i as k_reduce_final_cm does and writes red[i], the
// host's copy (host-mapped memory, so no copy follows), <sigma> for the next pass
// (red[R_SGALL..]) and the solid stress sums the next pass' mixed-form scale is taken from
// (red[R_SGSOL..]); block 0 also copies red[nvals, ncopy) (the flags the solve wrote) to the host.
struct PassFinal {
    const double* part;  // nvals columns of n values
    int n, nvals, ncopy;
    double* red;
    double* host;   // host-mapped, fine-grained
    double* sgsum;   // 9
    double* dscale;  // 2, or null
    unsigned* counter;
    long V;
};
__device__ __forceinline__ void pass_final_blk(const PassFinal& a, int i) {
    double s = 0.0;
    for (int b = threadIdx.x; b < a.n; b += BS) s += a.part[(size_t)i * a.n + b];
    __shared__ double sm[BS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (lane == 0) sm[wid] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = sm[0];
        for (int w = 1; w < BS / 64; w++) t += sm[w];
        a.red[i] = t;
        a.host[i] = t;
        if (i >= R_SGALL && i < R_SGALL + 9) a.sgsum[i - R_SGALL] = t;
    }
    if (i == 0)
        for (int k = a.nvals + threadIdx.x; k < a.ncopy; k += BS) a.host[k] = a.red[k];
    if (a.dscale) {  // the next pass' mixed-form scale, by the block that finishes last
        __shared__ int last;
        if (threadIdx.x == 0) {
            __threadfence();
            last = atomicAdd(a.counter, 1u) == (unsigned)(a.nvals - 1);
        }
        __syncthreads();
        if (last && threadIdx.x == 0) {
            __threadfence();
            double sg[9];
            for (int k = 0; k < 9; k++) sg[k] = __builtin_nontemporal_load(a.red + R_SGSOL + k);
            const double sc = mixed_scale(sg);
            a.dscale[0] = 1.0 / sc;
            a.dscale[1] = sc / (double)a.V;
            *a.counter = 0u;
        }
    }
    __threadfence_system();
}
__global__ __launch_bounds__(BS) void k_pass_final(PassFinal a) { pass_final_blk(a, blockIdx.x); }
__global__ __launch_bounds__(BS) void k_pass_final_bat(const FftWork* __restrict__ w, int nblocks, int nvals,
                                                       int ncopy, int nch) {
    const FftWork& e = w[blockIdx.y];
    PassFinal a{e.f.part, nblocks, nvals, ncopy, e.f.red, e.host, const_cast<double*>(e.sgsum), e.dscale, e.counter,
                e.f.V};
    if (nch > 1) {
        a.part = e.f.part + (size_t)nvals * nblocks;
        a.n = nch;
    }
    pass_final_blk(a, blockIdx.x);
}
// the pass sums of a group of RVEs with the same voxel count (as launch_pass_sums for each)
void launch_pass_sums_batch(const FftWork* w, int nrve, int nblocks, int nvals, int ncopy, hipStream_t s,
                            hipEvent_t done) {
    const int nch = (nblocks + RCH - 1) / RCH;
    if (nch > 1) hipLaunchKernelGGL(k_reduce_cols_bat, dim3(nch, nvals, nrve), dim3(BS), 0, s, w, nblocks, nvals);
    hipExtLaunchKernelGGL(k_pass_final_bat, dim3(nvals, nrve), dim3(BS), 0, s, nullptr, done, 0, w, nblocks, nvals,
                          ncopy, nch);
}
void launch_pass_sums(double* part, int nblocks, int nvals, int ncopy, double* red, double* host, double* sgsum,
                      double* dscale, unsigned* counter, long V, hipStream_t s, hipEvent_t done) {
    const int nch = (nblocks + RCH - 1) / RCH;
    PassFinal a{part, nblocks, nvals, ncopy, red, host, sgsum, dscale, counter, V};
    if (nch > 1) {
        double* tmp = part + (size_t)nvals * nblocks;
        hipLaunchKernelGGL(k_reduce_cols, dim3(nch, nvals), dim3(BS), 0, s, part, nblocks, tmp);
        a.part = tmp;
        a.n = nch;
    }
    hipExtLaunchKernelGGL(k_pass_final, dim3(nvals), dim3(BS), 0, s, nullptr, done, 0, a);
}
void launch_reduce_final(const double* part, int nblocks, int nvals, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_final, dim3(nvals), dim3(BS), 0, s, part, nblocks, nvals, out);
}

// max/min variants for the kinematic / diagnostic scans
template <class Op>
__global__ __launch_bounds__(BS) void k_reduce_final_op(const double* __restrict__ part, int nblocks, int nvals,
                                                        double* __restrict__ out) {
    const int i = blockIdx.x;
    double s = Op::init();
    for (int b = threadIdx.x; b < nblocks; b += BS) s = Op::f(s, part[(size_t)b * nvals + i]);
    __shared__ double sm[BS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1) s = Op::f(s, __shfl_down(s, o, 64));
    if (lane == 0) sm[wid] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = sm[0];
        for (int w = 1; w < BS / 64; w++) t = Op::f(t, sm[w]);
        out[i] = t;
    }
}

// ---------------------------------------------------------------------------
// <sigma> = sum_v w sigma  (LS-EVPFFT.f90:238-254)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BS) void k_sigma_sum(Fields f) {
    const long v = (long)blockIdx.x * BS + threadIdx.x;
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (v < f.V) {
        double s6[6], s[9];
        for (int c = 0; c < 6; c++) s6[c] = f.sg[c * f.ld + v];
        from6(s6, s);
        const double w = f.w[v];
        for (int c = 0; c < 9; c++) acc[c] = s[c] * w;
    }
    block_reduce_store<9, BS>(acc, f.part);
}
void launch_sigma_sum(const Fields& f, hipStream_t s, int* nb) {
    *nb = nblk(f.V);
    hipLaunchKernelGGL(k_sigma_sum, dim3(*nb), dim3(BS), 0, s, f);
}

// ---------------------------------------------------------------------------
// Fused pass: correction_to_current + dL_iter/mixing + solve_res
// ---------------------------------------------------------------------------

// Deviatoric b-basis components of sym(A (b x n) A^-1) = sym((A b) x (A^-T n)) for one
// slip system (update_schmid, various_functions.f90:472-539); A = Fe ag.
__device__ __forceinline__ void schmid5(const double* A, const double* Ai, const double* b3, const double* n3,
                                        double* s5) {
    double ab[3], an[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        ab[i] = A[3 * i] * b3[0] + A[3 * i + 1] * b3[1] + A[3 * i + 2] * b3[2];
        an[i] = Ai[i] * n3[0] + Ai[3 + i] * n3[1] + Ai[6 + i] * n3[2];
    }
    double m[9], s6[6];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) m[3 * i + j] = ab[i] * an[j];
    to6(m, s6);
#pragma unroll
    for (int c = 0; c < 5; c++) s5[c] = s6[c];
}

// Power-law slip rates at sigma (b-basis) -> edotp (5), dedotp (sym 5x5, 15), optional gdot store.
template <bool INTPOW, bool STORE>
__device__ __forceinline__ void slip_eval(const Fields& f, const double* __restrict__ tab, const PhaseDev& P,
                                          const Vox& x, const double* sg, double* ep, double* dd) {
    for (int i = 0; i < 5; i++) ep[i] = 0.0;
    for (int i = 0; i < 15; i++) dd[i] = 0.0;
    const double* nrs = tab + P.off_nrs;
    const double* g0 = tab + P.off_gamd0;
#pragma unroll 1
    for (int is = 0; is < P.nsyst; is++) {
        const double s0 = x.ld(f.sch, 5 * is + 0), s1 = x.ld(f.sch, 5 * is + 1), s2 = x.ld(f.sch, 5 * is + 2),
                     s3 = x.ld(f.sch, 5 * is + 3), s4 = x.ld(f.sch, 5 * is + 4);
        double rss = s0 * sg[0] + s1 * sg[1] + s2 * sg[2] + s3 * sg[3] + s4 * sg[4];
        const int isg = rss < 0.0 ? 1 : 0;
        const double tau = x.ld(f.ttau, 2 * is + isg);
        rss = rss / tau;
        const double n = nrs[is];
        const double dum = g0[is] * powm1<INTPOW>(fabs(rss), n);
        const double r1 = n * dum / tau;
        const double r2 = rss * dum;
        if (STORE) x.st(f.gdot, is, r2);
        ep[0] += s0 * r2; ep[1] += s1 * r2; ep[2] += s2 * r2; ep[3] += s3 * r2; ep[4] += s4 * r2;
        const double t0 = s0 * r1, t1 = s1 * r1, t2 = s2 * r1, t3 = s3 * r1, t4 = s4 * r1;
        dd[0] += s0 * t0; dd[1] += s0 * t1; dd[2] += s0 * t2; dd[3] += s0 * t3; dd[4] += s0 * t4;
        dd[5] += s1 * t1; dd[6] += s1 * t2; dd[7] += s1 * t3; dd[8] += s1 * t4;
        dd[9] += s2 * t2; dd[10] += s2 * t3; dd[11] += s2 * t4;
        dd[12] += s3 * t3; dd[13] += s3 * t4;
        dd[14] += s4 * t4;
    }
}
// slip_eval with the Schmid tensors rebuilt from the A, A^-1 that update_schmid stored (the
// same schmid5), for kernels of contexts that do not keep sch
template <bool INTPOW>
__device__ __forceinline__ void slip_eval_A(const Fields& f, const double* __restrict__ tab, const PhaseDev& P,
                                            const Vox& x, const double* sg, double* ep, double* dd) {
    double A[9], Ai[9];
    for (int k = 0; k < 9; k++) {
        A[k] = x.ld(f.schA, k);
        Ai[k] = x.ld(f.schA, 9 + k);
    }
    for (int i = 0; i < 5; i++) ep[i] = 0.0;
    for (int i = 0; i < 15; i++) dd[i] = 0.0;
    const double* nrs = tab + P.off_nrs;
    const double* g0 = tab + P.off_gamd0;
#pragma unroll 1
    for (int is = 0; is < P.nsyst; is++) {
        double sv[5];
        schmid5(A, Ai, tab + P.off_dbca + 3 * is, tab + P.off_dnca + 3 * is, sv);
        const double s0 = sv[0], s1 = sv[1], s2 = sv[2], s3 = sv[3], s4 = sv[4];
        double rss = s0 * sg[0] + s1 * sg[1] + s2 * sg[2] + s3 * sg[3] + s4 * sg[4];
        const int isg = rss < 0.0 ? 1 : 0;
        const double tau = x.ld(f.ttau, 2 * is + isg);
        rss = rss / tau;
        const double n = nrs[is];
        const double dum = g0[is] * powm1<INTPOW>(fabs(rss), n);
        const double r1 = n * dum / tau;
        const double r2 = rss * dum;
        ep[0] += s0 * r2; ep[1] += s1 * r2; ep[2] += s2 * r2; ep[3] += s3 * r2; ep[4] += s4 * r2;
        const double t0 = s0 * r1, t1 = s1 * r1, t2 = s2 * r1, t3 = s3 * r1, t4 = s4 * r1;
        dd[0] += s0 * t0; dd[1] += s0 * t1; dd[2] += s0 * t2; dd[3] += s0 * t3; dd[4] += s0 * t4;
        dd[5] += s1 * t1; dd[6] += s1 * t2; dd[7] += s1 * t3; dd[8] += s1 * t4;
        dd[9] += s2 * t2; dd[10] += s2 * t3; dd[11] += s2 * t4;
        dd[12] += s3 * t3; dd[13] += s3 * t4;
        dd[14] += s4 * t4;
    }
}
// slip_eval for the symmetric Newton solve: the Schmid tensors are rebuilt from A, A^-1
// (this thread's column of an LDS array [18][SBS]) and the wave-uniform crystal b, n, so
// an evaluation reads 2 n_s trial strengths from memory instead of 5 n_s Schmid
// components; the next system's strengths are loaded while the current one is computed.
// SYMTAU: the forward and backward trial strengths are equal (tau0f = tau0b for every
// system: init and the increment's final harden write the two columns bitwise equal), so
// the forward one is read, and the hardens inside the solve write only that one.
// ACC: accumulate edotp and its derivative; STORE, store: store gdot (the last evaluation's
// rates feed harden and the per-increment updates).
// Trial strengths are read for a group of TauGroup systems at a time, one group ahead:
// the loads of group g+1 are in flight while group g computes.  Group indices are
// clamped to the last system so that every load is unconditional and the wait counts
// stay exact.  tpn/tnn carry group 0 in and out: the caller loads it once after each
// change of the strengths, and the last group of an evaluation prefetches group 0 for
// the next evaluation.
template <bool SYMTAU>
struct TauGroup {
    static constexpr int G = SYMTAU ? 4 : 2;
};
template <bool SYMTAU>
__device__ __forceinline__ void load_tau_group(const Fields& f, const Vox& x, int ns, int g0, double* tpn,
                                               double* tnn) {
#pragma unroll
    for (int k = 0; k < TauGroup<SYMTAU>::G; k++) {
        const int c = __builtin_amdgcn_readfirstlane(g0 + k < ns ? g0 + k : ns - 1);
        tpn[k] = x.ld(f.ttau, 2 * c);
        if (!SYMTAU) tnn[k] = x.ld(f.ttau, 2 * c + 1);
    }
}
template <bool INTPOW, bool SYMTAU, bool ACC, bool STORE, class PH, class TB>
__device__ __forceinline__ void slip_eval_ab(const Fields& f, TB tab, const PH& P, const Vox& x, const double* sal,
                                             const double* sg, double* ep, double* dd, double* tpn, double* tnn,
                                             bool store = true) {
    double A[9], Ai[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        A[k] = sal[k * SBS];
        Ai[k] = sal[(9 + k) * SBS];
    }
    if (ACC) {
        for (int i = 0; i < 5; i++) ep[i] = 0.0;
        for (int i = 0; i < 15; i++) dd[i] = 0.0;
    }
    const int ns = P.nsyst;
    constexpr int G = TauGroup<SYMTAU>::G;
#pragma unroll 1
    for (int g0 = 0; g0 < ns; g0 += G) {
        const int gb = __builtin_amdgcn_readfirstlane(g0);
        const int nx = __builtin_amdgcn_readfirstlane(gb + G < ns ? gb + G : 0);  // next group (wraps to 0)
        double tp[G], tn[G];
#pragma unroll
        for (int k = 0; k < G; k++) {
            tp[k] = tpn[k];
            tn[k] = SYMTAU ? tpn[k] : tnn[k];
        }
        load_tau_group<SYMTAU>(f, x, ns, nx, tpn, tnn);
#pragma unroll
        for (int k = 0; k < G; k++) {
            const int is = gb + k;
            if (is >= ns) break;
            double sv[5], b3[3], n3[3];
            const int rb = __builtin_amdgcn_readfirstlane(P.off_sys + 8 * is);  // b, n, nrs, gamd0
            for (int q = 0; q < 3; q++) {
                b3[q] = tab[rb + q];
                n3[q] = tab[rb + 3 + q];
            }
            schmid5(A, Ai, b3, n3, sv);
            const double s0 = sv[0], s1 = sv[1], s2 = sv[2], s3 = sv[3], s4 = sv[4];
            double rss = s0 * sg[0] + s1 * sg[1] + s2 * sg[2] + s3 * sg[3] + s4 * sg[4];
            const double ti = recip(rss < 0.0 ? tn[k] : tp[k]);
            rss = rss * ti;
            const double n = tab[rb + 6];
            const double dum = tab[rb + 7] * powm1<INTPOW>(fabs(rss), n);
            const double r2 = rss * dum;
            if (STORE && store) x.st(f.gdot, is, r2);
            if (ACC) {
                const double r1 = n * dum * ti;
                ep[0] += s0 * r2; ep[1] += s1 * r2; ep[2] += s2 * r2; ep[3] += s3 * r2; ep[4] += s4 * r2;
                const double t0 = s0 * r1, t1 = s1 * r1, t2 = s2 * r1, t3 = s3 * r1, t4 = s4 * r1;
                dd[0] += s0 * t0; dd[1] += s0 * t1; dd[2] += s0 * t2; dd[3] += s0 * t3; dd[4] += s0 * t4;
                dd[5] += s1 * t1; dd[6] += s1 * t2; dd[7] += s1 * t3; dd[8] += s1 * t4;
                dd[9] += s2 * t2; dd[10] += s2 * t3; dd[11] += s2 * t4;
                dd[12] += s3 * t3; dd[13] += s3 * t4;
                dd[14] += s4 * t4;
            }
            // one system at a time: interleaving the unrolled systems raises the register
            // pressure past the two-wave budget (spills), and gains nothing at two waves per SIMD
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}
// gdot only (for harden and the final store): same arithmetic as slip_eval
template <bool INTPOW>
__device__ __forceinline__ void slip_rates(const Fields& f, const double* __restrict__ tab, const PhaseDev& P,
                                           const Vox& x, const double* sg) {
    const double* nrs = tab + P.off_nrs;
    const double* g0 = tab + P.off_gamd0;
    for (int is = 0; is < P.nsyst; is++) {
        const double s0 = x.ld(f.sch, 5 * is + 0), s1 = x.ld(f.sch, 5 * is + 1), s2 = x.ld(f.sch, 5 * is + 2),
                     s3 = x.ld(f.sch, 5 * is + 3), s4 = x.ld(f.sch, 5 * is + 4);
        double rss = s0 * sg[0] + s1 * sg[1] + s2 * sg[2] + s3 * sg[3] + s4 * sg[4];
        const int isg = rss < 0.0 ? 1 : 0;
        const double tau = x.ld(f.ttau, 2 * is + isg);
        rss = rss / tau;
        const double dum = g0[is] * powm1<INTPOW>(fabs(rss), nrs[is]);
        x.st(f.gdot, is, rss * dum);
    }
}
__device__ __forceinline__ int sym5(int i, int j) {  // packed index of the symmetric 5x5
    if (i > j) { int t = i; i = j; j = t; }
    return i * 5 - (i * (i - 1)) / 2 + (j - i);
}

// Voce hardening (harden, various_functions.f90:1439-1491) from f.gdot; writes the
// new trial strengths to ttau; returns ||ttau_old - ttau_new||^2.
template <bool UNI = false, class PH = PhaseDev, class TB = const double*>
__device__ __forceinline__ double harden_voxel(const Fields& f, TB tab, const PH& P, const Vox& x, double tdot,
                                               double* gamma1_out, bool store_crss) {
    const int ns = P.nsyst;
    constexpr int NR = 12;  // up to this many systems the |gdot| stay in registers
    double gr[NR];
    const bool inreg = ns <= NR;
    double deltgam = 0.0;
    if (inreg) {
#pragma unroll
        for (int js = 0; js < NR; js++) gr[js] = js < ns ? fabs(x.ld(f.gdot, js)) : 0.0;
#pragma unroll
        for (int js = 0; js < NR; js++)
            if (js < ns) deltgam += gr[js] * tdot;
    } else {
        for (int is = 0; is < ns; is++) deltgam += fabs(x.ld(f.gdot, is)) * tdot;
    }
    const double gamma0 = x.ld(f.gacc, 0);
    const double gamma1 = gamma0 + deltgam;
    *gamma1_out = gamma1;
    const TB hard = tab + P.off_hard;
    double e2 = 0.0, tn2 = 0.0;  // tn2: |new ttau|^2, columns in order
    int is = 0;
    for (int m0 = 0; m0 < P.nmodes; m0++) {
        const int m = uidx<UNI>(m0);
        const double tau1 = P.mode_tau1[m], th0 = P.mode_thet0[m], th1 = P.mode_thet1[m];
        const double fact = fabs(th0 / tau1);
        const double ex0 = exp(-fact * gamma0), ex1 = exp(-fact * gamma1);
        const double voce = deltgam * th1 + tau1 * (ex0 - ex1) + th1 * (gamma0 * ex0 - gamma1 * ex1);
        const double sc = voce * tdot / deltgam;
        for (int k = 0; k < P.nsm[m]; k++, is++) {
            is = uidx<UNI>(is);
            double dtau = 0.0;
            if (inreg) {
#pragma unroll
                for (int js = 0; js < NR; js++)
                    if (js < ns) dtau += hard[is * ns + js] * gr[js];
            } else {
                for (int js = 0; js < ns; js++) dtau += hard[is * ns + js] * fabs(x.ld(f.gdot, js));
            }
            const double t0 = x.ld(f.crss, 2 * is) + dtau * sc;
            const double t1 = x.ld(f.crss, 2 * is + 1) + dtau * sc;
            const double o0 = x.ld(f.ttau, 2 * is), o1 = x.ld(f.ttau, 2 * is + 1);
            e2 += (o0 - t0) * (o0 - t0) + (o1 - t1) * (o1 - t1);
            tn2 += t0 * t0;
            tn2 += t1 * t1;
            x.st(f.ttau, 2 * is, t0);
            x.st(f.ttau, 2 * is + 1, t1);
            if (store_crss) {
                x.st(f.crss, 2 * is, t0);
                x.st(f.crss, 2 * is + 1, t1);
            }
        }
    }
    x.st(f.taun, 0, sqrt(tn2));
    return e2;
}

// harden_voxel inside the solve for phases of one slip mode with at most 12 systems, written without
// branches on the system count: every load is unconditional (indices clamped to the last
// system, unused rates weighted by zero), so the rates, gacc and the first chunk's crss
// and old strengths are in flight together, and each later chunk of HC systems is read
// while the previous one computes.  The sums run over the same terms in the same order
// as harden_voxel (the added terms are exact zeros).  FINAL: the increment's last harden
// (update_hardening): crss and both strength columns take the new values, gacc the new
// accumulated shear, and the old strengths are not read.
template <bool SYMTAU, bool FINAL = false, class PH, class TB>
__device__ __forceinline__ double harden_small(const Fields& f, TB tab, const PH& P, const Vox& x, double tdot) {
    const int ns = __builtin_amdgcn_readfirstlane(P.nsyst);
    const int oh = __builtin_amdgcn_readfirstlane(P.off_hard);
    constexpr int NR = 12, HC = 4;
    auto clampi = [&](int i) { return __builtin_amdgcn_readfirstlane(i < ns ? i : ns - 1); };
    double cr[2 * HC], ot[2 * HC];
    auto load_chunk = [&](int c0, double* c2, double* o2) {
#pragma unroll
        for (int k = 0; k < HC; k++) {
            const int i = clampi(c0 + k);
            c2[2 * k] = x.ld(f.crss, 2 * i);
            o2[2 * k] = FINAL ? 0.0 : x.ld(f.ttau, 2 * i);
            // with equal forward and backward strengths the forward column stands for both
            c2[2 * k + 1] = SYMTAU ? c2[2 * k] : x.ld(f.crss, 2 * i + 1);
            o2[2 * k + 1] = SYMTAU || FINAL ? o2[2 * k] : x.ld(f.ttau, 2 * i + 1);
        }
    };
    load_chunk(0, cr, ot);
    double gr[NR];
#pragma unroll
    for (int js = 0; js < NR; js++) gr[js] = fabs(x.ld(f.gdot, clampi(js)));
#pragma unroll
    for (int js = 0; js < NR; js++) gr[js] = js < ns ? gr[js] : 0.0;
    const double gamma0 = x.ld(f.gacc, 0);
    double deltgam = 0.0;
#pragma unroll
    for (int js = 0; js < NR; js++) deltgam += gr[js] * tdot;
    const double gamma1 = gamma0 + deltgam;
    // Voce factor of the (single) mode
    double sc;
    {
        const double tau1 = P.mode_tau1[0], th0 = P.mode_thet0[0], th1 = P.mode_thet1[0];
        const double fact = fabs(th0 / tau1);
        const double ex0 = exp(-fact * gamma0), ex1 = exp(-fact * gamma1);
        const double voce = deltgam * th1 + tau1 * (ex0 - ex1) + th1 * (gamma0 * ex0 - gamma1 * ex1);
        sc = voce * tdot / deltgam;
    }
    double e2 = 0.0, tn2 = 0.0;  // tn2: |new ttau|^2, columns in order
#pragma unroll
    for (int c0 = 0; c0 < NR; c0 += HC) {
        if (c0 >= ns) break;
        double crn[2 * HC], otn[2 * HC];
        load_chunk(c0 + HC < ns ? c0 + HC : c0, crn, otn);
#pragma unroll
        for (int k = 0; k < HC; k++) {
            const int is = c0 + k;
            if (is >= ns) break;
            double dtau = 0.0;
#pragma unroll
            for (int js = 0; js < NR; js++)
                dtau += tab[__builtin_amdgcn_readfirstlane(oh + is * ns) + js] * gr[js];  // table padded
            const double t0 = cr[2 * k] + dtau * sc;
            const double t1 = cr[2 * k + 1] + dtau * sc;
            const double o0 = ot[2 * k], o1 = ot[2 * k + 1];
            if (!FINAL) e2 += (o0 - t0) * (o0 - t0) + (o1 - t1) * (o1 - t1);
            tn2 += t0 * t0;
            tn2 += t1 * t1;
            x.st(f.ttau, 2 * is, t0);
            if (!SYMTAU || FINAL) x.st(f.ttau, 2 * is + 1, t1);
            if (FINAL) {
                x.st(f.crss, 2 * is, t0);
                x.st(f.crss, 2 * is + 1, t1);
            }
        }
#pragma unroll
        for (int k = 0; k < 2 * HC; k++) {
            cr[k] = crn[k];
            ot[k] = otn[k];
        }
    }
    x.st(f.taun, 0, sqrt(tn2));
    if (FINAL) x.st(f.gacc, 0, gamma1);
    return e2;
}

// harden_voxel for phases with more than 12 systems (or several modes): dtau of 12 systems
// of a mode at a time, the rates read 4 at a time (in flight together) and each |gdot_js|
// once per chunk; the rows of the interaction matrix are read 4 wide (the table is padded;
// rates past the last system weigh zero).  dtau sums the same terms in the same order as
// harden_voxel.  SYMTAU: as harden_small.
template <bool SYMTAU, class PH, class TB>
__device__ __forceinline__ double harden_voxel_chunked(const Fields& f, TB tab, const PH& P, const Vox& x,
                                                       double tdot) {
    const int ns = __builtin_amdgcn_readfirstlane(P.nsyst);
    const int oh = __builtin_amdgcn_readfirstlane(P.off_hard);
    auto clampi = [&](int i) { return __builtin_amdgcn_readfirstlane(i < ns ? i : ns - 1); };
    constexpr int JB = 4, C = 12, SB = 4;
    auto rates = [&](int j0, double* g) {
#pragma unroll
        for (int jj = 0; jj < JB; jj++) g[jj] = fabs(x.ld(f.gdot, clampi(j0 + jj)));
#pragma unroll
        for (int jj = 0; jj < JB; jj++) g[jj] = j0 + jj < ns ? g[jj] : 0.0;
    };
    double deltgam = 0.0;
    for (int j00 = 0; j00 < ns; j00 += JB) {
        const int j0 = __builtin_amdgcn_readfirstlane(j00);
        double g[JB];
        rates(j0, g);
#pragma unroll
        for (int jj = 0; jj < JB; jj++) deltgam += g[jj] * tdot;
    }
    const double gamma0 = x.ld(f.gacc, 0);
    const double gamma1 = gamma0 + deltgam;
    double e2 = 0.0, tn2 = 0.0;  // tn2: |new ttau|^2, columns in order
    int is = 0;
    for (int m0 = 0; m0 < P.nmodes; m0++) {
        const int m = __builtin_amdgcn_readfirstlane(m0);
        const double tau1 = P.mode_tau1[m], th0 = P.mode_thet0[m], th1 = P.mode_thet1[m];
        const double fact = fabs(th0 / tau1);
        const double ex0 = exp(-fact * gamma0), ex1 = exp(-fact * gamma1);
        const double voce = deltgam * th1 + tau1 * (ex0 - ex1) + th1 * (gamma0 * ex0 - gamma1 * ex1);
        const double sc = voce * tdot / deltgam;
        const int nm = __builtin_amdgcn_readfirstlane(P.nsm[m]);
        for (int k0 = 0; k0 < nm; k0 += C) {
            const int i0 = __builtin_amdgcn_readfirstlane(is + k0);
            const int nk = __builtin_amdgcn_readfirstlane(nm - k0 < C ? nm - k0 : C);
            double dt[C];
#pragma unroll
            for (int k = 0; k < C; k++) dt[k] = 0.0;
            for (int j00 = 0; j00 < ns; j00 += JB) {
                const int j0 = __builtin_amdgcn_readfirstlane(j00);
                double g[JB];
                rates(j0, g);
#pragma unroll
                for (int k = 0; k < C; k++)
                    if (k < nk) {
                        const int row = __builtin_amdgcn_readfirstlane(oh + (i0 + k) * ns + j0);
#pragma unroll
                        for (int jj = 0; jj < JB; jj++) dt[k] += tab[row + jj] * g[jj];
                    }
            }
            // new strengths, SB systems at a time (their crss and old strengths in flight together)
#pragma unroll
            for (int s0 = 0; s0 < C; s0 += SB) {
                if (s0 >= nk) break;
                double cr[2 * SB], ot[2 * SB];
#pragma unroll
                for (int k = 0; k < SB; k++) {
                    const int i = __builtin_amdgcn_readfirstlane(i0 + (s0 + k < nk ? s0 + k : nk - 1));
                    cr[2 * k] = x.ld(f.crss, 2 * i);
                    cr[2 * k + 1] = SYMTAU ? cr[2 * k] : x.ld(f.crss, 2 * i + 1);
                    ot[2 * k] = x.ld(f.ttau, 2 * i);
                    ot[2 * k + 1] = SYMTAU ? ot[2 * k] : x.ld(f.ttau, 2 * i + 1);
                }
#pragma unroll
                for (int k = 0; k < SB; k++) {
                    if (s0 + k >= nk) break;
                    const int i = i0 + s0 + k;
                    const double t0 = cr[2 * k] + dt[s0 + k] * sc;
                    const double t1 = cr[2 * k + 1] + dt[s0 + k] * sc;
                    const double o0 = ot[2 * k], o1 = ot[2 * k + 1];
                    e2 += (o0 - t0) * (o0 - t0) + (o1 - t1) * (o1 - t1);
                    tn2 += t0 * t0;
                    tn2 += t1 * t1;
                    x.st(f.ttau, 2 * i, t0);
                    if (!SYMTAU) x.st(f.ttau, 2 * i + 1, t1);
                }
            }
        }
        is += nm;
    }
    x.st(f.taun, 0, sqrt(tn2));
    return e2;
}

template <bool INTPOW, bool SYMTAU, bool SMALLNS>
// p0 is wave-uniform (the caller iterates over the phases present in the wave), so the
// phase tables are read through scalar loads.
// cml: this thread's column of a block-shared LDS array [21][SBS] holding the L D L^T
// factors of cm^-1, read once per evaluation for the residual norm |cm R'| (keeping them out
// of the registers keeps the Newton state within two waves per SIMD).
// sal: this thread's column of an LDS array [18][SBS] that receives A, A^-1 (schA).
// pA, plam, psgt, pF, pS: A = Fe ag, lambda, sgt and the 42 operators of k_solve_ops, loaded by
// the caller before phase A (their loads overlap phase A's).
__device__ __forceinline__ void solve_solid_sym(const Fields& f, const PhaseTables& pt, const PassConst& pc,
                                                const Vox& x, int p0, double* cml, double* sal, double* sgf,
                                                double* acc, double w, const double* Lmix, const double* pA,
                                                const double* plam, const double* psgt, const double* pF,
                                                const double* pS) {
    const cphase& P = ((cphase*)pt.ph)[p0];
    cdbl* tab = (cdbl*)pt.data;
    const double tdot = pc.tdot;
    {  // A = Fe ag from memory; A^-1 by the same inv3 as update_schmid's
        double Am[9], Ai[9];
#pragma unroll
        for (int k = 0; k < 9; k++) Am[k] = pA[k];
        inv3(Am, Ai);
#pragma unroll
        for (int k = 0; k < 9; k++) {
            sal[k * SBS] = Am[k];
            sal[(9 + k) * SBS] = Ai[k];
        }
    }
    double A[21], b[6], sg[6];
    {
        double lam[6], sgt[6], D[6];
        for (int c = 0; c < 6; c++) {
            lam[c] = plam[c];
            sgt[c] = psgt[c];
            sg[c] = lam[c];  // the solve starts from lambda
        }
        to6(Lmix, D);
        // ithermo = 1, first increment of the process: edot gains the eigenstrain rate
        // (various_functions.f90:1226-1227, 1265); in R = sg - lambda + cm (edot - D) that is D - eth
        if (pc.thermo)
            for (int c = 0; c < 6; c++) D[c] -= x.ld(f.eth, c);
        // the operators of this increment (k_solve_ops): the L D L^T factors of cm^-1 (unit
        // L below the diagonal, 1/d on it) -> LDS, for the residual norm |cm R'|; cm^-1 =
        // L D L^T rebuilt here; A = cm^-1 + S, b = cm^-1 lambda + D + S sgt, S = sym(s66)/tdot
        {
            double F[21], dk[6];
#pragma unroll
            for (int i = 0; i < 21; i++) F[i] = pF[i];
#pragma unroll
            for (int i = 0; i < 21; i++) cml[i * SBS] = F[i];
#pragma unroll
            for (int k = 0; k < 6; k++) dk[k] = 1.0 / F[sp6(k, k)];
#pragma unroll
            for (int i = 0; i < 6; i++)
#pragma unroll
                for (int j = i; j < 6; j++) {
                    double s = 0.0;
#pragma unroll
                    for (int k = 0; k < i; k++) s += F[sp6(i, k)] * dk[k] * F[sp6(j, k)];
                    s += dk[i] * (j == i ? 1.0 : F[sp6(j, i)]);
                    A[sp6(i, j)] = s;
                }
        }
        spmv6(A, lam, b);
        for (int i = 0; i < 6; i++) b[i] += D[i];
#pragma unroll
        for (int i = 0; i < 6; i++)
#pragma unroll
            for (int j = i; j < 6; j++) {
                const double sij = pS[sp6(i, j)];
                A[sp6(i, j)] += sij;
                b[i] += sij * sgt[j];
                if (j != i) b[j] += sij * sgt[i];
            }
    }
    // The stress of the last evaluation is sgold after the line search, unless the
    // search ended on a cut (icut = 10); that rare stress is parked in the edotp
    // columns (written with the final edotp below) and flagged.
    double sgold[6] = {0, 0, 0, 0, 0, 0}, ep[5], dd[15], rp[6];
    bool parked = false;
    double sgnorm = 0.0, taun = 0.0;
    for (int c = 0; c < 6; c++) sgnorm += sg[c] * sg[c];
    sgnorm = sqrt(sgnorm);
    taun = x.ld(f.taun, 0);  // |ttau| as the last harden (or k_taun) left it
    const double tol = fmax(pc.error, pc.erroral);
    int itertau = 0;
    double errtau = 2.0 * tol;
    int nnr = 0, nev = 0, nout = 0;
    double tpn[TauGroup<SYMTAU>::G], tnn[TauGroup<SYMTAU>::G];
    while (itertau < 10 && errtau > tol) {
        itertau++;
        nout++;
        load_tau_group<SYMTAU>(f, x, P.nsyst, 0, tpn, tnn);  // the strengths changed (harden) or are new
        int iternr = 0;
        double erral = 2.0 * pc.erroral, resoldn = 1.0;
        while (iternr < 100 && erral > pc.erroral) {
            iternr++;
            nnr++;
            int icut = 0;
            double resn = 2.0 * resoldn;
            while (resn > resoldn && icut < 10) {
                nev++;
                parked = false;
                slip_eval_ab<INTPOW, SYMTAU, true, true>(f, tab, P, x, sal, sg, ep, dd, tpn, tnn, iternr > 1);
                spmv6(A, sg, rp);
#pragma unroll
                for (int i = 0; i < 6; i++) rp[i] += (i < 5 ? ep[i] : 0.0) - b[i];
                double rn = 0.0;
                {  // |cm R'| with cm = (L D L^T)^-1: y = L^-T D^-1 L^-1 R'
                    double y[6];
#pragma unroll
                    for (int i = 0; i < 6; i++) {
                        double s = rp[i];
#pragma unroll
                        for (int k = 0; k < i; k++) s -= cml[sp6(i, k) * SBS] * y[k];
                        y[i] = s;
                    }
#pragma unroll
                    for (int i = 0; i < 6; i++) y[i] *= cml[sp6(i, i) * SBS];
#pragma unroll
                    for (int i = 4; i >= 0; i--) {
                        double s = y[i];
#pragma unroll
                        for (int k = i + 1; k < 6; k++) s -= cml[sp6(k, i) * SBS] * y[k];
                        y[i] = s;
                    }
#pragma unroll
                    for (int i = 0; i < 6; i++) rn += y[i] * y[i];
                }
                resn = sqrt(rn);
                if (iternr == 1) resoldn = resn;
                if (resn > resoldn) {
                    for (int c = 0; c < 6; c++) x.st(f.edotp, c, sg[c]);
                    parked = true;
                    for (int c = 0; c < 6; c++) sg[c] = sgold[c] + (sg[c] - sgold[c]) * 0.5;
                }
                icut++;
            }
            resoldn = resn;
            for (int c = 0; c < 6; c++) sgold[c] = sg[c];
            double K[21];
#pragma unroll
            for (int i = 0; i < 6; i++)
#pragma unroll
                for (int j = i; j < 6; j++) K[sp6(i, j)] = A[sp6(i, j)] + (j < 5 ? dd[sym5(i, j)] : 0.0);
            {
                double rd[6];
                ldl6r(K, rd);
                ldl6_solve_r(K, rd, rp);
            }
            double d1 = 0.0;
            for (int c = 0; c < 6; c++) {
                sg[c] -= rp[c];
                d1 += (sg[c] - sgold[c]) * (sg[c] - sgold[c]);
            }
            erral = sqrt(d1) / sgnorm;
        }
        // The first Newton step's evaluation does not store its rates: a loop ends there
        // only when that step is already below the tolerance (its one evaluation, at
        // sgold, makes no line-search cut), and then the rates are evaluated again.
        if (iternr == 1) slip_eval_ab<INTPOW, SYMTAU, false, true>(f, tab, P, x, sal, sgold, ep, dd, tpn, tnn);
        // harden with the slip rates of the last evaluation
        if (pc.iuphard == 1) {
            double g1;
            const double e2 = SMALLNS ? harden_small<SYMTAU>(f, tab, P, x, tdot) : harden_voxel_chunked<SYMTAU>(f, tab, P, x, tdot);
            errtau = sqrt(e2) / taun;
        } else {
            errtau = 0.0;
        }
    }
    // errors: dsgnorm = |sg - lambda|^2, ddgnorm = |e|^2 with e = edot(last evaluation) - D
    // = ep + S (sev - sgt) - D.  R'(s) = A s - b + ep = cm^-1 (s - lambda) + e(s), so e is
    // R'(sev) - cm^-1 (sev - lambda), formed from A, b in registers and the L D L^T factors
    // of cm^-1 in LDS: of the inputs only lambda is read again.
    double dsgn = 0.0, dd2 = 0.0;
    {
        double lam[6], sev[6], e[6], q[6];
        for (int c = 0; c < 6; c++) lam[c] = x.ld(f.sg, c);
        for (int c = 0; c < 6; c++) dsgn += (sg[c] - lam[c]) * (sg[c] - lam[c]);
        for (int c = 0; c < 6; c++) sev[c] = parked ? x.ld(f.edotp, c) : sgold[c];
        spmv6(A, sev, e);
#pragma unroll
        for (int i = 0; i < 6; i++) {
            e[i] += (i < 5 ? ep[i] : 0.0) - b[i];
            q[i] = sev[i] - lam[i];
        }
        // q <- cm^-1 q = L (D (L^T q))
#pragma unroll
        for (int i = 0; i < 6; i++) {
            double s = q[i];
#pragma unroll
            for (int k = i + 1; k < 6; k++) s += cml[sp6(k, i) * SBS] * q[k];
            q[i] = s / cml[sp6(i, i) * SBS];
        }
#pragma unroll
        for (int i = 5; i >= 0; i--) {
            double s = q[i];
#pragma unroll
            for (int k = 0; k < i; k++) s += cml[sp6(i, k) * SBS] * q[k];
            q[i] = s;
        }
#pragma unroll
        for (int i = 0; i < 6; i++) dd2 += (e[i] - q[i]) * (e[i] - q[i]);
    }
    acc[R_ERRS] = dsgn * w;
    acc[R_ERRE] = dd2 * w;
    for (int c = 0; c < 6; c++) sgf[c] = sg[c];
    for (int c = 0; c < 5; c++) x.st(f.edotp, c, ep[c]);
    x.st(f.edotp, 5, 0.0);
    acc[R_NR] = (double)nnr;
    acc[R_EVAL] = (double)nev;
    acc[R_OUTER] = (double)nout;
}

// One 64-voxel block of the correction + constitutive kernel: block blk of the nblk blocks
// that cover this RVE.  pool: the block's LDS, per thread 21 (packed cm) + 18 (A, A^-1)
// doubles of the symmetric solve; the block reductions reuse it when no solve is using it.
template <bool MODAL, bool INTPOW, bool SYMTAU, bool SMALLNS>
__device__ __forceinline__ void correct_solve_block(const Fields& f, const PhaseTables& pt, const SolveArgs& a,
                                                    long blk, long nblk, double* pool) {
    const long V = f.V;
    const long v = blk * SBS + threadIdx.x;
    const bool act = v < V;
    const PassConst& pc = a.pc;
    const Vox x(v, f.ld);
    double* const cm_lds = pool;
    double* const sa_lds = pool + 21 * SBS;
    // Every load that depends on nothing is issued first, by every lane and in one basic block,
    // so that the loads of phase A and of the solve's inputs are in flight together: the
    // weight, the gas and phase bytes, L, F^-1 and dLref and, for the symmetric solve,
    // A = Fe ag, lambda and sgt; the 42 per-increment operators follow at the end of phase A's
    // arithmetic, ahead of its block reduction.  Lanes past V read column padding (every column
    // has at least 256 doubles past V) and bytes past V read 0 (the byte resources end at V);
    // they compute and store nothing.
    const double w0 = x.ld(f.w, 0);
    const int gasb = ld_byte(f.gas, V, v), ph = ld_byte(f.phase, V, v);
    double pA[9], plam[6], psgt[6], pF[21], pS[21];
    if (!MODAL) {
#pragma unroll
        for (int k = 0; k < 9; k++) pA[k] = x.ld(f.schA, k);
#pragma unroll
        for (int c = 0; c < 6; c++) {
            plam[c] = x.ld(f.sg, c);
            psgt[c] = x.ld(f.sgt, c);
        }
    }
    double L[9], Fi[9], t[9];
#pragma unroll
    for (int c = 0; c < 9; c++) {
        L[c] = x.ld(f.L, c);
        Fi[c] = x.ld(f.Fi, c);
        t[c] = x.ld(f.dLref, c);
    }
    double w = 0.0, Lmix[9];  // Lmix: this voxel's velocity gradient after the mixing (phase A)
    bool gas = false;
    // ---- phase A: correction_to_current + dL_iter/mixing; its 31 sums are reduced
    // before the constitutive solve so their accumulators are not live across it
    {
        double accA[R_NSOLVE - R_NPHASEB];
#pragma unroll
        for (int i = 0; i < R_NSOLVE - R_NPHASEB; i++) accA[i] = 0.0;
        if (act) {
            // dudot from the reduced fluctuation average
            double dud[9];
            {
                double dm[9];
                for (int i = 0; i < 9; i++) dm[i] = a.dvelgradavg[i] - a.fluct[i];
                mm3(dm, a.finv_inv, dud);
            }
            w = w0;
            gas = gasb != 0;
            double Lo[9];
            for (int c = 0; c < 9; c++) t[c] += dud[c];
            if (pc.lold_same) {
                for (int c = 0; c < 9; c++) Lo[c] = L[c];
            } else {
                for (int c = 0; c < 9; c++) Lo[c] = x.ld(f.Lold, c);
            }
            double q[9];
            mm3(t, Fi, q);
            for (int c = 0; c < 9; c++) L[c] += q[c];
            for (int c = 0; c < 9; c++) {
                accA[R_LALL - R_NPHASEB + c] = L[c] * w;
                accA[(gas ? R_LGAS : R_LSOL) - R_NPHASEB + c] = L[c] * w;
            }
            // dL_iter and mixing (LS-EVPFFT.f90:277-327)
            if (!pc.kin_ctrl) {
                double d2 = 0.0;
                for (int c = 0; c < 9; c++) d2 += (L[c] - Lo[c]) * (L[c] - Lo[c]);
                const double rel = sqrt(d2) / pc.lscale;
                accA[(gas ? R_DLGAS : R_DLSOL) - R_NPHASEB] = rel * rel * w;
                accA[(gas ? R_WGAS : R_WSOL) - R_NPHASEB] = w;
            }
            if (pc.do_mix) {
                for (int c = 0; c < 9; c++) L[c] = pc.mix * L[c] + (1.0 - pc.mix) * Lo[c];
            } else {
                for (int c = 0; c < 9; c++) L[c] = pc.mix * L[c];
            }
            for (int c = 0; c < 9; c++) {
                x.st(f.L, c, L[c]);
                Lmix[c] = L[c];
            }
        }
        if (!MODAL) {
#pragma unroll
            for (int i = 0; i < 21; i++) {
                pF[i] = x.ld(f.sop, i);
                pS[i] = x.ld(f.sop, 21 + i);
            }
        }
        if (MODAL) {
            block_reduce_store_cm<R_NSOLVE - R_NPHASEB, SBS>(accA, f.part, R_NPHASEB, blk, nblk);
        } else {  // scratch aliased onto the solve's LDS columns, free until phase B
            block_reduce_lds_cm<R_NSOLVE - R_NPHASEB, SBS>(accA, f.part, R_NPHASEB, pool, blk, nblk);
            __syncthreads();
        }
    }
    // ---- phase B: solve_res (various_functions.f90:1142-1364)
    double acc[R_NPHASEB];
#pragma unroll
    for (int i = 0; i < R_NPHASEB; i++) acc[i] = 0.0;
    if (act) {
        const double tdot = pc.tdot;
        double sgf[6];
        if (!gas && !MODAL) {
            // one pass per distinct phase among the wave's solid lanes (one for a
            // single-phase RVE); the phase id is then uniform inside the solve
            for (;;) {
                const int p0 = __builtin_amdgcn_readfirstlane(ph);
                if (ph == p0) {
                    solve_solid_sym<INTPOW, SYMTAU, SMALLNS>(f, pt, pc, x, p0, cm_lds + threadIdx.x, sa_lds + threadIdx.x, sgf,
                                                    acc, w, Lmix, pA, plam, psgt, pF, pS);
                    break;
                }
            }
        } else if (!gas) {
            double lam[6], sgt[6], D[6];
            {
                double L[9];
                for (int c = 0; c < 9; c++) L[c] = x.ld(f.L, c);
                to6(L, D);
                if (pc.thermo)  // eigenstrain rate of the first increment (ithermo = 1)
                    for (int c = 0; c < 6; c++) D[c] -= x.ld(f.eth, c);
            }
            for (int c = 0; c < 6; c++) {
                lam[c] = x.ld(f.sg, c);
                sgt[c] = x.ld(f.sgt, c);
            }
            const PhaseDev& P = pt.ph[f.phase[v]];
            const double* tab = pt.data;
            double cm[36], Jb[36], r0[6];
            for (int i = 0; i < 36; i++) cm[i] = x.ld(f.c66m, (long)i);
            {
                // Jb = I + cm s66 / tdot [- cGm];  r0 = -lambda - (cm s66/tdot) sgt - cm D [+ cGm lambda]
                double s66[36];
                for (int i = 0; i < 36; i++) s66[i] = x.ld(f.sg66, (long)i) / tdot;
                mm6(cm, s66, Jb);
                for (int i = 0; i < 6; i++) {
                    double s = -lam[i];
                    for (int k = 0; k < 6; k++) s -= Jb[6 * i + k] * sgt[k] + cm[6 * i + k] * D[k];
                    r0[i] = s;
                }
                for (int i = 0; i < 6; i++) Jb[7 * i] += 1.0;
                if (MODAL) {
                    for (int i = 0; i < 6; i++) {
                        double s = 0.0;
                        for (int k = 0; k < 6; k++) {
                            const double g = x.ld(f.cgm, (long)(6 * i + k));
                            Jb[6 * i + k] -= g;
                            s += g * lam[k];
                        }
                        r0[i] += s;
                    }
                }
            }
            double sg[6], sgold[6] = {0, 0, 0, 0, 0, 0}, sev[6], ep[5], dd[15];
            for (int c = 0; c < 6; c++) sg[c] = lam[c];
            double sgnorm = 0.0, taun = 0.0;
            for (int c = 0; c < 6; c++) sgnorm += sg[c] * sg[c];
            sgnorm = sqrt(sgnorm);
            for (int k = 0; k < 2 * P.nsyst; k++) {
                const double tv = x.ld(f.ttau, (long)k);
                taun += tv * tv;
            }
            taun = sqrt(taun);
            const double tol = fmax(pc.error, pc.erroral);
            int itertau = 0;
            double errtau = 2.0 * tol;
            double dsgn = 0.0;
            long nnr = 0, nev = 0, nout = 0;
            while (itertau < 10 && errtau > tol) {
                itertau++;
                nout++;
                int iternr = 0;
                double erral = 2.0 * pc.erroral, resoldn = 1.0;
                double res[6];
                while (iternr < 100 && erral > pc.erroral) {
                    iternr++;
                    nnr++;
                    int icut = 0;
                    double resn = 2.0 * resoldn;
                    while (resn > resoldn && icut < 10) {
                        nev++;
                        for (int c = 0; c < 6; c++) sev[c] = sg[c];
                        slip_eval<INTPOW, false>(f, tab, P, x, sg, ep, dd);
                        double rn = 0.0;
#pragma unroll
                        for (int i = 0; i < 6; i++) {
                            double s = r0[i];
#pragma unroll
                            for (int k = 0; k < 6; k++) s += Jb[6 * i + k] * sg[k];
#pragma unroll
                            for (int k = 0; k < 5; k++) s += cm[6 * i + k] * ep[k];
                            res[i] = s;
                            rn += s * s;
                        }
                        resn = sqrt(rn);
                        if (iternr == 1) resoldn = resn;
                        if (resn > resoldn)
                            for (int c = 0; c < 6; c++) sg[c] = sgold[c] + (sg[c] - sgold[c]) * 0.5;
                        icut++;
                    }
                    resoldn = resn;
                    for (int c = 0; c < 6; c++) sgold[c] = sg[c];
                    double Jm[36];
#pragma unroll
                    for (int i = 0; i < 6; i++)
#pragma unroll
                        for (int j = 0; j < 6; j++) {
                            double s = Jb[6 * i + j];
                            if (j < 5)
#pragma unroll
                                for (int k = 0; k < 5; k++) s += cm[6 * i + k] * dd[sym5(k, j)];
                            Jm[6 * i + j] = s;
                        }
                    if (!solve6(Jm, res)) flag_singular(f, v);  // lu_eqsystem (:1287)
                    double d1 = 0.0;
                    for (int c = 0; c < 6; c++) {
                        sg[c] -= res[c];
                        d1 += (sg[c] - sgold[c]) * (sg[c] - sgold[c]);
                    }
                    erral = sqrt(d1) / sgnorm;
                }
                // harden with the slip rates of the last evaluation
                slip_rates<INTPOW>(f, tab, P, x, sev);
                if (pc.iuphard == 1) {
                    double g1;
                    const double e2 = harden_voxel(f, tab, P, x, tdot, &g1, false);
                    errtau = sqrt(e2) / taun;
                } else {
                    errtau = 0.0;
                }
            }
            // errors: dsgnorm = |sg - lambda|^2, ddgnorm = |edot(last eval) - D|^2
            double s66sev[6], dd2 = 0.0;
            for (int c = 0; c < 6; c++) {
                dsgn += (sg[c] - lam[c]) * (sg[c] - lam[c]);
            }
            for (int i = 0; i < 6; i++) {
                double s = 0.0;
                for (int k = 0; k < 6; k++) s += x.ld(f.sg66, (long)(6 * i + k)) * ((sev[k] - sgt[k]) / tdot);
                s66sev[i] = s;
            }
            for (int i = 0; i < 6; i++) {
                const double e = (i < 5 ? ep[i] : 0.0) + s66sev[i] - D[i];
                dd2 += e * e;
            }
            acc[R_ERRS] = dsgn * w;
            acc[R_ERRE] = dd2 * w;
            for (int c = 0; c < 6; c++) sgf[c] = sg[c];
            for (int c = 0; c < 5; c++) x.st(f.edotp, c, ep[c]);
            x.st(f.edotp, 5, 0.0);
            acc[R_NR] = (double)nnr;
            acc[R_EVAL] = (double)nev;
            acc[R_OUTER] = (double)nout;
        } else {
            // gas voxel (various_functions.f90:1326-1345)
            double lam[6], sgt[6], D[6];
            {
                double L[9];
                for (int c = 0; c < 9; c++) L[c] = x.ld(f.L, c);
                to6(L, D);
            }
            for (int c = 0; c < 6; c++) {
                lam[c] = x.ld(f.sg, c);
                sgt[c] = x.ld(f.sgt, c);
            }
            double rr[6], tt[6];
            for (int i = 0; i < 6; i++) tt[i] = D[i] + sgt[i] / tdot * pc.complgas;
            for (int i = 0; i < 6; i++) {
                double s = lam[i];
                for (int k = 0; k < 6; k++) s += x.ld(f.c66m, (long)(6 * i + k)) * tt[k];
                if (MODAL)
                    for (int k = 0; k < 6; k++) s -= x.ld(f.cgm, (long)(6 * i + k)) * lam[k];
                rr[i] = s;
            }
            for (int i = 0; i < 6; i++) {
                double s = 0.0;
                for (int k = 0; k < 6; k++) s += x.ld(f.cgas, (long)(6 * i + k)) * rr[k];
                sgf[i] = s;
            }
            for (int c = 0; c < 6; c++) x.st(f.edotp, c, 0.0);
            for (int is = 0; is < f.ns; is++) x.st(f.gdot, is, 0.0);
        }
        for (int c = 0; c < 6; c++) x.st(f.sg, c, sgf[c]);
        double s3[9];
        from6(sgf, s3);
        for (int c = 0; c < 9; c++) {
            acc[R_SGALL + c] = s3[c] * w;
            if (!gas) acc[R_SGSOL + c] = s3[c] * w;
        }
    }
    __syncthreads();
    if (MODAL) block_reduce_store_cm<R_NPHASEB, SBS>(acc, f.part, 0, blk, nblk);
    else block_reduce_lds_cm<R_NPHASEB, SBS>(acc, f.part, 0, pool, blk, nblk);  // every solve has finished
}

template <bool MODAL, bool INTPOW, bool SYMTAU, bool SMALLNS>
__global__ __launch_bounds__(SBS) __attribute__((amdgpu_waves_per_eu(MODAL ? 1 : 2))) void k_correct_solve(Fields f, PhaseTables pt, SolveArgs a) {
    __shared__ double pool[MODAL ? 1 : 39 * SBS];
    correct_solve_block<MODAL, INTPOW, SYMTAU, SMALLNS>(f, pt, a, solve_block(blockIdx.x, gridDim.x), gridDim.x, pool);
}

// The same kernel over several RVEs in one launch: block b works for RVE blk_rve[b] (a
// per-RVE work table: its fields, phase tables, pass constants and first block), so small
// RVEs together fill the GPU.  Each RVE's blocks, sums and results are those of its own
// launch of k_correct_solve.
template <bool MODAL, bool INTPOW, bool SYMTAU, bool SMALLNS>
__global__ __launch_bounds__(SBS) __attribute__((amdgpu_waves_per_eu(MODAL ? 1 : 2))) void k_correct_solve_batch(
    const SolveWork* __restrict__ work, const int* __restrict__ blk_rve) {
    __shared__ double pool[MODAL ? 1 : 39 * SBS];
    const int r = __builtin_amdgcn_readfirstlane(blk_rve[blockIdx.x]);
    const SolveWork* w = work + r;
    correct_solve_block<MODAL, INTPOW, SYMTAU, SMALLNS>(w->f, w->pt, w->a, (long)blockIdx.x - w->blk0, w->nblk, pool);
}


// Public launcher with the full argument set.
void launch_correct_solve_full(const Fields& f, const PhaseTables& pt, const PassConst& pc,
                               const double* dvelgradavg, const double* finv_inv, const double* fluct_dev,
                               hipStream_t s, int* nb, int int_pow, int sym_tau, int small_ns, hipEvent_t e0,
                               hipEvent_t e1) {
    SolveArgs a;
    a.pc = pc;
    for (int i = 0; i < 9; i++) { a.dvelgradavg[i] = dvelgradavg[i]; a.finv_inv[i] = finv_inv[i]; }
    a.fluct = fluct_dev;
    *nb = (int)((f.V + SBS - 1) / SBS);
    if (pc.modal) {
        if (int_pow) hipExtLaunchKernelGGL((k_correct_solve<true, true, false, false>), dim3(*nb), dim3(SBS), 0, s, e0, e1, 0, f, pt, a);
        else hipExtLaunchKernelGGL((k_correct_solve<true, false, false, false>), dim3(*nb), dim3(SBS), 0, s, e0, e1, 0, f, pt, a);
        return;
    }
    // the symmetric-solve kernel is instantiated per (integer exponents, equal forward and
    // backward strengths, at most 12 systems per phase); e0, e1 (or null): events recorded at the
    // launch's start and end (no separate marker packets in the stream)
#define CG_SOLVE(IP, ST, SM) \
    hipExtLaunchKernelGGL((k_correct_solve<false, IP, ST, SM>), dim3(*nb), dim3(SBS), 0, s, e0, e1, 0, f, pt, a)
    switch ((int_pow ? 4 : 0) + (sym_tau ? 2 : 0) + (small_ns ? 1 : 0)) {
        case 0: CG_SOLVE(false, false, false); break;
        case 1: CG_SOLVE(false, false, true); break;
        case 2: CG_SOLVE(false, true, false); break;
        case 3: CG_SOLVE(false, true, true); break;
        case 4: CG_SOLVE(true, false, false); break;
        case 5: CG_SOLVE(true, false, true); break;
        case 6: CG_SOLVE(true, true, false); break;
        default: CG_SOLVE(true, true, true); break;
    }
#undef CG_SOLVE
}
int solve_blocks(long V) { return (int)((V + SBS - 1) / SBS); }
void launch_correct_solve_batch(const SolveWork* work, const int* blk_rve, int nblocks, int modal, int int_pow,
                                int sym_tau, int small_ns, hipStream_t s) {
    if (modal) {
        if (int_pow) hipLaunchKernelGGL((k_correct_solve_batch<true, true, false, false>), dim3(nblocks), dim3(SBS), 0, s, work, blk_rve);
        else hipLaunchKernelGGL((k_correct_solve_batch<true, false, false, false>), dim3(nblocks), dim3(SBS), 0, s, work, blk_rve);
        return;
    }
#define CG_SOLVE(IP, ST, SM) \
    hipLaunchKernelGGL((k_correct_solve_batch<false, IP, ST, SM>), dim3(nblocks), dim3(SBS), 0, s, work, blk_rve)
    switch ((int_pow ? 4 : 0) + (sym_tau ? 2 : 0) + (small_ns ? 1 : 0)) {
        case 0: CG_SOLVE(false, false, false); break;
        case 1: CG_SOLVE(false, false, true); break;
        case 2: CG_SOLVE(false, true, false); break;
        case 3: CG_SOLVE(false, true, true); break;
        case 4: CG_SOLVE(true, false, false); break;
        case 5: CG_SOLVE(true, false, true); break;
        case 6: CG_SOLVE(true, true, false); break;
        default: CG_SOLVE(true, true, true); break;
    }
#undef CG_SOLVE
}

// ---------------------------------------------------------------------------
// update_schmid (various_functions.f90:472-539): sch = b5(sym(A (b x n) A^-1)), A = Fe ag
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BS) void k_update_schmid(Fields f, PhaseTables pt, int write_sch) {
    const long V = f.V;
    const long v = (long)blockIdx.x * BS + threadIdx.x;
    if (v >= V || f.gas[v]) return;
    const PhaseDev& P = pt.ph[f.phase[v]];
    const double* dn = pt.data + P.off_dnca;
    const double* db = pt.data + P.off_dbca;
    double Fe[9], ag[9], A[9], Ai[9];
    for (int c = 0; c < 9; c++) { Fe[c] = f.Fe[c * f.ld + v]; ag[c] = f.ag[c * f.ld + v]; }
    mm3(Fe, ag, A);
    if (inv3(A, Ai) == 0.0) flag_singular(f, v);  // lu_inverse of A (:518-520)
    for (int c = 0; c < 9; c++) {
        f.schA[c * f.ld + v] = A[c];
        f.schA[(9 + c) * f.ld + v] = Ai[c];
    }
    if (!write_sch) return;
    for (int is = 0; is < P.nsyst; is++) {
        double s5[5];
        schmid5(A, Ai, db + 3 * is, dn + 3 * is, s5);
        for (int c = 0; c < 5; c++) f.sch[(5 * is + c) * f.ld + v] = s5[c];
    }
}
void launch_update_schmid(const Fields& f, const PhaseTables& pt, int write_sch, hipStream_t s) {
    hipLaunchKernelGGL(k_update_schmid, dim3(nblk(f.V)), dim3(BS), 0, s, f, pt, write_sch);
}
// sch from the A, A^-1 that update_schmid stored (what it would have written), for a host
// read of sch in a context whose solve does not keep it
__global__ __launch_bounds__(BS) void k_sch_from_A(Fields f, PhaseTables pt) {
    const long v = (long)blockIdx.x * BS + threadIdx.x;
    if (v >= f.V || f.gas[v]) return;
    const PhaseDev& P = pt.ph[f.phase[v]];
    double A[9], Ai[9];
    for (int c = 0; c < 9; c++) {
        A[c] = f.schA[c * f.ld + v];
        Ai[c] = f.schA[(9 + c) * f.ld + v];
    }
    for (int is = 0; is < P.nsyst; is++) {
        double s5[5];
        schmid5(A, Ai, pt.data + P.off_dbca + 3 * is, pt.data + P.off_dnca + 3 * is, s5);
        for (int c = 0; c < 5; c++) f.sch[(5 * is + c) * f.ld + v] = s5[c];
    }
}
void launch_sch_from_A(const Fields& f, const PhaseTables& pt, hipStream_t s) {
    hipLaunchKernelGGL(k_sch_from_A, dim3(nblk(f.V)), dim3(BS), 0, s, f, pt);
}

// ---------------------------------------------------------------------------
// calc_c066mod_Goperr066mod (various_functions.f90:812-906)
//   c066mod = M(F) c066 M(F)^T / J,  M(X)_{ca} = b_c : (b_a X^T)
//   cGm     = c066mod * [J Y_c : G0 : Y_d],  Y_c = b_c Fi^T
// ---------------------------------------------------------------------------
struct C066Args {
    double c066[36];
    double g0[81];
};
__global__ __launch_bounds__(BS) void k_c066mod(Fields f, C066Args a, int modal) {
    const long V = f.V;
    const long v = (long)blockIdx.x * BS + threadIdx.x;
    if (v >= V) return;
    double F[9], M[36], c[36];
    for (int i = 0; i < 9; i++) F[i] = f.F[i * f.ld + v];
    const double J = f.J[v];
    push_matrix(F, M);
    congr6(M, a.c066, c);
    for (int i = 0; i < 36; i++) {
        c[i] /= J;
        f.c66m[(long)i * f.ld + v] = c[i];
    }
    if (modal) {
        double Fi[9], Y[6][9], g66[36];
        for (int i = 0; i < 9; i++) Fi[i] = f.Fi[i * f.ld + v];
        for (int cc = 0; cc < 6; cc++) {
            double b[9];
            basis(cc, b);
            mmt3(b, Fi, Y[cc]);
        }
        for (int cc = 0; cc < 6; cc++) {
            double z[9];
            for (int kn = 0; kn < 9; kn++) {  // (G0 : Y_d) with index pair (i,m) free
                double s = 0.0;
                for (int im = 0; im < 9; im++) s += Y[cc][im] * a.g0[9 * im + kn];
                z[kn] = s;
            }
            for (int d = 0; d < 6; d++) {
                double s = 0.0;
                for (int kn = 0; kn < 9; kn++) s += z[kn] * Y[d][kn];
                g66[6 * cc + d] = s * J;
            }
        }
        double r[36];
        mm6(c, g66, r);
        for (int i = 0; i < 36; i++) f.cgm[(long)i * f.ld + v] = r[i];
    }
}
void launch_c066mod(const Fields& f, const double* c066, const double* goperr0, int modal, hipStream_t s) {
    C066Args a;
    for (int i = 0; i < 36; i++) a.c066[i] = c066[i];
    for (int i = 0; i < 81; i++) a.g0[i] = goperr0 ? goperr0[i] : 0.0;
    hipLaunchKernelGGL(k_c066mod, dim3(nblk(f.V)), dim3(BS), 0, s, f, a, modal);
}

// calc_gas_compl (various_functions.f90:909-951)
__global__ __launch_bounds__(BS) void k_gas_compl(Fields f, double tdot, double complgas, int modal) {
    const long V = f.V;
    const long v = (long)blockIdx.x * BS + threadIdx.x;
    if (v >= V) return;
    double a[36];
    if (!f.gas[v]) {
        for (int i = 0; i < 36; i++) f.cgas[(long)i * f.ld + v] = 0.0;
        return;
    }
    for (int i = 0; i < 36; i++) {
        a[i] = (i % 7 == 0 ? 1.0 : 0.0) + f.c66m[(long)i * f.ld + v] / tdot * complgas;
        if (modal) a[i] -= f.cgm[(long)i * f.ld + v];
    }
    if (!inv6(a)) flag_singular(f, v);
    for (int i = 0; i < 36; i++) f.cgas[(long)i * f.ld + v] = a[i];
}
void launch_gas_compl(const Fields& f, double tdot, double complgas, int modal, hipStream_t s) {
    hipLaunchKernelGGL(k_gas_compl, dim3(nblk(f.V)), dim3(BS), 0, s, f, tdot, complgas, modal);
}

// ---------------------------------------------------------------------------
// Goperr0 = (1/V) sum_xi Gamma(xi)  (the origin value of the inverse DFT of Gamma,
// which form_G obtains through an 81-component inverse FFT, :784-807).
// One thread per frequency of the full spectrum; 81 block partial sums.
// ---------------------------------------------------------------------------
struct GArgs { double c0[81]; double s0[81]; int igamma; };
__device__ void green_full(const GArgs& a, int kx, int ky, int kz, int n1, int n2, int n3, double* g) {
    // 0-based frequency indices; spatial_freq (fourier_functions.f90:340-371)
    const double tp = 2.0 * PI_D;
    double xi[3];
    // spatial_freq: 1-based index x = kx+1; x <= n/2 -> kx, else kx - n
    xi[0] = tp * (double)((kx + 1) <= n1 / 2 ? kx : kx - n1) / n1;
    xi[1] = tp * (double)((ky + 1) <= n2 / 2 ? ky : ky - n2) / n2;
    xi[2] = (n3 == 1) ? 0.0 : tp * (double)((kz + 1) <= n3 / 2 ? kz : kz - n3) / n3;
    const bool nx = (n1 % 2 == 0) && kx == n1 / 2, ny = (n2 % 2 == 0) && ky == n2 / 2,
               nz = (n3 % 2 == 0) && kz == n3 / 2;
    if (a.igamma == 0 && (nx || ny || (nz && n3 > 1))) {
        for (int i = 0; i < 81; i++) g[i] = -a.s0[i];
        return;
    }
    if (a.igamma >= 2 && ((nx && ny) || (nx && nz) || (ny && nz))) {
        for (int i = 0; i < 81; i++) g[i] = 0.0;
        return;
    }
    double D[9];
    if (a.igamma == 0) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) D[3 * i + j] = xi[i] * xi[j];
    } else if (a.igamma == 1) {
        const double xs = tp * kx / n1, ys = tp * ky / n2, zs = tp * kz / n3;
        D[0] = 2.0 * (cos(xs) - 1.0); D[4] = 2.0 * (cos(ys) - 1.0); D[8] = 2.0 * (cos(zs) - 1.0);
        D[1] = D[3] = -sin(xs) * sin(ys);
        D[2] = D[6] = -sin(xs) * sin(zs);
        D[5] = D[7] = -sin(ys) * sin(zs);
    } else {
        // kmod = i/4 prod_l (1 + e^{i xi_l}) * tan(xi/2);  D = Re(kmod kmod^H)
        double pr = 1.0, pim = 0.0;
        for (int l = 0; l < 3; l++) {
            const double cr = 1.0 + cos(xi[l]), ci = sin(xi[l]);
            const double nr = pr * cr - pim * ci, ni = pr * ci + pim * cr;
            pr = nr; pim = ni;
        }
        // f = i*0.25*(pr + i pim) = 0.25*(-pim + i pr); |f|^2 = 0.0625 (pr^2 + pim^2)
        const double f2 = 0.0625 * (pr * pr + pim * pim);
        double tn[3];
        for (int l = 0; l < 3; l++) tn[l] = tan(xi[l] * 0.5);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) D[3 * i + j] = tn[i] * tn[j] * f2;
    }
    double A[9], Ai[9];
    if (kx == 0 && ky == 0 && kz == 0) {
        for (int i = 0; i < 9; i++) Ai[i] = 0.0;
    } else {
        for (int i = 0; i < 3; i++)
            for (int k = 0; k < 3; k++) {
                double s = 0.0;
                for (int j = 0; j < 3; j++)
                    for (int l = 0; l < 3; l++) s += a.c0[27 * i + 9 * j + 3 * k + l] * D[3 * j + l];
                A[3 * i + k] = s;
            }
        inv3(A, Ai);
    }
    for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++)
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) g[27 * p + 9 * q + 3 * i + j] = -Ai[3 * p + i] * D[3 * q + j];
    if (a.igamma == 1) {  // G <- -G : C0 : G, four times (:744-780)
        double gt[81], aux[81];
        for (int i = 0; i < 81; i++) gt[i] = g[i];
        for (int it = 0; it < 4; it++) {
            for (int ij = 0; ij < 9; ij++) {
                double cg[9];
                for (int op = 0; op < 9; op++) {
                    double s = 0.0;
                    for (int mn = 0; mn < 9; mn++) s += g[9 * ij + mn] * a.c0[9 * mn + op];
                    cg[op] = s;
                }
                for (int kl = 0; kl < 9; kl++) {
                    double s = 0.0;
                    for (int op = 0; op < 9; op++) s += cg[op] * gt[9 * op + kl];
                    aux[9 * ij + kl] = -s;
                }
            }
            for (int i = 0; i < 81; i++) gt[i] = aux[i];
        }
        for (int i = 0; i < 81; i++) g[i] = gt[i];
    }
}
__global__ __launch_bounds__(BS) void k_goperr0(Fields f, GArgs a) {
    const long v = (long)blockIdx.x * BS + threadIdx.x;
    double acc[81];
    for (int i = 0; i < 81; i++) acc[i] = 0.0;
    if (v < f.V) {
        const int kx = v % f.n1, ky = (v / f.n1) % f.n2, kz = v / ((long)f.n1 * f.n2);
        green_full(a, kx, ky, kz, f.n1, f.n2, f.n3, acc);
    }
    block_reduce_store<81, BS>(acc, f.part);
}
void launch_goperr0(const Fields& f, const double* c0, const double* s0, int igamma, hipStream_t s, int* nb) {
    GArgs a;
    for (int i = 0; i < 81; i++) { a.c0[i] = c0[i]; a.s0[i] = s0[i]; }
    a.igamma = igamma;
    *nb = nblk(f.V);
    hipLaunchKernelGGL(k_goperr0, dim3(*nb), dim3(BS), 0, s, f, a);
}

// ---------------------------------------------------------------------------
// accepted-step accumulation (LS-EVPFFT.f90:431-446); max solid vm(sym L)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BS) void k_accumulate(Fields f, double tdot) {
    const long V = f.V;
    const long v = (long)blockIdx.x * BS + threadIdx.x;
    double acc[1] = {0.0};
    if (v < V) {
        double ep[6], L[9], d6[6];
        for (int c = 0; c < 6; c++) {
            ep[c] = f.edotp[c * f.ld + v];
            f.ept[c * f.ld + v] += ep[c] * tdot;
            f.sgt[c * f.ld + v] = f.sg[c * f.ld + v];
        }
        for (int c = 0; c < 9; c++) L[c] = f.L[c * f.ld + v];
        to6(L, d6);
        for (int c = 0; c < 6; c++) f.dgs[c * f.ld + v] += d6[c] * tdot;
        f.epvm[v] += vm_strain6(ep) * tdot;
        if (!f.gas[v]) acc[0] = vm_strain6(d6);
    }
    block_reduce_store<1, BS, OpMax>(acc, f.part);
}
void launch_accumulate(const Fields& f, double tdot, double* max_out, hipStream_t s, int* nb) {
    *nb = nblk(f.V);
    hipLaunchKernelGGL(k_accumulate, dim3(*nb), dim3(BS), 0, s, f, tdot);
    hipLaunchKernelGGL(k_reduce_final_op<OpMax>, dim3(1), dim3(BS), 0, s, f.part, *nb, 1, max_out);
}
// the 13 values of voxel v that eelavg takes from the last solid voxel: disgradsym (6), ept (6), w
__global__ void k_pick_last(Fields f, long v, double* __restrict__ out) {
    const int k = threadIdx.x;
    if (k < 6) out[k] = f.dgs[k * f.ld + v];
    else if (k < 12) out[k] = f.ept[(k - 6) * f.ld + v];
    else if (k == 12) out[k] = f.w[v];
}
void launch_pick_last(const Fields& f, long v, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_pick_last, dim3(1), dim3(64), 0, s, f, v, out);
}

// update_plastic_defgrad (various_functions.f90:1748-1827)
__global__ __launch_bounds__(BS) void k_plastic_defgrad(Fields f, PhaseTables pt, double tdot) {
    const long V = f.V;
    const long v = (long)blockIdx.x * BS + threadIdx.x;
    if (v >= V || f.gas[v]) return;
    const PhaseDev& P = pt.ph[f.phase[v]];
    const double* dn = pt.data + P.off_dnca;
    const double* db = pt.data + P.off_dbca;
    double ag[9], Lp[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = 0; c < 9; c++) ag[c] = f.ag[c * f.ld + v];
    for (int is = 0; is < P.nsyst; is++) {
        double n[3], b[3];
        for (int i = 0; i < 3; i++) {
            n[i] = ag[3 * i] * dn[3 * is] + ag[3 * i + 1] * dn[3 * is + 1] + ag[3 * i + 2] * dn[3 * is + 2];
            b[i] = ag[3 * i] * db[3 * is] + ag[3 * i + 1] * db[3 * is + 1] + ag[3 * i + 2] * db[3 * is + 2];
        }
        const double g = f.gdot[is * f.ld + v];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) Lp[3 * i + j] += b[i] * n[j] * g;
    }
    double d[9], E[9], Fp[9];
    for (int c = 0; c < 9; c++) { d[c] = Lp[c] * tdot; Fp[c] = f.Fp[c * f.ld + v]; }
    expm3(d, E);
    mm3(E, Fp, Fp);
    for (int c = 0; c < 9; c++) f.Fp[c * f.ld + v] = Fp[c];
}
void launch_plastic_defgrad(const Fields& f, const PhaseTables& pt, double tdot, hipStream_t s) {
    hipLaunchKernelGGL(k_plastic_defgrad, dim3(nblk(f.V)), dim3(BS), 0, s, f, pt, tdot);
}

// update_defgrad (various_functions.f90:2152-2282): 39 sums
//  0..8 sum w F, 9..17 sum wJ Fi, 18..26 solid, 27..35 gas, 36 sum_solid wJ, 37 sum_gas wJ, 38 sum wJ
__global__ __launch_bounds__(BS) void k_update_defgrad(Fields f, double tdot) {
    const long V = f.V;
    const long v = (long)blockIdx.x * BS + threadIdx.x;
    double acc[39];
    for (int i = 0; i < 39; i++) acc[i] = 0.0;
    if (v < V) {
        double L[9], F[9], d[9], E[9], Fi[9];
        for (int c = 0; c < 9; c++) { L[c] = f.L[c * f.ld + v]; F[c] = f.F[c * f.ld + v]; d[c] = L[c] * tdot; }
        expm3(d, E);
        mm3(E, F, F);
        // determinant33 (tensor_functions.f90:442-456) term order
        const double J = F[0] * F[4] * F[8] + F[1] * F[5] * F[6] + F[2] * F[3] * F[7] - F[6] * F[4] * F[2] -
                         F[7] * F[5] * F[0] - F[8] * F[3] * F[1];
        if (inv3(F, Fi) == 0.0) flag_singular(f, v);
        const double wg = 1.0 / (double)V;
        for (int c = 0; c < 9; c++) {
            f.F[c * f.ld + v] = F[c];
            f.Fi[c * f.ld + v] = Fi[c];
            f.Finc[c * f.ld + v] = E[c];
            acc[c] = F[c] * wg;
            acc[9 + c] = Fi[c] * wg * J;
            acc[(f.gas[v] ? 27 : 18) + c] = Fi[c] * wg * J;
        }
        f.J[v] = J;
        f.w[v] = wg * J;
        acc[f.gas[v] ? 37 : 36] = wg * J;
        acc[38] = J * wg;
    }
    block_reduce_store<39, BS>(acc, f.part);
}
void launch_update_defgrad(const Fields& f, double tdot, hipStream_t s, int* nb) {
    *nb = nblk(f.V);
    hipLaunchKernelGGL(k_update_defgrad, dim3(*nb), dim3(BS), 0, s, f, tdot);
}
__global__ __launch_bounds__(BS) void k_div_w(Fields f, double detfavg) {
    const long v = (long)blockIdx.x * BS + threadIdx.x;
    if (v < f.V) f.w[v] = f.w[v] / detfavg;
}
void launch_div_w(const Fields& f, double detfavg, hipStream_t s) {
    hipLaunchKernelGGL(k_div_w, dim3(nblk(f.V)), dim3(BS), 0, s, f, detfavg);
}

// update_elstic_defgrad (:2285-2324) + update_el_stiff (:2327-2402)
//   Fe = F (Fp F_ini)^-1 (F_ini = I without an initial displacement field), Fe_inc = Fe Fe_old^-1,
//   cg66 = T(A) C66cr T(A)^T / det Fe, sg66 = cg66^-1, A = Fe ag
__global__ __launch_bounds__(BS) void k_elastic_stiff(Fields f, PhaseTables pt) {
    const long V = f.V;
    const long v = (long)blockIdx.x * BS + threadIdx.x;
    if (v >= V || f.gas[v]) return;
    const PhaseDev& P = pt.ph[f.phase[v]];
    double F[9], Fp[9], Fpi[9], Feo[9], Feoi[9], Fe[9], Fei[9];
    for (int c = 0; c < 9; c++) { F[c] = f.F[c * f.ld + v]; Fp[c] = f.Fp[c * f.ld + v]; Feo[c] = f.Fe[c * f.ld + v]; }
    bool ok = inv3(Feo, Feoi) != 0.0;
    if (f.Fini) {  // (Fp F0)^-1 with the initial deformation gradient of init_defgrad
        double FpF0[9], F0[9];
        for (int c = 0; c < 9; c++) F0[c] = f.Fini[c * f.ld + v];
        mm3(Fp, F0, FpF0);
        ok = (inv3(FpF0, Fpi) != 0.0) && ok;
    } else {
        ok = (inv3(Fp, Fpi) != 0.0) && ok;
    }
    mm3(F, Fpi, Fe);
    mm3(Fe, Feoi, Fei);
    for (int c = 0; c < 9; c++) { f.Fe[c * f.ld + v] = Fe[c]; f.Feinc[c * f.ld + v] = Fei[c]; }
    double ag[9], A[9], T[36], c66[36];
    for (int c = 0; c < 9; c++) ag[c] = f.ag[c * f.ld + v];
    mm3(Fe, ag, A);
    rot_matrix(A, T);
    congr6(T, P.c66cr, c66);
    const double dFe = det3(Fe);
    for (int i = 0; i < 36; i++) c66[i] /= dFe;
    ok = inv6(c66) && ok;
    if (!ok) flag_singular(f, v);
    for (int i = 0; i < 36; i++) f.sg66[(long)i * f.ld + v] = c66[i];
}
void launch_elastic_stiff(const Fields& f, const PhaseTables& pt, hipStream_t s) {
    hipLaunchKernelGGL(k_elastic_stiff, dim3(nblk(f.V)), dim3(BS), 0, s, f, pt);
}

// update_tensor_config (:2405-2492), part a: rotations and L Finc^-1, 9 sums of w L
__global__ __launch_bounds__(BS) void k_tensor_config_a(Fields f) {
    const long V = f.V;
    const long v = (long)blockIdx.x * BS + threadIdx.x;
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (v < V) {
        double Fn[9], Fni[9], L[9];
        for (int c = 0; c < 9; c++) { Fn[c] = f.Finc[c * f.ld + v]; L[c] = f.L[c * f.ld + v]; }
        if (inv3(Fn, Fni) == 0.0) flag_singular(f, v);
        if (!f.gas[v]) {
            double R[9], lam[3], Fe[9];
            polar3(Fn, R, lam);
            for (int c = 0; c < 9; c++) Fe[c] = f.Feinc[c * f.ld + v];
            const double dfe = det3(Fe);
            double s6[6], s[9], t[9];
            // disgradsym, ept <- R (.) R^T
            for (int c = 0; c < 6; c++) s6[c] = f.dgs[c * f.ld + v];
            from6(s6, s); mm3(R, s, t); mmt3(t, R, s); to6(s, s6);
            for (int c = 0; c < 6; c++) f.dgs[c * f.ld + v] = s6[c];
            for (int c = 0; c < 6; c++) s6[c] = f.ept[c * f.ld + v];
            from6(s6, s); mm3(R, s, t); mmt3(t, R, s); to6(s, s6);
            for (int c = 0; c < 6; c++) f.ept[c * f.ld + v] = s6[c];
            // sgt <- Fe_inc sgt Fe_inc^T / det(Fe_inc)
            for (int c = 0; c < 6; c++) s6[c] = f.sgt[c * f.ld + v];
            from6(s6, s); mm3(Fe, s, t); mmt3(t, Fe, s);
            for (int c = 0; c < 9; c++) s[c] /= dfe;
            to6(s, s6);
            for (int c = 0; c < 6; c++) f.sgt[c * f.ld + v] = s6[c];
        }
        mm3(L, Fni, L);
        const double w = f.w[v];
        for (int c = 0; c < 9; c++) {
            f.L[c * f.ld + v] = L[c];
            acc[c] = L[c] * w;
        }
    }
    block_reduce_store<9, BS>(acc, f.part);
}
void launch_tensor_config_a(const Fields& f, hipStream_t s, int* nb) {
    *nb = nblk(f.V);
    hipLaunchKernelGGL(k_tensor_config_a, dim3(*nb), dim3(BS), 0, s, f);
}
struct Mat9 { double m[9]; };
__global__ __launch_bounds__(BS) void k_tensor_config_b(Fields f, Mat9 dud) {
    const long V = f.V;
    const long v = (long)blockIdx.x * BS + threadIdx.x;
    if (v >= V) return;
    double Fi[9], q[9];
    for (int c = 0; c < 9; c++) Fi[c] = f.Fi[c * f.ld + v];
    mm3(dud.m, Fi, q);
    for (int c = 0; c < 9; c++) f.L[c * f.ld + v] += q[c];
}
void launch_tensor_config_b(const Fields& f, const double* dud, hipStream_t s) {
    Mat9 m;
    for (int i = 0; i < 9; i++) m.m[i] = dud[i];
    hipLaunchKernelGGL(k_tensor_config_b, dim3(nblk(f.V)), dim3(BS), 0, s, f, m);
}

// update_hardening (:1493-1523)
// SMALLNS (one mode of at most 12 systems in every phase): the branch-free harden of the solve
// in its FINAL form; else the general harden_voxel.  Same sums in the same order.
template <bool SMALLNS, bool SYMTAU>
__global__ __launch_bounds__(BS) void k_hardening(Fields f, PhaseTables pt, double tdot) {
    const long V = f.V;
    const long v = (long)blockIdx.x * BS + threadIdx.x;
    if (v >= V || f.gas[v]) return;
    // one pass per distinct phase among the wave's lanes: inside it the phase is uniform and
    // its tables (modes, interaction matrix) are read with scalar loads
    const int ph = f.phase[v];
    for (;;) {
        const int p0 = __builtin_amdgcn_readfirstlane(ph);
        if (ph == p0) {
            if (SMALLNS) {
                harden_small<SYMTAU, true>(f, (cdbl*)pt.data, ((cphase*)pt.ph)[p0], Vox(v, f.ld), tdot);
            } else {
                double g1;
                harden_voxel<true>(f, (cdbl*)pt.data, ((cphase*)pt.ph)[p0], Vox(v, f.ld), tdot, &g1, true);
                f.gacc[v] = g1;
            }
            break;
        }
    }
}
// Operators of the symmetric Newton solve for one increment (constant while c066mod,
// sg66 and tdot are): f.sop = [L D L^T factors of cm^-1 (21: unit L below the diagonal,
// 1/d on it) | S = sym(s66)/tdot (21, packed)]; cm^-1 itself from the L D L^T of cm.
__global__ __launch_bounds__(BS) void k_solve_ops(Fields f, double tdot) {
    const long V = f.V;
    const long v = (long)blockIdx.x * BS + threadIdx.x;
    if (v >= V || f.gas[v]) return;
    const Vox x(v, f.ld);
    double Lf[21], Ci[21];
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = i; j < 6; j++)
            Lf[sp6(i, j)] = i == j ? x.ld(f.c66m, 7 * i) : 0.5 * (x.ld(f.c66m, 6 * i + j) + x.ld(f.c66m, 6 * j + i));
    ldl6(Lf);
#pragma unroll
    for (int j = 0; j < 6; j++) {  // columns of cm^-1 (upper part kept)
        double e[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        e[j] = 1.0;
        ldl6_solve(Lf, e);
#pragma unroll
        for (int i = 0; i <= j; i++) Ci[sp6(i, j)] = e[i];
    }
    ldl6(Ci);  // cm^-1 = L D L^T
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = 0; j <= i; j++) x.st(f.sop, sp6(i, j), i == j ? 1.0 / Ci[sp6(i, i)] : Ci[sp6(i, j)]);
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = i; j < 6; j++)
            x.st(f.sop, 21 + sp6(i, j),
                 i == j ? x.ld(f.sg66, 7 * i) / tdot : 0.5 * (x.ld(f.sg66, 6 * i + j) + x.ld(f.sg66, 6 * j + i)) / tdot);
}
// |ttau| per solid voxel (columns in order, as the solve's own norm) after the trial
// strengths were written by initialisation, a rollback or the host; sym: the forward
// column stands for both
__global__ __launch_bounds__(BS) void k_taun(Fields f, PhaseTables pt, int sym) {
    const long V = f.V;
    const long v = (long)blockIdx.x * BS + threadIdx.x;
    if (v >= V || f.gas[v]) return;
    const int ns = pt.ph[f.phase[v]].nsyst;
    double t2 = 0.0;
    for (int k = 0; k < 2 * ns; k++) {
        const double t = f.ttau[(long)(sym ? k & ~1 : k) * f.ld + v];
        t2 += t * t;
    }
    f.taun[v] = sqrt(t2);
}
void launch_taun(const Fields& f, const PhaseTables& pt, int sym, hipStream_t s) {
    hipLaunchKernelGGL(k_taun, dim3(nblk(f.V)), dim3(BS), 0, s, f, pt, sym);
}
void launch_solve_ops(const Fields& f, double tdot, hipStream_t s) {
    hipLaunchKernelGGL(k_solve_ops, dim3(nblk(f.V)), dim3(BS), 0, s, f, tdot);
}

void launch_hardening(const Fields& f, const PhaseTables& pt, double tdot, int small_ns, int sym_tau, hipStream_t s) {
    const dim3 g(nblk(f.V)), b(BS);
    if (small_ns && sym_tau) hipLaunchKernelGGL((k_hardening<true, true>), g, b, 0, s, f, pt, tdot);
    else if (small_ns) hipLaunchKernelGGL((k_hardening<true, false>), g, b, 0, s, f, pt, tdot);
    else hipLaunchKernelGGL((k_hardening<false, false>), g, b, 0, s, f, pt, tdot);
}

// macro averages (LS-EVPFFT.f90:460-523): 50 sums over solid voxels
//  0..8 w ept, 9..17 w disgradsym, 18..26 (unused), 27..35 w edotp, 36..44 w sgt,
//  45 gacumgr, 46 nsolid, 47 sum g, 48 sum |gdot|, 49 sum |gdot| g, (nsys is host-side)
__global__ __launch_bounds__(BS) void k_averages(Fields f, PhaseTables pt) {
    const long V = f.V;
    const long v = (long)blockIdx.x * BS + threadIdx.x;
    double acc[50];
    for (int i = 0; i < 50; i++) acc[i] = 0.0;
    if (v < V && !f.gas[v]) {
        const double w = f.w[v];
        double s6[6], s[9];
        for (int c = 0; c < 6; c++) s6[c] = f.ept[c * f.ld + v];
        from6(s6, s);
        for (int c = 0; c < 9; c++) acc[c] = s[c] * w;
        for (int c = 0; c < 6; c++) s6[c] = f.dgs[c * f.ld + v];
        from6(s6, s);
        for (int c = 0; c < 9; c++) acc[9 + c] = s[c] * w;
        for (int c = 0; c < 6; c++) s6[c] = f.edotp[c * f.ld + v];
        from6(s6, s);
        for (int c = 0; c < 9; c++) acc[27 + c] = s[c] * w;
        for (int c = 0; c < 6; c++) s6[c] = f.sgt[c * f.ld + v];
        from6(s6, s);
        for (int c = 0; c < 9; c++) acc[36 + c] = s[c] * w;
        acc[45] = f.gacc[v];
        acc[46] = 1.0;
        const PhaseDev& P = pt.ph[f.phase[v]];
        for (int is = 0; is < P.nsyst; is++) {
            const double g = 0.5 * (f.crss[(2 * is) * f.ld + v] + f.crss[(2 * is + 1) * f.ld + v]);
            const double a = fabs(f.gdot[is * f.ld + v]);
            acc[47] += g;
            acc[48] += a;
            acc[49] += a * g;
        }
    }
    block_reduce_store<50, BS>(acc, f.part);
}
void launch_averages(const Fields& f, const PhaseTables& pt, hipStream_t s, int* nb) {
    *nb = nblk(f.V);
    hipLaunchKernelGGL(k_averages, dim3(*nb), dim3(BS), 0, s, f, pt);
}

// ---------------------------------------------------------------------------
// calc_c0 (:2537-2649): per solid voxel
//   T = (s66 tdot/tdotref / tdot + dedotp)^-1  (both tangents from the tdotref-scaled one)
//   sum w T (36) and sum wgt N T N^T J (36),  N_{ca} = b_c : (b_a Fi^T)
// ---------------------------------------------------------------------------
template <bool INTPOW>
__global__ __launch_bounds__(SBS) void k_calc_c0(Fields f, PhaseTables pt, double tdot, double tdotref) {
    __shared__ double sm[24 * SBS];  // block sums, 24 of the 72 values at a time
    const long V = f.V;
    const long v = (long)blockIdx.x * SBS + threadIdx.x;
    double acc[72];
    for (int i = 0; i < 72; i++) acc[i] = 0.0;
    if (v < V && !f.gas[v]) {
        const PhaseDev& P = pt.ph[f.phase[v]];
        double sg[6], ep[5], dd[15], T[36];
        for (int c = 0; c < 6; c++) sg[c] = f.sg[c * f.ld + v];
        // plastic_strain_rate on a local copy (gdot is not modified), the Schmid tensors
        // rebuilt from the A, A^-1 of the increment's update_schmid
        slip_eval_A<INTPOW>(f, pt.data, P, Vox(v, f.ld), sg, ep, dd);
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) {
                double e = f.sg66[(long)(6 * i + j) * f.ld + v] / tdot * tdot / tdotref;
                if (i < 5 && j < 5) e += dd[sym5(i, j)];
                T[6 * i + j] = e;
            }
        if (!inv6(T)) flag_singular(f, v);
        double Fi[9], N[36], dP[36];
        for (int c = 0; c < 9; c++) Fi[c] = f.Fi[c * f.ld + v];
        push_matrix(Fi, N);
        congr6(N, T, dP);
        const double J = f.J[v], w = f.w[v], wg = 1.0 / (double)V;
        for (int i = 0; i < 36; i++) {
            acc[i] = T[i] * w;
            acc[36 + i] = dP[i] * J * wg;
        }
    }
    for (int k = 0; k < 3; k++) {
        block_reduce_lds<24, SBS>(acc + 24 * k, f.part, 72, 24 * k, sm);
        __syncthreads();  // the last readers of sm are done before the next chunk writes it
    }
}
void launch_calc_c0(const Fields& f, const PhaseTables& pt, double tdot, double tdotref, hipStream_t s, int* nb,
                    int int_pow) {
    *nb = (int)((f.V + SBS - 1) / SBS);
    if (int_pow) hipLaunchKernelGGL(k_calc_c0<true>, dim3(*nb), dim3(SBS), 0, s, f, pt, tdot, tdotref);
    else hipLaunchKernelGGL(k_calc_c0<false>, dim3(*nb), dim3(SBS), 0, s, f, pt, tdot, tdotref);
}

// ---------------------------------------------------------------------------
// kinematic_increments (:2766-2837): max vol / rot / stretch increments, min J next
// ---------------------------------------------------------------------------
__device__ __forceinline__ void kin_voxel(const double* Finc, double* vol, double* rot, double* str) {
    const double TINY = 2.2250738585072014e-308;
    const double Jinc = Finc[0] * Finc[4] * Finc[8] + Finc[1] * Finc[5] * Finc[6] + Finc[2] * Finc[3] * Finc[7] -
                        Finc[6] * Finc[4] * Finc[2] - Finc[7] * Finc[5] * Finc[0] - Finc[8] * Finc[3] * Finc[1];
    *vol = fabs(log(fmax(Jinc, TINY)));
    double R[9], lam[3];
    polar3(Finc, R, lam);
    double arg = 0.5 * (R[0] + R[4] + R[8] - 1.0);
    arg = fmax(-1.0, fmin(1.0, arg));
    *rot = acos(arg) * 180.0 / PI_D;
    double s = 0.0;
    for (int i = 0; i < 3; i++) {
        const double l = log(fmax(lam[i], TINY));
        s += l * l;
    }
    *str = sqrt(s);
}
// L: 9 columns of stride ld; J: V values; part: 4 nblk(V) block partials; red: 5 results
// (max volinc, max rotang, max stretchinc, min detF_next, nsub)
__global__ __launch_bounds__(BS) void k_kinematic(const double* __restrict__ L, long ld, const double* __restrict__ J,
                                                  long V, double tdot, double* __restrict__ part) {
    const long v = (long)blockIdx.x * BS + threadIdx.x;
    double mx[3] = {0.0, 0.0, 0.0}, mn[1] = {1.7976931348623157e308};
    if (v < V) {
        double d[9], E[9];
        for (int c = 0; c < 9; c++) d[c] = L[c * ld + v] * tdot;
        expm3(d, E);
        const double Jinc = E[0] * E[4] * E[8] + E[1] * E[5] * E[6] + E[2] * E[3] * E[7] - E[6] * E[4] * E[2] -
                            E[7] * E[5] * E[0] - E[8] * E[3] * E[1];
        kin_voxel(E, &mx[0], &mx[1], &mx[2]);
        mn[0] = J[v] * Jinc;
    }
    const int nb = gridDim.x;
    block_reduce_store<3, BS, OpMax>(mx, part);
    __syncthreads();
    block_reduce_store<1, BS, OpMin>(mn, part + 3 * (size_t)nb);
}
// kinematic_substep (various_functions.f90:2759-2760): nsub = ceiling of the largest ratio of
// an increment to its limit (5e-2 volume, 2 degrees, 5e-2 stretch), at least 1
__global__ void k_kinematic_nsub(double* red) {
    if (threadIdx.x == 0) {
        const double worst = fmax(1.0, fmax(red[0] / 5.0e-2, fmax(red[1] / 2.0, red[2] / 5.0e-2)));
        red[4] = fmax(1.0, ceil(worst));
    }
}
void launch_kinematic_raw(const double* L, long ld, const double* J, long V, double tdot, double* part, double* red,
                          hipStream_t s) {
    const int nb = nblk(V);
    hipLaunchKernelGGL(k_kinematic, dim3(nb), dim3(BS), 0, s, L, ld, J, V, tdot, part);
    hipLaunchKernelGGL(k_reduce_final_op<OpMax>, dim3(3), dim3(BS), 0, s, part, nb, 3, red);
    hipLaunchKernelGGL(k_reduce_final_op<OpMin>, dim3(1), dim3(BS), 0, s, part + 3 * (size_t)nb, nb, 1, red + 3);
    hipLaunchKernelGGL(k_kinematic_nsub, dim3(1), dim3(64), 0, s, red);
}
void launch_kinematic(const Fields& f, double tdot, hipStream_t s, int* nb) {
    *nb = nblk(f.V);
    launch_kinematic_raw(f.L, f.ld, f.J, f.V, tdot, f.part, f.red, s);
}
long kinematic_scratch_doubles(long V) { return 4L * nblk(V); }

// ---------------------------------------------------------------------------
// update_diagnostics (types.f90:345-758)
// sums: 0 s2all 1 wall 2 s2sol 3 wsol 4 s2gas 5 wgas
// max : 0 maxall 1 maxsol 2 maxgas 3 volsol 4 volgas 5 strsol 6 strgas 7 rotsol 8 rotgas 9 condsol 10 condgasc 11 condcgas
// min : 0 detF all ; argmin (value,index) solid / gas
// ---------------------------------------------------------------------------
__device__ double cond1_est6(const double* A) {
    // Hager-Higham 1-norm estimate of ||A^-1||_1 (Higham 1988, ACM TOMS 674) times ||A||_1
    const double HUGEV = 1.7976931348623157e308;
    double anorm = 0.0;
    for (int j = 0; j < 6; j++) {
        double s = 0.0;
        for (int i = 0; i < 6; i++) s += fabs(A[6 * i + j]);
        anorm = fmax(anorm, s);
    }
    double LU[36];
    int piv[6];
    for (int i = 0; i < 36; i++) LU[i] = A[i];
    if (!lu6(LU, piv) || anorm == 0.0) return HUGEV;
    double x[6], xs[6];
    for (int i = 0; i < 6; i++) x[i] = 1.0 / 6.0;
    lu6_solve(LU, piv, x);
    double est = 0.0;
    for (int i = 0; i < 6; i++) est += fabs(x[i]);
    for (int i = 0; i < 6; i++) { xs[i] = x[i] >= 0.0 ? 1.0 : -1.0; x[i] = xs[i]; }
    lu6_solve_t(LU, piv, x);
    int j = 0;
    for (int i = 1; i < 6; i++) if (fabs(x[i]) > fabs(x[j])) j = i;
    int iter = 2;
    for (;;) {
        for (int i = 0; i < 6; i++) x[i] = 0.0;
        x[j] = 1.0;
        lu6_solve(LU, piv, x);
        const double estold = est;
        est = 0.0;
        for (int i = 0; i < 6; i++) est += fabs(x[i]);
        bool same = true;
        for (int i = 0; i < 6; i++) if ((x[i] >= 0.0 ? 1.0 : -1.0) != xs[i]) { same = false; break; }
        if (same || est <= estold) break;
        for (int i = 0; i < 6; i++) { xs[i] = x[i] >= 0.0 ? 1.0 : -1.0; x[i] = xs[i]; }
        lu6_solve_t(LU, piv, x);
        const int jlast = j;
        j = 0;
        for (int i = 1; i < 6; i++) if (fabs(x[i]) > fabs(x[j])) j = i;
        if (x[jlast] != fabs(x[j]) && iter < 5) { iter++; continue; }
        break;
    }
    double alt = 1.0;
    for (int i = 0; i < 6; i++) { x[i] = alt * (1.0 + (double)i / 5.0); alt = -alt; }
    lu6_solve(LU, piv, x);
    double temp = 0.0;
    for (int i = 0; i < 6; i++) temp += fabs(x[i]);
    temp = 2.0 * (temp / 18.0);
    if (temp > est) est = temp;
    const double rcond = (1.0 / est) / anorm;
    if (!(rcond > 0.0)) return HUGEV;
    return 1.0 / rcond;
}

__global__ __launch_bounds__(BS) void k_diagnostics(Fields f, double lfloor, int scan_cond) {
    const long V = f.V;
    const long v = (long)blockIdx.x * BS + threadIdx.x;
    double sm[6] = {0, 0, 0, 0, 0, 0};
    double mx[12];
    for (int i = 0; i < 12; i++) mx[i] = 0.0;
    double mnall[1] = {1.7976931348623157e308};
    double js = 1.7976931348623157e308, jg = 1.7976931348623157e308;
    long is_ = -1, ig_ = -1;
    if (v < V) {
        double d2 = 0.0, l0 = 0.0;
        for (int c = 0; c < 9; c++) {
            const double a = f.L[c * f.ld + v], b = f.s_L[c * f.ld + v];
            d2 += (a - b) * (a - b);
            l0 += b * b;
        }
        const double rel = sqrt(d2) / fmax(sqrt(l0), lfloor);
        const double w = f.w[v], J = f.J[v];
        const bool gas = f.gas[v] != 0;
        sm[0] = rel * rel * w; sm[1] = w;
        mx[0] = rel;
        mnall[0] = J;
        if (gas) { sm[4] = rel * rel * w; sm[5] = w; mx[2] = rel; jg = J; ig_ = v; }
        else { sm[2] = rel * rel * w; sm[3] = w; mx[1] = rel; js = J; is_ = v; }
        double Fn[9], vol, rot, str;
        for (int c = 0; c < 9; c++) Fn[c] = f.Finc[c * f.ld + v];
        kin_voxel(Fn, &vol, &rot, &str);
        mx[gas ? 4 : 3] = vol;
        mx[gas ? 6 : 5] = str;
        mx[gas ? 8 : 7] = rot;
        if (scan_cond) {
            double A[36];
            for (int i = 0; i < 36; i++) A[i] = f.c66m[(long)i * f.ld + v];
            const double cA = cond1_est6(A);
            if (gas) {
                mx[10] = cA;
                for (int i = 0; i < 36; i++) A[i] = f.cgas[(long)i * f.ld + v];
                mx[11] = cond1_est6(A);
            } else {
                mx[9] = cA;
            }
        }
    }
    const int nb = gridDim.x;
    double* part = f.part;
    block_reduce_store<6, BS>(sm, part);
    __syncthreads();
    block_reduce_store<12, BS, OpMax>(mx, part + 6 * (size_t)nb);
    __syncthreads();
    block_reduce_store<1, BS, OpMin>(mnall, part + 18 * (size_t)nb);
    // arg-min of detF for solid and gas voxels: smallest value, ties -> larger index
    __shared__ double sv[2][BS];
    __shared__ long si[2][BS];
    __syncthreads();
    sv[0][threadIdx.x] = js; si[0][threadIdx.x] = is_;
    sv[1][threadIdx.x] = jg; si[1][threadIdx.x] = ig_;
    __syncthreads();
    for (int o = BS / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            for (int k = 0; k < 2; k++) {
                const double a = sv[k][threadIdx.x], b = sv[k][threadIdx.x + o];
                const long ia = si[k][threadIdx.x], ib = si[k][threadIdx.x + o];
                if (b < a || (b == a && ib > ia)) { sv[k][threadIdx.x] = b; si[k][threadIdx.x] = ib; }
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[19 * (size_t)nb + 4 * (size_t)blockIdx.x + 0] = sv[0][0];
        part[19 * (size_t)nb + 4 * (size_t)blockIdx.x + 1] = (double)si[0][0];
        part[19 * (size_t)nb + 4 * (size_t)blockIdx.x + 2] = sv[1][0];
        part[19 * (size_t)nb + 4 * (size_t)blockIdx.x + 3] = (double)si[1][0];
    }
}
// (minimum value, then largest index on ties) over the block candidates; the order is
// total, so the strided scan and the tree merge give the serial scan's answer
__device__ __forceinline__ bool argmin_better(double val, double idx, double best, double bi) {
    return idx >= 0.0 && (val < best || (val == best && idx > bi));
}
__global__ __launch_bounds__(BS) void k_argmin_final(const double* p, int nb, double* out) {
    __shared__ double sv[BS], si[BS];
    const int t = threadIdx.x;
    for (int k = 0; k < 2; k++) {
        double best = 1.7976931348623157e308, bi = -1.0;
        for (int b = t; b < nb; b += BS) {
            const double val = p[4 * b + 2 * k], idx = p[4 * b + 2 * k + 1];
            if (argmin_better(val, idx, best, bi)) { best = val; bi = idx; }
        }
        sv[t] = best;
        si[t] = bi;
        __syncthreads();
        for (int h = BS / 2; h > 0; h >>= 1) {
            if (t < h && argmin_better(sv[t + h], si[t + h], sv[t], si[t])) {
                sv[t] = sv[t + h];
                si[t] = si[t + h];
            }
            __syncthreads();
        }
        if (t == 0) {
            out[2 * k] = sv[0];
            out[2 * k + 1] = si[0];
        }
        __syncthreads();
    }
}
void launch_diagnostics(const Fields& f, double lfloor, int scan_cond, hipStream_t s, int* nb) {
    *nb = nblk(f.V);
    const int n = *nb;
    hipLaunchKernelGGL(k_diagnostics, dim3(n), dim3(BS), 0, s, f, lfloor, scan_cond);
    hipLaunchKernelGGL(k_reduce_final, dim3(6), dim3(BS), 0, s, f.part, n, 6, f.red);
    hipLaunchKernelGGL(k_reduce_final_op<OpMax>, dim3(12), dim3(BS), 0, s, f.part + 6 * (size_t)n, n, 12, f.red + 6);
    hipLaunchKernelGGL(k_reduce_final_op<OpMin>, dim3(1), dim3(BS), 0, s, f.part + 18 * (size_t)n, n, 1, f.red + 18);
    hipLaunchKernelGGL(k_argmin_final, dim3(1), dim3(BS), 0, s, f.part + 19 * (size_t)n, n, f.red + 19);
}

// take_snapshot / rollback (:2839-2895)
void launch_snapshot(const Fields& f, int restore, hipStream_t s, int full) {
    const size_t V = f.ld, ns = f.ns;  // whole columns, padding included
    struct { double* live; double* snap; size_t n; } T[] = {
        {f.L, f.s_L, 9 * V}, {f.Lold, f.s_Lold, 9 * V}, {f.sg, f.s_sg, 6 * V},
        {f.edotp, f.s_edotp, 6 * V}, {f.gdot, f.s_gdot, ns * V}, {f.ttau, f.s_ttau, 2 * ns * V}};
    for (auto& t : T) {
        if (!full && t.live != f.L) continue;
        if (restore) hipMemcpyAsync(t.live, t.snap, t.n * sizeof(double), hipMemcpyDeviceToDevice, s);
        else hipMemcpyAsync(t.snap, t.live, t.n * sizeof(double), hipMemcpyDeviceToDevice, s);
    }
}

// ---------------------------------------------------------------------------
// initial voxel state: data_grain (IO_functions.f90:759-809), Hall-Petch
// strengths (:100-199), initial_guess (various_functions.f90:8-53) and the
// 36 sums of w cg66 over solid voxels for the initial C0 (LS-EVPFFT.f90:156-166)
// ---------------------------------------------------------------------------
struct InitArgs { double udot[9]; double tdot; int keep_defgrad; };
__global__ __launch_bounds__(BS) void k_init_voxels(Fields f, PhaseTables pt, const double* __restrict__ euler,
                                                    const int* __restrict__ grain, const double* __restrict__ hp,
                                                    InitArgs a) {
    const long V = f.V;
    const long v = (long)blockIdx.x * BS + threadIdx.x;
    double acc[36];
    for (int i = 0; i < 36; i++) acc[i] = 0.0;
    if (v < V) {
        const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        // with an initial displacement field, F, F^-1, J and w_c are init_defgrad's (k_init_defgrad)
        const double wg = a.keep_defgrad ? f.w[v] : 1.0 / (double)V;
        if (!a.keep_defgrad) {
            f.w[v] = wg;
            f.J[v] = 1.0;
            for (int c = 0; c < 9; c++) { f.F[c * f.ld + v] = I3[c]; f.Fi[c * f.ld + v] = I3[c]; }
        }
        for (int c = 0; c < 9; c++) {
            f.Fe[c * f.ld + v] = I3[c]; f.Fp[c * f.ld + v] = I3[c];
            f.Finc[c * f.ld + v] = I3[c]; f.Feinc[c * f.ld + v] = I3[c];
            f.L[c * f.ld + v] = a.udot[c];
            f.Lold[c * f.ld + v] = 0.0;
            f.dLref[c * f.ld + v] = 0.0;
        }
        for (int c = 0; c < 6; c++) {
            f.sg[c * f.ld + v] = 0.0; f.sgt[c * f.ld + v] = 0.0; f.edotp[c * f.ld + v] = 0.0;
            f.ept[c * f.ld + v] = 0.0; f.dgs[c * f.ld + v] = 0.0;
        }
        f.epvm[v] = 0.0;
        f.gacc[v] = 0.0;
        for (int k = 0; k < f.ns; k++) f.gdot[(long)k * f.ld + v] = 0.0;
        for (int k = 0; k < 2 * f.ns; k++) { f.crss[(long)k * f.ld + v] = 0.0; f.ttau[(long)k * f.ld + v] = 0.0; }
        for (int k = 0; k < 5 * f.ns; k++) f.sch[(long)k * f.ld + v] = 0.0;
        for (int k = 0; k < 18; k++) f.schA[(long)k * f.ld + v] = 0.0;
        if (!f.gas[v]) {
            const PhaseDev& P = pt.ph[f.phase[v]];
            // Bunge angles -> a (sample -> crystal), tensor_functions.f90:525-569; ag = a^T
            const double d2r = PI_D / 180.0;
            const double ph = euler[3 * v] * d2r, th = euler[3 * v + 1] * d2r, tm = euler[3 * v + 2] * d2r;
            const double sph = sin(ph), cph = cos(ph), sth = sin(th), cth = cos(th), stm = sin(tm), ctm = cos(tm);
            double A[9];
            A[0] = ctm * cph - sph * stm * cth;  A[1] = ctm * sph + cph * stm * cth;  A[2] = sth * stm;
            A[3] = -stm * cph - sph * ctm * cth; A[4] = -sph * stm + cph * ctm * cth; A[5] = ctm * sth;
            A[6] = sph * sth;                    A[7] = -sth * cph;                   A[8] = cth;
            double ag[9];
            transp3(A, ag);
            for (int c = 0; c < 9; c++) f.ag[c * f.ld + v] = ag[c];
            double T[36], c66[36], s66[36];
            rot_matrix(ag, T);
            congr6(T, P.c66cr, c66);
            for (int i = 0; i < 36; i++) { s66[i] = c66[i]; acc[i] = c66[i] * wg; }
            if (!inv6(s66)) flag_singular(f, v);  // sg66 = cg66^-1 (data_grain)
            for (int i = 0; i < 36; i++) f.sg66[(long)i * f.ld + v] = s66[i];
            // initial_guess: sigma = cg66 (udot6 tdot)
            double u6[6], s6[6];
            to6(a.udot, u6);
            for (int i = 0; i < 6; i++) {
                double s = 0.0;
                for (int k = 0; k < 6; k++) s += c66[6 * i + k] * (u6[k] * a.tdot);
                s6[i] = s;
            }
            for (int c = 0; c < 6; c++) f.sg[c * f.ld + v] = s6[c];
            // Hall-Petch + directional term; trialtau(:,1) = crss(:,2)
            const double* tau = pt.data + P.off_tau;
            const double* hpf = pt.data + P.off_hpfac;
            const double* db = pt.data + P.off_dbca;
            const double hps = hp[v];  // Hall-Petch shift of this voxel's (grain, phase)
            for (int is = 0; is < P.nsyst; is++) {
                const double by = ag[3] * db[3 * is] + ag[4] * db[3 * is + 1] + ag[5] * db[3 * is + 2];
                const double dd = hpf[is] * sqrt(fabs(by));
                const double c1 = tau[3 * is] + hps + dd, c2 = tau[3 * is + 1] + hps + dd;
                f.crss[(2 * is) * f.ld + v] = c1;
                f.crss[(2 * is + 1) * f.ld + v] = c2;
                f.ttau[(2 * is) * f.ld + v] = c2;
                f.ttau[(2 * is + 1) * f.ld + v] = c2;
            }
        } else {
            for (int c = 0; c < 9; c++) f.ag[c * f.ld + v] = 0.0;
            for (int i = 0; i < 36; i++) f.sg66[(long)i * f.ld + v] = 0.0;
        }
    }
    block_reduce_store<36, BS>(acc, f.part);
}
void launch_init_voxels(const Fields& f, const PhaseTables& pt, const double* euler, const int* grain,
                        const double* hp_voxel, const double* udot, double tdot, int keep_defgrad, hipStream_t s,
                        int* nb) {
    InitArgs a;
    for (int i = 0; i < 9; i++) a.udot[i] = udot[i];
    a.tdot = tdot;
    a.keep_defgrad = keep_defgrad;
    *nb = nblk(f.V);
    hipLaunchKernelGGL(k_init_voxels, dim3(*nb), dim3(BS), 0, s, f, pt, euler, grain, hp_voxel, a);
}
void launch_copy(double* dst, const double* src, size_t n, hipStream_t s) {
    hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, s);
}
__global__ __launch_bounds__(BS) void k_fill(double* __restrict__ dst, double val, long n) {
    const long i = (long)blockIdx.x * BS + threadIdx.x;
    if (i < n) dst[i] = val;
}
void launch_fill(double* dst, double val, size_t n, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill, dim3(nblk((long)n)), dim3(BS), 0, s, dst, val, (long)n);
}

}  // namespace cg

namespace cg {
// ---------------------------------------------------------------------------
// nodal coordinates (init_micro types.f90:268-287; update_grid_velgrad_node
// various_functions.f90:2026-2054 with cell2node_tensor1 :2652-2691)
// ---------------------------------------------------------------------------
// init_defgrad (IO_functions.f90:815-936): F = grad(u0) + delt I, F0 = F, J, F^-1 and w = wgt J,
// with 30 block sums: w F^-1 (all, solid, gas voxels), w over solid and gas voxels, w.
// The reference's gradient_tensor1 (various_functions.f90:56-112) multiplies the displacement's
// spectrum by i sin(xi_m) (cell values, idispini 1) or by tan(xi_m/2) (i/4) prod_a (1 + e^{i xi_a})
// (node values, idispini 2); with e^{i xi_a} the shift by one voxel along a, these are exactly the
// periodic stencils (u(x + e_m) - u(x - e_m))/2 and, since (1 + e^{i xi}) tan(xi/2) = -i (e^{i xi} - 1),
// (1/4) sum over the four cell edges along m of u(edge end) - u(edge start), which this kernel
// applies in real space (the all-Nyquist term the reference zeroes is zero in the stencil).
__global__ __launch_bounds__(BS) void k_init_defgrad(Fields f, const double* __restrict__ disp, int idispini,
                                                     double delt) {
    const long V = f.V;
    const long v = (long)blockIdx.x * BS + threadIdx.x;
    double acc[30];
    for (int t = 0; t < 30; t++) acc[t] = 0.0;
    if (v < V) {
        const int n1 = f.n1, n2 = f.n2, n3 = f.n3;
        const int i = (int)(v % n1), j = (int)((v / n1) % n2), k = (int)(v / ((long)n1 * n2));
        auto wr = [](int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); };
        // cell (idispini 1) or node (2) value l at lattice point (a, b, c), periodic
        const long rs = idispini == 1 ? n1 : n1 + 1, ps = idispini == 1 ? (long)n1 * n2 : (long)(n1 + 1) * (n2 + 1);
        auto U = [&](int a, int b, int c, int l) { return disp[3 * (wr(a, n1) + rs * wr(b, n2) + ps * wr(c, n3)) + l]; };
        double F[9];
        for (int l = 0; l < 3; l++) {
            if (idispini == 1) {
                F[3 * l + 0] = 0.5 * (U(i + 1, j, k, l) - U(i - 1, j, k, l));
                F[3 * l + 1] = 0.5 * (U(i, j + 1, k, l) - U(i, j - 1, k, l));
                F[3 * l + 2] = 0.5 * (U(i, j, k + 1, l) - U(i, j, k - 1, l));
            } else {
                double g0 = 0.0, g1 = 0.0, g2 = 0.0;
                for (int p = 0; p < 2; p++)
                    for (int q = 0; q < 2; q++) {
                        g0 += U(i + 1, j + p, k + q, l) - U(i, j + p, k + q, l);
                        g1 += U(i + p, j + 1, k + q, l) - U(i + p, j, k + q, l);
                        g2 += U(i + p, j + q, k + 1, l) - U(i + p, j + q, k, l);
                    }
                F[3 * l + 0] = 0.25 * g0;
                F[3 * l + 1] = 0.25 * g1;
                F[3 * l + 2] = 0.25 * g2;
            }
        }
        F[0] += delt;
        F[4] += delt;
        F[8] += delt;
        // determinant33 (tensor_functions.f90:442-456) term order
        const double J = F[0] * F[4] * F[8] + F[1] * F[5] * F[6] + F[2] * F[3] * F[7] - F[6] * F[4] * F[2] -
                         F[7] * F[5] * F[0] - F[8] * F[3] * F[1];
        double Fi[9];
        if (inv3(F, Fi) == 0.0) flag_singular(f, v);  // lu_inverse (IO_functions.f90:907)
        const double wj = J / (double)V;
        const bool gas = f.gas[v] != 0;
        for (int c = 0; c < 9; c++) {
            f.F[c * f.ld + v] = F[c];
            f.Fini[c * f.ld + v] = F[c];
            f.Fi[c * f.ld + v] = Fi[c];
            acc[c] = Fi[c] * wj;
            acc[(gas ? 18 : 9) + c] = Fi[c] * wj;
        }
        f.J[v] = J;
        f.w[v] = wj;
        acc[gas ? 28 : 27] = wj;
        acc[29] = wj;
    }
    block_reduce_store<30, BS>(acc, f.part);
}
void launch_init_defgrad(const Fields& f, const double* disp, int idispini, double delt, hipStream_t s, int* nb) {
    *nb = nblk(f.V);
    hipLaunchKernelGGL(k_init_defgrad, dim3(*nb), dim3(BS), 0, s, f, disp, idispini, delt);
}
// xnode += the initial nodal displacement: the file's node values (idispini 2) or the mean of the
// eight cells around each node (idispini 1; cell2node_tensor1, various_functions.f90:2652-2691,
// summed in its order)
__global__ __launch_bounds__(BS) void k_xnode_disp(Fields f, const double* __restrict__ disp, int idispini,
                                                   int n3node) {
    const int n1 = f.n1, n2 = f.n2, n3 = f.n3;
    const long nn = (long)(n1 + 1) * (n2 + 1) * n3node;
    const long q = (long)blockIdx.x * BS + threadIdx.x;
    if (q >= nn) return;
    if (idispini == 2) {
        for (int l = 0; l < 3; l++) f.xnode[3 * q + l] += disp[3 * q + l];
        return;
    }
    const int a = (int)(q % (n1 + 1)), b = (int)((q / (n1 + 1)) % (n2 + 1)), c = (int)(q / ((long)(n1 + 1) * (n2 + 1)));
    const int i1 = a == 0 ? n1 - 1 : a - 1, i2 = a == n1 ? 0 : a;
    const int j1 = b == 0 ? n2 - 1 : b - 1, j2 = b == n2 ? 0 : b;
    const int k1 = c == 0 ? n3 - 1 : c - 1, k2 = c == n3 ? 0 : c;
    auto C = [&](int x, int y, int z, int l) { return disp[3 * (x + (long)n1 * (y + (long)n2 * z)) + l]; };
    for (int l = 0; l < 3; l++) {
        double s = C(i2, j2, k2, l);
        s += C(i1, j2, k2, l);
        s += C(i2, j1, k2, l);
        s += C(i2, j2, k1, l);
        s += C(i1, j1, k2, l);
        s += C(i1, j2, k1, l);
        s += C(i2, j1, k1, l);
        s += C(i1, j1, k1, l);
        f.xnode[3 * q + l] += s * 0.125;
    }
}
void launch_xnode_disp(const Fields& f, const double* disp, int idispini, int n3node, hipStream_t s) {
    const long nn = (long)(f.n1 + 1) * (f.n2 + 1) * n3node;
    hipLaunchKernelGGL(k_xnode_disp, dim3((unsigned)((nn + BS - 1) / BS)), dim3(BS), 0, s, f, disp, idispini, n3node);
}
__global__ __launch_bounds__(BS) void k_init_xnode(Fields f, int n3node) {
    const long nn = (long)(f.n1 + 1) * (f.n2 + 1) * n3node;
    const long q = (long)blockIdx.x * BS + threadIdx.x;
    if (q >= nn) return;
    const int i = q % (f.n1 + 1), j = (q / (f.n1 + 1)) % (f.n2 + 1), k = q / ((long)(f.n1 + 1) * (f.n2 + 1));
    f.xnode[3 * q] = (i + 1) - 0.5;
    f.xnode[3 * q + 1] = (j + 1) - 0.5;
    f.xnode[3 * q + 2] = n3node == 1 ? 0.0 : (k + 1) - 0.5;
}
struct NodeArgs { double lavg[9]; double tdot; int ideriv, n3node; };
__global__ __launch_bounds__(BS) void k_xnode(Fields f, const double* __restrict__ dv, NodeArgs a) {
    const int n1 = f.n1, n2 = f.n2, n3 = f.n3;
    const long V = f.V;
    const long nn = (long)(n1 + 1) * (n2 + 1) * a.n3node;
    const long q = (long)blockIdx.x * BS + threadIdx.x;
    if (q >= nn) return;
    const int i = q % (n1 + 1) + 1, j = (q / (n1 + 1)) % (n2 + 1) + 1, k = q / ((long)(n1 + 1) * (n2 + 1)) + 1;
    auto D = [&](int l, int ii, int jj, int kk) {
        return dv[l * V + (ii - 1) + (long)n1 * ((jj - 1) + (long)n2 * (kk - 1))];
    };
    double dn[3];
    if (a.ideriv != 2) {
        int k1 = k - 1, k2 = k, j1 = j - 1, j2 = j, i1 = i - 1, i2 = i;
        if (k1 == 0) k1 = n3;
        if (k2 == n3 + 1) k2 = 1;
        if (j1 == 0) j1 = n2;
        if (j2 == n2 + 1) j2 = 1;
        if (i1 == 0) i1 = n1;
        if (i2 == n2 + 1) i2 = 1;  // the reference wraps the x index with npts2 (:2678)
        for (int l = 0; l < 3; l++)
            dn[l] = (D(l, i2, j2, k2) + D(l, i1, j2, k2) + D(l, i2, j1, k2) + D(l, i2, j2, k1) + D(l, i1, j1, k2) +
                     D(l, i1, j2, k1) + D(l, i2, j1, k1) + D(l, i1, j1, k1)) * 0.125;
    } else {
        const int ii = i > n1 ? 1 : i, jj = j > n2 ? 1 : j, kk = k > n3 ? 1 : k;
        for (int l = 0; l < 3; l++) dn[l] = D(l, ii, jj, kk);
    }
    const double x[3] = {i - 0.5, j - 0.5, a.n3node == 1 ? 0.0 : k - 0.5};
    for (int l = 0; l < 3; l++) {
        const double lx = a.lavg[3 * l] * x[0] + a.lavg[3 * l + 1] * x[1] + a.lavg[3 * l + 2] * x[2];
        f.xnode[3 * q + l] += (lx + dn[l]) * a.tdot;
    }
}
void launch_init_xnode(const Fields& f, int n3node, hipStream_t s) {
    const long nn = (long)(f.n1 + 1) * (f.n2 + 1) * n3node;
    hipLaunchKernelGGL(k_init_xnode, dim3((unsigned)((nn + BS - 1) / BS)), dim3(BS), 0, s, f, n3node);
}
void launch_xnode(const Fields& f, const double* dv, const double* lavg, double tdot, int ideriv, int n3node,
                  hipStream_t s) {
    NodeArgs a;
    for (int i = 0; i < 9; i++) a.lavg[i] = lavg[i];
    a.tdot = tdot;
    a.ideriv = ideriv;
    a.n3node = n3node;
    const long nn = (long)(f.n1 + 1) * (f.n2 + 1) * n3node;
    hipLaunchKernelGGL(k_xnode, dim3((unsigned)((nn + BS - 1) / BS)), dim3(BS), 0, s, f, dv, a);
}
}  // namespace cg
