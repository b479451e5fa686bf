// fft_kernels.hip -- the FFT + Green-operator stage of one fixed-point pass.
//
// Reference: convolution_with_Goper (various_functions.f90:966-1059) with the
// FFTW plans of fourier_functions.f90:141-146 (9 in-place 3-D c2c transforms,
// forward e^{-i xi.x}, inverse unnormalised then Re()/V) and the Green operator
// of form_G (various_functions.f90:609-809).
//
// MI355X form: real-to-complex along x (Gamma is real and even in xi, so the
// Hermitian half spectrum carries the whole convolution), five LDS-staged
// Stockham passes over the 9 components:
//   x-fwd (fused PK1 = J (sigma - <sigma>) F^-T)  ->  y-fwd  ->
//   z-fwd + Gamma(xi) computed on the fly from C0 + z-inv  ->  y-inv  ->
//   x-inv (c2r, /V) writing dLref and the block partial sums of w dLref F^-1.
// The half spectrum is component-interleaved: spec[((k*n2 + j)*h1 + kx)*9 + c].
// Template parameter R selects the arithmetic of the spectral stage (double:
// FP64 form; float: mixed form with the input scaled by a stress norm).
#include <hip/hip_ext.h>
#include <string.h>

#include <vector>

#include "internal.h"
#include "reduce.h"
#include "tensor.h"
#include "fft_reg.h"

namespace cg {

constexpr int FBS = 256;

template <typename R> struct C2;
template <> struct C2<double> { using T = double2; };
template <> struct C2<float> { using T = float2; };

template <typename T2> __device__ __forceinline__ T2 cmul(T2 a, T2 b) {
    T2 r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
template <typename T2> __device__ __forceinline__ T2 cadd(T2 a, T2 b) { T2 r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename T2> __device__ __forceinline__ T2 csub(T2 a, T2 b) { T2 r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }

struct Axis {
    int n, np;
    int rad[12];
    int tw_off;  // offset of this axis' twiddle table (n entries) in the table buffer
};
struct Plan3 {
    Axis ax[3];
    const double2* tw64;
    const float2* tw32;
};

template <typename R> __device__ __forceinline__ typename C2<R>::T twid(const Plan3& p, int off, int idx);
template <> __device__ __forceinline__ double2 twid<double>(const Plan3& p, int off, int idx) { return p.tw64[off + idx]; }
template <> __device__ __forceinline__ float2 twid<float>(const Plan3& p, int off, int idx) { return p.tw32[off + idx]; }

// radix-R butterfly (forward: e^{-2 pi i rk/R}; inverse: conjugate)
template <typename T2, typename R>
__device__ __forceinline__ void butterfly(T2* x, int Rr, bool inv) {
    if (Rr == 2) {
        T2 a = x[0], b = x[1];
        x[0] = cadd(a, b);
        x[1] = csub(a, b);
    } else if (Rr == 4) {
        T2 a = cadd(x[0], x[2]), b = csub(x[0], x[2]), c = cadd(x[1], x[3]), d = csub(x[1], x[3]);
        // -i d (forward) / +i d (inverse)
        T2 md;
        if (!inv) { md.x = d.y; md.y = -d.x; } else { md.x = -d.y; md.y = d.x; }
        x[0] = cadd(a, c);
        x[2] = csub(a, c);
        x[1] = cadd(b, md);
        x[3] = csub(b, md);
    } else {  // 3
        const R h = (R)0.86602540378443864676;  // sqrt(3)/2
        T2 t1 = cadd(x[1], x[2]);
        T2 t2;
        t2.x = x[0].x - (R)0.5 * t1.x;
        t2.y = x[0].y - (R)0.5 * t1.y;
        T2 d = csub(x[1], x[2]);
        T2 s;  // -i h d (forward)
        if (!inv) { s.x = h * d.y; s.y = -h * d.x; } else { s.x = -h * d.y; s.y = h * d.x; }
        x[0] = cadd(x[0], t1);
        x[1] = cadd(t2, s);
        x[2] = csub(t2, s);
    }
}

// Stockham passes over nsig signals of length n in LDS (signal stride ss).
// Returns the buffer that holds the result.
template <typename R>
__device__ typename C2<R>::T* lds_fft(typename C2<R>::T* src, typename C2<R>::T* dst, int nsig, int ss,
                                      const Plan3& p, int axis, bool inv) {
    using T2 = typename C2<R>::T;
    const Axis& A = p.ax[axis];
    const int n = A.n;
    int Ns = 1;
    for (int ps = 0; ps < A.np; ps++) {
        const int Rr = A.rad[ps];
        const int m = n / Rr;
        const int stepN = n / (Ns * Rr);
        const int items = nsig * m;
        for (int it = threadIdx.x; it < items; it += blockDim.x) {
            const int sig = it / m, j = it - sig * m;
            const int k = j % Ns;
            T2 x[4];
            const T2* s = src + sig * ss;
#pragma unroll
            for (int r = 0; r < 4; r++)
                if (r < Rr) x[r] = s[j + r * m];
#pragma unroll
            for (int r = 1; r < 4; r++)
                if (r < Rr) {
                    T2 w = twid<R>(p, A.tw_off, (k * r * stepN) % n);
                    if (inv) w.y = -w.y;
                    x[r] = cmul(x[r], w);
                }
            butterfly<T2, R>(x, Rr, inv);
            const int d = (j / Ns) * Ns * Rr + k;
            T2* o = dst + sig * ss;
#pragma unroll
            for (int r = 0; r < 4; r++)
                if (r < Rr) o[d + r * Ns] = x[r];
        }
        __syncthreads();
        T2* t = src; src = dst; dst = t;
        Ns *= Rr;
    }
    return src;
}

// ---------------------------------------------------------------------------
// x forward with fused PK1 (LS-EVPFFT.f90:256-269; CauchyToPK1 :954-963)
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(FBS) void k_x_fwd_pk1(Fields f, Plan3 p, const double* __restrict__ sgsum, int rows,
                                                   double inv_scale) {
    using T2 = typename C2<R>::T;
    extern __shared__ __align__(16) unsigned char smem[];
    const int n1 = f.n1, h1 = f.h1;
    const int ss = n1 + 1;
    const int nsig = rows * 9;
    T2* b0 = (T2*)smem;
    T2* b1 = b0 + nsig * ss;
    const long V = f.V;
    double sa[9];
    for (int c = 0; c < 9; c++) sa[c] = sgsum[c];
    const long row0 = (long)blockIdx.x * rows;
    for (int e = threadIdx.x; e < rows * n1; e += blockDim.x) {
        const int r = e / n1, i = e - r * n1;
        const long v = (row0 + r) * n1 + i;
        double s6[6], s[9], Fi[9], pk[9];
        for (int c = 0; c < 6; c++) s6[c] = f.sg[c * f.ld + v];
        for (int c = 0; c < 9; c++) Fi[c] = f.Fi[c * f.ld + v];
        from6(s6, s);
        const double J = f.J[v];
        for (int c = 0; c < 9; c++) s[c] -= sa[c];
        mmt3(s, Fi, pk);  // (sigma - <sigma>) F^-T
        for (int c = 0; c < 9; c++) {
            T2 z;
            z.x = (R)(pk[c] * J * inv_scale);
            z.y = (R)0;
            b0[(r * 9 + c) * ss + i] = z;
        }
    }
    __syncthreads();
    T2* res = lds_fft<R>(b0, b1, nsig, ss, p, 0, false);
    T2* spec = (T2*)(sizeof(R) == 8 ? (void*)f.spec : (void*)f.specf);
    for (int e = threadIdx.x; e < rows * h1 * 9; e += blockDim.x) {
        const int r = e / (h1 * 9), q = e - r * h1 * 9;
        const int kx = q / 9, c = q - kx * 9;
        spec[((row0 + r) * h1 + kx) * 9 + c] = res[(r * 9 + c) * ss + kx];
    }
}

// ---------------------------------------------------------------------------
// y pass (forward or inverse) over tiles of T kx columns of one z plane
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(FBS) void k_y(Fields f, Plan3 p, int T, int inv, int nc) {
    using T2 = typename C2<R>::T;
    extern __shared__ __align__(16) unsigned char smem[];
    const int n2 = f.n2, h1 = f.h1;
    const int ntile = (h1 + T - 1) / T;
    const int k = blockIdx.x / ntile, tile = blockIdx.x - k * ntile;
    const int kx0 = tile * T;
    const int Te = min(T, h1 - kx0);
    const int nsig = Te * nc, ss = n2 + 1;
    T2* b0 = (T2*)smem;
    T2* b1 = b0 + T * 9 * ss;
    T2* spec = (T2*)(sizeof(R) == 8 ? (void*)f.spec : (void*)f.specf);
    const int w = Te * nc;
    for (int e = threadIdx.x; e < n2 * w; e += blockDim.x) {
        const int j = e / w, q = e - j * w;
        const int t = q / nc, c = q - t * nc;
        b0[q * ss + j] = spec[(((long)k * n2 + j) * h1 + kx0 + t) * 9 + c];
    }
    __syncthreads();
    T2* res = lds_fft<R>(b0, b1, nsig, ss, p, 1, inv != 0);
    for (int e = threadIdx.x; e < n2 * w; e += blockDim.x) {
        const int j = e / w, q = e - j * w;
        const int t = q / nc, c = q - t * nc;
        spec[(((long)k * n2 + j) * h1 + kx0 + t) * 9 + c] = res[q * ss + j];
    }
}

// ---------------------------------------------------------------------------
// z pass with the Green operator: forward z, Gamma(xi) : P^, inverse z
// ---------------------------------------------------------------------------

// full 81-component Gamma at 0-based frequency (kx, ky, kz) for igamma = 1, 2
__device__ void green81(const GreenDev& g, int kx, int ky, int kz, int n1, int n2, int n3, double* G) {
    const double tp = 2.0 * PI_D;
    double xi[3];
    xi[0] = tp * (double)((kx + 1) <= n1 / 2 ? kx : kx - n1) / n1;
    xi[1] = tp * (double)((ky + 1) <= n2 / 2 ? ky : ky - n2) / n2;
    xi[2] = (n3 == 1) ? 0.0 : tp * (double)((kz + 1) <= n3 / 2 ? kz : kz - n3) / n3;
    const bool nx = (n1 % 2 == 0) && kx == n1 / 2, ny = (n2 % 2 == 0) && ky == n2 / 2,
               nz = (n3 % 2 == 0) && kz == n3 / 2;
    if (g.igamma >= 2 && ((nx && ny) || (nx && nz) || (ny && nz))) {
        for (int i = 0; i < 81; i++) G[i] = 0.0;
        return;
    }
    double D[9];
    if (g.igamma == 1) {
        const double xs = tp * kx / n1, ys = tp * ky / n2, zs = tp * kz / n3;
        D[0] = 2.0 * (cos(xs) - 1.0); D[4] = 2.0 * (cos(ys) - 1.0); D[8] = 2.0 * (cos(zs) - 1.0);
        D[1] = D[3] = -sin(xs) * sin(ys);
        D[2] = D[6] = -sin(xs) * sin(zs);
        D[5] = D[7] = -sin(ys) * sin(zs);
    } else {
        double pr = 1.0, pim = 0.0;
        for (int l = 0; l < 3; l++) {
            const double cr = 1.0 + cos(xi[l]), ci = sin(xi[l]);
            const double nr = pr * cr - pim * ci, ni = pr * ci + pim * cr;
            pr = nr; pim = ni;
        }
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
                    for (int l = 0; l < 3; l++) s += g.c0[27 * i + 9 * j + 3 * k + l] * D[3 * j + l];
                A[3 * i + k] = s;
            }
        inv3(A, Ai);
    }
    for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++)
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) G[27 * p + 9 * q + 3 * i + j] = -Ai[3 * p + i] * D[3 * q + j];
    if (g.igamma == 1) {
        double gt[81], aux[81];
        for (int i = 0; i < 81; i++) gt[i] = G[i];
        for (int it = 0; it < 4; it++) {
            for (int ij = 0; ij < 9; ij++) {
                double cgv[9];
                for (int op = 0; op < 9; op++) {
                    double s = 0.0;
                    for (int mn = 0; mn < 9; mn++) s += G[9 * ij + mn] * g.c0[9 * mn + op];
                    cgv[op] = s;
                }
                for (int kl = 0; kl < 9; kl++) {
                    double s = 0.0;
                    for (int op = 0; op < 9; op++) s += cgv[op] * gt[9 * op + kl];
                    aux[9 * ij + kl] = -s;
                }
            }
            for (int i = 0; i < 81; i++) gt[i] = aux[i];
        }
        for (int i = 0; i < 81; i++) G[i] = gt[i];
    }
}

// general igamma (1, 2): the full 81-component operator, kept out of line so the
// igamma = 0 kernel stays small
template <typename R>
__device__ __noinline__ void green_apply_general(const GreenDev& g, int kx, int ky, int kz, int n1, int n2, int n3,
                                                 typename C2<R>::T* P, int st) {
    using T2 = typename C2<R>::T;
    T2 in[9];
    for (int c = 0; c < 9; c++) in[c] = P[c * st];
    // The reference keeps Re() of a full c2c inverse, i.e. it applies the even part
    // (Gamma(xi) + Gamma(-xi))/2 to the Hermitian spectrum; igamma = 2 is not even on
    // the Nyquist planes (xi = -pi is not mirrored), so both halves are evaluated.
    double G[81], Gm[81];
    green81(g, kx, ky, kz, n1, n2, n3, G);
    green81(g, (n1 - kx) % n1, (n2 - ky) % n2, (n3 - kz) % n3, n1, n2, n3, Gm);
    for (int i = 0; i < 81; i++) G[i] = 0.5 * (G[i] + Gm[i]);
    for (int ij = 0; ij < 9; ij++) {
        double sr = 0.0, si = 0.0;
        for (int kl = 0; kl < 9; kl++) {
            sr += G[9 * ij + kl] * (double)in[kl].x;
            si += G[9 * ij + kl] * (double)in[kl].y;
        }
        P[ij * st].x = (R)sr;
        P[ij * st].y = (R)si;
    }
}

// out_ij = sum_kl Gamma_ijkl(xi) P_kl at one frequency (9 complex in place, stride st)
template <typename R, bool GENERAL>
__device__ __forceinline__ void green_apply(const GreenDev& g, int kx, int ky, int kz, int n1, int n2, int n3,
                                            typename C2<R>::T* P, int st) {
    using T2 = typename C2<R>::T;
    if (kx == 0 && ky == 0 && kz == 0) {
        for (int c = 0; c < 9; c++) { P[c * st].x = (R)0; P[c * st].y = (R)0; }
        return;
    }
    if (GENERAL) {
        green_apply_general<R>(g, kx, ky, kz, n1, n2, n3, P, st);
        return;
    }
    const bool nx = (n1 % 2 == 0) && kx == n1 / 2, ny = (n2 % 2 == 0) && ky == n2 / 2,
               nz = (n3 % 2 == 0) && kz == n3 / 2 && n3 > 1;
    T2 in[9];
#pragma unroll
    for (int c = 0; c < 9; c++) in[c] = P[c * st];
    {
        if (nx || ny || nz) {  // Gamma = -S0 on any Nyquist plane (form_G :646-651)
            for (int ij = 0; ij < 9; ij++) {
                double sr = 0.0, si = 0.0;
                for (int kl = 0; kl < 9; kl++) {
                    sr -= g.s0[9 * ij + kl] * (double)in[kl].x;
                    si -= g.s0[9 * ij + kl] * (double)in[kl].y;
                }
                P[ij * st].x = (R)sr;
                P[ij * st].y = (R)si;
            }
            return;
        }
        const double tp = 2.0 * PI_D;
        double xi[3];
        xi[0] = tp * (double)((kx + 1) <= n1 / 2 ? kx : kx - n1) / n1;
        xi[1] = tp * (double)((ky + 1) <= n2 / 2 ? ky : ky - n2) / n2;
        xi[2] = (n3 == 1) ? 0.0 : tp * (double)((kz + 1) <= n3 / 2 ? kz : kz - n3) / n3;
        // A = C0 : (xi x xi), out = -(A^-1 (P xi)) x xi
        double A[9], Ai[9];
        const double d00 = xi[0] * xi[0], d11 = xi[1] * xi[1], d22 = xi[2] * xi[2], d01 = xi[0] * xi[1],
                     d02 = xi[0] * xi[2], d12 = xi[1] * xi[2];
        for (int i = 0; i < 3; i++)
            for (int k = 0; k < 3; k++) {
                const double* c = g.c0 + 27 * i + 3 * k;  // c[9*j + l] = c0_{i j k l}
                A[3 * i + k] = fma(c[11] + c[19], d12,
                                   fma(c[2] + c[18], d02,
                                       fma(c[1] + c[9], d01, fma(c[20], d22, fma(c[10], d11, c[0] * d00)))));
            }
        inv3(A, Ai);
        double qr[3], qi[3];
        for (int k = 0; k < 3; k++) {
            qr[k] = fma((double)in[3 * k + 2].x, xi[2], fma((double)in[3 * k + 1].x, xi[1], (double)in[3 * k].x * xi[0]));
            qi[k] = fma((double)in[3 * k + 2].y, xi[2], fma((double)in[3 * k + 1].y, xi[1], (double)in[3 * k].y * xi[0]));
        }
        for (int i = 0; i < 3; i++) {
            const double ur = -fma(Ai[3 * i + 2], qr[2], fma(Ai[3 * i + 1], qr[1], Ai[3 * i] * qr[0]));
            const double ui = -fma(Ai[3 * i + 2], qi[2], fma(Ai[3 * i + 1], qi[1], Ai[3 * i] * qi[0]));
            for (int j = 0; j < 3; j++) {
                P[(3 * i + j) * st].x = (R)(ur * xi[j]);
                P[(3 * i + j) * st].y = (R)(ui * xi[j]);
            }
        }
        return;
    }
}

template <typename R, bool GENERAL>
__global__ __launch_bounds__(FBS) void k_z_green(Fields f, Plan3 p, GreenDev g, int T) {
    using T2 = typename C2<R>::T;
    extern __shared__ __align__(16) unsigned char smem[];
    const int n2 = f.n2, n3 = f.n3, h1 = f.h1;
    const int ntile = (h1 + T - 1) / T;
    const int j = blockIdx.x / ntile, tile = blockIdx.x - j * ntile;
    const int kx0 = tile * T;
    const int Te = min(T, h1 - kx0);
    const int nsig = Te * 9, ss = n3 + 1;
    T2* b0 = (T2*)smem;
    T2* b1 = b0 + T * 9 * ss;
    T2* spec = (T2*)(sizeof(R) == 8 ? (void*)f.spec : (void*)f.specf);
    const int w = Te * 9;
    for (int e = threadIdx.x; e < n3 * w; e += blockDim.x) {
        const int k = e / w, q = e - k * w;
        b0[q * ss + k] = spec[(((long)k * n2 + j) * h1 + kx0) * 9 + q];
    }
    __syncthreads();
    T2* res = lds_fft<R>(b0, b1, nsig, ss, p, 2, false);
    for (int e = threadIdx.x; e < Te * n3; e += blockDim.x) {
        const int t = e / n3, kz = e - t * n3;
        green_apply<R, GENERAL>(g, kx0 + t, j, kz, f.n1, n2, n3, res + (t * 9) * ss + kz, ss);
    }
    __syncthreads();
    T2* other = (res == b0) ? b1 : b0;
    res = lds_fft<R>(res, other, nsig, ss, p, 2, true);
    for (int e = threadIdx.x; e < n3 * w; e += blockDim.x) {
        const int k = e / w, q = e - k * w;
        spec[(((long)k * n2 + j) * h1 + kx0) * 9 + q] = res[q * ss + k];
    }
}

// ---------------------------------------------------------------------------
// x inverse (c2r): dLref = Re(IFFT)/V, block partial sums of w dLref F^-1
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(FBS) void k_x_inv(Fields f, Plan3 p, int rows, double out_scale) {
    using T2 = typename C2<R>::T;
    extern __shared__ __align__(16) unsigned char smem[];
    const int n1 = f.n1, h1 = f.h1;
    const int ss = n1 + 1;
    const int nsig = rows * 9;
    T2* b0 = (T2*)smem;
    T2* b1 = b0 + nsig * ss;
    const long V = f.V;
    const long row0 = (long)blockIdx.x * rows;
    T2* spec = (T2*)(sizeof(R) == 8 ? (void*)f.spec : (void*)f.specf);
    for (int e = threadIdx.x; e < rows * h1 * 9; e += blockDim.x) {
        const int r = e / (h1 * 9), q = e - r * h1 * 9;
        const int kx = q / 9, c = q - kx * 9;
        T2 z = spec[((row0 + r) * h1 + kx) * 9 + c];
        T2* s = b0 + (r * 9 + c) * ss;
        s[kx] = z;
        if (kx > 0 && kx < n1 - kx) {  // Hermitian mirror X[n-k] = conj X[k]
            T2 cz;
            cz.x = z.x;
            cz.y = -z.y;
            s[n1 - kx] = cz;
        }
    }
    __syncthreads();
    T2* res = lds_fft<R>(b0, b1, nsig, ss, p, 0, true);
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int e = threadIdx.x; e < rows * n1; e += blockDim.x) {
        const int r = e / n1, i = e - r * n1;
        const long v = (row0 + r) * n1 + i;
        double d[9], Fi[9], q[9];
        for (int c = 0; c < 9; c++) {
            d[c] = (double)res[(r * 9 + c) * ss + i].x * out_scale;
            f.dLref[c * f.ld + v] = d[c];
            Fi[c] = f.Fi[c * f.ld + v];
        }
        mm3(d, Fi, q);
        const double w = f.w[v];
        for (int c = 0; c < 9; c++) acc[c] += q[c] * w;
    }
    block_reduce_store_cm<9, FBS>(acc, f.part, 0);
}

// ---------------------------------------------------------------------------
// host side: plans, twiddles, launches
// ---------------------------------------------------------------------------
static bool factor(int n, Axis& a) {
    a.n = n;
    a.np = 0;
    int m = n;
    while (m % 4 == 0 && m > 1) { a.rad[a.np++] = 4; m /= 4; }
    while (m % 3 == 0) { a.rad[a.np++] = 3; m /= 3; }
    while (m % 2 == 0) { a.rad[a.np++] = 2; m /= 2; }
    return m == 1 && a.np <= 12;
}

struct FftCtx {
    Plan3 plan;
    double2* tw64;
    float2* tw32;
    int rows, ty, tz;
    int use_reg;  // register-FFT kernels for the per-pass stage (all axis lengths supported)
    FftTune tune;  // block shapes and placement of the register passes
    // the Green operator constants the register z pass reads (device copy, its pinned staging)
    GreenDev* d_g = nullptr;
    GreenDev* h_g = nullptr;
    int g_valid = 0;
};
// The device copy of g, uploaded on stream s when it differs from the last one (g changes with C0,
// once per increment; the stream has passed the previous upload by then, as every pass waits
// for its sums).
static const GreenDev* green_dev(FftCtx* c, const GreenDev& g, hipStream_t s) {
    // compared member by member: the struct's tail padding is not part of the value
    const bool same = c->g_valid && c->h_g->igamma == g.igamma && memcmp(c->h_g->c0, g.c0, sizeof g.c0) == 0 &&
                      memcmp(c->h_g->s0, g.s0, sizeof g.s0) == 0;
    if (!same) {
        *c->h_g = g;
        hipMemcpyAsync(c->d_g, c->h_g, sizeof g, hipMemcpyHostToDevice, s);
        c->g_valid = 1;
    }
    return c->d_g;
}

FftCtx* fft_create(const Fields& f, char* err, int errlen) {
    FftCtx* c = new FftCtx();
    int ns[3] = {f.n1, f.n2, f.n3};
    int off = 0;
    for (int a = 0; a < 3; a++) {
        if (!factor(ns[a], c->plan.ax[a])) {
            snprintf(err, errlen, "grid size %d is not of the form 2^a 3^b", ns[a]);
            delete c;
            return nullptr;
        }
        c->plan.ax[a].tw_off = off;
        off += ns[a];
    }
    std::vector<double2> h64(off);
    std::vector<float2> h32(off);
    for (int a = 0; a < 3; a++) {
        const int n = ns[a];
        for (int k = 0; k < n; k++) {
            long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
            double2 z;
            z.x = (double)cosl(ang);
            z.y = (double)sinl(ang);
            h64[c->plan.ax[a].tw_off + k] = z;
            float2 zf;
            zf.x = (float)z.x;
            zf.y = (float)z.y;
            h32[c->plan.ax[a].tw_off + k] = zf;
        }
    }
    hipMalloc(&c->tw64, sizeof(double2) * off);
    hipMalloc(&c->tw32, sizeof(float2) * off);
    hipMemcpy(c->tw64, h64.data(), sizeof(double2) * off, hipMemcpyHostToDevice);
    hipMemcpy(c->tw32, h32.data(), sizeof(float2) * off, hipMemcpyHostToDevice);
    c->plan.tw64 = c->tw64;
    c->plan.tw32 = c->tw32;
    if (hipMalloc(&c->d_g, sizeof(GreenDev)) != hipSuccess || hipHostMalloc(&c->h_g, sizeof(GreenDev)) != hipSuccess) {
        snprintf(err, errlen, "fft_create: allocation of the Green operator constants failed");
        fft_destroy(c);
        return nullptr;
    }
    // rows per x block and kx tile widths (two complex buffers must fit the LDS)
    c->rows = 2;
    while ((long)(f.n2) * f.n3 % c->rows) c->rows /= 2;
    c->ty = 2;
    c->tz = 2;
    c->use_reg = 1;
    return c;
}
void fft_destroy(FftCtx* c) {
    if (!c) return;
    hipFree(c->tw64);
    hipFree(c->tw32);
    if (c->d_g) hipFree(c->d_g);
    if (c->h_g) hipHostFree(c->h_g);
    delete c;
}
int fft_rows(const FftCtx* c) { return c->rows; }

template <typename R>
static int run_green(const Fields& f, FftCtx* c, const double* sgsum, const GreenDev& g, double scale,
                     hipStream_t s) {
    using T2 = typename C2<R>::T;
    const int rows = c->rows;
    const int nrow = f.n2 * f.n3 / rows;
    const size_t shx = 2 * (size_t)rows * 9 * (f.n1 + 1) * sizeof(T2);
    const size_t shy = 2 * (size_t)c->ty * 9 * (f.n2 + 1) * sizeof(T2);
    const size_t shz = 2 * (size_t)c->tz * 9 * (f.n3 + 1) * sizeof(T2);
    const int ty_blocks = f.n3 * ((f.h1 + c->ty - 1) / c->ty);
    const int tz_blocks = f.n2 * ((f.h1 + c->tz - 1) / c->tz);
    hipLaunchKernelGGL(k_x_fwd_pk1<R>, dim3(nrow), dim3(FBS), shx, s, f, c->plan, sgsum, rows, 1.0 / scale);
    if (f.n2 > 1) hipLaunchKernelGGL(k_y<R>, dim3(ty_blocks), dim3(FBS), shy, s, f, c->plan, c->ty, 0, 9);
    if (g.igamma == 0) hipLaunchKernelGGL((k_z_green<R, false>), dim3(tz_blocks), dim3(FBS), shz, s, f, c->plan, g, c->tz);
    else hipLaunchKernelGGL((k_z_green<R, true>), dim3(tz_blocks), dim3(FBS), shz, s, f, c->plan, g, c->tz);
    if (f.n2 > 1) hipLaunchKernelGGL(k_y<R>, dim3(ty_blocks), dim3(FBS), shy, s, f, c->plan, c->ty, 1, 9);
    hipLaunchKernelGGL(k_x_inv<R>, dim3(nrow), dim3(FBS), shx, s, f, c->plan, rows, scale / (double)f.V);
    return nrow;
}

bool fft_reg_supported(const Fields& f);
// passes st0 .. st1-1 of the register form; returns the last pass' block count
template <typename R>
int run_green_reg(const Fields& f, const typename Cx<R>::T* tw1, const typename Cx<R>::T* tw2,
                  const typename Cx<R>::T* tw3, const double* sgsum, const GreenDev* g, int general,
                  const double* dscale, const FftTune& t, hipStream_t s, int st0, int st1, hipEvent_t ev0,
                  hipEvent_t ev1);
template <typename R>
int tune_reg(const Fields& f, const typename Cx<R>::T* tw1, const typename Cx<R>::T* tw2,
             const typename Cx<R>::T* tw3, const double* sgsum, const GreenDev* g, int general, const double* dscale,
             FftTune& t, hipStream_t s);

int fft_green(const Fields& f, FftCtx* c, const double* sgsum_dev, const double* c0, const double* s0, int igamma,
              int mixed, double scale, const double* dscale, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    GreenDev g;
    for (int i = 0; i < 81; i++) { g.c0[i] = c0[i]; g.s0[i] = s0[i]; }
    g.igamma = igamma;
    if (c->use_reg && fft_reg_supported(f)) {
        const Plan3& pl = c->plan;
        if (mixed)
            return run_green_reg<float>(f, c->tw32 + pl.ax[0].tw_off, c->tw32 + pl.ax[1].tw_off,
                                        c->tw32 + pl.ax[2].tw_off, sgsum_dev, green_dev(c, g, s), igamma != 0, dscale,
                                        c->tune, s, 0, 5, ev0, ev1);
        return run_green_reg<double>(f, c->tw64 + pl.ax[0].tw_off, c->tw64 + pl.ax[1].tw_off,
                                     c->tw64 + pl.ax[2].tw_off, sgsum_dev, green_dev(c, g, s), igamma != 0, dscale,
                                     c->tune, s, 0, 5, ev0, ev1);
    }
    if (mixed) return run_green<float>(f, c, sgsum_dev, g, scale, s);
    return run_green<double>(f, c, sgsum_dev, g, 1.0, s);
}

int fft_tune(const Fields& f, FftCtx* c, const double* sgsum_dev, const double* c0, const double* s0, int igamma,
             int mixed, const double* dscale, hipStream_t s) {
    if (!(c->use_reg && fft_reg_supported(f))) return 0;  // the LDS kernels have no shapes to pick
    GreenDev g;
    for (int i = 0; i < 81; i++) { g.c0[i] = c0[i]; g.s0[i] = s0[i]; }
    g.igamma = igamma;
    const Plan3& pl = c->plan;
    if (mixed)
        return tune_reg<float>(f, c->tw32 + pl.ax[0].tw_off, c->tw32 + pl.ax[1].tw_off, c->tw32 + pl.ax[2].tw_off,
                               sgsum_dev, green_dev(c, g, s), igamma != 0, dscale, c->tune, s);
    return tune_reg<double>(f, c->tw64 + pl.ax[0].tw_off, c->tw64 + pl.ax[1].tw_off, c->tw64 + pl.ax[2].tw_off,
                            sgsum_dev, green_dev(c, g, s), igamma != 0, dscale, c->tune, s);
}
int fft_green_passes(const Fields& f, FftCtx* c, const double* sgsum_dev, const double* c0, const double* s0,
                     int igamma, int mixed, const double* dscale, int st0, int st1, hipStream_t s, hipEvent_t ev0,
                     hipEvent_t ev1) {
    if (!(c->use_reg && fft_reg_supported(f))) return -1;
    GreenDev g;
    for (int i = 0; i < 81; i++) { g.c0[i] = c0[i]; g.s0[i] = s0[i]; }
    g.igamma = igamma;
    const Plan3& pl = c->plan;
    if (mixed)
        return run_green_reg<float>(f, c->tw32 + pl.ax[0].tw_off, c->tw32 + pl.ax[1].tw_off,
                                    c->tw32 + pl.ax[2].tw_off, sgsum_dev, green_dev(c, g, s), igamma != 0, dscale,
                                    c->tune, s, st0, st1, ev0, ev1);
    return run_green_reg<double>(f, c->tw64 + pl.ax[0].tw_off, c->tw64 + pl.ax[1].tw_off,
                                 c->tw64 + pl.ax[2].tw_off, sgsum_dev, green_dev(c, g, s), igamma != 0, dscale,
                                 c->tune, s, st0, st1, ev0, ev1);
}
int fft_green_batch(const FftWork* w, int nrve, const Fields& f0, int mixed, int general, hipStream_t s);
void fft_twiddles(const FftCtx* c, int mixed, const void** tw) {
    const Plan3& pl = c->plan;
    for (int k = 0; k < 3; k++)
        tw[k] = mixed ? (const void*)(c->tw32 + pl.ax[k].tw_off) : (const void*)(c->tw64 + pl.ax[k].tw_off);
}
FftTune fft_get_tune(const FftCtx* c) { return c->tune; }
void fft_set_tune(FftCtx* c, const FftTune& t) { c->tune = t; }
bool fft_tunable(const Fields& f, const FftCtx* c) { return c->use_reg && fft_reg_supported(f); }

}  // namespace cg

namespace cg {
// ---------------------------------------------------------------------------
// spectral integration of L F for the nodal coordinates
// (update_grid_velgrad_node :1970-2056, integrate_tensor2 :359-429):
// u_l = sum_m conj(k_m) T_lm / |k|^2 with k = i xi (ideriv 0), i sin(xi) (1) or the
// rotated operator (2).  The reference keeps Re() of a c2c inverse, i.e. the Hermitian
// part of the spectrum; on the Nyquist planes k(-xi) is not -k(xi), so the operator
// is averaged over xi and its mirror.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(FBS) void k_x_fwd_lf(Fields f, Plan3 p, int rows) {
    using T2 = double2;
    extern __shared__ __align__(16) unsigned char smem[];
    const int n1 = f.n1, h1 = f.h1, ss = n1 + 1, nsig = rows * 9;
    T2* b0 = (T2*)smem;
    T2* b1 = b0 + nsig * ss;
    const long V = f.V;
    const long row0 = (long)blockIdx.x * rows;
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int e = threadIdx.x; e < rows * n1; e += blockDim.x) {
        const int r = e / n1, i = e - r * n1;
        const long v = (row0 + r) * n1 + i;
        double L[9], F[9], q[9];
        for (int c = 0; c < 9; c++) { L[c] = f.L[c * f.ld + v]; F[c] = f.F[c * f.ld + v]; }
        mm3(L, F, q);
        for (int c = 0; c < 9; c++) {
            b0[(r * 9 + c) * ss + i] = make_double2(q[c], 0.0);
            acc[c] += q[c] * (1.0 / (double)V);
        }
    }
    __syncthreads();
    T2* res = lds_fft<double>(b0, b1, nsig, ss, p, 0, false);
    for (int e = threadIdx.x; e < rows * h1 * 9; e += blockDim.x) {
        const int r = e / (h1 * 9), q = e - r * h1 * 9;
        const int kx = q / 9, c = q - kx * 9;
        f.spec[((row0 + r) * h1 + kx) * 9 + c] = res[(r * 9 + c) * ss + kx];
    }
    block_reduce_store<9, FBS>(acc, f.part);
}

__device__ void kmod_at(int ideriv, int kx, int ky, int kz, int n1, int n2, int n3, double* kr, double* ki) {
    const double tp = 2.0 * PI_D;
    double xi[3];
    xi[0] = tp * (double)((kx + 1) <= n1 / 2 ? kx : kx - n1) / n1;
    xi[1] = tp * (double)((ky + 1) <= n2 / 2 ? ky : ky - n2) / n2;
    xi[2] = (n3 == 1) ? 0.0 : tp * (double)((kz + 1) <= n3 / 2 ? kz : kz - n3) / n3;
    if (ideriv == 2) {
        double pr = 1.0, pim = 0.0;
        for (int l = 0; l < 3; l++) {
            const double cr = 1.0 + cos(xi[l]), ci = sin(xi[l]);
            const double nr = pr * cr - pim * ci, ni = pr * ci + pim * cr;
            pr = nr; pim = ni;
        }
        const double fr = -0.25 * pim, fi = 0.25 * pr;  // i/4 * (pr + i pim)
        for (int l = 0; l < 3; l++) {
            const double t = tan(xi[l] * 0.5);
            kr[l] = t * fr;
            ki[l] = t * fi;
        }
    } else {
        for (int l = 0; l < 3; l++) {
            kr[l] = 0.0;
            ki[l] = ideriv == 0 ? xi[l] : sin(xi[l]);
        }
    }
}

__global__ __launch_bounds__(FBS) void k_z_integrate(Fields f, Plan3 p, int T, int ideriv) {
    using T2 = double2;
    extern __shared__ __align__(16) unsigned char smem[];
    const int n1 = f.n1, n2 = f.n2, n3 = f.n3, h1 = f.h1;
    const int ntile = (h1 + T - 1) / T;
    const int j = blockIdx.x / ntile, tile = blockIdx.x - j * ntile;
    const int kx0 = tile * T;
    const int Te = min(T, h1 - kx0);
    const int nsig = Te * 9, ss = n3 + 1;
    T2* b0 = (T2*)smem;
    T2* b1 = b0 + T * 9 * ss;
    const int w = Te * 9;
    for (int e = threadIdx.x; e < n3 * w; e += blockDim.x) {
        const int k = e / w, q = e - k * w;
        b0[q * ss + k] = f.spec[(((long)k * n2 + j) * h1 + kx0) * 9 + q];
    }
    __syncthreads();
    T2* res = lds_fft<double>(b0, b1, nsig, ss, p, 2, false);
    for (int e = threadIdx.x; e < Te * n3; e += blockDim.x) {
        const int t = e / n3, kz = e - t * n3;
        const int kx = kx0 + t;
        double Kr[3] = {0, 0, 0}, Ki[3] = {0, 0, 0};
        for (int side = 0; side < 2; side++) {
            double kr[3], ki[3];
            if (side == 0) kmod_at(ideriv, kx, j, kz, n1, n2, n3, kr, ki);
            else kmod_at(ideriv, (n1 - kx) % n1, (n2 - j) % n2, (n3 - kz) % n3, n1, n2, n3, kr, ki);
            const double k2 = kr[0] * kr[0] + ki[0] * ki[0] + kr[1] * kr[1] + ki[1] * ki[1] + kr[2] * kr[2] + ki[2] * ki[2];
            if (k2 > 0.0)
                for (int m = 0; m < 3; m++) {  // side 0: conj(k)/|k|^2 ; side 1 (mirror): k(-xi)/|k(-xi)|^2
                    Kr[m] += 0.5 * kr[m] / k2;
                    Ki[m] += 0.5 * (side == 0 ? -ki[m] : ki[m]) / k2;
                }
        }
        const bool allnyq = (n1 % 2 == 0 && kx == n1 / 2) && (n2 % 2 == 0 && j == n2 / 2) && (n3 % 2 == 0 && kz == n3 / 2) &&
                            n3 > 1;
        T2* P = res + (t * 9) * ss + kz;
        double ur[3], ui[3];
        for (int l = 0; l < 3; l++) {
            double sr = 0.0, si = 0.0;
            for (int m = 0; m < 3; m++) {
                const T2 z = P[(3 * l + m) * ss];
                sr += Kr[m] * z.x - Ki[m] * z.y;
                si += Kr[m] * z.y + Ki[m] * z.x;
            }
            ur[l] = (ideriv == 2 && allnyq) ? 0.0 : sr;
            ui[l] = (ideriv == 2 && allnyq) ? 0.0 : si;
        }
        for (int l = 0; l < 3; l++) P[l * ss] = make_double2(ur[l], ui[l]);
    }
    __syncthreads();
    T2* other = (res == b0) ? b1 : b0;
    res = lds_fft<double>(res, other, nsig, ss, p, 2, true);
    for (int e = threadIdx.x; e < n3 * w; e += blockDim.x) {
        const int k = e / w, q = e - k * w;
        if (q % 9 < 3) f.spec[(((long)k * n2 + j) * h1 + kx0) * 9 + q] = res[q * ss + k];
    }
}

__global__ __launch_bounds__(FBS) void k_x_inv3(Fields f, Plan3 p, int rows, double* __restrict__ dv) {
    using T2 = double2;
    extern __shared__ __align__(16) unsigned char smem[];
    const int n1 = f.n1, h1 = f.h1, ss = n1 + 1, nsig = rows * 3;
    T2* b0 = (T2*)smem;
    T2* b1 = b0 + nsig * ss;
    const long V = f.V;
    const long row0 = (long)blockIdx.x * rows;
    for (int e = threadIdx.x; e < rows * h1 * 3; e += blockDim.x) {
        const int r = e / (h1 * 3), q = e - r * h1 * 3;
        const int kx = q / 3, c = q - kx * 3;
        const T2 z = f.spec[((row0 + r) * h1 + kx) * 9 + c];
        T2* s = b0 + (r * 3 + c) * ss;
        s[kx] = z;
        if (kx > 0 && kx < n1 - kx) s[n1 - kx] = make_double2(z.x, -z.y);
    }
    __syncthreads();
    T2* res = lds_fft<double>(b0, b1, nsig, ss, p, 0, true);
    for (int e = threadIdx.x; e < rows * n1; e += blockDim.x) {
        const int r = e / n1, i = e - r * n1;
        const long v = (row0 + r) * n1 + i;
        for (int c = 0; c < 3; c++) dv[c * V + v] = res[(r * 3 + c) * ss + i].x / (double)V;
    }
}

static int integrate_reg(const Fields& f, const double2* const* tw, const FftTune& t, int ideriv, double* dv,
                         hipStream_t s);
int fft_integrate(const Fields& f, FftCtx* c, int ideriv, double* dv, hipStream_t s, int* cm) {
    if (fft_tunable(f, c)) {  // the register passes, partial sums column-major
        const void* tw[3];
        fft_twiddles(c, 0, tw);
        const double2* t2[3] = {(const double2*)tw[0], (const double2*)tw[1], (const double2*)tw[2]};
        *cm = 1;
        return integrate_reg(f, t2, c->tune, ideriv, dv, s);
    }
    *cm = 0;
    const int rows = c->rows;
    const int nrow = f.n2 * f.n3 / rows;
    const size_t shx = 2 * (size_t)rows * 9 * (f.n1 + 1) * sizeof(double2);
    const size_t shy = 2 * (size_t)c->ty * 9 * (f.n2 + 1) * sizeof(double2);
    const size_t shz = 2 * (size_t)c->tz * 9 * (f.n3 + 1) * sizeof(double2);
    const int ty_blocks = f.n3 * ((f.h1 + c->ty - 1) / c->ty);
    const int tz_blocks = f.n2 * ((f.h1 + c->tz - 1) / c->tz);
    hipLaunchKernelGGL(k_x_fwd_lf, dim3(nrow), dim3(FBS), shx, s, f, c->plan, rows);
    if (f.n2 > 1) hipLaunchKernelGGL(k_y<double>, dim3(ty_blocks), dim3(FBS), shy, s, f, c->plan, c->ty, 0, 9);
    hipLaunchKernelGGL(k_z_integrate, dim3(tz_blocks), dim3(FBS), shz, s, f, c->plan, c->tz, ideriv);
    if (f.n2 > 1) hipLaunchKernelGGL(k_y<double>, dim3(ty_blocks), dim3(FBS), shy, s, f, c->plan, c->ty, 1, 3);
    hipLaunchKernelGGL(k_x_inv3, dim3(nrow), dim3(FBS), shx, s, f, c->plan, rows, dv);
    return nrow;
}

}  // namespace cg

// ===========================================================================
// Register-FFT form of the FFT + Green stage (axis lengths with a Shape<N>):
// every transform is level-1 DFT in registers -> one LDS exchange -> level-2 DFT.
// Same five passes and half-spectrum layout as above; two real components share one
// complex transform in the x passes.
// ===========================================================================

namespace cg {

constexpr int YBS = 256;  // threads per block of the y passes (the tuner may halve it)


template <int N> struct RegDims {
    static constexpr int T = Shape<N>::T, E = N / T, U = (E + T - 1) / T;
    static constexpr int AN = U * T > E ? U * T : E;  // register slots per thread
    static constexpr int NS = YBS / T;                 // signals per y block of YBS threads
    // z pass: KT whole kx columns of 9 components per block, 24 signals' worth of threads
    // (216) for T = 4 and 8, one column (144 threads) for T = 16
    static constexpr int KT = T >= 16 ? 1 : 24 / T;    // kx columns per z block
    static constexpr int NSZ = 9 * KT;                 // signals per z block
    static constexpr int ZB = NSZ * T;                 // threads per z block
    static_assert(KT >= 1, "z block shape");
};
// y pass with YB threads per block (YBS or YBS / 2)
template <int N, int YB> struct YDims {
    static constexpr int T = Shape<N>::T, NS = YB / T;
};
// z pass with KT kx columns per block, or one column (KT1)
template <int N, bool KT1> struct ZDims {
    static constexpr int T = Shape<N>::T;
    static constexpr int KT = KT1 ? 1 : RegDims<N>::KT;
    static constexpr int NSZ = 9 * KT, ZB = NSZ * T;
};
// x passes: XB threads per block, RW rows of 9 components (5 signals a row).  FP64 rows run
// in one-wave blocks (more blocks per CU, and their barriers hold one wave); FP32 rows and
// T = 16 transforms in two-wave blocks.
template <int N, typename R> struct XDims {
    static constexpr int T = Shape<N>::T;
    static constexpr int XB = (T >= 16 || sizeof(R) < 8) ? 128 : 64;
    static constexpr int NSX = XB / T;  // signals per block
    static constexpr int RW = NSX / 5;  // rows per block
    static_assert(RW >= 1, "x block shape");
};

// Block b of a launch runs on XCD b % 8.  With xcd set, a pass gives each XCD a contiguous eighth
// of its tiles or rows, so neighbouring tiles (which share the 128-byte lines at their edges) meet
// in the same L2 (3% off the FP64 stage at 128^3, 5% off the FP32 one); else tiles go in launch
// order.  The tuner picks it per pass; the results do not depend on it.
#ifndef CG_XPAD
#define CG_XPAD 1
#endif
#ifndef CG_Y_HX
#define CG_Y_HX 1
#endif
#ifndef CG_XF_ATTR
#define CG_XF_ATTR
#endif
#ifndef CG_XI_ATTR
#define CG_XI_ATTR
#endif
#ifndef CG_Y_ATTR
#define CG_Y_ATTR
#endif
#ifndef CG_Z_ATTR
#define CG_Z_ATTR
#endif
__device__ __forceinline__ int fft_block(int xcd) {
    const int b = blockIdx.x, nb = gridDim.x;
    if (!xcd) return b;
    const int q = nb / 8;
    return b < 8 * q ? (b % 8) * q + b / 8 : b;
}
// The batched passes run over FftWork entries: blocks [r nbs, (r+1) nbs) of a launch work for
// entry r (one per RVE); an entry's blocks do what a context's own launch does (the same block
// functions), with xcd placed as fft_block places a launch's blocks.
__device__ __forceinline__ int entry_block(int nbs, int xcd, int* r) {
    *r = __builtin_amdgcn_readfirstlane(blockIdx.x / nbs);
    const int b = blockIdx.x - *r * nbs;
    if (!xcd) return b;
    const int q = nbs / 8;
    return b < 8 * q ? (b % 8) * q + b / 8 : b;
}

template <typename R> __device__ __forceinline__ typename Cx<R>::T* spec_of(const Fields& f) {
    return (typename Cx<R>::T*)(sizeof(R) == 8 ? (void*)f.spec : (void*)f.specf);
}

// x forward: PK1 = J (sigma - <sigma>) F^-T (LS-EVPFFT.f90:256-269), r2c along x.
// Components (2p, 2p+1) of a row form one complex signal; the two half spectra are
// separated with X[N-k] afterwards.
// LF (the nodal integration, update_grid_velgrad_node :1970-2056): the transformed field is
// L F instead, and the block's sums of L F / V go to f.part column-major (nbs blocks).
template <int N, typename R, bool LF = false>
__device__ __forceinline__ void x_fwd_blk(const Fields& f, const typename Cx<R>::T* __restrict__ tw,
                                          const double* __restrict__ sgsum, const double* __restrict__ dscale, int lb,
                                          int nbs = 0) {
    // 1 / the mixed form's stress scale (1 in FP64)
    const double inv_scale = sizeof(R) == 8 ? 1.0 : dscale[0];
    using T2 = typename Cx<R>::T;
    using D = RegDims<N>;
    using X = XDims<N, R>;
    constexpr int T = D::T, E = D::E, U = D::U, NS = X::NSX, RW = X::RW, XBS = X::XB;
    __shared__ T2 ex[N * NS];
    R* stage = (R*)ex;  // RW*9*N reals, dead before the exchange
    // component rows of the stage are SP = N + CG_XPAD reals apart, so the rows 2p that the five
    // signals of a row read together fall in different LDS banks
    constexpr int SP = N + CG_XPAD;
    static_assert(RW * 9 * SP * sizeof(R) <= N * NS * sizeof(T2), "stage fits the exchange");
    const long V = f.V;
    const long nrows = (long)f.n2 * f.n3;
    const long row0 = (long)lb * RW;  // the block's rows
    double sa[9], lfs[9];
    for (int c = 0; c < 9; c++) {
        sa[c] = LF ? 0.0 : sgsum[c];
        lfs[c] = 0.0;
    }
    // each thread's voxels: every load is issued before the products (rows past the grid
    // read the last row and are zeroed)
    constexpr int NE = (RW * N + XBS - 1) / XBS;
#pragma unroll
    for (int q = 0; q < NE; q++) {
        const int e = threadIdx.x + q * XBS;
        if (RW * N % XBS != 0 && e >= RW * N) break;
        const int r = e / N, i = e - r * N;
        const long row = row0 + r;
        const long v = (row < nrows ? row : nrows - 1) * N + i;
        if constexpr (LF) {
            double Lv[9], Fv[9], lf[9];
            for (int c = 0; c < 9; c++) {
                Lv[c] = f.L[c * f.ld + v];
                Fv[c] = f.F[c * f.ld + v];
            }
            mm3(Lv, Fv, lf);
            const bool in = row < nrows;
            for (int c = 0; c < 9; c++) {
                stage[(r * 9 + c) * SP + i] = in ? (R)lf[c] : (R)0;
                if (in) lfs[c] += lf[c] * (1.0 / (double)V);
            }
            continue;
        }
        double s6[6], s[9], Fi[9], pk[9];
        for (int c = 0; c < 6; c++) s6[c] = f.sg[c * f.ld + v];
        for (int c = 0; c < 9; c++) Fi[c] = f.Fi[c * f.ld + v];
        const double J = row < nrows ? f.J[v] : 0.0;
        from6(s6, s);
        for (int c = 0; c < 9; c++) s[c] -= sa[c];
        mmt3(s, Fi, pk);  // (sigma - <sigma>) F^-T
        for (int c = 0; c < 9; c++) stage[(r * 9 + c) * SP + i] = (R)(pk[c] * J * inv_scale);
    }
    __syncthreads();
    const int s = threadIdx.x % NS, t = threadIdx.x / NS;
    const int r = s / 5, p = s - 5 * r;
    const bool act = r < RW;
    T2 a[D::AN];
#pragma unroll
    for (int m = 0; m < E; m++) {
        const int i = t + T * m;
        R re = 0, im = 0;
        if (act) {
            re = stage[(r * 9 + 2 * p) * SP + i];
            if (p < 4) im = stage[(r * 9 + 2 * p + 1) * SP + i];
        }
        a[m] = T2{re, im};
    }
    __syncthreads();
    fft_dist<N, T, false>(a, s, t, NS, ex, tw);
    __syncthreads();
    // the transformed signals, SX = N + CG_XPAD apart (the nine components of a frequency
    // read five different signals: padded, they fall in different LDS banks)
    constexpr int SX = N + CG_XPAD;
    static_assert(5 * RW * SX <= N * NS, "padded signals fit the exchange");
    if (act)
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int q = t + T * u;
            if (U == 1 || q < E)
#pragma unroll
                for (int j = 0; j < T; j++) ex[s * SX + q + E * j] = a[u * T + j];
        }
    __syncthreads();
    T2* spec = spec_of<R>(f);
    constexpr int w = (N / 2 + 1) * 9;  // h1 = n1/2 + 1 complex values of 9 components a row
#pragma unroll
    for (int q = 0; q < (RW * w + XBS - 1) / XBS; q++) {
        const int e = threadIdx.x + q * XBS;
        if (RW * w % XBS != 0 && e >= RW * w) break;
        const int rr = e / w, rem = e - rr * w;
        const int kx = rem / 9, c = rem - kx * 9;
        const long row = row0 + rr;
        if (row >= nrows) continue;
        const int sp = rr * 5 + c / 2;
        const T2 Z = ex[sp * SX + kx], Zm = ex[sp * SX + (N - kx) % N];
        T2 o;
        if ((c & 1) == 0) o = T2{(R)0.5 * (Z.x + Zm.x), (R)0.5 * (Z.y - Zm.y)};  // (Z + conj Zm)/2
        else o = T2{(R)0.5 * (Z.y + Zm.y), (R)-0.5 * (Z.x - Zm.x)};           // (Z - conj Zm)/(2i)
        spec[row * w + rem] = o;
    }
    if constexpr (LF) block_reduce_store_cm<9, XBS>(lfs, f.part, 0, lb, nbs);
}
template <int N, typename R>
__global__ __launch_bounds__((XDims<N, R>::XB)) CG_XF_ATTR void k_x_fwd_reg(Fields f, const typename Cx<R>::T* __restrict__ tw,
                                                                          const double* __restrict__ sgsum,
                                                                          const double* __restrict__ dscale, int xcd) {
    x_fwd_blk<N, R>(f, tw, sgsum, dscale, fft_block(xcd));
}
template <int N, typename R>
__global__ __launch_bounds__((XDims<N, R>::XB)) void k_x_fwd_bat(const FftWork* __restrict__ w, int nbs) {
    int r;
    const int lb = entry_block(nbs, 0, &r);
    const FftWork& e = w[r];
    const Fields f = e.f;
    x_fwd_blk<N, R>(f, (const typename Cx<R>::T*)e.tw[0], e.sgsum, e.dscale, lb);
}

// y pass over every (kx, c) column of one z plane.  With two level-2 rounds per thread
// (U == 2: lengths 32, 96 and 128) the exchange runs in two rounds through half the LDS and each
// round's outputs are stored at once (fft_dist2), so more blocks fit a CU.
template <int N, bool INV, typename R, int YB>
__device__ __forceinline__ void y_blk(const Fields& f, const typename Cx<R>::T* __restrict__ tw, int ntile, int bid) {
    using T2 = typename Cx<R>::T;
    using D = RegDims<N>;
    constexpr int T = D::T, E = D::E, U = D::U, NS = YDims<N, YB>::NS;
    constexpr bool HX = CG_Y_HX && U == 2;
    __shared__ T2 ex[HX ? T * T * NS : N * NS];
    const int k = bid / ntile, tile = bid - k * ntile;
    const int ncol = f.h1 * 9;
    const int s = threadIdx.x % NS, t = threadIdx.x / NS;
    const int col = tile * NS + s;
    const bool act = col < ncol;
    T2* base = spec_of<R>(f) + (long)k * N * ncol + col;
    T2 a[D::AN];
#pragma unroll
    for (int m = 0; m < E; m++) a[m] = act ? base[(long)(t + T * m) * ncol] : T2{0, 0};
    if constexpr (HX) {
        fft_dist2<N, T, INV>(a, s, t, NS, ex, tw, [&](int, int q, const T2* d) {
            if (act)
#pragma unroll
                for (int j = 0; j < T; j++) base[(long)(q + E * j) * ncol] = d[j];
        });
        return;
    }
    fft_dist<N, T, INV>(a, s, t, NS, ex, tw);
    if (act)
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int q = t + T * u;
            if (U == 1 || q < E)
#pragma unroll
                for (int j = 0; j < T; j++) base[(long)(q + E * j) * ncol] = a[u * T + j];
        }
}
template <int N, bool INV, typename R, int YB>
__global__ __launch_bounds__(YB) CG_Y_ATTR void k_y_reg(Fields f, const typename Cx<R>::T* __restrict__ tw, int ntile, int xcd) {
    y_blk<N, INV, R, YB>(f, tw, ntile, fft_block(xcd));
}
template <int N, bool INV, typename R, int YB>
__global__ __launch_bounds__(YB) CG_Y_ATTR void k_y_bat(const FftWork* __restrict__ w, int ntile, int nbs) {
    int r;
    const int lb = entry_block(nbs, 0, &r);
    const FftWork& e = w[r];
    const Fields f = e.f;
    y_blk<N, INV, R, YB>(f, (const typename Cx<R>::T*)e.tw[1], ntile, lb);
}

// The integration operator of the nodal update at one frequency (integrate_tensor2
// :359-429 with the derivative operator of update_grid_velgrad_node): u_l = sum_m K_m T_lm
// with K = conj(k)/|k|^2 averaged with the mirror frequency's k(-xi)/|k(-xi)|^2 (on the
// Nyquist planes k(-xi) is not -k(xi)), k = i xi (ideriv 0), i sin(xi) (1) or the rotated
// operator (2, zero at the all-Nyquist point); u_l replaces component l < 3.  P: the nine
// components T_lm (3l + m) of the point, st apart.
template <typename T2>
__device__ __forceinline__ void integ_apply(int ideriv, int kx, int ky, int kz, int n1, int n2, int n3, T2* P, int st) {
    double Kr[3] = {0, 0, 0}, Ki[3] = {0, 0, 0};
    for (int side = 0; side < 2; side++) {
        double kr[3], ki[3];
        if (side == 0) kmod_at(ideriv, kx, ky, kz, n1, n2, n3, kr, ki);
        else kmod_at(ideriv, (n1 - kx) % n1, (n2 - ky) % n2, (n3 - kz) % n3, n1, n2, n3, kr, ki);
        const double k2 = kr[0] * kr[0] + ki[0] * ki[0] + kr[1] * kr[1] + ki[1] * ki[1] + kr[2] * kr[2] + ki[2] * ki[2];
        if (k2 > 0.0)
            for (int m = 0; m < 3; m++) {
                Kr[m] += 0.5 * kr[m] / k2;
                Ki[m] += 0.5 * (side == 0 ? -ki[m] : ki[m]) / k2;
            }
    }
    const bool allnyq = (n1 % 2 == 0 && kx == n1 / 2) && (n2 % 2 == 0 && ky == n2 / 2) && (n3 % 2 == 0 && kz == n3 / 2) &&
                        n3 > 1;
    double ur[3], ui[3];
    for (int l = 0; l < 3; l++) {
        double sr = 0.0, si = 0.0;
        for (int m = 0; m < 3; m++) {
            const T2 z = P[(3 * l + m) * st];
            sr += Kr[m] * z.x - Ki[m] * z.y;
            si += Kr[m] * z.y + Ki[m] * z.x;
        }
        ur[l] = (ideriv == 2 && allnyq) ? 0.0 : sr;
        ui[l] = (ideriv == 2 && allnyq) ? 0.0 : si;
    }
    for (int l = 0; l < 3; l++) P[l * st] = T2{ur[l], ui[l]};
}

// z pass with the Green operator: KT whole kx columns (9 components each) of one y row.
// INTEG: the nodal integration's operator (integ_apply) in place of Gamma.
template <int N, typename R, bool GENERAL, bool KT1, bool INTEG = false>
__device__ __forceinline__ void z_blk(const Fields& f, const typename Cx<R>::T* __restrict__ tw, const GreenDev* g,
                                      int ntile, int bid, int ideriv = 0) {
    using T2 = typename Cx<R>::T;
    using D = RegDims<N>;
    using Z = ZDims<N, KT1>;
    constexpr int T = D::T, E = D::E, U = D::U, NS = Z::NSZ, KT = Z::KT;
    __shared__ T2 ex[N * NS];
    const int h1 = f.h1;
    const int j = bid / ntile, tile = bid - j * ntile;
    const int kx0 = tile * KT;
    const int KTe = min(KT, h1 - kx0);
    const int s = threadIdx.x % NS, t = threadIdx.x / NS;
    const bool act = s < 9 * KTe;
    const long ps = (long)f.n2 * h1 * 9;
    T2* base = spec_of<R>(f) + ((long)j * h1 + kx0) * 9 + s;
    T2 a[D::AN];
#pragma unroll
    for (int m = 0; m < E; m++) a[m] = act ? base[(long)(t + T * m) * ps] : T2{0, 0};
    fft_dist<N, T, false>(a, s, t, NS, ex, tw);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int q = t + T * u;
        if (U == 1 || q < E)
#pragma unroll
            for (int jj = 0; jj < T; jj++) ex[(q + E * jj) * NS + s] = a[u * T + jj];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < KTe * N; e += Z::ZB) {
        const int kz = e / KTe, kxl = e - kz * KTe;
        if constexpr (INTEG) integ_apply(ideriv, kx0 + kxl, j, kz, f.n1, f.n2, N, ex + kz * NS + kxl * 9, 1);
        else green_apply<R, GENERAL>(*g, kx0 + kxl, j, kz, f.n1, f.n2, N, ex + kz * NS + kxl * 9, 1);
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < E; m++) a[m] = ex[(t + T * m) * NS + s];
    __syncthreads();
    fft_dist<N, T, true>(a, s, t, NS, ex, tw);
    if (act)
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int q = t + T * u;
            if (U == 1 || q < E)
#pragma unroll
                for (int jj = 0; jj < T; jj++) base[(long)(q + E * jj) * ps] = a[u * T + jj];
        }
}
template <int N, typename R, bool GENERAL, bool KT1>
__global__ __launch_bounds__((ZDims<N, KT1>::ZB)) CG_Z_ATTR void k_z_green_reg(Fields f, const typename Cx<R>::T* __restrict__ tw,
                                                                          const GreenDev* __restrict__ g, int ntile,
                                                                          int xcd) {
    z_blk<N, R, GENERAL, KT1>(f, tw, g, ntile, fft_block(xcd));
}
template <int N, typename R, bool GENERAL, bool KT1>
__global__ __launch_bounds__((ZDims<N, KT1>::ZB)) CG_Z_ATTR void k_z_bat(const FftWork* __restrict__ w, int ntile, int nbs) {
    int r;
    const int lb = entry_block(nbs, 0, &r);
    const FftWork& e = w[r];
    const Fields f = e.f;
    z_blk<N, R, GENERAL, KT1>(f, (const typename Cx<R>::T*)e.tw[2], &e.g, ntile, lb);
}

// x inverse (c2r): dLref = Re(IFFT)/V with the Hermitian continuation of the half
// spectrum (imaginary parts at kx = 0 and n1/2 dropped, as Re() of the reference's
// c2c inverse does); block partial sums of w dLref F^-1.
// LF (the nodal integration): components 0..2 of the result, /V, go to dv[c V + v]; no sums.
template <int N, typename R, bool LF = false>
__device__ __forceinline__ void x_inv_blk(const Fields& f, const typename Cx<R>::T* __restrict__ tw,
                                          const double* __restrict__ dscale, int lb, int nbs,
                                          double* __restrict__ dv = nullptr) {
    const double out_scale = sizeof(R) == 8 ? 1.0 / (double)f.V : dscale[1];  // stress scale / V
    using T2 = typename Cx<R>::T;
    using D = RegDims<N>;
    using X = XDims<N, R>;
    constexpr int T = D::T, E = D::E, U = D::U, NS = X::NSX, RW = X::RW, XBS = X::XB;
    constexpr int H = N / 2 + 1;
    __shared__ T2 ex[N * NS];
    static_assert(RW * H * 9 <= N * NS, "spectrum rows fit the exchange");
    constexpr int SP = N + CG_XPAD;  // component rows of the output stage (as in the x forward pass)
    static_assert(RW * 9 * SP * sizeof(R) <= N * NS * sizeof(T2), "output rows fit the exchange");
    const long V = f.V;
    const long nrows = (long)f.n2 * f.n3;
    const long row0 = (long)lb * RW;  // the block's rows (and its partial-sum slot of nbs)
    constexpr int w = H * 9;
    const T2* spec = spec_of<R>(f);
    // the block's spectrum rows: every load is issued before the first is written to LDS
    // (rows past the grid read element 0 and are zeroed)
    constexpr int NL = (RW * w + XBS - 1) / XBS;
    T2 sv[NL];
#pragma unroll
    for (int q = 0; q < NL; q++) {
        const int e = threadIdx.x + q * XBS;
        const int ec = RW * w % XBS == 0 || e < RW * w ? e : RW * w - 1;
        sv[q] = spec[row0 + ec / w < nrows ? row0 * w + ec : 0];
    }
    // then F^-1 and w of this thread's output voxels: their loads are in flight through the
    // transform (rows past the grid read the last row)
    constexpr int NE = (RW * N + XBS - 1) / XBS;
    double Fp[NE][9], wp[NE];
#pragma unroll
    for (int q = 0; q < NE; q++) {
        const int e = threadIdx.x + q * XBS;
        if (RW * N % XBS != 0 && e >= RW * N) break;
        if (LF) break;
        const int rr = e / N, i = e - rr * N;
        const long row = row0 + rr;
        const long v = (row < nrows ? row : nrows - 1) * N + i;
        for (int c = 0; c < 9; c++) Fp[q][c] = f.Fi[c * f.ld + v];
        wp[q] = row < nrows ? f.w[v] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < NL; q++) {
        const int e = threadIdx.x + q * XBS;
        if (RW * w % XBS == 0 || e < RW * w) ex[e] = row0 + e / w < nrows ? sv[q] : T2{0, 0};
    }
    __syncthreads();
    const int s = threadIdx.x % NS, t = threadIdx.x / NS;
    const int r = s / 5, p = s - 5 * r;
    const bool act = r < RW;
    T2 a[D::AN];
#pragma unroll
    for (int m = 0; m < E; m++) {
        const int k = t + T * m;
        const int kk = k <= N / 2 ? k : N - k;
        T2 A = T2{0, 0}, B = T2{0, 0};
        if (act) {
            A = ex[(r * H + kk) * 9 + 2 * p];
            if (p < 4) B = ex[(r * H + kk) * 9 + 2 * p + 1];
            if (kk == 0 || 2 * kk == N) { A.y = 0; B.y = 0; }
            if (k > N / 2) { A.y = -A.y; B.y = -B.y; }
        }
        a[m] = T2{A.x - B.y, A.y + B.x};  // A + i B
    }
    __syncthreads();
    fft_dist<N, T, true>(a, s, t, NS, ex, tw);
    __syncthreads();
    R* out = (R*)ex;
    if (act)
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int q = t + T * u;
            if (U == 1 || q < E)
#pragma unroll
                for (int jj = 0; jj < T; jj++) {
                    const T2 z = a[u * T + jj];
                    const int n = q + E * jj;
                    out[(r * 9 + 2 * p) * SP + n] = z.x;
                    if (p < 4) out[(r * 9 + 2 * p + 1) * SP + n] = z.y;
                }
        }
    __syncthreads();
    if constexpr (LF) {
#pragma unroll
        for (int q = 0; q < NE; q++) {
            const int e = threadIdx.x + q * XBS;
            if (RW * N % XBS != 0 && e >= RW * N) break;
            const int rr = e / N, i = e - rr * N;
            const long row = row0 + rr;
            if (row < nrows)
                for (int c = 0; c < 3; c++) dv[c * V + row * N + i] = (double)out[(rr * 9 + c) * SP + i] * out_scale;
        }
        return;
    }
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < NE; q++) {
        const int e = threadIdx.x + q * XBS;
        if (RW * N % XBS != 0 && e >= RW * N) break;
        const int rr = e / N, i = e - rr * N;
        const long row = row0 + rr;
        const bool in = row < nrows;
        const long v = (in ? row : nrows - 1) * N + i;
        double d[9], q9[9];
        for (int c = 0; c < 9; c++) {
            d[c] = (double)out[(rr * 9 + c) * SP + i] * out_scale;
            if (in) f.dLref[c * f.ld + v] = d[c];
        }
        mm3(d, Fp[q], q9);
        for (int c = 0; c < 9; c++) acc[c] += q9[c] * wp[q];
    }
    // partials in row order whatever the block placement, so the reduced sums do not depend on it
    block_reduce_store_cm<9, XBS>(acc, f.part, 0, lb, nbs);
}
template <int N, typename R>
__global__ __launch_bounds__((XDims<N, R>::XB)) CG_XI_ATTR void k_x_inv_reg(Fields f, const typename Cx<R>::T* __restrict__ tw,
                                                                          const double* __restrict__ dscale, int xcd) {
    x_inv_blk<N, R>(f, tw, dscale, fft_block(xcd), gridDim.x);
}
template <int N, typename R>
__global__ __launch_bounds__((XDims<N, R>::XB)) void k_x_inv_bat(const FftWork* __restrict__ w, int nbs) {
    int r;
    const int lb = entry_block(nbs, 0, &r);
    const FftWork& e = w[r];
    const Fields f = e.f;
    x_inv_blk<N, R>(f, (const typename Cx<R>::T*)e.tw[0], e.dscale, lb, nbs);
}

static bool reg_size(int n) { return n == 16 || n == 32 || n == 48 || n == 64 || n == 96 || n == 128 || n == 256; }
bool fft_reg_supported(const Fields& f) { return reg_size(f.n1) && reg_size(f.n2) && reg_size(f.n3); }

// e0, e1 (or null): events recorded at the launch's start and end (the stage's timers)
template <int N, typename R> static void launch_x_fwd_reg(const Fields& f, const typename Cx<R>::T* tw, const double* sgsum,
                                                          const double* dscale, int xcd, hipStream_t s, hipEvent_t e0,
                                                          hipEvent_t e1) {
    const long nrows = (long)f.n2 * f.n3;
    using X = XDims<N, R>;
    const int nb = (int)((nrows + X::RW - 1) / X::RW);
    hipExtLaunchKernelGGL((k_x_fwd_reg<N, R>), dim3(nb), dim3(X::XB), 0, s, e0, e1, 0, f, tw, sgsum, dscale, xcd);
}
template <int N, typename R> static int launch_x_inv_reg(const Fields& f, const typename Cx<R>::T* tw, const double* dscale,
                                                         int xcd, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    const long nrows = (long)f.n2 * f.n3;
    using X = XDims<N, R>;
    const int nb = (int)((nrows + X::RW - 1) / X::RW);
    hipExtLaunchKernelGGL((k_x_inv_reg<N, R>), dim3(nb), dim3(X::XB), 0, s, e0, e1, 0, f, tw, dscale, xcd);
    return nb;
}
template <int N, bool INV, typename R, int YB>
static void launch_y_reg_b(const Fields& f, const typename Cx<R>::T* tw, int xcd, hipStream_t s, hipEvent_t e0,
                           hipEvent_t e1) {
    constexpr int NS = YDims<N, YB>::NS;
    const int ntile = (f.h1 * 9 + NS - 1) / NS;
    hipExtLaunchKernelGGL((k_y_reg<N, INV, R, YB>), dim3(f.n3 * ntile), dim3(YB), 0, s, e0, e1, 0, f, tw, ntile, xcd);
}
template <int N, bool INV, typename R>
static void launch_y_reg(const Fields& f, const typename Cx<R>::T* tw, int half, int xcd, hipStream_t s, hipEvent_t e0,
                         hipEvent_t e1) {
    if (half && YBS / 2 >= 2 * Shape<N>::T) launch_y_reg_b<N, INV, R, YBS / 2>(f, tw, xcd, s, e0, e1);
    else launch_y_reg_b<N, INV, R, YBS>(f, tw, xcd, s, e0, e1);
}
template <int N, typename R, bool KT1>
static void launch_z_reg_b(const Fields& f, const typename Cx<R>::T* tw, const GreenDev* g, int general, int xcd,
                           hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    using Z = ZDims<N, KT1>;
    const int ntile = (f.h1 + Z::KT - 1) / Z::KT;
    if (!general)
        hipExtLaunchKernelGGL((k_z_green_reg<N, R, false, KT1>), dim3(f.n2 * ntile), dim3(Z::ZB), 0, s, e0, e1, 0, f, tw, g,
                              ntile, xcd);
    else
        hipExtLaunchKernelGGL((k_z_green_reg<N, R, true, KT1>), dim3(f.n2 * ntile), dim3(Z::ZB), 0, s, e0, e1, 0, f, tw, g,
                              ntile, xcd);
}
template <int N, typename R>
static void launch_z_reg(const Fields& f, const typename Cx<R>::T* tw, const GreenDev* g, int general, int kt1, int xcd,
                         hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    if constexpr (RegDims<N>::KT > 1) {
        if (kt1) {
            launch_z_reg_b<N, R, true>(f, tw, g, general, xcd, s, e0, e1);
            return;
        }
    }
    launch_z_reg_b<N, R, false>(f, tw, g, general, xcd, s, e0, e1);
}

#define CG_SWITCH_N(n, CALL)                                             \
    switch (n) {                                                         \
        case 16: CALL(16); break;                                        \
        case 32: CALL(32); break;                                        \
        case 48: CALL(48); break;                                        \
        case 64: CALL(64); break;                                        \
        case 96: CALL(96); break;                                        \
        case 128: CALL(128); break;                                      \
        case 256: CALL(256); break;                                      \
        default: break;                                                  \
    }

// One stage (0 x forward, 1 y forward, 2 z + Gamma, 3 y inverse, 4 x inverse) of the register
// form with the block shape and placement of t; returns the x inverse's block count.
template <typename R>
static int reg_stage(int stage, const Fields& f, const typename Cx<R>::T* tw1, const typename Cx<R>::T* tw2,
                     const typename Cx<R>::T* tw3, const double* sgsum, const GreenDev* g, int general,
                     const double* dscale, const FftTune& t, hipStream_t s, hipEvent_t e0 = nullptr,
                     hipEvent_t e1 = nullptr) {
    int nb = -1;
#define CG_XF(N) launch_x_fwd_reg<N, R>(f, tw1, sgsum, dscale, t.xcd[0], s, e0, e1)
#define CG_YF(N) launch_y_reg<N, false, R>(f, tw2, t.yhalf[0], t.xcd[1], s, e0, e1)
#define CG_Z(N) launch_z_reg<N, R>(f, tw3, g, general, t.kt1, t.xcd[2], s, e0, e1)
#define CG_YI(N) launch_y_reg<N, true, R>(f, tw2, t.yhalf[1], t.xcd[3], s, e0, e1)
#define CG_XI(N) nb = launch_x_inv_reg<N, R>(f, tw1, dscale, t.xcd[4], s, e0, e1)
    switch (stage) {
        case 0: CG_SWITCH_N(f.n1, CG_XF) break;
        case 1: CG_SWITCH_N(f.n2, CG_YF) break;
        case 2: CG_SWITCH_N(f.n3, CG_Z) break;
        case 3: CG_SWITCH_N(f.n2, CG_YI) break;
        default: CG_SWITCH_N(f.n1, CG_XI) break;
    }
#undef CG_XF
#undef CG_YF
#undef CG_Z
#undef CG_YI
#undef CG_XI
    return nb;
}

// The nodal integration on the register passes (FP64 spectrum): x forward of L F with its
// sums, the y passes of the stage, the z pass with integ_apply, x inverse to dv.
template <int N>
__global__ __launch_bounds__((XDims<N, double>::XB)) void k_x_fwd_lf_reg(Fields f, const double2* __restrict__ tw, int xcd) {
    x_fwd_blk<N, double, true>(f, tw, nullptr, nullptr, fft_block(xcd), gridDim.x);
}
template <int N>
__global__ __launch_bounds__((ZDims<N, false>::ZB)) void k_z_integ_reg(Fields f, const double2* __restrict__ tw, int ntile,
                                                                      int ideriv, int xcd) {
    z_blk<N, double, false, false, true>(f, tw, nullptr, ntile, fft_block(xcd), ideriv);
}
template <int N>
__global__ __launch_bounds__((XDims<N, double>::XB)) void k_x_inv_lf_reg(Fields f, const double2* __restrict__ tw,
                                                                       double* __restrict__ dv, int xcd) {
    x_inv_blk<N, double, true>(f, tw, nullptr, fft_block(xcd), gridDim.x, dv);
}
static int integrate_reg(const Fields& f, const double2* const* tw, const FftTune& t, int ideriv, double* dv,
                         hipStream_t s) {
    int nb = 0;
#define CG_IXF(N)                                                                                      \
    {                                                                                                  \
        using X = XDims<N, double>;                                                                    \
        nb = (int)(((long)f.n2 * f.n3 + X::RW - 1) / X::RW);                                          \
        hipLaunchKernelGGL((k_x_fwd_lf_reg<N>), dim3(nb), dim3(X::XB), 0, s, f, tw[0], t.xcd[0]);      \
    }
#define CG_IYF(N) launch_y_reg<N, false, double>(f, tw[1], t.yhalf[0], t.xcd[1], s, nullptr, nullptr)
#define CG_IZ(N)                                                                                                 \
    {                                                                                                            \
        using Z = ZDims<N, false>;                                                                               \
        const int ntile = (f.h1 + Z::KT - 1) / Z::KT;                                                            \
        hipLaunchKernelGGL((k_z_integ_reg<N>), dim3(f.n2 * ntile), dim3(Z::ZB), 0, s, f, tw[2], ntile, ideriv,   \
                           t.xcd[2]);                                                                            \
    }
#define CG_IYI(N) launch_y_reg<N, true, double>(f, tw[1], t.yhalf[1], t.xcd[3], s, nullptr, nullptr)
#define CG_IXI(N)                                                                                                \
    {                                                                                                            \
        using X = XDims<N, double>;                                                                              \
        const int nbi = (int)(((long)f.n2 * f.n3 + X::RW - 1) / X::RW);                                         \
        hipLaunchKernelGGL((k_x_inv_lf_reg<N>), dim3(nbi), dim3(X::XB), 0, s, f, tw[0], dv, t.xcd[4]);           \
    }
    CG_SWITCH_N(f.n1, CG_IXF)
    CG_SWITCH_N(f.n2, CG_IYF)
    CG_SWITCH_N(f.n3, CG_IZ)
    CG_SWITCH_N(f.n2, CG_IYI)
    CG_SWITCH_N(f.n1, CG_IXI)
#undef CG_IXF
#undef CG_IYF
#undef CG_IZ
#undef CG_IYI
#undef CG_IXI
    return nb;
}

// the batched passes over nrve FftWork entries of grids with f's dimensions (default shapes)
template <int N, typename R> static int bat_x_fwd(const FftWork* w, int nrve, const Fields& f, hipStream_t s) {
    using X = XDims<N, R>;
    const int nbs = (int)(((long)f.n2 * f.n3 + X::RW - 1) / X::RW);
    hipLaunchKernelGGL((k_x_fwd_bat<N, R>), dim3(nbs * nrve), dim3(X::XB), 0, s, w, nbs);
    return nbs;
}
template <int N, typename R> static int bat_x_inv(const FftWork* w, int nrve, const Fields& f, hipStream_t s) {
    using X = XDims<N, R>;
    const int nbs = (int)(((long)f.n2 * f.n3 + X::RW - 1) / X::RW);
    hipLaunchKernelGGL((k_x_inv_bat<N, R>), dim3(nbs * nrve), dim3(X::XB), 0, s, w, nbs);
    return nbs;
}
template <int N, bool INV, typename R> static void bat_y(const FftWork* w, int nrve, const Fields& f, hipStream_t s) {
    constexpr int NS = YDims<N, YBS>::NS;
    const int ntile = (f.h1 * 9 + NS - 1) / NS, nbs = f.n3 * ntile;
    hipLaunchKernelGGL((k_y_bat<N, INV, R, YBS>), dim3(nbs * nrve), dim3(YBS), 0, s, w, ntile, nbs);
}
template <int N, typename R> static void bat_z(const FftWork* w, int nrve, const Fields& f, int general, hipStream_t s) {
    using Z = ZDims<N, false>;
    const int ntile = (f.h1 + Z::KT - 1) / Z::KT, nbs = f.n2 * ntile;
    if (general) hipLaunchKernelGGL((k_z_bat<N, R, true, false>), dim3(nbs * nrve), dim3(Z::ZB), 0, s, w, ntile, nbs);
    else hipLaunchKernelGGL((k_z_bat<N, R, false, false>), dim3(nbs * nrve), dim3(Z::ZB), 0, s, w, ntile, nbs);
}
template <typename R> static int green_batch(const FftWork* w, int nrve, const Fields& f, int general, hipStream_t s) {
    int nb = -1;
#define CG_BXF(N) bat_x_fwd<N, R>(w, nrve, f, s)
#define CG_BYF(N) bat_y<N, false, R>(w, nrve, f, s)
#define CG_BZ(N) bat_z<N, R>(w, nrve, f, general, s)
#define CG_BYI(N) bat_y<N, true, R>(w, nrve, f, s)
#define CG_BXI(N) nb = bat_x_inv<N, R>(w, nrve, f, s)
    CG_SWITCH_N(f.n1, CG_BXF)
    CG_SWITCH_N(f.n2, CG_BYF)
    CG_SWITCH_N(f.n3, CG_BZ)
    CG_SWITCH_N(f.n2, CG_BYI)
    CG_SWITCH_N(f.n1, CG_BXI)
#undef CG_BXF
#undef CG_BYF
#undef CG_BZ
#undef CG_BYI
#undef CG_BXI
    return nb;
}
int fft_green_batch(const FftWork* w, int nrve, const Fields& f0, int mixed, int general, hipStream_t s) {
    if (!fft_reg_supported(f0) || nrve < 1) return -1;
    const int nb = mixed ? green_batch<float>(w, nrve, f0, general, s) : green_batch<double>(w, nrve, f0, general, s);
    if (nb > 0) launch_reduce_fluct_batch(w, nrve, nb, s);
    return nb;
}

template <typename R>
int run_green_reg(const Fields& f, const typename Cx<R>::T* tw1, const typename Cx<R>::T* tw2,
                  const typename Cx<R>::T* tw3, const double* sgsum, const GreenDev* g, int general,
                  const double* dscale, const FftTune& t, hipStream_t s, int st0, int st1, hipEvent_t ev0,
                  hipEvent_t ev1) {
    int nb = -1;
    for (int st = st0; st < st1; st++)
        nb = reg_stage<R>(st, f, tw1, tw2, tw3, sgsum, g, general, dscale, t, s, st == st0 ? ev0 : nullptr,
                          st == st1 - 1 ? ev1 : nullptr);
    return nb;
}

// The autotuner: each stage's candidate shapes (block placement on or off; the y passes with
// 256 or 128 threads a block; the z pass with its default kx columns or one) timed on the
// context's own stream, two warm-up and five timed launches each, the fastest kept.  Every
// candidate computes the same values in the same order, so the choice changes time only.
template <typename R>
int tune_reg(const Fields& f, const typename Cx<R>::T* tw1, const typename Cx<R>::T* tw2,
             const typename Cx<R>::T* tw3, const double* sgsum, const GreenDev* g, int general, const double* dscale,
             FftTune& t, hipStream_t s) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return -1;
    if (hipEventCreate(&e1) != hipSuccess) {
        hipEventDestroy(e0);
        return -1;
    }
    int rc = 0;
    for (int st = 0; st < 5 && !rc; st++) {
        const int nshape = (st == 1 || st == 3 || st == 2) ? 2 : 1;
        float best = 1e30f;
        FftTune keep = t;
        for (int sh = 0; sh < nshape && !rc; sh++)
            for (int x = 1; x >= 0 && !rc; x--) {
                FftTune c = keep;
                c.xcd[st] = x;
                if (st == 1) c.yhalf[0] = sh;
                if (st == 3) c.yhalf[1] = sh;
                if (st == 2) c.kt1 = sh;
                for (int w = 0; w < 2; w++) reg_stage<R>(st, f, tw1, tw2, tw3, sgsum, g, general, dscale, c, s);
                hipEventRecord(e0, s);
                for (int w = 0; w < 5; w++) reg_stage<R>(st, f, tw1, tw2, tw3, sgsum, g, general, dscale, c, s);
                hipEventRecord(e1, s);
                float ms = 0.0f;
                if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                    rc = -2;
                    break;
                }
                if (ms < best) {
                    best = ms;
                    t = c;
                }
            }
        t.us[st] = best * 1000.0f / 5.0f;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    t.tuned = rc ? 0 : 1;
    return rc;
}
template int run_green_reg<double>(const Fields&, const double2*, const double2*, const double2*, const double*,
                                   const GreenDev*, int, const double*, const FftTune&, hipStream_t, int, int, hipEvent_t,
                                   hipEvent_t);
template int run_green_reg<float>(const Fields&, const float2*, const float2*, const float2*, const double*,
                                  const GreenDev*, int, const double*, const FftTune&, hipStream_t, int, int, hipEvent_t,
                                   hipEvent_t);
template int tune_reg<double>(const Fields&, const double2*, const double2*, const double2*, const double*,
                              const GreenDev*, int, const double*, FftTune&, hipStream_t);
template int tune_reg<float>(const Fields&, const float2*, const float2*, const float2*, const double*,
                             const GreenDev*, int, const double*, FftTune&, hipStream_t);

}  // namespace cg
