// tensor.h -- small dense tensor algebra shared by the HIP kernels and the
// library's host-side macroscopic updates (FP64, __host__ __device__).
//
// Conventions: 3x3 tensors row-major a[3*i+j]; symmetric tensors as 6-vectors
// in the orthonormal b-basis of chg_basis (tensor_functions.f90:21-198):
//   b1=(e22-e11)/sqrt2, b2=(2e33-e11-e22)/sqrt6, b3=(e23+e32)/sqrt2,
//   b4=(e13+e31)/sqrt2, b5=(e12+e21)/sqrt2, b6=I/sqrt3.
// 6x6 operators act on b-basis vectors, row-major.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define CG_HD __host__ __device__ __forceinline__

namespace cg {

// 1/x for a positive normal x.  On the GPU: the hardware estimate refined by two Newton-Raphson
// steps r <- r + r (1 - x r), within an ulp or so of the rounded quotient in five instructions
// (the IEEE division takes ten).  CG_FAST_RCP=0: the division.
#ifndef CG_FAST_RCP
#define CG_FAST_RCP 1
#endif
CG_HD double recip(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (CG_FAST_RCP) {
        double r = __builtin_amdgcn_rcp(x);
        double e = fma(-x, r, 1.0);
        r = fma(r, e, r);
        e = fma(-x, r, 1.0);
        return fma(r, e, r);
    }
#endif
    return 1.0 / x;
}

constexpr double R2 = 0.70710678118654752440;   // 1/sqrt2
constexpr double R3 = 0.57735026918962576451;   // 1/sqrt3
constexpr double R6 = 0.40824829046386301637;   // 1/sqrt6
constexpr double R23 = 0.81649658092772603273;  // 2/sqrt6
constexpr double PI_D = 3.14159265358979323846;

// x^y for x >= 0 and a finite y > 0 (the power law's |rss/tau|^(n-1) with a non-integer n), as
// exp(y ln x) in FP64 with no table and no double-double steps (about 45 instructions; the
// library pow keeps extra precision and spills registers in the constitutive kernel):
//   ln x:  x = m 2^e with m in [1/sqrt2, sqrt2), ln m = 2 atanh(s), s = (m-1)/(m+1),
//          |s| <= 0.1716, by its Taylor series through s^23 (the next term is below 2^-60);
//   exp t: t = k ln2 + r, |r| <= ln2/2, k integer, ln2 in two parts (the first with 11 trailing
//          zero bits, so k ln2_hi is exact), e^r by its Taylor series through r^13/13! (the
//          next term is below 2^-57), then scaled by 2^k.
// The result is within a few ulp of x^y for |y ln x| up to ~40 (relative error ~ |y ln x| eps);
// x = 0 gives 0, y ln x < -745.2 gives 0.
CG_HD double xpow(double x, double y) {
    if (!(x > 0.0)) return x == 0.0 ? 0.0 : x;  // 0 -> 0; NaN propagates
    int e;
    double m = frexp(x, &e);  // m in [0.5, 1)
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        e -= 1;
    }
    const double s = (m - 1.0) * recip(m + 1.0);
    const double z = s * s;
    double p = 1.0 / 23.0;
    p = fma(p, z, 1.0 / 21.0);
    p = fma(p, z, 1.0 / 19.0);
    p = fma(p, z, 1.0 / 17.0);
    p = fma(p, z, 1.0 / 15.0);
    p = fma(p, z, 1.0 / 13.0);
    p = fma(p, z, 1.0 / 11.0);
    p = fma(p, z, 1.0 / 9.0);
    p = fma(p, z, 1.0 / 7.0);
    p = fma(p, z, 1.0 / 5.0);
    p = fma(p, z, 1.0 / 3.0);
    const double lnm = 2.0 * s + 2.0 * s * (z * p);
    constexpr double LN2 = 0.69314718055994530942;
    constexpr double LN2_HI = 6.93147180369123816490e-01;  // ln2 rounded to 42 significant bits
    constexpr double LN2_LO = 1.90821492927058770002e-10;  // ln2 - LN2_HI
    const double t = y * fma((double)e, LN2, lnm);
    if (t < -745.2) return 0.0;
    if (t > 709.8) return HUGE_VAL;
    const double k = rint(t * (1.0 / LN2));
    double r = fma(-k, LN2_HI, t);
    r = fma(-k, LN2_LO, r);
    double q = 1.0 / 6227020800.0;  // 1/13!
    q = fma(q, r, 1.0 / 479001600.0);
    q = fma(q, r, 1.0 / 39916800.0);
    q = fma(q, r, 1.0 / 3628800.0);
    q = fma(q, r, 1.0 / 362880.0);
    q = fma(q, r, 1.0 / 40320.0);
    q = fma(q, r, 1.0 / 5040.0);
    q = fma(q, r, 1.0 / 720.0);
    q = fma(q, r, 1.0 / 120.0);
    q = fma(q, r, 1.0 / 24.0);
    q = fma(q, r, 1.0 / 6.0);
    q = fma(q, r, 0.5);
    q = fma(q, r, 1.0);
    q = fma(q, r, 1.0);
    return ldexp(q, (int)k);
}

// ---- b-basis <-> 3x3 ------------------------------------------------------
// 6-vector of the symmetric part of a general 3x3 tensor.
CG_HD void to6(const double* a, double* s) {
    s[0] = R2 * (a[4] - a[0]);
    s[1] = R23 * a[8] - R6 * (a[0] + a[4]);
    s[2] = R2 * (a[5] + a[7]);
    s[3] = R2 * (a[2] + a[6]);
    s[4] = R2 * (a[1] + a[3]);
    s[5] = R3 * (a[0] + a[4] + a[8]);
}
// symmetric 3x3 from a 6-vector (kdim=6) or 5-vector (s[5] taken as 0).
CG_HD void from6(const double* s, double* a, bool dev = false) {
    const double h = dev ? 0.0 : R3 * s[5];
    a[0] = h - R2 * s[0] - R6 * s[1];
    a[4] = h + R2 * s[0] - R6 * s[1];
    a[8] = h + R23 * s[1];
    a[5] = a[7] = R2 * s[2];
    a[2] = a[6] = R2 * s[3];
    a[1] = a[3] = R2 * s[4];
}
// the six basis tensors as 3x3 (for contractions built from them)
CG_HD void basis(int a, double* b) {
    for (int i = 0; i < 9; i++) b[i] = 0.0;
    switch (a) {
        case 0: b[0] = -R2; b[4] = R2; break;
        case 1: b[0] = -R6; b[4] = -R6; b[8] = R23; break;
        case 2: b[5] = b[7] = R2; break;
        case 3: b[2] = b[6] = R2; break;
        case 4: b[1] = b[3] = R2; break;
        default: b[0] = b[4] = b[8] = R3; break;
    }
}

// ---- 3x3 ---------------------------------------------------------------------
CG_HD void mm3(const double* a, const double* b, double* c) {
    double t[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) t[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
#pragma unroll
    for (int i = 0; i < 9; i++) c[i] = t[i];
}
// c = a * b^T
CG_HD void mmt3(const double* a, const double* b, double* c) {
    double t[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) t[3 * i + j] = a[3 * i] * b[3 * j] + a[3 * i + 1] * b[3 * j + 1] + a[3 * i + 2] * b[3 * j + 2];
#pragma unroll
    for (int i = 0; i < 9; i++) c[i] = t[i];
}
CG_HD void transp3(const double* a, double* b) {
    double t[9] = {a[0], a[3], a[6], a[1], a[4], a[7], a[2], a[5], a[8]};
#pragma unroll
    for (int i = 0; i < 9; i++) b[i] = t[i];
}
CG_HD double det3(const double* m) {
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
}
// inverse by the adjugate; returns the determinant (0 -> singular)
CG_HD double inv3(const double* m, double* r) {
    double c0 = m[4] * m[8] - m[5] * m[7], c1 = m[5] * m[6] - m[3] * m[8], c2 = m[3] * m[7] - m[4] * m[6];
    double d = m[0] * c0 + m[1] * c1 + m[2] * c2;
    double id = 1.0 / d;
    double t[9];
    t[0] = c0 * id; t[1] = (m[2] * m[7] - m[1] * m[8]) * id; t[2] = (m[1] * m[5] - m[2] * m[4]) * id;
    t[3] = c1 * id; t[4] = (m[0] * m[8] - m[2] * m[6]) * id; t[5] = (m[2] * m[3] - m[0] * m[5]) * id;
    t[6] = c2 * id; t[7] = (m[1] * m[6] - m[0] * m[7]) * id; t[8] = (m[0] * m[4] - m[1] * m[3]) * id;
#pragma unroll
    for (int i = 0; i < 9; i++) r[i] = t[i];
    return d;
}
// von Mises equivalents of a general tensor (tensor_functions.f90:266-289)
CG_HD double vm_sq(const double* d) {
    double h = (d[0] + d[4] + d[8]) * (1.0 / 3.0);
    return (d[0] - h) * (d[0] - h) + (d[4] - h) * (d[4] - h) + (d[8] - h) * (d[8] - h) + d[1] * d[1] + d[2] * d[2] +
           d[5] * d[5] + d[3] * d[3] + d[6] * d[6] + d[7] * d[7];
}
CG_HD double vm_strain(const double* d) { return sqrt(vm_sq(d) * (2.0 / 3.0)); }
CG_HD double vm_stress(const double* d) { return sqrt(vm_sq(d) * 1.5); }
// the mixed form's stress scale: max(von Mises stress of the previous pass' solid stress sum, 1)
CG_HD double mixed_scale(const double* sgsum_solid) { return fmax(vm_stress(sgsum_solid), 1.0); }
// von Mises strain of a symmetric tensor given in the b-basis: the deviatoric
// 5-vector norm (the b-basis is orthonormal): sqrt(2/3 |dev|^2)
CG_HD double vm_strain6(const double* s) {
    return sqrt((2.0 / 3.0) * (s[0] * s[0] + s[1] * s[1] + s[2] * s[2] + s[3] * s[3] + s[4] * s[4]));
}

// ---- 6x6 dense ---------------------------------------------------------------
// Gaussian elimination with partial pivoting; solves A x = b in place (A destroyed).
// The pivot row is brought up by compare-and-swap over the rows below it, so every
// index is a compile-time constant after unrolling (the matrix stays in registers);
// the pivot is the first row of largest magnitude, as with an arg-max search.
// Returns false for an exactly singular pivot.
CG_HD bool solve6(double* A, double* b) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 6; k++) {
#pragma unroll
        for (int i = k + 1; i < 6; i++) {
            const bool sw = fabs(A[6 * i + k]) > fabs(A[6 * k + k]);
#pragma unroll
            for (int j = k; j < 6; j++) {
                const double t = A[6 * k + j];
                A[6 * k + j] = sw ? A[6 * i + j] : t;
                A[6 * i + j] = sw ? t : A[6 * i + j];
            }
            const double t = b[k];
            b[k] = sw ? b[i] : t;
            b[i] = sw ? t : b[i];
        }
        if (A[6 * k + k] == 0.0) ok = false;
        const double r = 1.0 / A[6 * k + k];
#pragma unroll
        for (int i = k + 1; i < 6; i++) {
            const double l = A[6 * i + k] * r;
#pragma unroll
            for (int j = k + 1; j < 6; j++) A[6 * i + j] -= l * A[6 * k + j];
            b[i] -= l * b[k];
        }
    }
#pragma unroll
    for (int i = 5; i >= 0; i--) {
        double s = b[i];
#pragma unroll
        for (int j = i + 1; j < 6; j++) s -= A[6 * i + j] * b[j];
        b[i] = s / A[6 * i + i];
    }
    return ok;
}
// ---- symmetric 6x6, packed upper triangle (21 entries, row by row) ----------
CG_HD constexpr int sp6(int i, int j) {
    return i <= j ? i * 6 - (i * (i - 1)) / 2 + (j - i) : j * 6 - (j * (j - 1)) / 2 + (i - j);
}
// packed symmetric part of a full row-major 6x6: (A + A^T)/2
CG_HD void sym_pack6(const double* A, double* P) {
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = i; j < 6; j++) P[sp6(i, j)] = i == j ? A[7 * i] : 0.5 * (A[6 * i + j] + A[6 * j + i]);
}
// y = P x
CG_HD void spmv6(const double* P, const double* x, double* y) {
#pragma unroll
    for (int i = 0; i < 6; i++) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < 6; j++) s += P[sp6(i, j)] * x[j];
        y[i] = s;
    }
}
// In-place L D L^T factorisation of a packed symmetric 6x6 (no pivoting; the
// matrices factored here are positive definite).  Afterwards P[sp6(i,i)] = d_i and
// P[sp6(i,j)] = L_ij (i > j).  Returns false if a pivot is not positive.
CG_HD bool ldl6(double* P) {
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 6; j++) {
        double d = P[sp6(j, j)];
#pragma unroll
        for (int k = 0; k < j; k++) d -= P[sp6(j, k)] * P[sp6(j, k)] * P[sp6(k, k)];
        if (!(d > 0.0)) ok = false;
        P[sp6(j, j)] = d;
        const double r = recip(d);  // K is SPD: d > 0
#pragma unroll
        for (int i = j + 1; i < 6; i++) {
            double s = P[sp6(i, j)];
#pragma unroll
            for (int k = 0; k < j; k++) s -= P[sp6(i, k)] * P[sp6(j, k)] * P[sp6(k, k)];
            P[sp6(i, j)] = s * r;
        }
    }
    return ok;
}
// solve (L D L^T) x = b in place with the factors of ldl6
CG_HD void ldl6_solve(const double* P, double* b) {
#pragma unroll
    for (int i = 1; i < 6; i++) {
        double s = b[i];
#pragma unroll
        for (int k = 0; k < i; k++) s -= P[sp6(i, k)] * b[k];
        b[i] = s;
    }
#pragma unroll
    for (int i = 0; i < 6; i++) b[i] /= P[sp6(i, i)];
#pragma unroll
    for (int i = 4; i >= 0; i--) {
        double s = b[i];
#pragma unroll
        for (int k = i + 1; k < 6; k++) s -= P[sp6(k, i)] * b[k];
        b[i] = s;
    }
}
// ldl6 that also returns the reciprocal pivots rd[i] = 1/d_i, and the solve that
// multiplies by them (no division in the solve)
CG_HD void ldl6r(double* P, double* rd) {
#pragma unroll
    for (int j = 0; j < 6; j++) {
        double d = P[sp6(j, j)];
#pragma unroll
        for (int k = 0; k < j; k++) d -= P[sp6(j, k)] * P[sp6(j, k)] * P[sp6(k, k)];
        P[sp6(j, j)] = d;
        const double r = recip(d);  // K is SPD: d > 0
        rd[j] = r;
#pragma unroll
        for (int i = j + 1; i < 6; i++) {
            double s = P[sp6(i, j)];
#pragma unroll
            for (int k = 0; k < j; k++) s -= P[sp6(i, k)] * P[sp6(j, k)] * P[sp6(k, k)];
            P[sp6(i, j)] = s * r;
        }
    }
}
CG_HD void ldl6_solve_r(const double* P, const double* rd, double* b) {
#pragma unroll
    for (int i = 1; i < 6; i++) {
        double s = b[i];
#pragma unroll
        for (int k = 0; k < i; k++) s -= P[sp6(i, k)] * b[k];
        b[i] = s;
    }
#pragma unroll
    for (int i = 0; i < 6; i++) b[i] *= rd[i];
#pragma unroll
    for (int i = 4; i >= 0; i--) {
        double s = b[i];
#pragma unroll
        for (int k = i + 1; k < 6; k++) s -= P[sp6(k, i)] * b[k];
        b[i] = s;
    }
}
// LU factorisation with partial pivoting of a 6x6 (in place; unit lower L)
// LU with partial pivoting.  Every loop is unrolled and the row exchanges are selects over the
// candidate rows, so A, piv and b stay in registers on the GPU (indexing them with the pivot
// would place them in scratch memory); the arithmetic and the pivot choice are those of the
// textbook loop.
CG_HD bool lu6(double* A, int* piv) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        int p = k;
        double mx = fabs(A[6 * k + k]);
#pragma unroll
        for (int i = k + 1; i < 6; i++) {
            double v = fabs(A[6 * i + k]);
            if (v > mx) { mx = v; p = i; }
        }
        piv[k] = p;
        if (mx == 0.0) { ok = false; continue; }
#pragma unroll
        for (int i = k + 1; i < 6; i++) {
            const bool sw = p == i;
#pragma unroll
            for (int j = 0; j < 6; j++) {
                const double t = A[6 * k + j];
                A[6 * k + j] = sw ? A[6 * i + j] : t;
                A[6 * i + j] = sw ? t : A[6 * i + j];
            }
        }
        double r = 1.0 / A[6 * k + k];
#pragma unroll
        for (int i = k + 1; i < 6; i++) {
            double l = A[6 * i + k] * r;
            A[6 * i + k] = l;
#pragma unroll
            for (int j = k + 1; j < 6; j++) A[6 * i + j] -= l * A[6 * k + j];
        }
    }
    return ok;
}
// b[k] <-> b[piv[k]] with the exchange as selects over the candidate rows
CG_HD void swap_piv6(const int* piv, double* b, int k) {
#pragma unroll
    for (int i = k + 1; i < 6; i++) {
        const bool sw = piv[k] == i;
        const double t = b[k];
        b[k] = sw ? b[i] : t;
        b[i] = sw ? t : b[i];
    }
}
CG_HD void lu6_solve(const double* LU, const int* piv, double* b) {
#pragma unroll
    for (int k = 0; k < 6; k++) swap_piv6(piv, b, k);
#pragma unroll
    for (int i = 1; i < 6; i++) {
        double s = b[i];
#pragma unroll
        for (int j = 0; j < i; j++) s -= LU[6 * i + j] * b[j];
        b[i] = s;
    }
#pragma unroll
    for (int i = 5; i >= 0; i--) {
        double s = b[i];
#pragma unroll
        for (int j = i + 1; j < 6; j++) s -= LU[6 * i + j] * b[j];
        b[i] = s / LU[6 * i + i];
    }
}
// solve A^T x = b given the LU of A
CG_HD void lu6_solve_t(const double* LU, const int* piv, double* b) {
#pragma unroll
    for (int i = 0; i < 6; i++) {
        double s = b[i];
#pragma unroll
        for (int j = 0; j < i; j++) s -= LU[6 * j + i] * b[j];
        b[i] = s / LU[6 * i + i];
    }
#pragma unroll
    for (int i = 4; i >= 0; i--) {
        double s = b[i];
#pragma unroll
        for (int j = i + 1; j < 6; j++) s -= LU[6 * j + i] * b[j];
        b[i] = s;
    }
#pragma unroll
    for (int k = 5; k >= 0; k--) swap_piv6(piv, b, k);
}
// inverse of a 6x6 through its LU factors (in place)
CG_HD bool inv6(double* A) {
    double LU[36];
    int piv[6];
#pragma unroll
    for (int i = 0; i < 36; i++) LU[i] = A[i];
    bool ok = lu6(LU, piv);
#pragma unroll
    for (int j = 0; j < 6; j++) {
        double e[6] = {0, 0, 0, 0, 0, 0};
        e[j] = 1.0;
        lu6_solve(LU, piv, e);
#pragma unroll
        for (int i = 0; i < 6; i++) A[6 * i + j] = e[i];
    }
    return ok;
}
CG_HD void mm6(const double* a, const double* b, double* c) {
    double t[36];
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            double s = 0.0;
            for (int k = 0; k < 6; k++) s += a[6 * i + k] * b[6 * k + j];
            t[6 * i + j] = s;
        }
    for (int i = 0; i < 36; i++) c[i] = t[i];
}
// c = m a m^T
CG_HD void congr6(const double* m, const double* a, double* c) {
    double t[36];
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            double s = 0.0;
            for (int k = 0; k < 6; k++) s += m[6 * i + k] * a[6 * k + j];
            t[6 * i + j] = s;
        }
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            double s = 0.0;
            for (int k = 0; k < 6; k++) s += t[6 * i + k] * m[6 * j + k];
            c[6 * i + j] = s;
        }
}
// M(X)_{ca} = b_c : (b_a X^T): the b-basis matrix of the map S -> sym(S X^T)
// used for the push-forwards  c0mod = sum_ab c_ab (b_a F^T)(x)(b_b F^T)/J.
CG_HD void push_matrix(const double* X, double* M) {
    double ba[9], y[9];
    for (int a = 0; a < 6; a++) {
        basis(a, ba);
        mmt3(ba, X, y);  // b_a X^T
        double ys[6];
        to6(y, ys);      // b_c : y for symmetric b_c equals b_c : sym(y)
        for (int c = 0; c < 6; c++) M[6 * c + a] = ys[c];
    }
}
// T(A)_{ca} = b_c : (A b_a A^T): the b-basis matrix of S -> A S A^T
CG_HD void rot_matrix(const double* A, double* T) {
    double ba[9], y[9];
    for (int a = 0; a < 6; a++) {
        basis(a, ba);
        mm3(A, ba, y);
        mmt3(y, A, y);
        double ys[6];
        to6(y, ys);
        for (int c = 0; c < 6; c++) T[6 * c + a] = ys[c];
    }
}

// ---- matrix exponential (various_functions.f90:1830-1967) --------------------
CG_HD double l1n3(const double* m) {
    double a = fabs(m[0]) + fabs(m[3]) + fabs(m[6]);
    double b = fabs(m[1]) + fabs(m[4]) + fabs(m[7]);
    double c = fabs(m[2]) + fabs(m[5]) + fabs(m[8]);
    return fmax(a, fmax(b, c));
}
CG_HD void expm3(const double* A0, double* E) {
    double n1 = l1n3(A0);
    if (n1 <= 0.1) {  // cubic Taylor
        double a2[9], a3[9];
        mm3(A0, A0, a2);
        mm3(a2, A0, a3);
        for (int i = 0; i < 9; i++) E[i] = (i % 4 == 0 ? 1.0 : 0.0) + A0[i] + a2[i] * 0.5 + a3[i] * (1.0 / 6.0);
        return;
    }
    // Pade [13/13] with scaling and squaring (Higham 2005 coefficients)
    const double b[14] = {64764752532480000.0, 32382376266240000.0, 7771770303897600.0, 1187353796428800.0,
                          129060195264000.0,   10559470521600.0,    670442572800.0,     33522128640.0,
                          1323241920.0,        40840800.0,          960960.0,           16380.0,
                          182.0,               1.0};
    const double theta13 = 5.371920351148152;
    // a finite norm needs at most 1022 squarings; a non-finite one (a diverged increment)
    // would turn the count into an unbounded loop, so it is capped (the result is then
    // non-finite, which the caller's error checks report)
    int s = 0;
    if (n1 > theta13) s = (int)fmin(ceil(log2(n1 / theta13)), 1100.0);
    double sc = ldexp(1.0, -s);
    double A[9], A2[9], A4[9], A6[9], U[9], V[9], T[9];
    for (int i = 0; i < 9; i++) A[i] = A0[i] * sc;
    mm3(A, A, A2);
    mm3(A2, A2, A4);
    mm3(A2, A4, A6);
    for (int i = 0; i < 9; i++) T[i] = b[13] * A6[i] + b[11] * A4[i] + b[9] * A2[i];
    mm3(A6, T, T);
    for (int i = 0; i < 9; i++) T[i] += b[7] * A6[i] + b[5] * A4[i] + b[3] * A2[i] + (i % 4 == 0 ? b[1] : 0.0);
    mm3(A, T, U);
    for (int i = 0; i < 9; i++) T[i] = b[12] * A6[i] + b[10] * A4[i] + b[8] * A2[i];
    mm3(A6, T, V);
    for (int i = 0; i < 9; i++) V[i] += b[6] * A6[i] + b[4] * A4[i] + b[2] * A2[i] + (i % 4 == 0 ? b[0] : 0.0);
    double P[9], Q[9], Qi[9];
    for (int i = 0; i < 9; i++) { P[i] = V[i] + U[i]; Q[i] = V[i] - U[i]; }
    inv3(Q, Qi);
    mm3(Qi, P, E);
    for (int k = 0; k < s; k++) mm3(E, E, E);
}

// ---- symmetric 3x3 eigenvalues / polar decomposition -------------------------
// eigen-decomposition of a symmetric 3x3 by cyclic Jacobi rotations
CG_HD void eig3(const double* S, double* w, double* Q) {
    double a[9];
    for (int i = 0; i < 9; i++) { a[i] = S[i]; Q[i] = (i % 4 == 0) ? 1.0 : 0.0; }
    for (int sweep = 0; sweep < 32; sweep++) {
        double off = a[1] * a[1] + a[2] * a[2] + a[5] * a[5];
        double dg = a[0] * a[0] + a[4] * a[4] + a[8] * a[8];
        if (!(off > 1e-36 * dg)) break;
        for (int p = 0; p < 2; p++)
            for (int r = p + 1; r < 3; r++) {
                double apr = a[3 * p + r];
                if (apr == 0.0) continue;
                double th = (a[3 * r + r] - a[3 * p + p]) / (2.0 * apr);
                double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; k++) {
                    double akp = a[3 * k + p], akr = a[3 * k + r];
                    a[3 * k + p] = c * akp - s * akr;
                    a[3 * k + r] = s * akp + c * akr;
                }
                for (int k = 0; k < 3; k++) {
                    double apk = a[3 * p + k], ark = a[3 * r + k];
                    a[3 * p + k] = c * apk - s * ark;
                    a[3 * r + k] = s * apk + c * ark;
                }
                for (int k = 0; k < 3; k++) {
                    double qkp = Q[3 * k + p], qkr = Q[3 * k + r];
                    Q[3 * k + p] = c * qkp - s * qkr;
                    Q[3 * k + r] = s * qkp + c * qkr;
                }
            }
    }
    w[0] = a[0]; w[1] = a[4]; w[2] = a[8];
}
// F = R U, U = sqrt(F^T F); also returns the principal stretches (eigenvalues of U)
CG_HD void polar3(const double* F, double* R, double* lam) {
    double C[9], w[3], Q[9];
    double Ft[9];
    transp3(F, Ft);
    mm3(Ft, F, C);
    eig3(C, w, Q);
    double is[3];
    for (int i = 0; i < 3; i++) {
        lam[i] = sqrt(fmax(w[i], 0.0));
        is[i] = 1.0 / lam[i];
    }
    double Ui[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            Ui[3 * i + j] = Q[3 * i] * is[0] * Q[3 * j] + Q[3 * i + 1] * is[1] * Q[3 * j + 1] + Q[3 * i + 2] * is[2] * Q[3 * j + 2];
    mm3(F, Ui, R);
}

}  // namespace cg
