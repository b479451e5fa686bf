/*
 * evpfft_oracle.c -- CPU restatement of the LS-EVPFFT increment loop.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libcreep_hip.so and the CPU
 * baseline.  Never linked into or called by the product.
 *
 * Conventions: 3x3 tensors row-major a[3*i+j] = A(i+1,j+1); 6x6 row-major;
 * 4th-order c[27*i+9*j+3*k+l]; voxel index v = i + n1*(j + n2*k) (x fastest,
 * the order of micro%voxel(i,j,k) loops with i innermost).
 */
#include "evpfft_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXPH 4
#define PI 3.141592653589793238462643383279502884

typedef double complex cplx;

/* ------------------------------------------------------------------------ */
/* small dense helpers (tensor_functions.f90)                                */
/* ------------------------------------------------------------------------ */

static const double ID3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};

static void mm33(const double *a, const double *b, double *c) {
    double t[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += a[3 * i + k] * b[3 * k + j];
            t[3 * i + j] = s;
        }
    memcpy(c, t, sizeof t);
}
static void tr33(const double *a, double *b) {
    double t[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) t[3 * i + j] = a[3 * j + i];
    memcpy(b, t, sizeof t);
}
static void mv33(const double *a, const double *x, double *y) {
    double t[3];
    for (int i = 0; i < 3; i++) t[i] = a[3 * i] * x[0] + a[3 * i + 1] * x[1] + a[3 * i + 2] * x[2];
    memcpy(y, t, sizeof t);
}
/* sym33 / antisym33 (tensor_functions.f90:459-509) */
static void sym33(const double *a, double *s) {
    double t[9];
    t[0] = a[0]; t[4] = a[4]; t[8] = a[8];
    t[1] = t[3] = 0.5 * (a[1] + a[3]);
    t[2] = t[6] = 0.5 * (a[2] + a[6]);
    t[5] = t[7] = 0.5 * (a[5] + a[7]);
    memcpy(s, t, sizeof t);
}
static void antisym33(const double *a, double *s) {
    double t[9];
    t[0] = t[4] = t[8] = 0;
    t[1] = 0.5 * (a[1] - a[3]); t[3] = -t[1];
    t[2] = 0.5 * (a[2] - a[6]); t[6] = -t[2];
    t[5] = 0.5 * (a[5] - a[7]); t[7] = -t[5];
    memcpy(s, t, sizeof t);
}
/* determinant33 (tensor_functions.f90:442-456) */
static double det33(const double *m) {
    return m[0] * m[4] * m[8] + m[1] * m[5] * m[6] + m[2] * m[3] * m[7] - m[6] * m[4] * m[2] -
           m[7] * m[5] * m[0] - m[8] * m[3] * m[1];
}
/* vm_stress / vm_strain (tensor_functions.f90:266-289) */
static double vm_core(const double *d) {
    double h = (d[0] + d[4] + d[8]) / 3.0;
    return (d[0] - h) * (d[0] - h) + (d[4] - h) * (d[4] - h) + (d[8] - h) * (d[8] - h) + d[1] * d[1] +
           d[2] * d[2] + d[5] * d[5] + d[3] * d[3] + d[6] * d[6] + d[7] * d[7];
}
static double vm_stress(const double *d) { return sqrt(1.5 * vm_core(d)); }
static double vm_strain(const double *d) { return sqrt(2.0 / 3.0 * vm_core(d)); }
static double norm2v(const double *a, int n) {
    double s = 0;
    for (int i = 0; i < n; i++) s += a[i] * a[i];
    return sqrt(s);
}

/* LU with partial pivoting (the dgetrf contract): returns 0 or k>0 if U(k,k)==0. */
static int lu_fact(double *a, int n, int *piv) {
    int info = 0;
    for (int k = 0; k < n; k++) {
        int p = k;
        double mx = fabs(a[k * n + k]);
        for (int i = k + 1; i < n; i++)
            if (fabs(a[i * n + k]) > mx) { mx = fabs(a[i * n + k]); p = i; }
        piv[k] = p;
        if (a[p * n + k] == 0.0) { if (!info) info = k + 1; continue; }
        if (p != k)
            for (int j = 0; j < n; j++) { double t = a[k * n + j]; a[k * n + j] = a[p * n + j]; a[p * n + j] = t; }
        double r = 1.0 / a[k * n + k];
        for (int i = k + 1; i < n; i++) {
            double l = a[i * n + k] * r;
            a[i * n + k] = l;
            for (int j = k + 1; j < n; j++) a[i * n + j] -= l * a[k * n + j];
        }
    }
    return info;
}
static void lu_solve(const double *lu, int n, const int *piv, double *b) {
    for (int k = 0; k < n; k++) { int p = piv[k]; if (p != k) { double t = b[k]; b[k] = b[p]; b[p] = t; } }
    for (int i = 0; i < n; i++) { double s = b[i]; for (int j = 0; j < i; j++) s -= lu[i * n + j] * b[j]; b[i] = s; }
    for (int i = n - 1; i >= 0; i--) { double s = b[i]; for (int j = i + 1; j < n; j++) s -= lu[i * n + j] * b[j]; b[i] = s / lu[i * n + i]; }
}
/* solve with A^T given the LU of A */
static void lu_solve_t(const double *lu, int n, const int *piv, double *b) {
    for (int i = 0; i < n; i++) { double s = b[i]; for (int j = 0; j < i; j++) s -= lu[j * n + i] * b[j]; b[i] = s / lu[i * n + i]; }
    for (int i = n - 1; i >= 0; i--) { double s = b[i]; for (int j = i + 1; j < n; j++) s -= lu[j * n + i] * b[j]; b[i] = s; }
    for (int k = n - 1; k >= 0; k--) { int p = piv[k]; if (p != k) { double t = b[k]; b[k] = b[p]; b[p] = t; } }
}

static int g_singular = 0; /* sticky flag: a singular LU was met (reference: stop 1) */

/* lu_inverse (tensor_functions.f90:313-339) */
static void lu_inverse(double *a, int n) {
    double lu[36];
    int piv[6];
    memcpy(lu, a, sizeof(double) * n * n);
    if (lu_fact(lu, n, piv)) { g_singular = 1; }
    for (int j = 0; j < n; j++) {
        double e[6] = {0};
        e[j] = 1.0;
        lu_solve(lu, n, piv, e);
        for (int i = 0; i < n; i++) a[i * n + j] = e[i];
    }
}
/* lu_eqsystem (tensor_functions.f90:344-359): b <- A^{-1} b */
static void lu_eqsystem(double *a, double *b, int n) {
    int piv[6];
    if (lu_fact(a, n, piv)) { g_singular = 1; }
    lu_solve(a, n, piv, b);
}

/* ------------------------------------------------------------------------ */
/* chg_basis (tensor_functions.f90:21-198) restated through the basis tensors */
/* ------------------------------------------------------------------------ */
static double BB[6][9];
static void init_basis(void) {
    static int done = 0;
    if (done) return;
    memset(BB, 0, sizeof BB);
    double r2 = 1.0 / sqrt(2.0), r6 = 1.0 / sqrt(6.0), r3 = 1.0 / sqrt(3.0);
    BB[0][0] = -r2; BB[0][4] = r2;                             /* (e22-e11)/sqrt2     */
    BB[1][0] = -r6; BB[1][4] = -r6; BB[1][8] = 2.0 * r6;        /* (2e33-e11-e22)/sqrt6 */
    BB[2][5] = BB[2][7] = r2;                                  /* (e23+e32)/sqrt2     */
    BB[3][2] = BB[3][6] = r2;                                  /* (e13+e31)/sqrt2     */
    BB[4][1] = BB[4][3] = r2;                                  /* (e12+e21)/sqrt2     */
    BB[5][0] = BB[5][4] = BB[5][8] = r3;                       /* I/sqrt3             */
    done = 1;
}
void orc_chg_basis(double *ce2, double *c2, double *ce4, double *c4, int iopt, int kdim) {
    init_basis();
    if (iopt == 1) {
        for (int p = 0; p < 9; p++) {
            double s = 0;
            for (int a = 0; a < kdim; a++) s += ce2[a] * BB[a][p];
            c2[p] = s;
        }
    } else if (iopt == 2) {
        for (int a = 0; a < kdim; a++) {
            double s = 0;
            for (int p = 0; p < 9; p++) s += BB[a][p] * c2[p];
            ce2[a] = s;
        }
    } else if (iopt == 3) {
        for (int p = 0; p < 9; p++)
            for (int q = 0; q < 9; q++) {
                double s = 0;
                for (int a = 0; a < kdim; a++)
                    for (int b = 0; b < kdim; b++) s += ce4[a * kdim + b] * BB[a][p] * BB[b][q];
                c4[p * 9 + q] = s;
            }
    } else if (iopt == 4) {
        for (int a = 0; a < kdim; a++)
            for (int b = 0; b < kdim; b++) {
                double s = 0;
                for (int p = 0; p < 9; p++) {
                    if (BB[a][p] == 0.0) continue;
                    for (int q = 0; q < 9; q++) s += BB[a][p] * c4[p * 9 + q] * BB[b][q];
                }
                ce4[a * kdim + b] = s;
            }
    }
}
#define B2(ce, c) orc_chg_basis((ce), (c), NULL, NULL, 2, 6)
#define B1(ce, c) orc_chg_basis((ce), (c), NULL, NULL, 1, 6)
#define B3(ce4, c4) orc_chg_basis(NULL, NULL, (ce4), (c4), 3, 6)
#define B4(ce4, c4) orc_chg_basis(NULL, NULL, (ce4), (c4), 4, 6)

/* voigt_vpsc (tensor_functions.f90:208-261), VPSC order 11,22,33,23,13,12 */
static const int IJV[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};
static void voigt_t1_to_t2(const double *t1, double *t2) {
    for (int i = 0; i < 6; i++) { t2[3 * IJV[i][0] + IJV[i][1]] = t1[i]; t2[3 * IJV[i][1] + IJV[i][0]] = t1[i]; }
}
static void voigt_t2_to_t1(const double *t2, double *t1) {
    for (int i = 0; i < 6; i++) t1[i] = t2[3 * IJV[i][0] + IJV[i][1]];
}
static void voigt_c4_to_c2(const double *c4, double *c2) {
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++)
            c2[i * 6 + j] = c4[27 * IJV[i][0] + 9 * IJV[i][1] + 3 * IJV[j][0] + IJV[j][1]];
}

/* euler iopt=2 (tensor_functions.f90:525-569): a transforms sample -> crystal */
static void euler2(double ph, double th, double tm, double *a) {
    double sph = sin(ph), cph = cos(ph), sth = sin(th), cth = cos(th), stm = sin(tm), ctm = cos(tm);
    a[0] = ctm * cph - sph * stm * cth;  /* a(1,1) */
    a[3] = -stm * cph - sph * ctm * cth; /* a(2,1) */
    a[6] = sph * sth;                    /* a(3,1) */
    a[1] = ctm * sph + cph * stm * cth;  /* a(1,2) */
    a[4] = -sph * stm + cph * ctm * cth; /* a(2,2) */
    a[7] = -sth * cph;                   /* a(3,2) */
    a[2] = sth * stm;                    /* a(1,3) */
    a[5] = ctm * sth;                    /* a(2,3) */
    a[8] = cth;                          /* a(3,3) */
}

/* ------------------------------------------------------------------------ */
/* symmetric 3x3 eigen-decomposition (cyclic Jacobi) and polar decomposition  */
/* ------------------------------------------------------------------------ */
static void jacobi33(const double *s, double *w, double *q) {
    double a[9];
    memcpy(a, s, sizeof a);
    memcpy(q, ID3, sizeof(double) * 9);
    for (int sweep = 0; sweep < 64; sweep++) {
        double off = a[1] * a[1] + a[2] * a[2] + a[5] * a[5];
        double dg = a[0] * a[0] + a[4] * a[4] + a[8] * a[8];
        if (off <= 1e-36 * dg || off == 0.0) break;
        for (int p = 0; p < 2; p++)
            for (int r = p + 1; r < 3; r++) {
                double apr = a[3 * p + r];
                if (apr == 0.0) continue;
                double theta = (a[3 * r + r] - a[3 * p + p]) / (2.0 * apr);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < 3; k++) { /* A <- J^T A J */
                    double akp = a[3 * k + p], akr = a[3 * k + r];
                    a[3 * k + p] = c * akp - sn * akr;
                    a[3 * k + r] = sn * akp + c * akr;
                }
                for (int k = 0; k < 3; k++) {
                    double apk = a[3 * p + k], ark = a[3 * r + k];
                    a[3 * p + k] = c * apk - sn * ark;
                    a[3 * r + k] = sn * apk + c * ark;
                }
                for (int k = 0; k < 3; k++) {
                    double qkp = q[3 * k + p], qkr = q[3 * k + r];
                    q[3 * k + p] = c * qkp - sn * qkr;
                    q[3 * k + r] = sn * qkp + c * qkr;
                }
            }
    }
    w[0] = a[0]; w[1] = a[4]; w[2] = a[8];
}
/* polar_dcmp (tensor_functions.f90:367-393): F = R U with U = (F^T F)^(1/2).
 * Restated through the eigen-decomposition of F^T F (same R, U as the SVD
 * route for non-singular F). */
void orc_polar(const double *F, double *R, double *U) {
    double ft[9], ftf[9], w[3], q[9], s[3], ui[9];
    tr33(F, ft);
    mm33(ft, F, ftf);
    jacobi33(ftf, w, q);
    for (int i = 0; i < 3; i++) s[i] = sqrt(w[i] > 0 ? w[i] : 0.0);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double u = 0, v = 0;
            for (int k = 0; k < 3; k++) { u += q[3 * i + k] * s[k] * q[3 * j + k]; v += q[3 * i + k] / s[k] * q[3 * j + k]; }
            U[3 * i + j] = u;
            ui[3 * i + j] = v;
        }
    mm33(F, ui, R);
}

/* ------------------------------------------------------------------------ */
/* matrix exponentials (various_functions.f90:1830-1967)                     */
/* ------------------------------------------------------------------------ */
static double l1_norm33(const double *m) {
    double l1 = 0;
    for (int j = 0; j < 3; j++) {
        double cs = fabs(m[j]) + fabs(m[3 + j]) + fabs(m[6 + j]);
        if (cs > l1) l1 = cs;
    }
    return l1;
}
static void expm_taylor3(const double *a, double *e) {
    double a2[9], a3[9];
    mm33(a, a, a2);
    mm33(a2, a, a3);
    for (int i = 0; i < 9; i++) e[i] = ID3[i] + a[i] + a2[i] / 2.0 + a3[i] / 6.0;
}
/* Pade(13) with scaling and squaring, Higham (2005) coefficients */
static void expm_pade13(const double *m, double *e) {
    static const double b[14] = {64764752532480000.0, 32382376266240000.0, 7771770303897600.0,
                                 1187353796428800.0, 129060195264000.0, 10559470521600.0,
                                 670442572800.0, 33522128640.0, 1323241920.0, 40840800.0, 960960.0,
                                 16380.0, 182.0, 1.0};
    const double theta13 = 5.371920351148152;
    double A[9], A2[9], A4[9], A6[9], U[9], V[9], T[9], P[9], Q[9];
    double nrm = l1_norm33(m);
    if (nrm <= 0.0) { memcpy(e, ID3, sizeof A); return; }
    int s = 0;
    /* capped for a non-finite norm, which would make the squaring loop unbounded */
    if (nrm > theta13) s = (int)fmin(ceil(log(nrm / theta13) / log(2.0)), 1100.0);
    double sc = pow(2.0, s);
    for (int i = 0; i < 9; i++) A[i] = m[i] / sc;
    mm33(A, A, A2);
    mm33(A2, A2, A4);
    mm33(A2, A4, A6);
    for (int i = 0; i < 9; i++) T[i] = b[13] * A6[i] + b[11] * A4[i] + b[9] * A2[i];
    mm33(A6, T, T);
    for (int i = 0; i < 9; i++) T[i] += b[7] * A6[i] + b[5] * A4[i] + b[3] * A2[i] + b[1] * ID3[i];
    mm33(A, T, U);
    for (int i = 0; i < 9; i++) T[i] = b[12] * A6[i] + b[10] * A4[i] + b[8] * A2[i];
    mm33(A6, T, V);
    for (int i = 0; i < 9; i++) V[i] += b[6] * A6[i] + b[4] * A4[i] + b[2] * A2[i] + b[0] * ID3[i];
    for (int i = 0; i < 9; i++) { P[i] = V[i] + U[i]; Q[i] = V[i] - U[i]; }
    int piv[3];
    if (lu_fact(Q, 3, piv)) g_singular = 1;
    for (int j = 0; j < 3; j++) {
        double col[3] = {P[j], P[3 + j], P[6 + j]};
        lu_solve(Q, 3, piv, col);
        e[j] = col[0]; e[3 + j] = col[1]; e[6 + j] = col[2];
    }
    for (int k = 0; k < s; k++) mm33(e, e, e);
}
void orc_expm_adaptive(const double *a, double *e) {
    if (l1_norm33(a) <= 0.1) expm_taylor3(a, e);
    else expm_pade13(a, e);
}

/* ------------------------------------------------------------------------ */
/* FFT: 3D unnormalised c2c over all axes (fourier_functions.f90:141-224)     */
/* ------------------------------------------------------------------------ */
typedef struct { int n; cplx *tw; } fftplan1;

static void fft_rec(const cplx *in, int sin, cplx *out, int n, int N, const cplx *tw) {
    if (n == 1) { out[0] = in[0]; return; }
    int p = (n % 2 == 0) ? 2 : (n % 3 == 0) ? 3 : n;
    int m = n / p;
    if (p == n) { /* direct DFT */
        for (int k = 0; k < n; k++) {
            cplx s = 0;
            for (int j = 0; j < n; j++) s += in[j * sin] * tw[((long)j * k % n) * (N / n)];
            out[k] = s;
        }
        return;
    }
    for (int r = 0; r < p; r++) fft_rec(in + r * sin, sin * p, out + r * m, m, N, tw);
    int stepN = N / n;
    for (int k = 0; k < m; k++) {
        cplx t[3];
        for (int r = 0; r < p; r++) t[r] = out[r * m + k] * tw[(r * k * stepN) % N];
        for (int q = 0; q < p; q++) {
            cplx s = 0;
            for (int r = 0; r < p; r++) s += t[r] * tw[((r * q) % p) * (N / p)];
            out[k + m * q] = s;
        }
    }
}
static cplx *make_tw(int n, int sign) {
    cplx *tw = malloc(sizeof(cplx) * n);
    for (int j = 0; j < n; j++) {
        double ang = 2.0 * PI * (double)j / (double)n;
        tw[j] = cos(ang) + sign * I * sin(ang);
    }
    return tw;
}
/* data: ncomp blocks of V complex (interleaved re,im), v = i + n1*(j + n2*k) */
int orc_fft3d(double *re_im, int n1, int n2, int n3, int ncomp, int sign) {
    cplx *d = (cplx *)re_im;
    long V = (long)n1 * n2 * n3;
    int ns[3] = {n1, n2, n3};
    long st[3] = {1, n1, (long)n1 * n2};
    for (int ax = 0; ax < 3; ax++) {
        int n = ns[ax];
        if (n == 1) continue;
        cplx *tw = make_tw(n, sign);
        long nlines = V / n;
        #pragma omp parallel
        {
            cplx *tmp = malloc(sizeof(cplx) * n), *o = malloc(sizeof(cplx) * n);
            #pragma omp for collapse(2) schedule(static)
            for (int c = 0; c < ncomp; c++)
                for (long ln = 0; ln < nlines; ln++) {
                    long base;
                    if (ax == 0) base = ln * n1;
                    else if (ax == 1) base = (ln % n1) + (ln / n1) * (long)n1 * n2;
                    else base = ln;
                    cplx *p = d + c * V + base;
                    for (int t = 0; t < n; t++) tmp[t] = p[t * st[ax]];
                    fft_rec(tmp, 1, o, n, n, tw);
                    for (int t = 0; t < n; t++) p[t * st[ax]] = o[t];
                }
            free(tmp); free(o);
        }
        free(tw);
    }
    return 0;
}

/* spatial_freq (fourier_functions.f90:340-371), 1-based indices */
void orc_spatial_freq(int x, int y, int z, int n1, int n2, int n3, double *xi) {
    const double twopi = 2.0 * PI;
    xi[0] = (x <= n1 / 2) ? twopi * (x - 1) / n1 : twopi * (x - n1 - 1) / n1;
    xi[1] = (y <= n2 / 2) ? twopi * (y - 1) / n2 : twopi * (y - n2 - 1) / n2;
    if (n3 == 1) xi[2] = 0.0;
    else xi[2] = (z <= n3 / 2) ? twopi * (z - 1) / n3 : twopi * (z - n3 - 1) / n3;
}

/* ------------------------------------------------------------------------ */
/* context                                                                   */
/* ------------------------------------------------------------------------ */
typedef struct {
    int igas, nsyst, nmodes, nsm[16];
    double *dnca, *dbca, *tau, *thet, *nrs, *gamd0, *hpfac, *hard;
    double C[81];
} phase_t;

struct orc_ctx {
    int n1, n2, n3;
    long V;
    int nph, modal, do_mix, nthreads;
    double complgas, wgt, wphsol, wphgas;
    phase_t ph[MAXPH];
    int ns; /* max systems over phases (per-voxel stride) */
    int nproc, proc_cap;
    orc_process *proc;
    char err[256];
    /* per voxel */
    int *grain, *phase;
    unsigned char *gas;
    double *euler;
    double *wgtc, *sg, *sgt, *velgrad, *velgradold, *dvelgradref, *defgrad, *defgradinv, *detF;
    double *defgrade, *defgradp, *ag, *cg66, *sg66, *c066mod, *cgm, *cgas;
    double *sch, *crss, *trialtau, *gamdot, *gacumgr;
    double *edotp, *ept, *disgradsym, *epvm, *defgradinc, *defgradeinc;
    double *eth; /* eigenstrain rate of the voxel's grain (ithermo = 1), 3x3 */
    int idispini;      /* initial displacement field (IO_functions.f90:147-154): 0 none, 1 cells, 2 nodes */
    double *dispini;   /* as read: [cell][3] (idispini 1) or [node][3] (2), i fastest */
    double *defgradini; /* F0 of init_defgrad, 3x3 per voxel (null: identity) */
    double *s_velgrad, *s_velgradold, *s_sg, *s_edotp, *s_gamdot, *s_trialtau;
    double *Goper; /* V*81 */
    double *xnode; /* 3*(n1+1)*(n2+1)*(n3node) */
    cplx *four;    /* 9*V */
    /* macro state (micro_type) */
    double c066[36], s066[36], c0[81], s0[81], Goperr0[81], de_ds[81];
    double wphcsol, wphcgas, defgradavg[9], defgradinvavgc[9], defgradinvavgcs[9], defgradinvavgcg[9];
    double dvelgradavg[9], velgradavg[9], velgradavgsol[9], velgradavggas[9];
    double sgavg[9], sgavgsol[9], sgsolavg[9], svmavg, dvmavg, edotvmmx, evmavg;
    double epavg[9], eavg[9], eelavg[9], edotpavg[9], evmpavg, dvmpavg, lvmavg;
    double dsim_snap[9], scauchy_snap[9];
    /* module globals (global.f90:19-27) */
    int ipr, imacroloop, imicro, iter;
    double erre, errs, errsbc;
    /* driver state (LS-EVPFFT.f90) */
    int initialised, started, nconverged, total_macrosteps, proc_active;
    double sim_time, tdot_plan, proc_time_target, proc_time_done;
    int kinematic_controlled;
    /* diagnostics prev values (diag_type) */
    double dprev_detF_min, dprev_detF_min_sol, dprev_detF_min_gas;
    double dprev_volinc_sol, dprev_volinc_gas, dprev_stretch_sol, dprev_stretch_gas;
    long long cnt_nr, cnt_eval, cnt_outer;
};

#define ALLOC(p, n) do { (p) = calloc((size_t)(n), sizeof(*(p))); if (!(p)) return NULL; } while (0)

orc_ctx *orc_create(int n1, int n2, int n3, int nph, double complgas, int modal, int do_mix,
                    int nthreads) {
    if (n1 < 1 || n2 < 1 || n3 < 1 || nph < 1 || nph > MAXPH) return NULL;
    orc_ctx *c = calloc(1, sizeof(orc_ctx));
    c->n1 = n1; c->n2 = n2; c->n3 = n3;
    c->V = (long)n1 * n2 * n3;
    c->nph = nph; c->complgas = complgas; c->modal = modal; c->do_mix = do_mix;
    c->nthreads = nthreads;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    c->wgt = 1.0 / (double)c->V;
    c->proc_cap = 8;
    c->proc = calloc(c->proc_cap, sizeof(orc_process));
    init_basis();
    g_singular = 0; /* a new run starts clean (the reference process would have stopped) */
    return c;
}

static void freeall(orc_ctx *c) {
    free(c->dispini);
    free(c->defgradini);
    void *ptrs[] = {c->grain, c->phase, c->gas, c->euler, c->wgtc, c->sg, c->sgt, c->velgrad,
                    c->velgradold, c->dvelgradref, c->defgrad, c->defgradinv, c->detF, c->defgrade,
                    c->defgradp, c->ag, c->cg66, c->sg66, c->c066mod, c->cgm, c->cgas, c->sch,
                    c->crss, c->trialtau, c->gamdot, c->gacumgr, c->edotp, c->ept, c->disgradsym,
                    c->epvm, c->defgradinc, c->defgradeinc, c->s_velgrad, c->s_velgradold, c->s_sg,
                    c->s_edotp, c->s_gamdot, c->s_trialtau, c->Goper, c->xnode, c->four, c->eth};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; i++) free(ptrs[i]);
}
void orc_destroy(orc_ctx *c) {
    if (!c) return;
    freeall(c);
    for (int p = 0; p < MAXPH; p++) {
        phase_t *ph = &c->ph[p];
        free(ph->dnca); free(ph->dbca); free(ph->tau); free(ph->thet); free(ph->nrs);
        free(ph->gamd0); free(ph->hpfac); free(ph->hard);
    }
    free(c->proc);
    free(c);
}
const char *orc_error(orc_ctx *c) { return c ? c->err : "null context"; }

static double *dup(const double *a, size_t n) {
    double *p = malloc(sizeof(double) * (n ? n : 1));
    if (a && n) memcpy(p, a, sizeof(double) * n);
    return p;
}

int orc_set_phase(orc_ctx *c, int iph, int igas, int nsyst, int nmodes, const int *nsm,
                  const double *dnca, const double *dbca, const double *tau, const double *thet,
                  const double *nrs, const double *gamd0, const double *hpfac, const double *hard,
                  const double *C) {
    if (iph < 1 || iph > c->nph) return -1;
    phase_t *p = &c->ph[iph - 1];
    p->igas = igas;
    if (igas) return 0;
    p->nsyst = nsyst; p->nmodes = nmodes;
    for (int m = 0; m < nmodes && m < 16; m++) p->nsm[m] = nsm[m];
    p->dnca = dup(dnca, 3 * nsyst); p->dbca = dup(dbca, 3 * nsyst);
    p->tau = dup(tau, 3 * nsyst); p->thet = dup(thet, 2 * nsyst);
    p->nrs = dup(nrs, nsyst); p->gamd0 = dup(gamd0, nsyst); p->hpfac = dup(hpfac, nsyst);
    p->hard = dup(hard, (size_t)nsyst * nsyst);
    memcpy(p->C, C, sizeof p->C);
    if (nsyst > c->ns) c->ns = nsyst;
    return 0;
}

int orc_set_voxels(orc_ctx *c, const double *euler, const int *grain, const int *phase) {
    long V = c->V;
    c->euler = dup(euler, 3 * V);
    c->grain = malloc(sizeof(int) * V);
    c->phase = malloc(sizeof(int) * V);
    memcpy(c->grain, grain, sizeof(int) * V);
    memcpy(c->phase, phase, sizeof(int) * V);
    return 0;
}

/* the initial displacement file's values (read_input, IO_functions.f90:147-154): before orc_init */
int orc_set_displacement(orc_ctx *c, int idispini, const double *disp) {
    if (idispini != 1 && idispini != 2) { snprintf(c->err, sizeof c->err, "idispini must be 1 or 2"); return -1; }
    long n = idispini == 1 ? c->V : (long)(c->n1 + 1) * (c->n2 + 1) * (c->n3 == 1 ? 1 : c->n3 + 1);
    free(c->dispini);
    c->dispini = dup(disp, 3 * n);
    c->idispini = idispini;
    return c->dispini ? 0 : -2;
}

int orc_add_process(orc_ctx *c, const orc_process *p) {
    if (c->nproc == c->proc_cap) {
        c->proc_cap *= 2;
        c->proc = realloc(c->proc, sizeof(orc_process) * c->proc_cap);
    }
    c->proc[c->nproc++] = *p;
    return 0;
}

static int alloc_state(orc_ctx *c) {
    long V = c->V;
    int ns = c->ns > 0 ? c->ns : 1;
    int n3node = c->n3 == 1 ? 1 : c->n3 + 1;
#define AL(p, n) do { (p) = calloc((size_t)(n), sizeof(*(p))); if (!(p)) return -1; } while (0)
    AL(c->gas, V); AL(c->wgtc, V); AL(c->sg, 9 * V); AL(c->sgt, 9 * V); AL(c->velgrad, 9 * V);
    AL(c->velgradold, 9 * V); AL(c->dvelgradref, 9 * V); AL(c->defgrad, 9 * V);
    AL(c->defgradinv, 9 * V); AL(c->detF, V); AL(c->defgrade, 9 * V); AL(c->defgradp, 9 * V);
    AL(c->ag, 9 * V); AL(c->cg66, 36 * V); AL(c->sg66, 36 * V); AL(c->c066mod, 36 * V);
    if (c->modal) AL(c->cgm, 36 * V);
    AL(c->cgas, 36 * V);
    AL(c->sch, (size_t)5 * ns * V); AL(c->crss, (size_t)2 * ns * V); AL(c->trialtau, (size_t)2 * ns * V);
    AL(c->gamdot, (size_t)ns * V); AL(c->gacumgr, V);
    AL(c->edotp, 9 * V); AL(c->ept, 9 * V); AL(c->disgradsym, 9 * V); AL(c->epvm, V); AL(c->eth, 9 * V);
    AL(c->defgradinc, 9 * V); AL(c->defgradeinc, 9 * V);
    AL(c->s_velgrad, 9 * V); AL(c->s_velgradold, 9 * V); AL(c->s_sg, 9 * V); AL(c->s_edotp, 9 * V);
    AL(c->s_gamdot, (size_t)ns * V); AL(c->s_trialtau, (size_t)2 * ns * V);
    AL(c->Goper, 81 * V);
    AL(c->xnode, 3L * (c->n1 + 1) * (c->n2 + 1) * n3node);
    AL(c->four, 9 * V);
#undef AL
    return 0;
}

/* deterministic weighted reductions: per z-plane serial sums, then planes in order */
#define PLANE_REDUCE(c, NV, ...)                                                           \
    do {                                                                                    \
        int n3_ = (c)->n3;                                                                  \
        long np_ = (long)(c)->n1 * (c)->n2;                                                 \
        double *part_ = calloc((size_t)n3_ * (NV), sizeof(double));                         \
        _Pragma("omp parallel for schedule(static)")                                        \
        for (int kz_ = 0; kz_ < n3_; kz_++) {                                               \
            double *acc = part_ + (size_t)kz_ * (NV);                                       \
            for (long q_ = 0; q_ < np_; q_++) {                                             \
                long v = kz_ * np_ + q_;                                                    \
                __VA_ARGS__                                                                 \
            }                                                                               \
        }                                                                                   \
        for (int kz_ = 0; kz_ < n3_; kz_++)                                                 \
            for (int t_ = 0; t_ < (NV); t_++) red_[t_] += part_[(size_t)kz_ * (NV) + t_];   \
        free(part_);                                                                        \
    } while (0)

/* gradient_tensor1 (various_functions.f90:56-112): u [v][3] -> grad [v][9], grad(l,m) = d_m u_l,
 * spectrally: the forward c2c transform of the three components, times i sin(xi_m) (ideriv 1) or
 * tan(xi_m/2) (i/4) prod_a (1 + e^{i xi_a}) (ideriv 2, zero at the all-Nyquist point of a 3D
 * grid), inverse, real part / V */
void orc_gradient_tensor1(const double *u, double *grad, int ideriv, int n1, int n2, int n3) {
    long V = (long)n1 * n2 * n3;
    cplx *f = malloc(sizeof(cplx) * 3 * V), *g = malloc(sizeof(cplx) * 9 * V);
    for (long v = 0; v < V; v++)
        for (int l = 0; l < 3; l++) f[l * V + v] = u[3 * v + l];
    orc_fft3d((double *)f, n1, n2, n3, 3, -1);
    for (int k = 1; k <= n3; k++)
        for (int j = 1; j <= n2; j++)
            for (int i = 1; i <= n1; i++) {
                long v = (i - 1) + (long)n1 * ((j - 1) + (long)n2 * (k - 1));
                double xk[3];
                orc_spatial_freq(i, j, k, n1, n2, n3, xk);
                cplx km[3];
                if (ideriv == 2) {
                    cplx base = I * 0.25 * (1.0 + cexp(I * xk[0])) * (1.0 + cexp(I * xk[1])) * (1.0 + cexp(I * xk[2]));
                    for (int l = 0; l < 3; l++) km[l] = tan(xk[l] / 2.0) * base;
                } else {
                    for (int l = 0; l < 3; l++) km[l] = I * sin(xk[l]);
                }
                int allnyq = (n1 % 2 == 0 && i == n1 / 2 + 1) && (n2 % 2 == 0 && j == n2 / 2 + 1) &&
                             (n3 % 2 == 0 && k == n3 / 2 + 1 && n3 > 1) && ideriv == 2;
                for (int l = 0; l < 3; l++)
                    for (int m = 0; m < 3; m++) g[(3 * l + m) * V + v] = allnyq ? 0.0 : f[l * V + v] * km[m];
            }
    orc_fft3d((double *)g, n1, n2, n3, 9, +1);
    for (long v = 0; v < V; v++)
        for (int t = 0; t < 9; t++) grad[9 * v + t] = creal(g[t * V + v]) / (double)V;
    free(f);
    free(g);
}

/* cell2node_tensor1 (various_functions.f90:2652-2691): node (i, j, k) = mean of the 8 cells
 * around it, periodic.  (The reference wraps the x index against npts2, which is the same as
 * against npts1 on the grids where its loop stays in bounds, n1 = n2.) */
static void cell2node3(const double *cell, double *node, int n1, int n2, int n3) {
    for (int k = 1; k <= n3 + 1; k++) {
        int k1 = k - 1 == 0 ? n3 : k - 1, k2 = k == n3 + 1 ? 1 : k;
        for (int j = 1; j <= n2 + 1; j++) {
            int j1 = j - 1 == 0 ? n2 : j - 1, j2 = j == n2 + 1 ? 1 : j;
            for (int i = 1; i <= n1 + 1; i++) {
                int i1 = i - 1 == 0 ? n1 : i - 1, i2 = i == n1 + 1 ? 1 : i;
#define CELL(a, b, cc) (cell + 3 * ((a - 1) + (long)n1 * ((b - 1) + (long)n2 * (cc - 1))))
                const double *p[8] = {CELL(i2, j2, k2), CELL(i1, j2, k2), CELL(i2, j1, k2), CELL(i2, j2, k1),
                                      CELL(i1, j1, k2), CELL(i1, j2, k1), CELL(i2, j1, k1), CELL(i1, j1, k1)};
#undef CELL
                double *o = node + 3 * ((i - 1) + (long)(n1 + 1) * ((j - 1) + (long)(n2 + 1) * (k - 1)));
                for (int l = 0; l < 3; l++) {
                    double s = p[0][l];
                    for (int q = 1; q < 8; q++) s += p[q][l];
                    o[l] = s * 0.125;
                }
            }
        }
    }
}

/* init_defgrad (IO_functions.f90:815-936): F = grad(u0) + diag(delt) with delt = (4, 4, 4) (the
 * value load_input gives it, :62-66), F0 = F, the J-weighted averages of F^-1 and the phase
 * weights, w_c = wgt J / <J>; the nodal coordinates move by the nodal displacement */
static void init_defgrad(orc_ctx *c) {
    long V = c->V;
    int n1 = c->n1, n2 = c->n2, n3 = c->n3, n3node = n3 == 1 ? 1 : n3 + 1;
    long nnode = (long)(n1 + 1) * (n2 + 1) * (n3 + 1);
    double *u = malloc(sizeof(double) * 3 * V), *node = calloc((size_t)3 * nnode, sizeof(double));
    double *grad = malloc(sizeof(double) * 9 * V);
    int ideriv;
    if (c->idispini == 1) {
        memcpy(u, c->dispini, sizeof(double) * 3 * V);
        cell2node3(u, node, n1, n2, n3);
        ideriv = 1;
    } else {
        long nn = (long)(n1 + 1) * (n2 + 1) * n3node;
        memcpy(node, c->dispini, sizeof(double) * 3 * nn);
        for (int k = 0; k < n3; k++)
            for (int j = 0; j < n2; j++)
                for (int i = 0; i < n1; i++)
                    for (int l = 0; l < 3; l++)
                        u[3 * (i + (long)n1 * (j + (long)n2 * k)) + l] =
                            node[3 * (i + (long)(n1 + 1) * (j + (long)(n2 + 1) * k)) + l];
        ideriv = 2;
    }
    orc_gradient_tensor1(u, grad, ideriv, n1, n2, n3);
    for (int k = 0; k < n3node; k++)
        for (int j = 0; j < n2 + 1; j++)
            for (int i = 0; i < n1 + 1; i++)
                for (int l = 0; l < 3; l++) {
                    long q = 3L * (i + (long)(n1 + 1) * (j + (long)(n2 + 1) * k)) + l;
                    c->xnode[q] += node[q];
                }
    memset(c->defgradavg, 0, sizeof c->defgradavg);
    c->defgradavg[0] = c->defgradavg[4] = c->defgradavg[8] = 4.0;
    double inv_c[9] = {0}, inv_s[9] = {0}, inv_g[9] = {0}, detFavg = 0.0;
    c->wphcsol = c->wphcgas = 0.0;
    if (!c->defgradini) c->defgradini = malloc(sizeof(double) * 9 * V);
    for (long v = 0; v < V; v++) {
        double *F = c->defgrad + 9 * v, Fi[9];
        for (int t = 0; t < 9; t++) F[t] = grad[9 * v + t] + c->defgradavg[t];
        memcpy(c->defgradini + 9 * v, F, sizeof Fi);
        c->detF[v] = det33(F);
        memcpy(Fi, F, sizeof Fi);
        lu_inverse(Fi, 3);
        memcpy(c->defgradinv + 9 * v, Fi, sizeof Fi);
        const double wj = c->wgt * c->detF[v];
        for (int t = 0; t < 9; t++) {
            inv_c[t] += Fi[t] * wj;
            if (c->gas[v]) inv_g[t] += Fi[t] * wj; else inv_s[t] += Fi[t] * wj;
        }
        if (c->gas[v]) c->wphcgas += wj; else c->wphcsol += wj;
        c->wgtc[v] = wj;
        detFavg += wj;
    }
    c->wphcsol /= detFavg;
    c->wphcgas /= detFavg;
    for (int t = 0; t < 9; t++) {
        c->defgradinvavgc[t] = inv_c[t] / detFavg;
        c->defgradinvavgcs[t] = inv_s[t] / detFavg / c->wphcsol;
        c->defgradinvavgcg[t] = c->wphcgas > 0.0 ? inv_g[t] / detFavg / c->wphcgas : ID3[t];
    }
    for (long v = 0; v < V; v++) c->wgtc[v] /= detFavg;
    free(u);
    free(node);
    free(grad);
}

/* ------------------------------------------------------------------------ */
/* load_input tail + data_grain (IO_functions.f90:100-199, 759-809, 347-363)  */
/* ------------------------------------------------------------------------ */
int orc_init(orc_ctx *c) {
    if (!c->euler || c->nproc < 1) { snprintf(c->err, sizeof c->err, "voxels/processes not set"); return -1; }
    if (alloc_state(c)) { snprintf(c->err, sizeof c->err, "out of memory"); return -2; }
    long V = c->V;
    int ns = c->ns;
    /* init_micro (types.f90:256-343) */
    for (long v = 0; v < V; v++) {
        c->wgtc[v] = c->wgt;
        memcpy(c->defgrad + 9 * v, ID3, sizeof ID3);
        memcpy(c->defgrade + 9 * v, ID3, sizeof ID3);
        memcpy(c->defgradp + 9 * v, ID3, sizeof ID3);
        memcpy(c->defgradinv + 9 * v, ID3, sizeof ID3);
        memcpy(c->defgradinc + 9 * v, ID3, sizeof ID3);
        memcpy(c->defgradeinc + 9 * v, ID3, sizeof ID3);
        c->detF[v] = 1.0;
    }
    int n3node = c->n3 == 1 ? 1 : c->n3 + 1;
    for (int k = 0; k < n3node; k++)
        for (int j = 0; j < c->n2 + 1; j++)
            for (int i = 0; i < c->n1 + 1; i++) {
                long q = 3L * (i + (long)(c->n1 + 1) * (j + (long)(c->n2 + 1) * k));
                c->xnode[q] = i + 1 - 0.5;
                c->xnode[q + 1] = j + 1 - 0.5;
                c->xnode[q + 2] = (n3node == 1) ? 0.0 : k + 1 - 0.5;
            }
    memcpy(c->defgradavg, ID3, sizeof ID3);
    memcpy(c->defgradinvavgc, ID3, sizeof ID3);
    memcpy(c->defgradinvavgcs, ID3, sizeof ID3);
    memcpy(c->defgradinvavgcg, ID3, sizeof ID3);

    /* data_grain: ag, cg66, sg66 (independent per voxel) */
    for (long v = 0; v < V; v++)
        if (c->phase[v] < 1 || c->phase[v] > c->nph) { snprintf(c->err, sizeof c->err, "phase id out of range"); return -3; }
    #pragma omp parallel for schedule(static)
    for (long v = 0; v < V; v++) {
        int jph = c->phase[v];
        phase_t *P = &c->ph[jph - 1];
        c->gas[v] = (unsigned char)(P->igas != 0);
        if (P->igas) continue;
        double aa[9];
        const double *e = c->euler + 3 * v;
        euler2(e[0] * PI / 180.0, e[1] * PI / 180.0, e[2] * PI / 180.0, aa);
        tr33(aa, c->ag + 9 * v);
        double cr[81], c66[36];
        for (int i1 = 0; i1 < 3; i1++) for (int j1 = 0; j1 < 3; j1++)
        for (int k1 = 0; k1 < 3; k1++) for (int l1 = 0; l1 < 3; l1++) {
            double dum = 0;
            for (int i2 = 0; i2 < 3; i2++) for (int j2 = 0; j2 < 3; j2++)
            for (int k2 = 0; k2 < 3; k2++) for (int l2 = 0; l2 < 3; l2++)
                dum += aa[3 * i2 + i1] * aa[3 * j2 + j1] * aa[3 * k2 + k1] * aa[3 * l2 + l1] *
                       P->C[27 * i2 + 9 * j2 + 3 * k2 + l2];
            cr[27 * i1 + 9 * j1 + 3 * k1 + l1] = dum;
        }
        B4(c66, cr);
        memcpy(c->cg66 + 36 * v, c66, sizeof c66);
        lu_inverse(c66, 6);
        memcpy(c->sg66 + 36 * v, c66, sizeof c66);
    }
    if (c->idispini > 0) init_defgrad(c); /* read_input, IO_functions.f90:152-154 */
    /* Hall-Petch startup shift (IO_functions.f90:107-199) */
    const double hp_k = 632.0 / 2.5, voxel_vol = 64.0;
    const double grain_dmin = 3.0 * pow(voxel_vol, 1.0 / 3.0);
    int gmax = 0;
    for (long v = 0; v < V; v++) if (c->grain[v] > gmax) gmax = c->grain[v];
    long *nvox = calloc((size_t)(gmax + 1) * c->nph, sizeof(long));
    for (long v = 0; v < V; v++)
        if (!c->gas[v]) {
            if (c->grain[v] < 1) { free(nvox); snprintf(c->err, sizeof c->err, "grain id out of bounds"); return -4; }
            nvox[(long)c->grain[v] * c->nph + c->phase[v] - 1]++;
        }
    for (long v = 0; v < V; v++) {
        c->gacumgr[v] = 0.0;
        if (c->gas[v]) continue;
        phase_t *P = &c->ph[c->phase[v] - 1];
        double hp = 0.0;
        long nv = nvox[(long)c->grain[v] * c->nph + c->phase[v] - 1];
        if (nv > 0) {
            double gd = pow((double)nv * voxel_vol, 1.0 / 3.0);
            if (gd < grain_dmin) gd = grain_dmin;
            hp = hp_k / sqrt(gd);
        }
        for (int i = 0; i < P->nsyst; i++) {
            double *cr = c->crss + ((size_t)v * ns + i) * 2, *tt = c->trialtau + ((size_t)v * ns + i) * 2;
            cr[0] = P->tau[3 * i] + hp;
            cr[1] = P->tau[3 * i + 1] + hp;
            double dbsa[3];
            mv33(c->ag + 9 * v, P->dbca + 3 * i, dbsa);
            double cosal = fabs(dbsa[1]);
            cr[0] += P->hpfac[i] * sqrt(cosal);
            cr[1] += P->hpfac[i] * sqrt(cosal);
            tt[0] = cr[1]; /* trialtau(:,1) from crss(:,2), IO_functions.f90:193 */
            tt[1] = cr[1];
        }
    }
    free(nvox);
    c->wphsol = c->wphgas = 0;
    for (long v = 0; v < V; v++) { if (c->gas[v]) c->wphgas += c->wgt; else c->wphsol += c->wgt; }
    if (c->idispini == 0) { /* IO_functions.f90:358-361 (else init_defgrad's J-weighted ones) */
        c->wphcgas = c->wphgas;
        c->wphcsol = c->wphsol;
    }

    /* initial_guess (various_functions.f90:8-53) */
    orc_process *p1 = &c->proc[0];
    for (long v = 0; v < V; v++) {
        memcpy(c->velgrad + 9 * v, p1->udot, sizeof(double) * 9);
        if (c->gas[v]) continue;
        double u6[6], s6[6];
        B2(u6, p1->udot);
        for (int a = 0; a < 6; a++) {
            double s = 0;
            for (int b = 0; b < 6; b++) s += c->cg66[36 * v + 6 * a + b] * (u6[b] * p1->tdot);
            s6[a] = s;
        }
        B1(s6, c->sg + 9 * v);
    }
    c->dvmavg = vm_strain(p1->dsim);
    c->lvmavg = c->dvmavg;

    /* initial reference stiffness (LS-EVPFFT.f90:156-173) */
    memset(c->c066, 0, sizeof c->c066);
    for (long v = 0; v < V; v++)
        if (!c->gas[v])
            for (int t = 0; t < 36; t++) c->c066[t] += c->cg66[36 * v + t] * c->wgtc[v];
    for (int t = 0; t < 36; t++) c->c066[t] = c->c066[t] / c->wphcsol * p1->tdot;
    memcpy(c->s066, c->c066, sizeof c->c066);
    lu_inverse(c->s066, 6);
    B3(c->c066, c->c0);
    B3(c->s066, c->s0);
    memcpy(c->de_ds, c->s0, sizeof c->s0);

    c->total_macrosteps = 0;
    for (int i = 0; i < c->nproc; i++) c->total_macrosteps += c->proc[i].nsteps;
    c->imacroloop = 1;
    c->ipr = -1;
    c->proc_active = 0;
    c->sim_time = 0;
    c->initialised = 1;
    if (g_singular) { snprintf(c->err, sizeof c->err, "Singular matrix in lu_inverse"); return 1; }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* per-increment operators                                                   */
/* ------------------------------------------------------------------------ */

/* take_snapshot / rollback (various_functions.f90:2839-2895) */
static void take_snapshot(orc_ctx *c) {
    orc_process *P = &c->proc[c->ipr];
    memcpy(c->dsim_snap, P->dsim, sizeof c->dsim_snap);
    memcpy(c->scauchy_snap, P->scauchy, sizeof c->scauchy_snap);
    size_t V = c->V, ns = c->ns;
    memcpy(c->s_velgrad, c->velgrad, 9 * V * sizeof(double));
    memcpy(c->s_velgradold, c->velgradold, 9 * V * sizeof(double));
    memcpy(c->s_sg, c->sg, 9 * V * sizeof(double));
    memcpy(c->s_edotp, c->edotp, 9 * V * sizeof(double));
    memcpy(c->s_gamdot, c->gamdot, ns * V * sizeof(double));
    memcpy(c->s_trialtau, c->trialtau, 2 * ns * V * sizeof(double));
}
static void rollback(orc_ctx *c) {
    orc_process *P = &c->proc[c->ipr];
    memcpy(P->dsim, c->dsim_snap, sizeof c->dsim_snap);
    memcpy(P->scauchy, c->scauchy_snap, sizeof c->scauchy_snap);
    size_t V = c->V, ns = c->ns;
    memcpy(c->velgrad, c->s_velgrad, 9 * V * sizeof(double));
    memcpy(c->velgradold, c->s_velgradold, 9 * V * sizeof(double));
    memcpy(c->sg, c->s_sg, 9 * V * sizeof(double));
    memcpy(c->edotp, c->s_edotp, 9 * V * sizeof(double));
    memcpy(c->gamdot, c->s_gamdot, ns * V * sizeof(double));
    memcpy(c->trialtau, c->s_trialtau, 2 * ns * V * sizeof(double));
}

/* update_schmid (various_functions.f90:472-539) */
static void update_schmid(orc_ctx *c) {
    long V = c->V;
    int ns = c->ns;
    #pragma omp parallel for schedule(static)
    for (long v = 0; v < V; v++) {
        if (c->gas[v]) continue;
        phase_t *P = &c->ph[c->phase[v] - 1];
        for (int is = 0; is < P->nsyst; is++) {
            double s[9], aa[9], ai[9], t[9], sy[9];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) s[3 * i + j] = P->dbca[3 * is + i] * P->dnca[3 * is + j];
            mm33(c->defgrade + 9 * v, c->ag + 9 * v, aa);
            memcpy(ai, aa, sizeof aa);
            lu_inverse(ai, 3);
            mm33(aa, s, t);
            mm33(t, ai, t);
            sym33(t, sy);
            orc_chg_basis(c->sch + ((size_t)v * ns + is) * 5, sy, NULL, NULL, 2, 5);
        }
    }
}

/* Green operator at one frequency (various_functions.f90:640-742), igamma=0 */
void orc_green_op(const double *c0, const double *xi, double *g1) {
    double D[9], a[9];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) D[3 * i + j] = xi[i] * xi[j];
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) {
            double s = 0;
            for (int j = 0; j < 3; j++) for (int l = 0; l < 3; l++) s += c0[27 * i + 9 * j + 3 * k + l] * D[3 * j + l];
            a[3 * i + k] = s;
        }
    lu_inverse(a, 3);
    for (int p = 0; p < 3; p++) for (int q = 0; q < 3; q++) for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++)
        g1[27 * p + 9 * q + 3 * i + j] = -a[3 * p + i] * D[3 * q + j];
}

static int is_nyq(int idx1, int n) { return (n % 2 == 0) && idx1 == n / 2 + 1; }

/* form_G (various_functions.f90:609-809) for igamma 0, 1, 2 at 1-based (kxx,kyy,kzz) */
static void green_at(orc_ctx *c, int igamma, int kxx, int kyy, int kzz, double *g1) {
    int n1 = c->n1, n2 = c->n2, n3 = c->n3;
    double xk[3];
    orc_spatial_freq(kxx, kyy, kzz, n1, n2, n3, xk);
    int nx = is_nyq(kxx, n1), ny = is_nyq(kyy, n2), nz = is_nyq(kzz, n3) && n3 > 1;
    if ((nx || ny || nz) && igamma == 0) {
        for (int t = 0; t < 81; t++) g1[t] = -c->s0[t];
    } else if (((nx && ny) || (nx && is_nyq(kzz, n3)) || (ny && is_nyq(kzz, n3))) && igamma >= 2) {
        memset(g1, 0, 81 * sizeof(double));
    } else {
        double D[9], a[9];
        if (igamma == 0) {
            for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) D[3 * i + j] = xk[i] * xk[j];
        } else if (igamma == 1) {
            double xs = 2.0 * PI * (kxx - 1) / n1, ys = 2.0 * PI * (kyy - 1) / n2, zs = 2.0 * PI * (kzz - 1) / n3;
            D[0] = 2.0 * (cos(xs) - 1.0); D[4] = 2.0 * (cos(ys) - 1.0); D[8] = 2.0 * (cos(zs) - 1.0);
            D[3] = D[1] = -sin(xs) * sin(ys);
            D[6] = D[2] = -sin(xs) * sin(zs);
            D[7] = D[5] = -sin(ys) * sin(zs);
        } else {
            cplx km[3], f = I * 0.25 * (1.0 + cexp(I * xk[0])) * (1.0 + cexp(I * xk[1])) * (1.0 + cexp(I * xk[2]));
            for (int i = 0; i < 3; i++) km[i] = tan(xk[i] / 2.0) * f;
            for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) D[3 * i + j] = creal(km[i] * conj(km[j]));
        }
        if (kxx == 1 && kyy == 1 && kzz == 1) {
            memset(a, 0, sizeof a);
        } else {
            for (int i = 0; i < 3; i++)
                for (int k = 0; k < 3; k++) {
                    double s = 0;
                    for (int j = 0; j < 3; j++) for (int l = 0; l < 3; l++) s += c->c0[27 * i + 9 * j + 3 * k + l] * D[3 * j + l];
                    a[3 * i + k] = s;
                }
            lu_inverse(a, 3);
        }
        for (int p = 0; p < 3; p++) for (int q = 0; q < 3; q++) for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++)
            g1[27 * p + 9 * q + 3 * i + j] = -a[3 * p + i] * D[3 * q + j];
    }
    if (igamma == 1) { /* four smoothing iterations G <- -G:C0:G (:744-780) */
        double gt[81], aux[81];
        memcpy(gt, g1, sizeof gt);
        for (int itt = 0; itt < 4; itt++) {
            for (int ij = 0; ij < 9; ij++)
                for (int kl = 0; kl < 9; kl++) {
                    double dum = 0;
                    for (int mn = 0; mn < 9; mn++)
                        for (int op = 0; op < 9; op++) dum -= g1[9 * ij + mn] * c->c0[9 * mn + op] * gt[9 * op + kl];
                    aux[9 * ij + kl] = dum;
                }
            memcpy(gt, aux, sizeof gt);
        }
        memcpy(g1, gt, sizeof gt);
    }
}
int orc_green_at(orc_ctx *c, int kx, int ky, int kz, double *g81) {
    green_at(c, c->proc[c->ipr < 0 ? 0 : c->ipr].igamma, kx, ky, kz, g81);
    return 0;
}

/* form_G + Goperr0 = real(IFFT(Gamma))(origin) = (1/V) sum_xi Gamma(xi) */
static void form_G(orc_ctx *c) {
    int igamma = c->proc[c->ipr].igamma;
    int n1 = c->n1, n2 = c->n2, n3 = c->n3;
    #pragma omp parallel for schedule(static)
    for (int kz = 1; kz <= n3; kz++)
        for (int ky = 1; ky <= n2; ky++)
            for (int kx = 1; kx <= n1; kx++) {
                long v = (kx - 1) + (long)n1 * ((ky - 1) + (long)n2 * (kz - 1));
                green_at(c, igamma, kx, ky, kz, c->Goper + 81 * v);
            }
    double red_[81] = {0};
    PLANE_REDUCE(c, 81, { for (int t = 0; t < 81; t++) acc[t] += c->Goper[81 * v + t]; });
    for (int t = 0; t < 81; t++) c->Goperr0[t] = red_[t] / (double)c->V;
}

/* calc_c066mod_Goperr066mod (various_functions.f90:812-906) */
static void calc_c066mod(orc_ctx *c) {
    long V = c->V;
    #pragma omp parallel for schedule(static)
    for (long v = 0; v < V; v++) {
        const double *F = c->defgrad + 9 * v, *Fi = c->defgradinv + 9 * v;
        double J = c->detF[v];
        double cm[81], cs[81], c66[36];
        for (int ii = 0; ii < 3; ii++) for (int jj = 0; jj < 3; jj++) for (int kk = 0; kk < 3; kk++) for (int ll = 0; ll < 3; ll++) {
            double dum = 0;
            for (int m = 0; m < 3; m++) for (int n = 0; n < 3; n++)
                dum += c->c0[27 * ii + 9 * m + 3 * kk + n] * F[3 * jj + m] * F[3 * ll + n];
            cm[27 * ii + 9 * jj + 3 * kk + ll] = dum / J;
        }
        for (int ii = 0; ii < 3; ii++) for (int jj = 0; jj < 3; jj++) for (int kk = 0; kk < 3; kk++) for (int ll = 0; ll < 3; ll++)
            cs[27 * ii + 9 * jj + 3 * kk + ll] = 0.25 * (cm[27 * ii + 9 * jj + 3 * kk + ll] + cm[27 * jj + 9 * ii + 3 * kk + ll] +
                                                        cm[27 * ii + 9 * jj + 3 * ll + kk] + cm[27 * jj + 9 * ii + 3 * ll + kk]);
        B4(c66, cs);
        memcpy(c->c066mod + 36 * v, c66, sizeof c66);
        if (c->modal) {
            double gm[81], gs[81], g66[36];
            for (int ii = 0; ii < 3; ii++) for (int jj = 0; jj < 3; jj++) for (int kk = 0; kk < 3; kk++) for (int ll = 0; ll < 3; ll++) {
                double dum = 0;
                for (int m = 0; m < 3; m++) for (int n = 0; n < 3; n++)
                    dum += c->Goperr0[27 * ii + 9 * m + 3 * kk + n] * Fi[3 * m + jj] * Fi[3 * n + ll];
                gm[27 * ii + 9 * jj + 3 * kk + ll] = dum * J;
            }
            for (int ii = 0; ii < 3; ii++) for (int jj = 0; jj < 3; jj++) for (int kk = 0; kk < 3; kk++) for (int ll = 0; ll < 3; ll++)
                gs[27 * ii + 9 * jj + 3 * kk + ll] = 0.25 * (gm[27 * ii + 9 * jj + 3 * kk + ll] + gm[27 * jj + 9 * ii + 3 * kk + ll] +
                                                            gm[27 * ii + 9 * jj + 3 * ll + kk] + gm[27 * jj + 9 * ii + 3 * ll + kk]);
            B4(g66, gs);
            for (int a = 0; a < 6; a++) for (int b = 0; b < 6; b++) {
                double s = 0;
                for (int k = 0; k < 6; k++) s += c66[6 * a + k] * g66[6 * k + b];
                c->cgm[36 * v + 6 * a + b] = s;
            }
        }
    }
}

/* calc_gas_compl (various_functions.f90:909-951) */
static void calc_gas_compl(orc_ctx *c) {
    double tdot = c->proc[c->ipr].tdot;
    for (long v = 0; v < c->V; v++) {
        double *cg = c->cgas + 36 * v;
        if (!c->gas[v]) { memset(cg, 0, 36 * sizeof(double)); continue; }
        double a[36] = {0};
        for (int i = 0; i < 6; i++) a[7 * i] = 1.0;
        for (int t = 0; t < 36; t++) a[t] += c->c066mod[36 * v + t] / tdot * c->complgas;
        if (c->modal) for (int t = 0; t < 36; t++) a[t] -= c->cgm[36 * v + t];
        lu_inverse(a, 6);
        memcpy(cg, a, sizeof a);
    }
}

/* ------------------------------------------------------------------------ */
/* constitutive solve (various_functions.f90:1142-1491)                       */
/* ------------------------------------------------------------------------ */
/* plastic_strain_rate (:1367-1419) */
void orc_plastic_strain_rate(const double *sg, const double *sch, const double *trialtau,
                             const double *gamd0, const double *nrs, int ns, double *edotp,
                             double *dedotp, double *gamdot) {
    memset(dedotp, 0, 36 * sizeof(double));
    memset(edotp, 0, 6 * sizeof(double));
    for (int is = 0; is < ns; is++) {
        const double *s = sch + 5 * is;
        double rss = s[0] * sg[0] + s[1] * sg[1] + s[2] * sg[2] + s[3] * sg[3] + s[4] * sg[4];
        int isign = (rss - 0.0 < 0.0) ? 1 : 0; /* kin == 0 */
        rss = (rss - 0.0) / trialtau[2 * is + isign];
        double dum = gamd0[is] * pow(fabs(rss), nrs[is] - 1.0);
        double rss1 = nrs[is] * dum / trialtau[2 * is + isign];
        double rss2 = rss * dum;
        gamdot[is] = rss2;
        for (int jj = 0; jj < 5; jj++) {
            edotp[jj] += s[jj] * rss2;
            for (int ii = 0; ii < 5; ii++) dedotp[6 * ii + jj] += s[ii] * s[jj] * rss1;
        }
    }
}

/* harden (:1439-1491); tau out (ns,2) */
static void harden(const orc_ctx *c, const phase_t *P, const double *gamdot, const double *crss,
                   double gacumgr, double tdot, double *tau, double *gamma1) {
    double deltgam = 0;
    for (int is = 0; is < P->nsyst; is++) deltgam += fabs(gamdot[is]) * tdot;
    double gamma0 = gacumgr;
    *gamma1 = gamma0 + deltgam;
    int is = 0, islast = 0;
    for (int t = 0; t < 2 * c->ns; t++) tau[t] = 0.0;
    for (int mode = 0; mode < P->nmodes; mode++) {
        islast += P->nsm[mode];
        double tau1 = P->tau[3 * (islast - 1) + 2];
        double thet0 = P->thet[2 * (islast - 1)], thet1 = P->thet[2 * (islast - 1) + 1];
        double fact = fabs(thet0 / tau1);
        double exp0 = exp(-fact * gamma0), exp1 = exp(-fact * (*gamma1));
        double voce = deltgam * thet1 + tau1 * (exp0 - exp1) + thet1 * (gamma0 * exp0 - (*gamma1) * exp1);
        for (int ism = 0; ism < P->nsm[mode]; ism++, is++) {
            double dtau = 0;
            for (int js = 0; js < P->nsyst; js++) dtau += P->hard[P->nsyst * is + js] * fabs(gamdot[js]);
            tau[2 * is] = crss[2 * is] + dtau * voce * tdot / deltgam;
            tau[2 * is + 1] = crss[2 * is + 1] + dtau * voce * tdot / deltgam;
        }
    }
}

typedef struct { double dsgnorm, ddgnorm; long nr, eval, outer; } vox_res;

/* solve_res for one voxel; updates sg, edotp, gamdot, trialtau in place */
static vox_res solve_voxel(orc_ctx *c, long v) {
    const orc_process *Pr = &c->proc[c->ipr];
    int ns = c->ns;
    vox_res out = {0, 0, 0, 0, 0};
    double lambda[6], sg[6], sgt[6], ls[9], vs[6];
    B2(lambda, c->sg + 9 * v);
    memcpy(sg, lambda, sizeof sg);
    B2(sgt, c->sgt + 9 * v);
    sym33(c->velgrad + 9 * v, ls);
    B2(vs, ls);
    const double *cm = c->c066mod + 36 * v, *cg = c->modal ? c->cgm + 36 * v : NULL;
    if (!c->gas[v]) {
        const phase_t *P = &c->ph[c->phase[v] - 1];
        const double *sch = c->sch + (size_t)v * ns * 5, *s66 = c->sg66 + 36 * v;
        double *gd = c->gamdot + (size_t)v * ns, *tt = c->trialtau + (size_t)v * ns * 2;
        const double *crss = c->crss + (size_t)v * ns * 2;
        double tdot = Pr->tdot;
        double sgnorm = norm2v(sg, 6), dgnorm = norm2v(vs, 6);
        (void)dgnorm;
        double taunrm = norm2v(tt, 2 * ns);
        double tol = Pr->error > Pr->erroral ? Pr->error : Pr->erroral;
        int itertau = 0;
        double errtau = 2.0 * tol;
        double edotp[6], ded[36], edote[6], edot[6], dedot[36], res[6], jac[36], sgold[6] = {0};
        double dsgnorm = 0, ddgnorm = 0;
        /* eigenstrain rate of the first increment of a process with ithermo = 1
           (various_functions.f90:1226-1227, added to edot at :1265) */
        const int thermo = c->imicro == 1 && Pr->ithermo == 1;
        double edotth[6] = {0, 0, 0, 0, 0, 0};
        if (thermo) B2(edotth, c->eth + 9 * v);
        while (itertau < 10 && errtau > tol) {
            itertau++;
            out.outer++;
            int iterNR = 0;
            double erral = 2.0 * Pr->erroral, resoldn = 1.0;
            while (iterNR < 100 && erral > Pr->erroral) {
                iterNR++;
                out.nr++;
                int icut = 0;
                double resn = 2.0 * resoldn;
                while (resn > resoldn && icut < 10) {
                    out.eval++;
                    orc_plastic_strain_rate(sg, sch, tt, P->gamd0, P->nrs, P->nsyst, edotp, ded, gd);
                    for (int a = 0; a < 6; a++) {
                        double s = 0;
                        for (int b = 0; b < 6; b++) s += s66[6 * a + b] * ((sg[b] - sgt[b]) / tdot);
                        edote[a] = s;
                    }
                    for (int a = 0; a < 6; a++) edot[a] = edotp[a] + edote[a];
                    if (thermo) for (int a = 0; a < 6; a++) edot[a] += edotth[a];
                    for (int t = 0; t < 36; t++) dedot[t] = ded[t] + s66[t] / tdot;
                    for (int a = 0; a < 6; a++) {
                        double s = 0;
                        for (int b = 0; b < 6; b++) s += cm[6 * a + b] * (edot[b] - vs[b]);
                        res[a] = sg[a] - lambda[a] + s;
                    }
                    if (cg)
                        for (int a = 0; a < 6; a++) {
                            double s = 0;
                            for (int b = 0; b < 6; b++) s += cg[6 * a + b] * (sg[b] - lambda[b]);
                            res[a] -= s;
                        }
                    resn = norm2v(res, 6);
                    if (iterNR == 1) resoldn = resn;
                    if (resn > resoldn)
                        for (int a = 0; a < 6; a++) sg[a] = sgold[a] + (sg[a] - sgold[a]) * 0.5;
                    icut++;
                }
                resoldn = resn;
                memcpy(sgold, sg, sizeof sg);
                for (int a = 0; a < 6; a++)
                    for (int b = 0; b < 6; b++) {
                        double s = (a == b) ? 1.0 : 0.0;
                        for (int k = 0; k < 6; k++) s += cm[6 * a + k] * dedot[6 * k + b];
                        jac[6 * a + b] = s;
                    }
                if (cg) for (int t = 0; t < 36; t++) jac[t] -= cg[t];
                lu_eqsystem(jac, res, 6);
                for (int a = 0; a < 6; a++) sg[a] -= res[a];
                double d1 = 0, d2 = 0, d3 = 0;
                for (int a = 0; a < 6; a++) {
                    d1 += (sg[a] - sgold[a]) * (sg[a] - sgold[a]);
                    d2 += (sg[a] - lambda[a]) * (sg[a] - lambda[a]);
                    d3 += (edot[a] - vs[a]) * (edot[a] - vs[a]);
                }
                dsgnorm = d2;
                ddgnorm = d3;
                erral = sqrt(d1) / sgnorm;
            }
            if (Pr->iuphard == 1) {
                double tnew[2 * 64 + 2], *tn = tnew, g1;
                double *heap = NULL;
                if (2 * ns > (int)(sizeof tnew / sizeof tnew[0])) tn = heap = malloc(sizeof(double) * 2 * ns);
                harden(c, P, gd, crss, c->gacumgr[v], tdot, tn, &g1);
                double e2 = 0;
                for (int t = 0; t < 2 * ns; t++) e2 += (tt[t] - tn[t]) * (tt[t] - tn[t]);
                errtau = sqrt(e2) / taunrm;
                memcpy(tt, tn, sizeof(double) * 2 * ns);
                free(heap);
            } else {
                errtau = 0.0;
            }
        }
        B1(sg, c->sg + 9 * v);
        B1(edotp, c->edotp + 9 * v);
        out.dsgnorm = dsgnorm;
        out.ddgnorm = ddgnorm;
    } else {
        double res[6], t6[6], s6[6];
        double tdot = Pr->tdot;
        for (int a = 0; a < 6; a++) t6[a] = vs[a] + sgt[a] / tdot * c->complgas;
        for (int a = 0; a < 6; a++) {
            double s = 0;
            for (int b = 0; b < 6; b++) s += cm[6 * a + b] * t6[b];
            res[a] = lambda[a] + s;
        }
        if (cg)
            for (int a = 0; a < 6; a++) {
                double s = 0;
                for (int b = 0; b < 6; b++) s += cg[6 * a + b] * lambda[b];
                res[a] -= s;
            }
        for (int a = 0; a < 6; a++) {
            double s = 0;
            for (int b = 0; b < 6; b++) s += c->cgas[36 * v + 6 * a + b] * res[b];
            s6[a] = s;
        }
        B1(s6, c->sg + 9 * v);
        memset(c->edotp + 9 * v, 0, 9 * sizeof(double));
        memset(c->gamdot + (size_t)v * ns, 0, ns * sizeof(double));
    }
    return out;
}
int orc_solve_voxel(orc_ctx *c, long v, double *sg6) {
    vox_res r = solve_voxel(c, v);
    (void)r;
    B2(sg6, c->sg + 9 * v);
    return 0;
}

static int solve_res(orc_ctx *c) {
    long long nr = 0, ev = 0, ou = 0;
    double red_[2] = {0, 0};
    int n3 = c->n3;
    long np = (long)c->n1 * c->n2;
    double *part = calloc((size_t)n3 * 2, sizeof(double));
    long long *cnt = calloc((size_t)n3 * 3, sizeof(long long));
    #pragma omp parallel for schedule(dynamic)
    for (int kz = 0; kz < n3; kz++) {
        for (long q = 0; q < np; q++) {
            long v = kz * np + q;
            vox_res r = solve_voxel(c, v);
            if (!c->gas[v]) {
                part[2 * kz] += r.dsgnorm * c->wgtc[v];
                part[2 * kz + 1] += r.ddgnorm * c->wgtc[v];
            }
            cnt[3 * kz] += r.nr; cnt[3 * kz + 1] += r.eval; cnt[3 * kz + 2] += r.outer;
        }
    }
    for (int kz = 0; kz < n3; kz++) {
        red_[0] += part[2 * kz]; red_[1] += part[2 * kz + 1];
        nr += cnt[3 * kz]; ev += cnt[3 * kz + 1]; ou += cnt[3 * kz + 2];
    }
    free(part); free(cnt);
    c->cnt_nr += nr; c->cnt_eval += ev; c->cnt_outer += ou;
    c->errs = red_[0];
    c->erre = red_[1];
    if (c->erre < 0 || c->errs < 0) { snprintf(c->err, sizeof c->err, "ERRE/ERRS negative"); return -5; }
    c->erre = sqrt(c->erre);
    c->errs = sqrt(c->errs);
    return 0;
}

/* state_6x6_evpsc (various_functions.f90:1638-1705) */
void orc_state_6x6(const int *idsim, double *dbar, const int *iscau, double *sbar,
                   const double *xmevptg, const double *devpzero, double *scauchy, double *dsim) {
    double profac[6], dbarv[6], sbarv[6], xm4[81], xmv[36], aux33[9], devv[6], aux6[6], aux6a[6], a66[36];
    for (int i = 0; i < 6; i++) profac[i] = 1.0 + ((i + 1) / 4);
    voigt_t2_to_t1(dsim, dbarv);
    voigt_t2_to_t1(scauchy, sbarv);
    B3((double *)xmevptg, xm4);
    voigt_c4_to_c2(xm4, xmv);
    B1((double *)devpzero, aux33);
    voigt_t2_to_t1(aux33, devv);
    for (int i = 0; i < 6; i++) aux6[i] = dbarv[i] - devv[i];
    for (int i = 0; i < 6; i++) {
        aux6a[i] = -1.0 * idsim[i] * aux6[i];
        for (int j = 0; j < 6; j++) {
            aux6a[i] += xmv[6 * i + j] * iscau[j] * sbarv[j] * profac[j];
            a66[6 * i + j] = iscau[j] * (i == j ? 1 : 0) - idsim[j] * xmv[6 * i + j] * profac[j];
        }
    }
    lu_eqsystem(a66, aux6a, 6);
    for (int i = 0; i < 6; i++) {
        aux6[i] = idsim[i] * aux6[i] + iscau[i] * aux6a[i];
        sbarv[i] = iscau[i] * sbarv[i] + idsim[i] * aux6a[i];
    }
    for (int i = 0; i < 6; i++) dbarv[i] = aux6[i] + devv[i];
    voigt_t1_to_t2(dbarv, dsim);
    B2(dbar, dsim);
    voigt_t1_to_t2(sbarv, scauchy);
    B2(sbar, scauchy);
}

/* calc_velgradavg_corr (various_functions.f90:1526-1635) */
static void calc_velgradavg_corr(orc_ctx *c) {
    orc_process *P = &c->proc[c->ipr];
    double red_[9] = {0};
    PLANE_REDUCE(c, 9, { if (!c->gas[v]) for (int t = 0; t < 9; t++) acc[t] += c->wgtc[v] * c->sg[9 * v + t]; });
    for (int t = 0; t < 9; t++) { c->sgavg[t] = red_[t]; c->sgavgsol[t] = red_[t] / c->wphcsol; }
    c->svmavg = vm_stress(c->sgavg);
    double aux33[9];
    for (int i = 0; i < 6; i++) {
        int ii = IJV[i][0], jj = IJV[i][1];
        aux33[3 * ii + jj] = P->iscau[i] * (P->scauchy[3 * ii + jj] - c->sgavg[3 * ii + jj]);
        aux33[3 * jj + ii] = P->iscau[i] * (P->scauchy[3 * jj + ii] - c->sgavg[3 * jj + ii]);
    }
    c->errsbc = norm2v(aux33, 9);
    double vs[9], va[9], dbar[6], sbar[6], mev[36], devp0[6];
    sym33(c->velgradavg, vs);
    antisym33(c->velgradavg, va);
    B2(dbar, vs);
    B2(sbar, c->sgavg);
    B4(mev, c->de_ds);
    for (int a = 0; a < 6; a++) {
        double s = 0;
        for (int b = 0; b < 6; b++) s += mev[6 * a + b] * sbar[b];
        devp0[a] = dbar[a] - s;
    }
    orc_state_6x6(P->idsim, dbar, P->iscau, sbar, mev, devp0, P->scauchy, P->dsim);
    double dsym[9], wasim[9] = {0}, dasym[9];
    for (int t = 0; t < 9; t++) dsym[t] = P->dsim[t] - vs[t];
    for (int ii = 0; ii < 3; ii++)
        for (int jj = 0; jj < 3; jj++) {
            if (ii == jj) { wasim[3 * ii + jj] = 0; continue; }
            int a = P->iudot[3 * ii + jj], b = P->iudot[3 * jj + ii];
            if (a == 1 && b == 0) {
                wasim[3 * ii + jj] = P->udot[3 * ii + jj] - P->dsim[3 * ii + jj];
                wasim[3 * jj + ii] = -wasim[3 * ii + jj];
            } else if (a == 0 && b == 1) {
                /* already assigned when (jj,ii) was visited or will be */
            } else if (a == 1 && b == 1) {
                wasim[3 * ii + jj] = 0.5 * (P->udot[3 * ii + jj] - P->udot[3 * jj + ii]);
            } else {
                wasim[3 * jj + ii] = 0.0;
            }
        }
    for (int t = 0; t < 9; t++) dasym[t] = wasim[t] - va[t];
    for (int t = 0; t < 9; t++) c->dvelgradavg[t] = dsym[t] + dasym[t];
}

/* one fixed-point pass (LS-EVPFFT.f90:238-358) */
/* correction_to_current (various_functions.f90:1062-1139): dL_fluct = dvelgradref F^-1, its
   weighted average, dudot = (dvelgradavg - <dL_fluct>) <F^-1>^-1, L += (dvelgradref + dudot) F^-1,
   and <L> over all / gas / solid voxels */
static void correction_to_current(orc_ctx *c) {
    double fl[9];
    {
        double red_[9] = {0};
        PLANE_REDUCE(c, 9, {
            double t9[9];
            mm33(c->dvelgradref + 9 * v, c->defgradinv + 9 * v, t9);
            for (int t = 0; t < 9; t++) acc[t] += c->wgtc[v] * t9[t];
        });
        memcpy(fl, red_, sizeof fl);
    }
    double finv[9], dud[9], dm[9];
    memcpy(finv, c->defgradinvavgc, sizeof finv);
    lu_inverse(finv, 3);
    for (int t = 0; t < 9; t++) dm[t] = c->dvelgradavg[t] - fl[t];
    mm33(dm, finv, dud);
    {
        double red_[27] = {0};
        PLANE_REDUCE(c, 27, {
            double t9[9], a9[9];
            for (int t = 0; t < 9; t++) a9[t] = c->dvelgradref[9 * v + t] + dud[t];
            mm33(a9, c->defgradinv + 9 * v, t9);
            for (int t = 0; t < 9; t++) c->velgrad[9 * v + t] += t9[t];
            for (int t = 0; t < 9; t++) acc[t] += c->velgrad[9 * v + t] * c->wgtc[v];
            if (c->gas[v]) for (int t = 0; t < 9; t++) acc[9 + t] += c->velgrad[9 * v + t] * c->wgtc[v];
            else for (int t = 0; t < 9; t++) acc[18 + t] += c->velgrad[9 * v + t] * c->wgtc[v];
        });
        for (int t = 0; t < 9; t++) {
            c->velgradavg[t] = red_[t];
            c->velgradavggas[t] = c->wphcgas > 0.0 ? red_[9 + t] / c->wphcgas : red_[9 + t];
            c->velgradavgsol[t] = red_[18 + t] / c->wphcsol;
        }
    }
}

static int fixed_point_pass(orc_ctx *c, double *mix) {
    long V = c->V;
    /* <sigma> over all voxels and PK1 (:238-269) */
    double red_[9] = {0};
    PLANE_REDUCE(c, 9, { for (int t = 0; t < 9; t++) acc[t] += c->sg[9 * v + t] * c->wgtc[v]; });
    memcpy(c->sgavg, red_, sizeof red_);
    #pragma omp parallel for schedule(static)
    for (long v = 0; v < V; v++) {
        double d[9], fit[9], pk[9];
        for (int t = 0; t < 9; t++) d[t] = c->sg[9 * v + t] - c->sgavg[t];
        tr33(c->defgradinv + 9 * v, fit);
        mm33(d, fit, pk);
        for (int t = 0; t < 9; t++) c->four[(size_t)t * V + v] = pk[t] * c->detF[v];
    }
    /* convolution_with_Goper (various_functions.f90:966-1059) */
    orc_fft3d((double *)c->four, c->n1, c->n2, c->n3, 9, -1);
    #pragma omp parallel for schedule(static)
    for (long v = 0; v < V; v++) {
        if (v == 0) { for (int t = 0; t < 9; t++) c->four[(size_t)t * V] = 0; continue; }
        cplx a[9], o[9];
        for (int t = 0; t < 9; t++) a[t] = c->four[(size_t)t * V + v];
        const double *G = c->Goper + 81 * v;
        for (int ij = 0; ij < 9; ij++) {
            cplx s = 0;
            for (int kl = 0; kl < 9; kl++) s += G[9 * ij + kl] * a[kl];
            o[ij] = s;
        }
        for (int t = 0; t < 9; t++) c->four[(size_t)t * V + v] = o[t];
    }
    orc_fft3d((double *)c->four, c->n1, c->n2, c->n3, 9, +1);
    #pragma omp parallel for schedule(static)
    for (long v = 0; v < V; v++)
        for (int t = 0; t < 9; t++) c->dvelgradref[9 * v + t] = creal(c->four[(size_t)t * V + v]) / (double)V;

    correction_to_current(c);

    /* dL_iter and mixing (LS-EVPFFT.f90:277-327) */
    double Lscale = c->lvmavg > 2.2250738585072014e-308 ? c->lvmavg : 2.2250738585072014e-308;
    double m = *mix;
    int kin = c->kinematic_controlled, dm_ = c->do_mix;
    {
        double red_[4] = {0};
        PLANE_REDUCE(c, 4, {
            double *L = c->velgrad + 9 * v, *Lo = c->velgradold + 9 * v;
            if (!kin) {
                double s = 0;
                for (int t = 0; t < 9; t++) s += (L[t] - Lo[t]) * (L[t] - Lo[t]);
                double rel = sqrt(s) / Lscale, w = c->wgtc[v];
                if (c->gas[v]) { acc[2] += rel * rel * w; acc[3] += w; }
                else { acc[0] += rel * rel * w; acc[1] += w; }
            }
            if (dm_) for (int t = 0; t < 9; t++) L[t] = m * L[t] + (1.0 - m) * Lo[t];
            else for (int t = 0; t < 9; t++) L[t] = m * L[t];
        });
        if (!kin) {
            double ds = red_[1] > 0 ? sqrt(red_[0] / red_[1]) : 0.0;
            double dg = red_[3] > 0 ? sqrt(red_[2] / red_[3]) : 0.0;
            double dl = ds > dg ? ds : dg;
            double r = 5.0e-2 / (dl > 5.0e-2 ? dl : 5.0e-2);
            if (r > 1.0) r = 1.0;
            *mix = r > 1.0e-8 ? r : 1.0e-8;
        }
    }
    /* solve_res, calc_velgradavg_corr, normalisation (:330-358) */
    int rc = solve_res(c);
    if (rc) return rc;
    calc_velgradavg_corr(c);
    c->erre = c->erre / c->lvmavg;
    c->errs = c->errs / c->svmavg;
    c->errsbc = c->errsbc / c->svmavg;
    memcpy(c->velgradold, c->velgrad, 9 * V * sizeof(double));
    return 0;
}

/* ------------------------------------------------------------------------ */
/* accepted-step updates (various_functions.f90:1748-2649)                    */
/* ------------------------------------------------------------------------ */
static void update_plastic_defgrad(orc_ctx *c, double tdot) {
    int ns = c->ns;
    #pragma omp parallel for schedule(static)
    for (long v = 0; v < c->V; v++) {
        if (c->gas[v]) continue;
        const phase_t *P = &c->ph[c->phase[v] - 1];
        double Lp[9] = {0}, e[9], d[9];
        for (int is = 0; is < P->nsyst; is++) {
            double dn[3], db[3];
            mv33(c->ag + 9 * v, P->dnca + 3 * is, dn);
            mv33(c->ag + 9 * v, P->dbca + 3 * is, db);
            double gd = c->gamdot[(size_t)v * ns + is];
            for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) Lp[3 * i + j] += db[i] * dn[j] * gd;
        }
        for (int t = 0; t < 9; t++) d[t] = Lp[t] * tdot;
        orc_expm_adaptive(d, e);
        mm33(e, c->defgradp + 9 * v, c->defgradp + 9 * v);
    }
}

/* integrate_tensor2 (:359-429) + cell2node_tensor1 (:2652-2691) + update_grid_velgrad_node (:1970-2056) */
static void update_grid_velgrad_node(orc_ctx *c, double tdot) {
    long V = c->V;
    int n1 = c->n1, n2 = c->n2, n3 = c->n3;
    int ideriv = c->proc[c->ipr].igamma;
    double red_[9] = {0};
    #pragma omp parallel for schedule(static)
    for (long v = 0; v < V; v++) {
        double t9[9];
        mm33(c->velgrad + 9 * v, c->defgrad + 9 * v, t9);
        for (int t = 0; t < 9; t++) c->four[(size_t)t * V + v] = t9[t];
    }
    PLANE_REDUCE(c, 9, { for (int t = 0; t < 9; t++) acc[t] += creal(c->four[(size_t)t * V + v]) * c->wgt; });
    double lavg[9];
    memcpy(lavg, red_, sizeof lavg);
    orc_fft3d((double *)c->four, n1, n2, n3, 9, -1);
    cplx *f3 = calloc((size_t)3 * V, sizeof(cplx));
    #pragma omp parallel for schedule(static)
    for (long v = 0; v < V; v++) {
        int i = v % n1 + 1, j = (v / n1) % n2 + 1, k = v / ((long)n1 * n2) + 1;
        double xk[3];
        orc_spatial_freq(i, j, k, n1, n2, n3, xk);
        cplx km[3];
        if (ideriv == 2) {
            cplx f = I * 0.25 * (1.0 + cexp(I * xk[0])) * (1.0 + cexp(I * xk[1])) * (1.0 + cexp(I * xk[2]));
            for (int l = 0; l < 3; l++) km[l] = tan(xk[l] / 2.0) * f;
        } else {
            for (int l = 0; l < 3; l++) km[l] = I * (ideriv == 0 ? xk[l] : sin(xk[l]));
        }
        /* Fortran dot_product(kmod, kmod) conjugates its first argument */
        cplx kk = conj(km[0]) * km[0] + conj(km[1]) * km[1] + conj(km[2]) * km[2];
        for (int l = 0; l < 3; l++) {
            cplx s = 0;
            if (creal(kk) > 0.0)
                for (int m = 0; m < 3; m++) s += conj(km[m]) * c->four[(size_t)(3 * l + m) * V + v] / kk;
            f3[(size_t)l * V + v] = s;
        }
        if (is_nyq(i, n1) && is_nyq(j, n2) && is_nyq(k, n3) && n3 > 1 && ideriv == 2)
            for (int l = 0; l < 3; l++) f3[(size_t)l * V + v] = 0;
    }
    orc_fft3d((double *)f3, n1, n2, n3, 3, +1);
    int n3node = n3 == 1 ? 1 : n3 + 1;
    #define DV(l, i, j, k) (creal(f3[(size_t)(l) * V + ((i) - 1) + (long)n1 * (((j) - 1) + (long)n2 * ((k) - 1))]) / (double)V)
    #pragma omp parallel for schedule(static)
    for (int k = 1; k <= n3node; k++)
        for (int j = 1; j <= n2 + 1; j++)
            for (int i = 1; i <= n1 + 1; i++) {
                double dvn[3];
                if (ideriv != 2) {
                    int k1 = k - 1, k2 = k, j1 = j - 1, j2 = j, i1 = i - 1, i2 = i;
                    if (k1 == 0) k1 = n3;
                    if (k2 == n3 + 1) k2 = 1;
                    if (j1 == 0) j1 = n2;
                    if (j2 == n2 + 1) j2 = 1;
                    if (i1 == 0) i1 = n1;
                    if (i2 == n2 + 1) i2 = 1; /* wraps with npts2 (various_functions.f90:2678) */
                    for (int l = 0; l < 3; l++)
                        dvn[l] = (DV(l, i2, j2, k2) + DV(l, i1, j2, k2) + DV(l, i2, j1, k2) + DV(l, i2, j2, k1) +
                                  DV(l, i1, j1, k2) + DV(l, i1, j2, k1) + DV(l, i2, j1, k1) + DV(l, i1, j1, k1)) * 0.125;
                } else {
                    int ii = i > n1 ? 1 : i, jj = j > n2 ? 1 : j, kk = k > n3 ? 1 : k;
                    for (int l = 0; l < 3; l++) dvn[l] = DV(l, ii, jj, kk);
                }
                double x[3] = {i - 0.5, j - 0.5, n3node == 1 ? 0.0 : k - 0.5}, lx[3];
                mv33(lavg, x, lx);
                double *xn = c->xnode + 3L * ((i - 1) + (long)(n1 + 1) * ((j - 1) + (long)(n2 + 1) * (k - 1)));
                for (int l = 0; l < 3; l++) xn[l] += (lx[l] + dvn[l]) * tdot;
            }
    #undef DV
    free(f3);
}

static void update_defgrad(orc_ctx *c, double tdot) {
    long V = c->V;
    double red_[39] = {0};
    PLANE_REDUCE(c, 39, {
        double d[9], e[9], old[9];
        memcpy(old, c->defgrad + 9 * v, sizeof old);
        for (int t = 0; t < 9; t++) d[t] = c->velgrad[9 * v + t] * tdot;
        orc_expm_adaptive(d, e);
        mm33(e, old, c->defgrad + 9 * v);
        c->detF[v] = det33(c->defgrad + 9 * v);
        double inv[9];
        memcpy(inv, c->defgrad + 9 * v, sizeof inv);
        lu_inverse(inv, 3);
        memcpy(c->defgradinv + 9 * v, inv, sizeof inv);
        double J = c->detF[v], w = c->wgt;
        for (int t = 0; t < 9; t++) acc[t] += c->defgrad[9 * v + t] * w;
        for (int t = 0; t < 9; t++) acc[9 + t] += inv[t] * w * J;
        if (c->gas[v]) { for (int t = 0; t < 9; t++) acc[27 + t] += inv[t] * w * J; acc[37] += w * J; }
        else { for (int t = 0; t < 9; t++) acc[18 + t] += inv[t] * w * J; acc[36] += w * J; }
        memcpy(c->defgradinc + 9 * v, e, sizeof e);
        c->wgtc[v] = w * J;
        acc[38] += J * w;
    });
    for (long v = 0; v < V; v++)
        if (c->detF[v] < 0.0) fprintf(stderr, " -> WARNING: detF=%13.6E in voxel %ld\n", c->detF[v], v);
    double detFavg = red_[38];
    memcpy(c->defgradavg, red_, 9 * sizeof(double));
    c->wphcsol = red_[36] / detFavg;
    c->wphcgas = red_[37] / detFavg;
    for (int t = 0; t < 9; t++) {
        c->defgradinvavgc[t] = red_[9 + t] / detFavg;
        c->defgradinvavgcs[t] = red_[18 + t] / detFavg / c->wphcsol;
        c->defgradinvavgcg[t] = c->wphcgas > 0.0 ? red_[27 + t] / detFavg / c->wphcgas : ID3[t];
    }
    #pragma omp parallel for schedule(static)
    for (long v = 0; v < V; v++) c->wgtc[v] /= detFavg;
}

static void update_elstic_defgrad(orc_ctx *c) {
    #pragma omp parallel for schedule(static)
    for (long v = 0; v < c->V; v++) {
        if (c->gas[v]) continue;
        double inv_old[9], fpi[9];
        memcpy(inv_old, c->defgrade + 9 * v, sizeof inv_old);
        lu_inverse(inv_old, 3);
        if (c->defgradini) mm33(c->defgradp + 9 * v, c->defgradini + 9 * v, fpi); /* Fp Fini */
        else memcpy(fpi, c->defgradp + 9 * v, sizeof fpi);                        /* Fini = I */
        lu_inverse(fpi, 3);
        mm33(c->defgrad + 9 * v, fpi, c->defgrade + 9 * v);
        mm33(c->defgrade + 9 * v, inv_old, c->defgradeinc + 9 * v);
    }
}

static void update_el_stiff(orc_ctx *c) {
    #pragma omp parallel for schedule(static)
    for (long v = 0; v < c->V; v++) {
        if (c->gas[v]) continue;
        const phase_t *P = &c->ph[c->phase[v] - 1];
        double aa[9], cc[81], c66[36];
        mm33(c->defgrade + 9 * v, c->ag + 9 * v, aa);
        double dFe = det33(c->defgrade + 9 * v);
        for (int i1 = 0; i1 < 3; i1++) for (int j1 = 0; j1 < 3; j1++) for (int k1 = 0; k1 < 3; k1++) for (int l1 = 0; l1 < 3; l1++) {
            double dum = 0;
            for (int i2 = 0; i2 < 3; i2++) for (int j2 = 0; j2 < 3; j2++) for (int k2 = 0; k2 < 3; k2++) for (int l2 = 0; l2 < 3; l2++)
                dum += aa[3 * i1 + i2] * aa[3 * j1 + j2] * P->C[27 * i2 + 9 * j2 + 3 * k2 + l2] * aa[3 * k1 + k2] * aa[3 * l1 + l2];
            cc[27 * i1 + 9 * j1 + 3 * k1 + l1] = dum / dFe;
        }
        B4(c66, cc);
        memcpy(c->cg66 + 36 * v, c66, sizeof c66);
        lu_inverse(c66, 6);
        memcpy(c->sg66 + 36 * v, c66, sizeof c66);
    }
}

static void update_tensor_config(orc_ctx *c) {
    long V = c->V;
    double red_[9] = {0};
    PLANE_REDUCE(c, 9, {
        double finc[9], feinc[9], fi[9], R[9], U[9], Rt[9], t[9];
        memcpy(finc, c->defgradinc + 9 * v, sizeof finc);
        memcpy(feinc, c->defgradeinc + 9 * v, sizeof feinc);
        double dfe = det33(feinc);
        memcpy(fi, finc, sizeof fi);
        lu_inverse(fi, 3);
        if (!c->gas[v]) {
            orc_polar(finc, R, U);
            tr33(R, Rt);
            mm33(R, c->disgradsym + 9 * v, t); mm33(t, Rt, c->disgradsym + 9 * v);
            mm33(R, c->ept + 9 * v, t); mm33(t, Rt, c->ept + 9 * v);
            double ft[9];
            tr33(feinc, ft);
            mm33(feinc, c->sgt + 9 * v, t); mm33(t, ft, t);
            for (int q = 0; q < 9; q++) c->sgt[9 * v + q] = t[q] / dfe;
        }
        mm33(c->velgrad + 9 * v, fi, c->velgrad + 9 * v);
        for (int q = 0; q < 9; q++) acc[q] += c->velgrad[9 * v + q] * c->wgtc[v];
    });
    double inv[9], dm[9], dud[9];
    memcpy(inv, c->defgradinvavgc, sizeof inv);
    lu_inverse(inv, 3);
    for (int t = 0; t < 9; t++) dm[t] = c->velgradavg[t] - red_[t];
    mm33(dm, inv, dud);
    #pragma omp parallel for schedule(static)
    for (long v = 0; v < V; v++) {
        double t9[9];
        mm33(dud, c->defgradinv + 9 * v, t9);
        for (int t = 0; t < 9; t++) c->velgrad[9 * v + t] += t9[t];
    }
}

static void update_hardening(orc_ctx *c, double tdot) {
    int ns = c->ns;
    #pragma omp parallel for schedule(static)
    for (long v = 0; v < c->V; v++) {
        if (c->gas[v]) continue;
        const phase_t *P = &c->ph[c->phase[v] - 1];
        double *tau = malloc(sizeof(double) * 2 * ns), g1;
        harden(c, P, c->gamdot + (size_t)v * ns, c->crss + (size_t)v * ns * 2, c->gacumgr[v], tdot, tau, &g1);
        memcpy(c->crss + (size_t)v * ns * 2, tau, sizeof(double) * 2 * ns);
        memcpy(c->trialtau + (size_t)v * ns * 2, tau, sizeof(double) * 2 * ns);
        c->gacumgr[v] = g1;
        free(tau);
    }
}

/* calc_c0 (various_functions.f90:2537-2649) */
static void calc_c0(orc_ctx *c) {
    const orc_process *Pr = &c->proc[c->ipr];
    int ns = c->ns;
    double red_[81 + 36] = {0};
    PLANE_REDUCE(c, 117, {
        if (!c->gas[v]) {
            const phase_t *P = &c->ph[c->phase[v] - 1];
            double sg6[6], sgt6[6], ep6[6], dep[36], ds[36], dst[36], dsg[81], dPg[81], dP66[36];
            double gtmp[64];
            double *gd = P->nsyst <= 64 ? gtmp : malloc(sizeof(double) * P->nsyst);
            B2(sg6, c->sg + 9 * v);
            B2(sgt6, c->sgt + 9 * v);
            orc_plastic_strain_rate(sg6, c->sch + (size_t)v * ns * 5, c->trialtau + (size_t)v * ns * 2,
                                    P->gamd0, P->nrs, P->nsyst, ep6, dep, gd);
            if (gd != gtmp) free(gd);
            const double *s66 = c->sg66 + 36 * v;
            for (int t = 0; t < 36; t++) {
                ds[t] = s66[t] / Pr->tdot + dep[t];
                dst[t] = s66[t] / Pr->tdot * Pr->tdot / Pr->tdotref + dep[t];
            }
            lu_inverse(ds, 6);
            lu_inverse(dst, 6);
            B3(dst, dsg); /* both tangents are built from the tdotref-scaled one (:2604-2605) */
            const double *Fi = c->defgradinv + 9 * v;
            double J = c->detF[v];
            for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) for (int k = 0; k < 3; k++) for (int l = 0; l < 3; l++) {
                double s = 0;
                for (int m = 0; m < 3; m++) for (int n = 0; n < 3; n++)
                    s += dsg[27 * i + 9 * m + 3 * k + n] * Fi[3 * j + m] * Fi[3 * l + n] * J;
                dPg[27 * i + 9 * j + 3 * k + l] = s;
            }
            B4(dP66, dPg);
            for (int t = 0; t < 81; t++) acc[t] += dsg[t] * c->wgtc[v];
            for (int t = 0; t < 36; t++) acc[81 + t] += dP66[t] * c->wgt;
        }
    });
    double dsde[81], aux66[36], dP[36];
    for (int t = 0; t < 81; t++) dsde[t] = red_[t] / c->wphcsol;
    for (int t = 0; t < 36; t++) dP[t] = red_[81 + t] / c->wphsol;
    B4(aux66, dsde);
    lu_inverse(aux66, 6);
    B3(aux66, c->de_ds);
    for (int t = 0; t < 36; t++) { c->c066[t] = dP[t] * Pr->xc0; aux66[t] = dP[t] * Pr->xc0; }
    lu_inverse(aux66, 6);
    B3(c->c066, c->c0);
    B3(aux66, c->s0);
}

/* kinematic_increments / kinematic_substep (various_functions.f90:2745-2837) */
static double stretch_log_norm(const double *U) {
    double w[3], q[9];
    jacobi33(U, w, q);
    double s = 0;
    for (int i = 0; i < 3; i++) { double l = log(w[i] > 2.2250738585072014e-308 ? w[i] : 2.2250738585072014e-308); s += l * l; }
    return sqrt(s);
}
static int kinematic_substep(orc_ctx *c, double tdot, double *mvol, double *mrot, double *mstr, double *mdet) {
    double red_[4] = {0, 0, 0, 0};
    int n3 = c->n3;
    long np = (long)c->n1 * c->n2;
    double *part = malloc(sizeof(double) * 4 * n3);
    #pragma omp parallel for schedule(static)
    for (int kz = 0; kz < n3; kz++) {
        double a0 = 0, a1 = 0, a2 = 0, a3 = 1.79769313486231570e308;
        for (long q = 0; q < np; q++) {
            long v = kz * np + q;
            double d[9], Fi[9], R[9], U[9];
            for (int t = 0; t < 9; t++) d[t] = c->velgrad[9 * v + t] * tdot;
            orc_expm_adaptive(d, Fi);
            double Jinc = det33(Fi), Jn = c->detF[v] * Jinc;
            double vol = fabs(log(Jinc > 2.2250738585072014e-308 ? Jinc : 2.2250738585072014e-308));
            orc_polar(Fi, R, U);
            double arg = 0.5 * (R[0] + R[4] + R[8] - 1.0);
            arg = arg > 1.0 ? 1.0 : (arg < -1.0 ? -1.0 : arg);
            double rot = acos(arg) * 180.0 / PI;
            double str = stretch_log_norm(U);
            if (vol > a0) a0 = vol;
            if (rot > a1) a1 = rot;
            if (str > a2) a2 = str;
            if (Jn < a3) a3 = Jn;
        }
        part[4 * kz] = a0; part[4 * kz + 1] = a1; part[4 * kz + 2] = a2; part[4 * kz + 3] = a3;
    }
    red_[3] = 1.79769313486231570e308;
    for (int kz = 0; kz < n3; kz++) {
        if (part[4 * kz] > red_[0]) red_[0] = part[4 * kz];
        if (part[4 * kz + 1] > red_[1]) red_[1] = part[4 * kz + 1];
        if (part[4 * kz + 2] > red_[2]) red_[2] = part[4 * kz + 2];
        if (part[4 * kz + 3] < red_[3]) red_[3] = part[4 * kz + 3];
    }
    free(part);
    *mvol = red_[0]; *mrot = red_[1]; *mstr = red_[2]; *mdet = red_[3];
    double worst = 1.0;
    if (red_[0] / 5.0e-2 > worst) worst = red_[0] / 5.0e-2;
    if (red_[1] / 2.0 > worst) worst = red_[1] / 2.0;
    if (red_[2] / 5.0e-2 > worst) worst = red_[2] / 5.0e-2;
    int nsub = (int)ceil(worst);
    return nsub > 1 ? nsub : 1;
}

/* ------------------------------------------------------------------------ */
/* update_diagnostics (types.f90:345-758) with a Hager-Higham 1-norm          */
/* estimator restated from Higham (1988), ACM TOMS Algorithm 674             */
/* ------------------------------------------------------------------------ */
double orc_cond1_est(const double *a66) {
    double lu[36];
    int piv[6], n = 6;
    memcpy(lu, a66, sizeof lu);
    double anorm = 0;
    for (int j = 0; j < n; j++) {
        double s = 0;
        for (int i = 0; i < n; i++) s += fabs(a66[6 * i + j]);
        if (s > anorm) anorm = s;
    }
    if (lu_fact(lu, n, piv)) return 1.79769313486231570e308;
    if (anorm == 0.0) return 1.79769313486231570e308;
    /* estimate ||A^{-1}||_1 */
    double x[6], xs[6], est, estold;
    int j, jlast, iter;
    for (int i = 0; i < n; i++) x[i] = 1.0 / n;
    lu_solve(lu, n, piv, x);
    est = 0; for (int i = 0; i < n; i++) est += fabs(x[i]);
    for (int i = 0; i < n; i++) { xs[i] = x[i] >= 0 ? 1.0 : -1.0; x[i] = xs[i]; }
    lu_solve_t(lu, n, piv, x);
    j = 0; for (int i = 1; i < n; i++) if (fabs(x[i]) > fabs(x[j])) j = i;
    iter = 2;
    for (;;) {
        for (int i = 0; i < n; i++) x[i] = 0.0;
        x[j] = 1.0;
        lu_solve(lu, n, piv, x);
        estold = est;
        est = 0; for (int i = 0; i < n; i++) est += fabs(x[i]);
        int same = 1;
        for (int i = 0; i < n; i++) { double s = x[i] >= 0 ? 1.0 : -1.0; if (s != xs[i]) { same = 0; break; } }
        if (same || est <= estold) break;
        for (int i = 0; i < n; i++) { xs[i] = x[i] >= 0 ? 1.0 : -1.0; x[i] = xs[i]; }
        lu_solve_t(lu, n, piv, x);
        jlast = j;
        j = 0; for (int i = 1; i < n; i++) if (fabs(x[i]) > fabs(x[j])) j = i;
        if (x[jlast] != fabs(x[j]) && iter < 5) { iter++; continue; }
        break;
    }
    /* alternating-sign test vector */
    double altsgn = 1.0;
    for (int i = 0; i < n; i++) { x[i] = altsgn * (1.0 + (double)i / (double)(n - 1)); altsgn = -altsgn; }
    lu_solve(lu, n, piv, x);
    double temp = 0; for (int i = 0; i < n; i++) temp += fabs(x[i]);
    temp = 2.0 * (temp / (3.0 * n));
    if (temp > est) est = temp;
    double rcond = (1.0 / est) / anorm;
    if (!(rcond > 0.0)) return 1.79769313486231570e308;
    return 1.0 / rcond;
}

static double safe_rel(double x, double xp) {
    return (xp <= 2.2250738585072014e-308 || xp >= 1.79769313486231570e308 * 0.5) ? 0.0 : fabs(x - xp) / xp;
}

static void update_diagnostics(orc_ctx *c, orc_row *r) {
    const double HUGEV = 1.79769313486231570e308, TINYV = 2.2250738585072014e-308;
    int n1 = c->n1, n2 = c->n2, n3 = c->n3;
    long np = (long)n1 * n2;
    double Lref = c->lvmavg > TINYV ? c->lvmavg : TINYV;
    double Lfloor = 1.0e-12 * Lref > TINYV ? 1.0e-12 * Lref : TINYV;
    /* per-plane partials: 0 detmin,1 detmin_sol,2 detmin_gas,3 s2all,4 wall,5 maxall,6 s2sol,7 wsol,
       8 maxsol,9 s2gas,10 wgas,11 maxgas,12 volsol,13 volgas,14 strsol,15 strgas,16 rotsol,17 rotgas,
       18 csol,19 cgasv,20 ccgas ; imin in ints */
    double *P = malloc(sizeof(double) * 21 * n3);
    long *IM = malloc(sizeof(long) * 2 * n3);
    #pragma omp parallel for schedule(static)
    for (int kz = 0; kz < n3; kz++) {
        double *a = P + 21 * kz;
        for (int t = 0; t < 21; t++) a[t] = 0;
        a[0] = a[1] = a[2] = HUGEV;
        long ims = -1, img = -1;
        for (long q = 0; q < np; q++) {
            long v = kz * np + q;
            double d2 = 0, l0 = 0;
            for (int t = 0; t < 9; t++) {
                double dl = c->velgrad[9 * v + t] - c->s_velgrad[9 * v + t];
                d2 += dl * dl;
                l0 += c->s_velgrad[9 * v + t] * c->s_velgrad[9 * v + t];
            }
            double den = sqrt(l0) > Lfloor ? sqrt(l0) : Lfloor;
            double rel = sqrt(d2) / den, w = c->wgtc[v], J = c->detF[v];
            a[3] += rel * rel * w; a[4] += w; if (rel > a[5]) a[5] = rel;
            if (J < a[0]) a[0] = J;
            if (c->gas[v]) {
                if (J <= a[2]) { a[2] = J; img = v; }
                a[9] += rel * rel * w; a[10] += w; if (rel > a[11]) a[11] = rel;
            } else {
                if (J <= a[1]) { a[1] = J; ims = v; }
                a[6] += rel * rel * w; a[7] += w; if (rel > a[8]) a[8] = rel;
            }
            const double *Fi = c->defgradinc + 9 * v;
            double Jinc = det33(Fi);
            double vol = fabs(log(Jinc > TINYV ? Jinc : TINYV));
            double R[9], U[9];
            orc_polar(Fi, R, U);
            double arg = 0.5 * ((R[0] + R[4] + R[8]) - 1.0);
            arg = arg > 1.0 ? 1.0 : (arg < -1.0 ? -1.0 : arg);
            double rot = acos(arg) * 180.0 / PI;
            double str = stretch_log_norm(U);
            int g = c->gas[v];
            if (vol > a[12 + g]) a[12 + g] = vol;
            if (rot > a[16 + g]) a[16 + g] = rot;
            if (str > a[14 + g]) a[14 + g] = str;
            double cA = orc_cond1_est(c->c066mod + 36 * v);
            if (g) {
                if (cA > a[19]) a[19] = cA;
                double cg = orc_cond1_est(c->cgas + 36 * v);
                if (cg > a[20]) a[20] = cg;
            } else if (cA > a[18]) a[18] = cA;
        }
        IM[2 * kz] = ims; IM[2 * kz + 1] = img;
    }
    double R_[21] = {0};
    R_[0] = R_[1] = R_[2] = HUGEV;
    long ims = -1, img = -1;
    for (int kz = 0; kz < n3; kz++) {
        double *a = P + 21 * kz;
        if (a[0] < R_[0]) R_[0] = a[0];
        /* combine with strict '<' between threads' partials, as the CRITICAL block does */
        if (a[1] < R_[1]) { R_[1] = a[1]; ims = IM[2 * kz]; }
        if (a[2] < R_[2]) { R_[2] = a[2]; img = IM[2 * kz + 1]; }
        for (int t = 3; t < 21; t++) {
            if (t == 5 || t == 8 || t == 11 || t >= 12) { if (a[t] > R_[t]) R_[t] = a[t]; }
            else R_[t] += a[t];
        }
    }
    free(P); free(IM);
    for (int t = 0; t < 3; t++) if (R_[t] >= HUGEV * 0.5) R_[t] = 0.0;
    r->detF_min_sol = R_[1]; r->detF_min_gas = R_[2];
    r->volinc_max_sol = R_[12]; r->volinc_max_gas = R_[13];
    r->stretchinc_max_sol = R_[14]; r->stretchinc_max_gas = R_[15];
    r->rotang_max_sol = R_[16]; r->rotang_max_gas = R_[17];
    r->cond_c066mod_max_sol = R_[18]; r->cond_c066mod_max_gas = R_[19]; r->cond_cgas_max = R_[20];
    r->dL_step_rms = R_[4] > 0 ? sqrt(R_[3] / R_[4]) : 0.0; r->dL_step_max = R_[5];
    r->dL_step_rms_sol = R_[7] > 0 ? sqrt(R_[6] / R_[7]) : 0.0; r->dL_step_max_sol = R_[8];
    r->dL_step_rms_gas = R_[10] > 0 ? sqrt(R_[9] / R_[10]) : 0.0; r->dL_step_max_gas = R_[11];
    for (int t = 0; t < 3; t++) { r->imin_sol[t] = 0; r->imin_gas[t] = 0; }
    if (ims >= 0) { r->imin_sol[0] = ims % n1 + 1; r->imin_sol[1] = (ims / n1) % n2 + 1; r->imin_sol[2] = ims / np + 1; }
    if (img >= 0) { r->imin_gas[0] = img % n1 + 1; r->imin_gas[1] = (img / n1) % n2 + 1; r->imin_gas[2] = img / np + 1; }
    r->dvelgradavg_mag = norm2v(c->dvelgradavg, 9);
    r->detF_min_sol_rel = safe_rel(r->detF_min_sol, c->dprev_detF_min_sol);
    r->detF_min_gas_rel = safe_rel(r->detF_min_gas, c->dprev_detF_min_gas);
    c->dprev_detF_min = R_[0];
    c->dprev_detF_min_sol = r->detF_min_sol;
    c->dprev_detF_min_gas = r->detF_min_gas;
    r->volinc_max_sol_rel = safe_rel(r->volinc_max_sol, c->dprev_volinc_sol);
    c->dprev_volinc_sol = r->volinc_max_sol;
    r->volinc_max_gas_rel = safe_rel(r->volinc_max_gas, c->dprev_volinc_gas);
    c->dprev_volinc_gas = r->volinc_max_gas;
    r->stretchinc_max_sol_rel = safe_rel(r->stretchinc_max_sol, c->dprev_stretch_sol);
    c->dprev_stretch_sol = r->stretchinc_max_sol;
    r->stretchinc_max_gas_rel = safe_rel(r->stretchinc_max_gas, c->dprev_stretch_gas);
    c->dprev_stretch_gas = r->stretchinc_max_gas;
}

/* ------------------------------------------------------------------------ */
/* driver (LS-EVPFFT.f90:180-590)                                            */
/* ------------------------------------------------------------------------ */
static void vgt6(const double *m, double *o) { /* 11,22,33,23,13,12 */
    o[0] = m[0]; o[1] = m[4]; o[2] = m[8]; o[3] = m[5]; o[4] = m[2]; o[5] = m[1];
}

int orc_step(orc_ctx *c, orc_row *row) {
    if (!c->initialised) return -1;
    memset(row, 0, sizeof *row);
    /* advance to the next process when the current one is done */
    while (!c->proc_active || c->imicro >= c->proc[c->ipr].nsteps) {
        if (c->ipr + 1 >= c->nproc) return 0;
        c->ipr++;
        orc_process *P = &c->proc[c->ipr];
        c->nconverged = 0;
        c->kinematic_controlled = (P->iscau[2] == 0);
        c->tdot_plan = P->tdot;
        c->proc_time_target = (double)P->nsteps * c->tdot_plan;
        c->proc_time_done = 0.0;
        c->imicro = 0;
        c->proc_active = 1;
    }
    orc_process *P = &c->proc[c->ipr];
    c->imicro++;
    if (P->ictrl == -1) {
        double rem = c->proc_time_target - c->proc_time_done;
        if (rem < 0) rem = 0;
        P->tdot = rem > 0.0 ? (c->tdot_plan < rem ? c->tdot_plan : rem) : c->tdot_plan;
    }
    take_snapshot(c);
    int did_converge = 0, accept = 0;
    if ((c->imicro == 1 && c->ipr == 0) || P->iupdate == 1) update_schmid(c);
    form_G(c);
    calc_c066mod(c);
    c->iter = 0;
    c->cnt_nr = c->cnt_eval = c->cnt_outer = 0;
    double mix = 1.0, mvol, mrot, mstr, mdet;
    long long passes = 0, rollbacks = 0;
    const long long cap = 100LL * (P->itmax > 0 ? P->itmax : 1) + 1000;
    while (!(did_converge || accept)) {
        if (c->iter == 0) {
            calc_gas_compl(c);
            memset(c->dvelgradavg, 0, sizeof c->dvelgradavg);
            c->edotvmmx = 0;
            mix = 1.0;
            c->erre = c->errs = c->errsbc = 2.0 * P->error;
        }
        c->iter++;
        passes++;
        if (passes > cap) { snprintf(c->err, sizeof c->err, "fixed-point loop did not terminate"); return -6; }
        int rc = fixed_point_pass(c, &mix);
        if (rc) return rc;
        if (g_singular) { snprintf(c->err, sizeof c->err, "Singular matrix in lu_inverse"); return -7; }
        if (P->ictrl == 0) {
            double vs[9];
            sym33(c->velgradavg, vs);
            double evm = vm_strain(vs);
            P->tdot = P->devmapp / evm;
            P->tdotref = P->tdot;
        }
        if (c->errs <= P->error && c->erre <= P->error && c->errsbc <= P->error) did_converge = 1;
        if (P->ictrl == -1) {
            if (P->tdot > P->tdotmin) {
                if (!did_converge && c->iter == P->itmax) {
                    double tdot = P->tdot;
                    int nsub = kinematic_substep(c, tdot, &mvol, &mrot, &mstr, &mdet);
                    tdot = tdot / nsub;
                    double h = 0.5 * P->tdot;
                    tdot = tdot < h ? tdot : h;
                    tdot = P->tdotmin > tdot ? P->tdotmin : tdot;
                    nsub = (int)ceil(P->tdot / tdot);
                    if (nsub > 1) {
                        double rem = c->proc_time_target - c->proc_time_done;
                        if (rem < 0) rem = 0;
                        c->tdot_plan = tdot;
                        int nnew = (c->imicro - 1) + (int)ceil(rem / c->tdot_plan);
                        c->total_macrosteps += nnew - P->nsteps;
                        P->nsteps = nnew;
                        P->tdot = c->tdot_plan < rem ? c->tdot_plan : rem;
                        c->nconverged = 0;
                        c->iter = 0;
                        rollback(c);
                        rollbacks++;
                    }
                } else if (did_converge) {
                    double tdot = P->tdot;
                    int nsub = kinematic_substep(c, tdot, &mvol, &mrot, &mstr, &mdet);
                    tdot = tdot / nsub;
                    if (mdet <= 1.0e-3) {
                        double h = 0.5 * P->tdot;
                        tdot = tdot < h ? tdot : h;
                    }
                    tdot = P->tdotmin > tdot ? P->tdotmin : tdot;
                    nsub = (int)ceil(P->tdot / tdot);
                    if (nsub > 1) {
                        double rem = c->proc_time_target - c->proc_time_done;
                        if (rem < 0) rem = 0;
                        c->tdot_plan = tdot;
                        int nnew = (c->imicro - 1) + (int)ceil(rem / c->tdot_plan);
                        c->total_macrosteps += nnew - P->nsteps;
                        P->nsteps = nnew;
                        P->tdot = c->tdot_plan < rem ? c->tdot_plan : rem;
                        did_converge = 0;
                        c->nconverged = 0;
                        c->iter = 0;
                        mix = 1.0;
                        rollback(c);
                        rollbacks++;
                    }
                }
            } else if (c->iter == P->itmax) {
                accept = 1;
            }
        }
    }
    double tdot = P->tdot;
    long V = c->V;
    /* accepted-step accumulation (:431-446) */
    double emx = c->edotvmmx;
    for (long v = 0; v < V; v++) {
        double vs[9];
        for (int t = 0; t < 9; t++) c->ept[9 * v + t] += c->edotp[9 * v + t] * tdot;
        memcpy(c->sgt + 9 * v, c->sg + 9 * v, 9 * sizeof(double));
        sym33(c->velgrad + 9 * v, vs);
        for (int t = 0; t < 9; t++) c->disgradsym[9 * v + t] += vs[t] * tdot;
        c->epvm[v] += vm_strain(c->edotp + 9 * v) * tdot;
        double e = vm_strain(vs);
        if (e > emx && !c->gas[v]) emx = e;
    }
    c->edotvmmx = emx;
    if (P->iupdate == 1) {
        update_plastic_defgrad(c, tdot);
        update_grid_velgrad_node(c, tdot);
        update_defgrad(c, tdot);
        update_elstic_defgrad(c);
        update_el_stiff(c);
        update_tensor_config(c);
    }
    if (P->iuphard == 1) update_hardening(c, tdot);
    /* macro averages (:460-523) */
    double vs[9];
    sym33(c->velgradavg, vs);
    c->lvmavg = vm_strain(c->velgradavg);
    c->dvmavg = vm_strain(vs);
    c->evmavg += c->dvmavg * tdot;
    double red_[45 + 5] = {0};
    int ns = c->ns;
    PLANE_REDUCE(c, 50, {
        if (!c->gas[v]) {
            double w = c->wgtc[v];
            for (int t = 0; t < 9; t++) {
                acc[t] += c->ept[9 * v + t] * w;
                acc[9 + t] += c->disgradsym[9 * v + t] * w;
                acc[18 + t] += (c->disgradsym[9 * v + t] - c->ept[9 * v + t]) * w;
                acc[27 + t] += c->edotp[9 * v + t] * w;
                acc[36 + t] += c->sgt[9 * v + t] * w;
            }
            acc[45] += c->gacumgr[v];
            acc[46] += 1.0;
            const phase_t *Ph = &c->ph[c->phase[v] - 1];
            for (int is = 0; is < Ph->nsyst; is++) {
                double gs = 0.5 * (c->crss[((size_t)v * ns + is) * 2] + c->crss[((size_t)v * ns + is) * 2 + 1]);
                double ag = fabs(c->gamdot[(size_t)v * ns + is]);
                acc[47] += gs;
                acc[48] += ag;
                acc[49] += ag * gs;
            }
        }
    });
    /* eelavg reproduces eelavg = eavg + (disgradsym - ept)*wgtc of the last solid
       voxel visited (LS-EVPFFT.f90:482) */
    long vlast = -1;
    for (long v = V - 1; v >= 0; v--) if (!c->gas[v]) { vlast = v; break; }
    double nsys_solid = 0;
    for (long v = 0; v < V; v++) if (!c->gas[v]) nsys_solid += c->ph[c->phase[v] - 1].nsyst;
    for (int t = 0; t < 9; t++) {
        c->epavg[t] = red_[t] / c->wphcsol;
        c->eavg[t] = red_[9 + t] / c->wphcsol;
        c->eelavg[t] = vlast >= 0 ? (red_[9 + t] + (c->disgradsym[9 * vlast + t] - c->ept[9 * vlast + t]) * c->wgtc[vlast]) / c->wphcsol : 0.0;
        c->edotpavg[t] = red_[27 + t] / c->wphcsol;
        c->sgsolavg[t] = red_[36 + t] / c->wphcsol;
    }
    double gammaavg = red_[46] > 0 ? red_[45] / red_[46] : 0.0;
    double gavg = nsys_solid > 0 ? red_[47] / nsys_solid : 0.0;
    double geffavg = red_[48] > 0 ? red_[49] / red_[48] : gavg;
    c->evmpavg = vm_strain(c->epavg);
    c->dvmpavg = vm_strain(c->edotpavg);

    update_diagnostics(c, row);
    row->erre = c->erre; row->errs = c->errs; row->errsbc = c->errsbc;
    row->diag_tdotref = P->tdotref;
    row->s33_target = P->scauchy[8];
    row->s33 = c->sgavg[8];
    row->s33_sol = c->sgsolavg[8];
    double tdot_used = P->tdot;
    if (P->ictrl == -1) {
        c->proc_time_done += tdot_used;
        P->tdot = c->tdot_plan;
    }
    row->devm_inc_max = c->edotvmmx * tdot_used;
    c->sim_time += tdot_used;
    row->iter = c->iter;
    row->imacroloop = c->imacroloop;
    row->ipr = c->ipr + 1;
    row->imicro = c->imicro;
    row->sim_time = c->sim_time;
    row->tdot = tdot_used;
    row->tdot_ref = P->tdotref;
    vgt6(c->eavg, row->eavg); vgt6(c->sgsolavg, row->sgsolavg); vgt6(c->epavg, row->epavg);
    vgt6(c->eelavg, row->eelavg); vgt6(c->velgradavg, row->velgradavg); vgt6(c->edotpavg, row->edotpavg);
    row->evm = c->evmavg; row->evmp = c->evmpavg; row->dvm = c->dvmavg; row->dvmp = c->dvmpavg;
    row->svm = c->svmavg; row->gamma = gammaavg; row->g = gavg; row->geff = geffavg;
    row->passes = passes; row->nr_steps = c->cnt_nr; row->evals = c->cnt_eval; row->outer = c->cnt_outer;
    row->rollbacks = rollbacks;
    /* ramp tdot back up (:557-581) */
    if (P->ictrl == -1) {
        if (c->tdot_plan < P->tdotref && c->nconverged >= 5) {
            c->nconverged = 0;
            c->tdot_plan = 2.0 * c->tdot_plan < P->tdotref ? 2.0 * c->tdot_plan : P->tdotref;
            double rem = c->proc_time_target - c->proc_time_done;
            if (rem < 0) rem = 0;
            int left = (int)ceil(rem / c->tdot_plan);
            int nnew = c->imicro + left;
            c->total_macrosteps += nnew - P->nsteps;
            P->nsteps = nnew;
            P->tdot = rem > 0.0 ? (c->tdot_plan < rem ? c->tdot_plan : rem) : c->tdot_plan;
        }
    } else if (P->tdot < P->tdotref && c->nconverged >= 5) {
        c->nconverged = 0;
        double t0 = P->tdot;
        P->tdot = 2.0 * P->tdot < P->tdotref ? 2.0 * P->tdot : P->tdotref;
        int left = P->nsteps - c->imicro;
        left = (int)ceil((double)left * t0 / P->tdot);
        c->total_macrosteps += left - (P->nsteps - c->imicro);
        P->nsteps = c->imicro + left;
    }
    if (did_converge) c->nconverged++;
    calc_c0(c);
    c->imacroloop++;
    if (g_singular) { snprintf(c->err, sizeof c->err, "Singular matrix in lu_inverse"); return -7; }
    return 1;
}

/* ------------------------------------------------------------------------ */
/* field access                                                              */
/* ------------------------------------------------------------------------ */
static double *field_ptr(orc_ctx *c, const char *name, int *ncomp) {
    struct { const char *n; double *p; int k; } T[] = {
        {"sg", c->sg, 9}, {"sgt", c->sgt, 9}, {"velgrad", c->velgrad, 9}, {"velgradold", c->velgradold, 9},
        {"dvelgradref", c->dvelgradref, 9}, {"defgrad", c->defgrad, 9}, {"defgradinv", c->defgradinv, 9},
        {"detF", c->detF, 1}, {"wgtc", c->wgtc, 1}, {"defgrade", c->defgrade, 9}, {"defgradp", c->defgradp, 9},
        {"ag", c->ag, 9}, {"cg66", c->cg66, 36}, {"sg66", c->sg66, 36}, {"c066mod", c->c066mod, 36},
        {"cgm", c->cgm, 36}, {"cgas", c->cgas, 36}, {"sch", c->sch, 5 * c->ns}, {"crss", c->crss, 2 * c->ns},
        {"trialtau", c->trialtau, 2 * c->ns}, {"gamdot", c->gamdot, c->ns}, {"gacumgr", c->gacumgr, 1},
        {"edotp", c->edotp, 9}, {"ept", c->ept, 9}, {"disgradsym", c->disgradsym, 9}, {"epvm", c->epvm, 1},
        {"eth", c->eth, 9},
        {"defgradinc", c->defgradinc, 9}, {"defgradeinc", c->defgradeinc, 9}, {"defgradini", c->defgradini, 9},
    };
    for (size_t i = 0; i < sizeof T / sizeof T[0]; i++)
        if (!strcmp(T[i].n, name)) { *ncomp = T[i].k; return T[i].p; }
    return NULL;
}
int orc_field_ncomp(orc_ctx *c, const char *name) {
    int k = 0;
    if (!strcmp(name, "xnode")) return 3;
    if (!strcmp(name, "macro")) return 1;
    return field_ptr(c, name, &k) ? k : -1;
}
int orc_get_field(orc_ctx *c, const char *name, double *out) {
    int k;
    if (!strcmp(name, "xnode")) {
        int n3node = c->n3 == 1 ? 1 : c->n3 + 1;
        memcpy(out, c->xnode, sizeof(double) * 3 * (c->n1 + 1) * (c->n2 + 1) * n3node);
        return 0;
    }
    if (!strcmp(name, "macro")) { /* c0, s0, de_ds, c066, Goperr0, sgavg, velgradavg, dvelgradavg, defgradinvavgc, scalars */
        double *o = out;
        memcpy(o, c->c0, 81 * 8); o += 81;
        memcpy(o, c->s0, 81 * 8); o += 81;
        memcpy(o, c->de_ds, 81 * 8); o += 81;
        memcpy(o, c->c066, 36 * 8); o += 36;
        memcpy(o, c->Goperr0, 81 * 8); o += 81;
        memcpy(o, c->sgavg, 72); o += 9;
        memcpy(o, c->velgradavg, 72); o += 9;
        memcpy(o, c->dvelgradavg, 72); o += 9;
        memcpy(o, c->defgradinvavgc, 72); o += 9;
        memcpy(o, c->defgradavg, 72); o += 9;
        o[0] = c->wphcsol; o[1] = c->wphcgas; o[2] = c->lvmavg; o[3] = c->svmavg; o[4] = c->erre;
        o[5] = c->errs; o[6] = c->errsbc; o[7] = c->wphsol;
        return 0;
    }
    double *p = field_ptr(c, name, &k);
    if (!p) return -1;
    memcpy(out, p, sizeof(double) * k * c->V);
    return 0;
}
int orc_set_field(orc_ctx *c, const char *name, const double *in) {
    int k;
    double *p = field_ptr(c, name, &k);
    if (!p) return -1;
    memcpy(p, in, sizeof(double) * k * c->V);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* test hooks                                                                */
/* ------------------------------------------------------------------------ */
/* form_G for process ipr0 (0-based) with the context's current c0/s0 */
int orc_form_G(orc_ctx *c, int ipr0) {
    int save = c->ipr;
    c->ipr = ipr0;
    form_G(c);
    c->ipr = save;
    return 0;
}
/* convolution_with_Goper on an arbitrary field: in/out (V,9) row-major tensors */
int orc_convolve(orc_ctx *c, const double *in, double *out) {
    long V = c->V;
    for (long v = 0; v < V; v++)
        for (int t = 0; t < 9; t++) c->four[(size_t)t * V + v] = in[9 * v + t];
    orc_fft3d((double *)c->four, c->n1, c->n2, c->n3, 9, -1);
    for (long v = 0; v < V; v++) {
        if (v == 0) { for (int t = 0; t < 9; t++) c->four[(size_t)t * V] = 0; continue; }
        cplx a[9], o[9];
        for (int t = 0; t < 9; t++) a[t] = c->four[(size_t)t * V + v];
        const double *G = c->Goper + 81 * v;
        for (int ij = 0; ij < 9; ij++) {
            cplx s = 0;
            for (int kl = 0; kl < 9; kl++) s += G[9 * ij + kl] * a[kl];
            o[ij] = s;
        }
        for (int t = 0; t < 9; t++) c->four[(size_t)t * V + v] = o[t];
    }
    orc_fft3d((double *)c->four, c->n1, c->n2, c->n3, 9, +1);
    for (long v = 0; v < V; v++)
        for (int t = 0; t < 9; t++) out[9 * v + t] = creal(c->four[(size_t)t * V + v]) / (double)V;
    return 0;
}
/* harden for voxel v of phase iph (1-based): tau_out (2*ns) and gamma1 */
int orc_harden(orc_ctx *c, long v, double *tau_out, double *gamma1) {
    const phase_t *P = &c->ph[c->phase[v] - 1];
    harden(c, P, c->gamdot + (size_t)v * c->ns, c->crss + (size_t)v * c->ns * 2, c->gacumgr[v],
           c->proc[c->ipr < 0 ? 0 : c->ipr].tdot, tau_out, gamma1);
    return 0;
}
/* set the macroscopic reference stiffness (c066 in the b-basis) and derived c0/s0 */
int orc_set_c066(orc_ctx *c, const double *c066) {
    memcpy(c->c066, c066, sizeof c->c066);
    memcpy(c->s066, c066, sizeof c->s066);
    lu_inverse(c->s066, 6);
    B3(c->c066, c->c0);
    B3(c->s066, c->s0);
    return 0;
}
/* set the iteration-time process index (for hooks that read process data) */
int orc_set_ipr(orc_ctx *c, int ipr0) { c->ipr = ipr0; return 0; }
/* calc_c0 with the current process and fields */
int orc_calc_c0(orc_ctx *c) { calc_c0(c); return 0; }
/* kinematic_increments at step tdot: out4 = max volinc, max rotang (deg), max stretchinc,
 * min detF_next; returns nsub */
int orc_kinematic(orc_ctx *c, double tdot, double *out4) {
    return kinematic_substep(c, tdot, out4, out4 + 1, out4 + 2, out4 + 3);
}

/* Run one named stage of the increment on the current fields (closed-form tests of each
 * restated procedure): "correction_to_current", "calc_c066mod", "calc_gas_compl",
 * "update_plastic_defgrad", "update_grid_velgrad_node", "update_defgrad",
 * "update_elstic_defgrad", "update_el_stiff", "update_tensor_config", "update_schmid".
 * Stages that read process data use process ipr (orc_set_ipr) and tdot. */
int orc_stage(orc_ctx *c, const char *name, double tdot) {
    if (c->ipr < 0) c->ipr = 0;
    orc_process *P = &c->proc[c->ipr];
    double save = P->tdot;
    P->tdot = tdot;
    int rc = 0;
    if (!strcmp(name, "correction_to_current")) correction_to_current(c);
    else if (!strcmp(name, "calc_c066mod")) calc_c066mod(c);
    else if (!strcmp(name, "calc_gas_compl")) calc_gas_compl(c);
    else if (!strcmp(name, "update_plastic_defgrad")) update_plastic_defgrad(c, tdot);
    else if (!strcmp(name, "update_grid_velgrad_node")) update_grid_velgrad_node(c, tdot);
    else if (!strcmp(name, "update_defgrad")) update_defgrad(c, tdot);
    else if (!strcmp(name, "update_elstic_defgrad")) update_elstic_defgrad(c);
    else if (!strcmp(name, "update_el_stiff")) update_el_stiff(c);
    else if (!strcmp(name, "update_tensor_config")) update_tensor_config(c);
    else if (!strcmp(name, "update_schmid")) update_schmid(c);
    else rc = -1;
    P->tdot = save;
    return rc;
}
/* Set one macroscopic quantity of micro_type: "velgradavg", "dvelgradavg", "defgradinvavgc"
 * (9 each, row-major), "Goperr0" (81) or "wph" (wphcsol, wphcgas). */
int orc_set_macro(orc_ctx *c, const char *name, const double *v) {
    if (!strcmp(name, "velgradavg")) memcpy(c->velgradavg, v, 72);
    else if (!strcmp(name, "dvelgradavg")) memcpy(c->dvelgradavg, v, 72);
    else if (!strcmp(name, "defgradinvavgc")) memcpy(c->defgradinvavgc, v, 72);
    else if (!strcmp(name, "Goperr0")) memcpy(c->Goperr0, v, 81 * 8);
    else if (!strcmp(name, "wph")) { c->wphcsol = v[0]; c->wphcgas = v[1]; }
    else return -1;
    return 0;
}
/* Macroscopic averages of correction_to_current: velgradavg, velgradavgsol, velgradavggas (27). */
int orc_get_lavg(orc_ctx *c, double *out27) {
    memcpy(out27, c->velgradavg, 72);
    memcpy(out27 + 9, c->velgradavgsol, 72);
    memcpy(out27 + 18, c->velgradavggas, 72);
    return 0;
}
