// fft_reg.h -- register-resident FFT building blocks for the spectral passes.
//
// An N-point transform is spread over T threads of E = N/T points each:
//   level 1: each thread takes x[t + T m] (m < E) and does an E-point DFT in registers,
//            then multiplies by w_N^{t q};
//   exchange through LDS once;
//   level 2: T-point DFTs over t, one per q (q < E), each done by one thread.
// The E- and T-point DFTs are compile-time Cooley-Tukey recursions on radices 4, 3, 2
// whose constant twiddles are exact literals (multiples of pi/48), so they fold into
// the instruction stream; the runtime twiddles w_N^{tq} come from a table computed on
// the host in long double.
#pragma once
#include <hip/hip_runtime.h>

namespace cg {

template <typename R> struct Cx;
template <> struct Cx<double> { using T = double2; };
template <> struct Cx<float> { using T = float2; };

// cos(k pi/48) for the k reached by the DFT sizes 2, 3, 4, 6, 8, 12, 16, 24
constexpr double cos_pi48(int k) {
    return k == 0    ? 1.0
           : k == 4  ? 0.96592582628906828675   // cos(pi/12)
           : k == 6  ? 0.92387953251128675613   // cos(pi/8)
           : k == 8  ? 0.86602540378443864676   // cos(pi/6)
           : k == 12 ? 0.70710678118654752440   // cos(pi/4)
           : k == 16 ? 0.5                      // cos(pi/3)
           : k == 18 ? 0.38268343236508977173   // cos(3pi/8)
           : k == 20 ? 0.25881904510252076235   // cos(5pi/12)
           : k == 24 ? 0.0
                     : 2.0;  // not reachable for the supported sizes
}
struct CS2 {
    double c, s;
};
// cos and sin of 2 pi m / N
constexpr CS2 cs2pi(int m, int N) {
    m = ((m % N) + N) % N;
    const int q = (4 * m) / N, rem = 4 * m - q * N;  // 2 pi m/N = (pi/2)(q + rem/N)
    const int k = (24 * rem) / N;                    // (pi/2) rem/N = k pi/48
    const double c = cos_pi48(k), s = cos_pi48(24 - k);
    return q == 0 ? CS2{c, s} : q == 1 ? CS2{-s, c} : q == 2 ? CS2{-c, -s} : CS2{s, -c};
}

template <typename T2> __host__ __device__ __forceinline__ T2 cxadd(T2 a, T2 b) { return T2{a.x + b.x, a.y + b.y}; }
template <typename T2> __host__ __device__ __forceinline__ T2 cxsub(T2 a, T2 b) { return T2{a.x - b.x, a.y - b.y}; }
template <typename T2> __host__ __device__ __forceinline__ T2 cxmul(T2 a, T2 b) {
    return T2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// x * w^m with w = e^{-2 pi i/N} (forward) or e^{+2 pi i/N} (inverse); m, N compile-time
template <bool INV, typename T2> __host__ __device__ __forceinline__ T2 rotc(T2 x, int m, int N) {
    m = ((m % N) + N) % N;
    if (m == 0) return x;
    if (2 * m == N) return T2{-x.x, -x.y};
    if (4 * m == N) return INV ? T2{-x.y, x.x} : T2{x.y, -x.x};
    if (4 * m == 3 * N) return INV ? T2{x.y, -x.x} : T2{-x.y, x.x};
    const CS2 w = cs2pi(m, N);
    using E = decltype(x.x);
    const E c = (E)w.c, s = (E)(INV ? w.s : -w.s);
    return T2{x.x * c - x.y * s, x.x * s + x.y * c};
}

// out[k] = sum_n in[n] w_N^{nk}, register arrays (constant indices after unrolling)
template <int N, bool INV> struct Dft {
    static constexpr int P = (N % 4 == 0) ? 4 : (N % 2 == 0 ? 2 : 3);
    static constexpr int Q = N / P;
    template <typename T2> __host__ __device__ static __forceinline__ void run(const T2* in, T2* out) {
        T2 y[P][Q];
#pragma unroll
        for (int r = 0; r < P; r++) {
            T2 sub[Q];
#pragma unroll
            for (int i = 0; i < Q; i++) sub[i] = in[r + P * i];
            Dft<Q, INV>::run(sub, y[r]);
        }
#pragma unroll
        for (int k2 = 0; k2 < Q; k2++) {
            T2 t[P], u[P];
#pragma unroll
            for (int r = 0; r < P; r++) t[r] = rotc<INV>(y[r][k2], r * k2, N);
            Dft<P, INV>::run(t, u);
#pragma unroll
            for (int k1 = 0; k1 < P; k1++) out[k2 + Q * k1] = u[k1];
        }
    }
};
template <bool INV> struct Dft<1, INV> {
    template <typename T2> __host__ __device__ static __forceinline__ void run(const T2* in, T2* out) { out[0] = in[0]; }
};
template <bool INV> struct Dft<2, INV> {
    template <typename T2> __host__ __device__ static __forceinline__ void run(const T2* in, T2* out) {
        out[0] = cxadd(in[0], in[1]);
        out[1] = cxsub(in[0], in[1]);
    }
};
template <bool INV> struct Dft<3, INV> {
    template <typename T2> __host__ __device__ static __forceinline__ void run(const T2* in, T2* out) {
        using E = decltype(in[0].x);
        const E h = (E)0.86602540378443864676;  // sqrt(3)/2
        const T2 t1 = cxadd(in[1], in[2]);
        const T2 t2 = T2{in[0].x - (E)0.5 * t1.x, in[0].y - (E)0.5 * t1.y};
        const T2 d = cxsub(in[1], in[2]);
        const T2 s = INV ? T2{-h * d.y, h * d.x} : T2{h * d.y, -h * d.x};  // -+ i h d
        out[0] = cxadd(in[0], t1);
        out[1] = cxadd(t2, s);
        out[2] = cxsub(t2, s);
    }
};
template <bool INV> struct Dft<4, INV> {
    template <typename T2> __host__ __device__ static __forceinline__ void run(const T2* in, T2* out) {
        const T2 a = cxadd(in[0], in[2]), b = cxsub(in[0], in[2]), c = cxadd(in[1], in[3]), d = cxsub(in[1], in[3]);
        const T2 md = INV ? T2{-d.y, d.x} : T2{d.y, -d.x};  // +- i d
        out[0] = cxadd(a, c);
        out[2] = cxsub(a, c);
        out[1] = cxadd(b, md);
        out[3] = cxsub(b, md);
    }
};

// Distributed N-point FFT (see the file comment).  Signal s, thread t of T.
// On entry a[m] = x[t + T m] (m < E).  On exit a[u T + j] = X[q + E j] for
// q = t + T u < E, j < T.  ex: LDS exchange of N*NS elements, element (i, s) at
// ex[i*NS + s]; the caller separates other uses of ex with barriers.  tw[k] =
// e^{-2 pi i k/N} (k < N), conjugated here for the inverse.
template <int N, int T, bool INV, typename T2, typename TW>
__device__ __forceinline__ void fft_dist(T2* a, int s, int t, int NS, T2* ex, const TW* tw) {
    constexpr int E = N / T;
    constexpr int U = (E + T - 1) / T;
    static_assert(E * T == N, "N = E T");
    static_assert(T <= E, "level 2 gives thread t the transforms q = t + T u < E: needs T <= E");
    T2 b[E];
    Dft<E, INV>::run(a, b);
#pragma unroll
    for (int q = 1; q < E; q++) {
        TW w = tw[(t * q) % N];
        if (INV) w.y = -w.y;
        b[q] = cxmul(b[q], T2{w.x, w.y});
    }
#pragma unroll
    for (int q = 0; q < E; q++) ex[(q * T + t) * NS + s] = b[q];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int q = t + T * u;
        if (U == 1 || q < E) {
            T2 c[T], d[T];
#pragma unroll
            for (int j = 0; j < T; j++) c[j] = ex[(q * T + j) * NS + s];
            Dft<T, INV>::run(c, d);
#pragma unroll
            for (int j = 0; j < T; j++) a[u * T + j] = d[j];
        }
    }
}

// fft_dist with the exchange in two rounds (U == 2: E = T + r, 0 < r <= T): round u passes the
// level-1 outputs q in [u T, min((u + 1) T, E)) through ex, which therefore holds T*T*NS
// elements (element (q - u T, j, s) at ex[((q - u T) * T + j) * NS + s]) instead of N*NS;
// thread t then does the level-2 transform q = t + T u (if q < E) and hands its outputs
// d[j] = X[q + E j] to sink(u, q, d) before the next round.  Same arithmetic as fft_dist.
// The caller separates earlier uses of ex with a barrier; ex is free again when this returns
// (after a barrier).
template <int N, int T, bool INV, typename T2, typename TW, class Sink>
__device__ __forceinline__ void fft_dist2(T2* a, int s, int t, int NS, T2* ex, const TW* tw, Sink sink) {
    constexpr int E = N / T;
    constexpr int U = (E + T - 1) / T;
    static_assert(E * T == N && U == 2, "two exchange rounds");
    T2 b[E];
    Dft<E, INV>::run(a, b);
#pragma unroll
    for (int q = 1; q < E; q++) {
        TW w = tw[(t * q) % N];
        if (INV) w.y = -w.y;
        b[q] = cxmul(b[q], T2{w.x, w.y});
    }
#pragma unroll
    for (int u = 0; u < 2; u++) {
        const int nq = u == 0 ? T : E - T;
#pragma unroll
        for (int qq = 0; qq < T; qq++)
            if (qq < nq) ex[(qq * T + t) * NS + s] = b[u * T + qq];
        __syncthreads();
        const int q = t + T * u;
        if (t < nq) {
            T2 c[T], d[T];
#pragma unroll
            for (int j = 0; j < T; j++) c[j] = ex[(t * T + j) * NS + s];
            Dft<T, INV>::run(c, d);
            sink(u, q, d);
        }
        __syncthreads();
    }
}

// Shapes of the distributed transforms for the sizes that have kernels
// (T threads per signal, E points per thread, U level-2 transforms per thread).
template <int N> struct Shape;
template <> struct Shape<16> { static constexpr int T = 4; };
template <> struct Shape<32> { static constexpr int T = 4; };
template <> struct Shape<48> { static constexpr int T = 4; };
template <> struct Shape<64> { static constexpr int T = 8; };
template <> struct Shape<96> { static constexpr int T = 8; };
template <> struct Shape<128> { static constexpr int T = 8; };
template <> struct Shape<256> { static constexpr int T = 16; };

}  // namespace cg
