// reduce.h -- deterministic block reductions (64-wide wavefront shuffles,
// then a fixed-order pass over the block's waves through LDS).
#pragma once
#include <hip/hip_runtime.h>

namespace cg {

struct OpSum { __device__ static double init() { return 0.0; } __device__ static double f(double a, double b) { return a + b; } };
struct OpMax { __device__ static double init() { return -1.7976931348623157e308; } __device__ static double f(double a, double b) { return a > b ? a : b; } };
struct OpMin { __device__ static double init() { return 1.7976931348623157e308; } __device__ static double f(double a, double b) { return a < b ? a : b; } };

__device__ __forceinline__ double wave_shfl_down(double v, int o) { return __shfl_down(v, o, 64); }

// Reduce acc[0..NV) over the block; thread 0..NV-1 write part[blockIdx.x*NV + i].
template <int NV, int BS, class Op = OpSum>
__device__ __forceinline__ void block_reduce_store(const double* acc, double* part) {
    static_assert(NV <= BS, "NV must not exceed the block size");
    __shared__ double sm[(BS / 64) * NV];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++) {
        double v = acc[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v = Op::f(v, wave_shfl_down(v, o));
        if (lane == 0) sm[wid * NV + i] = v;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = sm[threadIdx.x];
        for (int w = 1; w < BS / 64; w++) s = Op::f(s, sm[w * NV + threadIdx.x]);
        part[(size_t)blockIdx.x * NV + threadIdx.x] = s;
    }
}

// Same, writing part[blockIdx.x*stride + off + i] (several reductions sharing one row).
template <int NV, int BS, class Op = OpSum>
__device__ __forceinline__ void block_reduce_store_at(const double* acc, double* part, int stride, int off) {
    static_assert(NV <= BS, "NV must not exceed the block size");
    __shared__ double sm[(BS / 64) * NV];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++) {
        double v = acc[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v = Op::f(v, wave_shfl_down(v, o));
        if (lane == 0) sm[wid * NV + i] = v;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = sm[threadIdx.x];
        for (int w = 1; w < BS / 64; w++) s = Op::f(s, sm[w * NV + threadIdx.x]);
        part[(size_t)blockIdx.x * stride + off + threadIdx.x] = s;
    }
}

// Same, with caller-provided LDS scratch of (BS/64)*NV doubles (callers that alias it
// onto LDS they use elsewhere must separate the uses with barriers).
template <int NV, int BS, class Op = OpSum>
__device__ __forceinline__ void block_reduce_store_at(const double* acc, double* part, int stride, int off,
                                                      double* sm) {
    static_assert(NV <= BS, "NV must not exceed the block size");
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++) {
        double v = acc[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v = Op::f(v, wave_shfl_down(v, o));
        if (lane == 0) sm[wid * NV + i] = v;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = sm[threadIdx.x];
        for (int w = 1; w < BS / 64; w++) s = Op::f(s, sm[w * NV + threadIdx.x]);
        part[(size_t)blockIdx.x * stride + off + threadIdx.x] = s;
    }
}

// Block reduction through an LDS transpose (no cross-lane shuffles, so no lane-index
// registers): sm holds NV*BS doubles; fixed summation order (32-element chunks in
// order, then the BS/32 chunk sums in order), NV <= 32 and BS a multiple of 32.
// Ends with every thread past a barrier after its last read of sm... except the
// final NV readers: callers that reuse sm afterwards must add a barrier.
template <int NV, int BS>
__device__ __forceinline__ void block_reduce_lds(const double* acc, double* part, int stride, int off, double* sm) {
    static_assert(NV * (BS / 32) <= BS, "one thread per (value, chunk)");
    constexpr int NC = BS / 32;
    const int t = threadIdx.x;
#pragma unroll
    for (int i = 0; i < NV; i++) sm[i * BS + t] = acc[i];
    __syncthreads();
    double s = 0.0;
    if (t < NV * NC) {
        const double* c = sm + (t / NC) * BS + (t % NC) * 32;
        for (int k = 0; k < 32; k++) s += c[k];
    }
    __syncthreads();
    if (t < NV * NC) sm[(t / NC) * BS + (t % NC) * 32] = s;
    __syncthreads();
    if (t < NV) {
        double r = sm[t * BS];
        for (int c = 1; c < NC; c++) r += sm[t * BS + c * 32];
        part[(size_t)blockIdx.x * stride + off + t] = r;
    }
}

// Column-major variants: block b writes its sums to part[(off + i) * gridDim.x + b], so the
// final reduction of value i reads one contiguous column (launches with many blocks).
template <int NV, int BS>
__device__ __forceinline__ void block_reduce_store_cm(const double* acc, double* part, int off, long blk, long nblk) {
    static_assert(NV <= BS, "NV must not exceed the block size");
    __shared__ double sm[(BS / 64) * NV];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++) {
        double v = acc[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += wave_shfl_down(v, o);
        if (lane == 0) sm[wid * NV + i] = v;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = sm[threadIdx.x];
        for (int w = 1; w < BS / 64; w++) s += sm[w * NV + threadIdx.x];
        part[(size_t)(off + threadIdx.x) * nblk + blk] = s;
    }
}
template <int NV, int BS>
__device__ __forceinline__ void block_reduce_store_cm(const double* acc, double* part, int off) {
    block_reduce_store_cm<NV, BS>(acc, part, off, blockIdx.x, gridDim.x);
}
// (blk, nblk): this block's index and the block count of the launch it belongs to (a batched
// launch covers several of them)
template <int NV, int BS>
__device__ __forceinline__ void block_reduce_lds_cm(const double* acc, double* part, int off, double* sm, long blk,
                                                    long nblk) {
    static_assert(NV * (BS / 32) <= BS, "one thread per (value, chunk)");
    constexpr int NC = BS / 32;
    const int t = threadIdx.x;
#pragma unroll
    for (int i = 0; i < NV; i++) sm[i * BS + t] = acc[i];
    __syncthreads();
    double s = 0.0;
    if (t < NV * NC) {
        const double* c = sm + (t / NC) * BS + (t % NC) * 32;
        for (int k = 0; k < 32; k++) s += c[k];
    }
    __syncthreads();
    if (t < NV * NC) sm[(t / NC) * BS + (t % NC) * 32] = s;
    __syncthreads();
    if (t < NV) {
        double r = sm[t * BS];
        for (int c = 1; c < NC; c++) r += sm[t * BS + c * 32];
        part[(size_t)(off + t) * nblk + blk] = r;
    }
}
template <int NV, int BS>
__device__ __forceinline__ void block_reduce_lds_cm(const double* acc, double* part, int off, double* sm) {
    block_reduce_lds_cm<NV, BS>(acc, part, off, sm, blockIdx.x, gridDim.x);
}

}  // namespace cg
