// Device-side reduction helpers shared by the vector and PCG kernels.
// Wave64 shuffle tree, then an LDS tree across the workgroup's wavefronts;
// per-workgroup partials are stored with plain stores and combined, in a
// fixed order, by a later kernel (no atomics, no device-scope fences).
#pragma once

#include <hip/hip_runtime.h>

#include "gip_internal.h"

namespace gip {

enum { RED_SUM = 0, RED_MAXABS = 1, RED_MIN = 2 };

template <int OP>
struct RedOp;

template <>
struct RedOp<RED_SUM> {
    __device__ static double identity() { return 0.0; }
    __device__ static double combine(double a, double b) { return a + b; }
    __device__ static double term(const double *x, const double *y, int i) { return x[i] * y[i]; }
};

template <>
struct RedOp<RED_MAXABS> {
    __device__ static double identity() { return 0.0; }
    // NaN terms never replace the running value, as in GIP_VectorMulticore.cpp:151-153
    __device__ static double combine(double a, double b) { return b > a ? b : a; }
    __device__ static double term(const double *x, const double *, int i) { return fabs(x[i]); }
};

template <>
struct RedOp<RED_MIN> {
    __device__ static double identity() { return 1e20; }
    __device__ static double combine(double a, double b) { return b < a ? b : a; }
    __device__ static double term(const double *x, const double *, int i) { return x[i]; }
};

constexpr int kWaves = kBlock / 64;  // wavefronts per workgroup

// Sum over the wavefront (shuffle tree, no LDS or barrier); valid in lane 0.
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = v + __shfl_down(v, off, 64);
    return v;
}

// Reduce v over the workgroup; the result is valid in thread 0.
template <int OP>
__device__ inline double block_reduce(double v) {
    __shared__ double sh[kBlock / 64];
    __syncthreads();  // protect sh against a previous call
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = RedOp<OP>::combine(v, __shfl_down(v, off, 64));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        v = sh[0];
#pragma unroll
        for (int k = 1; k < kBlock / 64; ++k) v = RedOp<OP>::combine(v, sh[k]);
    }
    return v;
}

// M sums over the workgroup at once, valid in thread 0: each value takes
// block_reduce's shuffle tree and final order (so each total is bitwise
// block_reduce<RED_SUM>'s), and all M share one LDS exchange and its pair
// of barriers.
template <int M>
__device__ inline void block_reduce_n(double (&v)[M]) {
    __shared__ double sh[M][kBlock / 64];
    __syncthreads();  // protect sh against a previous call
#pragma unroll
    for (int m = 0; m < M; ++m)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[m] = v[m] + __shfl_down(v[m], off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int m = 0; m < M; ++m) sh[m][wid] = v[m];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int m = 0; m < M; ++m) {
            v[m] = sh[m][0];
#pragma unroll
            for (int k = 1; k < kBlock / 64; ++k) v[m] = v[m] + sh[m][k];
        }
}

// M sums of partials at once, for partial arrays of kMaxRedBlocks doubles
// whose entries past their producer's grid hold zeros.  Every load is
// issued unconditionally (a load under a bounds test would be turned into a
// branch with a wait at its join): a slot past n[m] is read from the array's
// last entry, a zero whenever n[m] < kMaxRedBlocks, so those reads are one
// broadcast cache line.  All loads are issued before the first is used
// (load(), early in a kernel) and reduce() sums them: the M sums cost one
// memory round trip.  The zero slots make each total bitwise the one
// sum_partials<RED_SUM> gives over the first n[m] entries.
template <int M>
struct PaddedSums {
    static constexpr int K = kMaxRedBlocks / kBlock;
    double v[M][K];
    __device__ inline void load(const double *const (&part)[M], const int (&n)[M]) {
#pragma unroll
        for (int m = 0; m < M; ++m)
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int i = threadIdx.x + k * kBlock;
                v[m][k] = part[m][i < n[m] ? i : kMaxRedBlocks - 1];
            }
    }
    __device__ inline void reduce(double (&out)[M]) {
        __shared__ double s_all[M];
        double acc[M];
#pragma unroll
        for (int m = 0; m < M; ++m) {
            acc[m] = 0.0;
#pragma unroll
            for (int k = 0; k < K; ++k) acc[m] = acc[m] + v[m][k];
        }
        block_reduce_n<M>(acc);
        if (threadIdx.x == 0)
#pragma unroll
            for (int m = 0; m < M; ++m) s_all[m] = acc[m];
        __syncthreads();
#pragma unroll
        for (int m = 0; m < M; ++m) out[m] = s_all[m];
    }
};

template <int M>
__device__ inline void sum_padded(const double *const (&part)[M], const int (&n)[M], double (&out)[M]) {
    PaddedSums<M> ps;
    ps.load(part, n);
    ps.reduce(out);
}

// Combine n partials in index order (thread t takes t, t+256, ..., then the
// fixed tree): the same value in every workgroup that calls it.  Used in
// kernel prologues to read the previous kernel's per-workgroup partials
// without atomics or cross-XCD fences (the kernel boundary orders them).
template <int OP>
__device__ inline double sum_partials(const double *__restrict__ part, int n) {
    __shared__ double s_all;
    // n <= kMaxRedBlocks: all of a thread's loads are issued before the first
    // is used (one memory round trip), then combined in index order
    constexpr int K = kMaxRedBlocks / kBlock;
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = threadIdx.x + k * kBlock;
        v[k] = i < n ? part[i] : RedOp<OP>::identity();
    }
    double acc = RedOp<OP>::identity();
#pragma unroll
    for (int k = 0; k < K; ++k) acc = RedOp<OP>::combine(acc, v[k]);
    acc = block_reduce<OP>(acc);
    if (threadIdx.x == 0) s_all = acc;
    __syncthreads();
    return s_all;
}

// M arrays of n partials summed together (n <= kMaxRedBlocks): only the
// slots below n are read, in a workgroup-uniform loop of ceil(n / 256)
// rounds, the M arrays' loads of a round issued together; the same total in
// every workgroup (thread t takes t, t+256, ..., then the fixed tree).
template <int M>
__device__ inline void sum_partials_n(const double *const (&part)[M], int n, double (&out)[M]) {
    __shared__ double s_all[M];
    double acc[M];
#pragma unroll
    for (int m = 0; m < M; ++m) acc[m] = 0.0;
    const int rounds = (n + kBlock - 1) / kBlock;
    for (int k = 0; k < rounds; ++k) {
        const int i = threadIdx.x + k * kBlock;
        double v[M];
#pragma unroll
        for (int m = 0; m < M; ++m) v[m] = i < n ? part[m][i] : 0.0;
#pragma unroll
        for (int m = 0; m < M; ++m) acc[m] = acc[m] + v[m];
    }
    block_reduce_n<M>(acc);
    if (threadIdx.x == 0)
#pragma unroll
        for (int m = 0; m < M; ++m) s_all[m] = acc[m];
    __syncthreads();
#pragma unroll
    for (int m = 0; m < M; ++m) out[m] = s_all[m];
}

int launch_reduce(gip_ctx *c, int op, int n, const double *x, const double *y, double *dres);

}  // namespace gip
