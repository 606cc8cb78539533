// Streaming-bandwidth calibration of libgip_hip: the practical HBM ceiling
// that the library's bandwidth-bound kernels (SpMV, the PCG's direction and
// update sweeps) are measured against, beside the 8 TB/s published peak.
//
// Four access mixes over n doubles (arrays far beyond the 256 MB Infinity
// Cache): copy (b = a), read (sum of a, one partial per workgroup), write
// (a = value), and the PCG direction sweep's mix (three reads, two writes:
// c += 0.5 a, a = b + 0.25 a).  Each in 8-byte or 16-byte accesses, with
// plain or nontemporal loads and stores, and one of two launch shapes: one
// access per thread over the whole array (workgroups in dispatch order), or
// `grid` workgroups striding over it in dispatch order.
#include "gip_internal.h"

namespace gip {

enum { ST_COPY = 0, ST_READ = 1, ST_WRITE = 2, ST_MIX = 3 };

typedef double dvec2 __attribute__((ext_vector_type(2)));  // one 16-byte access

template <int V>
struct Vec;
template <>
struct Vec<1> {
    using T = double;
};
template <>
struct Vec<2> {
    using T = dvec2;
};

template <class T, bool NT>
__device__ inline T ld(const T *p) {
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}
template <class T, bool NT>
__device__ inline void st(T *p, T v) {
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

__device__ inline double vsum(double a) { return a; }
__device__ inline double vsum(dvec2 a) { return a.x + a.y; }
__device__ inline dvec2 vmix(dvec2 a, double s, dvec2 b) { return a + s * b; }
__device__ inline double vmix(double a, double s, double b) { return a + s * b; }
template <class T>
__device__ inline T vfill(double v);
template <>
__device__ inline double vfill<double>(double v) { return v; }
template <>
__device__ inline dvec2 vfill<dvec2>(double v) { return dvec2{v, v}; }

template <int KIND, int V, bool NT>
__global__ __launch_bounds__(kBlock) void k_stream(size_t m, typename Vec<V>::T *__restrict__ a,
                                                   typename Vec<V>::T *__restrict__ b,
                                                   typename Vec<V>::T *__restrict__ c, double *__restrict__ part) {
    using T = typename Vec<V>::T;
    const size_t stride = (size_t)gridDim.x * kBlock;
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < m; i += stride) {
        if constexpr (KIND == ST_COPY) {
            st<T, NT>(b + i, ld<T, NT>(a + i));
        } else if constexpr (KIND == ST_READ) {
            acc += vsum(ld<T, NT>(a + i));
        } else if constexpr (KIND == ST_WRITE) {
            st<T, NT>(a + i, vfill<T>(1.0));
        } else {
            const T ai = ld<T, NT>(a + i), bi = ld<T, NT>(b + i), ci = ld<T, NT>(c + i);
            st<T, NT>(c + i, vmix(ci, 0.5, ai));
            st<T, NT>(a + i, vmix(bi, 0.25, ai));
        }
    }
    if constexpr (KIND == ST_READ) {
        // one partial per workgroup keeps the loads live
        __shared__ double sh[kBlock / 64];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
    }
}

template <int KIND, int V, bool NT>
static void launch_stream(hipStream_t s, unsigned grid, size_t m, double *a, double *b, double *c, double *part) {
    using T = typename Vec<V>::T;
    k_stream<KIND, V, NT><<<grid, kBlock, 0, s>>>(m, (T *)a, (T *)b, (T *)c, part);
}

template <int KIND>
static void dispatch_stream(hipStream_t s, int vec, bool nt, unsigned grid, size_t m, double *a, double *b, double *c,
                            double *part) {
    if (vec == 2)
        nt ? launch_stream<KIND, 2, true>(s, grid, m, a, b, c, part)
           : launch_stream<KIND, 2, false>(s, grid, m, a, b, c, part);
    else
        nt ? launch_stream<KIND, 1, true>(s, grid, m, a, b, c, part)
           : launch_stream<KIND, 1, false>(s, grid, m, a, b, c, part);
}

// A device-wide barrier, `iters` times, across a cooperative launch (every
// workgroup resident): each workgroup's thread 0 adds itself to a counter
// with a relaxed device-scope atomic and waits until the counter reaches
// the barrier's target.  FENCE adds a device-scope fence before the arrival
// and after the wait (what a barrier that publishes data written before it
// needs); PARTIALS also stores one value per workgroup before the barrier
// and sums every workgroup's value after it.  The wait is bounded: past
// `limit` polls (or once another workgroup has given up) a workgroup sets
// *failed and stops waiting, so every wave exits.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_barrier_probe(unsigned *count, double *part, int iters, int *failed,
                                                          unsigned limit) {
    double acc = 0.0;
    for (int k = 1; k <= iters; ++k) {
        if (MODE == 2 && threadIdx.x == 0) part[blockIdx.x] = (double)(k + blockIdx.x);
        __syncthreads();
        if (threadIdx.x == 0) {
            if (MODE >= 1) __threadfence();
            __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned target = (unsigned)k * gridDim.x;
            unsigned polls = 0;
            while (__hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                if (++polls > limit || __hip_atomic_load(failed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    __hip_atomic_store(failed, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (MODE >= 1) __threadfence();
        }
        __syncthreads();
        if (MODE == 2) {
            double v = 0.0;
            for (int i = threadIdx.x; i < (int)gridDim.x; i += kBlock)
                v += __hip_atomic_load(&part[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            acc += v;
            __syncthreads();
        }
    }
    if (MODE == 2 && blockIdx.x == 0 && threadIdx.x == 0) part[gridDim.x] = acc;
}

}  // namespace gip

using namespace gip;

extern "C" int gip_barrier_probe(gip_ctx *c, int32_t grid, int32_t iters, int32_t mode, float *us_per_barrier) {
    if (!c || !us_per_barrier || grid < 1 || iters < 1 || mode < 0 || mode > 2)
        return set_error(GIP_EINVAL, "gip_barrier_probe: bad argument");
    int per_cu = 0;
    const void *fn = mode == 0 ? (const void *)k_barrier_probe<0>
                     : mode == 1 ? (const void *)k_barrier_probe<1>
                                 : (const void *)k_barrier_probe<2>;
    GIP_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock, 0));
    if (grid > per_cu * c->num_cu) return set_error(GIP_EINVAL, "gip_barrier_probe: more workgroups than can be resident");
    unsigned *count = nullptr;
    int *failed = nullptr;
    double *part = nullptr;
    GIP_HIP_TRY(hipMalloc(&count, sizeof(unsigned)));
    GIP_HIP_TRY(hipMalloc(&failed, sizeof(int)));
    GIP_HIP_TRY(hipMalloc(&part, sizeof(double) * ((size_t)grid + 1)));
    hipError_t e = hipMemsetAsync(count, 0, sizeof(unsigned), c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(failed, 0, sizeof(int), c->stream);
    unsigned limit = 200000;
    void *args[] = {&count, &part, &iters, &failed, &limit};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
    if (e == hipSuccess) e = hipLaunchCooperativeKernel(fn, dim3(grid), dim3(kBlock), args, 0, c->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    int f = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess) e = hipMemcpy(&f, failed, sizeof(int), hipMemcpyDeviceToHost);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(count);
    hipFree(failed);
    hipFree(part);
    if (e != hipSuccess) return hip_fail(e, "gip_barrier_probe");
    if (f) return set_error(GIP_EHIP, "gip_barrier_probe: a barrier wait gave up");
    *us_per_barrier = 1e3f * ms / (float)iters;
    return GIP_OK;
}

extern "C" int gip_stream_probe(gip_ctx *c, int32_t kind, int32_t vec, int32_t nontemporal, int32_t grid, size_t n,
                                double *a, double *b, double *cc, double *part, size_t part_len, int32_t reps,
                                float *ms) {
    if (!c || !ms || kind < ST_COPY || kind > ST_MIX || (vec != 1 && vec != 2) || grid < 0 || reps < 1 || !a ||
        (kind != ST_READ && kind != ST_WRITE && !b) || (kind == ST_MIX && !cc) || n % 2)
        return set_error(GIP_EINVAL, "gip_stream_probe: bad argument");
    const size_t m = n / (size_t)vec;  // accesses
    const size_t per_thread_grid = (m + kBlock - 1) / kBlock;
    if (grid == 0 && per_thread_grid > 0x7fffffffu) return set_error(GIP_EINVAL, "gip_stream_probe: array too long");
    const unsigned g = grid ? (unsigned)grid : (unsigned)per_thread_grid;
    if (kind == ST_READ && (!part || g > part_len))
        return set_error(GIP_EINVAL, "gip_stream_probe: read needs a partial slot per workgroup");
    hipStream_t s = c->stream;
    hipEvent_t e0, e1;
    GIP_HIP_TRY(hipEventCreate(&e0));
    GIP_HIP_TRY(hipEventCreate(&e1));
    auto one = [&]() {
        switch (kind) {
            case ST_COPY: dispatch_stream<ST_COPY>(s, vec, nontemporal != 0, g, m, a, b, cc, part); break;
            case ST_READ: dispatch_stream<ST_READ>(s, vec, nontemporal != 0, g, m, a, b, cc, part); break;
            case ST_WRITE: dispatch_stream<ST_WRITE>(s, vec, nontemporal != 0, g, m, a, b, cc, part); break;
            default: dispatch_stream<ST_MIX>(s, vec, nontemporal != 0, g, m, a, b, cc, part); break;
        }
    };
    one();  // warm-up
    hipEventRecord(e0, s);
    for (int r = 0; r < reps; ++r) one();
    hipEventRecord(e1, s);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (e != hipSuccess) return hip_fail(e, "gip_stream_probe");
    *ms = t / (float)reps;
    return GIP_OK;
}
