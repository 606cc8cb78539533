// Vector kernels of libgip_hip: fill, axpy, axpby, dot, infinity norm, CFL.
// Semantics of GIP_VectorMulticore.cpp:44-195 and GIP_NonLinearMulticore.cpp:3-55;
// reductions are deterministic: fixed grid, fixed per-thread element sets,
// fixed wavefront/LDS combination tree, and the last workgroup to finish
// combines the per-workgroup partials in index order.
#include "gip_blas1.h"
#include "gip_internal.h"

namespace gip {

__global__ __launch_bounds__(kBlock) void k_fill(int n, double *__restrict__ x, double v) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) x[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_axpy(int n, double a, const double *__restrict__ x, double *__restrict__ y) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[i] = y[i] + a * x[i];
}

__global__ __launch_bounds__(kBlock) void k_axpby(int n, double a, const double *__restrict__ x, double b,
                                                  double *__restrict__ y) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * y[i];
}

// y[dst ? dst[i] : i] = x[src ? src[i] : i]: the halo pack and unpack
// (send list -> send block, senders' blocks -> halo slots)
__global__ __launch_bounds__(kBlock) void k_copy_indexed(int n, const int32_t *__restrict__ src,
                                                        const double *__restrict__ x, const int32_t *__restrict__ dst,
                                                        double *__restrict__ y) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[dst ? dst[i] : i] = x[src ? src[i] : i];
}

__global__ __launch_bounds__(kBlock) void k_copy(int n, const double *__restrict__ x, double *__restrict__ y) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[i] = x[i];
}

// Reductions run as two launches: every workgroup stores its partial, then
// one workgroup combines the partials in index order.  The kernel boundary
// orders the two, so no atomics or device-wide fences are needed (a fence
// per workgroup costs an L2 write-back on every XCD).
template <int OP>
__global__ __launch_bounds__(kBlock) void k_reduce(int n, const double *__restrict__ x, const double *__restrict__ y,
                                                   double *__restrict__ partials) {
    double acc = RedOp<OP>::identity();
    const int stride = gridDim.x * kBlock;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) acc = RedOp<OP>::combine(acc, RedOp<OP>::term(x, y, i));
    acc = block_reduce<OP>(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

// one workgroup: combine `np` partials; CFL = 1 applies GIP_NonLinearMulticore.cpp:47-50
template <int OP, int CFL>
__global__ __launch_bounds__(kBlock) void k_finalize(const double *__restrict__ partials, int np,
                                                    double *__restrict__ result) {
    const double acc = sum_partials<OP>(partials, np);
    if (threadIdx.x == 0) *result = CFL ? (0.8 * acc > 1e-1 ? 1e-1 : 0.8 * acc) : acc;
}

struct CflArgs {
    const double *u, *v, *w, *dxs;
    double gamma, rho;
};

__global__ __launch_bounds__(kBlock) void k_cfl(int n, CflArgs a, double *__restrict__ partials) {
    // GIP_NonLinearMulticore.cpp:14-43: the running minimum starts at 1e20 and
    // takes each admissible convective and diffusive limit.
    double acc = 1e20;
    const int stride = gridDim.x * kBlock;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        double ut = a.u[i], vt = a.v[i], wt = a.w[i], dx = a.dxs[i];
        double normit = fabs(ut) > fabs(vt) ? fabs(ut) : fabs(vt);
        if (normit < fabs(wt)) normit = fabs(wt);
        if (!(dx / normit < 1e-17)) {
            double c = 0.25 * dx / normit;
            if (acc > c) acc = c;
        }
        if (!(dx * dx * a.rho / a.gamma < 1e-17)) {
            double c = 0.2 * dx * dx * a.rho / a.gamma;
            if (acc > c) acc = c;
        }
    }
    acc = block_reduce<RED_MIN>(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

static inline int blocks_for(int n) { return (n + kBlock - 1) / kBlock; }

int launch_reduce(gip_ctx *c, int op, int n, const double *x, const double *y, double *dres) {
    int g = reduce_grid(n);
    double *part = c->d_partials;
    switch (op) {
        case RED_SUM:
            k_reduce<RED_SUM><<<g, kBlock, 0, c->stream>>>(n, x, y, part);
            k_finalize<RED_SUM, 0><<<1, kBlock, 0, c->stream>>>(part, g, dres);
            break;
        case RED_MAXABS:
            k_reduce<RED_MAXABS><<<g, kBlock, 0, c->stream>>>(n, x, y, part);
            k_finalize<RED_MAXABS, 0><<<1, kBlock, 0, c->stream>>>(part, g, dres);
            break;
        default:
            return set_error(GIP_EINVAL, "unknown reduction");
    }
    GIP_LAUNCH_CHECK("reduce kernel");
    return GIP_OK;
}

// A reduction whose result the host needs: the finalize kernel stores it
// straight into the context's pinned host slot (no device-to-host copy
// command; the reference's PCG sequence makes about 300 such round trips
// per step: C100K 5.22 -> 4.87 ms per step), read after the stream drains.
static int fetch_result(gip_ctx *c, double *out) {
    GIP_HIP_TRY(hipStreamSynchronize(c->stream));
    *out = c->h_result[0];
    return GIP_OK;
}

}  // namespace gip

using namespace gip;

extern "C" {

int gip_dfill(gip_ctx *c, int32_t n, double *x, double value) {
    if (!c || n < 0 || (n && !x)) return set_error(GIP_EINVAL, "gip_dfill: bad argument");
    if (!n) return GIP_OK;
    k_fill<<<blocks_for(n), kBlock, 0, c->stream>>>(n, x, value);
    GIP_LAUNCH_CHECK("gip_dfill");
    return GIP_OK;
}

int gip_daxpy(gip_ctx *c, int32_t n, double alpha, const double *x, double *y) {
    if (!c || n < 0 || (n && (!x || !y))) return set_error(GIP_EINVAL, "gip_daxpy: bad argument");
    if (!n) return GIP_OK;
    k_axpy<<<blocks_for(n), kBlock, 0, c->stream>>>(n, alpha, x, y);
    GIP_LAUNCH_CHECK("gip_daxpy");
    return GIP_OK;
}

int gip_daxpby(gip_ctx *c, int32_t n, double alpha, const double *x, double beta, double *y) {
    if (!c || n < 0 || (n && (!x || !y))) return set_error(GIP_EINVAL, "gip_daxpby: bad argument");
    if (!n) return GIP_OK;
    k_axpby<<<blocks_for(n), kBlock, 0, c->stream>>>(n, alpha, x, beta, y);
    GIP_LAUNCH_CHECK("gip_daxpby");
    return GIP_OK;
}

int gip_dcopy(gip_ctx *c, int32_t n, const double *x, double *y) {
    if (!c || n < 0 || (n && (!x || !y))) return set_error(GIP_EINVAL, "gip_dcopy: bad argument");
    if (!n) return GIP_OK;
    k_copy<<<blocks_for(n), kBlock, 0, c->stream>>>(n, x, y);
    GIP_LAUNCH_CHECK("gip_dcopy");
    return GIP_OK;
}

int gip_dcopy_indexed(gip_ctx *c, int32_t n, const int32_t *src, const double *x, const int32_t *dst, double *y) {
    if (!c || n < 0 || (n && (!x || !y))) return set_error(GIP_EINVAL, "gip_dcopy_indexed: bad argument");
    if (!n) return GIP_OK;
    k_copy_indexed<<<blocks_for(n), kBlock, 0, c->stream>>>(n, src, x, dst, y);
    GIP_LAUNCH_CHECK("gip_dcopy_indexed");
    return GIP_OK;
}

int gip_ddot_dev(gip_ctx *c, int32_t n, const double *x, const double *y, double *dres) {
    if (!c || n < 0 || (n && (!x || !y)) || !dres) return set_error(GIP_EINVAL, "gip_ddot: bad argument");
    return launch_reduce(c, RED_SUM, n, x, y, dres);
}

int gip_ddot(gip_ctx *c, int32_t n, const double *x, const double *y, double *result) {
    if (!result) return set_error(GIP_EINVAL, "gip_ddot: null result");
    int rc = gip_ddot_dev(c, n, x, y, c ? c->h_result_dev : nullptr);
    if (rc) return rc;
    return fetch_result(c, result);
}

int gip_dnrm_inf_dev(gip_ctx *c, int32_t n, const double *x, double *dres) {
    if (!c || n < 0 || (n && !x) || !dres) return set_error(GIP_EINVAL, "gip_dnrm_inf: bad argument");
    return launch_reduce(c, RED_MAXABS, n, x, x, dres);
}

int gip_dnrm_inf(gip_ctx *c, int32_t n, const double *x, double *result) {
    if (!result) return set_error(GIP_EINVAL, "gip_dnrm_inf: null result");
    int rc = gip_dnrm_inf_dev(c, n, x, c ? c->h_result_dev : nullptr);
    if (rc) return rc;
    return fetch_result(c, result);
}

int gip_cfl(gip_ctx *c, int32_t n, const double *u, const double *v, const double *w, const double *dxs,
            double gamma, double rho, double *dt) {
    if (!c || n < 0 || (n && (!u || !v || !w || !dxs)) || !dt) return set_error(GIP_EINVAL, "gip_cfl: bad argument");
    CflArgs a{u, v, w, dxs, gamma, rho};
    int g = reduce_grid(n);
    k_cfl<<<g, kBlock, 0, c->stream>>>(n, a, c->d_partials);
    k_finalize<RED_MIN, 1><<<1, kBlock, 0, c->stream>>>(c->d_partials, g, c->h_result_dev);
    GIP_LAUNCH_CHECK("gip_cfl");
    return fetch_result(c, dt);
}

}  // extern "C"
