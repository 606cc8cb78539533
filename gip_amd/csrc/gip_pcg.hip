// Fused, device-resident Jacobi-PCG of libgip_hip (GIP_LinearSolver.h:45-117).
//
// FP64 iteration k, two kernels, the reference recurrences and roundings:
//   K1  p_k = 1.0 s + beta p_{k-1}   (k >= 2; formed on the fly for every
//       gathered column, stored once for the own rows; the first iteration
//       uses p_1 from the setup), Ap = A p_k, dn = p_k.Ap -> alpha = r1/dn
//   K2  x += alpha p_k ; r += (-alpha) Ap ; s = Dinv r ; r.s, r.r
//       -> beta = r1'/r1, convergence, iteration count
// The direction lives in two buffers used alternately (K1 of iteration k
// reads p_{k-1} from one and writes p_k to the other), so no workgroup can
// see a half-updated direction.  Forming p on the fly in K1 computes
// exactly the expression the reference stores (GIP_LinearSolver.h:113).
// Every scalar lives in device memory; every kernel first reads the
// solver's state word and returns when the solve has finished, so the host
// queues many iterations and polls the state only now and then.
//
// MIXED: the same two kernels in FP32 (operator values, vectors and SpMV;
// dot products accumulate in FP64) on the correction equation, inside an
// FP64 defect-correction loop: when the FP32 residual has dropped by
// inner_tol (or reaches the target) the FP32 correction is added to the
// FP64 x, the FP64 true residual b - A x replaces the FP32 residual (the
// search direction is kept: residual replacement), and convergence is
// decided on that FP64 residual only.
#include <cmath>
#include <new>

#include "gip_blas1.h"
#include "gip_internal.h"
#include "gip_rows.h"
#include "gip_spmv.h"

namespace gip {

// Device scalar slots of a solver.
enum {
    S_RHO = 0,       // r.s of the current iterate
    S_ALPHA = 1,
    S_BETA = 2,
    S_RR = 3,        // r.r
    S_BB = 4,        // b.b
    S_DN = 5,
    S_TOL2 = 6,      // (tol ||b||)^2, negative when no tolerance
    S_TRUE_RR = 7,   // FP64 true residual of the last correction (MIXED)
    S_RHO_PREV = 8,  // r.s before the last update (MIXED)
    S_COUNT = 16
};
// State word: 0 running, 1 converged, 2 breakdown.
enum { ST_RUN = 0, ST_CONV = 1, ST_BREAK = 2 };

struct PcgDev {
    int n;
    double *x, *r, *s, *Ap;
    const double *b, *dinv;
    float *r32, *s32, *Ap32, *xc32;
    const float *dinv32;
    double *sc;  // S_COUNT scalars
    int *state, *iters;
    double *part;       // reduction partials
    unsigned int *cnt;  // counters
    int precond;
};

__device__ inline bool running(const PcgDev &d) { return *(volatile int *)d.state == ST_RUN; }

// ---- setup ---------------------------------------------------------------

// x = x0 (all), direction buffers = x0 outside the solved rows,
// Ap = (A x0)_i or x0_i on empty rows, r = b + (-1.0) Ap   (:55-70)
template <int W>
__global__ __launch_bounds__(kBlock) void k_setup_resid(EllView A, PcgDev d, const double *__restrict__ x0, int n_all,
                                                        double *__restrict__ p0, double *__restrict__ p1) {
    const int stride = gridDim.x * kBlock;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n_all; i += stride) {
        d.x[i] = x0[i];
        if (i >= d.n) {
            p0[i] = x0[i];
            p1[i] = x0[i];
        }
    }
    Operand<double> xo{x0, nullptr, 0.0, d.n, false};
    FOR_ROWS(d.n, row) {
        double sum;
        bool any = row_dot<W, double, double>(A, A.val, xo, row, sum);
        double t = any ? sum : x0[row];
        d.Ap[row] = t;
        d.r[row] = d.b[row] + -1.0 * t;
    }
}

// p = Dinv r (or r); r1 = r.p, b.b, r.r  (:72-80); state = converged when ||r|| <= tol ||b||
__global__ __launch_bounds__(kBlock) void k_setup_dir(PcgDev d, double *__restrict__ p, double tol) {
    double rs = 0.0, bb = 0.0, rr = 0.0;
    FOR_ROWS(d.n, row) {
        double ri = d.r[row];
        double pi;
        if (d.precond) {
            pi = 0.0;
            pi = pi + ri * d.dinv[row];
        } else {
            pi = ri;
        }
        p[row] = pi;
        rs = rs + ri * pi;
        bb = bb + d.b[row] * d.b[row];
        rr = rr + ri * ri;
    }
    rs = block_reduce<RED_SUM>(rs);
    double out0, out1;
    if (last_block_reduce<RED_SUM>(rs, d.part, d.cnt + 0, &out0)) {
        if (threadIdx.x == 0) d.sc[S_RHO] = out0;
    }
    bb = block_reduce<RED_SUM>(bb);
    rr = block_reduce<RED_SUM>(rr);
    if (last_block_reduce2(bb, rr, d.part + kMaxRedBlocks, d.cnt + 1, &out0, &out1)) {
        if (threadIdx.x == 0) {
            d.sc[S_BB] = out0;
            d.sc[S_RR] = out1;
            double tol2 = tol > 0.0 ? tol * tol * out0 : -1.0;
            d.sc[S_TOL2] = tol2;
            *d.state = (tol > 0.0 && out1 <= tol2) ? ST_CONV : ST_RUN;
        }
    }
}

// ---- FP64 iteration --------------------------------------------------------

template <int W, bool FUSE>
__global__ __launch_bounds__(kBlock) void k_iter_spmv(EllView A, PcgDev d, const double *__restrict__ p_old,
                                                      double *__restrict__ p_new) {
    if (!running(d)) return;
    const double beta = FUSE ? d.sc[S_BETA] : 0.0;
    Operand<double> po{FUSE ? p_old : p_new, d.s, beta, d.n, FUSE};
    double acc = 0.0;
    FOR_ROWS(d.n, row) {
        double pi = po.at(row);
        if (FUSE) p_new[row] = pi;
        double sum;
        if (row_dot<W, double, double>(A, A.val, po, row, sum)) d.Ap[row] = sum;
        acc = acc + pi * d.Ap[row];
    }
    acc = block_reduce<RED_SUM>(acc);
    double dn;
    if (last_block_reduce<RED_SUM>(acc, d.part, d.cnt + 0, &dn)) {
        if (threadIdx.x == 0) {
            double alpha = d.sc[S_RHO] / dn;  // :93
            d.sc[S_DN] = dn;
            d.sc[S_ALPHA] = alpha;
            if (dn == 0.0 || !isfinite(dn) || !isfinite(alpha)) *d.state = ST_BREAK;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_iter_update(PcgDev d, const double *__restrict__ p) {
    if (!running(d)) return;
    const double alpha = d.sc[S_ALPHA];
    const double malpha = -1.0 * alpha;
    double rs = 0.0, rr = 0.0;
    FOR_ROWS(d.n, row) {
        d.x[row] = d.x[row] + alpha * p[row];  // :95
        double ri = d.r[row] + malpha * d.Ap[row];  // :97
        d.r[row] = ri;
        double si;
        if (d.precond) {  // :99-105
            si = 0.0;
            si = si + ri * d.dinv[row];
        } else {
            si = ri;
        }
        d.s[row] = si;
        rs = rs + ri * si;
        rr = rr + ri * ri;
    }
    rs = block_reduce<RED_SUM>(rs);
    rr = block_reduce<RED_SUM>(rr);
    double r1, rrt;
    if (last_block_reduce2(rs, rr, d.part, d.cnt + 1, &r1, &rrt)) {
        if (threadIdx.x == 0) {
            double r0 = d.sc[S_RHO];  // :107-111
            d.sc[S_RHO] = r1;
            d.sc[S_BETA] = r1 / r0;
            d.sc[S_RR] = rrt;
            *d.iters += 1;
            double tol2 = d.sc[S_TOL2];
            if (tol2 >= 0.0 && rrt <= tol2) *d.state = ST_CONV;
        }
    }
}

// ---- MIXED iteration (FP32 vectors and operator, FP64 dot accumulation) --------

template <int W, bool FUSE>
__global__ __launch_bounds__(kBlock) void k_m_spmv(EllView A, PcgDev d, const float *__restrict__ p_old,
                                                   float *__restrict__ p_new) {
    if (!running(d)) return;
    const float beta = FUSE ? (float)d.sc[S_BETA] : 0.0f;
    Operand<float> po{FUSE ? p_old : p_new, d.s32, beta, d.n, FUSE};
    double acc = 0.0;
    FOR_ROWS(d.n, row) {
        float pi = po.at(row);
        if (FUSE) p_new[row] = pi;
        float sum;
        row_dot<W, float, float>(A, A.val32, po, row, sum);
        d.Ap32[row] = sum;
        acc = acc + (double)pi * (double)sum;
    }
    acc = block_reduce<RED_SUM>(acc);
    double dn;
    if (last_block_reduce<RED_SUM>(acc, d.part, d.cnt + 0, &dn)) {
        if (threadIdx.x == 0) {
            double alpha = d.sc[S_RHO] / dn;
            d.sc[S_ALPHA] = alpha;
            if (dn == 0.0 || !isfinite(dn) || !isfinite(alpha)) *d.state = ST_BREAK;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_m_update(PcgDev d, const float *__restrict__ p) {
    if (!running(d)) return;
    const float a32 = (float)d.sc[S_ALPHA];
    double rs = 0.0, rr = 0.0;
    FOR_ROWS(d.n, row) {
        d.xc32[row] = d.xc32[row] + a32 * p[row];
        float ri = d.r32[row] - a32 * d.Ap32[row];
        d.r32[row] = ri;
        float si = d.precond ? ri * d.dinv32[row] : ri;
        d.s32[row] = si;
        rs = rs + (double)ri * (double)si;
        rr = rr + (double)ri * (double)ri;
    }
    rs = block_reduce<RED_SUM>(rs);
    rr = block_reduce<RED_SUM>(rr);
    double r1, rrt;
    if (last_block_reduce2(rs, rr, d.part, d.cnt + 1, &r1, &rrt)) {
        if (threadIdx.x == 0) {
            double r0 = d.sc[S_RHO];
            d.sc[S_RHO_PREV] = r0;
            d.sc[S_RHO] = r1;
            d.sc[S_BETA] = r1 / r0;
            d.sc[S_RR] = rrt;
            *d.iters += 1;
        }
    }
}

// x += xc, xc = 0
__global__ __launch_bounds__(kBlock) void k_m_fold(PcgDev d) {
    FOR_ROWS(d.n, row) {
        d.x[row] = d.x[row] + (double)d.xc32[row];
        d.xc32[row] = 0.0f;
    }
}

// Defect correction: r = b - A x in FP64 (empty rows: x), r32 = r,
// s32 = Dinv r32, rho = r32.s32; the direction is kept and the next beta
// uses the replaced residual; convergence on the true ||r||.
template <int W>
__global__ __launch_bounds__(kBlock) void k_m_resid(EllView A, PcgDev d, int first) {
    double rs = 0.0, rr = 0.0;
    Operand<double> xo{d.x, nullptr, 0.0, d.n, false};
    FOR_ROWS(d.n, row) {
        double sum;
        bool any = row_dot<W, double, double>(A, A.val, xo, row, sum);
        double t = any ? sum : d.x[row];
        double ri = d.b[row] + -1.0 * t;
        d.r[row] = ri;
        float r32 = (float)ri;
        d.r32[row] = r32;
        float s32 = d.precond ? r32 * d.dinv32[row] : r32;
        d.s32[row] = s32;
        rs = rs + (double)r32 * (double)s32;
        rr = rr + ri * ri;
    }
    rs = block_reduce<RED_SUM>(rs);
    rr = block_reduce<RED_SUM>(rr);
    double r1, rrt;
    if (last_block_reduce2(rs, rr, d.part, d.cnt + 1, &r1, &rrt)) {
        if (threadIdx.x == 0) {
            // keep the direction (reliable update) unless the FP32 recursion
            // had drifted from the true residual by more than 10x: then restart
            if (!first) d.sc[S_BETA] = rrt > 100.0 * d.sc[S_RR] ? 0.0 : r1 / d.sc[S_RHO_PREV];
            d.sc[S_RHO] = r1;
            d.sc[S_TRUE_RR] = rrt;
            d.sc[S_RR] = rrt;
            double tol2 = d.sc[S_TOL2];
            if (tol2 >= 0.0 && rrt <= tol2) *d.state = ST_CONV;
        }
    }
}

// first direction p32 = s32, correction zero
__global__ __launch_bounds__(kBlock) void k_m_start(PcgDev d, float *__restrict__ p) {
    FOR_ROWS(d.n, row) {
        p[row] = d.s32[row];
        d.xc32[row] = 0.0f;
    }
}

__global__ __launch_bounds__(kBlock) void k_to32(int n, const double *__restrict__ a, float *__restrict__ b) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

}  // namespace gip

using namespace gip;

struct gip_pcg {
    gip_ctx *ctx;
    const gip_mat *A;
    int n, n_all, grid;
    double *r, *p[2], *s, *Ap, *dinv;
    float *r32, *p32[2], *s32, *Ap32, *xc32, *dinv32;
    double *sc, *part;
    unsigned int *cnt;
    int *state;  // state, iters
    int *h_state;  // pinned copy
    double *h_sc;  // pinned scalars
};

static int pcg_grid(int n) {
    int nchunk = (n + kBlock - 1) / kBlock;
    int g = ((nchunk + 7) / 8) * 8;
    if (g > kMaxRedBlocks) g = kMaxRedBlocks;
    if (g < 8) g = 8;
    return g;
}

// Launch KERNEL<W, ...> with W the operator's uniform ELL width (5..7) or 0.
#define DISPATCH_W(A, KERNEL, TAIL, ...)                                             \
    do {                                                                             \
        switch ((A)->format == GIP_FMT_ELL ? (A)->uniform_width : -1) {              \
            case 5: KERNEL<5 TAIL><<<S->grid, kBlock, 0, st>>>(__VA_ARGS__); break;  \
            case 6: KERNEL<6 TAIL><<<S->grid, kBlock, 0, st>>>(__VA_ARGS__); break;  \
            case 7: KERNEL<7 TAIL><<<S->grid, kBlock, 0, st>>>(__VA_ARGS__); break;  \
            default: KERNEL<0 TAIL><<<S->grid, kBlock, 0, st>>>(__VA_ARGS__); break; \
        }                                                                            \
    } while (0)
#define NOTAIL
#define FUSED_TAIL , true
#define FIRST_TAIL , false

static PcgDev make_dev(gip_pcg *S, const double *b, double *x, int precond) {
    PcgDev d;
    d.n = S->n;
    d.x = x;
    d.r = S->r;
    d.s = S->s;
    d.Ap = S->Ap;
    d.b = b;
    d.dinv = S->dinv;
    d.r32 = S->r32;
    d.s32 = S->s32;
    d.Ap32 = S->Ap32;
    d.xc32 = S->xc32;
    d.dinv32 = S->dinv32;
    d.sc = S->sc;
    d.state = S->state;
    d.iters = S->state + 1;
    d.part = S->part;
    d.cnt = S->cnt;
    d.precond = precond;
    return d;
}

extern "C" {

int gip_pcg_create(gip_ctx *c, const gip_mat *A, int32_t n_all, gip_pcg **out) {
    if (!c || !A || !out || n_all < A->nrows) return set_error(GIP_EINVAL, "gip_pcg_create: bad argument");
    if (n_all < A->cols_needed) return set_error(GIP_EINVAL, "gip_pcg_create: vectors shorter than the operator's columns");
    if (A->format != GIP_FMT_ELL && A->format != GIP_FMT_CSR)
        return set_error(GIP_EINVAL, "gip_pcg_create: operator must be ELL or CSR");
    gip_pcg *S = new (std::nothrow) gip_pcg();
    if (!S) return set_error(GIP_ENOMEM, "gip_pcg_create: allocation failed");
    S->ctx = c;
    S->A = A;
    S->n = A->nrows;
    S->n_all = n_all;
    S->grid = pcg_grid(A->nrows);
    size_t n = (size_t)std::max(1, A->nrows);
    size_t na = (size_t)std::max(1, n_all);  // directions are SpMV inputs: ghost columns included
    hipError_t e = hipSetDevice(c->device);
    auto dalloc = [&](void **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes);
    };
    dalloc((void **)&S->r, n * 8);
    dalloc((void **)&S->p[0], na * 8);
    dalloc((void **)&S->p[1], na * 8);
    dalloc((void **)&S->s, n * 8);
    dalloc((void **)&S->Ap, n * 8);
    dalloc((void **)&S->dinv, n * 8);
    dalloc((void **)&S->sc, S_COUNT * 8);
    dalloc((void **)&S->part, 4 * kMaxRedBlocks * 8);
    dalloc((void **)&S->cnt, 16 * sizeof(unsigned int));
    dalloc((void **)&S->state, 2 * sizeof(int));
    if (e == hipSuccess) e = hipHostMalloc((void **)&S->h_state, 2 * sizeof(int), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void **)&S->h_sc, S_COUNT * 8, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMemsetAsync(S->cnt, 0, 16 * sizeof(unsigned int), c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(S->sc, 0, S_COUNT * 8, c->stream);
    if (e != hipSuccess) {
        int rc = hip_fail(e, "gip_pcg_create");
        gip_pcg_destroy(S);
        return rc;
    }
    int rc = A->diag_missing ? GIP_OK : gip_mat_diag_inv(A, S->dinv);
    if (!rc) {
        e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hip_fail(e, "gip_pcg_create sync");
    }
    if (rc) {
        gip_pcg_destroy(S);
        return rc;
    }
    *out = S;
    return GIP_OK;
}

int gip_pcg_destroy(gip_pcg *S) {
    if (!S) return GIP_OK;
    hipSetDevice(S->ctx->device);
    hipStreamSynchronize(S->ctx->stream);
    void *ptrs[] = {S->r,     S->p[0],   S->p[1], S->s,    S->Ap,   S->dinv, S->r32, S->p32[0], S->p32[1],
                    S->s32,   S->Ap32,   S->xc32, S->dinv32, S->sc, S->part, S->cnt, S->state};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (S->h_state) hipHostFree(S->h_state);
    if (S->h_sc) hipHostFree(S->h_sc);
    delete S;
    return GIP_OK;
}

static int ensure_mixed(gip_pcg *S) {
    if (S->r32) return GIP_OK;
    const gip_mat *A = S->A;
    if ((A->format == GIP_FMT_ELL && !A->d_ell_val32) || (A->format == GIP_FMT_CSR && !A->d_val32))
        return set_error(GIP_EINVAL, "mixed precision needs an operator created with GIP_MAT_FP32");
    size_t n = (size_t)std::max(1, S->n), na = (size_t)std::max(1, S->n_all);
    float **bufs[] = {&S->r32, &S->s32, &S->Ap32, &S->xc32, &S->dinv32};
    for (float **b : bufs) GIP_HIP_TRY(hipMalloc(b, n * 4));
    // the FP32 directions are SpMV inputs: ghost columns included, and zero
    // there (the correction leaves the ghost values of x unchanged)
    for (int k = 0; k < 2; ++k) {
        GIP_HIP_TRY(hipMalloc(&S->p32[k], na * 4));
        GIP_HIP_TRY(hipMemsetAsync(S->p32[k], 0, na * 4, S->ctx->stream));
    }
    if (S->n) {
        k_to32<<<(S->n + kBlock - 1) / kBlock, kBlock, 0, S->ctx->stream>>>(S->n, S->dinv, S->dinv32);
        GIP_LAUNCH_CHECK("dinv32");
    }
    return GIP_OK;
}

static int poll(gip_pcg *S, bool scalars) {
    hipStream_t st = S->ctx->stream;
    GIP_HIP_TRY(hipMemcpyAsync(S->h_state, S->state, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    if (scalars) GIP_HIP_TRY(hipMemcpyAsync(S->h_sc, S->sc, S_COUNT * 8, hipMemcpyDeviceToHost, st));
    GIP_HIP_TRY(hipStreamSynchronize(st));
    return GIP_OK;
}

int gip_pcg_solve(gip_pcg *S, const double *b, double *x, const double *x0, const gip_pcg_opts *o, gip_pcg_stats *stats) {
    if (!S || !b || !x || !x0 || !o || o->maxit < 0) return set_error(GIP_EINVAL, "gip_pcg_solve: bad argument");
    if (o->precond && S->A->diag_missing) return set_error(GIP_EINVAL, "gip_pcg_solve: Jacobi needs every diagonal entry");
    gip_ctx *c = S->ctx;
    hipStream_t st = c->stream;
    const gip_mat *A = S->A;
    const bool mixed = o->precision == GIP_PCG_MIXED;
    if (mixed) {
        int rc = ensure_mixed(S);
        if (rc) return rc;
    }
    PcgDev d = make_dev(S, b, x, o->precond ? 1 : 0);
    EllView v = view(A);
    const int pollk = o->poll > 0 ? o->poll : 16;
    GIP_HIP_TRY(hipMemsetAsync(S->state, 0, 2 * sizeof(int), st));
    // setup, GIP_LinearSolver.h:55-80
    DISPATCH_W(A, k_setup_resid, NOTAIL, v, d, x0, S->n_all, S->p[0], S->p[1]);
    k_setup_dir<<<S->grid, kBlock, 0, st>>>(d, S->p[0], o->tol);
    GIP_LAUNCH_CHECK("pcg setup");
    int corrections = 0, rc;
    if ((rc = poll(S, true))) return rc;
    if (!mixed) {
        int k = 0;  // iterations queued
        while (k < o->maxit && S->h_state[0] == ST_RUN) {
            int chunk = std::min(pollk, o->maxit - k);
            for (int j = 0; j < chunk; ++j, ++k) {
                double *pn = S->p[k & 1];
                if (k == 0)
                    DISPATCH_W(A, k_iter_spmv, FIRST_TAIL, v, d, pn, pn);
                else
                    DISPATCH_W(A, k_iter_spmv, FUSED_TAIL, v, d, S->p[(k - 1) & 1], pn);
                k_iter_update<<<S->grid, kBlock, 0, st>>>(d, pn);
            }
            GIP_LAUNCH_CHECK("pcg iteration");
            if ((rc = poll(S, false))) return rc;
        }
    } else {
        // FP32 residual of x0 and the first direction
        DISPATCH_W(A, k_m_resid, NOTAIL, v, d, 1);
        k_m_start<<<S->grid, kBlock, 0, st>>>(d, S->p32[0]);
        GIP_LAUNCH_CHECK("mixed setup");
        if ((rc = poll(S, true))) return rc;
        double last_rr = S->h_sc[S_TRUE_RR];
        const double it = o->inner_tol > 0.0 ? o->inner_tol : 0.1;
        const double inner2 = it * it;
        int k = 0;
        while (k < o->maxit && S->h_state[0] == ST_RUN) {
            int chunk = std::min(pollk, o->maxit - k);
            for (int j = 0; j < chunk; ++j, ++k) {
                float *pn = S->p32[k & 1];
                if (k == 0)
                    DISPATCH_W(A, k_m_spmv, FIRST_TAIL, v, d, pn, pn);
                else
                    DISPATCH_W(A, k_m_spmv, FUSED_TAIL, v, d, S->p32[(k - 1) & 1], pn);
                k_m_update<<<S->grid, kBlock, 0, st>>>(d, pn);
            }
            GIP_LAUNCH_CHECK("mixed iteration");
            if ((rc = poll(S, true))) return rc;
            if (S->h_state[0] != ST_RUN) break;
            const double rr32 = S->h_sc[S_RR], tol2 = S->h_sc[S_TOL2];
            const bool near_target = tol2 >= 0.0 && rr32 <= tol2;
            if (rr32 <= inner2 * last_rr || near_target || k >= o->maxit) {
                k_m_fold<<<S->grid, kBlock, 0, st>>>(d);
                DISPATCH_W(A, k_m_resid, NOTAIL, v, d, 0);
                GIP_LAUNCH_CHECK("mixed correction");
                ++corrections;
                if ((rc = poll(S, true))) return rc;
                last_rr = S->h_sc[S_TRUE_RR];
            }
        }
    }
    if ((rc = poll(S, true))) return rc;
    int status = GIP_OK;
    if (S->h_state[0] == ST_BREAK)
        status = GIP_EBREAKDOWN;
    else if (o->tol > 0.0 && S->h_state[0] != ST_CONV)
        status = GIP_ENOCONV;
    if (stats) {
        stats->iters = S->h_state[1];
        stats->corrections = corrections;
        const double bb = S->h_sc[S_BB], rr = S->h_sc[S_RR];
        stats->relres = bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
        stats->status = status;
    }
    return status == GIP_EBREAKDOWN ? set_error(GIP_EBREAKDOWN, "PCG breakdown (p.Ap == 0 or non-finite step)") : GIP_OK;
}

}  // extern "C"
