// Fused, device-resident Jacobi-PCG of libgip_hip (GIP_LinearSolver.h:45-117).
//
// FP64 iteration (the reference recurrences, same operation order):
//   K1  Ap = A p ; dn = p.Ap                  -> alpha = r1/dn        (last workgroup)
//   K2  x += alpha p ; r += (-alpha) Ap ; s = Dinv r ; r.s, r.r -> beta = r1'/r1, convergence
//   K3  p = 1.0 s + beta p
// All scalars live in device memory; every kernel first reads the solver's
// state word and returns at once when the solve has finished, so the host
// can queue many iterations and poll the state only now and then.
//
// MIXED: the same three kernels in FP32 (operator values, vectors, SpMV;
// dots accumulate in FP64) on the correction equation, inside an FP64
// defect-correction loop: when the FP32 residual has dropped by inner_tol
// (or reaches the target) the FP32 correction is added to the FP64 x, the
// FP64 true residual b - A x replaces the FP32 residual (the search
// direction is kept), and convergence is decided on that FP64 residual.
#include <cmath>
#include <new>

#include "gip_blas1.h"
#include "gip_internal.h"
#include "gip_spmv.h"

namespace gip {

// Device scalar slots of a solver.
enum {
    S_RHO = 0,    // r.s of the current iterate
    S_ALPHA = 1,
    S_BETA = 2,
    S_RR = 3,     // r.r
    S_BB = 4,     // b.b
    S_DN = 5,
    S_TOL2 = 6,   // (tol ||b||)^2, 0 when no tolerance
    S_TRUE_RR = 7,
    S_COUNT = 8
};
// State word: 0 running, 1 converged, 2 breakdown.
enum { ST_RUN = 0, ST_CONV = 1, ST_BREAK = 2 };

struct EllView {
    int fmt;  // GIP_FMT_ELL / GIP_FMT_CSR / GIP_FMT_DIAG
    const int32_t *soff, *col, *rp;
    const double *val;
    const float *val32;
    const double *diag;
};

static EllView view(const gip_mat *A) {
    EllView v;
    v.fmt = A->format;
    v.soff = A->d_slice_off;
    v.col = A->format == GIP_FMT_ELL ? A->d_ell_col : A->d_col;
    v.rp = A->d_rowptr;
    v.val = A->format == GIP_FMT_ELL ? A->d_ell_val : A->d_val;
    v.val32 = A->format == GIP_FMT_ELL ? A->d_ell_val32 : A->d_val32;
    v.diag = A->d_diag;
    return v;
}

// Row sum in stored order; returns false for an empty row.
template <int W, class T, class V>
__device__ inline bool row_dot(const EllView &A, const V *__restrict__ val, const T *__restrict__ x, int row, T &sum) {
    sum = T(0);
    if (W > 0) {
        const int base = A.soff[row >> 6] + (row & 63);
        int c[W > 0 ? W : 1];
        V v[W > 0 ? W : 1];
#pragma unroll
        for (int k = 0; k < W; ++k) {
            c[k] = __builtin_nontemporal_load(&A.col[base + 64 * k]);
            v[k] = __builtin_nontemporal_load(&val[base + 64 * k]);
        }
#pragma unroll
        for (int k = 0; k < W; ++k) sum = sum + x[c[k]] * v[k];
        return true;
    } else if (A.fmt == GIP_FMT_ELL) {
        const int s = row >> 6, base = A.soff[s] + (row & 63), width = (A.soff[s + 1] - A.soff[s]) >> 6;
        bool any = false;
        for (int k = 0; k < width; ++k) {
            const int cc = A.col[base + 64 * k];
            if (cc < 0) break;
            sum = sum + x[cc] * val[base + 64 * k];
            any = true;
        }
        return any;
    } else {
        const int jb = A.rp[row], je = A.rp[row + 1];
        for (int j = jb; j < je; ++j) sum = sum + x[A.col[j]] * val[j];
        return jb != je;
    }
}

// Rows of workgroup b: each XCD (b % 8) sweeps a contiguous range of
// 256-row chunks, its workgroups taking the chunks in turn.
struct RowSweep {
    int c, c1, step;
    __device__ RowSweep(int n) {
        const int nchunk = (n + kBlock - 1) / kBlock;
        const int x = blockIdx.x & 7, local = blockIdx.x >> 3;
        step = gridDim.x >> 3;
        c = (int)((int64_t)nchunk * x / 8) + local;
        c1 = (int)((int64_t)nchunk * (x + 1) / 8);
    }
};

#define FOR_ROWS(n, row)                                                   \
    for (RowSweep rs__(n); rs__.c < rs__.c1; rs__.c += rs__.step)          \
        for (int row = rs__.c * kBlock + (int)threadIdx.x; row < (rs__.c + 1) * kBlock && row < (n); row = (n))

struct PcgDev {
    int n;
    double *x, *r, *p, *s, *Ap;
    const double *b, *dinv;
    float *r32, *p32, *s32, *Ap32, *xc32;
    const float *dinv32;
    double *sc;      // S_COUNT scalars
    int *state;
    double *part0, *part1;
    unsigned int *cnt;  // 4 counters
    int precond;
};

// ---- setup ---------------------------------------------------------------

// x = x0 (all), Ap = (A x)_i or x0_i on empty rows, r = b + (-1.0) Ap  (:65-70)
template <int W>
__global__ __launch_bounds__(kBlock) void k_setup_resid(EllView A, PcgDev d, const double *__restrict__ x0, int n_all) {
    const int stride = gridDim.x * kBlock;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n_all; i += stride) {
        d.x[i] = x0[i];
        if (i >= d.n) d.p[i] = x0[i];  // the direction keeps x0 outside the solved rows (:55-56)
    }
    // A reads x0 directly so no grid-wide ordering with the copy is needed
    FOR_ROWS(d.n, row) {
        double sum;
        bool any = row_dot<W, double, double>(A, A.val, x0, row, sum);
        double t = any ? sum : x0[row];
        d.Ap[row] = t;
        d.r[row] = d.b[row] + -1.0 * t;
    }
}

// p = Dinv r (or r); r1 = r.p, b.b, r.r  (:72-80)
__global__ __launch_bounds__(kBlock) void k_setup_dir(PcgDev d) {
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
        d.p[row] = pi;
        rs = rs + ri * pi;
        bb = bb + d.b[row] * d.b[row];
        rr = rr + ri * ri;
    }
    rs = block_reduce<RED_SUM>(rs);
    double out;
    if (last_block_reduce<RED_SUM>(rs, d.part0, d.cnt + 0, &out)) {
        if (threadIdx.x == 0) d.sc[S_RHO] = out;
    }
    bb = block_reduce<RED_SUM>(bb);
    if (last_block_reduce<RED_SUM>(bb, d.part1, d.cnt + 1, &out)) {
        if (threadIdx.x == 0) d.sc[S_BB] = out;
    }
    rr = block_reduce<RED_SUM>(rr);
    if (last_block_reduce<RED_SUM>(rr, d.part0 + kMaxRedBlocks, d.cnt + 2, &out)) {
        if (threadIdx.x == 0) d.sc[S_RR] = out;
    }
}

// state after setup: converged when ||r|| <= tol ||b||
__global__ void k_setup_state(PcgDev d, double tol) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double tol2 = tol > 0.0 ? tol * tol * d.sc[S_BB] : -1.0;
        d.sc[S_TOL2] = tol2;
        *d.state = (tol > 0.0 && d.sc[S_RR] <= tol2) ? ST_CONV : ST_RUN;
    }
}

// ---- FP64 iteration --------------------------------------------------------

template <int W>
__global__ __launch_bounds__(kBlock) void k_iter_spmv(EllView A, PcgDev d) {
    if (*(volatile int *)d.state != ST_RUN) return;
    double acc = 0.0;
    FOR_ROWS(d.n, row) {
        double sum;
        if (row_dot<W, double, double>(A, A.val, d.p, row, sum)) d.Ap[row] = sum;
        acc = acc + d.p[row] * d.Ap[row];
    }
    acc = block_reduce<RED_SUM>(acc);
    double dn;
    if (last_block_reduce<RED_SUM>(acc, d.part0, d.cnt + 0, &dn)) {
        if (threadIdx.x == 0) {
            double alpha = d.sc[S_RHO] / dn;
            d.sc[S_DN] = dn;
            d.sc[S_ALPHA] = alpha;
            if (dn == 0.0 || !isfinite(alpha)) *d.state = ST_BREAK;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_iter_update(PcgDev d) {
    if (*(volatile int *)d.state != ST_RUN) return;
    const double alpha = d.sc[S_ALPHA];
    if (!isfinite(alpha)) return;  // breakdown flagged by the previous kernel's last workgroup
    const double malpha = -1.0 * alpha;
    double rs = 0.0, rr = 0.0;
    FOR_ROWS(d.n, row) {
        d.x[row] = d.x[row] + alpha * d.p[row];
        double ri = d.r[row] + malpha * d.Ap[row];
        d.r[row] = ri;
        double si;
        if (d.precond) {
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
    double out;
    if (last_block_reduce<RED_SUM>(rs, d.part0, d.cnt + 0, &out)) {
        if (threadIdx.x == 0) {
            double r0 = d.sc[S_RHO];
            d.sc[S_RHO] = out;
            d.sc[S_BETA] = out / r0;
        }
    }
    rr = block_reduce<RED_SUM>(rr);
    if (last_block_reduce<RED_SUM>(rr, d.part1, d.cnt + 1, &out)) {
        if (threadIdx.x == 0) d.sc[S_RR] = out;
    }
}

// p = 1.0 s + beta p; the convergence decision of the iteration is taken
// here (after both reductions of k_iter_update have landed).
__global__ __launch_bounds__(kBlock) void k_iter_dir(PcgDev d, int *iters) {
    if (*(volatile int *)d.state != ST_RUN) return;
    const double beta = d.sc[S_BETA];
    FOR_ROWS(d.n, row) d.p[row] = 1.0 * d.s[row] + beta * d.p[row];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *iters += 1;
        double tol2 = d.sc[S_TOL2];
        if (tol2 >= 0.0 && d.sc[S_RR] <= tol2) *d.state = ST_CONV;
    }
}

// ---- MIXED iteration (FP32 vectors and operator, FP64 dot accumulation) --------

template <int W>
__global__ __launch_bounds__(kBlock) void k_m_spmv(EllView A, PcgDev d) {
    if (*(volatile int *)d.state != ST_RUN) return;
    double acc = 0.0;
    FOR_ROWS(d.n, row) {
        float sum;
        row_dot<W, float, float>(A, A.val32, d.p32, row, sum);
        d.Ap32[row] = sum;
        acc = acc + (double)d.p32[row] * (double)sum;
    }
    acc = block_reduce<RED_SUM>(acc);
    double dn;
    if (last_block_reduce<RED_SUM>(acc, d.part0, d.cnt + 0, &dn)) {
        if (threadIdx.x == 0) {
            double alpha = d.sc[S_RHO] / dn;
            d.sc[S_ALPHA] = alpha;
            if (dn == 0.0 || !isfinite(alpha)) *d.state = ST_BREAK;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_m_update(PcgDev d) {
    if (*(volatile int *)d.state != ST_RUN) return;
    const double alpha = d.sc[S_ALPHA];
    if (!isfinite(alpha)) return;
    const float a32 = (float)alpha;
    double rs = 0.0, rr = 0.0;
    FOR_ROWS(d.n, row) {
        d.xc32[row] = d.xc32[row] + a32 * d.p32[row];
        float ri = d.r32[row] - a32 * d.Ap32[row];
        d.r32[row] = ri;
        float si = d.precond ? ri * d.dinv32[row] : ri;
        d.s32[row] = si;
        rs = rs + (double)ri * (double)si;
        rr = rr + (double)ri * (double)ri;
    }
    rs = block_reduce<RED_SUM>(rs);
    double out;
    if (last_block_reduce<RED_SUM>(rs, d.part0, d.cnt + 0, &out)) {
        if (threadIdx.x == 0) {
            double r0 = d.sc[S_RHO];
            d.sc[S_RHO] = out;
            d.sc[S_BETA] = out / r0;
        }
    }
    rr = block_reduce<RED_SUM>(rr);
    if (last_block_reduce<RED_SUM>(rr, d.part1, d.cnt + 1, &out)) {
        if (threadIdx.x == 0) d.sc[S_RR] = out;
    }
}

__global__ __launch_bounds__(kBlock) void k_m_dir(PcgDev d, int *iters) {
    if (*(volatile int *)d.state != ST_RUN) return;
    const float beta = (float)d.sc[S_BETA];
    FOR_ROWS(d.n, row) d.p32[row] = d.s32[row] + beta * d.p32[row];
    if (blockIdx.x == 0 && threadIdx.x == 0) *iters += 1;
}

// Defect correction: x += xc, xc = 0, r = b - A x in FP64 (empty rows: x),
// r32 = r, s32 = Dinv r32, rho = r32.s32, true ||r||^2.
template <int W>
__global__ __launch_bounds__(kBlock) void k_m_fold(PcgDev d) {
    FOR_ROWS(d.n, row) {
        d.x[row] = d.x[row] + (double)d.xc32[row];
        d.xc32[row] = 0.0f;
    }
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_m_resid(EllView A, PcgDev d) {
    double rs = 0.0, rr = 0.0;
    FOR_ROWS(d.n, row) {
        double sum;
        bool any = row_dot<W, double, double>(A, A.val, d.x, row, sum);
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
    double out;
    if (last_block_reduce<RED_SUM>(rs, d.part0, d.cnt + 0, &out)) {
        if (threadIdx.x == 0) d.sc[S_RHO] = out;
    }
    rr = block_reduce<RED_SUM>(rr);
    if (last_block_reduce<RED_SUM>(rr, d.part1, d.cnt + 1, &out)) {
        if (threadIdx.x == 0) {
            d.sc[S_TRUE_RR] = out;
            d.sc[S_RR] = out;
            double tol2 = d.sc[S_TOL2];
            if (tol2 >= 0.0 && out <= tol2) *d.state = ST_CONV;
        }
    }
}

// p32 = s32 (first direction) and zero the correction
__global__ __launch_bounds__(kBlock) void k_m_start(PcgDev d) {
    FOR_ROWS(d.n, row) {
        d.p32[row] = d.s32[row];
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
    double *r, *p, *s, *Ap, *dinv;
    float *r32, *p32, *s32, *Ap32, *xc32, *dinv32;
    double *sc, *part;
    unsigned int *cnt;
    int *state, *iters;
    int *h_state;  // pinned: state, iters
};

static int pcg_grid(int n) {
    int nchunk = (n + kBlock - 1) / kBlock;
    int g = ((nchunk + 7) / 8) * 8;
    if (g > kMaxRedBlocks) g = kMaxRedBlocks;
    if (g < 8) g = 8;
    return g;
}

#define DISPATCH_W(A, KERNEL, ...)                                                    \
    do {                                                                              \
        switch ((A)->format == GIP_FMT_ELL ? (A)->uniform_width : -1) {               \
            case 5: KERNEL<5><<<S->grid, kBlock, 0, st>>>(__VA_ARGS__); break;        \
            case 6: KERNEL<6><<<S->grid, kBlock, 0, st>>>(__VA_ARGS__); break;        \
            case 7: KERNEL<7><<<S->grid, kBlock, 0, st>>>(__VA_ARGS__); break;        \
            default: KERNEL<0><<<S->grid, kBlock, 0, st>>>(__VA_ARGS__); break;       \
        }                                                                             \
    } while (0)

static PcgDev make_dev(gip_pcg *S, const double *b, double *x) {
    PcgDev d;
    d.n = S->n;
    d.x = x;
    d.r = S->r;
    d.p = S->p;
    d.s = S->s;
    d.Ap = S->Ap;
    d.b = b;
    d.dinv = S->dinv;
    d.r32 = S->r32;
    d.p32 = S->p32;
    d.s32 = S->s32;
    d.Ap32 = S->Ap32;
    d.xc32 = S->xc32;
    d.dinv32 = S->dinv32;
    d.sc = S->sc;
    d.state = S->state;
    d.part0 = S->part;
    d.part1 = S->part + 2 * kMaxRedBlocks;
    d.cnt = S->cnt;
    d.precond = 1;
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
    size_t na = (size_t)std::max(1, n_all);  // SpMV inputs also cover the ghost columns
    hipError_t e = hipSetDevice(c->device);
    auto dalloc = [&](void **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes);
    };
    dalloc((void **)&S->r, n * 8);
    dalloc((void **)&S->p, na * 8);
    dalloc((void **)&S->s, n * 8);
    dalloc((void **)&S->Ap, n * 8);
    dalloc((void **)&S->dinv, n * 8);
    dalloc((void **)&S->sc, S_COUNT * 8);
    dalloc((void **)&S->part, 4 * kMaxRedBlocks * 8);
    dalloc((void **)&S->cnt, 16 * sizeof(unsigned int));
    dalloc((void **)&S->state, 2 * sizeof(int));
    if (e == hipSuccess) e = hipHostMalloc((void **)&S->h_state, 2 * sizeof(int), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMemsetAsync(S->cnt, 0, 16 * sizeof(unsigned int), c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(S->sc, 0, S_COUNT * 8, c->stream);
    if (e != hipSuccess) {
        int rc = hip_fail(e, "gip_pcg_create");
        gip_pcg_destroy(S);
        return rc;
    }
    S->iters = S->state + 1;
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
    void *ptrs[] = {S->r, S->p, S->s, S->Ap, S->dinv, S->r32, S->p32, S->s32, S->Ap32, S->xc32, S->dinv32,
                    S->sc, S->part, S->cnt, S->state};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (S->h_state) hipHostFree(S->h_state);
    delete S;
    return GIP_OK;
}

static int ensure_mixed(gip_pcg *S) {
    if (S->r32) return GIP_OK;
    const gip_mat *A = S->A;
    if ((A->format == GIP_FMT_ELL && !A->d_ell_val32) || (A->format == GIP_FMT_CSR && !A->d_val32))
        return set_error(GIP_EINVAL, "mixed precision needs an operator created with GIP_MAT_FP32");
    size_t n = (size_t)std::max(1, S->n);
    float **bufs[] = {&S->r32, &S->s32, &S->Ap32, &S->xc32, &S->dinv32};
    for (float **b : bufs) GIP_HIP_TRY(hipMalloc(b, n * 4));
    // the FP32 direction is an SpMV input: ghost columns included, and zero there
    // (the correction leaves the ghost values of x unchanged)
    GIP_HIP_TRY(hipMalloc(&S->p32, (size_t)std::max(1, S->n_all) * 4));
    GIP_HIP_TRY(hipMemsetAsync(S->p32, 0, (size_t)std::max(1, S->n_all) * 4, S->ctx->stream));
    if (S->n) {
        k_to32<<<(S->n + kBlock - 1) / kBlock, kBlock, 0, S->ctx->stream>>>(S->n, S->dinv, S->dinv32);
        GIP_LAUNCH_CHECK("dinv32");
    }
    return GIP_OK;
}

static int poll_state(gip_pcg *S) {
    GIP_HIP_TRY(hipMemcpyAsync(S->h_state, S->state, 2 * sizeof(int), hipMemcpyDeviceToHost, S->ctx->stream));
    GIP_HIP_TRY(hipStreamSynchronize(S->ctx->stream));
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
    PcgDev d = make_dev(S, b, x);
    d.precond = o->precond ? 1 : 0;
    EllView v = view(A);
    const int poll = o->poll > 0 ? o->poll : 16;
    GIP_HIP_TRY(hipMemsetAsync(S->state, 0, 2 * sizeof(int), st));
    // setup, :55-80
    DISPATCH_W(A, k_setup_resid, v, d, x0, S->n_all);
    GIP_LAUNCH_CHECK("pcg setup residual");
    k_setup_dir<<<S->grid, kBlock, 0, st>>>(d);
    k_setup_state<<<1, 64, 0, st>>>(d, o->tol);
    GIP_LAUNCH_CHECK("pcg setup direction");
    int corrections = 0;
    double last_rr = 0.0;
    int rc = poll_state(S);
    if (rc) return rc;
    if (!mixed) {
        int done = 0;
        while (done < o->maxit && S->h_state[0] == ST_RUN) {
            int chunk = std::min(poll, o->maxit - done);
            for (int k = 0; k < chunk; ++k) {
                DISPATCH_W(A, k_iter_spmv, v, d);
                k_iter_update<<<S->grid, kBlock, 0, st>>>(d);
                k_iter_dir<<<S->grid, kBlock, 0, st>>>(d, S->iters);
            }
            GIP_LAUNCH_CHECK("pcg iteration");
            done += chunk;
            if ((rc = poll_state(S))) return rc;
        }
    } else {
        // r32 = r, s32 = Dinv r32, rho = r32.s32 from the FP64 residual of x0
        DISPATCH_W(A, k_m_resid, v, d);
        k_m_start<<<S->grid, kBlock, 0, st>>>(d);
        GIP_LAUNCH_CHECK("mixed setup");
        double h_sc[S_COUNT];
        GIP_HIP_TRY(hipMemcpyAsync(h_sc, S->sc, sizeof h_sc, hipMemcpyDeviceToHost, st));
        if ((rc = poll_state(S))) return rc;
        last_rr = h_sc[S_TRUE_RR];
        const double inner2 = (o->inner_tol > 0.0 ? o->inner_tol : 1e-3) * (o->inner_tol > 0.0 ? o->inner_tol : 1e-3);
        int done = 0;
        while (done < o->maxit && S->h_state[0] == ST_RUN) {
            int chunk = std::min(poll, o->maxit - done);
            for (int k = 0; k < chunk; ++k) {
                DISPATCH_W(A, k_m_spmv, v, d);
                k_m_update<<<S->grid, kBlock, 0, st>>>(d);
                k_m_dir<<<S->grid, kBlock, 0, st>>>(d, S->iters);
            }
            GIP_LAUNCH_CHECK("mixed iteration");
            done += chunk;
            GIP_HIP_TRY(hipMemcpyAsync(h_sc, S->sc, sizeof h_sc, hipMemcpyDeviceToHost, st));
            if ((rc = poll_state(S))) return rc;
            if (S->h_state[0] != ST_RUN) break;
            const double rr32 = h_sc[S_RR];
            const double tol2 = h_sc[S_TOL2];
            const bool near_target = tol2 >= 0.0 && rr32 <= tol2;
            if (rr32 <= inner2 * last_rr || near_target || done >= o->maxit) {
                k_m_fold<0><<<S->grid, kBlock, 0, st>>>(d);
                DISPATCH_W(A, k_m_resid, v, d);
                GIP_LAUNCH_CHECK("mixed correction");
                ++corrections;
                GIP_HIP_TRY(hipMemcpyAsync(h_sc, S->sc, sizeof h_sc, hipMemcpyDeviceToHost, st));
                if ((rc = poll_state(S))) return rc;
                last_rr = h_sc[S_TRUE_RR];
            }
        }
        if (S->h_state[0] == ST_RUN) {
            // fold any correction left (maxit reached between corrections)
            k_m_fold<0><<<S->grid, kBlock, 0, st>>>(d);
            DISPATCH_W(A, k_m_resid, v, d);
            GIP_LAUNCH_CHECK("mixed final correction");
            ++corrections;
        }
    }
    double h_sc[S_COUNT];
    GIP_HIP_TRY(hipMemcpyAsync(h_sc, S->sc, sizeof h_sc, hipMemcpyDeviceToHost, st));
    if ((rc = poll_state(S))) return rc;
    int status = GIP_OK;
    if (S->h_state[0] == ST_BREAK)
        status = GIP_EBREAKDOWN;
    else if (o->tol > 0.0 && S->h_state[0] != ST_CONV)
        status = GIP_ENOCONV;
    if (stats) {
        stats->iters = S->h_state[1];
        stats->corrections = corrections;
        stats->relres = h_sc[S_BB] > 0.0 ? std::sqrt(h_sc[S_RR] / h_sc[S_BB]) : std::sqrt(h_sc[S_RR]);
        stats->status = status;
    }
    return status == GIP_EBREAKDOWN ? set_error(GIP_EBREAKDOWN, "PCG breakdown (p.Ap == 0 or non-finite step)") : GIP_OK;
}

}  // extern "C"
