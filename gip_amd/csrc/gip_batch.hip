// Batched lid-driven cavities (include/gip_batch.h): one concatenated,
// block-diagonal system, per-instance scalars, segmented deterministic
// reductions, per-instance PCG convergence, hipGraph-captured step phases.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "gip_batch.h"
#include "gip_blas1.h"
#include "gip_internal.h"
#include "gip_rows.h"
#include "gip_spmv.h"

namespace gip {

constexpr int kChunk = kBlock;  // rows per chunk = threads per workgroup

// per-instance factors of a step (GIP_DNS.h:326-327, :345, :367, :381-382)
enum { F_A1 = 0, F_A2, F_SIGMA, F_CORR, F_F1, F_F2, F_RHO, F_COUNT = 8 };
enum { FAC_ONE = -1, FAC_ZERO = -2 };
// per-instance PCG scalars
enum { B_RHO = 0, B_ALPHA, B_BETA, B_RR, B_BB, B_TOL2, B_RHO_PREV, B_LAST_RR, B_NEED, B_ZAZ, B_C, B_COUNT = 16 };
enum { BS_RUN = 0, BS_CONV = 1, BS_BREAK = 2, BS_MAXED = 3 };

struct Seg {
    const int32_t *chunk_inst;  // chunk -> instance
    const int32_t *inst_c0;     // instance -> first chunk
    const int32_t *inst_nch;    // instance -> chunk count
    int nchunk;
};

__device__ inline double fac_of(const double *fac, int inst, int code) {
    return code == FAC_ONE ? 1.0 : code == FAC_ZERO ? 0.0 : fac[inst * F_COUNT + code];
}

// Segmented reductions run as two launches: the chunk kernel stores each
// chunk's workgroup totals (seg_store), then a finalize kernel with one
// workgroup per instance combines that instance's chunk totals in chunk
// order (seg_total) and applies the per-instance update.  The kernel
// boundary orders the stores before the loads: no atomics, no device-scope
// fences.
template <int K, int OP>
__device__ inline void seg_store(double (&v)[K], int c, double *part) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = block_reduce<OP>(v[k]);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) part[(size_t)c * K + k] = v[k];
    }
}

// Instance totals of the chunk totals (valid in thread 0).
template <int K, int OP>
__device__ inline void seg_total(const double *part, int inst, const Seg &S, double (&out)[K]) {
    const int c0 = S.inst_c0[inst], nch = S.inst_nch[inst];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double acc = RedOp<OP>::identity();
        for (int j = threadIdx.x; j < nch; j += kBlock) acc = RedOp<OP>::combine(acc, part[(size_t)(c0 + j) * K + k]);
        out[k] = block_reduce<OP>(acc);
    }
}

#define FOR_CHUNKS(S, c) for (int c = blockIdx.x; c < (S).nchunk; c += gridDim.x)

__global__ void k_b_prep(int B, const double *__restrict__ dt, const double *__restrict__ sigma, double rho,
                         double *__restrict__ fac) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    double *f = fac + (size_t)i * F_COUNT;
    f[F_A1] = 1.5 * dt[i];
    f[F_A2] = -0.5 * dt[i];
    f[F_SIGMA] = sigma[i];
    f[F_CORR] = -1.0 / (rho * sigma[i]);
    f[F_F1] = -1.0 / sigma[i];
    f[F_F2] = 1.0 / sigma[i];
    f[F_RHO] = rho;
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_b_spmv_seg(EllView A, int nrows, const double *__restrict__ x,
                                                       double *__restrict__ y, Seg S, const double *__restrict__ fac,
                                                       int ia, int ib) {
    Operand<double> xo{x};
    FOR_CHUNKS(S, c) {
        const int row = c * kChunk + threadIdx.x;
        if (row >= nrows) continue;
        double sum;
        if (!row_dot<W, double, double>(A, A.val, xo, row, sum)) continue;
        const int inst = S.chunk_inst[c];
        y[row] = fac_of(fac, inst, ia) * sum + fac_of(fac, inst, ib) * y[row];
    }
}

// u += a1 ax ; u += a2 ax0  (two axpys, GIP_DNS.h:329-334)
__global__ __launch_bounds__(kBlock) void k_b_ab2(int n, double *__restrict__ u, const double *__restrict__ ax,
                                                  const double *__restrict__ ax0, Seg S, const double *__restrict__ fac) {
    FOR_CHUNKS(S, c) {
        const int row = c * kChunk + threadIdx.x;
        if (row >= n) continue;
        const int inst = S.chunk_inst[c];
        double t = u[row] + fac[inst * F_COUNT + F_A1] * ax[row];
        u[row] = t + fac[inst * F_COUNT + F_A2] * ax0[row];
    }
}

__global__ __launch_bounds__(kBlock) void k_b_axpy_seg(int n, double *__restrict__ y, const double *__restrict__ x, Seg S,
                                                       const double *__restrict__ fac, int code) {
    FOR_CHUNKS(S, c) {
        const int row = c * kChunk + threadIdx.x;
        if (row >= n) continue;
        y[row] = y[row] + fac_of(fac, S.chunk_inst[c], code) * x[row];
    }
}

__global__ __launch_bounds__(kBlock) void k_b_copy(int n, double *__restrict__ dst, const double *__restrict__ src) {
    const int stride = gridDim.x * kBlock;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// massr_i = ||temp||_inf over instance i's inner rows; then time_i += dt_i (GIP_DNS.h:398-401)
__global__ __launch_bounds__(kBlock) void k_b_massr(const double *__restrict__ temp, Seg S, double *part) {
    FOR_CHUNKS(S, c) {
        double v[1] = {fabs(temp[c * kChunk + threadIdx.x])};
        seg_store<1, RED_MAXABS>(v, c, part);
    }
}

__global__ __launch_bounds__(kBlock) void k_b_massr_fin(const double *part, Seg S, double *__restrict__ massr,
                                                        double *__restrict__ time, const double *__restrict__ dt,
                                                        double *__restrict__ dt_used) {
    const int inst = blockIdx.x;
    double out[1];
    seg_total<1, RED_MAXABS>(part, inst, S, out);
    if (threadIdx.x == 0) {
        massr[inst] = out[0];
        dt_used[inst] = dt[inst];
        time[inst] += dt[inst];
    }
}

// dt_i = CFL over instance i's inner rows (GIP_NonLinearMulticore.cpp:3-55), sigma_i = 1/dt_i
__global__ __launch_bounds__(kBlock) void k_b_cfl(const double *__restrict__ u, const double *__restrict__ v,
                                                  const double *__restrict__ w, const double *__restrict__ dxs, Seg S,
                                                  const double *__restrict__ gamma, double rho, double *part) {
    FOR_CHUNKS(S, c) {
        const int inst = S.chunk_inst[c], i = c * kChunk + threadIdx.x;
        const double g = gamma[inst];
        double acc = 1e20;
        double ut = u[i], vt = v[i], wt = w[i], dx = dxs[i];
        double normit = fabs(ut) > fabs(vt) ? fabs(ut) : fabs(vt);
        if (normit < fabs(wt)) normit = fabs(wt);
        if (!(dx / normit < 1e-17)) {
            double cc = 0.25 * dx / normit;
            if (acc > cc) acc = cc;
        }
        if (!(dx * dx * rho / g < 1e-17)) {
            double cc = 0.2 * dx * dx * rho / g;
            if (acc > cc) acc = cc;
        }
        double vv[1] = {acc};
        seg_store<1, RED_MIN>(vv, c, part);
    }
}

__global__ __launch_bounds__(kBlock) void k_b_cfl_fin(const double *part, Seg S, double *__restrict__ dt,
                                                      double *__restrict__ sigma) {
    const int inst = blockIdx.x;
    double out[1];
    seg_total<1, RED_MIN>(part, inst, S, out);
    if (threadIdx.x == 0) {
        const double d = 0.8 * out[0] > 1e-1 ? 1e-1 : 0.8 * out[0];  // GIP_NonLinearMulticore.cpp:47-50
        dt[inst] = d;
        sigma[inst] = 1.0 / d;
    }
}

// The PCG product and update kernels instead run one workgroup per chunk in
// dispatch order (grid = the chunk count: the pace of a plain stream) and
// store one partial per wavefront, with no barrier (seg_store_w); their
// finalize kernels sum an instance's wavefront partials in chunk order
// (seg_total_w).  A chunk of an instance that has left the PCG returns
// after two loads.
template <int K>
__device__ inline void seg_store_w(double (&v)[K], int c, double *part) {
    const size_t w = (size_t)c * kWaves + (threadIdx.x >> 6);
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) part[w * K + k] = v[k];
    }
}

template <int K>
__device__ inline void seg_total_w(const double *part, int inst, const Seg &S, double (&out)[K]) {
    const size_t w0 = (size_t)S.inst_c0[inst] * kWaves;
    const int nw = S.inst_nch[inst] * kWaves;
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
#pragma unroll 8
    for (int j = threadIdx.x; j < nw; j += kBlock)
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = acc[k] + part[(w0 + j) * K + k];
    block_reduce_n<K>(acc);
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = acc[k];
}

// ---- batched PCG ------------------------------------------------------------

struct BPcg {
    int B, np, nall;
    const double *b, *dinv;
    double *x, *r, *s, *Ap;
    float *r32, *s32, *Ap32, *xc32;
    const float *dinv32;
    const double *one, *az;  // MIXED coarse step: 1 on real inner rows (0 on padding), A 1
    double *sc;              // B * B_COUNT
    int *state, *iters, *active;
    double *part;  // chunk totals or wavefront partials
    int maxit;
    double inner2;
};

__device__ inline double *bsc(const BPcg &d, int inst) { return d.sc + (size_t)inst * B_COUNT; }

__device__ inline void leave(const BPcg &d, int inst, int st) {
    d.state[inst] = st;
    atomicSub(d.active, 1);
}

__global__ void k_bp_init(BPcg d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < d.B) {
        d.state[i] = BS_RUN;
        d.iters[i] = 0;
    }
    if (i == 0) *d.active = d.B;
}

// x = x0 (all), direction buffers: x0 past the inner rows, zero on them;
// Ap = (A x0) or x0 on empty rows; r = b + (-1.0) Ap    (GIP_LinearSolver.h:55-70)
template <int W>
__global__ __launch_bounds__(kBlock) void k_bp_setup_resid(EllView A, BPcg d, Seg S, const double *__restrict__ x0,
                                                           double *__restrict__ p, float *__restrict__ q) {
    const int stride = gridDim.x * kBlock;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < d.nall; i += stride) {
        d.x[i] = x0[i];
        p[i] = i >= d.np ? x0[i] : 0.0;
        if (q) q[i] = 0.0f;
    }
    Operand<double> xo{x0};
    FOR_CHUNKS(S, c) {
        const int row = c * kChunk + threadIdx.x;
        double sum;
        bool any = row_dot<W, double, double>(A, A.val, xo, row, sum);
        double t = any ? sum : x0[row];
        d.Ap[row] = t;
        d.r[row] = d.b[row] + -1.0 * t;
    }
}

// s = Dinv r (the first direction, formed by the first fused iteration with
// beta = 0); rho = r.s, b.b, r.r; state (GIP_LinearSolver.h:72-80)
__global__ __launch_bounds__(kBlock) void k_bp_setup_dir(BPcg d, Seg S, int precond) {
    FOR_CHUNKS(S, c) {
        const int row = c * kChunk + threadIdx.x;
        const double ri = d.r[row];
        double si;
        if (precond) {
            si = 0.0;
            si = si + ri * d.dinv[row];
        } else {
            si = ri;
        }
        d.s[row] = si;
        double v[3] = {ri * si, d.b[row] * d.b[row], ri * ri};
        seg_store<3, RED_SUM>(v, c, d.part);
    }
}

__global__ __launch_bounds__(kBlock) void k_bp_setup_fin(BPcg d, Seg S, double tol) {
    const int inst = blockIdx.x;
    double out[3];
    seg_total<3, RED_SUM>(d.part, inst, S, out);
    if (threadIdx.x == 0) {
        double *q = bsc(d, inst);
        q[B_RHO] = out[0];
        q[B_BB] = out[1];
        q[B_RR] = out[2];
        q[B_TOL2] = tol * tol * out[1];
        q[B_BETA] = 0.0;
        if (out[2] <= q[B_TOL2]) leave(d, inst, BS_CONV);
    }
}

// direction of the running instances: p = 1.0 s + beta p (GIP_LinearSolver.h:113;
// the first iteration has beta = 0 and p = 0 on the solved rows, so p = s)
// (a grid-stride sweep: at 12-24 B per row a workgroup per chunk spends
// more on its launch than on its rows)
__global__ __launch_bounds__(kBlock) void k_bp_dir(BPcg d, Seg S, double *__restrict__ p) {
    const double *__restrict__ s = d.s;
    FOR_CHUNKS(S, c) {
        const int inst = S.chunk_inst[c];
        if (d.state[inst] != BS_RUN) continue;
        const int row = c * kChunk + threadIdx.x;
        const double beta = bsc(d, inst)[B_BETA];
        const double si = s[row], po = p[row];
        p[row] = 1.0 * si + beta * po;
    }
}

// Ap = A p (one gather of p per neighbour), chunk partials of p.Ap
template <int W>
__global__ __launch_bounds__(kBlock) void k_bp_spmv(EllView A, BPcg d, Seg S, const double *__restrict__ p) {
    const int c = blockIdx.x;
    const int inst = S.chunk_inst[c];
    if (d.state[inst] != BS_RUN) return;
    const int row = c * kChunk + threadIdx.x;
    Operand<double> po{p};
    const double pi = p[row];
    double sum;
    if (row_dot<W, double, double>(A, A.val, po, row, sum))
        d.Ap[row] = sum;
    else
        sum = d.Ap[row];
    double v[1] = {pi * sum};
    seg_store_w<1>(v, c, d.part);
}

// alpha = rho / p.Ap of the running instances (FP64 and MIXED)
__global__ __launch_bounds__(kBlock) void k_bp_spmv_fin(BPcg d, Seg S) {
    const int inst = blockIdx.x;
    if (d.state[inst] != BS_RUN) return;
    double out[1];
    seg_total_w<1>(d.part, inst, S, out);
    if (threadIdx.x == 0) {
        double *q = bsc(d, inst);
        const double alpha = q[B_RHO] / out[0];
        q[B_ALPHA] = alpha;
        if (out[0] == 0.0 || !isfinite(out[0]) || !isfinite(alpha)) leave(d, inst, BS_BREAK);
    }
}

__global__ __launch_bounds__(kBlock) void k_bp_update(BPcg d, Seg S, const double *__restrict__ p, int precond) {
    const int c = blockIdx.x;
    const int inst = S.chunk_inst[c];
    if (d.state[inst] != BS_RUN) return;
    const int row = c * kChunk + threadIdx.x;
    const double alpha = bsc(d, inst)[B_ALPHA];
    // every operand first: one memory round trip per row
    const double xo = d.x[row], pr = p[row], ro = d.r[row], apr = d.Ap[row];
    const double di = precond ? d.dinv[row] : 0.0;
    d.x[row] = xo + alpha * pr;
    const double ri = ro + (-1.0 * alpha) * apr;
    d.r[row] = ri;
    double si;
    if (precond) {
        si = 0.0;
        si = si + ri * di;
    } else {
        si = ri;
    }
    d.s[row] = si;
    double v[2] = {ri * si, ri * ri};
    seg_store_w<2>(v, c, d.part);
}

__global__ __launch_bounds__(kBlock) void k_bp_update_fin(BPcg d, Seg S) {
    const int inst = blockIdx.x;
    if (d.state[inst] != BS_RUN) return;
    double out[2];
    seg_total_w<2>(d.part, inst, S, out);
    if (threadIdx.x == 0) {
        double *q = bsc(d, inst);
        const double r0 = q[B_RHO];
        q[B_RHO] = out[0];
        q[B_BETA] = out[0] / r0;
        q[B_RR] = out[1];
        const int it = ++d.iters[inst];
        if (out[1] <= q[B_TOL2])
            leave(d, inst, BS_CONV);
        else if (it >= d.maxit)
            leave(d, inst, BS_MAXED);
    }
}

// MIXED coarse step of each instance (the single-cavity solver's, per
// instance): chunk partials of 1.r over the running instances (only those
// flagged for a correction when need_only) ...
__global__ __launch_bounds__(kBlock) void k_bpm_sumr(BPcg d, Seg S, int need_only) {
    FOR_CHUNKS(S, c) {
        const int inst = S.chunk_inst[c];
        if (d.state[inst] != BS_RUN || (need_only && bsc(d, inst)[B_NEED] == 0.0)) continue;
        const int row = c * kChunk + threadIdx.x;
        double v[1] = {d.r[row] * d.one[row]};
        seg_store<1, RED_SUM>(v, c, d.part);
    }
}

// ... c = (1.r)/(1.A1) per instance (0 when 1.A1 is not positive and finite)
__global__ __launch_bounds__(kBlock) void k_bpm_c_fin(BPcg d, Seg S, int need_only) {
    const int inst = blockIdx.x;
    if (d.state[inst] != BS_RUN || (need_only && bsc(d, inst)[B_NEED] == 0.0)) return;
    double out[1];
    seg_total<1, RED_SUM>(d.part, inst, S, out);
    if (threadIdx.x == 0) {
        double *q = bsc(d, inst);
        const double zaz = q[B_ZAZ];
        double cc = (zaz > 0.0 && isfinite(zaz)) ? out[0] / zaz : 0.0;
        q[B_C] = isfinite(cc) ? cc : 0.0;
    }
}

// mixed: the coarse step, then FP32 copies of the FP64 residual of x0; rho, b.b, true r.r
__global__ __launch_bounds__(kBlock) void k_bpm_setup(BPcg d, Seg S, int precond) {
    FOR_CHUNKS(S, c) {
        const int row = c * kChunk + threadIdx.x;
        const double cc = bsc(d, S.chunk_inst[c])[B_C];
        double ri = d.r[row];
        if (cc != 0.0) {
            d.x[row] = d.x[row] + cc * d.one[row];
            ri = ri - cc * d.az[row];
            d.r[row] = ri;
        }
        const float r32 = (float)ri;
        const float s32 = precond ? r32 * d.dinv32[row] : r32;
        d.r32[row] = r32;
        d.s32[row] = s32;
        d.xc32[row] = 0.0f;
        double v[3] = {(double)r32 * (double)s32, d.b[row] * d.b[row], ri * ri};
        seg_store<3, RED_SUM>(v, c, d.part);
    }
}

__global__ __launch_bounds__(kBlock) void k_bpm_setup_fin(BPcg d, Seg S, double tol) {
    const int inst = blockIdx.x;
    double out[3];
    seg_total<3, RED_SUM>(d.part, inst, S, out);
    if (threadIdx.x == 0) {
        double *q = bsc(d, inst);
        q[B_RHO] = out[0];
        q[B_BB] = out[1];
        q[B_RR] = out[2];
        q[B_LAST_RR] = out[2];
        q[B_TOL2] = tol * tol * out[1];
        q[B_BETA] = 0.0;
        q[B_NEED] = 0.0;
        if (out[2] <= q[B_TOL2]) leave(d, inst, BS_CONV);
    }
}

// (nontemporal accesses here and in the update: each array is streamed once
// per sweep; 64-cavity batch 133 -> 126 ms of PCG per step)
__global__ __launch_bounds__(kBlock) void k_bpm_dir(BPcg d, Seg S, float *__restrict__ p) {
    const float *__restrict__ s = d.s32;
    FOR_CHUNKS(S, c) {
        const int inst = S.chunk_inst[c];
        if (d.state[inst] != BS_RUN) continue;
        const int row = c * kChunk + threadIdx.x;
        const float beta = (float)bsc(d, inst)[B_BETA];
        const float si = __builtin_nontemporal_load(s + row), po = __builtin_nontemporal_load(p + row);
        __builtin_nontemporal_store(1.0f * si + beta * po, p + row);
    }
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_bpm_spmv(EllView A, BPcg d, Seg S, const float *__restrict__ p) {
    const int c = blockIdx.x;
    const int inst = S.chunk_inst[c];
    if (d.state[inst] != BS_RUN) return;
    const int row = c * kChunk + threadIdx.x;
    Operand<float> po{p};
    const float pi = p[row];
    float sum;
    row_dot<W, float, float>(A, A.val32, po, row, sum);
    d.Ap32[row] = sum;
    double v[1] = {(double)pi * (double)sum};
    seg_store_w<1>(v, c, d.part);
}

__global__ __launch_bounds__(kBlock) void k_bpm_update(BPcg d, Seg S, const float *__restrict__ p, int precond) {
    const int c = blockIdx.x;
    const int inst = S.chunk_inst[c];
    if (d.state[inst] != BS_RUN) return;
    const int row = c * kChunk + threadIdx.x;
    const float a32 = (float)bsc(d, inst)[B_ALPHA];
    const float xo = __builtin_nontemporal_load(d.xc32 + row), pr = __builtin_nontemporal_load(p + row),
                ro = __builtin_nontemporal_load(d.r32 + row), apr = __builtin_nontemporal_load(d.Ap32 + row);
    const float di = precond ? __builtin_nontemporal_load(d.dinv32 + row) : 0.0f;
    __builtin_nontemporal_store(xo + a32 * pr, d.xc32 + row);
    const float ri = ro - a32 * apr;
    __builtin_nontemporal_store(ri, d.r32 + row);
    const float si = precond ? ri * di : ri;
    __builtin_nontemporal_store(si, d.s32 + row);
    double v[2] = {(double)ri * (double)si, (double)ri * (double)ri};
    seg_store_w<2>(v, c, d.part);
}

__global__ __launch_bounds__(kBlock) void k_bpm_update_fin(BPcg d, Seg S) {
    const int inst = blockIdx.x;
    if (d.state[inst] != BS_RUN) return;
    double out[2];
    seg_total_w<2>(d.part, inst, S, out);
    if (threadIdx.x == 0) {
        double *q = bsc(d, inst);
        const double r0 = q[B_RHO];
        q[B_RHO_PREV] = r0;
        q[B_RHO] = out[0];
        q[B_BETA] = out[0] / r0;
        q[B_RR] = out[1];
        const int it = ++d.iters[inst];
        if (out[1] <= d.inner2 * q[B_LAST_RR] || out[1] <= q[B_TOL2] || it >= d.maxit) q[B_NEED] = 1.0;
    }
}

// FP64 defect correction of the flagged instances: x += xc, xc = 0
__global__ __launch_bounds__(kBlock) void k_bpm_fold(BPcg d, Seg S) {
    FOR_CHUNKS(S, c) {
        const int inst = S.chunk_inst[c];
        if (d.state[inst] != BS_RUN || bsc(d, inst)[B_NEED] == 0.0) continue;
        const int row = c * kChunk + threadIdx.x;
        d.x[row] = d.x[row] + (double)d.xc32[row];
        d.xc32[row] = 0.0f;
    }
}

// ... then r = b - A x (FP64) replaces the FP32 residual (after the coarse
// step); the direction is kept and beta uses the replaced residual;
// convergence on the true ||r||.
template <int W>
__global__ __launch_bounds__(kBlock) void k_bpm_resid(EllView A, BPcg d, Seg S) {
    Operand<double> xo{d.x};
    FOR_CHUNKS(S, c) {
        const int inst = S.chunk_inst[c];
        if (d.state[inst] != BS_RUN || bsc(d, inst)[B_NEED] == 0.0) continue;
        const int row = c * kChunk + threadIdx.x;
        double sum;
        const bool any = row_dot<W, double, double>(A, A.val, xo, row, sum);
        const double t = any ? sum : d.x[row];
        const double ri = d.b[row] + -1.0 * t;
        d.r[row] = ri;
        double v[1] = {ri * d.one[row]};
        seg_store<1, RED_SUM>(v, c, d.part);
    }
}

__global__ __launch_bounds__(kBlock) void k_bpm_coarse(BPcg d, Seg S, int precond) {
    FOR_CHUNKS(S, c) {
        const int inst = S.chunk_inst[c];
        if (d.state[inst] != BS_RUN || bsc(d, inst)[B_NEED] == 0.0) continue;
        const int row = c * kChunk + threadIdx.x;
        const double cc = bsc(d, inst)[B_C];
        double ri = d.r[row];
        if (cc != 0.0) {
            d.x[row] = d.x[row] + cc * d.one[row];
            ri = ri - cc * d.az[row];
            d.r[row] = ri;
        }
        const float r32 = (float)ri;
        const float s32 = precond ? r32 * d.dinv32[row] : r32;
        d.r32[row] = r32;
        d.s32[row] = s32;
        double v[2] = {(double)r32 * (double)s32, ri * ri};
        seg_store<2, RED_SUM>(v, c, d.part);
    }
}

__global__ __launch_bounds__(kBlock) void k_bpm_resid_fin(BPcg d, Seg S) {
    const int inst = blockIdx.x;
    if (d.state[inst] != BS_RUN || bsc(d, inst)[B_NEED] == 0.0) return;
    double out[2];
    seg_total<2, RED_SUM>(d.part, inst, S, out);
    if (threadIdx.x == 0) {
        double *q = bsc(d, inst);
        // keep the direction unless the FP32 recursion drifted >10x from the true residual
        q[B_BETA] = out[1] > 100.0 * q[B_RR] ? 0.0 : out[0] / q[B_RHO_PREV];
        q[B_RHO] = out[0];
        q[B_RR] = out[1];
        q[B_LAST_RR] = out[1];
        q[B_NEED] = 0.0;
        if (out[1] <= q[B_TOL2])
            leave(d, inst, BS_CONV);
        else if (d.iters[inst] >= d.maxit)
            leave(d, inst, BS_MAXED);
    }
}

}  // namespace gip

using namespace gip;

namespace {

enum { OP_D = 0, OP_EC, OP_L, OP_LB, OP_MX, OP_MY, OP_MZ, OP_GX, OP_GY, OP_GZ, OP_SFX, OP_SFY, OP_SFZ, OP_SO, OP_MRES, OP_N };
enum { ROWS_INNER, ROWS_ALL, ROWS_FACES, ROWS_EC };
enum { COLS_CELLS, COLS_FACES };

const gip_csr &op_of(const gip_cavity *c, int k) {
    const gip_csr *t[OP_N] = {&c->D, &c->Ec, &c->L, &c->Lb, &c->Mx, &c->My, &c->Mz, &c->Gx,
                              &c->Gy, &c->Gz, &c->SFx, &c->SFy, &c->SFz, &c->SO, &c->Mres};
    return *t[k];
}
const int kRows[OP_N] = {ROWS_INNER, ROWS_EC, ROWS_INNER, ROWS_ALL, ROWS_INNER, ROWS_INNER, ROWS_INNER, ROWS_ALL,
                         ROWS_ALL, ROWS_ALL, ROWS_FACES, ROWS_FACES, ROWS_FACES, ROWS_FACES, ROWS_INNER};
const int kCols[OP_N] = {COLS_CELLS, COLS_FACES, COLS_CELLS, COLS_CELLS, COLS_CELLS, COLS_CELLS, COLS_CELLS, COLS_CELLS,
                         COLS_CELLS, COLS_CELLS, COLS_CELLS, COLS_CELLS, COLS_CELLS, COLS_CELLS, COLS_FACES};

int64_t roundup(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

}  // namespace

struct gip_batch {
    gip_ctx *c = nullptr;
    int B = 0;
    double rho = 1.0;
    gip_batch_opts o{};
    std::vector<int64_t> N, G, F, off_in, off_gh, off_f, off_ec;
    int64_t NP = 0, ALLB = 0, FP = 0, GH = 0, ECR = 0;
    gip_mat *op[OP_N] = {};
    gip_mat *C = nullptr;
    // row-space segmentations (device)
    int32_t *d_seg_in[3] = {}, *d_seg_f[3] = {};
    Seg seg_in{}, seg_f{};
    // fields
    double *u = nullptr, *v = nullptr, *w = nullptr, *ax = nullptr, *ay = nullptr, *az = nullptr, *ax0 = nullptr,
           *ay0 = nullptr, *az0 = nullptr, *rhs = nullptr, *p = nullptr, *p0 = nullptr, *gpx = nullptr,
           *gpy = nullptr, *gpz = nullptr, *mf = nullptr, *dxs = nullptr, *temp = nullptr;
    // per instance
    double *dt = nullptr, *sigma = nullptr, *time = nullptr, *gamma = nullptr, *massr = nullptr, *dt_used = nullptr,
           *fac = nullptr;
    // pcg
    double *r = nullptr, *s = nullptr, *Ap = nullptr, *dinv = nullptr, *pp = nullptr;  // pp: the direction
    float *r32 = nullptr, *s32 = nullptr, *Ap32 = nullptr, *xc32 = nullptr, *dinv32 = nullptr, *pq = nullptr;
    double *cone = nullptr, *caz = nullptr;  // MIXED coarse step: 1 on inner rows, A 1
    double *sc = nullptr, *part = nullptr;
    int *state = nullptr, *iters = nullptr, *active = nullptr, *h_active = nullptr;
    bool started = false;
    hipGraphExec_t g_pre = nullptr, g_iter = nullptr, g_post = nullptr;
    // events around the step's three phases (recorded outside any capture)
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    float phase_ms[3] = {0.0f, 0.0f, 0.0f};
    int grid_in = 0, grid_f = 0, grid_all = 0;
    std::vector<void *> allocs;
};

namespace {

int dalloc(gip_batch *b, void **p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes ? bytes : 8);
    if (e != hipSuccess) return hip_fail(e, "batch allocation");
    b->allocs.push_back(*p);
    e = hipMemsetAsync(*p, 0, bytes ? bytes : 8, b->c->stream);
    if (e != hipSuccess) return hip_fail(e, "batch zero");
    return GIP_OK;
}

template <class T>
int dalloc_t(gip_batch *b, T **p, size_t count) {
    return dalloc(b, reinterpret_cast<void **>(p), sizeof(T) * count);
}

// Block-diagonal assembly of operator k over every instance.
void assemble(gip_batch *b, const gip_cavity *const *cavs, int k, std::vector<int32_t> &rp, std::vector<int32_t> &cl,
              std::vector<double> &vl, int64_t &nrows) {
    const int rows = kRows[k], cols = kCols[k];
    nrows = rows == ROWS_INNER ? b->NP : rows == ROWS_ALL ? b->ALLB : rows == ROWS_FACES ? b->FP : b->ECR;
    std::vector<int64_t> len((size_t)nrows + 1, 0);
    auto row_of = [&](int i, int64_t r) -> int64_t {
        switch (rows) {
            case ROWS_INNER: return b->off_in[i] + r;
            case ROWS_ALL: return r < b->N[i] ? b->off_in[i] + r : b->NP + b->off_gh[i] + (r - b->N[i]);
            case ROWS_FACES: return b->off_f[i] + r;
            default: return b->off_ec[i] + r;
        }
    };
    auto col_of = [&](int i, int64_t cidx) -> int64_t {
        if (cols == COLS_FACES) return b->off_f[i] + cidx;
        return cidx < b->N[i] ? b->off_in[i] + cidx : b->NP + b->off_gh[i] + (cidx - b->N[i]);
    };
    for (int i = 0; i < b->B; ++i) {
        const gip_csr &A = op_of(cavs[i], k);
        for (int64_t r = 0; r < A.nrows; ++r) len[(size_t)row_of(i, r) + 1] = A.rowptr[r + 1] - A.rowptr[r];
    }
    for (int64_t r = 0; r < nrows; ++r) len[(size_t)r + 1] += len[(size_t)r];
    rp.assign(len.begin(), len.end());
    cl.assign((size_t)len[(size_t)nrows], 0);
    vl.assign((size_t)len[(size_t)nrows], 0.0);
    for (int i = 0; i < b->B; ++i) {
        const gip_csr &A = op_of(cavs[i], k);
        for (int64_t r = 0; r < A.nrows; ++r) {
            int64_t q = len[(size_t)row_of(i, r)];
            for (int32_t j = A.rowptr[r]; j < A.rowptr[r + 1]; ++j, ++q) {
                cl[(size_t)q] = (int32_t)col_of(i, A.col[j]);
                vl[(size_t)q] = A.val[j];
            }
        }
    }
}

int make_seg(gip_batch *b, const std::vector<int64_t> &off, const std::vector<int64_t> &cnt, int64_t total,
             int32_t **d, Seg *S) {
    const int64_t nch = total / kChunk;
    std::vector<int32_t> ci((size_t)std::max<int64_t>(nch, 1)), c0((size_t)b->B), nc((size_t)b->B);
    for (int i = 0; i < b->B; ++i) {
        c0[(size_t)i] = (int32_t)(off[(size_t)i] / kChunk);
        nc[(size_t)i] = (int32_t)(roundup(cnt[(size_t)i], kChunk) / kChunk);
        for (int32_t k = 0; k < nc[(size_t)i]; ++k) ci[(size_t)(c0[(size_t)i] + k)] = i;
    }
    int rc = dalloc_t(b, &d[0], ci.size());
    if (!rc) rc = dalloc_t(b, &d[1], c0.size());
    if (!rc) rc = dalloc_t(b, &d[2], nc.size());
    if (rc) return rc;
    GIP_HIP_TRY(hipMemcpyAsync(d[0], ci.data(), ci.size() * 4, hipMemcpyHostToDevice, b->c->stream));
    GIP_HIP_TRY(hipMemcpyAsync(d[1], c0.data(), c0.size() * 4, hipMemcpyHostToDevice, b->c->stream));
    GIP_HIP_TRY(hipMemcpyAsync(d[2], nc.data(), nc.size() * 4, hipMemcpyHostToDevice, b->c->stream));
    GIP_HIP_TRY(hipStreamSynchronize(b->c->stream));
    *S = Seg{d[0], d[1], d[2], (int)nch};
    return GIP_OK;
}

// the shared operator view and kernel width (gip_rows.h)
EllView view_of(const gip_mat *A) { return view(A); }

#define BDISPATCH(A, GRID, KERNEL, ...)                                                          \
    do {                                                                                         \
        switch (kernel_width(A)) {                                                               \
            case 6: KERNEL<6><<<GRID, kBlock, 0, st>>>(__VA_ARGS__); break;                      \
            case 7: KERNEL<7><<<GRID, kBlock, 0, st>>>(__VA_ARGS__); break;                      \
            case -7: KERNEL<-7><<<GRID, kBlock, 0, st>>>(__VA_ARGS__); break;                    \
            case -8: KERNEL<-8><<<GRID, kBlock, 0, st>>>(__VA_ARGS__); break;                    \
            default: KERNEL<0><<<GRID, kBlock, 0, st>>>(__VA_ARGS__); break;                     \
        }                                                                                        \
    } while (0)

int seg_spmv(gip_batch *b, const gip_mat *A, const double *x, double *y, bool faces, int ia, int ib) {
    hipStream_t st = b->c->stream;
    const Seg &S = faces ? b->seg_f : b->seg_in;
    const int grid = faces ? b->grid_f : b->grid_in;
    BDISPATCH(A, grid, k_b_spmv_seg, view_of(A), A->nrows, x, y, S, b->fac, ia, ib);
    GIP_LAUNCH_CHECK("batch segmented spmv");
    return GIP_OK;
}

BPcg pcg_view(gip_batch *b) {
    BPcg d;
    d.B = b->B;
    d.np = (int)b->NP;
    d.nall = (int)b->ALLB;
    d.b = b->rhs;
    d.dinv = b->dinv;
    d.x = b->p;
    d.r = b->r;
    d.s = b->s;
    d.Ap = b->Ap;
    d.r32 = b->r32;
    d.s32 = b->s32;
    d.Ap32 = b->Ap32;
    d.xc32 = b->xc32;
    d.dinv32 = b->dinv32;
    d.one = b->cone;
    d.az = b->caz;
    d.sc = b->sc;
    d.state = b->state;
    d.iters = b->iters;
    d.active = b->active;
    d.part = b->part;
    d.maxit = b->o.pcg_maxit;
    const double it = b->o.inner_tol > 0 ? b->o.inner_tol : 0.1;
    d.inner2 = it * it;
    return d;
}

// predictor, right-hand side and PCG setup (GIP_DNS.h:308-347, GIP_LinearSolver.h:54-80)
int enqueue_pre(gip_batch *b) {
    hipStream_t st = b->c->stream;
    const int NP = (int)b->NP, ALLB = (int)b->ALLB;
    k_b_prep<<<(b->B + 63) / 64, 64, 0, st>>>(b->B, b->dt, b->sigma, b->rho, b->fac);
    int rc = gip_dcsrmv_vals(b->c, b->op[OP_EC], b->mf, b->C);
    const gip_mat *D = b->op[OP_D];
    if (!rc) rc = spmv_launch(b->c, D, b->u, b->ax, 1.0, 0.0, 0, nullptr);
    if (!rc) rc = spmv_launch(b->c, D, b->v, b->ay, 1.0, 0.0, 0, nullptr);
    if (!rc) rc = spmv_launch(b->c, D, b->w, b->az, 1.0, 0.0, 0, nullptr);
    if (!rc) rc = spmv_launch(b->c, b->C, b->u, b->ax, -1.0, 1.0, 1, nullptr);
    if (!rc) rc = spmv_launch(b->c, b->C, b->v, b->ay, -1.0, 1.0, 1, nullptr);
    if (!rc) rc = spmv_launch(b->c, b->C, b->w, b->az, -1.0, 1.0, 1, nullptr);
    if (rc) return rc;
    k_b_ab2<<<b->grid_in, kBlock, 0, st>>>(NP, b->u, b->ax, b->ax0, b->seg_in, b->fac);
    k_b_ab2<<<b->grid_in, kBlock, 0, st>>>(NP, b->v, b->ay, b->ay0, b->seg_in, b->fac);
    k_b_ab2<<<b->grid_in, kBlock, 0, st>>>(NP, b->w, b->az, b->az0, b->seg_in, b->fac);
    k_b_copy<<<b->grid_in, kBlock, 0, st>>>(NP, b->ax0, b->ax);
    k_b_copy<<<b->grid_in, kBlock, 0, st>>>(NP, b->ay0, b->ay);
    k_b_copy<<<b->grid_in, kBlock, 0, st>>>(NP, b->az0, b->az);
    GIP_LAUNCH_CHECK("batch predictor");
    rc = spmv_launch(b->c, b->op[OP_MX], b->u, b->rhs, 1.0, 0.0, 0, nullptr);
    if (!rc) rc = spmv_launch(b->c, b->op[OP_MY], b->v, b->rhs, 1.0, 1.0, 1, nullptr);
    if (!rc) rc = seg_spmv(b, b->op[OP_MZ], b->w, b->rhs, false, F_SIGMA, F_SIGMA);
    if (rc) return rc;
    // PCG setup: x = p, x0 = p0
    BPcg d = pcg_view(b);
    const gip_mat *L = b->op[OP_L];
    const bool mixed = b->o.precision == GIP_PCG_MIXED;
    k_bp_init<<<(b->B + 63) / 64, 64, 0, st>>>(d);
    BDISPATCH(L, b->grid_in, k_bp_setup_resid, view_of(L), d, b->seg_in, b->p0, b->pp, mixed ? b->pq : nullptr);
    if (mixed) {
        k_bpm_sumr<<<b->grid_in, kBlock, 0, st>>>(d, b->seg_in, 0);
        k_bpm_c_fin<<<b->B, kBlock, 0, st>>>(d, b->seg_in, 0);
        k_bpm_setup<<<b->grid_in, kBlock, 0, st>>>(d, b->seg_in, 1);
        k_bpm_setup_fin<<<b->B, kBlock, 0, st>>>(d, b->seg_in, b->o.pcg_tol);
    } else {
        k_bp_setup_dir<<<b->grid_in, kBlock, 0, st>>>(d, b->seg_in, 1);
        k_bp_setup_fin<<<b->B, kBlock, 0, st>>>(d, b->seg_in, b->o.pcg_tol);
    }
    (void)ALLB;
    GIP_LAUNCH_CHECK("batch pcg setup");
    return GIP_OK;
}

// `poll` iterations: direction, product, alpha, update, beta and convergence
int enqueue_iter(gip_batch *b) {
    hipStream_t st = b->c->stream;
    BPcg d = pcg_view(b);
    const gip_mat *L = b->op[OP_L];
    const bool mixed = b->o.precision == GIP_PCG_MIXED;
    const unsigned nch = (unsigned)b->seg_in.nchunk;  // one workgroup per chunk
    for (int k = 0; k < b->o.poll; ++k) {
        if (mixed) {
            k_bpm_dir<<<b->grid_in, kBlock, 0, st>>>(d, b->seg_in, b->pq);
            BDISPATCH(L, nch, k_bpm_spmv, view_of(L), d, b->seg_in, b->pq);
            k_bp_spmv_fin<<<b->B, kBlock, 0, st>>>(d, b->seg_in);
            k_bpm_update<<<nch, kBlock, 0, st>>>(d, b->seg_in, b->pq, 1);
            k_bpm_update_fin<<<b->B, kBlock, 0, st>>>(d, b->seg_in);
        } else {
            k_bp_dir<<<b->grid_in, kBlock, 0, st>>>(d, b->seg_in, b->pp);
            BDISPATCH(L, nch, k_bp_spmv, view_of(L), d, b->seg_in, b->pp);
            k_bp_spmv_fin<<<b->B, kBlock, 0, st>>>(d, b->seg_in);
            k_bp_update<<<nch, kBlock, 0, st>>>(d, b->seg_in, b->pp, 1);
            k_bp_update_fin<<<b->B, kBlock, 0, st>>>(d, b->seg_in);
        }
    }
    if (mixed) {
        k_bpm_fold<<<b->grid_in, kBlock, 0, st>>>(d, b->seg_in);
        BDISPATCH(L, b->grid_in, k_bpm_resid, view_of(L), d, b->seg_in);
        k_bpm_c_fin<<<b->B, kBlock, 0, st>>>(d, b->seg_in, 1);
        k_bpm_coarse<<<b->grid_in, kBlock, 0, st>>>(d, b->seg_in, 1);
        k_bpm_resid_fin<<<b->B, kBlock, 0, st>>>(d, b->seg_in);
    }
    GIP_LAUNCH_CHECK("batch pcg iterations");
    return GIP_OK;
}

// ghost shuffle, correction, fluxes, massr, CFL (GIP_DNS.h:348-414)
int enqueue_post(gip_batch *b) {
    hipStream_t st = b->c->stream;
    const int NP = (int)b->NP, ALLB = (int)b->ALLB;
    k_b_copy<<<b->grid_all, kBlock, 0, st>>>(NP, b->p0, b->p);
    k_b_copy<<<b->grid_all, kBlock, 0, st>>>(ALLB, b->p, b->p0);
    int rc = spmv_launch(b->c, b->op[OP_LB], b->p0, b->p, 1.0, 1.0, 1, nullptr);
    if (!rc) rc = spmv_launch(b->c, b->op[OP_GX], b->p, b->gpx, 1.0, 0.0, 0, nullptr);
    if (!rc) rc = spmv_launch(b->c, b->op[OP_GY], b->p, b->gpy, 1.0, 0.0, 0, nullptr);
    if (!rc) rc = spmv_launch(b->c, b->op[OP_GZ], b->p, b->gpz, 1.0, 0.0, 0, nullptr);
    if (rc) return rc;
    k_b_axpy_seg<<<b->grid_in, kBlock, 0, st>>>(NP, b->u, b->gpx, b->seg_in, b->fac, F_CORR);
    k_b_axpy_seg<<<b->grid_in, kBlock, 0, st>>>(NP, b->v, b->gpy, b->seg_in, b->fac, F_CORR);
    k_b_axpy_seg<<<b->grid_in, kBlock, 0, st>>>(NP, b->w, b->gpz, b->seg_in, b->fac, F_CORR);
    GIP_LAUNCH_CHECK("batch correction");
    rc = seg_spmv(b, b->op[OP_SO], b->p, b->mf, true, F_F1, FAC_ZERO);
    if (!rc) rc = seg_spmv(b, b->op[OP_SFX], b->gpx, b->mf, true, F_F2, FAC_ONE);
    if (!rc) rc = seg_spmv(b, b->op[OP_SFY], b->gpy, b->mf, true, F_F2, FAC_ONE);
    if (!rc) rc = seg_spmv(b, b->op[OP_SFZ], b->gpz, b->mf, true, F_F2, FAC_ONE);
    if (!rc) rc = spmv_launch(b->c, b->op[OP_SFX], b->u, b->mf, b->rho, 1.0, 1, nullptr);
    if (!rc) rc = spmv_launch(b->c, b->op[OP_SFY], b->v, b->mf, b->rho, 1.0, 1, nullptr);
    if (!rc) rc = spmv_launch(b->c, b->op[OP_SFZ], b->w, b->mf, b->rho, 1.0, 1, nullptr);
    if (!rc) rc = spmv_launch(b->c, b->op[OP_MRES], b->mf, b->temp, 1.0, 0.0, 0, nullptr);
    if (rc) return rc;
    k_b_massr<<<b->grid_in, kBlock, 0, st>>>(b->temp, b->seg_in, b->part);
    k_b_massr_fin<<<b->B, kBlock, 0, st>>>(b->part, b->seg_in, b->massr, b->time, b->dt, b->dt_used);
    k_b_cfl<<<b->grid_in, kBlock, 0, st>>>(b->u, b->v, b->w, b->dxs, b->seg_in, b->gamma, b->rho, b->part);
    k_b_cfl_fin<<<b->B, kBlock, 0, st>>>(b->part, b->seg_in, b->dt, b->sigma);
    GIP_LAUNCH_CHECK("batch fluxes and CFL");
    return GIP_OK;
}

int capture(gip_batch *b, int (*fn)(gip_batch *), hipGraphExec_t *out) {
    hipStream_t st = b->c->stream;
    hipGraph_t g = nullptr;
    GIP_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = fn(b);
    hipError_t e = hipStreamEndCapture(st, &g);
    if (rc) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    return GIP_OK;
}

int run_phase(gip_batch *b, int (*fn)(gip_batch *), hipGraphExec_t *g) {
    if (!b->o.use_graph) return fn(b);
    if (!*g) {
        int rc = capture(b, fn, g);
        if (rc) return rc;
    }
    GIP_HIP_TRY(hipGraphLaunch(*g, b->c->stream));
    return GIP_OK;
}

}  // namespace

extern "C" {

int gip_batch_create(gip_ctx *c, int32_t ninst, const gip_cavity *const *cavs, const double *gamma, double rho,
                     const gip_batch_opts *opts, gip_batch **out) {
    if (!c || ninst < 1 || !cavs || !gamma || !opts || !out) return set_error(GIP_EINVAL, "gip_batch_create: bad argument");
    if (opts->pcg_tol <= 0.0 || opts->pcg_maxit < 1) return set_error(GIP_EINVAL, "gip_batch_create: needs pcg_tol > 0 and pcg_maxit >= 1");
    for (int i = 0; i < ninst; ++i) {
        if (!cavs[i]) return set_error(GIP_EINVAL, "gip_batch_create: null cavity");
        for (int k = 0; k < OP_N; ++k)
            if (!op_of(cavs[i], k).rowptr) return set_error(GIP_EINVAL, "gip_batch_create: a cavity lacks an operator");
    }
    *out = nullptr;
    GIP_HIP_TRY(hipSetDevice(c->device));
    gip_batch *b = new (std::nothrow) gip_batch();
    if (!b) return set_error(GIP_ENOMEM, "gip_batch_create: allocation failed");
    b->c = c;
    b->B = ninst;
    b->rho = rho;
    b->o = *opts;
    if (b->o.poll <= 0) b->o.poll = 16;
    // layout
    const int B = ninst;
    b->N.resize(B);
    b->G.resize(B);
    b->F.resize(B);
    b->off_in.resize(B);
    b->off_gh.resize(B);
    b->off_f.resize(B);
    b->off_ec.resize(B);
    std::vector<int64_t> npad(B), fpad(B);
    for (int i = 0; i < B; ++i) {
        b->N[i] = cavs[i]->n_inner;
        b->G[i] = cavs[i]->n_ghost;
        b->F[i] = cavs[i]->n_faces;
        b->off_in[i] = b->NP;
        b->off_gh[i] = b->GH;
        b->off_f[i] = b->FP;
        b->off_ec[i] = b->ECR;
        b->NP += roundup(b->N[i], kChunk);
        b->GH += b->G[i];
        b->FP += roundup(b->F[i], kChunk);
        b->ECR += cavs[i]->Ec.nrows;
        npad[i] = b->N[i];
        fpad[i] = b->F[i];
    }
    b->ALLB = b->NP + b->GH;
    int rc = GIP_OK;
    if (b->ALLB > INT32_MAX / 2 || b->FP > INT32_MAX / 2) rc = set_error(GIP_EINVAL, "gip_batch_create: batch too large");
    if (!rc) rc = make_seg(b, b->off_in, npad, b->NP, b->d_seg_in, &b->seg_in);
    if (!rc) rc = make_seg(b, b->off_f, fpad, b->FP, b->d_seg_f, &b->seg_f);
    auto grid_for = [](int64_t nch) { return (int)std::min<int64_t>(std::max<int64_t>(nch, 1), 4096); };
    b->grid_in = grid_for(b->NP / kChunk);
    b->grid_f = grid_for(b->FP / kChunk);
    b->grid_all = grid_for((b->ALLB + kChunk - 1) / kChunk);
    // fields
    const size_t NA = (size_t)b->ALLB, FF = (size_t)b->FP, NPs = (size_t)b->NP;
    double **cells[] = {&b->u, &b->v, &b->w, &b->ax, &b->ay, &b->az, &b->ax0, &b->ay0, &b->az0,
                        &b->rhs, &b->p, &b->p0, &b->gpx, &b->gpy, &b->gpz, &b->dxs, &b->temp};
    for (double **f : cells)
        if (!rc) rc = dalloc_t(b, f, NA);
    if (!rc) rc = dalloc_t(b, &b->mf, FF);
    double **inst[] = {&b->dt, &b->sigma, &b->time, &b->gamma, &b->massr, &b->dt_used};
    for (double **f : inst)
        if (!rc) rc = dalloc_t(b, f, (size_t)B);
    if (!rc) rc = dalloc_t(b, &b->fac, (size_t)B * F_COUNT);
    // PCG workspace
    double **pw[] = {&b->r, &b->s, &b->Ap, &b->dinv};
    for (double **f : pw)
        if (!rc) rc = dalloc_t(b, f, NPs);
    if (!rc) rc = dalloc_t(b, &b->pp, NA);
    if (!rc && b->o.precision == GIP_PCG_MIXED) {
        float **fw[] = {&b->r32, &b->s32, &b->Ap32, &b->xc32, &b->dinv32};
        for (float **f : fw)
            if (!rc) rc = dalloc_t(b, f, NPs);
        if (!rc) rc = dalloc_t(b, &b->pq, NA);
        if (!rc) rc = dalloc_t(b, &b->cone, NPs);
        if (!rc) rc = dalloc_t(b, &b->caz, NPs);
    }
    if (!rc) rc = dalloc_t(b, &b->sc, (size_t)B * B_COUNT);
    // chunk totals (up to 3 per chunk) or wavefront partials (up to 2 per wavefront)
    if (!rc) rc = dalloc_t(b, &b->part, (size_t)std::max(b->NP, b->FP) / kChunk * (2 * kWaves) + 8);
    if (!rc) rc = dalloc_t(b, &b->state, (size_t)B);
    if (!rc) rc = dalloc_t(b, &b->iters, (size_t)B);
    if (!rc) rc = dalloc_t(b, &b->active, 1);
    if (!rc) {
        hipError_t e = hipHostMalloc((void **)&b->h_active, sizeof(int), hipHostMallocDefault);
        if (e != hipSuccess) rc = hip_fail(e, "hipHostMalloc");
        for (int k = 0; k < 4 && e == hipSuccess; ++k) e = hipEventCreate(&b->ev[k]);
        if (e != hipSuccess && !rc) rc = hip_fail(e, "hipEventCreate");
    }
    // fields from the cavities: cells (inner + ghosts), gamma
    if (!rc) {
        std::vector<double> hu(NA, 0.0), hv(NA, 0.0), hw(NA, 0.0), hd(NA, 0.0);
        for (int i = 0; i < B; ++i) {
            const gip_cavity *cv = cavs[i];
            for (int64_t r = 0; r < b->N[i] + b->G[i]; ++r) {
                int64_t q = r < b->N[i] ? b->off_in[i] + r : b->NP + b->off_gh[i] + (r - b->N[i]);
                hu[(size_t)q] = cv->u[r];
                hv[(size_t)q] = cv->v[r];
                hw[(size_t)q] = cv->w[r];
                hd[(size_t)q] = cv->dxs[r];
            }
        }
        hipStream_t st = c->stream;
        hipError_t e = hipMemcpyAsync(b->u, hu.data(), NA * 8, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(b->v, hv.data(), NA * 8, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(b->w, hw.data(), NA * 8, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(b->dxs, hd.data(), NA * 8, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(b->gamma, gamma, (size_t)B * 8, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = hip_fail(e, "batch field upload");
    }
    // operators (block diagonal), layout chosen by the autotuner when asked
    double *tx = nullptr, *ty = nullptr;
    if (!rc && b->o.autotune) {
        size_t m = (size_t)std::max(std::max(b->ALLB, b->FP), b->ECR);
        rc = dalloc_t(b, &tx, m);
        if (!rc) rc = dalloc_t(b, &ty, m);
    }
    for (int k = 0; k < OP_N && !rc; ++k) {
        std::vector<int32_t> rp, cl;
        std::vector<double> vl;
        int64_t nr = 0;
        assemble(b, cavs, k, rp, cl, vl, nr);
        const int32_t ncols = kCols[k] == COLS_FACES ? (int32_t)b->FP : (int32_t)b->ALLB;
        const int32_t flags = (k == OP_L && b->o.precision == GIP_PCG_MIXED) ? GIP_MAT_FP32 : 0;
        if (b->o.autotune && k != OP_L) {
            rc = gip_mat_autotune(c, (int32_t)nr, ncols, (int32_t)vl.size(), rp.data(), cl.data(), vl.data(), flags, tx,
                                  ty, 5, &b->op[k], nullptr);
        } else {
            rc = gip_mat_create(c, (int32_t)nr, ncols, (int32_t)vl.size(), rp.data(), cl.data(), vl.data(),
                                GIP_FMT_AUTO | flags, &b->op[k]);
        }
        if (!rc && k == OP_D) {
            // C takes D's pattern (GIP_DNS.h:150-151); its values are rebuilt each step
            rc = gip_mat_create(c, (int32_t)nr, ncols, (int32_t)vl.size(), rp.data(), cl.data(), vl.data(),
                                GIP_FMT_AUTO, &b->C);
        }
        if (!rc && k == OP_L) {
            // Jacobi inverse, zero on the padding rows (GIP_MatrixMulticore.cpp:177-210)
            std::vector<double> di((size_t)b->NP, 0.0);
            std::vector<float> di32((size_t)b->NP, 0.0f);
            for (int64_t r = 0; r < nr; ++r)
                for (int32_t j = rp[(size_t)r]; j < rp[(size_t)r + 1]; ++j)
                    if (cl[(size_t)j] == r) di[(size_t)r] = 1.0 / vl[(size_t)j];
            for (size_t r = 0; r < di.size(); ++r) di32[r] = (float)di[r];
            hipError_t e = hipMemcpyAsync(b->dinv, di.data(), di.size() * 8, hipMemcpyHostToDevice, c->stream);
            if (e == hipSuccess && b->dinv32)
                e = hipMemcpyAsync(b->dinv32, di32.data(), di32.size() * 4, hipMemcpyHostToDevice, c->stream);
            // MIXED coarse step: 1 on each instance's inner rows, A 1 in stored
            // order (columns outside the inner rows count 0), 1.A1 per instance
            std::vector<double> one((size_t)b->NP, 0.0), az((size_t)b->NP, 0.0), hsc((size_t)B * B_COUNT, 0.0);
            if (b->cone) {
                for (int i = 0; i < B; ++i)
                    for (int64_t r = 0; r < b->N[i]; ++r) one[(size_t)(b->off_in[i] + r)] = 1.0;
                for (int64_t r = 0; r < b->NP; ++r) {
                    double t = 0.0;
                    for (int32_t j = rp[(size_t)r]; j < rp[(size_t)r + 1]; ++j)
                        t = t + vl[(size_t)j] * (cl[(size_t)j] < b->NP ? one[(size_t)cl[(size_t)j]] : 0.0);
                    az[(size_t)r] = t;
                }
                for (int i = 0; i < B; ++i) {
                    double z = 0.0;
                    for (int64_t r = 0; r < b->N[i]; ++r) z = z + az[(size_t)(b->off_in[i] + r)];
                    hsc[(size_t)i * B_COUNT + B_ZAZ] = z;
                }
                if (e == hipSuccess) e = hipMemcpyAsync(b->cone, one.data(), one.size() * 8, hipMemcpyHostToDevice, c->stream);
                if (e == hipSuccess) e = hipMemcpyAsync(b->caz, az.data(), az.size() * 8, hipMemcpyHostToDevice, c->stream);
                if (e == hipSuccess) e = hipMemcpyAsync(b->sc, hsc.data(), hsc.size() * 8, hipMemcpyHostToDevice, c->stream);
            }
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) rc = hip_fail(e, "batch dinv upload");
        }
    }
    if (!rc && b->op[OP_EC]->nrows != b->C->nnz) rc = set_error(GIP_EINVAL, "gip_batch_create: Ec rows != nnz(D)");
    if (rc) {
        gip_batch_destroy(b);
        return rc;
    }
    *out = b;
    return GIP_OK;
}

int gip_batch_step(gip_batch *b) {
    if (!b) return set_error(GIP_EINVAL, "null batch");
    hipStream_t st = b->c->stream;
    int rc;
    if (!b->started) {  // initial CFL (GIP_DNS.h:101-102)
        k_b_cfl<<<b->grid_in, kBlock, 0, st>>>(b->u, b->v, b->w, b->dxs, b->seg_in, b->gamma, b->rho, b->part);
        k_b_cfl_fin<<<b->B, kBlock, 0, st>>>(b->part, b->seg_in, b->dt, b->sigma);
        GIP_LAUNCH_CHECK("batch initial CFL");
        b->started = true;
    }
    GIP_HIP_TRY(hipEventRecord(b->ev[0], st));
    if ((rc = run_phase(b, enqueue_pre, &b->g_pre))) return rc;
    GIP_HIP_TRY(hipEventRecord(b->ev[1], st));
    const int max_groups = (b->o.pcg_maxit + b->o.poll - 1) / b->o.poll + 2;
    for (int g = 0; g < max_groups; ++g) {
        if ((rc = run_phase(b, enqueue_iter, &b->g_iter))) return rc;
        GIP_HIP_TRY(hipMemcpyAsync(b->h_active, b->active, sizeof(int), hipMemcpyDeviceToHost, st));
        GIP_HIP_TRY(hipStreamSynchronize(st));
        if (*b->h_active <= 0) break;
    }
    GIP_HIP_TRY(hipEventRecord(b->ev[2], st));
    if ((rc = run_phase(b, enqueue_post, &b->g_post))) return rc;
    GIP_HIP_TRY(hipEventRecord(b->ev[3], st));
    GIP_HIP_TRY(hipStreamSynchronize(st));
    for (int k = 0; k < 3; ++k) GIP_HIP_TRY(hipEventElapsedTime(&b->phase_ms[k], b->ev[k], b->ev[k + 1]));
    return GIP_OK;
}

int gip_batch_phase_ms(gip_batch *b, float *ms) {
    if (!b || !ms) return set_error(GIP_EINVAL, "gip_batch_phase_ms: bad argument");
    for (int k = 0; k < 3; ++k) ms[k] = b->phase_ms[k];
    return GIP_OK;
}

int gip_batch_instance(gip_batch *b, int32_t i, gip_batch_info *info) {
    if (!b || !info || i < 0 || i >= b->B) return set_error(GIP_EINVAL, "gip_batch_instance: bad argument");
    hipStream_t st = b->c->stream;
    double t = 0, d = 0, m = 0;
    int it = 0, s = 0;
    GIP_HIP_TRY(hipMemcpyAsync(&t, b->time + i, 8, hipMemcpyDeviceToHost, st));
    GIP_HIP_TRY(hipMemcpyAsync(&d, b->dt_used + i, 8, hipMemcpyDeviceToHost, st));
    GIP_HIP_TRY(hipMemcpyAsync(&m, b->massr + i, 8, hipMemcpyDeviceToHost, st));
    GIP_HIP_TRY(hipMemcpyAsync(&it, b->iters + i, 4, hipMemcpyDeviceToHost, st));
    GIP_HIP_TRY(hipMemcpyAsync(&s, b->state + i, 4, hipMemcpyDeviceToHost, st));
    GIP_HIP_TRY(hipStreamSynchronize(st));
    info->time = t;
    info->dt = d;
    info->massr = m;
    info->pcg_iters = it;
    info->pcg_status = s == BS_CONV ? GIP_OK : s == BS_BREAK ? GIP_EBREAKDOWN : GIP_ENOCONV;
    info->n_inner = (int32_t)b->N[i];
    info->n_all = (int32_t)(b->N[i] + b->G[i]);
    info->n_faces = (int32_t)b->F[i];
    return GIP_OK;
}

int gip_batch_get_field(gip_batch *b, int32_t i, const char *field, double *host, int32_t n) {
    if (!b || !field || !host || i < 0 || i >= b->B) return set_error(GIP_EINVAL, "gip_batch_get_field: bad argument");
    hipStream_t st = b->c->stream;
    if (!std::strcmp(field, "mf")) {
        if (n != b->F[i]) return set_error(GIP_EINVAL, "gip_batch_get_field: length mismatch");
        GIP_HIP_TRY(hipMemcpyAsync(host, b->mf + b->off_f[i], (size_t)n * 8, hipMemcpyDeviceToHost, st));
        GIP_HIP_TRY(hipStreamSynchronize(st));
        return GIP_OK;
    }
    const double *src = !std::strcmp(field, "u") ? b->u : !std::strcmp(field, "v") ? b->v : !std::strcmp(field, "w") ? b->w
                        : !std::strcmp(field, "p") ? b->p : nullptr;
    if (!src) return set_error(GIP_EINVAL, "gip_batch_get_field: unknown field");
    if (n != b->N[i] + b->G[i]) return set_error(GIP_EINVAL, "gip_batch_get_field: length mismatch");
    GIP_HIP_TRY(hipMemcpyAsync(host, src + b->off_in[i], (size_t)b->N[i] * 8, hipMemcpyDeviceToHost, st));
    GIP_HIP_TRY(hipMemcpyAsync(host + b->N[i], src + b->NP + b->off_gh[i], (size_t)b->G[i] * 8, hipMemcpyDeviceToHost, st));
    GIP_HIP_TRY(hipStreamSynchronize(st));
    return GIP_OK;
}

int gip_batch_layouts(gip_batch *b, int32_t *format, int32_t *lanes) {
    if (!b || !format || !lanes) return set_error(GIP_EINVAL, "gip_batch_layouts: bad argument");
    for (int k = 0; k < OP_N; ++k) {
        format[k] = b->op[k]->format;
        lanes[k] = b->op[k]->format == GIP_FMT_ELL ? 1 << b->op[k]->ell_lg : b->op[k]->lanes;
    }
    return GIP_OK;
}

int gip_batch_destroy(gip_batch *b) {
    if (!b) return GIP_OK;
    hipSetDevice(b->c->device);
    hipStreamSynchronize(b->c->stream);
    if (b->g_pre) hipGraphExecDestroy(b->g_pre);
    if (b->g_iter) hipGraphExecDestroy(b->g_iter);
    if (b->g_post) hipGraphExecDestroy(b->g_post);
    for (int k = 0; k < OP_N; ++k)
        if (b->op[k]) gip_mat_destroy(b->op[k]);
    if (b->C) gip_mat_destroy(b->C);
    for (void *p : b->allocs) hipFree(p);
    if (b->h_active) hipHostFree(b->h_active);
    for (hipEvent_t e : b->ev)
        if (e) hipEventDestroy(e);
    delete b;
    return GIP_OK;
}

}  // extern "C"
