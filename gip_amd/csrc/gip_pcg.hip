// Device-resident Jacobi-PCG of libgip_hip (GIP_LinearSolver.h:45-117).
//
// FP64 iteration k, the reference recurrences and roundings
// (GIP_LinearSolver.h:85-116):
//   direction  x += alpha_{k-1} p (the previous iteration's update, :95,
//              moved here so that p is read once per iteration);
//              beta = rho_k / rho_{k-1}; p = 1.0 s + beta p   (:107-113)
//   product    Ap = A p, partials of p.Ap                     (:87-91)
//   update     alpha = rho_k / p.Ap; r += (-alpha) Ap; s = Dinv r;
//              partials of r.s and r.r                        (:93-105)
// Operators of at most 2^19 rows (2048 chunks of 256 rows) take two
// kernels per iteration: the direction is formed inside the product, and
// every workgroup of a kernel re-reduces the previous kernel's partials in
// its prologue (fence-free: plain stores, ordered by the kernel boundary;
// every workgroup computes bit-identical scalars), so a kernel costs about
// one memory round trip before it stores.  Up to kRowsMin rows, four
// kernels: grid-stride sweeps whose prologues re-reduce the partials, the
// product's partials folded first.  Longer operators sweep one row per
// thread in dispatch order (the pace of a plain stream of each kernel's
// access mix) and take their scalars from one-workgroup kernels (fold +
// finalize) instead, since that many workgroups cannot each re-reduce the
// partials; see the long-path section below.  Scalars that must survive a
// kernel (rho and alpha of the last two iterations, beta, the state word)
// are written by one workgroup: rho and alpha into the slots of their
// iteration's parity, which no workgroup of the same kernel reads.  Every
// kernel reads the solver's state word and returns when the solve has
// finished, so the host queues many iterations and polls the state only now
// and then.
//
// MIXED: the same iteration in FP32 (operator values, vectors and SpMV;
// dot products accumulate in FP64; on long operators three kernels whose
// grid-stride sweeps re-reduce the partials in their prologues: the
// row-per-thread form measured 18% slower per FP32 iteration at 256^3)
// on the correction equation, inside an
// FP64 defect-correction loop: when the FP32 residual has dropped by
// inner_tol (or reaches the target) the FP32 correction is added to the
// FP64 x, the FP64 true residual b - A x replaces the FP32 residual (the
// search direction is kept: reliable-update residual replacement, restarted
// with beta = 0 when the recursive residual had drifted more than 10x from
// the true one), and convergence is decided on that FP64 residual only.
// Each correction also takes the FP64 Galerkin step along the constant
// vector 1 of the solved rows, x += c 1 with c = (1.r)/(1.A1): for the
// pinned Neumann pressure Laplacian that vector is the near-null mode
// (eigenvalue ~ pin/N), which FP32 arithmetic cannot resolve; left to the
// FP32 iteration it re-enters through round-off and the corrections
// diverge.  For an SPD operator the step never increases the A-norm error.
// The FP32 iteration is deflated by the same vector (A p replaced by
// A p - A1 (1.Ap)/(1.A1)), so its residual stays orthogonal to 1 and the
// constant mode is left to the FP64 step.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <new>
#include <vector>

#include "gip_blas1.h"
#include "gip_internal.h"
#include "gip_rows.h"
#include "gip_spmv.h"

namespace gip {

// Device scalar slots of a solver.
enum {
    S_RHO0 = 0,     // rho of even iterations
    S_RHO1 = 1,     // rho of odd iterations
    S_RR = 3,       // r.r (FP64: last update; MIXED: last status or true residual)
    S_BB = 4,       // b.b
    S_TOL2 = 6,     // (tol ||b||)^2, negative when no tolerance
    S_TRUE_RR = 7,  // FP64 true residual of the last correction (MIXED)
    S_ZAZ = 8,      // 1.A1 over the solved rows (MIXED coarse step)
    S_ALPHA0 = 10,  // alpha of even iterations, as the update kernel formed it
    S_ALPHA1 = 11,  // alpha of odd iterations
    S_BETA = 12,    // beta of the next direction update (FP64, long operators)
    S_COUNT = 16
};
// State word: ST_RUN, ST_BREAK, or ST_CONV; the direction kernel of
// iteration k writes ST_CONV | k << 2, so readers look at the low two bits.
enum { ST_RUN = 0, ST_CONV = 1, ST_BREAK = 2, ST_MASK = 3 };
// Partial-sum arrays, kMaxRedBlocks doubles each.
// P_DN2: p.Ap partials of odd iterations on operators small enough for the
// fused direction-product kernel (which reads the previous iteration's
// partials while writing its own).
// P_SA, P_SA2: partials of 1.Ap of the mixed inner iteration (the constant
// vector's deflation), by iteration parity on small operators.
enum { P_DN = 0, P_RS = 1, P_RR = 2, P_BB = 3, P_RS2 = 4, P_RR2 = 5, P_SUMR = 6, P_DN2 = 7, P_SA = 8, P_SA2 = 9,
       P_COUNT = 10 };

// The nonzero entries of A 1 in FP32, when there are at most kAzList of
// them (count > kAzList: too many, the kernels read A 1 row by row).  A
// pinned Neumann Laplacian has zero row sums but at its pinned row, so its
// A 1 has one nonzero and the long update kernel need not stream A 1.
constexpr int kAzList = 8;
struct AzList {
    int count;
    int idx[kAzList];
    float val[kAzList];
};

struct PcgDev {
    int n, G;      // solved rows, workgroups of the sweeping kernels (<= kMaxRedBlocks)
    int nchunk;    // 256-row chunks: workgroups of the product kernels
    int Gdn;       // p.Ap partials the next kernels read (<= kMaxRedBlocks)
    double *chunk; // partials of the product: with the fold, kWaves per
                   // 256-row chunk (one per wavefront); without it, one per
                   // chunk, and these are the P_DN partials themselves
    double *chunk_s;  // long operators: a second array of per-wavefront partials (FP64: r.r; MIXED: 1.Ap)
    double *x, *r, *s, *Ap;
    const double *b, *dinv;
    float *r32, *s32, *Ap32, *xc32;
    const float *dinv32;
    const float *az32;  // A 1 in FP32 (MIXED deflation)
    const struct AzList *azl;  // its nonzeros, when few (MIXED deflation)
    int deflate;        // MIXED: deflate the constant vector in the inner iteration
    const double *az;  // A 1 on the solved rows (MIXED)
    double *sc;        // S_COUNT scalars
    int *state, *iters;
    double *part;  // P_COUNT x kMaxRedBlocks partials; each array has one
                   // producer grid size and stays zero past it
    int precond;
    __device__ double *P(int k) const { return part + k * kMaxRedBlocks; }
};

// The state word is written by workgroup 0 of a kernel while the other
// workgroups of that kernel may be reading it, so every access is an atomic
// one (relaxed, device scope): a reader sees either the word an earlier
// kernel left (ordered by the kernel boundary) or the one workgroup 0 of
// its own kernel is writing.  Each writer makes both outcomes give the same
// result: the direction kernel stamps a convergence with its iteration, and
// a workgroup that reads its own kernel's stamp takes the convergence path
// every workgroup computes; a breakdown leaves the iterates undefined, so
// skipping work is harmless.  An atomic load is an ordered memory
// reference: the compiler issues it where it is written, together with the
// kernel's other early loads.
__device__ inline int state_load(const int *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline void state_store(int *p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline bool running(const PcgDev &d) { return state_load(d.state) == ST_RUN; }
// A word read at the top of a kernel and used only after its prologue: an
// atomic load (workgroup scope, relaxed: a plain load instruction) is an
// ordered memory reference, so the compiler issues it where it is written,
// together with the other loads, instead of sinking it to its first use.
__device__ inline double early_load(const double *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline int conv_stamp(int k) { return ST_CONV | (k << 2); }
__device__ inline bool leader() { return blockIdx.x == 0 && threadIdx.x == 0; }

// v of lane i - (CTRL - 0x110) within the lane's row of 16 (0 past the
// row's start): a DPP row shift, moved as two 32-bit halves
template <int CTRL>
__device__ inline double dpp_row_shr(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ inline double read_lane(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum over the wavefront with DPP row shifts (an inclusive scan of each row
// of 16 lanes) and four lane reads, the same value in every lane; a fixed
// order, but not wave_sum's.  The wavefront partials of the long
// iterations' products and of the FP64 update use it (FP64 solve at 256^3
// 1163 -> 1155 ms against the shuffle sum).
__device__ inline double wave_sum_dpp(double v) {
    v = v + dpp_row_shr<0x111>(v);
    v = v + dpp_row_shr<0x112>(v);
    v = v + dpp_row_shr<0x114>(v);
    v = v + dpp_row_shr<0x118>(v);
    return ((read_lane(v, 15) + read_lane(v, 31)) + read_lane(v, 47)) + read_lane(v, 63);
}
__device__ inline double rho_slot(const PcgDev &d, int k) { return d.sc[S_RHO0 + (k & 1)]; }
// p.Ap partials of iteration k in the fused (small-operator) iteration
__device__ inline double *dn_part(const PcgDev &d, int k) { return d.P((k & 1) ? P_DN2 : P_DN); }
// 1.Ap partials of iteration k of the mixed fused iteration
__device__ inline double *sa_part(const PcgDev &d, int k) { return d.P((k & 1) ? P_SA2 : P_SA); }

// Deflation of the constant vector in the mixed inner iteration (deflated
// CG, Saad, Yeung, Erhel and Guyomarc'h, SIAM J. Sci. Comput. 21, 2000,
// with W = 1): the iteration uses w = P A p = A p - A1 (1.Ap)/(1.A1) in
// place of A p, so alpha = rho / (p.Ap - (1.Ap)^2 / (1.A1)) (1.Ap = A1.p:
// A symmetric on the solved rows).  The defect correction already applies
// the matching coarse step (x += c 1).  Off when 1.A1 is not positive and
// finite, like the coarse step.
__device__ inline bool deflating(const PcgDev &d, double zaz) { return d.deflate && zaz > 0.0 && isfinite(zaz); }
__device__ inline double m_denom(double dn, double sap, double zaz, bool dfl) {
    return dfl ? dn - sap * sap / zaz : dn;
}

// ---- setup ---------------------------------------------------------------

// x = x0 (all), direction buffers = x0 outside the solved rows,
// Ap = (A x0)_i or x0_i on empty rows, r = b + (-1.0) Ap   (:55-70)
template <int W>
__global__ __launch_bounds__(kBlock) void k_setup_resid(EllView A, PcgDev d, const double *__restrict__ x0, int n_all,
                                                        double *__restrict__ p, double *__restrict__ p2) {
    const int stride = gridDim.x * kBlock;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n_all; i += stride) {
        d.x[i] = x0[i];
        if (i >= d.n) {
            p[i] = x0[i];
            if (p2) p2[i] = x0[i];
        }
    }
    Operand<double> xo{x0};
    FOR_ROWS(d.n, row) {
        double sum;
        bool any = row_dot<W, double, double>(A, A.val, xo, row, sum);
        double t = any ? sum : x0[row];
        d.Ap[row] = t;
        d.r[row] = d.b[row] + -1.0 * t;
    }
}

// p = Dinv r (or r); partials of r.p, b.b, r.r  (:72-80)
__global__ __launch_bounds__(kBlock) void k_setup_dir(PcgDev d, double *__restrict__ p) {
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
    double red[3] = {rs, bb, rr};
    block_reduce_n<3>(red);
    if (threadIdx.x == 0) {
        d.P(P_RS)[blockIdx.x] = red[0];
        d.P(P_BB)[blockIdx.x] = red[1];
        d.P(P_RR)[blockIdx.x] = red[2];
    }
}

// one workgroup: rho_0, b.b, r.r, the tolerance, the state word and the iteration count
__global__ __launch_bounds__(kBlock) void k_setup_fin(PcgDev d, double tol) {
    const double rho = sum_partials<RED_SUM>(d.P(P_RS), d.G);
    const double bb = sum_partials<RED_SUM>(d.P(P_BB), d.G);
    const double rr = sum_partials<RED_SUM>(d.P(P_RR), d.G);
    if (threadIdx.x == 0) {
        const double tol2 = tol > 0.0 ? tol * tol * bb : -1.0;
        d.sc[S_RHO0] = rho;
        d.sc[S_BB] = bb;
        d.sc[S_RR] = rr;
        d.sc[S_TOL2] = tol2;
        state_store(d.state, (tol > 0.0 && rr <= tol2) ? ST_CONV : ST_RUN);
        *d.iters = 0;
    }
}

// ---- FP64 iteration --------------------------------------------------------

// alpha of iteration k, stored by workgroup 0 of K2 of k (every workgroup
// of K2 forms the same value from the same partials), so that a later
// kernel reads one scalar instead of re-reducing the p.Ap partials
__device__ inline double alpha_slot(const PcgDev &d, int k) { return d.sc[S_ALPHA0 + (k & 1)]; }
__device__ inline void store_alpha(const PcgDev &d, int k, double alpha) {
    if (leader()) d.sc[S_ALPHA0 + (k & 1)] = alpha;
}

// x += alpha_k p over the solved rows
__device__ inline void x_update(const PcgDev &d, const double *__restrict__ p, double alpha) {
    double *__restrict__ x = d.x;
    FOR_ROWS(d.n, row) {
        const double xo = x[row], po = p[row];
        x[row] = xo + alpha * po;  // :95
    }
}

// A product operand that also keeps the value it reads at the row's own
// column (the p_i of p.Ap, which the row's diagonal entry reads anyway), so
// that the product does not load it a second time (FP64 product: 1.8%
// faster at 256^3; the FP32 product keeps the separate load, which measured
// faster there).
template <class T>
struct SelfOperand {
    const T *__restrict__ p;
    int row;
    mutable T self;
    mutable bool found;
    __device__ inline T at(int j) const {
        const T v = p[j];
        if (j == row) {
            self = v;
            found = true;
        }
        return v;
    }
};

// Kernel prologues.  SMALL (operators of at most 2^19 rows, where every
// thread holds at most one row and a kernel lasts a few microseconds): the
// state word, the scalars, the partials and the row's operands are requested
// together, so the kernel spends one memory round trip before its first
// store, and the direction update is fused into the product (two kernels
// per iteration).  Otherwise (long kernels, where that latency is hidden by
// the sweep) the state word is tested first, with a scalar load, and the
// partial arrays are summed one at a time, which keeps the register count,
// and with it the occupancy of the sweep, low.

// x += alpha_k p of the last iteration k
__global__ __launch_bounds__(kBlock) void k_x_flush(PcgDev d, const double *__restrict__ p, int k) {
    if (!running(d)) return;
    x_update(d, p, alpha_slot(d, k));
}

// K1: Ap = A p (empty rows keep Ap), p.Ap partials; one row per thread.
// Long operators: workgroups in dispatch order, one partial per wavefront
// (with the direction and update sweeps in dispatch order too, 2% faster
// per FP64 iteration at 256^3 than the XCD-contiguous order, which keeps
// the neighbour gathers in one XCD's L2).  SMALL: XCD-contiguous order,
// one partial per chunk.  Both test the state word after the row's loads
// instead of before them, taking its latency off the kernel's critical
// path (FP64 solve at 256^3: 1256 -> 1238 ms); a finished solve then still
// reads the operator, but stores nothing.  (The same reordering in the
// long direction and update sweeps made the solve slower: 1303 ms.)
// SYM: the slot values from the operator's symmetric form (stencil layout).
template <int W, bool SMALL, bool SYM = false>
__device__ inline void iter_spmv(const EllView &A, const PcgDev &d, const double *__restrict__ p) {
    if constexpr (!SMALL) {
        const int st = state_load(d.state);  // tested after the row's loads, before any store
        const int c = blockIdx.x;
        const int row = c * kBlock + threadIdx.x;
        double acc = 0.0, sum = 0.0, pr = 0.0;
        bool any = false;
        if (row < d.n) {
            const SelfOperand<double> po{p, row, 0.0, false};
            if constexpr (SYM)
                any = row_dot_sym<W, double, double>(A, A.val, A.sym, po, row, sum);
            else
                any = row_dot<W, double, double>(A, A.val, po, row, sum);
            if (!any) sum = d.Ap[row];  // empty row: Ap keeps its value
            pr = po.found ? po.self : p[row];
        }
        if (st != ST_RUN) return;
        if (row < d.n) {
            if (any) d.Ap[row] = sum;
            acc = pr * sum;
        }
        // one partial per wavefront: no barrier, each wavefront retires alone
        acc = wave_sum_dpp(acc);
        if ((threadIdx.x & 63) == 0) d.chunk[c * kWaves + (threadIdx.x >> 6)] = acc;
        return;
    }
    const int st = state_load(d.state);
    const int c = xcd_block(blockIdx.x, gridDim.x);
    const int row = c * kBlock + threadIdx.x;
    double pr = 0.0, sum = 0.0;
    bool any = false;
    if (row < d.n) {
        Operand<double> po{p};
        pr = p[row];
        any = row_dot<W, double, double>(A, A.val, po, row, sum);
        if (!any) sum = d.Ap[row];  // empty row: Ap keeps its value
    }
    if (st != ST_RUN) return;
    if (any) d.Ap[row] = sum;
    double acc = row < d.n ? pr * sum : 0.0;
    acc = block_reduce<RED_SUM>(acc);
    if (threadIdx.x == 0) d.chunk[c] = acc;
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_iter_spmv(EllView A, PcgDev d, const double *__restrict__ p) {
    iter_spmv<W, false>(A, d, p);
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_iter_spmv_small(EllView A, PcgDev d, const double *__restrict__ p) {
    iter_spmv<W, true>(A, d, p);
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_iter_spmv_sym(EllView A, PcgDev d, const double *__restrict__ p) {
    iter_spmv<W, false, true>(A, d, p);
}

// The direction p_k = 1.0 s + beta p_{k-1} read at any column j of the
// product, computed exactly as the direction update computes it on the
// solved rows (:113, no contraction: the same bits); past the solved rows
// (ghost columns) the direction keeps its value.
struct DirOperand {
    const double *__restrict__ s, *__restrict__ p;
    double beta;
    int n;
    __device__ inline double at(int j) const {
        const double po = p[j], si = s[j < n ? j : 0];
        return j < n ? 1.0 * si + beta * po : po;
    }
};

// Small operators, k >= 1: K0 and K1 in one kernel over the product's grid
// (one row per thread).  rho_k, the convergence test, x += alpha_{k-1} p_{k-1}
// on the thread's row, p_k = 1.0 s + beta p_{k-1} into the other direction
// buffer, Ap = A p_k with p_k formed at every column read, and the p_k.Ap
// partials of iteration k (read by K2 of k while this kernel's own
// partials of k-1 stay intact: two partial arrays by parity).
// TOL: a solve with a tolerance (the r.r partials and the convergence
// test); without one (the reference's fixed iteration count) the prologue
// sums the rho partials only.
template <int W, bool TOL>
__device__ inline void iter_dirprod(const EllView &A, const PcgDev &d, const double *__restrict__ pold,
                                    double *__restrict__ pnew, int k) {
    const int st = state_load(d.state);
    const int c = xcd_block(blockIdx.x, gridDim.x);
    const int row = c * kBlock + threadIdx.x;
    const bool own = row < d.n;
    double xo = 0.0, si = 0.0, po = 0.0;
    if (own) {
        xo = d.x[row];
        si = d.s[row];
        po = pold[row];
    }
    const double tol2 = TOL ? early_load(d.sc + S_TOL2) : -1.0;
    const double rho_prev = early_load(d.sc + S_RHO0 + ((k - 1) & 1));
    const double alpha = early_load(d.sc + S_ALPHA0 + ((k - 1) & 1));  // alpha_{k-1}, as K2 of k-1 formed it (:93)
    constexpr int NS = TOL ? 2 : 1;
    PaddedSums<NS> ps;
    if constexpr (TOL)
        ps.load({d.P(P_RS), d.P(P_RR)}, {d.G, d.G});
    else
        ps.load({d.P(P_RS)}, {d.G});
    // stencil layout: the operator's loads (none depends on beta) go out
    // while the prologue's are in flight
    constexpr int SW = W < 0 ? -W : 1;
    StencilPre<SW, double> q;
    if (W < 0 && own) stencil_row_pre<SW, double>(A, A.val, row, q);
    double sums[NS];
    ps.reduce(sums);
    if (st != ST_RUN && st != conv_stamp(k)) return;
    if (TOL && tol2 >= 0.0 && sums[NS - 1] <= tol2) {  // every workgroup reaches the same decision
        if (own) d.x[row] = xo + alpha * po;
        if (leader()) state_store(d.state, conv_stamp(k));
        return;
    }
    const double beta = sums[0] / rho_prev;  // :107-109
    if (leader()) d.sc[S_RHO0 + (k & 1)] = sums[0];
    double pr = 0.0, sum = 0.0;
    if (own) {
        d.x[row] = xo + alpha * po;  // :95 of iteration k-1
        pr = 1.0 * si + beta * po;   // :113
        pnew[row] = pr;
        const DirOperand op{d.s, pold, beta, d.n};
        bool any;
        if constexpr (W < 0)
            any = stencil_row_fin<SW, double, double>(A, A.val, op, row, q, sum);  // :87
        else
            any = row_dot<W, double, double>(A, A.val, op, row, sum);
        if (any)
            d.Ap[row] = sum;
        else
            sum = d.Ap[row];  // empty row: Ap keeps its value
    }
    double acc = own ? pr * sum : 0.0;
    acc = block_reduce<RED_SUM>(acc);
    if (threadIdx.x == 0) dn_part(d, k)[c] = acc;
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_iter_dirprod(EllView A, PcgDev d, const double *__restrict__ pold,
                                                         double *__restrict__ pnew, int k) {
    iter_dirprod<W, true>(A, d, pold, pnew, k);
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_iter_dirprod_fixed(EllView A, PcgDev d, const double *__restrict__ pold,
                                                               double *__restrict__ pnew, int k) {
    iter_dirprod<W, false>(A, d, pold, pnew, k);
}

// the per-wavefront partials of the long products, folded in index order:
// thread i sums chunk i's kWaves partials, each workgroup its 256 chunks
// into one of the Gdn partials of P_DN
__global__ __launch_bounds__(kBlock) void k_fold_dn(PcgDev d) {
    if (!running(d)) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v = 0.0, sa = 0.0;
    if (i < d.nchunk) {
        const double *w = d.chunk + (size_t)i * kWaves;
#pragma unroll
        for (int j = 0; j < kWaves; ++j) v = v + w[j];
        if (d.deflate) {
            const double *ws = d.chunk_s + (size_t)i * kWaves;
#pragma unroll
            for (int j = 0; j < kWaves; ++j) sa = sa + ws[j];
        }
    }
    if (d.deflate) {
        double red[2] = {v, sa};
        block_reduce_n<2>(red);
        if (threadIdx.x == 0) {
            d.P(P_DN)[blockIdx.x] = red[0];
            d.P(P_SA)[blockIdx.x] = red[1];
        }
    } else {
        v = block_reduce<RED_SUM>(v);
        if (threadIdx.x == 0) d.P(P_DN)[blockIdx.x] = v;
    }
}

// K2 (operators of at most 2^19 rows): alpha_k, r += (-alpha) Ap, s = Dinv r,
// partials of r.s and r.r; each row's operands are loaded together (one
// round trip per row), requested before the prologue's sums
__global__ __launch_bounds__(kBlock) void k_iter_update(PcgDev d, int k) {
    double *__restrict__ r = d.r, *__restrict__ s = d.s;
    const double *__restrict__ Ap = d.Ap, *__restrict__ dinv = d.dinv;
    const bool precond = d.precond != 0;
    double rs = 0.0, rr = 0.0;
    {
        ThreadRows t(d.n);
        double ro = 0.0, apr = 0.0, di = 0.0;
        const int st = state_load(d.state);
        if (t.valid()) {
            ro = r[t.row];
            apr = Ap[t.row];
            if (precond) di = dinv[t.row];
        }
        const double rho = early_load(d.sc + S_RHO0 + (k & 1));
        double sums[1];
        sum_padded<1>({dn_part(d, k)}, {d.Gdn}, sums);
        const double dn = sums[0];
        if (st != ST_RUN) return;
        const double alpha = rho / dn;
        if (dn == 0.0 || !isfinite(dn) || !isfinite(alpha)) {
            if (leader()) state_store(d.state, ST_BREAK);
            return;
        }
        store_alpha(d, k, alpha);
        const double malpha = -1.0 * alpha;
        while (t.valid()) {
            const int row = t.row;
            const double ri = ro + malpha * apr;  // :97
            r[row] = ri;
            double si;
            if (precond) {  // :99-105
                si = 0.0;
                si = si + ri * di;
            } else {
                si = ri;
            }
            s[row] = si;
            rs = rs + ri * si;
            rr = rr + ri * ri;
            t.next();
            if (t.valid()) {
                ro = r[t.row];
                apr = Ap[t.row];
                if (precond) di = dinv[t.row];
            }
        }
    }
    double red[2] = {rs, rr};
    block_reduce_n<2>(red);
    if (threadIdx.x == 0) {
        d.P(P_RS)[blockIdx.x] = red[0];
        d.P(P_RR)[blockIdx.x] = red[1];
    }
    if (leader()) *d.iters = k + 1;
}

// ---- FP64 iteration on mid-size operators (2^19 to kRowsMin rows) ----------
//
// A sweep of 2048 (or fewer) workgroups striding over the rows in
// XCD-contiguous order, whose prologue re-reduces the previous kernel's
// partials (every workgroup forms the same scalars): four launches per
// iteration (direction, product, fold, update).  At these sizes the three
// extra launches of the row-per-thread iteration cost more than its faster
// sweeps save (FP64 solve to 1e-10, grid-stride against row-per-thread:
// 128^3 94.9 against 107.0 ms, 160^3 205.8 / 227.7, 192^3 419.2 / 424.6;
// past the switch 224^3 834.0 / 776.4, 256^3 1418 / 1286).
constexpr int kRowsMin = 1 << 23;

// K0 (k >= 1): rho_k, the convergence test, x += alpha_{k-1} p,
// beta, p = 1.0 s + beta p
__global__ __launch_bounds__(kBlock) void k_iter_dir(PcgDev d, double *__restrict__ p, int k) {
    double *__restrict__ x = d.x;
    const double *__restrict__ s = d.s;
    const int st = state_load(d.state);
    if (st != ST_RUN && st != conv_stamp(k)) return;
    const double rho = sum_partials<RED_SUM>(d.P(P_RS), d.G);  // :107
    const double alpha = alpha_slot(d, k - 1);
    const double tol2 = d.sc[S_TOL2];
    if (tol2 >= 0.0) {
        const double rr = sum_partials<RED_SUM>(d.P(P_RR), d.G);
        if (rr <= tol2) {  // every workgroup reaches the same decision
            x_update(d, p, alpha);
            if (leader()) state_store(d.state, conv_stamp(k));
            return;
        }
    }
    const double beta = rho / rho_slot(d, k - 1);  // :109
    if (leader()) d.sc[S_RHO0 + (k & 1)] = rho;
    FOR_ROWS(d.n, row) {
        const double xo = x[row], si = s[row], po = p[row];
        x[row] = xo + alpha * po;       // :95 of iteration k-1
        p[row] = 1.0 * si + beta * po;  // :113
    }
}

// K2: alpha_k from the folded p.Ap partials, r += (-alpha) Ap, s = Dinv r,
// partials of r.s and r.r (one per workgroup); each row's operands are
// loaded together (one round trip per row)
__global__ __launch_bounds__(kBlock) void k_iter_update_grid(PcgDev d, int k) {
    if (!running(d)) return;
    double *__restrict__ r = d.r, *__restrict__ s = d.s;
    const double *__restrict__ Ap = d.Ap, *__restrict__ dinv = d.dinv;
    const bool precond = d.precond != 0;
    double dsum[1];
    if (d.Gdn <= kBlock)  // one round of the uniform loop
        sum_partials_n<1>({d.P(P_DN)}, d.Gdn, dsum);
    else
        dsum[0] = sum_partials<RED_SUM>(d.P(P_DN), d.Gdn);
    const double dn = dsum[0];
    const double alpha = rho_slot(d, k) / dn;  // :93
    if (dn == 0.0 || !isfinite(dn) || !isfinite(alpha)) {
        if (leader()) state_store(d.state, ST_BREAK);
        return;
    }
    store_alpha(d, k, alpha);
    const double malpha = -1.0 * alpha;
    double rs = 0.0, rr = 0.0;
    FOR_ROWS(d.n, row) {
        const double ro = r[row], apr = Ap[row];
        const double di = precond ? dinv[row] : 0.0;
        const double ri = ro + malpha * apr;  // :97
        r[row] = ri;
        double si;
        if (precond) {  // :99-105
            si = 0.0;
            si = si + ri * di;
        } else {
            si = ri;
        }
        s[row] = si;
        rs = rs + ri * si;
        rr = rr + ri * ri;
    }
    double red[2] = {rs, rr};
    block_reduce_n<2>(red);
    if (threadIdx.x == 0) {
        d.P(P_RS)[blockIdx.x] = red[0];
        d.P(P_RR)[blockIdx.x] = red[1];
    }
    if (leader()) *d.iters = k + 1;
}

// ---- FP64 iteration on long operators (more than kRowsMin rows) -------------
//
// One row per thread over the whole operator, workgroups in dispatch order
// (consecutive 256-row chunks on consecutive XCDs, the whole chip streaming
// one narrow window of every array): the direction and update sweeps run
// at the pace of a plain stream of their access mix.  So many workgroups
// cannot each re-reduce the previous kernel's partials, so the scalars come
// from one-workgroup kernels instead:
//   K0  x += alpha_{k-1} p, p = 1.0 s + beta p        (beta, alpha from slots)
//   K1  Ap = A p, per-wavefront p.Ap partials          (iter_spmv)
//   fold + F1: alpha = rho_k / (p.Ap)                  (breakdown test)
//   K2  r += (-alpha) Ap, s = Dinv r, per-wavefront r.s and r.r partials
//   fold + F2: rho_{k+1}, r.r, the convergence test, beta
// F2 stamps a convergence with the next iteration's index; that
// iteration's K0 (or the final flush) then applies the last x update.

__device__ inline double ntload(const double *p) { return __builtin_nontemporal_load(p); }
__device__ inline void ntstore(double *p, double v) { __builtin_nontemporal_store(v, p); }

// K0 of iteration k >= 1
__global__ __launch_bounds__(kBlock) void k_iter_dir_rows(PcgDev d, double *__restrict__ p, int k) {
    const int st = state_load(d.state);
    const int row = blockIdx.x * kBlock + threadIdx.x;
    if (row >= d.n) return;
    const double alpha = alpha_slot(d, k - 1);
    double *__restrict__ x = d.x;
    if (st == conv_stamp(k)) {  // converged after iteration k-1: its x update only
        x[row] = x[row] + alpha * p[row];  // :95
        return;
    }
    if (st != ST_RUN) return;
    const double beta = d.sc[S_BETA];
    // nontemporal accesses: each array is streamed once per sweep (a plain
    // stream of this mix runs 8% faster with them on this device)
    const double xo = ntload(x + row), si = ntload(d.s + row), po = ntload(p + row);
    ntstore(x + row, xo + alpha * po);       // :95 of iteration k-1
    ntstore(p + row, 1.0 * si + beta * po);  // :113
}

// the per-wavefront partials of a and b, folded in index order into the
// Gdn partials of arrays ia and ib (b may be null)
__global__ __launch_bounds__(kBlock) void k_fold_two(PcgDev d, const double *__restrict__ a,
                                                    const double *__restrict__ b, int ia, int ib) {
    if (!running(d)) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double va = 0.0, vb = 0.0;
    if (i < d.nchunk) {
#pragma unroll
        for (int j = 0; j < kWaves; ++j) va = va + a[(size_t)i * kWaves + j];
        if (b)
#pragma unroll
            for (int j = 0; j < kWaves; ++j) vb = vb + b[(size_t)i * kWaves + j];
    }
    double red[2] = {va, vb};
    block_reduce_n<2>(red);
    if (threadIdx.x == 0) {
        d.P(ia)[blockIdx.x] = red[0];
        if (b) d.P(ib)[blockIdx.x] = red[1];
    }
}

// F1 (one workgroup): alpha_k = rho_k / p.Ap, or a breakdown
__device__ inline void alpha_finish(const PcgDev &d, int k, double dn) {
    const double alpha = rho_slot(d, k) / dn;  // :93
    if (dn == 0.0 || !isfinite(dn) || !isfinite(alpha))
        state_store(d.state, ST_BREAK);
    else
        d.sc[S_ALPHA0 + (k & 1)] = alpha;
}

__global__ __launch_bounds__(kBlock) void k_alpha_fin(PcgDev d, int k) {
    if (!running(d)) return;
    const double dn = sum_partials<RED_SUM>(d.P(P_DN), d.Gdn);
    if (threadIdx.x == 0) alpha_finish(d, k, dn);
}

// K2 of iteration k
__global__ __launch_bounds__(kBlock) void k_iter_update_rows(PcgDev d, int k) {
    if (!running(d)) return;
    const int row = blockIdx.x * kBlock + threadIdx.x;
    const double malpha = -1.0 * alpha_slot(d, k);
    double rs = 0.0, rr = 0.0;
    if (row < d.n) {
        const double ro = ntload(d.r + row), apr = ntload(d.Ap + row);
        const double di = d.precond ? ntload(d.dinv + row) : 0.0;
        const double ri = ro + malpha * apr;  // :97
        ntstore(d.r + row, ri);
        double si;
        if (d.precond) {  // :99-105
            si = 0.0;
            si = si + ri * di;
        } else {
            si = ri;
        }
        ntstore(d.s + row, si);
        rs = ri * si;
        rr = ri * ri;
    }
    // one partial per wavefront: no barrier
    rs = wave_sum_dpp(rs);
    rr = wave_sum_dpp(rr);
    const int wv = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if ((threadIdx.x & 63) == 0) {
        d.chunk[wv] = rs;
        d.chunk_s[wv] = rr;
    }
    if (leader()) *d.iters = k + 1;
}

// F2 (one workgroup) after K2 of iteration k - 1: rho_k, r.r, the
// convergence test (stamped with k), beta_k
__device__ inline void beta_finish(const PcgDev &d, int k, double rho, double rr) {
    const double tol2 = d.sc[S_TOL2];
    d.sc[S_RR] = rr;
    if (tol2 >= 0.0 && rr <= tol2) {
        state_store(d.state, conv_stamp(k));
        return;
    }
    d.sc[S_BETA] = rho / rho_slot(d, k - 1);  // :107-111
    d.sc[S_RHO0 + (k & 1)] = rho;
}

__global__ __launch_bounds__(kBlock) void k_beta_fin(PcgDev d, int k) {
    if (!running(d)) return;
    double sums[2];
    sum_partials_n<2>({d.P(P_RS), d.P(P_RR)}, d.Gdn, sums);
    if (threadIdx.x == 0) beta_finish(d, k, sums[0], sums[1]);
}


// x += alpha_k p of the last iteration k, unless a breakdown stopped the solve
__global__ __launch_bounds__(kBlock) void k_x_flush_rows(PcgDev d, const double *__restrict__ p, int k) {
    const int st = state_load(d.state);
    if (st != ST_RUN && st != conv_stamp(k + 1)) return;
    const int row = blockIdx.x * kBlock + threadIdx.x;
    if (row < d.n) d.x[row] = d.x[row] + alpha_slot(d, k) * p[row];
}

// one workgroup after the last iteration: the final convergence test on the
// r.r F2 left
__global__ __launch_bounds__(kBlock) void k_fin_rows(PcgDev d) {
    if (threadIdx.x != 0) return;
    const double rr = d.sc[S_RR], tol2 = d.sc[S_TOL2];
    const int st = state_load(d.state);
    if ((st == ST_RUN || (st & ST_MASK) == ST_CONV) && tol2 >= 0.0 && rr <= tol2) state_store(d.state, ST_CONV);
}

// one workgroup after the last iteration: r.r and the final convergence test
__global__ __launch_bounds__(kBlock) void k_fin(PcgDev d) {
    const double rr = sum_partials<RED_SUM>(d.P(P_RR), d.G);
    if (threadIdx.x == 0) {
        d.sc[S_RR] = rr;
        const double tol2 = d.sc[S_TOL2];
        if (state_load(d.state) == ST_RUN && tol2 >= 0.0 && rr <= tol2) state_store(d.state, ST_CONV);
    }
}

// ---- MIXED iteration (FP32 vectors and operator, FP64 dot accumulation) --------

// K0 (k >= 1): rho_k, the correction update xc += alpha_{k-1} p of the
// previous iteration (`pending`: not yet folded by a defect correction),
// beta (0 after a restart: the previous rho slot is +inf), p32 = s32 + beta p32
__global__ __launch_bounds__(kBlock) void k_m_dir(PcgDev d, float *__restrict__ p, int k, int pending) {
    float *__restrict__ xc = d.xc32;
    const float *__restrict__ s = d.s32;
    if (!running(d)) return;
    const double rho = sum_partials<RED_SUM>(d.P(P_RS), d.G);
    const float a32 = pending ? (float)alpha_slot(d, k - 1) : 0.0f;  // alpha_{k-1}, as K2 of k-1 formed it
    const float beta = (float)(rho / rho_slot(d, k - 1));
    if (leader()) d.sc[S_RHO0 + (k & 1)] = rho;
    // nontemporal accesses: each array is streamed once per sweep (mixed
    // solve at 256^3 over four runs: 392-394 ms with them, 403-405 without;
    // the sweeps' own times are unchanged within noise)
    FOR_ROWS(d.n, row) {
        const float xo = __builtin_nontemporal_load(xc + row), si = __builtin_nontemporal_load(s + row),
                    po = __builtin_nontemporal_load(p + row);
        if (pending) __builtin_nontemporal_store(xo + a32 * po, xc + row);
        __builtin_nontemporal_store(si + beta * po, p + row);
    }
}

// (as DirOperand, in FP32: the direction update's float expression)
struct DirOperand32 {
    const float *__restrict__ s, *__restrict__ p;
    float beta;
    int n;
    __device__ inline float at(int j) const {
        const float po = p[j], si = s[j < n ? j : 0];
        return j < n ? si + beta * po : po;
    }
};

// Small operators, k >= 1: k_m_dir and the FP32 product in one kernel (as
// k_iter_dirprod)
template <int W>
__global__ __launch_bounds__(kBlock) void k_m_dirprod(EllView A, PcgDev d, const float *__restrict__ pold,
                                                      float *__restrict__ pnew, int k, int pending) {
    const int st = state_load(d.state);
    const int c = xcd_block(blockIdx.x, gridDim.x);
    const int row = c * kBlock + threadIdx.x;
    const bool own = row < d.n;
    float xo = 0.0f, si = 0.0f, po = 0.0f;
    if (own) {
        xo = d.xc32[row];
        si = d.s32[row];
        po = pold[row];
    }
    const double rho_prev = early_load(d.sc + S_RHO0 + ((k - 1) & 1));
    const double alpha = early_load(d.sc + S_ALPHA0 + ((k - 1) & 1));  // alpha_{k-1}, as K2 of k-1 formed it
    const double zaz = early_load(d.sc + S_ZAZ);
    PaddedSums<1> ps;
    ps.load({d.P(P_RS)}, {d.G});
    constexpr int SW = W < 0 ? -W : 1;
    StencilPre<SW, float> q;
    if (W < 0 && own) stencil_row_pre<SW, float>(A, A.val32, row, q);
    double sums[1];
    ps.reduce(sums);
    if (st != ST_RUN) return;
    const bool dfl = deflating(d, zaz);
    const float a32 = pending ? (float)alpha : 0.0f;
    const float beta = (float)(sums[0] / rho_prev);
    if (leader()) d.sc[S_RHO0 + (k & 1)] = sums[0];
    double acc = 0.0, sa = 0.0;
    if (own) {
        if (pending) d.xc32[row] = xo + a32 * po;
        const float pr = si + beta * po;
        pnew[row] = pr;
        const DirOperand32 op{d.s32, pold, beta, d.n};
        float sum;
        if constexpr (W < 0)
            stencil_row_fin<SW, float, float>(A, A.val32, op, row, q, sum);
        else
            row_dot<W, float, float>(A, A.val32, op, row, sum);
        d.Ap32[row] = sum;
        acc = (double)pr * (double)sum;
        sa = (double)sum;
    }
    double red[2] = {acc, sa};
    block_reduce_n<2>(red);
    if (threadIdx.x == 0) {
        dn_part(d, k)[c] = red[0];
        if (dfl) sa_part(d, k)[c] = red[1];
    }
}

// K1 in FP32 (SMALL as in iter_spmv; the symmetric form of the slot values
// measured slower here, 121-131 against 118.5 us at 256^3, so the FP32
// product keeps the full slot values)
template <int W, bool SMALL>
__device__ inline void m_spmv(const EllView &A, const PcgDev &d, const float *__restrict__ p) {
    if constexpr (!SMALL) {
        // no state test: the product's stores (Ap32, its partials) are read
        // only by the fold and the update, which test it; and mixed
        // iterations are queued only while the solve runs (convergence is
        // decided at a defect correction, after which nothing is queued)
        const int c = xcd_block(blockIdx.x, gridDim.x);
        const int row = c * kBlock + threadIdx.x;
        double acc = 0.0, sa = 0.0;
        if (row < d.n) {
            // (a SelfOperand here measured 9% slower than the separate load)
            Operand<float> po{p};
            const float pr = p[row];
            float sum;
            row_dot<W, float, float>(A, A.val32, po, row, sum);
            d.Ap32[row] = sum;
            acc = (double)pr * (double)sum;
            sa = (double)sum;
        }
        acc = wave_sum_dpp(acc);
        const int wv = c * kWaves + (threadIdx.x >> 6);
        if ((threadIdx.x & 63) == 0) d.chunk[wv] = acc;
        if (d.deflate) {
            sa = wave_sum_dpp(sa);
            if ((threadIdx.x & 63) == 0) d.chunk_s[wv] = sa;
        }
        return;
    }
    const int st = state_load(d.state);
    const int c = xcd_block(blockIdx.x, gridDim.x);
    const int row = c * kBlock + threadIdx.x;
    float pr = 0.0f, sum = 0.0f;
    if (row < d.n) {
        Operand<float> po{p};
        pr = p[row];
        row_dot<W, float, float>(A, A.val32, po, row, sum);
    }
    if (st != ST_RUN) return;
    double acc = 0.0, sa = 0.0;
    if (row < d.n) {
        d.Ap32[row] = sum;
        acc = (double)pr * (double)sum;
        sa = (double)sum;
    }
    double red[2] = {acc, sa};
    block_reduce_n<2>(red);
    if (threadIdx.x == 0) {
        d.chunk[c] = red[0];
        if (d.deflate) sa_part(d, 0)[c] = red[1];
    }
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_m_spmv(EllView A, PcgDev d, const float *__restrict__ p) {
    m_spmv<W, false>(A, d, p);
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_m_spmv_small(EllView A, PcgDev d, const float *__restrict__ p) {
    m_spmv<W, true>(A, d, p);
}

template <bool SMALL>
__global__ __launch_bounds__(kBlock) void k_m_update(PcgDev d, int k) {
    float *__restrict__ r = d.r32, *__restrict__ s = d.s32;
    const float *__restrict__ Ap = d.Ap32, *__restrict__ dinv = d.dinv32;
    const bool precond = d.precond != 0;
    double rs = 0.0, rr = 0.0;
    if constexpr (!SMALL) {
        if (!running(d)) return;
        // every prologue load issued before the first is used: 1.A1, the
        // list of A 1's nonzeros, and the p.Ap and 1.Ap partials (summed
        // together; alpha goes to its scalar slot, so later kernels need
        // not reproduce this sum)
        const double zaz = d.sc[S_ZAZ];
        const int naz_all = d.azl->count;
        int li[kAzList];
        float lv[kAzList];
#pragma unroll
        for (int j = 0; j < kAzList; ++j) {
            li[j] = d.azl->idx[j];
            lv[j] = d.azl->val[j];
        }
        double sums[2];
        if (d.Gdn <= kBlock)  // one round of the uniform loop (up to 256^3)
            sum_partials_n<2>({d.P(P_DN), d.P(P_SA)}, d.Gdn, sums);
        else  // all slots in one round trip
            sum_padded<2>({d.P(P_DN), d.P(P_SA)}, {d.Gdn, d.Gdn}, sums);
        const bool dfl = deflating(d, zaz);
        const double sap = dfl ? sums[1] : 0.0;
        const double dn = m_denom(sums[0], sap, zaz, dfl);
        const double alpha = rho_slot(d, k) / dn;
        if (dn == 0.0 || !isfinite(dn) || !isfinite(alpha)) {
            if (leader()) state_store(d.state, ST_BREAK);
            return;
        }
        store_alpha(d, k, alpha);
        const float a32 = (float)alpha, c32 = dfl ? (float)(sap / zaz) : 0.0f;
        const float *__restrict__ az = d.az32;
        // A 1 from its list of nonzeros when it is short (the same values
        // as the dense array, without streaming it), else row by row
        const int naz = dfl ? naz_all : 0;
        const bool listed = naz <= kAzList;
#pragma unroll
        for (int j = 0; j < kAzList; ++j) li[j] = j < naz ? li[j] : -1;
        FOR_ROWS(d.n, row) {
            // every operand first, so the row costs one memory round trip
            const float ro = __builtin_nontemporal_load(r + row), apr = __builtin_nontemporal_load(Ap + row);
            const float di = precond ? __builtin_nontemporal_load(dinv + row) : 0.0f;
            float azr = 0.0f;
            if (dfl && !listed) {
                azr = az[row];
            } else {
                // rows compared with the list only in a chunk that holds a
                // listed row (a workgroup-uniform test per chunk)
                const int base = __builtin_amdgcn_readfirstlane(row - (int)threadIdx.x);
                bool hit = false;
#pragma unroll
                for (int j = 0; j < kAzList; ++j) hit = hit || (unsigned)(li[j] - base) < (unsigned)kBlock;
                if (hit)
#pragma unroll
                    for (int j = 0; j < kAzList; ++j) azr = row == li[j] ? lv[j] : azr;
            }
            const float wr = dfl ? apr - c32 * azr : apr;  // (P A p)_row
            const float ri = ro - a32 * wr;
            __builtin_nontemporal_store(ri, r + row);
            const float si = precond ? ri * di : ri;
            __builtin_nontemporal_store(si, s + row);
            rs = rs + (double)ri * (double)si;
            rr = rr + (double)ri * (double)ri;
        }
    } else {
        ThreadRows t(d.n);
        float ro = 0.0f, apr = 0.0f, di = 0.0f, azr = 0.0f;
        const float *__restrict__ az = d.az32;
        const int st = state_load(d.state);
        const bool want_az = d.deflate != 0;
        if (t.valid()) {
            ro = r[t.row];
            apr = Ap[t.row];
            if (precond) di = dinv[t.row];
            if (want_az) azr = az[t.row];
        }
        const double rho = early_load(d.sc + S_RHO0 + (k & 1));
        const double zaz = early_load(d.sc + S_ZAZ);
        double sums[2];
        sum_padded<2>({dn_part(d, k), sa_part(d, k)}, {d.Gdn, d.Gdn}, sums);
        if (st != ST_RUN) return;
        const bool dfl = deflating(d, zaz);
        const double dn = m_denom(sums[0], sums[1], zaz, dfl);
        const double alpha = rho / dn;
        if (dn == 0.0 || !isfinite(dn) || !isfinite(alpha)) {
            if (leader()) state_store(d.state, ST_BREAK);
            return;
        }
        store_alpha(d, k, alpha);
        const float a32 = (float)alpha, c32 = dfl ? (float)(sums[1] / zaz) : 0.0f;
        while (t.valid()) {
            const int row = t.row;
            const float wr = dfl ? apr - c32 * azr : apr;  // (P A p)_row
            const float ri = ro - a32 * wr;
            r[row] = ri;
            const float si = precond ? ri * di : ri;
            s[row] = si;
            rs = rs + (double)ri * (double)si;
            rr = rr + (double)ri * (double)ri;
            t.next();
            if (t.valid()) {
                ro = r[t.row];
                apr = Ap[t.row];
                if (precond) di = dinv[t.row];
                if (want_az) azr = az[t.row];
            }
        }
    }
    double red[2] = {rs, rr};
    block_reduce_n<2>(red);
    if (threadIdx.x == 0) {
        d.P(P_RS)[blockIdx.x] = red[0];
        d.P(P_RR)[blockIdx.x] = red[1];
    }
    if (leader()) *d.iters = k + 1;
}

// one workgroup: the FP32 recursive r.r for the host's correction test
__global__ __launch_bounds__(kBlock) void k_m_status(PcgDev d) {
    const double rr = sum_partials<RED_SUM>(d.P(P_RR), d.G);
    if (threadIdx.x == 0) d.sc[S_RR] = rr;
}

// x += (xc + alpha_k p) (the last iteration's pending update first), xc = 0
__global__ __launch_bounds__(kBlock) void k_m_fold(PcgDev d, const float *__restrict__ p, int k) {
    const float a32 = (float)alpha_slot(d, k);
    FOR_ROWS(d.n, row) {
        const float xcv = d.xc32[row] + a32 * p[row];
        d.x[row] = d.x[row] + (double)xcv;
        d.xc32[row] = 0.0f;
    }
}

// Defect correction, part 1: r = b - A x in FP64 (empty rows: x) and the
// partials of 1.r.
template <int W>
__global__ __launch_bounds__(kBlock) void k_m_resid(EllView A, PcgDev d) {
    double sr = 0.0;
    Operand<double> xo{d.x};
    FOR_ROWS(d.n, row) {
        double sum;
        bool any = row_dot<W, double, double>(A, A.val, xo, row, sum);
        double t = any ? sum : d.x[row];
        double ri = d.b[row] + -1.0 * t;
        d.r[row] = ri;
        sr = sr + ri;
    }
    sr = block_reduce<RED_SUM>(sr);
    if (threadIdx.x == 0) d.P(P_SUMR)[blockIdx.x] = sr;
}

// Part 2: the coarse step c = (1.r)/(1.A1) (skipped when off, or when 1.A1
// is not positive and finite): x += c, r -= c A1 on the solved rows; then
// r32 = r, s32 = Dinv r32, partials of r32.s32 and of the true r.r.
__global__ __launch_bounds__(kBlock) void k_m_coarse(PcgDev d, int use) {
    const double sr = sum_partials<RED_SUM>(d.P(P_SUMR), d.G);
    const double zaz = d.sc[S_ZAZ];
    double c = (use && zaz > 0.0 && isfinite(zaz)) ? sr / zaz : 0.0;
    if (!isfinite(c)) c = 0.0;
    double rs = 0.0, rr = 0.0;
    FOR_ROWS(d.n, row) {
        double ri = d.r[row];
        if (c != 0.0) {
            d.x[row] = d.x[row] + c;
            ri = ri - c * d.az[row];
            d.r[row] = ri;
        }
        float r32 = (float)ri;
        d.r32[row] = r32;
        float s32 = d.precond ? r32 * d.dinv32[row] : r32;
        d.s32[row] = s32;
        rs = rs + (double)r32 * (double)s32;
        rr = rr + ri * ri;
    }
    double red[2] = {rs, rr};
    block_reduce_n<2>(red);
    if (threadIdx.x == 0) {
        d.P(P_RS2)[blockIdx.x] = red[0];
        d.P(P_RR2)[blockIdx.x] = red[1];
    }
}

// One workgroup after k_m_resid: the replaced residual becomes the
// recursion's (its partials are copied over the FP32 ones); the direction is
// kept unless the recursive r.r had drifted below a hundredth of the true
// one, then the previous rho slot is set to +inf so the next beta is 0.
__global__ __launch_bounds__(kBlock) void k_m_corr_fin(PcgDev d, int k, int first) {
    const double rrt = sum_partials<RED_SUM>(d.P(P_RR2), d.G);
    const double rrec = first ? 0.0 : sum_partials<RED_SUM>(d.P(P_RR), d.G);
    __syncthreads();
    for (int i = threadIdx.x; i < d.G; i += kBlock) {
        d.P(P_RS)[i] = d.P(P_RS2)[i];
        d.P(P_RR)[i] = d.P(P_RR2)[i];
    }
    const double rs = first ? sum_partials<RED_SUM>(d.P(P_RS2), d.G) : 0.0;
    if (threadIdx.x == 0) {
        if (first) d.sc[S_RHO0] = rs;  // rho of iteration 0
        if (!first && rrt > 100.0 * rrec) d.sc[S_RHO0 + ((k - 1) & 1)] = INFINITY;
        d.sc[S_TRUE_RR] = rrt;
        d.sc[S_RR] = rrt;
        const double tol2 = d.sc[S_TOL2];
        if (state_load(d.state) == ST_RUN && tol2 >= 0.0 && rrt <= tol2) state_store(d.state, ST_CONV);
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

// z = 1 on rows [0, n), 0 on [n, n_all); az = 0 on [0, n)
__global__ __launch_bounds__(kBlock) void k_ones(int n, int n_all, double *__restrict__ z, double *__restrict__ az,
                                                 AzList *azl) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n_all) z[i] = i < n ? 1.0 : 0.0;
    if (i < n) az[i] = 0.0;
    if (i == 0) azl->count = 0;
}

// az32 = (float) az, and the list of its nonzeros (entries past kAzList
// are only counted; the order of the entries does not matter, each row
// matches at most one)
__global__ __launch_bounds__(kBlock) void k_az32(int n, const double *__restrict__ az, float *__restrict__ az32,
                                                 AzList *azl) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float v = (float)az[i];
    az32[i] = v;
    if (v != 0.0f) {
        const int slot = atomicAdd(&azl->count, 1);
        if (slot < kAzList) {
            azl->idx[slot] = i;
            azl->val[slot] = v;
        }
    }
}

// ---- the solve without allocations -------------------------------------------
//
// The workspace of one FP64 solve on an operator of at most kScratchRows
// rows lives in `scratch`, a file-scope device array (one copy per GPU, held
// until the process ends): r, s, Ap, Dinv (kScratchRows doubles each), the
// two direction buffers (kScratchAll each), the partial-sum arrays, the
// scalars and the state words.  Host code never passes `scratch` itself
// (in host code the name is a stand-in, not a device address): gip_init
// looks up its address on each device with hipGetSymbolAddress and keeps it
// in g_scratch, and the entry point builds its workspace from that pointer.
// Every call on a GPU shares its array; what must start at zero is cleared
// by a kernel at the start of each call, so calls that overlap on one GPU
// race.
constexpr int kScratchRows = 1 << 18;
constexpr int kScratchAll = (1 << 18) + (1 << 16);
constexpr size_t kScratchZero = (size_t)P_COUNT * kMaxRedBlocks + S_COUNT + 1;  // partials, scalars, state words
constexpr size_t kScratchDoubles = 4 * (size_t)kScratchRows + 2 * (size_t)kScratchAll + kScratchZero;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kScratchDevices = 64;
static double *g_scratch[kScratchDevices];
static std::mutex g_scratch_mu;

// Called by gip_init with `device` current.
int pcg_scratch_setup(int device) {
    if (device < 0 || device >= kScratchDevices) return GIP_OK;  // no no-alloc solves on this device
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return GIP_OK;
    double *ptr = nullptr;
    GIP_HIP_TRY(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
    g_scratch[device] = ptr;
    return GIP_OK;
}

__global__ __launch_bounds__(kBlock) void k_zero(size_t n, double *__restrict__ a) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

}  // namespace gip

using namespace gip;

struct gip_pcg {
    gip_ctx *ctx;
    const gip_mat *A;
    int n, n_all, grid;
    double *r, *p, *s, *Ap, *dinv;
    float *r32, *p32, *s32, *Ap32, *xc32, *dinv32;
    double *p2;   // second direction buffer of the fused iteration (small operators)
    float *p32b;  // its FP32 counterpart
    double *az;   // A 1 (MIXED coarse step)
    float *az32;  // its FP32 copy (MIXED deflation)
    AzList *azl;  // the FP32 copy's nonzeros, when few
    double *sc, *part, *chunk;
    int nchunk, gdn;
    bool fold;  // nchunk > kMaxRedBlocks: the chunk partials need the fold kernel;
                // otherwise the fused two-kernel iteration runs
    bool rows;  // fold and more than kRowsMin rows: the row-per-thread iteration
    int *state;    // state, iters
    int *h_state;  // pinned copy
    double *h_sc;  // pinned scalars
    hipEvent_t tev0, tev1;  // time_products: the sampled product launch
};

// Multiple of 8 (one share per XCD), at most kMaxRedBlocks (one partial each).
// The product kernels run one workgroup per 256-row chunk instead.
static int pcg_grid(int n) {
    int nchunk = (n + kBlock - 1) / kBlock;
    int g = ((nchunk + 7) / 8) * 8;
    if (g > kMaxRedBlocks) g = kMaxRedBlocks;
    if (g < 8) g = 8;
    return g;
}

// Launch KERNEL<W> with W the operator's kernel width: a uniform ELL width
// (5..7), a stencil width (-7, -8), or 0 for the generic loops.
#define DISPATCH_WG(A, GRID, KERNEL, ...)                                     \
    do {                                                                        \
        switch (kernel_width(A)) {                                              \
            case 5: KERNEL<5><<<GRID, kBlock, 0, st>>>(__VA_ARGS__); break;      \
            case 6: KERNEL<6><<<GRID, kBlock, 0, st>>>(__VA_ARGS__); break;      \
            case 7: KERNEL<7><<<GRID, kBlock, 0, st>>>(__VA_ARGS__); break;      \
            case -7: KERNEL<-7><<<GRID, kBlock, 0, st>>>(__VA_ARGS__); break;    \
            case -8: KERNEL<-8><<<GRID, kBlock, 0, st>>>(__VA_ARGS__); break;    \
            default: KERNEL<0><<<GRID, kBlock, 0, st>>>(__VA_ARGS__); break;     \
        }                                                                       \
    } while (0)
#define DISPATCH_W(A, KERNEL, ...) DISPATCH_WG(A, S->grid, KERNEL, __VA_ARGS__)
// the products over the symmetric form: one workgroup per chunk
#define SYM_LAUNCH(A, KERNEL, ...)                                                    \
    do {                                                                              \
        if (kernel_width(A) == -7)                                                    \
            KERNEL<-7><<<S->nchunk, kBlock, 0, st>>>(__VA_ARGS__);                    \
        else                                                                          \
            KERNEL<-8><<<S->nchunk, kBlock, 0, st>>>(__VA_ARGS__);                    \
    } while (0)

static PcgDev make_dev(gip_pcg *S, const double *b, double *x, int precond) {
    PcgDev d;
    d.n = S->n;
    d.G = S->grid;
    d.nchunk = S->nchunk;
    d.Gdn = S->gdn;
    d.chunk = S->fold ? S->chunk : S->part + P_DN * kMaxRedBlocks;
    d.chunk_s = S->fold ? S->chunk + (size_t)S->nchunk * kWaves : nullptr;
    d.az32 = S->az32;
    d.azl = S->azl;
    d.deflate = 0;
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
    d.az = S->az;
    d.sc = S->sc;
    d.state = S->state;
    d.iters = S->state + 1;
    d.part = S->part;
    d.precond = precond;
    return d;
}

extern "C" {

int gip_pcg_create(gip_ctx *c, const gip_mat *A, int32_t n_all, gip_pcg **out) {
    if (!c || !A || !out || n_all < A->nrows) return set_error(GIP_EINVAL, "gip_pcg_create: bad argument");
    // one stream orders the operator's value updates and the solves
    if (A->ctx != c) return set_error(GIP_EINVAL, "gip_pcg_create: operator belongs to another context");
    if (n_all < A->cols_needed) return set_error(GIP_EINVAL, "gip_pcg_create: vectors shorter than the operator's columns");
    if (A->format != GIP_FMT_ELL && A->format != GIP_FMT_CSR && A->format != GIP_FMT_STENCIL)
        return set_error(GIP_EINVAL, "gip_pcg_create: operator must be ELL, stencil or CSR");
    gip_pcg *S = new (std::nothrow) gip_pcg();
    if (!S) return set_error(GIP_ENOMEM, "gip_pcg_create: allocation failed");
    S->ctx = c;
    S->A = A;
    S->n = A->nrows;
    S->n_all = n_all;
    S->grid = pcg_grid(A->nrows);
    S->nchunk = std::max(1, (A->nrows + kBlock - 1) / kBlock);
    // up to kMaxRedBlocks chunks (2^19 rows) the next kernels read the
    // per-chunk partials directly; beyond, a fold kernel sums each run of
    // 256 chunks first
    S->fold = S->nchunk > kMaxRedBlocks;
    S->rows = S->fold && S->n > kRowsMin;
    S->gdn = S->fold ? (S->nchunk + kBlock - 1) / kBlock : S->nchunk;
    if (S->gdn > kMaxRedBlocks) {
        delete S;
        return set_error(GIP_EINVAL, "gip_pcg_create: more rows than the product partials cover (2^27)");
    }
    size_t n = (size_t)std::max(1, A->nrows);
    size_t na = (size_t)std::max(1, n_all);  // directions are SpMV inputs: ghost columns included
    hipError_t e = hipSetDevice(c->device);
    auto dalloc = [&](void **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes);
    };
    dalloc((void **)&S->r, n * 8);
    dalloc((void **)&S->p, na * 8);
    if (!S->fold) dalloc((void **)&S->p2, na * 8);
    dalloc((void **)&S->s, n * 8);
    dalloc((void **)&S->Ap, n * 8);
    dalloc((void **)&S->dinv, n * 8);
    dalloc((void **)&S->sc, S_COUNT * 8);
    dalloc((void **)&S->part, (size_t)P_COUNT * kMaxRedBlocks * 8);
    if (S->fold) dalloc((void **)&S->chunk, (size_t)S->nchunk * kWaves * 2 * 8);  // p.Ap, then 1.Ap
    dalloc((void **)&S->state, 2 * sizeof(int));
    if (e == hipSuccess) e = hipHostMalloc((void **)&S->h_state, 2 * sizeof(int), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void **)&S->h_sc, S_COUNT * 8, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMemsetAsync(S->sc, 0, S_COUNT * 8, c->stream);
    // partial slots past each producer's grid stay zero for sum_padded
    if (e == hipSuccess) e = hipMemsetAsync(S->part, 0, (size_t)P_COUNT * kMaxRedBlocks * 8, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(S->state, 0, 2 * sizeof(int), c->stream);
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
    void *ptrs[] = {S->r,    S->p,      S->s,  S->Ap, S->dinv, S->r32,   S->p32,  S->s32,  S->Ap32,
                    S->xc32, S->dinv32, S->az, S->sc, S->part, S->chunk, S->state, S->p2, S->p32b, S->az32, S->azl};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (S->h_state) hipHostFree(S->h_state);
    if (S->h_sc) hipHostFree(S->h_sc);
    if (S->tev0) hipEventDestroy(S->tev0);
    if (S->tev1) hipEventDestroy(S->tev1);
    delete S;
    return GIP_OK;
}

static int ensure_mixed(gip_pcg *S) {
    if (S->r32) return GIP_OK;
    const gip_mat *A = S->A;
    if (((A->format == GIP_FMT_ELL || A->format == GIP_FMT_STENCIL) && !A->d_ell_val32) ||
        (A->format == GIP_FMT_CSR && !A->d_val32))
        return set_error(GIP_EINVAL, "mixed precision needs an operator created with GIP_MAT_FP32");
    size_t n = (size_t)std::max(1, S->n), na = (size_t)std::max(1, S->n_all);
    float **bufs[] = {&S->r32, &S->s32, &S->Ap32, &S->xc32, &S->dinv32};
    for (float **b : bufs) GIP_HIP_TRY(hipMalloc(b, n * 4));
    GIP_HIP_TRY(hipMalloc(&S->az, n * 8));
    GIP_HIP_TRY(hipMalloc(&S->az32, n * 4));
    GIP_HIP_TRY(hipMalloc(&S->azl, sizeof(AzList)));
    GIP_HIP_TRY(hipMemsetAsync(S->azl, 0, sizeof(AzList), S->ctx->stream));
    // the FP32 direction is an SpMV input: ghost columns included, and zero
    // there (the correction leaves the ghost values of x unchanged)
    GIP_HIP_TRY(hipMalloc(&S->p32, na * 4));
    GIP_HIP_TRY(hipMemsetAsync(S->p32, 0, na * 4, S->ctx->stream));
    if (!S->fold) {
        GIP_HIP_TRY(hipMalloc(&S->p32b, na * 4));
        GIP_HIP_TRY(hipMemsetAsync(S->p32b, 0, na * 4, S->ctx->stream));
    }
    if (S->n) {
        k_to32<<<(S->n + kBlock - 1) / kBlock, kBlock, 0, S->ctx->stream>>>(S->n, S->dinv, S->dinv32);
        GIP_LAUNCH_CHECK("dinv32");
    }
    return GIP_OK;
}

// az = A 1 over the solved rows and 1.A1 into S_ZAZ, from A's current
// values (one product and one reduction per mixed solve); the FP64
// direction buffer serves as the vector 1 before the setup overwrites it.
static int coarse_setup(gip_pcg *S) {
    gip_ctx *c = S->ctx;
    if (S->n_all == 0) return GIP_OK;
    double *z = S->p;
    k_ones<<<(S->n_all + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(S->n, S->n_all, z, S->az, S->azl);
    GIP_LAUNCH_CHECK("coarse vector");
    int rc = spmv_launch(c, S->A, z, S->az, 1.0, 0.0, 0, nullptr);
    if (!rc) rc = launch_reduce(c, RED_SUM, S->n, S->az, z, S->sc + S_ZAZ);
    if (!rc && S->n) {
        k_az32<<<(S->n + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(S->n, S->az, S->az32, S->azl);
        GIP_LAUNCH_CHECK("coarse vector fp32");
    }
    return rc;
}

static int poll(gip_pcg *S, bool scalars) {
    hipStream_t st = S->ctx->stream;
    GIP_HIP_TRY(hipMemcpyAsync(S->h_state, S->state, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    if (scalars) GIP_HIP_TRY(hipMemcpyAsync(S->h_sc, S->sc, S_COUNT * 8, hipMemcpyDeviceToHost, st));
    GIP_HIP_TRY(hipStreamSynchronize(st));
    return GIP_OK;
}

// The solve on a solver's workspace (allocated by gip_pcg_create, or the
// no-alloc scratch); arguments already checked.
static int pcg_solve_impl(gip_pcg *S, const double *b, double *x, const double *x0, const gip_pcg_opts *o,
                          gip_pcg_stats *stats) {
    gip_ctx *c = S->ctx;
    hipStream_t st = c->stream;
    const gip_mat *A = S->A;
    const bool mixed = o->precision == GIP_PCG_MIXED;
    if (mixed) {
        int rc = ensure_mixed(S);
        if (rc) return rc;
    }
    if (mixed) {
        int rc = coarse_setup(S);
        if (rc) return rc;
    }
    // the symmetric form of a 7- or 8-slot stencil operator for the FP64
    // products of the long iterations (made and checked here after a change
    // of values; the no-alloc solve never has those iterations)
    bool sym = false;
    if (!mixed && S->fold && (kernel_width(A) == -7 || kernel_width(A) == -8)) {
        int rc = mat_sym_prepare(c, const_cast<gip_mat *>(A), &sym);
        if (rc) return rc;
    }
    PcgDev d = make_dev(S, b, x, o->precond ? 1 : 0);
    EllView v = view(A);
    const int pollk = o->poll > 0 ? o->poll : 16;
    const int use_coarse = o->coarse >= 0 ? 1 : 0;
    d.deflate = mixed && use_coarse;
    // direction of iteration k: one buffer with the fold, alternating by
    // parity in the fused iteration
    auto pdir = [&](int k) { return (!S->fold && (k & 1)) ? S->p2 : S->p; };
    auto pdir32 = [&](int k) { return (!S->fold && (k & 1)) ? S->p32b : S->p32; };
    // time_products: events around one product launch (k >= 1) per host
    // poll, read after the poll's synchronisation
    const bool timing = o->time_products != 0;
    if (timing && !S->tev0) {
        GIP_HIP_TRY(hipEventCreate(&S->tev0));
        GIP_HIP_TRY(hipEventCreate(&S->tev1));
    }
    std::vector<float> samples;
    bool armed = false;  // a sampled pair is recorded and not read yet
    auto sample_begin = [&](int k) {
        const bool on = timing && !armed && k > 0;
        if (on) hipEventRecord(S->tev0, st);
        return on;
    };
    auto sample_end = [&](bool on) {
        if (on) {
            hipEventRecord(S->tev1, st);
            armed = true;
        }
    };
    auto sample_read = [&]() {  // after a host synchronisation
        float ms = 0.0f;
        if (armed && hipEventElapsedTime(&ms, S->tev0, S->tev1) == hipSuccess) samples.push_back(ms);
        armed = false;
    };
    // setup, GIP_LinearSolver.h:55-80
    DISPATCH_W(A, k_setup_resid, v, d, x0, S->n_all, S->p, S->p2);
    k_setup_dir<<<S->grid, kBlock, 0, st>>>(d, S->p);
    k_setup_fin<<<1, kBlock, 0, st>>>(d, o->tol);
    GIP_LAUNCH_CHECK("pcg setup");
    int corrections = 0, rc;
    // FP64 with a fixed iteration count (tol <= 0): nothing can stop the
    // solve early but a breakdown, after which every queued kernel returns
    // at once, so all iterations are queued and the host reads the state
    // once, at the end (every 4096 iterations on long runs)
    const bool fixed = !mixed && !(o->tol > 0.0);
    if (fixed)
        S->h_state[0] = ST_RUN;
    else if ((rc = poll(S, true)))
        return rc;
    if (!mixed) {
        int k = 0;  // iterations queued
        while (k < o->maxit && S->h_state[0] == ST_RUN) {
            int chunk = std::min(fixed ? 4096 : pollk, o->maxit - k);
            for (int j = 0; j < chunk; ++j, ++k) {
                if (S->fold && !S->rows) {
                    if (k > 0) k_iter_dir<<<S->grid, kBlock, 0, st>>>(d, S->p, k);
                    const bool smp = sample_begin(k);
                    if (sym)
                        SYM_LAUNCH(A, k_iter_spmv_sym, v, d, S->p);
                    else
                        DISPATCH_WG(A, S->nchunk, k_iter_spmv, v, d, S->p);
                    sample_end(smp);
                    k_fold_dn<<<S->gdn, kBlock, 0, st>>>(d);
                    k_iter_update_grid<<<S->grid, kBlock, 0, st>>>(d, k);
                } else if (S->fold) {
                    // (partials: p.Ap in chunk, then r.s in chunk and r.r in chunk_s)
                    if (k > 0) k_iter_dir_rows<<<S->nchunk, kBlock, 0, st>>>(d, S->p, k);
                    const bool smp = sample_begin(k);
                    if (sym)
                        SYM_LAUNCH(A, k_iter_spmv_sym, v, d, S->p);
                    else
                        DISPATCH_WG(A, S->nchunk, k_iter_spmv, v, d, S->p);
                    sample_end(smp);
                    k_fold_two<<<S->gdn, kBlock, 0, st>>>(d, d.chunk, nullptr, P_DN, P_DN);
                    k_alpha_fin<<<1, kBlock, 0, st>>>(d, k);
                    k_iter_update_rows<<<S->nchunk, kBlock, 0, st>>>(d, k);
                    k_fold_two<<<S->gdn, kBlock, 0, st>>>(d, d.chunk, d.chunk_s, P_RS, P_RR);
                    k_beta_fin<<<1, kBlock, 0, st>>>(d, k + 1);
                } else {
                    const bool smp = sample_begin(k);
                    if (k > 0 && o->tol > 0.0)
                        DISPATCH_WG(A, S->nchunk, k_iter_dirprod, v, d, pdir(k - 1), pdir(k), k);
                    else if (k > 0)  // fixed iteration count: no r.r test
                        DISPATCH_WG(A, S->nchunk, k_iter_dirprod_fixed, v, d, pdir(k - 1), pdir(k), k);
                    else
                        DISPATCH_WG(A, S->nchunk, k_iter_spmv_small, v, d, S->p);
                    sample_end(smp);
                    k_iter_update<<<S->grid, kBlock, 0, st>>>(d, k);
                }
            }
            GIP_LAUNCH_CHECK("pcg iteration");
            if (!fixed) {
                if ((rc = poll(S, false))) return rc;
                sample_read();
            }
        }
        if (S->rows) {
            if (k > 0) k_x_flush_rows<<<S->nchunk, kBlock, 0, st>>>(d, S->p, k - 1);
            k_fin_rows<<<1, kBlock, 0, st>>>(d);
        } else {
            if (k > 0) k_x_flush<<<S->grid, kBlock, 0, st>>>(d, pdir(k - 1), k - 1);
            k_fin<<<1, kBlock, 0, st>>>(d);
        }
        GIP_LAUNCH_CHECK("pcg final test");
    } else {
        // FP64 residual of x0 (with the coarse step), FP32 copies, first direction
        DISPATCH_W(A, k_m_resid, v, d);
        k_m_coarse<<<S->grid, kBlock, 0, st>>>(d, use_coarse);
        k_m_corr_fin<<<1, kBlock, 0, st>>>(d, 0, 1);
        k_m_start<<<S->grid, kBlock, 0, st>>>(d, S->p32);
        GIP_LAUNCH_CHECK("mixed setup");
        if ((rc = poll(S, true))) return rc;
        double last_rr = S->h_sc[S_TRUE_RR];
        const double it = o->inner_tol > 0.0 ? o->inner_tol : 0.1;
        const double inner2 = it * it;
        int k = 0, pending = 0;  // pending: the last iteration's correction update is not yet applied
        while (k < o->maxit && S->h_state[0] == ST_RUN) {
            int chunk = std::min(pollk, o->maxit - k);
            for (int j = 0; j < chunk; ++j, ++k) {
                if (S->fold) {
                    if (k > 0) k_m_dir<<<S->grid, kBlock, 0, st>>>(d, S->p32, k, pending);
                    const bool smp = sample_begin(k);
                    DISPATCH_WG(A, S->nchunk, k_m_spmv, v, d, S->p32);
                    sample_end(smp);
                    k_fold_dn<<<S->gdn, kBlock, 0, st>>>(d);
                    k_m_update<false><<<S->grid, kBlock, 0, st>>>(d, k);
                } else {
                    const bool smp = sample_begin(k);
                    if (k > 0)
                        DISPATCH_WG(A, S->nchunk, k_m_dirprod, v, d, pdir32(k - 1), pdir32(k), k, pending);
                    else
                        DISPATCH_WG(A, S->nchunk, k_m_spmv_small, v, d, S->p32);
                    sample_end(smp);
                    k_m_update<true><<<S->grid, kBlock, 0, st>>>(d, k);
                }
                pending = 1;
            }
            k_m_status<<<1, kBlock, 0, st>>>(d);
            GIP_LAUNCH_CHECK("mixed iteration");
            if ((rc = poll(S, true))) return rc;
            sample_read();
            if (S->h_state[0] != ST_RUN) break;
            const double rr32 = S->h_sc[S_RR], tol2 = S->h_sc[S_TOL2];
            const bool near_target = tol2 >= 0.0 && rr32 <= tol2;
            if (rr32 <= inner2 * last_rr || near_target || k >= o->maxit) {
                k_m_fold<<<S->grid, kBlock, 0, st>>>(d, pdir32(k - 1), k - 1);
                pending = 0;
                DISPATCH_W(A, k_m_resid, v, d);
                k_m_coarse<<<S->grid, kBlock, 0, st>>>(d, use_coarse);
                k_m_corr_fin<<<1, kBlock, 0, st>>>(d, k, 0);
                GIP_LAUNCH_CHECK("mixed correction");
                ++corrections;
                if ((rc = poll(S, true))) return rc;
                last_rr = S->h_sc[S_TRUE_RR];
            }
        }
    }
    if ((rc = poll(S, true))) return rc;
    sample_read();
    int status = GIP_OK;
    const int fin = S->h_state[0] & ST_MASK;
    if (fin == ST_BREAK)
        status = GIP_EBREAKDOWN;
    else if (o->tol > 0.0 && fin != ST_CONV)
        status = GIP_ENOCONV;
    if (stats) {
        stats->iters = S->h_state[1];
        stats->corrections = corrections;
        const double bb = S->h_sc[S_BB], rr = S->h_sc[S_RR];
        stats->relres = bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
        stats->status = status;
        stats->product_samples = (int32_t)samples.size();
        stats->product_ms = 0.0;
        if (!samples.empty()) {  // median (a launch that found the solve finished is short)
            std::nth_element(samples.begin(), samples.begin() + samples.size() / 2, samples.end());
            stats->product_ms = samples[samples.size() / 2];
        }
    }
    return status == GIP_EBREAKDOWN ? set_error(GIP_EBREAKDOWN, "PCG breakdown (p.Ap == 0 or non-finite step)") : GIP_OK;
}

int gip_pcg_solve(gip_pcg *S, const double *b, double *x, const double *x0, const gip_pcg_opts *o, gip_pcg_stats *stats) {
    if (!S || !b || !x || !x0 || !o || o->maxit < 0) return set_error(GIP_EINVAL, "gip_pcg_solve: bad argument");
    if (o->precond && S->A->diag_missing) return set_error(GIP_EINVAL, "gip_pcg_solve: Jacobi needs every diagonal entry");
    return pcg_solve_impl(S, b, x, x0, o, stats);
}

int gip_pcg_solve_noalloc(gip_ctx *c, const gip_mat *A, int32_t n_all, const double *b, double *x, const double *x0,
                          const gip_pcg_opts *o, gip_pcg_stats *stats) {
    if (!c || !A || !b || !x || !x0 || !o || o->maxit < 0)
        return set_error(GIP_EINVAL, "gip_pcg_solve_noalloc: bad argument");
    if (A->ctx != c) return set_error(GIP_EINVAL, "gip_pcg_solve_noalloc: operator belongs to another context");
    if (o->precision != GIP_PCG_FP64) return set_error(GIP_EINVAL, "gip_pcg_solve_noalloc: FP64 only");
    if (o->time_products) return set_error(GIP_EINVAL, "gip_pcg_solve_noalloc: no product timing");
    if (A->format != GIP_FMT_ELL && A->format != GIP_FMT_CSR && A->format != GIP_FMT_STENCIL)
        return set_error(GIP_EINVAL, "gip_pcg_solve_noalloc: operator must be ELL, stencil or CSR");
    if (A->nrows > kScratchRows || n_all > kScratchAll)
        return set_error(GIP_EINVAL, "gip_pcg_solve_noalloc: at most 262144 rows and 327680 vector entries");
    if (n_all < A->nrows || n_all < A->cols_needed)
        return set_error(GIP_EINVAL, "gip_pcg_solve_noalloc: vectors shorter than the operator's columns");
    if (o->precond && A->diag_missing)
        return set_error(GIP_EINVAL, "gip_pcg_solve_noalloc: Jacobi needs every diagonal entry");
    double *base = (c->device >= 0 && c->device < kScratchDevices) ? g_scratch[c->device] : nullptr;
    if (!base) return set_error(GIP_EINVAL, "gip_pcg_solve_noalloc: no scratch on this device (gip_init sets it up)");
    GIP_HIP_TRY(hipSetDevice(c->device));
    int h_state[2] = {0, 0};
    double h_sc[S_COUNT] = {};
    gip_pcg S{};
    S.ctx = c;
    S.A = A;
    S.n = A->nrows;
    S.n_all = n_all;
    S.grid = pcg_grid(A->nrows);
    S.nchunk = std::max(1, (A->nrows + kBlock - 1) / kBlock);
    S.fold = false;  // at most 1024 chunks: the fused two-kernel iteration
    S.gdn = S.nchunk;
    S.r = base;
    S.s = S.r + kScratchRows;
    S.Ap = S.s + kScratchRows;
    S.dinv = S.Ap + kScratchRows;
    S.p = S.dinv + kScratchRows;
    S.p2 = S.p + kScratchAll;
    double *zero = S.p2 + kScratchAll;
    S.part = zero;
    S.sc = S.part + (size_t)P_COUNT * kMaxRedBlocks;
    S.state = reinterpret_cast<int *>(S.sc + S_COUNT);
    S.h_state = h_state;
    S.h_sc = h_sc;
    // partial slots past each producer's grid, the scalars and the state
    // words start at zero (a previous call may have left other values)
    k_zero<<<(unsigned)((kScratchZero + kBlock - 1) / kBlock), kBlock, 0, c->stream>>>(kScratchZero, zero);
    GIP_LAUNCH_CHECK("gip_pcg_solve_noalloc: clear");
    if (!A->diag_missing) {
        int rc = gip_mat_diag_inv(A, S.dinv);
        if (rc) return rc;
    }
    return pcg_solve_impl(&S, b, x, x0, o, stats);
}

}  // extern "C"
