// Device helpers shared by the SpMV, PCG and batch kernels: a read-only
// view of an operator, the product operand, the stored-order row sum, and
// the XCD-contiguous row sweep.
#pragma once

#include <hip/hip_runtime.h>

#include "gip_internal.h"

namespace gip {

struct EllView {
    int fmt;  // GIP_FMT_ELL / GIP_FMT_CSR / GIP_FMT_STENCIL
    const int32_t *soff, *col, *rp;
    const double *val;
    const float *val32;
    // stencil layout
    const int32_t *st_off, *st_tcol;
    const StencilHdr *st_hdr;
    const StencilTail *st_tail;
    int st_w, tail_base, cmax;
    int ell_lg;  // sliced ELL: log2 of the rows per slice
    // symmetric form of the slot values (null when not in use)
    const int8_t *st_code;
    const double *sym;
    int sym_nu, sym_stride;  // arrays of sym_stride values (every slice's rows), one after another
    int sym_d[kSymMaxD];
};

static EllView view(const gip_mat *A) {
    EllView v{};
    const bool packed = A->format == GIP_FMT_ELL || A->format == GIP_FMT_STENCIL;
    v.fmt = A->format;
    v.soff = A->d_slice_off;
    v.col = A->format == GIP_FMT_ELL ? A->d_ell_col : A->d_col;
    v.rp = A->d_rowptr;
    v.val = packed ? A->d_ell_val : A->d_val;
    v.val32 = packed ? A->d_ell_val32 : A->d_val32;
    v.st_off = A->d_st_off;
    v.st_tcol = A->d_st_tcol;
    v.st_hdr = A->d_st_hdr;
    v.st_tail = A->d_st_tail;
    v.st_w = A->st_w;
    v.tail_base = (int)A->st_tail_base;
    v.cmax = A->cols_needed > 0 ? A->cols_needed - 1 : 0;
    v.ell_lg = A->ell_lg;
    v.st_code = A->d_st_code;
    v.sym = A->sym_ok > 0 ? A->d_sym_val : nullptr;
    v.sym_nu = A->sym_nu;
    v.sym_stride = A->nslices * 64;
    for (int j = 0; j < kSymMaxD; ++j) v.sym_d[j] = A->sym_d[j];
    return v;
}

// Kernel template width of an operator: W > 0 uniform-width ELL, W < 0 the
// stencil layout with -W slots, 0 the generic loops.
static inline int kernel_width(const gip_mat *A) {
    if (A->format == GIP_FMT_ELL) return A->uniform_width;
    if (A->format == GIP_FMT_STENCIL) return -A->st_w;
    return 0;
}

// Workgroup b of G is placed on XCD b % 8 by the dispatcher; renumber the
// workgroups so that each XCD sweeps one contiguous range of rows and the
// neighbour gathers of a row range stay in that XCD's L2.
__device__ inline int xcd_block(int b, int G) {
    const int x = b & 7, idx = b >> 3, q = G >> 3, r = G & 7;
    return x * q + (x < r ? x : r) + idx;
}

// The input vector of a product.  The row sums take any operand type with
// at(j) (the PCG also forms a direction on the fly as its operand).
template <class T>
struct Operand {
    const T *__restrict__ p;
    __device__ inline T at(int j) const { return p[j]; }
};

// The tail of a row in a slice of kind 2 (empty rows, or rows of several
// tail entries), located through the slice header's lane masks; false for
// an empty row.
template <class T, class V, class X>
__device__ inline bool stencil_tail_hdr(const EllView &A, const V *__restrict__ val, const X &x, int s, int row,
                                        T &sum) {
    const StencilHdr *h = A.st_hdr + s;
    const uint64_t bit = 1ull << (row & 63), below = bit - 1ull;
    if (h->empty & bit) return false;  // empty row: left untouched
    int t = h->tsoff;
    for (int j = 0; j < kStencilMaxTail; ++j) {
        const uint64_t mj = h->m[j];
        if (!(mj & bit)) break;
        const int e = t + __popcll(mj & below);
        sum = sum + x.at(A.st_tcol[e]) * val[A.tail_base + e];
        t += __popcll(mj);
    }
    return true;
}

// A slice's only tail entry (record kind 3): its lane, column and position
// are wavefront-uniform, so its value and operand are loads every lane of
// the wavefront shares, issued without waiting for a per-lane index.
template <class T, class V, class X>
__device__ inline void stencil_tail_one(const EllView &A, const V *__restrict__ val, const X &x, const StencilTail &tl,
                                        int row, T &sum) {
    const int lane_t = (tl.kind >> 8) & 63;
    const V tv = val[A.tail_base + tl.tsoff];
    const T xt = x.at((int)(uint32_t)tl.m0);
    if ((row & 63) == lane_t) sum = sum + xt * tv;
}

// Stencil-layout row sum: the slots in order (column row + offset, clamped
// into the vector for slots the row lacks, whose value is 0), then the
// row's tail in stored order.  The wavefront's 64 rows are one slice, so
// the slice's offsets and header are wavefront-uniform (scalar loads); a
// tail entry's position follows from the header's lane masks.
template <int SW, class T, class V, class X>
__device__ inline bool stencil_row(const EllView &A, const V *__restrict__ val, const X &x, int row, T &sum) {
    const int w = SW > 0 ? SW : A.st_w;
    const int s = __builtin_amdgcn_readfirstlane(row >> 6);
    const int32_t *off = A.st_off + (size_t)s * w;
    const V *vs = val + (size_t)s * w * 64 + (row & 63);
    if (SW > 0) {
        int c[SW > 0 ? SW : 1];
        V v[SW > 0 ? SW : 1];
#pragma unroll
        for (int k = 0; k < SW; ++k) {
            const int cc = row + off[k];
            c[k] = cc < 0 ? 0 : (cc > A.cmax ? A.cmax : cc);
            v[k] = __builtin_nontemporal_load(&vs[64 * k]);
        }
#pragma unroll
        for (int k = 0; k < SW; ++k) sum = sum + x.at(c[k]) * v[k];
    } else {
        for (int k = 0; k < w; ++k) {
            const int cc = row + off[k];
            sum = sum + x.at(cc < 0 ? 0 : (cc > A.cmax ? A.cmax : cc)) * vs[64 * k];
        }
    }
    // compact 16-byte record per slice; the 80-byte header only for the
    // rare slices with empty rows or rows of several tail entries
    const StencilTail tl = A.st_tail[s];
    const int kind = tl.kind & 0xff;
    if (kind == 3) {
        stencil_tail_one(A, val, x, tl, row, sum);
    } else if (kind == 1) {
        const uint64_t bit = 1ull << (row & 63);
        if (tl.m0 & bit) {
            const int e = tl.tsoff + __popcll(tl.m0 & (bit - 1ull));
            sum = sum + x.at(A.st_tcol[e]) * val[A.tail_base + e];
        }
    } else if (kind == 2) {
        return stencil_tail_hdr(A, val, x, s, row, sum);
    }
    return true;
}

// Stencil-layout row sum over the symmetric form of the slot values
// (gip_layout.h, plan_symmetric): the same values in the same order as
// stencil_row, a slot of offset -d taking its value from row - d's slot of
// offset +d.  The slice's 8 slot codes are one scalar load.
template <int SW, class T, class V, class X>
__device__ inline bool stencil_row_sym(const EllView &A, const V *__restrict__ val, const V *__restrict__ sym,
                                       const X &x, int row, T &sum) {
    const int s = __builtin_amdgcn_readfirstlane(row >> 6);
    const int lane = row & 63;
    const size_t stride = (size_t)A.sym_stride;
    const int32_t *off = A.st_off + (size_t)s * SW;
    const uint64_t codes = *reinterpret_cast<const uint64_t *>(A.st_code + (size_t)s * 8);
    int c[SW];
    V v[SW];
#pragma unroll
    for (int k = 0; k < SW; ++k) {
        const int cc = row + off[k];
        c[k] = cc < 0 ? 0 : (cc > A.cmax ? A.cmax : cc);
        // array and row of the slot's value from the (scalar) code, so that
        // every slot is one unconditional load, issued before any sum
        const int code = (int)(int8_t)(codes >> (8 * k));
        const bool lower = code >= kSymLower;
        const int j = lower ? code - kSymLower : (code < 0 ? 0 : code);
        const int d = !lower ? 0 : j == 0 ? A.sym_d[0] : j == 1 ? A.sym_d[1] : j == 2 ? A.sym_d[2] : A.sym_d[3];
        const int pr = row - d, prc = pr < 0 ? 0 : pr;
        const V u = sym[(size_t)j * stride + prc];
        v[k] = (code < 0 || pr < 0) ? V(0) : u;
    }
#pragma unroll
    for (int k = 0; k < SW; ++k) sum = sum + x.at(c[k]) * v[k];
    const StencilTail tl = A.st_tail[s];
    const int kind = tl.kind & 0xff;
    if (kind == 3) {
        stencil_tail_one(A, val, x, tl, row, sum);
    } else if (kind == 1) {
        const uint64_t bit = 1ull << lane;
        if (tl.m0 & bit) {
            const int e = tl.tsoff + __popcll(tl.m0 & (bit - 1ull));
            sum = sum + x.at(A.st_tcol[e]) * val[A.tail_base + e];
        }
    } else if (kind == 2) {
        return stencil_tail_hdr(A, val, x, s, row, sum);
    }
    return true;
}

template <int W, class T, class V, class X>
__device__ inline bool row_dot_sym(const EllView &A, const V *__restrict__ val, const V *__restrict__ sym,
                                   const X &x, int row, T &sum) {
    static_assert(W < 0, "the symmetric form is a stencil layout");
    sum = T(0);
    return stencil_row_sym<-W, T, V, X>(A, val, sym, x, row, sum);
}

// The operator's part of a stencil row (slot columns and values, and for a
// slice of kind 0, 1 or 3 its one possible tail entry), loaded apart from the
// operand so that a kernel can request it before its prologue; the operand
// reads follow in stencil_row_fin, in the order stencil_row uses.
template <int SW, class V>
struct StencilPre {
    int c[SW];
    V v[SW];
    int kind, slice, tcol;  // tcol < 0: no tail entry (kinds 0, 1 and 3)
    V tval;
};

template <int SW, class V>
__device__ inline void stencil_row_pre(const EllView &A, const V *__restrict__ val, int row, StencilPre<SW, V> &q) {
    const int s = __builtin_amdgcn_readfirstlane(row >> 6);
    const int32_t *off = A.st_off + (size_t)s * SW;
    const V *vs = val + (size_t)s * SW * 64 + (row & 63);
#pragma unroll
    for (int k = 0; k < SW; ++k) {
        const int cc = row + off[k];
        q.c[k] = cc < 0 ? 0 : (cc > A.cmax ? A.cmax : cc);
        q.v[k] = __builtin_nontemporal_load(&vs[64 * k]);
    }
    const StencilTail tl = A.st_tail[s];  // the 16-byte record in one load
    q.slice = s;
    q.kind = tl.kind & 0xff;
    q.tcol = -1;
    q.tval = V(0);
    if (q.kind == 3) {
        if ((row & 63) == ((tl.kind >> 8) & 63)) {
            q.tcol = (int)(uint32_t)tl.m0;
            q.tval = val[A.tail_base + tl.tsoff];
        }
    } else if (q.kind == 1) {
        const uint64_t bit = 1ull << (row & 63);
        if (tl.m0 & bit) {
            const int e = tl.tsoff + __popcll(tl.m0 & (bit - 1ull));
            q.tcol = A.st_tcol[e];
            q.tval = val[A.tail_base + e];
        }
    }
}

template <int SW, class T, class V, class X>
__device__ inline bool stencil_row_fin(const EllView &A, const V *__restrict__ val, const X &x, int row,
                                       const StencilPre<SW, V> &q, T &sum) {
    sum = T(0);
#pragma unroll
    for (int k = 0; k < SW; ++k) sum = sum + x.at(q.c[k]) * q.v[k];
    if (q.kind == 1 || q.kind == 3) {
        if (q.tcol >= 0) sum = sum + x.at(q.tcol) * q.tval;
    } else if (q.kind == 2) {
        return stencil_tail_hdr(A, val, x, q.slice, row, sum);
    }
    return true;
}

// Row sum in stored order; returns false for an empty row.
template <int W, class T, class V, class X>
__device__ inline bool row_dot(const EllView &A, const V *__restrict__ val, const X &x, int row, T &sum) {
    sum = T(0);
    if (W < 0) {
        return stencil_row<(W < 0 ? -W : 0), T, V, X>(A, val, x, row, sum);
    } else if (W == 0 && A.fmt == GIP_FMT_STENCIL) {
        return stencil_row<0, T, V, X>(A, val, x, row, sum);
    } else if (W > 0) {
        const int lg = A.ell_lg;
        const int base = A.soff[row >> lg] + (row & ((1 << lg) - 1));
        int c[W > 0 ? W : 1];
        V v[W > 0 ? W : 1];
#pragma unroll
        for (int k = 0; k < W; ++k) {
            c[k] = __builtin_nontemporal_load(&A.col[base + (k << lg)]);
            v[k] = __builtin_nontemporal_load(&val[base + (k << lg)]);
        }
#pragma unroll
        for (int k = 0; k < W; ++k) sum = sum + x.at(c[k]) * v[k];
        return true;
    } else if (A.fmt == GIP_FMT_ELL) {
        const int lg = A.ell_lg;
        const int s = row >> lg, base = A.soff[s] + (row & ((1 << lg) - 1));
        const int width = (A.soff[s + 1] - A.soff[s]) >> lg;
        bool any = false;
        for (int k = 0; k < width; ++k) {
            const int cc = A.col[base + (k << lg)];
            if (cc < 0) break;
            sum = sum + x.at(cc) * val[base + (k << lg)];
            any = true;
        }
        return any;
    } else {
        const int jb = A.rp[row], je = A.rp[row + 1];
        for (int j = jb; j < je; ++j) sum = sum + x.at(A.col[j]) * val[j];
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

#define FOR_ROWS(n, row)                                          \
    for (RowSweep rs__(n); rs__.c < rs__.c1; rs__.c += rs__.step) \
        for (int row = rs__.c * kBlock + (int)threadIdx.x; row < (rs__.c + 1) * kBlock && row < (n); row = (n))

// The rows FOR_ROWS gives this thread, as a cursor, so that a kernel can
// request its first row's operands before its prologue and each next row's
// right after storing the current one.  Only the last chunk of the matrix
// holds rows >= n, so the first such row ends the sweep.
struct ThreadRows {
    RowSweep rs;
    int row, n;
    __device__ explicit ThreadRows(int n_) : rs(n_), n(n_) { row = rs.c * kBlock + (int)threadIdx.x; }
    __device__ bool valid() const { return rs.c < rs.c1 && row < n; }
    __device__ void next() {
        rs.c += rs.step;
        row = rs.c * kBlock + (int)threadIdx.x;
    }
};

}  // namespace gip
