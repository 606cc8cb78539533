// Sparse operators of libgip_hip: device layouts (sliced ELL, CSR,
// diagonal) and the SpMV family of GIP_MatrixMulticore.cpp:64-143.
//
// Every kernel sums a row's products in stored (CSR) order with separately
// rounded multiplies and adds (the library is built without contraction),
// starting from 0.0, and leaves empty rows untouched: results are bitwise
// those of the CPU specification.
//
// Sliced ELL: slice s holds rows [H s, H s + H), H = 2^lg = 32, 64 (one
// wavefront, the default) or 128 rows; entry k of row H s + l is stored at
// slice_off[s] + H k + l, so the lanes of a wavefront read consecutive
// values and indices per entry.  Rows shorter than the slice width are
// padded with column -1.
//
// Stencil: the rows of a slice share per-slot column offsets (see
// gip_layout.h), so a slot stores only values; the wavefront reads 64
// consecutive values per slot and gathers x at row + offset.
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "gip_internal.h"
#include "gip_layout.h"
#include "gip_rows.h"
#include "gip_spmv.h"

namespace gip {

template <int MODE, int W>
__global__ __launch_bounds__(kBlock) void k_ell(int nrows, int lg, const int32_t *__restrict__ soff,
                                               const int32_t *__restrict__ col, const double *__restrict__ val,
                                               const double *__restrict__ x, double *__restrict__ y, double a,
                                               double b, const int32_t *__restrict__ scatter) {
    const int row = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (row >= nrows) return;
    const int s = row >> lg, l = row & ((1 << lg) - 1);
    const int base = soff[s] + l;
    double sum = 0.0;
    bool any;
    if (W > 0) {
        int c[W > 0 ? W : 1];
        double v[W > 0 ? W : 1];
#pragma unroll
        for (int k = 0; k < W; ++k) {
            c[k] = __builtin_nontemporal_load(&col[base + (k << lg)]);
            v[k] = __builtin_nontemporal_load(&val[base + (k << lg)]);
        }
#pragma unroll
        for (int k = 0; k < W; ++k) sum = sum + x[c[k]] * v[k];
        any = true;
    } else {
        const int width = (soff[s + 1] - soff[s]) >> lg;
        any = false;
        for (int k = 0; k < width; ++k) {
            const int cc = col[base + (k << lg)];
            if (cc < 0) break;
            sum = sum + x[cc] * val[base + (k << lg)];
            any = true;
        }
    }
    if (!any) return;
    const int out = scatter ? scatter[row] : row;
    if (MODE == 0)
        y[out] = sum;
    else
        y[out] = a * sum + b * y[out];
}

// Stencil-layout product (W < 0: -W slots unrolled; 0: runtime width).
// Workgroups in dispatch order (consecutive 256-row chunks on consecutive
// XCDs, so the whole chip streams one narrow window of the value array):
// 4.4% faster than the XCD-contiguous order on the 256^3 Laplacian, whose
// L2 reuse of the neighbour gathers it gives up.  The FP32 product and the
// PCG's products keep the XCD-contiguous order (measured faster or equal).
template <int MODE, int W>
__global__ __launch_bounds__(kBlock) void k_stencil(EllView A, int nrows, const double *__restrict__ x,
                                                   double *__restrict__ y, double a, double b,
                                                   const int32_t *__restrict__ scatter) {
    const int row = blockIdx.x * kBlock + threadIdx.x;
    if (row >= nrows) return;
    Operand<double> xo{x};
    double sum;
    if (!row_dot<W, double, double>(A, A.val, xo, row, sum)) return;
    const int out = scatter ? scatter[row] : row;
    if (MODE == 0)
        y[out] = sum;
    else
        y[out] = a * sum + b * y[out];
}

template <int W>
__global__ __launch_bounds__(kBlock) void k_stencil32(EllView A, int nrows, const float *__restrict__ x,
                                                     float *__restrict__ y) {
    const int row = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (row >= nrows) return;
    Operand<float> xo{x};
    float sum;
    row_dot<W, float, float>(A, A.val32, xo, row, sum);
    y[row] = sum;
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_csr(int nrows, const int32_t *__restrict__ rp,
                                               const int32_t *__restrict__ col, const double *__restrict__ val,
                                               const double *__restrict__ x, double *__restrict__ y, double a,
                                               double b, const int32_t *__restrict__ scatter) {
    const int row = blockIdx.x * kBlock + threadIdx.x;
    if (row >= nrows) return;
    const int jb = rp[row], je = rp[row + 1];
    if (jb == je) return;
    double sum = 0.0;
    for (int j = jb; j < je; ++j) sum = sum + x[col[j]] * val[j];
    const int out = scatter ? scatter[row] : row;
    if (MODE == 0)
        y[out] = sum;
    else
        y[out] = a * sum + b * y[out];
}

// CSR with R lanes per row (64/R rows per wavefront): the lanes load a
// row's entries cooperatively (coalesced) and form the products; the
// group's first lane then adds the products in stored order, so the sum
// is the same as the thread-per-row kernel's.
template <int MODE, int R>
__global__ __launch_bounds__(kBlock) void k_csr_vec(int nrows, const int32_t *__restrict__ rp,
                                                   const int32_t *__restrict__ col, const double *__restrict__ val,
                                                   const double *__restrict__ x, double *__restrict__ y, double a,
                                                   double b, const int32_t *__restrict__ scatter) {
    const int gid = blockIdx.x * kBlock + threadIdx.x;
    const int row = gid / R, lane = threadIdx.x % R, lead = threadIdx.x & ~(R - 1) & 63;
    const bool live = row < nrows;
    const int jb = live ? rp[row] : 0, je = live ? rp[row + 1] : 0;
    double sum = 0.0;
    // every lane of the wavefront runs the same number of rounds
    int rounds = (je - jb + R - 1) / R;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) rounds = max(rounds, __shfl_xor(rounds, off, 64));
    for (int t = 0; t < rounds; ++t) {
        const int j = jb + t * R + lane;
        const double prod = j < je ? x[col[j]] * val[j] : 0.0;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const double pq = __shfl(prod, lead + q, 64);
            if (jb + t * R + q < je) sum = sum + pq;
        }
    }
    if (!live || lane != 0 || jb == je) return;
    const int out = scatter ? scatter[row] : row;
    if (MODE == 0)
        y[out] = sum;
    else
        y[out] = a * sum + b * y[out];
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_diag(int nrows, const double *__restrict__ d, const double *__restrict__ x,
                                                double *__restrict__ y, double a, double b,
                                                const int32_t *__restrict__ scatter) {
    const int row = blockIdx.x * kBlock + threadIdx.x;
    if (row >= nrows) return;
    double sum = 0.0;
    sum = sum + x[row] * d[row];
    const int out = scatter ? scatter[row] : row;
    if (MODE == 0)
        y[out] = sum;
    else
        y[out] = a * sum + b * y[out];
}

// FP32 products for the mixed-precision inner solver (no empty-row rule needed:
// rows without entries produce 0).
__global__ __launch_bounds__(kBlock) void k_ell32(int nrows, int lg, const int32_t *__restrict__ soff,
                                                 const int32_t *__restrict__ col, const float *__restrict__ val,
                                                 const float *__restrict__ x, float *__restrict__ y) {
    const int row = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (row >= nrows) return;
    const int s = row >> lg, l = row & ((1 << lg) - 1);
    const int base = soff[s] + l, width = (soff[s + 1] - soff[s]) >> lg;
    float sum = 0.0f;
    for (int k = 0; k < width; ++k) {
        const int cc = col[base + (k << lg)];
        if (cc < 0) break;
        sum = sum + x[cc] * val[base + (k << lg)];
    }
    y[row] = sum;
}

__global__ __launch_bounds__(kBlock) void k_csr32(int nrows, const int32_t *__restrict__ rp,
                                                 const int32_t *__restrict__ col, const float *__restrict__ val,
                                                 const float *__restrict__ x, float *__restrict__ y) {
    const int row = blockIdx.x * kBlock + threadIdx.x;
    if (row >= nrows) return;
    float sum = 0.0f;
    for (int j = rp[row]; j < rp[row + 1]; ++j) sum = sum + x[col[j]] * val[j];
    y[row] = sum;
}

__global__ __launch_bounds__(kBlock) void k_gather_vals(int n, const double *__restrict__ src,
                                                       const int32_t *__restrict__ map, double *__restrict__ dst) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) dst[i] = src[map[i]];
}

__global__ __launch_bounds__(kBlock) void k_scatter_vals(int n, const double *__restrict__ src,
                                                        const int32_t *__restrict__ map, double *__restrict__ dst) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) dst[map[i]] = src[i];
}

__global__ __launch_bounds__(kBlock) void k_inv(int n, const double *__restrict__ d, double *__restrict__ out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) out[i] = 1.0 / d[i];
}

__global__ __launch_bounds__(kBlock) void k_to_f32(int64_t n, const double *__restrict__ src, float *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) dst[i] = (float)src[i];
}

static inline int nblocks(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

// The Jacobi diagonal from the stored values after they change: the value
// at each row's diagonal position (its last diagonal entry).
__global__ __launch_bounds__(kBlock) void k_diag_refresh(int n, const int32_t *__restrict__ src,
                                                         const double *__restrict__ store, double *__restrict__ diag) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const int32_t j = src[i];
        diag[i] = j >= 0 ? store[j] : 0.0;
    }
}

// The symmetric form's arrays from the slot values: array j (values
// [j rows_pad, (j+1) rows_pad)) holds at row r the value of the row's slot
// of code j, 0 when its slice has none.  One thread per row, padding rows
// of the last slice included.
__global__ __launch_bounds__(kBlock) void k_sym_make(int64_t rows_pad, int w, int nu, const int8_t *__restrict__ code,
                                                     const double *__restrict__ val, double *__restrict__ sym) {
    const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (r >= rows_pad) return;
    const int64_t sl = r >> 6;
    const int lane = (int)(r & 63);
    const uint64_t codes = *reinterpret_cast<const uint64_t *>(code + sl * 8);
    for (int j = 0; j < nu; ++j) {
        double v = 0.0;
        for (int k = 0; k < w; ++k)
            if ((int)(int8_t)(codes >> (8 * k)) == j) v = val[(sl * w + k) * 64 + lane];
        sym[j * rows_pad + r] = v;
    }
}

// Every slot the symmetric form derives must hold the bits it derives: a
// slot of offset -d_j the value of array j at row - d_j (0 before row 0),
// an unused slot +0.  Any mismatch raises the flag.
__global__ __launch_bounds__(kBlock) void k_sym_check(int nrows, int64_t rows_pad, int w, int d0, int d1, int d2, int d3,
                                                      const int8_t *__restrict__ code, const double *__restrict__ val,
                                                      const double *__restrict__ sym, int32_t *__restrict__ flag) {
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= nrows) return;
    const int64_t sl = r >> 6;
    const int lane = r & 63;
    const uint64_t codes = *reinterpret_cast<const uint64_t *>(code + sl * 8);
    bool bad = false;
    for (int k = 0; k < w; ++k) {
        const int cd = (int)(int8_t)(codes >> (8 * k));
        if (cd >= 0 && cd < kSymLower) continue;  // stored as is
        double want = 0.0;
        if (cd >= kSymLower) {
            const int j = cd - kSymLower;
            const int d = j == 0 ? d0 : j == 1 ? d1 : j == 2 ? d2 : d3;
            const int pr = r - d;
            if (pr >= 0) want = sym[j * rows_pad + pr];
        }
        const double got = val[(sl * w + k) * 64 + lane];
        if (__double_as_longlong(got) != __double_as_longlong(want)) bad = true;
    }
    if (bad) atomicOr(flag, 1);
}

template <class T>
static int upload(gip_ctx *c, T **dptr, const T *host, size_t count) {
    GIP_HIP_TRY(hipMalloc(dptr, sizeof(T) * (count ? count : 1)));
    if (count) GIP_HIP_TRY(hipMemcpyAsync(*dptr, host, sizeof(T) * count, hipMemcpyHostToDevice, c->stream));
    return GIP_OK;
}

static void free_mat_arrays(gip_mat *A) {
    void *ptrs[] = {A->d_rowptr,   A->d_col,     A->d_val,      A->d_val32,    A->d_slice_off, A->d_ell_col,
                    A->d_ell_val,  A->d_ell_val32, A->d_csr2ell, A->d_diag,     A->d_tsrc,      A->d_tdst,
                    A->d_st_off,   A->d_st_hdr,   A->d_st_tail,  A->d_st_tcol,  A->d_st_code,   A->d_sym_val,
                    A->d_sym_flag, A->d_diag_src};
    for (void *p : ptrs)
        if (p) hipFree(p);
}

// Dispatch y (= or +=) A x in FP64.  scatter: optional output index per row.
int spmv_launch(gip_ctx *c, const gip_mat *A, const double *x, double *y, double a, double b, int mode,
                const int32_t *scatter) {
    if (A->nrows == 0) return GIP_OK;
    const int g = nblocks(A->nrows);
    hipStream_t st = c->stream;
    if (A->format == GIP_FMT_ELL) {
#define GIP_ELL_CASE(Wc)                                                                                        \
    case Wc:                                                                                                    \
        if (mode == 0)                                                                                          \
            k_ell<0, Wc><<<g, kBlock, 0, st>>>(A->nrows, A->ell_lg, A->d_slice_off, A->d_ell_col, A->d_ell_val, x,  \
                                               y, a, b, scatter);                                               \
        else                                                                                                    \
            k_ell<1, Wc><<<g, kBlock, 0, st>>>(A->nrows, A->ell_lg, A->d_slice_off, A->d_ell_col, A->d_ell_val, x,  \
                                               y, a, b, scatter);                                               \
        break;
        switch (A->uniform_width) {
            GIP_ELL_CASE(1)
            GIP_ELL_CASE(2)
            GIP_ELL_CASE(3)
            GIP_ELL_CASE(4)
            GIP_ELL_CASE(5)
            GIP_ELL_CASE(6)
            GIP_ELL_CASE(7)
            GIP_ELL_CASE(8)
            default:
                if (mode == 0)
                    k_ell<0, 0><<<g, kBlock, 0, st>>>(A->nrows, A->ell_lg, A->d_slice_off, A->d_ell_col, A->d_ell_val,
                                                     x, y, a, b, scatter);
                else
                    k_ell<1, 0><<<g, kBlock, 0, st>>>(A->nrows, A->ell_lg, A->d_slice_off, A->d_ell_col, A->d_ell_val,
                                                     x, y, a, b, scatter);
        }
#undef GIP_ELL_CASE
    } else if (A->format == GIP_FMT_STENCIL) {
        const EllView v = view(A);
#define GIP_ST_CASE(Wc)                                                                   \
    case Wc:                                                                              \
        if (mode == 0)                                                                    \
            k_stencil<0, Wc><<<g, kBlock, 0, st>>>(v, A->nrows, x, y, a, b, scatter);     \
        else                                                                              \
            k_stencil<1, Wc><<<g, kBlock, 0, st>>>(v, A->nrows, x, y, a, b, scatter);     \
        break;
        switch (-A->st_w) {
            GIP_ST_CASE(-2)
            GIP_ST_CASE(-3)
            GIP_ST_CASE(-4)
            GIP_ST_CASE(-6)
            GIP_ST_CASE(-7)
            GIP_ST_CASE(-8)
            default:
                if (mode == 0)
                    k_stencil<0, 0><<<g, kBlock, 0, st>>>(v, A->nrows, x, y, a, b, scatter);
                else
                    k_stencil<1, 0><<<g, kBlock, 0, st>>>(v, A->nrows, x, y, a, b, scatter);
        }
#undef GIP_ST_CASE
    } else if (A->format == GIP_FMT_DIAG) {
        if (mode == 0)
            k_diag<0><<<g, kBlock, 0, st>>>(A->nrows, A->d_diag, x, y, a, b, scatter);
        else
            k_diag<1><<<g, kBlock, 0, st>>>(A->nrows, A->d_diag, x, y, a, b, scatter);
    } else if (A->lanes > 1) {
        const int gv = (int)(((int64_t)A->nrows * A->lanes + kBlock - 1) / kBlock);
#define GIP_VEC_CASE(Rc)                                                                                            \
    case Rc:                                                                                                        \
        if (mode == 0)                                                                                              \
            k_csr_vec<0, Rc><<<gv, kBlock, 0, st>>>(A->nrows, A->d_rowptr, A->d_col, A->d_val, x, y, a, b, scatter); \
        else                                                                                                        \
            k_csr_vec<1, Rc><<<gv, kBlock, 0, st>>>(A->nrows, A->d_rowptr, A->d_col, A->d_val, x, y, a, b, scatter); \
        break;
        switch (A->lanes) {
            GIP_VEC_CASE(2)
            GIP_VEC_CASE(4)
            GIP_VEC_CASE(8)
            default:
                return set_error(GIP_EINVAL, "unsupported lanes per row");
        }
#undef GIP_VEC_CASE
    } else {
        if (mode == 0)
            k_csr<0><<<g, kBlock, 0, st>>>(A->nrows, A->d_rowptr, A->d_col, A->d_val, x, y, a, b, scatter);
        else
            k_csr<1><<<g, kBlock, 0, st>>>(A->nrows, A->d_rowptr, A->d_col, A->d_val, x, y, a, b, scatter);
    }
    GIP_LAUNCH_CHECK("spmv kernel");
    return GIP_OK;
}

int spmv32_launch(gip_ctx *c, const gip_mat *A, const float *x, float *y) {
    if (A->nrows == 0) return GIP_OK;
    if (A->format == GIP_FMT_ELL) {
        if (!A->d_ell_val32) return set_error(GIP_EINVAL, "operator has no FP32 values");
        k_ell32<<<nblocks(A->nrows), kBlock, 0, c->stream>>>(A->nrows, A->ell_lg, A->d_slice_off, A->d_ell_col,
                                                             A->d_ell_val32, x, y);
    } else if (A->format == GIP_FMT_STENCIL) {
        if (!A->d_ell_val32) return set_error(GIP_EINVAL, "operator has no FP32 values");
        const EllView v = view(A);
        const int g = nblocks(A->nrows);
        switch (-A->st_w) {
            case -7: k_stencil32<-7><<<g, kBlock, 0, c->stream>>>(v, A->nrows, x, y); break;
            case -8: k_stencil32<-8><<<g, kBlock, 0, c->stream>>>(v, A->nrows, x, y); break;
            default: k_stencil32<0><<<g, kBlock, 0, c->stream>>>(v, A->nrows, x, y); break;
        }
    } else if (A->format == GIP_FMT_CSR) {
        if (!A->d_val32) return set_error(GIP_EINVAL, "operator has no FP32 values");
        k_csr32<<<nblocks(A->nrows), kBlock, 0, c->stream>>>(A->nrows, A->d_rowptr, A->d_col, A->d_val32, x, y);
    } else {
        return set_error(GIP_EINVAL, "FP32 SpMV needs an ELL, stencil or CSR operator");
    }
    GIP_LAUNCH_CHECK("spmv32 kernel");
    return GIP_OK;
}

int mat_sym_prepare(gip_ctx *c, gip_mat *A, bool *ok) {
    *ok = false;
    if (A->format != GIP_FMT_STENCIL || A->sym_nu <= 0 || !A->d_st_code || A->nrows <= 0) return GIP_OK;
    if (A->sym_version == A->val_version + 1) {
        *ok = A->sym_ok > 0;
        return GIP_OK;
    }
    const int64_t rows_pad = (int64_t)A->nslices * 64, n = rows_pad * A->sym_nu;
    if (!A->d_sym_val) GIP_HIP_TRY(hipMalloc(&A->d_sym_val, sizeof(double) * n));
    if (!A->d_sym_flag) GIP_HIP_TRY(hipMalloc(&A->d_sym_flag, sizeof(int32_t)));
    k_sym_make<<<nblocks(rows_pad), kBlock, 0, c->stream>>>(rows_pad, A->st_w, A->sym_nu, A->d_st_code, A->d_ell_val,
                                                           A->d_sym_val);
    GIP_HIP_TRY(hipMemsetAsync(A->d_sym_flag, 0, sizeof(int32_t), c->stream));
    k_sym_check<<<nblocks(A->nrows), kBlock, 0, c->stream>>>(A->nrows, rows_pad, A->st_w, A->sym_d[0], A->sym_d[1],
                                                            A->sym_d[2], A->sym_d[3], A->d_st_code, A->d_ell_val,
                                                            A->d_sym_val, A->d_sym_flag);
    GIP_LAUNCH_CHECK("symmetric slot values");
    int32_t flag = 1;
    GIP_HIP_TRY(hipMemcpyAsync(&flag, A->d_sym_flag, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    GIP_HIP_TRY(hipStreamSynchronize(c->stream));
    A->sym_version = A->val_version + 1;
    if (flag) {
        // not symmetric: the arrays go until the values change again
        GIP_HIP_TRY(hipFree(A->d_sym_val));
        A->d_sym_val = nullptr;
        A->sym_ok = -1;
        return GIP_OK;
    }
    A->sym_ok = 1;
    *ok = true;
    return GIP_OK;
}

// After a change of an operator's stored values: the Jacobi diagonal
// follows them (a DIAG operator stores it directly).
static int diag_refresh(gip_ctx *c, gip_mat *A) {
    if (A->format == GIP_FMT_DIAG || !A->d_diag_src || A->nrows <= 0) return GIP_OK;
    const double *store = (A->format == GIP_FMT_ELL || A->format == GIP_FMT_STENCIL) ? A->d_ell_val : A->d_val;
    k_diag_refresh<<<nblocks(A->nrows), kBlock, 0, c->stream>>>(A->nrows, A->d_diag_src, store, A->d_diag);
    GIP_LAUNCH_CHECK("diagonal refresh");
    return GIP_OK;
}

}  // namespace gip

using namespace gip;

extern "C" {

int gip_mat_create(gip_ctx *c, int32_t nrows, int32_t ncols, int32_t nnz, const int32_t *rowptr, const int32_t *col,
                   const double *val, int32_t flags, gip_mat **out) {
    if (!c || !out || nrows < 0 || nnz < 0 || !rowptr || (nnz && (!col || !val)))
        return set_error(GIP_EINVAL, "gip_mat_create: bad argument");
    *out = nullptr;
    if (rowptr[0] != 0 || rowptr[nrows] != nnz) return set_error(GIP_EINVAL, "gip_mat_create: rowptr does not span nnz");
    for (int32_t i = 0; i < nrows; ++i)
        if (rowptr[i + 1] < rowptr[i]) return set_error(GIP_EINVAL, "gip_mat_create: rowptr not monotone");
    int32_t maxc = -1;
    for (int32_t j = 0; j < nnz; ++j) {
        if (col[j] < 0) return set_error(GIP_EINVAL, "gip_mat_create: negative column index");
        maxc = std::max(maxc, col[j]);
    }
    GIP_HIP_TRY(hipSetDevice(c->device));
    gip_mat *A = new (std::nothrow) gip_mat();
    if (!A) return set_error(GIP_ENOMEM, "gip_mat_create: allocation failed");
    A->ctx = c;
    A->nrows = nrows;
    A->ncols = ncols;
    A->nnz = nnz;
    A->cols_needed = maxc + 1;
    A->lanes = 1;
    const int32_t want = flags & 0xff;
    const bool fp32 = (flags & GIP_MAT_FP32) != 0;

    // diagonal value per row (last diagonal entry wins, GIP_MatrixMulticore.cpp:191-204)
    std::vector<double> diag((size_t)nrows, 0.0);
    std::vector<int32_t> diag_src((size_t)nrows, -1);  // CSR index of the row's last diagonal entry
    A->diag_missing = 0;
    bool is_diag = (nnz == nrows);
    for (int32_t i = 0; i < nrows; ++i) {
        bool found = false;
        for (int32_t j = rowptr[i]; j < rowptr[i + 1]; ++j)
            if (col[j] == i) {
                diag[(size_t)i] = val[j];
                diag_src[(size_t)i] = j;
                found = true;
            }
        if (!found) A->diag_missing = 1;
        if (is_diag && !(rowptr[i] == i && col[i] == i)) is_diag = false;
    }
    // slice widths (H rows per slice)
    const int lg = (flags & GIP_MAT_ELL_SLICE32) ? 5 : (flags & GIP_MAT_ELL_SLICE128) ? 7 : 6;
    const int32_t H = 1 << lg;
    A->ell_lg = lg;
    int32_t ns = (nrows + H - 1) / H;
    std::vector<int32_t> soff((size_t)ns + 1, 0);
    int32_t uw = -1;
    bool uniform = true;
    int64_t stored = 0;
    for (int32_t s = 0; s < ns; ++s) {
        int32_t w = 0;
        for (int32_t r = s * H; r < std::min(nrows, (s + 1) * H); ++r) {
            int32_t len = rowptr[r + 1] - rowptr[r];
            w = std::max(w, len);
            if (uw < 0) uw = len;
            if (len != uw) uniform = false;
        }
        stored += (int64_t)w * H;
        if (stored > INT32_MAX) uniform = false;
        soff[(size_t)s + 1] = (int32_t)std::min<int64_t>(stored, INT32_MAX);
    }
    // stencil layout: planned when automatic or asked for; chosen automatically
    // when it streams clearly fewer bytes than ELL and CSR would
    StencilPlan plan;
    bool stencil_ok = false;
    if ((want == GIP_FMT_AUTO && !(is_diag && nrows > 0)) || want == GIP_FMT_STENCIL)
        stencil_ok = plan_stencil(nrows, nnz, rowptr, col, plan);
    int32_t fmt = want;
    if (fmt == GIP_FMT_STENCIL && !stencil_ok) fmt = GIP_FMT_AUTO;
    if (fmt == GIP_FMT_AUTO) {
        const bool ell_ok = stored <= INT32_MAX && stored <= (int64_t)(1.2 * nnz) + H * 8;
        const int64_t csr_bytes = 12 * (int64_t)nnz + 8 * (int64_t)nrows;
        const int64_t ell_bytes = ell_ok ? 12 * stored + 4 * (int64_t)ns : csr_bytes;
        if (is_diag && nrows > 0)
            fmt = GIP_FMT_DIAG;
        else if (stencil_ok && (double)stencil_bytes(plan, nrows) < 0.85 * (double)std::min(ell_bytes, csr_bytes))
            fmt = GIP_FMT_STENCIL;
        else if (ell_ok)
            fmt = GIP_FMT_ELL;
        else
            fmt = GIP_FMT_CSR;
    }
    if (fmt == GIP_FMT_DIAG && !is_diag) fmt = GIP_FMT_CSR;
    if (fmt == GIP_FMT_ELL && stored > INT32_MAX) fmt = GIP_FMT_CSR;
    A->format = fmt;
    int rc = GIP_OK;
    // CSR structure always kept (transposes, value maps)
    rc = upload(c, &A->d_rowptr, rowptr, (size_t)nrows + 1);
    if (!rc) rc = upload(c, &A->d_diag, diag.data(), diag.size());
    // the diagonal's positions in the value storage, to follow value changes
    if (!rc && fmt != GIP_FMT_DIAG) {
        for (int32_t i = 0; i < nrows; ++i) {
            int32_t &j = diag_src[(size_t)i];
            if (j < 0) continue;
            if (fmt == GIP_FMT_ELL)
                j = (int32_t)(soff[(size_t)(i / H)] + (int64_t)H * (j - rowptr[i]) + i % H);
            else if (fmt == GIP_FMT_STENCIL)
                j = plan.pos[(size_t)j];
        }
        rc = upload(c, &A->d_diag_src, diag_src.data(), diag_src.size());
    }
    if (!rc && fmt == GIP_FMT_ELL) {
        A->nslices = ns;
        A->ell_stored = stored;
        A->uniform_width = (uniform && uw > 0 && uw <= 8 && (int64_t)uw * H * ns == stored) ? uw : 0;
        if (uniform && uw > 0 && uw <= 8 && nrows % H != 0) A->uniform_width = uw;  // tail rows never run
        std::vector<int32_t> ecol((size_t)stored, -1);
        std::vector<double> evalv((size_t)stored, 0.0);
        std::vector<int32_t> c2e((size_t)nnz);
        for (int32_t r = 0; r < nrows; ++r) {
            int32_t s = r / H, l = r % H;
            for (int32_t j = rowptr[r]; j < rowptr[r + 1]; ++j) {
                int64_t pos = soff[(size_t)s] + (int64_t)H * (j - rowptr[r]) + l;
                ecol[(size_t)pos] = col[j];
                evalv[(size_t)pos] = val[j];
                c2e[(size_t)j] = (int32_t)pos;
            }
        }
        // padding slots of the tail slice must not be read by a uniform kernel: rows >= nrows return early
        rc = upload(c, &A->d_slice_off, soff.data(), soff.size());
        if (!rc) rc = upload(c, &A->d_ell_col, ecol.data(), ecol.size());
        if (!rc) rc = upload(c, &A->d_ell_val, evalv.data(), evalv.size());
        if (!rc) rc = upload(c, &A->d_csr2ell, c2e.data(), c2e.size());
        if (!rc) {
            // the host arrays above are uploaded asynchronously: wait before they go
            hipError_t e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) rc = hip_fail(e, "ELL upload");
        }
        if (!rc && fp32) {
            hipError_t e = hipMalloc(&A->d_ell_val32, sizeof(float) * (size_t)std::max<int64_t>(stored, 1));
            if (e != hipSuccess) rc = hip_fail(e, "hipMalloc fp32 values");
            else if (stored) k_to_f32<<<nblocks(stored), kBlock, 0, c->stream>>>(stored, A->d_ell_val, A->d_ell_val32);
        }
        // CSR columns for transposes
        if (!rc) rc = upload(c, &A->d_col, col, (size_t)nnz);
    } else if (!rc && fmt == GIP_FMT_STENCIL) {
        A->nslices = plan.ns;
        A->ell_stored = plan.stored;
        A->st_w = plan.w;
        A->st_tail_base = plan.tail_base;
        A->st_tail_nnz = (int32_t)plan.tcol.size();
        std::vector<double> sval((size_t)plan.stored, 0.0);
        for (int32_t j = 0; j < nnz; ++j) sval[(size_t)plan.pos[(size_t)j]] = val[j];
        rc = upload(c, &A->d_ell_val, sval.data(), sval.size());
        if (!rc) rc = upload(c, &A->d_csr2ell, plan.pos.data(), plan.pos.size());
        if (!rc) rc = upload(c, &A->d_st_off, plan.off.data(), plan.off.size());
        if (!rc) rc = upload(c, &A->d_st_hdr, plan.hdr.data(), plan.hdr.size());
        if (!rc) rc = upload(c, &A->d_st_tail, plan.tail.data(), plan.tail.size());
        if (!rc) rc = upload(c, &A->d_st_tcol, plan.tcol.data(), plan.tcol.size());
        if (!rc) rc = upload(c, &A->d_col, col, (size_t)nnz);
        SymPlan sym;
        if (!rc && !(flags & GIP_MAT_NO_SYM) && plan_symmetric(plan, sym)) {
            rc = upload(c, &A->d_st_code, sym.code.data(), sym.code.size());
            A->sym_nu = sym.nu;
            A->sym_own_slots = sym.own_slots;
            for (int j = 0; j < kSymMaxD; ++j) A->sym_d[j] = sym.d[j];
        }
        if (!rc) {
            // the host arrays above are uploaded asynchronously: wait before they go
            hipError_t e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) rc = hip_fail(e, "stencil upload");
        }
        if (!rc && fp32) {
            hipError_t e = hipMalloc(&A->d_ell_val32, sizeof(float) * (size_t)plan.stored);
            if (e != hipSuccess) rc = hip_fail(e, "hipMalloc fp32 values");
            else k_to_f32<<<nblocks(plan.stored), kBlock, 0, c->stream>>>(plan.stored, A->d_ell_val, A->d_ell_val32);
        }
    } else if (!rc) {
        A->ell_stored = nnz;
        rc = upload(c, &A->d_col, col, (size_t)nnz);
        if (!rc) rc = upload(c, &A->d_val, val, (size_t)nnz);
        if (!rc && fp32 && fmt == GIP_FMT_CSR) {
            GIP_HIP_TRY(hipStreamSynchronize(c->stream));
            hipError_t e = hipMalloc(&A->d_val32, sizeof(float) * (size_t)std::max(nnz, 1));
            if (e != hipSuccess) rc = hip_fail(e, "hipMalloc fp32 values");
            else if (nnz) k_to_f32<<<nblocks(nnz), kBlock, 0, c->stream>>>(nnz, A->d_val, A->d_val32);
        }
    }
    if (!rc) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hip_fail(e, "gip_mat_create upload");
    }
    if (rc) {
        free_mat_arrays(A);
        delete A;
        return rc;
    }
    *out = A;
    return GIP_OK;
}

int gip_mat_destroy(gip_mat *A) {
    if (!A) return GIP_OK;
    hipSetDevice(A->ctx->device);
    hipStreamSynchronize(A->ctx->stream);
    if (A->transpose) gip_mat_destroy(A->transpose);
    free_mat_arrays(A);
    delete A;
    return GIP_OK;
}

int gip_mat_info(const gip_mat *A, int32_t *nrows, int32_t *nnz, int32_t *format, int64_t *stored) {
    if (!A) return set_error(GIP_EINVAL, "null operator");
    if (nrows) *nrows = A->nrows;
    if (nnz) *nnz = A->nnz;
    if (format) *format = A->format;
    if (stored) *stored = A->ell_stored;
    return GIP_OK;
}

int gip_mat_product_bytes(const gip_mat *A, int32_t fp32, int64_t *bytes) {
    if (!A || !bytes) return set_error(GIP_EINVAL, "gip_mat_product_bytes: bad argument");
    const int64_t vb = fp32 ? 4 : 8, rows = A->nrows, nnz = A->nnz, cols = A->cols_needed;
    const int64_t vecs = vb * cols + vb * rows;  // the input's columns once, the output once
    int64_t b;
    switch (A->format) {
        case GIP_FMT_STENCIL: {
            const int64_t slots = (int64_t)A->nslices * A->st_w;
            b = slots * 64 * vb + slots * 4                        // slot values, slot offsets
                + (int64_t)A->nslices * (int64_t)sizeof(StencilTail)  // per-slice tail records
                + (int64_t)A->st_tail_nnz * (vb + 4) + vecs;        // tail values and columns
            break;
        }
        case GIP_FMT_ELL:
            b = A->ell_stored * (vb + 4) + 4 * ((int64_t)A->nslices + 1) + vecs;
            break;
        case GIP_FMT_DIAG:
            b = vb * rows + vecs;
            break;
        default:
            b = nnz * (vb + 4) + 4 * (rows + 1) + vecs;
    }
    *bytes = b;
    return GIP_OK;
}

int gip_mat_sym_state(const gip_mat *A, int32_t *state, int64_t *bytes) {
    if (!A || !state) return set_error(GIP_EINVAL, "gip_mat_sym_state: bad argument");
    const bool current = A->sym_version == A->val_version + 1;
    *state = A->format != GIP_FMT_STENCIL || A->sym_nu <= 0 ? 0 : current ? A->sym_ok : 2;
    // the symmetric product's own bytes: the values of the slots of offsets
    // >= 0, the slot offsets and codes, the tail records, values and
    // columns, the input once, the output once (the values of the slots of
    // negative offset are re-reads of the same arrays)
    if (bytes)
        *bytes = *state == 0 ? 0
                             : A->sym_own_slots * 64 * 8 +
                                   (int64_t)A->nslices * (4 * A->st_w + 8 + (int64_t)sizeof(StencilTail)) +
                                   (int64_t)A->st_tail_nnz * 12 + 8 * ((int64_t)A->cols_needed + A->nrows);
    return GIP_OK;
}

int gip_mat_autotune(gip_ctx *c, int32_t nrows, int32_t ncols, int32_t nnz, const int32_t *rowptr, const int32_t *col,
                     const double *val, int32_t flags, const double *x, double *y, int32_t reps, gip_mat **out,
                     gip_tune_report *rep) {
    if (!c || !out || !x || !y || reps < 1) return set_error(GIP_EINVAL, "gip_mat_autotune: bad argument");
    *out = nullptr;
    struct Cand { int fmt, lanes, slice_flag; };
    const Cand cands[8] = {{GIP_FMT_ELL, 1, GIP_MAT_ELL_SLICE32}, {GIP_FMT_ELL, 1, 0},
                           {GIP_FMT_ELL, 1, GIP_MAT_ELL_SLICE128}, {GIP_FMT_CSR, 1, 0},
                           {GIP_FMT_CSR, 2, 0}, {GIP_FMT_CSR, 4, 0},
                           {GIP_FMT_CSR, 8, 0}, {GIP_FMT_STENCIL, 1, 0}};
    const int32_t base_flags = flags & ~0xff & ~(GIP_MAT_ELL_SLICE32 | GIP_MAT_ELL_SLICE128);
    gip_tune_report r{};
    r.chosen_format = -1;
    float best = 0.0f;
    gip_mat *keep = nullptr, *csr = nullptr;
    hipEvent_t e0, e1;
    GIP_HIP_TRY(hipEventCreate(&e0));
    GIP_HIP_TRY(hipEventCreate(&e1));
    int rc = GIP_OK;
    for (int k = 0; k < 8 && !rc; ++k) {
        r.us[k] = -1.0f;
        gip_mat *A = nullptr;
        if (cands[k].fmt == GIP_FMT_ELL || cands[k].fmt == GIP_FMT_STENCIL) {
            rc = gip_mat_create(c, nrows, ncols, nnz, rowptr, col, val, cands[k].fmt | base_flags | cands[k].slice_flag,
                                &A);
            if (rc) break;
            // not a candidate: the layout does not fit, or ELL would pad heavily
            if (A->format != cands[k].fmt ||
                (A->format == GIP_FMT_ELL && A->ell_stored > (int64_t)(1.2 * nnz) + (8 << A->ell_lg))) {
                gip_mat_destroy(A);
                continue;
            }
        } else {
            if (!csr) rc = gip_mat_create(c, nrows, ncols, nnz, rowptr, col, val, GIP_FMT_CSR | base_flags, &csr);
            if (rc) break;
            A = csr;
            A->lanes = cands[k].lanes;
        }
        for (int w = 0; w < 2 && !rc; ++w) rc = spmv_launch(c, A, x, y, 1.0, 0.0, 0, nullptr);
        hipEventRecord(e0, c->stream);
        for (int w = 0; w < reps && !rc; ++w) rc = spmv_launch(c, A, x, y, 1.0, 0.0, 0, nullptr);
        hipEventRecord(e1, c->stream);
        float ms = 0.0f;
        if (!rc && hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) {
            r.us[k] = 1e3f * ms / reps;
            if (r.chosen_format < 0 || r.us[k] < best) {
                best = r.us[k];
                r.chosen_format = cands[k].fmt;
                r.chosen_lanes = cands[k].lanes;
                r.chosen_slice = cands[k].fmt == GIP_FMT_ELL ? 1 << A->ell_lg : 0;
                if (keep && keep != csr) gip_mat_destroy(keep);
                keep = A;
            } else if (A != csr) {
                gip_mat_destroy(A);
            }
        } else if (A != csr) {
            gip_mat_destroy(A);
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (csr && keep != csr) gip_mat_destroy(csr);
    if (rc || !keep) {
        if (keep) gip_mat_destroy(keep);
        return rc ? rc : set_error(GIP_EHIP, "gip_mat_autotune: no candidate could be timed");
    }
    if (keep == csr) keep->lanes = r.chosen_lanes;
    if (rep) *rep = r;
    *out = keep;
    return GIP_OK;
}

int gip_mat_set_lanes(gip_mat *A, int32_t lanes) {
    if (!A || !(lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8))
        return set_error(GIP_EINVAL, "gip_mat_set_lanes: lanes must be 1, 2, 4 or 8");
    if (A->format != GIP_FMT_CSR && lanes != 1) return set_error(GIP_EINVAL, "gip_mat_set_lanes: CSR operators only");
    A->lanes = lanes;
    return GIP_OK;
}

int gip_mat_cols_needed(const gip_mat *A, int32_t *ncols) {
    if (!A || !ncols) return set_error(GIP_EINVAL, "null argument");
    *ncols = A->cols_needed;
    return GIP_OK;
}

// Device pointer to the values in CSR order, or an ELL map.
static int values_csr_order(const gip_mat *A, double *dst_dev) {
    gip_ctx *c = A->ctx;
    if (A->nnz == 0) return GIP_OK;
    if (A->format == GIP_FMT_ELL || A->format == GIP_FMT_STENCIL) {
        k_gather_vals<<<nblocks(A->nnz), kBlock, 0, c->stream>>>(A->nnz, A->d_ell_val, A->d_csr2ell, dst_dev);
        GIP_LAUNCH_CHECK("gather values");
    } else if (A->format == GIP_FMT_DIAG) {
        GIP_HIP_TRY(hipMemcpyAsync(dst_dev, A->d_diag, sizeof(double) * A->nnz, hipMemcpyDeviceToDevice, c->stream));
    } else {
        GIP_HIP_TRY(hipMemcpyAsync(dst_dev, A->d_val, sizeof(double) * A->nnz, hipMemcpyDeviceToDevice, c->stream));
    }
    return GIP_OK;
}

int gip_mat_get_values(const gip_mat *A, double *host) {
    if (!A || (A->nnz && !host)) return set_error(GIP_EINVAL, "gip_mat_get_values: bad argument");
    if (!A->nnz) return GIP_OK;
    gip_ctx *c = A->ctx;
    double *tmp = nullptr;
    GIP_HIP_TRY(hipMalloc(&tmp, sizeof(double) * A->nnz));
    int rc = values_csr_order(A, tmp);
    if (!rc) {
        hipError_t e = hipMemcpyAsync(host, tmp, sizeof(double) * A->nnz, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hip_fail(e, "gip_mat_get_values");
    }
    hipFree(tmp);
    return rc;
}

int gip_mat_set_values(gip_mat *A, const double *host) {
    if (!A || (A->nnz && !host)) return set_error(GIP_EINVAL, "gip_mat_set_values: bad argument");
    if (!A->nnz) return GIP_OK;
    gip_ctx *c = A->ctx;
    ++A->val_version;
    if (A->format == GIP_FMT_ELL || A->format == GIP_FMT_STENCIL) {
        double *tmp = nullptr;
        GIP_HIP_TRY(hipMalloc(&tmp, sizeof(double) * A->nnz));
        hipError_t e = hipMemcpyAsync(tmp, host, sizeof(double) * A->nnz, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) {
            k_scatter_vals<<<nblocks(A->nnz), kBlock, 0, c->stream>>>(A->nnz, tmp, A->d_csr2ell, A->d_ell_val);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        hipFree(tmp);
        if (e != hipSuccess) return hip_fail(e, "gip_mat_set_values");
        if (A->d_ell_val32) k_to_f32<<<nblocks(A->ell_stored), kBlock, 0, c->stream>>>(A->ell_stored, A->d_ell_val, A->d_ell_val32);
    } else {
        double *dst = A->format == GIP_FMT_DIAG ? A->d_diag : A->d_val;
        GIP_HIP_TRY(hipMemcpyAsync(dst, host, sizeof(double) * A->nnz, hipMemcpyHostToDevice, c->stream));
        GIP_HIP_TRY(hipStreamSynchronize(c->stream));
        if (A->d_val32) k_to_f32<<<nblocks(A->nnz), kBlock, 0, c->stream>>>(A->nnz, A->d_val, A->d_val32);
    }
    GIP_LAUNCH_CHECK("gip_mat_set_values");
    return diag_refresh(c, A);
}

int gip_mat_diag_inv(const gip_mat *A, double *dinv) {
    if (!A || (A->nrows && !dinv)) return set_error(GIP_EINVAL, "gip_mat_diag_inv: bad argument");
    if (A->diag_missing) return set_error(GIP_EINVAL, "gip_mat_diag_inv: a row has no diagonal entry");
    if (!A->nrows) return GIP_OK;
    k_inv<<<nblocks(A->nrows), kBlock, 0, A->ctx->stream>>>(A->nrows, A->d_diag, dinv);
    GIP_LAUNCH_CHECK("gip_mat_diag_inv");
    return GIP_OK;
}

int gip_dcsrmv(gip_ctx *c, const gip_mat *A, const double *x, double *y) {
    if (!c || !A || (A->nrows && (!x || !y))) return set_error(GIP_EINVAL, "gip_dcsrmv: bad argument");
    return spmv_launch(c, A, x, y, 1.0, 0.0, 0, nullptr);
}

int gip_dcsrmv_ab(gip_ctx *c, const gip_mat *A, const double *x, double *y, double alpha, double beta) {
    if (!c || !A || (A->nrows && (!x || !y))) return set_error(GIP_EINVAL, "gip_dcsrmv_ab: bad argument");
    return spmv_launch(c, A, x, y, alpha, beta, 1, nullptr);
}

int gip_dcsrmv_vals(gip_ctx *c, const gip_mat *A, const double *x, gip_mat *B) {
    if (!c || !A || !B || (A->nrows && !x)) return set_error(GIP_EINVAL, "gip_dcsrmv_vals: bad argument");
    if (A->nrows != B->nnz) return set_error(GIP_EINVAL, "gip_dcsrmv_vals: rows of A must equal nnz of B");
    ++B->val_version;
    int rc;
    if (B->format == GIP_FMT_ELL || B->format == GIP_FMT_STENCIL)
        rc = spmv_launch(c, A, x, B->d_ell_val, 1.0, 0.0, 0, B->d_csr2ell);
    else if (B->format == GIP_FMT_DIAG)
        rc = spmv_launch(c, A, x, B->d_diag, 1.0, 0.0, 0, nullptr);
    else
        rc = spmv_launch(c, A, x, B->d_val, 1.0, 0.0, 0, nullptr);
    if (rc) return rc;
    if (B->d_ell_val32) k_to_f32<<<nblocks(B->ell_stored), kBlock, 0, c->stream>>>(B->ell_stored, B->d_ell_val, B->d_ell_val32);
    if (B->d_val32) k_to_f32<<<nblocks(B->nnz), kBlock, 0, c->stream>>>(B->nnz, B->d_val, B->d_val32);
    GIP_LAUNCH_CHECK("gip_dcsrmv_vals");
    return diag_refresh(c, B);
}

// B's values (CSR order) <- alpha A x + beta B's values, rows of A = nnz of
// B, empty rows of A leave their value (GIP_MatrixMulticore.cpp:125-143):
// the product's alpha/beta epilogue reads and writes B's stored values
// through the CSR -> storage map.
int gip_dcsrmv_vals_ab(gip_ctx *c, const gip_mat *A, const double *x, gip_mat *B, double alpha, double beta) {
    if (!c || !A || !B || (A->nrows && !x)) return set_error(GIP_EINVAL, "gip_dcsrmv_vals_ab: bad argument");
    if (A->nrows != B->nnz) return set_error(GIP_EINVAL, "gip_dcsrmv_vals_ab: rows of A must equal nnz of B");
    ++B->val_version;
    int rc;
    if (B->format == GIP_FMT_ELL || B->format == GIP_FMT_STENCIL)
        rc = spmv_launch(c, A, x, B->d_ell_val, alpha, beta, 1, B->d_csr2ell);
    else if (B->format == GIP_FMT_DIAG)
        rc = spmv_launch(c, A, x, B->d_diag, alpha, beta, 1, nullptr);
    else
        rc = spmv_launch(c, A, x, B->d_val, alpha, beta, 1, nullptr);
    if (rc) return rc;
    if (B->d_ell_val32) k_to_f32<<<nblocks(B->ell_stored), kBlock, 0, c->stream>>>(B->ell_stored, B->d_ell_val, B->d_ell_val32);
    if (B->d_val32) k_to_f32<<<nblocks(B->nnz), kBlock, 0, c->stream>>>(B->nnz, B->d_val, B->d_val32);
    GIP_LAUNCH_CHECK("gip_dcsrmv_vals_ab");
    return diag_refresh(c, B);
}

int gip_scsrmv(gip_ctx *c, const gip_mat *A, const float *x, float *y) {
    if (!c || !A || (A->nrows && (!x || !y))) return set_error(GIP_EINVAL, "gip_scsrmv: bad argument");
    return spmv32_launch(c, A, x, y);
}

__global__ __launch_bounds__(kBlock) void k_fill0(int n, double *__restrict__ y) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[i] = 0.0;
}

static double *value_store(const gip_mat *A) {
    return (A->format == GIP_FMT_ELL || A->format == GIP_FMT_STENCIL) ? A->d_ell_val
           : A->format == GIP_FMT_DIAG                                 ? A->d_diag
                                                                       : A->d_val;
}

__global__ __launch_bounds__(kBlock) void k_refresh(int n, const double *__restrict__ src, const int32_t *__restrict__ sidx,
                                                   double *__restrict__ dst, const int32_t *__restrict__ didx) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) dst[didx[i]] = src[sidx[i]];
}

// Build the transpose of A once: structure by a counting sort on the host
// (a row of A^T lists the rows of A in ascending order), plus the maps that
// let its values be refreshed from A's current values on the device.
static int build_transpose(gip_ctx *c, gip_mat *A) {
    std::vector<int32_t> rp((size_t)A->nrows + 1), cl((size_t)A->nnz), c2e;
    hipError_t e = hipMemcpyAsync(rp.data(), A->d_rowptr, sizeof(int32_t) * rp.size(), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && A->nnz)
        e = hipMemcpyAsync(cl.data(), A->d_col, sizeof(int32_t) * cl.size(), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && (A->format == GIP_FMT_ELL || A->format == GIP_FMT_STENCIL) && A->nnz) {
        c2e.resize((size_t)A->nnz);
        e = hipMemcpyAsync(c2e.data(), A->d_csr2ell, sizeof(int32_t) * c2e.size(), hipMemcpyDeviceToHost, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(e, "transpose structure");
    int32_t nc = 0;
    for (int32_t v : cl) nc = std::max(nc, v + 1);
    std::vector<int32_t> trp((size_t)nc + 1, 0), tcl((size_t)A->nnz), tsrc((size_t)A->nnz);
    for (int32_t v : cl) trp[(size_t)v + 1]++;
    for (int32_t k = 0; k < nc; ++k) trp[(size_t)k + 1] += trp[(size_t)k];
    std::vector<int32_t> fillp(trp.begin(), trp.end() - 1);
    for (int32_t r = 0; r < A->nrows; ++r)
        for (int32_t j = rp[(size_t)r]; j < rp[(size_t)r + 1]; ++j) {
            int32_t p = fillp[(size_t)cl[(size_t)j]]++;
            tcl[(size_t)p] = r;
            tsrc[(size_t)p] = c2e.empty() ? j : c2e[(size_t)j];
        }
    std::vector<double> zeros((size_t)A->nnz, 0.0);
    gip_mat *T = nullptr;
    int rc = gip_mat_create(c, nc, A->nrows, A->nnz, trp.data(), tcl.data(), zeros.data(), GIP_FMT_AUTO, &T);
    if (rc) return rc;
    std::vector<int32_t> tdst((size_t)A->nnz);
    if ((T->format == GIP_FMT_ELL || T->format == GIP_FMT_STENCIL) && A->nnz) {
        e = hipMemcpy(tdst.data(), T->d_csr2ell, sizeof(int32_t) * tdst.size(), hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            gip_mat_destroy(T);
            return hip_fail(e, "transpose map");
        }
    } else {
        for (int32_t p = 0; p < A->nnz; ++p) tdst[(size_t)p] = p;
    }
    rc = upload(c, &A->d_tsrc, tsrc.data(), tsrc.size());
    if (!rc) rc = upload(c, &A->d_tdst, tdst.data(), tdst.size());
    if (rc) {
        gip_mat_destroy(T);
        return rc;
    }
    A->transpose = T;
    return GIP_OK;
}

int gip_dcsrmv_t(gip_ctx *c, const gip_mat *Ac, const double *x, double *y, int32_t ncols_out) {
    gip_mat *A = const_cast<gip_mat *>(Ac);
    if (!c || !A || ncols_out < 0 || (ncols_out && !y) || (A->nrows && !x))
        return set_error(GIP_EINVAL, "gip_dcsrmv_t: bad argument");
    if (!A->transpose) {
        int rc = build_transpose(c, A);
        if (rc) return rc;
    }
    gip_mat *T = A->transpose;
    if (T->nrows > ncols_out) return set_error(GIP_EINVAL, "gip_dcsrmv_t: ncols_out smaller than the largest column + 1");
    if (A->nnz) {
        k_refresh<<<nblocks(A->nnz), kBlock, 0, c->stream>>>(A->nnz, value_store(A), A->d_tsrc, value_store(T), A->d_tdst);
        GIP_LAUNCH_CHECK("gip_dcsrmv_t refresh");
    }
    if (ncols_out) {
        k_fill0<<<nblocks(ncols_out), kBlock, 0, c->stream>>>(ncols_out, y);
        GIP_LAUNCH_CHECK("gip_dcsrmv_t fill");
    }
    return spmv_launch(c, T, x, y, 1.0, 0.0, 0, nullptr);
}

}  // extern "C"
