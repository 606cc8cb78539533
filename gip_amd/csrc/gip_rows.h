// Device helpers shared by the PCG and batch kernels: a read-only view of an
// operator, the product operand (plain, or the PCG direction formed on the
// fly), the stored-order row sum, and the XCD-contiguous row sweep.
#pragma once

#include <hip/hip_runtime.h>

#include "gip_internal.h"

namespace gip {

struct EllView {
    int fmt;  // GIP_FMT_ELL / GIP_FMT_CSR
    const int32_t *soff, *col, *rp;
    const double *val;
    const float *val32;
};

static EllView view(const gip_mat *A) {
    EllView v;
    v.fmt = A->format;
    v.soff = A->d_slice_off;
    v.col = A->format == GIP_FMT_ELL ? A->d_ell_col : A->d_col;
    v.rp = A->d_rowptr;
    v.val = A->format == GIP_FMT_ELL ? A->d_ell_val : A->d_val;
    v.val32 = A->format == GIP_FMT_ELL ? A->d_ell_val32 : A->d_val32;
    return v;
}

// The operand of a product: either a plain vector, or the direction formed
// on the fly as 1.0 s_j + beta p_j on the solved rows (p_j beyond them).
template <class T>
struct Operand {
    const T *p;
    const T *s;
    T beta;
    int n;
    bool fused;
    __device__ inline T at(int j) const {
        if (!fused || j >= n) return p[j];
        return T(1) * s[j] + beta * p[j];
    }
};

// Row sum in stored order; returns false for an empty row.
template <int W, class T, class V>
__device__ inline bool row_dot(const EllView &A, const V *__restrict__ val, const Operand<T> &x, int row, T &sum) {
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
        for (int k = 0; k < W; ++k) sum = sum + x.at(c[k]) * v[k];
        return true;
    } else if (A.fmt == GIP_FMT_ELL) {
        const int s = row >> 6, base = A.soff[s] + (row & 63), width = (A.soff[s + 1] - A.soff[s]) >> 6;
        bool any = false;
        for (int k = 0; k < width; ++k) {
            const int cc = A.col[base + 64 * k];
            if (cc < 0) break;
            sum = sum + x.at(cc) * val[base + 64 * k];
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

}  // namespace gip
