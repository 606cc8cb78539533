// Internal declarations shared by the HIP translation units of libgip_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "gip_hip.h"
#include "gip_layout.h"

namespace gip {

int set_error(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define GIP_HIP_TRY(expr)                                   \
    do {                                                    \
        hipError_t e__ = (expr);                            \
        if (e__ != hipSuccess) return gip::hip_fail(e__, #expr); \
    } while (0)

#define GIP_LAUNCH_CHECK(name)                                         \
    do {                                                               \
        hipError_t e__ = hipGetLastError();                            \
        if (e__ != hipSuccess) return gip::hip_fail(e__, name);        \
    } while (0)

constexpr int kBlock = 256;        // threads per workgroup for streaming kernels
constexpr int kMaxRedBlocks = 2048; // partial sums per reduction

}  // namespace gip

struct gip_ctx {
    int device;
    hipStream_t stream;
    double *d_partials;     // kMaxRedBlocks * 4 doubles of reduction partials (one per workgroup)
    double *h_result;       // pinned host landing slot for reduction results
    double *h_result_dev;   // its device address: the finalize kernels store the result there
    int num_cu;
};

struct gip_mat {
    gip_ctx *ctx;
    int32_t nrows, ncols, nnz;
    int32_t cols_needed;  // largest column index + 1
    int32_t format;
    int32_t lanes;  // CSR: lanes per row (1, 2, 4, 8)
    // CSR (kept for every format: value get/set, transposed products, diagonal)
    int32_t *d_rowptr;
    int32_t *d_col;
    double *d_val;
    float *d_val32;
    // sliced ELL with H = 2^ell_lg rows per slice (32, 64 or 128): slice s
    // covers rows [H s, H s + H); entry k of row H s + l sits at
    // slice_off[s] + H k + l; padding has col = -1.
    int32_t ell_lg;
    int32_t nslices;
    int32_t uniform_width; // > 0 when every slice has this width and no padding
    int64_t ell_stored;
    int32_t *d_slice_off; // nslices + 1
    int32_t *d_ell_col;
    double *d_ell_val;
    float *d_ell_val32;
    int32_t *d_csr2ell;   // CSR nnz index -> position in d_ell_val (ELL and stencil)
    // stencil layout (gip_layout.h): slice s, slot k holds the entry of
    // column row + off[s W + k] of each of its 64 rows, value at
    // d_ell_val[(s W + k) 64 + lane] (0 where a row lacks that column); a
    // row's remaining entries, in stored order, form its tail, located
    // through the slice header (lane masks per tail rank); tail values at
    // d_ell_val[tail_base + t], columns d_st_tcol[t].
    int32_t st_w;
    int64_t st_tail_base;
    int32_t st_tail_nnz;
    int32_t *d_st_off;        // nslices * st_w
    gip::StencilHdr *d_st_hdr;    // nslices (read for kind-2 slices only)
    gip::StencilTail *d_st_tail;  // nslices
    int32_t *d_st_tcol;    // st_tail_nnz
    // symmetric form of the slot values (gip_layout.h, plan_symmetric), used
    // by the FP64 PCG's product: sym_nu arrays of nslices * 64 values (0: the
    // layout has none), sym_d the paired offsets, made and checked on the
    // device from the slot values at the first solve after a change of the
    // values (val_version)
    int32_t sym_nu;
    int32_t sym_d[gip::kSymMaxD];
    int64_t sym_own_slots;  // slots of offsets >= 0 over all slices
    int8_t *d_st_code;  // nslices * 8 slot codes
    uint64_t val_version;
    double *d_sym_val;
    int32_t *d_sym_flag;
    uint64_t sym_version;  // val_version + 1 the symmetric values were made from (0: never)
    int32_t sym_ok;        // 1: in use, -1: the values are not symmetric
    // a_ii per row (the stored values of a DIAG operator), for Jacobi; for
    // the other layouts refreshed from d_diag_src (each row's last diagonal
    // entry's position in the value storage, -1 for none) when values change
    double *d_diag;
    int32_t *d_diag_src;
    int32_t diag_missing;
    // lazily built transpose: structure once, values refreshed per use
    gip_mat *transpose;
    int32_t *d_tsrc;  // per transposed entry: position in this operator's value storage
    int32_t *d_tdst;  // per transposed entry: position in the transpose's value storage
};

namespace gip {
int reduce_grid(int n);
// Makes (once per version of the values) and checks the symmetric form of
// a stencil operator's slot values; *ok is true when the PCG's product may
// read them.  Allocates on first use.
int mat_sym_prepare(gip_ctx *c, gip_mat *A, bool *ok);
// Looks up the no-alloc PCG scratch array's address on `device` (current);
// called once per device by gip_init.
int pcg_scratch_setup(int device);
}
