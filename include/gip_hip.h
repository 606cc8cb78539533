/*
 * gip_hip.h — C ABI of the MI355X (gfx950) kernel library behind the HIP
 * specification of GIP's algebraic fractional-step Navier–Stokes path.
 *
 * It replaces the reference's CUDA kernel library interface
 * (src/Specifications/Cuda/Headers/GIP_gpu.h:1-47, kernels in
 * src/Specifications/Cuda/Libs/libgpu.cu), with one C name per overload,
 * status returns instead of silent failures, and the CPU semantics of the
 * Multicore specification where the CUDA one differs (empty CSR rows are
 * left untouched; the infinity norm covers every element; the CFL seed is
 * 1e20).  Floating-point work is FP64 unless a name says otherwise and is
 * compiled without multiply-add contraction: every SpMV row sums its
 * products in stored order with separately rounded multiplies and adds,
 * exactly as GIP_MatrixMulticore.cpp:64-103 does.  Reductions are
 * deterministic (fixed grid, fixed combination order).
 *
 * All calls are ordered on the context's stream.  Calls that return a host
 * scalar (gip_ddot, gip_dnrm_inf, gip_cfl) wait for it; the rest are
 * asynchronous.  Device pointers are plain pointers from gip_malloc.
 * Every entry returns GIP_OK (0) or a negative gip_status; the message of
 * the last failure on the calling thread is gip_last_error().
 */
#ifndef GIP_HIP_H
#define GIP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gip_status {
    GIP_OK = 0,
    GIP_EINVAL = -1,     /* bad argument or shape                          */
    GIP_EHIP = -2,       /* HIP runtime error                              */
    GIP_ENOMEM = -3,     /* device or host allocation failed               */
    GIP_EBREAKDOWN = -4, /* PCG breakdown: p.Ap == 0 or a non-finite step  */
    GIP_ENOCONV = -5     /* PCG stopped at maxit above tolerance           */
} gip_status;

typedef struct gip_ctx gip_ctx; /* one device + one stream + reduction scratch */
typedef struct gip_mat gip_mat; /* a device-resident sparse operator           */
typedef struct gip_pcg gip_pcg; /* a device-resident fused PCG solver          */

const char *gip_last_error(void);

/* ---- device and context ------------------------------------------------
 * gip_init replaces GIP_cudaSetDevice (Headers/GIP_gpu.h:43) as called by
 * GIP_ArchGPU::setUp (GIP_ArchGPU.cpp:19-26); the context owns a
 * default-priority stream and the scratch of the reductions. */
int gip_init(int device, gip_ctx **ctx);
int gip_finalize(gip_ctx *ctx);
int gip_sync(gip_ctx *ctx);
/* Make `device` the calling host thread's current device (threads that
 * drive contexts they did not create, e.g. one thread per partition). */
int gip_use_device(int device);
void *gip_stream(gip_ctx *ctx); /* the hipStream_t the context launches on */
int gip_device_count(int *count);

/* ---- memory -------------------------------------------------------------
 * GIP_cudaMallocDouble/Int (Headers/GIP_gpu.h:4,6), GIP_cudaFree (:8),
 * GIP_cudaMemcpyDToGpu/IToGpu/IToCpu/DToCpu (:10-16), GIP_cudaDcopy (:20).
 * Copies are stream-ordered; d2h waits for the copy. */
int gip_malloc(gip_ctx *ctx, size_t bytes, void **dptr);
int gip_free(gip_ctx *ctx, void *dptr);
int gip_memcpy_h2d(gip_ctx *ctx, void *dst, const void *src, size_t bytes);
int gip_memcpy_d2h(gip_ctx *ctx, void *dst, const void *src, size_t bytes);
int gip_memcpy_d2d(gip_ctx *ctx, void *dst, const void *src, size_t bytes);

/* ---- events (GIP_cudaEvent* of Headers/GIP_gpu.h:32-41, GIP_TimerCuda.cpp) */
typedef struct gip_event gip_event;
int gip_event_create(gip_ctx *ctx, gip_event **ev);
int gip_event_record(gip_ctx *ctx, gip_event *ev);
int gip_event_elapsed_ms(gip_event *start, gip_event *stop, float *ms); /* waits for stop */
int gip_event_destroy(gip_event *ev);

/* ---- sparse operators ----------------------------------------------------
 * Device copy of a CSR operator (GIP_MatrixCuda.cpp:14-53).  The host
 * arrays are read once and may be freed afterwards.  The layout is chosen
 * per operator: a diagonal vector when the operator is diagonal; the
 * stencil layout when the rows of each 64-row slice share their column
 * offsets (structured meshes: per slice and entry slot one offset col - row,
 * values column-major, the few entries that do not fit kept in a small
 * CSR tail summed after the slots; no column index per entry); sliced ELL
 * (64-row slices, column-major inside a slice) when padding stays small;
 * CSR otherwise.  `format` forces one of GIP_FMT_*; a forced stencil layout
 * that the operator does not fit falls back to the automatic choice.
 * Every layout sums each row in stored order (bitwise CSR results for
 * finite inputs; the stencil layout adds +-0 for slots a row lacks).
 * `ncols` is informational (columns may exceed it, as in the reference). */
enum { GIP_FMT_AUTO = 0, GIP_FMT_CSR = 1, GIP_FMT_ELL = 2, GIP_FMT_DIAG = 3, GIP_FMT_STENCIL = 4 };
enum { GIP_MAT_FP32 = 1 << 8 }; /* also keep an FP32 copy of the values (mixed precision) */
/* Rows per sliced-ELL slice (default 64, one wavefront): 32 pads less on
 * irregular rows (two slices per wavefront), 128 amortises the slice
 * offsets; the autotuner tries all three. */
enum { GIP_MAT_ELL_SLICE32 = 1 << 9, GIP_MAT_ELL_SLICE128 = 1 << 10 };
/* A stencil-layout operator whose values are bitwise symmetric
 * (a_{r,r+d} == a_{r+d,r}) is multiplied inside the FP64 PCG from the
 * values of its slots of offsets d >= 0 only, each slot of offset -d
 * reading row r - d's slot of offset +d: the same values in the same
 * order, so the same bits, with about 4/7 of the value traffic for a
 * 7-point stencil (the FP32 product of the mixed solve keeps all slots).
 * The symmetric arrays are made and checked against every slot on the
 * device at the first large solve (more than 2^19 rows) after the values
 * change (gip_mat_set_values, gip_dcsrmv_vals*); an operator that fails
 * the check keeps the full product.  That first solve allocates the arrays
 * (one value per row for each stored offset: 6 for the cavity Laplacian's
 * 7 slots) and must not overlap another solve on the same operator.
 * GIP_MAT_NO_SYM turns this off. */
enum { GIP_MAT_NO_SYM = 1 << 11 };
int gip_mat_create(gip_ctx *ctx, int32_t nrows, int32_t ncols, int32_t nnz, const int32_t *rowptr, const int32_t *col,
                   const double *val, int32_t flags, gip_mat **A);
int gip_mat_destroy(gip_mat *A);
/* nrows, nnz, chosen format, stored (padded) entries */
int gip_mat_info(const gip_mat *A, int32_t *nrows, int32_t *nnz, int32_t *format, int64_t *stored);
/* Bytes one product y = A x must move in the operator's chosen layout
 * (fp32 != 0: the FP32 values of a GIP_MAT_FP32 operator): the stored
 * values with their per-entry or per-slot indices and per-slice metadata,
 * the input vector once over its columns, the output once over the rows.
 * For CSR that is the CSR count 12 nnz + 4 (rows+1) + 8 rows + 8 cols; the
 * stencil layout stores no column per entry (one offset per slot of 64
 * rows), so it moves fewer bytes than the CSR count of the same operator.
 * No counterpart in the reference: the unit of a layout's roofline. */
int gip_mat_product_bytes(const gip_mat *A, int32_t fp32, int64_t *bytes);
/* State of the symmetric form (GIP_MAT_NO_SYM above): 0 none (not a
 * stencil-layout operator, turned off, or offsets it cannot pair), 2 not
 * yet made for the current values, 1 in use, -1 the values are not
 * symmetric.  bytes (optional): what one FP64 product moves with it (the
 * values of the slots of offsets >= 0, the slot metadata, the tails, the
 * input and output once; 0 for state 0).  No counterpart in the reference. */
int gip_mat_sym_state(const gip_mat *A, int32_t *state, int64_t *bytes);
/* Largest column index + 1: the length an input vector of a product must have. */
int gip_mat_cols_needed(const gip_mat *A, int32_t *ncols);
/* CSR operators: lanes per row (1, 2, 4, 8 = 64, 32, 16, 8 rows per
 * wavefront); the lanes load a row cooperatively and the sum keeps the
 * stored order. */
int gip_mat_set_lanes(gip_mat *A, int32_t lanes);
/* In-project SpMV autotuner: builds the operator in each candidate layout
 * (sliced ELL with 32, 64 or 128 rows per slice; CSR with 1, 2, 4 or 8
 * lanes per row; stencil), times `reps` products y = A x on the device
 * (x, y: device vectors covering the operator) and returns the fastest; the
 * rest are freed.  `report` (optional) gets each candidate's mean time in
 * microseconds (-1: not a candidate) in that order, and for an ELL choice
 * its rows per slice. */
typedef struct gip_tune_report {
    int32_t chosen_format, chosen_lanes, chosen_slice;
    float us[8];
} gip_tune_report;
int gip_mat_autotune(gip_ctx *ctx, int32_t nrows, int32_t ncols, int32_t nnz, const int32_t *rowptr, const int32_t *col,
                     const double *val, int32_t flags, const double *x, double *y, int32_t reps, gip_mat **A,
                     gip_tune_report *report);
/* Values in CSR order, device <-> host. */
int gip_mat_get_values(const gip_mat *A, double *host_vals);
int gip_mat_set_values(gip_mat *A, const double *host_vals);
/* 1/a_ii per row into a device vector of nrows values, as
 * GIP_MatrixMulticore::setUpDiagonal (GIP_MatrixMulticore.cpp:177-210),
 * from the current values (after gip_mat_set_values or gip_dcsrmv_vals*
 * too; a solver takes them at gip_pcg_create, as the reference's solver
 * at its setUp); GIP_EINVAL if a row has no diagonal entry. */
int gip_mat_diag_inv(const gip_mat *A, double *dinv);

/* ---- SpMV ----------------------------------------------------------------
 * gip_dcsrmv:      y_i = (Ax)_i for every non-empty row           GIP_cudaDCsrSpMV 8-arg (Headers/GIP_gpu.h:22),
 *                                                                  semantics GIP_MatrixMulticore.cpp:64-83
 * gip_dcsrmv_ab:   y_i = alpha*(Ax)_i + beta*y_i, non-empty rows   GIP_cudaDCsrSpMV 10-arg (:24), :85-103
 * gip_dcsrmv_vals: B's values (CSR order) <- Ax, non-empty rows   GIP_MatrixMulticore.cpp:105-123
 *                  (nrows(A) must equal nnz(B))
 * gip_dcsrmv_vals_ab: B's values <- alpha*(Ax)_i + beta*B's values, non-empty rows
 *                  (nrows(A) must equal nnz(B))                    GIP_MatrixMulticore.cpp:125-143
 * gip_dcsrmv_t:    y = A^T x over all columns < ncols_out, y zeroed first (adjoint operators)
 * gip_scsrmv:      FP32 y = A32 x32 (needs GIP_MAT_FP32)                                               */
int gip_dcsrmv(gip_ctx *ctx, const gip_mat *A, const double *x, double *y);
int gip_dcsrmv_ab(gip_ctx *ctx, const gip_mat *A, const double *x, double *y, double alpha, double beta);
int gip_dcsrmv_vals(gip_ctx *ctx, const gip_mat *A, const double *x, gip_mat *B);
int gip_dcsrmv_vals_ab(gip_ctx *ctx, const gip_mat *A, const double *x, gip_mat *B, double alpha, double beta);
int gip_dcsrmv_t(gip_ctx *ctx, const gip_mat *A, const double *x, double *y, int32_t ncols_out);
int gip_scsrmv(gip_ctx *ctx, const gip_mat *A, const float *x, float *y);

/* ---- vector operations ---------------------------------------------------
 * GIP_cudaDaxpy 6/7-arg (Headers/GIP_gpu.h:26,28), GIP_cudaDdot (:30),
 * GIP_gpuNormi (:46), GIP_gpuGetCFL (:48); semantics
 * GIP_VectorMulticore.cpp:44-195, GIP_NonLinearMulticore.cpp:3-55. */
int gip_dfill(gip_ctx *ctx, int32_t n, double *x, double value);                      /* x = value     */
int gip_daxpy(gip_ctx *ctx, int32_t n, double alpha, const double *x, double *y);      /* y += alpha x  */
int gip_daxpby(gip_ctx *ctx, int32_t n, double alpha, const double *x, double beta, double *y); /* y = ax+by */
int gip_dcopy(gip_ctx *ctx, int32_t n, const double *x, double *y);                  /* y = x, copyTo (GIP_VectorMulticore.cpp:167-181) */
/* y[dst ? dst[i] : i] = x[src ? src[i] : i] for i < n: the halo pack
 * (send list -> send buffer) and unpack (receive buffer -> halo slots) of
 * GIP_Topo.cpp:101-122, done on the device (the CUDA specification stages
 * them on the host, GIP_VectorCuda.cpp:224-247).  Indices are trusted. */
int gip_dcopy_indexed(gip_ctx *ctx, int32_t n, const int32_t *src, const double *x, const int32_t *dst, double *y);
int gip_ddot(gip_ctx *ctx, int32_t n, const double *x, const double *y, double *result);         /* host result */
int gip_dnrm_inf(gip_ctx *ctx, int32_t n, const double *x, double *result);                    /* max |x_i|  */
int gip_cfl(gip_ctx *ctx, int32_t n, const double *u, const double *v, const double *w, const double *dxs,
            double gamma, double rho, double *dt);
/* Same reductions with the result left in device memory (no host wait). */
int gip_ddot_dev(gip_ctx *ctx, int32_t n, const double *x, const double *y, double *dresult);
int gip_dnrm_inf_dev(gip_ctx *ctx, int32_t n, const double *x, double *dresult);

/* ---- streaming calibration -----------------------------------------------
 * No counterpart in the reference: the practical HBM ceiling the
 * bandwidth-bound kernels are measured against.  Times `reps` launches of
 * one access mix over n doubles (n even; a, b, c device arrays of n
 * doubles, b and c unused where the mix needs fewer): GIP_STREAM_COPY
 * (b = a), GIP_STREAM_READ (sum of a; one partial per workgroup into part,
 * part_len slots), GIP_STREAM_WRITE (a = 1), GIP_STREAM_MIX (three reads,
 * two writes: c += 0.5 a, a = b + 0.25 a).  vec: 1 (8-byte) or 2 (16-byte)
 * accesses; nontemporal: 1 for nontemporal loads and stores; grid: 0 for
 * one access per thread over the array, else that many workgroups striding
 * over it.  *ms receives the mean time of one launch. */
enum { GIP_STREAM_COPY = 0, GIP_STREAM_READ = 1, GIP_STREAM_WRITE = 2, GIP_STREAM_MIX = 3 };
int gip_stream_probe(gip_ctx *ctx, int32_t kind, int32_t vec, int32_t nontemporal, int32_t grid, size_t n, double *a,
                     double *b, double *c, double *part, size_t part_len, int32_t reps, float *ms);
/* Mean cost of a device-wide barrier among `grid` resident workgroups
 * (cooperative launch, `iters` barriers): mode 0 a relaxed atomic counter,
 * 1 with a device-scope fence before arriving and after leaving (what a
 * barrier that publishes data needs), 2 also one value stored per
 * workgroup before and every value summed after.  The measurement behind
 * keeping the PCG's global reductions at kernel boundaries. */
int gip_barrier_probe(gip_ctx *ctx, int32_t grid, int32_t iters, int32_t mode, float *us_per_barrier);

/* ---- fused Jacobi-PCG ----------------------------------------------------
 * GIP_LinearSolver::solve (GIP_LinearSolver.h:45-117) with every scalar
 * kept on the device: per iteration a direction kernel, a product kernel
 * (with its p.Ap partials) and an update kernel (r, s = Dinv r and the
 * partials of r.s and r.r); on operators of at most 2^19 rows the direction
 * is formed inside the product (two kernels per iteration).  The host only
 * reads the solver's state word every `poll` iterations.
 * precision GIP_PCG_FP64: the reference recurrences in FP64.
 * precision GIP_PCG_MIXED: FP32 operator values, vectors and SpMV inside an
 *   FP64 defect-correction loop: the FP64 true residual r = b - A x is
 *   recomputed whenever the FP32 residual has dropped by `inner_tol`, the
 *   correction is accumulated into the FP64 x, and convergence is decided
 *   on the FP64 true residual only.  Dots and norms accumulate in FP64.
 * tol > 0 stops once ||r||_2 <= tol ||b||_2 (FP64 recursive residual for
 * FP64, FP64 true residual for MIXED); tol <= 0 runs maxit iterations.  */
enum { GIP_PCG_FP64 = 0, GIP_PCG_MIXED = 1 };
typedef struct gip_pcg_opts {
    int32_t maxit;     /* iteration limit (inner iterations for MIXED)          */
    double tol;        /* relative residual target, <= 0: fixed maxit           */
    int32_t precond;   /* 1: Jacobi, 0: none                                    */
    int32_t precision; /* GIP_PCG_FP64 or GIP_PCG_MIXED                         */
    double inner_tol;  /* MIXED: residual drop that triggers an FP64 correction */
    int32_t poll;      /* iterations queued between host polls (0: 16)          */
    int32_t coarse;    /* MIXED: at every FP64 correction, the Galerkin step
                          x += c 1, c = (1.r)/(1.A1), along the constant vector
                          of the solved rows (the near-null mode of a pinned
                          Neumann Laplacian, which FP32 cannot resolve), and
                          the FP32 inner iteration deflated by that vector
                          (A p replaced by A p - A1 (1.Ap)/(1.A1)):
                          0 on (default), -1 off                                  */
    int32_t time_products; /* 1: time one product launch per host poll with events on
                              the solver's stream (read at the poll's synchronisation);
                              the median is returned in gip_pcg_stats.product_ms      */
} gip_pcg_opts;
typedef struct gip_pcg_stats {
    int32_t iters;       /* iterations done                                     */
    int32_t corrections; /* FP64 true-residual corrections (MIXED)              */
    double relres;       /* ||r||_2/||b||_2 at exit (recursive for FP64)        */
    int32_t status;      /* GIP_OK, GIP_EBREAKDOWN or GIP_ENOCONV               */
    double product_ms;   /* time_products: median duration of the sampled product
                            launches (the SpMV kernel of an iteration; with the
                            direction fused in on operators of <= 2^19 rows)     */
    int32_t product_samples; /* launches sampled                                  */
} gip_pcg_stats;
/* The operator must belong to ctx (GIP_EINVAL otherwise): its value updates
 * and the solves are ordered on the one stream. */
int gip_pcg_create(gip_ctx *ctx, const gip_mat *A, int32_t n_all, gip_pcg **solver);
/* Solves rows [0,nrows(A)) of A x = b; x0 is copied into x first; vectors are
 * n_all long (entries past nrows(A) are carried over from x0). */
int gip_pcg_solve(gip_pcg *solver, const double *b, double *x, const double *x0, const gip_pcg_opts *opts,
                  gip_pcg_stats *stats);
int gip_pcg_destroy(gip_pcg *solver);

/* The FP64 solve of gip_pcg_solve without any allocation: no counterpart in
 * the reference; for callers that cannot allocate device or pinned memory
 * (code under stream capture, tight serving loops).  The solver workspace
 * (r, s, Ap, Dinv, two directions, partial sums, scalars, state) is a
 * device array of about 13.8 MB, one per GPU, set up by gip_init and held
 * until the process ends; its limit is 262,144 rows (nrows(A)) and 327,680
 * vector entries (n_all): larger operators, GIP_PCG_MIXED and
 * time_products are rejected with GIP_EINVAL.  Same kernels, iterates and
 * stopping rule as gip_pcg_solve on a solver created for A.  Every call on
 * a GPU shares that GPU's workspace: calls that overlap on one GPU (other
 * host threads, other streams) race. */
int gip_pcg_solve_noalloc(gip_ctx *ctx, const gip_mat *A, int32_t n_all, const double *b, double *x,
                          const double *x0, const gip_pcg_opts *opts, gip_pcg_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
