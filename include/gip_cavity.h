/*
 * gip_cavity.h — C ABI of the operator side of the GIP algebraic
 * fractional-step Navier–Stokes path: the lid-driven-cavity operator
 * generator for hexahedral grids, and readers/writers of the reference's
 * on-disk operator formats.
 *
 * The reference ships only a pre-generated tetrahedral case (C100K, four
 * partitions) and has no generator (README.md:24).  The operator relations
 * this generator follows are the ones the shipped C100K operators satisfy:
 *   L    = -Mres.SO           (pressure Laplacian, GIP_DNS.h:340-354)
 *   M{c} = -Mres.SF{c}        (divergence, GIP_DNS.h:342-344)
 *   C    =  Ec.mf             (convection rebuilt from face fluxes, GIP_DNS.h:311)
 *   D                         (diffusion with Dirichlet ghost coefficients)
 *   G{c} =  Omega^-1 M{c}^T   (gradient, GIP_DNS.h:358-360)
 *   Lb                        (ghost <- adjacent cell, GIP_DNS.h:351)
 *   SO, SF{c}                 (face operators, GIP_DNS.h:381-389)
 * On-disk formats:
 *   .csr   GIP_MatrixMulticore.cpp:12-27  int32[5]{nnz,ncols,nrows,ceil(nrows/1024),1024},
 *          float64[nnz] val, int32[nnz] col, int32[nrows+1] rowptr
 *   .vec   GIP_VectorMulticore.cpp:28-40  int32 n, float64[n]
 *   .topo  GIP_Topo.cpp:12-49             9 ints, send_indx, send_row, recv_indx, recv_row (ASCII)
 *   paths  GIP_DNS.h:149-285              InputFiles/{Operators,Scalarfields,Topos}/<name>/<name>_<np>p/...
 *
 * All entry points return 0 on success and a negative GIP_CAV_E* code on
 * failure; gip_cavity_last_error() returns a message for the last failure
 * of the calling thread.
 */
#ifndef GIP_CAVITY_H
#define GIP_CAVITY_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GIP_CAV_OK = 0,
    GIP_CAV_EINVAL = -1, /* bad argument                         */
    GIP_CAV_EIO = -2,    /* file missing, short or malformed      */
    GIP_CAV_ENOMEM = -3, /* host allocation failed                */
    GIP_CAV_ETOPO = -4   /* inconsistent partition topology       */
};

/* One CSR operator, layout of GIP_Matrix.h:60-66 (host arrays). */
typedef struct gip_csr {
    int32_t nnz;
    int32_t ncols; /* header value only; not a bound (columns may reach past it) */
    int32_t nrows;
    double *val;
    int32_t *col;
    int32_t *rowptr; /* nrows+1 entries */
} gip_csr;

/* Operator selection bits for gip_cavity_generate(). */
enum {
    GIP_OP_D = 1u << 0,
    GIP_OP_EC = 1u << 1,
    GIP_OP_L = 1u << 2,
    GIP_OP_LB = 1u << 3,
    GIP_OP_M = 1u << 4,   /* Mx, My, Mz */
    GIP_OP_G = 1u << 5,   /* Gx, Gy, Gz */
    GIP_OP_SF = 1u << 6,  /* SFx, SFy, SFz */
    GIP_OP_SO = 1u << 7,
    GIP_OP_MRES = 1u << 8,
    GIP_OP_ALL = 0x1ffu
};

/* Hexahedral lid-driven cavity on the unit cube. */
typedef struct gip_cavity_desc {
    int32_t nx, ny, nz;       /* cells per axis (>= 2)                                 */
    double beta;              /* 0: uniform; >0: tanh wall clustering, x_i=(1+tanh(beta(2i/n-1))/tanh beta)/2 */
    double gamma;             /* kinematic viscosity 1/Re baked into D           */
    double lid_u;             /* lid ghost u value where lid_profile is NULL      */
    const double *lid_profile;/* optional nx*ny lid u values, index i+nx*j       */
    int32_t pin;              /* 1: scale L(0,0) by 1.01 (pressure pin)           */
    uint32_t ops;             /* GIP_OP_* bits to build                           */
} gip_cavity_desc;

/* One partition of a cavity case: layout [inner | boundary ghosts], no halo. */
typedef struct gip_cavity {
    int32_t n_inner, n_ghost, n_all, n_faces;
    int32_t nx, ny, nz; /* 0 for cases read from files */
    gip_csr D, Ec, L, Lb, Mx, My, Mz, Gx, Gy, Gz, SFx, SFy, SFz, SO, Mres;
    double *u, *v, *w, *dxs; /* n_all values each                              */
    double *vol;             /* n_inner cell volumes (NULL for file cases)      */
    int32_t *lid_ghost;      /* nx*ny ghost indices of the lid, i+nx*j (hex only) */
} gip_cavity;

const char *gip_cavity_last_error(void);

/* Build the operators selected by desc->ops plus the initial fields. */
int gip_cavity_generate(const gip_cavity_desc *desc, gip_cavity **out);
void gip_cavity_free(gip_cavity *cav);

/* Write/read a one-partition case in the reference's InputFiles layout
 * under <root>/InputFiles/... with file suffix <name>_1p-0 (GIP_DNS.h:149). */
int gip_cavity_write(const gip_cavity *cav, const char *root, const char *name);
int gip_cavity_read(const char *root, const char *name, gip_cavity **out);

/* Read an <nparts>-partition case (<name>_<nparts>p-<rank> files) and merge
 * it into one partition: rows keep their per-rank order, halo columns are
 * redirected through the paired send/recv lists of GIP_Topo.cpp:150-173,
 * face spaces stay per rank (concatenated). */
int gip_cavity_merge(const char *root, const char *name, int32_t nparts, gip_cavity **out);

/* Write a one-partition case as an <nparts>-partition case
 * (<name>_<nparts>p-<rank> files), the decomposition the reference's
 * driver reads with one MPI rank per partition (GIP_DNS.h:149,
 * DrivenCavity.cpp:30-33): inner cells in contiguous index blocks (z slabs
 * of a hex grid), each boundary ghost with its adjacent cell, a rank's
 * faces those of its cells (faces between two ranks on both sides), local
 * order [interiors | interfaces | ghosts | halo grouped by source rank],
 * halo lists paired as GIP_Topo.cpp:124-181 exchanges them.  Operator rows
 * keep their stored column order, so every row sums its terms as in the
 * undivided case.  gip_cavity_merge reads the result back (with the cells
 * renumbered rank by rank). */
int gip_cavity_split(const gip_cavity *cav, int32_t nparts, const char *root, const char *name);

/* Seeded zero-mean N(0,1) field over n values (splitmix64 + Box-Muller,
 * then the mean is removed): the Poisson right-hand side of the PCG
 * workloads.  Deterministic for a given (n, seed). */
int gip_cavity_rhs_normal(int32_t n, uint64_t seed, double *out);

/* Single-file helpers (formats above). */
int gip_csr_read(const char *path, gip_csr *A);
int gip_csr_write(const char *path, const gip_csr *A);
void gip_csr_release(gip_csr *A);
int gip_vec_read(const char *path, double **vals, int32_t *n);
int gip_vec_write(const char *path, const double *vals, int32_t n);

#ifdef __cplusplus
}
#endif
#endif
