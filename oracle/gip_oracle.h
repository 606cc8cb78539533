/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's CPU path for the algebraic
 * fractional-step Navier–Stokes integrator, collapsed to one partition:
 *   - the Multicore specification (GIP_MatrixMulticore.cpp, GIP_VectorMulticore.cpp,
 *     GIP_NonLinearMulticore.cpp),
 *   - the templated Jacobi-PCG (GIP_LinearSolver.h:45-117),
 *   - the templated integrator (GIP_DNS.h:98-414).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * use it, and only as the checker / CPU baseline; the product path never
 * links, loads or calls it.
 *
 * Every function states the reference lines it restates.  Arithmetic is
 * done in the same order as the reference (serial loops, stored CSR order)
 * so that, built without FMA contraction, its results are bitwise equal to
 * the reference binary run on one partition.
 *
 * One addition beyond the reference: a tolerance stop for the PCG
 * (tol > 0: stop once ||r||_2 <= tol*||b||_2 over the inner range, checked at
 * the end of each iteration; the reference reads tol but never uses it,
 * GIP_LinearSolver.h:34).  tol <= 0 gives the reference's fixed-iteration
 * behaviour.
 */
#ifndef GIP_ORACLE_H
#define GIP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_csr {
    int nnz, ncols, nrows;
    double *val;
    int *col;
    int *rowptr;
} orc_csr;

/* GIP_MatrixMulticore.cpp:5-31 (binary .csr read). 0 on success. */
int orc_csr_read(const char *path, orc_csr *A);
void orc_csr_free(orc_csr *A);

/* GIP_MatrixMulticore.cpp:64-83: y_i = sum_j a_ij x_col, stored order; empty rows untouched. */
void orc_spmv(const orc_csr *A, const double *x, double *y);
/* GIP_MatrixMulticore.cpp:85-103: y_i = alpha*(Ax)_i + beta*y_i for non-empty rows. */
void orc_spmv_ab(const orc_csr *A, const double *x, double *y, double alpha, double beta);
/* GIP_MatrixMulticore.cpp:105-123: result written into another matrix's value array. */
void orc_spmv_vals(const orc_csr *A, const double *x, double *bval);
/* GIP_MatrixMulticore.cpp:125-143: B's values = alpha*(Ax)_i + beta*B's values, non-empty rows. */
void orc_spmv_vals_ab(const orc_csr *A, const double *x, double *bval, double alpha, double beta);
/* GIP_MatrixMulticore.cpp:177-210: Dinv = diag(1/a_ii) as an nrows-nnz CSR.
 * Returns -1 when a row has no diagonal entry (the reference leaves it uninitialised). */
int orc_diag_setup(const orc_csr *A, orc_csr *Dinv);

/* Threads for the row loops (products, axpy, axpby) above 65,536 rows.
 * Each row is computed by one thread with the serial expression and the
 * reductions stay serial, so results are bitwise the same at any count. */
void orc_set_threads(int n);
/* Blocked dot products (default off): sums of 4096-term blocks, each in
 * index order, added in block order; independent of the thread count,
 * equal to the serial sum up to rounding.  Off, orc_dot is the reference's
 * serial sum (GIP_VectorMulticore.cpp:93-128). */
void orc_set_blocked_dot(int on);

/* GIP_VectorMulticore.cpp:44-67 */
void orc_axpy(double *y, const double *x, double alpha, int n);
/* GIP_VectorMulticore.cpp:69-92: y = alpha*x + beta*y */
void orc_axpby(double *y, const double *x, double alpha, double beta, int n);
/* GIP_VectorMulticore.cpp:93-128 (serial sum, one partition) */
double orc_dot(const double *y, const double *x, int n);
/* GIP_VectorMulticore.cpp:130-164 */
double orc_norm_inf(const double *x, int n);
/* GIP_NonLinearMulticore.cpp:3-55 */
double orc_cfl(int inner, const double *u, const double *v, const double *w, const double *dxs, double gamma, double rho);

/* GIP_LinearSolver.h:45-117.  n_inner rows are solved, vectors are n_all long.
 * x receives the solution (x0 is copied in first).  Returns 0, or -2 on a
 * breakdown (p.Ap == 0 or a non-finite step; iteration stops there).
 * iters/relres (optional) receive the iteration count and ||r||_2/||b||_2. */
int orc_pcg(const orc_csr *A, const orc_csr *Dinv, int n_inner, int n_all, const double *b, double *x, const double *x0,
            int maxit, double tol, int precond, int *iters, double *relres);

/* The integrator state, GIP_DNS.h:13-96. */
typedef struct orc_dns {
    int inner, all, faces;
    orc_csr C, Ec, D, L, Lb, Mx, My, Mz, Gx, Gy, Gz, SFx, SFy, SFz, SO, Mres, Dinv;
    double *u, *v, *w, *ax, *ay, *az, *ax0, *ay0, *az0, *rhs, *p, *p0, *gpx, *gpy, *gpz, *mf, *dxs, *temp;
    double gamma, rho, sigma, dt, time;
    int it;
    int pcg_maxit, precond;
    double pcg_tol;
    int last_iters;
    double last_relres;
    int started;
    int borrowed; /* operators owned by the caller (orc_dns_open_mem) */
} orc_dns;

/* Reads <root>/InputFiles/.../<name>_1p-0.* (GIP_DNS.h:115-306, one partition). */
int orc_dns_open(const char *root, const char *name, orc_dns **out);
/* The same state from operators and fields in memory (a generated case):
 * ops holds D, Ec, L, Lb, Mx, My, Mz, Gx, Gy, Gz, SFx, SFy, SFz, SO, Mres,
 * borrowed (they must outlive the state); u, v, w, dxs (all values) are copied. */
int orc_dns_open_mem(int inner, int all, int faces, const orc_csr *ops, const double *u, const double *v,
                     const double *w, const double *dxs, orc_dns **out);
void orc_dns_close(orc_dns *s);
/* Parameters the reference hardcodes (GIP_DNS.h:132-137, :180-185). */
void orc_dns_params(orc_dns *s, double gamma, double rho, int pcg_maxit, double pcg_tol, int precond);
/* One time step (GIP_DNS.h:98-113 loop body + :393-414).  The first call also
 * computes the initial CFL dt.  Outputs: time after the step, dt used by the
 * step, massr, PCG iterations. Returns the PCG status. */
int orc_dns_step(orc_dns *s, double *time, double *dt_used, double *massr, int *iters);
/* State vector by name ("u","v","w","p","p0","mf","ax0","ay0","az0","rhs"); *n receives its length. */
double *orc_dns_field(orc_dns *s, const char *name, int *n);

#ifdef __cplusplus
}
#endif
#endif
