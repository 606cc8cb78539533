/*
 * ORACLE — TEST INFRASTRUCTURE ONLY (see gip_oracle.h).
 * Plain-C restatement of the reference CPU path, one partition.
 */
#include "gip_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Row loops (products, axpy) run on OpenMP threads above this many rows:
 * each row's value is computed by one thread with the serial expression, so
 * results do not depend on the thread count.  Reductions (dot, norm, CFL)
 * stay serial, in index order, as in the reference. */
#define ORC_PAR_MIN 65536

void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

/* Blocked dot products (off by default): partial sums of ORC_DOT_BLOCK
 * consecutive terms, each in index order, computed in parallel and added in
 * block order.  The result does not depend on the thread count; it differs
 * from the serial sum by rounding only (a different association). */
#define ORC_DOT_BLOCK 4096
static int g_blocked_dot = 0;
void orc_set_blocked_dot(int on) { g_blocked_dot = on != 0; }

static double dot_blocked(const double *y, const double *x, int n) {
    const int nb = (n + ORC_DOT_BLOCK - 1) / ORC_DOT_BLOCK;
    double *part = (double *)malloc(sizeof(double) * (size_t)(nb > 0 ? nb : 1));
#pragma omp parallel for schedule(static)
    for (int b = 0; b < nb; b++) {
        const int i1 = (b + 1) * ORC_DOT_BLOCK < n ? (b + 1) * ORC_DOT_BLOCK : n;
        double t = 0.0;
        for (int i = b * ORC_DOT_BLOCK; i < i1; i++) t += y[i] * x[i];
        part[b] = t;
    }
    double t = 0.0;
    for (int b = 0; b < nb; b++) t += part[b];
    free(part);
    return t;
}

/* ---- files --------------------------------------------------------------- */

/* GIP_MatrixMulticore.cpp:5-31 */
int orc_csr_read(const char *path, orc_csr *A) {
    memset(A, 0, sizeof(*A));
    FILE *fp = fopen(path, "rb");
    if (!fp) return -1;
    int sizes[5];
    if (fread(sizes, sizeof(int), 5, fp) != 5) { fclose(fp); return -1; }
    A->nnz = sizes[0];
    A->ncols = sizes[1];
    A->nrows = sizes[2];
    A->val = (double *)malloc(sizeof(double) * (size_t)(A->nnz > 0 ? A->nnz : 1));
    A->col = (int *)malloc(sizeof(int) * (size_t)(A->nnz > 0 ? A->nnz : 1));
    A->rowptr = (int *)malloc(sizeof(int) * (size_t)(A->nrows + 1));
    int ok = fread(A->val, sizeof(double), (size_t)A->nnz, fp) == (size_t)A->nnz &&
             fread(A->col, sizeof(int), (size_t)A->nnz, fp) == (size_t)A->nnz &&
             fread(A->rowptr, sizeof(int), (size_t)A->nrows + 1, fp) == (size_t)A->nrows + 1;
    fclose(fp);
    if (!ok) { orc_csr_free(A); return -1; }
    return 0;
}

void orc_csr_free(orc_csr *A) {
    free(A->val);
    free(A->col);
    free(A->rowptr);
    memset(A, 0, sizeof(*A));
}

/* GIP_VectorMulticore.cpp:16-42 */
static double *read_vec(const char *path, int size) {
    FILE *fp = fopen(path, "rb");
    if (!fp) return NULL;
    int n = 0;
    double *v = (double *)calloc((size_t)size, sizeof(double));
    if (fread(&n, sizeof(int), 1, fp) != 1 || n > size || fread(v, sizeof(double), (size_t)n, fp) != (size_t)n) {
        fclose(fp);
        free(v);
        return NULL;
    }
    fclose(fp);
    return v;
}

/* GIP_Topo.cpp:3-57 (sizes only; one partition has no halo lists) */
static int read_topo(const char *path, int sizes[9]) {
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    for (int i = 0; i < 9; ++i)
        if (fscanf(fp, " %d", &sizes[i]) != 1) { fclose(fp); return -1; }
    fclose(fp);
    return 0;
}

/* ---- Multicore matrix ---------------------------------------------------- */

/* GIP_MatrixMulticore.cpp:64-83 */
void orc_spmv(const orc_csr *A, const double *x, double *y) {
#pragma omp parallel for schedule(static) if (A->nrows > ORC_PAR_MIN)
    for (int i = 0; i < A->nrows; i++) {
        double sum = 0;
        int jb = A->rowptr[i], je = A->rowptr[i + 1];
        for (int j = jb; j < je; j++) sum += x[A->col[j]] * A->val[j];
        if (jb != je) y[i] = sum;
    }
}

/* GIP_MatrixMulticore.cpp:85-103 */
void orc_spmv_ab(const orc_csr *A, const double *x, double *y, double alpha, double beta) {
#pragma omp parallel for schedule(static) if (A->nrows > ORC_PAR_MIN)
    for (int i = 0; i < A->nrows; i++) {
        double sum = 0;
        int jb = A->rowptr[i], je = A->rowptr[i + 1];
        for (int j = jb; j < je; j++) sum += x[A->col[j]] * A->val[j];
        if (jb != je) y[i] = alpha * sum + beta * y[i];
    }
}

/* GIP_MatrixMulticore.cpp:105-123 */
void orc_spmv_vals(const orc_csr *A, const double *x, double *bval) { orc_spmv(A, x, bval); }

/* GIP_MatrixMulticore.cpp:125-143: B's values = alpha*(Ax) + beta*B's values, non-empty rows */
void orc_spmv_vals_ab(const orc_csr *A, const double *x, double *bval, double alpha, double beta) {
    orc_spmv_ab(A, x, bval, alpha, beta);
}

/* GIP_MatrixMulticore.cpp:177-210 */
int orc_diag_setup(const orc_csr *A, orc_csr *D) {
    int n = A->nrows, missing = 0;
    D->nrows = D->ncols = D->nnz = n;
    D->val = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    D->col = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    D->rowptr = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    for (int i = 0; i < n; i++) {
        int found = 0;
        for (int j = A->rowptr[i]; j < A->rowptr[i + 1]; j++)
            if (A->col[j] == i) {
                D->val[i] = 1.0 / A->val[j];
                found = 1;
            }
        if (!found) { D->val[i] = 0.0; missing = 1; }
        D->col[i] = i;
        D->rowptr[i] = i;
    }
    D->rowptr[n] = n;
    return missing ? -1 : 0;
}

/* ---- Multicore vector ---------------------------------------------------- */

void orc_axpy(double *y, const double *x, double alpha, int n) {
#pragma omp parallel for schedule(static) if (n > ORC_PAR_MIN)
    for (int i = 0; i < n; i++) y[i] += alpha * x[i];
}

void orc_axpby(double *y, const double *x, double alpha, double beta, int n) {
#pragma omp parallel for schedule(static) if (n > ORC_PAR_MIN)
    for (int i = 0; i < n; i++) y[i] = alpha * x[i] + beta * y[i];
}

double orc_dot(const double *y, const double *x, int n) {
    if (g_blocked_dot && n > ORC_PAR_MIN) return dot_blocked(y, x, n);
    double t = 0.0;
    for (int i = 0; i < n; i++) t += y[i] * x[i];
    return t;
}

double orc_norm_inf(const double *x, int n) {
    double t = 0.0;
    for (int i = 0; i < n; i++)
        if (fabs(x[i]) > t) t = fabs(x[i]);
    return t;
}

/* GIP_NonLinearMulticore.cpp:3-55 */
double orc_cfl(int inner, const double *u, const double *v, const double *w, const double *dxs, double gamma, double rho) {
    double dt = 1e20, conv = 0.25, normit;
    for (int i = 0; i < inner; i++) {
        double ut = u[i], vt = v[i], wt = w[i], dx = dxs[i];
        normit = fabs(ut) > fabs(vt) ? fabs(ut) : fabs(vt);
        if (normit < fabs(wt)) normit = fabs(wt);
        if (!(dx / normit < 1e-17))
            if (dt > conv * dx / normit) dt = conv * dx / normit;
        if (!(dx * dx * rho / gamma < 1e-17))
            if (dt > 0.2 * dx * dx * rho / gamma) dt = 0.2 * dx * dx * rho / gamma;
    }
    return 0.8 * dt > 1e-1 ? 1e-1 : 0.8 * dt;
}

/* ---- Jacobi-PCG, GIP_LinearSolver.h:45-117 ------------------------------- */

int orc_pcg(const orc_csr *A, const orc_csr *Dinv, int n, int n_all, const double *b, double *x, const double *x0,
            int maxit, double tol, int precond, int *iters, double *relres) {
    double *dp = (double *)malloc(sizeof(double) * (size_t)n_all);
    double *dAx = (double *)malloc(sizeof(double) * (size_t)n_all);
    double *dr = (double *)calloc((size_t)n_all, sizeof(double));
    double *ds = (double *)calloc((size_t)n_all, sizeof(double));
    int status = 0, k = 0;
    memcpy(dp, x0, sizeof(double) * (size_t)n_all);   /* :55-56 */
    memcpy(dr, b, sizeof(double) * (size_t)n);        /* :65 */
    memcpy(dAx, x0, sizeof(double) * (size_t)n_all);  /* :66 */
    memcpy(x, x0, sizeof(double) * (size_t)n_all);    /* :67 */
    orc_spmv(A, x, dAx);                              /* :69 */
    orc_axpy(dr, dAx, -1.0, n);                       /* :70 */
    if (precond) orc_spmv(Dinv, dr, dp);              /* :72-79 */
    else memcpy(dp, dr, sizeof(double) * (size_t)n);
    double r1 = orc_dot(dr, dp, n);                   /* :80 */
    double bnorm = 0.0, rnorm = 0.0;
    if (tol > 0.0) {
        bnorm = sqrt(orc_dot(b, b, n));
        rnorm = sqrt(orc_dot(dr, dr, n));
    }
    if (!(tol > 0.0 && rnorm <= tol * bnorm)) {
        for (k = 1; k <= maxit; k++) {                /* :85 */
            orc_spmv(A, dp, dAx);                     /* :89 */
            double dn = orc_dot(dp, dAx, n);          /* :91 */
            double alpha = r1 / dn;                   /* :93 */
            if (dn == 0.0 || !isfinite(dn) || !isfinite(alpha)) { status = -2; k--; break; }
            orc_axpy(x, dp, alpha, n);                /* :95 */
            orc_axpy(dr, dAx, -1.0 * alpha, n);       /* :97 */
            if (precond) orc_spmv(Dinv, dr, ds);      /* :99-105 */
            else memcpy(ds, dr, sizeof(double) * (size_t)n);
            double r0 = r1;                           /* :107 */
            r1 = orc_dot(dr, ds, n);                  /* :109 */
            double beta = r1 / r0;                    /* :111 */
            orc_axpby(dp, ds, 1.0, beta, n);          /* :113 */
            if (tol > 0.0) {
                rnorm = sqrt(orc_dot(dr, dr, n));
                if (rnorm <= tol * bnorm) break;
            }
        }
        if (k > maxit) k = maxit;
    }
    if (iters) *iters = k;
    if (relres) {
        if (!(tol > 0.0)) {
            bnorm = sqrt(orc_dot(b, b, n));
            rnorm = sqrt(orc_dot(dr, dr, n));
        }
        *relres = bnorm > 0.0 ? rnorm / bnorm : rnorm;
    }
    free(dp); free(dAx); free(dr); free(ds);
    return status;
}

/* ---- integrator, GIP_DNS.h ----------------------------------------------- */

static int path_op(char *buf, size_t sz, const char *root, const char *group, const char *op, const char *name) {
    return snprintf(buf, sz, "%s/InputFiles/Operators/%s/%s/%s_1p/%s%s_1p-0.csr", root, group, name, name, op, name);
}

int orc_dns_open(const char *root, const char *name, orc_dns **out) {
    char path[4096];
    int cs[9], fs[9];
    *out = NULL;
    snprintf(path, sizeof path, "%s/InputFiles/Topos/%s/%s_1p/topoCols%s_1p-0.topo", root, name, name, name);
    if (read_topo(path, cs)) return -1;
    snprintf(path, sizeof path, "%s/InputFiles/Topos/%s/%s_1p/topoFaces%s_1p-0.topo", root, name, name, name);
    if (read_topo(path, fs)) return -1;
    if (cs[8] != 0) return -1; /* one partition: no halo */
    orc_dns *s = (orc_dns *)calloc(1, sizeof(orc_dns));
    s->inner = cs[5] + cs[6];
    s->all = cs[4];
    s->faces = fs[4];
    struct { const char *g, *o; orc_csr *A; } ops[] = {
        {"CD", "D", &s->D}, {"CD", "D", &s->C}, {"CD", "Ec", &s->Ec}, {"L", "L", &s->L}, {"L", "Lb", &s->Lb},
        {"M", "Mx", &s->Mx}, {"M", "My", &s->My}, {"M", "Mz", &s->Mz}, {"G", "Gx", &s->Gx}, {"G", "Gy", &s->Gy},
        {"G", "Gz", &s->Gz}, {"S", "SFx", &s->SFx}, {"S", "SFy", &s->SFy}, {"S", "SFz", &s->SFz}, {"S", "SO", &s->SO},
        {"M", "Mres", &s->Mres}};
    for (size_t i = 0; i < sizeof ops / sizeof ops[0]; ++i) {
        path_op(path, sizeof path, root, ops[i].g, ops[i].o, name);
        if (orc_csr_read(path, ops[i].A)) { orc_dns_close(s); return -1; }
    }
    double **cellv[] = {&s->ax, &s->ay, &s->az, &s->ax0, &s->ay0, &s->az0, &s->rhs, &s->p, &s->p0,
                        &s->gpx, &s->gpy, &s->gpz, &s->temp};
    for (size_t i = 0; i < sizeof cellv / sizeof cellv[0]; ++i) *cellv[i] = (double *)calloc((size_t)s->all, sizeof(double));
    s->mf = (double *)calloc((size_t)s->faces, sizeof(double));
    const char *fn[] = {"u", "v", "w", "dxs"};
    double **fv[] = {&s->u, &s->v, &s->w, &s->dxs};
    for (int f = 0; f < 4; ++f) {
        snprintf(path, sizeof path, "%s/InputFiles/Scalarfields/%s/%s_1p/%s%s_1p-0.vec", root, name, name, fn[f], name);
        *fv[f] = read_vec(path, s->all);
        if (!*fv[f]) { orc_dns_close(s); return -1; }
    }
    orc_diag_setup(&s->L, &s->Dinv);
    /* GIP_DNS.h:132-137, :180-185 */
    s->rho = 1.0;
    s->gamma = 1.0 / 1.0e3;
    s->dt = 1e-3;
    s->time = 0.0;
    s->pcg_maxit = 100;
    s->pcg_tol = 0.0;
    s->precond = 1;
    *out = s;
    return 0;
}

/* The operator set in the order of orc_dns_open_mem's ops argument. */
static orc_csr *const *mem_ops(orc_dns *s, orc_csr *slots[15]) {
    orc_csr *t[15] = {&s->D,  &s->Ec, &s->L,   &s->Lb,  &s->Mx,  &s->My, &s->Mz, &s->Gx,
                      &s->Gy, &s->Gz, &s->SFx, &s->SFy, &s->SFz, &s->SO, &s->Mres};
    memcpy(slots, t, sizeof t);
    return slots;
}

/* GIP_DNS.h:115-306 with the operators and fields taken from memory: the
 * operators are borrowed (they must outlive the state), C gets its own
 * value array over D's structure (its values are rewritten every step,
 * GIP_DNS.h:312), the fields are copied. */
int orc_dns_open_mem(int inner, int all, int faces, const orc_csr *ops, const double *u, const double *v,
                     const double *w, const double *dxs, orc_dns **out) {
    *out = NULL;
    if (inner <= 0 || all < inner || faces < 0 || !ops || !u || !v || !w || !dxs) return -1;
    orc_dns *s = (orc_dns *)calloc(1, sizeof(orc_dns));
    s->inner = inner;
    s->all = all;
    s->faces = faces;
    s->borrowed = 1;
    orc_csr *slots[15];
    mem_ops(s, slots);
    for (int i = 0; i < 15; ++i) *slots[i] = ops[i];
    s->C = s->D;
    s->C.val = (double *)malloc(sizeof(double) * (size_t)(s->D.nnz > 0 ? s->D.nnz : 1));
    memcpy(s->C.val, s->D.val, sizeof(double) * (size_t)s->D.nnz);
    double **cellv[] = {&s->ax, &s->ay, &s->az, &s->ax0, &s->ay0, &s->az0, &s->rhs, &s->p, &s->p0,
                        &s->gpx, &s->gpy, &s->gpz, &s->temp};
    for (size_t i = 0; i < sizeof cellv / sizeof cellv[0]; ++i) *cellv[i] = (double *)calloc((size_t)s->all, sizeof(double));
    s->mf = (double *)calloc((size_t)s->faces > 0 ? (size_t)s->faces : 1, sizeof(double));
    const double *src[] = {u, v, w, dxs};
    double **fv[] = {&s->u, &s->v, &s->w, &s->dxs};
    for (int f = 0; f < 4; ++f) {
        *fv[f] = (double *)malloc(sizeof(double) * (size_t)s->all);
        memcpy(*fv[f], src[f], sizeof(double) * (size_t)s->all);
    }
    orc_diag_setup(&s->L, &s->Dinv);
    s->rho = 1.0;
    s->gamma = 1.0 / 1.0e3;
    s->dt = 1e-3;
    s->time = 0.0;
    s->pcg_maxit = 100;
    s->pcg_tol = 0.0;
    s->precond = 1;
    *out = s;
    return 0;
}

void orc_dns_close(orc_dns *s) {
    if (!s) return;
    if (s->borrowed) {
        free(s->C.val);
        orc_csr_free(&s->Dinv);
    } else {
        orc_csr *ops[] = {&s->C, &s->Ec, &s->D, &s->L, &s->Lb, &s->Mx, &s->My, &s->Mz, &s->Gx,
                          &s->Gy, &s->Gz, &s->SFx, &s->SFy, &s->SFz, &s->SO, &s->Mres, &s->Dinv};
        for (size_t i = 0; i < sizeof ops / sizeof ops[0]; ++i) orc_csr_free(ops[i]);
    }
    double *vs[] = {s->u, s->v, s->w, s->ax, s->ay, s->az, s->ax0, s->ay0, s->az0, s->rhs, s->p,
                    s->p0, s->gpx, s->gpy, s->gpz, s->mf, s->dxs, s->temp};
    for (size_t i = 0; i < sizeof vs / sizeof vs[0]; ++i) free(vs[i]);
    free(s);
}

double *orc_dns_field(orc_dns *s, const char *name, int *n) {
    struct { const char *nm; double *v; } t[] = {{"u", s->u},     {"v", s->v},     {"w", s->w},     {"p", s->p},
                                                 {"p0", s->p0},   {"ax0", s->ax0}, {"ay0", s->ay0}, {"az0", s->az0},
                                                 {"rhs", s->rhs}, {"mf", s->mf}};
    for (size_t i = 0; i < sizeof t / sizeof t[0]; ++i)
        if (!strcmp(name, t[i].nm)) {
            if (n) *n = strcmp(name, "mf") ? s->all : s->faces;
            return t[i].v;
        }
    return NULL;
}

void orc_dns_params(orc_dns *s, double gamma, double rho, int pcg_maxit, double pcg_tol, int precond) {
    s->gamma = gamma;
    s->rho = rho;
    s->pcg_maxit = pcg_maxit;
    s->pcg_tol = pcg_tol;
    s->precond = precond;
}

int orc_dns_step(orc_dns *s, double *time, double *dt_used, double *massr, int *iters) {
    const int n = s->inner;
    if (!s->started) { /* GIP_DNS.h:101-102 */
        s->dt = orc_cfl(n, s->u, s->v, s->w, s->dxs, s->gamma, s->rho);
        s->sigma = 1.0 / s->dt;
        s->started = 1;
    }
    /* PredictorVelocity, GIP_DNS.h:308-338 */
    orc_spmv_vals(&s->Ec, s->mf, s->C.val);
    orc_spmv(&s->D, s->u, s->ax);
    orc_spmv(&s->D, s->v, s->ay);
    orc_spmv(&s->D, s->w, s->az);
    orc_spmv_ab(&s->C, s->u, s->ax, -1.0, 1.0);
    orc_spmv_ab(&s->C, s->v, s->ay, -1.0, 1.0);
    orc_spmv_ab(&s->C, s->w, s->az, -1.0, 1.0);
    double a1 = 1.5 * s->dt, a2 = -0.5 * s->dt;
    orc_axpy(s->u, s->ax, a1, n);
    orc_axpy(s->u, s->ax0, a2, n);
    orc_axpy(s->v, s->ay, a1, n);
    orc_axpy(s->v, s->ay0, a2, n);
    orc_axpy(s->w, s->az, a1, n);
    orc_axpy(s->w, s->az0, a2, n);
    memcpy(s->ax0, s->ax, sizeof(double) * (size_t)n);
    memcpy(s->ay0, s->ay, sizeof(double) * (size_t)n);
    memcpy(s->az0, s->az, sizeof(double) * (size_t)n);
    /* PoissonEquation, GIP_DNS.h:340-354 */
    orc_spmv(&s->Mx, s->u, s->rhs);
    orc_spmv_ab(&s->My, s->v, s->rhs, 1.0, 1.0);
    orc_spmv_ab(&s->Mz, s->w, s->rhs, s->sigma, s->sigma);
    int st = orc_pcg(&s->L, &s->Dinv, n, s->all, s->rhs, s->p, s->p0, s->pcg_maxit, s->pcg_tol, s->precond,
                     &s->last_iters, &s->last_relres);
    memcpy(s->p0, s->p, sizeof(double) * (size_t)n);
    memcpy(s->p, s->p0, sizeof(double) * (size_t)s->all); /* OWNED == ALL on one partition */
    orc_spmv_ab(&s->Lb, s->p0, s->p, 1.0, 1.0);
    /* VelocityCorrection, GIP_DNS.h:356-376 */
    orc_spmv(&s->Gx, s->p, s->gpx);
    orc_spmv(&s->Gy, s->p, s->gpy);
    orc_spmv(&s->Gz, s->p, s->gpz);
    double fact = -1.0 / (s->rho * s->sigma);
    orc_axpy(s->u, s->gpx, fact, n);
    orc_axpy(s->v, s->gpy, fact, n);
    orc_axpy(s->w, s->gpz, fact, n);
    /* MassFluxes, GIP_DNS.h:378-391 */
    double f1 = -1.0 / s->sigma, f2 = 1.0 / s->sigma;
    orc_spmv_ab(&s->SO, s->p, s->mf, f1, 0.0);
    orc_spmv_ab(&s->SFx, s->gpx, s->mf, f2, 1.0);
    orc_spmv_ab(&s->SFy, s->gpy, s->mf, f2, 1.0);
    orc_spmv_ab(&s->SFz, s->gpz, s->mf, f2, 1.0);
    orc_spmv_ab(&s->SFx, s->u, s->mf, s->rho, 1.0);
    orc_spmv_ab(&s->SFy, s->v, s->mf, s->rho, 1.0);
    orc_spmv_ab(&s->SFz, s->w, s->mf, s->rho, 1.0);
    /* CFLCondition, GIP_DNS.h:393-414 */
    orc_spmv(&s->Mres, s->mf, s->temp);
    double mr = orc_norm_inf(s->temp, n);
    s->time += s->dt;
    if (dt_used) *dt_used = s->dt;
    if (time) *time = s->time;
    if (massr) *massr = mr;
    if (iters) *iters = s->last_iters;
    s->dt = orc_cfl(n, s->u, s->v, s->w, s->dxs, s->gamma, s->rho);
    s->sigma = 1.0 / s->dt;
    s->it++;
    return st;
}
