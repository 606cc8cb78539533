// Reverse-mode derivative of the time integration (include/gip_adjoint.h).
//
// Forward step n (dt_n frozen, sigma = 1/dt_n), as GIP_DNS.h:308-391:
//   c   = Ec mf                      (C's values)
//   ax  = D u - C(c) u               (ay, az alike; inner rows)
//   u*  = u + a1 ax + a2 ax0         (inner rows; ghosts unchanged), ax0' = ax
//   rhs = sigma (Mx u* + My v* + Mz w*)
//   p   = L^-1 rhs ; pe = p on the inner cells, Lb p on the ghosts
//   g   = G pe ; u' = u* + f g       (f = -1/(rho sigma))
//   mf' = f1 SO pe + f2 SF g + rho SF u'   on interior faces (f1 = -1/sigma, f2 = 1/sigma);
//         boundary faces keep mf
// Reverse step, with adjoints (u~, ax0~, mf~) of step n+1's inputs:
//   u~ += rho SF^T mf~ ; g~ = f2 SF^T mf~ ; pe~ = f1 SO^T mf~ ; mf~_old = mf~ on boundary faces
//   g~ += f u~ (inner) ; pe~ += G^T g~ ; p~ = pe~ + Lb^T pe~ (inner)
//   rhs~ = L^-1 p~ (L symmetric) ; u~ += sigma M^T rhs~
//   ax~ = ax0~ + a1 u~ (inner) ; ax0~ = a2 u~ (inner)
//   u~ += D^T ax~ - C^T ax~ ; c~ = -(ax~ (x) u + ay~ (x) v + az~ (x) w) on C's pattern ; mf~_old += Ec^T c~
// Ghost entries of u~ pass through every step: the lid ghosts hold the same
// parameter at every step, so at the end they carry its gradient.
#include <cmath>
#include <new>
#include <vector>

#include "gip_adjoint.h"
#include "gip_blas1.h"
#include "gip_internal.h"

namespace gip {

__global__ __launch_bounds__(kBlock) void k_energy(int n, const double *__restrict__ vol, const double *__restrict__ u,
                                                   const double *__restrict__ v, const double *__restrict__ w,
                                                   double *__restrict__ part) {
    double acc = 0.0;
    const int stride = gridDim.x * kBlock;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
        acc = acc + vol[i] * (u[i] * u[i] + v[i] * v[i] + w[i] * w[i]);
    acc = block_reduce<RED_SUM>(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// one workgroup: J = 0.5 * (sum of the partials in index order)
__global__ __launch_bounds__(kBlock) void k_energy_fin(const double *__restrict__ part, int np, double *__restrict__ out) {
    const double tot = sum_partials<RED_SUM>(part, np);
    if (threadIdx.x == 0) *out = 0.5 * tot;
}

// seed: y_i = V_i x_i on the inner cells
__global__ __launch_bounds__(kBlock) void k_vmul(int n, const double *__restrict__ vol, const double *__restrict__ x,
                                                 double *__restrict__ y) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[i] = vol[i] * x[i];
}

// c~_e = -(ax~_r u_c + ay~_r v_c + az~_r w_c) for every entry e = (r, c) of C
__global__ __launch_bounds__(kBlock) void k_cbar(int nrows, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                 const double *__restrict__ bax, const double *__restrict__ bay,
                                                 const double *__restrict__ baz, const double *__restrict__ u,
                                                 const double *__restrict__ v, const double *__restrict__ w,
                                                 double *__restrict__ cbar) {
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= nrows) return;
    const double a = bax[r], b = bay[r], c = baz[r];
    for (int e = rp[r]; e < rp[r + 1]; ++e) {
        const int j = col[e];
        cbar[e] = -(a * u[j] + b * v[j] + c * w[j]);
    }
}

// y_i = mask_i ? x_i : 0
__global__ __launch_bounds__(kBlock) void k_mask(int n, const int32_t *__restrict__ mask, const double *__restrict__ x,
                                                 double *__restrict__ y) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[i] = mask[i] ? x[i] : 0.0;
}

// u[g_k] = lid_k
__global__ void k_scatter(int n, const int32_t *__restrict__ idx, const double *__restrict__ vals, double *__restrict__ y) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) y[idx[k]] = vals[k];
}

// out_k = y[g_k]
__global__ void k_gather(int n, const int32_t *__restrict__ idx, const double *__restrict__ y, double *__restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = y[idx[k]];
}

}  // namespace gip

using namespace gip;

namespace {
enum { A_D = 0, A_EC, A_L, A_LB, A_MX, A_MY, A_MZ, A_GX, A_GY, A_GZ, A_SFX, A_SFY, A_SFZ, A_SO, A_N };
}

struct gip_adjoint {
    gip_ctx *c = nullptr;
    gip_adjoint_opts o{};
    int N = 0, NA = 0, F = 0, NLID = 0, NNZC = 0;
    gip_mat *op[A_N] = {};
    gip_mat *C = nullptr;
    gip_pcg *pcg = nullptr;
    int32_t *d_lid = nullptr, *d_crp = nullptr, *d_ccol = nullptr, *d_bmask = nullptr;
    double *d_vol = nullptr, *d_base[3] = {}, *d_lidv = nullptr, *d_dxs = nullptr;
    // forward state and intermediates
    double *u = nullptr, *v = nullptr, *w = nullptr, *ax0 = nullptr, *ay0 = nullptr, *az0 = nullptr, *mf = nullptr,
           *p0 = nullptr, *p = nullptr, *ax = nullptr, *ay = nullptr, *az = nullptr, *rhs = nullptr, *gx = nullptr,
           *gy = nullptr, *gz = nullptr, *scal = nullptr, *part = nullptr;
    // adjoint state and work
    double *bu = nullptr, *bv = nullptr, *bw = nullptr, *bax0 = nullptr, *bay0 = nullptr, *baz0 = nullptr,
           *bmf = nullptr, *bgx = nullptr, *bgy = nullptr, *bgz = nullptr, *bpe = nullptr, *bp = nullptr,
           *brhs = nullptr, *bax = nullptr, *bay = nullptr, *baz = nullptr, *cbar = nullptr, *tmp = nullptr,
           *zero = nullptr, *bmfo = nullptr;
    std::vector<double> dts;
    bool frozen = false, have_forward = false;
    std::vector<double *> ckpt;  // packed states
    double *seg = nullptr;       // segment inputs: K * (3 NA + F)
    std::vector<void *> allocs;
};

namespace {

int alloc(gip_adjoint *a, double **p, size_t n) {
    hipError_t e = hipMalloc(p, sizeof(double) * (n ? n : 1));
    if (e != hipSuccess) return hip_fail(e, "adjoint allocation");
    a->allocs.push_back(*p);
    e = hipMemsetAsync(*p, 0, sizeof(double) * (n ? n : 1), a->c->stream);
    return e == hipSuccess ? GIP_OK : hip_fail(e, "adjoint zero");
}

template <class T>
int upload(gip_adjoint *a, T **p, const T *host, size_t n) {
    hipError_t e = hipMalloc(p, sizeof(T) * (n ? n : 1));
    if (e != hipSuccess) return hip_fail(e, "adjoint allocation");
    a->allocs.push_back(*p);
    if (n) e = hipMemcpyAsync(*p, host, sizeof(T) * n, hipMemcpyHostToDevice, a->c->stream);
    return e == hipSuccess ? GIP_OK : hip_fail(e, "adjoint upload");
}

inline int nb(int n) { return (n + kBlock - 1) / kBlock; }

#define TRY(x)                    \
    do {                          \
        int rc__ = (x);           \
        if (rc__) return rc__;    \
    } while (0)

int copy(gip_adjoint *a, double *dst, const double *src, int n) {
    return gip_memcpy_d2d(a->c, dst, src, sizeof(double) * (size_t)n);
}

// y[0:n] += s * (A^T x)[0:n]
int tmul(gip_adjoint *a, gip_mat *A, const double *x, int n, double s, double *y) {
    TRY(gip_dcsrmv_t(a->c, A, x, a->tmp, n));
    return gip_daxpy(a->c, n, s, a->tmp, y);
}

// one forward step with frozen dt (GIP_DNS.h:308-391)
int fwd_step(gip_adjoint *a, double dt) {
    gip_ctx *c = a->c;
    const double sigma = 1.0 / dt, rho = a->o.rho;
    const double a1 = 1.5 * dt, a2 = -0.5 * dt;
    const int N = a->N, NA = a->NA;
    TRY(gip_dcsrmv_vals(c, a->op[A_EC], a->mf, a->C));
    TRY(gip_dcsrmv(c, a->op[A_D], a->u, a->ax));
    TRY(gip_dcsrmv(c, a->op[A_D], a->v, a->ay));
    TRY(gip_dcsrmv(c, a->op[A_D], a->w, a->az));
    TRY(gip_dcsrmv_ab(c, a->C, a->u, a->ax, -1.0, 1.0));
    TRY(gip_dcsrmv_ab(c, a->C, a->v, a->ay, -1.0, 1.0));
    TRY(gip_dcsrmv_ab(c, a->C, a->w, a->az, -1.0, 1.0));
    TRY(gip_daxpy(c, N, a1, a->ax, a->u));
    TRY(gip_daxpy(c, N, a2, a->ax0, a->u));
    TRY(gip_daxpy(c, N, a1, a->ay, a->v));
    TRY(gip_daxpy(c, N, a2, a->ay0, a->v));
    TRY(gip_daxpy(c, N, a1, a->az, a->w));
    TRY(gip_daxpy(c, N, a2, a->az0, a->w));
    TRY(copy(a, a->ax0, a->ax, N));
    TRY(copy(a, a->ay0, a->ay, N));
    TRY(copy(a, a->az0, a->az, N));
    TRY(gip_dcsrmv(c, a->op[A_MX], a->u, a->rhs));
    TRY(gip_dcsrmv_ab(c, a->op[A_MY], a->v, a->rhs, 1.0, 1.0));
    TRY(gip_dcsrmv_ab(c, a->op[A_MZ], a->w, a->rhs, sigma, sigma));
    gip_pcg_opts po{a->o.pcg_maxit, a->o.pcg_tol, 1, GIP_PCG_FP64, 0.0, 32, 0};
    gip_pcg_stats ps{};
    TRY(gip_pcg_solve(a->pcg, a->rhs, a->p, a->p0, &po, &ps));
    TRY(copy(a, a->p0, a->p, N));
    TRY(copy(a, a->p, a->p0, NA));
    TRY(gip_dcsrmv_ab(c, a->op[A_LB], a->p0, a->p, 1.0, 1.0));
    TRY(gip_dcsrmv(c, a->op[A_GX], a->p, a->gx));
    TRY(gip_dcsrmv(c, a->op[A_GY], a->p, a->gy));
    TRY(gip_dcsrmv(c, a->op[A_GZ], a->p, a->gz));
    const double fact = -1.0 / (rho * sigma);
    TRY(gip_daxpy(c, N, fact, a->gx, a->u));
    TRY(gip_daxpy(c, N, fact, a->gy, a->v));
    TRY(gip_daxpy(c, N, fact, a->gz, a->w));
    const double f1 = -1.0 / sigma, f2 = 1.0 / sigma;
    TRY(gip_dcsrmv_ab(c, a->op[A_SO], a->p, a->mf, f1, 0.0));
    TRY(gip_dcsrmv_ab(c, a->op[A_SFX], a->gx, a->mf, f2, 1.0));
    TRY(gip_dcsrmv_ab(c, a->op[A_SFY], a->gy, a->mf, f2, 1.0));
    TRY(gip_dcsrmv_ab(c, a->op[A_SFZ], a->gz, a->mf, f2, 1.0));
    TRY(gip_dcsrmv_ab(c, a->op[A_SFX], a->u, a->mf, rho, 1.0));
    TRY(gip_dcsrmv_ab(c, a->op[A_SFY], a->v, a->mf, rho, 1.0));
    TRY(gip_dcsrmv_ab(c, a->op[A_SFZ], a->w, a->mf, rho, 1.0));
    return GIP_OK;
}

// one reverse step (see the file header); (u, v, w, mf) hold step n's inputs
int adj_step(gip_adjoint *a, double dt) {
    gip_ctx *c = a->c;
    const double sigma = 1.0 / dt, rho = a->o.rho;
    const double a1 = 1.5 * dt, a2 = -0.5 * dt;
    const double f = -1.0 / (rho * sigma), f1 = -1.0 / sigma, f2 = 1.0 / sigma;
    const int N = a->N, NA = a->NA, F = a->F;
    TRY(gip_dcsrmv_vals(c, a->op[A_EC], a->mf, a->C));
    // mass fluxes
    TRY(tmul(a, a->op[A_SFX], a->bmf, NA, rho, a->bu));
    TRY(tmul(a, a->op[A_SFY], a->bmf, NA, rho, a->bv));
    TRY(tmul(a, a->op[A_SFZ], a->bmf, NA, rho, a->bw));
    TRY(gip_dfill(c, NA, a->bgx, 0.0));
    TRY(gip_dfill(c, NA, a->bgy, 0.0));
    TRY(gip_dfill(c, NA, a->bgz, 0.0));
    TRY(gip_dfill(c, NA, a->bpe, 0.0));
    TRY(tmul(a, a->op[A_SFX], a->bmf, NA, f2, a->bgx));
    TRY(tmul(a, a->op[A_SFY], a->bmf, NA, f2, a->bgy));
    TRY(tmul(a, a->op[A_SFZ], a->bmf, NA, f2, a->bgz));
    TRY(tmul(a, a->op[A_SO], a->bmf, NA, f1, a->bpe));
    k_mask<<<nb(F), kBlock, 0, c->stream>>>(F, a->d_bmask, a->bmf, a->bmfo);
    GIP_LAUNCH_CHECK("adjoint boundary faces");
    // velocity correction
    TRY(gip_daxpy(c, N, f, a->bu, a->bgx));
    TRY(gip_daxpy(c, N, f, a->bv, a->bgy));
    TRY(gip_daxpy(c, N, f, a->bw, a->bgz));
    TRY(tmul(a, a->op[A_GX], a->bgx, NA, 1.0, a->bpe));
    TRY(tmul(a, a->op[A_GY], a->bgy, NA, 1.0, a->bpe));
    TRY(tmul(a, a->op[A_GZ], a->bgz, NA, 1.0, a->bpe));
    // ghost extension pe = (p, Lb p), then the adjoint pressure solve on the symmetric L
    TRY(gip_dfill(c, NA, a->bp, 0.0));
    TRY(copy(a, a->bp, a->bpe, N));
    TRY(tmul(a, a->op[A_LB], a->bpe, N, 1.0, a->bp));
    gip_pcg_opts po{a->o.pcg_maxit, a->o.pcg_tol, 1, GIP_PCG_FP64, 0.0, 32, 0};
    gip_pcg_stats ps{};
    TRY(gip_pcg_solve(a->pcg, a->bp, a->brhs, a->zero, &po, &ps));
    TRY(tmul(a, a->op[A_MX], a->brhs, NA, sigma, a->bu));
    TRY(tmul(a, a->op[A_MY], a->brhs, NA, sigma, a->bv));
    TRY(tmul(a, a->op[A_MZ], a->brhs, NA, sigma, a->bw));
    // Adams-Bashforth history
    TRY(copy(a, a->bax, a->bax0, N));
    TRY(copy(a, a->bay, a->bay0, N));
    TRY(copy(a, a->baz, a->baz0, N));
    TRY(gip_daxpy(c, N, a1, a->bu, a->bax));
    TRY(gip_daxpy(c, N, a1, a->bv, a->bay));
    TRY(gip_daxpy(c, N, a1, a->bw, a->baz));
    TRY(gip_daxpby(c, N, a2, a->bu, 0.0, a->bax0));
    TRY(gip_daxpby(c, N, a2, a->bv, 0.0, a->bay0));
    TRY(gip_daxpby(c, N, a2, a->bw, 0.0, a->baz0));
    // convection and diffusion: u~ += D^T ax~ - C^T ax~, then C's values through Ec
    double *bx[3] = {a->bax, a->bay, a->baz}, *bu3[3] = {a->bu, a->bv, a->bw};
    for (int k = 0; k < 3; ++k) {
        TRY(tmul(a, a->op[A_D], bx[k], NA, 1.0, bu3[k]));
        TRY(tmul(a, a->C, bx[k], NA, -1.0, bu3[k]));
    }
    k_cbar<<<nb(N), kBlock, 0, c->stream>>>(N, a->d_crp, a->d_ccol, a->bax, a->bay, a->baz, a->u, a->v, a->w, a->cbar);
    GIP_LAUNCH_CHECK("adjoint convection");
    TRY(gip_dcsrmv_t(c, a->op[A_EC], a->cbar, a->bmf, F));
    TRY(gip_daxpy(c, F, 1.0, a->bmfo, a->bmf));
    return GIP_OK;
}

size_t state_len(const gip_adjoint *a) { return 7 * (size_t)a->NA + (size_t)a->F; }

int save_state(gip_adjoint *a, double *dst) {
    double *src[7] = {a->u, a->v, a->w, a->ax0, a->ay0, a->az0, a->p0};
    for (int k = 0; k < 7; ++k) TRY(copy(a, dst + (size_t)k * a->NA, src[k], a->NA));
    return copy(a, dst + 7 * (size_t)a->NA, a->mf, a->F);
}

int load_state(gip_adjoint *a, const double *s) {
    double *dst[7] = {a->u, a->v, a->w, a->ax0, a->ay0, a->az0, a->p0};
    for (int k = 0; k < 7; ++k) TRY(copy(a, dst[k], s + (size_t)k * a->NA, a->NA));
    return copy(a, a->mf, s + 7 * (size_t)a->NA, a->F);
}

}  // namespace

extern "C" {

int gip_adjoint_create(gip_ctx *c, const gip_cavity *cav, const gip_adjoint_opts *o, gip_adjoint **out) {
    if (!c || !cav || !o || !out || o->steps < 1 || o->checkpoint_every < 1 || o->pcg_maxit < 1 || !(o->pcg_tol > 0))
        return set_error(GIP_EINVAL, "gip_adjoint_create: bad argument");
    if (!cav->lid_ghost || !cav->vol || cav->nx < 1) return set_error(GIP_EINVAL, "gip_adjoint_create: needs a generated hexahedral cavity");
    *out = nullptr;
    gip_adjoint *a = new (std::nothrow) gip_adjoint();
    if (!a) return set_error(GIP_ENOMEM, "gip_adjoint_create: allocation failed");
    a->c = c;
    a->o = *o;
    if (a->o.rho <= 0) a->o.rho = 1.0;
    a->N = cav->n_inner;
    a->NA = cav->n_all;
    a->F = cav->n_faces;
    a->NLID = cav->nx * cav->ny;
    a->NNZC = cav->D.nnz;
    const gip_csr *ops[A_N] = {&cav->D, &cav->Ec, &cav->L, &cav->Lb, &cav->Mx, &cav->My, &cav->Mz,
                               &cav->Gx, &cav->Gy, &cav->Gz, &cav->SFx, &cav->SFy, &cav->SFz, &cav->SO};
    int rc = GIP_OK;
    for (int k = 0; k < A_N && !rc; ++k) {
        if (!ops[k]->rowptr) rc = set_error(GIP_EINVAL, "gip_adjoint_create: a cavity operator is missing");
        if (!rc) rc = gip_mat_create(c, ops[k]->nrows, ops[k]->ncols, ops[k]->nnz, ops[k]->rowptr, ops[k]->col,
                                     ops[k]->val, GIP_FMT_AUTO, &a->op[k]);
    }
    if (!rc) rc = gip_mat_create(c, cav->D.nrows, cav->D.ncols, cav->D.nnz, cav->D.rowptr, cav->D.col, cav->D.val,
                                 GIP_FMT_AUTO, &a->C);
    if (!rc) rc = gip_pcg_create(c, a->op[A_L], a->NA, &a->pcg);
    const size_t NA = (size_t)a->NA, F = (size_t)a->F;
    double **cells[] = {&a->u, &a->v, &a->w, &a->ax0, &a->ay0, &a->az0, &a->p0, &a->p, &a->ax, &a->ay, &a->az,
                        &a->rhs, &a->gx, &a->gy, &a->gz, &a->bu, &a->bv, &a->bw, &a->bax0, &a->bay0, &a->baz0,
                        &a->bgx, &a->bgy, &a->bgz, &a->bpe, &a->bp, &a->brhs, &a->bax, &a->bay, &a->baz, &a->zero};
    for (double **f : cells)
        if (!rc) rc = alloc(a, f, NA);
    if (!rc) rc = alloc(a, &a->mf, F);
    if (!rc) rc = alloc(a, &a->bmf, F);
    if (!rc) rc = alloc(a, &a->bmfo, F);
    if (!rc) rc = alloc(a, &a->cbar, (size_t)a->NNZC);
    if (!rc) rc = alloc(a, &a->tmp, std::max(std::max(NA, F), (size_t)a->NNZC));
    if (!rc) rc = alloc(a, &a->scal, 8);
    if (!rc) rc = alloc(a, &a->part, kMaxRedBlocks);
    // boundary faces: rows of SO without entries
    std::vector<int32_t> bmask(F, 0);
    for (size_t fi = 0; fi < F; ++fi) bmask[fi] = cav->SO.rowptr[fi + 1] == cav->SO.rowptr[fi];
    if (!rc) rc = upload(a, &a->d_bmask, bmask.data(), F);
    if (!rc) rc = upload(a, &a->d_lid, cav->lid_ghost, (size_t)a->NLID);
    if (!rc) rc = upload(a, &a->d_crp, cav->D.rowptr, (size_t)cav->D.nrows + 1);
    if (!rc) rc = upload(a, &a->d_ccol, cav->D.col, (size_t)a->NNZC);
    if (!rc) rc = upload(a, &a->d_vol, cav->vol, (size_t)a->N);
    if (!rc) rc = upload(a, &a->d_base[0], cav->u, NA);
    if (!rc) rc = upload(a, &a->d_base[1], cav->v, NA);
    if (!rc) rc = upload(a, &a->d_base[2], cav->w, NA);
    if (!rc) rc = upload(a, &a->d_dxs, cav->dxs, NA);
    if (!rc) rc = alloc(a, &a->d_lidv, (size_t)a->NLID);
    const int K = o->checkpoint_every, nseg = (o->steps + K - 1) / K;
    for (int s = 0; s < nseg && !rc; ++s) {
        double *ck = nullptr;
        rc = alloc(a, &ck, state_len(a));
        a->ckpt.push_back(ck);
    }
    if (!rc) rc = alloc(a, &a->seg, (size_t)K * (3 * NA + F));
    if (!rc) rc = gip_sync(c);
    a->dts.assign((size_t)o->steps, 0.0);
    if (rc) {
        gip_adjoint_destroy(a);
        return rc;
    }
    *out = a;
    return GIP_OK;
}

int gip_adjoint_forward(gip_adjoint *a, const double *lid, const double *u0, const double *v0, const double *w0,
                        double *J) {
    if (!a || !lid) return set_error(GIP_EINVAL, "gip_adjoint_forward: bad argument");
    gip_ctx *c = a->c;
    const int N = a->N, NA = a->NA;
    // initial state: cavity fields, inner velocity and lid ghosts from the parameters
    double *dst[3] = {a->u, a->v, a->w};
    const double *init[3] = {u0, v0, w0};
    for (int k = 0; k < 3; ++k) {
        TRY(copy(a, dst[k], a->d_base[k], NA));
        if (init[k])
            TRY(gip_memcpy_h2d(c, dst[k], init[k], sizeof(double) * (size_t)N));
        else
            TRY(gip_dfill(c, N, dst[k], 0.0));
    }
    TRY(gip_memcpy_h2d(c, a->d_lidv, lid, sizeof(double) * (size_t)a->NLID));
    k_scatter<<<(a->NLID + 255) / 256, 256, 0, c->stream>>>(a->NLID, a->d_lid, a->d_lidv, a->u);
    GIP_LAUNCH_CHECK("lid parameters");
    double *zeros[] = {a->ax0, a->ay0, a->az0, a->p0, a->p};
    for (double *z : zeros) TRY(gip_dfill(c, NA, z, 0.0));
    TRY(gip_dfill(c, a->F, a->mf, 0.0));
    const int K = a->o.checkpoint_every;
    double dt = 0.0;
    // the first run takes each step's dt from the CFL rule (GIP_DNS.h:101, :408); later runs reuse them
    if (!a->frozen) TRY(gip_cfl(c, N, a->u, a->v, a->w, a->d_dxs, a->o.gamma, a->o.rho, &dt));
    for (int n = 0; n < a->o.steps; ++n) {
        if (n % K == 0) TRY(save_state(a, a->ckpt[(size_t)(n / K)]));
        if (!a->frozen) a->dts[(size_t)n] = dt;
        TRY(fwd_step(a, a->dts[(size_t)n]));
        if (!a->frozen && n + 1 < a->o.steps) TRY(gip_cfl(c, N, a->u, a->v, a->w, a->d_dxs, a->o.gamma, a->o.rho, &dt));
    }
    a->frozen = true;
    k_energy<<<reduce_grid(N), kBlock, 0, c->stream>>>(N, a->d_vol, a->u, a->v, a->w, a->part);
    k_energy_fin<<<1, kBlock, 0, c->stream>>>(a->part, reduce_grid(N), a->scal);
    GIP_LAUNCH_CHECK("objective");
    double h = 0.0;
    TRY(gip_memcpy_d2h(c, &h, a->scal, sizeof(double)));
    if (J) *J = h;
    a->have_forward = true;
    return GIP_OK;
}

int gip_adjoint_backward(gip_adjoint *a, double *glid, double *gu0, double *gv0, double *gw0) {
    if (!a || !a->have_forward) return set_error(GIP_EINVAL, "gip_adjoint_backward: run gip_adjoint_forward first");
    gip_ctx *c = a->c;
    const int N = a->N, NA = a->NA, F = a->F, K = a->o.checkpoint_every, steps = a->o.steps;
    // seeds: dJ/du = V u on the inner cells
    double *bu3[3] = {a->bu, a->bv, a->bw}, *u3[3] = {a->u, a->v, a->w};
    for (int k = 0; k < 3; ++k) {
        TRY(gip_dfill(c, NA, bu3[k], 0.0));
        k_vmul<<<nb(N), kBlock, 0, c->stream>>>(N, a->d_vol, u3[k], bu3[k]);
    }
    GIP_LAUNCH_CHECK("adjoint seed");
    double *ba0[3] = {a->bax0, a->bay0, a->baz0};
    for (double *z : ba0) TRY(gip_dfill(c, NA, z, 0.0));
    TRY(gip_dfill(c, F, a->bmf, 0.0));
    const size_t slot = 3 * (size_t)NA + (size_t)F;
    for (int s = (steps + K - 1) / K - 1; s >= 0; --s) {
        const int n0 = s * K, n1 = std::min(steps, n0 + K);
        TRY(load_state(a, a->ckpt[(size_t)s]));
        for (int n = n0; n < n1; ++n) {  // recompute the segment, keeping each step's inputs
            double *in = a->seg + (size_t)(n - n0) * slot;
            TRY(copy(a, in, a->u, NA));
            TRY(copy(a, in + NA, a->v, NA));
            TRY(copy(a, in + 2 * (size_t)NA, a->w, NA));
            TRY(copy(a, in + 3 * (size_t)NA, a->mf, F));
            TRY(fwd_step(a, a->dts[(size_t)n]));
        }
        for (int n = n1 - 1; n >= n0; --n) {
            const double *in = a->seg + (size_t)(n - n0) * slot;
            TRY(copy(a, a->u, in, NA));
            TRY(copy(a, a->v, in + NA, NA));
            TRY(copy(a, a->w, in + 2 * (size_t)NA, NA));
            TRY(copy(a, a->mf, in + 3 * (size_t)NA, F));
            TRY(adj_step(a, a->dts[(size_t)n]));
        }
    }
    if (glid) {
        k_gather<<<(a->NLID + 255) / 256, 256, 0, c->stream>>>(a->NLID, a->d_lid, a->bu, a->d_lidv);
        GIP_LAUNCH_CHECK("lid gradient");
        TRY(gip_memcpy_d2h(c, glid, a->d_lidv, sizeof(double) * (size_t)a->NLID));
    }
    double *outs[3] = {gu0, gv0, gw0};
    for (int k = 0; k < 3; ++k)
        if (outs[k]) TRY(gip_memcpy_d2h(c, outs[k], bu3[k], sizeof(double) * (size_t)N));
    return GIP_OK;
}

int gip_adjoint_dts(gip_adjoint *a, double *dts) {
    if (!a || !dts || !a->frozen) return set_error(GIP_EINVAL, "gip_adjoint_dts: no forward run yet");
    for (size_t k = 0; k < a->dts.size(); ++k) dts[k] = a->dts[k];
    return GIP_OK;
}

int gip_adjoint_destroy(gip_adjoint *a) {
    if (!a) return GIP_OK;
    hipStreamSynchronize(a->c->stream);
    if (a->pcg) gip_pcg_destroy(a->pcg);
    for (int k = 0; k < A_N; ++k)
        if (a->op[k]) gip_mat_destroy(a->op[k]);
    if (a->C) gip_mat_destroy(a->C);
    for (void *p : a->allocs) hipFree(p);
    delete a;
    return GIP_OK;
}

}  // extern "C"
