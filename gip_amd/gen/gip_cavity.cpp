// Lid-driven-cavity operator generator for hexahedral grids and the
// reference's operator file formats (reader, writer, partition merger).
// C ABI in include/gip_cavity.h; relations and formats are cited there.
//
// Grid conventions (one partition):
//   cell (i,j,k)  -> i + nx*(j + ny*k)                       [0, N)
//   x-face (i,j,k), i in [0,nx]  -> i + (nx+1)*(j + ny*k)      [0, Fx)
//   y-face (i,j,k), j in [0,ny]  -> Fx + i + nx*(j + (ny+1)*k)
//   z-face (i,j,k), k in [0,nz]  -> Fx + Fy + i + nx*(j + ny*k)
//   face normals point towards +x/+y/+z; Mres(i,f) = +1 when the normal
//   leaves cell i, -1 when it enters it.
//   one boundary ghost per boundary face, numbered in face order, stored
//   after the inner cells: [inner | ghosts] (GIP_Topo.h:8-18 with no halo).
#include "gip_cavity.h"

#include <sys/mman.h>
#include <sys/stat.h>
#include <sys/types.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

// Large operator arrays are 2 MiB aligned and advised as huge pages: the
// first touch of a 256^3 case otherwise costs millions of 4 KiB page faults.
template <class T>
T *alloc_n(size_t n) {
    if (n == 0) n = 1;
    size_t bytes = n * sizeof(T);
    if (bytes < (size_t(8) << 20)) return static_cast<T *>(std::malloc(bytes));
    const size_t huge = size_t(2) << 20;
    void *p = nullptr;
    if (posix_memalign(&p, huge, (bytes + huge - 1) / huge * huge) != 0) return nullptr;
    madvise(p, (bytes + huge - 1) / huge * huge, MADV_HUGEPAGE);
    return static_cast<T *>(p);
}

void csr_zero(gip_csr *A) { std::memset(A, 0, sizeof(*A)); }

// Allocate a CSR whose row lengths are given; fills rowptr by prefix sum.
int csr_alloc(gip_csr *A, int32_t nrows, int32_t ncols, const std::vector<int32_t> &len) {
    int64_t nnz = 0;
    A->rowptr = alloc_n<int32_t>((size_t)nrows + 1);
    if (!A->rowptr) return fail(GIP_CAV_ENOMEM, "rowptr allocation failed");
    for (int32_t r = 0; r < nrows; ++r) {
        A->rowptr[r] = (int32_t)nnz;
        nnz += len[r];
    }
    if (nnz > INT32_MAX) return fail(GIP_CAV_EINVAL, "nnz exceeds int32 range");
    A->rowptr[nrows] = (int32_t)nnz;
    A->nnz = (int32_t)nnz;
    A->nrows = nrows;
    A->ncols = ncols;
    A->val = alloc_n<double>((size_t)nnz);
    A->col = alloc_n<int32_t>((size_t)nnz);
    if (!A->val || !A->col) return fail(GIP_CAV_ENOMEM, "csr value allocation failed");
    return GIP_CAV_OK;
}

struct Hex {
    int nx, ny, nz;
    int64_t N, Fx, Fy, Fz, F, Gx, Gy, Gz, G;
    std::vector<double> xf, yf, zf, hx, hy, hz, xc, yc, zc;

    int64_t cell(int i, int j, int k) const { return i + (int64_t)nx * (j + (int64_t)ny * k); }
    int64_t fx(int i, int j, int k) const { return i + (int64_t)(nx + 1) * (j + (int64_t)ny * k); }
    int64_t fy(int i, int j, int k) const { return Fx + i + (int64_t)nx * (j + (int64_t)(ny + 1) * k); }
    int64_t fz(int i, int j, int k) const { return Fx + Fy + i + (int64_t)nx * (j + (int64_t)ny * k); }
    int64_t gx(int j, int k, int side) const { return 2 * (j + (int64_t)ny * k) + side; }
    int64_t gy(int i, int k, int side) const { return Gx + 2 * (int64_t)nx * k + (int64_t)side * nx + i; }
    int64_t gz(int i, int j, int side) const { return Gx + Gy + (int64_t)side * nx * ny + i + (int64_t)nx * j; }
};

void axis_coords(int n, double beta, std::vector<double> &f, std::vector<double> &h, std::vector<double> &c) {
    f.resize(n + 1);
    h.resize(n);
    c.resize(n);
    for (int i = 0; i <= n; ++i) {
        double s = (double)i / n;
        if (beta > 0.0)
            f[i] = 0.5 * (1.0 + std::tanh(beta * (2.0 * s - 1.0)) / std::tanh(beta));
        else
            f[i] = s;
    }
    f[0] = 0.0;
    f[n] = 1.0;
    for (int i = 0; i < n; ++i) {
        h[i] = f[i + 1] - f[i];
        c[i] = 0.5 * (f[i] + f[i + 1]);
    }
}

// One side of a cell: s = 0..5 for -x,+x,-y,+y,-z,+z.
struct Side {
    int64_t face;   // global face id
    int sign;       // Mres(i, face)
    double area;    // face area
    bool bnd;       // boundary face
    int64_t nb;     // neighbour cell (if !bnd) or ghost index in [0,G) (if bnd)
    double dist;    // centre-centre distance, or centre-wall distance on the boundary
};

void cell_sides(const Hex &h, int i, int j, int k, Side s[6]) {
    const int64_t c = h.cell(i, j, k);
    double ax = h.hy[j] * h.hz[k], ay = h.hx[i] * h.hz[k], az = h.hx[i] * h.hy[j];
    // -x
    s[0].face = h.fx(i, j, k); s[0].sign = -1; s[0].area = ax; s[0].bnd = (i == 0);
    s[0].nb = s[0].bnd ? h.gx(j, k, 0) : c - 1;
    s[0].dist = s[0].bnd ? h.xc[i] - h.xf[i] : h.xc[i] - h.xc[i - 1];
    // +x
    s[1].face = h.fx(i + 1, j, k); s[1].sign = +1; s[1].area = ax; s[1].bnd = (i == h.nx - 1);
    s[1].nb = s[1].bnd ? h.gx(j, k, 1) : c + 1;
    s[1].dist = s[1].bnd ? h.xf[i + 1] - h.xc[i] : h.xc[i + 1] - h.xc[i];
    // -y
    s[2].face = h.fy(i, j, k); s[2].sign = -1; s[2].area = ay; s[2].bnd = (j == 0);
    s[2].nb = s[2].bnd ? h.gy(i, k, 0) : c - h.nx;
    s[2].dist = s[2].bnd ? h.yc[j] - h.yf[j] : h.yc[j] - h.yc[j - 1];
    // +y
    s[3].face = h.fy(i, j + 1, k); s[3].sign = +1; s[3].area = ay; s[3].bnd = (j == h.ny - 1);
    s[3].nb = s[3].bnd ? h.gy(i, k, 1) : c + h.nx;
    s[3].dist = s[3].bnd ? h.yf[j + 1] - h.yc[j] : h.yc[j + 1] - h.yc[j];
    // -z
    s[4].face = h.fz(i, j, k); s[4].sign = -1; s[4].area = az; s[4].bnd = (k == 0);
    s[4].nb = s[4].bnd ? h.gz(i, j, 0) : c - (int64_t)h.nx * h.ny;
    s[4].dist = s[4].bnd ? h.zc[k] - h.zf[k] : h.zc[k] - h.zc[k - 1];
    // +z
    s[5].face = h.fz(i, j, k + 1); s[5].sign = +1; s[5].area = az; s[5].bnd = (k == h.nz - 1);
    s[5].nb = s[5].bnd ? h.gz(i, j, 1) : c + (int64_t)h.nx * h.ny;
    s[5].dist = s[5].bnd ? h.zf[k + 1] - h.zc[k] : h.zc[k + 1] - h.zc[k];
}

// Ascending column order of the 7-entry cell row (L/D/C pattern): neighbour
// cells below, the cell itself, neighbour cells above, then ghost columns in
// ghost order.  Entry = side index, or -1 for the diagonal.
int row_order(const Side s[6], int order[7]) {
    int n = 0;
    const int below[3] = {4, 2, 0}, above[3] = {1, 3, 5};
    for (int t = 0; t < 3; ++t)
        if (!s[below[t]].bnd) order[n++] = below[t];
    order[n++] = -1;
    for (int t = 0; t < 3; ++t)
        if (!s[above[t]].bnd) order[n++] = above[t];
    for (int t = 0; t < 6; ++t)
        if (s[t].bnd) order[n++] = t;
    return n;
}

template <class F>
void for_cells(const Hex &h, F fn) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < h.nz; ++k)
        for (int j = 0; j < h.ny; ++j)
            for (int i = 0; i < h.nx; ++i) fn(i, j, k);
}

void free_contents(gip_cavity *c) {
    gip_csr *ops[] = {&c->D, &c->Ec, &c->L, &c->Lb, &c->Mx, &c->My, &c->Mz, &c->Gx,
                      &c->Gy, &c->Gz, &c->SFx, &c->SFy, &c->SFz, &c->SO, &c->Mres};
    for (gip_csr *A : ops) gip_csr_release(A);
    std::free(c->u); std::free(c->v); std::free(c->w); std::free(c->dxs);
    std::free(c->vol); std::free(c->lid_ghost);
}

gip_cavity *new_cavity() {
    gip_cavity *c = static_cast<gip_cavity *>(std::calloc(1, sizeof(gip_cavity)));
    return c;
}

int gen_cell7(const Hex &h, gip_csr *A, int which, double gamma, int pin) {
    // which: 0 = L, 1 = D
    const int64_t N = h.N;
    std::vector<int32_t> len((size_t)N, 7);
    int rc = csr_alloc(A, (int32_t)N, (int32_t)(N + h.G), len);
    if (rc) return rc;
    for_cells(h, [&](int i, int j, int k) {
        const int64_t c = h.cell(i, j, k);
        Side s[6];
        cell_sides(h, i, j, k, s);
        int order[7];
        row_order(s, order);
        const double V = h.hx[i] * h.hy[j] * h.hz[k];
        double coef[6];
        double diag = 0.0;
        for (int t = 0; t < 6; ++t) {
            if (which == 0) {
                coef[t] = s[t].bnd ? 0.0 : -s[t].area / s[t].dist;
                if (!s[t].bnd) diag += s[t].area / s[t].dist;
            } else {
                coef[t] = gamma * s[t].area / (s[t].dist * V);
                diag -= coef[t];
            }
        }
        if (which == 0 && pin && c == 0) diag *= 1.01;
        int64_t p = 7 * c;
        for (int e = 0; e < 7; ++e, ++p) {
            int t = order[e];
            if (t < 0) {
                A->col[p] = (int32_t)c;
                A->val[p] = diag;
            } else {
                A->col[p] = (int32_t)(s[t].bnd ? N + s[t].nb : s[t].nb);
                A->val[p] = coef[t];
            }
        }
    });
    return GIP_CAV_OK;
}

int gen_ec(const Hex &h, gip_csr *A) {
    const int64_t N = h.N;
    std::vector<int32_t> len((size_t)(7 * N), 1);
    // diagonal rows hold one (zero) weight per face of the cell
    for_cells(h, [&](int i, int j, int k) {
        Side s[6];
        cell_sides(h, i, j, k, s);
        int order[7];
        row_order(s, order);
        for (int e = 0; e < 7; ++e)
            if (order[e] < 0) len[(size_t)(7 * h.cell(i, j, k) + e)] = 6;
    });
    int rc = csr_alloc(A, (int32_t)(7 * N), (int32_t)h.F, len);
    if (rc) return rc;
    for_cells(h, [&](int i, int j, int k) {
        const int64_t c = h.cell(i, j, k);
        Side s[6];
        cell_sides(h, i, j, k, s);
        int order[7];
        row_order(s, order);
        const double V = h.hx[i] * h.hy[j] * h.hz[k];
        int64_t p = 12 * c;
        for (int e = 0; e < 7; ++e) {
            int t = order[e];
            if (t < 0) {
                for (int q = 0; q < 6; ++q, ++p) {  // faces ascend with side index
                    A->col[p] = (int32_t)s[q].face;
                    A->val[p] = 0.0;
                }
            } else {
                A->col[p] = (int32_t)s[t].face;
                A->val[p] = 0.5 * s[t].sign / V;
                ++p;
            }
        }
    });
    return GIP_CAV_OK;
}

int gen_mres(const Hex &h, gip_csr *A) {
    std::vector<int32_t> len((size_t)h.N, 6);
    int rc = csr_alloc(A, (int32_t)h.N, (int32_t)h.F, len);
    if (rc) return rc;
    for_cells(h, [&](int i, int j, int k) {
        Side s[6];
        cell_sides(h, i, j, k, s);
        int64_t p = 6 * h.cell(i, j, k);
        for (int t = 0; t < 6; ++t, ++p) {
            A->col[p] = (int32_t)s[t].face;
            A->val[p] = (double)s[t].sign;
        }
    });
    return GIP_CAV_OK;
}

// Divergence M{axis} (N rows) and gradient G{axis} (N+G rows, ghost rows
// empty); structural zeros of the hexahedral stencil are not stored.
int gen_m_g(const Hex &h, int axis, gip_csr *M, gip_csr *G) {
    const int64_t N = h.N;
    const int lo = 2 * axis, hi = 2 * axis + 1;
    auto entries = [&](int i, int j, int k, int64_t col[3], double mv[3], double gv[3]) {
        Side s[6];
        cell_sides(h, i, j, k, s);
        const int64_t c = h.cell(i, j, k);
        const double V = h.hx[i] * h.hy[j] * h.hz[k];
        const double A = s[lo].area;  // both faces of the axis share the area
        int n = 0;
        double aplus = s[hi].bnd ? 1.0 : 0.5, aminus = s[lo].bnd ? 1.0 : 0.5;
        double mdiag = (s[hi].bnd ? 0.0 : -0.5 * A) + (s[lo].bnd ? 0.0 : 0.5 * A);
        if (!s[lo].bnd) { col[n] = s[lo].nb; mv[n] = 0.5 * A; gv[n] = -(0.5 * A) / V; ++n; }
        if (aplus != aminus) { col[n] = c; mv[n] = mdiag; gv[n] = A * (aplus - aminus) / V; ++n; }
        if (!s[hi].bnd) { col[n] = s[hi].nb; mv[n] = -0.5 * A; gv[n] = (0.5 * A) / V; ++n; }
        return n;
    };
    std::vector<int32_t> len((size_t)N);
    for_cells(h, [&](int i, int j, int k) {
        int64_t col[3]; double mv[3], gv[3];
        len[(size_t)h.cell(i, j, k)] = entries(i, j, k, col, mv, gv);
    });
    int rc = csr_alloc(M, (int32_t)N, (int32_t)N, len);
    if (rc) return rc;
    std::vector<int32_t> glen((size_t)(N + h.G), 0);
    std::copy(len.begin(), len.end(), glen.begin());
    rc = csr_alloc(G, (int32_t)(N + h.G), (int32_t)(N + h.G), glen);
    if (rc) return rc;
    for_cells(h, [&](int i, int j, int k) {
        int64_t col[3]; double mv[3], gv[3];
        int n = entries(i, j, k, col, mv, gv);
        int64_t c = h.cell(i, j, k);
        int64_t p = M->rowptr[c];
        for (int e = 0; e < n; ++e) {
            M->col[p + e] = G->col[p + e] = (int32_t)col[e];
            M->val[p + e] = mv[e];
            G->val[p + e] = gv[e];
        }
    });
    return GIP_CAV_OK;
}

// Face operators: SF{axis} (axis>=0) or SO (axis<0); F rows, two entries
// per interior face (back cell, front cell), boundary rows empty.
int gen_face(const Hex &h, int axis, gip_csr *A) {
    const int64_t F = h.F;
    std::vector<int32_t> len((size_t)F, 0);
    // face -> (back, front, coefficient pair) computed per face block
    auto visit = [&](auto fn) {
#pragma omp parallel for schedule(static)
        for (int k = 0; k < h.nz; ++k)
            for (int j = 0; j < h.ny; ++j)
                for (int i = 1; i < h.nx; ++i) {  // interior x-faces
                    double Af = h.hy[j] * h.hz[k];
                    fn(h.fx(i, j, k), h.cell(i - 1, j, k), h.cell(i, j, k), 0, Af, h.xc[i] - h.xc[i - 1]);
                }
#pragma omp parallel for schedule(static)
        for (int k = 0; k < h.nz; ++k)
            for (int j = 1; j < h.ny; ++j)
                for (int i = 0; i < h.nx; ++i) {
                    double Af = h.hx[i] * h.hz[k];
                    fn(h.fy(i, j, k), h.cell(i, j - 1, k), h.cell(i, j, k), 1, Af, h.yc[j] - h.yc[j - 1]);
                }
#pragma omp parallel for schedule(static)
        for (int k = 1; k < h.nz; ++k)
            for (int j = 0; j < h.ny; ++j)
                for (int i = 0; i < h.nx; ++i) {
                    double Af = h.hx[i] * h.hy[j];
                    fn(h.fz(i, j, k), h.cell(i, j, k - 1), h.cell(i, j, k), 2, Af, h.zc[k] - h.zc[k - 1]);
                }
    };
    visit([&](int64_t f, int64_t, int64_t, int ax, double, double) {
        if (axis < 0 || ax == axis) len[(size_t)f] = 2;
    });
    int rc = csr_alloc(A, (int32_t)F, (int32_t)(h.N + h.G), len);
    if (rc) return rc;
    visit([&](int64_t f, int64_t b, int64_t c, int ax, double Af, double d) {
        if (!(axis < 0 || ax == axis)) return;
        int64_t p = A->rowptr[f];
        A->col[p] = (int32_t)b;
        A->col[p + 1] = (int32_t)c;
        if (axis < 0) {
            A->val[p] = -Af / d;
            A->val[p + 1] = Af / d;
        } else {
            A->val[p] = 0.5 * Af;
            A->val[p + 1] = 0.5 * Af;
        }
    });
    return GIP_CAV_OK;
}

// ghost index (0-based within ghosts) -> adjacent inner cell
int64_t ghost_adjacent(const Hex &h, int64_t g) {
    if (g < h.Gx) {
        int64_t side = g & 1, jk = g >> 1;
        int64_t j = jk % h.ny, k = jk / h.ny;
        return h.cell(side ? h.nx - 1 : 0, (int)j, (int)k);
    }
    g -= h.Gx;
    if (g < h.Gy) {
        int64_t k = g / (2 * (int64_t)h.nx), rem = g % (2 * (int64_t)h.nx);
        int64_t side = rem / h.nx, i = rem % h.nx;
        return h.cell((int)i, side ? h.ny - 1 : 0, (int)k);
    }
    g -= h.Gy;
    int64_t side = g / ((int64_t)h.nx * h.ny), ij = g % ((int64_t)h.nx * h.ny);
    return h.cell((int)(ij % h.nx), (int)(ij / h.nx), side ? h.nz - 1 : 0);
}

int gen_lb(const Hex &h, gip_csr *A) {
    std::vector<int32_t> len((size_t)(h.N + h.G), 0);
    for (int64_t g = 0; g < h.G; ++g) len[(size_t)(h.N + g)] = 1;
    int rc = csr_alloc(A, (int32_t)(h.N + h.G), (int32_t)(h.N + h.G), len);
    if (rc) return rc;
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < h.G; ++g) {
        int64_t p = A->rowptr[h.N + g];
        A->col[p] = (int32_t)ghost_adjacent(h, g);
        A->val[p] = 1.0;
    }
    return GIP_CAV_OK;
}

std::string join_path(const std::string &a, const std::string &b) { return a.empty() ? b : a + "/" + b; }

int mkdirs(const std::string &path) {
    std::string cur;
    size_t pos = 0;
    while (pos != std::string::npos) {
        pos = path.find('/', pos + 1);
        cur = path.substr(0, pos);
        if (cur.empty()) continue;
        if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) return fail(GIP_CAV_EIO, "cannot create directory " + cur);
    }
    return GIP_CAV_OK;
}

struct CasePaths {
    std::string root, name, tag;  // tag = <name>_<np>p
    std::string op(const char *group, const char *op, int rank) const {
        return join_path(root, "InputFiles/Operators/" + std::string(group) + "/" + name + "/" + tag + "/" + op + name + "_" +
                                   tag.substr(name.size() + 1) + "-" + std::to_string(rank) + ".csr");
    }
    std::string vec(const char *field, int rank) const {
        return join_path(root, "InputFiles/Scalarfields/" + name + "/" + tag + "/" + field + name + "_" + tag.substr(name.size() + 1) +
                                   "-" + std::to_string(rank) + ".vec");
    }
    std::string topo(const char *kind, int rank) const {
        return join_path(root, "InputFiles/Topos/" + name + "/" + tag + "/" + kind + name + "_" + tag.substr(name.size() + 1) + "-" +
                                   std::to_string(rank) + ".topo");
    }
    std::string dir(const char *kind, const char *group) const {
        return join_path(root, "InputFiles/" + std::string(kind) + "/" + (group[0] ? std::string(group) + "/" : "") + name + "/" + tag);
    }
};

CasePaths make_paths(const char *root, const char *name, int np) {
    CasePaths p;
    p.root = root ? root : "";
    p.name = name;
    p.tag = p.name + "_" + std::to_string(np) + "p";
    return p;
}

struct Topo {
    int sizes[9];
    std::vector<int32_t> send_indx, send_row, recv_indx, recv_row;
    int inner() const { return sizes[5] + sizes[6]; }
    int owned() const { return sizes[3]; }
    int all() const { return sizes[4]; }
};

int read_topo(const std::string &path, Topo *t) {
    FILE *fp = std::fopen(path.c_str(), "r");
    if (!fp) return fail(GIP_CAV_EIO, "cannot open topology " + path);
    bool ok = true;
    for (int i = 0; i < 9 && ok; ++i) ok = std::fscanf(fp, " %d", &t->sizes[i]) == 1;
    auto rd = [&](std::vector<int32_t> &v, int n) {
        if (!ok || n < 0) { ok = false; return; }
        v.resize((size_t)n);
        for (int i = 0; i < n && ok; ++i) ok = std::fscanf(fp, " %d", &v[(size_t)i]) == 1;
    };
    if (ok) {
        rd(t->send_indx, t->sizes[0]);
        rd(t->send_row, t->sizes[2]);
        rd(t->recv_indx, t->sizes[1]);
        rd(t->recv_row, t->sizes[2]);
    }
    std::fclose(fp);
    if (!ok) return fail(GIP_CAV_EIO, "malformed topology " + path);
    return GIP_CAV_OK;
}

int write_topo(const std::string &path, const int sizes[9]) {
    FILE *fp = std::fopen(path.c_str(), "w");
    if (!fp) return fail(GIP_CAV_EIO, "cannot write " + path);
    for (int i = 0; i < 9; ++i) std::fprintf(fp, " %d", sizes[i]);
    // no send/recv entries; send_row and recv_row are {0,0} for cols_size 2
    std::fprintf(fp, " \n \n 0 0 \n \n 0 0 \n");
    std::fclose(fp);
    return GIP_CAV_OK;
}

// A topology file with halo lists (format of GIP_Topo.cpp:12-49).
int write_topo_lists(const std::string &path, const int sizes[9], const std::vector<int32_t> &send_indx,
                     const std::vector<int32_t> &send_row, const std::vector<int32_t> &recv_indx,
                     const std::vector<int32_t> &recv_row) {
    FILE *fp = std::fopen(path.c_str(), "w");
    if (!fp) return fail(GIP_CAV_EIO, "cannot write " + path);
    for (int i = 0; i < 9; ++i) std::fprintf(fp, " %d", sizes[i]);
    auto put = [&](const std::vector<int32_t> &v) {
        std::fprintf(fp, " \n");
        for (int32_t x : v) std::fprintf(fp, " %d", x);
    };
    put(send_indx);
    put(send_row);
    put(recv_indx);
    put(recv_row);
    std::fprintf(fp, " \n");
    return std::fclose(fp) == 0 ? GIP_CAV_OK : fail(GIP_CAV_EIO, "write failed " + path);
}

}  // namespace

extern "C" {

const char *gip_cavity_last_error(void) { return g_err.c_str(); }

void gip_csr_release(gip_csr *A) {
    if (!A) return;
    std::free(A->val);
    std::free(A->col);
    std::free(A->rowptr);
    csr_zero(A);
}

int gip_csr_read(const char *path, gip_csr *A) {
    csr_zero(A);
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return fail(GIP_CAV_EIO, std::string("cannot open operator ") + path);
    int32_t h[5];
    if (std::fread(h, sizeof(int32_t), 5, fp) != 5 || h[0] < 0 || h[2] < 0) {
        std::fclose(fp);
        return fail(GIP_CAV_EIO, std::string("short operator header ") + path);
    }
    A->nnz = h[0];
    A->ncols = h[1];
    A->nrows = h[2];
    A->val = alloc_n<double>((size_t)A->nnz);
    A->col = alloc_n<int32_t>((size_t)A->nnz);
    A->rowptr = alloc_n<int32_t>((size_t)A->nrows + 1);
    if (!A->val || !A->col || !A->rowptr) {
        std::fclose(fp);
        gip_csr_release(A);
        return fail(GIP_CAV_ENOMEM, "operator allocation failed");
    }
    bool ok = std::fread(A->val, sizeof(double), (size_t)A->nnz, fp) == (size_t)A->nnz &&
              std::fread(A->col, sizeof(int32_t), (size_t)A->nnz, fp) == (size_t)A->nnz &&
              std::fread(A->rowptr, sizeof(int32_t), (size_t)A->nrows + 1, fp) == (size_t)A->nrows + 1;
    std::fclose(fp);
    if (!ok || A->rowptr[0] != 0 || A->rowptr[A->nrows] != A->nnz) {
        gip_csr_release(A);
        return fail(GIP_CAV_EIO, std::string("truncated or inconsistent operator ") + path);
    }
    return GIP_CAV_OK;
}

int gip_csr_write(const char *path, const gip_csr *A) {
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return fail(GIP_CAV_EIO, std::string("cannot write ") + path);
    int32_t h[5] = {A->nnz, A->ncols, A->nrows, (A->nrows + 1023) / 1024, 1024};
    bool ok = std::fwrite(h, sizeof(int32_t), 5, fp) == 5 &&
              std::fwrite(A->val, sizeof(double), (size_t)A->nnz, fp) == (size_t)A->nnz &&
              std::fwrite(A->col, sizeof(int32_t), (size_t)A->nnz, fp) == (size_t)A->nnz &&
              std::fwrite(A->rowptr, sizeof(int32_t), (size_t)A->nrows + 1, fp) == (size_t)A->nrows + 1;
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? GIP_CAV_OK : fail(GIP_CAV_EIO, std::string("write failed ") + path);
}

int gip_vec_read(const char *path, double **vals, int32_t *n) {
    *vals = nullptr;
    *n = 0;
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return fail(GIP_CAV_EIO, std::string("cannot open scalar field ") + path);
    int32_t m = 0;
    if (std::fread(&m, sizeof(int32_t), 1, fp) != 1 || m < 0) {
        std::fclose(fp);
        return fail(GIP_CAV_EIO, std::string("short scalar field header ") + path);
    }
    double *v = alloc_n<double>((size_t)m);
    if (!v) {
        std::fclose(fp);
        return fail(GIP_CAV_ENOMEM, "field allocation failed");
    }
    bool ok = std::fread(v, sizeof(double), (size_t)m, fp) == (size_t)m;
    std::fclose(fp);
    if (!ok) {
        std::free(v);
        return fail(GIP_CAV_EIO, std::string("truncated scalar field ") + path);
    }
    *vals = v;
    *n = m;
    return GIP_CAV_OK;
}

int gip_vec_write(const char *path, const double *vals, int32_t n) {
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return fail(GIP_CAV_EIO, std::string("cannot write ") + path);
    bool ok = std::fwrite(&n, sizeof(int32_t), 1, fp) == 1 && std::fwrite(vals, sizeof(double), (size_t)n, fp) == (size_t)n;
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? GIP_CAV_OK : fail(GIP_CAV_EIO, std::string("write failed ") + path);
}

void gip_cavity_free(gip_cavity *cav) {
    if (!cav) return;
    free_contents(cav);
    std::free(cav);
}

int gip_cavity_generate(const gip_cavity_desc *d, gip_cavity **out) {
    *out = nullptr;
    if (!d || d->nx < 2 || d->ny < 2 || d->nz < 2) return fail(GIP_CAV_EINVAL, "grid needs at least 2 cells per axis");
    if (!(d->gamma >= 0.0) || d->beta < 0.0) return fail(GIP_CAV_EINVAL, "gamma and beta must be non-negative");
    Hex h;
    h.nx = d->nx; h.ny = d->ny; h.nz = d->nz;
    h.N = (int64_t)h.nx * h.ny * h.nz;
    h.Fx = (int64_t)(h.nx + 1) * h.ny * h.nz;
    h.Fy = (int64_t)h.nx * (h.ny + 1) * h.nz;
    h.Fz = (int64_t)h.nx * h.ny * (h.nz + 1);
    h.F = h.Fx + h.Fy + h.Fz;
    h.Gx = 2 * (int64_t)h.ny * h.nz;
    h.Gy = 2 * (int64_t)h.nx * h.nz;
    h.Gz = 2 * (int64_t)h.nx * h.ny;
    h.G = h.Gx + h.Gy + h.Gz;
    if (12 * h.N > INT32_MAX || h.F > INT32_MAX) return fail(GIP_CAV_EINVAL, "grid too large for int32 indices");
    axis_coords(h.nx, d->beta, h.xf, h.hx, h.xc);
    axis_coords(h.ny, d->beta, h.yf, h.hy, h.yc);
    axis_coords(h.nz, d->beta, h.zf, h.hz, h.zc);

    gip_cavity *c = new_cavity();
    if (!c) return fail(GIP_CAV_ENOMEM, "cavity allocation failed");
    c->n_inner = (int32_t)h.N;
    c->n_ghost = (int32_t)h.G;
    c->n_all = (int32_t)(h.N + h.G);
    c->n_faces = (int32_t)h.F;
    c->nx = h.nx; c->ny = h.ny; c->nz = h.nz;
    const uint32_t ops = d->ops ? d->ops : (uint32_t)GIP_OP_ALL;
    int rc = GIP_CAV_OK;
    if (!rc && (ops & GIP_OP_D)) rc = gen_cell7(h, &c->D, 1, d->gamma, 0);
    if (!rc && (ops & GIP_OP_L)) rc = gen_cell7(h, &c->L, 0, 0.0, d->pin);
    if (!rc && (ops & GIP_OP_EC)) rc = gen_ec(h, &c->Ec);
    if (!rc && (ops & GIP_OP_LB)) rc = gen_lb(h, &c->Lb);
    if (!rc && (ops & (GIP_OP_M | GIP_OP_G))) {
        gip_csr *M[3] = {&c->Mx, &c->My, &c->Mz}, *G[3] = {&c->Gx, &c->Gy, &c->Gz};
        for (int a = 0; a < 3 && !rc; ++a) rc = gen_m_g(h, a, M[a], G[a]);
        if (!(ops & GIP_OP_M)) for (int a = 0; a < 3; ++a) gip_csr_release(M[a]);
        if (!(ops & GIP_OP_G)) for (int a = 0; a < 3; ++a) gip_csr_release(G[a]);
    }
    if (!rc && (ops & GIP_OP_SF)) {
        rc = gen_face(h, 0, &c->SFx);
        if (!rc) rc = gen_face(h, 1, &c->SFy);
        if (!rc) rc = gen_face(h, 2, &c->SFz);
    }
    if (!rc && (ops & GIP_OP_SO)) rc = gen_face(h, -1, &c->SO);
    if (!rc && (ops & GIP_OP_MRES)) rc = gen_mres(h, &c->Mres);
    if (!rc) {
        size_t na = (size_t)c->n_all;
        c->u = static_cast<double *>(std::calloc(na, sizeof(double)));
        c->v = static_cast<double *>(std::calloc(na, sizeof(double)));
        c->w = static_cast<double *>(std::calloc(na, sizeof(double)));
        c->dxs = static_cast<double *>(std::calloc(na, sizeof(double)));
        c->vol = alloc_n<double>((size_t)h.N);
        c->lid_ghost = alloc_n<int32_t>((size_t)h.nx * h.ny);
        if (!c->u || !c->v || !c->w || !c->dxs || !c->vol || !c->lid_ghost) rc = fail(GIP_CAV_ENOMEM, "field allocation failed");
    }
    if (!rc) {
        for_cells(h, [&](int i, int j, int k) {
            int64_t cc = h.cell(i, j, k);
            // CFL length: smallest edge / sqrt(3).  With the reference's
            // diffusive limit 0.8*0.2*dx^2/gamma this stays inside the
            // explicit Adams-Bashforth-2 stability bound h^2/(12 gamma) of
            // the 7-point diffusion operator (dx = h would exceed it 2x).
            c->dxs[cc] = std::min(h.hx[i], std::min(h.hy[j], h.hz[k])) / std::sqrt(3.0);
            c->vol[cc] = h.hx[i] * h.hy[j] * h.hz[k];
        });
        for (int j = 0; j < h.ny; ++j)
            for (int i = 0; i < h.nx; ++i) {
                int64_t g = h.N + h.gz(i, j, 1);
                c->lid_ghost[i + (size_t)h.nx * j] = (int32_t)g;
                c->u[g] = d->lid_profile ? d->lid_profile[i + (size_t)h.nx * j] : d->lid_u;
            }
    }
    if (rc) {
        gip_cavity_free(c);
        return rc;
    }
    *out = c;
    return GIP_CAV_OK;
}

int gip_cavity_write(const gip_cavity *c, const char *root, const char *name) {
    if (!c || !name || !name[0]) return fail(GIP_CAV_EINVAL, "cavity and name required");
    CasePaths p = make_paths(root, name, 1);
    const char *groups[] = {"CD", "L", "M", "G", "S"};
    for (const char *g : groups) {
        int rc = mkdirs(p.dir("Operators", g));
        if (rc) return rc;
    }
    int rc = mkdirs(p.dir("Scalarfields", ""));
    if (!rc) rc = mkdirs(p.dir("Topos", ""));
    if (rc) return rc;
    struct Item { const char *group, *op; const gip_csr *A; int32_t ncols; };
    const int32_t owned = c->n_all, inner = c->n_inner, F = c->n_faces;
    Item items[] = {{"CD", "D", &c->D, owned},     {"CD", "Ec", &c->Ec, F},     {"L", "L", &c->L, owned},
                    {"L", "Lb", &c->Lb, owned},    {"M", "Mx", &c->Mx, inner},  {"M", "My", &c->My, inner},
                    {"M", "Mz", &c->Mz, inner},    {"M", "Mres", &c->Mres, F},  {"G", "Gx", &c->Gx, owned},
                    {"G", "Gy", &c->Gy, owned},    {"G", "Gz", &c->Gz, owned},  {"S", "SFx", &c->SFx, owned},
                    {"S", "SFy", &c->SFy, owned},  {"S", "SFz", &c->SFz, owned}, {"S", "SO", &c->SO, owned}};
    for (const Item &it : items) {
        if (!it.A->rowptr) return fail(GIP_CAV_EINVAL, std::string("operator not built: ") + it.op);
        gip_csr A = *it.A;
        A.ncols = it.ncols;
        rc = gip_csr_write(p.op(it.group, it.op, 0).c_str(), &A);
        if (rc) return rc;
    }
    const char *fields[] = {"u", "v", "w", "dxs"};
    const double *fv[] = {c->u, c->v, c->w, c->dxs};
    for (int f = 0; f < 4; ++f) {
        rc = gip_vec_write(p.vec(fields[f], 0).c_str(), fv[f], c->n_all);
        if (rc) return rc;
    }
    int cols[9] = {0, 0, 2, c->n_all, c->n_all, c->n_inner, 0, c->n_ghost, 0};
    int faces[9] = {0, 0, 2, c->n_faces, c->n_faces, c->n_inner, 0, c->n_ghost, 0};
    rc = write_topo(p.topo("topoCols", 0), cols);
    if (!rc) rc = write_topo(p.topo("topoFaces", 0), faces);
    return rc;
}

int gip_cavity_read(const char *root, const char *name, gip_cavity **out) { return gip_cavity_merge(root, name, 1, out); }

int gip_cavity_rhs_normal(int32_t n, uint64_t seed, double *out) {
    if (n < 0 || (n > 0 && !out)) return fail(GIP_CAV_EINVAL, "bad rhs arguments");
    uint64_t s = seed;
    auto next = [&s]() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    };
    const double two_pi = 6.283185307179586476925286766559;
    double mean = 0.0;
    for (int32_t i = 0; i < n; ++i) {
        double u1 = ((double)(next() >> 11) + 1.0) * (1.0 / 9007199254740993.0);  // (0,1]
        double u2 = (double)(next() >> 11) * (1.0 / 9007199254740992.0);          // [0,1)
        out[i] = std::sqrt(-2.0 * std::log(u1)) * std::cos(two_pi * u2);
        mean += out[i];
    }
    if (n > 0) mean /= n;
    for (int32_t i = 0; i < n; ++i) out[i] -= mean;
    return GIP_CAV_OK;
}

int gip_cavity_merge(const char *root, const char *name, int32_t np, gip_cavity **out) {
    *out = nullptr;
    if (!name || np < 1) return fail(GIP_CAV_EINVAL, "name and a positive partition count required");
    CasePaths p = make_paths(root, name, np);
    std::vector<Topo> tc((size_t)np), tf((size_t)np);
    for (int r = 0; r < np; ++r) {
        int rc = read_topo(p.topo("topoCols", r), &tc[(size_t)r]);
        if (!rc) rc = read_topo(p.topo("topoFaces", r), &tf[(size_t)r]);
        if (rc) return rc;
        if (tc[(size_t)r].sizes[2] != np + 1) return fail(GIP_CAV_ETOPO, "topology column count does not match partition count");
    }
    std::vector<int64_t> off_in((size_t)np + 1, 0), off_gh((size_t)np + 1, 0), off_f((size_t)np + 1, 0);
    for (int r = 0; r < np; ++r) {
        const Topo &t = tc[(size_t)r];
        off_in[(size_t)r + 1] = off_in[(size_t)r] + t.inner();
        off_gh[(size_t)r + 1] = off_gh[(size_t)r] + (t.owned() - t.inner());
        off_f[(size_t)r + 1] = off_f[(size_t)r] + tf[(size_t)r].all();
    }
    const int64_t N = off_in[(size_t)np], G = off_gh[(size_t)np], F = off_f[(size_t)np];
    // local cell index -> merged index, per rank
    std::vector<std::vector<int32_t>> map((size_t)np);
    for (int r = 0; r < np; ++r) {
        const Topo &t = tc[(size_t)r];
        std::vector<int32_t> &m = map[(size_t)r];
        m.assign((size_t)t.all(), -1);
        for (int l = 0; l < t.owned(); ++l)
            m[(size_t)l] = (int32_t)(l < t.inner() ? off_in[(size_t)r] + l : N + off_gh[(size_t)r] + (l - t.inner()));
    }
    for (int r = 0; r < np; ++r) {
        const Topo &t = tc[(size_t)r];
        for (int q = 0; q < np; ++q) {
            const Topo &s = tc[(size_t)q];
            int nrecv = t.recv_row[(size_t)q + 1] - t.recv_row[(size_t)q];
            int nsend = s.send_row[(size_t)r + 1] - s.send_row[(size_t)r];
            if (nrecv != nsend) return fail(GIP_CAV_ETOPO, "send/recv list sizes disagree between partitions");
            for (int kk = 0; kk < nrecv; ++kk) {
                int slot = t.recv_indx[(size_t)(t.recv_row[(size_t)q] + kk)];
                int src = s.send_indx[(size_t)(s.send_row[(size_t)r] + kk)];
                if (slot < t.owned() || slot >= t.all() || src < 0 || src >= s.inner())
                    return fail(GIP_CAV_ETOPO, "halo slot or send index out of range");
                map[(size_t)r][(size_t)slot] = (int32_t)(off_in[(size_t)q] + src);
            }
        }
        for (int l = t.owned(); l < t.all(); ++l)
            if (map[(size_t)r][(size_t)l] < 0) return fail(GIP_CAV_ETOPO, "halo slot without a sender");
    }

    gip_cavity *c = new_cavity();
    if (!c) return fail(GIP_CAV_ENOMEM, "cavity allocation failed");
    c->n_inner = (int32_t)N;
    c->n_ghost = (int32_t)G;
    c->n_all = (int32_t)(N + G);
    c->n_faces = (int32_t)F;

    // kind: 0 = rows are inner cells (concat), 1 = rows are owned cells
    // (inner rows first, ghost rows after), 2 = rows are faces/C entries (concat)
    // colmap: 0 = cells, 1 = faces
    auto merge_op = [&](const char *group, const char *op, int kind, int colmap, int32_t ncols_out, gip_csr *dst) -> int {
        std::vector<gip_csr> parts((size_t)np);
        int rc = GIP_CAV_OK;
        for (int r = 0; r < np && !rc; ++r) rc = gip_csr_read(p.op(group, op, r).c_str(), &parts[(size_t)r]);
        if (!rc) {
            // row sources: (rank, local row) in merged row order
            std::vector<std::pair<int, int32_t>> src;
            for (int r = 0; r < np; ++r) {
                int32_t nr = parts[(size_t)r].nrows;
                int32_t lim = kind == 1 ? tc[(size_t)r].inner() : nr;
                if (kind == 1 && nr != tc[(size_t)r].owned()) rc = fail(GIP_CAV_ETOPO, std::string(op) + ": row count is not the owned size");
                for (int32_t l = 0; l < std::min(lim, nr); ++l) src.emplace_back(r, l);
            }
            if (kind == 1)
                for (int r = 0; r < np; ++r)
                    for (int32_t l = tc[(size_t)r].inner(); l < parts[(size_t)r].nrows; ++l) src.emplace_back(r, l);
            std::vector<int32_t> len(src.size());
            for (size_t e = 0; e < src.size(); ++e) {
                const gip_csr &A = parts[(size_t)src[e].first];
                len[e] = A.rowptr[src[e].second + 1] - A.rowptr[src[e].second];
            }
            if (!rc) rc = csr_alloc(dst, (int32_t)src.size(), ncols_out, len);
            for (size_t e = 0; e < src.size() && !rc; ++e) {
                int r = src[e].first;
                const gip_csr &A = parts[(size_t)r];
                int32_t b = A.rowptr[src[e].second], en = A.rowptr[src[e].second + 1];
                int32_t q = dst->rowptr[e];
                for (int32_t t = b; t < en; ++t, ++q) {
                    int32_t cl = A.col[t];
                    int32_t mc;
                    if (colmap == 0) {
                        if (cl < 0 || cl >= (int32_t)map[(size_t)r].size()) { rc = fail(GIP_CAV_ETOPO, std::string(op) + ": column out of range"); break; }
                        mc = map[(size_t)r][(size_t)cl];
                    } else {
                        if (cl < 0 || cl >= tf[(size_t)r].all()) { rc = fail(GIP_CAV_ETOPO, std::string(op) + ": face column out of range"); break; }
                        mc = (int32_t)(off_f[(size_t)r] + cl);
                    }
                    dst->col[q] = mc;
                    dst->val[q] = A.val[t];
                }
            }
        }
        for (auto &A : parts) gip_csr_release(&A);
        return rc;
    };
    const int32_t owned = c->n_all;
    int rc = merge_op("CD", "D", 0, 0, owned, &c->D);
    if (!rc) rc = merge_op("CD", "Ec", 2, 1, (int32_t)F, &c->Ec);
    if (!rc) rc = merge_op("L", "L", 0, 0, owned, &c->L);
    if (!rc) rc = merge_op("L", "Lb", 1, 0, owned, &c->Lb);
    if (!rc) rc = merge_op("M", "Mx", 0, 0, (int32_t)N, &c->Mx);
    if (!rc) rc = merge_op("M", "My", 0, 0, (int32_t)N, &c->My);
    if (!rc) rc = merge_op("M", "Mz", 0, 0, (int32_t)N, &c->Mz);
    if (!rc) rc = merge_op("M", "Mres", 0, 1, (int32_t)F, &c->Mres);
    if (!rc) rc = merge_op("G", "Gx", 1, 0, owned, &c->Gx);
    if (!rc) rc = merge_op("G", "Gy", 1, 0, owned, &c->Gy);
    if (!rc) rc = merge_op("G", "Gz", 1, 0, owned, &c->Gz);
    if (!rc) rc = merge_op("S", "SFx", 2, 0, owned, &c->SFx);
    if (!rc) rc = merge_op("S", "SFy", 2, 0, owned, &c->SFy);
    if (!rc) rc = merge_op("S", "SFz", 2, 0, owned, &c->SFz);
    if (!rc) rc = merge_op("S", "SO", 2, 0, owned, &c->SO);
    const char *fields[] = {"u", "v", "w", "dxs"};
    double **fv[] = {&c->u, &c->v, &c->w, &c->dxs};
    for (int f = 0; f < 4 && !rc; ++f) {
        *fv[f] = static_cast<double *>(std::calloc((size_t)c->n_all, sizeof(double)));
        if (!*fv[f]) { rc = fail(GIP_CAV_ENOMEM, "field allocation failed"); break; }
        for (int r = 0; r < np && !rc; ++r) {
            double *vals = nullptr;
            int32_t n = 0;
            rc = gip_vec_read(p.vec(fields[f], r).c_str(), &vals, &n);
            if (!rc && n < tc[(size_t)r].owned()) rc = fail(GIP_CAV_EIO, "scalar field shorter than the owned range");
            for (int l = 0; l < tc[(size_t)r].owned() && !rc; ++l) (*fv[f])[map[(size_t)r][(size_t)l]] = vals[l];
            std::free(vals);
        }
    }
    if (rc) {
        gip_cavity_free(c);
        return rc;
    }
    *out = c;
    return GIP_CAV_OK;
}

int gip_cavity_split(const gip_cavity *c, int32_t np, const char *root, const char *name) {
    if (!c || !name || !name[0] || np < 1) return fail(GIP_CAV_EINVAL, "cavity, name and a positive partition count required");
    const gip_csr *cellops[] = {&c->D, &c->L, &c->Mx, &c->My, &c->Mz, &c->Gx, &c->Gy, &c->Gz};
    const gip_csr *faceops[] = {&c->SFx, &c->SFy, &c->SFz, &c->SO};
    const gip_csr *all[] = {&c->D, &c->Ec, &c->L, &c->Lb, &c->Mx, &c->My, &c->Mz, &c->Gx,
                            &c->Gy, &c->Gz, &c->SFx, &c->SFy, &c->SFz, &c->SO, &c->Mres};
    for (const gip_csr *A : all)
        if (!A->rowptr) return fail(GIP_CAV_EINVAL, "gip_cavity_split: every operator must be built");
    const int64_t N = c->n_inner, NA = c->n_all, F = c->n_faces;
    if (np > N) return fail(GIP_CAV_EINVAL, "gip_cavity_split: more partitions than cells");
    if (c->Ec.nrows != c->D.nnz) return fail(GIP_CAV_ETOPO, "gip_cavity_split: Ec rows != nnz(D)");
    // owners: inner cells in contiguous index blocks (slabs of a hex grid),
    // each boundary ghost with the cell it is adjacent to (its Lb entry)
    std::vector<int32_t> owner((size_t)NA, -1);
    for (int r = 0; r < np; ++r)  // rank r: cells [r N / np, (r+1) N / np)
        for (int64_t i = (int64_t)r * N / np; i < (int64_t)(r + 1) * N / np; ++i) owner[(size_t)i] = r;
    for (int64_t g = N; g < NA; ++g) {
        const int32_t b = c->Lb.rowptr[g], e = c->Lb.rowptr[g + 1];
        if (e - b != 1 || c->Lb.col[b] < 0 || c->Lb.col[b] >= N)
            return fail(GIP_CAV_ETOPO, "gip_cavity_split: a ghost row of Lb does not name one adjacent cell");
        owner[(size_t)g] = owner[(size_t)c->Lb.col[b]];
    }
    // faces of each rank: the faces of its cells (Mres); interface faces on both sides
    std::vector<std::vector<int32_t>> faces((size_t)np);
    for (int64_t i = 0; i < N; ++i)
        for (int32_t t = c->Mres.rowptr[i]; t < c->Mres.rowptr[i + 1]; ++t) {
            const int32_t f = c->Mres.col[t];
            if (f < 0 || f >= F) return fail(GIP_CAV_ETOPO, "gip_cavity_split: Mres column out of range");
            faces[(size_t)owner[(size_t)i]].push_back(f);
        }
    for (auto &v : faces) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    // cells each rank reads but does not own (its halo), and the cells that
    // border another rank (reading or read by it): the interface cells
    std::vector<std::vector<int32_t>> halo((size_t)np);
    std::vector<char> iface((size_t)N, 0);
    for (int r = 0; r < np; ++r) {
        std::vector<int32_t> &h = halo[(size_t)r];
        auto touch = [&](int32_t row_cell, int32_t col) -> int {
            if (col < 0 || col >= NA) return fail(GIP_CAV_ETOPO, "gip_cavity_split: cell column out of range");
            if (owner[(size_t)col] == r) return GIP_CAV_OK;
            if (col >= N) return fail(GIP_CAV_ETOPO, "gip_cavity_split: a partition would read another's ghost");
            h.push_back(col);
            iface[(size_t)col] = 1;
            if (row_cell >= 0) iface[(size_t)row_cell] = 1;
            return GIP_CAV_OK;
        };
        const int64_t lo = (int64_t)r * N / np, hi = (int64_t)(r + 1) * N / np;
        for (const gip_csr *A : cellops)
            for (int64_t i = lo; i < hi; ++i)
                for (int32_t t = A->rowptr[i]; t < A->rowptr[i + 1]; ++t)
                    if (int rc = touch((int32_t)i, A->col[t])) return rc;
        for (const gip_csr *A : faceops)
            for (int32_t f : faces[(size_t)r]) {
                int32_t mine = -1;  // the face's cell in this rank marks the interface
                for (int32_t t = A->rowptr[f]; t < A->rowptr[f + 1]; ++t)
                    if (A->col[t] >= 0 && A->col[t] < N && owner[(size_t)A->col[t]] == r) mine = A->col[t];
                for (int32_t t = A->rowptr[f]; t < A->rowptr[f + 1]; ++t)
                    if (int rc = touch(mine, A->col[t])) return rc;
            }
        std::sort(h.begin(), h.end(), [&](int32_t a, int32_t b) {
            return owner[(size_t)a] != owner[(size_t)b] ? owner[(size_t)a] < owner[(size_t)b] : a < b;
        });
        h.erase(std::unique(h.begin(), h.end()), h.end());
    }
    CasePaths p = make_paths(root, name, np);
    const char *groups[] = {"CD", "L", "M", "G", "S"};
    for (const char *g : groups)
        if (int rc = mkdirs(p.dir("Operators", g))) return rc;
    if (int rc = mkdirs(p.dir("Scalarfields", ""))) return rc;
    if (int rc = mkdirs(p.dir("Topos", ""))) return rc;
    // local numbering of every rank: [interiors | interfaces | ghosts | halo by source rank]
    std::vector<std::vector<int32_t>> cells((size_t)np);  // local -> global
    std::vector<int32_t> ninter((size_t)np), nifc((size_t)np), nghost((size_t)np);
    for (int r = 0; r < np; ++r) {
        const int64_t lo = (int64_t)r * N / np, hi = (int64_t)(r + 1) * N / np;
        std::vector<int32_t> &v = cells[(size_t)r];
        for (int64_t i = lo; i < hi; ++i)
            if (!iface[(size_t)i]) v.push_back((int32_t)i);
        ninter[(size_t)r] = (int32_t)v.size();
        for (int64_t i = lo; i < hi; ++i)
            if (iface[(size_t)i]) v.push_back((int32_t)i);
        nifc[(size_t)r] = (int32_t)(v.size() - ninter[(size_t)r]);
        for (int64_t g = N; g < NA; ++g)
            if (owner[(size_t)g] == r) v.push_back((int32_t)g);
        nghost[(size_t)r] = (int32_t)(v.size() - ninter[(size_t)r] - nifc[(size_t)r]);
        v.insert(v.end(), halo[(size_t)r].begin(), halo[(size_t)r].end());
    }
    std::vector<int32_t> loc((size_t)NA), floc((size_t)F);
    for (int r = 0; r < np; ++r) {
        const std::vector<int32_t> &cl = cells[(size_t)r];
        const std::vector<int32_t> &fl = faces[(size_t)r];
        const int32_t inner = ninter[(size_t)r] + nifc[(size_t)r], owned = inner + nghost[(size_t)r];
        const int32_t alln = (int32_t)cl.size(), nf = (int32_t)fl.size();
        std::fill(loc.begin(), loc.end(), -1);
        std::fill(floc.begin(), floc.end(), -1);
        for (int32_t l = 0; l < alln; ++l) loc[(size_t)cl[(size_t)l]] = l;
        for (int32_t l = 0; l < nf; ++l) floc[(size_t)fl[(size_t)l]] = l;
        // the rows of an operator in this rank, in its order, columns mapped
        // (stored order kept, so every row sums its terms as the whole case does)
        auto emit = [&](const char *group, const char *op, const gip_csr &A, const std::vector<int32_t> &rows,
                        bool face_cols, int32_t ncols) -> int {
            std::vector<int32_t> len(rows.size());
            for (size_t e = 0; e < rows.size(); ++e) len[e] = A.rowptr[rows[e] + 1] - A.rowptr[rows[e]];
            gip_csr B;
            csr_zero(&B);
            int rc = csr_alloc(&B, (int32_t)rows.size(), ncols, len);
            for (size_t e = 0; e < rows.size() && !rc; ++e) {
                int32_t q = B.rowptr[e];
                for (int32_t t = A.rowptr[rows[e]]; t < A.rowptr[rows[e] + 1]; ++t, ++q) {
                    const int32_t m = face_cols ? floc[(size_t)A.col[t]] : loc[(size_t)A.col[t]];
                    if (m < 0) {
                        rc = fail(GIP_CAV_ETOPO, std::string("gip_cavity_split: ") + op + " reads outside its partition");
                        break;
                    }
                    B.col[q] = m;
                    B.val[q] = A.val[t];
                }
            }
            if (!rc) rc = gip_csr_write(p.op(group, op, r).c_str(), &B);
            gip_csr_release(&B);
            return rc;
        };
        std::vector<int32_t> inner_rows(cl.begin(), cl.begin() + inner), owned_rows(cl.begin(), cl.begin() + owned);
        // Ec: one row per entry of this rank's D rows, in their order
        std::vector<int32_t> ec_rows;
        for (int32_t i : inner_rows)
            for (int32_t t = c->D.rowptr[i]; t < c->D.rowptr[i + 1]; ++t) ec_rows.push_back(t);
        int rc = emit("CD", "D", c->D, inner_rows, false, owned);
        if (!rc) rc = emit("CD", "Ec", c->Ec, ec_rows, true, nf);
        if (!rc) rc = emit("L", "L", c->L, inner_rows, false, owned);
        if (!rc) rc = emit("L", "Lb", c->Lb, owned_rows, false, owned);
        if (!rc) rc = emit("M", "Mx", c->Mx, inner_rows, false, inner);
        if (!rc) rc = emit("M", "My", c->My, inner_rows, false, inner);
        if (!rc) rc = emit("M", "Mz", c->Mz, inner_rows, false, inner);
        if (!rc) rc = emit("M", "Mres", c->Mres, inner_rows, true, nf);
        if (!rc) rc = emit("G", "Gx", c->Gx, owned_rows, false, owned);
        if (!rc) rc = emit("G", "Gy", c->Gy, owned_rows, false, owned);
        if (!rc) rc = emit("G", "Gz", c->Gz, owned_rows, false, owned);
        if (!rc) rc = emit("S", "SFx", c->SFx, fl, false, owned);
        if (!rc) rc = emit("S", "SFy", c->SFy, fl, false, owned);
        if (!rc) rc = emit("S", "SFz", c->SFz, fl, false, owned);
        if (!rc) rc = emit("S", "SO", c->SO, fl, false, owned);
        if (rc) return rc;
        const char *fields[] = {"u", "v", "w", "dxs"};
        const double *fv[] = {c->u, c->v, c->w, c->dxs};
        std::vector<double> buf((size_t)alln);
        for (int f = 0; f < 4 && !rc; ++f) {
            for (int32_t l = 0; l < alln; ++l) buf[(size_t)l] = fv[f][cl[(size_t)l]];
            rc = gip_vec_write(p.vec(fields[f], r).c_str(), buf.data(), alln);
        }
        if (rc) return rc;
        // halo lists: this rank's slots grouped by source rank; every other
        // rank's send block to this rank lists the same cells in the same order
        std::vector<int32_t> send_indx, send_row(1, 0), recv_indx, recv_row(1, 0);
        for (int q = 0; q < np; ++q) {
            for (int32_t l = owned; l < alln; ++l)
                if (owner[(size_t)cl[(size_t)l]] == q) recv_indx.push_back(l);
            recv_row.push_back((int32_t)recv_indx.size());
            for (int32_t g : halo[(size_t)q])
                if (owner[(size_t)g] == r) send_indx.push_back(loc[(size_t)g]);
            send_row.push_back((int32_t)send_indx.size());
        }
        for (int32_t l : send_indx)
            if (l < ninter[(size_t)r] || l >= inner)
                return fail(GIP_CAV_ETOPO, "gip_cavity_split: a sent cell is not an interface cell");
        const int sizes[9] = {(int)send_indx.size(), (int)recv_indx.size(), np + 1, owned, alln, ninter[(size_t)r],
                              nifc[(size_t)r], nghost[(size_t)r], alln - owned};
        rc = write_topo_lists(p.topo("topoCols", r), sizes, send_indx, send_row, recv_indx, recv_row);
        // the faces topology: only its `all` (the face count) is read (GIP_DNS.h:271)
        const int fsizes[9] = {0, 0, np + 1, nf, nf, ninter[(size_t)r], nifc[(size_t)r], nghost[(size_t)r], alln - owned};
        const std::vector<int32_t> none, zeros((size_t)np + 1, 0);
        if (!rc) rc = write_topo_lists(p.topo("topoFaces", r), fsizes, none, zeros, none, zeros);
        if (rc) return rc;
    }
    return GIP_CAV_OK;
}

}  // extern "C"
