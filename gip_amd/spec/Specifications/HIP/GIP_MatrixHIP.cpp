#include "GIP_MatrixHIP.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>

GIP_MatrixHIP::~GIP_MatrixHIP() {
    if (pcg) gip_pcg_destroy(pcg);
    if (dmat) gip_mat_destroy(dmat);
    if (sharedHost) {  // the caller owns the host arrays
        csrValA = nullptr;
        csrColIndA = nullptr;
        csrRowIndA = nullptr;
    }
}

void GIP_MatrixHIP::postConstructShared(const gip_csr &A, GIP_Topo *topo, GIP_Arch *arch) {
    nnz = A.nnz;
    num_cols = A.ncols;
    num_rows = A.nrows;
    csrValA = A.val;
    csrColIndA = A.col;
    csrRowIndA = A.rowptr;
    sharedHost = true;
    myTopo = topo;
    myArch = arch;
    createDevice(GIP_FMT_AUTO);
}

void GIP_MatrixHIP::createDevice(int flags) {
    if (dmat) gip_mat_destroy(dmat);
    dmat = nullptr;
    devFlags = flags;
    gip_check(gip_mat_create(ctx(), num_rows, num_cols, nnz, csrRowIndA, csrColIndA, csrValA, flags, &dmat),
              "gip_mat_create");
    gip_check(gip_mat_cols_needed(dmat, &colsNeeded), "gip_mat_cols_needed");
}

void GIP_MatrixHIP::checkShapes(GIP_Vector *x, int outLen) {
    // the kernels trust the vector lengths: check them on the host first
    if (x->getSize() < colsNeeded || outLen < num_rows) {
        std::fprintf(stderr, "gip_hip: operator (%d rows, columns up to %d) applied to vectors of %d -> %d entries\n",
                     num_rows, colsNeeded - 1, x->getSize(), outLen);
        std::abort();
    }
}

void GIP_MatrixHIP::postConstruct(std::string name, GIP_Topo *topo, GIP_Arch *arch) {
    gip_csr A;
    if (gip_csr_read(name.c_str(), &A) != 0)
        throw std::runtime_error("error opening operator file " + name + ": " + gip_cavity_last_error());
    postConstruct(A, topo, arch);
    gip_csr_release(&A);
}

void GIP_MatrixHIP::postConstruct(const gip_csr &A, GIP_Topo *topo, GIP_Arch *arch) {
    nnz = A.nnz;
    num_cols = A.ncols;
    num_rows = A.nrows;
    csrValA = new double[nnz > 0 ? nnz : 1];
    csrColIndA = new int[nnz > 0 ? nnz : 1];
    csrRowIndA = new int[num_rows + 1];
    if (nnz) {
        std::memcpy(csrValA, A.val, sizeof(double) * (size_t)nnz);
        std::memcpy(csrColIndA, A.col, sizeof(int) * (size_t)nnz);
    }
    std::memcpy(csrRowIndA, A.rowptr, sizeof(int) * (size_t)(num_rows + 1));
    myTopo = topo;
    myArch = arch;
    createDevice(GIP_FMT_AUTO);
}

void GIP_MatrixHIP::spmv(GIP_Vector *x, GIP_Vector *y) {
    checkShapes(x, y->getSize());
    gip_check(gip_dcsrmv(ctx(), dmat, x->getDevicePtr(), y->getDevicePtr()), "spmv");
}

void GIP_MatrixHIP::spmv(GIP_Vector *x, GIP_Vector *y, double alpha, double beta) {
    checkShapes(x, y->getSize());
    gip_check(gip_dcsrmv_ab(ctx(), dmat, x->getDevicePtr(), y->getDevicePtr(), alpha, beta), "spmv_ab");
}

void GIP_MatrixHIP::spmv(GIP_Vector *x, GIP_Matrix *B) {
    GIP_MatrixHIP *Bh = dynamic_cast<GIP_MatrixHIP *>(B);
    if (!Bh) {
        std::fprintf(stderr, "gip_hip: spmv into the values of a non-HIP operator\n");
        std::abort();
    }
    checkShapes(x, Bh->nnz);
    gip_check(gip_dcsrmv_vals(ctx(), dmat, x->getDevicePtr(), Bh->dmat), "spmv_vals");
}

void GIP_MatrixHIP::spmv(GIP_Vector *x, GIP_Matrix *B, double alpha, double beta) {
    // B.values = alpha A x + beta B.values on the device, through B's
    // values in CSR order (GIP_MatrixMulticore.cpp:125-143)
    GIP_MatrixHIP *Bh = dynamic_cast<GIP_MatrixHIP *>(B);
    if (!Bh) {
        std::fprintf(stderr, "gip_hip: spmv into the values of a non-HIP operator\n");
        std::abort();
    }
    checkShapes(x, Bh->nnz);
    gip_check(gip_dcsrmv_vals_ab(ctx(), dmat, x->getDevicePtr(), Bh->dmat, alpha, beta), "spmv_vals_ab");
}

void GIP_MatrixHIP::downloadValues() {
    if (sharedHost) {  // never write into the caller's arrays: take a private copy first
        double *v = new double[nnz > 0 ? nnz : 1];
        int *c = new int[nnz > 0 ? nnz : 1];
        int *r = new int[num_rows + 1];
        std::memcpy(c, csrColIndA, sizeof(int) * (size_t)nnz);
        std::memcpy(r, csrRowIndA, sizeof(int) * (size_t)(num_rows + 1));
        csrValA = v;
        csrColIndA = c;
        csrRowIndA = r;
        sharedHost = false;
    }
    gip_check(gip_mat_get_values(dmat, csrValA), "get values");
}

// GIP_MatrixMulticore.cpp:177-210: the last diagonal entry of each row wins.
void GIP_MatrixHIP::setUpDiagonal(GIP_Matrix *A) {
    num_rows = A->getNumRows();
    num_cols = A->getNumRows();
    nnz = A->getNumRows();
    double *valA = A->getCsrValA();
    int *colA = A->getCsrColIndA();
    int *rowA = A->getCsrRowIndA();
    csrValA = new double[nnz > 0 ? nnz : 1];
    csrColIndA = new int[nnz > 0 ? nnz : 1];
    csrRowIndA = new int[num_rows + 1];
    for (int i = 0; i < num_rows; i++) {
        bool found = false;
        for (int j = rowA[i]; j < rowA[i + 1]; j++)
            if (i == colA[j]) {
                csrValA[i] = 1.0 / valA[j];
                found = true;
            }
        if (!found) {
            std::fprintf(stderr, "gip_hip: setUpDiagonal: row %d has no diagonal entry\n", i);
            std::abort();
        }
        csrColIndA[i] = i;
        csrRowIndA[i] = i;
    }
    csrRowIndA[num_rows] = nnz;
    myTopo = A->getMyTopo();
    myArch = A->getMyArch();
    createDevice(GIP_FMT_AUTO);
}

void GIP_MatrixHIP::printRows(int n) {
    downloadValues();
    GIP_Matrix::printRows(n);
}

int GIP_MatrixHIP::pcgSolve(GIP_Vector *b, GIP_Vector *x, GIP_Vector *x0, const gip_pcg_opts &opts,
                            gip_pcg_stats *stats) {
    if (opts.precision == GIP_PCG_MIXED && !(devFlags & GIP_MAT_FP32)) {
        if (pcg) gip_pcg_destroy(pcg);
        pcg = nullptr;
        createDevice(GIP_FMT_AUTO | GIP_MAT_FP32);
    }
    if (!pcg) gip_check(gip_pcg_create(ctx(), dmat, myTopo->getAllSize(), &pcg), "gip_pcg_create");
    return gip_pcg_solve(pcg, b->getDevicePtr(), x->getDevicePtr(), x0->getDevicePtr(), &opts, stats);
}
