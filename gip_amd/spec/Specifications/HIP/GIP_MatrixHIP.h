// HIP operator specification: the host CSR arrays of the base class stay
// resident (GIP_LinearSolver::setUp reads them through setUpDiagonal) and a
// device copy in the layout the kernel library picks (sliced ELL, diagonal
// or CSR) serves every product.  Semantics of GIP_MatrixMulticore.cpp:64-210.
#ifndef _gip_matrixhip_
#define _gip_matrixhip_

#include <string>

#include "GIP_ArchHIP.h"
#include "../../AbstractInterfaces/GIP_Matrix.h"
#include "GIP_VectorHIP.h"
#include "gip_cavity.h"

class GIP_MatrixHIP : public GIP_Matrix {
   public:
    GIP_MatrixHIP() {}
    ~GIP_MatrixHIP();
    GIP_MatrixHIP(const GIP_MatrixHIP &) = delete;
    GIP_MatrixHIP &operator=(const GIP_MatrixHIP &) = delete;

    // binary .csr file (GIP_MatrixMulticore.cpp:5-31)
    void postConstruct(std::string name, GIP_Topo *topo, GIP_Arch *arch);
    // in-memory CSR (copied)
    void postConstruct(const gip_csr &A, GIP_Topo *topo, GIP_Arch *arch);
    // in-memory CSR whose host arrays stay owned by the caller and must
    // outlive this operator (no host copy of large generated cases)
    void postConstructShared(const gip_csr &A, GIP_Topo *topo, GIP_Arch *arch);

    void spmv(GIP_Vector *x, GIP_Vector *y);                        // y = A x, empty rows untouched
    void spmv(GIP_Vector *x, GIP_Vector *y, double alpha, double beta);
    void spmv(GIP_Vector *x, GIP_Matrix *B);                        // B's values = A x (rows of A = nnz of B)
    void spmv(GIP_Vector *x, GIP_Matrix *B, double alpha, double beta);
    void setUpDiagonal(GIP_Matrix *A);                              // 1/a_ii as an nrows-nnz operator
    void printRows(int n);

    gip_mat *handle() { return dmat; }
    // Fused device-resident PCG on this operator (see gip_pcg_solve).
    int pcgSolve(GIP_Vector *b, GIP_Vector *x, GIP_Vector *x0, const gip_pcg_opts &opts, gip_pcg_stats *stats);
    // Refresh the host CSR values from the device copy.
    void downloadValues();

   protected:
    void createDevice(int flags);
    void checkShapes(GIP_Vector *x, int outLen);
    int colsNeeded = 0;
    gip_ctx *ctx() { return GIP_ArchHIP::of(myArch)->ctx(); }
    gip_mat *dmat = nullptr;
    gip_pcg *pcg = nullptr;
    int devFlags = 0;
    bool sharedHost = false;
};

#endif
