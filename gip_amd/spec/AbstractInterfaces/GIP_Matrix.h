// Sparse-operator abstraction (interface of GIP_Matrix.h:11-51 in the
// reference): host CSR arrays plus the virtual SpMV family a
// specification provides.
#ifndef _gip_matrix_
#define _gip_matrix_

#include <string>

#include "GIP_Arch.h"
#include "GIP_Vector.h"

class GIP_Matrix {
   public:
    GIP_Matrix() {}
    virtual ~GIP_Matrix();
    virtual void postConstruct(std::string name, GIP_Topo *, GIP_Arch *) {}

    virtual void spmv(GIP_Vector *, GIP_Vector *);                   // y = A x
    virtual void spmv(GIP_Vector *, GIP_Vector *, double, double) {} // y = a A x + b y
    virtual void spmv(GIP_Vector *, GIP_Matrix *) {}                 // B.values = A x
    virtual void spmv(GIP_Vector *, GIP_Matrix *, double, double) {} // B.values = a A x + b B.values
    virtual void setUpDiagonal(GIP_Matrix *) {}
    virtual void printRows(int);

    virtual double *getCsrValADevice() { return nullptr; }
    int getNumNnz() { return nnz; }
    int getNumRows() { return num_rows; }
    int getNumCols() { return num_cols; }
    double *getCsrValA() { return csrValA; }
    int *getCsrColIndA() { return csrColIndA; }
    int *getCsrRowIndA() { return csrRowIndA; }
    GIP_Topo *getMyTopo() { return myTopo; }
    GIP_Arch *getMyArch() { return myArch; }

   protected:
    double *csrValA = nullptr;
    int *csrColIndA = nullptr;
    int *csrRowIndA = nullptr;
    int num_rows = 0;
    int num_cols = 0;
    int nnz = 0;
    GIP_Topo *myTopo = nullptr;
    GIP_Arch *myArch = nullptr;
};

#endif
