// Out-of-line members of the abstract classes.
#include <cmath>
#include <iostream>

#include "GIP_Matrix.h"
#include "GIP_Vector.h"

GIP_Vector::GIP_Vector(int n, GIP_Arch *arch) : vec(new double[n > 0 ? n : 1]()), size(n), myArch(arch) {}

GIP_Vector::GIP_Vector(GIP_Topo *topo, GIP_Arch *arch)
    : vec(new double[topo->getAllSize() > 0 ? topo->getAllSize() : 1]()), size(topo->getAllSize()), myTopo(topo),
      myArch(arch) {}

GIP_Vector::~GIP_Vector() { delete[] vec; }

void GIP_Vector::operator=(double value) {
    for (int i = 0; i < size; i++) vec[i] = value;
}

void GIP_Vector::FillRandom() {
    for (int i = 0; i < size; i++) vec[i] = std::cos(1.0 * i) + std::cos(10.0 * i) + std::cos(100.0 * i) + std::cos(1000.0 * i);
}

GIP_Matrix::~GIP_Matrix() {
    delete[] csrValA;
    delete[] csrColIndA;
    delete[] csrRowIndA;
}

void GIP_Matrix::spmv(GIP_Vector *x, GIP_Vector *b) {
    // host CSR product over the host arrays; empty rows untouched
    double *in = x->getHostPtr(), *out = b->getHostPtr();
    for (int i = 0; i < num_rows; i++) {
        double sum = 0;
        for (int j = csrRowIndA[i]; j < csrRowIndA[i + 1]; j++) sum += in[csrColIndA[j]] * csrValA[j];
        if (csrRowIndA[i] != csrRowIndA[i + 1]) out[i] = sum;
    }
}

void GIP_Matrix::printRows(int n) {
    std::cout << "Matrix Info:" << std::endl
              << " ROWS : " << num_rows << " COLS : " << num_cols << " NNZ : " << nnz << std::endl;
    for (int i = 0; i < n && i < num_rows; i++) {
        std::cout << " " << i << " :";
        for (int j = csrRowIndA[i]; j < csrRowIndA[i + 1]; j++) std::cout << " " << csrValA[j];
        std::cout << std::endl;
    }
}
