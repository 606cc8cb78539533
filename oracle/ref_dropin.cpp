// TEST INFRASTRUCTURE ONLY — drop-in check of the HIP specification.
//
// The reference's DrivenCavity.cpp:1-43 with a third backend branch: the
// reference's own, unmodified templates (src/Algorithms/GIP_DNS.h,
// src/Algorithms/GIP_LinearSolver.h) and abstract/basic interfaces are
// instantiated with this repository's HIP specification classes
// (gip_amd/spec/Specifications/HIP), compiled against the reference's
// abstract class headers.  Built by oracle/build_ref.sh into
// oracle/_ref/DrivenCavityHIP_reftemplates; tests compare its output with
// the oracle.  Usage is the reference's: run from a directory holding
// InputFiles/, `DrivenCavityHIP_reftemplates <name>` (one rank).
#include <stdio.h>

#include <iostream>

#include "src/Algorithms/GIP_DNS.h"
#include "HIP.h"

typedef GIP_MatrixHIP Matrix;
typedef GIP_VectorHIP Vector;
typedef GIP_NonLinearHIP NonLinear;
typedef GIP_ArchHIP Architecture;

int main(int argc, char **argv) {
    if (argc - 1 != 1) {
        std::cout << "a.out <Mesh_file>  " << std::endl;
        return 0;
    }
    int prov;
    MPI_Init_thread(0, 0, MPI_THREAD_SINGLE, &prov);
    {
        GIP_DNS<Matrix, Vector, Architecture, NonLinear> driven(argv[1]);
        driven.setUp();
        driven.integrate();
    }
    MPI_Finalize();
    return 0;
}
