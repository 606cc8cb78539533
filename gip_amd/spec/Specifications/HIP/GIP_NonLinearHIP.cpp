#include "GIP_NonLinearHIP.h"

#include <cstdio>
#include <cstdlib>


double GIP_NonLinearHIP::getCFL(GIP_Topo *topo, GIP_Vector *u, GIP_Vector *v, GIP_Vector *w, GIP_Vector *dxs,
                                GIP_Vector *, double gamma, double rho) {
    double dt = 0.0;
    gip_ctx *c = GIP_ArchHIP::of(u->getMyArch())->ctx();
    const int n = topo->getInnerSize();
    if (u->getSize() < n || v->getSize() < n || w->getSize() < n || dxs->getSize() < n) {
        std::fprintf(stderr, "gip_hip: getCFL on vectors shorter than the inner range\n");
        std::abort();
    }
    gip_check(gip_cfl(c, topo->getInnerSize(), u->getDevicePtr(), v->getDevicePtr(), w->getDevicePtr(),
                      dxs->getDevicePtr(), gamma, rho, &dt),
              "gip_cfl");
    // the smallest bound over the partitions (MPI_Allreduce MIN,
    // GIP_NonLinearMulticore.cpp:45): the local value is 0.8 min(...) capped
    // at 0.1, and the cap commutes with the minimum
    return GIP_ArchHIP::of(u->getMyArch())->allreduce(dt, GIP_RankComm::MIN);
}
