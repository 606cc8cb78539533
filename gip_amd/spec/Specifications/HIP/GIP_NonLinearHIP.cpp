#include "GIP_NonLinearHIP.h"

double GIP_NonLinearHIP::getCFL(GIP_Topo *topo, GIP_Vector *u, GIP_Vector *v, GIP_Vector *w, GIP_Vector *dxs,
                                GIP_Vector *, double gamma, double rho) {
    double dt = 0.0;
    gip_ctx *c = GIP_ArchHIP::of(u->getMyArch())->ctx();
    gip_check(gip_cfl(c, topo->getInnerSize(), u->getDevicePtr(), v->getDevicePtr(), w->getDevicePtr(),
                      dxs->getDevicePtr(), gamma, rho, &dt),
              "gip_cfl");
    return dt;
}
