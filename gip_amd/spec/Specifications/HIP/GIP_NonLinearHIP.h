// HIP specification of the non-linear operations: the self-adaptive CFL
// time step of GIP_NonLinearMulticore.cpp:3-55 as one fused min-reduction
// kernel over the inner cells.
#ifndef _gip_nlinearhip_
#define _gip_nlinearhip_

#include "../../AbstractInterfaces/GIP_NonLinear.h"
#include "GIP_VectorHIP.h"

class GIP_NonLinearHIP : public GIP_NonLinear {
   public:
    GIP_NonLinearHIP() {}
    ~GIP_NonLinearHIP() {}
    double getCFL(GIP_Topo *topo, GIP_Vector *u, GIP_Vector *v, GIP_Vector *w, GIP_Vector *dxs, GIP_Vector *temp,
                  double gamma, double rho);
};

#endif
