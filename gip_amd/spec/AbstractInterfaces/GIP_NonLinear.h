// Non-linear operations abstraction (interface of GIP_NonLinear.h:6-14 in the reference).
#ifndef _gip_nlinear_
#define _gip_nlinear_

#include "GIP_Vector.h"

class GIP_NonLinear {
   public:
    GIP_NonLinear() {}
    virtual ~GIP_NonLinear() {}
    // self-adaptive time step (convective and diffusive limits) over the inner cells
    virtual double getCFL(GIP_Topo *, GIP_Vector *, GIP_Vector *, GIP_Vector *, GIP_Vector *, GIP_Vector *, double,
                          double) {
        return 0;
    }
};

#endif
