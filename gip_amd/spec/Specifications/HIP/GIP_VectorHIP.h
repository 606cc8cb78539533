// HIP vector specification: a device-resident field (FP64) over a
// GIP_Topo layout.  Operations follow GIP_VectorMulticore.cpp:44-216 and
// run as gfx950 kernels through the C ABI; dot() and norm() return host
// scalars (the template contract), everything else stays on the device.
// The base class's host array is a mirror used for file input and
// getHostPtr() (which copies the device values into it first).
#ifndef _gip_vectorhip_
#define _gip_vectorhip_

#include <string>

#include "GIP_ArchHIP.h"
#include "../../AbstractInterfaces/GIP_Vector.h"

class GIP_VectorHIP : public GIP_Vector {
   public:
    GIP_VectorHIP() {}
    GIP_VectorHIP(int n, GIP_Arch *arch);
    GIP_VectorHIP(GIP_Topo *topo, GIP_Arch *arch);
    ~GIP_VectorHIP();
    GIP_VectorHIP(const GIP_VectorHIP &) = delete;
    GIP_VectorHIP &operator=(const GIP_VectorHIP &) = delete;

    void postConstruct(GIP_Topo *topo, GIP_Arch *arch);
    void postConstruct(std::string name, GIP_Topo *topo, GIP_Arch *arch);
    // in-memory variant: n values copied from host memory (n <= size, rest zero)
    void postConstruct(const double *values, int n, GIP_Topo *topo, GIP_Arch *arch);

    void axpy(GIP_Vector *x, double alpha, enum RUN_DOMAIN rdom);               // this += alpha x
    void axpy(GIP_Vector *x, double alpha, double beta, enum RUN_DOMAIN rdom);  // this = alpha x + beta this
    double dot(GIP_Vector *x, GIP_Vector *temp, enum RUN_DOMAIN rdom);
    double norm(GIP_Vector *temp, enum RUN_DOMAIN rdom);                        // max |this_i|
    void copyTo(GIP_Vector *x, enum RUN_DOMAIN rdom);                           // x[0:n] = this[0:n]
    void update();
    void operator=(int value);
    void operator=(double value);

    double *getDevicePtr() { return dvec; }
    double *getHostPtr();                 // device -> host mirror, then the mirror
    void uploadHost();                    // host mirror -> device
    void TransferToDevice(enum RUN_DOMAIN) { uploadHost(); }
    void TransferToHost(enum RUN_DOMAIN) { getHostPtr(); }

   protected:
    int range(enum RUN_DOMAIN d);
    void allocDevice();
    gip_ctx *ctx() { return GIP_ArchHIP::of(myArch)->ctx(); }
    double *dvec = nullptr;
};

#endif
