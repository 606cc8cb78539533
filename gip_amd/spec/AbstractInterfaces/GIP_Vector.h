// Vector abstraction (interface of GIP_Vector.h:13-53 in the reference):
// a host array `vec` of `size` values laid out by a GIP_Topo, plus the
// virtual operations a specification provides.
#ifndef _gip_vector_
#define _gip_vector_

#include <string>

#include "../AbstractInterfaces/GIP_Arch.h"
#include "../BasicInterfaces/GIP_Topo.h"

class GIP_Vector {
   public:
    GIP_Vector() {}
    GIP_Vector(int n, GIP_Arch *arch);
    GIP_Vector(GIP_Topo *topo, GIP_Arch *arch);
    virtual ~GIP_Vector();

    virtual void postConstruct(GIP_Topo *, GIP_Arch *) {}
    virtual void postConstruct(std::string name, GIP_Topo *, GIP_Arch *) {}

    virtual void operator=(double);

    virtual double *getDevicePtr() { return nullptr; }
    double *getHostPtr() { return vec; }
    int getSize() { return size; }
    virtual void axpy(GIP_Vector *, double) {}
    virtual void axpy(GIP_Vector *, double, double) {}
    virtual double dot(GIP_Vector *, GIP_Vector *) { return 0; }
    virtual void copyTo(GIP_Vector *) {}
    virtual void update() {}

    virtual void FillRandom();
    virtual void TransferToDevice(enum RUN_DOMAIN) {}
    virtual void TransferToHost(enum RUN_DOMAIN) {}

    GIP_Topo *getMyTopo() { return myTopo; }
    GIP_Arch *getMyArch() { return myArch; }

   protected:
    double *vec = nullptr;
    int size = 0;
    GIP_Topo *myTopo = nullptr;
    GIP_Arch *myArch = nullptr;
};

#endif
