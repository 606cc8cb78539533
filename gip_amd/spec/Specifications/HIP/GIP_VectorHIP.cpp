#include "GIP_VectorHIP.h"

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>

#include "gip_cavity.h"

GIP_VectorHIP::GIP_VectorHIP(int n, GIP_Arch *arch) : GIP_Vector(n, arch) { allocDevice(); }

GIP_VectorHIP::GIP_VectorHIP(GIP_Topo *topo, GIP_Arch *arch) : GIP_Vector(topo, arch) { allocDevice(); }

GIP_VectorHIP::~GIP_VectorHIP() {
    if (dvec) GIP_ArchHIP::of(myArch)->release(dvec, sizeof(double) * (size_t)(size > 0 ? size : 1));
}

void GIP_VectorHIP::allocDevice() {
    dvec = static_cast<double *>(GIP_ArchHIP::of(myArch)->acquire(sizeof(double) * (size_t)(size > 0 ? size : 1)));
    gip_check(gip_dfill(ctx(), size, dvec, 0.0), "vector zero");
}

void GIP_VectorHIP::postConstruct(GIP_Topo *topo, GIP_Arch *arch) {
    size = topo->getAllSize();
    vec = new double[size > 0 ? size : 1]();
    myTopo = topo;
    myArch = arch;
    allocDevice();
}

// .vec format of GIP_VectorMulticore.cpp:16-42
void GIP_VectorHIP::postConstruct(std::string name, GIP_Topo *topo, GIP_Arch *arch) {
    postConstruct(topo, arch);
    double *vals = nullptr;
    int32_t n = 0;
    if (gip_vec_read(name.c_str(), &vals, &n) != 0)
        throw std::runtime_error("error opening scalarfield file " + name + ": " + gip_cavity_last_error());
    if (n > size) {
        std::free(vals);
        throw std::runtime_error("scalarfield " + name + " longer than the topology");
    }
    for (int i = 0; i < n; ++i) vec[i] = vals[i];
    std::free(vals);
    uploadHost();
}

void GIP_VectorHIP::postConstruct(const double *values, int n, GIP_Topo *topo, GIP_Arch *arch) {
    postConstruct(topo, arch);
    for (int i = 0; i < n && i < size; ++i) vec[i] = values[i];
    uploadHost();
}

int GIP_VectorHIP::range(enum RUN_DOMAIN d) {
    int n = 0;
    switch (d) {
        case _ALL_: n = size; break;
        case _OWNED_: n = myTopo->getOwnedSize(); break;
        case _INNER_: n = myTopo->getInnerSize(); break;
        default: n = 0;
    }
    if (n > size) {
        std::fprintf(stderr, "gip_hip: range of %d entries on a vector of %d\n", n, size);
        std::abort();
    }
    return n;
}

static int covered(GIP_Vector *x, int n) {
    // the other operand must cover the range too (host-side bounds check)
    if (x->getSize() < n) {
        std::fprintf(stderr, "gip_hip: vector of %d entries used over %d\n", x->getSize(), n);
        std::abort();
    }
    return n;
}

void GIP_VectorHIP::axpy(GIP_Vector *x, double alpha, enum RUN_DOMAIN rdom) {
    gip_check(gip_daxpy(ctx(), covered(x, range(rdom)), alpha, x->getDevicePtr(), dvec), "axpy");
}

void GIP_VectorHIP::axpy(GIP_Vector *x, double alpha, double beta, enum RUN_DOMAIN rdom) {
    gip_check(gip_daxpby(ctx(), covered(x, range(rdom)), alpha, x->getDevicePtr(), beta, dvec), "axpby");
}

// the partition's sum, combined over the partitions (MPI_Allreduce SUM in
// GIP_VectorMulticore.cpp:125)
double GIP_VectorHIP::dot(GIP_Vector *x, GIP_Vector *, enum RUN_DOMAIN rdom) {
    double r = 0.0;
    gip_check(gip_ddot(ctx(), covered(x, range(rdom)), dvec, x->getDevicePtr(), &r), "dot");
    return GIP_ArchHIP::of(myArch)->allreduce(r, GIP_RankComm::SUM);
}

// (MPI_Allreduce MAX in GIP_VectorMulticore.cpp:161)
double GIP_VectorHIP::norm(GIP_Vector *, enum RUN_DOMAIN rdom) {
    double r = 0.0;
    gip_check(gip_dnrm_inf(ctx(), range(rdom), dvec, &r), "norm");
    return GIP_ArchHIP::of(myArch)->allreduce(r, GIP_RankComm::MAX);
}

void GIP_VectorHIP::copyTo(GIP_Vector *x, enum RUN_DOMAIN rdom) {
    gip_check(gip_memcpy_d2d(ctx(), x->getDevicePtr(), dvec, sizeof(double) * (size_t)covered(x, range(rdom))),
              "copyTo");
}

// the halo refresh of GIP_VectorMulticore.cpp:197-216 (GIP_Topo.cpp:175-181):
// device pack and unpack kernels, among the partitions of one device or
// staged through the host between processes; a lone partition has no halo
void GIP_VectorHIP::update() { GIP_ArchHIP::of(myArch)->exchange(myTopo, dvec); }

void GIP_VectorHIP::operator=(int value) { gip_check(gip_dfill(ctx(), size, dvec, (double)value), "fill"); }

void GIP_VectorHIP::operator=(double value) { gip_check(gip_dfill(ctx(), size, dvec, value), "fill"); }

double *GIP_VectorHIP::getHostPtr() {
    gip_check(gip_memcpy_d2h(ctx(), vec, dvec, sizeof(double) * (size_t)size), "download");
    return vec;
}

void GIP_VectorHIP::uploadHost() {
    gip_check(gip_memcpy_h2d(ctx(), dvec, vec, sizeof(double) * (size_t)size), "upload");
}
