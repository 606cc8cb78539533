// HIP node specification: one partition on one MI355X device.
// Replaces GIP_ArchGPU (GIP_ArchGPU.h, GIP_ArchGPU.cpp:19-26): setUp()
// selects the device and creates the gip_hip context (stream + reduction
// scratch) every vector, operator and solver of this node launches on.
// It also keeps a pool of device buffers so that the solver temporaries
// (GIP_LinearSolver.h:54-61) do not allocate on every solve.
#ifndef _gip_archhip_
#define _gip_archhip_

#include <cstddef>
#include <map>
#include <vector>

#include "../../AbstractInterfaces/GIP_Arch.h"
#include "gip_hip.h"

class GIP_ArchHIP : public GIP_Arch {
   public:
    GIP_ArchHIP() {}
    GIP_ArchHIP(int cpus, int cores, int devs);
    ~GIP_ArchHIP();
    GIP_ArchHIP(const GIP_ArchHIP &) = delete;
    GIP_ArchHIP &operator=(const GIP_ArchHIP &) = delete;

    void postConstruct(int cpus, int cores, int devs);
    void setUp();

    gip_ctx *ctx();
    int device() const { return devId; }
    void setDevice(int d) { devId = d; }

    // membership in a group of partitions stepped together on one device
    // (GIP_PartitionHIP.h); without one the node is the only partition
    void setPartition(class GIP_PartitionGroup *group, int rank) {
        group_ = group;
        rank_ = rank;
    }
    class GIP_PartitionGroup *partition() { return group_; }
    int partRank() const { return rank_; }
    int partCount() const;

    // device buffer pool
    void *acquire(size_t bytes);
    void release(void *ptr, size_t bytes);

    // the HIP node behind a GIP_Arch* (aborts when it is another kind of node)
    static GIP_ArchHIP *of(GIP_Arch *a);

   protected:
    int devId = 0;
    gip_ctx *ctx_ = nullptr;
    class GIP_PartitionGroup *group_ = nullptr;
    int rank_ = 0;
    std::multimap<size_t, void *> pool_;
};

// Abort with the C ABI's last error message when a call fails.
void gip_check(int rc, const char *what);

#endif
