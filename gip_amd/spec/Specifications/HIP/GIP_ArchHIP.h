// HIP node specification: one partition on one MI355X device.
// Replaces GIP_ArchGPU (GIP_ArchGPU.h, GIP_ArchGPU.cpp:19-26): setUp()
// selects the device and creates the gip_hip context (stream + reduction
// scratch) every vector, operator and solver of this node launches on.
// It also keeps a pool of device buffers so that the solver temporaries
// (GIP_LinearSolver.h:54-61) do not allocate on every solve.
//
// A node is the only partition, one of a group of partitions stepped by
// threads of this process (GIP_PartitionHIP.h), or one of several processes
// (MPI ranks, GIP_RanksMPI.cpp).  allreduce() and exchange() are the
// collectives the vectors and the CFL bound make across the partitions in
// either case, and the identity for a lone partition.
#ifndef _gip_archhip_
#define _gip_archhip_

#include <cstddef>
#include <map>
#include <vector>

#include "../../AbstractInterfaces/GIP_Arch.h"
#include "gip_hip.h"

class GIP_Topo;

// Collectives across processes, one partition per process.
class GIP_RankComm {
   public:
    enum { SUM = 0, MAX = 1, MIN = 2 };
    virtual ~GIP_RankComm() {}
    virtual int rank() const = 0;
    virtual int size() const = 0;
    virtual int localRank() const = 0;  // among the processes of this host
    virtual double allreduce(double v, int op) = 0;
    // refresh the halo of a device vector laid out by `topo`
    virtual void exchange(GIP_Topo *topo, gip_ctx *ctx, double *dvec) = 0;
};

// Set, before main, by a translation unit that links MPI (GIP_RanksMPI.cpp,
// compiled in the MPI build of this specification; the library build leaves
// it null).  Returns the communicator of this process when more than one
// process runs, else nullptr.
extern GIP_RankComm *(*gip_rank_comm_factory)();

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
    GIP_RankComm *ranks() { return ranks_; }
    int partRank() const;
    int partCount() const;

    // the partitions' combined value (op: GIP_RankComm::SUM, MAX or MIN)
    double allreduce(double v, int op);
    // refresh the halo of a device vector of this node laid out by `topo`
    void exchange(GIP_Topo *topo, double *dvec);

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
    GIP_RankComm *ranks_ = nullptr;
    std::multimap<size_t, void *> pool_;
};

// Abort with the C ABI's last error message when a call fails.
void gip_check(int rc, const char *what);

#endif
