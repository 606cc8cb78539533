// Several partitions of one case resident on one device, each stepped by
// its own host thread with its own node (context and stream): the
// in-process counterpart of the reference's MPI ranks (one rank per
// partition, DrivenCavity.cpp:30-33).  The group provides what the ranks
// exchange through MPI in the reference:
//   - the halo refresh of a vector (GIP_Topo.cpp:101-181, host-staged in
//     GIP_VectorCuda.cpp:224-247): every partition packs its send list into
//     its block of one device buffer, then every partition gathers its halo
//     slots from the senders' blocks, both as device kernels
//     (gip_dcopy_indexed), with a rendezvous of the threads after each;
//   - the global sums, maxima and minima of dot products, norms and the CFL
//     bound (MPI_Allreduce, GIP_VectorMulticore.cpp:125,161,
//     GIP_NonLinearMulticore.cpp:45), combined in rank order, so every
//     partition gets the same value.
// Every partition must make the same sequence of collective calls (as MPI
// ranks do).  A partition that fails calls abort(), which releases the
// others from any rendezvous with an exception instead of a deadlock.  The
// rendezvous spins on an atomic generation counter (yielding after a few
// thousand polls): a wait on a condition variable costs tens of
// microseconds per wake-up, and the reference sequence makes about three
// hundred rendezvous per step.
#ifndef _gip_partitionhip_
#define _gip_partitionhip_

#include <atomic>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "gip_hip.h"

class GIP_Topo;

class GIP_PartitionGroup {
   public:
    enum { SUM = 0, MAX = 1, MIN = 2 };
    GIP_PartitionGroup(int nparts, int device);
    ~GIP_PartitionGroup();
    GIP_PartitionGroup(const GIP_PartitionGroup &) = delete;
    GIP_PartitionGroup &operator=(const GIP_PartitionGroup &) = delete;

    int size() const { return n_; }
    void barrier();
    double allreduce(int rank, double v, int op);
    // refresh the halo of a device vector laid out by `topo` (the cell
    // topology of `rank`; the first call of every rank registers it)
    void exchange(int rank, GIP_Topo *topo, gip_ctx *ctx, double *dvec);
    void abort();

   private:
    void registerTopo(int rank, GIP_Topo *topo, gip_ctx *ctx);
    int n_, device_;
    gip_ctx *ctx_ = nullptr;  // owns the shared send buffer
    std::atomic<int> arrived_{0};
    std::atomic<unsigned> generation_{0};
    std::atomic<bool> aborted_{false};
    // allreduce operands, two sets used alternately: a set is rewritten only
    // two rendezvous after it was read, so one rendezvous per allreduce
    std::vector<double> vals_[2];
    std::vector<unsigned> parity_;  // per rank: the set its next allreduce writes
    // halo exchange: the senders' blocks of one device buffer
    std::vector<GIP_Topo *> topo_;
    std::vector<int64_t> send_off_;
    double *d_send_ = nullptr;
    struct Maps {
        int32_t *send_indx = nullptr, *recv_indx = nullptr, *src = nullptr;  // device
        int nsend = 0, nrecv = 0;
        gip_ctx *ctx = nullptr;
    };
    std::vector<Maps> maps_;
};

#endif
