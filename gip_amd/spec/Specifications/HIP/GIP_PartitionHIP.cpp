#include "GIP_PartitionHIP.h"

#include <string>
#include <thread>

#include "GIP_ArchHIP.h"
#include "../../BasicInterfaces/GIP_Topo.h"

GIP_PartitionGroup::GIP_PartitionGroup(int nparts, int device)
    : n_(nparts), device_(device), topo_((size_t)nparts, nullptr), send_off_((size_t)nparts + 1, 0),
      maps_((size_t)nparts) {
    if (nparts < 1) throw std::runtime_error("GIP_PartitionGroup: at least one partition");
    vals_[0].assign((size_t)nparts, 0.0);
    vals_[1].assign((size_t)nparts, 0.0);
    parity_.assign((size_t)nparts, 0u);
    gip_check(gip_init(device, &ctx_), "gip_init (partition group)");
}

GIP_PartitionGroup::~GIP_PartitionGroup() {
    for (Maps &m : maps_) {
        if (m.send_indx) gip_free(ctx_, m.send_indx);
        if (m.recv_indx) gip_free(ctx_, m.recv_indx);
        if (m.src) gip_free(ctx_, m.src);
    }
    if (d_send_) gip_free(ctx_, d_send_);
    gip_finalize(ctx_);
}

void GIP_PartitionGroup::abort() { aborted_.store(true, std::memory_order_release); }

// Central counter with a generation number: the last partition to arrive
// resets the counter and then advances the generation (release), which the
// others wait for (acquire), so everything written before the rendezvous is
// visible after it.
void GIP_PartitionGroup::barrier() {
    if (aborted_.load(std::memory_order_acquire)) throw std::runtime_error("another partition failed");
    if (n_ == 1) return;
    const unsigned gen = generation_.load(std::memory_order_acquire);
    if (arrived_.fetch_add(1, std::memory_order_acq_rel) + 1 == n_) {
        arrived_.store(0, std::memory_order_relaxed);
        generation_.store(gen + 1, std::memory_order_release);
        return;
    }
    for (unsigned polls = 0; generation_.load(std::memory_order_acquire) == gen; ++polls) {
        if (aborted_.load(std::memory_order_acquire)) throw std::runtime_error("another partition failed");
        if (polls > 4096) std::this_thread::yield();
    }
}

double GIP_PartitionGroup::allreduce(int rank, double v, int op) {
    if (n_ == 1) return v;
    std::vector<double> &vals = vals_[parity_[(size_t)rank]];
    parity_[(size_t)rank] ^= 1u;
    vals[(size_t)rank] = v;
    barrier();
    double r = vals[0];
    for (int q = 1; q < n_; ++q) {
        const double x = vals[(size_t)q];
        if (op == SUM)
            r = r + x;
        else if (op == MAX)
            r = x > r ? x : r;  // a NaN never replaces the running value
        else
            r = x < r ? x : r;
    }
    return r;
}

static int32_t *upload_ints(gip_ctx *c, const int *v, size_t n) {
    void *p = nullptr;
    gip_check(gip_malloc(c, sizeof(int32_t) * (n ? n : 1), &p), "gip_malloc (halo map)");
    if (n) gip_check(gip_memcpy_h2d(c, p, v, sizeof(int32_t) * n), "halo map upload");
    return static_cast<int32_t *>(p);
}

// Pairs every halo slot of every partition with its sender's entry: block q
// of rank r's receive list and block r of rank q's send list name the same
// cells in the same order (GIP_Topo.cpp:124-181).
void GIP_PartitionGroup::registerTopo(int rank, GIP_Topo *topo, gip_ctx *ctx) {
    topo_[(size_t)rank] = topo;
    barrier();
    if (rank == 0) {
        for (int q = 0; q < n_; ++q) send_off_[(size_t)q + 1] = send_off_[(size_t)q] + topo_[(size_t)q]->getSendSize();
        void *p = nullptr;
        gip_check(gip_malloc(ctx_, sizeof(double) * (size_t)(send_off_[(size_t)n_] ? send_off_[(size_t)n_] : 1), &p),
                  "gip_malloc (halo send buffer)");
        d_send_ = static_cast<double *>(p);
    }
    barrier();
    GIP_Topo *t = topo;
    const int *rrow = t->getRecvRow();
    std::vector<int> src((size_t)t->getRecvSize());
    for (int q = 0; q < n_; ++q) {
        const int *srow = topo_[(size_t)q]->getSendRow();
        const int nrecv = rrow[q + 1] - rrow[q], nsend = srow[rank + 1] - srow[rank];
        if (nrecv != nsend) {
            abort();
            throw std::runtime_error("partition " + std::to_string(rank) + ": halo block from partition " +
                                     std::to_string(q) + " has " + std::to_string(nrecv) + " slots, its sender lists " +
                                     std::to_string(nsend));
        }
        for (int k = 0; k < nrecv; ++k) src[(size_t)(rrow[q] + k)] = (int)(send_off_[(size_t)q] + srow[rank] + k);
    }
    Maps &m = maps_[(size_t)rank];
    m.ctx = ctx;
    m.nsend = t->getSendSize();
    m.nrecv = t->getRecvSize();
    m.send_indx = upload_ints(ctx_, t->getSendIndx(), (size_t)m.nsend);
    m.recv_indx = upload_ints(ctx_, t->getRecvIndx(), (size_t)m.nrecv);
    m.src = upload_ints(ctx_, src.data(), src.size());
    gip_check(gip_sync(ctx_), "halo maps");
}

void GIP_PartitionGroup::exchange(int rank, GIP_Topo *topo, gip_ctx *ctx, double *dvec) {
    if (n_ == 1) return;
    if (!maps_[(size_t)rank].ctx) registerTopo(rank, topo, ctx);  // every partition's first exchange
    if (topo != topo_[(size_t)rank]) throw std::runtime_error("halo exchange of a vector on another topology");
    const Maps &m = maps_[(size_t)rank];
    // pack this partition's send list into its block
    gip_check(gip_dcopy_indexed(ctx, m.nsend, m.send_indx, dvec, nullptr, d_send_ + send_off_[(size_t)rank]),
              "halo pack");
    gip_check(gip_sync(ctx), "halo pack");
    barrier();
    // gather the halo slots from the senders' blocks
    gip_check(gip_dcopy_indexed(ctx, m.nrecv, m.src, d_send_, m.recv_indx, dvec), "halo unpack");
    gip_check(gip_sync(ctx), "halo unpack");
    barrier();  // nobody repacks before every partition has gathered
}
