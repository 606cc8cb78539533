#include "GIP_ArchHIP.h"

#include <cstdio>
#include <cstdlib>

#include "GIP_PartitionHIP.h"

GIP_RankComm *(*gip_rank_comm_factory)() = nullptr;

static_assert(int(GIP_RankComm::SUM) == int(GIP_PartitionGroup::SUM) &&
                  int(GIP_RankComm::MAX) == int(GIP_PartitionGroup::MAX) &&
                  int(GIP_RankComm::MIN) == int(GIP_PartitionGroup::MIN),
              "one numbering of the reduction operations");

int GIP_ArchHIP::partCount() const { return group_ ? group_->size() : ranks_ ? ranks_->size() : 1; }

int GIP_ArchHIP::partRank() const { return group_ ? rank_ : ranks_ ? ranks_->rank() : 0; }

double GIP_ArchHIP::allreduce(double v, int op) {
    if (group_) return group_->allreduce(rank_, v, op);
    if (ranks_) return ranks_->allreduce(v, op);
    return v;
}

void GIP_ArchHIP::exchange(GIP_Topo *topo, double *dvec) {
    if (!topo) return;
    if (group_)
        group_->exchange(rank_, topo, ctx(), dvec);
    else if (ranks_)
        ranks_->exchange(topo, ctx(), dvec);
}

void gip_check(int rc, const char *what) {
    if (rc != GIP_OK) {
        std::fprintf(stderr, "gip_hip: %s failed (%d): %s\n", what, rc, gip_last_error());
        std::abort();
    }
}

GIP_ArchHIP::GIP_ArchHIP(int cpus, int cores, int devs) : GIP_Arch(cpus, cores, devs), devId(0) {}

GIP_ArchHIP::~GIP_ArchHIP() {
    delete ranks_;  // its device buffers belong to ctx_, still alive here
    if (ctx_) {
        for (auto &kv : pool_) gip_free(ctx_, kv.second);
        gip_finalize(ctx_);
    }
}

void GIP_ArchHIP::postConstruct(int cpus, int cores, int devs) {
    num_cpus = cpus;
    num_cores = cores;
    num_devs = devs;
}

// Several processes: one per partition, each on device localRank mod the
// device count (the reference takes rank mod the node's CPU count,
// GIP_ArchGPU.cpp:3-8); several processes may share one device.
void GIP_ArchHIP::setUp() {
    if (ctx_) return;
    if (!group_ && gip_rank_comm_factory && (ranks_ = gip_rank_comm_factory())) {
        int ndev = 0;
        gip_check(gip_device_count(&ndev), "gip_device_count");
        if (ndev > 0) devId = ranks_->localRank() % ndev;
    }
    gip_check(gip_init(devId, &ctx_), "gip_init");
}

gip_ctx *GIP_ArchHIP::ctx() {
    if (!ctx_) setUp();
    return ctx_;
}

void *GIP_ArchHIP::acquire(size_t bytes) {
    auto it = pool_.find(bytes);
    if (it != pool_.end()) {
        void *p = it->second;
        pool_.erase(it);
        return p;
    }
    void *p = nullptr;
    gip_check(gip_malloc(ctx(), bytes, &p), "gip_malloc");
    return p;
}

void GIP_ArchHIP::release(void *ptr, size_t bytes) {
    if (ptr) pool_.emplace(bytes, ptr);
}

GIP_ArchHIP *GIP_ArchHIP::of(GIP_Arch *a) {
    GIP_ArchHIP *h = dynamic_cast<GIP_ArchHIP *>(a);
    if (!h) {
        std::fprintf(stderr, "gip_hip: the node passed to a HIP vector/operator is not a GIP_ArchHIP\n");
        std::abort();
    }
    return h;
}
