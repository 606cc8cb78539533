#include "GIP_ArchHIP.h"

#include <cstdio>
#include <cstdlib>

#include "GIP_PartitionHIP.h"

int GIP_ArchHIP::partCount() const { return group_ ? group_->size() : 1; }

void gip_check(int rc, const char *what) {
    if (rc != GIP_OK) {
        std::fprintf(stderr, "gip_hip: %s failed (%d): %s\n", what, rc, gip_last_error());
        std::abort();
    }
}

GIP_ArchHIP::GIP_ArchHIP(int cpus, int cores, int devs) : GIP_Arch(cpus, cores, devs), devId(0) {}

GIP_ArchHIP::~GIP_ArchHIP() {
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

void GIP_ArchHIP::setUp() {
    if (ctx_) return;
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
