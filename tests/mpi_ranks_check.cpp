// CPU check of the HIP specification's MPI ranks (GIP_RanksMPI.cpp) under
// mpiexec: one process per partition of a case written by the partition
// writer, whose cell field carries every cell's global number.  The device
// entry points the node and the rank communicator call are replaced here by
// host versions (device memory is host memory, the indexed copy a loop), so
// only the rank logic is under test: device selection, the three
// all-reductions, and the halo exchange, which must leave every halo slot
// holding the number the writer stored there.
//
//   mpi_ranks_check <topology file pattern> <label vector pattern>
// Each pattern holds one %d, replaced by the rank.  Prints one line per rank
// ("rank R of N: ...") and exits 0 when every check holds; a failed check
// aborts the process.
#include <mpi.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "GIP_ArchHIP.h"
#include "GIP_Topo.h"
#include "gip_cavity.h"
#include "gip_hip.h"

struct gip_ctx {
    int device;
};

extern "C" {
const char *gip_last_error(void) { return "host stand-in"; }
int gip_device_count(int *count) {
    *count = 3;  // fewer devices than ranks: ranks share them
    return GIP_OK;
}
int gip_init(int device, gip_ctx **ctx) {
    *ctx = new gip_ctx{device};
    return GIP_OK;
}
int gip_finalize(gip_ctx *ctx) {
    delete ctx;
    return GIP_OK;
}
int gip_sync(gip_ctx *) { return GIP_OK; }
int gip_malloc(gip_ctx *, size_t bytes, void **p) {
    *p = std::malloc(bytes);
    return *p ? GIP_OK : GIP_ENOMEM;
}
int gip_free(gip_ctx *, void *p) {
    std::free(p);
    return GIP_OK;
}
int gip_memcpy_h2d(gip_ctx *, void *dst, const void *src, size_t bytes) {
    std::memcpy(dst, src, bytes);
    return GIP_OK;
}
int gip_memcpy_d2h(gip_ctx *, void *dst, const void *src, size_t bytes) {
    std::memcpy(dst, src, bytes);
    return GIP_OK;
}
int gip_dcopy_indexed(gip_ctx *, int32_t n, const int32_t *src, const double *x, const int32_t *dst, double *y) {
    for (int32_t i = 0; i < n; ++i) y[dst ? dst[i] : i] = x[src ? src[i] : i];
    return GIP_OK;
}
}

static void require(bool ok, int rank, const char *what) {
    if (!ok) {
        std::fprintf(stderr, "rank %d: check failed: %s\n", rank, what);
        std::fflush(stderr);
        std::abort();
    }
}

static std::string fmt(const char *pattern, int rank) {
    char buf[4096];
    std::snprintf(buf, sizeof buf, pattern, rank);
    return buf;
}

int main(int argc, char **argv) {
    int prov = 0;
    MPI_Init_thread(nullptr, nullptr, MPI_THREAD_SINGLE, &prov);
    int rank = 0, size = 1;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    require(argc == 3, rank, "usage: mpi_ranks_check <topology pattern> <label pattern>");
    {
        GIP_ArchHIP node(1, 4, 4);
        node.setUp();
        if (size == 1) {
            require(node.ranks() == nullptr && node.partCount() == 1, rank, "a lone process has no communicator");
            require(node.allreduce(2.5, GIP_RankComm::SUM) == 2.5, rank, "lone all-reduce is the identity");
            std::printf("rank 0 of 1: no communicator\n");
        } else {
            require(node.ranks() != nullptr, rank, "communicator");
            require(node.partRank() == rank && node.partCount() == size, rank, "rank and size");
            require(node.device() == node.ranks()->localRank() % 3, rank, "device = local rank mod devices");
            const double r1 = rank + 1.0;
            require(node.allreduce(r1, GIP_RankComm::SUM) == size * (size + 1) / 2.0, rank, "SUM");
            require(node.allreduce(r1, GIP_RankComm::MAX) == (double)size, rank, "MAX");
            require(node.allreduce(r1, GIP_RankComm::MIN) == 1.0, rank, "MIN");
        }

        GIP_Topo topo;
        topo.postConstruct(fmt(argv[1], rank));
        double *labels = nullptr;
        int32_t n = 0;
        require(gip_vec_read(fmt(argv[2], rank).c_str(), &labels, &n) == 0, rank, "label vector read");
        require(n == topo.getAllSize(), rank, "label vector covers the topology");
        std::vector<double> want(labels, labels + n);
        std::free(labels);

        void *p = nullptr;
        gip_malloc(nullptr, sizeof(double) * (size_t)n, &p);
        double *dvec = static_cast<double *>(p);
        const int nrecv = topo.getRecvSize();
        const int *recv = topo.getRecvIndx();
        for (int pass = 0; pass < 2; ++pass) {  // the second pass reuses the uploaded lists
            std::memcpy(dvec, want.data(), sizeof(double) * (size_t)n);
            for (int k = 0; k < nrecv; ++k) dvec[recv[k]] = -1.0;
            node.exchange(&topo, dvec);
            for (int i = 0; i < n; ++i) require(dvec[i] == want[(size_t)i], rank, "cell after the exchange");
        }
        gip_free(nullptr, dvec);
        require(size == 1 || nrecv > 0, rank, "a partition of a split case receives a halo");
        std::printf("rank %d of %d: %d halo slots refreshed, %d cells checked\n", rank, size, nrecv, (int)n);
        std::fflush(stdout);
        MPI_Barrier(MPI_COMM_WORLD);
    }
    MPI_Finalize();
    return 0;
}
