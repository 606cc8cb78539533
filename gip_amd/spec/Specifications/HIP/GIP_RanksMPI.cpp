// The MPI build of the HIP specification: one partition per process, as the
// reference runs its partitioned cases (mpirun -np N, DrivenCavity.cpp:30-33;
// files named per rank, GIP_DNS.h:143-150).  This file is compiled only
// where MPI is (the build inside the reference tree, whose GIP_Topo needs
// <mpi.h> as well); the library build of the specification leaves it out,
// and its nodes then never have a communicator.  With one process nothing
// changes: the factory returns no communicator.
//   - allreduce: MPI_Allreduce with the reference's operations
//     (GIP_VectorMulticore.cpp:125,161, GIP_NonLinearMulticore.cpp:45);
//   - exchange: the send list gathered on the device (gip_dcopy_indexed) into
//     one contiguous block, one copy to the host, the reference's
//     point-to-point pattern (GIP_Topo.cpp:124-149: block q of the send list
//     goes to rank q, block q of the receive list comes from rank q), one
//     copy back and the receive list scattered on the device.  Only the send
//     and receive entries cross the host link, not the whole interface and
//     halo ranges that GIP_VectorCuda.cpp:224-247 copies.
// The communicator is a duplicate of MPI_COMM_WORLD, so the exchanges never
// match messages of the application's own.
#include <mpi.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>

#include "GIP_ArchHIP.h"
#include "../../BasicInterfaces/GIP_Topo.h"

namespace {

[[noreturn]] void fail(const char *what) {
    std::fprintf(stderr, "gip_hip (MPI ranks): %s\n", what);
    std::abort();
}

void mpi_check(int rc, const char *what) {
    if (rc != MPI_SUCCESS) fail(what);
}

// the send or receive list of one topology: rows[q] .. rows[q+1] is the
// block exchanged with rank q
struct Side {
    int n = 0;
    int32_t *d_indx = nullptr;  // device copy of the cell indices
    double *d_buf = nullptr;    // device staging, n values
    std::vector<double> h_buf;  // host staging
    std::vector<int> rows;
};

struct Halo {
    gip_ctx *ctx = nullptr;
    Side send, recv;
    std::vector<MPI_Request> req;
};

class MpiRanks : public GIP_RankComm {
   public:
    explicit MpiRanks(MPI_Comm world) {
        mpi_check(MPI_Comm_dup(world, &comm_), "MPI_Comm_dup");
        mpi_check(MPI_Comm_rank(comm_, &rank_), "MPI_Comm_rank");
        mpi_check(MPI_Comm_size(comm_, &size_), "MPI_Comm_size");
        MPI_Comm host;
        mpi_check(MPI_Comm_split_type(comm_, MPI_COMM_TYPE_SHARED, rank_, MPI_INFO_NULL, &host),
                  "MPI_Comm_split_type");
        mpi_check(MPI_Comm_rank(host, &local_), "MPI_Comm_rank (host)");
        mpi_check(MPI_Comm_free(&host), "MPI_Comm_free (host)");
    }

    ~MpiRanks() override {
        for (auto &kv : halos_) {
            Halo &h = kv.second;
            for (Side *s : {&h.send, &h.recv}) {
                if (s->d_indx) gip_free(h.ctx, s->d_indx);
                if (s->d_buf) gip_free(h.ctx, s->d_buf);
            }
        }
        // the reference driver finalizes MPI before its objects go out of
        // scope (DrivenCavity.cpp:40)
        int finalized = 0;
        MPI_Finalized(&finalized);
        if (!finalized) MPI_Comm_free(&comm_);
    }

    int rank() const override { return rank_; }
    int size() const override { return size_; }
    int localRank() const override { return local_; }

    double allreduce(double v, int op) override {
        const MPI_Op mop = op == SUM ? MPI_SUM : op == MAX ? MPI_MAX : MPI_MIN;
        double r = 0.0;
        mpi_check(MPI_Allreduce(&v, &r, 1, MPI_DOUBLE, mop, comm_), "MPI_Allreduce");
        return r;
    }

    void exchange(GIP_Topo *topo, gip_ctx *ctx, double *dvec) override {
        Halo &h = halo(topo, ctx);
        if (h.send.n) {
            gip_check(gip_dcopy_indexed(ctx, h.send.n, h.send.d_indx, dvec, nullptr, h.send.d_buf), "halo pack");
            gip_check(gip_memcpy_d2h(ctx, h.send.h_buf.data(), h.send.d_buf, sizeof(double) * (size_t)h.send.n),
                      "halo download");
        }
        int nreq = 0;
        for (int q = 0; q < size_; ++q) {
            const int n = h.recv.rows[(size_t)q + 1] - h.recv.rows[(size_t)q];
            if (n)
                mpi_check(MPI_Irecv(h.recv.h_buf.data() + h.recv.rows[(size_t)q], n, MPI_DOUBLE, q, kTag, comm_,
                                    &h.req[(size_t)nreq++]),
                          "MPI_Irecv");
        }
        for (int q = 0; q < size_; ++q) {
            const int n = h.send.rows[(size_t)q + 1] - h.send.rows[(size_t)q];
            if (n)
                mpi_check(MPI_Isend(h.send.h_buf.data() + h.send.rows[(size_t)q], n, MPI_DOUBLE, q, kTag, comm_,
                                    &h.req[(size_t)nreq++]),
                          "MPI_Isend");
        }
        mpi_check(MPI_Waitall(nreq, h.req.data(), MPI_STATUSES_IGNORE), "MPI_Waitall");
        if (h.recv.n) {
            gip_check(gip_memcpy_h2d(ctx, h.recv.d_buf, h.recv.h_buf.data(), sizeof(double) * (size_t)h.recv.n),
                      "halo upload");
            gip_check(gip_dcopy_indexed(ctx, h.recv.n, nullptr, h.recv.d_buf, h.recv.d_indx, dvec), "halo unpack");
        }
    }

   private:
    static constexpr int kTag = 0x6970;

    // the lists of a topology, checked and uploaded at its first exchange:
    // its row tables hold one block per rank (size + 1 offsets, the
    // topology file's third header field for a case partitioned for this
    // many ranks), and every index names a cell of the topology
    Halo &halo(GIP_Topo *topo, gip_ctx *ctx) {
        auto it = halos_.find(topo);
        if (it != halos_.end()) {
            if (it->second.ctx != ctx) fail("one topology exchanged on two nodes");
            return it->second;
        }
        Halo &h = halos_[topo];
        h.ctx = ctx;
        const int cells = topo->getAllSize();
        fill(h.send, ctx, topo->getSendSize(), topo->getSendIndx(), topo->getSendRow(), cells, "send");
        fill(h.recv, ctx, topo->getRecvSize(), topo->getRecvIndx(), topo->getRecvRow(), cells, "receive");
        h.req.resize(2 * (size_t)size_);
        return h;
    }

    void fill(Side &s, gip_ctx *ctx, int n, const int *indx, const int *rows, int cells, const char *which) {
        if (n < 0 || (n && (!indx || !rows)) || (!n && !rows)) fail("a topology without its exchange lists");
        s.n = n;
        s.rows.assign(rows, rows + size_ + 1);
        if (s.rows[0] != 0 || s.rows[(size_t)size_] != n) {
            std::fprintf(stderr, "gip_hip (MPI ranks): the %s table does not hold %d blocks covering %d entries\n",
                         which, size_, n);
            fail("topology partitioned for another number of ranks");
        }
        for (int q = 0; q < size_; ++q)
            if (s.rows[(size_t)q + 1] < s.rows[(size_t)q]) fail("decreasing exchange table");
        for (int i = 0; i < n; ++i)
            if (indx[i] < 0 || indx[i] >= cells) fail("exchange index outside the topology");
        s.h_buf.assign((size_t)(n ? n : 1), 0.0);
        void *p = nullptr;
        gip_check(gip_malloc(ctx, sizeof(int32_t) * (size_t)(n ? n : 1), &p), "gip_malloc (halo list)");
        s.d_indx = static_cast<int32_t *>(p);
        gip_check(gip_malloc(ctx, sizeof(double) * (size_t)(n ? n : 1), &p), "gip_malloc (halo staging)");
        s.d_buf = static_cast<double *>(p);
        if (n) {
            std::vector<int32_t> v(indx, indx + n);
            gip_check(gip_memcpy_h2d(ctx, s.d_indx, v.data(), sizeof(int32_t) * (size_t)n), "halo list upload");
        }
    }

    MPI_Comm comm_ = MPI_COMM_NULL;
    int rank_ = 0, size_ = 1, local_ = 0;
    std::map<GIP_Topo *, Halo> halos_;
};

GIP_RankComm *make_ranks() {
    int flag = 0;
    MPI_Initialized(&flag);
    if (!flag) return nullptr;
    MPI_Finalized(&flag);
    if (flag) return nullptr;
    int n = 1;
    mpi_check(MPI_Comm_size(MPI_COMM_WORLD, &n), "MPI_Comm_size");
    return n > 1 ? new MpiRanks(MPI_COMM_WORLD) : nullptr;
}

[[maybe_unused]] const bool registered = (gip_rank_comm_factory = &make_ranks, true);

}  // namespace
