// ORACLE — TEST INFRASTRUCTURE ONLY.
// Drives the reference's own, unmodified templates and Multicore
// specification (compiled from the reference sources by oracle/build_ref.sh)
// so that tests and the CPU-baseline leg can read full-precision results
// and timings from the real reference code.
//
//   gip_ref_harness dns <name> <steps> <gamma> <pcg_maxit> [dumpdir]
//       runs GIP_DNS<Multicore...> exactly as GIP_DNS::integrate()
//       (GIP_DNS.h:98-113), printing each step's line at 17 digits; rank r
//       dumps u,v,w,p to dumpdir/<field>-<r>.vec.  Needs ./InputFiles.
//   gip_ref_harness pcg <n> <iters> <seed>
//       builds the n^3 cavity Laplacian (pinned) with libgip_cavity, a
//       zero-mean N(0,1) right-hand side, and times the reference
//       GIP_LinearSolver::solve (GIP_LinearSolver.h:45-117) for <iters>
//       fixed iterations on one rank.
//
// Private members are reached by widening access around the reference
// headers; no reference source is changed.
#include <omp.h>
#include <mpi.h>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <map>
#include <string>

#define private public
#define protected public
#include "src/Algorithms/GIP_DNS.h"
#include "src/Specifications/Multicore/Multicore.h"
#undef private
#undef protected

#include "gip_cavity.h"

typedef GIP_DNS<GIP_MatrixMulticore, GIP_VectorMulticore, GIP_ArchCPU, GIP_NonLinearMulticore> DNS;

static void dump(const char *dir, const char *nm, int rank, GIP_VectorMulticore &v) {
    char path[4096];
    std::snprintf(path, sizeof path, "%s/%s-%d.vec", dir, nm, rank);
    gip_vec_write(path, v.getHostPtr(), v.getSize());
}

static int run_dns(int argc, char **argv) {
    if (argc < 6) return 1;
    int rank;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    DNS *d = new DNS(argv[2]);
    d->setUp();
    d->maxIt = std::atoi(argv[3]);
    d->gamma = std::atof(argv[4]);
    int pcg_maxit = std::atoi(argv[5]);
    if (pcg_maxit != 100) {
        GIP_Parameters param;
        param.setDouble("tol", 1e-10);
        param.setBool("precond", true);
        param.setInt("maxIt", pcg_maxit);
        d->cg.setUp(&d->L, &param);
    }
    std::cout.precision(17);
    MPI_Barrier(MPI_COMM_WORLD);
    double t0 = MPI_Wtime();
    d->integrate();
    double t1 = MPI_Wtime();
    if (rank == 0) std::fprintf(stderr, "integrate_ms %.3f\n", (t1 - t0) * 1e3);
    if (argc > 6) {
        dump(argv[6], "u", rank, d->u);
        dump(argv[6], "v", rank, d->v);
        dump(argv[6], "w", rank, d->w);
        dump(argv[6], "p", rank, d->p);
    }
    return 0;
}

static int run_pcg(int argc, char **argv) {
    if (argc < 5) return 1;
    int n = std::atoi(argv[2]), iters = std::atoi(argv[3]);
    uint64_t seed = std::strtoull(argv[4], nullptr, 10);
    gip_cavity_desc desc{};
    desc.nx = desc.ny = desc.nz = n;
    desc.gamma = 1e-3;
    desc.pin = 1;
    desc.ops = GIP_OP_L;
    gip_cavity *cav = nullptr;
    if (gip_cavity_generate(&desc, &cav)) {
        std::fprintf(stderr, "generate: %s\n", gip_cavity_last_error());
        return 2;
    }
    // one-partition topology, sizes as GIP_Topo.cpp:17-26 would read them
    GIP_Topo topo;
    topo.send_size = topo.recv_size = 0;
    topo.cols_size = 2;
    topo.ownedsize = topo.allsize = cav->n_all;
    topo.interiors = cav->n_inner;
    topo.interfaces = 0;
    topo.boundaries = cav->n_ghost;
    topo.halos = 0;
    topo.innersize = cav->n_inner;
    topo.send_indx = topo.recv_indx = nullptr;
    topo.send_row = new int[2]{0, 0};
    topo.recv_row = new int[2]{0, 0};
    topo.send_buffer = topo.recv_buffer = nullptr;
    GIP_ArchCPU node(1, 1, 1);
    node.setUp();
    GIP_MatrixMulticore L;
    L.nnz = cav->L.nnz;
    L.num_cols = cav->n_all;
    L.num_rows = cav->L.nrows;
    L.csrValA = cav->L.val;
    L.csrColIndA = cav->L.col;
    L.csrRowIndA = cav->L.rowptr;
    L.myTopo = &topo;
    L.myArch = &node;
    GIP_VectorMulticore b, x, x0;
    b.postConstruct(&topo, &node);
    x.postConstruct(&topo, &node);
    x0.postConstruct(&topo, &node);
    b = 0;
    x = 0;
    x0 = 0;
    // zero-mean N(0,1) right-hand side over the inner cells
    double *bp = b.getHostPtr();
    gip_cavity_rhs_normal(cav->n_inner, seed, bp);
    GIP_LinearSolver<GIP_MatrixMulticore, GIP_VectorMulticore> cg;
    GIP_Parameters param;
    param.setDouble("tol", 1e-10);
    param.setBool("precond", true);
    param.setInt("maxIt", iters);
    cg.setUp(&L, &param);
    double t0 = MPI_Wtime();
    cg.solve(&b, &x, &x0);
    double t1 = MPI_Wtime();
    // true residual ||b - Ax||_2 / ||b||_2 over the inner cells
    GIP_VectorMulticore ax;
    ax.postConstruct(&topo, &node);
    ax = 0;
    L.spmv(&x, &ax);
    double rr = 0.0, bb = 0.0;
    for (int i = 0; i < cav->n_inner; ++i) {
        double r = bp[i] - ax.getHostPtr()[i];
        rr += r * r;
        bb += bp[i] * bp[i];
    }
    std::printf("{\"n\": %d, \"iters\": %d, \"solve_ms\": %.3f, \"ms_per_iter\": %.6f, \"relres\": %.6e}\n", n, iters,
                (t1 - t0) * 1e3, (t1 - t0) * 1e3 / iters, std::sqrt(rr / bb));
    return 0;
}

int main(int argc, char **argv) {
    int prov;
    MPI_Init_thread(0, 0, MPI_THREAD_SINGLE, &prov);
    int rc = 1;
    if (argc > 1 && !std::strcmp(argv[1], "dns")) rc = run_dns(argc, argv);
    else if (argc > 1 && !std::strcmp(argv[1], "pcg")) rc = run_pcg(argc, argv);
    if (rc == 1)
        std::fprintf(stderr, "usage: gip_ref_harness dns <name> <steps> <gamma> <pcg_maxit> [dumpdir]\n"
                             "       gip_ref_harness pcg <n> <iters> <seed>\n");
    MPI_Finalize();
    return rc;
}
