// Jacobi-preconditioned conjugate gradients as a template over the
// Matrix/Vector specifications (contract of GIP_LinearSolver.h:11-29 in the
// reference: setUp(Matrix*, GIP_Parameters*), solve(b, x, x0), calcResid).
//
// Parameters: "maxIt" (int), "precond" (bool), "tol" (double) and two
// additions of this backend:
//   tol > 0   stops once ||r||_2 <= tol ||b||_2 over the inner range,
//             checked at the end of each iteration (the reference reads tol
//             but always runs maxIt iterations, GIP_LinearSolver.h:34,85);
//             tol <= 0 keeps the reference's fixed-iteration behaviour.
//   "mode"    "reference" (default): the operation sequence of
//             GIP_LinearSolver.h:54-116 through the Vector/Matrix methods;
//             "fused" / "mixed": the specification's device-resident solver
//             (Matrix::pcgSolve), FP64 or FP32-inner mixed precision.
#ifndef _gip_linsolver_
#define _gip_linsolver_

#include <cmath>
#include <string>
#include <type_traits>
#include <utility>

#include "GIP_Arch.h"
#include "GIP_Parameters.h"
#include "GIP_Topo.h"
#include "gip_hip.h"

namespace gip_detail {
template <class M, class V, class = void>
struct has_pcg : std::false_type {};
template <class M, class V>
struct has_pcg<M, V,
               std::void_t<decltype(std::declval<M &>().pcgSolve((V *)nullptr, (V *)nullptr, (V *)nullptr,
                                                                  std::declval<const gip_pcg_opts &>(),
                                                                  (gip_pcg_stats *)nullptr))>> : std::true_type {};
}  // namespace gip_detail

template <class Matrix, class Vector>
class GIP_LinearSolver {
   public:
    GIP_LinearSolver() {}
    ~GIP_LinearSolver() {}
    void setUp(Matrix *A, GIP_Parameters *param);
    void solve(Vector *b, Vector *x, Vector *x0);
    double calcResid(Vector *x, Vector *b);

    int lastIterations() const { return lastIters; }
    double lastRelResidual() const { return lastRel; }
    int lastStatus() const { return lastStat; }
    int lastCorrections() const { return lastCorr; }

   protected:
    Matrix *dA = nullptr;
    Matrix diagonal;
    double tol = 0.0;
    int maxIt = 0;
    bool precond = true;
    std::string mode = "reference";
    double innerTol = 0.1;
    GIP_Topo *myTopo = nullptr;
    GIP_Arch *myNode = nullptr;
    int lastIters = 0, lastStat = 0, lastCorr = 0;
    double lastRel = 0.0;
};

template <class Matrix, class Vector>
void GIP_LinearSolver<Matrix, Vector>::setUp(Matrix *A, GIP_Parameters *param) {
    tol = param->hasDouble("tol") ? param->getDouble("tol") : 0.0;
    maxIt = param->getInt("maxIt");
    precond = param->hasBool("precond") ? param->getBool("precond") : true;
    mode = param->hasString("mode") ? param->getString("mode") : std::string("reference");
    innerTol = param->hasDouble("innerTol") ? param->getDouble("innerTol") : 0.1;
    dA = A;
    myTopo = A->getMyTopo();
    if (precond) diagonal.setUpDiagonal(A);
    myNode = A->getMyArch();
}

template <class Matrix, class Vector>
void GIP_LinearSolver<Matrix, Vector>::solve(Vector *db, Vector *dx, Vector *dx0) {
    if constexpr (gip_detail::has_pcg<Matrix, Vector>::value) {
        if (mode == "fused" || mode == "mixed") {
            gip_pcg_opts o{};
            o.maxit = maxIt;
            o.tol = tol;
            o.precond = precond ? 1 : 0;
            o.precision = mode == "mixed" ? GIP_PCG_MIXED : GIP_PCG_FP64;
            o.inner_tol = innerTol;
            o.poll = 16;
            gip_pcg_stats st{};
            dA->pcgSolve(db, dx, dx0, o, &st);
            lastIters = st.iters;
            lastRel = st.relres;
            lastStat = st.status;
            lastCorr = st.corrections;
            return;
        }
    }
    // the operation sequence of GIP_LinearSolver.h:54-116
    Vector dp(myTopo, myNode);
    dx0->update();
    dx0->copyTo(&dp, _ALL_);
    Vector dAx(myTopo, myNode);
    Vector dr(myTopo, myNode);
    Vector ds(myTopo, myNode);
    Vector dtemp(myTopo, myNode);
    dtemp = 0;
    db->copyTo(&dr, _INNER_);
    dx0->copyTo(&dAx, _ALL_);
    dx0->copyTo(dx, _ALL_);
    dA->spmv(dx, &dAx);
    dr.axpy(&dAx, -1.0, _INNER_);
    if (precond)
        diagonal.spmv(&dr, &dp);
    else
        dr.copyTo(&dp, _INNER_);
    double r1 = dr.dot(&dp, &dtemp, _INNER_);
    double bnorm = 0.0, rnorm = 0.0;
    if (tol > 0.0) {
        bnorm = std::sqrt(db->dot(db, &dtemp, _INNER_));
        rnorm = std::sqrt(dr.dot(&dr, &dtemp, _INNER_));
    }
    int k = 0;
    lastStat = 0;
    if (!(tol > 0.0 && rnorm <= tol * bnorm)) {
        for (k = 1; k <= maxIt; ++k) {
            dp.update();
            dA->spmv(&dp, &dAx);
            double dn = dp.dot(&dAx, &dtemp, _INNER_);
            double alpha = r1 / dn;
            if (dn == 0.0 || !std::isfinite(dn) || !std::isfinite(alpha)) {  // breakdown: stop before the update
                lastStat = GIP_EBREAKDOWN;
                --k;
                break;
            }
            dx->axpy(&dp, alpha, _INNER_);
            dr.axpy(&dAx, -1.0 * alpha, _INNER_);
            if (precond)
                diagonal.spmv(&dr, &ds);
            else
                dr.copyTo(&ds, _INNER_);
            double r0 = r1;
            r1 = dr.dot(&ds, &dtemp, _INNER_);
            double beta = r1 / r0;
            dp.axpy(&ds, 1.0, beta, _INNER_);
            if (tol > 0.0) {
                rnorm = std::sqrt(dr.dot(&dr, &dtemp, _INNER_));
                if (rnorm <= tol * bnorm) break;
            }
        }
        if (k > maxIt) k = maxIt;
    }
    if (!(tol > 0.0)) {
        bnorm = std::sqrt(db->dot(db, &dtemp, _INNER_));
        rnorm = std::sqrt(dr.dot(&dr, &dtemp, _INNER_));
    }
    lastIters = k;
    lastRel = bnorm > 0.0 ? rnorm / bnorm : rnorm;
    lastCorr = 0;
    if (lastStat == 0 && tol > 0.0 && !(rnorm <= tol * bnorm)) lastStat = GIP_ENOCONV;
}

// ||A x - b||_2 over the owned range (GIP_LinearSolver.h:120-133)
template <class Matrix, class Vector>
double GIP_LinearSolver<Matrix, Vector>::calcResid(Vector *dx, Vector *db) {
    Vector db2(myTopo, myNode);
    Vector dtemp(myTopo, myNode);
    dtemp = 0;
    dx->update();
    dA->spmv(dx, &db2);
    db2.axpy(db, 1.0, -1.0, _OWNED_);
    return std::sqrt(db2.dot(&db2, &dtemp, _OWNED_));
}

#endif
