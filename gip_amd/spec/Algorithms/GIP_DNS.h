// Algebraic fractional-step integrator of the incompressible Navier–Stokes
// equations as a template over the specifications (contract of
// GIP_DNS.h:13-96 in the reference: GIP_DNS(name), setUp(), integrate()).
// Each time step runs the five phases of GIP_DNS.h:308-414 in the same
// order with the same operators:
//   predictor   C.values = Ec.mf ; a = D u - C u (x3) ; u += dt(1.5 a - 0.5 a0) on the inner cells
//   Poisson     rhs = sigma (Mx u + My v + Mz w) ; PCG on L ; ghost p <- adjacent p (Lb)
//   correction  gp = G p (x3) ; u -= (dt/rho) gp
//   mass flux   mf = -dt SO p + dt SF gp + rho SF u
//   CFL         massr = ||Mres mf||_inf ; time += dt ; dt = CFL(u, v, w) ; sigma = 1/dt
// The dead stores u0 = u of the reference (GIP_DNS.h:410-412) are left out.
//
// Additions: the constants the reference hardcodes (GIP_DNS.h:132-137,
// :180-185) come from configure(); setUp(const gip_cavity*) takes the
// operators from memory instead of InputFiles; step() runs one step and
// reports its numbers; an optional hook is called at the six phase
// boundaries of a step (0 before the predictor ... 5 after the CFL phase),
// for timing.  Files are read relative to root() for the node's partition
// (<name>_<np>p/...-<rank>; one partition: <name>_1p/...-0).
#ifndef _gip_dns_
#define _gip_dns_

#include <cstdio>
#include <functional>
#include <iostream>
#include <stdexcept>
#include <string>

#include "GIP_LinearSolver.h"
#include "GIP_Parameters.h"
#include "GIP_Topo.h"
#include "gip_cavity.h"

struct GIP_StepInfo {
    double time, dt, massr, pcg_relres;
    int step, pcg_iters, pcg_status, pcg_corrections;
};

template <class Matrix, class Vector, class Node, class NonLinear>
class GIP_DNS {
   public:
    GIP_DNS() : nodo(1, 4, 4) {}
    explicit GIP_DNS(const char *name) : nodo(1, 4, 4), name(name) {}
    ~GIP_DNS() {}

    // keys: steps, gamma, rho, pcg_tol, pcg_maxIt, pcg_precond, pcg_mode, pcg_innerTol, print, device
    void configure(const GIP_Parameters &p);
    void setRoot(const std::string &r) { root = r; }
    void setUp();
    void setUp(const gip_cavity *cav);
    void integrate();
    void step(GIP_StepInfo *info);
    // called with 0..5 at the phase boundaries of every step (empty: none)
    void setPhaseHook(std::function<void(int)> h) { phaseHook = std::move(h); }

    Vector &field(const std::string &f);
    int innerSize() { return topoCols.getInnerSize(); }
    int allSize() { return topoCols.getAllSize(); }
    int ownedSize() { return topoCols.getOwnedSize(); }
    int facesSize() { return topoFaces.getAllSize(); }
    double currentDt() const { return dt; }
    double currentTime() const { return time; }
    Node &node() { return nodo; }
    Matrix &op(const std::string &o);

   protected:
    Node nodo;  // first member: destroyed last, after every vector and operator
    std::string name, root;
    Matrix C, Ec, D, L, Lb, Mx, My, Mz, Gx, Gy, Gz, SFx, SFy, SFz, SO, Mres;
    Vector u, v, w, ax, ay, az, ax0, ay0, az0, rhs, p, p0, gpx, gpy, gpz, mf, dxs, temp, temp2;
    GIP_LinearSolver<Matrix, Vector> cg;
    GIP_Topo topoCols, topoFaces;
    NonLinear noLin;

    double gamma = 1.0 / 1.0e3, rho = 1.0, sigma = 0.0, dt = 1e-3, time = 0.0;
    int maxIt = 10, it = 0;
    bool started = false;
    int print = 1;  // 0: silent, 1: reference format, 2: 17 digits + PCG iterations
    GIP_Parameters solverParam;

    std::string path(const char *kind, const char *group, const char *op, const char *ext);
    void allocateFields();
    void setUpSolver();
    std::function<void(int)> phaseHook;
    void mark(int phase) {
        if (phaseHook) phaseHook(phase);
    }
    void PredictorVelocity();
    void PoissonEquation();
    void VelocityCorrection();
    void MassFluxes();
    void CFLCondition(GIP_StepInfo *info);
};

template <class Matrix, class Vector, class Node, class NonLinear>
void GIP_DNS<Matrix, Vector, Node, NonLinear>::configure(const GIP_Parameters &p) {
    if (p.hasInt("steps")) maxIt = p.getInt("steps");
    if (p.hasDouble("gamma")) gamma = p.getDouble("gamma");
    if (p.hasDouble("rho")) rho = p.getDouble("rho");
    if (p.hasInt("print")) print = p.getInt("print");
    if (p.hasInt("device")) nodo.setDevice(p.getInt("device"));
    solverParam.setDouble("tol", p.hasDouble("pcg_tol") ? p.getDouble("pcg_tol") : 0.0);
    solverParam.setInt("maxIt", p.hasInt("pcg_maxIt") ? p.getInt("pcg_maxIt") : 100);
    solverParam.setBool("precond", p.hasBool("pcg_precond") ? p.getBool("pcg_precond") : true);
    solverParam.setString("mode", p.hasString("pcg_mode") ? p.getString("pcg_mode") : std::string("reference"));
    solverParam.setDouble("innerTol", p.hasDouble("pcg_innerTol") ? p.getDouble("pcg_innerTol") : 0.1);
    // the device-resident solvers work on one partition; several partitions
    // take the reference sequence (halo refresh and global dot products
    // through the node's partition group)
    if (nodo.partCount() > 1 && solverParam.getString("mode") != "reference")
        throw std::runtime_error("several partitions need the reference PCG sequence (pcg_mode reference)");
    if (nodo.partRank() != 0) print = 0;  // rank 0 prints, as GIP_DNS.h:403
}

template <class Matrix, class Vector, class Node, class NonLinear>
std::string GIP_DNS<Matrix, Vector, Node, NonLinear>::path(const char *kind, const char *group, const char *op,
                                                          const char *ext) {
    // <name>_<np>p/<op><name>_<np>p-<rank> (GIP_DNS.h:149): np and rank are
    // those of the node's partition group (1 and 0 without one)
    std::string base = root.empty() ? std::string() : root + "/";
    std::string g = group[0] ? std::string(group) + "/" : std::string();
    const std::string np = std::to_string(nodo.partCount()) + "p";
    return base + "InputFiles/" + kind + "/" + g + name + "/" + name + "_" + np + "/" + op + name + "_" + np + "-" +
           std::to_string(nodo.partRank()) + ext;
}

template <class Matrix, class Vector, class Node, class NonLinear>
void GIP_DNS<Matrix, Vector, Node, NonLinear>::allocateFields() {
    nodo.setUp();
    Vector *cells[] = {&p, &p0, &ax, &ax0, &ay, &ay0, &az, &az0, &gpx, &gpy, &gpz, &rhs, &temp, &temp2};
    for (Vector *f : cells) f->postConstruct(&topoCols, &nodo);
    mf.postConstruct(&topoFaces, &nodo);
}

template <class Matrix, class Vector, class Node, class NonLinear>
void GIP_DNS<Matrix, Vector, Node, NonLinear>::setUpSolver() {
    if (!solverParam.hasInt("maxIt")) {
        solverParam.setDouble("tol", 0.0);
        solverParam.setInt("maxIt", 100);
        solverParam.setBool("precond", true);
    }
    cg.setUp(&L, &solverParam);
}

// GIP_DNS.h:115-306 on one partition
template <class Matrix, class Vector, class Node, class NonLinear>
void GIP_DNS<Matrix, Vector, Node, NonLinear>::setUp() {
    topoCols.postConstruct(path("Topos", "", "topoCols", ".topo"));
    topoFaces.postConstruct(path("Topos", "", "topoFaces", ".topo"));
    allocateFields();
    u.postConstruct(path("Scalarfields", "", "u", ".vec"), &topoCols, &nodo);
    v.postConstruct(path("Scalarfields", "", "v", ".vec"), &topoCols, &nodo);
    w.postConstruct(path("Scalarfields", "", "w", ".vec"), &topoCols, &nodo);
    dxs.postConstruct(path("Scalarfields", "", "dxs", ".vec"), &topoCols, &nodo);
    D.postConstruct(path("Operators", "CD", "D", ".csr"), &topoCols, &nodo);
    C.postConstruct(path("Operators", "CD", "D", ".csr"), &topoCols, &nodo);  // C takes D's pattern
    Ec.postConstruct(path("Operators", "CD", "Ec", ".csr"), &topoFaces, &nodo);
    L.postConstruct(path("Operators", "L", "L", ".csr"), &topoCols, &nodo);
    Lb.postConstruct(path("Operators", "L", "Lb", ".csr"), &topoCols, &nodo);
    Mx.postConstruct(path("Operators", "M", "Mx", ".csr"), &topoCols, &nodo);
    My.postConstruct(path("Operators", "M", "My", ".csr"), &topoCols, &nodo);
    Mz.postConstruct(path("Operators", "M", "Mz", ".csr"), &topoCols, &nodo);
    Gx.postConstruct(path("Operators", "G", "Gx", ".csr"), &topoCols, &nodo);
    Gy.postConstruct(path("Operators", "G", "Gy", ".csr"), &topoCols, &nodo);
    Gz.postConstruct(path("Operators", "G", "Gz", ".csr"), &topoCols, &nodo);
    SFx.postConstruct(path("Operators", "S", "SFx", ".csr"), &topoFaces, &nodo);
    SFy.postConstruct(path("Operators", "S", "SFy", ".csr"), &topoFaces, &nodo);
    SFz.postConstruct(path("Operators", "S", "SFz", ".csr"), &topoFaces, &nodo);
    SO.postConstruct(path("Operators", "S", "SO", ".csr"), &topoFaces, &nodo);
    Mres.postConstruct(path("Operators", "M", "Mres", ".csr"), &topoCols, &nodo);
    setUpSolver();
}

template <class Matrix, class Vector, class Node, class NonLinear>
void GIP_DNS<Matrix, Vector, Node, NonLinear>::setUp(const gip_cavity *c) {
    topoCols.postConstructSizes(c->n_inner, c->n_ghost);
    topoFaces.postConstructSizes(c->n_faces, 0);
    allocateFields();
    u.postConstruct(c->u, c->n_all, &topoCols, &nodo);
    v.postConstruct(c->v, c->n_all, &topoCols, &nodo);
    w.postConstruct(c->w, c->n_all, &topoCols, &nodo);
    dxs.postConstruct(c->dxs, c->n_all, &topoCols, &nodo);
    // operators keep pointing at the cavity's host arrays (the cavity outlives
    // the integrator); C gets its own copy since its values are rebuilt
    D.postConstructShared(c->D, &topoCols, &nodo);
    C.postConstruct(c->D, &topoCols, &nodo);
    Ec.postConstructShared(c->Ec, &topoFaces, &nodo);
    L.postConstructShared(c->L, &topoCols, &nodo);
    Lb.postConstructShared(c->Lb, &topoCols, &nodo);
    Mx.postConstructShared(c->Mx, &topoCols, &nodo);
    My.postConstructShared(c->My, &topoCols, &nodo);
    Mz.postConstructShared(c->Mz, &topoCols, &nodo);
    Gx.postConstructShared(c->Gx, &topoCols, &nodo);
    Gy.postConstructShared(c->Gy, &topoCols, &nodo);
    Gz.postConstructShared(c->Gz, &topoCols, &nodo);
    SFx.postConstructShared(c->SFx, &topoFaces, &nodo);
    SFy.postConstructShared(c->SFy, &topoFaces, &nodo);
    SFz.postConstructShared(c->SFz, &topoFaces, &nodo);
    SO.postConstructShared(c->SO, &topoFaces, &nodo);
    Mres.postConstructShared(c->Mres, &topoCols, &nodo);
    setUpSolver();
}

template <class Matrix, class Vector, class Node, class NonLinear>
void GIP_DNS<Matrix, Vector, Node, NonLinear>::integrate() {
    for (int k = 0; k < maxIt; ++k) step(nullptr);
}

template <class Matrix, class Vector, class Node, class NonLinear>
void GIP_DNS<Matrix, Vector, Node, NonLinear>::step(GIP_StepInfo *info) {
    if (!started) {  // GIP_DNS.h:101-102
        dt = noLin.getCFL(&topoCols, &u, &v, &w, &dxs, &temp, gamma, rho);
        sigma = 1.0 / dt;
        started = true;
    }
    mark(0);
    PredictorVelocity();
    mark(1);
    PoissonEquation();
    mark(2);
    VelocityCorrection();
    mark(3);
    MassFluxes();
    mark(4);
    CFLCondition(info);
    mark(5);
    ++it;
}

template <class Matrix, class Vector, class Node, class NonLinear>
void GIP_DNS<Matrix, Vector, Node, NonLinear>::PredictorVelocity() {
    Ec.spmv(&mf, &C);
    D.spmv(&u, &ax);
    D.spmv(&v, &ay);
    D.spmv(&w, &az);
    C.spmv(&u, &ax, -1.0, 1.0);
    C.spmv(&v, &ay, -1.0, 1.0);
    C.spmv(&w, &az, -1.0, 1.0);
    const double a1 = 1.5 * dt, a2 = -0.5 * dt;
    u.axpy(&ax, a1, _INNER_);
    u.axpy(&ax0, a2, _INNER_);
    v.axpy(&ay, a1, _INNER_);
    v.axpy(&ay0, a2, _INNER_);
    w.axpy(&az, a1, _INNER_);
    w.axpy(&az0, a2, _INNER_);
    u.update();
    v.update();
    w.update();
    ax.copyTo(&ax0, _INNER_);
    ay.copyTo(&ay0, _INNER_);
    az.copyTo(&az0, _INNER_);
}

template <class Matrix, class Vector, class Node, class NonLinear>
void GIP_DNS<Matrix, Vector, Node, NonLinear>::PoissonEquation() {
    Mx.spmv(&u, &rhs);
    My.spmv(&v, &rhs, 1.0, 1.0);
    Mz.spmv(&w, &rhs, sigma, sigma);
    cg.solve(&rhs, &p, &p0);
    p.update();
    p.copyTo(&p0, _INNER_);
    p0.copyTo(&p, _OWNED_);
    Lb.spmv(&p0, &p, 1.0, 1.0);
}

template <class Matrix, class Vector, class Node, class NonLinear>
void GIP_DNS<Matrix, Vector, Node, NonLinear>::VelocityCorrection() {
    Gx.spmv(&p, &gpx);
    Gy.spmv(&p, &gpy);
    Gz.spmv(&p, &gpz);
    gpx.update();
    gpy.update();
    gpz.update();
    const double fact = -1.0 / (rho * sigma);
    u.axpy(&gpx, fact, _INNER_);
    v.axpy(&gpy, fact, _INNER_);
    w.axpy(&gpz, fact, _INNER_);
    u.update();
    v.update();
    w.update();
}

template <class Matrix, class Vector, class Node, class NonLinear>
void GIP_DNS<Matrix, Vector, Node, NonLinear>::MassFluxes() {
    const double f1 = -1.0 / sigma, f2 = 1.0 / sigma;
    SO.spmv(&p, &mf, f1, 0.0);
    SFx.spmv(&gpx, &mf, f2, 1.0);
    SFy.spmv(&gpy, &mf, f2, 1.0);
    SFz.spmv(&gpz, &mf, f2, 1.0);
    SFx.spmv(&u, &mf, rho, 1.0);
    SFy.spmv(&v, &mf, rho, 1.0);
    SFz.spmv(&w, &mf, rho, 1.0);
}

template <class Matrix, class Vector, class Node, class NonLinear>
void GIP_DNS<Matrix, Vector, Node, NonLinear>::CFLCondition(GIP_StepInfo *info) {
    Mres.spmv(&mf, &temp);
    const double massres = temp.norm(&temp2, _INNER_);
    const double used = dt;
    time += dt;
    if (print == 1) {
        std::printf("time %e [dt::%e] massr %e ite %d\n", time, dt, massres, it + 1);
    } else if (print == 2) {
        std::printf("time %.17e [dt::%.17e] massr %.17e ite %d pcg %d status %d\n", time, dt, massres, it + 1,
                    cg.lastIterations(), cg.lastStatus());
    }
    if (print) std::fflush(stdout);
    dt = noLin.getCFL(&topoCols, &u, &v, &w, &dxs, &temp, gamma, rho);
    sigma = 1.0 / dt;
    if (info) {
        info->time = time;
        info->dt = used;
        info->massr = massres;
        info->step = it + 1;
        info->pcg_iters = cg.lastIterations();
        info->pcg_status = cg.lastStatus();
        info->pcg_relres = cg.lastRelResidual();
        info->pcg_corrections = cg.lastCorrections();
    }
}

template <class Matrix, class Vector, class Node, class NonLinear>
Vector &GIP_DNS<Matrix, Vector, Node, NonLinear>::field(const std::string &f) {
    if (f == "u") return u;
    if (f == "v") return v;
    if (f == "w") return w;
    if (f == "p") return p;
    if (f == "mf") return mf;
    if (f == "rhs") return rhs;
    if (f == "ax0") return ax0;
    if (f == "ay0") return ay0;
    if (f == "az0") return az0;
    if (f == "p0") return p0;
    if (f == "dxs") return dxs;
    throw std::out_of_range("unknown field " + f);
}

template <class Matrix, class Vector, class Node, class NonLinear>
Matrix &GIP_DNS<Matrix, Vector, Node, NonLinear>::op(const std::string &o) {
    Matrix *all[] = {&C, &Ec, &D, &L, &Lb, &Mx, &My, &Mz, &Gx, &Gy, &Gz, &SFx, &SFy, &SFz, &SO, &Mres};
    const char *names[] = {"C", "Ec", "D", "L", "Lb", "Mx", "My", "Mz", "Gx", "Gy", "Gz", "SFx", "SFy", "SFz", "SO", "Mres"};
    for (int i = 0; i < 16; ++i)
        if (o == names[i]) return *all[i];
    throw std::out_of_range("unknown operator " + o);
}

#endif
