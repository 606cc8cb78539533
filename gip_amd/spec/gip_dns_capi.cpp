// C ABI of GIP_DNS instantiated with the HIP specification (include/gip_dns.h).
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "GIP_DNS.h"
#include "GIP_PartitionHIP.h"
#include "HIP.h"
#include "gip_dns.h"
#include "gip_markers.h"

typedef GIP_DNS<GIP_MatrixHIP, GIP_VectorHIP, GIP_ArchHIP, GIP_NonLinearHIP> HipDNS;

struct gip_dns {
    HipDNS sim;
    gip_event *ev[6] = {};  // device events at the phase boundaries of the last step
    bool timed = false;
    explicit gip_dns(const char *n) : sim(n) {}
    ~gip_dns() {
        for (gip_event *e : ev)
            if (e) gip_event_destroy(e);
    }
    // events recorded on the node's stream at each phase boundary: no
    // synchronisation inside the step; read after it (each phase is also a
    // profiler range)
    void attach_timer() {
        gip_ctx *c = sim.node().ctx();
        for (gip_event *&e : ev)
            if (gip_event_create(c, &e) != GIP_OK) throw std::runtime_error("event create failed");
        sim.setPhaseHook([this, c](int ph) {
            gip_phase_ranges(ph);
            gip_event_record(c, ev[ph]);
            if (ph == 5) timed = true;
        });
    }
};

static thread_local std::string g_err;

static int fail(int code, const std::string &m) {
    g_err = m;
    return code;
}

static GIP_Parameters params_of(const gip_dns_config *c) {
    GIP_Parameters p;
    gip_dns_config d{};
    if (!c) {
        d.gamma = 1e-3;
        d.rho = 1.0;
        d.pcg_maxit = 100;
        c = &d;
    }
    p.setDouble("gamma", c->gamma);
    p.setDouble("rho", c->rho > 0 ? c->rho : 1.0);
    p.setInt("pcg_maxIt", c->pcg_maxit);
    p.setDouble("pcg_tol", c->pcg_tol);
    p.setBool("pcg_precond", true);
    p.setString("pcg_mode", c->pcg_mode == 1 ? "fused" : c->pcg_mode == 2 ? "mixed" : "reference");
    p.setDouble("pcg_innerTol", c->pcg_inner_tol > 0 ? c->pcg_inner_tol : 0.1);
    p.setInt("print", c->print);
    p.setInt("device", c->device);
    p.setInt("steps", 0);
    return p;
}

extern "C" {

const char *gip_dns_last_error(void) { return g_err.c_str(); }

int gip_dns_open(const char *root, const char *name, const gip_dns_config *cfg, gip_dns **out) {
    if (!name || !out) return fail(GIP_EINVAL, "name and out required");
    try {
        std::unique_ptr<gip_dns> s(new gip_dns(name));
        s->sim.setRoot(root ? root : "");
        s->sim.configure(params_of(cfg));
        s->sim.setUp();
        s->attach_timer();
        *out = s.release();
        return GIP_OK;
    } catch (const std::exception &e) {
        return fail(GIP_EINVAL, e.what());
    }
}

int gip_dns_from_cavity(const gip_cavity *cav, const gip_dns_config *cfg, gip_dns **out) {
    if (!cav || !out) return fail(GIP_EINVAL, "cavity and out required");
    try {
        std::unique_ptr<gip_dns> s(new gip_dns("memory"));
        s->sim.configure(params_of(cfg));
        s->sim.setUp(cav);
        s->attach_timer();
        *out = s.release();
        return GIP_OK;
    } catch (const std::exception &e) {
        return fail(GIP_EINVAL, e.what());
    }
}

int gip_dns_step(gip_dns *s, gip_dns_step_info *info) {
    if (!s) return fail(GIP_EINVAL, "null simulation");
    GIP_StepInfo st{};
    s->sim.step(&st);
    if (info) {
        info->time = st.time;
        info->dt = st.dt;
        info->massr = st.massr;
        info->pcg_relres = st.pcg_relres;
        info->step = st.step;
        info->pcg_iters = st.pcg_iters;
        info->pcg_status = st.pcg_status;
        info->pcg_corrections = st.pcg_corrections;
    }
    return GIP_OK;
}

int gip_dns_get_field(gip_dns *s, const char *field, double *host, int32_t n) {
    if (!s || !field || !host) return fail(GIP_EINVAL, "null argument");
    try {
        GIP_VectorHIP &v = s->sim.field(field);
        if (n != v.getSize()) return fail(GIP_EINVAL, "field length mismatch");
        std::memcpy(host, v.getHostPtr(), sizeof(double) * (size_t)n);
        return GIP_OK;
    } catch (const std::exception &e) {
        return fail(GIP_EINVAL, e.what());
    }
}

int gip_dns_set_field(gip_dns *s, const char *field, const double *host, int32_t n) {
    if (!s || !field || !host) return fail(GIP_EINVAL, "null argument");
    try {
        GIP_VectorHIP &v = s->sim.field(field);
        if (n != v.getSize()) return fail(GIP_EINVAL, "field length mismatch");
        double *h = v.getHostPtr();
        std::memcpy(h, host, sizeof(double) * (size_t)n);
        v.uploadHost();
        return GIP_OK;
    } catch (const std::exception &e) {
        return fail(GIP_EINVAL, e.what());
    }
}

int gip_dns_sizes(gip_dns *s, int32_t *inner, int32_t *all, int32_t *faces) {
    if (!s) return fail(GIP_EINVAL, "null simulation");
    if (inner) *inner = s->sim.innerSize();
    if (all) *all = s->sim.allSize();
    if (faces) *faces = s->sim.facesSize();
    return GIP_OK;
}

int gip_dns_sync(gip_dns *s) {
    if (!s) return fail(GIP_EINVAL, "null simulation");
    return gip_sync(s->sim.node().ctx());
}

int gip_dns_phase_ms(gip_dns *s, double *ms) {
    if (!s || !ms) return fail(GIP_EINVAL, "null argument");
    if (!s->timed) return fail(GIP_EINVAL, "no step taken yet");
    for (int i = 0; i < 5; ++i) {
        float t = 0.0f;
        int rc = gip_event_elapsed_ms(s->ev[i], s->ev[i + 1], &t);
        if (rc) return fail(rc, "event elapsed failed");
        ms[i] = t;
    }
    return GIP_OK;
}

int gip_dns_close(gip_dns *s) {
    delete s;
    return GIP_OK;
}

}  // extern "C"

// ---- several partitions on one device -----------------------------------------

struct gip_dns_parts {
    int n, device;
    std::unique_ptr<GIP_PartitionGroup> group;  // destroyed after the partitions
    std::vector<std::unique_ptr<HipDNS>> sims;
    ~gip_dns_parts() {
        sims.clear();
        group.reset();
    }
    // fn(rank) on one host thread per partition (device current on each);
    // the first failure aborts the group so that no partition waits forever
    int run(const std::function<void(int)> &fn) {
        std::vector<std::string> err((size_t)n);
        std::vector<std::thread> th;
        for (int r = 0; r < n; ++r)
            th.emplace_back([&, r] {
                try {
                    gip_check(gip_use_device(device), "gip_use_device");
                    fn(r);
                } catch (const std::exception &e) {
                    err[(size_t)r] = e.what();
                    group->abort();
                }
            });
        for (auto &t : th) t.join();
        for (int r = 0; r < n; ++r)
            if (!err[(size_t)r].empty()) return fail(GIP_EINVAL, "partition " + std::to_string(r) + ": " + err[(size_t)r]);
        return GIP_OK;
    }
};

extern "C" {

int gip_dns_open_parts(const char *root, const char *name, int32_t nparts, const gip_dns_config *cfg,
                       gip_dns_parts **out) {
    if (!name || !out || nparts < 1) return fail(GIP_EINVAL, "name, out and a positive partition count required");
    *out = nullptr;
    try {
        std::unique_ptr<gip_dns_parts> s(new gip_dns_parts());
        s->n = nparts;
        s->device = cfg ? cfg->device : 0;
        s->group.reset(new GIP_PartitionGroup(nparts, s->device));
        for (int r = 0; r < nparts; ++r) s->sims.emplace_back(new HipDNS(name));
        const GIP_Parameters prm = params_of(cfg);
        int rc = s->run([&](int r) {
            HipDNS &d = *s->sims[(size_t)r];
            d.node().setPartition(s->group.get(), r);
            d.setRoot(root ? root : "");
            d.configure(prm);
            d.setUp();
        });
        if (rc) return rc;
        *out = s.release();
        return GIP_OK;
    } catch (const std::exception &e) {
        return fail(GIP_EINVAL, e.what());
    }
}

int gip_dns_parts_step(gip_dns_parts *s, gip_dns_step_info *info) {
    if (!s) return fail(GIP_EINVAL, "null simulation");
    GIP_StepInfo st{};
    int rc = s->run([&](int r) {
        GIP_StepInfo mine{};
        s->sims[(size_t)r]->step(&mine);
        if (r == 0) st = mine;  // time, dt and massr are global: every rank holds them
    });
    if (rc) return rc;
    if (info) {
        info->time = st.time;
        info->dt = st.dt;
        info->massr = st.massr;
        info->pcg_relres = st.pcg_relres;
        info->step = st.step;
        info->pcg_iters = st.pcg_iters;
        info->pcg_status = st.pcg_status;
        info->pcg_corrections = st.pcg_corrections;
    }
    return GIP_OK;
}

int gip_dns_parts_sizes(gip_dns_parts *s, int32_t rank, int32_t *inner, int32_t *owned, int32_t *all,
                        int32_t *faces) {
    if (!s || rank < 0 || rank >= s->n) return fail(GIP_EINVAL, "bad simulation or rank");
    HipDNS &d = *s->sims[(size_t)rank];
    if (inner) *inner = d.innerSize();
    if (owned) *owned = d.ownedSize();
    if (all) *all = d.allSize();
    if (faces) *faces = d.facesSize();
    return GIP_OK;
}

int gip_dns_parts_get_field(gip_dns_parts *s, int32_t rank, const char *field, double *host, int32_t n) {
    if (!s || rank < 0 || rank >= s->n || !field || !host) return fail(GIP_EINVAL, "bad argument");
    try {
        HipDNS &d = *s->sims[(size_t)rank];
        GIP_VectorHIP &v = d.field(field);
        if (n != v.getSize()) return fail(GIP_EINVAL, "field length mismatch");
        if (int rc = gip_use_device(d.node().device())) return fail(rc, gip_last_error());
        std::memcpy(host, v.getHostPtr(), sizeof(double) * (size_t)n);
        return GIP_OK;
    } catch (const std::exception &e) {
        return fail(GIP_EINVAL, e.what());
    }
}

int gip_dns_parts_close(gip_dns_parts *s) {
    delete s;
    return GIP_OK;
}

}  // extern "C"
