// C ABI of GIP_DNS instantiated with the HIP specification (include/gip_dns.h).
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>

#include "GIP_DNS.h"
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
