/*
 * gip_dns.h — C ABI of the templated integrator instantiated with the HIP
 * specification: GIP_DNS<GIP_MatrixHIP, GIP_VectorHIP, GIP_ArchHIP,
 * GIP_NonLinearHIP>.  The reference drives its integrator from
 * DrivenCavity.cpp:21-43 (GIP_DNS(name); setUp(); integrate()); this entry
 * set exposes the same object step by step, so that tests and benchmarks
 * can compare each step with the CPU path.
 *
 * Parameters the reference hardcodes (GIP_DNS.h:132-137, :180-185) are
 * passed in gip_dns_config.  pcg_mode: 0 = the reference operation sequence
 * of GIP_LinearSolver.h:54-116 through the HIP Vector/Matrix methods,
 * 1 = fused device-resident FP64 PCG, 2 = mixed-precision PCG.
 * Returns 0 or a negative gip_status (include/gip_hip.h).
 */
#ifndef GIP_DNS_H
#define GIP_DNS_H

#include <stdint.h>

#include "gip_cavity.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gip_dns gip_dns;

typedef struct gip_dns_config {
    double gamma;       /* 1/Re used by the CFL diffusive limit (reference: 1e-3) */
    double rho;         /* density (reference: 1)                                   */
    int32_t pcg_maxit;  /* reference: 100                                           */
    double pcg_tol;     /* <= 0: fixed pcg_maxit iterations (reference behaviour)   */
    int32_t pcg_mode;   /* 0 reference sequence, 1 fused FP64, 2 mixed precision    */
    double pcg_inner_tol;
    int32_t print;      /* 0 silent, 1 reference line per step, 2 full precision    */
    int32_t device;
} gip_dns_config;

typedef struct gip_dns_step_info {
    double time, dt, massr, pcg_relres;
    int32_t step, pcg_iters, pcg_status, pcg_corrections;
} gip_dns_step_info;

/* One-partition case from <root>/InputFiles/... (<name>_1p-0 files). */
int gip_dns_open(const char *root, const char *name, const gip_dns_config *cfg, gip_dns **out);
/* Case from an in-memory cavity (gip_cavity_generate / gip_cavity_read). */
int gip_dns_from_cavity(const gip_cavity *cav, const gip_dns_config *cfg, gip_dns **out);
int gip_dns_step(gip_dns *s, gip_dns_step_info *info);
/* Device time of the five phases of the last step, in ms: predictor,
 * Poisson (right-hand side, PCG, ghost shuffle), velocity correction, mass
 * fluxes, CFL (GIP_DNS.h:308-414), from events recorded on the stream at
 * the phase boundaries (no synchronisation inside the step).  Not in the
 * reference, which has no per-phase timer; GIP_Timer (GIP_Timer.h:6-26)
 * times whole channels. */
int gip_dns_phase_ms(gip_dns *s, double *ms);
/* Copy a field ("u","v","w","p","mf","p0","ax0","ay0","az0","rhs") to host memory; n = its length. */
int gip_dns_get_field(gip_dns *s, const char *field, double *host, int32_t n);
int gip_dns_set_field(gip_dns *s, const char *field, const double *host, int32_t n);
int gip_dns_sizes(gip_dns *s, int32_t *inner, int32_t *all, int32_t *faces);
int gip_dns_sync(gip_dns *s);
int gip_dns_close(gip_dns *s);
const char *gip_dns_last_error(void);

/* Several partitions of one case (<name>_<nparts>p-<rank> files, the
 * decomposition the reference's driver runs with one MPI rank per partition,
 * DrivenCavity.cpp:30-33) resident on one device, each stepped by its own
 * host thread with its own context and stream.  Halo refreshes
 * (GIP_Topo.cpp:124-181) are device pack/unpack kernels among the
 * partitions; dot products, norms and the CFL bound are combined over the
 * partitions in rank order (MPI_Allreduce in the reference).  pcg_mode must
 * be 0 (the reference sequence).  A step returns rank 0's numbers (time,
 * dt and massr are global); fields are per partition (inner cells, boundary
 * ghosts, halo). */
typedef struct gip_dns_parts gip_dns_parts;
int gip_dns_open_parts(const char *root, const char *name, int32_t nparts, const gip_dns_config *cfg,
                       gip_dns_parts **out);
int gip_dns_parts_step(gip_dns_parts *s, gip_dns_step_info *info);
int gip_dns_parts_sizes(gip_dns_parts *s, int32_t rank, int32_t *inner, int32_t *owned, int32_t *all,
                        int32_t *faces);
int gip_dns_parts_get_field(gip_dns_parts *s, int32_t rank, const char *field, double *host, int32_t n);
int gip_dns_parts_close(gip_dns_parts *s);

#ifdef __cplusplus
}
#endif
#endif
