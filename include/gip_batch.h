/*
 * gip_batch.h — many independent lid-driven cavities advanced together on
 * one device (part of libgip_hip.so).
 *
 * Each instance is the reference's time step (GIP_DNS.h:98-414) with its own
 * operators, fields, time step, Reynolds number and pressure solve.  The
 * instances are concatenated into one layout:
 *   cells  [inner_0 | inner_1 | ... | ghosts_0 | ghosts_1 | ...]
 *   faces  [faces_0 | faces_1 | ...]
 * with each instance's inner and face ranges padded to whole 256-row chunks
 * (padding rows are empty), so that every workgroup-sized chunk belongs to
 * one instance.  Operators become block-diagonal, so one launch covers every
 * instance.  Per-instance scalars (dt, sigma = 1/dt, the PCG's alpha, beta
 * and convergence) sit in device arrays indexed by each chunk's instance.
 * Per-instance reductions (dots, massr, CFL) are deterministic.  Each chunk
 * writes a partial, and the chunk that completes an instance combines that
 * instance's partials in index order.
 *
 * A step runs as three captured phases: predictor + right-hand side + PCG
 * setup, `poll` PCG iterations, and projection + fluxes + CFL.  When
 * use_graph is set they are captured once as hipGraphs and replayed.  The
 * host only reads the number of still-running solves between iteration
 * launches.
 */
#ifndef GIP_BATCH_H
#define GIP_BATCH_H

#include <stdint.h>

#include "gip_cavity.h"
#include "gip_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gip_batch gip_batch;

typedef struct gip_batch_opts {
    int32_t pcg_maxit;   /* iteration limit of each solve                                 */
    double pcg_tol;      /* relative residual target of each instance's solve (> 0)     */
    int32_t precision;   /* GIP_PCG_FP64 or GIP_PCG_MIXED                                */
    double inner_tol;    /* MIXED: FP32 residual drop that triggers an FP64 correction    */
    int32_t poll;        /* iterations per launch group (0: 16)                           */
    int32_t use_graph;   /* capture the step phases as hipGraphs and replay them          */
    int32_t autotune;    /* choose each operator's SpMV layout with gip_mat_autotune      */
} gip_batch_opts;

typedef struct gip_batch_info {
    double time;        /* simulated time after the last step             */
    double dt;          /* time step used by the last step                */
    double massr;       /* ||Mres mf||_inf after the last step            */
    int32_t pcg_iters;  /* iterations of the last pressure solve          */
    int32_t pcg_status; /* 0 converged, GIP_ENOCONV, GIP_EBREAKDOWN       */
    int32_t n_inner, n_all, n_faces;
} gip_batch_info;

/* cavs[i] must hold every operator (GIP_OP_ALL); gamma[i] is instance i's
 * 1/Re for the CFL limit (its D already carries it).  The cavities are read
 * during the call only. */
int gip_batch_create(gip_ctx *ctx, int32_t ninst, const gip_cavity *const *cavs, const double *gamma, double rho,
                     const gip_batch_opts *opts, gip_batch **out);
/* One time step of every instance. */
int gip_batch_step(gip_batch *b);
/* Device time (ms, HIP events on the library stream) of the last step's three
 * phases: ms[0] predictor and Poisson right-hand side, ms[1] the PCG
 * iterations including the host's reads of the running count between poll
 * groups, ms[2] correction, mass fluxes and CFL.  No counterpart in the
 * reference (its GIP_Timer brackets whole phases on the host,
 * GIP_Timer.h:6-26); the batch has no per-instance phase boundary. */
int gip_batch_phase_ms(gip_batch *b, float *ms);
int gip_batch_instance(gip_batch *b, int32_t inst, gip_batch_info *info);
/* Instance field in its own layout ("u","v","w","p": n_all values; "mf": n_faces). */
int gip_batch_get_field(gip_batch *b, int32_t inst, const char *field, double *host, int32_t n);
/* Chosen SpMV layout per operator in the order D, Ec, L, Lb, Mx, My, Mz,
 * Gx, Gy, Gz, SFx, SFy, SFz, SO, Mres: format, and in `lanes` the CSR lanes
 * per row, or for sliced ELL its rows per slice (32, 64 or 128). */
int gip_batch_layouts(gip_batch *b, int32_t *format, int32_t *lanes);
int gip_batch_destroy(gip_batch *b);

#ifdef __cplusplus
}
#endif
#endif
