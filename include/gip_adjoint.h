/*
 * gip_adjoint.h — reverse-mode derivative of the fractional-step time
 * integration (GIP_DNS.h:98-414) on one device (part of libgip_hip.so).
 *
 * The forward map runs `steps` time steps of a lid-driven cavity with the
 * time steps frozen: the first forward run takes them from the CFL rule
 * (GIP_NonLinearMulticore.cpp:3-55), and later runs reuse that sequence.
 * The objective is the kinetic energy at the final step,
 *   J = 1/2 sum_i V_i (u_i^2 + v_i^2 + w_i^2)   over the inner cells.
 * Parameters are the lid ghost values of u (the per-face lid profile) and
 * the initial inner velocity field.
 *
 * gip_adjoint_backward() runs the hand-written adjoint of every phase in
 * reverse order:
 *   - transposed products of D, C, M, G, Lb, SO, SF and Ec;
 *   - an adjoint pressure solve on L with the same Jacobi-PCG, since L is
 *     symmetric, including its pin;
 *   - the adjoint of the convective operator C(mf) through both the
 *     velocity and the face fluxes;
 *   - the Adams–Bashforth history.
 * The forward state is checkpointed every `checkpoint_every` steps, and each
 * segment is recomputed from its checkpoint during the reverse sweep.
 * Pressure solves (forward and adjoint) run to `pcg_tol`.
 */
#ifndef GIP_ADJOINT_H
#define GIP_ADJOINT_H

#include <stdint.h>

#include "gip_cavity.h"
#include "gip_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gip_adjoint gip_adjoint;

typedef struct gip_adjoint_opts {
    int32_t steps;            /* time steps of the forward map                 */
    int32_t checkpoint_every; /* forward state kept every this many steps     */
    double gamma;             /* 1/Re for the CFL rule of the first forward run */
    double rho;
    double pcg_tol;           /* relative residual of every pressure solve    */
    int32_t pcg_maxit;
} gip_adjoint_opts;

/* cav: a hexahedral cavity with every operator and the lid ghost list
 * (gip_cavity_generate); it is read during the call only. */
int gip_adjoint_create(gip_ctx *ctx, const gip_cavity *cav, const gip_adjoint_opts *opts, gip_adjoint **out);
/* Forward run from the given parameters: lid (nx*ny values, index i+nx*j),
 * u0/v0/w0 (n_inner values each; NULL: zero).  The first call fixes the
 * time-step sequence; J receives the objective. */
int gip_adjoint_forward(gip_adjoint *a, const double *lid, const double *u0, const double *v0, const double *w0,
                        double *J);
/* Gradient of J at the parameters of the last forward run (any output may be NULL). */
int gip_adjoint_backward(gip_adjoint *a, double *grad_lid, double *grad_u0, double *grad_v0, double *grad_w0);
/* The frozen time steps (steps values) of the first forward run. */
int gip_adjoint_dts(gip_adjoint *a, double *dts);
int gip_adjoint_destroy(gip_adjoint *a);

#ifdef __cplusplus
}
#endif
#endif
