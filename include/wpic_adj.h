/*
 * wpic_adj.h -- adjoint of the 2D periodic step loop on the device (FP64):
 * gradients of the final magnetic energy W_B = sum (Bx^2+By^2+Bz^2)/(8 pi)
 * (the energy_history sum of 2d/proj/weibel/app.f90:521-534) with respect to
 * the initial particles and fields, and from those with respect to the beam
 * parameters of wpic_ic_beams2d and the amplitude of a Bz perturbation.
 *
 * The reference has no adjoint; these entry points have no counterpart there.
 * The forward is the fused loop of wpic_step2d (periodic boundaries, an FP64
 * mode).  wpic_adj_forward2d keeps a device checkpoint every `every` steps;
 * wpic_adj_backward2d replays each segment from its checkpoint and runs the
 * hand-written adjoint kernels of the sort permutation, the implicit field
 * update (the CG solves transposed: the operator is symmetric) and the fused
 * push + Esirkepov deposit, step by step in reverse.
 */
#ifndef WPIC_ADJ_H
#define WPIC_ADJ_H

#include "wpic.h"

#ifdef __cplusplus
extern "C" {
#endif

/* relative CG stopping tolerance of field__fdtd_i (the reference's err =
 * 1e-6, 2d/common/field.f90:339); finite-difference checks of gradients need a
 * much tighter one so that the forward is a smooth function of its inputs */
int wpic_set_cg_tol2d(wpic_ctx *ctx, double err);

/* Bz(i, j) += amp * sin(2 pi mode (i - nxgs) / nx) on every node, ghosts
 * included (a seeded sinusoidal perturbation of the initial field) */
int wpic_perturb_bz2d(wpic_ctx *ctx, double amp, int mode);

/* nsteps of the fused loop from the current state, recording a checkpoint
 * every `every` steps; *wb = W_B of the final fields */
int wpic_adj_forward2d(wpic_ctx *ctx, int nsteps, int every, double *wb);

/* the adjoint of the recorded run, seeded with dL/dW_B = seed; afterwards
 * the adjoint of the initial state is held on the device and the context's
 * particles and fields are those at the end of the first checkpoint segment */
int wpic_adj_backward2d(wpic_ctx *ctx, double seed);

/* dL/d(udrift), dL/d(uth_e) of the wpic_ic_beams2d state the recorded run
 * started from, and dL/d(amp) of a wpic_perturb_bz2d(amp, mode) applied to it */
int wpic_adj_param_grads2d(wpic_ctx *ctx, double udrift, double uth_e, int bz_mode, double out[3]);

/* the adjoint of the initial state in the reference's layouts: lam_up like
 * up(6, np, nys:nye, nsp) (slot 6 zero), lam_uf like uf(6, ...) with the
 * ghost nodes zero (their adjoint is folded into the interior) */
int wpic_adj_download2d(wpic_ctx *ctx, double *lam_up, double *lam_uf, int np);

#ifdef __cplusplus
}
#endif
#endif
