/*
 * wpic3d.h -- C-ABI of the MI355X-native WumingPIC step loop in 3D, periodic
 * in x, y and z at one rank (the 3d/common modules in 3d/proj/weibel's loop).
 *
 * As wpic.h, one opaque context owns device-resident particles (structure of
 * arrays, every (j, k) column sorted by x cell), fields and solver arrays on
 * one GPU and one HIP stream.  Host arrays use the reference's 3D layouts:
 *
 *   up(7, np, nys:nye, nzs:nze, nsp)              real(8), id bit-cast in slot 7
 *   uf(6, nxgs-2:nxge+2, nys-2:nye+2, nzs-2:nze+2) real(8)  Bx By Bz Ex Ey Ez
 *   cumcnt(nxgs:nxge+1, nys:nye, nzs:nze, nsp)     integer
 *   np2(nys:nye, nzs:nze, nsp)                     integer
 *
 * Return codes, wpic_last_error() and the WPIC_MODE_* arithmetic modes are
 * those of wpic.h.  Calls on one context are not thread-safe.
 */
#ifndef WPIC3D_H
#define WPIC3D_H

#include "wpic.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wpic_ctx3d wpic_ctx3d;

/* Union of the arguments of particle__init (3d/common/particle.f90:18-49),
 * field__init (3d/common/field.f90:22-67), sort__init (3d/common/sort.f90:16-37)
 * and boundary_periodic__init (3d/common/boundary_periodic.f90:26-64) at one
 * rank: nys = nygs, nye = nyge, nzs = nzgs, nze = nzge. */
typedef struct {
  int ndim; /* 7 (3d/proj/weibel/app.f90:60,170-173) */
  int nsp;
  int nxgs, nxge, nygs, nyge, nzgs, nzge;
  double delx, delt, c, gfac;
  double q[WPIC_MAXSP], r[WPIC_MAXSP];
  int mode;           /* WPIC_MODE_* */
  long long capacity; /* particles per species held on the device */
} wpic_params3d;

/* context ----------------------------------------------------------------- */
int wpic_create3d(wpic_ctx3d **ctx, const wpic_params3d *p);
int wpic_destroy3d(wpic_ctx3d *ctx);
int wpic_set_mode3d(wpic_ctx3d *ctx, int mode);
int wpic_sync3d(wpic_ctx3d *ctx);

/* reference-layout transfers --------------------------------------------- */
/* up after sort__bucket with its cumcnt (3d/proj/weibel/app.f90:108) */
int wpic_upload_up3d(wpic_ctx3d *ctx, const double *up, const int *np2, const int *cumcnt, int np);
int wpic_download_up3d(wpic_ctx3d *ctx, double *up, int *np2, int *cumcnt, int np);
/* gp between particle__solv, the boundaries and sort__bucket (app.f90:103-108) */
int wpic_upload_gp3d(wpic_ctx3d *ctx, const double *gp, const int *np2, int np);
int wpic_download_gp3d(wpic_ctx3d *ctx, double *gp, int *np2, int np);
int wpic_upload_fields3d(wpic_ctx3d *ctx, const double *uf);
int wpic_download_fields3d(wpic_ctx3d *ctx, double *uf);
/* field.f90's saved uj (after set_boundary_curre) and df (after the last
 * set_boundary_dfield of field__fdtd_i), 3d/common/field.f90:99-124 */
int wpic_download_current3d(wpic_ctx3d *ctx, double *uj);
int wpic_download_dfield3d(wpic_ctx3d *ctx, double *df);
int wpic_upload_dfield3d(wpic_ctx3d *ctx, const double *df);

/* reference entry points ------------------------------------------------- */
/* particle__solv(gp, up, uf, cumcnt, nxs, nxe): 3d/common/particle.f90:52 */
int wpic_particle_solv3d(wpic_ctx3d *ctx);
/* particle__solv_vay(gp, up, uf, cumcnt, nxs, nxe): 3d/common/particle.f90:236
 * -- the Vay kick with particle__solv's staging, gather and move; FP64 modes */
int wpic_particle_solv_vay3d(wpic_ctx3d *ctx);
/* the kick of wpic_particle_solv3d and wpic_step3d: WPIC_PUSHER_BORIS (the
 * default) or WPIC_PUSHER_VAY; FP64 modes (moments keep mom_calc's Boris kick) */
int wpic_set_pusher3d(wpic_ctx3d *ctx, int pusher);
/* field__fdtd_i(uf, up, gp, cumcnt, nxs, nxe, bc_dfield, bc_curre, bc_phi):
 * 3d/common/field.f90:70, with boundary_periodic's procedures on the device */
int wpic_field_fdtd_i3d(wpic_ctx3d *ctx);
/* boundary_periodic__particle_x(gp, np2): 3d/common/boundary_periodic.f90:68 */
int wpic_bc_particle_x3d(wpic_ctx3d *ctx);
/* boundary_periodic__particle_yz(gp, np2): 3d/common/boundary_periodic.f90:104 */
int wpic_bc_particle_yz3d(wpic_ctx3d *ctx);
/* sort__bucket(up, gp, cumcnt, np2, nxs, nxe): 3d/common/sort.f90:40 */
int wpic_sort_bucket3d(wpic_ctx3d *ctx);
int wpic_cg_iters3d(wpic_ctx3d *ctx, int iters[3]);
/* relative CG stopping tolerance of field__fdtd_i (the reference's err =
 * 1e-6, 3d/common/field.f90:429); no counterpart in the reference, which
 * fixes it: for callers that need a tighter solve (closed-form checks,
 * finite differences) */
int wpic_set_cg_tol3d(wpic_ctx3d *ctx, double err);

/* the loop body of 3d/proj/weibel/app.f90:103-108, fused on the device */
int wpic_step3d(wpic_ctx3d *ctx, int nsteps);

/* diagnostics: energy sums {sum r(gamma-1), sum E^2/8pi, sum B^2/8pi}
 * (3d/proj/weibel/app.f90:510-560); Gauss residual G(nx, ny, nz) column-major
 * (no reference counterpart); particle counts per species */
int wpic_energy3d(wpic_ctx3d *ctx, double out[3]);
/* energy_history's columns of the 3D app (3d/proj/weibel/app.f90:509-577):
 * out[0..nsp-1] the kinetic energy of each species, out[nsp] E^2/8pi,
 * out[nsp+1] B^2/8pi */
int wpic_energy_species3d(wpic_ctx3d *ctx, double *out);
int wpic_gauss3d(wpic_ctx3d *ctx, double *G);
int wpic_counts3d(wpic_ctx3d *ctx, long long *counts);

/* Moments mom(7, nxgs-1:nxge+1, nys-1:nye+1, nzs-1:nze+1, nsp) real(8), as
 * the 2D calls of wpic.h: mom_calc__accl (3d/common/mom_calc.f90:49-216),
 * mom_calc__nvt (:219-333), boundary_periodic__mom
 * (3d/common/boundary_periodic.f90:1102-1235) and the three in one pass
 * (the output sequence of 3d/proj/weibel/app.f90:122-124). */
int wpic_mom_accl3d(wpic_ctx3d *ctx);
int wpic_mom_nvt3d(wpic_ctx3d *ctx);
int wpic_bc_mom3d(wpic_ctx3d *ctx);
int wpic_moments3d(wpic_ctx3d *ctx);
int wpic_download_mom3d(wpic_ctx3d *ctx, double *mom);
int wpic_upload_mom3d(wpic_ctx3d *ctx, const double *mom);

/* seeded two-species beam state, as wpic_ic_beams2d in 3D (positions uniform
 * in each cell in x, y and z; ions and electrons co-located) */
int wpic_ic_beams3d(wpic_ctx3d *ctx, unsigned long long seed, int n_ppc, double udrift,
                    double uth_e, double uth_i, double b0);

/* HIP-event timing on the context's stream: names "push", "deposit", "field",
 * "cg", "exchange", "sort" */
/* The reference 3D Weibel app's state (3d/proj/weibel/app.f90 set_initial_condition,
 * set_particle_ids): Bz = b0, n0 particles per cell per species on a regular x
 * lattice along each column, y and z uniform in the cell and shared by the
 * species; Maxwellian momenta with spread v_thi (ions) / v_the (electrons),
 * t_ani times that along z; ids -1, -2, ... column by column per species; the
 * OS-seeded draws replaced by counter-based draws from `seed`.  FP64 modes. */
int wpic_ic_weibel3d(wpic_ctx3d *ctx, unsigned long long seed, int n0, double v_thi, double v_the, double t_ani,
                     double b0);
int wpic_set_timing3d(wpic_ctx3d *ctx, int on);
int wpic_timing3d(wpic_ctx3d *ctx, const char *name, double *ms, long long *count);

#ifdef __cplusplus
}
#endif
#endif
