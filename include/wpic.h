/*
 * wpic.h -- C-ABI of the MI355X-native WumingPIC step loop (2D).
 *
 * One opaque context per simulation owns device-resident particles (structure
 * of arrays, sorted by cell), fields and solver work arrays on one GPU and one
 * HIP stream.  Host arrays cross this boundary only in the reference's own
 * Fortran layouts, so the reference's module entry points can be rebound onto
 * these calls (see wumingpic_amd/fortran/ for the ISO_C_BINDING shim):
 *
 *   up(ndim, np, nys:nye, nsp)        real(8), id bit-cast in slot 6
 *   uf(6, nxgs-2:nxge+2, nys-2:nye+2) real(8)  Bx By Bz Ex Ey Ez
 *   cumcnt(nxgs:nxge+1, nys:nye, nsp) integer
 *   np2(nys:nye, nsp)                 integer
 *
 * Every call returns 0 on success and a negative WPIC_E* code on failure;
 * wpic_last_error() gives the message (the reference's own text where the
 * reference stops with one).  Calls on one context are not thread-safe.
 */
#ifndef WPIC_H
#define WPIC_H

#ifdef __cplusplus
extern "C" {
#endif

#define WPIC_MAXSP 4

/* boundary kinds */
#define WPIC_BC_PERIODIC 0 /* 2d/common/boundary_periodic.f90 */
#define WPIC_BC_WALLX 1    /* 2d/proj/reconnection/boundary_reconnection.f90 */
/* 2d/proj/shock/boundary_shock.f90 at the active edge nxe (wpic_set_nxe2d,
 * wpic_relocate2d; nxe = nxge unless set): the reconnection's left wall, the
 * injection edge at the right (boundary_shock__injection as the step's x
 * boundary, zero field and phi ghosts beyond nxe); FP64 modes */
#define WPIC_BC_SHOCK 2

/* particle kicks of the FP64 modes (wpic_set_pusher2d) */
#define WPIC_PUSHER_BORIS 0 /* particle__solv: 2d/common/particle.f90:48-179 */
#define WPIC_PUSHER_VAY 1   /* particle__solv_vay: 2d/common/particle.f90:182-315 */

/* arithmetic modes of the FP64 path */
#define WPIC_MODE_FAST 0   /* tiled deposit with atomics, tree dot products */
#define WPIC_MODE_PARITY 1 /* reference summation order: bit-exact vs 1 thread */
/* Mixed precision, fixed at wpic_create2d / wpic_create3d: particles stored
 * as an int32 x cell plus FP32 in-cell offsets and FP32 momenta (32 B with the
 * id in 2D, 36 B in 3D), push and deposit in FP32 (the tile's current
 * accumulated in FP32 in 2D, FP64 in 3D), FP64 global current sum and FP64
 * field solve.  Runs the fused loop
 * (wpic_step2d / wpic_step3d) with the up/fields transfers, diagnostics and
 * seeded state; the single-routine entry points need an FP64 mode. */
#define WPIC_MODE_MIXED 2

/* error codes */
#define WPIC_OK 0
#define WPIC_EINVAL -1   /* bad argument or call order */
#define WPIC_EHIP -2     /* HIP runtime error */
#define WPIC_ECG -3      /* "stop at cgm after ite_max" (field.f90:427-430) */
#define WPIC_EMEM -4     /* "memory over (np2 > np)" (boundary_periodic.f90:231-234) */
#define WPIC_EDOMAIN -5  /* particle outside the sortable domain */

typedef struct wpic_ctx wpic_ctx;

/* Union of the arguments of particle__init (2d/common/particle.f90:18-23),
 * field__init (2d/common/field.f90:22-29), sort__init (2d/common/sort.f90:16)
 * and boundary_*__init (2d/common/boundary_periodic.f90:26-33) at one rank:
 * nys = nygs, nye = nyge. */
typedef struct {
  int ndim; /* 6 */
  int nsp;
  int nxgs, nxge, nygs, nyge;
  double delx, delt, c, gfac;
  double q[WPIC_MAXSP], r[WPIC_MAXSP];
  int bc;   /* WPIC_BC_* */
  int mode; /* WPIC_MODE_* */
  long long capacity; /* particles per species held on the device */
} wpic_params2d;

/* context ----------------------------------------------------------------- */
int wpic_create2d(wpic_ctx **ctx, const wpic_params2d *p);
int wpic_destroy(wpic_ctx *ctx);
const char *wpic_last_error(void);
int wpic_set_mode(wpic_ctx *ctx, int mode);
int wpic_sync(wpic_ctx *ctx);

/* reference-layout transfers --------------------------------------------- */
/* up: sorted particles with their cumcnt (the app's `up` after sort__bucket,
 * 2d/proj/weibel/app.f90:107); only slots 1..np2(j,isp) of each row are read */
int wpic_upload_up2d(wpic_ctx *ctx, const double *up, const int *np2, const int *cumcnt, int np);
int wpic_download_up2d(wpic_ctx *ctx, double *up, int *np2, int *cumcnt, int np);
/* gp: the particle array handed between particle__solv, bc__particle_x/y and
 * sort__bucket (2d/proj/weibel/app.f90:102-107).  Uploading gp starts the
 * state "rows ready for sort__bucket" (the restart path, app.f90:351-352). */
int wpic_upload_gp2d(wpic_ctx *ctx, const double *gp, const int *np2, int np);
int wpic_download_gp2d(wpic_ctx *ctx, double *gp, int *np2, int np);
int wpic_upload_fields2d(wpic_ctx *ctx, const double *uf);
int wpic_download_fields2d(wpic_ctx *ctx, double *uf);
/* field.f90's saved module arrays: uj(3,...) after set_boundary_curre and
 * df(6,...) after the last set_boundary_dfield of field__fdtd_i */
int wpic_download_current2d(wpic_ctx *ctx, double *uj);
int wpic_download_dfield2d(wpic_ctx *ctx, double *df);
int wpic_upload_dfield2d(wpic_ctx *ctx, const double *df);

/* reference entry points ------------------------------------------------- */
/* particle__solv(gp, up, uf, cumcnt, nxs, nxe): 2d/common/particle.f90:48 */
int wpic_particle_solv2d(wpic_ctx *ctx);
/* particle__solv_vay(gp, up, uf, cumcnt, nxs, nxe): 2d/common/particle.f90:182
 * -- the Vay kick (Phys. Plasmas 15, 056701 (2008)) with particle__solv's
 * field staging, gather and move; FP64 modes. */
int wpic_particle_solv_vay2d(wpic_ctx *ctx);
/* The kick that wpic_particle_solv2d and wpic_step2d use from now on:
 * WPIC_PUSHER_BORIS (the default, the reference apps' choice) or
 * WPIC_PUSHER_VAY; FP64 modes.  Moments keep mom_calc__accl's Boris kick, as
 * the reference's mom_calc.f90 does. */
int wpic_set_pusher2d(wpic_ctx *ctx, int pusher);
/* field__fdtd_i(uf, up, gp, cumcnt, nxs, nxe, bc_dfield, bc_curre, bc_phi):
 * 2d/common/field.f90:66.  The boundary procedures are applied on the device
 * according to the context's boundary kind. */
int wpic_field_fdtd_i2d(wpic_ctx *ctx);
/* boundary_periodic__particle_x(gp, np2): boundary_periodic.f90:61, or
 * boundary_reconnection__particle_x(gp, np2, nxs, nxe): :61 */
int wpic_bc_particle_x2d(wpic_ctx *ctx);
/* boundary_shock__injection(up, np2, nxs, nxe, u0):
 * 2d/proj/shock/boundary_shock.f90:255 on gp, with nxs = nxgs, nxe = nxge;
 * also sets the u0 the step uses.  WPIC_BC_SHOCK only. */
int wpic_bc_injection2d(wpic_ctx *ctx, double u0);
/* the injection 4-velocity u0 (< 0: inflow from the right, the shock app's
 * -|u_inject|, 2d/proj/shock/app.f90:319) of the step's injection edge */
int wpic_set_injection2d(wpic_ctx *ctx, double u0);
/* inject (2d/proj/shock/app.f90:697-856) on the sorted state at the active
 * edge nxe: nl[r] particles per species join row r at x = nxe dx + (m-1/2)/nl
 * |v0| dt moved by (v0 + ux) dt, y uniform in the row (shared by the
 * species), Maxwellian momenta (v_thi ions, v_the electrons) boosted by the
 * inflow profile vprofile(x) = v0/2 (1 + tanh((x - l_damp - nxgs dx) /
 * (l_damp/10))), ids continuing each species' count; they close cell nxe-1
 * (cumcnt(nxe) and every entry after it grow by nl[r]); the upstream B and E
 * are set at nxe-1 and nxe.  The split of the injected number over rows
 * (random in the reference) is the caller's; `it` keys the draws of this
 * call.  nl must hold ny = nyge - nygs + 1 entries (one per row); the call
 * reads that many. */
/* The shock app's initial state (2d/proj/shock/app.f90:405-470) over the
 * active box nxs..nxe (wpic_set_nxe2d; nxe = nxge unless set):
 * n0 (nxe - nxs - 1) particles per row per species on a regular x
 * lattice over [nxs, nxe) dx, y uniform in the row and shared, Maxwellian
 * momenta (v_thi, v_the) boosted by vprofile(x); B = b0 (cos theta,
 * sin theta cos phi, sin theta sin phi) and E = -vprofile(i dx) x B / c at
 * every node; sorted as the reference sorts them; counter-based draws. */
int wpic_ic_shock2d(wpic_ctx *ctx, unsigned long long seed, int n0, double v0, double v_thi, double v_the,
                    double l_damp, double b0, double theta_bn, double phi_bn);
int wpic_inject2d(wpic_ctx *ctx, unsigned long long seed, int it, const int *nl, double v0, double v_thi,
                  double v_the, double l_damp, double b0, double theta_bn, double phi_bn);
/* The shock's expanding box.  The app keeps the active right edge nxe
 * (app.f90:276-277: nxe = nxgs + n_x_ini) and passes nxs, nxe to every
 * routine; here the context holds it: the injection edge (xend = nxe dx +
 * v0 dt), the shock's particle_x wall at nxe-1, the field solve (its right
 * hand side, CG, E update and boundaries over nodes nxs..nxe, zeros at
 * nxe+1, uf += df over nxs-2..nxe+2), the initial state and inject follow it.
 * The cells beyond nxe hold no particles; their cumcnt entries carry the
 * row's total.  WPIC_BC_SHOCK only; nxgs+3 <= nxe <= nxge. */
int wpic_set_nxe2d(wpic_ctx *ctx, int nxe);
int wpic_get_nxe2d(wpic_ctx *ctx, int *nxe);
/* relocate (2d/proj/shock/app.f90:615-692) on the sorted state: nothing when
 * nxe = nxge; else nxe grows by one and each row of each species gains n0
 * particles at its end on the lattice x = (nxe-1) dx + (m-1/2)/n0 dx, y
 * uniform in the row (shared by the species), momenta as inject's, ids
 * -(m + (j - nygs) n0 + nptotal) continuing each species' count; cumcnt(nxe)
 * = cumcnt(nxe-1) + n0 as the reference sets it; the upstream B and E are set
 * at nxe-1 and nxe.  `it` keys the draws of this call. */
int wpic_relocate2d(wpic_ctx *ctx, unsigned long long seed, int it, int n0, double v0, double v_thi, double v_the,
                    double l_damp, double b0, double theta_bn, double phi_bn);
/* boundary_periodic__particle_y(gp, np2): boundary_periodic.f90:99 */
int wpic_bc_particle_y2d(wpic_ctx *ctx);
/* sort__bucket(up, gp, cumcnt, np2, nxs, nxe): 2d/common/sort.f90:36 */
int wpic_sort_bucket2d(wpic_ctx *ctx);
/* CG iteration counts of the last field__fdtd_i, one per component */
int wpic_cg_iters2d(wpic_ctx *ctx, int iters[3]);

/* the whole step loop in the app's order (2d/proj/weibel/app.f90:102-107 or
 * 2d/proj/reconnection/app.f90:101-106), fused on the device */
int wpic_step2d(wpic_ctx *ctx, int nsteps);

/* diagnostics -------------------------------------------------------------- */
/* energy_history sums (2d/proj/weibel/app.f90:503-534):
 * out = {sum r(gamma-1), sum E^2/8pi, sum B^2/8pi} over interior cells */
int wpic_energy2d(wpic_ctx *ctx, double out[3]);
/* Gauss residual div E - 4 pi rho at cell centres, rho with the deposit's
 * quadratic shape; G(nx, ny) column-major.  No reference counterpart. */
int wpic_gauss2d(wpic_ctx *ctx, double *G);
/* wpic_gauss2d without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The charge plane
 * and the residual live in a fixed device array of the library (5 * 2^20
 * doubles, 40 MiB, held on every GPU the library runs on until the process
 * ends), so (nx+4)*(ny+4) + nx*ny may not exceed 5 * 2^20: a larger grid is
 * refused with WPIC_EINVAL and a message.  Every call on one GPU uses that
 * GPU's array: calls that overlap on one GPU (from several threads or
 * contexts) race. */
int wpic_gauss2d_noalloc(wpic_ctx *ctx, double *G);
/* particle counts per species on the device */
int wpic_counts2d(wpic_ctx *ctx, long long *counts);

/* Moments mom(7, nxgs-1:nxge+1, nys-1:nye+1, nsp) real(8): N, Vx, Vy, Vz,
 * Txx, Tyy, Tzz at the nodes (i+1/2, j+1/2), held on the device.
 * mom_calc__init (2d/common/mom_calc.f90:18-45) takes particle__init's
 * arguments, so the context carries its state.
 * wpic_mom_accl2d:  mom_calc__accl (mom_calc.f90:48-164): gp := up with the
 *                   momenta kicked over delt/2, positions kept, slot-aligned.
 * wpic_mom_nvt2d:   mom_calc__nvt (mom_calc.f90:167-249) of gp.  Bit-identical
 *                   to the reference at one thread in WPIC_MODE_PARITY; LDS
 *                   and HBM atomics (summation order free) in WPIC_MODE_FAST.
 * wpic_bc_mom2d:    boundary_periodic__mom (boundary_periodic.f90:571-636) or
 *                   boundary_reconnection__mom
 *                   (2d/proj/reconnection/boundary_reconnection.f90:582-650),
 *                   by the context's boundary kind.
 * wpic_moments2d:   the app's output sequence accl, nvt, bc__mom
 *                   (2d/proj/weibel/app.f90:121-123) in one pass over up
 *                   (FAST: the kick fused into the deposit); gp is left
 *                   undefined, as mom_calc__accl leaves it overwritten.
 *                   WPIC_MODE_MIXED: the same sequence in FP32 from the mixed
 *                   record (FP32 LDS node tiles, FP64 global sum); the
 *                   particles are not touched.
 * Particles whose lower node falls outside their row or the grid give
 * WPIC_EDOMAIN.  The single routines accl, nvt and bc need an FP64 mode. */
int wpic_mom_accl2d(wpic_ctx *ctx);
int wpic_mom_nvt2d(wpic_ctx *ctx);
int wpic_bc_mom2d(wpic_ctx *ctx);
int wpic_moments2d(wpic_ctx *ctx);
int wpic_download_mom2d(wpic_ctx *ctx, double *mom);
int wpic_upload_mom2d(wpic_ctx *ctx, const double *mom);

/* seeded initial state ----------------------------------------------------- */
/* Counter-based seeded two-species beam state: n_ppc particles per cell per
 * species at uniformly random in-cell positions, ions and electrons
 * co-located, electrons u_x = +/-udrift (alternating by slot) + uth_e N(0,1),
 * u_y, u_z = uth_e N(0,1); ions uth_i N(0,1); ids -1,-2,... per species as
 * 2d/proj/weibel/app.f90:442-474 numbers them; uf = 0 except Bz = b0. */
int wpic_ic_beams2d(wpic_ctx *ctx, unsigned long long seed, int n_ppc, double udrift,
                    double uth_e, double uth_i, double b0);

/* The reference apps' own initial states (FP64 modes), with its OS-seeded
 * random draws replaced by counter-based draws from `seed`:
 * wpic_ic_weibel2d: 2d/proj/weibel/app.f90:366-438 -- Bz = b0 everywhere,
 *   n0 particles per cell per species on a regular x lattice along each row,
 *   y uniform in the row, shared by both species; Maxwellian momenta with
 *   spread v_thi (ions) / v_the (electrons), t_ani times that along z.
 * wpic_ic_harris2d: 2d/proj/reconnection/app.f90:260-452 (walls in x) --
 *   the Harris sheet b0 tanh((x-x0)/lcs) along y with the X-point
 *   perturbation e1 (the reference hard-codes 0.12), nbg background and
 *   ncs sheet particles per cell, drift carrying jz, temperatures vti, vte,
 *   Te/Ti = rtemp; lcs in cells (the reference's lcs*c/wpi); the loaded
 *   particles are sorted as the reference sorts them.
 * Ids are -1, -2, ... row by row per species (app.f90:455-470). */
int wpic_ic_weibel2d(wpic_ctx *ctx, unsigned long long seed, int n0, double v_thi, double v_the, double t_ani,
                     double b0);
int wpic_ic_harris2d(wpic_ctx *ctx, unsigned long long seed, int nbg, int ncs, double lcs, double x0, double y0,
                     double b0, double e1, double rtemp, double vti, double vte);

/* timing of named kernels with HIP events on the context's stream ---------- */
int wpic_set_timing(wpic_ctx *ctx, int on);
/* name: "push_deposit", "sort", "exchange", "field", "cg"; returns the summed
 * milliseconds and the number of launches since timing was switched on */
int wpic_timing(wpic_ctx *ctx, const char *name, double *ms, long long *count);

#ifdef __cplusplus
}
#endif
#endif
