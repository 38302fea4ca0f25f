/*
 * wpic_oracle.h -- CPU restatement of WumingPIC's per-timestep particle loop.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product path (the HIP library,
 * the Fortran shim, bench.py's timed region) may include, link or call this.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
 * as the checker and as the timed CPU baseline.
 *
 * Every function restates one reference routine, in the reference's own
 * operation order (Fortran left-to-right evaluation, parentheses respected,
 * no FMA contraction: build with -ffp-contract=off), on arrays in the
 * reference's own Fortran (column-major) layout:
 *
 *   up(ndim, np, nys:nye, nsp)        particles, padded rows of capacity np
 *   uf(6, nxgs-2:nxge+2, nys-2:nye+2) fields Bx By Bz Ex Ey Ez
 *   uj(3, nxgs-2:nxge+2, nys-2:nye+2) current Jx Jy Jz
 *   cumcnt(nxgs:nxge+1, nys:nye, nsp) in-row cell offsets, np2(nys:nye, nsp)
 *
 * Pinning status of each function (what checks it, see tests/):
 *   or2d_particle_solv    bit-exact vs the reference particle.f90 built with
 *                         amdflang (oracle/_ref), plus closed forms.
 *   or2d_particle_solv_vay bit-exact vs the reference particle__solv_vay
 *                         built with amdflang (oracle/_ref, golden fixture),
 *                         plus closed forms: B = 0 gives u + fac2 E, pure B
 *                         keeps |u|, the E x B drift velocity is stationary.
 *   OR_BC_SHOCK parts (or2d_bc_injection; the right-edge zeros of bc_dfield
 *                         and bc_phi): the left edge is the reconnection wall
 *                         (pinned with it); the injection is an involution and
 *                         with u0 = 0 a static mirror at nxe; the zero ghosts
 *                         are checked after a field solve.  No reference build
 *                         (boundary_shock.f90 needs MPI): parity unpinned
 *                         beyond those properties.
 *   or2d_sort_bucket      bit-exact vs the reference sort.f90 (oracle/_ref),
 *                         plus brute-force histogram / stable permutation.
 *   or2d_ele_cur          discrete continuity (rho^{n+1}-rho^n+dt div J = 0),
 *                         single-particle brute force, J=0 at rest.
 *   or2d_field_fdtd_i     CG solution vs an exact FFT solve of the periodic
 *                         operator; vacuum wave energy; div B / Gauss law.
 *   or2d_bc_*             brute-force periodic indexing on tiny grids, the
 *                         round-down wrap edge case, wall sign tables.
 *   or2d_mom_accl, or3d_mom_accl, or2d_mom_nvt, or3d_mom_nvt
 *                         bit-exact vs the reference mom_calc.f90 (2D and 3D)
 *                         built with amdflang (oracle/_ref); nvt also against
 *                         a brute-force periodic CIC deposit and sum rules.
 *   or2d_bc_mom, or3d_bc_mom
 *                         the folded moments equal a deposit with periodic
 *                         (or mirrored-wall) node indices computed directly;
 *                         interior totals conserve the ghost sums.
 *   or2d_bc_particle_y    particle multiset / id conservation, row membership;
 *                         the exact in-row order after hole filling is
 *                         "parity unpinned" (only a reference build with MPI
 *                         could pin it, and MPI is absent from the image).
 */
#ifndef WPIC_ORACLE_H
#define WPIC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXSP 4

/* boundary kinds: boundary_periodic (2d/common/boundary_periodic.f90) and
 * boundary_reconnection (2d/proj/reconnection/boundary_reconnection.f90) */
enum { OR_BC_PERIODIC = 0, OR_BC_WALLX = 1, OR_BC_SHOCK = 2 };

typedef struct {
  int ndim, np, nsp;
  int nxgs, nxge, nygs, nyge, nys, nye;
  double delx, delt, c, gfac;
  double q[OR_MAXSP], r[OR_MAXSP];
  int bc;
  double u0; /* OR_BC_SHOCK: the injection 4-velocity (negative: inflow from the right) */
  int nxe;   /* OR_BC_SHOCK: the active right edge while the box expands (0: nxge) */
} or2d_params;

/* persistent module state of field.f90 (df, gkl, uj are `save`d) */
typedef struct {
  double *df;  /* df(6, nxgs-2:nxge+2, nys-2:nye+2) */
  double *gkl; /* gkl(3, nxgs:nxge, nys:nye) */
  double *uj;  /* uj(3, nxgs-2:nxge+2, nys-2:nye+2) */
  int cg_iters[3];
} or2d_fstate;

int or2d_fstate_init(const or2d_params *p, or2d_fstate *st);
void or2d_fstate_free(or2d_fstate *st);

void or_set_threads(int n);
/* cgm's relative stopping tolerance (default 1e-6, field.f90:339) */
void or_set_cg_err(double e);
double or_cg_err(void);
const char *or_last_error(void);

/* field.f90:22-63 */
void or2d_field_consts(const or2d_params *p, double f[5]);
/* particle.f90:48-179 */
void or2d_particle_solv(const or2d_params *p, double *gp, const double *up,
                        const double *uf, const int *cumcnt);
void or2d_particle_solv_vay(const or2d_params *p, double *gp, const double *up, const double *uf,
                            const int *cumcnt);
/* boundary_shock__injection: 2d/proj/shock/boundary_shock.f90:255-297 */
void or2d_bc_injection(const or2d_params *p, double *up, const int *np2);
/* field.f90:189-316 */
void or2d_ele_cur(const or2d_params *p, double *uj, const double *up,
                  const double *gp, const int *cumcnt);
/* boundary_periodic.f90:357-508 / boundary_reconnection.f90:364-502 */
void or2d_bc_curre(const or2d_params *p, double *uj);
/* boundary_periodic.f90:251-354 / boundary_reconnection.f90:254-361 */
void or2d_bc_dfield(const or2d_params *p, double *df);
/* boundary_periodic.f90:511-568 / boundary_reconnection.f90:505-579 */
void or2d_bc_phi(const or2d_params *p, double *phi, int l);
/* field.f90:66-186 (calls ele_cur, bc_curre, cgm, bc_dfield) */
int or2d_field_fdtd_i(const or2d_params *p, or2d_fstate *st, double *uf,
                      const double *up, const double *gp, const int *cumcnt);
/* field.f90:319-461 on st->df, st->gkl */
int or2d_cgm(const or2d_params *p, or2d_fstate *st);
/* field.f90:124-184 from the folded current already in st->uj */
int or2d_field_from_current(const or2d_params *p, or2d_fstate *st, double *uf);
/* boundary_periodic.f90:61-96 / boundary_reconnection.f90:61-99 */
void or2d_bc_particle_x(const or2d_params *p, double *up, const int *np2);
/* boundary_periodic.f90:99-248 (identical in boundary_reconnection.f90) */
int or2d_bc_particle_y(const or2d_params *p, double *up, int *np2);
/* sort.f90:36-82 */
void or2d_sort_bucket(const or2d_params *p, double *gp, const double *up,
                      int *cumcnt, const int *np2);
/* one step in the app's order: 2d/proj/weibel/app.f90:102-107 (periodic) or
 * 2d/proj/reconnection/app.f90:101-106 (walls) */
int or2d_step(const or2d_params *p, or2d_fstate *st, double *up, double *gp,
              double *uf, int *cumcnt, int *np2, int nsteps);
/* energy_history sums: 2d/proj/weibel/app.f90:503-534.
 * out = {kinetic(isp=1..nsp summed), E^2/8pi, B^2/8pi} */
void or2d_energy(const or2d_params *p, const double *up, const double *uf,
                 const int *np2, double out[3]);
/* Gauss-law residual G(i,j) = div E - 4 pi rho at cell centres, rho with the
 * same quadratic shape as the deposit (s0 x s0).  G(nxgs:nxge, nys:nye). */
void or2d_gauss(const or2d_params *p, double *G, const double *uf,
                const double *up, const int *np2);
/* moments: mom(7, nxgs-1:nxge+1, nys-1:nye+1, nsp) = N, Vx, Vy, Vz, Txx,
 * Tyy, Tzz.  mom_calc__accl (2d/common/mom_calc.f90:48-164), mom_calc__nvt
 * (:167-249), boundary_periodic__mom (boundary_periodic.f90:571-636) or
 * boundary_reconnection__mom (boundary_reconnection.f90:582-650) by p->bc. */
void or2d_mom_accl(const or2d_params *p, double *gp, const double *up, const double *uf, const int *cumcnt);
void or2d_mom_nvt(const or2d_params *p, double *mom, const double *up, const int *np2);
void or2d_bc_mom(const or2d_params *p, double *mom);

/* ---------------------------------------------------------------------------
 * 3D (the 3d/common modules), periodic in x, y and z at one rank.
 *   up(ndim=7, np, nys:nye, nzs:nze, nsp), uf(6, nxgs-2:nxge+2, nys-2:nye+2,
 *   nzs-2:nze+2), cumcnt(nxgs:nxge+1, nys:nye, nzs:nze, nsp), np2(nys:nye,
 *   nzs:nze, nsp).
 * Pinning: or3d_particle_solv, or3d_particle_solv_vay and or3d_sort_bucket are bit-exact against the
 * reference's own 3d/common/particle.f90 and sort.f90 (oracle/_ref); the
 * deposit by discrete continuity; the CG by an FFT solve; the step by Gauss's
 * law; the y-z exchange order (boundary_periodic__particle_yz) is parity
 * unpinned.
 * ------------------------------------------------------------------------- */
typedef struct {
  int ndim, np, nsp;
  int nxgs, nxge, nygs, nyge, nzgs, nzge, nys, nye, nzs, nze;
  double delx, delt, c, gfac;
  double q[OR_MAXSP], r[OR_MAXSP];
} or3d_params;

typedef struct {
  double *df, *gkl, *uj;
  int cg_iters[3];
} or3d_fstate;

int or3d_fstate_init(const or3d_params *p, or3d_fstate *st);
void or3d_fstate_free(or3d_fstate *st);
void or3d_field_consts(const or3d_params *p, double f[5]);
/* 3d/common/particle.f90:52-233 */
void or3d_particle_solv_vay(const or3d_params *p, double *gp, const double *up, const double *uf,
                            const int *cumcnt);
void or3d_particle_solv(const or3d_params *p, double *gp, const double *up, const double *uf, const int *cumcnt);
/* 3d/common/field.f90:211-406 */
void or3d_ele_cur(const or3d_params *p, double *uj, const double *up, const double *gp, const int *cumcnt);
/* 3d/common/boundary_periodic.f90:676-978, 458-673, 981-1099 */
void or3d_bc_curre(const or3d_params *p, double *uj);
void or3d_bc_dfield(const or3d_params *p, double *df);
void or3d_bc_phi(const or3d_params *p, double *phi, int l);
/* 3d/common/field.f90:409-560 */
int or3d_cgm(const or3d_params *p, or3d_fstate *st);
/* 3d/common/field.f90:127-206 from the folded current already in st->uj */
int or3d_field_from_current(const or3d_params *p, or3d_fstate *st, double *uf);
/* 3d/common/field.f90:70-208 */
int or3d_field_fdtd_i(const or3d_params *p, or3d_fstate *st, double *uf, const double *up, const double *gp,
                      const int *cumcnt);
/* 3d/common/boundary_periodic.f90:68-101 (round to nearest) */
void or3d_bc_particle_x(const or3d_params *p, double *up, const int *np2);
/* 3d/common/boundary_periodic.f90:104-455 */
int or3d_bc_particle_yz(const or3d_params *p, double *up, int *np2);
/* 3d/common/sort.f90:40-88 */
void or3d_sort_bucket(const or3d_params *p, double *gp, const double *up, int *cumcnt, const int *np2);
/* 3d/proj/weibel/app.f90:103-108 */
int or3d_step(const or3d_params *p, or3d_fstate *st, double *up, double *gp, double *uf, int *cumcnt, int *np2,
              int nsteps);
/* out = {sum r(gamma-1), sum E^2/8pi, sum B^2/8pi} over interior cells */
void or3d_energy(const or3d_params *p, const double *up, const double *uf, const int *np2, double out[3]);
/* G(nx, ny, nz) = div E - 4 pi rho at cell centres */
void or3d_gauss(const or3d_params *p, double *G, const double *uf, const double *up, const int *np2);
/* moments: mom(7, nxgs-1:nxge+1, nys-1:nye+1, nzs-1:nze+1, nsp).
 * 3d/common/mom_calc.f90:49-216 (accl), :219-333 (nvt),
 * 3d/common/boundary_periodic.f90:1102-1235 (bc) */
void or3d_mom_accl(const or3d_params *p, double *gp, const double *up, const double *uf, const int *cumcnt);
void or3d_mom_nvt(const or3d_params *p, double *mom, const double *up, const int *np2);
void or3d_bc_mom(const or3d_params *p, double *mom);

/* round-toward-minus-infinity arithmetic used by the periodic wraps
 * (ieee_set_rounding_mode(ieee_down), boundary_periodic.f90:74,124) */
double or_add_rd(double a, double b);
double or_mul_rd(double a, double b);
double or_div_rd(double a, double b);

#ifdef __cplusplus
}
#endif
#endif
