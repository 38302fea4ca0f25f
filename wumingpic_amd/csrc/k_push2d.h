// k_push2d.h -- the tile geometry, the staged cell-centred fields and the
// bit-exact Buneman-Boris and Vay kicks shared by the 2D push (k_push2d.hip)
// and the moment kernels (k_mom2d.hip).
#pragma once

#include "wpic_internal.h"

namespace wpic {

constexpr int TX = 32, TY = 4, NT = 256;
constexpr int SFW = TX + 2, SFH = TY + 2;  // staged fields: one-cell halo
constexpr int ACW = TX + 4, ACH = TY + 4;  // current accumulator: two-cell halo

// cell-centred field component c at (i, j) (particle.f90:74-80)
__device__ inline double cell_field(const Grid2D& g, const double* uf, int c, int i, int j) {
  const size_t pl = (size_t)g.fw * g.fh;
  const double* u = uf + pl * c;
  switch (c) {
    case 0: return 0.5 * (+u[fidx(g, i, j)] + u[fidx(g, i, j + 1)]);
    case 1: return 0.5 * (+u[fidx(g, i, j)] + u[fidx(g, i + 1, j)]);
    case 2:
      return 0.25 * (+u[fidx(g, i, j)] + u[fidx(g, i + 1, j)] + u[fidx(g, i, j + 1)] + u[fidx(g, i + 1, j + 1)]);
    case 3: return 0.5 * (+u[fidx(g, i, j)] + u[fidx(g, i + 1, j)]);
    case 4: return 0.5 * (+u[fidx(g, i, j)] + u[fidx(g, i, j + 1)]);
    default: return u[fidx(g, i, j)];
  }
}

// The six fields at a particle of cell (i, j) with the second-order shape
// (particle.f90:96-133), from the tile's staged cell-centred values.
__device__ inline void gather6(const double (*sf)[SFW * SFH], int li, int lj, int i, int j, double d_delx, double x,
                               double y, double* fl) {
  double sx0, sx1, sx2, sy0, sy1, sy2;
  shape3(x * d_delx - 0.5 - i, sx0, sx1, sx2);
  shape3(y * d_delx - 0.5 - j, sy0, sy1, sy2);
  // re-read the cell's staged fields from LDS for every particle (the group's
  // lanes read the same words: a broadcast) instead of holding all 54 in
  // registers across the particle loop
  int soff = (lj + 1) * SFW + (li + 1);
  asm volatile("" : "+v"(soff));
#pragma unroll
  for (int cmp = 0; cmp < 6; cmp++) {
    const double* f = &sf[cmp][soff];
    fl[cmp] = +(+f[-SFW - 1] * sx0 + f[-SFW] * sx1 + f[-SFW + 1] * sx2) * sy0 +
              (+f[-1] * sx0 + f[0] * sx1 + f[1] * sx2) * sy1 +
              (+f[SFW - 1] * sx0 + f[SFW] * sx1 + f[SFW + 1] * sx2) * sy2;
  }
}

// Buneman-Boris push of one particle in cell (i, j) (particle.f90:98-163),
// fields gathered from the tile's staged cell-centred values.  Shared by the
// parity and fast kernels, so both produce bit-identical gp.  gmv returns the
// 1/gamma of the move.
__device__ inline void boris_push(const double (*sf)[SFW * SFH], int li, int lj, int i, int j, double d_delx,
                                  double delt, double c, bool unit_c, double fac1, double txxx, double fac2,
                                  double x, double y, double ux, double uy, double uz, double& nxp, double& nyp,
                                  double& g3, double& g4, double& g5, double& gmv) {
  double fl[6];
  gather6(sf, li, lj, i, j, d_delx, x, y, fl);
  const double bpx = fl[0], bpy = fl[1], bpz = fl[2], epx = fl[3], epy = fl[4], epz = fl[5];
  double uvm1 = ux + fac1 * epx;
  double uvm2 = uy + fac1 * epy;
  double uvm3 = uz + fac1 * epz;
  double gam = sqrt(c * c + uvm1 * uvm1 + uvm2 * uvm2 + uvm3 * uvm3);
  const double igam = 1. / gam;
  const double fac1r = fac1 * igam;
  const double fac2r = fac2 / (gam + txxx * (bpx * bpx + bpy * bpy + bpz * bpz) * igam);
  const double uvm4 = uvm1 + fac1r * (+uvm2 * bpz - uvm3 * bpy);
  const double uvm5 = uvm2 + fac1r * (+uvm3 * bpx - uvm1 * bpz);
  const double uvm6 = uvm3 + fac1r * (+uvm1 * bpy - uvm2 * bpx);
  uvm1 = uvm1 + fac2r * (+uvm5 * bpz - uvm6 * bpy);
  uvm2 = uvm2 + fac2r * (+uvm6 * bpx - uvm4 * bpz);
  uvm3 = uvm3 + fac2r * (+uvm4 * bpy - uvm5 * bpx);
  g3 = uvm1 + fac1 * epx;
  g4 = uvm2 + fac1 * epy;
  g5 = uvm3 + fac1 * epz;
  const double u2 = +g3 * g3 + g4 * g4 + g5 * g5;
  gmv = 1.0 / sqrt(1.0 + (unit_c ? u2 : u2 / (c * c)));
  nxp = x + g3 * delt * gmv;
  nyp = y + g4 * delt * gmv;
}

// Vay push of one particle in cell (i, j) (particle__solv_vay,
// particle.f90:262-301; Vay, Phys. Plasmas 15, 056701 (2008)): the same
// gather, u' = u + fac2 E + (fac1/gam) u x B, tau = fac1 B / c, the new gamma
// from sigma and u'.tau, then the implicit rotation; the move uses that gamma.
// Same arguments as boris_push (txxx unused).
__device__ inline void vay_push(const double (*sf)[SFW * SFH], int li, int lj, int i, int j, double d_delx,
                                double delt, double c, double fac1, double fac2, double x, double y, double ux,
                                double uy, double uz, double& nxp, double& nyp, double& g3, double& g4, double& g5,
                                double& gmv) {
  double fl[6];
  gather6(sf, li, lj, i, j, d_delx, x, y, fl);
  const double bpx = fl[0], bpy = fl[1], bpz = fl[2], epx = fl[3], epy = fl[4], epz = fl[5];
  double gam = sqrt(c * c + ux * ux + uy * uy + uz * uz);
  const double fac1r = fac1 / gam;
  const double uvm4 = ux + fac2 * epx + fac1r * (+uy * bpz - uz * bpy);
  const double uvm5 = uy + fac2 * epy + fac1r * (+uz * bpx - ux * bpz);
  const double uvm6 = uz + fac2 * epz + fac1r * (+ux * bpy - uy * bpx);
  const double taux = fac1 * bpx / c, tauy = fac1 * bpy / c, tauz = fac1 * bpz / c;
  const double tau2 = taux * taux + tauy * tauy + tauz * tauz;
  const double ua = (uvm4 * taux + uvm5 * tauy + uvm6 * tauz) / c;
  const double sigma = 1.0 + (uvm4 * uvm4 + uvm5 * uvm5 + uvm6 * uvm6) / (c * c) - tau2;
  const double gam2 = 0.5 * (sigma + sqrt(sigma * sigma + 4.0 * (tau2 + ua * ua)));
  gam = sqrt(gam2);
  const double s = 1.0 / (tau2 + gam2);
  g3 = s * (gam2 * uvm4 + c * ua * taux + gam * (uvm5 * tauz - uvm6 * tauy));
  g4 = s * (gam2 * uvm5 + c * ua * tauy + gam * (uvm6 * taux - uvm4 * tauz));
  g5 = s * (gam2 * uvm6 + c * ua * tauz + gam * (uvm4 * tauy - uvm5 * taux));
  gmv = 1.0 / gam;
  nxp = x + g3 * delt * gmv;
  nyp = y + g4 * delt * gmv;
}

}  // namespace wpic
