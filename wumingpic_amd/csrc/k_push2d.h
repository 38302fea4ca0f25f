// k_push2d.h -- the tile geometry, the staged cell-centred fields and the
// bit-exact Buneman-Boris kick shared by the 2D push (k_push2d.hip) and the
// moment kernels (k_mom2d.hip).
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

// Buneman-Boris push of one particle in cell (i, j) (particle.f90:98-163),
// fields gathered from the tile's staged cell-centred values.  Shared by the
// parity and fast kernels, so both produce bit-identical gp.  gmv returns the
// 1/gamma of the move.
__device__ inline void boris_push(const double (*sf)[SFW * SFH], int li, int lj, int i, int j, double d_delx,
                                  double delt, double c, bool unit_c, double fac1, double txxx, double fac2,
                                  double x, double y, double ux, double uy, double uz, double& nxp, double& nyp,
                                  double& g3, double& g4, double& g5, double& gmv) {
  double sx0, sx1, sx2, sy0, sy1, sy2;
  shape3(x * d_delx - 0.5 - i, sx0, sx1, sx2);
  shape3(y * d_delx - 0.5 - j, sy0, sy1, sy2);
  // re-read the cell's staged fields from LDS for every particle (the group's
  // lanes read the same words: a broadcast) instead of holding all 54 in
  // registers across the particle loop
  int soff = (lj + 1) * SFW + (li + 1);
  asm volatile("" : "+v"(soff));
  double fl[6];
#pragma unroll
  for (int cmp = 0; cmp < 6; cmp++) {
    const double* f = &sf[cmp][soff];
    fl[cmp] = +(+f[-SFW - 1] * sx0 + f[-SFW] * sx1 + f[-SFW + 1] * sx2) * sy0 +
              (+f[-1] * sx0 + f[0] * sx1 + f[1] * sx2) * sy1 +
              (+f[SFW - 1] * sx0 + f[SFW] * sx1 + f[SFW + 1] * sx2) * sy2;
  }
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

}  // namespace wpic
