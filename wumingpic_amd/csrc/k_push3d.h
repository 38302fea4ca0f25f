// k_push3d.h -- the 3D tile geometry, the staged cell-centred fields and
// the bit-exact Buneman-Boris kick shared by the 3D push (k_push3d.hip) and
// the 3D moment kernels (k_mom3d.hip).
#pragma once

#include "wpic_internal3d.h"

namespace wpic {

namespace p3 {
constexpr int TX = 8, TY = 4, TZ = 4, NT = 256, G = 16;
constexpr int SW = TX + 2, SH = TY + 2, SD = TZ + 2, SN = SW * SH * SD;  // staged tmpf
constexpr int AW = TX + 4, AH = TY + 4, AD = TZ + 4, AN = AW * AH * AD;  // current tile
constexpr int NB = 16;   // particles per shape batch of a half-wave
constexpr int NSV = 42;  // s0x s0y s0z dsx dsy dsz (5 each); fast mode adds cx cy cz (4 each)
}  // namespace p3

// tmpf(c, i, j, k), particle.f90:74-91
__device__ inline double tmpf3(const Grid3D& g, const double* uf, int c, int i, int j, int k) {
  const size_t pl = plane3(g);
  const double* u = uf + pl * c;
  const size_t o = fidx3(g, i, j, k);
  const size_t dx = 1, dy = (size_t)g.fw, dz = (size_t)g.fw * g.fh;
  switch (c) {
    case 0: return 2.5e-1 * (+u[o] + u[o + dy] + u[o + dz] + u[o + dy + dz]);
    case 1: return 2.5e-1 * (+u[o] + u[o + dx] + u[o + dz] + u[o + dx + dz]);
    case 2: return 2.5e-1 * (+u[o] + u[o + dx] + u[o + dy] + u[o + dx + dy]);
    case 3: return 5e-1 * (+u[o] + u[o + dx]);
    case 4: return 5e-1 * (+u[o] + u[o + dy]);
    default: return 5e-1 * (+u[o] + u[o + dz]);
  }
}

__device__ inline void tile_origin(const Grid3D& g, int tile, int ntx, int nty, int& i0, int& j0, int& k0) {
  const int tx = tile % ntx, rest = tile / ntx;
  i0 = g.nxgs + tx * p3::TX;
  j0 = g.nys + (rest % nty) * p3::TY;
  k0 = g.nzs + (rest / nty) * p3::TZ;
}

// The six fields at a particle of cell (i, j, k) from the staged tmpf (so is
// the LDS offset of the cell) with the second-order shape
// (particle.f90:93-197); sh returns the particle's old shape (x, y, z
// triplets), which the deposit reuses.
__device__ inline void gather6_3(const double (*sf)[p3::SN], int so, int i, int j, int k, double d_delx, double x,
                                 double y, double z, double* fl, double sh[9]) {
  using namespace p3;
  double shxm, shx, shxp, shym, shy, shyp, shzm, shz, shzp;
  shape3(x * d_delx - 5e-1 - i, shxm, shx, shxp);
  shape3(y * d_delx - 5e-1 - j, shym, shy, shyp);
  shape3(z * d_delx - 5e-1 - k, shzm, shz, shzp);
  // re-read the cell's staged fields from LDS for every particle instead of
  // holding all 162 in registers across the particle loop
  asm volatile("" : "+v"(so));
#pragma unroll
  for (int cm = 0; cm < 6; cm++) {
    const double* f = &sf[cm][so];
#define RX(o) (+f[(o)-1] * shxm + f[o] * shx + f[(o) + 1] * shxp)
#define RY(o) (+RX((o)-SW) * shym + RX(o) * shy + RX((o) + SW) * shyp)
    fl[cm] = RY(-SW * SH) * shzm + RY(0) * shz + RY(SW * SH) * shzp;
#undef RY
#undef RX
  }
  sh[0] = shxm; sh[1] = shx; sh[2] = shxp;
  sh[3] = shym; sh[4] = shy; sh[5] = shyp;
  sh[6] = shzm; sh[7] = shz; sh[8] = shzp;
}

// Buneman-Boris push of one particle of cell (i, j, k) (particle.f90:93-217).
// Shared by k_push3d and the fused fast kernel, so both give the same gp bit
// for bit.
__device__ inline void boris_push3(const double (*sf)[p3::SN], int so, int i, int j, int k, double d_delx, double delt,
                                   double c, double fac1, double txxx, double fac2, double x, double y, double z,
                                   double ux, double uy, double uz, double& nxp, double& nyp, double& nzp, double& g4,
                                   double& g5, double& g6, double sh[9]) {
  double fl[6];
  gather6_3(sf, so, i, j, k, d_delx, x, y, z, fl, sh);
  const double bpx = fl[0], bpy = fl[1], bpz = fl[2], epx = fl[3], epy = fl[4], epz = fl[5];
  double uvm1 = ux + fac1 * epx;
  double uvm2 = uy + fac1 * epy;
  double uvm3 = uz + fac1 * epz;
  double gam = sqrt(c * c + uvm1 * uvm1 + uvm2 * uvm2 + uvm3 * uvm3);
  const double igam = 1e0 / gam;
  const double fac1r = fac1 * igam;
  const double fac2r = fac2 / (gam + txxx * (bpx * bpx + bpy * bpy + bpz * bpz) * igam);
  const double uvm4 = uvm1 + fac1r * (+uvm2 * bpz - uvm3 * bpy);
  const double uvm5 = uvm2 + fac1r * (+uvm3 * bpx - uvm1 * bpz);
  const double uvm6 = uvm3 + fac1r * (+uvm1 * bpy - uvm2 * bpx);
  uvm1 = uvm1 + fac2r * (+uvm5 * bpz - uvm6 * bpy);
  uvm2 = uvm2 + fac2r * (+uvm6 * bpx - uvm4 * bpz);
  uvm3 = uvm3 + fac2r * (+uvm4 * bpy - uvm5 * bpx);
  g4 = uvm1 + fac1 * epx;
  g5 = uvm2 + fac1 * epy;
  g6 = uvm3 + fac1 * epz;
  gam = 1e0 / sqrt(1e0 + (+g4 * g4 + g5 * g5 + g6 * g6) / (c * c));
  nxp = x + g4 * delt * gam;
  nyp = y + g5 * delt * gam;
  nzp = z + g6 * delt * gam;
}

// Vay push of one particle of cell (i, j, k) (particle__solv_vay,
// 3d/common/particle.f90:370-403; Vay, Phys. Plasmas 15, 056701 (2008)): the
// same gather, u' = u + fac2 E + (fac1/gam) u x B, tau = fac1 B / c, the new
// gamma from sigma and u'.tau, the implicit rotation, the move with that gamma.
__device__ inline void vay_push3(const double (*sf)[p3::SN], int so, int i, int j, int k, double d_delx, double delt,
                                 double c, double fac1, double fac2, double x, double y, double z, double ux,
                                 double uy, double uz, double& nxp, double& nyp, double& nzp, double& g4, double& g5,
                                 double& g6, double sh[9]) {
  double fl[6];
  gather6_3(sf, so, i, j, k, d_delx, x, y, z, fl, sh);
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
  g4 = s * (gam2 * uvm4 + c * ua * taux + gam * (uvm5 * tauz - uvm6 * tauy));
  g5 = s * (gam2 * uvm5 + c * ua * tauy + gam * (uvm6 * taux - uvm4 * tauz));
  g6 = s * (gam2 * uvm6 + c * ua * tauz + gam * (uvm4 * tauy - uvm5 * taux));
  gam = 1.0 / gam;
  nxp = x + g4 * delt * gam;
  nyp = y + g5 * delt * gam;
  nzp = z + g6 * delt * gam;
}

}  // namespace wpic
