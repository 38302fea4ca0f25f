// k_mom2d.hip -- velocity moments of the 2D particles (2d/common/mom_calc.f90)
// and their boundary folds (boundary_periodic.f90:571-636,
// 2d/proj/reconnection/boundary_reconnection.f90:582-650).
//
// mom(7, nxgs-1:nxge+1, nys-1:nye+1, nsp) is kept in the reference's own
// layout (the seven moments of a node contiguous), so transfers are plain
// copies.  Three deposits:
//   k_mom_tile<KICK>  fast: one workgroup per tile of cells accumulates the
//                     tile's nodes (one-node halo) in LDS and adds them to
//                     HBM once per node and species; with KICK it applies
//                     mom_calc__accl's half-step kick on the fly (the fused
//                     accl + nvt of the app's output sequence) and never
//                     writes gp;
//   k_mom_any         fast, for a gp in any slot order: HBM atomics;
//   k_mom_parity      one thread per node sums its particles in the
//                     reference's (j, ii) order: bit-identical to
//                     mom_calc__nvt at one thread.
#include "k_push2d.h"

namespace wpic {

struct MomArgs {
  Grid2D g;
  Parts A, B;  // A: up (sorted, read through ap), B: gp
  const int *rowbaseA, *np2A, *rowbaseB, *np2B;
  const int* cumcnt;
  const int* ap;
  const double* uf;
  double* mom;
  int* err;
  int ntx, nty;
  double fac1[WPIC_MAXSP], txxx[WPIC_MAXSP], fac2[WPIC_MAXSP];  // mom_calc.f90:90-92 with delt/2
};

// stage the tile's cell-centred fields plus a one-cell halo (mom_calc.f90:68-81)
__device__ inline void stage_fields(const Grid2D& g, const double* uf, double (*sf)[SFW * SFH], int i0, int j0) {
  for (int k = threadIdx.x; k < 6 * SFW * SFH; k += NT) {
    const int c = k / (SFW * SFH), rem = k - c * (SFW * SFH);
    const int i = i0 - 1 + rem % SFW, j = j0 - 1 + rem / SFW;
    if (i <= g.nxge + 1 && j <= g.nye + 1) sf[c][rem] = cell_field(g, uf, c, i, j);
  }
}

// the seven moments of one particle times the CIC weight wx*wy, in the
// reference's product order (mom_calc.f90:203-243)
struct MomTerms {
  double u[3], gam;
  __device__ void set(double ux, double uy, double uz, double c) {
    u[0] = ux;
    u[1] = uy;
    u[2] = uz;
    gam = 1. / sqrt(1.0 + (+ux * ux + uy * uy + uz * uz) / (c * c));
  }
  __device__ double term(int k, double wx, double wy) const {
    if (k == 0) return wx * wy;
    if (k < 4) return u[k - 1] * gam * wx * wy;
    return u[k - 4] * u[k - 4] * gam * wx * wy;
  }
};

// ---------------------------------------------------------------------------
// mom_calc__accl: gp(1:2) = up(1:2), gp(3:5) = up(3:5) kicked over delt/2
// (mom_calc.f90:131-155); gp is slot-aligned with up, the id travels along.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_mom_accl(MomArgs a) {
  __shared__ double sf[6][SFW * SFH];
  const Grid2D& g = a.g;
  const int ntiles = a.ntx * a.nty;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  if (tile >= ntiles) return;
  const int i0 = g.nxgs + (tile % a.ntx) * TX;
  const int j0 = g.nys + (tile / a.ntx) * TY;
  stage_fields(g, a.uf, sf, i0, j0);
  __syncthreads();
  constexpr int G = 16, NGRP = NT / G;
  const int grp = threadIdx.x / G, lg = threadIdx.x % G;
  const double hdelt = g.delt * 0.5;
  for (int cc = grp; cc < TX * TY; cc += NGRP) {
    const int li = cc % TX, lj = cc / TX;
    const int i = i0 + li, j = j0 + lj;
    if (i > g.nxge || j > g.nye) continue;
    for (int s = 0; s < g.nsp; s++) {
      const int rw = s * g.ny + (j - g.nys);
      const int* cr = a.cumcnt + (size_t)rw * (g.nx + 1);
      const int base = a.rowbaseA[rw];
      const int pend = base + cr[i - g.nxgs + 1];
      for (int pi = base + cr[i - g.nxgs] + lg; pi < pend; pi += G) {
        const int ia = a.ap ? a.ap[pi] : pi;
        const double x = a.A.x[ia], y = a.A.y[ia];
        double nxp, nyp, g3, g4, g5, gmv;
        boris_push(sf, li, lj, i, j, g.d_delx, hdelt, g.c, true, a.fac1[s], a.txxx[s], a.fac2[s], x, y, a.A.ux[ia],
                   a.A.uy[ia], a.A.uz[ia], nxp, nyp, g3, g4, g5, gmv);
        a.B.x[pi] = x;
        a.B.y[pi] = y;
        a.B.ux[pi] = g3;
        a.B.uy[pi] = g4;
        a.B.uz[pi] = g5;
        a.B.id[pi] = a.A.id[ia];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// fast deposit over tiles: particles of the tile (A kicked on the fly, or B
// in A's rows) into an LDS node tile, then one HBM atomic per node.  The node
// tile covers i0-1 .. i0+TX and j0-1 .. j0+TY: a particle of cell i has its
// lower node int(x/dx - 1/2) in {i-1, i}.  A gp particle that is not in the
// cell its slot belongs to (gp not sorted by cell) is added to HBM directly.
// ---------------------------------------------------------------------------
constexpr int MTW = TX + 2, MTH = TY + 2;
#ifndef MOM_G
#define MOM_G 8
#endif

// the seven moments of a particle rotated by r (out[q] = m[(q + r) % 7]),
// r in 0..6, with three static conditional rotations
__device__ inline void rotate7(double (&m)[7], int r) {
  double t[7];
#pragma unroll
  for (int q = 0; q < 7; q++) t[q] = (r & 1) ? m[(q + 1) % 7] : m[q];
#pragma unroll
  for (int q = 0; q < 7; q++) m[q] = (r & 2) ? t[(q + 2) % 7] : t[q];
#pragma unroll
  for (int q = 0; q < 7; q++) t[q] = (r & 4) ? m[(q + 4) % 7] : m[q];
#pragma unroll
  for (int q = 0; q < 7; q++) m[q] = t[q];
}

template <bool KICK>
__global__ __launch_bounds__(NT) void k_mom_tile(MomArgs a) {
  constexpr int G = MOM_G, NGRP = NT / G;  // lanes per cell, cells in flight
  __shared__ double sf[KICK ? 6 : 1][SFW * SFH];
  __shared__ double mt[7][MTW * MTH];
  const Grid2D& g = a.g;
  const int ntiles = a.ntx * a.nty;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  if (tile >= ntiles) return;
  const int i0 = g.nxgs + (tile % a.ntx) * TX;
  const int j0 = g.nys + (tile / a.ntx) * TY;
  const int ncx = min(TX, g.nxge + 1 - i0), ncy = min(TY, g.nye + 1 - j0);
  const int tid = threadIdx.x;
  if constexpr (KICK) stage_fields(g, a.uf, sf, i0, j0);
  const double hdelt = g.delt * 0.5, d_delx = g.d_delx, c = g.c;
  const int grp = tid / G, lg = tid % G;
  // The lanes of a group work on one cell (and neighbouring groups on
  // neighbouring cells), so their particles' nodes coincide; each lane walks
  // the seven moment planes in an order rotated by its lane index, so that
  // lanes meeting on a node at the same instruction mostly add to different
  // planes.
  const int rot = tid % 7;

  for (int s = 0; s < g.nsp; s++) {
    for (int k = tid; k < 7 * MTW * MTH; k += NT) (&mt[0][0])[k] = 0.0;
    __syncthreads();
    for (int cc = grp; cc < TX * TY; cc += NGRP) {
      const int li = cc % TX, lj = cc / TX;
      if (li >= ncx || lj >= ncy) continue;  // uniform within a group
      const int i = i0 + li, j = j0 + lj;
      const int rw = s * g.ny + (j - g.nys);
      const int* cr = a.cumcnt + (size_t)rw * (g.nx + 1);
      const int base = a.rowbaseA[rw];
      const int pe = base + cr[i - g.nxgs + 1];
      for (int pi = base + cr[i - g.nxgs] + lg; pi < pe; pi += G) {
        double x, y, ux, uy, uz;
        if constexpr (KICK) {
          const int ia = a.ap ? a.ap[pi] : pi;
          x = a.A.x[ia];
          y = a.A.y[ia];
          double nxp, nyp, gmv;
          boris_push(sf, li, lj, i, j, d_delx, hdelt, c, true, a.fac1[s], a.txxx[s], a.fac2[s], x, y, a.A.ux[ia],
                     a.A.uy[ia], a.A.uz[ia], nxp, nyp, ux, uy, uz, gmv);
        } else {
          x = a.B.x[pi];
          y = a.B.y[pi];
          ux = a.B.ux[pi];
          uy = a.B.uy[pi];
          uz = a.B.uz[pi];
        }
        const int ih = (int)(x * d_delx - 0.5), jh = (int)(y * d_delx - 0.5);
        if (ih < g.nxgs - 1 || ih > g.nxge || jh < j - 1 || jh > j) {
          atomicOr(a.err, ERR_DOMAIN);
          continue;
        }
        const double dx = x * d_delx - 0.5 - ih, dxm = 1. - dx;
        const double dy = y * d_delx - 0.5 - jh, dym = 1. - dy;
        MomTerms m;
        m.set(ux, uy, uz, c);
        const int lx = ih - (i0 - 1), ly = jh - (j0 - 1);
        if (lx < 0 || lx > ncx) {
          // a gp particle whose node lies off this tile (gp not sorted by
          // cell): straight to HBM
          double* p00 = a.mom + midx(g, ih, jh, s);
          double* p01 = a.mom + midx(g, ih, jh + 1, s);
#pragma unroll
          for (int k = 0; k < 7; k++) {
            unsafeAtomicAdd(p00 + k, m.term(k, dxm, dym));
            unsafeAtomicAdd(p00 + 7 + k, m.term(k, dx, dym));
            unsafeAtomicAdd(p01 + k, m.term(k, dxm, dy));
            unsafeAtomicAdd(p01 + 7 + k, m.term(k, dx, dy));
          }
          continue;
        }
        double mv[7] = {1.0,
                        m.u[0] * m.gam,
                        m.u[1] * m.gam,
                        m.u[2] * m.gam,
                        m.u[0] * m.u[0] * m.gam,
                        m.u[1] * m.u[1] * m.gam,
                        m.u[2] * m.u[2] * m.gam};
        rotate7(mv, rot);
        const double w00 = dxm * dym, w10 = dx * dym, w01 = dxm * dy, w11 = dx * dy;
        double* node = &mt[0][ly * MTW + lx];
#pragma unroll
        for (int q = 0; q < 7; q++) {
          const int plane = q + rot < 7 ? q + rot : q + rot - 7;
          double* p = node + plane * (MTW * MTH);
          atomicAdd(p, mv[q] * w00);
          atomicAdd(p + 1, mv[q] * w10);
          atomicAdd(p + MTW, mv[q] * w01);
          atomicAdd(p + MTW + 1, mv[q] * w11);
        }
      }
    }
    __syncthreads();
    for (int k = tid; k < 7 * MTW * MTH; k += NT) {
      const int node = k / 7, comp = k % 7;
      const double v = mt[comp][node];
      if (v == 0.0) continue;
      const int I = i0 - 1 + node % MTW, J = j0 - 1 + node / MTW;
      if (I > g.nxge + 1 || J > g.nye + 1) continue;
      unsafeAtomicAdd(a.mom + midx(g, I, J, s) + comp, v);
    }
    __syncthreads();
  }
}

// gp in any slot order (rows of B): one thread per particle, HBM atomics
__global__ __launch_bounds__(NT) void k_mom_any(MomArgs a) {
  const Grid2D& g = a.g;
  const int rw = blockIdx.y, s = rw / g.ny, j = g.nys + rw % g.ny;
  const int base = a.rowbaseB[rw], n = a.np2B[rw];
  for (int t = blockIdx.x * NT + threadIdx.x; t < n; t += gridDim.x * NT) {
    const int pi = base + t;
    const double x = a.B.x[pi], y = a.B.y[pi];
    const int ih = (int)(x * g.d_delx - 0.5), jh = (int)(y * g.d_delx - 0.5);
    if (ih < g.nxgs - 1 || ih > g.nxge || jh < j - 1 || jh > j) {
      atomicOr(a.err, ERR_DOMAIN);
      continue;
    }
    const double dx = x * g.d_delx - 0.5 - ih, dxm = 1. - dx;
    const double dy = y * g.d_delx - 0.5 - jh, dym = 1. - dy;
    MomTerms m;
    m.set(a.B.ux[pi], a.B.uy[pi], a.B.uz[pi], g.c);
    double* p00 = a.mom + midx(g, ih, jh, s);
    double* p10 = a.mom + midx(g, ih + 1, jh, s);
    double* p01 = a.mom + midx(g, ih, jh + 1, s);
    double* p11 = a.mom + midx(g, ih + 1, jh + 1, s);
#pragma unroll
    for (int k = 0; k < 7; k++) {
      unsafeAtomicAdd(p00 + k, m.term(k, dxm, dym));
      unsafeAtomicAdd(p10 + k, m.term(k, dx, dym));
      unsafeAtomicAdd(p01 + k, m.term(k, dxm, dy));
      unsafeAtomicAdd(p11 + k, m.term(k, dx, dy));
    }
  }
}

// ---------------------------------------------------------------------------
// parity: thread per node (n, m) and species.  The particles that reach node
// row m live in rows m-1..m+1 (lower node jh in {j-1, j}); with gp
// slot-aligned with the sorted up only cells n-1..n+1 of those rows can reach
// node n, a contiguous slot range.  Each node adds its terms in the
// reference's (j, ii) order, starting from +0 like the OMP REDUCTION copy.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_mom_parity(MomArgs a, int aligned) {
  const Grid2D& g = a.g;
  const int n = g.nxgs - 1 + blockIdx.x * NT + threadIdx.x;
  if (n > g.nxge + 1) return;
  const int mh = g.ny + 2;
  const int s = blockIdx.y / mh, m = g.nys - 1 + blockIdx.y % mh;
  double acc[7];
#pragma unroll
  for (int k = 0; k < 7; k++) acc[k] = 0.0;
  for (int j = max(g.nys, m - 1); j <= min(g.nye, m + 1); j++) {
    const int rw = s * g.ny + (j - g.nys);
    const int base = a.rowbaseB[rw];
    int pb = 0, pe = a.np2B[rw];
    if (aligned) {
      const int lo = max(g.nxgs, n - 1), hi = min(g.nxge, n + 1);
      const int* cr = a.cumcnt + (size_t)rw * (g.nx + 1);
      pb = cr[lo - g.nxgs];
      pe = cr[hi + 1 - g.nxgs];
    }
    for (int p = pb; p < pe; p++) {
      const int pi = base + p;
      const double x = a.B.x[pi], y = a.B.y[pi];
      const int ih = (int)(x * g.d_delx - 0.5), jh = (int)(y * g.d_delx - 0.5);
      if (ih < g.nxgs - 1 || ih > g.nxge || jh < j - 1 || jh > j) {
        atomicOr(a.err, ERR_DOMAIN);
        continue;
      }
      double wx, wy;
      const double dx = x * g.d_delx - 0.5 - ih, dy = y * g.d_delx - 0.5 - jh;
      if (ih == n)
        wx = 1. - dx;
      else if (ih + 1 == n)
        wx = dx;
      else
        continue;
      if (jh == m)
        wy = 1. - dy;
      else if (jh + 1 == m)
        wy = dy;
      else
        continue;
      MomTerms t;
      t.set(a.B.ux[pi], a.B.uy[pi], a.B.uz[pi], g.c);
#pragma unroll
      for (int k = 0; k < 7; k++) acc[k] = acc[k] + t.term(k, wx, wy);
    }
  }
  double* o = a.mom + midx(g, n, m, s);
#pragma unroll
  for (int k = 0; k < 7; k++) o[k] = acc[k];
}

// ---------------------------------------------------------------------------
// boundary folds at one rank: x over all node rows (periodic: nxge+1 onto
// nxgs, nxgs-1 onto nxge; walls: each ghost onto its own edge), then y
// (nys-1 onto nye, then nye+1 onto nys) over all x nodes.
// ---------------------------------------------------------------------------
__global__ void k_mom_bc_x(Grid2D g, double* mom) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int mh = g.ny + 2;
  if (t >= 7 * mh * g.nsp) return;
  const int k = t % 7, jj = (t / 7) % mh, s = t / (7 * mh);
  const int j = g.nys - 1 + jj;
  const bool wall = walls(g);
  const int xa = wall ? g.nxgs - 1 : g.nxge + 1, xb = wall ? g.nxge + 1 : g.nxgs - 1;
  mom[midx(g, g.nxgs, j, s) + k] = mom[midx(g, g.nxgs, j, s) + k] + mom[midx(g, xa, j, s) + k];
  mom[midx(g, g.nxge, j, s) + k] = mom[midx(g, g.nxge, j, s) + k] + mom[midx(g, xb, j, s) + k];
}

__global__ void k_mom_bc_y(Grid2D g, double* mom) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int mw = g.nx + 2;
  if (t >= 7 * mw * g.nsp) return;
  const int k = t % 7, ii = (t / 7) % mw, s = t / (7 * mw);
  const int i = g.nxgs - 1 + ii;
  mom[midx(g, i, g.nye, s) + k] = mom[midx(g, i, g.nye, s) + k] + mom[midx(g, i, g.nys - 1, s) + k];
  mom[midx(g, i, g.nys, s) + k] = mom[midx(g, i, g.nys, s) + k] + mom[midx(g, i, g.nye + 1, s) + k];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static MomArgs mom_args(wpic_ctx* c) {
  MomArgs a;
  a.g = c->g;
  a.A = c->A;
  a.B = c->B;
  a.rowbaseA = c->rowbaseA;
  a.np2A = c->np2A;
  a.rowbaseB = c->rowbaseB;
  a.np2B = c->np2B;
  a.cumcnt = c->cumcnt;
  a.ap = c->aperm;
  a.uf = c->uf;
  a.mom = c->mom;
  a.err = c->err;
  a.ntx = (c->g.nx + TX - 1) / TX;
  a.nty = (c->g.ny + TY - 1) / TY;
  // mom_calc__init halves delt (mom_calc.f90:34); fac1, txxx, fac2 as
  // mom_calc.f90:90-92
  const double delt = c->g.delt * 0.5;
  for (int s = 0; s < WPIC_MAXSP; s++) {
    const double q = c->g.q[s], r = c->g.r[s];
    a.fac1[s] = s < c->g.nsp ? q / r * 0.5 * delt : 0.0;
    a.txxx[s] = a.fac1[s] * a.fac1[s];
    a.fac2[s] = s < c->g.nsp ? q * delt / r : 0.0;
  }
  return a;
}

hipError_t launch_mom_accl(wpic_ctx* c) {
  MomArgs a = mom_args(c);
  hipLaunchKernelGGL(k_mom_accl, dim3(xcd_grid(a.ntx * a.nty)), dim3(NT), 0, c->st, a);
  return hipGetLastError();
}

hipError_t launch_mom_deposit(wpic_ctx* c, int kind) {
  MomArgs a = mom_args(c);
  const size_t nm = (size_t)7 * (c->g.nx + 2) * (c->g.ny + 2) * c->g.nsp;
  if (kind != MOM_PARITY_ALIGNED && kind != MOM_PARITY_ANY) {
    hipError_t e = launch_zero(c, c->mom, nm);
    if (e != hipSuccess) return e;
  }
  switch (kind) {
    case MOM_TILE_KICK:
      hipLaunchKernelGGL(k_mom_tile<true>, dim3(xcd_grid(a.ntx * a.nty)), dim3(NT), 0, c->st, a);
      break;
    case MOM_TILE_GP:
      hipLaunchKernelGGL(k_mom_tile<false>, dim3(xcd_grid(a.ntx * a.nty)), dim3(NT), 0, c->st, a);
      break;
    case MOM_ANY:
      hipLaunchKernelGGL(k_mom_any, dim3(8, c->g.nsp * c->g.ny), dim3(NT), 0, c->st, a);
      break;
    default:
      hipLaunchKernelGGL(k_mom_parity, dim3((c->g.nx + 2 + NT - 1) / NT, c->g.nsp * (c->g.ny + 2)), dim3(NT), 0,
                         c->st, a, kind == MOM_PARITY_ALIGNED ? 1 : 0);
      break;
  }
  return hipGetLastError();
}

hipError_t launch_mom_bc(wpic_ctx* c) {
  const Grid2D& g = c->g;
  const int nx_t = 7 * (g.ny + 2) * g.nsp, ny_t = 7 * (g.nx + 2) * g.nsp;
  hipLaunchKernelGGL(k_mom_bc_x, dim3((nx_t + 255) / 256), dim3(256), 0, c->st, g, c->mom);
  hipLaunchKernelGGL(k_mom_bc_y, dim3((ny_t + 255) / 256), dim3(256), 0, c->st, g, c->mom);
  return hipGetLastError();
}

}  // namespace wpic
