// k_push2d.hip -- Buneman-Boris push (2d/common/particle.f90:48-179) and
// Esirkepov charge-conserving deposit (2d/common/field.f90:189-316) on tiles.
//
// One workgroup owns a TX x TY tile of cells.  It stages the cell-centred
// fields of the tile plus a one-cell halo in LDS (computed from uf as
// particle.f90:70-83 does) and, in fast mode, a current accumulator for the
// tile plus a two-cell halo.  Lane groups of G lanes walk one cell at a time
// (particles of a cell are contiguous, sorted by cell); each lane keeps the
// cell's 65 current weights in registers, the group reduces them once per cell
// and adds them to LDS; the tile is added to HBM with one atomic per node.
//
// Parity mode (G = 1): one lane per cell accumulates in the reference's slot
// order and writes its 5x5x3 block to a per-cell buffer; k_gather_pj then adds
// the blocks in the reference's (j, i) cell order, reproducing uj bit for bit.
//
// All arithmetic is written in the reference's evaluation order and the
// library is built with -ffp-contract=off, so the push is bit-identical to
// the reference at any thread count.
#include "k_push2d.h"

namespace wpic {

struct PushArgs {
  Grid2D g;
  Parts A, B;
  const int* rowbase;
  const int* cumcnt;
  const int* ap;  // A's slot -> record (null: identity)
  const double* uf;
  double* uj;
  double* pjbuf;
  int* code;
  int* err;
  int ntx, nty;
  int vay;  // the Vay kick instead of Buneman-Boris (particle__solv_vay)
};

// register layout of one cell's current block:
//   [0,20)  pjx(ip,jp), ip = -1..2, jp = -2..2  -> (jp+2)*4 + ip+1
//   [20,40) pjy(ip,jp), ip = -2..2, jp = -1..2  -> 20 + (jp+1)*5 + ip+2
//   [40,65) pjz(ip,jp)                          -> 40 + (jp+2)*5 + ip+2
// pjx(-2,.) and pjy(.,-2) are identically +0 in the reference and adding +0
// to a never-negative-zero sum is the identity, so they are not carried.
constexpr int NPJ = 65;

// Esirkepov shape arrays of one particle (field.f90:223-268): s0 at the old
// position relative to cell (i, j), ds = (shifted new shape) - s0
__device__ inline void esirkepov_shapes(const Grid2D& g, double x, double y, double nxp, double nyp, int i, int j,
                                        double s0x[5], double s0y[5], double dsx[5], double dsy[5]) {
  const double d_delx = g.d_delx;
  double t0, t1, t2;
  shape3(x * d_delx - 0.5 - i, t0, t1, t2);
  s0x[0] = 0.0; s0x[1] = t0; s0x[2] = t1; s0x[3] = t2; s0x[4] = 0.0;
  shape3(y * d_delx - 0.5 - j, t0, t1, t2);
  s0y[0] = 0.0; s0y[1] = t0; s0y[2] = t1; s0y[3] = t2; s0y[4] = 0.0;
#pragma unroll
  for (int d = 0; d < 2; d++) {
    const double gp = d == 0 ? nxp : nyp;
    const int cell = d == 0 ? i : j;
    const int i2 = (int)(gp * d_delx);
    const double dh = gp * d_delx - 0.5 - i2;
    const int inc = i2 - cell;
    const double s1_1 = 0.5 * (0.5 - dh) * (0.5 - dh);
    const double s1_2 = 0.75 - dh * dh;
    const double s1_3 = 0.5 * (0.5 + dh) * (0.5 + dh);
    // smo_1 = -(inc-abs(inc))*0.5+0 etc. (field.f90:244-246): whole numbers,
    // never -0; |inc| = 1 at most except where a reflecting wall moves a
    // particle by more, and then the same formula applies
    const double smo_1 = inc < 0 ? (double)(-inc) : 0.0;
    const double smo_2 = (double)(-abs(inc) + 1);
    const double smo_3 = inc > 0 ? (double)inc : 0.0;
    double* ds = d == 0 ? dsx : dsy;
    const double* s0 = d == 0 ? s0x : s0y;
    ds[0] = s1_1 * smo_1;
    ds[1] = s1_1 * smo_2 + s1_2 * smo_1;
    ds[2] = s1_2 * smo_2 + s1_3 * smo_1 + s1_1 * smo_3;
    ds[3] = s1_3 * smo_2 + s1_2 * smo_3;
    ds[4] = s1_3 * smo_3;
#pragma unroll
    for (int k = 0; k < 5; k++) ds[k] = ds[k] - s0[k];
  }
}

// The shapes of a particle that stays in cell (i, j): inc = 0 makes smo =
// (0, 1, 0), so ds(-2) = ds(2) = 0 and ds(k) = s1(k) - s0(k) for k = -1..1,
// the values esirkepov_shapes gives there (x * 1 and + 0 are exact), with
// three of its ten live at a time
__device__ inline void stayer_shapes(double d_delx, double x, double y, double nxp, double nyp, int i, int j,
                                     double s0x[3], double s0y[3], double dsx[3], double dsy[3]) {
  shape3(x * d_delx - 0.5 - i, s0x[0], s0x[1], s0x[2]);
  shape3(y * d_delx - 0.5 - j, s0y[0], s0y[1], s0y[2]);
  double t0, t1, t2;
  shape3(nxp * d_delx - 0.5 - i, t0, t1, t2);
  dsx[0] = t0 - s0x[0]; dsx[1] = t1 - s0x[1]; dsx[2] = t2 - s0x[2];
  shape3(nyp * d_delx - 0.5 - j, t0, t1, t2);
  dsy[0] = t0 - s0y[0]; dsy[1] = t1 - s0y[1]; dsy[2] = t2 - s0y[2];
}

// One cell's N accumulators: parity mode stores them to the cell's block at
// offset OFF; fast mode reduces them over the G lanes of the group and adds
// them to the tile accumulator, spreading the atomics over the lanes.
template <int N, int OFF, int G, bool PARITY>
__device__ inline void flush_block(double (&v)[N], double* pjbuf, double (*acc)[ACW * ACH], int cellidx, int li,
                                   int lj, int lg) {
  if (PARITY) {
    double* dst = pjbuf + (size_t)cellidx * NPJ + OFF;
#pragma unroll
    for (int k = 0; k < N; k++) dst[k] = v[k];
    return;
  }
  if (G > 1) {
#pragma unroll
    for (int k = 0; k < N; k++) {
#pragma unroll
      for (int off = G / 2; off >= 1; off >>= 1) v[k] += __shfl_xor(v[k], off, G);
    }
  }
#pragma unroll
  for (int kk = 0; kk < N; kk++) {
    if ((kk % G) != lg) continue;
    const int k = kk + OFF;
    int comp, ip, jp;
    if (k < 20) {
      comp = 0; jp = k / 4 - 2; ip = k % 4 - 1;
    } else if (k < 40) {
      comp = 1; jp = (k - 20) / 5 - 1; ip = (k - 20) % 5 - 2;
    } else {
      comp = 2; jp = (k - 40) / 5 - 2; ip = (k - 40) % 5 - 2;
    }
    atomicAdd(&acc[comp][(lj + 2 + jp) * ACW + (li + 2 + ip)], v[kk]);
  }
}

// boundary_periodic__particle_x (:74-88, x wrap under round-down, periodic
// only) and particle_y's row test and y wrap (:120-154) for a particle of row
// j; returns the exchange code (new x cell << 2 | row step + 1).
template <int BC>
__device__ inline int wrap_code(const Grid2D& g, double& nxp, double& nyp, int j, int* err) {
  if (BC == WPIC_BC_PERIODIC) {
    const int ipos = cell_rd(nxp, g.delx);
    if (ipos < g.nxgs)
      nxp = rd_add(nxp, g.spanx);
    else if (ipos >= g.nxge + 1)
      nxp = rd_add(nxp, -g.spanx);
  }
  const int jpos = cell_rd(nyp, g.delx);
  if (jpos != j) {
    if (jpos <= g.nygs - 1)
      nyp = rd_add(nyp, g.spany);
    else if (jpos >= g.nyge + 1)
      nyp = rd_add(nyp, -g.spany);
  }
  int ni = (int)nxp, dy = jpos - j;
  if (ni < g.nxgs || ni > g.nxge || dy < -1 || dy > 1) {
    atomicOr(err, ERR_DOMAIN);
    ni = ni < g.nxgs ? g.nxgs : (ni > g.nxge ? g.nxge : ni);
    dy = dy < -1 ? -1 : (dy > 1 ? 1 : dy);
  }
  return ((ni - g.nxgs) << 2) | (dy + 1);
}

// One pass over a cell's particles accumulating the current components
// selected by DZ/DX/DY from the old (A) and new (B) positions, then flushing
// them.  With WRAP it also applies boundary_periodic__particle_x/_y's
// round-down wraps (:74-88, :124-154) and writes the exchange code.
template <bool DZ, bool DX, bool DY, bool WRAP, int G, bool PARITY, int BC>
__device__ inline void deposit_pass(const PushArgs& a, double (*acc)[ACW * ACH], int i, int j, int li, int lj, int lg,
                                    int cellidx) {
  const Grid2D& g = a.g;
  const double cc2 = g.c * g.c;
  const bool unit_c = cc2 == 1.0;
  const double fac = 1.0 / 3.0;
  double az[DZ ? 25 : 1], ax[DX ? 20 : 1], ay[DY ? 20 : 1];
#pragma unroll
  for (int k = 0; k < (DZ ? 25 : 1); k++) az[k] = 0.0;
#pragma unroll
  for (int k = 0; k < (DX ? 20 : 1); k++) ax[k] = 0.0;
#pragma unroll
  for (int k = 0; k < (DY ? 20 : 1); k++) ay[k] = 0.0;
  for (int s = 0; s < g.nsp; s++) {
    const int rw = s * g.ny + (j - g.nys);
    const int* cr = a.cumcnt + (size_t)rw * (g.nx + 1);
    const int base = a.rowbase[rw];
    const int pbeg = base + cr[i - g.nxgs], pend = base + cr[i - g.nxgs + 1];
    const double q = g.q[s];
    const double qk = q * g.delx * g.d_delt;
    for (int pi = pbeg + lg; pi < pend; pi += G) {
      double nxp = a.B.x[pi], nyp = a.B.y[pi];
      const int ia = a.ap ? a.ap[pi] : pi;
      double s0x[5], s0y[5], dsx[5], dsy[5];
      esirkepov_shapes(g, a.A.x[ia], a.A.y[ia], nxp, nyp, i, j, s0x, s0y, dsx, dsy);
      if (DZ) {
        const double g3 = a.B.ux[pi], g4 = a.B.uy[pi], g5 = a.B.uz[pi];
        const double u2 = +g3 * g3 + g4 * g4 + g5 * g5;
        const double gvz = g5 / sqrt(1. + (unit_c ? u2 : u2 / cc2));
        const double qgvz = q * gvz;
#pragma unroll
        for (int jp = 0; jp < 5; jp++)
#pragma unroll
          for (int ip = 0; ip < 5; ip++)
            az[jp * 5 + ip] = az[jp * 5 + ip] + qgvz * (+s0x[ip] * s0y[jp] + 0.5 * dsx[ip] * s0y[jp] +
                                                         0.5 * s0x[ip] * dsy[jp] + fac * dsx[ip] * dsy[jp]);
      }
      if (DX) {
#pragma unroll
        for (int jp = 0; jp < 5; jp++) {
          const double wy = s0y[jp] + 0.5 * dsy[jp];
          double run = 0.0;
#pragma unroll
          for (int ip = 0; ip < 4; ip++) {
            run = run - qk * dsx[ip] * wy;
            ax[jp * 4 + ip] = ax[jp * 4 + ip] + run;
          }
        }
      }
      if (DY) {
#pragma unroll
        for (int ip = 0; ip < 5; ip++) {
          const double wx = s0x[ip] + 0.5 * dsx[ip];
          double run = 0.0;
#pragma unroll
          for (int jp = 0; jp < 4; jp++) {
            run = run - qk * dsy[jp] * wx;
            ay[jp * 5 + ip] = ay[jp * 5 + ip] + run;
          }
        }
      }
      if (WRAP) {
        const int code = wrap_code<BC>(g, nxp, nyp, j, a.err);
        a.B.x[pi] = nxp;
        a.B.y[pi] = nyp;
        a.code[pi] = code;
      }
    }
  }
  if constexpr (DZ) flush_block<25, 40, G, PARITY>(az, a.pjbuf, acc, cellidx, li, lj, lg);
  if constexpr (DX) flush_block<20, 0, G, PARITY>(ax, a.pjbuf, acc, cellidx, li, lj, lg);
  if constexpr (DY) flush_block<20, 20, G, PARITY>(ay, a.pjbuf, acc, cellidx, li, lj, lg);
}

// Per cell, three passes over its particles keep few values live at once:
//   A: push and reflecting walls (the cell's 54 staged field values stay in
//      registers), gp stored;  B: Jz and Jx;  C: Jy, then the periodic wraps
//   and the exchange code.  B and C re-read the cell's particles while they
//   are still in L2.  Each accumulator still receives the cell's particles in
//   slot order, so G = 1 reproduces the reference's sums.
template <int G, bool PUSH, bool DEPOSIT, bool FUSED, bool PARITY, int BC>
__global__ __launch_bounds__(NT) void k_push_deposit(PushArgs a) {
  __shared__ double sf[6][SFW * SFH];
  __shared__ double acc[(DEPOSIT && !PARITY) ? 3 : 1][ACW * ACH];

  const Grid2D& g = a.g;
  const int ntiles = a.ntx * a.nty;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  if (tile >= ntiles) return;
  const int i0 = g.nxgs + (tile % a.ntx) * TX;
  const int j0 = g.nys + (tile / a.ntx) * TY;
  const int tid = threadIdx.x;

  if (PUSH) {
    for (int k = tid; k < 6 * SFW * SFH; k += NT) {
      const int c = k / (SFW * SFH), rem = k - c * (SFW * SFH);
      const int i = i0 - 1 + rem % SFW, j = j0 - 1 + rem / SFW;
      if (i <= g.nxge + 1 && j <= g.nye + 1) sf[c][rem] = cell_field(g, a.uf, c, i, j);
    }
  }
  if (DEPOSIT && !PARITY) {
    for (int k = tid; k < 3 * ACW * ACH; k += NT) (&acc[0][0])[k] = 0.0;
  }
  __syncthreads();

  const int grp = tid / G, lg = tid % G;
  constexpr int NGRP = NT / G;
  const double d_delx = g.d_delx, delt = g.delt, c = g.c;
  const double cc2 = c * c;
  const bool unit_c = cc2 == 1.0;  // x/(c*c) == x exactly: skip the division
  const double fac = 1.0 / 3.0;

  for (int cc = grp; cc < TX * TY; cc += NGRP) {
    const int li = cc % TX, lj = cc / TX;
    const int i = i0 + li, j = j0 + lj;
    if (i > g.nxge || j > g.nye) continue;  // uniform within a group
    const int cellidx = cidx(g, i, j);

    // ---- pass A: push and reflecting walls; gp stored unwrapped ----
    for (int s = 0; s < g.nsp; s++) {
      const int rw = s * g.ny + (j - g.nys);
      const int* cr = a.cumcnt + (size_t)rw * (g.nx + 1);
      const int base = a.rowbase[rw];
      const int pbeg = base + cr[i - g.nxgs], pend = base + cr[i - g.nxgs + 1];
      const double q = g.q[s], rm = g.r[s];
      const double fac1 = q / rm * 0.5 * delt;
      const double txxx = fac1 * fac1;
      const double fac2 = q * delt / rm;
      for (int pi = pbeg + lg; pi < pend; pi += G) {
        const int ia = a.ap ? a.ap[pi] : pi;
        const double x = a.A.x[ia], y = a.A.y[ia];
        double nxp, nyp, g3, g4, g5;
        if (PUSH) {
          double gmv;
          if (a.vay)
            vay_push(sf, li, lj, i, j, d_delx, delt, c, fac1, fac2, x, y, a.A.ux[ia], a.A.uy[ia], a.A.uz[ia], nxp,
                     nyp, g3, g4, g5, gmv);
          else
            boris_push(sf, li, lj, i, j, d_delx, delt, c, unit_c, fac1, txxx, fac2, x, y, a.A.ux[ia], a.A.uy[ia],
                       a.A.uz[ia], nxp, nyp, g3, g4, g5, gmv);
        } else {
          nxp = a.B.x[pi];
          nyp = a.B.y[pi];
          g3 = a.B.ux[pi];
          g4 = a.B.uy[pi];
          g5 = a.B.uz[pi];
        }
        // reflecting walls act before the deposit in the reconnection order
        // (2d/proj/reconnection/app.f90:101-103; boundary_reconnection.f90:79-91)
        if (FUSED) reflect_x<BC>(g, nxp, g3, g4, g5);
        if (PUSH) {
          a.B.x[pi] = nxp;
          a.B.y[pi] = nyp;
          a.B.ux[pi] = g3;
          a.B.uy[pi] = g4;
          a.B.uz[pi] = g5;
          a.B.id[pi] = a.A.id[ia];  // particle.f90:173-177
        }
      }
    }
    if (!DEPOSIT) continue;

    // ---- deposit passes: Jz (field.f90:292-298), Jx (:274-281), Jy (:283-290);
    //      the last pass also applies the wraps and writes the exchange code ----
    deposit_pass<true, false, false, false, G, PARITY, BC>(a, acc, i, j, li, lj, lg, cellidx);
    deposit_pass<false, true, false, false, G, PARITY, BC>(a, acc, i, j, li, lj, lg, cellidx);
    deposit_pass<false, false, true, FUSED, G, PARITY, BC>(a, acc, i, j, li, lj, lg, cellidx);
  }

  if (DEPOSIT && !PARITY) {
    __syncthreads();
    const size_t pl = (size_t)g.fw * g.fh;
    for (int k = tid; k < 3 * ACW * ACH; k += NT) {
      const int comp = k / (ACW * ACH), rem = k - comp * (ACW * ACH);
      const double v = acc[comp][rem];
      if (v == 0.0) continue;
      const int I = i0 - 2 + rem % ACW, J = j0 - 2 + rem / ACW;
      if (I > g.nxge + 2 || J > g.nye + 2) continue;
      unsafeAtomicAdd(a.uj + pl * comp + fidx(g, I, J), v);
    }
  }
}

// Current of one particle as the full 5x5 Esirkepov block in the factored
// form below, added to the LDS tile at cell (li, lj) with atomics (zeros
// skipped): the path of a particle that changes cell.
__device__ inline void deposit2_full(double (*acc)[ACW * ACH], int li, int lj, const double s0x[5], const double s0y[5],
                                     const double dsx[5], const double dsy[5], double qk, double qvz) {
  double cx = 0.0, cy = 0.0;
#pragma unroll
  for (int m = 0; m < 4; m++) {
    cx -= dsx[m];
    cy -= dsy[m];
#pragma unroll
    for (int n = 0; n < 5; n++) {
      const double vx = cx * (qk * fma(0.5, dsy[n], s0y[n]));  // x face m - 1, y node n - 2
      const double vy = cy * (qk * fma(0.5, dsx[n], s0x[n]));  // x node n - 2, y face m - 1
      if (vx != 0.0) atomicAdd(&acc[0][(lj + n) * ACW + (li + 1 + m)], vx);
      if (vy != 0.0) atomicAdd(&acc[1][(lj + 1 + m) * ACW + (li + n)], vy);
    }
  }
#pragma unroll
  for (int jp = 0; jp < 5; jp++) {
    const double A = qvz * fma(0.5, dsy[jp], s0y[jp]), B = qvz * fma(1.0 / 3.0, dsy[jp], 0.5 * s0y[jp]);
#pragma unroll
    for (int ip = 0; ip < 5; ip++) {
      double v = dsx[ip] * B;
      if (ip >= 1 && ip <= 3) v = fma(s0x[ip], A, v);
      if (v != 0.0) atomicAdd(&acc[2][(lj + jp) * ACW + (li + ip)], v);
    }
  }
}

// Push of one particle of cell (i, j) and, with reflecting walls, the
// reflection that precedes the deposit in the reconnection order
// (2d/proj/reconnection/app.f90:101-103; boundary_reconnection.f90:79-91).
template <int BC, bool VAY>
__device__ inline void push_reflect(const Grid2D& g, const double (*sf)[SFW * SFH], int li, int lj, int i, int j,
                                    bool unit_c, const double* kc, double x, double y, double ux, double uy, double uz,
                                    double& nxp, double& nyp, double& g3, double& g4, double& g5, double& gmv) {
  if (VAY)
    vay_push(sf, li, lj, i, j, g.d_delx, g.delt, g.c, kc[0], kc[2], x, y, ux, uy, uz, nxp, nyp, g3, g4, g5, gmv);
  else
    boris_push(sf, li, lj, i, j, g.d_delx, g.delt, g.c, unit_c, kc[0], kc[1], kc[2], x, y, ux, uy, uz, nxp, nyp, g3,
               g4, g5, gmv);
  reflect_x<BC>(g, nxp, g3, g4, g5);
  // the injection edge sends ux to 2 u0 - ux, which changes gamma: vz of the
  // deposit comes from the momenta after the boundary (field.f90:270-272)
  if (BC == WPIC_BC_SHOCK) {
    const double u2 = +g3 * g3 + g4 * g4 + g5 * g5;
    gmv = 1.0 / sqrt(1.0 + (unit_c ? u2 : u2 / (g.c * g.c)));
  }
}

// Fast mode, push and deposit in one pass over each particle.  The push is the
// shared bit-exact boris_push; the current is the same Esirkepov weights in a
// factored form that reorders the reference's sums (so uj agrees to rounding,
// not bit for bit):
//   Jx(ip, jp) += cx(ip) * wy(jp),  cx(m) = -sum_{k<=m} dsx(k),  wy = qk (s0y + ds_y/2)
//   Jy(ip, jp) += cy(jp) * wx(ip),  likewise with x and y exchanged
//   Jz(ip, jp) += s0x(ip) * A(jp) + dsx(ip) * B(jp),
//                 A = q vz (s0y + dsy/2),  B = q vz (s0y/2 + dsy/3)
// which expands to field.f90:274-298 term by term.  A particle that stays in
// its cell has ds = 0 on the outer nodes and cx = 0 on the outer faces (the
// last up to rounding), so its current lives on 6 + 6 + 9 nodes: each lane
// accumulates those 21 values in registers over the particles of the cell it
// handles and the group reduces them once per cell.  A particle that changes
// cell (about one in eight here) is queued in its wave's LDS queue with what
// its current needs (old and unwrapped new position, q vz); after each round
// of the wave's cells all its lanes add the queued particles' full blocks
// (deposit2_full), and beyond the queue's capacity a particle adds its block at
// once.  The queues are per wave, so no barrier stops the workgroup between
// the staging and the final flush.
//
// Work mapping: wave w walks row w of the 32 x 4 tile, four x-adjacent cells
// per round (one per 16-lane group), so a cell's x-neighbours share its wave
// and its y-neighbours run in the sibling waves at about the same time: the
// records of particles that crossed a cell or row boundary are read while the
// lines they share with their old neighbours are still in L2.  Each lane walks
// its cell's particles (all species as one list) with a two-deep software
// pipeline: the permutation entry of element t + 2G and the record of element
// t + G are in flight while element t is pushed.  With FUSED the wraps and the
// exchange code follow in the same pass; otherwise gp keeps the unwrapped
// positions as particle__solv leaves them.
constexpr int QW = 160;  // queued cell-changing particles per wave and round
template <int G, bool FUSED, int BC, bool VAY>
__global__ __launch_bounds__(NT, 2) void k_push_deposit_fast(PushArgs a) {
  constexpr int NW = NT / 64;
  static_assert(NW == TY && 64 / G * (TX / (64 / G)) == TX, "one wave per tile row, whole rounds");
  __shared__ double sf[6][SFW * SFH];
  __shared__ double acc[3][ACW * ACH];
  __shared__ double spc[WPIC_MAXSP][5];  // fac1, txxx, fac2, qk, q
  __shared__ double xqd[NW][QW][5];      // x, y, unwrapped new x, y, q vz
  __shared__ int xqc[NW][QW];            // cell in the tile | species << 8
  __shared__ int nxq[NW];

  const Grid2D& g = a.g;
  const int ntiles = a.ntx * a.nty;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  if (tile >= ntiles) return;
  const int i0 = g.nxgs + (tile % a.ntx) * TX;
  const int j0 = g.nys + (tile / a.ntx) * TY;
  const int tid = threadIdx.x;

  for (int k = tid; k < 6 * SFW * SFH; k += NT) {
    const int c = k / (SFW * SFH), rem = k - c * (SFW * SFH);
    const int i = i0 - 1 + rem % SFW, j = j0 - 1 + rem / SFW;
    if (i <= g.nxge + 1 && j <= g.nye + 1) sf[c][rem] = cell_field(g, a.uf, c, i, j);
  }
  for (int k = tid; k < 3 * ACW * ACH; k += NT) (&acc[0][0])[k] = 0.0;
  if (tid < NW) nxq[tid] = 0;
  // per-species constants (particle.f90:92-94, field.f90:274)
#pragma unroll
  for (int s = 0; s < WPIC_MAXSP; s++) {
    if (tid == s && s < g.nsp) {
      const double q = g.q[s], rm = g.r[s];
      const double fac1 = q / rm * 0.5 * g.delt;
      spc[s][0] = fac1;
      spc[s][1] = fac1 * fac1;
      spc[s][2] = q * g.delt / rm;
      spc[s][3] = q * g.delx * g.d_delt;
      spc[s][4] = q;
    }
  }
  __syncthreads();

  constexpr int RX = 64 / G;  // cells per wave round
  const int w = tid >> 6, lane = tid & 63;
  const int lg = lane % G;
  const int lj = w, j = j0 + lj;
  const bool unit_c = g.c * g.c == 1.0;
  int* const nq = &nxq[w];
  double(*const qd)[5] = xqd[w];
  int* const qc = xqc[w];

  for (int rnd = 0; rnd < TX / RX; rnd++) {
    const int li = rnd * RX + lane / G;
    const int i = i0 + li;
    if (i <= g.nxge && j <= g.nye) {  // uniform within a group
    // stayers: sx[m][u] Jx at x face m, y node u - 1; sy[m][u] Jy at y face
    // m, x node u - 1; sz[v][u] Jz at x node u - 1, y node v - 1
    double sx[2][3], sy[2][3], sz[3][3];
#pragma unroll
    for (int u = 0; u < 3; u++) {
      sx[0][u] = sx[1][u] = sy[0][u] = sy[1][u] = 0.0;
      sz[0][u] = sz[1][u] = sz[2][u] = 0.0;
    }

    // the cell's particles over all species as one list
    CellList cl;
    int tot = 0;
#pragma unroll
    for (int s = 0; s < WPIC_MAXSP; s++) {
      if (s < g.nsp) {
        const int rw = s * g.ny + (j - g.nys);
        const int* cr = a.cumcnt + (size_t)rw * (g.nx + 1);
        const int c0 = cr[i - g.nxgs];
        cl.set(s, a.rowbase[rw] + c0, tot);
        tot += cr[i - g.nxgs + 1] - c0;
      } else {
        cl.set(s, 0, 0x7fffffff);
      }
    }
    int ia_nn = 0;
    double px = 0.0, py = 0.0, pux = 0.0, puy = 0.0, puz = 0.0;
    uint64_t pid = 0;
    {
      int s_, p_;
      int ia_n = 0;
      if (lg < tot) {
        cl.slot(lg, s_, p_);
        ia_n = a.ap ? a.ap[p_] : p_;
      }
      if (lg + G < tot) {
        cl.slot(lg + G, s_, p_);
        ia_nn = a.ap ? a.ap[p_] : p_;
      }
      if (lg < tot) {
        px = a.A.x[ia_n]; py = a.A.y[ia_n];
        pux = a.A.ux[ia_n]; puy = a.A.uy[ia_n]; puz = a.A.uz[ia_n];
        pid = a.A.id[ia_n];
      }
    }
    for (int t = lg; t < tot; t += G) {
      int s, pi;
      cl.slot(t, s, pi);
      const double x = px, y = py, ux = pux, uy = puy, uz = puz;
      const uint64_t id = pid;
      if (t + G < tot) {
        const int ia = ia_nn;
        px = a.A.x[ia]; py = a.A.y[ia];
        pux = a.A.ux[ia]; puy = a.A.uy[ia]; puz = a.A.uz[ia];
        pid = a.A.id[ia];
      }
      if (t + 2 * G < tot) {
        int s_, p_;
        cl.slot(t + 2 * G, s_, p_);
        ia_nn = a.ap ? a.ap[p_] : p_;
      }
      const double* kc = spc[s];
      double nxp, nyp, g3, g4, g5, gmv;
      push_reflect<FUSED ? BC : WPIC_BC_PERIODIC, VAY>(g, sf, li, lj, i, j, unit_c, kc, x, y, ux, uy, uz, nxp, nyp,
                                                       g3, g4, g5, gmv);
      a.B.ux[pi] = g3;
      a.B.uy[pi] = g4;
      a.B.uz[pi] = g5;
      a.B.id[pi] = id;  // particle.f90:173-177

      const double qvz = kc[4] * (g5 * gmv);
      const int i2 = (int)(nxp * g.d_delx), j2 = (int)(nyp * g.d_delx);
      if (i2 == i && j2 == j) {
        const double qk = kc[3];
        double s0x[3], s0y[3], dsx[3], dsy[3];
        stayer_shapes(g.d_delx, x, y, nxp, nyp, i, j, s0x, s0y, dsx, dsy);
        const double cx1 = 0.0 - dsx[0], cx2 = cx1 - dsx[1];
        const double cy1 = 0.0 - dsy[0], cy2 = cy1 - dsy[1];
#pragma unroll
        for (int u = 0; u < 3; u++) {
          const double wy = qk * fma(0.5, dsy[u], s0y[u]);
          const double wx = qk * fma(0.5, dsx[u], s0x[u]);
          sx[0][u] = fma(cx1, wy, sx[0][u]);
          sx[1][u] = fma(cx2, wy, sx[1][u]);
          sy[0][u] = fma(cy1, wx, sy[0][u]);
          sy[1][u] = fma(cy2, wx, sy[1][u]);
        }
#pragma unroll
        for (int v = 0; v < 3; v++) {
          const double A = qvz * fma(0.5, dsy[v], s0y[v]);
          const double B = qvz * fma(1.0 / 3.0, dsy[v], 0.5 * s0y[v]);
#pragma unroll
          for (int u = 0; u < 3; u++) sz[v][u] = fma(s0x[u], A, fma(dsx[u], B, sz[v][u]));
        }
      } else {
        const int slot = wave_enqueue(nq);
        if (slot < QW) {
          qd[slot][0] = x;
          qd[slot][1] = y;
          qd[slot][2] = nxp;
          qd[slot][3] = nyp;
          qd[slot][4] = qvz;
          qc[slot] = (lj * TX + li) | (s << 8);
        } else {
          double s0x[5], s0y[5], dsx[5], dsy[5];
          esirkepov_shapes(g, x, y, nxp, nyp, i, j, s0x, s0y, dsx, dsy);
          deposit2_full(acc, li, lj, s0x, s0y, dsx, dsy, kc[3], qvz);
        }
      }

      if (FUSED) {
        const int code = wrap_code<BC>(g, nxp, nyp, j, a.err);
        a.code[pi] = code;
      }
      a.B.x[pi] = nxp;
      a.B.y[pi] = nyp;
    }
    // reduce the stayers' 21 values over the group, spread the LDS atomics
#pragma unroll
    for (int q = 0; q < 21; q++) {
      double v = q < 6 ? sx[q / 3][q % 3] : (q < 12 ? sy[(q - 6) / 3][(q - 6) % 3] : sz[(q - 12) / 3][(q - 12) % 3]);
      static_assert(G == 16, "the group is one DPP row");
      v = row16_sum(v);
      if ((q % G) == lg && v != 0.0) {
        int comp, xo, yo;
        if (q < 6) {
          comp = 0; xo = q / 3; yo = q % 3 - 1;
        } else if (q < 12) {
          comp = 1; yo = (q - 6) / 3; xo = (q - 6) % 3 - 1;
        } else {
          comp = 2; yo = (q - 12) / 3 - 1; xo = (q - 12) % 3 - 1;
        }
        atomicAdd(&acc[comp][(lj + 2 + yo) * ACW + (li + 2 + xo)], v);
      }
    }
    }  // the cell
    // the wave's queued particles that changed cell: the queue was written by
    // this wave's own lanes, and LDS operations of one wave complete in order
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int n = *nq < QW ? *nq : QW;
    for (int q = lane; q < n; q += 64) {
      const int e = qc[q];
      const int qcc = e & 255, s = e >> 8;
      const int qli = qcc % TX, qlj = qcc / TX;
      double s0x[5], s0y[5], dsx[5], dsy[5];
      esirkepov_shapes(g, qd[q][0], qd[q][1], qd[q][2], qd[q][3], i0 + qli, j0 + qlj, s0x, s0y, dsx, dsy);
      deposit2_full(acc, qli, qlj, s0x, s0y, dsx, dsy, spc[s][3], qd[q][4]);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane == 0) *nq = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }

  __syncthreads();
  const size_t pl = (size_t)g.fw * g.fh;
  for (int k = tid; k < 3 * ACW * ACH; k += NT) {
    const int comp = k / (ACW * ACH), rem = k - comp * (ACW * ACH);
    const double v = acc[comp][rem];
    if (v == 0.0) continue;
    const int I = i0 - 2 + rem % ACW, J = j0 - 2 + rem / ACW;
    if (I > g.nxge + 2 || J > g.nye + 2) continue;
    unsafeAtomicAdd(a.uj + pl * comp + fidx(g, I, J), v);
  }
}

// uj(c, I, J) = sum over source cells in the reference's (j, i) loop order of
// their block entry (field.f90:212-310 at one thread), starting from +0.
__global__ __launch_bounds__(256) void k_gather_pj(Grid2D g, const double* __restrict__ pjbuf,
                                                  double* __restrict__ uj) {
  const int n = g.fw * g.fh;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int I = g.nxgs - 2 + k % g.fw, J = g.nys - 2 + k / g.fw;
  const size_t pl = (size_t)g.fw * g.fh;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  const int jlo = J - 2 > g.nys ? J - 2 : g.nys, jhi = J + 2 < g.nye ? J + 2 : g.nye;
  const int ilo = I - 2 > g.nxgs ? I - 2 : g.nxgs, ihi = I + 2 < g.nxge ? I + 2 : g.nxge;
  for (int j = jlo; j <= jhi; j++)
    for (int i = ilo; i <= ihi; i++) {
      const int ip = I - i, jp = J - j;
      const double* b = pjbuf + (size_t)cidx(g, i, j) * NPJ;
      if (ip >= -1) sx = sx + b[(jp + 2) * 4 + ip + 1];
      if (jp >= -1) sy = sy + b[20 + (jp + 1) * 5 + ip + 2];
      sz = sz + b[40 + (jp + 2) * 5 + ip + 2];
    }
  const int f = fidx(g, I, J);
  uj[f] = sx;
  uj[pl + f] = sy;
  uj[2 * pl + f] = sz;
}

template <int G, bool PARITY, int BC>
static hipError_t launch_t(wpic_ctx* c, bool push, bool deposit, bool fused, PushArgs& a, int nblk) {
  dim3 grid(nblk), block(NT);
#define L(P, D, F) hipLaunchKernelGGL((k_push_deposit<G, P, D, F, PARITY, BC>), grid, block, 0, c->st, a)
  if (push && deposit && fused)
    L(true, true, true);
  else if (push && !deposit && !fused)
    L(true, false, false);
  else if (!push && deposit && !fused)
    L(false, true, false);
  else if (push && deposit && !fused)
    L(true, true, false);
  else
    return hipErrorInvalidValue;
#undef L
  return hipGetLastError();
}

hipError_t launch_push_deposit(wpic_ctx* c, bool push, bool deposit, bool fused) {
  PushArgs a;
  a.g = c->g;
  a.A = c->A;
  a.B = c->B;
  a.rowbase = c->rowbaseA;
  a.cumcnt = c->cumcnt;
  a.ap = c->aperm;
  a.uf = c->uf;
  a.uj = c->uj;
  a.pjbuf = c->pjbuf;
  a.code = c->code;
  a.err = c->err;
  a.ntx = (c->g.nx + TX - 1) / TX;
  a.nty = (c->g.ny + TY - 1) / TY;
  a.vay = c->pusher == WPIC_PUSHER_VAY;
  const int nblk = xcd_grid(a.ntx * a.nty);
  const bool parity = c->mode == WPIC_MODE_PARITY && deposit;
  if (parity) {
    if (c->g.bc == WPIC_BC_SHOCK) return launch_t<1, true, WPIC_BC_SHOCK>(c, push, deposit, fused, a, nblk);
    return c->g.bc == WPIC_BC_WALLX ? launch_t<1, true, WPIC_BC_WALLX>(c, push, deposit, fused, a, nblk)
                                    : launch_t<1, true, WPIC_BC_PERIODIC>(c, push, deposit, fused, a, nblk);
  }
  if (push && deposit && fused) {
    constexpr int GF = 16;  // a group's 16 lanes move 128 bytes of each record array per pass (one DPP row)
    const bool wall = c->g.bc == WPIC_BC_WALLX;
    if (c->g.bc == WPIC_BC_SHOCK && a.vay)
      hipLaunchKernelGGL((k_push_deposit_fast<GF, true, WPIC_BC_SHOCK, true>), dim3(nblk), dim3(NT), 0, c->st, a);
    else if (c->g.bc == WPIC_BC_SHOCK)
      hipLaunchKernelGGL((k_push_deposit_fast<GF, true, WPIC_BC_SHOCK, false>), dim3(nblk), dim3(NT), 0, c->st, a);
    else if (wall && a.vay)
      hipLaunchKernelGGL((k_push_deposit_fast<GF, true, WPIC_BC_WALLX, true>), dim3(nblk), dim3(NT), 0, c->st, a);
    else if (wall)
      hipLaunchKernelGGL((k_push_deposit_fast<GF, true, WPIC_BC_WALLX, false>), dim3(nblk), dim3(NT), 0, c->st, a);
    else if (a.vay)
      hipLaunchKernelGGL((k_push_deposit_fast<GF, true, WPIC_BC_PERIODIC, true>), dim3(nblk), dim3(NT), 0, c->st, a);
    else
      hipLaunchKernelGGL((k_push_deposit_fast<GF, true, WPIC_BC_PERIODIC, false>), dim3(nblk), dim3(NT), 0, c->st,
                         a);
    return hipGetLastError();
  }
  if (c->g.bc == WPIC_BC_SHOCK) return launch_t<8, false, WPIC_BC_SHOCK>(c, push, deposit, fused, a, nblk);
  return c->g.bc == WPIC_BC_WALLX ? launch_t<8, false, WPIC_BC_WALLX>(c, push, deposit, fused, a, nblk)
                                  : launch_t<8, false, WPIC_BC_PERIODIC>(c, push, deposit, fused, a, nblk);
}

hipError_t launch_gather_pj(wpic_ctx* c) {
  const int n = c->g.fw * c->g.fh;
  hipLaunchKernelGGL(k_gather_pj, dim3((n + 255) / 256), dim3(256), 0, c->st, c->g, c->pjbuf, c->uj);
  return hipGetLastError();
}

}  // namespace wpic
