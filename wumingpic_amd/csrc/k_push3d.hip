// k_push3d.hip -- 3D Buneman-Boris push (3d/common/particle.f90:52-233) and
// Esirkepov deposit (3d/common/field.f90:211-406) on tiles of cells.
//
// Push: one workgroup per TX x TY x TZ tile stages the cell-centred fields
// (tmpf, particle.f90:74-91) of the tile plus a one-cell halo in LDS; groups of
// G lanes walk one cell at a time and each lane pushes every G-th particle of
// the cell with the reference's operation order (27-point gather nested as
// ((x row) y) z, particle.f90:126-185), so gp is bit-identical to the
// reference's at any thread count.
//
// Deposit: each half-wave (32 lanes) owns one cell at a time.  Its lanes first
// compute the old/new shape arrays s0 and ds of up to 16 particles into LDS;
// then lane (jp, kp) (25 of the 32) walks those particles in slot order and
// accumulates the cell's pjx(:, jp, kp), pjy(:, jp, kp), pjz(:, jp, kp) with
// the reference's dstmp / running pjtmp arithmetic (field.f90:335-380) in
// parity mode, or in a factored form with FMA (the running sums of ds
// precomputed per particle) in fast mode.
// Parity mode writes the 300-entry block per cell; k_gather_pj3d then adds the
// blocks into uj in the reference's (k, j, i) cell order with the scatter
// permutation uj(2) += pjy(jp, ip, kp), uj(3) += pjz(kp, ip, jp) (:390-398),
// bit for bit.  Fast mode adds the blocks into an LDS tile accumulator and
// the tile into uj with atomics.
#include "k_push3d.h"

namespace wpic {

struct Push3Args {
  Grid3D g;
  Parts3 A, B;
  const int* colbase;
  const int* cumcnt;
  const int* ap;  // A's slot -> record in A's buffer (null: identity)
  int* nov;       // fused fast kernel: count of the overflow list below
  int *ovslot, *ovcell;
  const double* uf;
  double* uj;
  double* pjbuf;
  int ntx, nty, ntz;
  int vay;  // the Vay kick instead of Buneman-Boris (particle__solv_vay)
  int* code;  // fused fast kernel: the exchange codes of gp
  int* err;
};

__global__ __launch_bounds__(p3::NT) void k_push3d(Push3Args a) {
  using namespace p3;
  __shared__ double sf[6][SN];
  const Grid3D& g = a.g;
  const int ntiles = a.ntx * a.nty * a.ntz;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  if (tile >= ntiles) return;
  int i0, j0, k0;
  tile_origin(g, tile, a.ntx, a.nty, i0, j0, k0);
  const int tid = threadIdx.x;
  for (int q = tid; q < 6 * SN; q += NT) {
    const int c = q / SN, rem = q - c * SN;
    const int i = i0 - 1 + rem % SW, j = j0 - 1 + (rem / SW) % SH, k = k0 - 1 + rem / (SW * SH);
    if (i <= g.nxge + 1 && j <= g.nye + 1 && k <= g.nze + 1) sf[c][rem] = tmpf3(g, a.uf, c, i, j, k);
  }
  __syncthreads();

  const int grp = tid / G, lg = tid % G;
  constexpr int NGRP = NT / G;
  const double d_delx = g.d_delx, delt = g.delt, c = g.c;
  for (int cc = grp; cc < TX * TY * TZ; cc += NGRP) {
    const int li = cc % TX, lj = (cc / TX) % TY, lk = cc / (TX * TY);
    const int i = i0 + li, j = j0 + lj, k = k0 + lk;
    if (i > g.nxge || j > g.nye || k > g.nze) continue;  // uniform within a group
    const int col = (j - g.nys) + g.ny * (k - g.nzs);
    for (int s = 0; s < g.nsp; s++) {
      const int rw = s * g.ncol + col;
      const int* cr = a.cumcnt + (size_t)rw * (g.nx + 1);
      const int base = a.colbase[rw];
      const int pbeg = base + cr[i - g.nxgs], pend = base + cr[i - g.nxgs + 1];
      const double q = g.q[s], rm = g.r[s];
      const double fac1 = q / rm * 5e-1 * delt;
      const double txxx = fac1 * fac1;
      const double fac2 = q * delt / rm;
      for (int pi = pbeg + lg; pi < pend; pi += G) {
        const int ia = a.ap ? a.ap[pi] : pi;
        const double x = a.A.x[ia], y = a.A.y[ia], z = a.A.z[ia];
        double nxp, nyp, nzp, g4, g5, g6, sh[9];
        const int so = (lk + 1) * (SW * SH) + (lj + 1) * SW + (li + 1);
        if (a.vay)
          vay_push3(sf, so, i, j, k, d_delx, delt, c, fac1, fac2, x, y, z, a.A.ux[ia], a.A.uy[ia], a.A.uz[ia], nxp,
                    nyp, nzp, g4, g5, g6, sh);
        else
          boris_push3(sf, so, i, j, k, d_delx, delt, c, fac1, txxx, fac2, x, y, z, a.A.ux[ia], a.A.uy[ia],
                      a.A.uz[ia], nxp, nyp, nzp, g4, g5, g6, sh);
        a.B.x[pi] = nxp;
        a.B.y[pi] = nyp;
        a.B.z[pi] = nzp;
        a.B.ux[pi] = g4;
        a.B.uy[pi] = g5;
        a.B.uz[pi] = g6;
        a.B.id[pi] = a.A.id[ia];  // particle.f90:227-231
      }
    }
  }
}

// s0 and ds of one particle (field.f90:253-329) for cell (i, j, k):
// v[0..4] s0x, [5..9] s0y, [10..14] s0z, [15..19] dsx, [20..24] dsy, [25..29] dsz
__device__ inline void shapes3(const Grid3D& g, const double old[3], const double nw[3], const int cell[3],
                               double* v) {
  const double d_delx = g.d_delx;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    double s0m, s00, s0p;
    shape3(old[d] * d_delx - 5e-1 - cell[d], s0m, s00, s0p);
    const int i2 = (int)(nw[d] * d_delx);
    const double dh = nw[d] * d_delx - 5e-1 - i2;
    const int inc = i2 - cell[d];
    double s1_1, s1_2, s1_3;
    shape3(dh, s1_1, s1_2, s1_3);
    // smo_* = -(inc-abs(inc))*5d-1+0 etc.: whole numbers, never -0
    const double smo_1 = inc < 0 ? (double)(-inc) : 0.0;
    const double smo_2 = (double)(-abs(inc) + 1);
    const double smo_3 = inc > 0 ? (double)inc : 0.0;
    double* s0 = v + 5 * d;
    double* ds = v + 15 + 5 * d;
    s0[0] = 0e0;
    s0[1] = s0m;
    s0[2] = s00;
    s0[3] = s0p;
    s0[4] = 0e0;
    ds[0] = s1_1 * smo_1 - 0e0;
    ds[1] = s1_1 * smo_2 + s1_2 * smo_1 - s0m;
    ds[2] = s1_2 * smo_2 + s1_3 * smo_1 + s1_1 * smo_3 - s00;
    ds[3] = s1_3 * smo_2 + s1_2 * smo_3 - s0p;
    ds[4] = s1_3 * smo_3 - 0e0;
  }
}

template <bool PARITY>
__global__ __launch_bounds__(p3::NT) void k_deposit3d(Push3Args a) {
  using namespace p3;
  __shared__ double shp[NT / 32][NB][NSV];
  __shared__ double acc[PARITY ? 1 : 3][PARITY ? 1 : AN];
  const Grid3D& g = a.g;
  const int ntiles = a.ntx * a.nty * a.ntz;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  if (tile >= ntiles) return;
  int i0, j0, k0;
  tile_origin(g, tile, a.ntx, a.nty, i0, j0, k0);
  const int tid = threadIdx.x;
  if (!PARITY) {
    for (int q = tid; q < 3 * AN; q += NT) (&acc[0][0])[q] = 0.0;
    __syncthreads();
  }
  const int hw = tid >> 5, hl = tid & 31;      // half-wave and its lane
  const int ja = hl % 5, kb = hl / 5;          // this lane's (jp, kp) + 2 when hl < 25
  const bool accl = hl < 25;
  const double fac = 1e0 / 3e0;
  double(*sh)[NSV] = shp[hw];

  for (int cc = hw; cc < TX * TY * TZ; cc += NT / 32) {
    const int li = cc % TX, lj = (cc / TX) % TY, lk = cc / (TX * TY);
    const int i = i0 + li, j = j0 + lj, k = k0 + lk;
    const bool live = i <= g.nxge && j <= g.nye && k <= g.nze;
    const int cell[3] = {i, j, k};
    const int col = (j - g.nys) + g.ny * (k - g.nzs);
    double px[4] = {0, 0, 0, 0}, py[4] = {0, 0, 0, 0}, pz[4] = {0, 0, 0, 0};
    for (int s = 0; s < g.nsp; s++) {
      int pbeg = 0, pend = 0;
      if (live) {
        const int rw = s * g.ncol + col;
        const int* cr = a.cumcnt + (size_t)rw * (g.nx + 1);
        const int base = a.colbase[rw];
        pbeg = base + cr[i - g.nxgs];
        pend = base + cr[i - g.nxgs + 1];
      }
      const double qdxdt = g.q[s] * g.delx * g.d_delt;
      for (int b0 = pbeg; b0 < pend; b0 += NB) {
        const int nb = pend - b0 < NB ? pend - b0 : NB;
        if (hl < nb) {
          const int pi = b0 + hl;
          const int ia = a.ap ? a.ap[pi] : pi;
          const double old[3] = {a.A.x[ia], a.A.y[ia], a.A.z[ia]};
          const double nw[3] = {a.B.x[pi], a.B.y[pi], a.B.z[pi]};
          double v[NSV];
          shapes3(g, old, nw, cell, v);
          if (!PARITY) {
            // running sums c(m) = -sum_{k<=m} ds(k) of the factored form
#pragma unroll
            for (int d = 0; d < 3; d++) {
              double run = 0.0;
#pragma unroll
              for (int m = 0; m < 4; m++) {
                run -= v[15 + 5 * d + m];
                v[30 + 4 * d + m] = run;
              }
            }
          }
#pragma unroll
          for (int q = 0; q < (PARITY ? 30 : NSV); q++) sh[hl][q] = v[q];
        }
        // the shapes were written by other lanes of this wave
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (accl && !PARITY) {
          // fast mode: pj(m-1, jp, kp) += c(m) * w(jp, kp), the dstmp of the
          // reference written as w = (s0 + ds/2) s0' + (s0/2 + ds/3) ds' (FMA)
          for (int p = 0; p < nb; p++) {
            const double* v = sh[p];
            const double *s0x = v, *s0y = v + 5, *s0z = v + 10, *dsx = v + 15, *dsy = v + 20, *dsz = v + 25;
            const double *cxs = v + 30, *cys = v + 34, *czs = v + 38;
            const double axj = fma(0.5, dsx[ja], s0x[ja]), bxj = fma(fac, dsx[ja], 0.5 * s0x[ja]);
            const double ayj = fma(0.5, dsy[ja], s0y[ja]), byj = fma(fac, dsy[ja], 0.5 * s0y[ja]);
            const double wx = fma(ayj, s0z[kb], byj * dsz[kb]) * qdxdt;
            const double wy = fma(axj, s0z[kb], bxj * dsz[kb]) * qdxdt;
            const double wz = fma(axj, s0y[kb], bxj * dsy[kb]) * qdxdt;
#pragma unroll
            for (int m = 0; m < 4; m++) {
              px[m] = fma(cxs[m], wx, px[m]);
              py[m] = fma(cys[m], wy, py[m]);
              pz[m] = fma(czs[m], wz, pz[m]);
            }
          }
        }
        if (accl && PARITY) {
          for (int p = 0; p < nb; p++) {
            const double* v = sh[p];
            const double *s0x = v, *s0y = v + 5, *s0z = v + 10, *dsx = v + 15, *dsy = v + 20, *dsz = v + 25;
            double dstmp, pjtmp;
            // pjx(:, jp, kp)
            dstmp = ((s0y[ja] + 5e-1 * dsy[ja]) * s0z[kb] + (5e-1 * s0y[ja] + fac * dsy[ja]) * dsz[kb]) * qdxdt;
            pjtmp = -dsx[0] * dstmp;
            px[0] = px[0] + pjtmp;
            pjtmp = pjtmp - dsx[1] * dstmp;
            px[1] = px[1] + pjtmp;
            pjtmp = pjtmp - dsx[2] * dstmp;
            px[2] = px[2] + pjtmp;
            pjtmp = pjtmp - dsx[3] * dstmp;
            px[3] = px[3] + pjtmp;
            // pjy(:, jp, kp)
            dstmp = ((s0x[ja] + 5e-1 * dsx[ja]) * s0z[kb] + (5e-1 * s0x[ja] + fac * dsx[ja]) * dsz[kb]) * qdxdt;
            pjtmp = -dsy[0] * dstmp;
            py[0] = py[0] + pjtmp;
            pjtmp = pjtmp - dsy[1] * dstmp;
            py[1] = py[1] + pjtmp;
            pjtmp = pjtmp - dsy[2] * dstmp;
            py[2] = py[2] + pjtmp;
            pjtmp = pjtmp - dsy[3] * dstmp;
            py[3] = py[3] + pjtmp;
            // pjz(:, jp, kp)
            dstmp = ((s0x[ja] + 5e-1 * dsx[ja]) * s0y[kb] + (5e-1 * s0x[ja] + fac * dsx[ja]) * dsy[kb]) * qdxdt;
            pjtmp = -dsz[0] * dstmp;
            pz[0] = pz[0] + pjtmp;
            pjtmp = pjtmp - dsz[1] * dstmp;
            pz[1] = pz[1] + pjtmp;
            pjtmp = pjtmp - dsz[2] * dstmp;
            pz[2] = pz[2] + pjtmp;
            pjtmp = pjtmp - dsz[3] * dstmp;
            pz[3] = pz[3] + pjtmp;
          }
        }
        // before the next batch overwrites the shapes
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
    if (!live || !accl) continue;
    if (PARITY) {
      double* b = a.pjbuf + cidx3(g, i, j, k) * 300 + hl * 4;
#pragma unroll
      for (int m = 0; m < 4; m++) {
        b[m] = px[m];
        b[100 + m] = py[m];
        b[200 + m] = pz[m];
      }
    } else {
      // pjx(m-1, jp, kp) -> node (i+m-1, j+jp, k+kp); pjy -> (i+jp, j+m-1, k+kp);
      // pjz -> (i+jp, j+kp, k+m-1), with jp = ja-2, kp = kb-2
      const int bx = (li + 2), by = (lj + 2), bz = (lk + 2);
#pragma unroll
      for (int m = 0; m < 4; m++) {
        atomicAdd(&acc[0][((bz + kb - 2) * AH + (by + ja - 2)) * AW + (bx + m - 1)], px[m]);
        atomicAdd(&acc[1][((bz + kb - 2) * AH + (by + m - 1)) * AW + (bx + ja - 2)], py[m]);
        atomicAdd(&acc[2][((bz + m - 1) * AH + (by + kb - 2)) * AW + (bx + ja - 2)], pz[m]);
      }
    }
  }
  if (!PARITY) {
    __syncthreads();
    const size_t pl = plane3(g);
    for (int q = tid; q < 3 * AN; q += NT) {
      const int comp = q / AN, rem = q - comp * AN;
      const double v = acc[comp][rem];
      if (v == 0.0) continue;
      const int I = i0 - 2 + rem % AW, J = j0 - 2 + (rem / AW) % AH, K = k0 - 2 + rem / (AW * AH);
      if (I > g.nxge + 2 || J > g.nye + 2 || K > g.nze + 2) continue;
      unsafeAtomicAdd(a.uj + pl * comp + fidx3(g, I, J, K), v);
    }
  }
}

// uj(c, I, J, K) = sum over source cells in the reference's (k, j, i) order of
// their block entry, from +0 (field.f90:231-403 at one thread).  pjx(-2,.,.),
// pjy(-2,.,.) and pjz(-2,.,.) are structurally +0 and are skipped.
__global__ __launch_bounds__(256) void k_gather_pj3d(Grid3D g, const double* __restrict__ pjbuf,
                                                    double* __restrict__ uj) {
  const size_t n = plane3(g);
  const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int I = g.nxgs - 2 + (int)(q % g.fw), J = g.nys - 2 + (int)((q / g.fw) % g.fh),
            K = g.nzs - 2 + (int)(q / ((size_t)g.fw * g.fh));
  double sx = 0.0, sy = 0.0, sz = 0.0;
  const int klo = K - 2 > g.nzs ? K - 2 : g.nzs, khi = K + 2 < g.nze ? K + 2 : g.nze;
  const int jlo = J - 2 > g.nys ? J - 2 : g.nys, jhi = J + 2 < g.nye ? J + 2 : g.nye;
  const int ilo = I - 2 > g.nxgs ? I - 2 : g.nxgs, ihi = I + 2 < g.nxge ? I + 2 : g.nxge;
  for (int k = klo; k <= khi; k++)
    for (int j = jlo; j <= jhi; j++)
      for (int i = ilo; i <= ihi; i++) {
        const int ox = I - i, oy = J - j, oz = K - k;
        const double* b = pjbuf + cidx3(g, i, j, k) * 300;
        // lane (ja, kb) = (a + 2, b + 2) holds pj(m - 1, a, b) at [lane*4 + m]
        if (ox >= -1) sx = sx + b[((oz + 2) * 5 + (oy + 2)) * 4 + ox + 1];
        if (oy >= -1) sy = sy + b[100 + ((oz + 2) * 5 + (ox + 2)) * 4 + oy + 1];
        if (oz >= -1) sz = sz + b[200 + ((oy + 2) * 5 + (ox + 2)) * 4 + oz + 1];
      }
  uj[q] = sx;
  uj[n + q] = sy;
  uj[2 * n + q] = sz;
}

// Current of one particle that changes cell (field.f90:253-380, factored as
// below): the 4-node window [lo, lo+3] (of the 5 nodes -2..2) per axis holds
// every non-zero s0 and ds when |inc| <= 1; smo_* as field.f90:258-266.
// add(comp, xo, yo, zo, v) receives each non-zero value with its node offset
// from the particle's cell (i, j, k).
template <class Add>
__device__ inline void deposit3_window(const Grid3D& g, const Parts3& A, const Parts3& B, const int* ap, int pc,
                                       const int cell[3], double qd, Add add) {
  const int ia = ap ? ap[pc] : pc;
  const double old[3] = {A.x[ia], A.y[ia], A.z[ia]};
  const double nw[3] = {B.x[pc], B.y[pc], B.z[pc]};
  const double d_delx = g.d_delx;
  double S0[3][4], DS[3][4], cq[3][3];
  int lo[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    double s0m, s00, s0p, s11, s12, s13;
    shape3(old[d] * d_delx - 5e-1 - cell[d], s0m, s00, s0p);
    const int i2 = (int)(nw[d] * d_delx);
    const double dh = nw[d] * d_delx - 5e-1 - i2;
    const int in = i2 - cell[d];
    shape3(dh, s11, s12, s13);
    const double smo1 = in < 0 ? (double)(-in) : 0.0;
    const double smo2 = (double)(-abs(in) + 1);
    const double smo3 = in > 0 ? (double)in : 0.0;
    const double f0 = s11 * smo1 - 0e0;
    const double f1 = s11 * smo2 + s12 * smo1 - s0m;
    const double f2 = s12 * smo2 + s13 * smo1 + s11 * smo3 - s00;
    const double f3 = s13 * smo2 + s12 * smo3 - s0p;
    const double f4 = s13 * smo3 - 0e0;
    const bool left = in < 0;
    lo[d] = left ? 0 : 1;
    S0[d][0] = left ? 0.0 : s0m;
    S0[d][1] = left ? s0m : s00;
    S0[d][2] = left ? s00 : s0p;
    S0[d][3] = left ? s0p : 0.0;
    DS[d][0] = left ? f0 : f1;
    DS[d][1] = left ? f1 : f2;
    DS[d][2] = left ? f2 : f3;
    DS[d][3] = left ? f3 : f4;
    double run = 0.0;
#pragma unroll
    for (int m = 0; m < 3; m++) {
      run = run - DS[d][m];
      cq[d][m] = run * qd;
    }
  }
  // window node u of an axis is node lo + u - 2 of the cell; face m lies
  // between window nodes m and m + 1, at offset lo + m - 1
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const double ax = fma(0.5, DS[0][u], S0[0][u]), bxu = fma(1.0 / 3.0, DS[0][u], 0.5 * S0[0][u]);
    const double ay = fma(0.5, DS[1][u], S0[1][u]), byu = fma(1.0 / 3.0, DS[1][u], 0.5 * S0[1][u]);
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const double wx = fma(ay, S0[2][v], byu * DS[2][v]);  // y node u, z node v
      const double wy = fma(ax, S0[2][v], bxu * DS[2][v]);  // x node u, z node v
      const double wz = fma(ax, S0[1][v], bxu * DS[1][v]);  // x node u, y node v
#pragma unroll
      for (int m = 0; m < 3; m++) {
        const double vx = cq[0][m] * wx, vy = cq[1][m] * wy, vz = cq[2][m] * wz;
        if (vx != 0.0) add(0, lo[0] + m - 1, lo[1] + u - 2, lo[2] + v - 2, vx);
        if (vy != 0.0) add(1, lo[0] + u - 2, lo[1] + m - 1, lo[2] + v - 2, vy);
        if (vz != 0.0) add(2, lo[0] + u - 2, lo[1] + v - 2, lo[2] + m - 1, vz);
      }
    }
  }
}

// The same window from the particle's old and new in-cell shape arguments
// (dh = x/dx - 0.5 - cell per axis, field.f90:253-266) and its cell increments.
template <class Add>
__device__ inline void deposit3_window_dh(const double dh0[3], const double dh1[3], const int inc[3], double qd,
                                          Add add) {
  double S0[3][4], DS[3][4], cq[3][3];
  int lo[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    double s0m, s00, s0p, s11, s12, s13;
    shape3(dh0[d], s0m, s00, s0p);
    shape3(dh1[d], s11, s12, s13);
    const int in = inc[d];
    const double smo1 = in < 0 ? (double)(-in) : 0.0;
    const double smo2 = (double)(-abs(in) + 1);
    const double smo3 = in > 0 ? (double)in : 0.0;
    const double f0 = s11 * smo1 - 0e0;
    const double f1 = s11 * smo2 + s12 * smo1 - s0m;
    const double f2 = s12 * smo2 + s13 * smo1 + s11 * smo3 - s00;
    const double f3 = s13 * smo2 + s12 * smo3 - s0p;
    const double f4 = s13 * smo3 - 0e0;
    const bool left = in < 0;
    lo[d] = left ? 0 : 1;
    S0[d][0] = left ? 0.0 : s0m;
    S0[d][1] = left ? s0m : s00;
    S0[d][2] = left ? s00 : s0p;
    S0[d][3] = left ? s0p : 0.0;
    DS[d][0] = left ? f0 : f1;
    DS[d][1] = left ? f1 : f2;
    DS[d][2] = left ? f2 : f3;
    DS[d][3] = left ? f3 : f4;
    double run = 0.0;
#pragma unroll
    for (int m = 0; m < 3; m++) {
      run = run - DS[d][m];
      cq[d][m] = run * qd;
    }
  }
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const double ax = fma(0.5, DS[0][u], S0[0][u]), bxu = fma(1.0 / 3.0, DS[0][u], 0.5 * S0[0][u]);
    const double ay = fma(0.5, DS[1][u], S0[1][u]), byu = fma(1.0 / 3.0, DS[1][u], 0.5 * S0[1][u]);
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const double wx = fma(ay, S0[2][v], byu * DS[2][v]);  // y node u, z node v
      const double wy = fma(ax, S0[2][v], bxu * DS[2][v]);  // x node u, z node v
      const double wz = fma(ax, S0[1][v], bxu * DS[1][v]);  // x node u, y node v
#pragma unroll
      for (int m = 0; m < 3; m++) {
        const double vx = cq[0][m] * wx, vy = cq[1][m] * wy, vz = cq[2][m] * wz;
        if (vx != 0.0) add(0, lo[0] + m - 1, lo[1] + u - 2, lo[2] + v - 2, vx);
        if (vy != 0.0) add(1, lo[0] + u - 2, lo[1] + m - 1, lo[2] + v - 2, vy);
        if (vz != 0.0) add(2, lo[0] + u - 2, lo[1] + v - 2, lo[2] + m - 1, vz);
      }
    }
  }
}

// Fast mode, push and deposit in one pass over each particle.  The push is the
// shared bit-exact boris_push3; the deposit uses the factored weights of
// k_deposit3d's fast path,
//   Jx(m, j, k) += cx(m) (ay(j) s0z(k) + by(j) dsz(k)),  cx(m) = -sum_{n<=m} dsx(n),
//   ay = s0y + dsy/2,  by = s0y/2 + dsy/3,
// and Jy, Jz with the axes permuted (field.f90:335-380 expanded term by term).
// A particle that stays in its cell along every axis has ds = 0 on the outer
// nodes and cx = 0 on the outer faces (the last one up to rounding), so its
// current lives on 2 x 3 x 3 nodes per component: every lane accumulates
// those 54 values in registers over the particles of the cell it handles and
// the lane group reduces them once per cell.  A particle that changes cell is
// queued in LDS, private to the wave; after the wave's cells of a round, its
// lanes add the queued particles' 3 x 4 x 4 windows per component
// (deposit3_window) to the LDS tile.  The tile is added to uj with one HBM
// atomic per node.
namespace p3f {
constexpr int G = 8;     // lanes per cell
constexpr int QW = 128;  // queued cell-changing particles per wave and round (beyond: a global list)
}

template <bool VAY>
__global__ __launch_bounds__(p3::NT, 2) void k_push_deposit3d_fast(Push3Args a) {
  using namespace p3;
  constexpr int GF = p3f::G;
  __shared__ double sf[6][SN];
  __shared__ double acc[3][AN];
  __shared__ double spc[WPIC_MAXSP][4];  // fac1, txxx, fac2, q delx / delt
  // each wave's cell-changing particles of the round: old and new shape
  // arguments per axis, and cell in the tile | species << 7 | (inc + 1) per
  // axis << 9, 11, 13
  constexpr int NW = NT / 64;
  using p3f::QW;
  __shared__ double xqd[NW][QW][6];
  __shared__ int xqc[NW][QW];
  __shared__ int nxq[NW];
  const Grid3D& g = a.g;
  const int ntiles = a.ntx * a.nty * a.ntz;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  if (tile >= ntiles) return;
  int i0, j0, k0;
  tile_origin(g, tile, a.ntx, a.nty, i0, j0, k0);
  const int tid = threadIdx.x;
  for (int q = tid; q < 6 * SN; q += NT) {
    const int c = q / SN, rem = q - c * SN;
    const int i = i0 - 1 + rem % SW, j = j0 - 1 + (rem / SW) % SH, k = k0 - 1 + rem / (SW * SH);
    if (i <= g.nxge + 1 && j <= g.nye + 1 && k <= g.nze + 1) sf[c][rem] = tmpf3(g, a.uf, c, i, j, k);
  }
  for (int q = tid; q < 3 * AN; q += NT) (&acc[0][0])[q] = 0.0;
  if (tid < NW) nxq[tid] = 0;
#pragma unroll
  for (int s = 0; s < WPIC_MAXSP; s++) {
    if (tid == s && s < g.nsp) {
      const double q = g.q[s], rm = g.r[s];
      const double fac1 = q / rm * 5e-1 * g.delt;
      spc[s][0] = fac1;
      spc[s][1] = fac1 * fac1;
      spc[s][2] = q * g.delt / rm;
      spc[s][3] = q * g.delx * g.d_delt;
    }
  }
  __syncthreads();

  const int grp = tid / GF, lg = tid % GF;
  constexpr int NGRP = NT / GF;
  const double d_delx = g.d_delx, delt = g.delt, c = g.c;
  const int w = tid >> 6, lane = tid & 63;
  int* const nq_r = &nxq[w];
  // rounds of NGRP cells (one per group, 64 / G per wave), each wave
  // draining its own queue after its cells of the round
  for (int rnd = 0; rnd < (TX * TY * TZ) / NGRP; rnd++) {
    const int cc = rnd * NGRP + grp;
    const int li = cc % TX, lj = (cc / TX) % TY, lk = cc / (TX * TY);
    const int i = i0 + li, j = j0 + lj, k = k0 + lk;
    if (i <= g.nxge && j <= g.nye && k <= g.nze) {  // uniform within a group
    const int col = (j - g.nys) + g.ny * (k - g.nzs);
    CellList cl;
    int ntot = 0;
#pragma unroll
    for (int s = 0; s < WPIC_MAXSP; s++) {
      if (s < g.nsp) {
        const int rw = s * g.ncol + col;
        const int* cr = a.cumcnt + (size_t)rw * (g.nx + 1);
        const int c0 = cr[i - g.nxgs], c1 = cr[i - g.nxgs + 1];
        cl.set(s, a.colbase[rw] + c0, ntot);
        ntot += c1 - c0;
      } else {
        cl.set(s, 0, 0x7fffffff);
      }
    }
    // stayers' current: [m][a][b] with m the face (offset m along the
    // component's axis) and a, b the nodes (offsets a-1, b-1) of the other two
    // axes in the order (y, z) for Jx, (x, z) for Jy, (x, y) for Jz
    double jx[18], jy[18], jz[18];
#pragma unroll
    for (int q = 0; q < 18; q++) jx[q] = jy[q] = jz[q] = 0.0;
    const int so = (lk + 1) * (SW * SH) + (lj + 1) * SW + (li + 1);
    const int bx = li + 2, by = lj + 2, bz = lk + 2;  // the cell in the current tile

    // two-deep load pipeline: the record of element t + G and the
    // permutation entry of element t + 2G are in flight while element t is
    // pushed
    int ia_nn = 0;
    double px = 0.0, py = 0.0, pz = 0.0, pux = 0.0, puy = 0.0, puz = 0.0;
    uint64_t pid = 0;
    {
      int s_, p_, ia_n = 0;
      if (lg < ntot) {
        cl.slot(lg, s_, p_);
        ia_n = a.ap ? a.ap[p_] : p_;
      }
      if (lg + GF < ntot) {
        cl.slot(lg + GF, s_, p_);
        ia_nn = a.ap ? a.ap[p_] : p_;
      }
      if (lg < ntot) {
        px = a.A.x[ia_n]; py = a.A.y[ia_n]; pz = a.A.z[ia_n];
        pux = a.A.ux[ia_n]; puy = a.A.uy[ia_n]; puz = a.A.uz[ia_n];
        pid = a.A.id[ia_n];
      }
    }
    for (int t = lg; t < ntot; t += GF) {
      int sc, pc;
      cl.slot(t, sc, pc);
      const double x = px, y = py, z = pz, ux = pux, uy = puy, uz = puz;
      const uint64_t id = pid;
      if (t + GF < ntot) {
        const int ia = ia_nn;
        px = a.A.x[ia]; py = a.A.y[ia]; pz = a.A.z[ia];
        pux = a.A.ux[ia]; puy = a.A.uy[ia]; puz = a.A.uz[ia];
        pid = a.A.id[ia];
      }
      if (t + 2 * GF < ntot) {
        int s_, p_;
        cl.slot(t + 2 * GF, s_, p_);
        ia_nn = a.ap ? a.ap[p_] : p_;
      }
      const double* kc = spc[sc];
      double nw[3], g4, g5, g6, sh[9];
      if (VAY)
        vay_push3(sf, so, i, j, k, d_delx, delt, c, kc[0], kc[2], x, y, z, ux, uy, uz, nw[0], nw[1], nw[2], g4, g5,
                  g6, sh);
      else
        boris_push3(sf, so, i, j, k, d_delx, delt, c, kc[0], kc[1], kc[2], x, y, z, ux, uy, uz, nw[0], nw[1], nw[2],
                    g4, g5, g6, sh);
      a.B.ux[pc] = g4;
      a.B.uy[pc] = g5;
      a.B.uz[pc] = g6;
      a.B.id[pc] = id;

      // new shape and cell increment per axis (field.f90:253-329)
      const double qd = kc[3];
      const int cell[3] = {i, j, k};
      double s1[3][3], dh1[3];
      int inc[3];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        const int i2 = (int)(nw[d] * d_delx);
        dh1[d] = nw[d] * d_delx - 5e-1 - i2;
        inc[d] = i2 - cell[d];
        shape3(dh1[d], s1[d][0], s1[d][1], s1[d][2]);
      }
      bool over = false;
      if ((inc[0] | inc[1] | inc[2]) == 0) {
        // nodes -1..1 of each axis: s0, ds = s1 - s0, a = s0 + ds/2, b = s0/2 + ds/3
        double s0[3][3], ds[3][3], A[3][3], B[3][3], cq[3][2];
#pragma unroll
        for (int d = 0; d < 3; d++) {
#pragma unroll
          for (int n = 0; n < 3; n++) {
            s0[d][n] = sh[3 * d + n];
            ds[d][n] = s1[d][n] - s0[d][n];
            A[d][n] = fma(0.5, ds[d][n], s0[d][n]);
            B[d][n] = fma(1.0 / 3.0, ds[d][n], 0.5 * s0[d][n]);
          }
          cq[d][0] = (0.0 - ds[d][0]) * qd;
          cq[d][1] = ((0.0 - ds[d][0]) - ds[d][1]) * qd;
        }
#pragma unroll
        for (int u = 0; u < 3; u++)
#pragma unroll
          for (int v = 0; v < 3; v++) {
            const double wx = fma(A[1][u], s0[2][v], B[1][u] * ds[2][v]);  // y node u, z node v
            const double wy = fma(A[0][u], s0[2][v], B[0][u] * ds[2][v]);  // x node u, z node v
            const double wz = fma(A[0][u], s0[1][v], B[0][u] * ds[1][v]);  // x node u, y node v
#pragma unroll
            for (int m = 0; m < 2; m++) {
              const int q = (m * 3 + u) * 3 + v;
              jx[q] = fma(cq[0][m], wx, jx[q]);
              jy[q] = fma(cq[1][m], wy, jy[q]);
              jz[q] = fma(cq[2][m], wz, jz[q]);
            }
          }
      } else {
        // a particle that changes cell: its window is added after the cell loop
        const int slot = wave_enqueue(nq_r);
        if (slot < QW) {
          const double o[3] = {x, y, z};
#pragma unroll
          for (int d = 0; d < 3; d++) {
            xqd[w][slot][d] = o[d] * d_delx - 5e-1 - cell[d];
            xqd[w][slot][3 + d] = dh1[d];
          }
          xqc[w][slot] = cc | (sc << 7) | ((inc[0] + 1) << 9) | ((inc[1] + 1) << 11) | ((inc[2] + 1) << 13);
        } else {
          // beyond the LDS queue: listed for k_deposit3d_overflow, which
          // reads its new position unwrapped and wraps it after the deposit
          const int o = atomicAdd(a.nov, 1);
          a.ovslot[o] = pc;
          a.ovcell[o] = (int)cidx3(g, i, j, k) | (sc << 28);
          over = true;
        }
      }
      // the x wrap, the y-z wraps and the exchange code: particle_x and
      // particle_yz act on gp after the deposit, which uses nw unwrapped
      double wx = nw[0], wy = nw[1], wz = nw[2];
      if (!over) {
        bool moved;
        a.code[pc] = wrap3_code<true>(g, wx, wy, wz, j, k, a.err, moved);
      }
      a.B.x[pc] = wx;
      a.B.y[pc] = wy;
      a.B.z[pc] = wz;

    }
    // reduce the stayers' current over the group, spread the LDS atomics
#pragma unroll
    for (int q = 0; q < 54; q++) {
      double v = q < 18 ? jx[q] : (q < 36 ? jy[q - 18] : jz[q - 36]);
      static_assert(GF == 8, "the group is one half of a DPP row");
      v = row8_sum(v);
      if ((q % GF) == lg && v != 0.0) {
        const int comp = q / 18, r = q % 18, m = r / 9, u = (r / 3) % 3, b = r % 3;
        int xo, yo, zo;
        if (comp == 0) {
          xo = m; yo = u - 1; zo = b - 1;
        } else if (comp == 1) {
          xo = u - 1; yo = m; zo = b - 1;
        } else {
          xo = u - 1; yo = b - 1; zo = m;
        }
        atomicAdd(&acc[comp][((bz + zo) * AH + (by + yo)) * AW + (bx + xo)], v);
      }
    }
    }  // the cell

    // the wave's queued particles of the round: the queue was written by this
    // wave's own lanes, and LDS operations of one wave complete in order
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int nq = *nq_r < QW ? *nq_r : QW;
    for (int q = lane; q < nq; q += 64) {
      const int e = xqc[w][q];
      const int qcc = e & 127, s = (e >> 7) & 3;
      const int inc[3] = {((e >> 9) & 3) - 1, ((e >> 11) & 3) - 1, ((e >> 13) & 3) - 1};
      const int qbx = qcc % TX + 2, qby = (qcc / TX) % TY + 2, qbz = qcc / (TX * TY) + 2;
      const double dh0[3] = {xqd[w][q][0], xqd[w][q][1], xqd[w][q][2]};
      const double dh1[3] = {xqd[w][q][3], xqd[w][q][4], xqd[w][q][5]};
      deposit3_window_dh(dh0, dh1, inc, spc[s][3], [&](int comp, int xo, int yo, int zo, double v) {
        atomicAdd(&acc[comp][((qbz + zo) * AH + (qby + yo)) * AW + (qbx + xo)], v);
      });
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane == 0) *nq_r = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  __syncthreads();
  const size_t pl = plane3(g);
  for (int q = tid; q < 3 * AN; q += NT) {
    const int comp = q / AN, rem = q - comp * AN;
    const double v = acc[comp][rem];
    if (v == 0.0) continue;
    const int I = i0 - 2 + rem % AW, J = j0 - 2 + (rem / AW) % AH, K = k0 - 2 + rem / (AW * AH);
    if (I > g.nxge + 2 || J > g.nye + 2 || K > g.nze + 2) continue;
    unsafeAtomicAdd(a.uj + pl * comp + fidx3(g, I, J, K), v);
  }
}

// The cell-changing particles that did not fit a tile's LDS queue, added to
// uj with HBM atomics (ghost nodes included; the current boundary folds them).
__global__ __launch_bounds__(256) void k_deposit3d_overflow(Push3Args a) {
  const Grid3D& g = a.g;
  const int n = *a.nov;
  const size_t pl = plane3(g);
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const int pc = a.ovslot[q], pk = a.ovcell[q];
    const int s = pk >> 28, cix = pk & ((1 << 28) - 1);
    const int cell[3] = {g.nxgs + cix % g.nx, g.nys + (cix / g.nx) % g.ny, g.nzs + cix / (g.nx * g.ny)};
    double qs = g.q[0];
#pragma unroll
    for (int k = 1; k < WPIC_MAXSP; k++)
      if (s == k) qs = g.q[k];
    const double qd = qs * g.delx * g.d_delt;
    deposit3_window(g, a.A, a.B, a.ap, pc, cell, qd, [&](int comp, int xo, int yo, int zo, double v) {
      unsafeAtomicAdd(a.uj + pl * comp + fidx3(g, cell[0] + xo, cell[1] + yo, cell[2] + zo), v);
    });
    // the wraps and the code the fused kernel left to this deposit
    double x = a.B.x[pc], y = a.B.y[pc], z = a.B.z[pc];
    bool moved;
    a.code[pc] = wrap3_code<true>(g, x, y, z, cell[1], cell[2], a.err, moved);
    a.B.x[pc] = x;
    a.B.y[pc] = y;
    a.B.z[pc] = z;
  }
}

static Push3Args push_args(wpic_ctx3d* c) {
  Push3Args a;
  a.g = c->g;
  a.A = c->A;
  a.B = c->B;
  a.colbase = c->colbaseA;
  a.cumcnt = c->cumcnt;
  a.ap = c->aperm;
  a.nov = nullptr;
  a.ovslot = a.ovcell = nullptr;
  a.uf = c->uf;
  a.uj = c->uj;
  a.pjbuf = c->pjbuf;
  a.ntx = (c->g.nx + p3::TX - 1) / p3::TX;
  a.nty = (c->g.ny + p3::TY - 1) / p3::TY;
  a.ntz = (c->g.nz + p3::TZ - 1) / p3::TZ;
  a.vay = c->pusher == WPIC_PUSHER_VAY;
  a.code = c->code;
  a.err = c->err;
  return a;
}

hipError_t launch_push3d(wpic_ctx3d* c) {
  Push3Args a = push_args(c);
  const int nblk = xcd_grid(a.ntx * a.nty * a.ntz);
  hipLaunchKernelGGL(k_push3d, dim3(nblk), dim3(p3::NT), 0, c->st, a);
  return hipGetLastError();
}

hipError_t launch_deposit3d(wpic_ctx3d* c) {
  Push3Args a = push_args(c);
  const int nblk = xcd_grid(a.ntx * a.nty * a.ntz);
  if (c->mode == WPIC_MODE_PARITY)
    hipLaunchKernelGGL(k_deposit3d<true>, dim3(nblk), dim3(p3::NT), 0, c->st, a);
  else
    hipLaunchKernelGGL(k_deposit3d<false>, dim3(nblk), dim3(p3::NT), 0, c->st, a);
  return hipGetLastError();
}

__global__ void k_zero_int3(int* p) { *p = 0; }

// The exchange scratch (R, holes) is free between the sort and the next
// exchange; the fused kernel lists its queue overflow there.
hipError_t launch_push_deposit3d(wpic_ctx3d* c) {
  Push3Args a = push_args(c);
  a.nov = c->nov;
  a.ovslot = c->R;
  a.ovcell = c->holes;
  const int nblk = xcd_grid(a.ntx * a.nty * a.ntz);
  hipLaunchKernelGGL(k_zero_int3, dim3(1), dim3(1), 0, c->st, c->nov);
  if (a.vay)
    hipLaunchKernelGGL(k_push_deposit3d_fast<true>, dim3(nblk), dim3(p3::NT), 0, c->st, a);
  else
    hipLaunchKernelGGL(k_push_deposit3d_fast<false>, dim3(nblk), dim3(p3::NT), 0, c->st, a);
  hipLaunchKernelGGL(k_deposit3d_overflow, dim3(256), dim3(256), 0, c->st, a);
  return hipGetLastError();
}

hipError_t launch_gather_pj3d(wpic_ctx3d* c) {
  const size_t n = plane3(c->g);
  hipLaunchKernelGGL(k_gather_pj3d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->st, c->g, c->pjbuf, c->uj);
  return hipGetLastError();
}

}  // namespace wpic
