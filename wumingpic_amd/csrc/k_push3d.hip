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
// the reference's dstmp / running pjtmp arithmetic (field.f90:335-380).
// Parity mode writes the 300-entry block per cell; k_gather_pj3d then adds the
// blocks into uj in the reference's (k, j, i) cell order with the scatter
// permutation uj(2) += pjy(jp, ip, kp), uj(3) += pjz(kp, ip, jp) (:390-398),
// bit for bit.  Fast mode adds the blocks into an LDS tile accumulator and
// the tile into uj with atomics.
#include "wpic_internal3d.h"

namespace wpic {

namespace p3 {
constexpr int TX = 8, TY = 4, TZ = 4, NT = 256, G = 4;
constexpr int SW = TX + 2, SH = TY + 2, SD = TZ + 2, SN = SW * SH * SD;  // staged tmpf
constexpr int AW = TX + 4, AH = TY + 4, AD = TZ + 4, AN = AW * AH * AD;  // current tile
constexpr int NB = 16;   // particles per shape batch of a half-wave
constexpr int NSV = 30;  // s0x s0y s0z dsx dsy dsz, 5 each
}  // namespace p3

struct Push3Args {
  Grid3D g;
  Parts3 A, B;
  const int* colbase;
  const int* cumcnt;
  const double* uf;
  double* uj;
  double* pjbuf;
  int ntx, nty, ntz;
};

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
        const double x = a.A.x[pi], y = a.A.y[pi], z = a.A.z[pi];
        double shxm, shx, shxp, shym, shy, shyp, shzm, shz, shzp;
        shape3(x * d_delx - 5e-1 - i, shxm, shx, shxp);
        shape3(y * d_delx - 5e-1 - j, shym, shy, shyp);
        shape3(z * d_delx - 5e-1 - k, shzm, shz, shzp);
        int so = (lk + 1) * (SW * SH) + (lj + 1) * SW + (li + 1);
        asm volatile("" : "+v"(so));
        double fl[6];
#pragma unroll
        for (int cm = 0; cm < 6; cm++) {
          const double* f = &sf[cm][so];
#define RX(o) (+f[(o)-1] * shxm + f[o] * shx + f[(o) + 1] * shxp)
#define RY(o) (+RX((o)-SW) * shym + RX(o) * shy + RX((o) + SW) * shyp)
          fl[cm] = RY(-SW * SH) * shzm + RY(0) * shz + RY(SW * SH) * shzp;
#undef RY
#undef RX
        }
        const double bpx = fl[0], bpy = fl[1], bpz = fl[2], epx = fl[3], epy = fl[4], epz = fl[5];
        double uvm1 = a.A.ux[pi] + fac1 * epx;
        double uvm2 = a.A.uy[pi] + fac1 * epy;
        double uvm3 = a.A.uz[pi] + fac1 * epz;
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
        const double g4 = uvm1 + fac1 * epx;
        const double g5 = uvm2 + fac1 * epy;
        const double g6 = uvm3 + fac1 * epz;
        gam = 1e0 / sqrt(1e0 + (+g4 * g4 + g5 * g5 + g6 * g6) / (c * c));
        a.B.x[pi] = x + g4 * delt * gam;
        a.B.y[pi] = y + g5 * delt * gam;
        a.B.z[pi] = z + g6 * delt * gam;
        a.B.ux[pi] = g4;
        a.B.uy[pi] = g5;
        a.B.uz[pi] = g6;
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
          const double old[3] = {a.A.x[pi], a.A.y[pi], a.A.z[pi]};
          const double nw[3] = {a.B.x[pi], a.B.y[pi], a.B.z[pi]};
          double v[NSV];
          shapes3(g, old, nw, cell, v);
#pragma unroll
          for (int q = 0; q < NSV; q++) sh[hl][q] = v[q];
        }
        // the shapes were written by other lanes of this wave
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (accl) {
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

static Push3Args push_args(wpic_ctx3d* c) {
  Push3Args a;
  a.g = c->g;
  a.A = c->A;
  a.B = c->B;
  a.colbase = c->colbaseA;
  a.cumcnt = c->cumcnt;
  a.uf = c->uf;
  a.uj = c->uj;
  a.pjbuf = c->pjbuf;
  a.ntx = (c->g.nx + p3::TX - 1) / p3::TX;
  a.nty = (c->g.ny + p3::TY - 1) / p3::TY;
  a.ntz = (c->g.nz + p3::TZ - 1) / p3::TZ;
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

hipError_t launch_gather_pj3d(wpic_ctx3d* c) {
  const size_t n = plane3(c->g);
  hipLaunchKernelGGL(k_gather_pj3d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->st, c->g, c->pjbuf, c->uj);
  return hipGetLastError();
}

}  // namespace wpic
