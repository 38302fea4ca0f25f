// k_mom3d.hip -- velocity moments of the 3D particles (3d/common/mom_calc.f90)
// and their periodic boundary folds (3d/common/boundary_periodic.f90:
// 1102-1235).
//
// mom(7, nxgs-1:nxge+1, nys-1:nye+1, nzs-1:nze+1, nsp) in the reference's
// own layout.  The same three deposits as in 2D:
//   k_mom3d_tile<KICK>  fast: per tile of cells an LDS node tile (one-node
//                       halo), lane groups per cell; with KICK
//                       the half-step kick of mom_calc__accl is applied on the
//                       fly and gp is never written;
//   k_mom3d_any         fast, any slot order: HBM atomics;
//   k_mom3d_parity      one thread per node in the reference's serial
//                       (k, j, ii) order: bit-identical to mom_calc__nvt.
// The 3D nvt takes floor() for the lower node and the offset from the
// position itself, up(1)-5d-1-ih (mom_calc.f90:243-251).
#include "k_push3d.h"

namespace wpic {

struct Mom3Args {
  Grid3D g;
  Parts3 A, B;
  const int *colbaseA, *np2A, *colbaseB, *np2B;
  const int* cumcnt;
  const int* ap;
  const double* uf;
  double* mom;
  int* err;
  int ntx, nty, ntz;
  double fac1[WPIC_MAXSP], txxx[WPIC_MAXSP], fac2[WPIC_MAXSP];  // mom_calc.f90:99-101 with delt/2
};

__host__ __device__ inline size_t midx3(const Grid3D& g, int i, int j, int k, int s) {
  return (size_t)7 *
         ((size_t)(i - (g.nxgs - 1)) +
          (size_t)(g.nx + 2) * ((size_t)(j - (g.nys - 1)) + (size_t)(g.ny + 2) * ((size_t)(k - (g.nzs - 1)) + (size_t)(g.nz + 2) * s)));
}

__device__ inline void stage_fields3(const Grid3D& g, const double* uf, double (*sf)[p3::SN], int i0, int j0, int k0) {
  using namespace p3;
  for (int q = threadIdx.x; q < 6 * SN; q += NT) {
    const int c = q / SN, rem = q - c * SN;
    const int i = i0 - 1 + rem % SW, j = j0 - 1 + (rem / SW) % SH, k = k0 - 1 + rem / (SW * SH);
    if (i <= g.nxge + 1 && j <= g.nye + 1 && k <= g.nze + 1) sf[c][rem] = tmpf3(g, uf, c, i, j, k);
  }
}

// the seven moments of one particle times wx*wy*wz, in the reference's
// product order (mom_calc.f90:253-324)
struct Mom3Terms {
  double u[3], gam;
  __device__ void set(double ux, double uy, double uz, double c) {
    u[0] = ux;
    u[1] = uy;
    u[2] = uz;
    gam = 1e0 / sqrt(1e0 + (+ux * ux + uy * uy + uz * uz) / (c * c));
  }
  __device__ double term(int k, double wx, double wy, double wz) const {
    if (k == 0) return wx * wy * wz;
    if (k < 4) return u[k - 1] * gam * wx * wy * wz;
    return u[k - 4] * u[k - 4] * gam * wx * wy * wz;
  }
};

// lower node and offsets of one particle (mom_calc.f90:240-251)
struct Cic3 {
  int ih, jh, kh;
  double d[3];
  __device__ void set(const Grid3D& g, double x, double y, double z) {
    ih = (int)floor(x * g.d_delx - 5e-1);
    jh = (int)floor(y * g.d_delx - 5e-1);
    kh = (int)floor(z * g.d_delx - 5e-1);
    d[0] = x - 5e-1 - ih;
    d[1] = y - 5e-1 - jh;
    d[2] = z - 5e-1 - kh;
  }
  // the particle's nodes lie in the grid and its column reaches them
  __device__ bool ok(const Grid3D& g, int j, int k) const {
    return ih >= g.nxgs - 1 && ih <= g.nxge && jh >= j - 1 && jh <= j && kh >= k - 1 && kh <= k;
  }
  __device__ double w(int a, int o) const { return o ? d[a] : 1e0 - d[a]; }
};

// ---------------------------------------------------------------------------
// mom_calc__accl (mom_calc.f90:49-216): gp(1:3) = up(1:3), gp(4:6) kicked over
// delt/2; slot-aligned with up, the id travels along
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(p3::NT) void k_mom_accl3d(Mom3Args a) {
  using namespace p3;
  __shared__ double sf[6][SN];
  const Grid3D& g = a.g;
  const int ntiles = a.ntx * a.nty * a.ntz;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  if (tile >= ntiles) return;
  int i0, j0, k0;
  tile_origin(g, tile, a.ntx, a.nty, i0, j0, k0);
  stage_fields3(g, a.uf, sf, i0, j0, k0);
  __syncthreads();
  constexpr int NGRP = NT / G;
  const int grp = threadIdx.x / G, lg = threadIdx.x % G;
  const double hdelt = g.delt * 5e-1;
  for (int cc = grp; cc < TX * TY * TZ; cc += NGRP) {
    const int li = cc % TX, lj = (cc / TX) % TY, lk = cc / (TX * TY);
    const int i = i0 + li, j = j0 + lj, k = k0 + lk;
    if (i > g.nxge || j > g.nye || k > g.nze) continue;
    const int so = ((lk + 1) * SH + (lj + 1)) * SW + (li + 1);
    const int col = (j - g.nys) + g.ny * (k - g.nzs);
    for (int s = 0; s < g.nsp; s++) {
      const int cs = s * g.ncol + col;
      const int* cr = a.cumcnt + (size_t)cs * (g.nx + 1);
      const int base = a.colbaseA[cs];
      const int pend = base + cr[i - g.nxgs + 1];
      for (int pi = base + cr[i - g.nxgs] + lg; pi < pend; pi += G) {
        const int ia = a.ap ? a.ap[pi] : pi;
        const double x = a.A.x[ia], y = a.A.y[ia], z = a.A.z[ia];
        double nxp, nyp, nzp, g4, g5, g6, sh[9];
        boris_push3(sf, so, i, j, k, g.d_delx, hdelt, g.c, a.fac1[s], a.txxx[s], a.fac2[s], x, y, z, a.A.ux[ia],
                    a.A.uy[ia], a.A.uz[ia], nxp, nyp, nzp, g4, g5, g6, sh);
        a.B.x[pi] = x;
        a.B.y[pi] = y;
        a.B.z[pi] = z;
        a.B.ux[pi] = g4;
        a.B.uy[pi] = g5;
        a.B.uz[pi] = g6;
        a.B.id[pi] = a.A.id[ia];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// fast deposit over tiles.  The node tile covers i0-1 .. i0+TX (and so on in
// y, z).  A gp particle outside the cells its slot belongs to goes to HBM
// directly.
// ---------------------------------------------------------------------------
namespace m3 {
constexpr int MW = p3::TX + 2, MH = p3::TY + 2, MD = p3::TZ + 2, MN = MW * MH * MD;
}  // namespace m3

#ifndef MOM3_G
#define MOM3_G 2
#endif

// the seven moments rotated by r (out[q] = m[(q + r) % 7]), r in 0..6
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
__global__ __launch_bounds__(p3::NT) void k_mom3d_tile(Mom3Args a) {
  using namespace p3;
  using namespace m3;
  constexpr int GM = MOM3_G, NGRP = NT / GM;  // lanes per cell, cells in flight
  __shared__ double sf[KICK ? 6 : 1][SN];
  __shared__ double mt[7][MN];
  const Grid3D& g = a.g;
  const int ntiles = a.ntx * a.nty * a.ntz;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  if (tile >= ntiles) return;
  int i0, j0, k0;
  tile_origin(g, tile, a.ntx, a.nty, i0, j0, k0);
  const int ncx = min(TX, g.nxge + 1 - i0), ncy = min(TY, g.nye + 1 - j0), ncz = min(TZ, g.nze + 1 - k0);
  const int tid = threadIdx.x;
  if constexpr (KICK) stage_fields3(g, a.uf, sf, i0, j0, k0);
  const double hdelt = g.delt * 5e-1, c = g.c;
  const int grp = tid / GM, lg = tid % GM;
  // lanes of a group share a cell and so its nodes: each walks the seven
  // moment planes in an order rotated by its lane index to spread the LDS
  // atomics
  const int rot = tid % 7;

  for (int s = 0; s < g.nsp; s++) {
    for (int q = tid; q < 7 * MN; q += NT) (&mt[0][0])[q] = 0.0;
    __syncthreads();
    for (int cc = grp; cc < TX * TY * TZ; cc += NGRP) {
      const int li = cc % TX, lj = (cc / TX) % TY, lk = cc / (TX * TY);
      if (li >= ncx || lj >= ncy || lk >= ncz) continue;  // uniform within a group
      const int i = i0 + li, j = j0 + lj, k = k0 + lk;
      const int so = ((lk + 1) * SH + (lj + 1)) * SW + (li + 1);
      const int cs = s * g.ncol + (j - g.nys) + g.ny * (k - g.nzs);
      const int* cr = a.cumcnt + (size_t)cs * (g.nx + 1);
      const int base = a.colbaseA[cs];
      const int pe = base + cr[i - g.nxgs + 1];
      for (int pi = base + cr[i - g.nxgs] + lg; pi < pe; pi += GM) {
        double x, y, z, ux, uy, uz;
        if constexpr (KICK) {
          const int ia = a.ap ? a.ap[pi] : pi;
          x = a.A.x[ia];
          y = a.A.y[ia];
          z = a.A.z[ia];
          double nxp, nyp, nzp, sh[9];
          boris_push3(sf, so, i, j, k, g.d_delx, hdelt, c, a.fac1[s], a.txxx[s], a.fac2[s], x, y, z, a.A.ux[ia],
                      a.A.uy[ia], a.A.uz[ia], nxp, nyp, nzp, ux, uy, uz, sh);
        } else {
          x = a.B.x[pi];
          y = a.B.y[pi];
          z = a.B.z[pi];
          ux = a.B.ux[pi];
          uy = a.B.uy[pi];
          uz = a.B.uz[pi];
        }
        Cic3 w;
        w.set(g, x, y, z);
        if (!w.ok(g, j, k)) {
          atomicOr(a.err, ERR_DOMAIN);
          continue;
        }
        Mom3Terms m;
        m.set(ux, uy, uz, c);
        const int lx = w.ih - (i0 - 1), ly = w.jh - (j0 - 1), lz = w.kh - (k0 - 1);
        if (lx < 0 || lx > ncx) {
#pragma unroll
          for (int n = 0; n < 8; n++) {
            const int ox = n & 1, oy = (n >> 1) & 1, oz = n >> 2;
            double* p = a.mom + midx3(g, w.ih + ox, w.jh + oy, w.kh + oz, s);
            const double wx = w.w(0, ox), wy = w.w(1, oy), wz = w.w(2, oz);
#pragma unroll
            for (int q = 0; q < 7; q++) unsafeAtomicAdd(p + q, m.term(q, wx, wy, wz));
          }
          continue;
        }
        double mv[7] = {1e0,
                        m.u[0] * m.gam,
                        m.u[1] * m.gam,
                        m.u[2] * m.gam,
                        m.u[0] * m.u[0] * m.gam,
                        m.u[1] * m.u[1] * m.gam,
                        m.u[2] * m.u[2] * m.gam};
        rotate7(mv, rot);
        double wn[8];
#pragma unroll
        for (int n = 0; n < 8; n++) wn[n] = w.w(0, n & 1) * w.w(1, (n >> 1) & 1) * w.w(2, n >> 2);
        double* node = &mt[0][(lz * MH + ly) * MW + lx];
#pragma unroll
        for (int q = 0; q < 7; q++) {
          const int plane = q + rot < 7 ? q + rot : q + rot - 7;
          double* p = node + plane * MN;
#pragma unroll
          for (int n = 0; n < 8; n++) atomicAdd(p + ((n >> 2) * MH + ((n >> 1) & 1)) * MW + (n & 1), mv[q] * wn[n]);
        }
      }
    }
    __syncthreads();
    for (int q = tid; q < 7 * MN; q += NT) {
      const int node = q / 7, comp = q % 7;
      const double v = mt[comp][node];
      if (v == 0.0) continue;
      const int I = i0 - 1 + node % MW, J = j0 - 1 + (node / MW) % MH, K = k0 - 1 + node / (MW * MH);
      if (I > g.nxge + 1 || J > g.nye + 1 || K > g.nze + 1) continue;
      unsafeAtomicAdd(a.mom + midx3(g, I, J, K, s) + comp, v);
    }
    __syncthreads();
  }
}

// gp in any slot order: one block per (species, column), HBM atomics
__global__ __launch_bounds__(p3::NT) void k_mom3d_any(Mom3Args a) {
  const Grid3D& g = a.g;
  const int cs = blockIdx.x, s = cs / g.ncol, col = cs % g.ncol;
  const int j = g.nys + col % g.ny, k = g.nzs + col / g.ny;
  const int base = a.colbaseB[cs], n = a.np2B[cs];
  for (int t = threadIdx.x; t < n; t += p3::NT) {
    const int pi = base + t;
    Cic3 w;
    w.set(g, a.B.x[pi], a.B.y[pi], a.B.z[pi]);
    if (!w.ok(g, j, k)) {
      atomicOr(a.err, ERR_DOMAIN);
      continue;
    }
    Mom3Terms m;
    m.set(a.B.ux[pi], a.B.uy[pi], a.B.uz[pi], g.c);
#pragma unroll
    for (int nn = 0; nn < 8; nn++) {
      const int ox = nn & 1, oy = (nn >> 1) & 1, oz = nn >> 2;
      double* p = a.mom + midx3(g, w.ih + ox, w.jh + oy, w.kh + oz, s);
      const double wx = w.w(0, ox), wy = w.w(1, oy), wz = w.w(2, oz);
#pragma unroll
      for (int q = 0; q < 7; q++) unsafeAtomicAdd(p + q, m.term(q, wx, wy, wz));
    }
  }
}

// ---------------------------------------------------------------------------
// parity: one thread per node (n, m, l) and species, adding the particles of
// columns (m-1..m+1, l-1..l+1) in the reference's (k, j, ii) order; with gp
// sorted like up only cells n-1..n+1 of each column are scanned.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(p3::NT) void k_mom3d_parity(Mom3Args a, int aligned) {
  const Grid3D& g = a.g;
  const long long t = (long long)blockIdx.x * p3::NT + threadIdx.x;
  const int mw = g.nx + 2, mh = g.ny + 2, md = g.nz + 2;
  if (t >= (long long)mw * mh * md * g.nsp) return;
  const int n = g.nxgs - 1 + (int)(t % mw);
  const int m = g.nys - 1 + (int)((t / mw) % mh);
  const int l = g.nzs - 1 + (int)((t / ((long long)mw * mh)) % md);
  const int s = (int)(t / ((long long)mw * mh * md));
  double acc[7];
#pragma unroll
  for (int q = 0; q < 7; q++) acc[q] = 0e0;
  for (int k = max(g.nzs, l - 1); k <= min(g.nze, l + 1); k++)
    for (int j = max(g.nys, m - 1); j <= min(g.nye, m + 1); j++) {
      const int cs = s * g.ncol + (j - g.nys) + g.ny * (k - g.nzs);
      const int base = a.colbaseB[cs];
      int pb = 0, pe = a.np2B[cs];
      if (aligned) {
        const int* cr = a.cumcnt + (size_t)cs * (g.nx + 1);
        pb = cr[max(g.nxgs, n - 1) - g.nxgs];
        pe = cr[min(g.nxge, n + 1) + 1 - g.nxgs];
      }
      for (int p = pb; p < pe; p++) {
        const int pi = base + p;
        Cic3 w;
        w.set(g, a.B.x[pi], a.B.y[pi], a.B.z[pi]);
        if (!w.ok(g, j, k)) {
          atomicOr(a.err, ERR_DOMAIN);
          continue;
        }
        const int ox = n - w.ih, oy = m - w.jh, oz = l - w.kh;
        if (ox < 0 || ox > 1 || oy < 0 || oy > 1 || oz < 0 || oz > 1) continue;
        Mom3Terms tm;
        tm.set(a.B.ux[pi], a.B.uy[pi], a.B.uz[pi], g.c);
        const double wx = w.w(0, ox), wy = w.w(1, oy), wz = w.w(2, oz);
#pragma unroll
        for (int q = 0; q < 7; q++) acc[q] = acc[q] + tm.term(q, wx, wy, wz);
      }
    }
  double* o = a.mom + midx3(g, n, m, l, s);
#pragma unroll
  for (int q = 0; q < 7; q++) o[q] = acc[q];
}

// ---------------------------------------------------------------------------
// periodic folds at one rank: x over everything; y (nys-1 onto nye, then
// nye+1 onto nys) over all x nodes and z planes; z (nzs-1 onto nze, then
// nze+1 onto nzs) over all x nodes and interior y
// ---------------------------------------------------------------------------
__global__ void k_mom3d_bc_x(Grid3D g, double* mom) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int mh = g.ny + 2, md = g.nz + 2;
  if (t >= 7LL * mh * md * g.nsp) return;
  const int q = (int)(t % 7), j = g.nys - 1 + (int)((t / 7) % mh), k = g.nzs - 1 + (int)((t / (7 * mh)) % md);
  const int s = (int)(t / (7LL * mh * md));
  double* lo = mom + midx3(g, g.nxgs, j, k, s) + q;
  double* hi = mom + midx3(g, g.nxge, j, k, s) + q;
  *lo = *lo + mom[midx3(g, g.nxge + 1, j, k, s) + q];
  *hi = *hi + mom[midx3(g, g.nxgs - 1, j, k, s) + q];
}

__global__ void k_mom3d_bc_y(Grid3D g, double* mom) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int mw = g.nx + 2, md = g.nz + 2;
  if (t >= 7LL * mw * md * g.nsp) return;
  const int q = (int)(t % 7), i = g.nxgs - 1 + (int)((t / 7) % mw), k = g.nzs - 1 + (int)((t / (7 * mw)) % md);
  const int s = (int)(t / (7LL * mw * md));
  double* hi = mom + midx3(g, i, g.nye, k, s) + q;
  *hi = *hi + mom[midx3(g, i, g.nys - 1, k, s) + q];
  double* lo = mom + midx3(g, i, g.nys, k, s) + q;
  *lo = *lo + mom[midx3(g, i, g.nye + 1, k, s) + q];
}

__global__ void k_mom3d_bc_z(Grid3D g, double* mom) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int mw = g.nx + 2;
  if (t >= 7LL * mw * g.ny * g.nsp) return;
  const int q = (int)(t % 7), i = g.nxgs - 1 + (int)((t / 7) % mw), j = g.nys + (int)((t / (7 * mw)) % g.ny);
  const int s = (int)(t / (7LL * mw * g.ny));
  double* hi = mom + midx3(g, i, j, g.nze, s) + q;
  *hi = *hi + mom[midx3(g, i, j, g.nzs - 1, s) + q];
  double* lo = mom + midx3(g, i, j, g.nzs, s) + q;
  *lo = *lo + mom[midx3(g, i, j, g.nze + 1, s) + q];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static Mom3Args mom3_args(wpic_ctx3d* c) {
  Mom3Args a;
  a.g = c->g;
  a.A = c->A;
  a.B = c->B;
  a.colbaseA = c->colbaseA;
  a.np2A = c->np2A;
  a.colbaseB = c->colbaseB;
  a.np2B = c->np2B;
  a.cumcnt = c->cumcnt;
  a.ap = c->aperm;
  a.uf = c->uf;
  a.mom = c->mom;
  a.err = c->err;
  a.ntx = (c->g.nx + p3::TX - 1) / p3::TX;
  a.nty = (c->g.ny + p3::TY - 1) / p3::TY;
  a.ntz = (c->g.nz + p3::TZ - 1) / p3::TZ;
  // mom_calc__init: delt = delt_in*5d-1 (mom_calc.f90:37); :99-101
  const double delt = c->g.delt * 5e-1;
  for (int s = 0; s < WPIC_MAXSP; s++) {
    const double q = c->g.q[s], r = c->g.r[s];
    a.fac1[s] = s < c->g.nsp ? q / r * 5e-1 * delt : 0.0;
    a.txxx[s] = a.fac1[s] * a.fac1[s];
    a.fac2[s] = s < c->g.nsp ? q * delt / r : 0.0;
  }
  return a;
}

size_t mom3_size(const Grid3D& g) { return (size_t)7 * (g.nx + 2) * (g.ny + 2) * (g.nz + 2) * g.nsp; }

hipError_t launch_mom_accl3d(wpic_ctx3d* c) {
  Mom3Args a = mom3_args(c);
  hipLaunchKernelGGL(k_mom_accl3d, dim3(xcd_grid(a.ntx * a.nty * a.ntz)), dim3(p3::NT), 0, c->st, a);
  return hipGetLastError();
}

hipError_t launch_mom_deposit3d(wpic_ctx3d* c, int kind) {
  Mom3Args a = mom3_args(c);
  const Grid3D& g = c->g;
  if (kind != MOM_PARITY_ALIGNED && kind != MOM_PARITY_ANY) {
    hipError_t e = launch_fill3d(c, c->mom, mom3_size(g), 0.0);
    if (e != hipSuccess) return e;
  }
  const int nt = a.ntx * a.nty * a.ntz;
  switch (kind) {
    case MOM_TILE_KICK:
      hipLaunchKernelGGL(k_mom3d_tile<true>, dim3(xcd_grid(nt)), dim3(p3::NT), 0, c->st, a);
      break;
    case MOM_TILE_GP:
      hipLaunchKernelGGL(k_mom3d_tile<false>, dim3(xcd_grid(nt)), dim3(p3::NT), 0, c->st, a);
      break;
    case MOM_ANY:
      hipLaunchKernelGGL(k_mom3d_any, dim3(g.nsp * g.ncol), dim3(p3::NT), 0, c->st, a);
      break;
    default: {
      const long long nn = (long long)(g.nx + 2) * (g.ny + 2) * (g.nz + 2) * g.nsp;
      hipLaunchKernelGGL(k_mom3d_parity, dim3((unsigned)((nn + p3::NT - 1) / p3::NT)), dim3(p3::NT), 0, c->st, a,
                         kind == MOM_PARITY_ALIGNED ? 1 : 0);
      break;
    }
  }
  return hipGetLastError();
}

hipError_t launch_mom_bc3d(wpic_ctx3d* c) {
  const Grid3D& g = c->g;
  const long long tx = 7LL * (g.ny + 2) * (g.nz + 2) * g.nsp, ty = 7LL * (g.nx + 2) * (g.nz + 2) * g.nsp,
                  tz = 7LL * (g.nx + 2) * g.ny * g.nsp;
  hipLaunchKernelGGL(k_mom3d_bc_x, dim3((unsigned)((tx + 255) / 256)), dim3(256), 0, c->st, g, c->mom);
  hipLaunchKernelGGL(k_mom3d_bc_y, dim3((unsigned)((ty + 255) / 256)), dim3(256), 0, c->st, g, c->mom);
  hipLaunchKernelGGL(k_mom3d_bc_z, dim3((unsigned)((tz + 255) / 256)), dim3(256), 0, c->st, g, c->mom);
  return hipGetLastError();
}

}  // namespace wpic
