// k_mixed3d.hip -- the mixed-precision form of the 3D fused push + deposit +
// x wrap (the physics of 3d/common/particle.f90:52-233 and
// 3d/common/field.f90:211-406 followed by the particle boundaries).
//
// A particle is an int32 x cell plus FP32 offsets (x - ix, y - j, z - k) in
// [0, 1) and FP32 momenta, 36 B with the id: positions keep ~1e-7 of a cell
// everywhere in the box and the shape functions see offsets directly.  Push
// and deposit run in FP32 (FMA contraction on); each particle's current is
// formed in FP32 and accumulated in an FP64 LDS tile, the tile is added to
// the FP64 uj with FP64 atomics and the field solve is the FP64 one.  The work mapping is the FP64
// fused kernel's (k_push_deposit3d_fast): 8x4x4-cell tiles, 8 lanes per cell,
// the 54 current values of a particle that stays in its cell in registers, a
// particle that changes cell queued (per wave) with its offsets and added with
// its 3 x 4 x 4 window per component after the wave's round of cells.  The push writes the exchange code
// ((new x cell - nxgs) << 4 | y-z direction) itself: the x wrap is a cell
// wrap and the y-z offsets are already relative to the destination column.
#include "wpic_internal3d.h"

namespace wpic {
// the tile accumulates in FP64: LDS FP32 atomics serialize badly when the
// lanes of a wave add to the same node, FP64 ones do not
__device__ inline void tile_add(double* p, float v) { atomicAdd(p, (double)v); }

namespace m3 {
constexpr int TX = 8, TY = 4, TZ = 4, NT = 256, G = 8;
constexpr int SW = TX + 2, SH = TY + 2, SD = TZ + 2, SN = SW * SH * SD;  // staged tmpf
constexpr int AW = TX + 4, AH = TY + 4, AD = TZ + 4, AN = AW * AH * AD;  // current tile
constexpr int QW = 128;    // queued cell-changing particles per wave and round (beyond: added at once)
}  // namespace m3

struct Mix3Args {
  Grid3D g;
  PartsM3 A, B;
  const int* colbase;
  const int* cumcnt;
  const int* ap;  // A's slot -> record (null: identity)
  const double* uf;
  double* uj;
  int* code;
  int* err;
  int ntx, nty, ntz;
};

// tmpf(c, i, j, k) of particle.f90:74-91 in FP64, rounded to FP32 by the caller
__device__ inline double tmpf3m(const Grid3D& g, const double* uf, int c, int i, int j, int k) {
  const size_t pl = plane3(g);
  const double* u = uf + pl * c;
  const size_t o = fidx3(g, i, j, k);
  const size_t dx = 1, dy = (size_t)g.fw, dz = (size_t)g.fw * g.fh;
  switch (c) {
    case 0: return 0.25 * (u[o] + u[o + dy] + u[o + dz] + u[o + dy + dz]);
    case 1: return 0.25 * (u[o] + u[o + dx] + u[o + dz] + u[o + dx + dz]);
    case 2: return 0.25 * (u[o] + u[o + dx] + u[o + dy] + u[o + dx + dy]);
    case 3: return 0.5 * (u[o] + u[o + dx]);
    case 4: return 0.5 * (u[o] + u[o + dy]);
    default: return 0.5 * (u[o] + u[o + dz]);
  }
}

__device__ inline void shape3m(float dh, float& s0, float& s1, float& s2) {
  s0 = 0.5f * (0.5f - dh) * (0.5f - dh);
  s1 = 0.75f - dh * dh;
  s2 = 0.5f * (0.5f + dh) * (0.5f + dh);
}

struct Mix3Const {
  float delt, ic2, cc2;  // dt, 1/c^2, c^2
};

// The push of one particle in FP32 (particle.f90:93-217 in the cell + offset
// form): old shape sh (x, y, z triplets), new offsets f and cell increments inc.
struct Mix3Push {
  float sh[9];
  float f[3];
  int inc[3];
  float g4, g5, g6, gmv;
};

__device__ inline Mix3Push mixed_push3(const Mix3Const& mc, const float (*sf)[m3::SN], int so, const float* kc,
                                       float ox, float oy, float oz, float ux, float uy, float uz) {
#pragma clang fp contract(fast)
  using namespace m3;
  const float fac1 = kc[0], txxx = kc[1], fac2 = kc[2];
  Mix3Push r;
  shape3m(ox - 0.5f, r.sh[0], r.sh[1], r.sh[2]);
  shape3m(oy - 0.5f, r.sh[3], r.sh[4], r.sh[5]);
  shape3m(oz - 0.5f, r.sh[6], r.sh[7], r.sh[8]);
  asm volatile("" : "+v"(so));
  float fl[6];
#pragma unroll
  for (int cm = 0; cm < 6; cm++) {
    const float* f = &sf[cm][so];
#define RX(o) (f[(o)-1] * r.sh[0] + f[o] * r.sh[1] + f[(o) + 1] * r.sh[2])
#define RY(o) (RX((o)-SW) * r.sh[3] + RX(o) * r.sh[4] + RX((o) + SW) * r.sh[5])
    fl[cm] = RY(-SW * SH) * r.sh[6] + RY(0) * r.sh[7] + RY(SW * SH) * r.sh[8];
#undef RY
#undef RX
  }
  const float bpx = fl[0], bpy = fl[1], bpz = fl[2], epx = fl[3], epy = fl[4], epz = fl[5];
  float uvm1 = ux + fac1 * epx;
  float uvm2 = uy + fac1 * epy;
  float uvm3 = uz + fac1 * epz;
  const float gam = sqrtf(mc.cc2 + uvm1 * uvm1 + uvm2 * uvm2 + uvm3 * uvm3);
  const float igam = 1.0f / gam;
  const float fac1r = fac1 * igam;
  const float fac2r = fac2 / (gam + txxx * (bpx * bpx + bpy * bpy + bpz * bpz) * igam);
  const float uvm4 = uvm1 + fac1r * (uvm2 * bpz - uvm3 * bpy);
  const float uvm5 = uvm2 + fac1r * (uvm3 * bpx - uvm1 * bpz);
  const float uvm6 = uvm3 + fac1r * (uvm1 * bpy - uvm2 * bpx);
  uvm1 = uvm1 + fac2r * (uvm5 * bpz - uvm6 * bpy);
  uvm2 = uvm2 + fac2r * (uvm6 * bpx - uvm4 * bpz);
  uvm3 = uvm3 + fac2r * (uvm4 * bpy - uvm5 * bpx);
  r.g4 = uvm1 + fac1 * epx;
  r.g5 = uvm2 + fac1 * epy;
  r.g6 = uvm3 + fac1 * epz;
  r.gmv = 1.0f / sqrtf(1.0f + (r.g4 * r.g4 + r.g5 * r.g5 + r.g6 * r.g6) * mc.ic2);
  const float o[3] = {ox, oy, oz}, u[3] = {r.g4, r.g5, r.g6};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const float np = o[d] + u[d] * mc.delt * r.gmv;
    const float fi = floorf(np);
    r.inc[d] = (int)fi;
    r.f[d] = np - fi;
    if (r.f[d] >= 1.0f) {
      r.f[d] -= 1.0f;
      r.inc[d] += 1;
    }
  }
  return r;
}

// Current of one particle that changes cell, with the 4-node window per axis
// of k_push_deposit3d_fast (|inc| <= 1): add(comp, xo, yo, zo, v) receives
// each non-zero value with its node offset from the particle's old cell.
template <class Add>
__device__ inline void deposit3_window_m(const Mix3Push& r, float qd, Add add) {
#pragma clang fp contract(fast)
  float S0[3][4], DS[3][4], cq[3][3];
  int lo[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int in = r.inc[d];
    float s11, s12, s13;
    shape3m(r.f[d] - 0.5f, s11, s12, s13);
    const float a = (float)abs(in);
    const float smo1 = in < 0 ? a : 0.0f, smo2 = 1.0f - a, smo3 = in > 0 ? a : 0.0f;
    const float s0m = r.sh[3 * d], s00 = r.sh[3 * d + 1], s0p = r.sh[3 * d + 2];
    const float f0 = s11 * smo1;
    const float f1 = s11 * smo2 + s12 * smo1 - s0m;
    const float f2 = s12 * smo2 + s13 * smo1 + s11 * smo3 - s00;
    const float f3 = s13 * smo2 + s12 * smo3 - s0p;
    const float f4 = s13 * smo3;
    const bool left = in < 0;
    lo[d] = left ? 0 : 1;
    S0[d][0] = left ? 0.0f : s0m;
    S0[d][1] = left ? s0m : s00;
    S0[d][2] = left ? s00 : s0p;
    S0[d][3] = left ? s0p : 0.0f;
    DS[d][0] = left ? f0 : f1;
    DS[d][1] = left ? f1 : f2;
    DS[d][2] = left ? f2 : f3;
    DS[d][3] = left ? f3 : f4;
    float run = 0.0f;
#pragma unroll
    for (int m = 0; m < 3; m++) {
      run -= DS[d][m];
      cq[d][m] = run * qd;
    }
  }
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const float ax = S0[0][u] + 0.5f * DS[0][u], bx = 0.5f * S0[0][u] + (1.0f / 3.0f) * DS[0][u];
    const float ay = S0[1][u] + 0.5f * DS[1][u], by = 0.5f * S0[1][u] + (1.0f / 3.0f) * DS[1][u];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const float wx = ay * S0[2][v] + by * DS[2][v];  // y node u, z node v
      const float wy = ax * S0[2][v] + bx * DS[2][v];  // x node u, z node v
      const float wz = ax * S0[1][v] + bx * DS[1][v];  // x node u, y node v
#pragma unroll
      for (int m = 0; m < 3; m++) {
        const float vx = cq[0][m] * wx, vy = cq[1][m] * wy, vz = cq[2][m] * wz;
        if (vx != 0.0f) add(0, lo[0] + m - 1, lo[1] + u - 2, lo[2] + v - 2, vx);
        if (vy != 0.0f) add(1, lo[0] + u - 2, lo[1] + m - 1, lo[2] + v - 2, vy);
        if (vz != 0.0f) add(2, lo[0] + u - 2, lo[1] + v - 2, lo[2] + m - 1, vz);
      }
    }
  }
}

__device__ inline void tile_origin_m(const Grid3D& g, int tile, int ntx, int nty, int& i0, int& j0, int& k0) {
  const int tx = tile % ntx, rest = tile / ntx;
  i0 = g.nxgs + tx * m3::TX;
  j0 = g.nys + (rest % nty) * m3::TY;
  k0 = g.nzs + (rest / nty) * m3::TZ;
}

__global__ __launch_bounds__(m3::NT, 3) void k_push_deposit_mixed3d(Mix3Args a) {
#pragma clang fp contract(fast)
  using namespace m3;
  __shared__ float sf[6][SN];
  __shared__ double acc[3][AN];
  __shared__ float spc[WPIC_MAXSP][4];  // fac1, txxx, fac2, q delx / delt
  // each wave's cell-changing particles of the round: old and new in-cell
  // offsets, and cell in the tile | species << 7 | (inc + 1) per axis << 9,
  // 11, 13
  constexpr int NW = NT / 64;
  __shared__ float xqo[NW][QW][6];
  __shared__ int xqc[NW][QW];
  __shared__ int nxq[NW];
  const Grid3D& g = a.g;
  const int ntiles = a.ntx * a.nty * a.ntz;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  if (tile >= ntiles) return;
  int i0, j0, k0;
  tile_origin_m(g, tile, a.ntx, a.nty, i0, j0, k0);
  const int tid = threadIdx.x;
  for (int q = tid; q < 6 * SN; q += NT) {
    const int c = q / SN, rem = q - c * SN;
    const int i = i0 - 1 + rem % SW, j = j0 - 1 + (rem / SW) % SH, k = k0 - 1 + rem / (SW * SH);
    if (i <= g.nxge + 1 && j <= g.nye + 1 && k <= g.nze + 1) sf[c][rem] = (float)tmpf3m(g, a.uf, c, i, j, k);
  }
  for (int q = tid; q < 3 * AN; q += NT) (&acc[0][0])[q] = 0.0f;
  if (tid < NW) nxq[tid] = 0;
#pragma unroll
  for (int s = 0; s < WPIC_MAXSP; s++) {
    if (tid == s && s < g.nsp) {
      const float fac1 = (float)(g.q[s] / g.r[s] * 0.5 * g.delt);
      spc[s][0] = fac1;
      spc[s][1] = fac1 * fac1;
      spc[s][2] = (float)(g.q[s] * g.delt / g.r[s]);
      spc[s][3] = (float)(g.q[s] * g.delx * g.d_delt);
    }
  }
  __syncthreads();

  const int grp = tid / G, lg = tid % G;
  constexpr int NGRP = NT / G;
  Mix3Const mc;
  mc.delt = (float)g.delt;
  mc.ic2 = (float)(1.0 / (g.c * g.c));
  mc.cc2 = (float)(g.c * g.c);
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
    const int so = (lk + 1) * (SW * SH) + (lj + 1) * SW + (li + 1);
    const int bx = li + 2, by = lj + 2, bz = lk + 2;
    // stayers' current: [m][u][v], face m, nodes u - 1, v - 1 of the other two
    // axes in the order (y, z) for Jx, (x, z) for Jy, (x, y) for Jz
    float jx[18], jy[18], jz[18];
#pragma unroll
    for (int q = 0; q < 18; q++) jx[q] = jy[q] = jz[q] = 0.0f;
    for (int s = 0; s < g.nsp; s++) {
      const int rw = s * g.ncol + col;
      const int* cr = a.cumcnt + (size_t)rw * (g.nx + 1);
      const int base = a.colbase[rw];
      const int pbeg = base + cr[i - g.nxgs], pend = base + cr[i - g.nxgs + 1];
      const float* kc = spc[s];
      const float qd = kc[3];
      for (int pi = pbeg + lg; pi < pend; pi += G) {
        const int ia = a.ap ? a.ap[pi] : pi;
        const float o[3] = {a.A.ox[ia], a.A.oy[ia], a.A.oz[ia]};
        const Mix3Push r = mixed_push3(mc, sf, so, kc, o[0], o[1], o[2], a.A.ux[ia], a.A.uy[ia], a.A.uz[ia]);
        a.B.ux[pi] = r.g4;
        a.B.uy[pi] = r.g5;
        a.B.uz[pi] = r.g6;
        a.B.id[pi] = a.A.id[ia];
        int dj = r.inc[1], dk = r.inc[2];
        if (r.inc[0] < -1 || r.inc[0] > 1 || dj < -1 || dj > 1 || dk < -1 || dk > 1) {
          atomicOr(a.err, ERR_DOMAIN);
          dj = dj < -1 ? -1 : (dj > 1 ? 1 : dj);
          dk = dk < -1 ? -1 : (dk > 1 ? 1 : dk);
        }
        if ((r.inc[0] | r.inc[1] | r.inc[2]) == 0) {
          float s0[3][3], ds[3][3], A[3][3], B[3][3], cq[3][2];
#pragma unroll
          for (int d = 0; d < 3; d++) {
            float s1[3];
            shape3m(r.f[d] - 0.5f, s1[0], s1[1], s1[2]);
#pragma unroll
            for (int n = 0; n < 3; n++) {
              s0[d][n] = r.sh[3 * d + n];
              ds[d][n] = s1[n] - s0[d][n];
              A[d][n] = s0[d][n] + 0.5f * ds[d][n];
              B[d][n] = 0.5f * s0[d][n] + (1.0f / 3.0f) * ds[d][n];
            }
            cq[d][0] = -ds[d][0] * qd;
            cq[d][1] = (-ds[d][0] - ds[d][1]) * qd;
          }
#pragma unroll
          for (int u = 0; u < 3; u++)
#pragma unroll
            for (int v = 0; v < 3; v++) {
              const float wx = A[1][u] * s0[2][v] + B[1][u] * ds[2][v];
              const float wy = A[0][u] * s0[2][v] + B[0][u] * ds[2][v];
              const float wz = A[0][u] * s0[1][v] + B[0][u] * ds[1][v];
#pragma unroll
              for (int m = 0; m < 2; m++) {
                const int q = (m * 3 + u) * 3 + v;
                jx[q] += cq[0][m] * wx;
                jy[q] += cq[1][m] * wy;
                jz[q] += cq[2][m] * wz;
              }
            }
        } else {
          const int slot = wave_enqueue(nq_r);
          if (slot < QW) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
              xqo[w][slot][d] = o[d];
              xqo[w][slot][3 + d] = r.f[d];
            }
            xqc[w][slot] = cc | (s << 7) | ((r.inc[0] + 1) << 9) | ((r.inc[1] + 1) << 11) | ((r.inc[2] + 1) << 13);
          } else {
            deposit3_window_m(r, qd, [&](int comp, int xo, int yo, int zo, float v) {
              tile_add(&acc[comp][((bz + zo) * AH + (by + yo)) * AW + (bx + xo)], v);
            });
          }
        }
        // the x wrap is a cell wrap; the y-z step goes to the exchange
        int cx = i + r.inc[0];
        if (cx < g.nxgs) cx += g.nx;
        else if (cx > g.nxge) cx -= g.nx;
        if (cx < g.nxgs || cx > g.nxge) {
          atomicOr(a.err, ERR_DOMAIN);
          cx = cx < g.nxgs ? g.nxgs : g.nxge;
        }
        a.B.ix[pi] = cx;
        a.B.ox[pi] = r.f[0];
        a.B.oy[pi] = r.f[1];
        a.B.oz[pi] = r.f[2];
        a.code[pi] = ((cx - g.nxgs) << CODE_SHIFT3) | ((dk + 1) * 3 + (dj + 1));
      }
    }
    // reduce the stayers' current over the group, spread the LDS atomics
#pragma unroll
    for (int q = 0; q < 54; q++) {
      float v = q < 18 ? jx[q] : (q < 36 ? jy[q - 18] : jz[q - 36]);
      static_assert(G == 8, "the group is one half of a DPP row");
      v = row8_sum(v);
      if ((q % G) == lg && v != 0.0f) {
        const int comp = q / 18, rr = q % 18, m = rr / 9, u = (rr / 3) % 3, b = rr % 3;
        int xo, yo, zo;
        if (comp == 0) {
          xo = m; yo = u - 1; zo = b - 1;
        } else if (comp == 1) {
          xo = u - 1; yo = m; zo = b - 1;
        } else {
          xo = u - 1; yo = b - 1; zo = m;
        }
        tile_add(&acc[comp][((bz + zo) * AH + (by + yo)) * AW + (bx + xo)], v);
      }
    }
    }  // the cell

    // the wave's queued particles of the round, from their offsets: the queue
    // was written by this wave's own lanes, and LDS operations of one wave
    // complete in order
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int nq = *nq_r < QW ? *nq_r : QW;
    for (int q = lane; q < nq; q += 64) {
      const int e = xqc[w][q];
      const int qcc = e & 127, s = (e >> 7) & 3;
      const int qbx = qcc % TX + 2, qby = (qcc / TX) % TY + 2, qbz = qcc / (TX * TY) + 2;
      Mix3Push r;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        shape3m(xqo[w][q][d] - 0.5f, r.sh[3 * d], r.sh[3 * d + 1], r.sh[3 * d + 2]);
        r.f[d] = xqo[w][q][3 + d];
        r.inc[d] = ((e >> (9 + 2 * d)) & 3) - 1;
      }
      deposit3_window_m(r, spc[s][3], [&](int comp, int xo, int yo, int zo, float v) {
        tile_add(&acc[comp][((qbz + zo) * AH + (qby + yo)) * AW + (qbx + xo)], v);
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

hipError_t launch_push_deposit_mixed3d(wpic_ctx3d* c) {
  Mix3Args a;
  a.g = c->g;
  a.A = c->MA;
  a.B = c->MB;
  a.colbase = c->colbaseA;
  a.cumcnt = c->cumcnt;
  a.ap = c->aperm;
  a.uf = c->uf;
  a.uj = c->uj;
  a.code = c->code;
  a.err = c->err;
  a.ntx = (c->g.nx + m3::TX - 1) / m3::TX;
  a.nty = (c->g.ny + m3::TY - 1) / m3::TY;
  a.ntz = (c->g.nz + m3::TZ - 1) / m3::TZ;
  const int nblk = xcd_grid(a.ntx * a.nty * a.ntz);
  hipLaunchKernelGGL(k_push_deposit_mixed3d, dim3(nblk), dim3(m3::NT), 0, c->st, a);
  return hipGetLastError();
}

}  // namespace wpic
