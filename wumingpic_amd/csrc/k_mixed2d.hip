// k_mixed2d.hip -- the mixed-precision form of the fused push + deposit +
// wrap (the same physics as 2d/common/particle.f90:48-179 and
// 2d/common/field.f90:189-316 followed by the particle boundaries).
//
// A particle is an int32 x cell plus FP32 offsets (x - ix, y - row) in [0, 1)
// and FP32 momenta: positions keep ~1e-7 of a cell everywhere in the box, the
// shape functions see offsets directly, and a particle record is 32 B instead
// of 48 B.  Push and deposit run in FP32 (FMA contraction on) with the tile's
// current accumulated in FP32 in LDS; the tile is added to the FP64 uj with
// FP64 atomics (the global current reduction stays FP64), and the field solve
// is the FP64 one.  Work mapping as the FP64 fast kernel (one workgroup per
// 32x4-cell tile) with 32-lane groups per cell, so that a group's loads and
// stores of each 4-byte record array cover 128 contiguous bytes per pass as
// the FP64 kernel's 16-lane groups do.
#include "wpic_internal.h"

namespace wpic {

namespace mx {
constexpr int TX = 32, TY = 4, NT = 256, G = 32;
constexpr int SFW = TX + 2, SFH = TY + 2;
constexpr int ACW = TX + 4, ACH = TY + 4;
}  // namespace mx

struct MixArgs {
  Grid2D g;
  PartsM A, B;
  const int* rowbase;
  const int* cumcnt;
  const int* ap;  // A's slot -> record (null: identity)
  const double* uf;
  double* uj;
  int* code;
  int* err;
  int ntx, nty;
  int* rcnt;  // particles per (species, row, x cell) after the step, or null
};

// cell-centred field component c at (i, j) (particle.f90:74-80), FP64
__device__ inline double cell_field_m(const Grid2D& g, const double* uf, int c, int i, int j) {
  const size_t pl = (size_t)g.fw * g.fh;
  const double* u = uf + pl * c;
  switch (c) {
    case 0: return 0.5 * (u[fidx(g, i, j)] + u[fidx(g, i, j + 1)]);
    case 1: return 0.5 * (u[fidx(g, i, j)] + u[fidx(g, i + 1, j)]);
    case 2: return 0.25 * (u[fidx(g, i, j)] + u[fidx(g, i + 1, j)] + u[fidx(g, i, j + 1)] + u[fidx(g, i + 1, j + 1)]);
    case 3: return 0.5 * (u[fidx(g, i, j)] + u[fidx(g, i + 1, j)]);
    case 4: return 0.5 * (u[fidx(g, i, j)] + u[fidx(g, i, j + 1)]);
    default: return u[fidx(g, i, j)];
  }
}

__device__ inline void shape3f(float dh, float& s0, float& s1, float& s2) {
  s0 = 0.5f * (0.5f - dh) * (0.5f - dh);
  s1 = 0.75f - dh * dh;
  s2 = 0.5f * (0.5f + dh) * (0.5f + dh);
}

// ds = (new shape shifted by inc) - s0, index 0..4 <-> -2..2, with the
// reference's smo_* weights (field.f90:244-266); |inc| <= 1 except for a
// particle that a reflecting wall sends across two cells, where the formula is
// applied as the reference applies it
__device__ inline void ds_of(float s0m, float s00, float s0p, float fnew, int inc, float ds[5]) {
  float s1m, s10, s1p;
  shape3f(fnew - 0.5f, s1m, s10, s1p);
  const float a = (float)abs(inc);
  const float smo1 = inc < 0 ? a : 0.0f, smo2 = 1.0f - a, smo3 = inc > 0 ? a : 0.0f;
  ds[0] = s1m * smo1;
  ds[1] = s1m * smo2 + s10 * smo1 - s0m;
  ds[2] = s10 * smo2 + s1p * smo1 + s1m * smo3 - s00;
  ds[3] = s1p * smo2 + s10 * smo3 - s0p;
  ds[4] = s1p * smo3;
}

// FP32 reciprocal and reciprocal square root of the hardware (about one ulp):
// the mixed form's arithmetic is FP32 throughout, so correctly rounded
// division sequences buy nothing its 1e-7 offsets can hold
__device__ inline float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float rsqf(float x) { return __builtin_amdgcn_rsqf(x); }

// Every lane keeps the full 5x5x3 FP32 block of the cell it walks (no LDS
// traffic per particle), the group reduces it once per cell.  Work mapping:
// wave w walks row w of the tile two x-adjacent cells at a time (one per
// 32-lane group), so the records of particles that crossed a cell or row
// boundary are read while the lines they share with their old neighbours
// are in L2.
template <int BC>
__global__ __launch_bounds__(mx::NT, 2) void k_push_deposit_mixed(MixArgs a) {
#pragma clang fp contract(fast)
  using namespace mx;
  __shared__ float sf[6][SFW * SFH];
  __shared__ float acc[3][ACW * ACH];
  __shared__ float spc[WPIC_MAXSP][5];  // fac1, txxx, fac2, qk, q
  // particles per (species, row, cell) after the step for the tile's cells
  // and their one-cell halo (rcnt)
  __shared__ int hcn[WPIC_MAXSP][TY + 2][TX + 2];
  const bool counting = a.rcnt != nullptr;

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
    if (i <= g.nxge + 1 && j <= g.nye + 1) sf[c][rem] = (float)cell_field_m(g, a.uf, c, i, j);
  }
  for (int k = tid; k < 3 * ACW * ACH; k += NT) (&acc[0][0])[k] = 0.0f;
  if (counting)
    for (int k = tid; k < WPIC_MAXSP * (TY + 2) * (TX + 2); k += NT) (&hcn[0][0][0])[k] = 0;
  // per-species constants (particle.f90:92-94, field.f90:274), formed in FP64
  // and rounded once
  if (tid < g.nsp && tid < WPIC_MAXSP) {
    const double q = g.q[tid], rm = g.r[tid];
    const float fac1 = (float)(q / rm * 0.5 * g.delt);
    spc[tid][0] = fac1;
    spc[tid][1] = fac1 * fac1;
    spc[tid][2] = (float)(q * g.delt / rm);
    spc[tid][3] = (float)(q * g.delx * g.d_delt);
    spc[tid][4] = (float)q;
  }
  __syncthreads();

  const int grp = tid / G, lg = tid % G;
  constexpr int NGRP = NT / G;
  const float delt = (float)g.delt;
  const float ic2 = (float)(1.0 / (g.c * g.c));
  const float cc2 = (float)(g.c * g.c);

  constexpr int RX = NGRP / TY;  // a round is an RX x TY block of cells
  static_assert(RX * TY == NGRP && TX % RX == 0, "rounds must tile the tile");
  for (int rnd = 0; rnd < TX / RX; rnd++) {
    const int li = rnd * RX + grp % RX, lj = grp / RX;
    const int i = i0 + li, j = j0 + lj;
    if (i > g.nxge || j > g.nye) continue;  // uniform within a group
    float ax[20], ay[20], az[25];
#pragma unroll
    for (int k = 0; k < 20; k++) ax[k] = ay[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 25; k++) az[k] = 0.0f;

    // the cell's particles over all species as one list, walked with a
    // two-deep software pipeline: the record of element t + G and the
    // permutation entry of element t + 2G are in flight while element t is
    // pushed
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
    float pox = 0.0f, poy = 0.0f, pux = 0.0f, puy = 0.0f, puz = 0.0f;
    uint64_t pid = 0;
    {
      int s_, p_, ia_n = 0;
      if (lg < tot) {
        cl.slot(lg, s_, p_);
        ia_n = a.ap ? a.ap[p_] : p_;
      }
      if (lg + G < tot) {
        cl.slot(lg + G, s_, p_);
        ia_nn = a.ap ? a.ap[p_] : p_;
      }
      if (lg < tot) {
        pox = a.A.ox[ia_n]; poy = a.A.oy[ia_n];
        pux = a.A.ux[ia_n]; puy = a.A.uy[ia_n]; puz = a.A.uz[ia_n];
        pid = a.A.id[ia_n];
      }
    }
    {
      for (int t = lg; t < tot; t += G) {
        int s, pi;
        cl.slot(t, s, pi);
        const float ox = pox, oy = poy, u1 = pux, u2 = puy, u3 = puz;
        const uint64_t id = pid;
        if (t + G < tot) {
          const int ia = ia_nn;
          pox = a.A.ox[ia]; poy = a.A.oy[ia];
          pux = a.A.ux[ia]; puy = a.A.uy[ia]; puz = a.A.uz[ia];
          pid = a.A.id[ia];
        }
        if (t + 2 * G < tot) {
          int s_, p_;
          cl.slot(t + 2 * G, s_, p_);
          ia_nn = a.ap ? a.ap[p_] : p_;
        }
        const float* kc = spc[s];
        const float fac1 = kc[0], txxx = kc[1], fac2 = kc[2], qk = kc[3], qf = kc[4];
        float sx0, sx1, sx2, sy0, sy1, sy2;
        shape3f(ox - 0.5f, sx0, sx1, sx2);
        shape3f(oy - 0.5f, sy0, sy1, sy2);
        int so = (lj + 1) * SFW + (li + 1);
        asm volatile("" : "+v"(so));
        float fl[6];
#pragma unroll
        for (int cm = 0; cm < 6; cm++) {
          const float* f = &sf[cm][so];
          fl[cm] = (f[-SFW - 1] * sx0 + f[-SFW] * sx1 + f[-SFW + 1] * sx2) * sy0 +
                   (f[-1] * sx0 + f[0] * sx1 + f[1] * sx2) * sy1 +
                   (f[SFW - 1] * sx0 + f[SFW] * sx1 + f[SFW + 1] * sx2) * sy2;
        }
        const float bpx = fl[0], bpy = fl[1], bpz = fl[2], epx = fl[3], epy = fl[4], epz = fl[5];
        float uvm1 = u1 + fac1 * epx;
        float uvm2 = u2 + fac1 * epy;
        float uvm3 = u3 + fac1 * epz;
        const float gam2 = cc2 + uvm1 * uvm1 + uvm2 * uvm2 + uvm3 * uvm3;
        const float igam = rsqf(gam2);
        const float gam = gam2 * igam;
        const float fac1r = fac1 * igam;
        const float fac2r = fac2 * rcpf(gam + txxx * (bpx * bpx + bpy * bpy + bpz * bpz) * igam);
        const float uvm4 = uvm1 + fac1r * (uvm2 * bpz - uvm3 * bpy);
        const float uvm5 = uvm2 + fac1r * (uvm3 * bpx - uvm1 * bpz);
        const float uvm6 = uvm3 + fac1r * (uvm1 * bpy - uvm2 * bpx);
        uvm1 = uvm1 + fac2r * (uvm5 * bpz - uvm6 * bpy);
        uvm2 = uvm2 + fac2r * (uvm6 * bpx - uvm4 * bpz);
        uvm3 = uvm3 + fac2r * (uvm4 * bpy - uvm5 * bpx);
        float g3 = uvm1 + fac1 * epx, g4 = uvm2 + fac1 * epy, g5 = uvm3 + fac1 * epz;
        const float gmv = rsqf(1.0f + (g3 * g3 + g4 * g4 + g5 * g5) * ic2);
        // new position as cell + offset
        const float nx_ = ox + g3 * delt * gmv, ny_ = oy + g4 * delt * gmv;
        float fxi = floorf(nx_), fyi = floorf(ny_);
        int incx = (int)fxi, incy = (int)fyi;
        float fx = nx_ - fxi, fy = ny_ - fyi;
        if (fx >= 1.0f) { fx -= 1.0f; incx += 1; }
        if (fy >= 1.0f) { fy -= 1.0f; incy += 1; }
        int cx = i + incx;
        // reflecting walls act before the deposit (reconnection order,
        // boundary_reconnection.f90:79-91): x' = 2(nxs+1) - x or 2(nxe-1) - x
        if (BC == WPIC_BC_WALLX && (cx < g.nxgs + 1 || cx >= g.nxge - 1)) {
          const int W = cx < g.nxgs + 1 ? 2 * (g.nxgs + 1) : 2 * (g.nxge - 1);
          cx = W - cx - 1;
          fx = 1.0f - fx;
          if (fx >= 1.0f) { fx -= 1.0f; cx += 1; }
          g3 = -g3; g4 = -g4; g5 = -g5;
          incx = cx - i;
        }
        a.B.ux[pi] = g3;
        a.B.uy[pi] = g4;
        a.B.uz[pi] = g5;
        a.B.id[pi] = id;
        if (incy < -1 || incy > 1) {
          atomicOr(a.err, ERR_DOMAIN);
          incy = incy < -1 ? -1 : 1;
        }

        // Esirkepov weights from offsets, factored as the FP64 fast kernel
        float dsx[5], dsy[5];
        ds_of(sx0, sx1, sx2, fx, incx, dsx);
        ds_of(sy0, sy1, sy2, fy, incy, dsy);
        const float s0x[5] = {0.0f, sx0, sx1, sx2, 0.0f}, s0y[5] = {0.0f, sy0, sy1, sy2, 0.0f};
        const float qvz = qf * (g5 * gmv);
        float wx[5], wy[5], A[5], B[5];
#pragma unroll
        for (int k = 0; k < 5; k++) {
          const float ty = s0y[k] + 0.5f * dsy[k];
          wy[k] = qk * ty;
          A[k] = qvz * ty;
          B[k] = qvz * (0.5f * s0y[k] + (1.0f / 3.0f) * dsy[k]);
          wx[k] = qk * (s0x[k] + 0.5f * dsx[k]);
        }
        float cxs = 0.0f, cys = 0.0f;
#pragma unroll
        for (int m = 0; m < 4; m++) {
          cxs -= dsx[m];
          cys -= dsy[m];
#pragma unroll
          for (int jp = 0; jp < 5; jp++) ax[jp * 4 + m] += cxs * wy[jp];
#pragma unroll
          for (int ip = 0; ip < 5; ip++) ay[m * 5 + ip] += cys * wx[ip];
        }
#pragma unroll
        for (int jp = 0; jp < 5; jp++)
#pragma unroll
          for (int ip = 0; ip < 5; ip++) az[jp * 5 + ip] += s0x[ip] * A[jp] + dsx[ip] * B[jp];

        // periodic x wrap of the cell; the row step goes to the exchange
        if (BC == WPIC_BC_PERIODIC) {
          if (cx < g.nxgs) cx += g.nx;
          else if (cx > g.nxge) cx -= g.nx;
        }
        if (cx < g.nxgs || cx > g.nxge) {
          atomicOr(a.err, ERR_DOMAIN);
          cx = cx < g.nxgs ? g.nxgs : g.nxge;
        }
        a.B.ix[pi] = cx;
        a.B.ox[pi] = fx;
        a.B.oy[pi] = fy;
        const int code = ((cx - g.nxgs) << 2) | (incy + 1);
        a.code[pi] = code;
        if (counting && code != (((i - g.nxgs) << 2) | 1)) count_leaver<TX, TY>(g, hcn, a.rcnt, i0, j0, li, lj, s, code);
      }
    }
    // reduce the cell's 65 weights over the group, spread the LDS atomics
#pragma unroll
    for (int k = 0; k < 65; k++) {
      float v = k < 20 ? ax[k] : (k < 40 ? ay[k - 20] : az[k - 40]);
      static_assert(G == 32, "the group is two DPP rows");
      // each row's sum, then row_bcast:15 adds lane 15 of the group's lower
      // row to the lanes of its upper row, which then hold the group's total
      v = row16_sum(v);
      v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), DPP_ROW_BCAST15, 0xa, 0xf, false));
      if ((k % 16) + 16 == lg) {
        int comp, ip, jp;
        if (k < 20) {
          comp = 0; jp = k / 4 - 2; ip = k % 4 - 1;
        } else if (k < 40) {
          comp = 1; jp = (k - 20) / 5 - 1; ip = (k - 20) % 5 - 2;
        } else {
          comp = 2; jp = (k - 40) / 5 - 2; ip = (k - 40) % 5 - 2;
        }
        atomicAdd(&acc[comp][(lj + 2 + jp) * ACW + (li + 2 + ip)], v);
      }
    }
  }

  __syncthreads();
  if (counting) flush_counts<TX, TY>(g, hcn, a.rcnt, a.cumcnt, i0, j0);
  const size_t pl = (size_t)g.fw * g.fh;
  for (int k = tid; k < 3 * ACW * ACH; k += NT) {
    const int comp = k / (ACW * ACH), rem = k - comp * (ACW * ACH);
    const float v = acc[comp][rem];
    if (v == 0.0f) continue;
    const int I = i0 - 2 + rem % ACW, J = j0 - 2 + rem / ACW;
    if (I > g.nxge + 2 || J > g.nye + 2) continue;
    unsafeAtomicAdd(a.uj + pl * comp + fidx(g, I, J), (double)v);
  }
}

// The moments of the output sequence in the mixed form: mom_calc__accl's
// half-step kick (2d/common/mom_calc.f90:84-155, delt/2, no move) and
// mom_calc__nvt's CIC deposit of the seven moments (:167-249) fused, FP32 per
// particle from the int32 cell + FP32 offset record, accumulated per species
// in an FP64 LDS node tile (one-node halo; the lanes of a group share their
// cell's nodes, and FP32 LDS atomics that meet on an address serialize far
// worse than FP64 ones) and added to the FP64 mom array with one FP64 atomic
// per node and moment.  The lower CIC node of a particle
// at cell i + offset o is i when o >= 1/2 (weight o - 1/2 on i + 1), else
// i - 1 (weight o + 1/2 on i): int(x/dx - 1/2) and x/dx - 1/2 - ih of the
// reference, with delx = 1.
__global__ __launch_bounds__(mx::NT) void k_mom_mixed(MixArgs a, double* mom) {
#pragma clang fp contract(fast)
  using namespace mx;
  constexpr int GM = 8, NGRP = NT / GM;
  constexpr int MTW = TX + 2, MTH = TY + 2;
  __shared__ float sf[6][SFW * SFH];
  __shared__ double mt[7][MTW * MTH];
  const Grid2D& g = a.g;
  const int ntiles = a.ntx * a.nty;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  if (tile >= ntiles) return;
  const int i0 = g.nxgs + (tile % a.ntx) * TX;
  const int j0 = g.nys + (tile / a.ntx) * TY;
  const int ncx = min(TX, g.nxge + 1 - i0), ncy = min(TY, g.nye + 1 - j0);
  const int tid = threadIdx.x;
  for (int k = tid; k < 6 * SFW * SFH; k += NT) {
    const int c = k / (SFW * SFH), rem = k - c * (SFW * SFH);
    const int i = i0 - 1 + rem % SFW, j = j0 - 1 + rem / SFW;
    if (i <= g.nxge + 1 && j <= g.nye + 1) sf[c][rem] = (float)cell_field_m(g, a.uf, c, i, j);
  }
  const int grp = tid / GM, lg = tid % GM;
  const float ic2 = (float)(1.0 / (g.c * g.c));
  const float cc2 = (float)(g.c * g.c);
  const int rot = tid % 7;  // lanes meeting on a node add to different planes first

  for (int s = 0; s < g.nsp; s++) {
    for (int k = tid; k < 7 * MTW * MTH; k += NT) (&mt[0][0])[k] = 0.0;
    __syncthreads();
    const double hdelt = g.delt * 0.5;
    const float fac1 = (float)(g.q[s] / g.r[s] * 0.5 * hdelt);
    const float txxx = fac1 * fac1;
    const float fac2 = (float)(g.q[s] * hdelt / g.r[s]);
    for (int cc = grp; cc < TX * TY; cc += NGRP) {
      const int li = cc % TX, lj = cc / TX;
      if (li >= ncx || lj >= ncy) continue;  // uniform within a group
      const int i = i0 + li, j = j0 + lj;
      const int rw = s * g.ny + (j - g.nys);
      const int* cr = a.cumcnt + (size_t)rw * (g.nx + 1);
      const int base = a.rowbase[rw];
      const int pe = base + cr[i - g.nxgs + 1];
      for (int pi = base + cr[i - g.nxgs] + lg; pi < pe; pi += GM) {
        const int ia = a.ap ? a.ap[pi] : pi;
        const float ox = a.A.ox[ia], oy = a.A.oy[ia];
        float sx0, sx1, sx2, sy0, sy1, sy2;
        shape3f(ox - 0.5f, sx0, sx1, sx2);
        shape3f(oy - 0.5f, sy0, sy1, sy2);
        const int so = (lj + 1) * SFW + (li + 1);
        float fl[6];
#pragma unroll
        for (int cm = 0; cm < 6; cm++) {
          const float* f = &sf[cm][so];
          fl[cm] = (f[-SFW - 1] * sx0 + f[-SFW] * sx1 + f[-SFW + 1] * sx2) * sy0 +
                   (f[-1] * sx0 + f[0] * sx1 + f[1] * sx2) * sy1 +
                   (f[SFW - 1] * sx0 + f[SFW] * sx1 + f[SFW + 1] * sx2) * sy2;
        }
        const float bpx = fl[0], bpy = fl[1], bpz = fl[2], epx = fl[3], epy = fl[4], epz = fl[5];
        float uvm1 = a.A.ux[ia] + fac1 * epx;
        float uvm2 = a.A.uy[ia] + fac1 * epy;
        float uvm3 = a.A.uz[ia] + fac1 * epz;
        const float gam = sqrtf(cc2 + uvm1 * uvm1 + uvm2 * uvm2 + uvm3 * uvm3);
        const float igam = 1.0f / gam;
        const float fac1r = fac1 * igam;
        const float fac2r = fac2 / (gam + txxx * (bpx * bpx + bpy * bpy + bpz * bpz) * igam);
        const float uvm4 = uvm1 + fac1r * (uvm2 * bpz - uvm3 * bpy);
        const float uvm5 = uvm2 + fac1r * (uvm3 * bpx - uvm1 * bpz);
        const float uvm6 = uvm3 + fac1r * (uvm1 * bpy - uvm2 * bpx);
        uvm1 = uvm1 + fac2r * (uvm5 * bpz - uvm6 * bpy);
        uvm2 = uvm2 + fac2r * (uvm6 * bpx - uvm4 * bpz);
        uvm3 = uvm3 + fac2r * (uvm4 * bpy - uvm5 * bpx);
        const float ux = uvm1 + fac1 * epx, uy = uvm2 + fac1 * epy, uz = uvm3 + fac1 * epz;
        const float gi = 1.0f / sqrtf(1.0f + (ux * ux + uy * uy + uz * uz) * ic2);
        const bool hx = ox >= 0.5f, hy = oy >= 0.5f;
        const float dx = hx ? ox - 0.5f : ox + 0.5f, dxm = 1.0f - dx;
        const float dy = hy ? oy - 0.5f : oy + 0.5f, dym = 1.0f - dy;
        const int lx = li + (hx ? 1 : 0), ly = lj + (hy ? 1 : 0);  // lower node in the tile (halo at 0)
        float mv[7] = {1.0f, ux * gi, uy * gi, uz * gi, ux * ux * gi, uy * uy * gi, uz * uz * gi};
        const float w00 = dxm * dym, w10 = dx * dym, w01 = dxm * dy, w11 = dx * dy;
        double* node = &mt[0][ly * MTW + lx];
#pragma unroll
        for (int q = 0; q < 7; q++) {
          const int plane = q + rot < 7 ? q + rot : q + rot - 7;
          const float v = mv[plane];
          double* p = node + plane * (MTW * MTH);
          atomicAdd(p, (double)(v * w00));
          atomicAdd(p + 1, (double)(v * w10));
          atomicAdd(p + MTW, (double)(v * w01));
          atomicAdd(p + MTW + 1, (double)(v * w11));
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
      unsafeAtomicAdd(mom + midx(g, I, J, s) + comp, v);
    }
    __syncthreads();
  }
}

hipError_t launch_mom_mixed(wpic_ctx* c) {
  MixArgs a;
  a.g = c->g;
  a.A = c->MA;
  a.B = c->MB;
  a.rowbase = c->rowbaseA;
  a.cumcnt = c->cumcnt;
  a.ap = c->aperm;
  a.uf = c->uf;
  a.uj = c->uj;
  a.code = c->code;
  a.err = c->err;
  a.ntx = (c->g.nx + mx::TX - 1) / mx::TX;
  a.nty = (c->g.ny + mx::TY - 1) / mx::TY;
  hipError_t e = launch_zero(c, c->mom, (size_t)7 * (c->g.nx + 2) * (c->g.ny + 2) * c->g.nsp);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_mom_mixed, dim3(xcd_grid(a.ntx * a.nty)), dim3(mx::NT), 0, c->st, a, c->mom);
  return hipGetLastError();
}

hipError_t launch_push_deposit_mixed(wpic_ctx* c) {
  MixArgs a;
  a.g = c->g;
  a.A = c->MA;
  a.B = c->MB;
  a.rowbase = c->rowbaseA;
  a.cumcnt = c->cumcnt;
  a.ap = c->aperm;
  a.uf = c->uf;
  a.uj = c->uj;
  a.code = c->code;
  a.err = c->err;
  a.ntx = (c->g.nx + mx::TX - 1) / mx::TX;
  a.nty = (c->g.ny + mx::TY - 1) / mx::TY;
  a.rcnt = c->rcnt_fill ? c->rcnt : nullptr;
  const int nblk = xcd_grid(a.ntx * a.nty);
  if (c->g.bc == WPIC_BC_WALLX)
    hipLaunchKernelGGL(k_push_deposit_mixed<WPIC_BC_WALLX>, dim3(nblk), dim3(mx::NT), 0, c->st, a);
  else
    hipLaunchKernelGGL(k_push_deposit_mixed<WPIC_BC_PERIODIC>, dim3(nblk), dim3(mx::NT), 0, c->st, a);
  return hipGetLastError();
}

}  // namespace wpic
