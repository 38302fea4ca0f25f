// k_field3d.hip -- the 3D implicit FDTD update of 3d/common/field.f90:70-208:
// current ghost folds (boundary_periodic__curre, 3d/common/boundary_periodic.f90:
// 676-978), the right-hand side gkl (:128-158), the three CG solves of cgm
// (:409-560) in lock step with per-component stopping, the E update (:164-187),
// the field ghost fills (boundary_periodic__dfield, :458-673) and uf += df.
//
// CG vectors hold the interior only; the 7-point stencil maps an out-of-range
// face neighbour to its periodic image, the value boundary_periodic__phi
// (:981-1099) copies into that ghost.  Dot products: fast mode sums fixed
// per-block partials in a fixed tree (deterministic), parity mode sums
// serially in the reference's (k, j, i) order.
#include "wpic_internal3d.h"

namespace wpic {

namespace {

constexpr double kPi = 3.141592653589793;  // 4d0*atan(1d0)

__device__ inline double bsum256(double v, double* sh) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) t = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return t;
}

__device__ inline size_t tid1() { return blockIdx.x * (size_t)blockDim.x + threadIdx.x; }

}  // namespace

// ---------------------------------------------------------------------------
// current ghosts, in the order of boundary_periodic__curre
// ---------------------------------------------------------------------------
// y: fold the two ghost rows of every (i, k), ghosts included
__global__ void k_curre3_yfold(Grid3D g, double* uj) {
  const size_t t = tid1();
  const size_t per = (size_t)g.fw * g.fd;
  if (t >= 3 * per) return;
  const int comp = (int)(t / per);
  const int I = g.nxgs - 2 + (int)(t % g.fw), K = g.nzs - 2 + (int)((t % per) / g.fw);
  double* u = uj + plane3(g) * comp;
  const int nys = g.nys, nye = g.nye;
  const double a = u[fidx3(g, I, nys - 2, K)], b = u[fidx3(g, I, nys - 1, K)];
  u[fidx3(g, I, nye - 1, K)] = u[fidx3(g, I, nye - 1, K)] + a;
  u[fidx3(g, I, nye, K)] = u[fidx3(g, I, nye, K)] + b;
  const double c = u[fidx3(g, I, nye + 1, K)], d = u[fidx3(g, I, nye + 2, K)];
  u[fidx3(g, I, nys, K)] = u[fidx3(g, I, nys, K)] + c;
  u[fidx3(g, I, nys + 1, K)] = u[fidx3(g, I, nys + 1, K)] + d;
}

// z: fold the two ghost planes for interior j, all i
__global__ void k_curre3_zfold(Grid3D g, double* uj) {
  const size_t t = tid1();
  const size_t per = (size_t)g.fw * g.ny;
  if (t >= 3 * per) return;
  const int comp = (int)(t / per);
  const int I = g.nxgs - 2 + (int)(t % g.fw), J = g.nys + (int)((t % per) / g.fw);
  double* u = uj + plane3(g) * comp;
  const int nzs = g.nzs, nze = g.nze;
  const double a = u[fidx3(g, I, J, nzs - 2)], b = u[fidx3(g, I, J, nzs - 1)];
  u[fidx3(g, I, J, nze - 1)] = u[fidx3(g, I, J, nze - 1)] + a;
  u[fidx3(g, I, J, nze)] = u[fidx3(g, I, J, nze)] + b;
  const double c = u[fidx3(g, I, J, nze + 1)], d = u[fidx3(g, I, J, nze + 2)];
  u[fidx3(g, I, J, nzs)] = u[fidx3(g, I, J, nzs)] + c;
  u[fidx3(g, I, J, nzs + 1)] = u[fidx3(g, I, J, nzs + 1)] + d;
}

// one ghost row copied back in y, for interior k
__global__ void k_curre3_ycopy(Grid3D g, double* uj) {
  const size_t t = tid1();
  const size_t per = (size_t)g.fw * g.nz;
  if (t >= 3 * per) return;
  const int comp = (int)(t / per);
  const int I = g.nxgs - 2 + (int)(t % g.fw), K = g.nzs + (int)((t % per) / g.fw);
  double* u = uj + plane3(g) * comp;
  u[fidx3(g, I, g.nye + 1, K)] = u[fidx3(g, I, g.nys, K)];
  u[fidx3(g, I, g.nys - 1, K)] = u[fidx3(g, I, g.nye, K)];
}

// one ghost plane copied back in z, for j in nys-1..nye+1
__global__ void k_curre3_zcopy(Grid3D g, double* uj) {
  const size_t t = tid1();
  const size_t per = (size_t)g.fw * (g.ny + 2);
  if (t >= 3 * per) return;
  const int comp = (int)(t / per);
  const int I = g.nxgs - 2 + (int)(t % g.fw), J = g.nys - 1 + (int)((t % per) / g.fw);
  double* u = uj + plane3(g) * comp;
  u[fidx3(g, I, J, g.nze + 1)] = u[fidx3(g, I, J, g.nzs)];
  u[fidx3(g, I, J, g.nzs - 1)] = u[fidx3(g, I, J, g.nze)];
}

// x: fold then copy two layers, every (j, k)
__global__ void k_curre3_x(Grid3D g, double* uj) {
  const size_t t = tid1();
  const size_t per = (size_t)g.fh * g.fd;
  if (t >= 3 * per) return;
  const int comp = (int)(t / per);
  const int J = g.nys - 2 + (int)(t % g.fh), K = g.nzs - 2 + (int)((t % per) / g.fh);
  double* u = uj + plane3(g) * comp;
  const int nxs = g.nxgs, nxe = g.nxge;
#define U(i) u[fidx3(g, i, J, K)]
  U(nxe - 1) = U(nxe - 1) + U(nxs - 2);
  U(nxe) = U(nxe) + U(nxs - 1);
  U(nxs) = U(nxs) + U(nxe + 1);
  U(nxs + 1) = U(nxs + 1) + U(nxe + 2);
  U(nxs - 2) = U(nxe - 1);
  U(nxs - 1) = U(nxe);
  U(nxe + 1) = U(nxs);
  U(nxe + 2) = U(nxs + 1);
#undef U
}

// ---------------------------------------------------------------------------
// field ghosts (boundary_periodic__dfield): y rows, z planes, x columns
// ---------------------------------------------------------------------------
__global__ void k_dfield3_y(Grid3D g, double* df) {
  const size_t t = tid1();
  const size_t per = (size_t)g.nx * g.nz;
  if (t >= 6 * per) return;
  const int comp = (int)(t / per);
  const int I = g.nxgs + (int)(t % g.nx), K = g.nzs + (int)((t % per) / g.nx);
  double* d = df + plane3(g) * comp;
  const double a = d[fidx3(g, I, g.nys, K)], b = d[fidx3(g, I, g.nys + 1, K)];
  d[fidx3(g, I, g.nye + 1, K)] = a;
  d[fidx3(g, I, g.nye + 2, K)] = b;
  const double c = d[fidx3(g, I, g.nye - 1, K)], e = d[fidx3(g, I, g.nye, K)];
  d[fidx3(g, I, g.nys - 2, K)] = c;
  d[fidx3(g, I, g.nys - 1, K)] = e;
}

__global__ void k_dfield3_z(Grid3D g, double* df) {
  const size_t t = tid1();
  const size_t per = (size_t)g.nx * g.fh;
  if (t >= 6 * per) return;
  const int comp = (int)(t / per);
  const int I = g.nxgs + (int)(t % g.nx), J = g.nys - 2 + (int)((t % per) / g.nx);
  double* d = df + plane3(g) * comp;
  const double a = d[fidx3(g, I, J, g.nzs)], b = d[fidx3(g, I, J, g.nzs + 1)];
  d[fidx3(g, I, J, g.nze + 1)] = a;
  d[fidx3(g, I, J, g.nze + 2)] = b;
  const double c = d[fidx3(g, I, J, g.nze - 1)], e = d[fidx3(g, I, J, g.nze)];
  d[fidx3(g, I, J, g.nzs - 2)] = c;
  d[fidx3(g, I, J, g.nzs - 1)] = e;
}

__global__ void k_dfield3_x(Grid3D g, double* df) {
  const size_t t = tid1();
  const size_t per = (size_t)g.fh * g.fd;
  if (t >= 6 * per) return;
  const int comp = (int)(t / per);
  const int J = g.nys - 2 + (int)(t % g.fh), K = g.nzs - 2 + (int)((t % per) / g.fh);
  double* d = df + plane3(g) * comp;
  const int nxs = g.nxgs, nxe = g.nxge;
  d[fidx3(g, nxs - 2, J, K)] = d[fidx3(g, nxe - 1, J, K)];
  d[fidx3(g, nxs - 1, J, K)] = d[fidx3(g, nxe, J, K)];
  d[fidx3(g, nxe + 1, J, K)] = d[fidx3(g, nxs, J, K)];
  d[fidx3(g, nxe + 2, J, K)] = d[fidx3(g, nxs + 1, J, K)];
}

// ---------------------------------------------------------------------------
// CG
// ---------------------------------------------------------------------------
struct Cg3Args {
  Grid3D g;
  const double *uf, *uj;
  double* df;
  double *phi, *pold, *pnew, *r, *b, *ap;  // [3][nx*ny*nz]
  double* part;                            // [3][NPB3][2]
  CgScal* sc;
  int* err;
  int parity;
};

// in 32-bit arithmetic: a context holds fewer than 2^31 cells
__device__ inline void decode3(const Grid3D& g, size_t q, int& i, int& j, int& k) {
  const unsigned qq = (unsigned)q, nx = (unsigned)g.nx, ny = (unsigned)g.ny;
  const unsigned row = qq / nx;
  i = g.nxgs + (int)(qq - row * nx);
  j = g.nys + (int)(row % ny);
  k = g.nzs + (int)(row / ny);
}

// interior index of the periodic image of (i, j, k), one coordinate at most
// one cell outside
__device__ inline size_t nbr3(const Grid3D& g, int i, int j, int k) {
  if (i < g.nxgs) i = g.nxge;
  else if (i > g.nxge) i = g.nxgs;
  if (j < g.nys) j = g.nye;
  else if (j > g.nye) j = g.nys;
  if (k < g.nzs) k = g.nze;
  else if (k > g.nze) k = g.nzs;
  return cidx3(g, i, j, k);
}

__device__ inline double serial_dot3(const double* x, const double* y, size_t n) {
  double s = 0.0;
  for (size_t k = 0; k < n; k++) s = s + x[k] * y[k];
  return s;
}

// the NPB3 block partials: each thread adds NPB3/256 consecutive ones in
// order, then a fixed tree over the block
__device__ inline double part_sum3(const Cg3Args& a, int l, int slot, double* sh) {
  constexpr int PER = NPB3 / 256;
  double v = 0.0;
  for (int q = 0; q < PER; q++) v = v + a.part[((size_t)l * NPB3 + threadIdx.x * PER + q) * 2 + slot];
  return bsum256(v, sh);
}

// gkl (field.f90:128-158), b = f5*gkl, warm start phi = df(l), sum b^2
__global__ __launch_bounds__(256) void k_cg3_init(Cg3Args a) {
  __shared__ double sh[4];
  const Grid3D& g = a.g;
  const size_t n = (size_t)g.nx * g.ny * g.nz, pl = plane3(g);
  const double *U1 = a.uf, *U2 = a.uf + pl, *U3 = a.uf + 2 * pl, *U4 = a.uf + 3 * pl, *U5 = a.uf + 4 * pl,
               *U6 = a.uf + 5 * pl;
  const double *J1 = a.uj, *J2 = a.uj + pl, *J3 = a.uj + 2 * pl;
  const double f1 = g.f1, f2 = g.f2, f3 = g.f3, f5 = g.f5;
  double s[3] = {0.0, 0.0, 0.0};
  for (size_t q = blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)NPB3 * 256) {
    int i, j, k;
    decode3(g, q, i, j, k);
#define F(P, ii, jj, kk) P[fidx3(g, ii, jj, kk)]
    const double gk1 = +f2 * (+F(U1, i, j, k - 1) + F(U1, i, j - 1, k) + F(U1, i - 1, j, k) - 6e0 * F(U1, i, j, k) +
                              F(U1, i + 1, j, k) + F(U1, i, j + 1, k) + F(U1, i, j, k + 1) +
                              f3 * (-F(J3, i, j - 1, k) + F(J3, i, j, k) + F(J2, i, j, k - 1) - F(J2, i, j, k))) -
                       f1 * (-F(U6, i, j - 1, k) + F(U6, i, j, k) + F(U5, i, j, k - 1) - F(U5, i, j, k));
    const double gk2 = +f2 * (+F(U2, i, j, k - 1) + F(U2, i, j - 1, k) + F(U2, i - 1, j, k) - 6e0 * F(U2, i, j, k) +
                              F(U2, i + 1, j, k) + F(U2, i, j + 1, k) + F(U2, i, j, k + 1) +
                              f3 * (-F(J1, i, j, k - 1) + F(J1, i, j, k) + F(J3, i - 1, j, k) - F(J3, i, j, k))) -
                       f1 * (-F(U4, i, j, k - 1) + F(U4, i, j, k) + F(U6, i - 1, j, k) - F(U6, i, j, k));
    const double gk3 = +f2 * (+F(U3, i, j, k - 1) + F(U3, i, j - 1, k) + F(U3, i - 1, j, k) - 6e0 * F(U3, i, j, k) +
                              F(U3, i + 1, j, k) + F(U3, i, j + 1, k) + F(U3, i, j, k + 1) +
                              f3 * (-F(J2, i - 1, j, k) + F(J2, i, j, k) + F(J1, i, j - 1, k) - F(J1, i, j, k))) -
                       f1 * (-F(U5, i - 1, j, k) + F(U5, i, j, k) + F(U4, i, j - 1, k) - F(U4, i, j, k));
#undef F
    const double gk[3] = {gk1, gk2, gk3};
    const size_t f = fidx3(g, i, j, k);
#pragma unroll
    for (int l = 0; l < 3; l++) {
      const double bv = f5 * gk[l];
      a.b[(size_t)l * n + q] = bv;
      a.phi[(size_t)l * n + q] = a.df[pl * l + f];
      s[l] = s[l] + bv * bv;
    }
  }
  if (!a.parity) {
    for (int l = 0; l < 3; l++) {
      const double t = bsum256(s[l], sh);
      if (threadIdx.x == 0) a.part[((size_t)l * NPB3 + blockIdx.x) * 2] = t;
    }
  }
}

// sum_g = |b|^2, eps = sqrt(sum_g)*err (field.f90:454-456)
__global__ __launch_bounds__(256) void k_cg3_fin0(Cg3Args a) {
  __shared__ double sh[4];
  const int l = blockIdx.x;
  const size_t n = (size_t)a.g.nx * a.g.ny * a.g.nz;
  double t;
  if (a.parity)
    t = threadIdx.x == 0 ? serial_dot3(a.b + (size_t)l * n, a.b + (size_t)l * n, n) : 0.0;
  else
    t = part_sum3(a, l, 0, sh);
  if (threadIdx.x == 0) {
    a.sc[l].sum_g = t;
    a.sc[l].eps = sqrt(t) * a.g.cg_err;
  }
}

// r = b + phi(k-1) + phi(j-1) + phi(i-1) - f4 phi + phi(i+1) + phi(j+1) + phi(k+1); p = r
__global__ __launch_bounds__(256) void k_cg3_r0(Cg3Args a) {
  __shared__ double sh[4];
  const Grid3D& g = a.g;
  const size_t n = (size_t)g.nx * g.ny * g.nz;
  const double f4 = g.f4;
  double s[3] = {0.0, 0.0, 0.0};
  for (size_t q = blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)NPB3 * 256) {
    int i, j, k;
    decode3(g, q, i, j, k);
#pragma unroll
    for (int l = 0; l < 3; l++) {
      const double* ph = a.phi + (size_t)l * n;
      const double rv = a.b[(size_t)l * n + q] + ph[nbr3(g, i, j, k - 1)] + ph[nbr3(g, i, j - 1, k)] +
                        ph[nbr3(g, i - 1, j, k)] - f4 * ph[q] + ph[nbr3(g, i + 1, j, k)] + ph[nbr3(g, i, j + 1, k)] +
                        ph[nbr3(g, i, j, k + 1)];
      a.r[(size_t)l * n + q] = rv;
      a.pold[(size_t)l * n + q] = rv;
      s[l] = s[l] + rv * rv;
    }
  }
  if (!a.parity) {
    for (int l = 0; l < 3; l++) {
      const double t = bsum256(s[l], sh);
      if (threadIdx.x == 0) a.part[((size_t)l * NPB3 + blockIdx.x) * 2] = t;
    }
  }
}

__global__ __launch_bounds__(256) void k_cg3_fin1(Cg3Args a) {
  __shared__ double sh[4];
  const int l = blockIdx.x;
  const size_t n = (size_t)a.g.nx * a.g.ny * a.g.nz;
  double t;
  if (a.parity)
    t = threadIdx.x == 0 ? serial_dot3(a.r + (size_t)l * n, a.r + (size_t)l * n, n) : 0.0;
  else
    t = part_sum3(a, l, 0, sh);
  if (threadIdx.x == 0) {
    CgScal& c = a.sc[l];
    c.sumr_g = t;
    c.ite = 0;
    c.bv = 0.0;
    c.active = (sqrt(t) > c.eps) && (c.sum_g > c.eps);
  }
}

// p = r + bv p (the previous iteration's update, fused); ap = -p(k-1) - p(j-1)
// - p(i-1) + f4 p - p(i+1) - p(j+1) - p(k+1); sum r^2 and p.ap (:483-505)
// p = r + bv p (the previous iteration's update, fused here), then ap = f4 p -
// sum of the six neighbours; sum r^2 and sum p.ap (3d/common/field.f90's cgm).
// A block marches a tile of 32 x 8 columns through a slab of z levels: p at
// levels k-1, k, k+1 of a column stay in registers, level k of the tile (with
// its one-column halo) goes through LDS, so r and the old p are read from HBM
// once per node.  The work units (tile, slab) are dealt to the NPB3 blocks;
// a block with none writes zero partials.  Every value is the per-node
// expression of the reference's order, so pn and ap do not depend on the
// tiling (the fast mode's partial sums do, in a fixed order).
__global__ __launch_bounds__(256) void k_cg3_a(Cg3Args a) {
  constexpr int TXC = 32, TYC = 8, SW3 = TXC + 2;
  __shared__ double sh[4];
  __shared__ double sm[(TYC + 2) * SW3];
  const int l = blockIdx.y;
  if (!a.sc[l].active) return;
  const Grid3D& g = a.g;
  const size_t n = (size_t)g.nx * g.ny * g.nz;
  const double f4 = g.f4, bv = a.sc[l].bv;
  const double *r = a.r + (size_t)l * n, *po = a.pold + (size_t)l * n;
  double* pn = a.pnew + (size_t)l * n;
  double* ap = a.ap + (size_t)l * n;
  const int ntx = (g.nx + TXC - 1) / TXC, nty = (g.ny + TYC - 1) / TYC, ntile = ntx * nty;
  int nsl = NPB3 / ntile;
  if (nsl < 1) nsl = 1;
  if (nsl > g.nz) nsl = g.nz;
  const int len = (g.nz + nsl - 1) / nsl;
  const int units = ntile * nsl;
  const int tx = threadIdx.x % TXC, ty = threadIdx.x / TXC;
  double s_r = 0.0, s_2 = 0.0;
  for (int u = blockIdx.x; u < units; u += NPB3) {
    const int sl = u / ntile, t = u - sl * ntile;
    const int ks = g.nzs + sl * len;
    const int ke = ks + len - 1 < g.nze ? ks + len - 1 : g.nze;
    if (ks > g.nze) continue;  // uniform in the block
    const int i = g.nxgs + (t % ntx) * TXC + tx, j = g.nys + (t / ntx) * TYC + ty;
    const bool act = i <= g.nxge && j <= g.nye;
    size_t m;
#define P(ii, jj, kk) (m = nbr3(g, ii, jj, kk), r[m] + bv * po[m])
    double pB = act ? P(i, j, ks - 1) : 0.0, pC = act ? P(i, j, ks) : 0.0;
    for (int k = ks; k <= ke; k++) {
      const double pT = act ? P(i, j, k + 1) : 0.0;
      if (act) {
        sm[(ty + 1) * SW3 + tx + 1] = pC;
        if (tx == 0) sm[(ty + 1) * SW3] = P(i - 1, j, k);
        if (tx == TXC - 1 || i == g.nxge) sm[(ty + 1) * SW3 + tx + 2] = P(i + 1, j, k);
        if (ty == 0) sm[tx + 1] = P(i, j - 1, k);
        if (ty == TYC - 1 || j == g.nye) sm[(ty + 2) * SW3 + tx + 1] = P(i, j + 1, k);
      }
      __syncthreads();
      if (act) {
        const double pS = sm[ty * SW3 + tx + 1], pW = sm[(ty + 1) * SW3 + tx];
        const double pE = sm[(ty + 1) * SW3 + tx + 2], pN = sm[(ty + 2) * SW3 + tx + 1];
        const double av = -pB - pS - pW + f4 * pC - pE - pN - pT;
        const size_t q = cidx3(g, i, j, k);
        pn[q] = pC;
        ap[q] = av;
        s_r = s_r + r[q] * r[q];
        s_2 = s_2 + pC * av;
      }
      __syncthreads();
      pB = pC;
      pC = pT;
    }
#undef P
  }
  if (!a.parity) {
    const double t1 = bsum256(s_r, sh);
    const double t2 = bsum256(s_2, sh);
    if (threadIdx.x == 0) {
      a.part[((size_t)l * NPB3 + blockIdx.x) * 2] = t1;
      a.part[((size_t)l * NPB3 + blockIdx.x) * 2 + 1] = t2;
    }
  }
}

__global__ __launch_bounds__(256) void k_cg3_fa(Cg3Args a) {
  __shared__ double sh[4];
  const int l = blockIdx.x;
  if (!a.sc[l].active) return;
  const size_t n = (size_t)a.g.nx * a.g.ny * a.g.nz;
  double t1, t2;
  if (a.parity) {
    t1 = threadIdx.x == 0 ? serial_dot3(a.r + (size_t)l * n, a.r + (size_t)l * n, n) : 0.0;
    t2 = threadIdx.x == 0 ? serial_dot3(a.pnew + (size_t)l * n, a.ap + (size_t)l * n, n) : 0.0;
  } else {
    t1 = part_sum3(a, l, 0, sh);
    t2 = part_sum3(a, l, 1, sh);
  }
  if (threadIdx.x == 0) {
    CgScal& c = a.sc[l];
    c.sumr_g = t1;
    c.sum2_g = t2;
    c.av = t1 / t2;
    c.ite = c.ite + 1;
    c.sum_g = sqrt(t1);
  }
}

__global__ __launch_bounds__(256) void k_cg3_b(Cg3Args a) {
  __shared__ double sh[4];
  const int l = blockIdx.y;
  if (!a.sc[l].active) return;
  const Grid3D& g = a.g;
  const size_t n = (size_t)g.nx * g.ny * g.nz;
  const double av = a.sc[l].av;
  double *ph = a.phi + (size_t)l * n, *r = a.r + (size_t)l * n;
  const double *p = a.pnew + (size_t)l * n, *ap = a.ap + (size_t)l * n;
  double s = 0.0;
  for (size_t q = blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)NPB3 * 256) {
    ph[q] = ph[q] + av * p[q];
    const double rv = r[q] - av * ap[q];
    r[q] = rv;
    s = s + rv * rv;
  }
  if (!a.parity) {
    const double t = bsum256(s, sh);
    if (threadIdx.x == 0) a.part[((size_t)l * NPB3 + blockIdx.x) * 2] = t;
  }
}

__global__ __launch_bounds__(256) void k_cg3_fb(Cg3Args a) {
  __shared__ double sh[4];
  const int l = blockIdx.x;
  if (!a.sc[l].active) return;
  const size_t n = (size_t)a.g.nx * a.g.ny * a.g.nz;
  double t;
  if (a.parity)
    t = threadIdx.x == 0 ? serial_dot3(a.r + (size_t)l * n, a.r + (size_t)l * n, n) : 0.0;
  else
    t = part_sum3(a, l, 0, sh);
  if (threadIdx.x == 0) {
    CgScal& c = a.sc[l];
    if (c.ite >= 100) {
      atomicOr(a.err, ERR_CG);
      c.active = 0;
      return;
    }
    c.bv = t / c.sumr_g;
    c.active = c.sum_g > c.eps;
  }
}

__global__ void k_cg3_store(Cg3Args a) {
  const Grid3D& g = a.g;
  const size_t n = (size_t)g.nx * g.ny * g.nz;
  const size_t q = tid1();
  if (q >= n) return;
  int i, j, k;
  decode3(g, q, i, j, k);
  const size_t pl = plane3(g), f = fidx3(g, i, j, k);
  for (int l = 0; l < 3; l++) a.df[pl * l + f] = a.phi[(size_t)l * n + q];
}

// E update (field.f90:164-187)
__global__ void k_e3_update(Grid3D g, const double* __restrict__ uf, const double* __restrict__ uj,
                            double* __restrict__ df) {
  const size_t n = (size_t)g.nx * g.ny * g.nz;
  const size_t q = tid1();
  if (q >= n) return;
  int i, j, k;
  decode3(g, q, i, j, k);
  const size_t pl = plane3(g);
  const double f1 = g.f1, gfac = g.gfac, pdt = g.pi4dt;
#define D(c, ii, jj, kk) df[pl * ((c)-1) + fidx3(g, ii, jj, kk)]
#define U(c, ii, jj, kk) uf[pl * ((c)-1) + fidx3(g, ii, jj, kk)]
#define J(c) uj[pl * ((c)-1) + fidx3(g, i, j, k)]
  const double e4 = +f1 * (+gfac * (-D(3, i, j, k) + D(3, i, j + 1, k) + D(2, i, j, k) - D(2, i, j, k + 1)) +
                           (-U(3, i, j, k) + U(3, i, j + 1, k) + U(2, i, j, k) - U(2, i, j, k + 1))) -
                    pdt * J(1);
  const double e5 = +f1 * (+gfac * (-D(1, i, j, k) + D(1, i, j, k + 1) + D(3, i, j, k) - D(3, i + 1, j, k)) +
                           (-U(1, i, j, k) + U(1, i, j, k + 1) + U(3, i, j, k) - U(3, i + 1, j, k))) -
                    pdt * J(2);
  const double e6 = +f1 * (+gfac * (-D(2, i, j, k) + D(2, i + 1, j, k) + D(1, i, j, k) - D(1, i, j + 1, k)) +
                           (-U(2, i, j, k) + U(2, i + 1, j, k) + U(1, i, j, k) - U(1, i, j + 1, k))) -
                    pdt * J(3);
  D(4, i, j, k) = e4;
  D(5, i, j, k) = e5;
  D(6, i, j, k) = e6;
#undef D
#undef U
#undef J
}

__global__ void k_add3(size_t n, double* __restrict__ uf, const double* __restrict__ df) {
  for (size_t q = tid1(); q < n; q += (size_t)gridDim.x * blockDim.x) uf[q] = uf[q] + df[q];
}

__global__ void k_fill3(double* p, size_t n, double v) {
  for (size_t q = tid1(); q < n; q += (size_t)gridDim.x * blockDim.x) p[q] = v;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static Cg3Args cg3_args(wpic_ctx3d* c) {
  Cg3Args a;
  a.g = c->g;
  a.uf = c->uf;
  a.uj = c->uj;
  a.df = c->df;
  a.phi = c->phi;
  a.pold = c->pa;
  a.pnew = c->pb;
  a.r = c->rr;
  a.b = c->bb;
  a.ap = c->ap;
  a.part = c->part;
  a.sc = c->scal;
  a.err = c->err;
  a.parity = c->mode == WPIC_MODE_PARITY;
  return a;
}

static unsigned grid_n(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  return (unsigned)(b ? b : 1);
}

static unsigned grid_all(size_t n) { return (unsigned)((n + 255) / 256 ? (n + 255) / 256 : 1); }

hipError_t launch_fill3d(wpic_ctx3d* c, double* p, size_t n, double v) {
  hipLaunchKernelGGL(k_fill3, dim3(grid_n(n)), dim3(256), 0, c->st, p, n, v);
  return hipGetLastError();
}

hipError_t launch_curre_bc3d(wpic_ctx3d* c) {
  const Grid3D& g = c->g;
  hipLaunchKernelGGL(k_curre3_yfold, dim3(grid_all(3 * (size_t)g.fw * g.fd)), dim3(256), 0, c->st, g, c->uj);
  hipLaunchKernelGGL(k_curre3_zfold, dim3(grid_all(3 * (size_t)g.fw * g.ny)), dim3(256), 0, c->st, g, c->uj);
  hipLaunchKernelGGL(k_curre3_ycopy, dim3(grid_all(3 * (size_t)g.fw * g.nz)), dim3(256), 0, c->st, g, c->uj);
  hipLaunchKernelGGL(k_curre3_zcopy, dim3(grid_all(3 * (size_t)g.fw * (g.ny + 2))), dim3(256), 0, c->st, g, c->uj);
  hipLaunchKernelGGL(k_curre3_x, dim3(grid_all(3 * (size_t)g.fh * g.fd)), dim3(256), 0, c->st, g, c->uj);
  return hipGetLastError();
}

hipError_t launch_dfield_bc3d(wpic_ctx3d* c) {
  const Grid3D& g = c->g;
  hipLaunchKernelGGL(k_dfield3_y, dim3(grid_all(6 * (size_t)g.nx * g.nz)), dim3(256), 0, c->st, g, c->df);
  hipLaunchKernelGGL(k_dfield3_z, dim3(grid_all(6 * (size_t)g.nx * g.fh)), dim3(256), 0, c->st, g, c->df);
  hipLaunchKernelGGL(k_dfield3_x, dim3(grid_all(6 * (size_t)g.fh * g.fd)), dim3(256), 0, c->st, g, c->df);
  return hipGetLastError();
}

hipError_t launch_cg_init3d(wpic_ctx3d* c) {
  Cg3Args a = cg3_args(c);
  hipLaunchKernelGGL(k_cg3_init, dim3(NPB3), dim3(256), 0, c->st, a);
  hipLaunchKernelGGL(k_cg3_fin0, dim3(3), dim3(256), 0, c->st, a);
  hipLaunchKernelGGL(k_cg3_r0, dim3(NPB3), dim3(256), 0, c->st, a);
  hipLaunchKernelGGL(k_cg3_fin1, dim3(3), dim3(256), 0, c->st, a);
  return hipGetLastError();
}

hipError_t launch_cg_iter3d(wpic_ctx3d* c, int it) {
  Cg3Args a = cg3_args(c);
  if (it & 1) {
    a.pold = c->pb;
    a.pnew = c->pa;
  }
  hipLaunchKernelGGL(k_cg3_a, dim3(NPB3, 3), dim3(256), 0, c->st, a);
  hipLaunchKernelGGL(k_cg3_fa, dim3(3), dim3(256), 0, c->st, a);
  hipLaunchKernelGGL(k_cg3_b, dim3(NPB3, 3), dim3(256), 0, c->st, a);
  hipLaunchKernelGGL(k_cg3_fb, dim3(3), dim3(256), 0, c->st, a);
  return hipGetLastError();
}

hipError_t launch_cg_store3d(wpic_ctx3d* c) {
  Cg3Args a = cg3_args(c);
  hipLaunchKernelGGL(k_cg3_store, dim3(grid_all((size_t)c->g.nx * c->g.ny * c->g.nz)), dim3(256), 0, c->st, a);
  return hipGetLastError();
}

hipError_t launch_e_update3d(wpic_ctx3d* c) {
  hipLaunchKernelGGL(k_e3_update, dim3(grid_all((size_t)c->g.nx * c->g.ny * c->g.nz)), dim3(256), 0, c->st, c->g,
                     c->uf, c->uj, c->df);
  return hipGetLastError();
}

hipError_t launch_uf_add3d(wpic_ctx3d* c) {
  const size_t n = 6 * plane3(c->g);
  hipLaunchKernelGGL(k_add3, dim3(grid_n(n)), dim3(256), 0, c->st, n, c->uf, c->df);
  return hipGetLastError();
}

}  // namespace wpic
