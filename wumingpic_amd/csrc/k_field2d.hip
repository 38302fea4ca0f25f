// k_field2d.hip -- implicit FDTD update (Hoshino 2013) of 2d/common/field.f90:
// current ghost folds, the right-hand side gkl (:125-146), the three CG solves
// of cgm (:319-461) run in lock step with per-component stopping, the E update
// (:154-171), the field ghost fills and uf += df (:176-184).
//
// All kernels are coalesced sweeps over x-fastest planes.  Ghost values of the
// CG vectors are not stored: the 5-point stencil maps an out-of-range
// neighbour to the interior value the boundary routine would have copied
// there (boundary_periodic.f90:511-568, boundary_reconnection.f90:505-579),
// with the same sign, so the stencil sees the same bits.
//
// Dot products: fast mode sums fixed per-block partials in a fixed tree
// (deterministic); parity mode sums serially in the reference's (j, i) order.
#include "wpic_internal.h"

namespace wpic {

constexpr double kPi = 3.141592653589793;  // 4.0D0*atan(1.0D0)

// ---------------------------------------------------------------------------
// deterministic block reduction (256 threads)
// ---------------------------------------------------------------------------
__device__ inline double block_sum256(double v, double* sh) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) t = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return t;  // valid in thread 0
}

// ---------------------------------------------------------------------------
// current ghost folds at one rank: boundary_periodic.f90:369-506
// ---------------------------------------------------------------------------
__global__ void k_curre_y(Grid2D g, double* uj) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * g.fw) return;
  const int comp = t / g.fw, I = g.nxgs - 2 + t % g.fw;
  double* u = uj + (size_t)g.fw * g.fh * comp;
  const int nys = g.nys, nye = g.nye;
  const double a = u[fidx(g, I, nys - 2)], b = u[fidx(g, I, nys - 1)];
  u[fidx(g, I, nye - 1)] = u[fidx(g, I, nye - 1)] + a;
  u[fidx(g, I, nye)] = u[fidx(g, I, nye)] + b;
  const double c = u[fidx(g, I, nye + 1)], d = u[fidx(g, I, nye + 2)];
  u[fidx(g, I, nys)] = u[fidx(g, I, nys)] + c;
  u[fidx(g, I, nys + 1)] = u[fidx(g, I, nys + 1)] + d;
  u[fidx(g, I, nye + 1)] = u[fidx(g, I, nys)];
  u[fidx(g, I, nye + 2)] = u[fidx(g, I, nys + 1)];
  u[fidx(g, I, nys - 2)] = u[fidx(g, I, nye - 1)];
  u[fidx(g, I, nys - 1)] = u[fidx(g, I, nye)];
}

__global__ void k_curre_x(Grid2D g, double* uj) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * g.fh) return;
  const int comp = t / g.fh, J = g.nys - 2 + t % g.fh;
  double* u = uj + (size_t)g.fw * g.fh * comp;
  const int nxs = g.nxgs, nxe = g.nxge;
  u[fidx(g, nxe - 1, J)] = u[fidx(g, nxe - 1, J)] + u[fidx(g, nxs - 2, J)];
  u[fidx(g, nxe, J)] = u[fidx(g, nxe, J)] + u[fidx(g, nxs - 1, J)];
  u[fidx(g, nxs, J)] = u[fidx(g, nxs, J)] + u[fidx(g, nxe + 1, J)];
  u[fidx(g, nxs + 1, J)] = u[fidx(g, nxs + 1, J)] + u[fidx(g, nxe + 2, J)];
  u[fidx(g, nxs - 2, J)] = u[fidx(g, nxe - 1, J)];
  u[fidx(g, nxs - 1, J)] = u[fidx(g, nxe, J)];
  u[fidx(g, nxe + 1, J)] = u[fidx(g, nxs, J)];
  u[fidx(g, nxe + 2, J)] = u[fidx(g, nxs + 1, J)];
}

// ---------------------------------------------------------------------------
// field ghost fills: boundary_periodic.f90:263-352; walls :350-359 of
// boundary_reconnection.f90
// ---------------------------------------------------------------------------
__global__ void k_dfield_y(Grid2D g, double* df) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * g.nx) return;
  const int comp = t / g.nx, I = g.nxgs + t % g.nx;
  double* d = df + (size_t)g.fw * g.fh * comp;
  const int nys = g.nys, nye = g.nye;
  const double a = d[fidx(g, I, nys)], b = d[fidx(g, I, nys + 1)];
  d[fidx(g, I, nye + 1)] = a;
  d[fidx(g, I, nye + 2)] = b;
  const double c = d[fidx(g, I, nye - 1)], e = d[fidx(g, I, nye)];
  d[fidx(g, I, nys - 2)] = c;
  d[fidx(g, I, nys - 1)] = e;
}

__global__ void k_dfield_x(Grid2D g, double* df) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.fh) return;
  const int J = g.nys - 2 + t;
  const size_t pl = (size_t)g.fw * g.fh;
  const int nxs = g.nxgs, nxe = g.nxge;
  if (g.bc == WPIC_BC_SHOCK) {
    // boundary_shock.f90:395-405: the left wall as the reconnection's, zeros
    // beyond the injection edge
    df[fidx(g, nxs - 1, J)] = -df[fidx(g, nxs, J)];
    for (int c = 1; c <= 3; c++) df[pl * c + fidx(g, nxs - 1, J)] = df[pl * c + fidx(g, nxs + 1, J)];
    for (int c = 4; c <= 5; c++) df[pl * c + fidx(g, nxs - 1, J)] = -df[pl * c + fidx(g, nxs, J)];
    for (int c = 0; c < 6; c++) df[pl * c + fidx(g, nxe + 1, J)] = 0.0;
    return;
  }
  if (g.bc == WPIC_BC_WALLX) {
    double* d1 = df;
    d1[fidx(g, nxs - 1, J)] = -d1[fidx(g, nxs, J)];
    for (int c = 1; c <= 3; c++) df[pl * c + fidx(g, nxs - 1, J)] = df[pl * c + fidx(g, nxs + 1, J)];
    for (int c = 4; c <= 5; c++) df[pl * c + fidx(g, nxs - 1, J)] = -df[pl * c + fidx(g, nxs, J)];
    d1[fidx(g, nxe, J)] = -d1[fidx(g, nxe - 1, J)];
    for (int c = 1; c <= 3; c++) df[pl * c + fidx(g, nxe + 1, J)] = df[pl * c + fidx(g, nxe - 1, J)];
    for (int c = 4; c <= 5; c++) df[pl * c + fidx(g, nxe, J)] = -df[pl * c + fidx(g, nxe - 1, J)];
    return;
  }
  for (int c = 0; c < 6; c++) {
    double* d = df + pl * c;
    d[fidx(g, nxs - 2, J)] = d[fidx(g, nxe - 1, J)];
    d[fidx(g, nxs - 1, J)] = d[fidx(g, nxe, J)];
    d[fidx(g, nxe + 1, J)] = d[fidx(g, nxs, J)];
    d[fidx(g, nxe + 2, J)] = d[fidx(g, nxs + 1, J)];
  }
}

// ---------------------------------------------------------------------------
// CG helpers
// ---------------------------------------------------------------------------
struct CgArgs {
  Grid2D g;
  const double *uf, *uj;
  double* df;
  double *phi, *pold, *pnew, *r, *b, *ap;  // each [3][nx*ny]
  double* part;                            // [3][NPB][2]
  CgScal* sc;
  int* err;
  int parity;
};

// value the boundary routine would have put at (i, j) for component l (0..2);
// (i, j) is at most one cell outside the interior, in x or in y
// (sgn = 0: the boundary routine puts a zero there)
__device__ inline int nbr_map(const Grid2D& g, int i, int j, int l, double& sgn) {
  sgn = 1.0;
  if (j < g.nys) j = g.nye;
  else if (j > g.nye) j = g.nys;
  if (g.bc == WPIC_BC_SHOCK) {
    // boundary_shock.f90:603-621
    if (i < g.nxgs) {
      if (l == 0) { sgn = -1.0; i = g.nxgs; } else i = g.nxgs + 1;
    } else if (i > g.nxge) {
      sgn = 0.0;
      i = g.nxge;
    }
  } else if (g.bc == WPIC_BC_WALLX) {
    if (i < g.nxgs) {
      if (l == 0) { sgn = -1.0; i = g.nxgs; } else i = g.nxgs + 1;
    } else if (i > g.nxge) {
      if (l == 0) { sgn = -1.0; i = g.nxge - 2; } else i = g.nxge - 1;
    }
  } else {
    if (i < g.nxgs) i = g.nxge;
    else if (i > g.nxge) i = g.nxgs;
  }
  return cidx(g, i, j);
}

// the ghost value for the sign nbr_map returned: v, -v or a +0 zero
__device__ inline double signed_val(double sgn, double v) { return sgn > 0.0 ? v : (sgn < 0.0 ? -v : 0.0); }

__device__ inline double nbr(const Grid2D& g, const double* v, int i, int j, int l) {
  double s;
  const int k = nbr_map(g, i, j, l, s);
  return signed_val(s, v[k]);
}

// gkl (field.f90:125-146), b = f5*gkl, warm start phi = df(l), sum b^2.
// Component l = blockIdx.y: three blocks per cell range, each thread's sum
// over its cells in the same order as with one block for all three.
__global__ __launch_bounds__(256) void k_cg_init(CgArgs a) {
  __shared__ double sh[4];
  const Grid2D& g = a.g;
  const int n = g.nx * g.ny;
  const size_t pl = (size_t)g.fw * g.fh;
  const double *U1 = a.uf, *U2 = a.uf + pl, *U3 = a.uf + 2 * pl, *U4 = a.uf + 3 * pl, *U5 = a.uf + 4 * pl,
               *U6 = a.uf + 5 * pl;
  const double *J1 = a.uj, *J2 = a.uj + pl, *J3 = a.uj + 2 * pl;
  const double f1 = g.f1, f2 = g.f2, f3 = g.f3, f5 = g.f5;
  const int l = blockIdx.y;
  double s = 0.0;
  for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += NPB * 256) {
    const int i = g.nxgs + k % g.nx, j = g.nys + k / g.nx;
#define F(P, ii, jj) P[fidx(g, ii, jj)]
    double gk;
    if (l == 0)
      gk = +f2 * (+F(U1, i, j - 1) + F(U1, i - 1, j) - 4. * F(U1, i, j) + F(U1, i + 1, j) + F(U1, i, j + 1) +
                  f3 * (-F(J3, i, j - 1) + F(J3, i, j))) -
           f1 * (-F(U6, i, j - 1) + F(U6, i, j));
    else if (l == 1)
      gk = +f2 * (+F(U2, i, j - 1) + F(U2, i - 1, j) - 4. * F(U2, i, j) + F(U2, i + 1, j) + F(U2, i, j + 1) -
                  f3 * (-F(J3, i - 1, j) + F(J3, i, j))) +
           f1 * (-F(U6, i - 1, j) + F(U6, i, j));
    else
      gk = +f2 * (+F(U3, i, j - 1) + F(U3, i - 1, j) - 4. * F(U3, i, j) + F(U3, i + 1, j) + F(U3, i, j + 1) +
                  f3 * (-F(J2, i - 1, j) + F(J2, i, j) + F(J1, i, j - 1) - F(J1, i, j))) -
           f1 * (-F(U5, i - 1, j) + F(U5, i, j) + F(U4, i, j - 1) - F(U4, i, j));
#undef F
    const double bv = f5 * gk;
    a.b[(size_t)l * n + k] = bv;
    a.phi[(size_t)l * n + k] = a.df[pl * l + fidx(g, i, j)];
    s = s + bv * bv;
  }
  if (!a.parity) {
    const double t = block_sum256(s, sh);
    if (threadIdx.x == 0) a.part[((size_t)l * NPB + blockIdx.x) * 2] = t;
  }
}

// serial sum over interior cells in (j, i) order of x*y (parity mode)
__device__ inline double serial_dot(const double* x, const double* y, int n) {
  double s = 0.0;
  for (int k = 0; k < n; k++) s = s + x[k] * y[k];
  return s;
}

__device__ inline double part_sum(const CgArgs& a, int l, int slot, double* sh) {
  const double v = a.part[((size_t)l * NPB + threadIdx.x) * 2 + slot];
  return block_sum256(v, sh);
}

// sum_g = |b|^2, eps = sqrt(sum_g)*err (field.f90:351-364)
__global__ __launch_bounds__(256) void k_cg_fin0(CgArgs a) {
  __shared__ double sh[4];
  const int l = blockIdx.x, n = a.g.nx * a.g.ny;
  double t;
  if (a.parity) {
    t = threadIdx.x == 0 ? serial_dot(a.b + (size_t)l * n, a.b + (size_t)l * n, n) : 0.0;
  } else {
    t = part_sum(a, l, 0, sh);
  }
  if (threadIdx.x == 0) {
    a.sc[l].sum_g = t;
    a.sc[l].eps = sqrt(t) * a.g.cg_err;
  }
}

// r = b + (sum of neighbours) - f4 phi; p = r; sum r^2 (field.f90:366-381);
// component l = blockIdx.y, as in k_cg_init
__global__ __launch_bounds__(256) void k_cg_r0(CgArgs a) {
  __shared__ double sh[4];
  const Grid2D& g = a.g;
  const int n = g.nx * g.ny;
  const double f4 = g.f4;
  const int l = blockIdx.y;
  const double* ph = a.phi + (size_t)l * n;
  double s = 0.0;
  for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += NPB * 256) {
    const int i = g.nxgs + k % g.nx, j = g.nys + k / g.nx;
    const double rv = a.b[(size_t)l * n + k] + nbr(g, ph, i, j - 1, l) + nbr(g, ph, i - 1, j, l) - f4 * ph[k] +
                      nbr(g, ph, i + 1, j, l) + nbr(g, ph, i, j + 1, l);
    a.r[(size_t)l * n + k] = rv;
    a.pold[(size_t)l * n + k] = rv;
    s = s + rv * rv;
  }
  if (!a.parity) {
    const double t = block_sum256(s, sh);
    if (threadIdx.x == 0) a.part[((size_t)l * NPB + blockIdx.x) * 2] = t;
  }
}

// sumr_g; enter the loop iff sqrt(sumr_g) > eps and sum_g > eps (:383-387)
__global__ __launch_bounds__(256) void k_cg_fin1(CgArgs a) {
  __shared__ double sh[4];
  const int l = blockIdx.x, n = a.g.nx * a.g.ny;
  double t;
  if (a.parity)
    t = threadIdx.x == 0 ? serial_dot(a.r + (size_t)l * n, a.r + (size_t)l * n, n) : 0.0;
  else
    t = part_sum(a, l, 0, sh);
  if (threadIdx.x == 0) {
    CgScal& c = a.sc[l];
    c.sumr_g = t;
    c.ite = 0;
    c.bv = 0.0;
    c.active = (sqrt(t) > c.eps) && (c.sum_g > c.eps);
  }
}

// p = r + bv p (previous iteration's :444-450, fused here), then
// ap = f4 p - sum of neighbours; sum r^2 and sum p.ap (:391-407)
__global__ __launch_bounds__(256) void k_cg_a(CgArgs a) {
  __shared__ double sh[4];
  const int l = blockIdx.y;
  if (!a.sc[l].active) return;
  const Grid2D& g = a.g;
  const int n = g.nx * g.ny;
  const double f4 = g.f4, bv = a.sc[l].bv;
  const double *r = a.r + (size_t)l * n, *po = a.pold + (size_t)l * n;
  double* pn = a.pnew + (size_t)l * n;
  double* ap = a.ap + (size_t)l * n;
  double s_r = 0.0, s_2 = 0.0;
  for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += NPB * 256) {
    const int i = g.nxgs + k % g.nx, j = g.nys + k / g.nx;
    double sg;
    int m;
    m = nbr_map(g, i, j - 1, l, sg);
    const double pS = signed_val(sg, r[m] + bv * po[m]);
    m = nbr_map(g, i - 1, j, l, sg);
    const double pW = signed_val(sg, r[m] + bv * po[m]);
    m = nbr_map(g, i + 1, j, l, sg);
    const double pE = signed_val(sg, r[m] + bv * po[m]);
    m = nbr_map(g, i, j + 1, l, sg);
    const double pN = signed_val(sg, r[m] + bv * po[m]);
    const double pC = r[k] + bv * po[k];
    const double av = -pS - pW + f4 * pC - pE - pN;
    pn[k] = pC;
    ap[k] = av;
    s_r = s_r + r[k] * r[k];
    s_2 = s_2 + pC * av;
  }
  if (!a.parity) {
    const double t1 = block_sum256(s_r, sh);
    const double t2 = block_sum256(s_2, sh);
    if (threadIdx.x == 0) {
      a.part[((size_t)l * NPB + blockIdx.x) * 2] = t1;
      a.part[((size_t)l * NPB + blockIdx.x) * 2 + 1] = t2;
    }
  }
}

// av = sumr_g / sum2_g (:409-415); sum_g = sqrt(sumr_g) (:426)
__global__ __launch_bounds__(256) void k_cg_fa(CgArgs a) {
  __shared__ double sh[4];
  const int l = blockIdx.x;
  if (!a.sc[l].active) return;
  const int n = a.g.nx * a.g.ny;
  double t1, t2;
  if (a.parity) {
    t1 = threadIdx.x == 0 ? serial_dot(a.r + (size_t)l * n, a.r + (size_t)l * n, n) : 0.0;
    t2 = threadIdx.x == 0 ? serial_dot(a.pnew + (size_t)l * n, a.ap + (size_t)l * n, n) : 0.0;
  } else {
    t1 = part_sum(a, l, 0, sh);
    t2 = part_sum(a, l, 1, sh);
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

// phi += av p; r -= av ap (:417-424); sum r^2 (:432-439)
__global__ __launch_bounds__(256) void k_cg_b(CgArgs a) {
  __shared__ double sh[4];
  const int l = blockIdx.y;
  if (!a.sc[l].active) return;
  const Grid2D& g = a.g;
  const int n = g.nx * g.ny;
  const double av = a.sc[l].av;
  double *ph = a.phi + (size_t)l * n, *r = a.r + (size_t)l * n;
  const double *p = a.pnew + (size_t)l * n, *ap = a.ap + (size_t)l * n;
  double s = 0.0;
  for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += NPB * 256) {
    ph[k] = ph[k] + av * p[k];
    const double rv = r[k] - av * ap[k];
    r[k] = rv;
    s = s + rv * rv;
  }
  if (!a.parity) {
    const double t = block_sum256(s, sh);
    if (threadIdx.x == 0) a.part[((size_t)l * NPB + blockIdx.x) * 2] = t;
  }
}

// ite >= ite_max stops (:427-430); bv = sum1/sumr_g (:441-442); loop test
__global__ __launch_bounds__(256) void k_cg_fb(CgArgs a) {
  __shared__ double sh[4];
  const int l = blockIdx.x;
  if (!a.sc[l].active) return;
  const int n = a.g.nx * a.g.ny;
  double t;
  if (a.parity)
    t = threadIdx.x == 0 ? serial_dot(a.r + (size_t)l * n, a.r + (size_t)l * n, n) : 0.0;
  else
    t = part_sum(a, l, 0, sh);
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

// db(l, interior) = phi (:455-457)
__global__ void k_cg_store(CgArgs a) {
  const Grid2D& g = a.g;
  const int n = g.nx * g.ny;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = g.nxgs + k % g.nx, j = g.nys + k / g.nx;
  const size_t pl = (size_t)g.fw * g.fh;
  for (int l = 0; l < 3; l++) a.df[pl * l + fidx(g, i, j)] = a.phi[(size_t)l * n + k];
}

// E update (field.f90:154-171)
__global__ void k_e_update(Grid2D g, const double* __restrict__ uf, const double* __restrict__ uj,
                           double* __restrict__ df) {
  const int n = g.nx * g.ny;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = g.nxgs + k % g.nx, j = g.nys + k / g.nx;
  const size_t pl = (size_t)g.fw * g.fh;
  const double f1 = g.f1, gfac = g.gfac, pdt = g.pi4dt;
#define D(c, ii, jj) df[pl * ((c)-1) + fidx(g, ii, jj)]
#define U(c, ii, jj) uf[pl * ((c)-1) + fidx(g, ii, jj)]
#define J(c, ii, jj) uj[pl * ((c)-1) + fidx(g, ii, jj)]
  const double e4 = +f1 * (+gfac * (-D(3, i, j) + D(3, i, j + 1)) + (-U(3, i, j) + U(3, i, j + 1))) - pdt * J(1, i, j);
  const double e5 = -f1 * (+gfac * (-D(3, i, j) + D(3, i + 1, j)) + (-U(3, i, j) + U(3, i + 1, j))) - pdt * J(2, i, j);
  const double e6 = +f1 * (+gfac * (-D(2, i, j) + D(2, i + 1, j) + D(1, i, j) - D(1, i, j + 1)) +
                           (-U(2, i, j) + U(2, i + 1, j) + U(1, i, j) - U(1, i, j + 1))) -
                    pdt * J(3, i, j);
  D(4, i, j) = e4;
  D(5, i, j) = e5;
  D(6, i, j) = e6;
#undef D
#undef U
#undef J
}

__global__ void k_uf_add(size_t n, double* __restrict__ uf, const double* __restrict__ df) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
    uf[k] = uf[k] + df[k];
}

// uf += df over columns nxgs-2..nxge+2 of a cut grid (field.f90:176-184 with
// nxe below the planes' edge); the columns beyond keep their values
__global__ void k_uf_add_cols(Grid2D g, double* __restrict__ uf, const double* __restrict__ df) {
  const int w = g.nx + 4;
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (t >= (size_t)6 * w * g.fh) return;
  const int i = (int)(t % w);
  const size_t rest = t / w;
  const int J = (int)(rest % g.fh), comp = (int)(rest / g.fh);
  const size_t k = (size_t)g.fw * g.fh * comp + (size_t)g.fw * J + i;
  uf[k] = uf[k] + df[k];
}

__global__ void k_fill(double* p, size_t n, double v) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) p[k] = v;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// The field solve's grid.  While the shock's box expands the solve spans
// nodes nxgs..nxa (field.f90 runs every loop over nxs..nxe, the CG's vectors
// and sums included, and its boundary routines act at nxe+1): the same grid
// with nx and nxge cut there, over the full-size planes (fidx keeps the
// storage width fw).
static Grid2D field_grid(const wpic_ctx* c) {
  Grid2D g = c->g;
  if (g.nxa != g.nxge) {
    g.nx = g.nxa - g.nxgs + 1;
    g.nxge = g.nxa;
  }
  return g;
}

static CgArgs cg_args(wpic_ctx* c) {
  CgArgs a;
  a.g = field_grid(c);
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

static int grid_for(size_t n, int bs = 256) {
  size_t g = (n + bs - 1) / bs;
  if (g > 65535u * 16u) g = 65535u * 16u;
  return (int)(g ? g : 1);
}

hipError_t launch_fill(wpic_ctx* c, double* p, size_t n, double v) {
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n) < 8192 ? grid_for(n) : 8192), dim3(256), 0, c->st, p, n, v);
  return hipGetLastError();
}

hipError_t launch_zero(wpic_ctx* c, double* p, size_t n) { return launch_fill(c, p, n, 0.0); }

// (the y fold runs over every column of the plane: beyond nxa+2 the current
// is zero while the shock's box expands)
hipError_t launch_curre_bc(wpic_ctx* c) {
  hipLaunchKernelGGL(k_curre_y, dim3(grid_for(3 * c->g.fw)), dim3(256), 0, c->st, c->g, c->uj);
  if (c->g.bc == WPIC_BC_PERIODIC)
    hipLaunchKernelGGL(k_curre_x, dim3(grid_for(3 * c->g.fh)), dim3(256), 0, c->st, c->g, c->uj);
  return hipGetLastError();
}

hipError_t launch_dfield_bc(wpic_ctx* c) {
  const Grid2D g = field_grid(c);
  hipLaunchKernelGGL(k_dfield_y, dim3(grid_for(6 * g.nx)), dim3(256), 0, c->st, g, c->df);
  hipLaunchKernelGGL(k_dfield_x, dim3(grid_for(g.fh)), dim3(256), 0, c->st, g, c->df);
  return hipGetLastError();
}

hipError_t launch_cg_init(wpic_ctx* c) {
  CgArgs a = cg_args(c);
  hipLaunchKernelGGL(k_cg_init, dim3(NPB, 3), dim3(256), 0, c->st, a);
  hipLaunchKernelGGL(k_cg_fin0, dim3(3), dim3(256), 0, c->st, a);
  hipLaunchKernelGGL(k_cg_r0, dim3(NPB, 3), dim3(256), 0, c->st, a);
  hipLaunchKernelGGL(k_cg_fin1, dim3(3), dim3(256), 0, c->st, a);
  return hipGetLastError();
}

// the solve's start after b and phi are set elsewhere (the adjoint's transposed solves)
hipError_t launch_cg_start(wpic_ctx* c) {
  CgArgs a = cg_args(c);
  hipLaunchKernelGGL(k_cg_fin0, dim3(3), dim3(256), 0, c->st, a);
  hipLaunchKernelGGL(k_cg_r0, dim3(NPB, 3), dim3(256), 0, c->st, a);
  hipLaunchKernelGGL(k_cg_fin1, dim3(3), dim3(256), 0, c->st, a);
  return hipGetLastError();
}

hipError_t launch_cg_iter(wpic_ctx* c, int it) {
  CgArgs a = cg_args(c);
  // p alternates between the two buffers: iteration it reads (it even ? pa : pb)
  if (it & 1) {
    a.pold = c->pb;
    a.pnew = c->pa;
  }
  hipLaunchKernelGGL(k_cg_a, dim3(NPB, 3), dim3(256), 0, c->st, a);
  hipLaunchKernelGGL(k_cg_fa, dim3(3), dim3(256), 0, c->st, a);
  hipLaunchKernelGGL(k_cg_b, dim3(NPB, 3), dim3(256), 0, c->st, a);
  hipLaunchKernelGGL(k_cg_fb, dim3(3), dim3(256), 0, c->st, a);
  return hipGetLastError();
}

hipError_t launch_cg_store(wpic_ctx* c) {
  CgArgs a = cg_args(c);
  hipLaunchKernelGGL(k_cg_store, dim3(grid_for((size_t)a.g.nx * a.g.ny)), dim3(256), 0, c->st, a);
  return hipGetLastError();
}

hipError_t launch_e_update(wpic_ctx* c) {
  const Grid2D g = field_grid(c);
  hipLaunchKernelGGL(k_e_update, dim3(grid_for((size_t)g.nx * g.ny)), dim3(256), 0, c->st, g, c->uf, c->uj, c->df);
  return hipGetLastError();
}

hipError_t launch_uf_add(wpic_ctx* c) {
  if (c->g.nxa != c->g.nxge) {
    const Grid2D g = field_grid(c);
    const size_t n = (size_t)6 * (g.nx + 4) * g.fh;
    hipLaunchKernelGGL(k_uf_add_cols, dim3(grid_for(n)), dim3(256), 0, c->st, g, c->uf, c->df);
    return hipGetLastError();
  }
  const size_t n = (size_t)6 * c->g.fw * c->g.fh;
  hipLaunchKernelGGL(k_uf_add, dim3(grid_for(n) < 8192 ? grid_for(n) : 8192), dim3(256), 0, c->st, n, c->uf, c->df);
  return hipGetLastError();
}

}  // namespace wpic
