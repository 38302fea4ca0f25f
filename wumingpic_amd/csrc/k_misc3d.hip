// k_misc3d.hip -- 3D layout conversion at the boundary (reference AoS of
// ndim = 7 doubles <-> device SoA), the seeded initial state and diagnostics.
#include "wpic_internal3d.h"

namespace wpic {

struct Xfer3Args {
  Grid3D g;
  Parts3 P;
  const int *colbase, *np2;
  const int* R;              // non-null: records seen through a permutation (the exchange's R, or A's aperm)
  const long long* aosbase;  // [nsp*ncol] start of each column in the AoS buffer
};

__global__ void k_aos_to_soa3(Xfer3Args a, const double* __restrict__ aos) {
  const int rw = blockIdx.x;
  const int base = a.colbase[rw], n = a.np2[rw];
  const double* src = aos + a.aosbase[rw] * 7;
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    const double* s = src + (size_t)q * 7;
    const int d = base + q;
    a.P.x[d] = s[0];
    a.P.y[d] = s[1];
    a.P.z[d] = s[2];
    a.P.ux[d] = s[3];
    a.P.uy[d] = s[4];
    a.P.uz[d] = s[5];
    uint64_t b;
    __builtin_memcpy(&b, &s[6], 8);
    a.P.id[d] = b;
  }
}

__global__ void k_soa_to_aos3(Xfer3Args a, double* __restrict__ aos) {
  const int rw = blockIdx.x;
  const int base = a.colbase[rw], n = a.np2[rw];
  double* dst = aos + a.aosbase[rw] * 7;
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    int s = a.R ? a.R[base + q] : base + q;
    if ((unsigned)s >= (unsigned)(a.g.ncap * a.g.nsp)) s = 0;
    double* d = dst + (size_t)q * 7;
    d[0] = a.P.x[s];
    d[1] = a.P.y[s];
    d[2] = a.P.z[s];
    d[3] = a.P.ux[s];
    d[4] = a.P.uy[s];
    d[5] = a.P.uz[s];
    const uint64_t b = a.P.id[s];
    __builtin_memcpy(&d[6], &b, 8);
  }
}

__global__ void k_planes_from_aos3(size_t n, int ncomp, const double* __restrict__ aos, double* __restrict__ pl) {
  const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (q >= n) return;
  for (int c = 0; c < ncomp; c++) pl[(size_t)c * n + q] = aos[q * ncomp + c];
}

__global__ void k_planes_to_aos3(size_t n, int ncomp, const double* __restrict__ pl, double* __restrict__ aos) {
  const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (q >= n) return;
  for (int c = 0; c < ncomp; c++) aos[q * ncomp + c] = pl[(size_t)c * n + q];
}

hipError_t launch_planes_from_aos3d(wpic_ctx3d* c, const double* aos, double* planes, int ncomp) {
  const size_t n = plane3(c->g);
  hipLaunchKernelGGL(k_planes_from_aos3, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->st, n, ncomp, aos,
                     planes);
  return hipGetLastError();
}

hipError_t launch_planes_to_aos3d(wpic_ctx3d* c, const double* planes, double* aos, int ncomp) {
  const size_t n = plane3(c->g);
  hipLaunchKernelGGL(k_planes_to_aos3, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->st, n, ncomp, planes,
                     aos);
  return hipGetLastError();
}

// which: 0 = A columns, 1 = B seen through R (exchange pending), 2 = B columns
hipError_t launch_up_to_soa3d(wpic_ctx3d* c, const double* aos, long long n, int which) {
  (void)n;
  Xfer3Args a;
  a.g = c->g;
  a.P = which == 0 ? c->A : c->B;
  a.colbase = which == 0 ? c->colbaseA : c->colbaseB;
  a.np2 = which == 0 ? c->np2A : c->np2B;
  a.R = nullptr;
  a.aosbase = c->aosbase;
  hipLaunchKernelGGL(k_aos_to_soa3, dim3((unsigned)(c->g.nsp * c->g.ncol)), dim3(256), 0, c->st, a, aos);
  return hipGetLastError();
}

hipError_t launch_soa_to_aos3d(wpic_ctx3d* c, double* aos, int which) {
  Xfer3Args a;
  a.g = c->g;
  a.P = which == 0 ? c->A : c->B;
  if (which == 0) {
    a.colbase = c->colbaseA;
    a.np2 = c->np2A;
    a.R = c->aperm;
  } else if (which == 1) {
    a.colbase = c->colbaseN;
    a.np2 = c->np2N;
    a.R = c->R;
  } else {
    a.colbase = c->colbaseB;
    a.np2 = c->np2B;
    a.R = nullptr;
  }
  a.aosbase = c->aosbase;
  hipLaunchKernelGGL(k_soa_to_aos3, dim3((unsigned)(c->g.nsp * c->g.ncol)), dim3(256), 0, c->st, a, aos);
  return hipGetLastError();
}

// mixed-precision transfers: the reference's FP64 AoS columns <-> x cell +
// FP32 offsets (y, z relative to the column) and momenta
__global__ void k_aos_to_mixed3(Xfer3Args a, PartsM3 M, const double* __restrict__ aos) {
  const int rw = blockIdx.x;
  const int col = rw % a.g.ncol;
  const int j = a.g.nys + col % a.g.ny, k = a.g.nzs + col / a.g.ny;
  const int base = a.colbase[rw], n = a.np2[rw];
  const double* src = aos + a.aosbase[rw] * 7;
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    const double* s = src + (size_t)q * 7;
    const int d = base + q;
    const int ix = (int)s[0];
    M.ix[d] = ix;
    M.ox[d] = to_offset(s[0] - ix);
    M.oy[d] = to_offset(s[1] - j);
    M.oz[d] = to_offset(s[2] - k);
    M.ux[d] = (float)s[3];
    M.uy[d] = (float)s[4];
    M.uz[d] = (float)s[5];
    uint64_t b;
    __builtin_memcpy(&b, &s[6], 8);
    M.id[d] = b;
  }
}

__global__ void k_mixed_to_aos3(Xfer3Args a, PartsM3 M, double* __restrict__ aos) {
  const int rw = blockIdx.x;
  const int col = rw % a.g.ncol;
  const int j = a.g.nys + col % a.g.ny, k = a.g.nzs + col / a.g.ny;
  const int base = a.colbase[rw], n = a.np2[rw];
  double* dst = aos + a.aosbase[rw] * 7;
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    int s = a.R ? a.R[base + q] : base + q;
    if ((unsigned)s >= (unsigned)(a.g.ncap * a.g.nsp)) s = 0;
    double* d = dst + (size_t)q * 7;
    d[0] = (double)M.ix[s] + (double)M.ox[s];
    d[1] = (double)j + (double)M.oy[s];
    d[2] = (double)k + (double)M.oz[s];
    d[3] = M.ux[s];
    d[4] = M.uy[s];
    d[5] = M.uz[s];
    const uint64_t b = M.id[s];
    __builtin_memcpy(&d[6], &b, 8);
  }
}

hipError_t launch_up_to_mixed3d(wpic_ctx3d* c, const double* aos) {
  Xfer3Args a;
  a.g = c->g;
  a.P = c->A;
  a.colbase = c->colbaseA;
  a.np2 = c->np2A;
  a.R = nullptr;
  a.aosbase = c->aosbase;
  hipLaunchKernelGGL(k_aos_to_mixed3, dim3((unsigned)(c->g.nsp * c->g.ncol)), dim3(256), 0, c->st, a, c->MA, aos);
  return hipGetLastError();
}

hipError_t launch_mixed_to_aos3d(wpic_ctx3d* c, double* aos) {
  Xfer3Args a;
  a.g = c->g;
  a.P = c->A;
  a.colbase = c->colbaseA;
  a.np2 = c->np2A;
  a.R = c->aperm;
  a.aosbase = c->aosbase;
  hipLaunchKernelGGL(k_mixed_to_aos3, dim3((unsigned)(c->g.nsp * c->g.ncol)), dim3(256), 0, c->st, a, c->MA, aos);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// seeded initial state: counter-based splitmix64 draws keyed by (seed, stream,
// column, cell, slot); Box-Muller normals
// ---------------------------------------------------------------------------
namespace {
__device__ inline uint64_t smix3(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__device__ inline uint64_t rkey3(uint64_t seed, uint64_t stream, uint64_t a, uint64_t b, uint64_t c) {
  return smix3(smix3(smix3(smix3(seed ^ 0x3c6ef372fe94f82bull) ^ stream) ^ a) ^ ((b << 20) ^ c));
}
__device__ inline double u01_3(uint64_t h) { return (double)(h >> 11) * 0x1.0p-53; }
__device__ inline void normal2_3(uint64_t h1, uint64_t h2, double& n0, double& n1) {
  const double u1 = 1.0 - u01_3(h1);
  const double u2 = u01_3(h2);
  const double rr = sqrt(-2.0 * log(u1));
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  n0 = rr * c;
  n1 = rr * s;
}
// a coordinate uniform in cell `cell`, kept inside it after rounding
__device__ inline double in_cell(int cell, double u, double delx, double d_delx) {
  double x = ((double)cell + u) * delx;
  if ((int)(x * d_delx) != cell) x = next_down((double)(cell + 1) * delx);
  return x;
}
}  // namespace

struct Ic3Args {
  Grid3D g;
  Parts3 A;
  PartsM3 M;
  unsigned long long seed;
  int n_ppc;
  double udrift, uth_e, uth_i;
};

// MIXED: the same draws stored as x cell + FP32 offsets and momenta, so the
// FP64 and mixed forms start from the same state up to FP32 rounding
template <bool MIXED>
__global__ void k_ic_beams3(Ic3Args a) {
  const Grid3D& g = a.g;
  const long long colen = (long long)a.n_ppc * g.nx;
  const long long per = colen * g.ncol;
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= per * g.nsp) return;
  const int s = (int)(t / per);
  const long long q = t - s * per;
  const int col = (int)(q / colen);
  const long long m = q - col * colen;
  const int ci = (int)(m / a.n_ppc), sl = (int)(m - (long long)ci * a.n_ppc);
  const long long d = s * g.ncap + q;
  const int i = g.nxgs + ci, j = g.nys + col % g.ny, k = g.nzs + col / g.ny;
  const double x = in_cell(i, u01_3(rkey3(a.seed, 1, col, ci, sl)), g.delx, g.d_delx);
  const double y = in_cell(j, u01_3(rkey3(a.seed, 2, col, ci, sl)), g.delx, g.d_delx);
  const double z = in_cell(k, u01_3(rkey3(a.seed, 3, col, ci, sl)), g.delx, g.d_delx);
  double n0, n1, n2, n3;
  normal2_3(rkey3(a.seed, 10 + 4 * s, col, ci, sl), rkey3(a.seed, 11 + 4 * s, col, ci, sl), n0, n1);
  normal2_3(rkey3(a.seed, 12 + 4 * s, col, ci, sl), rkey3(a.seed, 13 + 4 * s, col, ci, sl), n2, n3);
  const bool electron = s == g.nsp - 1 && g.nsp > 1;
  const double th = electron ? a.uth_e : a.uth_i;
  double ux = th * n0;
  if (electron) ux += (sl & 1) ? -a.udrift : a.udrift;
  if (MIXED) {
    a.M.ix[d] = i;
    a.M.ox[d] = to_offset(x - i);
    a.M.oy[d] = to_offset(y - j);
    a.M.oz[d] = to_offset(z - k);
    a.M.ux[d] = (float)ux;
    a.M.uy[d] = (float)(th * n1);
    a.M.uz[d] = (float)(th * n2);
    a.M.id[d] = (uint64_t)(-(q + 1));
  } else {
    a.A.x[d] = x;
    a.A.y[d] = y;
    a.A.z[d] = z;
    a.A.ux[d] = ux;
    a.A.uy[d] = th * n1;
    a.A.uz[d] = th * n2;
    a.A.id[d] = (uint64_t)(-(q + 1));
  }
}

hipError_t launch_ic_beams3d(wpic_ctx3d* c, unsigned long long seed, int n_ppc, double udrift, double uth_e,
                             double uth_i) {
  Ic3Args a;
  a.g = c->g;
  a.A = c->A;
  a.M = c->MA;
  a.seed = seed;
  a.n_ppc = n_ppc;
  a.udrift = udrift;
  a.uth_e = uth_e;
  a.uth_i = uth_i;
  const long long n = (long long)n_ppc * c->g.nx * c->g.ncol * c->g.nsp;
  if (c->mixed)
    hipLaunchKernelGGL(k_ic_beams3<true>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->st, a);
  else
    hipLaunchKernelGGL(k_ic_beams3<false>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->st, a);
  return hipGetLastError();
}

// 3d/proj/weibel/app.f90 set_initial_condition + set_particle_ids: Bz = b0,
// n0 particles per cell on a regular x lattice along each column, y and z
// uniform in the column's cell, shared by the species; Maxwellian momenta with
// spread sd, t_ani times that along z; ids -1, -2, ... column by column (j
// fastest) per species.  FP64 form.
struct IcW3Args {
  Grid3D g;
  Parts3 A;
  unsigned long long seed;
  int n0;
  double sd[WPIC_MAXSP], t_ani;
};

__global__ void k_ic_weibel3(IcW3Args a) {
  const Grid3D& g = a.g;
  const long long colen = (long long)a.n0 * g.nx;
  const long long per = colen * g.ncol;
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= per * g.nsp) return;
  const int s = (int)(t / per);
  const long long q = t - s * per;
  const int col = (int)(q / colen);
  const int ii = (int)(q - col * colen) + 1;  // the reference's slot index
  const long long d = s * g.ncap + q;
  const int j = g.nys + col % g.ny, k = g.nzs + col / g.ny;
  const int np2 = (int)colen;
  a.A.x[d] = (g.nxgs + (g.nxge - g.nxgs + 1) * (ii - 5e-1) / np2) * g.delx;
  a.A.y[d] = (j + u01_3(rkey3(a.seed, 4, col, 0, ii))) * g.delx;
  a.A.z[d] = (k + u01_3(rkey3(a.seed, 5, col, 0, ii))) * g.delx;
  double n0, n1, n2, n3;
  normal2_3(rkey3(a.seed, 20 + 4 * s, col, 0, ii), rkey3(a.seed, 21 + 4 * s, col, 0, ii), n0, n1);
  normal2_3(rkey3(a.seed, 22 + 4 * s, col, 0, ii), rkey3(a.seed, 23 + 4 * s, col, 0, ii), n2, n3);
  a.A.ux[d] = a.sd[s] * n0;
  a.A.uy[d] = a.sd[s] * n1;
  a.A.uz[d] = a.t_ani * a.sd[s] * n2;
  a.A.id[d] = (uint64_t)(-(q + 1));
}

hipError_t launch_ic_weibel3d(wpic_ctx3d* c, unsigned long long seed, int n0, const double* sd, double t_ani) {
  IcW3Args a;
  a.g = c->g;
  a.A = c->A;
  a.seed = seed;
  a.n0 = n0;
  for (int s = 0; s < WPIC_MAXSP; s++) a.sd[s] = s < c->g.nsp ? sd[s] : 0.0;
  a.t_ani = t_ani;
  const long long n = (long long)n0 * c->g.nx * c->g.ncol * c->g.nsp;
  hipLaunchKernelGGL(k_ic_weibel3, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->st, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// diagnostics
// ---------------------------------------------------------------------------
namespace {
constexpr int RB3 = 256;
__device__ inline double bsum3(double v, double* sh) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return t;
}
}  // namespace

template <bool MIXED>
__global__ __launch_bounds__(256) void k_energy3(Grid3D g, Parts3 A, PartsM3 M, const int* colbase, const int* np2,
                                                 const int* ap, const double* uf, double* out, int sp) {
  __shared__ double sh[4];
  double kin = 0.0, e2 = 0.0, b2 = 0.0;
  const int ncols = g.nsp * g.ncol;
  for (int rw = blockIdx.x; rw < ncols; rw += gridDim.x) {
    const int s = rw / g.ncol;
    if (sp >= 0 && s != sp) continue;  // one species' kinetic energy
    const int base = colbase[rw], n = np2[rw];
    for (int q = threadIdx.x; q < n; q += 256) {
      const int p = ap ? ap[base + q] : base + q;
      const double ux = MIXED ? (double)M.ux[p] : A.ux[p];
      const double uy = MIXED ? (double)M.uy[p] : A.uy[p];
      const double uz = MIXED ? (double)M.uz[p] : A.uz[p];
      const double u2 = ux * ux + uy * uy + uz * uz;
      kin = kin + g.r[s] * (sqrt(1 + u2 / (g.c * g.c)) - 1);
    }
  }
  const size_t pl = plane3(g);
  const size_t ncell = (size_t)g.nx * g.ny * g.nz;
  for (size_t q = blockIdx.x * 256 + threadIdx.x; q < ncell; q += (size_t)gridDim.x * 256) {
    const int i = g.nxgs + (int)(q % g.nx), j = g.nys + (int)((q / g.nx) % g.ny),
              k = g.nzs + (int)(q / ((size_t)g.nx * g.ny));
    const size_t f = fidx3(g, i, j, k);
    b2 = b2 + uf[f] * uf[f] + uf[pl + f] * uf[pl + f] + uf[2 * pl + f] * uf[2 * pl + f];
    e2 = e2 + uf[3 * pl + f] * uf[3 * pl + f] + uf[4 * pl + f] * uf[4 * pl + f] + uf[5 * pl + f] * uf[5 * pl + f];
  }
  kin = bsum3(kin, sh);
  e2 = bsum3(e2, sh);
  b2 = bsum3(b2, sh);
  if (threadIdx.x == 0) {
    out[blockIdx.x * 3 + 0] = kin;
    out[blockIdx.x * 3 + 1] = e2;
    out[blockIdx.x * 3 + 2] = b2;
  }
}

hipError_t launch_energy3d(wpic_ctx3d* c, double* dout, int sp) {
  if (c->mixed)
    hipLaunchKernelGGL(k_energy3<true>, dim3(RB3), dim3(256), 0, c->st, c->g, c->A, c->MA, c->colbaseA, c->np2A,
                       c->aperm, c->uf, dout, sp);
  else
    hipLaunchKernelGGL(k_energy3<false>, dim3(RB3), dim3(256), 0, c->st, c->g, c->A, c->MA, c->colbaseA, c->np2A,
                       c->aperm, c->uf, dout, sp);
  return hipGetLastError();
}

// rho at cell centres with the quadratic shape, wrapped periodically, on the
// interior array [nx*ny*nz]
template <bool MIXED>
__global__ __launch_bounds__(256) void k_rho3(Grid3D g, Parts3 A, PartsM3 M, const int* colbase, const int* np2,
                                              const int* ap, double* rho) {
  const int rw = blockIdx.x;
  const int s = rw / g.ncol;
  const int base = colbase[rw], n = np2[rw];
  for (int q = threadIdx.x; q < n; q += 256) {
    const int p = ap ? ap[base + q] : base + q;
    const int jc = g.nys + (rw % g.ncol) % g.ny, kc = g.nzs + (rw % g.ncol) / g.ny;
    const double x = MIXED ? (double)M.ix[p] + (double)M.ox[p] : A.x[p];
    const double y = MIXED ? (double)jc + (double)M.oy[p] : A.y[p];
    const double z = MIXED ? (double)kc + (double)M.oz[p] : A.z[p];
    const int i0 = (int)(x * g.d_delx), j0 = (int)(y * g.d_delx), k0 = (int)(z * g.d_delx);
    if (i0 < g.nxgs || i0 > g.nxge || j0 < g.nys || j0 > g.nye || k0 < g.nzs || k0 > g.nze) continue;
    double sx[3], sy[3], sz[3];
    shape3(x * g.d_delx - 0.5 - i0, sx[0], sx[1], sx[2]);
    shape3(y * g.d_delx - 0.5 - j0, sy[0], sy[1], sy[2]);
    shape3(z * g.d_delx - 0.5 - k0, sz[0], sz[1], sz[2]);
    for (int c = 0; c < 3; c++)
      for (int b = 0; b < 3; b++)
        for (int a = 0; a < 3; a++) {
          int ii = i0 - 1 + a, jj = j0 - 1 + b, kk = k0 - 1 + c;
          if (ii < g.nxgs) ii += g.nx;
          if (ii > g.nxge) ii -= g.nx;
          if (jj < g.nys) jj += g.ny;
          if (jj > g.nye) jj -= g.ny;
          if (kk < g.nzs) kk += g.nz;
          if (kk > g.nze) kk -= g.nz;
          unsafeAtomicAdd(&rho[cidx3(g, ii, jj, kk)], g.q[s] * sx[a] * sy[b] * sz[c]);
        }
  }
}

__global__ void k_gauss3(Grid3D g, const double* uf, const double* rho, double* G) {
  const size_t n = (size_t)g.nx * g.ny * g.nz;
  const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int i = g.nxgs + (int)(q % g.nx), j = g.nys + (int)((q / g.nx) % g.ny), k = g.nzs + (int)(q / ((size_t)g.nx * g.ny));
  const size_t pl = plane3(g);
  const double *ex = uf + 3 * pl, *ey = uf + 4 * pl, *ez = uf + 5 * pl;
  const double dive = (ex[fidx3(g, i + 1, j, k)] - ex[fidx3(g, i, j, k)]) * g.d_delx +
                      (ey[fidx3(g, i, j + 1, k)] - ey[fidx3(g, i, j, k)]) * g.d_delx +
                      (ez[fidx3(g, i, j, k + 1)] - ez[fidx3(g, i, j, k)]) * g.d_delx;
  G[q] = dive - 4.0 * 3.141592653589793 * rho[q];
}

hipError_t launch_gauss3d(wpic_ctx3d* c, double* G, double* rho) {
  const size_t n = (size_t)c->g.nx * c->g.ny * c->g.nz;
  hipError_t e = launch_fill3d(c, rho, n, 0.0);
  if (e != hipSuccess) return e;
  if (c->mixed)
    hipLaunchKernelGGL(k_rho3<true>, dim3((unsigned)(c->g.nsp * c->g.ncol)), dim3(256), 0, c->st, c->g, c->A, c->MA,
                       c->colbaseA, c->np2A, c->aperm, rho);
  else
    hipLaunchKernelGGL(k_rho3<false>, dim3((unsigned)(c->g.nsp * c->g.ncol)), dim3(256), 0, c->st, c->g, c->A, c->MA,
                       c->colbaseA, c->np2A, c->aperm, rho);
  hipLaunchKernelGGL(k_gauss3, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->st, c->g, c->uf, rho, G);
  return hipGetLastError();
}

}  // namespace wpic
