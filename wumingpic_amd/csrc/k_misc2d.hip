// k_misc2d.hip -- layout conversion at the boundary (reference AoS <-> device
// SoA), the seeded initial state, and diagnostics (energy sums, Gauss law).
#include "wpic_internal.h"

namespace wpic {

// ---------------------------------------------------------------------------
// particles: compact AoS rows (ndim = 6 doubles per particle, the reference's
// up(:, ii, j, isp) contiguous) <-> SoA, over all rows of A (which = 0) or B
// (which = 1).  The AoS buffer holds species after species, rows in order.
// ---------------------------------------------------------------------------
struct XferArgs {
  Grid2D g;
  Parts P;
  const int *rowbase, *np2;
  const int* R;  // non-null: records seen through a permutation (the exchange's R, or A's aperm)
  const long long* aosbase;  // [nsp*ny] start of each row in the AoS buffer
};

__global__ void k_aos_to_soa(XferArgs a, const double* __restrict__ aos) {
  const int rw = blockIdx.x;
  const int base = a.rowbase[rw], n = a.np2[rw];
  const double* src = aos + a.aosbase[rw] * 6;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const double* s = src + (size_t)k * 6;
    const int d = base + k;
    a.P.x[d] = s[0];
    a.P.y[d] = s[1];
    a.P.ux[d] = s[2];
    a.P.uy[d] = s[3];
    a.P.uz[d] = s[4];
    uint64_t b;
    __builtin_memcpy(&b, &s[5], 8);
    a.P.id[d] = b;
  }
}

__global__ void k_soa_to_aos(XferArgs a, double* __restrict__ aos) {
  const int rw = blockIdx.x;
  const int base = a.rowbase[rw], n = a.np2[rw];
  double* dst = aos + a.aosbase[rw] * 6;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    int s = a.R ? a.R[base + k] : base + k;
    if ((unsigned)s >= (unsigned)(a.g.ncap * a.g.nsp)) s = 0;
    double* d = dst + (size_t)k * 6;
    d[0] = a.P.x[s];
    d[1] = a.P.y[s];
    d[2] = a.P.ux[s];
    d[3] = a.P.uy[s];
    d[4] = a.P.uz[s];
    const uint64_t b = a.P.id[s];
    __builtin_memcpy(&d[5], &b, 8);
  }
}

// fields: reference uf(ncomp, i, j) AoS <-> planes
__global__ void k_planes_from_aos(int n, int ncomp, const double* __restrict__ aos, double* __restrict__ pl) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  for (int c = 0; c < ncomp; c++) pl[(size_t)c * n + k] = aos[(size_t)k * ncomp + c];
}

__global__ void k_planes_to_aos(int n, int ncomp, const double* __restrict__ pl, double* __restrict__ aos) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  for (int c = 0; c < ncomp; c++) aos[(size_t)k * ncomp + c] = pl[(size_t)c * n + k];
}

hipError_t launch_planes_from_aos(wpic_ctx* c, const double* aos, double* planes, int ncomp) {
  const int n = c->g.fw * c->g.fh;
  hipLaunchKernelGGL(k_planes_from_aos, dim3((n + 255) / 256), dim3(256), 0, c->st, n, ncomp, aos, planes);
  return hipGetLastError();
}

hipError_t launch_planes_to_aos(wpic_ctx* c, const double* planes, double* aos, int ncomp) {
  const int n = c->g.fw * c->g.fh;
  hipLaunchKernelGGL(k_planes_to_aos, dim3((n + 255) / 256), dim3(256), 0, c->st, n, ncomp, planes, aos);
  return hipGetLastError();
}

// which: 0 = A rows, 1 = B rows (through R when the exchange is pending),
// 2 = B rows in B's own order.  aosbase lives in the staging tail.
hipError_t launch_up_to_soa(wpic_ctx* c, const double* aos, long long n, int which) {
  (void)n;
  XferArgs a;
  a.g = c->g;
  a.P = which == 0 ? c->A : c->B;
  a.rowbase = which == 0 ? c->rowbaseA : c->rowbaseB;
  a.np2 = which == 0 ? c->np2A : c->np2B;
  a.R = nullptr;
  a.aosbase = c->aosbase;
  hipLaunchKernelGGL(k_aos_to_soa, dim3(c->g.nsp * c->g.ny), dim3(256), 0, c->st, a, aos);
  return hipGetLastError();
}

hipError_t launch_soa_to_aos(wpic_ctx* c, double* aos, int which) {
  XferArgs a;
  a.g = c->g;
  a.P = which == 0 ? c->A : c->B;
  if (which == 0) {
    a.rowbase = c->rowbaseA;
    a.np2 = c->np2A;
    a.R = c->aperm;
  } else if (which == 1) {
    a.rowbase = c->rowbaseN;
    a.np2 = c->np2N;
    a.R = c->R;
  } else {
    a.rowbase = c->rowbaseB;
    a.np2 = c->np2B;
    a.R = nullptr;
  }
  a.aosbase = c->aosbase;
  hipLaunchKernelGGL(k_soa_to_aos, dim3(c->g.nsp * c->g.ny), dim3(256), 0, c->st, a, aos);
  return hipGetLastError();
}

// mixed-precision transfers: the reference's FP64 AoS rows <-> x cell + FP32
// offsets and momenta (y offset relative to the row)
__global__ void k_aos_to_mixed(XferArgs a, PartsM M, const double* __restrict__ aos) {
  const int rw = blockIdx.x;
  const int base = a.rowbase[rw], n = a.np2[rw];
  const int j = a.g.nys + rw % a.g.ny;
  const double* src = aos + a.aosbase[rw] * 6;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const double* s = src + (size_t)k * 6;
    const int d = base + k;
    const int ix = (int)s[0];
    M.ix[d] = ix;
    M.ox[d] = to_offset(s[0] - ix);
    M.oy[d] = to_offset(s[1] - j);
    M.ux[d] = (float)s[2];
    M.uy[d] = (float)s[3];
    M.uz[d] = (float)s[4];
    uint64_t b;
    __builtin_memcpy(&b, &s[5], 8);
    M.id[d] = b;
  }
}

__global__ void k_mixed_to_aos(XferArgs a, PartsM M, double* __restrict__ aos) {
  const int rw = blockIdx.x;
  const int base = a.rowbase[rw], n = a.np2[rw];
  const int j = a.g.nys + rw % a.g.ny;
  double* dst = aos + a.aosbase[rw] * 6;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    int s = a.R ? a.R[base + k] : base + k;
    if ((unsigned)s >= (unsigned)(a.g.ncap * a.g.nsp)) s = 0;
    double* d = dst + (size_t)k * 6;
    d[0] = (double)M.ix[s] + (double)M.ox[s];
    d[1] = (double)j + (double)M.oy[s];
    d[2] = M.ux[s];
    d[3] = M.uy[s];
    d[4] = M.uz[s];
    const uint64_t b = M.id[s];
    __builtin_memcpy(&d[5], &b, 8);
  }
}

hipError_t launch_up_to_mixed(wpic_ctx* c, const double* aos) {
  XferArgs a;
  a.g = c->g;
  a.P = c->A;
  a.rowbase = c->rowbaseA;
  a.np2 = c->np2A;
  a.R = nullptr;
  a.aosbase = c->aosbase;
  hipLaunchKernelGGL(k_aos_to_mixed, dim3(c->g.nsp * c->g.ny), dim3(256), 0, c->st, a, c->MA, aos);
  return hipGetLastError();
}

hipError_t launch_mixed_to_aos(wpic_ctx* c, double* aos) {
  XferArgs a;
  a.g = c->g;
  a.P = c->A;
  a.rowbase = c->rowbaseA;
  a.np2 = c->np2A;
  a.R = c->aperm;
  a.aosbase = c->aosbase;
  hipLaunchKernelGGL(k_mixed_to_aos, dim3(c->g.nsp * c->g.ny), dim3(256), 0, c->st, a, c->MA, aos);
  return hipGetLastError();
}

// the sorted state in slot order: dst[slot] = src[ap[slot]] over the active
// slots of the rows (rowbase then np2)
__global__ void k_gather_rows(Grid2D g, Parts src, const int* ap, Parts dst, const int* rows) {
  const int nr = g.nsp * g.ny;
  const int rw = blockIdx.x;
  const int base = rows[rw], n = rows[nr + rw];
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const int d = base + k;
    const int s = ap ? ap[d] : d;
    dst.x[d] = src.x[s];
    dst.y[d] = src.y[s];
    dst.ux[d] = src.ux[s];
    dst.uy[d] = src.uy[s];
    dst.uz[d] = src.uz[s];
    if (dst.id) dst.id[d] = src.id[s];
  }
}

hipError_t launch_gather_rows(wpic_ctx* c, const Parts& src, const int* ap, const Parts& dst, const int* rows) {
  hipLaunchKernelGGL(k_gather_rows, dim3(c->g.nsp * c->g.ny), dim3(256), 0, c->st, c->g, src, ap, dst, rows);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// seeded initial state: counter-based (splitmix64) draws keyed by
// (seed, stream, row, cell, slot), so the state is independent of the launch
// shape; Box-Muller normals.
// ---------------------------------------------------------------------------
__device__ inline uint64_t smix(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__device__ inline uint64_t rkey(uint64_t seed, uint64_t stream, uint64_t a, uint64_t b, uint64_t c) {
  return smix(smix(smix(smix(seed ^ 0x5851f42d4c957f2dull) ^ stream) ^ a) ^ ((b << 20) ^ c));
}
__device__ inline double u01(uint64_t h) { return (double)(h >> 11) * 0x1.0p-53; }
__device__ inline void normal2(uint64_t h1, uint64_t h2, double& n0, double& n1) {
  const double u1 = 1.0 - u01(h1);  // (0, 1]
  const double u2 = u01(h2);
  const double rr = sqrt(-2.0 * log(u1));
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  n0 = rr * c;
  n1 = rr * s;
}

struct IcArgs {
  Grid2D g;
  Parts A;
  PartsM M;
  uint64_t* id;
  unsigned long long seed;
  int n_ppc;
  double udrift, uth_e, uth_i;
};

// MIXED: the same draws stored as x cell + FP32 offsets and momenta, so the
// FP64 and mixed forms start from the same state up to FP32 rounding
template <bool MIXED>
__global__ void k_ic_beams(IcArgs a) {
  const Grid2D& g = a.g;
  const long long per = (long long)a.n_ppc * g.nx * g.ny;
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= per * g.nsp) return;
  const int s = (int)(t / per);
  const long long k = t - s * per;
  const long long rowlen = (long long)a.n_ppc * g.nx;
  const int r = (int)(k / rowlen);
  const long long m = k - r * rowlen;
  const int ci = (int)(m / a.n_ppc), sl = (int)(m - (long long)ci * a.n_ppc);
  const long long d = s * g.ncap + k;
  // positions shared by all species (co-located: zero charge density)
  const double u0 = u01(rkey(a.seed, 1, r, ci, sl)), u1 = u01(rkey(a.seed, 2, r, ci, sl));
  const int i = g.nxgs + ci, j = g.nys + r;
  double x = ((double)i + u0) * g.delx, y = ((double)j + u1) * g.delx;
  if ((int)(x * g.d_delx) != i) x = next_down((double)(i + 1) * g.delx);
  if ((int)(y * g.d_delx) != j) y = next_down((double)(j + 1) * g.delx);
  double n0, n1, n2, n3;
  normal2(rkey(a.seed, 10 + 4 * s, r, ci, sl), rkey(a.seed, 11 + 4 * s, r, ci, sl), n0, n1);
  normal2(rkey(a.seed, 12 + 4 * s, r, ci, sl), rkey(a.seed, 13 + 4 * s, r, ci, sl), n2, n3);
  const bool electron = s == g.nsp - 1 && g.nsp > 1;
  const double th = electron ? a.uth_e : a.uth_i;
  double ux = th * n0;
  if (electron) ux += (sl & 1) ? -a.udrift : a.udrift;
  if (MIXED) {
    a.M.ix[d] = i;
    a.M.ox[d] = to_offset(x - i);
    a.M.oy[d] = to_offset(y - j);
    a.M.ux[d] = (float)ux;
    a.M.uy[d] = (float)(th * n1);
    a.M.uz[d] = (float)(th * n2);
  } else {
    a.A.x[d] = x;
    a.A.y[d] = y;
    a.A.ux[d] = ux;
    a.A.uy[d] = th * n1;
    a.A.uz[d] = th * n2;
  }
  // id = -(pid), pid counted row by row from 1 (app.f90:455-470)
  const long long pid = k + 1;
  a.id[d] = (uint64_t)(-pid);
}

static IcArgs ic_args(wpic_ctx* c, unsigned long long seed, int n_ppc, double udrift, double uth_e,
                      double uth_i) {
  IcArgs a;
  a.g = c->g;
  a.A = c->A;
  a.M = c->MA;
  a.id = c->mixed ? c->MA.id : c->A.id;
  a.seed = seed;
  a.n_ppc = n_ppc;
  a.udrift = udrift;
  a.uth_e = uth_e;
  a.uth_i = uth_i;
  return a;
}

hipError_t launch_ic_beams(wpic_ctx* c, unsigned long long seed, int n_ppc, double udrift, double uth_e,
                           double uth_i) {
  IcArgs a = ic_args(c, seed, n_ppc, udrift, uth_e, uth_i);
  const long long n = (long long)n_ppc * c->g.nx * c->g.ny * c->g.nsp;
  hipLaunchKernelGGL(k_ic_beams<false>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->st, a);
  return hipGetLastError();
}

hipError_t launch_ic_beams_mixed(wpic_ctx* c, unsigned long long seed, int n_ppc, double udrift, double uth_e,
                                 double uth_i) {
  IcArgs a = ic_args(c, seed, n_ppc, udrift, uth_e, uth_i);
  const long long n = (long long)n_ppc * c->g.nx * c->g.ny * c->g.nsp;
  hipLaunchKernelGGL(k_ic_beams<true>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->st, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The reference apps' own initial states, with the OS-seeded random draws of
// the reference replaced by the counter-based draws above.
// ---------------------------------------------------------------------------
// 2d/proj/weibel/app.f90:366-438: n0 particles per cell per species, x on a
// regular lattice along each row, (nxgs + nx*(ii-0.5)/np2)*delx, y uniform in
// the row, shared by the species; Maxwellian momenta with spread sd and t_ani
// times that along z; ids -(row-major count) per species (app.f90:455-470).
struct IcWeibelArgs {
  Grid2D g;
  Parts A;
  unsigned long long seed;
  int n0;
  double sd[WPIC_MAXSP], t_ani;
};

__global__ void k_ic_weibel(IcWeibelArgs a) {
  const Grid2D& g = a.g;
  const long long rowlen = (long long)a.n0 * g.nx, per = rowlen * g.ny;
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= per * g.nsp) return;
  const int s = (int)(t / per);
  const long long k = t - s * per;
  const int r = (int)(k / rowlen);
  const int ii = (int)(k - r * rowlen) + 1;  // the reference's slot index
  const long long d = s * g.ncap + k;
  const int j = g.nys + r;
  const int np2 = (int)rowlen;
  a.A.x[d] = (g.nxgs + (g.nxge - g.nxgs + 1) * (ii - 0.5) / np2) * g.delx;
  a.A.y[d] = (j + u01(rkey(a.seed, 3, r, 0, ii))) * g.delx;
  double n0, n1, n2, n3;
  normal2(rkey(a.seed, 20 + 4 * s, r, 0, ii), rkey(a.seed, 21 + 4 * s, r, 0, ii), n0, n1);
  normal2(rkey(a.seed, 22 + 4 * s, r, 0, ii), rkey(a.seed, 23 + 4 * s, r, 0, ii), n2, n3);
  a.A.ux[d] = a.sd[s] * n0;
  a.A.uy[d] = a.sd[s] * n1;
  a.A.uz[d] = a.t_ani * a.sd[s] * n2;
  a.A.id[d] = (uint64_t)(-(k + 1));
}

// 2d/proj/reconnection/app.f90:373-452: the Harris sheet with the X-point
// perturbation.  Per row, ibg background particles uniform in
// [(nxgs+1)dx, (nxge-1)dx) and np2 - ibg sheet particles from the inverse of
// the sheet's cosh^-2 profile, shared by the species; Maxwellian momenta plus
// the drift that carries jz; written unsorted into B (the reference sorts
// them next, app.f90:346-348).
struct IcHarrisArgs {
  Grid2D g;
  Parts B;
  unsigned long long seed;
  int np2, ibg;
  double x0, y0, lcs, b0, e1, ncs, nbg, f[2], sd[2];
};

__device__ inline double harris_arg(const IcHarrisArgs& a, double x, double y) {
  return ((x - a.x0) * (x - a.x0) + (y - a.y0) * (y - a.y0)) / ((2 * a.lcs) * (2 * a.lcs));
}

__global__ void k_ic_harris(IcHarrisArgs a) {
  const Grid2D& g = a.g;
  const long long per = (long long)a.np2 * g.ny;
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= per) return;
  const int r = (int)(t / a.np2);
  const int ii = (int)(t - (long long)r * a.np2) + 1;
  const int j = g.nys + r;
  const double pi = 3.141592653589793;
  double x;
  if (ii <= a.ibg) {
    x = (g.nxgs + 1) * g.delx + u01(rkey(a.seed, 4, r, 0, ii)) * g.delx * (g.nxge - g.nxgs - 2);
  } else {
    double r1 = u01(rkey(a.seed, 5, r, 0, ii));
    r1 = (2.0 * r1 - 1.0) * tanh(0.5 * (g.nxge - g.nxgs - 2) * g.delx / a.lcs);
    x = a.lcs * 0.5 * (log(1.0 + r1) - log(1.0 - r1)) + a.x0;
  }
  const double y = (double)j * g.delx + u01(rkey(a.seed, 6, r, 0, ii)) * g.delx;
  const double ch = cosh((x - a.x0) / a.lcs);
  const double sech2 = 1.0 / (ch * ch);
  const double e = harris_arg(a, x, y);
  const double jz = +a.b0 / (4 * pi * a.lcs) * sech2 - 2 * a.e1 * a.b0 / (4 * pi * a.lcs) * (1.0 - e) * exp(-e);
  const double den = a.ncs * sech2 + a.nbg;
  for (int s = 0; s < 2; s++) {
    double n0, n1, n2, n3;
    normal2(rkey(a.seed, 30 + 4 * s, r, 0, ii), rkey(a.seed, 31 + 4 * s, r, 0, ii), n0, n1);
    normal2(rkey(a.seed, 32 + 4 * s, r, 0, ii), rkey(a.seed, 33 + 4 * s, r, 0, ii), n2, n3);
    const long long d = s * g.ncap + t;
    a.B.x[d] = x;
    a.B.y[d] = y;
    a.B.ux[d] = a.sd[s] * n0;
    a.B.uy[d] = a.sd[s] * n1;
    a.B.uz[d] = a.sd[s] * n2 + a.f[s] * jz / den;
    a.B.id[d] = (uint64_t)(-(t + 1));
  }
}

// the Harris field with the perturbation at the nodes (i dx, j dx), ghosts
// included (app.f90:397-412)
__global__ void k_ic_harris_field(IcHarrisArgs a, double* uf) {
  const Grid2D& g = a.g;
  const size_t pl = (size_t)g.fw * g.fh;
  const size_t n = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (n >= pl) return;
  const int i = g.nxgs - 2 + (int)(n % g.fw), j = g.nys - 2 + (int)(n / g.fw);
  const double x = i * g.delx, y = j * g.delx;
  const double e = exp(-harris_arg(a, x, y));
  uf[n] = 0.0 + a.e1 * a.b0 * ((y - a.y0) / a.lcs) * e;
  uf[pl + n] = a.b0 * tanh((x - a.x0) / a.lcs) + -a.e1 * a.b0 * ((x - a.x0) / a.lcs) * e;
  for (int c = 2; c < 6; c++) uf[c * pl + n] = 0.0;
}

hipError_t launch_ic_weibel(wpic_ctx* c, unsigned long long seed, int n0, const double* sd, double t_ani) {
  IcWeibelArgs a;
  a.g = c->g;
  a.A = c->A;
  a.seed = seed;
  a.n0 = n0;
  for (int s = 0; s < WPIC_MAXSP; s++) a.sd[s] = s < c->g.nsp ? sd[s] : 0.0;
  a.t_ani = t_ani;
  const long long n = (long long)n0 * c->g.nx * c->g.ny * c->g.nsp;
  hipLaunchKernelGGL(k_ic_weibel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->st, a);
  return hipGetLastError();
}

hipError_t launch_ic_harris(wpic_ctx* c, unsigned long long seed, int np2, int ibg, const double* hp) {
  IcHarrisArgs a;
  a.g = c->g;
  a.B = c->B;
  a.seed = seed;
  a.np2 = np2;
  a.ibg = ibg;
  a.x0 = hp[0];
  a.y0 = hp[1];
  a.lcs = hp[2];
  a.b0 = hp[3];
  a.e1 = hp[4];
  a.ncs = hp[5];
  a.nbg = hp[6];
  a.f[0] = hp[7];
  a.f[1] = hp[8];
  a.sd[0] = hp[9];
  a.sd[1] = hp[10];
  const long long n = (long long)np2 * c->g.ny;
  hipLaunchKernelGGL(k_ic_harris, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->st, a);
  const size_t pl = (size_t)c->g.fw * c->g.fh;
  hipLaunchKernelGGL(k_ic_harris_field, dim3((unsigned)((pl + 255) / 256)), dim3(256), 0, c->st, a, c->uf);
  return hipGetLastError();
}

// The shock app's initial state (2d/proj/shock/app.f90:405-470) over the
// active box (nxs = nxgs, nxe = nxa): per row np2 particles on a regular
// x lattice over [nxs, nxe) dx, y uniform in the row and shared by the
// species, Maxwellian momenta boosted by the inflow profile; written unsorted
// into B (the reference sorts them next); the upstream field with
// E = -vprofile(i dx) x B / c at every node.
struct IcShockArgs {
  Grid2D g;
  Parts B;
  unsigned long long seed;
  int np2;
  double v0, xc, xs, sd[2];
};

__device__ inline double vprofile(const IcShockArgs& a, double x) {
  return 0.5 * a.v0 * (1 + tanh((x - a.xc) / a.xs));
}

__global__ void k_ic_shock(IcShockArgs a) {
  const Grid2D& g = a.g;
  const long long per = (long long)a.np2 * g.ny;
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= per) return;
  const int r = (int)(t / a.np2);
  const int ii = (int)(t - (long long)r * a.np2) + 1;
  const int j = g.nys + r;
  const double c = g.c;
  const double x = (g.nxgs + (g.nxa - g.nxgs) * (ii - 0.5) / a.np2) * g.delx;
  const double y = (j + u01(rkey(a.seed, 7, r, 0, ii))) * g.delx;
  for (int s = 0; s < 2; s++) {
    double n0, n1, n2, n3;
    normal2(rkey(a.seed, 50 + 4 * s, r, 0, ii), rkey(a.seed, 51 + 4 * s, r, 0, ii), n0, n1);
    normal2(rkey(a.seed, 52 + 4 * s, r, 0, ii), rkey(a.seed, 53 + 4 * s, r, 0, ii), n2, n3);
    double ux = a.sd[s] * n0;
    const double uy = a.sd[s] * n1, uz = a.sd[s] * n2;
    const double v1 = vprofile(a, x);
    const double gam1 = 1 / sqrt(1 - (v1 / c) * (v1 / c));
    const double gamp = sqrt(1 + (ux * ux + uy * uy + uz * uz) / (c * c));
    ux = gam1 * (ux + v1 * gamp);
    const long long d = s * g.ncap + t;
    a.B.x[d] = x;
    a.B.y[d] = y;
    a.B.ux[d] = ux;
    a.B.uy[d] = uy;
    a.B.uz[d] = uz;
    // set_particle_ids (app.f90:500-524): the id is -pid, made positive
    // (a tracer written by io__orb) for x in [xrs, xre] = [450, 500]
    const bool tracer = x >= 450.0 && x <= 500.0;
    a.B.id[d] = tracer ? (uint64_t)(t + 1) : (uint64_t)(-(t + 1));
  }
}

__global__ void k_ic_shock_field(IcShockArgs a, double* uf, double bx, double by, double bz) {
  const Grid2D& g = a.g;
  const size_t pl = (size_t)g.fw * g.fh;
  const size_t n = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (n >= pl) return;
  const int i = g.nxgs - 2 + (int)(n % g.fw);
  uf[n] = bx;
  uf[pl + n] = by;
  uf[2 * pl + n] = bz;
  uf[3 * pl + n] = 0.0;
  uf[4 * pl + n] = +vprofile(a, i * g.delx) * uf[2 * pl + n] / g.c;
  uf[5 * pl + n] = -vprofile(a, i * g.delx) * uf[pl + n] / g.c;
}

hipError_t launch_ic_shock(wpic_ctx* c, unsigned long long seed, int np2, double v0, const double* sd, double l_damp,
                           const double* b) {
  IcShockArgs a;
  a.g = c->g;
  a.B = c->B;
  a.seed = seed;
  a.np2 = np2;
  a.v0 = v0;
  a.xc = l_damp + c->g.nxgs * c->g.delx;
  a.xs = l_damp * 0.1;
  a.sd[0] = sd[0];
  a.sd[1] = sd[1];
  const long long n = (long long)np2 * c->g.ny;
  hipLaunchKernelGGL(k_ic_shock, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->st, a);
  const size_t pl = (size_t)c->g.fw * c->g.fh;
  hipLaunchKernelGGL(k_ic_shock_field, dim3((unsigned)((pl + 255) / 256)), dim3(256), 0, c->st, a, c->uf, b[0], b[1],
                     b[2]);
  return hipGetLastError();
}

// The shock app's inject (2d/proj/shock/app.f90:697-856) at the active edge
// nxe = nxa, fused with the re-layout it needs: one block per destination
// row copies the row's sorted particles (read through the permutation) into
// B at the packed new base, then draws the row's nl injected particles after
// them: positions spread over |v0| dt beyond nxe, moved by (v0 + ux) dt, y
// uniform in the row and shared by the species, Maxwellian momenta boosted
// by the inflow profile v1(x), ids continuing the species' count.
// cumcnt(nxe) and every entry after it grow by nl: the injected particles
// close cell nxe-1's range, where the reference counts them, and the cells
// from nxe on keep their (empty) ranges.
//
// With `reloc` the same kernel is relocate (app.f90:615-692), called after
// nxa has grown by one: nl = n0 particles per row on a regular x lattice over
// cell nxe-1 (not moved), y uniform in the row and shared by the species,
// the same momenta and ids (pid = ii + (j - nygs) n0 + nptotal); cumcnt(nxe)
// = cumcnt(nxe-1) + n0 as the reference sets it, and the entries after it
// the row's new total (the cells beyond the new edge stay empty).
struct InjArgs {
  Grid2D g;
  Parts A, B;
  const int* ap;
  const int *base_old, *np2_old, *base_new;
  const int* nl;             // [ny]
  const long long* idoff;    // [nsp*ny]
  int* cumcnt;
  unsigned long long seed;
  int it, reloc;
  double v0, x0, sd[2], xc, xs;
};

__global__ __launch_bounds__(256) void k_inject(InjArgs a) {
  const Grid2D& g = a.g;
  const int rw = blockIdx.x;
  const int s = rw / g.ny, r = rw - s * g.ny;
  const int n_old = a.np2_old[rw], nl = a.nl[r];
  const int bo = a.base_old[rw], bn = a.base_new[rw];
  for (int k = threadIdx.x; k < n_old; k += 256) {
    const int ia = a.ap ? a.ap[bo + k] : bo + k;
    const int d = bn + k;
    a.B.x[d] = a.A.x[ia];
    a.B.y[d] = a.A.y[ia];
    a.B.ux[d] = a.A.ux[ia];
    a.B.uy[d] = a.A.uy[ia];
    a.B.uz[d] = a.A.uz[ia];
    a.B.id[d] = a.A.id[ia];
  }
  const double delx = g.delx, delt = g.delt, c = g.c;
  const int j = g.nys + r;
  const int tag = a.reloc ? 70 : 40;  // relocate draws from streams of its own
  for (int m = 1 + threadIdx.x; m <= nl; m += 256) {
    const int d = bn + n_old + m - 1;
    double x = a.reloc ? (g.nxa - 1) * delx + (m - 0.5) / nl * delx : g.nxa * delx + (m - 0.5) / nl * a.x0;
    const double y = (j + u01(rkey(a.seed, tag, r, (uint64_t)a.it, m))) * delx;
    double n0, n1, n2, n3;
    normal2(rkey(a.seed, tag + 1 + 2 * s, r, (uint64_t)a.it, m), rkey(a.seed, tag + 2 + 2 * s, r, (uint64_t)a.it, m),
            n0, n1);
    normal2(rkey(a.seed, tag + 5 + 2 * s, r, (uint64_t)a.it, m), rkey(a.seed, tag + 6 + 2 * s, r, (uint64_t)a.it, m),
            n2, n3);
    double ux = a.sd[s] * n0;
    const double uy = a.sd[s] * n1, uz = a.sd[s] * n2;
    if (!a.reloc) x = x + (a.v0 + ux) * delt;
    const double v1 = 0.5 * a.v0 * (1 + tanh((x - a.xc) / a.xs));
    const double gam1 = 1 / sqrt(1 - (v1 / c) * (v1 / c));
    const double gamp = sqrt(1 + (ux * ux + uy * uy + uz * uz) / (c * c));
    ux = gam1 * (ux + v1 * gamp);
    a.B.x[d] = x;
    a.B.y[d] = y;
    a.B.ux[d] = ux;
    a.B.uy[d] = uy;
    a.B.uz[d] = uz;
    a.B.id[d] = (uint64_t)(-((long long)m + a.idoff[rw]));
  }
  if (threadIdx.x == 0 && nl > 0) {
    int* cr = a.cumcnt + (size_t)rw * (g.nx + 1);
    const int e = g.nxa - g.nxgs;  // cumcnt(nxe)
    if (a.reloc) {
      cr[e] = cr[e - 1] + nl;
      for (int k = e + 1; k <= g.nx; k++) cr[k] = n_old + nl;
    } else {
      for (int k = e; k <= g.nx; k++) cr[k] += nl;
    }
  }
}

// the upstream field at i = nxe-1 and nxe over every row (app.f90:840-850,
// relocate's :681-690)
__global__ void k_inject_fields(Grid2D g, double* uf, double by, double bz, double v0) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.fh) return;
  const int J = g.nys - 2 + t;
  const size_t pl = (size_t)g.fw * g.fh;
  const int f1 = fidx(g, g.nxa - 1, J), f2 = fidx(g, g.nxa, J);
  uf[pl + f1] = by;
  uf[2 * pl + f1] = bz;
  uf[4 * pl + f1] = +v0 * uf[2 * pl + f1] / g.c;
  uf[5 * pl + f1] = -v0 * uf[pl + f1] / g.c;
  uf[pl + f2] = by;
  uf[2 * pl + f2] = bz;
}

hipError_t launch_inject(wpic_ctx* c, unsigned long long seed, int it, const int* base_new, const int* nl,
                         const long long* idoff, double v0, const double* sd, double l_damp, double by, double bz,
                         bool reloc) {
  InjArgs a;
  a.g = c->g;
  a.A = c->A;
  a.B = c->B;
  a.ap = c->aperm;
  a.base_old = c->rowbaseA;
  a.np2_old = c->np2A;
  a.base_new = base_new;
  a.nl = nl;
  a.idoff = idoff;
  a.cumcnt = c->cumcnt;
  a.seed = seed;
  a.it = it;
  a.reloc = reloc ? 1 : 0;
  a.v0 = v0;
  a.x0 = fabs(v0) * c->g.delt;
  a.sd[0] = sd[0];
  a.sd[1] = sd[1];
  a.xc = l_damp + c->g.nxgs * c->g.delx;
  a.xs = l_damp * 0.1;
  hipLaunchKernelGGL(k_inject, dim3(c->g.nsp * c->g.ny), dim3(256), 0, c->st, a);
  hipLaunchKernelGGL(k_inject_fields, dim3((c->g.fh + 255) / 256), dim3(256), 0, c->st, c->g, c->uf, by, bz, v0);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// diagnostics
// ---------------------------------------------------------------------------
constexpr int RB = 256;   // reduction blocks
__device__ inline double bsum(double v, double* sh) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return t;
}

// out[blk*3 + {0,1,2}] = partial {kinetic, E^2, B^2}
template <bool MIXED>
__global__ __launch_bounds__(256) void k_energy(Grid2D g, Parts A, PartsM M, const int* rowbase, const int* np2,
                                                const int* ap, const double* uf, double* out) {
  __shared__ double sh[4];
  double kin = 0.0, e2 = 0.0, b2 = 0.0;
  const int nrow = g.nsp * g.ny;
  for (int rw = blockIdx.x; rw < nrow; rw += gridDim.x) {
    const int s = rw / g.ny;
    const int base = rowbase[rw], n = np2[rw];
    for (int k = threadIdx.x; k < n; k += 256) {
      const int p = ap ? ap[base + k] : base + k;
      const double ux = MIXED ? (double)M.ux[p] : A.ux[p];
      const double uy = MIXED ? (double)M.uy[p] : A.uy[p];
      const double uz = MIXED ? (double)M.uz[p] : A.uz[p];
      const double u2 = ux * ux + uy * uy + uz * uz;
      const double gam = sqrt(1 + u2 / (g.c * g.c));
      kin = kin + g.r[s] * (gam - 1);
    }
  }
  const size_t pl = (size_t)g.fw * g.fh;
  const int ncell = g.nx * g.ny;
  for (int k = blockIdx.x * 256 + threadIdx.x; k < ncell; k += gridDim.x * 256) {
    const int f = fidx(g, g.nxgs + k % g.nx, g.nys + k / g.nx);
    b2 = b2 + uf[f] * uf[f] + uf[pl + f] * uf[pl + f] + uf[2 * pl + f] * uf[2 * pl + f];
    e2 = e2 + uf[3 * pl + f] * uf[3 * pl + f] + uf[4 * pl + f] * uf[4 * pl + f] + uf[5 * pl + f] * uf[5 * pl + f];
  }
  kin = bsum(kin, sh);
  e2 = bsum(e2, sh);
  b2 = bsum(b2, sh);
  if (threadIdx.x == 0) {
    out[blockIdx.x * 3 + 0] = kin;
    out[blockIdx.x * 3 + 1] = e2;
    out[blockIdx.x * 3 + 2] = b2;
  }
}

hipError_t launch_energy(wpic_ctx* c, double* dout) {
  hipLaunchKernelGGL(k_energy<false>, dim3(RB), dim3(256), 0, c->st, c->g, c->A, c->MA, c->rowbaseA, c->np2A,
                     c->aperm, c->uf, dout);
  return hipGetLastError();
}

hipError_t launch_energy_mixed(wpic_ctx* c, double* dout) {
  hipLaunchKernelGGL(k_energy<true>, dim3(RB), dim3(256), 0, c->st, c->g, c->A, c->MA, c->rowbaseA, c->np2A,
                     c->aperm, c->uf, dout);
  return hipGetLastError();
}

// rho with the quadratic shape at cell centres, then G = div E - 4 pi rho
template <bool MIXED>
__global__ __launch_bounds__(256) void k_rho(Grid2D g, Parts A, PartsM M, const int* rowbase, const int* np2,
                                             const int* ap, double* rho) {
  const int rw = blockIdx.x;
  const int s = rw / g.ny;
  const int base = rowbase[rw], n = np2[rw];
  const int gw = g.fw;
  const int jrow = g.nys + rw % g.ny;
  for (int k = threadIdx.x; k < n; k += 256) {
    const int p = ap ? ap[base + k] : base + k;
    const double x = MIXED ? (double)M.ix[p] + (double)M.ox[p] : A.x[p];
    const double y = MIXED ? (double)jrow + (double)M.oy[p] : A.y[p];
    const int i0 = (int)(x * g.d_delx), j0 = (int)(y * g.d_delx);
    if (i0 < g.nxgs || i0 > g.nxge || j0 < g.nys || j0 > g.nye) continue;
    double sx[3], sy[3];
    shape3(x * g.d_delx - 0.5 - i0, sx[0], sx[1], sx[2]);
    shape3(y * g.d_delx - 0.5 - j0, sy[0], sy[1], sy[2]);
    for (int b = 0; b < 3; b++)
      for (int a = 0; a < 3; a++)
        unsafeAtomicAdd(&rho[(i0 - 1 + a - (g.nxgs - 2)) + gw * (j0 - 1 + b - (g.nys - 2))],
                        g.q[s] * sx[a] * sy[b]);
  }
}

__global__ void k_rho_fold_y(Grid2D g, double* rho) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int I = g.nxgs - 1 + t;
  if (I > g.nxge + 1) return;
  rho[fidx(g, I, g.nye)] += rho[fidx(g, I, g.nys - 1)];
  rho[fidx(g, I, g.nys)] += rho[fidx(g, I, g.nye + 1)];
}

__global__ void k_rho_fold_x(Grid2D g, double* rho) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int J = g.nys + t;
  if (J > g.nye) return;
  rho[fidx(g, g.nxge, J)] += rho[fidx(g, g.nxgs - 1, J)];
  rho[fidx(g, g.nxgs, J)] += rho[fidx(g, g.nxge + 1, J)];
}

__global__ void k_gauss(Grid2D g, const double* uf, const double* rho, double* G) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= g.nx * g.ny) return;
  const int i = g.nxgs + k % g.nx, j = g.nys + k / g.nx;
  const size_t pl = (size_t)g.fw * g.fh;
  const double* ex = uf + 3 * pl;
  const double* ey = uf + 4 * pl;
  const double dive = (ex[fidx(g, i + 1, j)] - ex[fidx(g, i, j)]) * g.d_delx +
                      (ey[fidx(g, i, j + 1)] - ey[fidx(g, i, j)]) * g.d_delx;
  G[k] = dive - 4.0 * 3.141592653589793 * rho[fidx(g, i, j)];
}

hipError_t launch_gauss(wpic_ctx* c, double* G, double* rho) {
  const size_t n = (size_t)c->g.fw * c->g.fh;
  hipError_t e = launch_zero(c, rho, n);
  if (e != hipSuccess) return e;
  if (c->mixed)
    hipLaunchKernelGGL(k_rho<true>, dim3(c->g.nsp * c->g.ny), dim3(256), 0, c->st, c->g, c->A, c->MA, c->rowbaseA,
                       c->np2A, c->aperm, rho);
  else
    hipLaunchKernelGGL(k_rho<false>, dim3(c->g.nsp * c->g.ny), dim3(256), 0, c->st, c->g, c->A, c->MA, c->rowbaseA,
                       c->np2A, c->aperm, rho);
  // y is periodic under every boundary kind; x only in the periodic box
  hipLaunchKernelGGL(k_rho_fold_y, dim3((c->g.nx + 2 + 255) / 256), dim3(256), 0, c->st, c->g, rho);
  if (c->g.bc == WPIC_BC_PERIODIC)
    hipLaunchKernelGGL(k_rho_fold_x, dim3((c->g.ny + 255) / 256), dim3(256), 0, c->st, c->g, rho);
  hipLaunchKernelGGL(k_gauss, dim3((c->g.nx * c->g.ny + 255) / 256), dim3(256), 0, c->st, c->g, c->uf, rho, G);
  return hipGetLastError();
}

}  // namespace wpic
