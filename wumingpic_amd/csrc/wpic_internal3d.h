// wpic_internal3d.h -- grid, context and launchers of the 3D step loop
// (periodic in x, y and z at one rank).  Only the library's own translation
// units include this; the public boundary is include/wpic3d.h.
#pragma once

#include "../../include/wpic3d.h"
#include "wpic_internal.h"

namespace wpic {

// Field arrays carry the reference's two ghost layers in every direction:
// i in nxgs-2..nxge+2, j in nys-2..nye+2, k in nzs-2..nze+2, x fastest.
struct Grid3D {
  int nx, ny, nz, nxgs, nxge, nygs, nyge, nzgs, nzge, nys, nye, nzs, nze;
  int fw, fh, fd;  // nx+4, ny+4, nz+4
  int nsp, ncol;   // ncol = ny*nz (j, k) columns per species
  double delx, delt, c, gfac, d_delx, d_delt;
  double f1, f2, f3, f4, f5, pi4dt;  // 3d/common/field.f90:57-61; 4*pi*delt
  double spanx, spany, spanz;        // (nxge-nxgs+1)*delx etc., round to nearest
  double q[WPIC_MAXSP], r[WPIC_MAXSP];
  long long ncap;
  double cg_err;  // CG relative tolerance (3d/common/field.f90:430: 1e-6)
};

__host__ __device__ inline size_t fidx3(const Grid3D& g, int i, int j, int k) {
  return (size_t)(i - (g.nxgs - 2)) + (size_t)g.fw * ((size_t)(j - (g.nys - 2)) + (size_t)g.fh * (k - (g.nzs - 2)));
}
__host__ __device__ inline size_t cidx3(const Grid3D& g, int i, int j, int k) {
  return (size_t)(i - g.nxgs) + (size_t)g.nx * ((size_t)(j - g.nys) + (size_t)g.ny * (k - g.nzs));
}
__host__ __device__ inline size_t plane3(const Grid3D& g) { return (size_t)g.fw * g.fh * g.fd; }

// a particle buffer: the record and its id travel together
struct Parts3 {
  double *x, *y, *z, *ux, *uy, *uz;
  uint64_t* id;
};

// mixed-precision particles: x cell index, FP32 in-cell offsets (x - ix,
// y - j, z - k, all in [0, 1)) and FP32 momenta; 36 B per particle with the id
struct PartsM3 {
  int* ix;
  float *ox, *oy, *oz, *ux, *uy, *uz;
  uint64_t* id;
};

// Exchange direction of a particle leaving column (j, k) for (j+dj, k+dk):
// dir = (dk+1)*3 + (dj+1); DIR_STAY = 4.  The exchange code of a particle is
// (new x cell - nxgs) << 4 | dir.
constexpr int DIR_STAY = 4;
constexpr int CODE_SHIFT3 = 4;

// The boundaries a pushed FP64 particle of column (j, k) meets before the
// sort, and its exchange code: (WRAPX) boundary_periodic__particle_x
// (3d/common/boundary_periodic.f90:80-95, ipos = int(x*d_delx)), then
// particle_yz's y and z wraps (:147-165) and the destination.  A step beyond
// one column or an x cell outside the box flags ERR_DOMAIN and is clamped.
// Returns the code; `moved` says whether y or z left the column.
template <bool WRAPX>
__device__ inline int wrap3_code(const Grid3D& g, double& x, double& y, double& z, int j, int k, int* err,
                                 bool& moved) {
  if (WRAPX) {
    const int ipos = (int)(x * g.d_delx);
    if (ipos < g.nxgs)
      x = x + g.spanx;
    else if (ipos >= g.nxge + 1)
      x = x - g.spanx;
  }
  const int jpos = (int)(y * g.d_delx), kpos = (int)(z * g.d_delx);
  int dj = jpos - j, dk = kpos - k;
  moved = !(jpos == j && kpos == k);
  if (moved) {
    if (jpos <= g.nygs - 1)
      y = y + g.spany;
    else if (jpos >= g.nyge + 1)
      y = y - g.spany;
    if (kpos <= g.nzgs - 1)
      z = z + g.spanz;
    else if (kpos >= g.nzge + 1)
      z = z - g.spanz;
  }
  int ni = (int)x;
  if (ni < g.nxgs || ni > g.nxge || dj < -1 || dj > 1 || dk < -1 || dk > 1) {
    atomicOr(err, ERR_DOMAIN);
    ni = ni < g.nxgs ? g.nxgs : (ni > g.nxge ? g.nxge : ni);
    dj = dj < -1 ? -1 : (dj > 1 ? 1 : dj);
    dk = dk < -1 ? -1 : (dk > 1 ? 1 : dk);
  }
  return ((ni - g.nxgs) << CODE_SHIFT3) | ((dk + 1) * 3 + (dj + 1));
}
// CG partial-sum blocks in 3D: enough blocks to cover 256^3 cells with
// few cells per thread; the partials are summed in a fixed order
constexpr int NPB3 = 2048;

}  // namespace wpic

struct wpic_ctx3d {
  wpic_params3d P;
  wpic::Grid3D g;
  int mode;
  hipStream_t st;
  // the step's exchange and sort run on st2 while the field solve runs on st
  // (they share no buffer); ev_push / ev_sort order them around the push
  hipStream_t st2;
  hipEvent_t ev_push, ev_sort;
  int device;

  // particles: species s occupies [s*ncap, (s+1)*ncap); columns compact within.
  // A holds the sorted particles (the reference's up) as aperm[k] -> slot of
  // A's buffer (null: identity); B (gp) shares A's slots after the push, which
  // moves each record and its id.  The sort writes the permutation and swaps
  // the buffers (ids included).
  wpic::Parts3 A, B;
  bool mixed;  // WPIC_MODE_MIXED context: MA/MB hold the particles
  wpic::PartsM3 MA, MB;
  int* aperm;    // A's slot -> record in A's buffer, or null
  int* permbuf;  // [N] the sort's permutation
  int* nov;      // [1] overflow count of the fused push+deposit
  uint64_t* id[2];  // the two id buffers (A's and B's)
  int *code, *R, *holes, *mlist;  // [N]
  int* cols[3];                   // column structures: base [nsp*ncol] then np2 [nsp*ncol]
  int ca, cb, cn;                 // which column structure belongs to A, B, the exchange
  int *colbaseA, *np2A, *cumcnt;  // cumcnt [nsp*ncol*(nx+1)]
  int *colbaseB, *np2B;
  int *colbaseN, *np2N;
  int* dcnt;    // [nsp*ncol][9] movers per direction (DIR_STAY: stayers)
  int* dstart;  // [nsp*ncol][9] exclusive start of each direction in mlist
  int* colkm;   // [nsp*ncol][2] leavers, arrivals
  int* csum;    // [nsp * ceil(ncol / 1024)] column-count workgroup totals
  int* seg;     // [ncol][8] arrival segments (source column << 4 | dir), in order
  int* nseg;    // [ncol]
  bool b_valid, b_aligned, exch_pending;
  bool b_coded;  // gp holds its exchange codes and its x, y, z wraps (the fused push wrote them)
  int pusher;   // WPIC_PUSHER_BORIS or WPIC_PUSHER_VAY
  bool b_accl;  // B is mom_calc__accl's gp: A's positions, A's cell order
  std::vector<int> h_cols[3];
  bool h_cols_ok[3];

  double *uf, *df, *uj;  // 6, 6, 3 components of (nx+4)(ny+4)(nz+4)
  double* pjbuf;         // parity deposit blocks [nx*ny*nz][300] (lazy)

  double *phi, *pa, *pb, *rr, *bb, *ap;  // CG vectors [3][nx*ny*nz]
  double* part;
  wpic::CgScal* scal;
  int* err;
  int cg_iters[3];

  double* staging;
  size_t staging_n;
  long long* aosbase;  // [nsp*ncol]
  double* gbuf;
  double* red;
  double* mom;  // mom(7, nxgs-1:nxge+1, nys-1:nye+1, nzs-1:nze+1, nsp) (lazy)

  bool timing;
  std::vector<wpic::TimerRec> timers;
  std::vector<std::pair<std::string, std::pair<double, long long>>> tsum;
};

namespace wpic {

int set_error(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

// moments (MOM_* deposit kinds of wpic_internal.h)
size_t mom3_size(const Grid3D& g);
hipError_t launch_mom_accl3d(wpic_ctx3d* c);
hipError_t launch_mom_deposit3d(wpic_ctx3d* c, int kind);
hipError_t launch_mom_bc3d(wpic_ctx3d* c);
// push / deposit
hipError_t launch_push3d(wpic_ctx3d* c);
hipError_t launch_deposit3d(wpic_ctx3d* c);
hipError_t launch_push_deposit3d(wpic_ctx3d* c);
hipError_t launch_push_deposit_mixed3d(wpic_ctx3d* c);
hipError_t launch_gather_pj3d(wpic_ctx3d* c);
// fields
hipError_t launch_fill3d(wpic_ctx3d* c, double* p, size_t n, double v);
hipError_t launch_curre_bc3d(wpic_ctx3d* c);
hipError_t launch_dfield_bc3d(wpic_ctx3d* c);
hipError_t launch_cg_init3d(wpic_ctx3d* c);
hipError_t launch_cg_iter3d(wpic_ctx3d* c, int it);
hipError_t launch_cg_store3d(wpic_ctx3d* c);
hipError_t launch_e_update3d(wpic_ctx3d* c);
hipError_t launch_uf_add3d(wpic_ctx3d* c);
hipError_t launch_energy3d(wpic_ctx3d* c, double* dout, int sp = -1);
hipError_t launch_ic_weibel3d(wpic_ctx3d* c, unsigned long long seed, int n0, const double* sd, double t_ani);
hipError_t launch_gauss3d(wpic_ctx3d* c, double* G, double* rho);
// particles
hipError_t launch_wrapx3d(wpic_ctx3d* c);
hipError_t launch_colscan3d(wpic_ctx3d* c, int wrap = 1);
hipError_t launch_colcount3d(wpic_ctx3d* c);
hipError_t launch_colbuild3d(wpic_ctx3d* c);
hipError_t launch_colsort3d(wpic_ctx3d* c);
hipError_t launch_colsort3d_fast(wpic_ctx3d* c);
hipError_t launch_codes_identity3d(wpic_ctx3d* c);
// transfers and initial state
hipError_t launch_up_to_soa3d(wpic_ctx3d* c, const double* aos, long long n, int which);
hipError_t launch_soa_to_aos3d(wpic_ctx3d* c, double* aos, int which);
hipError_t launch_planes_from_aos3d(wpic_ctx3d* c, const double* aos, double* planes, int ncomp);
hipError_t launch_planes_to_aos3d(wpic_ctx3d* c, const double* planes, double* aos, int ncomp);
hipError_t launch_ic_beams3d(wpic_ctx3d* c, unsigned long long seed, int n_ppc, double udrift, double uth_e,
                             double uth_i);
hipError_t launch_up_to_mixed3d(wpic_ctx3d* c, const double* aos);
hipError_t launch_mixed_to_aos3d(wpic_ctx3d* c, double* aos);

}  // namespace wpic
