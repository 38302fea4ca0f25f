// wpic_internal.h -- device helpers, the context and the kernel launchers of
// the 2D MI355X step loop.  Only the library's own translation units include
// this; the public boundary is include/wpic.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/wpic.h"

namespace wpic {

// ---------------------------------------------------------------------------
// Grid and physics constants, passed by value to every kernel.
// Field planes carry the reference's two ghost layers:
// i in nxgs-2..nxge+2, j in nys-2..nye+2, x fastest (coalesced along x).
// ---------------------------------------------------------------------------
struct Grid2D {
  int nx, ny, nxgs, nxge, nygs, nyge, nys, nye;
  int nxa;     // the active right edge node nxe (= nxge; below it while the shock's box expands)
  int fw, fh;  // plane width nx+4 and height ny+4
  int bc, nsp;
  double delx, delt, c, gfac, d_delx, d_delt;
  double f1, f2, f3, f4, f5, pi4dt;  // field.f90:53-57; 4*pi*delt
  double spanx, spany;               // nx*delx, ny*delx rounded down
  double wall_lo, wall_hi;           // 2*(nxs+1)*delx, 2*(nxa-1)*delx
  double u0, xend;                   // WPIC_BC_SHOCK: injection 4-velocity, moving mirror position
  double q[WPIC_MAXSP], r[WPIC_MAXSP];
  long long ncap;                    // particle capacity per species
  double cg_err;                     // CG relative tolerance (field.f90:339: 1e-6)
};

// reflecting x walls: the reconnection's pair, or the shock's left wall and
// injection edge (their current, moment and y boundaries are the same)
__host__ __device__ inline bool walls(const Grid2D& g) { return g.bc != WPIC_BC_PERIODIC; }

__host__ __device__ inline int fidx(const Grid2D& g, int i, int j) {
  return (i - (g.nxgs - 2)) + g.fw * (j - (g.nys - 2));
}
// interior cell index (CG vectors, parity deposit blocks)
__host__ __device__ inline int cidx(const Grid2D& g, int i, int j) {
  return (i - g.nxgs) + g.nx * (j - g.nys);
}

// moments mom(7, nxgs-1:nxge+1, nys-1:nye+1, nsp): the seven moments of a
// node contiguous, as the reference's array
__host__ __device__ inline size_t midx(const Grid2D& g, int i, int j, int s) {
  return (size_t)7 * ((size_t)(i - (g.nxgs - 1)) + (size_t)(g.nx + 2) * ((size_t)(j - (g.nys - 1)) + (size_t)(g.ny + 2) * s));
}

// a particle buffer: the record and the reference's id (raw 64 bits) travel
// together
struct Parts {
  double *x, *y, *ux, *uy, *uz;
  uint64_t* id;
};

// mixed-precision particles: x cell index, FP32 in-cell offsets (x - ix and
// y - row, both in [0, 1)) and FP32 momenta; 32 B per particle with the id
struct PartsM {
  int* ix;
  float *ox, *oy, *ux, *uy, *uz;
  uint64_t* id;
};

// an in-cell offset in [0, 1) as FP32: rounding up to 1 is pulled back below
__host__ __device__ inline float to_offset(double f) {
  float o = (float)f;
  if (o >= 1.0f) o = 0x1.fffffep-1f;
  if (o < 0.0f) o = 0.0f;
  return o;
}

// device error bits, OR-ed by kernels and checked by the host at sync points
enum : int { ERR_DOMAIN = 1, ERR_CG = 2, ERR_ROW = 4 };

// ---------------------------------------------------------------------------
// Round-toward-minus-infinity double arithmetic (the periodic particle wraps
// run under ieee_down: boundary_periodic.f90:74,124).  Correctly rounded-down
// results from round-to-nearest operations plus an exact error term.
// ---------------------------------------------------------------------------
__host__ __device__ inline double next_down(double s) {
  long long b;
  __builtin_memcpy(&b, &s, 8);
  if (s > 0.0)
    b -= 1;
  else
    b += 1;
  double r;
  __builtin_memcpy(&r, &b, 8);
  return r;
}

__host__ __device__ inline double rd_add(double a, double b) {
  double s = a + b;
  if (s == 0.0) {
    if (a == 0.0 && b == 0.0 && !__builtin_signbit(a) && !__builtin_signbit(b)) return 0.0;
    return -0.0;
  }
  double bb = s - a;
  double err = (a - (s - bb)) + (b - bb);
  if (err < 0.0) s = next_down(s);
  return s;
}

__host__ __device__ inline double rd_mul(double a, double b) {
  double p = a * b;
  double err = __builtin_fma(a, b, -p);
  if (err < 0.0) p = next_down(p);
  return p;
}

__host__ __device__ inline double rd_div(double a, double b) {
  double q = a / b;
  double r = __builtin_fma(-q, b, a);
  if ((r < 0.0 && b > 0.0) || (r > 0.0 && b < 0.0)) q = next_down(q);
  return q;
}

// int(x / delx) of the particle boundaries, under round-down
// (boundary_periodic.f90:74) or round-to-nearest: with delx = 1, the only
// cell width a context accepts (sort__bucket bins by int(x)), the quotient is
// x itself in every rounding mode
__host__ __device__ inline int cell_rd(double x, double delx) { return (int)(delx == 1.0 ? x : rd_div(x, delx)); }
__host__ __device__ inline int cell_rn(double x, double delx) { return (int)(delx == 1.0 ? x : x / delx); }

// The x boundary that acts on a pushed particle before the deposit:
// reconnection walls (boundary_reconnection.f90:79-91, WPIC_BC_WALLX) or the
// shock's left wall and the mirror moving with the inflow at xend, which sends
// ux to 2 u0 - ux (boundary_shock__injection, boundary_shock.f90:270-293,
// WPIC_BC_SHOCK).  INJ false gives boundary_shock__particle_x's wall pair.
template <int BC, bool INJ = true>
__host__ __device__ inline void reflect_x(const Grid2D& g, double& x, double& ux, double& uy, double& uz) {
  if (BC == WPIC_BC_PERIODIC) return;
  const int ipos = cell_rn(x, g.delx);
  if (ipos < g.nxgs + 1) {
    x = g.wall_lo - x;
    ux = -ux; uy = -uy; uz = -uz;
  } else if (BC == WPIC_BC_SHOCK && INJ) {
    if (x > g.xend) {
      x = +2. * g.xend - x;
      ux = +2. * g.u0 - ux;
      uy = -uy; uz = -uz;
    }
  } else if (ipos >= g.nxa - 1) {
    x = g.wall_hi - x;
    ux = -ux; uy = -uy; uz = -uz;
  }
}


// quadratic spline with nodes at cell centres (particle.f90:99-107)
__host__ __device__ inline void shape3(double dh, double& s0, double& s1, double& s2) {
  s0 = 0.5 * (0.5 - dh) * (0.5 - dh);
  s1 = 0.75 - dh * dh;
  s2 = 0.5 * (0.5 + dh) * (0.5 + dh);
}

// The particles of one cell over all species, as one list: element t is slot
// beg[s] + t - cum[s] of species s, the last s with cum[s] <= t (species
// beyond nsp carry cum = INT_MAX).  Kept as named scalars and resolved with
// masks: a chain of selects over an array is turned by the compiler into an
// indexed table in scratch memory.
struct CellList {
  int o0, d1, d2, d3, c1, c2, c3;  // slot offset of species 0, its steps to 1..3, their first elements
  int ob;                          // the offset of the species set last (building d1..d3)
  __device__ void set(int s, int beg, int cum) {
    const int o = beg - cum;
    if (s == 0) { o0 = o; }
    if (s == 1) { d1 = o - ob; c1 = cum; }
    if (s == 2) { d2 = o - ob; c2 = cum; }
    if (s == 3) { d3 = o - ob; c3 = cum; }
    ob = o;
  }
  __device__ void slot(int t, int& s, int& p) const {
    static_assert(WPIC_MAXSP == 4, "CellList holds four species");
    const int m1 = -(int)(t >= c1), m2 = -(int)(t >= c2), m3 = -(int)(t >= c3);
    s = -(m1 + m2 + m3);
    p = t + o0 + (m1 & d1) + (m2 & d2) + (m3 & d3);
  }
};

// Sum over the 16 lanes of a DPP row (lanes 16r .. 16r+15), every lane of the
// row receiving it: quad permutes for the partners 1 and 2 apart, then row
// rotations by 4 and 8.  Cross-lane moves on the VALU (DPP) instead of
// ds_bpermute through the LDS crossbar.  Each row must be wholly active or
// wholly inactive (rows never exchange data).
template <int CTRL>
__device__ inline float dpp_f32(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ inline double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
constexpr int DPP_QUAD_XOR1 = 0xB1;  // quad_perm [1, 0, 3, 2]
constexpr int DPP_QUAD_XOR2 = 0x4E;  // quad_perm [2, 3, 0, 1]
constexpr int DPP_ROW_ROR4 = 0x124;
constexpr int DPP_ROW_ROR8 = 0x128;
constexpr int DPP_ROW_BCAST15 = 0x142;  // lane 15 of each row to the next row's lanes
template <class T>
__device__ inline T row16_sum(T v) {
  if constexpr (sizeof(T) == 8) {
    v += dpp_f64<DPP_QUAD_XOR1>(v);
    v += dpp_f64<DPP_QUAD_XOR2>(v);
    v += dpp_f64<DPP_ROW_ROR4>(v);
    v += dpp_f64<DPP_ROW_ROR8>(v);
  } else {
    v += dpp_f32<DPP_QUAD_XOR1>(v);
    v += dpp_f32<DPP_QUAD_XOR2>(v);
    v += dpp_f32<DPP_ROW_ROR4>(v);
    v += dpp_f32<DPP_ROW_ROR8>(v);
  }
  return v;
}

// Sum over the 8 lanes of a half DPP row, every lane of the half receiving
// it: quad permutes for the partners 1 and 2 apart, then the half-row mirror
// (lane l and 7 - l), which pairs each quad with the other.
constexpr int DPP_ROW_HALF_MIRROR = 0x141;
template <class T>
__device__ inline T row8_sum(T v) {
  if constexpr (sizeof(T) == 8) {
    v += dpp_f64<DPP_QUAD_XOR1>(v);
    v += dpp_f64<DPP_QUAD_XOR2>(v);
    v += dpp_f64<DPP_ROW_HALF_MIRROR>(v);
  } else {
    v += dpp_f32<DPP_QUAD_XOR1>(v);
    v += dpp_f32<DPP_QUAD_XOR2>(v);
    v += dpp_f32<DPP_ROW_HALF_MIRROR>(v);
  }
  return v;
}

// The sort's per-cell counts kept by a fused push: every cell of the tile
// keeps its particles (cumcnt, added when the tile's counts are flushed)
// except those that leave it, each of which moves one count from its own cell
// to the cell and row of its exchange code ((new x cell - nxgs) << 2 | row
// step + 1): in the tile's LDS counts (one-cell halo), or straight to rcnt
// when the destination lies beyond (a wrap across the box).  rcnt is
// [nsp*ny][nx].
template <int TXC, int TYC>
__device__ inline void count_leaver(const Grid2D& g, int (*hcn)[TYC + 2][TXC + 2], int* rcnt, int i0, int j0, int li,
                                    int lj, int s, int code) {
  atomicAdd(&hcn[s][lj + 1][li + 1], -1);
  const int cn = (code >> 2) + g.nxgs, dy = (code & 3) - 1;
  const int hx = cn - i0 + 1;
  if (hx >= 0 && hx <= TXC + 1) {
    atomicAdd(&hcn[s][lj + 1 + dy][hx], 1);
  } else {
    int jn = j0 + lj + dy;
    if (jn < g.nys) jn += g.ny;
    else if (jn > g.nye) jn -= g.ny;
    atomicAdd(&rcnt[((size_t)s * g.ny + (jn - g.nys)) * g.nx + (cn - g.nxgs)], 1);
  }
}
// the tile's counts, with its cells' own particles, added to rcnt (rows wrap
// periodically in y)
template <int TXC, int TYC>
__device__ inline void flush_counts(const Grid2D& g, int (*hcn)[TYC + 2][TXC + 2], int* rcnt, const int* cumcnt,
                                    int i0, int j0) {
  constexpr int N = (TYC + 2) * (TXC + 2);
  for (int k = threadIdx.x; k < g.nsp * N; k += blockDim.x) {
    const int s = k / N, rem = k - s * N;
    const int hr = rem / (TXC + 2), hx = rem % (TXC + 2);
    int v = hcn[s][hr][hx];
    const int cx = i0 - 1 + hx;
    int jn = j0 - 1 + hr;
    if (hr >= 1 && hr <= TYC && hx >= 1 && hx <= TXC && cx <= g.nxge && jn <= g.nye) {
      const int* cr = cumcnt + ((size_t)s * g.ny + (jn - g.nys)) * (g.nx + 1) + (cx - g.nxgs);
      v += cr[1] - cr[0];
    }
    if (v == 0) continue;
    if (jn < g.nys) jn += g.ny;
    else if (jn > g.nye) jn -= g.ny;
    atomicAdd(&rcnt[((size_t)s * g.ny + (jn - g.nys)) * g.nx + (cx - g.nxgs)], v);
  }
}

// Slots in an LDS queue for the active lanes of a wave with one atomic on the
// counter *n: the active lanes get consecutive slots in lane order.
__device__ inline int wave_enqueue(int* n) {
  const uint64_t m = __ballot(1);
  const int lane = __lane_id();
  const int leader = __ffsll((unsigned long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(n, __popcll(m));
  base = __shfl(base, leader);
  return base + __popcll(m & ((1ull << lane) - 1));
}

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous band of logical tiles so neighbouring tiles (which share halo
// atomics) meet in the same L2.  Launch 8*ceil(n/8) groups; extra ones exit.
__device__ inline int xcd_tile(int bid, int ntiles) {
  const int per = (ntiles + 7) >> 3;
  return (bid & 7) * per + (bid >> 3);
}
inline int xcd_grid(int ntiles) { return ((ntiles + 7) >> 3) << 3; }

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
struct CgScal {
  double sum_g, eps, sumr_g, sum2_g, av, bv;
  int ite, active;
};

struct TimerRec {
  std::string name;
  hipEvent_t a, b;
};

}  // namespace wpic

struct wpic_ctx {
  wpic_params2d P;
  wpic::Grid2D g;
  int mode;
  hipStream_t st;
  // the step's exchange and sort run on st2 while the field solve runs on st
  // (they share no buffer); ev_push / ev_sort order them around the push
  hipStream_t st2;
  hipEvent_t ev_push, ev_sort;
  int device;

  // particles: species s occupies [s*ncap, (s+1)*ncap); rows compact within
  // A holds the sorted particles (the reference's up) as aperm[k] -> slot of
  // A's buffer (null: identity); B holds gp, slot-aligned with A after the
  // push, which moves each record and its id.  The sort writes only the
  // permutation, then A and B swap buffers (ids included): the records are
  // read through aperm by the next push.
  wpic::Parts A, B;
  bool mixed;            // WPIC_MODE_MIXED context: MA/MB hold the particles
  wpic::PartsM MA, MB;
  int* aperm;            // A's slot -> record in A's buffer, or null
  int* permbuf;          // [N] the sort's permutation
  uint64_t* id[2];  // the two id buffers (A's and B's)
  int *code, *R, *holes, *ulist, *dlist;
  int* rows[3];                   // row structures: rowbase [nsp*ny] then np2 [nsp*ny]
  int ra, rb, rn;                 // which row structure belongs to A, B, the exchange
  int *rowbaseA, *np2A, *cumcnt;  // [nsp*ny], [nsp*ny], [nsp*ny*(nx+1)]
  int *rowbaseB, *np2B;           // rows of B (gp)
  int *rowbaseN, *np2N;           // rows after the y exchange
  int* rowcnt;                    // [nsp*ny][4] leavers, up, down, incoming
  // [nsp*ny][nx] particles per (species, row, x cell) after the step's push:
  // the mixed push counts them (rcnt_fill) and the fast sort takes the rows'
  // cell counts from them instead of counting the codes again
  int* rcnt;
  bool rcnt_fill;
  bool b_valid;       // B holds a gp array
  bool b_aligned;     // B's slots are A's slots (after particle__solv)
  int pusher;         // WPIC_PUSHER_BORIS or WPIC_PUSHER_VAY
  bool b_accl;        // B is mom_calc__accl's gp: A's positions, A's cell order
  bool exch_pending;  // R / N rows describe B after bc__particle_y
  std::vector<int> h_rows[3];  // host mirrors of the row structures
  bool h_rows_ok[3];

  // fields: planes of (nx+4)*(ny+4)
  double *uf, *df, *uj;
  double* pjbuf;  // parity-mode per-cell current blocks [nx*ny][65]

  // CG: interior vectors per component
  double *phi, *pa, *pb, *rr, *bb, *ap;  // [3][nx*ny]
  double* part;                          // block partial sums [3][NPB][2]
  wpic::CgScal* scal;                    // [3]
  int* err;                              // device error bits
  int* hflags;                           // pinned host mirror: err, active[3], ite[3]
  int cg_iters[3];
  int cg_hint;  // the largest iteration count of the last solve (the first batch's size)

  double* staging;  // transfer staging (lazy)
  size_t staging_n;
  long long* aosbase;  // [nsp*ny] row starts in a compact AoS transfer buffer
  double* gbuf;        // Gauss residual / rho scratch (lazy)
  double* mom;         // moments mom(7, nxgs-1:nxge+1, nys-1:nye+1, nsp) (lazy)
  double* red;      // diagnostics reduction scratch
  size_t red_n;

  bool timing;
  std::vector<wpic::TimerRec> timers;
  std::vector<std::pair<std::string, std::pair<double, long long>>> tsum;

  int* perm_out;          // when set, the sort records the gp slot of every sorted slot
  struct AdjState* adj;   // recorded run and adjoint arrays (wpic_adj.h), lazy
};

namespace wpic {

constexpr int NPB = 256;  // fixed number of partial-sum blocks in CG reductions

int set_error(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

// launchers (each enqueues on ctx->st) -------------------------------------
// push / deposit
hipError_t launch_push_deposit(wpic_ctx* c, bool push, bool deposit, bool fused);
hipError_t launch_gather_pj(wpic_ctx* c);
// fields
hipError_t launch_zero(wpic_ctx* c, double* p, size_t n);
hipError_t launch_curre_bc(wpic_ctx* c);
hipError_t launch_dfield_bc(wpic_ctx* c);
hipError_t launch_cg_init(wpic_ctx* c);
hipError_t launch_cg_iter(wpic_ctx* c, int it_in_batch);
hipError_t launch_cg_store(wpic_ctx* c);
hipError_t launch_e_update(wpic_ctx* c);
hipError_t launch_uf_add(wpic_ctx* c);
hipError_t launch_energy(wpic_ctx* c, double* dout);
hipError_t launch_gauss(wpic_ctx* c, double* G, double* rho);
hipError_t launch_codes_identity(wpic_ctx* c);
// particles: boundaries, exchange, sort
hipError_t launch_wrapx(wpic_ctx* c, bool injection);
hipError_t launch_ic_shock(wpic_ctx* c, unsigned long long seed, int np2, double v0, const double* sd, double l_damp,
                           const double* b);
hipError_t launch_inject(wpic_ctx* c, unsigned long long seed, int it, const int* base_new, const int* nl,
                         const long long* idoff, double v0, const double* sd, double l_damp, double by, double bz,
                         bool reloc);
hipError_t launch_rowscan(wpic_ctx* c, bool ywrap);
hipError_t launch_rowcount(wpic_ctx* c);
hipError_t launch_rowbuild(wpic_ctx* c);
hipError_t launch_rowsort(wpic_ctx* c);
hipError_t launch_rowsort_fast(wpic_ctx* c);
// transfers and initial state
hipError_t launch_up_to_soa(wpic_ctx* c, const double* aos, long long n, int which);
hipError_t launch_soa_to_aos(wpic_ctx* c, double* aos, int which);
hipError_t launch_planes_from_aos(wpic_ctx* c, const double* aos, double* planes, int ncomp);
hipError_t launch_planes_to_aos(wpic_ctx* c, const double* planes, double* aos, int ncomp);
hipError_t launch_ic_beams(wpic_ctx* c, unsigned long long seed, int n_ppc, double udrift, double uth_e,
                           double uth_i);
hipError_t launch_fill(wpic_ctx* c, double* p, size_t n, double v);
hipError_t launch_ic_weibel(wpic_ctx* c, unsigned long long seed, int n0, const double* sd, double t_ani);
// hp: x0, y0, lcs, b0, e1, ncs, nbg, f1, f2, sd1, sd2
hipError_t launch_ic_harris(wpic_ctx* c, unsigned long long seed, int np2, int ibg, const double* hp);
hipError_t launch_gather_rows(wpic_ctx* c, const Parts& src, const int* ap, const Parts& dst, const int* rows);
// moments (mom_calc.f90, boundary_*__mom): deposit kinds
enum { MOM_TILE_KICK = 0, MOM_TILE_GP = 1, MOM_ANY = 2, MOM_PARITY_ALIGNED = 3, MOM_PARITY_ANY = 4 };
hipError_t launch_mom_accl(wpic_ctx* c);
hipError_t launch_mom_deposit(wpic_ctx* c, int kind);
hipError_t launch_mom_bc(wpic_ctx* c);
// mixed-precision form
hipError_t launch_push_deposit_mixed(wpic_ctx* c);
hipError_t launch_up_to_mixed(wpic_ctx* c, const double* aos);
hipError_t launch_mixed_to_aos(wpic_ctx* c, double* aos);
hipError_t launch_ic_beams_mixed(wpic_ctx* c, unsigned long long seed, int n_ppc, double udrift, double uth_e,
                                 double uth_i);
hipError_t launch_energy_mixed(wpic_ctx* c, double* dout);
hipError_t launch_mom_mixed(wpic_ctx* c);

// one fused step (wpic_step2d's body) and the CG iteration loop of field__fdtd_i
int step_once(wpic_ctx* c);
int cg_loop(wpic_ctx* c);
int check_err_ctx(wpic_ctx* c);
hipError_t launch_cg_start(wpic_ctx* c);
// adjoint
hipError_t launch_adj_unsort(wpic_ctx* c, const int* rows, const int* perm, const double* lamA, double* lamB,
                             size_t N);
hipError_t launch_adj_field1(wpic_ctx* c, const double* lamF, double* lamF2, double* lamJ);
hipError_t launch_adj_field2(wpic_ctx* c, double* lamF2, double* lamJ);
hipError_t launch_adj_particles(wpic_ctx* c, const Parts& P, const int* rows, const double* uf, const double* lamB,
                                const double* lamJ, double* lamA, double* lamT, size_t N);
hipError_t launch_adj_t2f(wpic_ctx* c, const double* lamT, double* lamF);
hipError_t launch_adj_seed(wpic_ctx* c, double* lamF, double seed);
hipError_t launch_adj_params(wpic_ctx* c, const Parts& P, const uint64_t* id, const int* rows, const double* lamA,
                             size_t N, const double* lamF, double udrift, double uth_e, int mode, int n_ppc,
                             double* out);
hipError_t launch_perturb_bz(wpic_ctx* c, double amp, int mode);
hipError_t launch_adj_lamF_to_planes(wpic_ctx* c, const double* lamF, double* planes);
void adj_free(wpic_ctx* c);

void timer_begin(wpic_ctx* c, const char* name, hipEvent_t* a);
void timer_end(wpic_ctx* c, hipEvent_t a);

}  // namespace wpic
