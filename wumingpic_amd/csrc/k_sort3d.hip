// k_sort3d.hip -- 3D particle boundaries, the y-z column exchange and the
// per-column counting sort at one rank (3d/common/boundary_periodic.f90:68-455,
// 3d/common/sort.f90:40-88).
//
// Particles of species s live in (j, k) columns, each a contiguous segment of
// the species' block.  boundary_periodic__particle_yz appends every leaver to
// the bucket of its destination column in the scan order (k, j, slot), then
// moves the ghost buckets across the periodic faces (y first, then z) by
// appending whole buckets.  The arrivals of a column are therefore a fixed
// sequence of "segments" -- all movers of one source column in one direction,
// in slot order -- and that sequence depends only on the grid, so the host
// computes it once per context (seg / nseg).  Then, in parallel:
//   k_colscan3d   per column: y/z wraps, direction, exchange code; leaver slots
//                 (holes) and the movers of each direction in slot order;
//   k_colcount3d  arrivals, new column lengths and bases (block-wise
//                 prefix sums, then the blocks' offsets);
//   k_colbuild3d  R = for each slot of the post-exchange column the particle
//                 that lands there (hole filling as :418-455 does it);
//   k_colsort3d   stable counting sort of R by x cell: for every sorted slot
//                 the gp slot its particle comes from (the next push reads
//                 the record and its id through it).
// The fused fast and mixed steps skip k_colbuild3d: k_colsort3d_fast sorts
// each column's stayers and arrivals straight from the mover lists.
#include "wpic_internal3d.h"

namespace wpic {

namespace {
constexpr int SB = 256;
__device__ inline uint64_t lt_mask(int lane) { return lane == 0 ? 0ull : (~0ull >> (64 - lane)); }
}  // namespace

struct Sort3Args {
  Grid3D g;
  Parts3 A, B;
  int* perm;
  int *code, *R, *holes, *mlist;
  int *colbaseB, *np2B, *colbaseN, *np2N, *cumcnt;
  int *dcnt, *dstart, *colkm;
  int* csum;  // [nsp * ceil(ncol / 1024)] column-count workgroup totals
  const int *seg, *nseg;
  int* err;
};

// boundary_periodic__particle_x (:80-95): ipos = int(x*d_delx), round to nearest
__global__ __launch_bounds__(SB) void k_wrapx3d(Sort3Args a) {
  const Grid3D& g = a.g;
  const int rw = blockIdx.x;
  const int base = a.colbaseB[rw], n = a.np2B[rw];
  for (int q = threadIdx.x; q < n; q += SB) {
    const int p = base + q;
    const double x = a.B.x[p];
    const int ipos = (int)(x * g.d_delx);
    if (ipos < g.nxgs)
      a.B.x[p] = x + g.spanx;
    else if (ipos >= g.nxge + 1)
      a.B.x[p] = x - g.spanx;
  }
}

// per column: the y/z wraps of particle_yz (:147-165) and the exchange code;
// pass 2 lists the holes and the movers of each direction in slot order.
// Each wave owns a contiguous quarter of the column: pass 1 counts its
// quarter's directions with ballots (SF chunks of 64 per round, their loads
// issued before the first is used), one barrier turns the four waves' counts
// into offsets, and pass 2 writes the wave's holes and movers from them, so
// the block waits at three barriers whatever the column's length.
__global__ __launch_bounds__(SB) void k_colscan3d(Sort3Args a, int wrap) {
  constexpr int SF = 4, NW = SB / 64;
  __shared__ int wt[NW][10];
  __shared__ int dst[9];
  const Grid3D& g = a.g;
  const int rw = blockIdx.x;
  const int col = rw % g.ncol;
  const int j = g.nys + col % g.ny, k = g.nzs + col / g.ny;
  const int base = a.colbaseB[rw], n = a.np2B[rw];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint64_t lt = lt_mask(lane);
  const int lo = (int)(((long long)n * w) / NW), hi = (int)(((long long)n * (w + 1)) / NW);
  // pass 1: codes and the quarter's per-direction counts (9) and leavers
  int cnt[10];
#pragma unroll
  for (int d = 0; d < 10; d++) cnt[d] = 0;
  for (int k0 = lo; k0 < hi; k0 += 64 * SF) {
    int dir[SF];
    if (wrap) {
#pragma unroll
      for (int u = 0; u < SF; u++) {
        const int q = k0 + u * 64 + lane;
        dir[u] = -1;
        if (q < hi) {
          const int p = base + q;
          double x = a.B.x[p], y = a.B.y[p], z = a.B.z[p];
          bool moved;
          const int code = wrap3_code<false>(g, x, y, z, j, k, a.err, moved);
          if (moved) {
            a.B.y[p] = y;
            a.B.z[p] = z;
          }
          a.code[p] = code;
          dir[u] = code & 15;
        }
      }
    } else {
      int v[SF];
#pragma unroll
      for (int u = 0; u < SF; u++) {
        const int q = k0 + u * 64 + lane;
        v[u] = a.code[base + (q < hi ? q : lo)];
      }
#pragma unroll
      for (int u = 0; u < SF; u++) dir[u] = k0 + u * 64 + lane < hi ? (v[u] & 15) : -1;
    }
#pragma unroll
    for (int u = 0; u < SF; u++) {
#pragma unroll
      for (int d = 0; d < 9; d++) cnt[d] += __popcll(__ballot(dir[u] == d));
      cnt[9] += __popcll(__ballot(dir[u] >= 0 && dir[u] != DIR_STAY));
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int d = 0; d < 10; d++) wt[w][d] = cnt[d];
  }
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int d = 0; d < 9; d++) {
      int t = 0;
      for (int v = 0; v < NW; v++) t += wt[v][d];
      dst[d] = acc;
      a.dcnt[rw * 9 + d] = t;
      a.dstart[rw * 9 + d] = acc;
      if (d != DIR_STAY) acc += t;
    }
    a.colkm[rw * 2] = acc;  // leavers
  }
  __syncthreads();
  // this wave's running offsets: the earlier waves' counts, then its own
  int oh = 0, od[9];
#pragma unroll
  for (int d = 0; d < 9; d++) od[d] = dst[d];
  for (int v = 0; v < w; v++) {
    oh += wt[v][9];
#pragma unroll
    for (int d = 0; d < 9; d++) od[d] += wt[v][d];
  }
  // pass 2: holes (1-based slots of the leavers) and per-direction mover lists
  for (int k0 = lo; k0 < hi; k0 += 64 * SF) {
    int v[SF];
#pragma unroll
    for (int u = 0; u < SF; u++) {
      const int q = k0 + u * 64 + lane;
      v[u] = a.code[base + (q < hi ? q : lo)];
    }
#pragma unroll
    for (int u = 0; u < SF; u++) {
      const int q = k0 + u * 64 + lane;
      const int dir = q < hi ? (v[u] & 15) : -1;
      const bool lv = dir >= 0 && dir != DIR_STAY;
      const uint64_t bl = __ballot(lv);
      if (lv) a.holes[base + oh + __popcll(bl & lt)] = q + 1;
      oh += __popcll(bl);
#pragma unroll
      for (int d = 0; d < 9; d++) {
        if (d == DIR_STAY) continue;
        const uint64_t bd = __ballot(dir == d);
        if (dir == d) a.mlist[base + od[d] + __popcll(bd & lt)] = base + q;
        od[d] += __popcll(bd);
      }
    }
  }
}

// arrivals m of every column, new length L = np2 - leavers + m, and its base:
// one workgroup per 1024 columns of a species takes the exclusive prefix of L
// within its columns (written as colbaseN for now) and its total; a single
// workgroup then turns the totals into offsets, and a last pass adds them.
constexpr int CC = 1024;  // columns per counting workgroup
__global__ __launch_bounds__(CC) void k_colcount3d(Sort3Args a) {
  __shared__ int wsum[CC / 64];
  const Grid3D& g = a.g;
  const int ncol = g.ncol, tid = threadIdx.x;
  const int nbs = (ncol + CC - 1) / CC;
  const int s = blockIdx.x / nbs, r = (blockIdx.x % nbs) * CC + tid;
  int L = 0;
  if (r < ncol) {
    const int rw = s * ncol + r;
    int m = 0;
    for (int e = 0; e < a.nseg[r]; e++) {
      const int sg = a.seg[r * 8 + e];
      m += a.dcnt[(s * ncol + (sg >> 4)) * 9 + (sg & 15)];
    }
    a.colkm[rw * 2 + 1] = m;
    L = a.np2B[rw] - a.colkm[rw * 2] + m;
    a.np2N[rw] = L;
  }
  const int lane = tid & 63, w = tid >> 6;
  int v = L;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  if (lane == 63) wsum[w] = v;
  __syncthreads();
  int pre = 0;
  for (int q = 0; q < w; q++) pre += wsum[q];
  const int excl = pre + v - L;
  if (r < ncol) a.colbaseN[s * ncol + r] = excl;
  if (tid == CC - 1) a.csum[blockIdx.x] = excl + L;
}

// exclusive scan of the counting workgroups' totals, per species
__global__ void k_colcount3d_scan(Sort3Args a) {
  const Grid3D& g = a.g;
  const int nbs = (g.ncol + CC - 1) / CC;
  const int s = threadIdx.x;
  if (s >= g.nsp) return;
  int acc = 0;
  for (int b = 0; b < nbs; b++) {
    const int t = a.csum[s * nbs + b];
    a.csum[s * nbs + b] = acc;
    acc += t;
  }
}

// column bases: the species' block plus the counting workgroup's offset
__global__ __launch_bounds__(256) void k_colcount3d_base(Sort3Args a) {
  const Grid3D& g = a.g;
  const int nbs = (g.ncol + CC - 1) / CC;
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= g.nsp * g.ncol) return;
  const int s = q / g.ncol, r = q - s * g.ncol;
  a.colbaseN[q] += (int)(s * g.ncap) + a.csum[s * nbs + r / CC];
}

// R of one column: slot -> source particle (boundary_periodic.f90:418-455)
__global__ __launch_bounds__(SB) void k_colbuild3d(Sort3Args a) {
  __shared__ int wt[SB / 64][2];
  __shared__ int run[2];
  __shared__ int sh_a;
  __shared__ int scnt[8], sptr[8];
  __shared__ int nsg;
  const Grid3D& g = a.g;
  const int rw = blockIdx.x;
  const int s = rw / g.ncol, col = rw % g.ncol;
  const int base = a.colbaseB[rw], n = a.np2B[rw];
  const int kk = a.colkm[rw * 2], m = a.colkm[rw * 2 + 1];
  const int L = a.np2N[rw], nb = a.colbaseN[rw];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (tid == 0) {
    nsg = a.nseg[col];
    for (int e = 0; e < nsg; e++) {
      const int sg = a.seg[col * 8 + e];
      const int src = s * g.ncol + (sg >> 4), d = sg & 15;
      scnt[e] = a.dcnt[src * 9 + d];
      sptr[e] = a.colbaseB[src] + a.dstart[src * 9 + d];
    }
  }
  __syncthreads();
  // the t-th arrival (0-based) of this column
  auto in = [&](int t) -> int {
    for (int e = 0; e < nsg; e++) {
      if (t < scnt[e]) return a.mlist[sptr[e] + t];
      t -= scnt[e];
    }
    return -1;
  };
  const int* holes = a.holes + base;

  const int keep = n < L ? n : L;
  for (int q = tid; q < keep; q += SB)
    if ((a.code[base + q] & 15) == DIR_STAY) a.R[nb + q] = base + q;
  const int kf = kk < m ? kk : m;
  for (int t = tid; t < kf; t += SB) {
    const int h = holes[t];
    if (h >= 1 && h <= L) a.R[nb + h - 1] = in(t);
  }
  for (int t = tid; t < m - kk; t += SB) a.R[nb + n + t] = in(kk + t);
  if (m >= kk) return;

  // fewer arrivals than holes: tail particles fill the remaining holes, the
  // last one into the first free hole (loop1).  Tail = slots (L, n].
  if (tid == 0) {
    int lo = 0, hi = kk;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (holes[mid] <= L)
        lo = mid + 1;
      else
        hi = mid;
    }
    sh_a = lo;
    run[0] = 0;
    run[1] = 0;
  }
  __syncthreads();
  const int aa = sh_a;
  const int T = n - L;
  int totalq = 0;
  for (int pass = 0; pass < 2; pass++) {
    for (int c0 = 0; c0 < T; c0 += SB) {
      const int q = c0 + tid;
      bool lv = false;
      if (q < T) lv = (a.code[base + L + q] & 15) != DIR_STAY;
      const uint64_t bl = __ballot(lv);
      if (lane == 0) wt[w][0] = __popcll(bl);
      __syncthreads();
      int ol = run[0];
      for (int v = 0; v < w; v++) ol += wt[v][0];
      const int rank = aa + ol + __popcll(bl & lt_mask(lane)) + 1;
      const bool nonh = q < T && (!lv || rank <= m);
      const uint64_t bn = __ballot(nonh);
      if (lane == 0) wt[w][1] = __popcll(bn);
      __syncthreads();
      int on = run[1];
      for (int v = 0; v < w; v++) on += wt[v][1];
      if (pass == 1 && nonh) {
        const int incl = on + __popcll(bn & lt_mask(lane)) + 1;
        const int rr = totalq - incl + 1;
        const int tgt = m + rr;
        const int src = lv ? in(rank - 1) : base + L + q;
        const int h = (tgt >= 1 && tgt <= kk) ? holes[tgt - 1] : 0;
        if (h >= 1 && h <= L && src >= 0)
          a.R[nb + h - 1] = src;
        else
          atomicOr(a.err, ERR_ROW);
      }
      __syncthreads();
      if (tid == 0) {
        int t0 = 0, t1 = 0;
        for (int v = 0; v < SB / 64; v++) {
          t0 += wt[v][0];
          t1 += wt[v][1];
        }
        run[0] += t0;
        run[1] += t1;
      }
      __syncthreads();
    }
    totalq = run[1];
    __syncthreads();
    if (tid == 0) {
      run[0] = 0;
      run[1] = 0;
    }
    __syncthreads();
  }
}

__device__ inline uint64_t peers_bits(int key, bool act, int nbits) {
  uint64_t peers = __ballot(act);
  for (int b = 0; b < nbits; b++) {
    const bool bit = (key >> b) & 1;
    const uint64_t msk = __ballot(act && bit);
    peers &= bit ? msk : ~msk;
  }
  return peers;
}

// lanes whose key equals this lane's.  Slots in cell order carry keys within
// a few cells of each other: when every active key lies within [k0 - 8,
// k0 + 7] of the first active lane's k0, four ballots on key - k0 + 8 decide
// equality; otherwise all nbits.
__device__ inline uint64_t peers_of(int key, bool act, int nbits) {
  const uint64_t am = __ballot(act);
  if (am == 0) return 0;
  const int k0 = __builtin_amdgcn_readlane(key, __builtin_ctzll(am));
  const int rel = key - k0 + 8;
  if (__ballot(act && (rel < 0 || rel > 15)) != 0) return peers_bits(key, act, nbits);
  uint64_t peers = am;
#pragma unroll
  for (int b = 0; b < 4; b++) {
    const bool bit = (rel >> b) & 1;
    const uint64_t msk = __ballot(act && bit);
    peers &= bit ? msk : ~msk;
  }
  return peers;
}

// stable counting sort of one column by x cell (sort.f90:56-81), B -> A.
// Each wave owns a contiguous quarter of the column: its per-key histogram,
// then (after a scan over keys and waves) its per-key write cursors.  A round
// takes SF chunks of 64: their loads are issued before the first is used, the
// cursors move by LDS atomics (served in issue order), and the lowest lane of
// each key broadcasts the key's base by one permute.
__global__ __launch_bounds__(SB) void k_colsort3d(Sort3Args a, int nbits) {
  extern __shared__ int hist[];  // [SB/64][nx]
  __shared__ int wsum[SB / 64];
  __shared__ int carry;
  constexpr int SF = 8, NW = SB / 64;
  const Grid3D& g = a.g;
  const int nx = g.nx;
  const int rw = blockIdx.x;
  const int n = a.np2N[rw], nb = a.colbaseN[rw];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const unsigned ntot = (unsigned)(g.ncap * g.nsp);
  const uint64_t lt = lt_mask(lane);
  for (int q = tid; q < NW * nx; q += SB) hist[q] = 0;
  __syncthreads();
  const int lo = (int)(((long long)n * w) / NW), hi = (int)(((long long)n * (w + 1)) / NW);
  int* hw = hist + w * nx;
  int* cc = a.cumcnt + (size_t)rw * (nx + 1);
  bool bad = false;
  for (int pass = 0; pass < 2; pass++) {
    for (int k0 = lo; k0 < hi; k0 += 64 * SF) {
      int src[SF];
#pragma unroll
      for (int u = 0; u < SF; u++) {
        const int q = k0 + u * 64 + lane;
        src[u] = a.R[nb + (q < hi ? q : lo)];
      }
      bool act[SF];
#pragma unroll
      for (int u = 0; u < SF; u++) {
        const bool in = k0 + u * 64 + lane < hi;
        act[u] = in && (unsigned)src[u] < ntot;
        bad |= in && !act[u];
      }
      int key[SF];
#pragma unroll
      for (int u = 0; u < SF; u++) key[u] = a.code[act[u] ? src[u] : 0];
      uint64_t peers[SF];
#pragma unroll
      for (int u = 0; u < SF; u++) {
        key[u] = act[u] ? (key[u] >> CODE_SHIFT3) : 0;
        peers[u] = peers_of(key[u], act[u], nbits);
      }
      if (pass == 0) {
#pragma unroll
        for (int u = 0; u < SF; u++)
          if (act[u] && __popcll(peers[u] & lt) == 0) atomicAdd(&hw[key[u]], __popcll(peers[u]));
      } else {
        int old[SF];
#pragma unroll
        for (int u = 0; u < SF; u++) {
          old[u] = 0;
          if (act[u] && __popcll(peers[u] & lt) == 0) old[u] = atomicAdd(&hw[key[u]], __popcll(peers[u]));
        }
        int pos[SF];
#pragma unroll
        for (int u = 0; u < SF; u++) {
          const int lead = act[u] ? __builtin_ctzll(peers[u]) : lane;
          pos[u] = __shfl(old[u], lead, 64) + __popcll(peers[u] & lt);
        }
#pragma unroll
        for (int u = 0; u < SF; u++)
          if (act[u]) a.perm[nb + pos[u]] = src[u];
      }
    }
    __syncthreads();
    if (pass == 1) break;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (int k0 = 0; k0 < nx; k0 += SB) {
      const int key = k0 + tid;
      int tot = 0;
      if (key < nx)
        for (int v = 0; v < NW; v++) tot += hist[v * nx + key];
      int v = tot;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
      }
      if (lane == 63) wsum[w] = v;
      __syncthreads();
      int pre = carry;
      for (int q = 0; q < w; q++) pre += wsum[q];
      const int excl = pre + v - tot;
      if (key < nx) {
        cc[key] = excl;
        int st = excl;
        for (int q = 0; q < NW; q++) {
          const int h = hist[q * nx + key];
          hist[q * nx + key] = st;
          st += h;
        }
      }
      __syncthreads();
      if (tid == SB - 1) carry = excl + tot;
      __syncthreads();
    }
    if (tid == 0) cc[nx] = n;
    __syncthreads();
  }
  if (bad) atomicOr(a.err, ERR_ROW);
}

// The fused step's exchange and sort in the fast and mixed modes, in one
// pass per destination column and without the reference's hole filling: the
// column's particles are its own stayers (direction DIR_STAY) in slot order
// and the arrivals of its segments (k_colscan3d's per-direction mover lists),
// each wave taking a quarter of the stayer slots followed by a quarter of the
// arrivals; a stable counting sort by x cell of that sequence, written as a
// permutation.  Each cell receives the particles particle_yz and sort__bucket
// put there (np2 and cumcnt are the reference's), in an order of its own, so
// only the summation order of the deposit changes.
__global__ __launch_bounds__(SB) void k_colsort3d_fast(Sort3Args a) {
  extern __shared__ int hist[];  // [SB/64][nx]
  __shared__ int wsum[SB / 64];
  __shared__ int carry;
  __shared__ int scnt[8], sptr[8];
  __shared__ int nsg;
  constexpr int SF = 8, NW = SB / 64;
  const Grid3D& g = a.g;
  const int nx = g.nx;
  const int rw = blockIdx.x;
  const int s = rw / g.ncol, col = rw % g.ncol;
  const int base = a.colbaseB[rw], n = a.np2B[rw];
  const int L = a.np2N[rw], nb = a.colbaseN[rw];
  const int m = a.colkm[rw * 2 + 1];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const unsigned ntot = (unsigned)(g.ncap * g.nsp);
  if (tid == 0) {
    nsg = a.nseg[col];
    for (int e = 0; e < nsg; e++) {
      const int sg = a.seg[col * 8 + e];
      const int src = s * g.ncol + (sg >> 4), d = sg & 15;
      scnt[e] = a.dcnt[src * 9 + d];
      sptr[e] = a.colbaseB[src] + a.dstart[src * 9 + d];
    }
  }
  for (int q = tid; q < NW * nx; q += SB) hist[q] = 0;
  __syncthreads();
  // this wave's stayer slots [lo, hi) of the column and arrivals [alo, ahi)
  const int lo = (int)(((long long)n * w) / NW), hi = (int)(((long long)n * (w + 1)) / NW);
  const int alo = (int)(((long long)m * w) / NW), ahi = (int)(((long long)m * (w + 1)) / NW);
  const int nsl = hi - lo;  // chunk index space: stayer slots then arrivals
  const int tot = nsl + (ahi - alo);
  // the segments in registers: arrival t of segment q (first arrival sb[q])
  // is list entry sp[q] + t; an empty tail of segments starts at infinity
  int sb[8], sp[8];
  {
    const int nseg = nsg;
    int acc = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const bool in = q < nseg;
      const int cnt = in ? scnt[q] : 0;
      sb[q] = in ? acc : 0x7fffffff;
      sp[q] = in ? sptr[q] - acc : 0;
      acc += cnt;
    }
  }
  int* hw = hist + w * nx;
  int* cc = a.cumcnt + (size_t)rw * (nx + 1);
  bool bad = false;
  for (int pass = 0; pass < 2; pass++) {
    for (int e0 = 0; e0 < tot; e0 += 64 * SF) {
      // a stayer slot's code (direction and key in one word), or an
      // arrival's list entry (its source slot); every load of the chunk is
      // issued before the first is used
      int v[SF];
#pragma unroll
      for (int u = 0; u < SF; u++) {
        const int e = e0 + u * 64 + lane;
        const int t = alo + (e - nsl);
        int d = sp[0];
#pragma unroll
        for (int q = 1; q < 8; q++) d = t >= sb[q] ? sp[q] : d;
        const int* ptr = e < nsl ? a.code + (base + lo + e) : (e < tot ? a.mlist + (d + t) : a.code);
        v[u] = *ptr;
      }
      int src[SF];
#pragma unroll
      for (int u = 0; u < SF; u++) {
        const int e = e0 + u * 64 + lane;
        const bool arr = e >= nsl && e < tot;
        const bool okarr = arr && (unsigned)v[u] < ntot;
        bad |= arr && !okarr;
        src[u] = e < nsl ? (((v[u] & 15) == DIR_STAY) ? base + lo + e : -1) : (okarr ? v[u] : -1);
      }
      // an arrival's code (the stayers already hold theirs)
      int c2[SF];
#pragma unroll
      for (int u = 0; u < SF; u++) {
        const int e = e0 + u * 64 + lane;
        c2[u] = a.code[(e >= nsl && src[u] >= 0) ? src[u] : 0];
      }
      int key[SF];
#pragma unroll
      for (int u = 0; u < SF; u++) {
        const int e = e0 + u * 64 + lane;
        key[u] = src[u] < 0 ? 0 : ((e < nsl ? v[u] : c2[u]) >> CODE_SHIFT3);
      }
      // one LDS atomic per particle: lanes of a key meet on its counter and
      // the LDS serves them one after another (the order within a cell is
      // the fast modes' own)
      if (pass == 0) {
#pragma unroll
        for (int u = 0; u < SF; u++)
          if (src[u] >= 0) atomicAdd(&hw[key[u]], 1);
      } else {
        int pos[SF];
#pragma unroll
        for (int u = 0; u < SF; u++) pos[u] = src[u] >= 0 ? atomicAdd(&hw[key[u]], 1) : 0;
#pragma unroll
        for (int u = 0; u < SF; u++) {
          const bool ok = src[u] >= 0 && pos[u] < L;
          bad |= src[u] >= 0 && !ok;
          if (ok) a.perm[nb + pos[u]] = src[u];
        }
      }
    }
    __syncthreads();
    if (pass == 1) break;
    // exclusive scan over keys of the column totals -> cumcnt; per-wave cursors
    if (tid == 0) carry = 0;
    __syncthreads();
    for (int k0 = 0; k0 < nx; k0 += SB) {
      const int key = k0 + tid;
      int kt = 0;
      if (key < nx)
        for (int q = 0; q < NW; q++) kt += hist[q * nx + key];
      int v = kt;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
      }
      if (lane == 63) wsum[w] = v;
      __syncthreads();
      int pre = carry;
      for (int q = 0; q < w; q++) pre += wsum[q];
      const int excl = pre + v - kt;
      if (key < nx) {
        cc[key] = excl;
        int st = excl;
        for (int q = 0; q < NW; q++) {
          const int h = hist[q * nx + key];
          hist[q * nx + key] = st;
          st += h;
        }
      }
      __syncthreads();
      if (tid == SB - 1) carry = excl + kt;
      __syncthreads();
    }
    if (tid == 0) cc[nx] = L;
    __syncthreads();
  }
  if (bad) atomicOr(a.err, ERR_ROW);
}

// sort__bucket on a gp that did not go through particle_yz: identity R
__global__ __launch_bounds__(SB) void k_codes_identity3d(Sort3Args a) {
  const Grid3D& g = a.g;
  const int rw = blockIdx.x;
  const int base = a.colbaseB[rw], n = a.np2B[rw], nb = a.colbaseN[rw];
  for (int q = threadIdx.x; q < n; q += SB) {
    int ni = (int)a.B.x[base + q];
    if (ni < g.nxgs || ni > g.nxge) {
      atomicOr(a.err, ERR_DOMAIN);
      ni = ni < g.nxgs ? g.nxgs : g.nxge;
    }
    a.code[base + q] = ((ni - g.nxgs) << CODE_SHIFT3) | DIR_STAY;
    a.R[nb + q] = base + q;
  }
}

static Sort3Args sort3_args(wpic_ctx3d* c) {
  Sort3Args a;
  a.g = c->g;
  a.A = c->A;
  a.B = c->B;
  a.perm = c->permbuf;
  a.code = c->code;
  a.R = c->R;
  a.holes = c->holes;
  a.mlist = c->mlist;
  a.colbaseB = c->colbaseB;
  a.np2B = c->np2B;
  a.colbaseN = c->colbaseN;
  a.np2N = c->np2N;
  a.cumcnt = c->cumcnt;
  a.dcnt = c->dcnt;
  a.dstart = c->dstart;
  a.colkm = c->colkm;
  a.csum = c->csum;
  a.seg = c->seg;
  a.nseg = c->nseg;
  a.err = c->err;
  return a;
}

static dim3 ncols(wpic_ctx3d* c) { return dim3((unsigned)(c->g.nsp * c->g.ncol)); }

hipError_t launch_wrapx3d(wpic_ctx3d* c) {
  hipLaunchKernelGGL(k_wrapx3d, ncols(c), dim3(SB), 0, c->st, sort3_args(c));
  return hipGetLastError();
}

// wrap = 0: the codes come from the mixed-precision push (cell wrap done)
hipError_t launch_colscan3d(wpic_ctx3d* c, int wrap) {
  hipLaunchKernelGGL(k_colscan3d, ncols(c), dim3(SB), 0, c->st, sort3_args(c), wrap);
  return hipGetLastError();
}

hipError_t launch_colcount3d(wpic_ctx3d* c) {
  const Sort3Args a = sort3_args(c);
  const int nbs = (c->g.ncol + CC - 1) / CC, n = c->g.nsp * c->g.ncol;
  hipLaunchKernelGGL(k_colcount3d, dim3((unsigned)(c->g.nsp * nbs)), dim3(CC), 0, c->st, a);
  hipLaunchKernelGGL(k_colcount3d_scan, dim3(1), dim3(64), 0, c->st, a);
  hipLaunchKernelGGL(k_colcount3d_base, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->st, a);
  return hipGetLastError();
}

hipError_t launch_colbuild3d(wpic_ctx3d* c) {
  hipLaunchKernelGGL(k_colbuild3d, ncols(c), dim3(SB), 0, c->st, sort3_args(c));
  return hipGetLastError();
}

hipError_t launch_colsort3d(wpic_ctx3d* c) {
  int nbits = 0;
  while ((1 << nbits) < c->g.nx) nbits++;
  const size_t shm = sizeof(int) * (SB / 64) * (size_t)c->g.nx;
  hipLaunchKernelGGL(k_colsort3d, ncols(c), dim3(SB), shm, c->st, sort3_args(c), nbits);
  return hipGetLastError();
}

hipError_t launch_colsort3d_fast(wpic_ctx3d* c) {
  const size_t shm = sizeof(int) * (SB / 64) * (size_t)c->g.nx;
  hipLaunchKernelGGL(k_colsort3d_fast, ncols(c), dim3(SB), shm, c->st, sort3_args(c));
  return hipGetLastError();
}

hipError_t launch_codes_identity3d(wpic_ctx3d* c) {
  hipLaunchKernelGGL(k_codes_identity3d, ncols(c), dim3(SB), 0, c->st, sort3_args(c));
  return hipGetLastError();
}

}  // namespace wpic
