// k_sort2d.hip -- particle boundaries, row exchange and per-row counting sort
// (2d/common/boundary_periodic.f90:61-248, 2d/common/sort.f90:36-82,
// 2d/proj/reconnection/boundary_reconnection.f90:61-99) at one rank.
//
// Particles of species s live in rows j = nys..nye, each row a contiguous
// segment of the species' structure-of-arrays block.  The reference's in-row
// order after boundary_periodic__particle_y (leavers' slots refilled first by
// arriving particles, then from the tail) followed by its stable bucket sort
// is reproduced exactly, but computed in parallel:
//   k_rowscan   one workgroup per row: leaver slots ("holes") and the up/down
//               movers in slot order, by wavefront ballots and prefix counts;
//   k_rowcount  new row lengths and bases (one scan over rows);
//   k_rowbuild  one workgroup per row: R = for each slot of the post-exchange
//               row, the index of the particle that lands there;
//   k_rowsort   one workgroup per row: stable counting sort of R by cell
//               (per-wave histograms, then a wave multi-split by ballots on
//               the key bits), written as a permutation.
#include "wpic_internal.h"

namespace wpic {

constexpr int SB = 256;  // threads per row workgroup

struct SortArgs {
  Grid2D g;
  Parts A, B;
  int *code, *R, *holes, *ulist, *dlist;
  int *rowbaseB, *np2B, *rowbaseN, *np2N, *rowbaseA, *np2A, *cumcnt, *rowcnt;
  int* err;
  const int* rcnt;  // particles per (species, row, x cell) from the fused push, or null
};

__device__ inline uint64_t lanemask_lt(int lane) { return lane == 0 ? 0ull : (~0ull >> (64 - lane)); }

// boundary_periodic__particle_x (round-down wrap), the walls of
// boundary_reconnection__particle_x / boundary_shock__particle_x, or (inj)
// boundary_shock__injection, on B
__global__ __launch_bounds__(SB) void k_wrapx(SortArgs a, int inj) {
  const Grid2D& g = a.g;
  const int rw = blockIdx.x;
  const int base = a.rowbaseB[rw], n = a.np2B[rw];
  for (int k = threadIdx.x; k < n; k += SB) {
    const int p = base + k;
    double x = a.B.x[p];
    if (walls(g)) {
      double ux = a.B.ux[p], uy = a.B.uy[p], uz = a.B.uz[p];
      if (inj)
        reflect_x<WPIC_BC_SHOCK, true>(g, x, ux, uy, uz);
      else
        reflect_x<WPIC_BC_SHOCK, false>(g, x, ux, uy, uz);
      a.B.x[p] = x;
      a.B.ux[p] = ux;
      a.B.uy[p] = uy;
      a.B.uz[p] = uz;
    } else {
      const int ipos = cell_rd(x, g.delx);
      if (ipos < g.nxgs)
        a.B.x[p] = rd_add(x, g.spanx);
      else if (ipos >= g.nxge + 1)
        a.B.x[p] = rd_add(x, -g.spanx);
    }
  }
}

// Per row: (YWRAP) the y wrap and destination code as particle_y computes
// them (:143-167), then holes / up movers / down movers in slot order.
template <bool YWRAP>
__global__ __launch_bounds__(SB) void k_rowscan(SortArgs a) {
  __shared__ int wt[SB / 64][3];
  __shared__ int run[3];
  const Grid2D& g = a.g;
  const int rw = blockIdx.x;
  const int j = g.nys + rw % g.ny;
  const int base = a.rowbaseB[rw], n = a.np2B[rw];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (tid < 3) run[tid] = 0;
  __syncthreads();
  for (int c0 = 0; c0 < n; c0 += SB) {
    const int k = c0 + tid;
    int dy = 0;
    if (k < n) {
      const int p = base + k;
      if (YWRAP) {
        double y = a.B.y[p];
        const int jpos = cell_rd(y, g.delx);
        if (jpos != j) {
          if (jpos <= g.nygs - 1)
            a.B.y[p] = rd_add(y, g.spany);
          else if (jpos >= g.nyge + 1)
            a.B.y[p] = rd_add(y, -g.spany);
        }
        int ni = (int)a.B.x[p];
        dy = jpos - j;
        if (ni < g.nxgs || ni > g.nxge || dy < -1 || dy > 1) {
          atomicOr(a.err, ERR_DOMAIN);
          ni = ni < g.nxgs ? g.nxgs : (ni > g.nxge ? g.nxge : ni);
          dy = dy < -1 ? -1 : (dy > 1 ? 1 : dy);
        }
        a.code[p] = ((ni - g.nxgs) << 2) | (dy + 1);
      } else {
        dy = (a.code[p] & 3) - 1;
      }
    }
    const uint64_t bl = __ballot(dy != 0), bu = __ballot(dy == 1), bd = __ballot(dy == -1);
    const uint64_t lt = lanemask_lt(lane);
    if (lane == 0) {
      wt[w][0] = __popcll(bl);
      wt[w][1] = __popcll(bu);
      wt[w][2] = __popcll(bd);
    }
    __syncthreads();
    int ol = run[0], ou = run[1], od = run[2];
    for (int v = 0; v < w; v++) {
      ol += wt[v][0];
      ou += wt[v][1];
      od += wt[v][2];
    }
    if (dy != 0) a.holes[base + ol + __popcll(bl & lt)] = k + 1;
    if (dy == 1) a.ulist[base + ou + __popcll(bu & lt)] = base + k;
    if (dy == -1) a.dlist[base + od + __popcll(bd & lt)] = base + k;
    __syncthreads();
    if (tid < 3) {
      int t = 0;
      for (int v = 0; v < SB / 64; v++) t += wt[v][tid];
      run[tid] += t;
    }
    __syncthreads();
  }
  if (tid < 3) a.rowcnt[rw * 4 + tid] = run[tid];
}

// incoming count m_J and new row length L = np2 - k + m; bases by one scan
// per species.  bff(J) receives row J-1's up movers then row J+1's down
// movers; at J = nys the down movers of nys+1 come first and nye's up movers
// (sent through the periodic transfer, :173-189) are appended.
__global__ __launch_bounds__(1024) void k_rowcount(SortArgs a) {
  __shared__ int wsum[16];
  __shared__ int carry;
  const Grid2D& g = a.g;
  const int ny = g.ny, tid = threadIdx.x;
  for (int s = 0; s < g.nsp; s++) {
    if (tid == 0) carry = 0;
    __syncthreads();
    for (int r0 = 0; r0 < ny; r0 += 1024) {
      const int r = r0 + tid;
      int L = 0;
      if (r < ny) {
        const int rw = s * ny + r;
        const int rp = s * ny + (r == 0 ? ny - 1 : r - 1), rn = s * ny + (r == ny - 1 ? 0 : r + 1);
        const int m = a.rowcnt[rp * 4 + 1] + a.rowcnt[rn * 4 + 2];
        a.rowcnt[rw * 4 + 3] = m;
        L = a.np2B[rw] - a.rowcnt[rw * 4 + 0] + m;
        a.np2N[rw] = L;
      }
      // block exclusive scan of L
      const int lane = tid & 63, w = tid >> 6;
      int v = L;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
      }
      if (lane == 63) wsum[w] = v;
      __syncthreads();
      if (tid == 0) {
        int acc = 0;
        for (int q = 0; q < 16; q++) {
          const int t = wsum[q];
          wsum[q] = acc;
          acc += t;
        }
      }
      __syncthreads();
      const int excl = carry + wsum[w] + v - L;
      if (r < ny) a.rowbaseN[s * ny + r] = (int)(s * g.ncap) + excl;
      __syncthreads();
      if (tid == 1023) carry = excl + L;
      __syncthreads();
    }
  }
}

struct Incoming {
  const int *l1, *l2;
  int c1;
  __device__ int operator()(int t) const { return t < c1 ? l1[t] : l2[t - c1]; }
};

// R for one row: slot -> source particle (boundary_periodic.f90:193-236)
__global__ __launch_bounds__(SB) void k_rowbuild(SortArgs a) {
  __shared__ int wt[SB / 64][2];
  __shared__ int run[2];
  __shared__ int sh_a;
  const Grid2D& g = a.g;
  const int rw = blockIdx.x;
  const int s = rw / g.ny, r = rw % g.ny, ny = g.ny;
  const int base = a.rowbaseB[rw], n = a.np2B[rw];
  const int kk = a.rowcnt[rw * 4 + 0], m = a.rowcnt[rw * 4 + 3];
  const int L = a.np2N[rw], nb = a.rowbaseN[rw];
  const int rp = s * ny + (r == 0 ? ny - 1 : r - 1), rn = s * ny + (r == ny - 1 ? 0 : r + 1);
  Incoming in;
  if (r == 0) {
    in.l1 = a.dlist + a.rowbaseB[rn];
    in.c1 = a.rowcnt[rn * 4 + 2];
    in.l2 = a.ulist + a.rowbaseB[rp];
  } else {
    in.l1 = a.ulist + a.rowbaseB[rp];
    in.c1 = a.rowcnt[rp * 4 + 1];
    in.l2 = a.dlist + a.rowbaseB[rn];
  }
  const int* holes = a.holes + base;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;

  // non-leavers keep their slot when it survives
  const int keep = n < L ? n : L;
  for (int k = tid; k < keep; k += SB)
    if ((a.code[base + k] & 3) == 1) a.R[nb + k] = base + k;
  // the first min(k, m) holes take the arrivals in order
  const int kf = kk < m ? kk : m;
  for (int t = tid; t < kf; t += SB) {
    const int h = holes[t];
    if (h >= 1 && h <= L) a.R[nb + h - 1] = in(t);
  }
  // more arrivals than holes: the rest are appended
  for (int t = tid; t < m - kk; t += SB) a.R[nb + n + t] = in(kk + t);
  if (m >= kk) return;

  // fewer arrivals than holes: tail particles fill the remaining holes, the
  // last one into the first free hole (loop1, :197-216).  Tail = slots (L, n].
  if (tid == 0) {
    int lo = 0, hi = kk;  // a = number of holes <= L
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (holes[mid] <= L)
        lo = mid + 1;
      else
        hi = mid;
    }
    sh_a = lo;
    run[0] = 0;  // leavers seen in the tail so far
    run[1] = 0;  // non-H' slots seen so far
  }
  __syncthreads();
  const int aa = sh_a;
  const int T = n - L;
  // pass 1: count the tail slots that are not unfilled holes (non-H')
  int totalq = 0;
  for (int pass = 0; pass < 2; pass++) {
    for (int c0 = 0; c0 < T; c0 += SB) {
      const int k = c0 + tid;  // tail offset, slot = L + 1 + k
      bool lv = false;
      if (k < T) lv = (a.code[base + L + k] & 3) != 1;
      const uint64_t bl = __ballot(lv);
      if (lane == 0) wt[w][0] = __popcll(bl);
      __syncthreads();
      int ol = run[0];
      for (int v = 0; v < w; v++) ol += wt[v][0];
      const int rank = aa + ol + __popcll(bl & lanemask_lt(lane)) + 1;  // 1-based hole rank
      const bool nonh = k < T && (!lv || rank <= m);
      const uint64_t bn = __ballot(nonh);
      if (lane == 0) wt[w][1] = __popcll(bn);
      __syncthreads();
      int on = run[1];
      for (int v = 0; v < w; v++) on += wt[v][1];
      if (pass == 1 && nonh) {
        // descending position among non-H' tail slots
        const int incl = on + __popcll(bn & lanemask_lt(lane)) + 1;
        const int rr = totalq - incl + 1;
        const int tgt = m + rr;  // 1-based hole index
        const int src = lv ? in(rank - 1) : base + L + k;
        const int h = (tgt >= 1 && tgt <= kk) ? holes[tgt - 1] : 0;
        if (h >= 1 && h <= L)
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

// peers of this lane among the active lanes: lanes holding the same key
__device__ inline uint64_t key_peers_bits(int key, bool act, int nbits) {
  uint64_t peers = __ballot(act);
  for (int b = 0; b < nbits; b++) {
    const bool bit = (key >> b) & 1;
    const uint64_t m = __ballot(act && bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

// The same mask, cheaper for the usual chunk: slots in cell order carry keys
// within a few cells of each other, so when every active key lies within
// [k0 - 8, k0 + 7] of the first active lane's k0, four ballots on key - k0
// + 8 decide equality; otherwise all nbits.
__device__ inline uint64_t key_peers(int key, bool act, int nbits) {
  const uint64_t am = __ballot(act);
  if (am == 0) return 0;
  const int k0 = __builtin_amdgcn_readlane(key, __builtin_ctzll(am));
  const int rel = key - k0 + 8;
  if (__ballot(act && (rel < 0 || rel > 15)) != 0) return key_peers_bits(key, act, nbits);
  uint64_t peers = am;
#pragma unroll
  for (int b = 0; b < 4; b++) {
    const bool bit = (rel >> b) & 1;
    const uint64_t m = __ballot(act && bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

// stable counting sort of one row by cell (sort.f90:52-78) as a permutation.
// Each wave owns a contiguous quarter of the row: its per-key histogram, then
// (after a scan over keys and waves) its per-key write cursors.  Within a
// wave, lanes with equal keys are ranked by one ballot per key bit.  Chunks
// of 64 are processed U at a time so their loads overlap.
constexpr int SU = 4;

// The sort moves no particle record: for every sorted slot it writes the gp
// slot its particle comes from; the next push reads the record and its id
// through it.
__global__ __launch_bounds__(SB) void k_rowsort(SortArgs a, int* perm, int nbits) {
  extern __shared__ int hist[];  // [SB/64][nx]
  __shared__ int wsum[SB / 64];
  __shared__ int carry;
  const Grid2D& g = a.g;
  const int nx = g.nx;
  const int rw = blockIdx.x;
  const int n = a.np2N[rw], nb = a.rowbaseN[rw];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const unsigned ntot = (unsigned)(g.ncap * g.nsp);
  const uint64_t lt = lanemask_lt(lane);
  constexpr int NW = SB / 64;
  for (int k = tid; k < NW * nx; k += SB) hist[k] = 0;
  __syncthreads();
  const int lo = (int)(((long long)n * w) / NW), hi = (int)(((long long)n * (w + 1)) / NW);
  int* hw = hist + w * nx;
  // per-wave histogram; one lane per distinct key adds its group's count
  for (int k0 = lo; k0 < hi; k0 += 64 * SU) {
    int key[SU];
    bool act[SU];
#pragma unroll
    for (int u = 0; u < SU; u++) {
      const int k = k0 + u * 64 + lane;
      act[u] = k < hi;
      int src = act[u] ? a.R[nb + k] : 0;
      if ((unsigned)src >= ntot) src = 0;  // reported in the scatter pass
      key[u] = act[u] ? (a.code[src] >> 2) : 0;
    }
#pragma unroll
    for (int u = 0; u < SU; u++) {
      const uint64_t peers = key_peers(key[u], act[u], nbits);
      if (act[u] && __popcll(peers & lt) == 0) hw[key[u]] += __popcll(peers);
    }
  }
  __syncthreads();
  // exclusive scan over keys of the row totals -> cumcnt; per-wave cursors
  if (tid == 0) carry = 0;
  __syncthreads();
  int* cc = a.cumcnt + (size_t)rw * (nx + 1);
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
  // stable scatter: each wave walks its part in order
  for (int k0 = lo; k0 < hi; k0 += 64 * SU) {
    int src[SU], key[SU];
    bool act[SU];
#pragma unroll
    for (int u = 0; u < SU; u++) {
      const int k = k0 + u * 64 + lane;
      act[u] = k < hi;
      src[u] = act[u] ? a.R[nb + k] : 0;
      if ((unsigned)src[u] >= ntot) {
        atomicOr(a.err, ERR_ROW);
        act[u] = false;
        src[u] = 0;
      }
    }
#pragma unroll
    for (int u = 0; u < SU; u++) key[u] = act[u] ? (a.code[src[u]] >> 2) : 0;
#pragma unroll
    for (int u = 0; u < SU; u++) {
      const uint64_t peers = key_peers(key[u], act[u], nbits);
      const int rank = __popcll(peers & lt);
      const int cnt = __popcll(peers);
      int pos = 0;
      if (act[u]) pos = hw[key[u]] + rank;
      if (act[u] && rank == 0) hw[key[u]] = pos + cnt;
      if (act[u]) perm[nb + pos] = src[u];
    }
  }
}

// The fused step's exchange and sort in the fast and mixed modes, in one pass
// per destination row and without the reference's hole filling: the row's
// particles are its own stayers (row step 0) in slot order and the arrivals
// of the rows below and above (the up/down lists of k_rowscan), each wave
// taking a quarter of the stayer slots followed by a quarter of the
// arrivals; a stable counting sort by cell of that sequence, written as a
// permutation.  Each cell receives the particles the reference's exchange and
// sort__bucket put there (cumcnt and np2 are the reference's), in a fixed
// order of its own, so only the summation order of the deposit changes.
__global__ __launch_bounds__(SB) void k_rowsort_fast(SortArgs a, int* perm, int nbits) {
  extern __shared__ int hist[];  // [SB/64][nx]
  __shared__ int wsum[SB / 64];
  __shared__ int carry;
  const Grid2D& g = a.g;
  const int nx = g.nx, ny = g.ny;
  const int rw = blockIdx.x;
  const int s = rw / ny, r = rw % ny;
  const int base = a.rowbaseB[rw], n = a.np2B[rw];
  const int L = a.np2N[rw], nb = a.rowbaseN[rw];
  const int rp = s * ny + (r == 0 ? ny - 1 : r - 1), rn = s * ny + (r == ny - 1 ? 0 : r + 1);
  const int* ul = a.ulist + a.rowbaseB[rp];
  const int nu = a.rowcnt[rp * 4 + 1];
  const int* dl = a.dlist + a.rowbaseB[rn];
  const int m = nu + a.rowcnt[rn * 4 + 2];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint64_t lt = lanemask_lt(lane);
  constexpr int NW = SB / 64;
  for (int k = tid; k < NW * nx; k += SB) hist[k] = 0;
  __syncthreads();
  // this wave's stayer slots [lo, hi) of the row and arrivals [alo, ahi)
  const int lo = (int)(((long long)n * w) / NW), hi = (int)(((long long)n * (w + 1)) / NW);
  const int alo = (int)(((long long)m * w) / NW), ahi = (int)(((long long)m * (w + 1)) / NW);
  const int nsl = hi - lo;  // chunk index space: stayer slots then arrivals
  const int tot = nsl + (ahi - alo);
  // with the push's per-cell counts (rcnt) the cursors come from them and
  // one pass places the particles, the four waves sharing one cursor per key
  // (the order within a cell is then the LDS's, not the slots'); otherwise a
  // counting pass builds per-wave histograms first
  const bool counted = a.rcnt != nullptr;
  int* hw = counted ? hist : hist + w * nx;
  const unsigned ntot = (unsigned)(g.ncap * g.nsp);
  const int* codes = a.code;
  if (counted) {
    if (tid == 0) carry = 0;
    __syncthreads();
    int* cc = a.cumcnt + (size_t)rw * (nx + 1);
    const int* rc = a.rcnt + (size_t)rw * nx;
    for (int k0 = 0; k0 < nx; k0 += SB) {
      const int key = k0 + tid;
      const int t = key < nx ? rc[key] : 0;
      int v = t;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(v, off, 64);
        if (lane >= off) v += y;
      }
      if (lane == 63) wsum[w] = v;
      __syncthreads();
      int pre = carry;
      for (int q = 0; q < w; q++) pre += wsum[q];
      const int excl = pre + v - t;
      if (key < nx) {
        cc[key] = excl;
        hist[key] = excl;
      }
      __syncthreads();
      if (tid == SB - 1) carry = excl + t;
      __syncthreads();
    }
    if (tid == 0) {
      cc[nx] = L;
      if (carry != L) atomicOr(a.err, ERR_ROW);
    }
    __syncthreads();
  }
  // SF chunks of 64 per wave and round.  Every load of a round is issued
  // before the first is used (addresses are selected, not branched on), and
  // the per-key cursors move by LDS atomics, whose returns the LDS serves in
  // issue order: a round waits for memory twice and for the LDS once.
  constexpr int SF = 8;
  bool bad = false;
  for (int pass = counted ? 1 : 0; pass < 2; pass++) {
    for (int e0 = 0; e0 < tot; e0 += 64 * SF) {
      // a stayer slot's code (row step and key in one word), or an arrival's
      // list entry (its source slot)
      int v[SF];
#pragma unroll
      for (int u = 0; u < SF; u++) {
        const int e = e0 + u * 64 + lane;
        const int t = alo + (e - nsl);
        const int* ptr = e < nsl ? codes + (base + lo + e) : (e < tot ? (t < nu ? ul + t : dl + (t - nu)) : codes);
        v[u] = *ptr;
      }
      int src[SF];
#pragma unroll
      for (int u = 0; u < SF; u++) {
        const int e = e0 + u * 64 + lane;
        const bool arr = e >= nsl && e < tot;
        const bool okarr = arr && (unsigned)v[u] < ntot;
        bad |= arr && !okarr;
        src[u] = e < nsl ? (((v[u] & 3) == 1) ? base + lo + e : -1) : (okarr ? v[u] : -1);
      }
      // an arrival's code (the stayers already hold theirs)
      int c2[SF];
#pragma unroll
      for (int u = 0; u < SF; u++) {
        const int e = e0 + u * 64 + lane;
        c2[u] = codes[(e >= nsl && src[u] >= 0) ? src[u] : 0];
      }
      int key[SF];
      uint64_t peers[SF];
#pragma unroll
      for (int u = 0; u < SF; u++) {
        const int e = e0 + u * 64 + lane;
        key[u] = src[u] < 0 ? 0 : ((e < nsl ? v[u] : c2[u]) >> 2);
        peers[u] = key_peers(key[u], src[u] >= 0, nbits);
      }
      if (pass == 0) {
#pragma unroll
        for (int u = 0; u < SF; u++)
          if (src[u] >= 0 && __popcll(peers[u] & lt) == 0) atomicAdd(&hw[key[u]], __popcll(peers[u]));
      } else {
        // the lowest lane of each key takes the key's cursor for all of them
        int old[SF];
#pragma unroll
        for (int u = 0; u < SF; u++) {
          old[u] = 0;
          if (src[u] >= 0 && __popcll(peers[u] & lt) == 0) old[u] = atomicAdd(&hw[key[u]], __popcll(peers[u]));
        }
        int pos[SF];
#pragma unroll
        for (int u = 0; u < SF; u++) {
          const int lead = src[u] >= 0 ? __builtin_ctzll(peers[u]) : lane;
          pos[u] = __shfl(old[u], lead, 64) + __popcll(peers[u] & lt);
        }
#pragma unroll
        for (int u = 0; u < SF; u++) {
          const bool ok = src[u] >= 0 && pos[u] < L;
          bad |= src[u] >= 0 && !ok;
          if (ok) perm[nb + pos[u]] = src[u];
        }
      }
    }
    __syncthreads();
    if (pass == 1) break;
    // exclusive scan over keys of the row totals -> cumcnt; per-wave cursors
    if (tid == 0) carry = 0;
    __syncthreads();
    int* cc = a.cumcnt + (size_t)rw * (nx + 1);
    for (int k0 = 0; k0 < nx; k0 += SB) {
      const int key = k0 + tid;
      int t = 0;
      if (key < nx)
        for (int v = 0; v < NW; v++) t += hist[v * nx + key];
      int v = t;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(v, off, 64);
        if (lane >= off) v += y;
      }
      if (lane == 63) wsum[w] = v;
      __syncthreads();
      int pre = carry;
      for (int q = 0; q < w; q++) pre += wsum[q];
      const int excl = pre + v - t;
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
      if (tid == SB - 1) carry = excl + t;
      __syncthreads();
    }
    if (tid == 0) {
      cc[nx] = L;
      if (carry != L) atomicOr(a.err, ERR_ROW);
    }
    __syncthreads();
  }
  if (bad) atomicOr(a.err, ERR_ROW);
}

// sort__bucket on a gp that did not go through bc__particle_y (the restart
// path, app.f90:351-352): identity permutation, key int(x)
__global__ __launch_bounds__(SB) void k_codes_identity(SortArgs a) {
  const Grid2D& g = a.g;
  const int rw = blockIdx.x;
  const int base = a.rowbaseB[rw], n = a.np2B[rw], nb = a.rowbaseN[rw];
  for (int k = threadIdx.x; k < n; k += SB) {
    int ni = (int)a.B.x[base + k];
    if (ni < g.nxgs || ni > g.nxge) {
      atomicOr(a.err, ERR_DOMAIN);
      ni = ni < g.nxgs ? g.nxgs : g.nxge;
    }
    a.code[base + k] = ((ni - g.nxgs) << 2) | 1;
    a.R[nb + k] = base + k;
  }
}

static SortArgs sort_args(wpic_ctx* c) {
  SortArgs a;
  a.g = c->g;
  a.A = c->A;
  a.B = c->B;
  a.code = c->code;
  a.R = c->R;
  a.holes = c->holes;
  a.ulist = c->ulist;
  a.dlist = c->dlist;
  a.rowbaseB = c->rowbaseB;
  a.np2B = c->np2B;
  a.rowbaseN = c->rowbaseN;
  a.np2N = c->np2N;
  a.rowbaseA = c->rowbaseA;
  a.np2A = c->np2A;
  a.cumcnt = c->cumcnt;
  a.rowcnt = c->rowcnt;
  a.err = c->err;
  a.rcnt = c->rcnt_fill ? c->rcnt : nullptr;
  return a;
}

hipError_t launch_wrapx(wpic_ctx* c, bool injection) {
  SortArgs a = sort_args(c);
  hipLaunchKernelGGL(k_wrapx, dim3(c->g.nsp * c->g.ny), dim3(SB), 0, c->st, a, injection ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_rowscan(wpic_ctx* c, bool ywrap) {
  SortArgs a = sort_args(c);
  if (ywrap)
    hipLaunchKernelGGL(k_rowscan<true>, dim3(c->g.nsp * c->g.ny), dim3(SB), 0, c->st, a);
  else
    hipLaunchKernelGGL(k_rowscan<false>, dim3(c->g.nsp * c->g.ny), dim3(SB), 0, c->st, a);
  return hipGetLastError();
}

hipError_t launch_codes_identity(wpic_ctx* c) {
  SortArgs a = sort_args(c);
  hipLaunchKernelGGL(k_codes_identity, dim3(c->g.nsp * c->g.ny), dim3(SB), 0, c->st, a);
  return hipGetLastError();
}

hipError_t launch_rowcount(wpic_ctx* c) {
  SortArgs a = sort_args(c);
  hipLaunchKernelGGL(k_rowcount, dim3(1), dim3(1024), 0, c->st, a);
  return hipGetLastError();
}

hipError_t launch_rowbuild(wpic_ctx* c) {
  SortArgs a = sort_args(c);
  hipLaunchKernelGGL(k_rowbuild, dim3(c->g.nsp * c->g.ny), dim3(SB), 0, c->st, a);
  return hipGetLastError();
}

hipError_t launch_rowsort_fast(wpic_ctx* c) {
  SortArgs a = sort_args(c);
  int nbits = 0;
  while ((1 << nbits) < c->g.nx) nbits++;
  const size_t shm = sizeof(int) * (SB / 64) * (size_t)c->g.nx;
  int* perm = c->perm_out ? c->perm_out : c->permbuf;
  hipLaunchKernelGGL(k_rowsort_fast, dim3(c->g.nsp * c->g.ny), dim3(SB), shm, c->st, a, perm, nbits);
  return hipGetLastError();
}

hipError_t launch_rowsort(wpic_ctx* c) {
  SortArgs a = sort_args(c);
  int nbits = 0;
  while ((1 << nbits) < c->g.nx) nbits++;
  const size_t shm = sizeof(int) * (SB / 64) * (size_t)c->g.nx;
  int* perm = c->perm_out ? c->perm_out : c->permbuf;
  hipLaunchKernelGGL(k_rowsort, dim3(c->g.nsp * c->g.ny), dim3(SB), shm, c->st, a, perm, nbits);
  return hipGetLastError();
}

}  // namespace wpic
