// wpic_api3d.hip -- the C-ABI of include/wpic3d.h: the 3D context, the
// reference entry points as a state machine over device-resident arrays,
// layout conversion at the boundary and the fused step loop.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "wpic_internal3d.h"

using namespace wpic;

#define HIPCHK(expr)                                  \
  do {                                                \
    hipError_t e_ = (expr);                           \
    if (e_ != hipSuccess) return hip_fail(e_, #expr); \
  } while (0)

namespace {

template <class T>
hipError_t dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  return hipMalloc((void**)p, n * sizeof(T));
}

void fold_timers3(wpic_ctx3d* c) {
  for (auto& t : c->timers) {
    if (!t.b) continue;
    float ms = 0.f;
    (void)hipEventSynchronize(t.b);
    (void)hipEventElapsedTime(&ms, t.a, t.b);
    bool found = false;
    for (auto& s : c->tsum)
      if (s.first == t.name) {
        s.second.first += ms;
        s.second.second += 1;
        found = true;
      }
    if (!found) c->tsum.push_back({t.name, {ms, 1}});
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  c->timers.clear();
}

struct Timed3 {
  wpic_ctx3d* c;
  hipEvent_t a = nullptr;
  Timed3(wpic_ctx3d* cc, const char* name) : c(cc) {
    if (!c->timing) return;
    (void)hipEventCreate(&a);
    (void)hipEventRecord(a, c->st);
    TimerRec t;
    t.name = name;
    t.a = a;
    t.b = nullptr;
    c->timers.push_back(t);
  }
  ~Timed3() {
    if (!c->timing || !a) return;
    for (auto it = c->timers.rbegin(); it != c->timers.rend(); ++it)
      if (it->a == a) {
        (void)hipEventCreate(&it->b);
        (void)hipEventRecord(it->b, c->st);
        break;
      }
    if (c->timers.size() > 2048) fold_timers3(c);
  }
};

int pick_free3(int a, int b) {
  for (int k = 0; k < 3; k++)
    if (k != a && k != b) return k;
  return 0;
}

void set_cols(wpic_ctx3d* c) {
  const int nr = c->g.nsp * c->g.ncol;
  c->colbaseA = c->cols[c->ca];
  c->np2A = c->cols[c->ca] + nr;
  c->colbaseB = c->cols[c->cb];
  c->np2B = c->cols[c->cb] + nr;
  c->colbaseN = c->cols[c->cn];
  c->np2N = c->cols[c->cn] + nr;
}

int fetch_cols(wpic_ctx3d* c, int which) {
  if (c->h_cols_ok[which]) return 0;
  const int nr = c->g.nsp * c->g.ncol;
  c->h_cols[which].resize(2 * (size_t)nr);
  HIPCHK(hipMemcpyAsync(c->h_cols[which].data(), c->cols[which], sizeof(int) * 2 * nr, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  c->h_cols_ok[which] = true;
  return 0;
}

int ensure_staging3(wpic_ctx3d* c, size_t n) {
  if (c->staging_n >= n) return 0;
  if (c->staging) HIPCHK(hipFree(c->staging));
  c->staging = nullptr;
  HIPCHK(dalloc(&c->staging, n));
  c->staging_n = n;
  return 0;
}

int check_err3(wpic_ctx3d* c) {
  int e = 0;
  HIPCHK(hipMemcpyAsync(&e, c->err, sizeof(int), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  if (e) {
    HIPCHK(hipMemsetAsync(c->err, 0, sizeof(int), c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (e & ERR_CG) return set_error(WPIC_ECG, "********** stop at cgm after ite_max **********");
    if (e & ERR_DOMAIN)
      return set_error(WPIC_EDOMAIN, "particle outside the domain or moved more than one cell in a step");
    return set_error(WPIC_EINVAL, "inconsistent particle columns on the device (exchange bookkeeping)");
  }
  return 0;
}

// The arrival order of boundary_periodic__particle_yz at one rank
// (3d/common/boundary_periodic.f90:147-455), computed once on tokens: every
// (source column, direction) pair lands in the bucket of (jpos, kpos) during
// the (k, j) scan; then the y ghost buckets are appended to nye / nys (for
// every kb, ghosts included) and the z ghost buckets to nze / nzs (interior
// j).  The bucket of an interior column is then its ordered segment list.
void segment_table(const Grid3D& g, std::vector<int>& seg, std::vector<int>& nseg) {
  const int ny = g.ny, nz = g.nz, bw = ny + 2, bh = nz + 2;
  std::vector<std::vector<int>> B((size_t)bw * bh);
  auto bk = [&](int jb, int kb) -> std::vector<int>& { return B[(size_t)(jb + 1) + (size_t)bw * (kb + 1)]; };
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int d = 0; d < 9; d++) {
        if (d == DIR_STAY) continue;
        const int dj = d % 3 - 1, dk = d / 3 - 1;
        bk(j + dj, k + dk).push_back(((j + ny * k) << 4) | d);
      }
  auto append = [](std::vector<int>& dst, std::vector<int>& src) {
    dst.insert(dst.end(), src.begin(), src.end());
    src.clear();
  };
  for (int kb = -1; kb <= nz; kb++) append(bk(ny - 1, kb), bk(-1, kb));
  for (int kb = -1; kb <= nz; kb++) append(bk(0, kb), bk(ny, kb));
  for (int j = 0; j < ny; j++) append(bk(j, nz - 1), bk(j, -1));
  for (int j = 0; j < ny; j++) append(bk(j, 0), bk(j, nz));
  seg.assign((size_t)ny * nz * 8, 0);
  nseg.assign((size_t)ny * nz, 0);
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++) {
      const std::vector<int>& v = bk(j, k);
      const int col = j + ny * k;
      nseg[col] = (int)v.size();
      for (size_t e = 0; e < v.size() && e < 8; e++) seg[(size_t)col * 8 + e] = v[e];
    }
}

int ensure_pjbuf3(wpic_ctx3d* c) {
  if (c->pjbuf) return 0;
  HIPCHK(dalloc(&c->pjbuf, (size_t)c->g.nx * c->g.ny * c->g.nz * 300));
  return 0;
}

int field_after_deposit3(wpic_ctx3d* c) {
  Timed3 t(c, "field");
  HIPCHK(launch_curre_bc3d(c));
  {
    Timed3 tc(c, "cg");
    HIPCHK(launch_cg_init3d(c));
    const int NB = 4;
    int it = 0;
    struct {
      CgScal s[3];
    } h;
    int herr = 0;
    for (;;) {
      for (int b = 0; b < NB && it < 100; b++, it++) HIPCHK(launch_cg_iter3d(c, it));
      HIPCHK(hipMemcpyAsync(h.s, c->scal, sizeof(h.s), hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipMemcpyAsync(&herr, c->err, sizeof(int), hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipStreamSynchronize(c->st));
      if (herr) return check_err3(c);
      if (!h.s[0].active && !h.s[1].active && !h.s[2].active) break;
      if (it >= 100) return set_error(WPIC_ECG, "********** stop at cgm after ite_max **********");
    }
    for (int l = 0; l < 3; l++) c->cg_iters[l] = h.s[l].ite;
    HIPCHK(launch_cg_store3d(c));
  }
  HIPCHK(launch_dfield_bc3d(c));
  HIPCHK(launch_e_update3d(c));
  HIPCHK(launch_dfield_bc3d(c));
  HIPCHK(launch_uf_add3d(c));
  return 0;
}

int deposit3(wpic_ctx3d* c) {
  const bool parity = c->mode == WPIC_MODE_PARITY;
  if (parity) {
    int rc = ensure_pjbuf3(c);
    if (rc) return rc;
  } else {
    HIPCHK(launch_fill3d(c, c->uj, 3 * plane3(c->g), 0.0));
  }
  {
    Timed3 t(c, "deposit");
    HIPCHK(launch_deposit3d(c));
  }
  if (parity) HIPCHK(launch_gather_pj3d(c));
  return 0;
}

int push3(wpic_ctx3d* c) {
  c->cb = c->ca;
  if (c->cn == c->ca) c->cn = pick_free3(c->ca, c->cb);
  set_cols(c);
  {
    Timed3 t(c, "push");
    HIPCHK(launch_push3d(c));
  }
  c->b_valid = true;
  c->b_coded = false;
  c->b_accl = false;
  c->b_aligned = true;
  c->exch_pending = false;
  return 0;
}

// fast and mixed modes: particle__solv and ele_cur in one fused pass, which
// also writes the exchange codes with the x cell wrapped (and, FP64, the x,
// y and z wraps of gp: particle_x and particle_yz's position changes)
int push_deposit3(wpic_ctx3d* c) {
  c->cb = c->ca;
  if (c->cn == c->ca) c->cn = pick_free3(c->ca, c->cb);
  set_cols(c);
  HIPCHK(launch_fill3d(c, c->uj, 3 * plane3(c->g), 0.0));
  {
    Timed3 t(c, "push_deposit");
    if (c->mixed)
      HIPCHK(launch_push_deposit_mixed3d(c));
    else
      HIPCHK(launch_push_deposit3d(c));
  }
  c->b_valid = true;
  c->b_accl = false;
  c->b_aligned = true;
  c->b_coded = true;
  c->exch_pending = false;
  return 0;
}

// particle_yz on gp (columns B -> columns N, permutation R); with
// `lists_only` the column counts and mover lists alone, for the fused fast
// sort (k_colsort3d_fast builds each new column from them)
int exchange3(wpic_ctx3d* c, bool lists_only = false) {
  Timed3 t(c, "exchange");
  HIPCHK(launch_colscan3d(c, c->b_coded ? 0 : 1));
  HIPCHK(launch_colcount3d(c));
  if (!lists_only) HIPCHK(launch_colbuild3d(c));
  c->h_cols_ok[c->cn] = false;
  c->exch_pending = true;
  return 0;
}

// sort__bucket: gp through R into A, by x cell within columns (fast: the
// fused step's one-pass exchange and sort, k_colsort3d_fast)
int sort_into_a3(wpic_ctx3d* c, bool fast = false) {
  {
    Timed3 t(c, "sort");
    HIPCHK(fast ? launch_colsort3d_fast(c) : launch_colsort3d(c));
  }
  const int old_a = c->ca;
  c->ca = c->cn;
  c->cn = old_a;
  if (c->cb == c->ca) c->cb = pick_free3(c->ca, c->cn);
  // the sorted state is gp's buffer seen through the new permutation; the
  // old sorted buffer becomes the next gp
  c->aperm = c->permbuf;
  std::swap(c->A, c->B);
  std::swap(c->MA, c->MB);
  set_cols(c);
  c->b_valid = false;
  c->b_accl = false;
  c->b_aligned = false;
  c->exch_pending = false;
  c->h_cols_ok[c->ca] = false;
  return 0;
}

int mixed_only_step3() {
  return set_error(WPIC_EINVAL,
                   "the mixed-precision form runs the fused loop (wpic_step3d); single routines need an FP64 mode");
}

int cols_from_np2(wpic_ctx3d* c, const int* np2, std::vector<int>& hcols, std::vector<long long>& aosb,
                  long long& total) {
  const int ncol = c->g.ncol, nsp = c->g.nsp, nr = nsp * ncol;
  hcols.assign(2 * (size_t)nr, 0);
  aosb.assign(nr, 0);
  total = 0;
  for (int s = 0; s < nsp; s++) {
    long long off = 0;
    for (int r = 0; r < ncol; r++) {
      const int rw = s * ncol + r;
      if (np2[rw] < 0) return set_error(WPIC_EINVAL, "negative np2");
      hcols[rw] = (int)(s * c->g.ncap + off);
      hcols[nr + rw] = np2[rw];
      aosb[rw] = total;
      off += np2[rw];
      total += np2[rw];
    }
    if (off > c->g.ncap) return set_error(WPIC_EMEM, "memory over (np2 > np): species exceeds the device capacity");
  }
  return 0;
}

int upload_cols(wpic_ctx3d* c, int which, const double* up, const int* np2, int np, int kind) {
  std::vector<int> hcols;
  std::vector<long long> aosb;
  long long total = 0;
  int rc = cols_from_np2(c, np2, hcols, aosb, total);
  if (rc) return rc;
  const int nr = c->g.nsp * c->g.ncol;
  for (int rw = 0; rw < nr; rw++)
    if (np2[rw] > np) return set_error(WPIC_EINVAL, "np2 exceeds the column capacity np");
  rc = ensure_staging3(c, (size_t)total * 7);
  if (rc) return rc;
  std::vector<double> h((size_t)total * 7);
  for (int rw = 0; rw < nr; rw++)
    std::memcpy(&h[(size_t)aosb[rw] * 7], up + (size_t)7 * np * rw, sizeof(double) * 7 * (size_t)np2[rw]);
  HIPCHK(hipMemcpyAsync(c->staging, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->cols[which], hcols.data(), sizeof(int) * 2 * nr, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->aosbase, aosb.data(), sizeof(long long) * nr, hipMemcpyHostToDevice, c->st));
  if (kind == 3)
    HIPCHK(launch_up_to_mixed3d(c, c->staging));
  else
    HIPCHK(launch_up_to_soa3d(c, c->staging, total, kind));
  HIPCHK(hipStreamSynchronize(c->st));
  c->h_cols[which] = hcols;
  c->h_cols_ok[which] = true;
  return 0;
}

int download_cols(wpic_ctx3d* c, int which, int kind, double* up, int* np2, int np) {
  int rc = fetch_cols(c, which);
  if (rc) return rc;
  const int nr = c->g.nsp * c->g.ncol;
  const std::vector<int>& hc = c->h_cols[which];
  std::vector<long long> aosb(nr);
  long long total = 0;
  for (int rw = 0; rw < nr; rw++) {
    aosb[rw] = total;
    total += hc[nr + rw];
    if (hc[nr + rw] > np) return set_error(WPIC_EMEM, "memory over (np2 > np)");
  }
  rc = ensure_staging3(c, (size_t)total * 7);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(c->aosbase, aosb.data(), sizeof(long long) * nr, hipMemcpyHostToDevice, c->st));
  if (kind == 3)
    HIPCHK(launch_mixed_to_aos3d(c, c->staging));
  else
    HIPCHK(launch_soa_to_aos3d(c, c->staging, kind));
  std::vector<double> h((size_t)total * 7);
  HIPCHK(hipMemcpyAsync(h.data(), c->staging, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  for (int rw = 0; rw < nr; rw++) {
    std::memcpy(up + (size_t)7 * np * rw, &h[(size_t)aosb[rw] * 7], sizeof(double) * 7 * (size_t)hc[nr + rw]);
    if (np2) np2[rw] = hc[nr + rw];
  }
  return 0;
}

int download_planes3(wpic_ctx3d* c, const double* planes, double* out, int ncomp) {
  const size_t n = (size_t)ncomp * plane3(c->g);
  int rc = ensure_staging3(c, n);
  if (rc) return rc;
  HIPCHK(launch_planes_to_aos3d(c, planes, c->staging, ncomp));
  HIPCHK(hipMemcpyAsync(out, c->staging, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int upload_planes3(wpic_ctx3d* c, const double* in, double* planes, int ncomp) {
  const size_t n = (size_t)ncomp * plane3(c->g);
  int rc = ensure_staging3(c, n);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(c->staging, in, sizeof(double) * n, hipMemcpyHostToDevice, c->st));
  HIPCHK(launch_planes_from_aos3d(c, c->staging, planes, ncomp));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

}  // namespace

extern "C" {

int wpic_create3d(wpic_ctx3d** out, const wpic_params3d* p) {
  if (!out || !p) return set_error(WPIC_EINVAL, "null argument");
  *out = nullptr;
  if (p->ndim != 7) return set_error(WPIC_EINVAL, "Error: ndim must be 7");
  if (p->nsp < 1 || p->nsp > WPIC_MAXSP) return set_error(WPIC_EINVAL, "nsp out of range");
  const int nx = p->nxge - p->nxgs + 1, ny = p->nyge - p->nygs + 1, nz = p->nzge - p->nzgs + 1;
  if (nx < 4 || ny < 3 || nz < 3) return set_error(WPIC_EINVAL, "grid must be at least 4x3x3 cells");
  if (nx > 4096) return set_error(WPIC_EINVAL, "nx above 4096 cells per column is not supported");
  if (p->nxgs < 2 || p->nygs < 2 || p->nzgs < 2)
    return set_error(WPIC_EINVAL, "nxgs, nygs and nzgs must be >= 2 (positive coordinates)");
  if (p->delx != 1.0)
    return set_error(WPIC_EINVAL, "delx must be 1: sort__bucket bins by int(x) (3d/common/sort.f90:65)");
  if (!(p->c * p->delt <= p->delx))
    return set_error(WPIC_EINVAL, "c*delt must not exceed delx (particles move less than one cell per step)");
  if (p->capacity <= 0 || p->capacity * p->nsp >= (1ll << 31))
    return set_error(WPIC_EINVAL, "capacity must be positive and nsp*capacity below 2^31");
  if ((long long)ny * nz >= (1ll << 27)) return set_error(WPIC_EINVAL, "too many (j, k) columns");
  if ((long long)nx * ny * nz >= (1ll << 31)) return set_error(WPIC_EINVAL, "too many cells (2^31 at most)");
  if (p->mode != WPIC_MODE_FAST && p->mode != WPIC_MODE_PARITY && p->mode != WPIC_MODE_MIXED)
    return set_error(WPIC_EINVAL, "bad mode");

  wpic_ctx3d* c = new wpic_ctx3d();
  std::memset((void*)&c->g, 0, sizeof(c->g));
  c->P = *p;
  c->mode = p->mode;
  c->pusher = WPIC_PUSHER_BORIS;
  c->mixed = p->mode == WPIC_MODE_MIXED;
  Grid3D& g = c->g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.nxgs = p->nxgs;
  g.nxge = p->nxge;
  g.nygs = p->nygs;
  g.nyge = p->nyge;
  g.nzgs = p->nzgs;
  g.nzge = p->nzge;
  g.nys = p->nygs;
  g.nye = p->nyge;
  g.nzs = p->nzgs;
  g.nze = p->nzge;
  g.fw = nx + 4;
  g.fh = ny + 4;
  g.fd = nz + 4;
  g.nsp = p->nsp;
  g.ncol = ny * nz;
  g.delx = p->delx;
  g.delt = p->delt;
  g.c = p->c;
  g.gfac = p->gfac;
  g.d_delx = 1e0 / p->delx;
  g.d_delt = 1e0 / p->delt;
  const double pi = 3.141592653589793;
  // 3d/common/field.f90:57-61
  g.f1 = p->c * p->delt / p->delx;
  g.f2 = p->gfac * g.f1 * g.f1;
  g.f3 = 4e0 * pi * p->delx / p->c;
  const double t = p->delx / (p->c * p->delt * p->gfac);
  g.f4 = 6e0 + t * t;
  g.f5 = t * t;
  g.pi4dt = 4e0 * pi * p->delt;
  g.cg_err = 1e-6;
  // (nxge-nxgs+1)*delx etc. (boundary_periodic.f90:89-91,150-163)
  g.spanx = (double)nx * p->delx;
  g.spany = (double)ny * p->delx;
  g.spanz = (double)nz * p->delx;
  for (int s = 0; s < WPIC_MAXSP; s++) {
    g.q[s] = p->q[s];
    g.r[s] = p->r[s];
  }
  g.ncap = p->capacity;

  std::vector<int> hseg, hnseg;
  segment_table(g, hseg, hnseg);

  hipError_t e = hipGetDevice(&c->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_push, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_sort, hipEventDisableTiming);
  const size_t N = (size_t)g.ncap * g.nsp;
  const int nr = g.nsp * g.ncol;
  const size_t pl = plane3(g), nc = (size_t)nx * ny * nz;
  if (c->mixed) {
    for (PartsM3* m : {&c->MA, &c->MB}) {
      if (e == hipSuccess) e = dalloc(&m->ix, N);
      for (float** q : {&m->ox, &m->oy, &m->oz, &m->ux, &m->uy, &m->uz})
        if (e == hipSuccess) e = dalloc(q, N);
    }
  } else {
    double** dp[] = {&c->A.x, &c->A.y, &c->A.z, &c->A.ux, &c->A.uy, &c->A.uz,
                     &c->B.x, &c->B.y, &c->B.z, &c->B.ux, &c->B.uy, &c->B.uz};
    for (auto q : dp)
      if (e == hipSuccess) e = dalloc(q, N);
  }
  for (int k = 0; k < 2 && e == hipSuccess; k++) e = dalloc(&c->id[k], N);
  int** ip[] = {&c->code, &c->R, &c->holes, &c->mlist, &c->permbuf};
  for (auto q : ip)
    if (e == hipSuccess) e = dalloc(q, N);
  for (int k = 0; k < 3 && e == hipSuccess; k++) {
    e = dalloc(&c->cols[k], (size_t)2 * nr);
    if (e == hipSuccess) e = hipMemset(c->cols[k], 0, sizeof(int) * 2 * nr);
  }
  if (e == hipSuccess) e = dalloc(&c->cumcnt, (size_t)nr * (nx + 1));
  if (e == hipSuccess) e = hipMemset(c->cumcnt, 0, sizeof(int) * (size_t)nr * (nx + 1));
  if (e == hipSuccess) e = dalloc(&c->dcnt, (size_t)9 * nr);
  if (e == hipSuccess) e = dalloc(&c->dstart, (size_t)9 * nr);
  if (e == hipSuccess) e = dalloc(&c->colkm, (size_t)2 * nr);
  if (e == hipSuccess) e = dalloc(&c->csum, (size_t)c->g.nsp * ((c->g.ncol + 1023) / 1024));
  if (e == hipSuccess) e = dalloc(&c->seg, hseg.size());
  if (e == hipSuccess) e = dalloc(&c->nseg, hnseg.size());
  if (e == hipSuccess) e = hipMemcpy(c->seg, hseg.data(), sizeof(int) * hseg.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->nseg, hnseg.data(), sizeof(int) * hnseg.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = dalloc(&c->aosbase, (size_t)nr);
  if (e == hipSuccess) e = dalloc(&c->uf, 6 * pl);
  if (e == hipSuccess) e = dalloc(&c->df, 6 * pl);
  if (e == hipSuccess) e = dalloc(&c->uj, 3 * pl);
  double** cg[] = {&c->phi, &c->pa, &c->pb, &c->rr, &c->bb, &c->ap};
  for (auto q : cg)
    if (e == hipSuccess) e = dalloc(q, 3 * nc);
  if (e == hipSuccess) e = dalloc(&c->part, (size_t)3 * NPB3 * 2);
  if (e == hipSuccess) e = dalloc(&c->scal, 3);
  if (e == hipSuccess) e = dalloc(&c->err, 1);
  if (e == hipSuccess) e = dalloc(&c->nov, 1);
  if (e == hipSuccess) e = dalloc(&c->red, (size_t)256 * 3);
  if (e == hipSuccess) e = hipMemset(c->err, 0, sizeof(int));
  if (e == hipSuccess) e = hipMemset(c->uf, 0, sizeof(double) * 6 * pl);
  if (e == hipSuccess) e = hipMemset(c->df, 0, sizeof(double) * 6 * pl);
  if (e == hipSuccess) e = hipMemset(c->uj, 0, sizeof(double) * 3 * pl);
  if (e == hipSuccess) e = hipMemset(c->R, 0, sizeof(int) * N);
  if (e == hipSuccess) e = hipMemset(c->code, 0, sizeof(int) * N);
  if (e == hipSuccess) e = hipMemset(c->permbuf, 0, sizeof(int) * N);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    wpic_destroy3d(c);
    return hip_fail(e, "wpic_create3d");
  }
  c->ca = 0;
  c->cb = 1;
  c->cn = 2;
  set_cols(c);
  c->A.id = c->MA.id = c->id[0];
  c->B.id = c->MB.id = c->id[1];
  for (int k = 0; k < 3; k++) c->h_cols_ok[k] = false;
  *out = c;
  return 0;
}

int wpic_destroy3d(wpic_ctx3d* c) {
  if (!c) return 0;
  if (c->st) (void)hipStreamSynchronize(c->st);
  for (auto& t : c->timers) {
    if (t.a) (void)hipEventDestroy(t.a);
    if (t.b) (void)hipEventDestroy(t.b);
  }
  void* ps[] = {c->A.x,    c->A.y,   c->A.z,      c->A.ux,    c->A.uy,   c->A.uz,    c->B.x,    c->B.y,
                c->B.z,    c->B.ux,  c->B.uy,     c->B.uz,    c->id[0],  c->id[1],   c->code,   c->R,
                c->holes,  c->mlist, c->cols[0],  c->cols[1], c->cols[2], c->cumcnt, c->dcnt,   c->dstart,
                c->colkm,  c->csum,  c->seg,   c->nseg,     c->aosbase, c->uf,     c->df,      c->uj,     c->pjbuf,
                c->phi,    c->pa,    c->pb,       c->rr,      c->bb,     c->ap,      c->part,   c->scal,
                c->err,    c->staging, c->gbuf,   c->red,     c->permbuf, c->nov,    c->mom,
                c->MA.ix,  c->MA.ox, c->MA.oy,    c->MA.oz,   c->MA.ux,  c->MA.uy,  c->MA.uz,
                c->MB.ix,  c->MB.ox, c->MB.oy,    c->MB.oz,   c->MB.ux,  c->MB.uy,  c->MB.uz};
  for (void* q : ps)
    if (q) (void)hipFree(q);
  if (c->ev_push) (void)hipEventDestroy(c->ev_push);
  if (c->ev_sort) (void)hipEventDestroy(c->ev_sort);
  if (c->st2) (void)hipStreamDestroy(c->st2);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
  return 0;
}

int wpic_set_mode3d(wpic_ctx3d* c, int mode) {
  if (!c || (mode != WPIC_MODE_FAST && mode != WPIC_MODE_PARITY && mode != WPIC_MODE_MIXED))
    return set_error(WPIC_EINVAL, "bad mode");
  if (c->mixed != (mode == WPIC_MODE_MIXED))
    return set_error(WPIC_EINVAL, "the mixed-precision form is fixed when the context is created");
  c->mode = mode;
  return 0;
}

int wpic_sync3d(wpic_ctx3d* c) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  HIPCHK(hipStreamSynchronize(c->st));
  return check_err3(c);
}

int wpic_upload_up3d(wpic_ctx3d* c, const double* up, const int* np2, const int* cumcnt, int np) {
  if (!c || !up || !np2 || !cumcnt) return set_error(WPIC_EINVAL, "null argument");
  const int nx = c->g.nx, nr = c->g.nsp * c->g.ncol;
  for (int rw = 0; rw < nr; rw++) {
    const int* cr = cumcnt + (size_t)rw * (nx + 1);
    if (cr[0] != 0 || cr[nx] != np2[rw])
      return set_error(WPIC_EINVAL, "cumcnt must run from 0 to np2 in every column (a sorted up)");
    for (int i = 0; i < nx; i++)
      if (cr[i + 1] < cr[i]) return set_error(WPIC_EINVAL, "cumcnt must be non-decreasing");
  }
  c->aperm = nullptr;
  int rc = upload_cols(c, c->ca, up, np2, np, c->mixed ? 3 : 0);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(c->cumcnt, cumcnt, sizeof(int) * (size_t)nr * (nx + 1), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  c->b_valid = false;
  c->b_accl = false;
  c->b_aligned = false;
  c->exch_pending = false;
  return 0;
}

int wpic_download_up3d(wpic_ctx3d* c, double* up, int* np2, int* cumcnt, int np) {
  if (!c || !up) return set_error(WPIC_EINVAL, "null argument");
  int rc = download_cols(c, c->ca, c->mixed ? 3 : 0, up, np2, np);
  if (rc) return rc;
  if (cumcnt) {
    const size_t ncc = (size_t)c->g.nsp * c->g.ncol * (c->g.nx + 1);
    HIPCHK(hipMemcpy(cumcnt, c->cumcnt, sizeof(int) * ncc, hipMemcpyDeviceToHost));
  }
  return 0;
}

int wpic_upload_gp3d(wpic_ctx3d* c, const double* gp, const int* np2, int np) {
  if (!c || !gp || !np2) return set_error(WPIC_EINVAL, "null argument");
  if (c->mixed) return mixed_only_step3();
  if (c->cb == c->ca) c->cb = pick_free3(c->ca, c->cn);
  set_cols(c);
  int rc = upload_cols(c, c->cb, gp, np2, np, 1);
  if (rc) return rc;
  c->b_valid = true;
  c->b_coded = false;
  c->b_accl = false;
  c->exch_pending = false;
  rc = fetch_cols(c, c->ca);
  if (rc) return rc;
  c->b_aligned = c->h_cols[c->ca] == c->h_cols[c->cb];
  return 0;
}

int wpic_download_gp3d(wpic_ctx3d* c, double* gp, int* np2, int np) {
  if (!c || !gp) return set_error(WPIC_EINVAL, "null argument");
  if (c->mixed) return mixed_only_step3();
  if (!c->b_valid) return set_error(WPIC_EINVAL, "no gp on the device (call particle__solv first)");
  if (c->exch_pending) return download_cols(c, c->cn, 1, gp, np2, np);
  return download_cols(c, c->cb, 2, gp, np2, np);
}

int wpic_upload_fields3d(wpic_ctx3d* c, const double* uf) {
  if (!c || !uf) return set_error(WPIC_EINVAL, "null argument");
  return upload_planes3(c, uf, c->uf, 6);
}

int wpic_download_fields3d(wpic_ctx3d* c, double* uf) {
  if (!c || !uf) return set_error(WPIC_EINVAL, "null argument");
  return download_planes3(c, c->uf, uf, 6);
}

int wpic_download_current3d(wpic_ctx3d* c, double* uj) {
  if (!c || !uj) return set_error(WPIC_EINVAL, "null argument");
  return download_planes3(c, c->uj, uj, 3);
}

int wpic_download_dfield3d(wpic_ctx3d* c, double* df) {
  if (!c || !df) return set_error(WPIC_EINVAL, "null argument");
  return download_planes3(c, c->df, df, 6);
}

int wpic_upload_dfield3d(wpic_ctx3d* c, const double* df) {
  if (!c || !df) return set_error(WPIC_EINVAL, "null argument");
  return upload_planes3(c, df, c->df, 6);
}

int wpic_particle_solv3d(wpic_ctx3d* c) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (c->mixed) return mixed_only_step3();
  int rc = push3(c);
  if (rc) return rc;
  return check_err3(c);
}

int wpic_particle_solv_vay3d(wpic_ctx3d* c) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  const int saved = c->pusher;
  c->pusher = WPIC_PUSHER_VAY;
  const int rc = wpic_particle_solv3d(c);
  c->pusher = saved;
  return rc;
}

int wpic_set_pusher3d(wpic_ctx3d* c, int pusher) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (pusher != WPIC_PUSHER_BORIS && pusher != WPIC_PUSHER_VAY) return set_error(WPIC_EINVAL, "bad pusher");
  if (c->mixed && pusher != WPIC_PUSHER_BORIS)
    return set_error(WPIC_EINVAL, "the Vay kick runs in the FP64 modes (the mixed form pushes with Buneman-Boris)");
  c->pusher = pusher;
  return 0;
}

int wpic_field_fdtd_i3d(wpic_ctx3d* c) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (c->mixed) return mixed_only_step3();
  if (!c->b_valid || !c->b_aligned || c->exch_pending)
    return set_error(WPIC_EINVAL, "field__fdtd_i needs gp slot-aligned with up (call particle__solv first)");
  int rc = deposit3(c);
  if (rc) return rc;
  rc = field_after_deposit3(c);
  if (rc) return rc;
  return check_err3(c);
}

int wpic_bc_particle_x3d(wpic_ctx3d* c) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (c->mixed) return mixed_only_step3();
  if (!c->b_valid || c->exch_pending) return set_error(WPIC_EINVAL, "bc__particle_x needs gp before the exchange");
  c->b_accl = false;
  HIPCHK(launch_wrapx3d(c));
  return check_err3(c);
}

int wpic_bc_particle_yz3d(wpic_ctx3d* c) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (c->mixed) return mixed_only_step3();
  if (!c->b_valid || c->exch_pending) return set_error(WPIC_EINVAL, "bc__particle_yz needs gp before the exchange");
  if (c->cn == c->cb || c->cn == c->ca) c->cn = pick_free3(c->ca, c->cb);
  set_cols(c);
  int rc = exchange3(c);
  if (rc) return rc;
  return check_err3(c);
}

int wpic_sort_bucket3d(wpic_ctx3d* c) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (c->mixed) return mixed_only_step3();
  if (!c->b_valid) return set_error(WPIC_EINVAL, "sort__bucket needs gp on the device");
  if (!c->exch_pending) {
    if (c->cn == c->cb || c->cn == c->ca) c->cn = pick_free3(c->ca, c->cb);
    set_cols(c);
    const int nr = c->g.nsp * c->g.ncol;
    HIPCHK(hipMemcpyAsync(c->cols[c->cn], c->cols[c->cb], sizeof(int) * 2 * nr, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(launch_codes_identity3d(c));
    c->exch_pending = true;
  }
  int rc = sort_into_a3(c);
  if (rc) return rc;
  return check_err3(c);
}

int wpic_cg_iters3d(wpic_ctx3d* c, int iters[3]) {
  if (!c || !iters) return set_error(WPIC_EINVAL, "null argument");
  for (int l = 0; l < 3; l++) iters[l] = c->cg_iters[l];
  return 0;
}

int wpic_set_cg_tol3d(wpic_ctx3d* c, double err) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (!(err > 0.0 && err < 1.0)) return set_error(WPIC_EINVAL, "the CG tolerance must lie in (0, 1)");
  c->g.cg_err = err;
  return 0;
}

namespace {
// launches made while it lives go to the context's second stream
struct OnSecondStream3 {
  wpic_ctx3d* c;
  hipStream_t saved;
  explicit OnSecondStream3(wpic_ctx3d* c_) : c(c_), saved(c_->st) { c->st = c->st2; }
  ~OnSecondStream3() { c->st = saved; }
};
}  // namespace

int wpic_step3d(wpic_ctx3d* c, int nsteps) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  for (int it = 0; it < nsteps; it++) {
    c->cn = (c->ca + 1) % 3;
    int rc;
    if (c->mode == WPIC_MODE_FAST || c->mixed) {
      rc = push_deposit3(c);
      if (rc) return rc;
    } else {
      rc = push3(c);
      if (rc) return rc;
      c->cn = (c->ca + 1) % 3;
      set_cols(c);
      rc = deposit3(c);
      if (rc) return rc;
    }
    // the x wrap, exchange and sort work on gp and its columns, the field
    // solve on the current and the fields: the former run on the second
    // stream beside the latter, and the first stream waits for them
    HIPCHK(hipEventRecord(c->ev_push, c->st));
    HIPCHK(hipStreamWaitEvent(c->st2, c->ev_push, 0));
    {
      OnSecondStream3 on2(c);
      if (!c->b_coded) {
        Timed3 t(c, "exchange");
        HIPCHK(launch_wrapx3d(c));
      }
      const bool fast = c->mode == WPIC_MODE_FAST || c->mixed;
      rc = exchange3(c, fast);
      if (rc) return rc;
      rc = sort_into_a3(c, fast);
      if (rc) return rc;
      HIPCHK(hipEventRecord(c->ev_sort, c->st));
    }
    rc = field_after_deposit3(c);
    HIPCHK(hipStreamWaitEvent(c->st, c->ev_sort, 0));
    if (rc) return rc;
  }
  return check_err3(c);
}

int wpic_energy3d(wpic_ctx3d* c, double out[3]) {
  if (!c || !out) return set_error(WPIC_EINVAL, "null argument");
  HIPCHK(launch_energy3d(c, c->red));
  std::vector<double> h(256 * 3);
  HIPCHK(hipMemcpyAsync(h.data(), c->red, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  double k = 0, e = 0, b = 0;
  for (int i = 0; i < 256; i++) {
    k += h[i * 3];
    e += h[i * 3 + 1];
    b += h[i * 3 + 2];
  }
  const double pi = 3.141592653589793;
  out[0] = k;
  out[1] = e / (8 * pi);
  out[2] = b / (8 * pi);
  return 0;
}

int wpic_energy_species3d(wpic_ctx3d* c, double* out) {
  if (!c || !out) return set_error(WPIC_EINVAL, "null argument");
  const double pi = 3.141592653589793;
  std::vector<double> h(256 * 3);
  for (int s = 0; s < c->g.nsp; s++) {
    HIPCHK(launch_energy3d(c, c->red, s));
    HIPCHK(hipMemcpyAsync(h.data(), c->red, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    double k = 0, e = 0, b = 0;
    for (int i = 0; i < 256; i++) {
      k += h[i * 3];
      e += h[i * 3 + 1];
      b += h[i * 3 + 2];
    }
    out[s] = k;
    out[c->g.nsp] = e / (8 * pi);
    out[c->g.nsp + 1] = b / (8 * pi);
  }
  return 0;
}

int wpic_ic_weibel3d(wpic_ctx3d* c, unsigned long long seed, int n0, double v_thi, double v_the, double t_ani,
                     double b0) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (c->mixed) return set_error(WPIC_EINVAL, "the Weibel lattice state is built in the FP64 modes");
  const Grid3D& g = c->g;
  if (g.nsp != 2) return set_error(WPIC_EINVAL, "the Weibel state has two species (ions, electrons)");
  if (n0 <= 0 || (long long)n0 * g.nx * g.ncol > g.ncap)
    return set_error(WPIC_EMEM, "Error: Too large number of particles");
  const int nr = g.nsp * g.ncol;
  std::vector<int> hcols(2 * (size_t)nr);
  std::vector<int> cc((size_t)nr * (g.nx + 1));
  for (int s = 0; s < g.nsp; s++)
    for (int r = 0; r < g.ncol; r++) {
      const int rw = s * g.ncol + r;
      hcols[rw] = (int)(s * g.ncap + (long long)r * n0 * g.nx);
      hcols[nr + rw] = n0 * g.nx;
      // the x lattice puts n0 particles in each cell, in order
      for (int i = 0; i <= g.nx; i++) cc[(size_t)rw * (g.nx + 1) + i] = i * n0;
    }
  HIPCHK(hipMemcpyAsync(c->cols[c->ca], hcols.data(), sizeof(int) * 2 * nr, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->cumcnt, cc.data(), sizeof(int) * cc.size(), hipMemcpyHostToDevice, c->st));
  c->aperm = nullptr;
  const double sd[WPIC_MAXSP] = {v_thi, v_the, 0.0, 0.0};
  HIPCHK(launch_ic_weibel3d(c, seed, n0, sd, t_ani));
  const size_t pl = plane3(g);
  HIPCHK(launch_fill3d(c, c->uf, 6 * pl, 0.0));
  HIPCHK(launch_fill3d(c, c->uf + 2 * pl, pl, b0));
  HIPCHK(launch_fill3d(c, c->df, 6 * pl, 0.0));
  HIPCHK(hipStreamSynchronize(c->st));
  c->h_cols[c->ca] = hcols;
  c->h_cols_ok[c->ca] = true;
  c->b_valid = false;
  c->b_accl = false;
  c->b_aligned = false;
  c->exch_pending = false;
  return 0;
}

int wpic_gauss3d(wpic_ctx3d* c, double* G) {
  if (!c || !G) return set_error(WPIC_EINVAL, "null argument");
  const size_t nc = (size_t)c->g.nx * c->g.ny * c->g.nz;
  if (!c->gbuf) HIPCHK(dalloc(&c->gbuf, 2 * nc));
  HIPCHK(launch_gauss3d(c, c->gbuf + nc, c->gbuf));
  HIPCHK(hipMemcpyAsync(G, c->gbuf + nc, sizeof(double) * nc, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int wpic_counts3d(wpic_ctx3d* c, long long* counts) {
  if (!c || !counts) return set_error(WPIC_EINVAL, "null argument");
  int rc = fetch_cols(c, c->ca);
  if (rc) return rc;
  const int ncol = c->g.ncol, nr = c->g.nsp * ncol;
  for (int s = 0; s < c->g.nsp; s++) {
    long long t = 0;
    for (int r = 0; r < ncol; r++) t += c->h_cols[c->ca][nr + s * ncol + r];
    counts[s] = t;
  }
  return 0;
}

int wpic_ic_beams3d(wpic_ctx3d* c, unsigned long long seed, int n_ppc, double udrift, double uth_e, double uth_i,
                    double b0) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  const Grid3D& g = c->g;
  if (n_ppc <= 0 || (long long)n_ppc * g.nx * g.ncol > g.ncap)
    return set_error(WPIC_EMEM, "Error: Too large number of particles");
  const int nr = g.nsp * g.ncol;
  std::vector<int> hcols(2 * (size_t)nr);
  std::vector<int> cc((size_t)nr * (g.nx + 1));
  for (int s = 0; s < g.nsp; s++)
    for (int r = 0; r < g.ncol; r++) {
      const int rw = s * g.ncol + r;
      hcols[rw] = (int)(s * g.ncap + (long long)r * n_ppc * g.nx);
      hcols[nr + rw] = n_ppc * g.nx;
      for (int i = 0; i <= g.nx; i++) cc[(size_t)rw * (g.nx + 1) + i] = i * n_ppc;
    }
  HIPCHK(hipMemcpyAsync(c->cols[c->ca], hcols.data(), sizeof(int) * 2 * nr, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->cumcnt, cc.data(), sizeof(int) * cc.size(), hipMemcpyHostToDevice, c->st));
  c->aperm = nullptr;
  HIPCHK(launch_ic_beams3d(c, seed, n_ppc, udrift, uth_e, uth_i));
  const size_t pl = plane3(g);
  HIPCHK(launch_fill3d(c, c->uf, 6 * pl, 0.0));
  HIPCHK(launch_fill3d(c, c->uf + 2 * pl, pl, b0));
  HIPCHK(launch_fill3d(c, c->df, 6 * pl, 0.0));
  HIPCHK(hipStreamSynchronize(c->st));
  c->h_cols[c->ca] = hcols;
  c->h_cols_ok[c->ca] = true;
  c->b_valid = false;
  c->b_accl = false;
  c->b_aligned = false;
  c->exch_pending = false;
  return 0;
}

int wpic_set_timing3d(wpic_ctx3d* c, int on) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  fold_timers3(c);
  c->tsum.clear();
  c->timing = on != 0;
  return 0;
}

int wpic_timing3d(wpic_ctx3d* c, const char* name, double* ms, long long* count) {
  if (!c || !name || !ms || !count) return set_error(WPIC_EINVAL, "null argument");
  HIPCHK(hipStreamSynchronize(c->st));
  fold_timers3(c);
  *ms = 0;
  *count = 0;
  for (auto& s : c->tsum)
    if (s.first == name) {
      *ms = s.second.first;
      *count = s.second.second;
    }
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// moments: mom_calc__accl / mom_calc__nvt (3d/common/mom_calc.f90:49-333) and
// boundary_periodic__mom (3d/common/boundary_periodic.f90:1102-1235)
// ---------------------------------------------------------------------------
static int ensure_mom3(wpic_ctx3d* c) {
  if (c->mom) return 0;
  HIPCHK(dalloc(&c->mom, mom3_size(c->g)));
  HIPCHK(launch_fill3d(c, c->mom, mom3_size(c->g), 0.0));
  return 0;
}

int wpic_mom_accl3d(wpic_ctx3d* c) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (c->mixed) return mixed_only_step3();
  c->cb = c->ca;
  if (c->cn == c->ca) c->cn = pick_free3(c->ca, c->cb);
  set_cols(c);
  {
    Timed3 t(c, "mom_accl");
    HIPCHK(launch_mom_accl3d(c));
  }
  c->b_valid = true;
  c->b_coded = false;
  c->b_accl = true;  // gp holds up's positions: sorted by cell like up
  c->b_aligned = true;
  c->exch_pending = false;
  return check_err3(c);
}

int wpic_mom_nvt3d(wpic_ctx3d* c) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (c->mixed) return mixed_only_step3();
  if (!c->b_valid || c->exch_pending)
    return set_error(WPIC_EINVAL, "mom_calc__nvt needs gp on the device (call mom_calc__accl first)");
  int rc = ensure_mom3(c);
  if (rc) return rc;
  set_cols(c);
  int kind;
  if (c->mode == WPIC_MODE_PARITY)
    kind = c->b_accl && c->b_aligned ? MOM_PARITY_ALIGNED : MOM_PARITY_ANY;
  else
    kind = c->b_aligned ? MOM_TILE_GP : MOM_ANY;
  {
    Timed3 t(c, "mom_nvt");
    HIPCHK(launch_mom_deposit3d(c, kind));
  }
  return check_err3(c);
}

int wpic_bc_mom3d(wpic_ctx3d* c) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  int rc = ensure_mom3(c);
  if (rc) return rc;
  HIPCHK(launch_mom_bc3d(c));
  return check_err3(c);
}

int wpic_moments3d(wpic_ctx3d* c) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (c->mixed) return mixed_only_step3();
  int rc = ensure_mom3(c);
  if (rc) return rc;
  {
    Timed3 t(c, "moments");
    if (c->mode == WPIC_MODE_PARITY) {
      c->cb = c->ca;
      if (c->cn == c->ca) c->cn = pick_free3(c->ca, c->cb);
      set_cols(c);
      HIPCHK(launch_mom_accl3d(c));
      HIPCHK(launch_mom_deposit3d(c, MOM_PARITY_ALIGNED));
    } else {
      set_cols(c);
      HIPCHK(launch_mom_deposit3d(c, MOM_TILE_KICK));
    }
    HIPCHK(launch_mom_bc3d(c));
  }
  // the reference's sequence leaves gp overwritten by mom_calc__accl
  c->b_valid = false;
  c->b_accl = false;
  c->b_aligned = false;
  c->exch_pending = false;
  return check_err3(c);
}

int wpic_download_mom3d(wpic_ctx3d* c, double* mom) {
  if (!c || !mom) return set_error(WPIC_EINVAL, "null argument");
  int rc = ensure_mom3(c);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(mom, c->mom, sizeof(double) * mom3_size(c->g), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int wpic_upload_mom3d(wpic_ctx3d* c, const double* mom) {
  if (!c || !mom) return set_error(WPIC_EINVAL, "null argument");
  int rc = ensure_mom3(c);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(c->mom, mom, sizeof(double) * mom3_size(c->g), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}
