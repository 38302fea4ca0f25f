// wpic_api2d.cpp -- the C-ABI of include/wpic.h: context, the reference
// entry points as a small state machine over device-resident arrays, layout
// conversion at the boundary, and the fused step loop.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "wpic_internal.h"

using namespace wpic;

namespace {
thread_local std::string g_err;
}

namespace wpic {

int set_error(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return set_error(WPIC_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

void timer_begin(wpic_ctx* c, const char* name, hipEvent_t* a) {
  *a = nullptr;
  if (!c->timing) return;
  (void)hipEventCreate(a);
  (void)hipEventRecord(*a, c->st);
  TimerRec t;
  t.name = name;
  t.a = *a;
  t.b = nullptr;
  c->timers.push_back(t);
}

static void fold_timers(wpic_ctx* c) {
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

void timer_end(wpic_ctx* c, hipEvent_t a) {
  if (!c->timing || !a) return;
  for (auto it = c->timers.rbegin(); it != c->timers.rend(); ++it)
    if (it->a == a) {
      (void)hipEventCreate(&it->b);
      (void)hipEventRecord(it->b, c->st);
      break;
    }
  if (c->timers.size() > 2048) fold_timers(c);
}

}  // namespace wpic

#define HIPCHK(expr)                                  \
  do {                                                \
    hipError_t e_ = (expr);                           \
    if (e_ != hipSuccess) return hip_fail(e_, #expr); \
  } while (0)

namespace {

struct Timed {
  wpic_ctx* c;
  hipEvent_t a;
  Timed(wpic_ctx* cc, const char* name) : c(cc) { timer_begin(c, name, &a); }
  ~Timed() { timer_end(c, a); }
};

template <class T>
hipError_t dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  return hipMalloc((void**)p, n * sizeof(T));
}

// the row-structure slot that is neither a nor b
int pick_free(int a, int b) {
  for (int k = 0; k < 3; k++)
    if (k != a && k != b) return k;
  return 0;
}

void set_rows(wpic_ctx* c) {
  const int nr = c->g.nsp * c->g.ny;
  c->rowbaseA = c->rows[c->ra];
  c->np2A = c->rows[c->ra] + nr;
  c->rowbaseB = c->rows[c->rb];
  c->np2B = c->rows[c->rb] + nr;
  c->rowbaseN = c->rows[c->rn];
  c->np2N = c->rows[c->rn] + nr;
}

// host mirror of a row structure (rowbase then np2)
int fetch_rows(wpic_ctx* c, int which) {
  if (c->h_rows_ok[which]) return 0;
  const int nr = c->g.nsp * c->g.ny;
  c->h_rows[which].resize(2 * nr);
  HIPCHK(hipMemcpyAsync(c->h_rows[which].data(), c->rows[which], sizeof(int) * 2 * nr, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  c->h_rows_ok[which] = true;
  return 0;
}

int ensure_staging(wpic_ctx* c, size_t n) {
  if (c->staging_n >= n) return 0;
  if (c->staging) HIPCHK(hipFree(c->staging));
  c->staging = nullptr;
  HIPCHK(dalloc(&c->staging, n));
  c->staging_n = n;
  return 0;
}

int check_err(wpic_ctx* c) {
  int e = 0;
  HIPCHK(hipMemcpyAsync(&e, c->err, sizeof(int), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  if (e) {
    HIPCHK(hipMemsetAsync(c->err, 0, sizeof(int), c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (e & ERR_CG) return set_error(WPIC_ECG, "********** stop at cgm after ite_max **********");
    if (e & ERR_DOMAIN)
      return set_error(WPIC_EDOMAIN, "particle outside the domain or moved more than one cell in a step");
    return set_error(WPIC_EINVAL, "inconsistent particle rows on the device (exchange bookkeeping)");
  }
  return 0;
}

int ensure_pjbuf(wpic_ctx* c) {
  if (c->pjbuf) return 0;
  HIPCHK(dalloc(&c->pjbuf, (size_t)c->g.nx * c->g.ny * 65));
  return 0;
}

// field__fdtd_i after the deposit: current BCs, RHS, CG, E, BCs, update
int cg_loop_impl(wpic_ctx* c) {
  // iterations are launched a batch at a time; the three components stop on
  // their own device flags (an iteration past convergence changes nothing),
  // the host polls once per batch.  The first batch is the last solve's
  // count plus one, so a step whose count does not grow polls once.
  const int NB = 4;
  int it = 0;
  struct {
    CgScal s[3];
    int err;
  } h;
  int nb = c->cg_hint > 0 ? c->cg_hint + 1 : NB;
  for (;;) {
    for (int b = 0; b < nb && it < 100; b++, it++) HIPCHK(launch_cg_iter(c, it));
    nb = NB;
    HIPCHK(hipMemcpyAsync(h.s, c->scal, sizeof(h.s), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(&h.err, c->err, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (h.err) return check_err(c);
    if (!h.s[0].active && !h.s[1].active && !h.s[2].active) break;
    if (it >= 100) return set_error(WPIC_ECG, "********** stop at cgm after ite_max **********");
  }
  c->cg_hint = 0;
  for (int l = 0; l < 3; l++) {
    c->cg_iters[l] = h.s[l].ite;
    if (h.s[l].ite > c->cg_hint) c->cg_hint = h.s[l].ite;
  }
  return 0;
}

// field__fdtd_i after the deposit: current BCs, RHS, CG, E, BCs, update
int field_after_deposit(wpic_ctx* c) {
  Timed t(c, "field");
  HIPCHK(launch_curre_bc(c));
  {
    Timed tc(c, "cg");
    HIPCHK(launch_cg_init(c));
    int rc = cg_loop_impl(c);
    if (rc) return rc;
    HIPCHK(launch_cg_store(c));
  }
  HIPCHK(launch_dfield_bc(c));
  HIPCHK(launch_e_update(c));
  HIPCHK(launch_dfield_bc(c));
  HIPCHK(launch_uf_add(c));
  return 0;
}

int deposit(wpic_ctx* c, bool push, bool fused) {
  if (c->rcnt_fill) HIPCHK(hipMemsetAsync(c->rcnt, 0, sizeof(int) * (size_t)c->g.nsp * c->g.ny * c->g.nx, c->st));
  if (c->mixed) {
    HIPCHK(launch_zero(c, c->uj, (size_t)3 * c->g.fw * c->g.fh));
    Timed t(c, "push_deposit");
    HIPCHK(launch_push_deposit_mixed(c));
    return 0;
  }
  const bool parity = c->mode == WPIC_MODE_PARITY;
  if (parity) {
    int rc = ensure_pjbuf(c);
    if (rc) return rc;
  } else {
    HIPCHK(launch_zero(c, c->uj, (size_t)3 * c->g.fw * c->g.fh));
  }
  {
    Timed t(c, "push_deposit");
    HIPCHK(launch_push_deposit(c, push, true, fused));
  }
  if (parity) HIPCHK(launch_gather_pj(c));
  return 0;
}

// bc__particle_y's exchange on B (rows B -> rows N, permutation R); with
// `lists_only` the row counts and up/down lists alone, for the fused fast
// sort (k_rowsort_fast builds each new row from them)
int exchange(wpic_ctx* c, bool ywrap, bool lists_only = false) {
  Timed t(c, "exchange");
  HIPCHK(launch_rowscan(c, ywrap));
  HIPCHK(launch_rowcount(c));
  if (!lists_only) HIPCHK(launch_rowbuild(c));
  c->h_rows_ok[c->rn] = false;
  c->exch_pending = true;
  return 0;
}

int mixed_only_step() {
  return set_error(WPIC_EINVAL,
                   "the mixed-precision form runs the fused loop (wpic_step2d); single routines need an FP64 mode");
}

// sort__bucket: gather B through R into A, by cell within rows (fast: the
// fused step's one-pass exchange and sort, see k_rowsort_fast)
int sort_into_a(wpic_ctx* c, bool fast = false) {
  {
    Timed t(c, "sort");
    HIPCHK(fast ? launch_rowsort_fast(c) : launch_rowsort(c));
  }
  const int old_a = c->ra;
  c->ra = c->rn;
  c->rn = old_a;
  if (c->rb == c->ra) c->rb = pick_free(c->ra, c->rn);
  // the sorted state is gp's buffer seen through the new permutation; the
  // old sorted buffer becomes the next gp
  c->aperm = c->perm_out ? c->perm_out : c->permbuf;
  std::swap(c->A, c->B);
  std::swap(c->MA, c->MB);
  set_rows(c);
  c->b_valid = false;
  c->b_accl = false;
  c->b_aligned = false;
  c->exch_pending = false;
  return 0;
}

}  // namespace

namespace wpic {

// launches made while it lives go to the context's second stream
struct OnSecondStream {
  wpic_ctx* c;
  hipStream_t saved;
  explicit OnSecondStream(wpic_ctx* c_) : c(c_), saved(c_->st) { c->st = c->st2; }
  ~OnSecondStream() { c->st = saved; }
};

int step_once(wpic_ctx* c) {
  // B shares A's slots and rows; the exchange writes rows N
  c->rb = c->ra;
  c->rn = (c->ra + 1) % 3;
  set_rows(c);
  // the mixed push counts each row's particles per x cell for the sort (in the
  // FP64 fast push the same counting costs more than the sort's counting pass)
  c->rcnt_fill = c->mixed;
  int rc = deposit(c, true, true);
  if (rc) {
    c->rcnt_fill = false;
    return rc;
  }
  c->b_valid = true;
  c->b_accl = false;
  c->b_aligned = true;
  // the exchange and sort read the push's codes and rows, the field solve its
  // current and the fields: independent, so the exchange and sort run on the
  // second stream beside the field solve (whose CG polls only the first);
  // the first stream waits for them before the next step
  HIPCHK(hipEventRecord(c->ev_push, c->st));
  HIPCHK(hipStreamWaitEvent(c->st2, c->ev_push, 0));
  {
    OnSecondStream on2(c);
    // the parity mode keeps the reference's hole filling and bucket order; the
    // fast and mixed modes put each cell's particles in an order of their own
    const bool fast = c->mode != WPIC_MODE_PARITY;
    rc = exchange(c, false, fast);
    if (rc) {
      c->rcnt_fill = false;
      return rc;
    }
    rc = sort_into_a(c, fast);
    c->rcnt_fill = false;
    if (rc) return rc;
    HIPCHK(hipEventRecord(c->ev_sort, c->st));
  }
  rc = field_after_deposit(c);
  HIPCHK(hipStreamWaitEvent(c->st, c->ev_sort, 0));
  if (rc) return rc;
  c->h_rows_ok[c->ra] = false;
  return 0;
}

int cg_loop(wpic_ctx* c) { return cg_loop_impl(c); }
int check_err_ctx(wpic_ctx* c) { return check_err(c); }

}  // namespace wpic

extern "C" {

const char* wpic_last_error(void) { return g_err.c_str(); }

int wpic_create2d(wpic_ctx** out, const wpic_params2d* p) {
  if (!out || !p) return set_error(WPIC_EINVAL, "null argument");
  *out = nullptr;
  if (p->ndim != 6) return set_error(WPIC_EINVAL, "Error: ndim must be 6");
  if (p->nsp < 1 || p->nsp > WPIC_MAXSP) return set_error(WPIC_EINVAL, "nsp out of range");
  const int nx = p->nxge - p->nxgs + 1, ny = p->nyge - p->nygs + 1;
  if (nx < 4 || ny < 4) return set_error(WPIC_EINVAL, "grid must be at least 4x4 cells");
  if (nx > 4096) return set_error(WPIC_EINVAL, "nx above 4096 cells per row is not supported");
  if (p->nxgs < 2 || p->nygs < 2) return set_error(WPIC_EINVAL, "nxgs and nygs must be >= 2 (positive coordinates)");
  if (p->delx != 1.0)
    return set_error(WPIC_EINVAL, "delx must be 1: sort__bucket bins by int(x) (2d/common/sort.f90:60)");
  if (!(p->c * p->delt <= p->delx))
    return set_error(WPIC_EINVAL, "c*delt must not exceed delx (particles move less than one cell per step)");
  if (p->bc != WPIC_BC_PERIODIC && p->bc != WPIC_BC_WALLX && p->bc != WPIC_BC_SHOCK)
    return set_error(WPIC_EINVAL, "unknown boundary kind");
  if (p->bc == WPIC_BC_SHOCK && p->mode == WPIC_MODE_MIXED)
    return set_error(WPIC_EINVAL, "the shock boundaries run in the FP64 modes");
  if (p->mode != WPIC_MODE_FAST && p->mode != WPIC_MODE_PARITY && p->mode != WPIC_MODE_MIXED)
    return set_error(WPIC_EINVAL, "bad mode");
  if (p->capacity <= 0 || (long long)p->capacity * p->nsp >= (1ll << 31))
    return set_error(WPIC_EINVAL, "capacity must be positive and nsp*capacity below 2^31");

  wpic_ctx* c = new wpic_ctx();
  std::memset((void*)&c->g, 0, sizeof(c->g));
  c->P = *p;
  c->mode = p->mode;
  c->pusher = WPIC_PUSHER_BORIS;
  c->mixed = p->mode == WPIC_MODE_MIXED;
  Grid2D& g = c->g;
  g.nx = nx;
  g.ny = ny;
  g.nxgs = p->nxgs;
  g.nxge = p->nxge;
  g.nxa = p->nxge;
  g.nygs = p->nygs;
  g.nyge = p->nyge;
  g.nys = p->nygs;
  g.nye = p->nyge;
  g.fw = nx + 4;
  g.fh = ny + 4;
  g.bc = p->bc;
  g.nsp = p->nsp;
  g.delx = p->delx;
  g.delt = p->delt;
  g.c = p->c;
  g.gfac = p->gfac;
  g.d_delx = 1. / p->delx;
  g.d_delt = 1. / p->delt;
  const double pi = 3.141592653589793;
  // field.f90:53-57
  g.f1 = p->c * p->delt / p->delx;
  g.f2 = p->gfac * g.f1 * g.f1;
  g.f3 = 4.0 * pi * p->delx / p->c;
  const double t = p->delx / (p->c * p->delt * p->gfac);
  g.f4 = 4.0 + t * t;
  g.f5 = t * t;
  g.pi4dt = 4. * pi * p->delt;
  g.spanx = rd_mul((double)nx, p->delx);
  g.spany = rd_mul((double)ny, p->delx);
  g.wall_lo = (double)(2.0f * (float)(p->nxgs + 1)) * p->delx;
  g.wall_hi = (double)(2.0f * (float)(p->nxge - 1)) * p->delx;
  g.u0 = 0.0;
  g.xend = p->nxge * p->delx;
  for (int s = 0; s < WPIC_MAXSP; s++) {
    g.q[s] = p->q[s];
    g.r[s] = p->r[s];
  }
  g.ncap = p->capacity;
  g.cg_err = 1e-6;

  hipError_t e = hipGetDevice(&c->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_push, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_sort, hipEventDisableTiming);
  const size_t N = (size_t)g.ncap * g.nsp;
  const int nr = g.nsp * ny;
  const size_t pl = (size_t)g.fw * g.fh, nc = (size_t)nx * ny;
  if (c->mixed) {
    for (PartsM* m : {&c->MA, &c->MB}) {
      if (e == hipSuccess) e = dalloc(&m->ix, N);
      for (float** q : {&m->ox, &m->oy, &m->ux, &m->uy, &m->uz})
        if (e == hipSuccess) e = dalloc(q, N);
    }
  } else {
    double** dp[] = {&c->A.x, &c->A.y, &c->A.ux, &c->A.uy, &c->A.uz, &c->B.x, &c->B.y, &c->B.ux, &c->B.uy, &c->B.uz};
    for (auto q : dp)
      if (e == hipSuccess) e = dalloc(q, N);
  }
  for (int k = 0; k < 2 && e == hipSuccess; k++) e = dalloc(&c->id[k], N);
  int** ip[] = {&c->code, &c->R, &c->holes, &c->ulist, &c->dlist, &c->permbuf};
  for (auto q : ip)
    if (e == hipSuccess) e = dalloc(q, N);
  for (int k = 0; k < 3 && e == hipSuccess; k++) {
    e = dalloc(&c->rows[k], (size_t)2 * nr);
    if (e == hipSuccess) e = hipMemset(c->rows[k], 0, sizeof(int) * 2 * nr);
  }
  if (e == hipSuccess) e = dalloc(&c->cumcnt, (size_t)nr * (nx + 1));
  if (e == hipSuccess) e = hipMemset(c->cumcnt, 0, sizeof(int) * (size_t)nr * (nx + 1));
  if (e == hipSuccess) e = dalloc(&c->rowcnt, (size_t)4 * nr);
  if (e == hipSuccess) e = dalloc(&c->rcnt, (size_t)nr * c->g.nx);
  if (e == hipSuccess) e = dalloc(&c->aosbase, (size_t)nr);
  if (e == hipSuccess) e = dalloc(&c->uf, 6 * pl);
  if (e == hipSuccess) e = dalloc(&c->df, 6 * pl);
  if (e == hipSuccess) e = dalloc(&c->uj, 3 * pl);
  double** cg[] = {&c->phi, &c->pa, &c->pb, &c->rr, &c->bb, &c->ap};
  for (auto q : cg)
    if (e == hipSuccess) e = dalloc(q, 3 * nc);
  if (e == hipSuccess) e = dalloc(&c->part, (size_t)3 * NPB * 2);
  if (e == hipSuccess) e = dalloc(&c->scal, 3);
  if (e == hipSuccess) e = dalloc(&c->err, 1);
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
    wpic_destroy(c);
    return hip_fail(e, "wpic_create2d");
  }
  c->ra = 0;
  c->rb = 1;
  c->rn = 2;
  set_rows(c);
  c->A.id = c->MA.id = c->id[0];
  c->B.id = c->MB.id = c->id[1];
  for (int k = 0; k < 3; k++) c->h_rows_ok[k] = false;
  *out = c;
  return 0;
}

int wpic_destroy(wpic_ctx* c) {
  if (!c) return 0;
  if (c->st) (void)hipStreamSynchronize(c->st);
  adj_free(c);
  for (auto& t : c->timers) {
    if (t.a) (void)hipEventDestroy(t.a);
    if (t.b) (void)hipEventDestroy(t.b);
  }
  void* ps[] = {c->A.x, c->A.y, c->A.ux, c->A.uy, c->A.uz, c->B.x, c->B.y, c->B.ux, c->B.uy, c->B.uz,
                c->id[0], c->id[1], c->code, c->R, c->holes, c->ulist, c->dlist, c->rows[0], c->rows[1],
                c->rows[2], c->cumcnt, c->rowcnt, c->rcnt, c->aosbase, c->uf, c->df, c->uj, c->pjbuf, c->phi, c->pa,
                c->pb, c->rr, c->bb, c->ap, c->part, c->scal, c->err, c->staging, c->gbuf, c->red, c->mom,
                c->MA.ix, c->MA.ox, c->MA.oy, c->MA.ux, c->MA.uy, c->MA.uz,
                c->MB.ix, c->MB.ox, c->MB.oy, c->MB.ux, c->MB.uy, c->MB.uz, c->permbuf};
  for (void* q : ps)
    if (q) (void)hipFree(q);
  if (c->ev_push) (void)hipEventDestroy(c->ev_push);
  if (c->ev_sort) (void)hipEventDestroy(c->ev_sort);
  if (c->st2) (void)hipStreamDestroy(c->st2);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
  return 0;
}

int wpic_set_mode(wpic_ctx* c, int mode) {
  if (!c || (mode != WPIC_MODE_FAST && mode != WPIC_MODE_PARITY && mode != WPIC_MODE_MIXED))
    return set_error(WPIC_EINVAL, "bad mode");
  if (c->mixed != (mode == WPIC_MODE_MIXED))
    return set_error(WPIC_EINVAL, "the mixed-precision form is fixed when the context is created");
  c->mode = mode;
  return 0;
}

int wpic_sync(wpic_ctx* c) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  HIPCHK(hipStreamSynchronize(c->st));
  return check_err(c);
}

// ---------------------------------------------------------------------------
// transfers
// ---------------------------------------------------------------------------
static int rows_from_np2(wpic_ctx* c, const int* np2, std::vector<int>& hrows, std::vector<long long>& aosb,
                         long long& total) {
  const int ny = c->g.ny, nsp = c->g.nsp, nr = nsp * ny;
  hrows.assign(2 * nr, 0);
  aosb.assign(nr, 0);
  total = 0;
  for (int s = 0; s < nsp; s++) {
    long long off = 0;
    for (int r = 0; r < ny; r++) {
      const int rw = s * ny + r;
      if (np2[rw] < 0) return set_error(WPIC_EINVAL, "negative np2");
      hrows[rw] = (int)(s * c->g.ncap + off);
      hrows[nr + rw] = np2[rw];
      aosb[rw] = total;
      off += np2[rw];
      total += np2[rw];
    }
    if (off > c->g.ncap) return set_error(WPIC_EMEM, "memory over (np2 > np): species exceeds the device capacity");
  }
  return 0;
}

static int upload_rows(wpic_ctx* c, int which, const double* up, const int* np2, int np, int kind) {
  std::vector<int> hrows;
  std::vector<long long> aosb;
  long long total = 0;
  int rc = rows_from_np2(c, np2, hrows, aosb, total);
  if (rc) return rc;
  const int nr = c->g.nsp * c->g.ny;
  for (int rw = 0; rw < nr; rw++)
    if (np2[rw] > np) return set_error(WPIC_EINVAL, "np2 exceeds the row capacity np");
  rc = ensure_staging(c, (size_t)total * 6);
  if (rc) return rc;
  std::vector<double> h((size_t)total * 6);
  for (int rw = 0; rw < nr; rw++)
    std::memcpy(&h[(size_t)aosb[rw] * 6], up + (size_t)6 * np * rw, sizeof(double) * 6 * (size_t)np2[rw]);
  HIPCHK(hipMemcpyAsync(c->staging, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->rows[which], hrows.data(), sizeof(int) * 2 * nr, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->aosbase, aosb.data(), sizeof(long long) * nr, hipMemcpyHostToDevice, c->st));
  if (c->mixed)
    HIPCHK(launch_up_to_mixed(c, c->staging));
  else
    HIPCHK(launch_up_to_soa(c, c->staging, total, kind));
  HIPCHK(hipStreamSynchronize(c->st));
  c->h_rows[which] = hrows;
  c->h_rows_ok[which] = true;
  return 0;
}

static int download_rows(wpic_ctx* c, int which, int kind, double* up, int* np2, int np) {
  int rc = fetch_rows(c, which);
  if (rc) return rc;
  const int nr = c->g.nsp * c->g.ny;
  const std::vector<int>& hr = c->h_rows[which];
  std::vector<long long> aosb(nr);
  long long total = 0;
  for (int rw = 0; rw < nr; rw++) {
    aosb[rw] = total;
    total += hr[nr + rw];
    if (hr[nr + rw] > np) return set_error(WPIC_EMEM, "memory over (np2 > np)");
  }
  rc = ensure_staging(c, (size_t)total * 6);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(c->aosbase, aosb.data(), sizeof(long long) * nr, hipMemcpyHostToDevice, c->st));
  if (c->mixed)
    HIPCHK(launch_mixed_to_aos(c, c->staging));
  else
    HIPCHK(launch_soa_to_aos(c, c->staging, kind));
  std::vector<double> h((size_t)total * 6);
  HIPCHK(hipMemcpyAsync(h.data(), c->staging, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  for (int rw = 0; rw < nr; rw++) {
    std::memcpy(up + (size_t)6 * np * rw, &h[(size_t)aosb[rw] * 6], sizeof(double) * 6 * (size_t)hr[nr + rw]);
    if (np2) np2[rw] = hr[nr + rw];
  }
  return 0;
}

int wpic_upload_up2d(wpic_ctx* c, const double* up, const int* np2, const int* cumcnt, int np) {
  if (!c || !up || !np2 || !cumcnt) return set_error(WPIC_EINVAL, "null argument");
  {
    // the push and deposit walk cumcnt ranges: require sort__bucket's invariant
    const int nx = c->g.nx, nr = c->g.nsp * c->g.ny;
    for (int rw = 0; rw < nr; rw++) {
      const int* cr = cumcnt + (size_t)rw * (nx + 1);
      if (cr[0] != 0 || cr[nx] != np2[rw])
        return set_error(WPIC_EINVAL, "cumcnt must run from 0 to np2 in every row (a sorted up)");
      for (int i = 0; i < nx; i++)
        if (cr[i + 1] < cr[i]) return set_error(WPIC_EINVAL, "cumcnt must be non-decreasing");
    }
  }
  c->aperm = nullptr;
  int rc = upload_rows(c, c->ra, up, np2, np, 0);
  if (rc) return rc;
  const size_t ncc = (size_t)c->g.nsp * c->g.ny * (c->g.nx + 1);
  HIPCHK(hipMemcpyAsync(c->cumcnt, cumcnt, sizeof(int) * ncc, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  c->b_valid = false;
  c->b_accl = false;
  c->b_aligned = false;
  c->exch_pending = false;
  return 0;
}

int wpic_download_up2d(wpic_ctx* c, double* up, int* np2, int* cumcnt, int np) {
  if (!c || !up) return set_error(WPIC_EINVAL, "null argument");
  int rc = download_rows(c, c->ra, 0, up, np2, np);
  if (rc) return rc;
  if (cumcnt) {
    const size_t ncc = (size_t)c->g.nsp * c->g.ny * (c->g.nx + 1);
    HIPCHK(hipMemcpy(cumcnt, c->cumcnt, sizeof(int) * ncc, hipMemcpyDeviceToHost));
  }
  return 0;
}

int wpic_upload_gp2d(wpic_ctx* c, const double* gp, const int* np2, int np) {
  if (c && c->mixed) return mixed_only_step();
  if (!c || !gp || !np2) return set_error(WPIC_EINVAL, "null argument");
  if (c->rb == c->ra) c->rb = pick_free(c->ra, c->rn);
  set_rows(c);
  int rc = upload_rows(c, c->rb, gp, np2, np, 1);
  if (rc) return rc;
  c->b_valid = true;
  c->b_accl = false;
  c->exch_pending = false;
  // gp sits slot for slot on up when the rows agree (a field__fdtd_i input);
  // both use compact row bases, so equal counts mean equal bases
  rc = fetch_rows(c, c->ra);
  if (rc) return rc;
  c->b_aligned = c->h_rows[c->ra] == c->h_rows[c->rb];
  return 0;
}

int wpic_download_gp2d(wpic_ctx* c, double* gp, int* np2, int np) {
  if (c && c->mixed) return mixed_only_step();
  if (!c || !gp) return set_error(WPIC_EINVAL, "null argument");
  if (!c->b_valid) return set_error(WPIC_EINVAL, "no gp on the device (call particle__solv first)");
  if (c->exch_pending) return download_rows(c, c->rn, 1, gp, np2, np);
  return download_rows(c, c->rb, 2, gp, np2, np);
}

int wpic_upload_fields2d(wpic_ctx* c, const double* uf) {
  if (!c || !uf) return set_error(WPIC_EINVAL, "null argument");
  const size_t n = (size_t)6 * c->g.fw * c->g.fh;
  int rc = ensure_staging(c, n);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(c->staging, uf, sizeof(double) * n, hipMemcpyHostToDevice, c->st));
  HIPCHK(launch_planes_from_aos(c, c->staging, c->uf, 6));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

static int download_planes(wpic_ctx* c, const double* planes, double* out, int ncomp) {
  const size_t n = (size_t)ncomp * c->g.fw * c->g.fh;
  int rc = ensure_staging(c, n);
  if (rc) return rc;
  HIPCHK(launch_planes_to_aos(c, planes, c->staging, ncomp));
  HIPCHK(hipMemcpyAsync(out, c->staging, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int wpic_download_fields2d(wpic_ctx* c, double* uf) {
  if (!c || !uf) return set_error(WPIC_EINVAL, "null argument");
  return download_planes(c, c->uf, uf, 6);
}

int wpic_download_current2d(wpic_ctx* c, double* uj) {
  if (!c || !uj) return set_error(WPIC_EINVAL, "null argument");
  return download_planes(c, c->uj, uj, 3);
}

int wpic_download_dfield2d(wpic_ctx* c, double* df) {
  if (!c || !df) return set_error(WPIC_EINVAL, "null argument");
  return download_planes(c, c->df, df, 6);
}

int wpic_upload_dfield2d(wpic_ctx* c, const double* df) {
  if (!c || !df) return set_error(WPIC_EINVAL, "null argument");
  const size_t n = (size_t)6 * c->g.fw * c->g.fh;
  int rc = ensure_staging(c, n);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(c->staging, df, sizeof(double) * n, hipMemcpyHostToDevice, c->st));
  HIPCHK(launch_planes_from_aos(c, c->staging, c->df, 6));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

// ---------------------------------------------------------------------------
// reference entry points
// ---------------------------------------------------------------------------
int wpic_particle_solv2d(wpic_ctx* c) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (c->mixed) return mixed_only_step();
  c->rb = c->ra;
  if (c->rn == c->ra) c->rn = pick_free(c->ra, c->rb);
  set_rows(c);
  {
    Timed t(c, "push");
    HIPCHK(launch_push_deposit(c, true, false, false));
  }
  c->b_valid = true;
  c->b_accl = false;
  c->b_aligned = true;
  c->exch_pending = false;
  return check_err(c);
}

int wpic_particle_solv_vay2d(wpic_ctx* c) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (c->mixed) return mixed_only_step();
  const int saved = c->pusher;
  c->pusher = WPIC_PUSHER_VAY;
  const int rc = wpic_particle_solv2d(c);
  c->pusher = saved;
  return rc;
}

int wpic_set_pusher2d(wpic_ctx* c, int pusher) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (pusher != WPIC_PUSHER_BORIS && pusher != WPIC_PUSHER_VAY) return set_error(WPIC_EINVAL, "bad pusher");
  if (c->mixed && pusher != WPIC_PUSHER_BORIS)
    return set_error(WPIC_EINVAL, "the Vay kick runs in the FP64 modes (the mixed form pushes with Buneman-Boris)");
  c->pusher = pusher;
  return 0;
}

int wpic_field_fdtd_i2d(wpic_ctx* c) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (c->mixed) return mixed_only_step();
  if (!c->b_valid || !c->b_aligned || c->exch_pending)
    return set_error(WPIC_EINVAL, "field__fdtd_i needs gp slot-aligned with up (call particle__solv first)");
  int rc = deposit(c, false, false);
  if (rc) return rc;
  rc = field_after_deposit(c);
  if (rc) return rc;
  return check_err(c);
}

int wpic_bc_particle_x2d(wpic_ctx* c) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (c->mixed) return mixed_only_step();
  if (!c->b_valid || c->exch_pending) return set_error(WPIC_EINVAL, "bc__particle_x needs gp before the y exchange");
  c->b_accl = false;
  HIPCHK(launch_wrapx(c, false));
  return check_err(c);
}

namespace {
// xend = nxe dx + v0 dt with v0 = u0 / gamma0 (boundary_shock.f90:269)
void set_u0(wpic_ctx* c, double u0) {
  const double delx = c->g.delx, cc = c->g.c;
  c->g.u0 = u0;
  c->g.xend = c->g.nxa * delx + u0 / sqrt(1 + (u0 * u0) / (cc * cc)) * c->g.delt;
}

// the active right edge nxe: the injection edge and the shock's particle_x
// wall move with it (boundary_shock.f90:87-88,269)
void set_nxa(wpic_ctx* c, int nxe) {
  c->g.nxa = nxe;
  c->g.wall_hi = (double)(2.0f * (float)(nxe - 1)) * c->g.delx;
  set_u0(c, c->g.u0);
}
}  // namespace

int wpic_set_injection2d(wpic_ctx* c, double u0) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (c->g.bc != WPIC_BC_SHOCK) return set_error(WPIC_EINVAL, "the injection edge needs WPIC_BC_SHOCK");
  if (!std::isfinite(u0)) return set_error(WPIC_EINVAL, "u0 must be finite");
  set_u0(c, u0);
  return 0;
}

int wpic_ic_shock2d(wpic_ctx* c, unsigned long long seed, int n0, double v0, double v_thi, double v_the,
                    double l_damp, double b0, double theta_bn, double phi_bn) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (c->mixed) return mixed_only_step();
  if (c->g.bc != WPIC_BC_SHOCK) return set_error(WPIC_EINVAL, "the shock state needs WPIC_BC_SHOCK");
  const Grid2D& g = c->g;
  if (g.nsp != 2) return set_error(WPIC_EINVAL, "the shock state has two species (ions, electrons)");
  if (!(std::fabs(v0) < g.c) || !(l_damp > 0.0)) return set_error(WPIC_EINVAL, "need |v0| < c and l_damp_ini > 0");
  // np2 = n0 (nxe - nxs - 1) per row (app.f90:334)
  const long long np2l = (long long)n0 * (g.nxa - g.nxgs - 1);
  if (n0 <= 0 || np2l * g.ny > g.ncap) return set_error(WPIC_EMEM, "Error: Too large number of particles");
  const int np2 = (int)np2l;
  const int nr = g.nsp * g.ny;
  std::vector<int> hrows(2 * nr);
  for (int s = 0; s < g.nsp; s++)
    for (int r = 0; r < g.ny; r++) {
      hrows[s * g.ny + r] = (int)(s * g.ncap + (long long)r * np2);
      hrows[nr + s * g.ny + r] = np2;
    }
  if (c->rb == c->ra) c->rb = pick_free(c->ra, c->rn);
  set_rows(c);
  HIPCHK(hipMemcpyAsync(c->rows[c->rb], hrows.data(), sizeof(int) * 2 * nr, hipMemcpyHostToDevice, c->st));
  const double sd[2] = {v_thi, v_the};
  const double b[3] = {b0 * cos(theta_bn), b0 * sin(theta_bn) * cos(phi_bn), b0 * sin(theta_bn) * sin(phi_bn)};
  HIPCHK(launch_ic_shock(c, seed, np2, v0, sd, l_damp, b));
  const size_t pl = (size_t)g.fw * g.fh;
  HIPCHK(launch_zero(c, c->df, 6 * pl));
  c->h_rows[c->rb] = hrows;
  c->h_rows_ok[c->rb] = true;
  c->b_valid = true;
  c->b_accl = false;
  c->b_aligned = false;
  c->exch_pending = false;
  return wpic_sort_bucket2d(c);
}

namespace {
// checks shared by inject and relocate
int shock_rows_check(wpic_ctx* c, const char* what, double v0, double l_damp) {
  if (c->mixed) return mixed_only_step();
  if (c->g.bc != WPIC_BC_SHOCK) return set_error(WPIC_EINVAL, std::string(what) + " needs WPIC_BC_SHOCK");
  if (c->g.nsp != 2) return set_error(WPIC_EINVAL, std::string(what) + " loads two species (ions, electrons)");
  if (c->exch_pending)
    return set_error(WPIC_EINVAL, std::string(what) + " needs the sorted state (after sort__bucket)");
  if (!(std::fabs(v0) < c->g.c) || !(l_damp > 0.0)) return set_error(WPIC_EINVAL, "need |v0| < c and l_damp_ini > 0");
  return 0;
}

// inject and relocate: row r of each species grows by nl[r] particles at its
// end, the rows re-laid at their new packed bases (k_inject)
int grow_rows(wpic_ctx* c, unsigned long long seed, int it, const int* nl, double v0, double v_thi, double v_the,
              double l_damp, double b0, double theta_bn, double phi_bn, bool reloc) {
  const Grid2D& g = c->g;
  const int nr = g.nsp * g.ny;
  if (fetch_rows(c, c->ra)) return WPIC_EHIP;
  const std::vector<int> old = c->h_rows[c->ra];
  std::vector<int> hrows(2 * nr);
  std::vector<long long> idoff(nr);
  for (int s = 0; s < g.nsp; s++) {
    long long total = 0, base = 0, ninj = 0;
    for (int r = 0; r < g.ny; r++) total += old[nr + s * g.ny + r];
    for (int r = 0; r < g.ny; r++) {
      if (nl[r] < 0) return set_error(WPIC_EINVAL, "negative injection count");
      const int rw = s * g.ny + r;
      const long long n = (long long)old[nr + rw] + nl[r];
      hrows[rw] = (int)(s * g.ncap + base);
      hrows[nr + rw] = (int)n;
      // pid = ii - np2 + ncinj_grid(j) + nptotal(isp) (app.f90:833)
      idoff[rw] = ninj + total;
      base += n;
      ninj += nl[r];
      if (base > g.ncap) return set_error(WPIC_EMEM, "injection beyond the particle capacity");
    }
  }
  // scratch: idoff (8-byte aligned at the start of rowcnt), then nl
  long long* d_idoff = reinterpret_cast<long long*>(c->rowcnt);
  int* d_nl = c->rowcnt + 2 * nr;
  HIPCHK(hipMemcpyAsync(d_idoff, idoff.data(), sizeof(long long) * nr, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(d_nl, nl, sizeof(int) * g.ny, hipMemcpyHostToDevice, c->st));
  if (c->rb == c->ra) c->rb = pick_free(c->ra, c->rn);
  HIPCHK(hipMemcpyAsync(c->rows[c->rb], hrows.data(), sizeof(int) * 2 * nr, hipMemcpyHostToDevice, c->st));
  const double sd[2] = {v_thi, v_the};
  const double by = b0 * sin(theta_bn) * cos(phi_bn), bz = b0 * sin(theta_bn) * sin(phi_bn);
  HIPCHK(launch_inject(c, seed, it, c->rows[c->rb], d_nl, d_idoff, v0, sd, l_damp, by, bz, reloc));
  HIPCHK(hipStreamSynchronize(c->st));
  // B holds the new sorted state in identity order
  std::swap(c->A, c->B);
  std::swap(c->MA, c->MB);
  std::swap(c->ra, c->rb);
  c->aperm = nullptr;
  c->h_rows[c->ra] = hrows;
  c->h_rows_ok[c->ra] = true;
  c->h_rows_ok[c->rb] = false;
  set_rows(c);
  c->b_valid = false;
  c->b_accl = false;
  c->b_aligned = false;
  c->exch_pending = false;
  return check_err(c);
}
}  // namespace

int wpic_inject2d(wpic_ctx* c, unsigned long long seed, int it, const int* nl, double v0, double v_thi, double v_the,
                  double l_damp, double b0, double theta_bn, double phi_bn) {
  if (!c || !nl) return set_error(WPIC_EINVAL, "null argument");
  int rc = shock_rows_check(c, "inject", v0, l_damp);
  if (rc) return rc;
  return grow_rows(c, seed, it, nl, v0, v_thi, v_the, l_damp, b0, theta_bn, phi_bn, false);
}

int wpic_relocate2d(wpic_ctx* c, unsigned long long seed, int it, int n0, double v0, double v_thi, double v_the,
                    double l_damp, double b0, double theta_bn, double phi_bn) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  int rc = shock_rows_check(c, "relocate", v0, l_damp);
  if (rc) return rc;
  if (n0 < 0) return set_error(WPIC_EINVAL, "negative n0");
  if (c->g.nxa == c->g.nxge) return 0;  // app.f90:622
  const int old = c->g.nxa;
  set_nxa(c, old + 1);
  const std::vector<int> nl(c->g.ny, n0);
  rc = grow_rows(c, seed, it, nl.data(), v0, v_thi, v_the, l_damp, b0, theta_bn, phi_bn, true);
  if (rc) set_nxa(c, old);
  return rc;
}

int wpic_set_nxe2d(wpic_ctx* c, int nxe) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  const Grid2D& g = c->g;
  if (nxe == g.nxa) return 0;
  if (g.bc != WPIC_BC_SHOCK) return set_error(WPIC_EINVAL, "only the shock's box has an active edge below nxge");
  if (nxe < g.nxgs + 3 || nxe > g.nxge)
    return set_error(WPIC_EINVAL, "nxe must lie in [nxgs+3, nxge], got " + std::to_string(nxe));
  set_nxa(c, nxe);
  return 0;
}

int wpic_get_nxe2d(wpic_ctx* c, int* nxe) {
  if (!c || !nxe) return set_error(WPIC_EINVAL, "null argument");
  *nxe = c->g.nxa;
  return 0;
}

int wpic_bc_injection2d(wpic_ctx* c, double u0) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (c->mixed) return mixed_only_step();
  if (c->g.bc != WPIC_BC_SHOCK) return set_error(WPIC_EINVAL, "bc__injection needs WPIC_BC_SHOCK");
  if (!std::isfinite(u0)) return set_error(WPIC_EINVAL, "u0 must be finite");
  if (!c->b_valid || c->exch_pending) return set_error(WPIC_EINVAL, "bc__injection needs gp before the y exchange");
  set_u0(c, u0);
  c->b_accl = false;
  HIPCHK(launch_wrapx(c, true));
  return check_err(c);
}

int wpic_bc_particle_y2d(wpic_ctx* c) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (c->mixed) return mixed_only_step();
  if (!c->b_valid || c->exch_pending) return set_error(WPIC_EINVAL, "bc__particle_y needs gp before the y exchange");
  if (c->rn == c->rb || c->rn == c->ra) c->rn = pick_free(c->ra, c->rb);
  set_rows(c);
  int rc = exchange(c, true);
  if (rc) return rc;
  return check_err(c);
}

int wpic_sort_bucket2d(wpic_ctx* c) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (c->mixed) return mixed_only_step();
  if (!c->b_valid) return set_error(WPIC_EINVAL, "sort__bucket needs gp on the device");
  if (!c->exch_pending) {
    // the restart path: sort gp as it stands
    if (c->rn == c->rb || c->rn == c->ra) c->rn = pick_free(c->ra, c->rb);
    set_rows(c);
    const int nr = c->g.nsp * c->g.ny;
    HIPCHK(hipMemcpyAsync(c->rows[c->rn], c->rows[c->rb], sizeof(int) * 2 * nr, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(launch_codes_identity(c));
    c->exch_pending = true;
  }
  int rc = sort_into_a(c);
  if (rc) return rc;
  c->h_rows_ok[c->ra] = false;
  return check_err(c);
}

int wpic_cg_iters2d(wpic_ctx* c, int iters[3]) {
  if (!c || !iters) return set_error(WPIC_EINVAL, "null argument");
  for (int l = 0; l < 3; l++) iters[l] = c->cg_iters[l];
  return 0;
}

int wpic_step2d(wpic_ctx* c, int nsteps) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  for (int it = 0; it < nsteps; it++) {
    int rc = step_once(c);
    if (rc) return rc;
  }
  return check_err(c);
}

// ---------------------------------------------------------------------------
// diagnostics
// ---------------------------------------------------------------------------
int wpic_energy2d(wpic_ctx* c, double out[3]) {
  if (!c || !out) return set_error(WPIC_EINVAL, "null argument");
  if (c->mixed)
    HIPCHK(launch_energy_mixed(c, c->red));
  else
    HIPCHK(launch_energy(c, c->red));
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

int wpic_gauss2d(wpic_ctx* c, double* G) {
  if (!c || !G) return set_error(WPIC_EINVAL, "null argument");
  const size_t pl = (size_t)c->g.fw * c->g.fh, nc = (size_t)c->g.nx * c->g.ny;
  if (!c->gbuf) HIPCHK(dalloc(&c->gbuf, pl + nc));
  HIPCHK(launch_gauss(c, c->gbuf + pl, c->gbuf));
  HIPCHK(hipMemcpyAsync(G, c->gbuf + pl, sizeof(double) * nc, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int wpic_counts2d(wpic_ctx* c, long long* counts) {
  if (!c || !counts) return set_error(WPIC_EINVAL, "null argument");
  int rc = fetch_rows(c, c->ra);
  if (rc) return rc;
  const int ny = c->g.ny, nr = c->g.nsp * ny;
  for (int s = 0; s < c->g.nsp; s++) {
    long long t = 0;
    for (int r = 0; r < ny; r++) t += c->h_rows[c->ra][nr + s * ny + r];
    counts[s] = t;
  }
  return 0;
}

int wpic_ic_beams2d(wpic_ctx* c, unsigned long long seed, int n_ppc, double udrift, double uth_e, double uth_i,
                    double b0) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  const Grid2D& g = c->g;
  if (n_ppc <= 0 || (long long)n_ppc * g.nx * g.ny > g.ncap)
    return set_error(WPIC_EMEM, "Error: Too large number of particles");
  const int nr = g.nsp * g.ny;
  std::vector<int> hrows(2 * nr);
  std::vector<int> cc((size_t)nr * (g.nx + 1));
  for (int s = 0; s < g.nsp; s++)
    for (int r = 0; r < g.ny; r++) {
      const int rw = s * g.ny + r;
      hrows[rw] = (int)(s * g.ncap + (long long)r * n_ppc * g.nx);
      hrows[nr + rw] = n_ppc * g.nx;
      for (int i = 0; i <= g.nx; i++) cc[(size_t)rw * (g.nx + 1) + i] = i * n_ppc;
    }
  HIPCHK(hipMemcpyAsync(c->rows[c->ra], hrows.data(), sizeof(int) * 2 * nr, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->cumcnt, cc.data(), sizeof(int) * cc.size(), hipMemcpyHostToDevice, c->st));
  c->aperm = nullptr;
  if (c->mixed)
    HIPCHK(launch_ic_beams_mixed(c, seed, n_ppc, udrift, uth_e, uth_i));
  else
    HIPCHK(launch_ic_beams(c, seed, n_ppc, udrift, uth_e, uth_i));
  const size_t pl = (size_t)g.fw * g.fh;
  HIPCHK(launch_zero(c, c->uf, 6 * pl));
  HIPCHK(launch_fill(c, c->uf + 2 * pl, pl, b0));
  HIPCHK(launch_zero(c, c->df, 6 * pl));
  HIPCHK(hipStreamSynchronize(c->st));
  c->h_rows[c->ra] = hrows;
  c->h_rows_ok[c->ra] = true;
  c->b_valid = false;
  c->b_accl = false;
  c->b_aligned = false;
  c->exch_pending = false;
  return 0;
}

int wpic_ic_weibel2d(wpic_ctx* c, unsigned long long seed, int n0, double v_thi, double v_the, double t_ani,
                     double b0) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (c->mixed) return mixed_only_step();
  const Grid2D& g = c->g;
  if (n0 <= 0 || (long long)n0 * g.nx * g.ny > g.ncap) return set_error(WPIC_EMEM, "Error: Too large number of particles");
  if (g.nsp != 2) return set_error(WPIC_EINVAL, "the Weibel state has two species (ions, electrons)");
  const int nr = g.nsp * g.ny;
  std::vector<int> hrows(2 * nr);
  std::vector<int> cc((size_t)nr * (g.nx + 1));
  for (int s = 0; s < g.nsp; s++)
    for (int r = 0; r < g.ny; r++) {
      const int rw = s * g.ny + r;
      hrows[rw] = (int)(s * g.ncap + (long long)r * n0 * g.nx);
      hrows[nr + rw] = n0 * g.nx;
      // x on the lattice: n0 per cell in order (app.f90:338-351)
      for (int i = 0; i <= g.nx; i++) cc[(size_t)rw * (g.nx + 1) + i] = i * n0;
    }
  HIPCHK(hipMemcpyAsync(c->rows[c->ra], hrows.data(), sizeof(int) * 2 * nr, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->cumcnt, cc.data(), sizeof(int) * cc.size(), hipMemcpyHostToDevice, c->st));
  c->aperm = nullptr;
  const double sd[WPIC_MAXSP] = {v_thi, v_the, 0.0, 0.0};
  HIPCHK(launch_ic_weibel(c, seed, n0, sd, t_ani));
  const size_t pl = (size_t)g.fw * g.fh;
  HIPCHK(launch_zero(c, c->uf, 6 * pl));
  HIPCHK(launch_fill(c, c->uf + 2 * pl, pl, b0));
  HIPCHK(launch_zero(c, c->df, 6 * pl));
  HIPCHK(hipStreamSynchronize(c->st));
  c->h_rows[c->ra] = hrows;
  c->h_rows_ok[c->ra] = true;
  c->b_valid = false;
  c->b_accl = false;
  c->b_aligned = false;
  c->exch_pending = false;
  return check_err(c);
}

int wpic_ic_harris2d(wpic_ctx* c, unsigned long long seed, int nbg, int ncs, double lcs, double x0, double y0,
                     double b0, double e1, double rtemp, double vti, double vte) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (c->mixed) return mixed_only_step();
  const Grid2D& g = c->g;
  if (g.nsp != 2) return set_error(WPIC_EINVAL, "the Harris state has two species (ions, electrons)");
  if (!(lcs > 0.0)) return set_error(WPIC_EINVAL, "lcs must be positive");
  // particles per row and background particles per row (app.f90:304, 383)
  const double np2d = nbg * (g.nxge - g.nxgs) + ncs * 2 * lcs;
  if (!(np2d >= 0.0) || np2d > 2.0e9) return set_error(WPIC_EMEM, "Error: Too large number of particles");
  const int np2 = (int)np2d;
  if ((long long)np2 * g.ny > g.ncap) return set_error(WPIC_EMEM, "Error: Too large number of particles");
  const int ibg = nbg * (g.nxge - g.nxgs - 2);
  if (ibg > np2 || ibg < 0) return set_error(WPIC_EINVAL, "background particles exceed the particles per row");
  // the Box-Muller factor 1/sqrt(2.) with a default-real literal (app.f90:380-381)
  const double s2 = (double)std::sqrt(2.0f);
  const double hp[11] = {x0, y0, lcs, b0, e1, (double)ncs, (double)nbg,
                         1.0 / ((1.0 + rtemp) * g.q[0]), rtemp / ((1.0 + rtemp) * g.q[1]), vti / s2, vte / s2};
  const int nr = g.nsp * g.ny;
  std::vector<int> hrows(2 * nr);
  for (int s = 0; s < g.nsp; s++)
    for (int r = 0; r < g.ny; r++) {
      hrows[s * g.ny + r] = (int)(s * g.ncap + (long long)r * np2);
      hrows[nr + s * g.ny + r] = np2;
    }
  if (c->rb == c->ra) c->rb = pick_free(c->ra, c->rn);
  set_rows(c);
  HIPCHK(hipMemcpyAsync(c->rows[c->rb], hrows.data(), sizeof(int) * 2 * nr, hipMemcpyHostToDevice, c->st));
  HIPCHK(launch_ic_harris(c, seed, np2, ibg, hp));
  const size_t pl = (size_t)g.fw * g.fh;
  HIPCHK(launch_zero(c, c->df, 6 * pl));
  c->h_rows[c->rb] = hrows;
  c->h_rows_ok[c->rb] = true;
  c->b_valid = true;
  c->b_accl = false;
  c->b_aligned = false;
  c->exch_pending = false;
  // sort__bucket(gp, up, ...) of the loaded particles (app.f90:346-348)
  return wpic_sort_bucket2d(c);
}

int wpic_set_timing(wpic_ctx* c, int on) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  fold_timers(c);
  c->tsum.clear();
  c->timing = on != 0;
  return 0;
}

int wpic_timing(wpic_ctx* c, const char* name, double* ms, long long* count) {
  if (!c || !name) return set_error(WPIC_EINVAL, "null argument");
  HIPCHK(hipStreamSynchronize(c->st));
  fold_timers(c);
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
// moments: mom_calc__accl / mom_calc__nvt (2d/common/mom_calc.f90:48-249) and
// boundary_periodic__mom / boundary_reconnection__mom
// ---------------------------------------------------------------------------
static int ensure_mom(wpic_ctx* c) {
  if (c->mom) return 0;
  HIPCHK(dalloc(&c->mom, (size_t)7 * (c->g.nx + 2) * (c->g.ny + 2) * c->g.nsp));
  HIPCHK(launch_zero(c, c->mom, (size_t)7 * (c->g.nx + 2) * (c->g.ny + 2) * c->g.nsp));
  return 0;
}

int wpic_mom_accl2d(wpic_ctx* c) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (c->mixed) return mixed_only_step();
  c->rb = c->ra;
  if (c->rn == c->ra) c->rn = pick_free(c->ra, c->rb);
  set_rows(c);
  {
    Timed t(c, "mom_accl");
    HIPCHK(launch_mom_accl(c));
  }
  c->b_valid = true;
  c->b_accl = true;  // gp holds up's positions: sorted by cell like up
  c->b_aligned = true;
  c->exch_pending = false;
  return check_err(c);
}

int wpic_mom_nvt2d(wpic_ctx* c) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (c->mixed) return mixed_only_step();
  if (!c->b_valid || c->exch_pending)
    return set_error(WPIC_EINVAL, "mom_calc__nvt needs gp on the device (call mom_calc__accl first)");
  int rc = ensure_mom(c);
  if (rc) return rc;
  set_rows(c);
  const bool parity = c->mode == WPIC_MODE_PARITY;
  // the tile deposit needs gp in up's rows (particles whose node lies off
  // their tile go to HBM directly); the parity gather's cell-range shortcut
  // needs gp sorted by cell like up, as mom_calc__accl leaves it
  int kind;
  if (parity)
    kind = c->b_accl && c->b_aligned ? MOM_PARITY_ALIGNED : MOM_PARITY_ANY;
  else
    kind = c->b_aligned ? MOM_TILE_GP : MOM_ANY;
  {
    Timed t(c, "mom_nvt");
    HIPCHK(launch_mom_deposit(c, kind));
  }
  return check_err(c);
}

int wpic_bc_mom2d(wpic_ctx* c) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  int rc = ensure_mom(c);
  if (rc) return rc;
  HIPCHK(launch_mom_bc(c));
  return check_err(c);
}

int wpic_moments2d(wpic_ctx* c) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  int rc = ensure_mom(c);
  if (rc) return rc;
  if (c->mixed) {
    // the fused output sequence in the mixed form; the state is not touched
    {
      Timed t(c, "moments");
      set_rows(c);
      HIPCHK(launch_mom_mixed(c));
      HIPCHK(launch_mom_bc(c));
    }
    return check_err(c);
  }
  {
    Timed t(c, "moments");
    if (c->mode == WPIC_MODE_PARITY) {
      c->rb = c->ra;
      if (c->rn == c->ra) c->rn = pick_free(c->ra, c->rb);
      set_rows(c);
      HIPCHK(launch_mom_accl(c));
      HIPCHK(launch_mom_deposit(c, MOM_PARITY_ALIGNED));
    } else {
      set_rows(c);
      HIPCHK(launch_mom_deposit(c, MOM_TILE_KICK));
    }
    HIPCHK(launch_mom_bc(c));
  }
  // the reference's sequence leaves gp overwritten by mom_calc__accl
  c->b_valid = false;
  c->b_accl = false;
  c->b_aligned = false;
  c->exch_pending = false;
  return check_err(c);
}

int wpic_download_mom2d(wpic_ctx* c, double* mom) {
  if (!c || !mom) return set_error(WPIC_EINVAL, "null argument");
  int rc = ensure_mom(c);
  if (rc) return rc;
  const size_t n = (size_t)7 * (c->g.nx + 2) * (c->g.ny + 2) * c->g.nsp;
  HIPCHK(hipMemcpyAsync(mom, c->mom, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int wpic_upload_mom2d(wpic_ctx* c, const double* mom) {
  if (!c || !mom) return set_error(WPIC_EINVAL, "null argument");
  int rc = ensure_mom(c);
  if (rc) return rc;
  const size_t n = (size_t)7 * (c->g.nx + 2) * (c->g.ny + 2) * c->g.nsp;
  HIPCHK(hipMemcpyAsync(c->mom, mom, sizeof(double) * n, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}
