// wpic_adj2d.hip -- the C-ABI of include/wpic_adj.h: a recorded forward run
// with device checkpoints, and its adjoint by segment replay.  No reference
// counterpart (the reference has no adjoint).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/wpic_adj.h"
#include "wpic_internal.h"

using namespace wpic;

#define HIPCHK(expr)                                  \
  do {                                                \
    hipError_t e_ = (expr);                           \
    if (e_ != hipSuccess) return hip_fail(e_, #expr); \
  } while (0)

// device storage of a recorded run: a checkpoint of the sorted state every
// `every` steps, one segment's per-step start states for the replay, and the
// adjoint arrays
struct AdjState {
  int nsteps = 0, every = 0, nseg = 0;
  size_t N = 0;  // particle slots (capacity x species)
  int nr = 0, ncc = 0;
  size_t pl = 0, nc = 0;
  std::vector<double*> ck_part, ck_uf, ck_df;
  std::vector<uint64_t*> ck_id;
  std::vector<int*> ck_rows, ck_cumcnt;
  std::vector<double*> st_part, st_uf;
  std::vector<int*> st_rows, st_perm;
  double *lamA = nullptr, *lamA2 = nullptr, *lamB = nullptr;
  double *lamF = nullptr, *lamF2 = nullptr, *lamJ = nullptr, *lamT = nullptr;
  double* red = nullptr;
  bool recorded = false, done = false;
};

namespace wpic {

void adj_free(wpic_ctx* c) {
  AdjState* a = c->adj;
  if (!a) return;
  auto fr = [](void* p) {
    if (p) (void)hipFree(p);
  };
  for (auto p : a->ck_part) fr(p);
  for (auto p : a->ck_uf) fr(p);
  for (auto p : a->ck_df) fr(p);
  for (auto p : a->ck_id) fr(p);
  for (auto p : a->ck_rows) fr(p);
  for (auto p : a->ck_cumcnt) fr(p);
  for (auto p : a->st_part) fr(p);
  for (auto p : a->st_uf) fr(p);
  for (auto p : a->st_rows) fr(p);
  for (auto p : a->st_perm) fr(p);
  for (void* p : {(void*)a->lamA, (void*)a->lamA2, (void*)a->lamB, (void*)a->lamF, (void*)a->lamF2, (void*)a->lamJ,
                  (void*)a->lamT, (void*)a->red})
    fr(p);
  delete a;
  c->adj = nullptr;
}

}  // namespace wpic

namespace {

template <class T>
hipError_t dal(T** p, size_t n) {
  *p = nullptr;
  return hipMalloc((void**)p, (n ? n : 1) * sizeof(T));
}

Parts parts_of(double* base, uint64_t* id, size_t N) {
  return Parts{base, base + N, base + 2 * N, base + 3 * N, base + 4 * N, id};
}

int copy_d(wpic_ctx* c, void* dst, const void* src, size_t bytes) {
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->st));
  return 0;
}

// the sorted state (A, its ids, rows, cumcnt) and the field arrays
int save_state(wpic_ctx* c, double* part, uint64_t* id, int* rows, int* cumcnt, double* uf, double* df) {
  AdjState& a = *c->adj;
  const size_t N = a.N;
  // the sorted particles and their ids in slot order (A is read through its
  // permutation)
  if (part) HIPCHK(launch_gather_rows(c, c->A, c->aperm, parts_of(part, id, N), c->rows[c->ra]));
  if (rows && copy_d(c, rows, c->rows[c->ra], sizeof(int) * 2 * a.nr)) return WPIC_EHIP;
  if (cumcnt && copy_d(c, cumcnt, c->cumcnt, sizeof(int) * a.ncc)) return WPIC_EHIP;
  if (uf && copy_d(c, uf, c->uf, sizeof(double) * 6 * a.pl)) return WPIC_EHIP;
  if (df && copy_d(c, df, c->df, sizeof(double) * 6 * a.pl)) return WPIC_EHIP;
  return 0;
}

int restore_state(wpic_ctx* c, int s) {
  AdjState& a = *c->adj;
  const size_t N = a.N;
  double* dst[5] = {c->A.x, c->A.y, c->A.ux, c->A.uy, c->A.uz};
  for (int k = 0; k < 5; k++)
    if (copy_d(c, dst[k], a.ck_part[s] + k * N, sizeof(double) * N)) return WPIC_EHIP;
  c->aperm = nullptr;
  if (copy_d(c, c->A.id, a.ck_id[s], sizeof(uint64_t) * N)) return WPIC_EHIP;
  if (copy_d(c, c->rows[c->ra], a.ck_rows[s], sizeof(int) * 2 * a.nr)) return WPIC_EHIP;
  if (copy_d(c, c->cumcnt, a.ck_cumcnt[s], sizeof(int) * a.ncc)) return WPIC_EHIP;
  if (copy_d(c, c->uf, a.ck_uf[s], sizeof(double) * 6 * a.pl)) return WPIC_EHIP;
  if (copy_d(c, c->df, a.ck_df[s], sizeof(double) * 6 * a.pl)) return WPIC_EHIP;
  c->h_rows_ok[c->ra] = false;
  c->b_valid = false;
  c->b_accl = false;
  c->b_aligned = false;
  c->exch_pending = false;
  return 0;
}

int alloc_adj(wpic_ctx* c, int nsteps, int every) {
  adj_free(c);
  c->adj = new AdjState();
  AdjState& a = *c->adj;
  const Grid2D& g = c->g;
  a.nsteps = nsteps;
  a.every = every;
  a.nseg = (nsteps + every - 1) / every;
  a.N = (size_t)g.ncap * g.nsp;
  a.nr = g.nsp * g.ny;
  a.ncc = a.nr * (g.nx + 1);
  a.pl = (size_t)g.fw * g.fh;
  a.nc = (size_t)g.nx * g.ny;
  hipError_t e = hipSuccess;
  a.ck_part.assign(a.nseg, nullptr);
  a.ck_uf.assign(a.nseg, nullptr);
  a.ck_df.assign(a.nseg, nullptr);
  a.ck_id.assign(a.nseg, nullptr);
  a.ck_rows.assign(a.nseg, nullptr);
  a.ck_cumcnt.assign(a.nseg, nullptr);
  for (int s = 0; s < a.nseg && e == hipSuccess; s++) {
    e = dal(&a.ck_part[s], 5 * a.N);
    if (e == hipSuccess) e = dal(&a.ck_uf[s], 6 * a.pl);
    if (e == hipSuccess) e = dal(&a.ck_df[s], 6 * a.pl);
    if (e == hipSuccess) e = dal(&a.ck_id[s], a.N);
    if (e == hipSuccess) e = dal(&a.ck_rows[s], (size_t)2 * a.nr);
    if (e == hipSuccess) e = dal(&a.ck_cumcnt[s], (size_t)a.ncc);
  }
  const int K = every < nsteps ? every : nsteps;
  a.st_part.assign(K, nullptr);
  a.st_uf.assign(K, nullptr);
  a.st_perm.assign(K, nullptr);
  a.st_rows.assign(K + 1, nullptr);
  for (int t = 0; t < K && e == hipSuccess; t++) {
    e = dal(&a.st_part[t], 5 * a.N);
    if (e == hipSuccess) e = dal(&a.st_uf[t], 6 * a.pl);
    if (e == hipSuccess) e = dal(&a.st_perm[t], a.N);
  }
  for (int t = 0; t <= K && e == hipSuccess; t++) e = dal(&a.st_rows[t], (size_t)2 * a.nr);
  for (double** p : {&a.lamA, &a.lamA2, &a.lamB})
    if (e == hipSuccess) e = dal(p, 5 * a.N);
  for (double** p : {&a.lamF, &a.lamF2, &a.lamT})
    if (e == hipSuccess) e = dal(p, 6 * a.nc);
  if (e == hipSuccess) e = dal(&a.lamJ, 3 * a.nc);
  if (e == hipSuccess) e = dal(&a.red, (size_t)256 * 3);
  if (e != hipSuccess) {
    adj_free(c);
    return set_error(WPIC_EMEM, std::string("device memory for the recorded run: ") + hipGetErrorString(e));
  }
  return 0;
}

int w_b(wpic_ctx* c, double* wb) {
  double out[3];
  int rc = wpic_energy2d(c, out);
  if (rc) return rc;
  *wb = out[2];
  return 0;
}

}  // namespace

extern "C" {

int wpic_set_cg_tol2d(wpic_ctx* c, double err) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (!(err > 0.0 && err < 1.0)) return set_error(WPIC_EINVAL, "the CG tolerance must lie in (0, 1)");
  c->g.cg_err = err;
  return 0;
}

int wpic_perturb_bz2d(wpic_ctx* c, double amp, int mode) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  HIPCHK(launch_perturb_bz(c, amp, mode));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int wpic_adj_forward2d(wpic_ctx* c, int nsteps, int every, double* wb) {
  if (!c || !wb) return set_error(WPIC_EINVAL, "null argument");
  if (c->mixed || c->g.bc != WPIC_BC_PERIODIC)
    return set_error(WPIC_EINVAL, "the adjoint covers the periodic FP64 loop");
  if (c->pusher != WPIC_PUSHER_BORIS) return set_error(WPIC_EINVAL, "the adjoint covers the Buneman-Boris kick");
  if (nsteps < 1 || every < 1) return set_error(WPIC_EINVAL, "nsteps and every must be positive");
  int rc = alloc_adj(c, nsteps, every);
  if (rc) return rc;
  AdjState& a = *c->adj;
  for (int it = 0; it < nsteps; it++) {
    if (it % every == 0) {
      const int s = it / every;
      rc = save_state(c, a.ck_part[s], a.ck_id[s], a.ck_rows[s], a.ck_cumcnt[s], a.ck_uf[s], a.ck_df[s]);
      if (rc) return rc;
    }
    rc = step_once(c);
    if (rc) return rc;
  }
  rc = check_err_ctx(c);
  if (rc) return rc;
  rc = w_b(c, wb);
  if (rc) return rc;
  a.recorded = true;
  a.done = false;
  return 0;
}

int wpic_adj_backward2d(wpic_ctx* c, double seed) {
  if (!c) return set_error(WPIC_EINVAL, "null context");
  if (!c->adj || !c->adj->recorded) return set_error(WPIC_EINVAL, "no recorded run (call wpic_adj_forward2d)");
  AdjState& a = *c->adj;
  const size_t N = a.N;
  // the final state is the current one: dW_B/duf seeds the field adjoint
  HIPCHK(launch_adj_seed(c, a.lamF, seed));
  HIPCHK(hipMemsetAsync(a.lamA, 0, sizeof(double) * 5 * N, c->st));
  for (int s = a.nseg - 1; s >= 0; s--) {
    const int t0 = s * a.every;
    const int K = (a.nsteps - t0) < a.every ? (a.nsteps - t0) : a.every;
    int rc = restore_state(c, s);
    if (rc) return rc;
    for (int t = 0; t < K; t++) {
      rc = save_state(c, a.st_part[t], nullptr, a.st_rows[t], nullptr, a.st_uf[t], nullptr);
      if (rc) return rc;
      c->perm_out = a.st_perm[t];
      rc = step_once(c);
      c->perm_out = nullptr;
      if (rc) return rc;
    }
    rc = save_state(c, nullptr, nullptr, a.st_rows[K], nullptr, nullptr, nullptr);
    if (rc) return rc;
    for (int t = K - 1; t >= 0; t--) {
      HIPCHK(hipMemsetAsync(a.lamB, 0, sizeof(double) * 5 * N, c->st));
      HIPCHK(launch_adj_unsort(c, a.st_rows[t + 1], a.st_perm[t], a.lamA, a.lamB, N));
      HIPCHK(launch_adj_field1(c, a.lamF, a.lamF2, a.lamJ));
      HIPCHK(launch_cg_start(c));
      rc = cg_loop(c);
      if (rc) return rc;
      HIPCHK(launch_adj_field2(c, a.lamF2, a.lamJ));
      HIPCHK(hipMemsetAsync(a.lamT, 0, sizeof(double) * 6 * a.nc, c->st));
      HIPCHK(launch_adj_particles(c, parts_of(a.st_part[t], nullptr, N), a.st_rows[t], a.st_uf[t], a.lamB, a.lamJ, a.lamA2,
                                  a.lamT, N));
      HIPCHK(launch_adj_t2f(c, a.lamT, a.lamF2));
      std::swap(a.lamF, a.lamF2);
      std::swap(a.lamA, a.lamA2);
    }
  }
  // the context now holds the state at the end of the first segment's replay
  a.done = true;
  return check_err_ctx(c);
}

int wpic_adj_param_grads2d(wpic_ctx* c, double udrift, double uth_e, int bz_mode, double out[3]) {
  if (!c || !out) return set_error(WPIC_EINVAL, "null argument");
  if (!c->adj || !c->adj->done) return set_error(WPIC_EINVAL, "no adjoint (call wpic_adj_backward2d)");
  AdjState& a = *c->adj;
  // the seeded state's slot layout: n_ppc particles per cell in every row
  std::vector<int> rows(2 * a.nr);
  HIPCHK(hipMemcpy(rows.data(), a.ck_rows[0], sizeof(int) * rows.size(), hipMemcpyDeviceToHost));
  const int n_ppc = rows[a.nr] / c->g.nx;
  if (n_ppc < 1 || rows[a.nr] != n_ppc * c->g.nx)
    return set_error(WPIC_EINVAL, "the recorded run did not start from a wpic_ic_beams2d state");
  HIPCHK(launch_adj_params(c, parts_of(a.ck_part[0], a.ck_id[0], a.N), a.ck_id[0], a.ck_rows[0], a.lamA, a.N, a.lamF, udrift,
                           uth_e, bz_mode, n_ppc, a.red));
  std::vector<double> h(256 * 3);
  HIPCHK(hipMemcpyAsync(h.data(), a.red, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  out[0] = out[1] = out[2] = 0.0;
  for (int b = 0; b < 256; b++)
    for (int k = 0; k < 3; k++) out[k] += h[b * 3 + k];
  return 0;
}

int wpic_adj_download2d(wpic_ctx* c, double* lam_up, double* lam_uf, int np) {
  if (!c || !lam_up || !lam_uf) return set_error(WPIC_EINVAL, "null argument");
  if (!c->adj || !c->adj->done) return set_error(WPIC_EINVAL, "no adjoint (call wpic_adj_backward2d)");
  AdjState& a = *c->adj;
  const Grid2D& g = c->g;
  std::vector<int> rows(2 * a.nr);
  HIPCHK(hipMemcpy(rows.data(), a.ck_rows[0], sizeof(int) * rows.size(), hipMemcpyDeviceToHost));
  std::vector<double> lam(5 * a.N);
  HIPCHK(hipMemcpy(lam.data(), a.lamA, sizeof(double) * lam.size(), hipMemcpyDeviceToHost));
  for (int rw = 0; rw < a.nr; rw++) {
    const int base = rows[rw], n = rows[a.nr + rw];
    if (n > np) return set_error(WPIC_EMEM, "memory over (np2 > np)");
    double* dst = lam_up + (size_t)6 * np * rw;
    std::memset(dst, 0, sizeof(double) * 6 * (size_t)np);
    for (int k = 0; k < n; k++)
      for (int q = 0; q < 5; q++) dst[(size_t)6 * k + q] = lam[q * a.N + base + k];
  }
  double* planes = nullptr;
  HIPCHK(hipMalloc((void**)&planes, sizeof(double) * 6 * a.pl));
  hipError_t e = launch_adj_lamF_to_planes(c, a.lamF, planes);
  std::vector<double> h(6 * a.pl);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), planes, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->st);
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  (void)hipFree(planes);
  if (e != hipSuccess) return hip_fail(e, "wpic_adj_download2d");
  // planes [c][node] -> uf(c, node) layout
  const size_t npl = a.pl;
  for (size_t k = 0; k < npl; k++)
    for (int q = 0; q < 6; q++) lam_uf[k * 6 + q] = h[q * npl + k];
  (void)g;
  return 0;
}

}  // extern "C"
